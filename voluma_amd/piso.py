"""One PISO time step of the reference's incompressible solver on the GPU.

``PisoFlow`` mirrors ``IncompressibleFlow`` (reference
src/fvm/solvers/IncompressibleFlow.{h,cpp}) for laminar flow with upwind convection,
Euler or backward time stepping, fixedValue / zeroGradient boundary conditions and no
body force or turbulence: each ``step(dt)`` stores the old levels, assembles the
momentum matrix per component (ddt + div(φ, U) − laplacian(ν, U), :183-185), solves the
momentum predictor (:212-227), then runs the pressure correctors (:242-362): A and H of
the momentum matrix, rAU and HbyA, rAUf, the skew-corrected phiHbyA, the ddtPhiCorr
coupling, the pressure equation with its reference cell, the face flux update and the
velocity correction. Every operator runs on the GPU through the library's entry points
(FvMatrix assembly, relax/A/H/flux, the fvc operators, the corrector steps and the
linear solves); the host only evaluates the boundary values of U and p
(``correctBoundaryConditions``) and keeps the field arrays between the calls.

Meshes come from ``voluma_amd.mesh``; a 2-D mesh (an empty patch) solves the two
in-plane velocity components, the reference's ``solvedComponents``. Results are those
of the reference's ``step``: with reproducible sums (``Context.set_reproducible``) the
two use the same partition-independent solvers and agree bit for bit.
"""
from __future__ import annotations

import numpy as np

import voluma_amd as va


class PisoFlow:
    def __init__(self, ctx, mesh, nu, u_bc, p_bc, *, n_corr=2, n_non_orth=0, momentum_predictor=True, ddt="Euler",
                 grad_p=va.GRAD_GAUSS_LINEAR, nonorth_corr=True, p_ref_cell=0, p_ref_value=0.0, ddt_phi_coeff=0.0,
                 u_controls=None, p_controls=None, p_final_controls=None):
        """u_bc / p_bc: per patch of ``mesh.computational()`` (empty patches excluded),
        None for zeroGradient or the fixed value (a 3-vector for U, a scalar for p).
        Solver controls default to PimpleControls' (IncompressibleFlow.h:31-33)."""
        if ddt not in ("Euler", "backward"):
            raise ValueError("ddt must be 'Euler' or 'backward'")
        comp = mesh.computational()
        f = comp["faces"]
        g = mesh.geometry()
        self.ctx, self.nu = ctx, float(nu)
        self.n, self.ni = mesh.n_cells, mesh.n_internal
        self.nF = len(f)
        self.nb = self.nF - self.ni
        own = np.ascontiguousarray(mesh.owner[f]).astype(np.int32)
        self.own_b = own[self.ni:]
        pick = lambda k: np.ascontiguousarray(g[k][f])
        self.V = np.ascontiguousarray(g["V"])
        self.Sf, self.magSf, self.w, self.d = pick("Sf"), pick("magSf"), pick("w"), pick("d")
        self.ndc, self.corr, self.skew = pick("ndc"), pick("corr"), pick("skew")
        self.skewed = bool(g["skewed"])
        self.empty_dir = comp["empty_dir"]
        self.solved = comp["solved_components"]
        self.lm = va.LduMesh(ctx, self.n, own[: self.ni], mesh.neighbour, self.magSf[: self.ni] * self.ndc[: self.ni])
        self.lm.set_boundary(self.own_b)
        self.lm.set_geometry(self.V, self.Sf, self.magSf, self.w, self.d, self.skew, self.skewed)
        self.lm.set_empty_dir(self.empty_dir)
        self.UEqn = [va.LduMatrix(self.lm) for _ in range(3)]
        self.pEqn = va.LduMatrix(self.lm)
        # boundary conditions per kept boundary face
        nbp = len(comp["patch_names"])
        if len(u_bc) != nbp or len(p_bc) != nbp:
            raise ValueError(f"u_bc and p_bc need one entry per patch ({nbp})")
        self.u_fix = np.zeros(self.nb, np.int32)
        self.u_val = np.zeros((self.nb, 3))
        self.p_fix = np.zeros(self.nb, np.int32)
        self.p_val = np.zeros(self.nb)
        for q, (s, z) in enumerate(zip(comp["patch_start"], comp["patch_size"])):
            sl = slice(s - self.ni, s - self.ni + z)
            if u_bc[q] is not None:
                self.u_fix[sl] = 1
                self.u_val[sl] = np.asarray(u_bc[q], dtype=np.float64)
            if p_bc[q] is not None:
                self.p_fix[sl] = 1
                self.p_val[sl] = float(p_bc[q])
        self.n_corr, self.n_non_orth = int(n_corr), int(n_non_orth)
        self.momentum_predictor, self.ddt = bool(momentum_predictor), ddt
        self.grad_p, self.nonorth_corr = int(grad_p), bool(nonorth_corr)
        self.p_ref_cell, self.p_ref_value = int(p_ref_cell), float(p_ref_value)
        self.ddt_phi_coeff = float(ddt_phi_coeff)
        self.u_controls = u_controls or va.SolverControls("PBiCGStab", "DILU", 1e-8, 0.0, 1000)
        self.p_controls = p_controls or va.SolverControls("PCG", "GAMG", 1e-8, 0.01, 2000,
                                                          gamgSmoother=va.GAMG_GAUSS_SEIDEL)
        self.p_final_controls = p_final_controls or va.SolverControls("PCG", "GAMG", 1e-8, 0.0, 2000,
                                                                      gamgSmoother=va.GAMG_GAUSS_SEIDEL)
        # fields and time levels
        self.U = np.zeros((self.n, 3))
        self.p = np.zeros(self.n)
        self.phi = np.zeros(self.nF)
        self.U0 = self.U00 = None
        self.phi_old = self.phi_old_old = None
        self.n_old = 0
        self.dt, self.time_index = -1.0, 0
        self.perf = []

    # ---------------------------------------------------------------- boundary values
    def _U_b(self):
        """U on the boundary faces: the fixed value, or the owner cell's (zeroGradient)."""
        return np.where(self.u_fix[:, None] == 1, self.u_val, self.U[self.own_b])

    def _p_b(self):
        return np.where(self.p_fix == 1, self.p_val, self.p[self.own_b])

    def _grad_p(self):
        return self.lm.fvc_grad(self.grad_p, self.p, self.p_fix, self._p_b())

    def _nonorth(self, psi):
        if not self.nonorth_corr:
            return None
        return {"psi": psi, "w": self.w, "Sf": self.Sf, "corr": self.corr, "V": self.V}

    # ---------------------------------------------------------------- one time step
    def step(self, dt):
        """IncompressibleFlow::step(dt) (IncompressibleFlow.cpp:144-370), one outer
        iteration (PISO)."""
        lm, V, nu_f = self.lm, self.V, np.full(self.nF, self.nu)
        self.dt, dt0 = float(dt), (self.dt if self.time_index > 0 else -1.0)
        self.time_index += 1
        if self.n_old >= 1:
            self.U00 = self.U0
        self.U0 = self.U.copy()
        self.n_old = min(self.n_old + 1, 2)
        self.phi_old_old, self.phi_old = self.phi_old, self.phi.copy()
        U00 = self.U00 if self.n_old >= 2 else self.U0  # VolField::oldOld()
        need_ref = not bool(self.p_fix.any())
        self.perf = []

        # momentum matrix, one component at a time (:183-185), relax(1) leaves it (:210-211)
        k_u = va.ddt_coeffs(self.ddt, self.dt, dt0, self.n_old >= 2)
        for k, A in enumerate(self.UEqn):
            A.fvm_reset()
            A.fvm_ddt(+1, k_u, V, self.U0[:, k], U00[:, k])
            A.fvm_div_upwind(+1, self.phi, self.u_fix, self.u_val[:, k])
            A.fvm_laplacian(-1, self.magSf, self.ndc, self.u_fix, self.u_val[:, k], face_gamma=nu_f,
                            nonorth=self._nonorth(np.ascontiguousarray(self.U[:, k])))
        if self.momentum_predictor:  # M = UEqn; M −= −∇p; solve (:212-227)
            ngp = -self._grad_p()
            for k in self.solved:
                A = self.UEqn[k]
                b = A.fvm_finish_source(np.ascontiguousarray(ngp[:, k]), V)
                x = np.ascontiguousarray(self.U[:, k])
                self.perf.append(va.solve_ldu(A, b, x, self.u_controls, field=f"U{k}"))
                self.U[:, k] = x

        for corr in range(self.n_corr):
            final_corr = corr == self.n_corr - 1
            Acell = self.UEqn[0].fvm_A(V, 3)
            H = np.stack([self.UEqn[k].fvm_H(np.ascontiguousarray(self.U[:, k]), V, 3) for k in range(3)], axis=1)
            rAU, HbyA = lm.piso_rau_hbya(Acell, H)
            zb, zi = np.zeros((self.nb,)), np.zeros(self.nb, np.int32)
            if self.skewed:
                rAUf = lm.fvc_interpolate(rAU, grad=lm.fvc_grad(va.GRAD_LEAST_SQUARES, rAU, zi, zb))
            else:
                rAUf = lm.fvc_interpolate(rAU)
            Ub = self._U_b()
            gH = None
            if self.skewed:
                b_val = lm.piso_hbya_boundary(self.u_fix, Ub, rAU, self._grad_p())
                gH = lm.fvc_grad(va.GRAD_LEAST_SQUARES, HbyA, self.u_fix, b_val)
            phiHbyA = lm.fvc_flux(HbyA, gH, self.u_fix, Ub)
            if self.ddt_phi_coeff != 0.0:
                k_phi = va.ddt_coeffs(self.ddt, self.dt, dt0, self.n_old >= 2 and self.phi_old_old is not None)
                phiHbyA = lm.piso_ddt_phi_corr(self.ddt_phi_coeff, k_phi[1], k_phi[2], self.U0, U00, self.phi_old,
                                               self.phi_old_old, rAUf, phiHbyA)
            div_phi = lm.fvc_div(phiHbyA)
            for no in range(self.n_non_orth + 1):
                final_no = no == self.n_non_orth
                P = self.pEqn
                P.fvm_reset()
                P.fvm_laplacian(+1, self.magSf, self.ndc, self.p_fix, self.p_val, face_gamma=rAUf,
                                nonorth=self._nonorth(self.p))
                P.fvm_source(div_phi, V)  # pEqn −= divPhi
                if need_ref:
                    P.fvm_set_reference(self.p_ref_cell, self.p_ref_value)
                b = P.fvm_finish()
                x = np.ascontiguousarray(self.p.copy())
                ctl = self.p_final_controls if (final_corr and final_no) else self.p_controls
                self.perf.append(va.solve_ldu(P, b, x, ctl, field="p"))
                self.p = x
                if final_no:
                    self.phi = lm.piso_face_op(va.FACE_A_MINUS_B, phiHbyA, P.fvm_flux(self.p))
            self.U = lm.piso_correct_U(HbyA, rAU, self._grad_p(), -1, self.empty_dir)
        return self
