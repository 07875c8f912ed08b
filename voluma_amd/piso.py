"""Incompressible flow time steps of the reference's solver, resident on the GPU.

``PisoFlow`` is the Python face of the library's ``vol_flow`` (C ABI, ``flow.cpp``): the
reference's ``IncompressibleFlow`` (src/fvm/solvers/IncompressibleFlow.{h,cpp}) for
laminar flow with upwind convection, Euler or backward time stepping, fixedValue /
zeroGradient boundary conditions and no body force or turbulence. ``step(dt)`` stores
the old levels, assembles the momentum matrix per component (ddt + div(φ, U) −
laplacian(ν, U), :183-185), solves the momentum predictor (:212-227), then runs the
pressure correctors (:242-362): A and H of the momentum matrix, rAU and HbyA, rAUf, the
skew-corrected phiHbyA, the ddtPhiCorr coupling, the pressure equation with its
reference cell, the face flux update and the velocity correction. U, p, φ, their old
levels and the matrices stay on the GPU between and during steps; ``U``, ``p`` and
``phi`` copy the fields in or out.

Meshes come from ``voluma_amd.mesh``; cyclic patch pairs become internal faces
(``Mesh.computational``) and a 2-D mesh (an empty patch) solves the two in-plane
velocity components, the reference's ``solvedComponents``. With reproducible
sums (``Context.set_reproducible``) a step uses the reference's partition-independent
solvers and reproduces its fields bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

import voluma_amd as va
from voluma_amd import _Controls, _FlowOptions, _Perf, _check, _p, lib


class PisoFlow:
    def __init__(self, ctx, mesh, nu, u_bc, p_bc, *, n_corr=2, n_non_orth=0, momentum_predictor=True, ddt="Euler",
                 grad_p=va.GRAD_GAUSS_LINEAR, nonorth_corr=True, p_ref_cell=0, p_ref_value=0.0, ddt_phi_coeff=0.0,
                 u_controls=None, p_controls=None, p_final_controls=None):
        """u_bc / p_bc: per patch of ``mesh.computational()`` (empty patches excluded),
        None for zeroGradient or the fixed value (a 3-vector for U, a scalar for p).
        Solver controls default to PimpleControls' (IncompressibleFlow.h:31-33), the
        pressure's GAMG smoothed by Gauss–Seidel as the reference's."""
        if ddt not in ("Euler", "backward"):
            raise ValueError("ddt must be 'Euler' or 'backward'")
        comp = mesh.computational()
        g = comp["geometry"]
        self.ctx = ctx
        self.n, self.ni, self.nF = mesh.n_cells, comp["n_internal"], len(comp["faces"])
        nb = self.nF - self.ni
        own = comp["owner"]
        magSf, ndc = g["magSf"], g["ndc"]
        self.mesh = va.LduMesh(ctx, self.n, own[: self.ni], comp["neighbour"], magSf[: self.ni] * ndc[: self.ni])
        self.mesh.set_boundary(own[self.ni:])
        self.mesh.set_geometry(g["V"], g["Sf"], magSf, g["w"], g["d"], g["skew"], g["skewed"])
        self.mesh.set_empty_dir(comp["empty_dir"])
        nbp = len(comp["patch_names"])
        if len(u_bc) != nbp or len(p_bc) != nbp:
            raise ValueError(f"u_bc and p_bc need one entry per patch ({nbp})")
        u_fix, u_val = np.zeros(nb, np.int32), np.zeros((max(nb, 1), 3))
        p_fix, p_val = np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1))
        for q, (s, z) in enumerate(zip(comp["patch_start"], comp["patch_size"])):
            sl = slice(s - self.ni, s - self.ni + z)
            if u_bc[q] is not None:
                u_fix[sl], u_val[sl] = 1, np.asarray(u_bc[q], dtype=np.float64)
            if p_bc[q] is not None:
                p_fix[sl], p_val[sl] = 1, float(p_bc[q])
        o = _FlowOptions()
        lib().vol_flow_options_default(ctypes.byref(o))
        o.nu, o.nCorr, o.nNonOrthCorr = float(nu), int(n_corr), int(n_non_orth)
        o.momentumPredictor, o.ddtScheme = int(bool(momentum_predictor)), va.DDT_EULER if ddt == "Euler" else va.DDT_BACKWARD
        o.gradPScheme, o.laplacianNonOrthCorr = int(grad_p), int(bool(nonorth_corr))
        o.pRefCell, o.pRefValue, o.ddtPhiCoeff = int(p_ref_cell), float(p_ref_value), float(ddt_phi_coeff)
        self._keep = []
        for name, c in (("UControls", u_controls), ("pControls", p_controls), ("pFinalControls", p_final_controls)):
            if c is not None:
                cc = c._c()
                self._keep.append((c, cc))
                setattr(o, name, cc)
        self._opt = o
        h = ctypes.c_void_p()
        u_fix_c = np.zeros(max(nb, 1), np.int32)
        u_fix_c[:nb] = u_fix
        _check(lib().vol_flow_create(self.mesh._h, ctypes.byref(o), _p(ndc), _p(g["corr"]), _p(u_fix_c),
                                     _p(np.ascontiguousarray(u_val)), _p(p_fix), _p(p_val), va.VOL_HOST,
                                     ctypes.byref(h)))
        self._h = h
        self.perf = []

    def close(self):
        if getattr(self, "_h", None):
            lib().vol_flow_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    # ---------------------------------------------------------------- fields
    def _get(self):
        U, p, phi = np.zeros((self.n, 3)), np.zeros(self.n), np.zeros(max(self.nF, 1))
        _check(lib().vol_flow_get_fields(self._h, _p(U), _p(p), _p(phi), va.VOL_HOST))
        return U, p, phi[: self.nF]

    def _set(self, U=None, p=None, phi=None):
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        U, p, phi = c(U), c(p), c(phi)
        _check(lib().vol_flow_set_fields(self._h, va._q(U), va._q(p), va._q(phi), va.VOL_HOST))

    U = property(lambda self: self._get()[0], lambda self, v: self._set(U=v))
    p = property(lambda self: self._get()[1], lambda self, v: self._set(p=v))
    phi = property(lambda self: self._get()[2], lambda self, v: self._set(phi=v))

    # ---------------------------------------------------------------- one time step
    def step(self, dt):
        """IncompressibleFlow::step(dt) (IncompressibleFlow.cpp:144-370), one outer
        iteration (PISO), on the GPU. ``perf`` then holds the solves' performances."""
        cap = 64
        pf = (_Perf * cap)()
        cnt = ctypes.c_int()
        _check(lib().vol_flow_step(self._h, float(dt), ctypes.cast(pf, ctypes.c_void_p), cap, ctypes.byref(cnt)))
        self.perf = [va.SolverPerformance(pf[i].solver.decode(), "", pf[i].initialResidual, pf[i].finalResidual,
                                          pf[i].residualSum, pf[i].normFactor, pf[i].iterations,
                                          bool(pf[i].converged), pf[i].outerIterations)
                     for i in range(min(cnt.value, cap))]
        return self
