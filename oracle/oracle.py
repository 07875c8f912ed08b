"""ctypes binding of oracle/liboracle.so, the C restatement of the reference's
LDU linear-solve path (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the CPU baseline. The product
package (voluma_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "liboracle.so")
_LIB = None

MODE_PLAIN, MODE_REPRODUCIBLE, MODE_EXACT = 0, 1, 2
SMOOTHER_GS, SMOOTHER_JACOBI = 0, 1


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"{_PATH} not built (make -C oracle oracle)")
        lib = ctypes.CDLL(_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        lib.oracle_exact_sum.argtypes = [P, ctypes.c_long]
        lib.oracle_exact_sum.restype = D
        lib.oracle_plain_sum.argtypes = [P, ctypes.c_long]
        lib.oracle_plain_sum.restype = D
        lib.oracle_mesh_new.argtypes = [I, I, P, P, P, I]
        lib.oracle_mesh_new.restype = P
        lib.oracle_mesh_free.argtypes = [P]
        lib.oracle_gamg_nlevels.argtypes = [P]
        lib.oracle_gamg_nlevels.restype = I
        lib.oracle_gamg_level_sizes.argtypes = [P, I, P, P]
        lib.oracle_gamg_level.argtypes = [P, I, P, P, P, P, P, P]
        lib.oracle_amul.argtypes = [P] * 6
        lib.oracle_gamg_coeffs.argtypes = [P, P, P, P, I, P, P, P]
        lib.oracle_gamg_apply.argtypes = [P, P, P, P, P, P, I, D, I, I]
        lib.oracle_gamg_apply.restype = I
        lib.oracle_solve.argtypes = [P, P, P, P, I, P, P, ctypes.c_char_p, ctypes.c_char_p, D, D, I, I, I, I, I, D, I,
                                     I, P, P, P, P, ctypes.c_char_p, I]
        lib.oracle_solve.restype = I
        lib.oracle_laplacian.argtypes = [I, I, I] + [P] * 10
        lib.oracle_laplacian_nonorth.argtypes = [I, I] + [P] * 9
        lib.oracle_grad_gauss.argtypes = [I, I, I] + [P] * 7
        lib.oracle_div_upwind.argtypes = [I, I, I] + [P] * 10
        lib.oracle_ddt_euler.argtypes = [I, P, D, P, P, P]
        lib.oracle_fold.argtypes = [I, I, I] + [P] * 7
        lib.oracle_fvc_interpolate.argtypes = [I, I, I, P, P, P, P, I, P, P, P]
        lib.oracle_fvc_grad_ls.argtypes = [I, I, I, P, P, P, P, P, I, P, P, P, P]
        lib.oracle_fvc_grad_gauss_skew.argtypes = [I, I, I] + [P] * 12
        lib.oracle_fvc_div.argtypes = [I, I, I, P, P, P, P, P]
        lib.oracle_fvc_surface_integrate.argtypes = [I, I, I, P, P, P, I, P, P]
        lib.oracle_fvc_flux.argtypes = [I, I, I] + [P] * 10
        lib.oracle_piso_rau_hbya.argtypes = [I, P, P, P, P]
        lib.oracle_piso_hbya_boundary.argtypes = [I, I, P, P, P, P, P, P]
        lib.oracle_piso_face_op.argtypes = [I, I, P, P, P, P]
        lib.oracle_piso_ddt_phi_corr.argtypes = [I, P, P, P, P, D, D, D, P, P, P, P, P, P]
        lib.oracle_piso_constrain_pressure.argtypes = [I, I, I, P, P, P, P, P, P]
        lib.oracle_fvc_reconstruct.argtypes = [I, I, I, P, P, P, P, P, I, P]
        lib.oracle_piso_correct_U.argtypes = [I, P, P, P, I, I, P]
        lib.oracle_relax.argtypes = [I, I, I] + [P] * 6 + [D, P, P]
        lib.oracle_set_reference.argtypes = [I, ctypes.c_long, D, P, P]
        lib.oracle_A.argtypes = [I, I, I, P, P, P, P, I, P]
        lib.oracle_H.argtypes = [I, I, I] + [P] * 9 + [I, P]
        lib.oracle_flux.argtypes = [I, I] + [P] * 9
        _LIB = lib
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def exact_sum(v) -> float:
    v = _f64(v)
    return float(_lib().oracle_exact_sum(_p(v), v.size))


def plain_sum(v) -> float:
    v = _f64(v)
    return float(_lib().oracle_plain_sum(_p(v), v.size))


class OracleError(RuntimeError):
    pass


class LduMesh:
    """LDU addressing (+ optional GAMG weights) for the oracle's solver."""

    def __init__(self, n_cells, owner, neighbour, weights=None, merge_levels=2):
        self.n = int(n_cells)
        self.own = _i32(owner)
        self.nei = _i32(neighbour)
        self.nf = int(self.own.size)
        w = _f64(weights)
        self._h = _lib().oracle_mesh_new(self.n, self.nf, _p(self.own), _p(self.nei), _p(w), int(merge_levels))

    def __del__(self):
        if getattr(self, "_h", None):
            _lib().oracle_mesh_free(self._h)
            self._h = None

    # ---------------------------------------------------------------- GAMG
    def gamg_levels(self):
        lib = _lib()
        out = []
        for k in range(lib.oracle_gamg_nlevels(self._h)):
            n, nf = ctypes.c_int(), ctypes.c_int()
            lib.oracle_gamg_level_sizes(self._h, k, ctypes.byref(n), ctypes.byref(nf))
            n, nf = n.value, nf.value
            d = {"n": n, "agg": np.zeros(n, np.int32), "faceMap": np.zeros(nf, np.int32),
                 "faceFlip": np.zeros(nf, np.int32), "l": np.zeros(nf, np.int32), "u": np.zeros(nf, np.int32),
                 "weight": np.zeros(nf)}
            lib.oracle_gamg_level(self._h, k, *(_p(d[x]) for x in ("agg", "faceMap", "faceFlip", "l", "u", "weight")))
            out.append(d)
        return out

    def gamg_coeffs(self, diag, upper, lower, k):
        lv = self.gamg_levels()[k]
        cd, cu, cl = np.zeros(lv["n"]), np.zeros(len(lv["l"])), np.zeros(len(lv["l"]))
        upper = _f64(upper)
        lower = upper if lower is None else _f64(lower)
        _lib().oracle_gamg_coeffs(self._h, _p(_f64(diag)), _p(upper), _p(lower), k, _p(cd), _p(cu), _p(cl))
        return cd, cu, cl

    def gamg_apply(self, diag, upper, lower, r, smoother=SMOOTHER_GS, omega=1.0, n_pre=2, n_post=2):
        upper = _f64(upper)
        lower = upper if lower is None else _f64(lower)
        w = np.zeros(self.n)
        rc = _lib().oracle_gamg_apply(self._h, _p(_f64(diag)), _p(upper), _p(lower), _p(_f64(r)), _p(w),
                                      int(smoother), float(omega), int(n_pre), int(n_post))
        if rc:
            raise OracleError(f"gamg_apply failed ({rc})")
        return w

    # ---------------------------------------------------------------- solve
    def amul(self, diag, upper, lower, x):
        upper = _f64(upper)
        lower = upper if lower is None else _f64(lower)
        y = np.zeros(self.n)
        _lib().oracle_amul(self._h, _p(_f64(diag)), _p(upper), _p(lower), _p(_f64(x)), _p(y))
        return y

    def solve(self, diag, upper, lower, b, x0=None, solver="PCG", preconditioner="diagonal", tolerance=1e-8,
              rel_tol=0.0, max_iter=1000, min_iter=0, n_sweeps=1, mode=MODE_PLAIN, smoother=SMOOTHER_GS,
              omega=1.0, n_pre=2, n_post=2, history=False):
        """solveLdu on (diag (boundary-folded), upper, lower or None=symmetric, b, x0).

        Returns (x, perf dict[, residual history]). Raises OracleError where the
        reference throws.
        """
        diag, upper, b = _f64(diag), _f64(upper), _f64(b)
        sym = lower is None
        lower = upper if sym else _f64(lower)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
        perf = np.zeros(4)
        it, cv = ctypes.c_int(), ctypes.c_int()
        hist = np.full(max_iter + max(n_sweeps, 1) + 1, np.nan) if history else None
        err = ctypes.create_string_buffer(256)
        rc = _lib().oracle_solve(self._h, _p(diag), _p(upper), _p(lower), int(sym), _p(b), _p(x), solver.encode(),
                                 preconditioner.encode(), float(tolerance), float(rel_tol), int(max_iter),
                                 int(min_iter), int(n_sweeps), int(mode), int(smoother), float(omega), int(n_pre),
                                 int(n_post), _p(perf), ctypes.byref(it), ctypes.byref(cv), _p(hist), err, 256)
        if rc:
            raise OracleError(err.value.decode() or f"oracle_solve failed ({rc})")
        p = {"initialResidual": perf[0], "finalResidual": perf[1], "residualSum": perf[2], "normFactor": perf[3],
             "iterations": it.value, "converged": bool(cv.value)}
        if history:
            return x, p, hist[: it.value + 1]
        return x, p


# -------------------------------------------------------------------- assembly
def laplacian(mesh, bc_fixed, bc_value=None, gamma=None):
    """fvm::laplacian coefficients: (diag, upper, internalCoeffs, boundarySource)."""
    g = mesh.geometry()
    n, ni, nb = mesh.n_cells, mesh.n_internal, mesh.n_boundary
    diag, upper, ic, bs = np.zeros(n), np.zeros(ni), np.zeros(nb), np.zeros(nb)
    bcv = np.zeros(nb) if bc_value is None else _f64(bc_value)
    _lib().oracle_laplacian(n, ni, nb, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(g["magSf"]), _p(g["ndc"]),
                            _p(_f64(gamma)), _p(_i32(bc_fixed)), _p(bcv), _p(diag), _p(upper), _p(ic), _p(bs))
    return diag, upper, ic, bs


def laplacian_nonorth(mesh, grad, gamma=None):
    g = mesh.geometry()
    n, ni = mesh.n_cells, mesh.n_internal
    source, ffc = np.zeros(n), np.zeros(ni)
    _lib().oracle_laplacian_nonorth(n, ni, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(g["magSf"]),
                                    _p(g["w"]), _p(np.ascontiguousarray(g["corr"])), _p(_f64(gamma)),
                                    _p(np.ascontiguousarray(grad, dtype=np.float64)), _p(source), _p(ffc))
    return source, ffc


def grad_gauss(mesh, psi, bval):
    g = mesh.geometry()
    n = mesh.n_cells
    out = np.zeros((n, 3))
    _lib().oracle_grad_gauss(n, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)),
                             _p(g["w"]), _p(np.ascontiguousarray(g["Sf"])), _p(g["V"]), _p(_f64(psi)),
                             _p(_f64(bval)), _p(out))
    return out


def div_upwind(mesh, F, bc_fixed, bc_value=None):
    n, ni, nb = mesh.n_cells, mesh.n_internal, mesh.n_boundary
    diag, upper, lower, ic, bs = np.zeros(n), np.zeros(ni), np.zeros(ni), np.zeros(nb), np.zeros(nb)
    bcv = np.zeros(nb) if bc_value is None else _f64(bc_value)
    _lib().oracle_div_upwind(n, ni, nb, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(_f64(F)),
                             _p(_i32(bc_fixed)), _p(bcv), _p(diag), _p(upper), _p(lower), _p(ic), _p(bs))
    return diag, upper, lower, ic, bs


def ddt_euler(mesh, dt, psi0):
    n = mesh.n_cells
    diag, source = np.zeros(n), np.zeros(n)
    _lib().oracle_ddt_euler(n, _p(mesh.geometry()["V"]), float(dt), _p(_f64(psi0)), _p(diag), _p(source))
    return diag, source


def fold(mesh, diag, source, ic, bs):
    n = mesh.n_cells
    d, b = np.zeros(n), np.zeros(n)
    _lib().oracle_fold(n, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)), _p(_f64(diag)), _p(_f64(source)),
                       _p(_f64(ic)), _p(_f64(bs)), _p(d), _p(b))
    return d, b


def boundary_values(mesh, psi, bc_fixed, bc_value=None):
    """Boundary face values after correctBoundaryConditions: the fixed value on
    fixedValue faces, the owner cell's value on zeroGradient faces
    (BoundaryCondition.h:96-137)."""
    ni = mesh.n_internal
    own_b = np.asarray(mesh.owner[ni:])
    bcv = np.zeros(mesh.n_boundary) if bc_value is None else _f64(bc_value)
    return np.where(np.asarray(bc_fixed) != 0, bcv, _f64(psi)[own_b])


def transport(mesh, dt, nu, psi0, F, bc_fixed, bc_value=None, non_orth=True):
    """One component of the momentum matrix as IncompressibleFlow.cpp:183-185 composes
    it, ddt(Euler) + div(F, upwind) − laplacian(nu), restated with the FvMatrix algebra
    of FvMatrix.h:57-80,232-239 (each term built alone, then added with s = ±1; the
    lower part appears when the div term, which has one, is added). psi0 is both the
    old-time and the current field. Returns the FvMatrix parts and the fold
    (FvMatrix.h:154-163) as a dict."""
    n, ni, nb = mesh.n_cells, mesh.n_internal, mesh.n_boundary
    psi0 = _f64(psi0)
    # ddt: FvMatrix(psi) zeros, then diag += c·V, source += V(c0 ψ⁰ − c00 ψ⁰⁰)
    diag, source = ddt_euler(mesh, dt, psi0)
    upper = np.zeros(ni)
    ic, bs = np.zeros(nb), np.zeros(nb)
    # += div: the div term has a lower part, so lower() is first copied from upper
    dd, du, dl, dic, dbs = div_upwind(mesh, F, bc_fixed, bc_value)
    diag = diag + dd
    lower = upper.copy() + 1.0 * dl
    upper = upper + 1.0 * du
    source = source + np.zeros(n)
    ic, bs = ic + dic, bs + dbs
    # −= laplacian(nu) (symmetric: its lowerPtr is its upper)
    gam = np.full(mesh.n_faces, float(nu))
    ld, lu, lic, lbs = laplacian(mesh, bc_fixed, bc_value, gamma=gam)
    lsrc = np.zeros(n)
    lffc = None
    if non_orth and np.any(np.einsum("ij,ij->i", mesh.geometry()["corr"][:ni], mesh.geometry()["corr"][:ni]) > 1e-20):
        g = grad_gauss(mesh, psi0, boundary_values(mesh, psi0, bc_fixed, bc_value))
        lsrc, lffc = laplacian_nonorth(mesh, g, gam)
    diag = diag - ld
    lower = lower + (-1.0) * lu
    upper = upper + (-1.0) * lu
    source = source - lsrc
    ic, bs = ic - lic, bs - lbs
    ffc = np.zeros(mesh.n_faces)
    if lffc is not None:
        ffc[:ni] = ffc[:ni] + (-1.0) * lffc
    dF, bF = fold(mesh, diag, source, ic, bs)
    return {"diag": diag, "upper": upper, "lower": lower, "source": source, "ic": ic, "bs": bs, "ffc": ffc,
            "has_ffc": lffc is not None, "dF": dF, "bF": bF}


def pressure_matrix(mesh, gamma, psi, bc_fixed, bc_value=None, non_orth=True):
    """fvm::laplacian(γ, ψ) on its own (Fvm.h:142-181), as a pressure equation is
    built: FvMatrix parts with lower = upper and the correction flux when the mesh has
    non-orthogonal faces."""
    n, ni = mesh.n_cells, mesh.n_internal
    gam = np.full(mesh.n_faces, float(gamma))
    ld, lu, lic, lbs = laplacian(mesh, bc_fixed, bc_value, gamma=gam)
    source = np.zeros(n)
    ffc = np.zeros(mesh.n_faces)
    has = bool(non_orth and np.any(np.einsum("ij,ij->i", mesh.geometry()["corr"][:ni],
                                             mesh.geometry()["corr"][:ni]) > 1e-20))
    if has:
        g = grad_gauss(mesh, psi, boundary_values(mesh, psi, bc_fixed, bc_value))
        source, lffc = laplacian_nonorth(mesh, g, gam)
        ffc[:ni] = lffc
    return {"diag": ld, "upper": lu, "lower": lu.copy(), "source": source, "ic": lic, "bs": lbs, "ffc": ffc,
            "has_ffc": has}


# ---------------------------------------------------------------- pressure-velocity coupling
def relax(mesh, M, psi, alpha):
    """FvMatrix::relax(alpha) (FvMatrix.h:111-131): (diag, source) of the relaxed matrix."""
    diag, source = _f64(M["diag"]).copy(), _f64(M["source"]).copy()
    _lib().oracle_relax(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                        _p(_i32(mesh.neighbour)), _p(_f64(M["upper"])), _p(_f64(M["lower"])), _p(_f64(M["ic"])),
                        _p(_f64(psi)), float(alpha), _p(diag), _p(source))
    return diag, source


def set_reference(diag, source, cell, value):
    """FvMatrix::setReference (FvMatrix.h:134-141): (diag, source) afterwards."""
    d, s = _f64(diag).copy(), _f64(source).copy()
    _lib().oracle_set_reference(len(d), int(cell), float(value), _p(d), _p(s))
    return d, s


def A(mesh, diag, ic, n_cmpt=1):
    """FvMatrix::A() (FvMatrix.h:179-187); n_cmpt 3 for a vector equation."""
    out = np.zeros(mesh.n_cells)
    _lib().oracle_A(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)), _p(_f64(diag)),
                    _p(_f64(ic)), _p(mesh.geometry()["V"]), int(n_cmpt), _p(out))
    return out


def H(mesh, M, psi, n_cmpt=1):
    """FvMatrix::H() (FvMatrix.h:189-208) of one component."""
    out = np.zeros(mesh.n_cells)
    _lib().oracle_H(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)),
                    _p(_f64(M["upper"])), _p(_f64(M["lower"])), _p(_f64(M["source"])), _p(_f64(M["ic"])),
                    _p(_f64(M["bs"])), _p(_f64(psi)), _p(mesh.geometry()["V"]), int(n_cmpt), _p(out))
    return out


def flux(mesh, M, psi):
    """FvMatrix::flux() (FvMatrix.h:210-222) of one component."""
    out = np.zeros(mesh.n_faces)
    ffc = _f64(M["ffc"]) if M.get("has_ffc") else None
    _lib().oracle_flux(mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)),
                       _p(_f64(M["upper"])), _p(_f64(M["lower"])), _p(_f64(M["ic"])), _p(_f64(M["bs"])),
                       _p(ffc) if ffc is not None else None, _p(_f64(psi)), _p(out))
    return out


def transport_post(mesh, dt, nu, psi0, F, bc_fixed, bc_value, alpha, ref_cell, ref_value, non_orth=True,
                   pressure=False):
    """The momentum matrix of :func:`transport` (or the pressure matrix of
    :func:`pressure_matrix` when `pressure`), relax(alpha), A(), H() and flux() of the
    relaxed matrix at ψ = psi0, then setReference. psi0 (n,) for a scalar field or (3, n)
    for a vector field (bc_value then (3, nBoundary)); the same dict as the reference
    driver's transport_post."""
    psi0 = np.asarray(psi0, dtype=np.float64)
    vec = psi0.ndim == 2
    comps = psi0 if vec else psi0[None, :]
    bvals = (np.asarray(bc_value, dtype=np.float64) if vec else np.asarray(bc_value, dtype=np.float64)[None, :])
    nc = 3 if vec else 1
    res = {k: [] for k in ("source_relaxed", "H", "flux", "source_ref")}
    for k in range(comps.shape[0]):
        if pressure:
            M = pressure_matrix(mesh, nu, comps[k], bc_fixed, bvals[k], non_orth)
        else:
            M = transport(mesh, dt, nu, comps[k], F, bc_fixed, bvals[k], non_orth)
        d, s = relax(mesh, M, comps[k], alpha)
        R = dict(M, diag=d, source=s)
        res["source_relaxed"].append(s)
        res["H"].append(H(mesh, R, comps[k], nc))
        res["flux"].append(flux(mesh, R, comps[k]))
        dr, sr = set_reference(d, s, ref_cell, ref_value)
        res["source_ref"].append(sr)
        res["diag_relaxed"], res["A"], res["diag_ref"] = d, A(mesh, d, M["ic"], nc), dr
    for k in ("source_relaxed", "H", "flux", "source_ref"):
        res[k] = np.array(res[k]) if vec else res[k][0]
    return res


# ---------------------------------------------------------------- explicit corrector operators
def _geo3(mesh):
    g = mesh.geometry()
    return g, np.ascontiguousarray(g["skew"]) if g["skewed"] else None


def fvc_interpolate(mesh, c, grad=None):
    """fvc::interpolate(m, c[, grad]) (Fvc.h:24-31,72-81); c (n,) or (n, 3); grad (n, 3)
    or (n, 9) for the skew correction (applied on skewed meshes)."""
    g, sk = _geo3(mesh)
    c = np.ascontiguousarray(c, dtype=np.float64)
    nc = 1 if c.ndim == 1 else 3
    out = np.zeros((mesh.n_faces,) if nc == 1 else (mesh.n_faces, 3))
    gr = None if grad is None else np.ascontiguousarray(grad, dtype=np.float64)
    _lib().oracle_fvc_interpolate(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                                  _p(_i32(mesh.neighbour)), _p(g["w"]), None if sk is None else _p(sk), nc, _p(c),
                                  None if gr is None else _p(gr), _p(out))
    return out


def fvc_grad_ls(mesh, c, b_fixed, b_value):
    """fvc::gradLS(m, c, bFixesValue, bValue) (Fvc.h:185-206)."""
    g = mesh.geometry()
    c = np.ascontiguousarray(c, dtype=np.float64)
    nc = 1 if c.ndim == 1 else 3
    out = np.zeros((mesh.n_cells, 3 * nc))
    _lib().oracle_fvc_grad_ls(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                              _p(_i32(mesh.neighbour)), _p(np.ascontiguousarray(g["d"])),
                              _p(np.ascontiguousarray(g["Sf"])), _p(g["magSf"]), nc, _p(c), _p(_i32(b_fixed)),
                              _p(np.ascontiguousarray(b_value, dtype=np.float64)), _p(out))
    return out


def fvc_grad_gauss(mesh, psi, b_fixed, b_value):
    """fvc::grad(ψ, GaussLinear) with fixedValue (b_value) / zeroGradient faces."""
    bval = np.where(np.asarray(b_fixed) != 0, _f64(b_value), _f64(psi)[np.asarray(mesh.owner[mesh.n_internal:])])
    return grad_gauss(mesh, psi, bval)


def fvc_grad_gauss_skew(mesh, psi, b_fixed, b_value):
    """fvc::grad(ψ, GaussSkew) (Fvc.h:223-259); it applies the skew vectors whether or not
    the mesh counts as skewed."""
    g = mesh.geometry()
    sk = np.ascontiguousarray(g["skew"])
    out = np.zeros((mesh.n_cells, 3))
    _lib().oracle_fvc_grad_gauss_skew(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                                      _p(_i32(mesh.neighbour)), _p(g["w"]), _p(sk), _p(np.ascontiguousarray(g["d"])),
                                      _p(np.ascontiguousarray(g["Sf"])), _p(g["magSf"]), _p(g["V"]), _p(_f64(psi)),
                                      _p(_i32(b_fixed)), _p(_f64(b_value)), _p(out))
    return out


def fvc_div(mesh, F):
    """fvc::div(m, F) (Fvc.h:294-305)."""
    out = np.zeros(mesh.n_cells)
    _lib().oracle_fvc_div(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                          _p(_i32(mesh.neighbour)), _p(mesh.geometry()["V"]), _p(_f64(F)), _p(out))
    return out


def fvc_surface_integrate(mesh, F):
    """fvc::surfaceIntegrate(m, F) (Fvc.h:307-318); F (nFaces,) or (nFaces, 3)."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    nc = 1 if F.ndim == 1 else 3
    out = np.zeros((mesh.n_cells,) if nc == 1 else (mesh.n_cells, 3))
    _lib().oracle_fvc_surface_integrate(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                                        _p(_i32(mesh.neighbour)), _p(mesh.geometry()["V"]), nc, _p(F), _p(out))
    return out


def fvc_flux(mesh, U, grad, b_fixed, Ub):
    """phiHbyA-style face flux of a cell vector field (IncompressibleFlow.cpp:276-291)."""
    g, sk = _geo3(mesh)
    out = np.zeros(mesh.n_faces)
    gr = None if grad is None else np.ascontiguousarray(grad, dtype=np.float64)
    _lib().oracle_fvc_flux(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                           _p(_i32(mesh.neighbour)), _p(g["w"]), None if sk is None else _p(sk),
                           _p(np.ascontiguousarray(g["Sf"])), _p(np.ascontiguousarray(U, dtype=np.float64)),
                           None if gr is None else _p(gr), _p(_i32(b_fixed)),
                           _p(np.ascontiguousarray(Ub, dtype=np.float64)), _p(out))
    return out


# ---------------------------------------------------------------- pressure-corrector steps
# mesh: n_cells, n_internal, n_boundary, owner (all faces), neighbour and geometry()
# (Sf, magSf, w over the faces the solve uses).
def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def piso_rau_hbya(A, H):
    """rAU = 1/A, HbyA = rAU·H (IncompressibleFlow.cpp:245-249)."""
    A, H = _f64(A), _f64(H)
    rAU, HbyA = np.zeros(len(A)), np.zeros((len(A), 3))
    _lib().oracle_piso_rau_hbya(len(A), _p(A), _p(H), _p(rAU), _p(HbyA))
    return rAU, HbyA


def piso_hbya_boundary(mesh, b_fixed, Ub, rAU, gradp):
    """U_b + rAU_P ∇p_P on fixed-flux boundary faces, 0 elsewhere (IncompressibleFlow.cpp:261-270)."""
    out = np.zeros((mesh.n_boundary, 3))
    _lib().oracle_piso_hbya_boundary(mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)), _p(_i32(b_fixed)),
                                     _p(_f64(Ub)), _p(_f64(rAU)), _p(_f64(gradp)), _p(out))
    return out


def piso_face_op(op, a, b, c=None):
    """op 0: a + b·c, 1: a − b, 2: a − b/c (IncompressibleFlow.cpp:219-220,294-295,338-339,351-352)."""
    a, b, c = _f64(a), _f64(b), _f64(c)
    out = np.zeros(len(a))
    _lib().oracle_piso_face_op(len(a), int(op), _p(a), _p(b), None if c is None else _p(c), _p(out))
    return out


def piso_ddt_phi_corr(mesh, coeff, c0, c00, U0, U00, phi_old, phi_old_old, rAUf, phiHbyA):
    """phiHbyA after the ddtPhiCorr coupling (IncompressibleFlow.cpp:296-316)."""
    g = mesh.geometry()
    out = _f64(phiHbyA).copy()
    _lib().oracle_piso_ddt_phi_corr(mesh.n_internal, _p(_i32(mesh.owner)), _p(_i32(mesh.neighbour)), _p(_f64(g["w"])),
                                    _p(_f64(g["Sf"])), float(coeff), float(c0), float(c00), _p(_f64(U0)),
                                    None if U00 is None else _p(_f64(U00)), _p(_f64(phi_old)),
                                    None if phi_old_old is None else _p(_f64(phi_old_old)), _p(_f64(rAUf)), _p(out))
    return out


def piso_constrain_pressure(mesh, start, size, phiHbyA, Ub, rAUf):
    """fixedFluxPressure gradient of boundary faces [start, start + size) (IncompressibleFlow.cpp:17-29)."""
    g = mesh.geometry()
    out = np.zeros(max(size, 1))
    _lib().oracle_piso_constrain_pressure(mesh.n_internal, int(start), int(size), _p(_f64(g["Sf"])),
                                          _p(_f64(g["magSf"])), _p(_f64(phiHbyA)), _p(_f64(Ub)), _p(_f64(rAUf)),
                                          _p(out))
    return out[:size]


def fvc_reconstruct(mesh, ssf, empty_dir=-1):
    """fvc::reconstruct of a face flux (Fvc.h:33-67): nCells × 3."""
    g = mesh.geometry()
    out = np.zeros((mesh.n_cells, 3))
    _lib().oracle_fvc_reconstruct(mesh.n_cells, mesh.n_internal, mesh.n_boundary, _p(_i32(mesh.owner)),
                                  _p(_i32(mesh.neighbour)), _p(_f64(g["Sf"])), _p(_f64(g["magSf"])), _p(_f64(ssf)),
                                  int(empty_dir), _p(out))
    return out


def piso_correct_U(HbyA, rAU, v, sign, empty_dir=-1):
    """U = HbyA − rAU·v (sign < 0) or HbyA + rAU·v, empty direction zeroed (IncompressibleFlow.cpp:347-360)."""
    HbyA = _f64(HbyA)
    out = np.zeros_like(HbyA)
    _lib().oracle_piso_correct_U(len(HbyA), _p(HbyA), _p(_f64(rAU)), _p(_f64(v)), int(sign), int(empty_dir), _p(out))
    return out

