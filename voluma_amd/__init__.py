"""voluma_amd — MI355X (gfx950) LDU linear-solve path with Voluma's linalg interface.

Python mirror of the reference's solver API (src/fvm/linalg/LinearSolver.h:15-45),
bound through the C ABI of include/voluma_amd.h (``libvoluma_amd.so``):

* :class:`SolverControls` / :class:`SolverPerformance` — same fields and defaults;
* :class:`LduMesh` — owner/neighbour addressing and the cached GAMG hierarchy;
* :class:`LduMatrix` — diag (boundary-folded) / upper / lower resident on the GPU;
* :func:`solve_ldu` (``solveLdu``) and :func:`amul`.

The native library is required: importing the solver API without it raises.
Mesh synthesis lives in :mod:`voluma_amd.mesh`, benchmark systems in
:mod:`voluma_amd.workloads`.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvoluma_amd.so")

VOL_OK = 0
VOL_HOST, VOL_DEVICE = 0, 1
SUMS_CONTEXT, SUMS_FAST, SUMS_REPRODUCIBLE, SUMS_EXACT = -1, 0, 1, 2
FP64, MIXED = 0, 1
GAMG_JACOBI, GAMG_GAUSS_SEIDEL = 0, 1
GRAD_GAUSS_LINEAR, GRAD_LEAST_SQUARES, GRAD_GAUSS_SKEW = 0, 1, 2
FACE_A_PLUS_B_TIMES_C, FACE_A_MINUS_B, FACE_A_MINUS_B_OVER_C = 0, 1, 2
DDT_EULER, DDT_BACKWARD = 1, 2

_ERRORS = {
    -1: "invalid argument", -2: "unknown solver", -3: "unknown preconditioner", -4: "PCG on nonsymmetric matrix",
    -5: "singular coarsest matrix", -6: "unsupported", -7: "HIP error", -8: "no GAMG weights", -9: "timeout",
}


class VolumaError(RuntimeError):
    """A non-zero status of the C ABI (the reference would have thrown)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code} {_ERRORS.get(code, '?')}] {msg}")
        self.code = code


class _Controls(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_char_p), ("preconditioner", ctypes.c_char_p), ("tolerance", ctypes.c_double),
                ("relTol", ctypes.c_double), ("maxIter", ctypes.c_int), ("minIter", ctypes.c_int),
                ("nSweeps", ctypes.c_int), ("sums", ctypes.c_int), ("precision", ctypes.c_int),
                ("gamgPreSweeps", ctypes.c_int), ("gamgPostSweeps", ctypes.c_int), ("gamgOmega", ctypes.c_double),
                ("innerRelTol", ctypes.c_double), ("gamgSmoother", ctypes.c_int), ("reserved", ctypes.c_int)]


class _NonOrth(ctypes.Structure):
    _fields_ = [("psi", ctypes.c_void_p), ("w", ctypes.c_void_p), ("Sf", ctypes.c_void_p), ("corrVecs", ctypes.c_void_p),
                ("V", ctypes.c_void_p)]


class _FlowOptions(ctypes.Structure):
    _fields_ = [("nu", ctypes.c_double), ("nCorr", ctypes.c_int), ("nNonOrthCorr", ctypes.c_int),
                ("momentumPredictor", ctypes.c_int), ("ddtScheme", ctypes.c_int), ("gradPScheme", ctypes.c_int),
                ("laplacianNonOrthCorr", ctypes.c_int), ("pRefCell", ctypes.c_int64), ("pRefValue", ctypes.c_double),
                ("ddtPhiCoeff", ctypes.c_double), ("UControls", _Controls), ("pControls", _Controls),
                ("pFinalControls", _Controls)]


class _Perf(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_char * 32), ("initialResidual", ctypes.c_double),
                ("finalResidual", ctypes.c_double), ("residualSum", ctypes.c_double),
                ("normFactor", ctypes.c_double), ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("outerIterations", ctypes.c_int), ("reserved", ctypes.c_int)]


class _Request(ctypes.Structure):
    _fields_ = [("mesh", ctypes.c_int), ("diag", ctypes.c_void_p), ("upper", ctypes.c_void_p),
                ("lower", ctypes.c_void_p), ("nRhs", ctypes.c_int), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("controls", _Controls), ("perf", ctypes.c_void_p)]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f64n(a):
    return None if a is None else _f64(a)


def _q(a):
    return None if a is None else _p(a)


def _sized(a, count, name, dtype=np.float64):
    """Contiguous copy of a with `count` elements of dtype, else ValueError: the C ABI
    copies exactly that many elements and cannot check a caller's array."""
    if a is None:
        raise ValueError(f"{name} is required")
    arr = np.ascontiguousarray(a, dtype=dtype)
    if arr.size != count:
        raise ValueError(f"{name}: expected {count} elements, got {arr.size} (shape {arr.shape})")
    return arr


_LIB = None
EXPORTED = ("vol_controls_default", "vol_last_error", "vol_context_create", "vol_context_destroy",
            "vol_set_reproducible", "vol_synchronize", "vol_device_alloc", "vol_device_free", "vol_memcpy_to_device",
            "vol_memcpy_to_host", "vol_mesh_create", "vol_mesh_destroy", "vol_mesh_gamg_levels",
            "vol_mesh_gamg_level", "vol_matrix_create", "vol_matrix_set", "vol_matrix_destroy", "vol_amul",
            "vol_solve", "vol_timing_enable", "vol_timing_read", "vol_mesh_set_boundary", "vol_assemble_laplacian",
            "vol_assemble_div_upwind", "vol_matrix_get", "vol_device_copy", "vol_fvm_reset", "vol_fvm_ddt_euler",
            "vol_fvm_div_upwind", "vol_fvm_laplacian", "vol_fvm_source", "vol_fvm_get", "vol_fvm_finish",
            "vol_batch_create", "vol_batch_destroy", "vol_batch_info", "vol_batch_set", "vol_batch_solve",
            "vol_server_create", "vol_server_add_mesh", "vol_server_submit", "vol_server_wait", "vol_server_stats",
            "vol_server_destroy", "vol_fvm_finish_source", "vol_fvm_relax", "vol_fvm_set_reference", "vol_fvm_A",
            "vol_fvm_H", "vol_fvm_flux", "vol_mesh_set_geometry", "vol_fvc_interpolate", "vol_fvc_grad", "vol_fvc_div",
            "vol_fvc_surface_integrate", "vol_fvc_flux", "vol_geometry", "vol_piso_rau_hbya",
            "vol_piso_hbya_boundary", "vol_piso_face_op", "vol_piso_ddt_phi_corr", "vol_piso_constrain_pressure",
            "vol_fvc_reconstruct", "vol_piso_correct_U", "vol_fvm_ddt",
            "vol_mesh_set_empty_dir", "vol_boundary_values", "vol_copy_strided", "vol_flow_options_default",
            "vol_flow_create", "vol_flow_set_fields", "vol_flow_get_fields", "vol_flow_step", "vol_flow_destroy",
            "vol_solve_noalloc", "vol_mesh_gamg_setup_info", "vol_batch_graph_stats", "vol_server_prefetch")


def lib():
    """The loaded native library (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.vol_last_error.restype = ctypes.c_char_p
        L.vol_controls_default.argtypes = [ctypes.POINTER(_Controls)]
        L.vol_context_create.argtypes = [I, ctypes.POINTER(P)]
        L.vol_context_destroy.argtypes = [P]
        L.vol_set_reproducible.argtypes = [P, I]
        L.vol_synchronize.argtypes = [P]
        L.vol_device_alloc.argtypes = [P, ctypes.c_uint64, ctypes.POINTER(P)]
        L.vol_device_free.argtypes = [P, P]
        L.vol_memcpy_to_device.argtypes = [P, P, P, ctypes.c_uint64]
        L.vol_memcpy_to_host.argtypes = [P, P, P, ctypes.c_uint64]
        L.vol_mesh_create.argtypes = [P, I, I, P, P, P, I, ctypes.POINTER(P)]
        L.vol_mesh_destroy.argtypes = [P]
        L.vol_mesh_gamg_levels.argtypes = [P, ctypes.POINTER(I), P]
        L.vol_mesh_gamg_level.argtypes = [P, I, ctypes.POINTER(I), P, P, P]
        L.vol_matrix_create.argtypes = [P, ctypes.POINTER(P)]
        L.vol_matrix_set.argtypes = [P, P, P, P, I]
        L.vol_matrix_destroy.argtypes = [P]
        L.vol_amul.argtypes = [P, P, P, I]
        L.vol_solve.argtypes = [P, P, P, ctypes.POINTER(_Controls), ctypes.POINTER(_Perf), I]
        L.vol_solve_noalloc.argtypes = [P, P, P, ctypes.POINTER(_Controls), ctypes.POINTER(_Perf)]
        L.vol_mesh_gamg_setup_info.argtypes = [P, ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I)]
        L.vol_batch_graph_stats.argtypes = [P, ctypes.POINTER(I), ctypes.POINTER(I), I]
        L.vol_timing_enable.argtypes = [P, I]
        L.vol_timing_read.argtypes = [P, ctypes.POINTER(D), ctypes.POINTER(ctypes.c_int64), I]
        L.vol_mesh_set_boundary.argtypes = [P, I, P]
        L.vol_matrix_get.argtypes = [P, P, P, P, I]
        L.vol_device_copy.argtypes = [P, P, P, ctypes.c_int64, I, ctypes.POINTER(D)]
        L.vol_assemble_laplacian.argtypes = [P, D, P, P, P, P, P, P, P, P, P, I]
        L.vol_assemble_div_upwind.argtypes = [P, P, P, P, P, P, P, P, I]
        L.vol_fvm_reset.argtypes = [P]
        L.vol_fvm_ddt_euler.argtypes = [P, I, D, P, P, I]
        L.vol_fvm_ddt.argtypes = [P, I, D, D, D, P, P, P, I]
        L.vol_mesh_set_empty_dir.argtypes = [P, I]
        L.vol_boundary_values.argtypes = [P, I, P, P, P, P, I]
        L.vol_copy_strided.argtypes = [P, ctypes.c_int64, P, I, P, I, D]
        L.vol_flow_options_default.argtypes = [ctypes.POINTER(_FlowOptions)]
        L.vol_flow_options_default.restype = None
        L.vol_flow_create.argtypes = [P, ctypes.POINTER(_FlowOptions), P, P, P, P, P, P, I, ctypes.POINTER(P)]
        L.vol_flow_set_fields.argtypes = [P, P, P, P, I]
        L.vol_flow_get_fields.argtypes = [P, P, P, P, I]
        L.vol_flow_step.argtypes = [P, D, P, I, ctypes.POINTER(I)]
        L.vol_flow_destroy.argtypes = [P]
        L.vol_fvm_div_upwind.argtypes = [P, I, P, P, P, I]
        L.vol_fvm_laplacian.argtypes = [P, I, D, P, P, P, P, P, ctypes.POINTER(_NonOrth), I]
        L.vol_fvm_source.argtypes = [P, P, P, I]
        L.vol_fvm_get.argtypes = [P, P, P, P, P, P, I]
        L.vol_fvm_finish.argtypes = [P, P, I]
        L.vol_fvm_finish_source.argtypes = [P, P, P, P, I]
        L.vol_fvm_relax.argtypes = [P, D, P, I]
        L.vol_fvm_set_reference.argtypes = [P, ctypes.c_int64, D]
        L.vol_fvm_A.argtypes = [P, P, I, P, I]
        L.vol_fvm_H.argtypes = [P, P, P, I, P, I]
        L.vol_fvm_flux.argtypes = [P, P, P, I]
        L.vol_geometry.argtypes = [P, I, P, I, P, P, I, P, P, I] + [P] * 11 + [I]
        L.vol_mesh_set_geometry.argtypes = [P, P, P, P, P, P, P, I, I]
        L.vol_fvc_interpolate.argtypes = [P, I, P, P, P, P, I]
        L.vol_fvc_grad.argtypes = [P, I, I, P, P, P, P, I]
        L.vol_fvc_div.argtypes = [P, P, P, I]
        L.vol_fvc_surface_integrate.argtypes = [P, I, P, P, I]
        L.vol_fvc_flux.argtypes = [P, P, P, P, P, P, I]
        L.vol_piso_rau_hbya.argtypes = [P, P, P, P, P, I]
        L.vol_piso_hbya_boundary.argtypes = [P, P, P, P, P, P, I]
        L.vol_piso_face_op.argtypes = [P, I, P, P, P, P, I]
        L.vol_piso_ddt_phi_corr.argtypes = [P, D, D, D, P, P, P, P, P, P, I]
        L.vol_piso_constrain_pressure.argtypes = [P, I, I, P, P, P, P, I]
        L.vol_fvc_reconstruct.argtypes = [P, P, I, P, I]
        L.vol_piso_correct_U.argtypes = [P, P, P, P, I, I, P, I]
        L.vol_batch_create.argtypes = [P, I, P, P, P, P, P, ctypes.POINTER(P)]
        L.vol_batch_destroy.argtypes = [P]
        L.vol_batch_info.argtypes = [P, P, P, P]
        L.vol_batch_set.argtypes = [P, P, P, P, I]
        L.vol_batch_solve.argtypes = [P, P, P, ctypes.POINTER(_Controls), P, I]
        L.vol_server_create.argtypes = [I, I, ctypes.c_uint64, ctypes.POINTER(P)]
        L.vol_server_add_mesh.argtypes = [P, I, I, P, P, P, P, I, ctypes.POINTER(I)]
        L.vol_server_submit.argtypes = [P, ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_int64)]
        L.vol_server_wait.argtypes = [P, ctypes.c_int64, D, ctypes.POINTER(I), P]
        L.vol_server_prefetch.argtypes = [P, I, D]
        L.vol_server_stats.argtypes = [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(I), ctypes.POINTER(I)]
        L.vol_server_destroy.argtypes = [P]
        for name in EXPORTED:
            if name not in ("vol_controls_default", "vol_last_error"):
                getattr(L, name).restype = I
        _LIB = L
    return _LIB


def _check(rc):
    if rc != VOL_OK:
        raise VolumaError(rc, lib().vol_last_error().decode())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclasses.dataclass
class SolverControls:
    """SolverControls (LinearSolver.h:15-23) + GPU extensions (sums, GAMG smoother)."""
    solver: str = "PCG"
    preconditioner: str = "diagonal"
    tolerance: float = 1e-8
    relTol: float = 0.0
    maxIter: int = 1000
    minIter: int = 0
    nSweeps: int = 1
    sums: int = SUMS_CONTEXT
    precision: int = FP64
    gamgPreSweeps: int = 2
    gamgPostSweeps: int = 2
    gamgOmega: float = 0.8
    innerRelTol: float = 1e-2
    gamgSmoother: int = 0  # GAMG_JACOBI | GAMG_GAUSS_SEIDEL

    def _c(self):
        self._keep = (self.solver.encode(), self.preconditioner.encode())
        return _Controls(self._keep[0], self._keep[1], self.tolerance, self.relTol, self.maxIter, self.minIter,
                         self.nSweeps, self.sums, self.precision, self.gamgPreSweeps, self.gamgPostSweeps,
                         self.gamgOmega, self.innerRelTol, self.gamgSmoother, 0)


@dataclasses.dataclass
class SolverPerformance:
    """SolverPerformance (LinearSolver.h:25-31)."""
    solver: str = ""
    field: str = ""
    initialResidual: float = 0.0
    finalResidual: float = 0.0
    residualSum: float = 0.0
    normFactor: float = 0.0
    iterations: int = 0
    converged: bool = False
    outerIterations: int = 0


class Context:
    """One GPU and one HIP stream (vol_context)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().vol_context_create(int(device), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().vol_context_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_reproducible(self, on=True):
        _check(lib().vol_set_reproducible(self._h, int(bool(on))))

    def synchronize(self):
        _check(lib().vol_synchronize(self._h))

    def timing(self, on=True):
        _check(lib().vol_timing_enable(self._h, int(bool(on))))

    def read_timing(self, reset=True):
        ms, cnt = ctypes.c_double(), ctypes.c_int64()
        _check(lib().vol_timing_read(self._h, ctypes.byref(ms), ctypes.byref(cnt), int(bool(reset))))
        return ms.value, cnt.value


class DeviceArray:
    """float64 buffer in the context's GPU memory (for VOL_DEVICE calls)."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self._ctxh = ctx._h  # the allocation holds its own reference to the context
        p = ctypes.c_void_p()
        _check(lib().vol_device_alloc(self._ctxh, self.n * self.dtype.itemsize, ctypes.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, ctx, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(ctx, a.size, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(lib().vol_memcpy_to_device(self._ctxh, self.ptr, _p(a), a.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().vol_memcpy_to_host(self._ctxh, _p(out), self.ptr, out.nbytes))
        return out

    def copy_from(self, src: "DeviceArray", repeats=1):
        """self ← src on the GPU (streaming kernel); returns seconds per copy."""
        t = ctypes.c_double()
        _check(lib().vol_device_copy(self._ctxh, self.ptr, src.ptr, min(self.n, src.n), int(repeats), ctypes.byref(t)))
        return t.value

    def free(self):
        if getattr(self, "ptr", None):
            lib().vol_device_free(self._ctxh, self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class LduMesh:
    """LDU addressing (internal faces) and, with weights, the GAMG hierarchy (vol_mesh)."""

    def __init__(self, ctx: Context, n_cells, owner, neighbour, gamg_weights=None, merge_levels=2):
        self.ctx = ctx
        own = np.ascontiguousarray(owner, dtype=np.int32)
        nei = np.ascontiguousarray(neighbour, dtype=np.int32)
        if own.size != nei.size:
            raise ValueError("owner/neighbour sizes differ")
        self.n_cells, self.n_faces = int(n_cells), int(own.size)
        self.n_boundary = 0  # until set_boundary
        w = None if gamg_weights is None else np.ascontiguousarray(gamg_weights, dtype=np.float64)
        h = ctypes.c_void_p()
        _check(lib().vol_mesh_create(ctx._h, self.n_cells, self.n_faces, _p(own), _p(nei),
                                     None if w is None else _p(w), int(merge_levels), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().vol_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def set_boundary(self, boundary_owner):
        """Owner cells of the boundary faces (Mesh::owner past nInternalFaces), for assembly."""
        bo = np.ascontiguousarray(boundary_owner, dtype=np.int32)
        _check(lib().vol_mesh_set_boundary(self._h, int(bo.size), _p(bo)))
        self.n_boundary = int(bo.size)

    # ---- explicit corrector operators (Fvc.h) on the GPU, host arrays
    def set_geometry(self, V, Sf, magSf, w, d, skew=None, skewed=False):
        """Mesh geometry for the fvc operators (face arrays over internal + boundary faces)."""
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        arrs = [c(V), c(Sf), c(magSf), c(w), c(d), None if skew is None else c(skew)]
        _check(lib().vol_mesh_set_geometry(self._h, *(_q(a) for a in arrs), int(bool(skewed)), VOL_HOST))

    def set_empty_dir(self, empty_dir):
        """2-D mesh: its empty direction (−1 for 3-D) for the gradients."""
        _check(lib().vol_mesh_set_empty_dir(self._h, int(empty_dir)))

    def _nf_total(self):
        return self.n_faces + getattr(self, "n_boundary", 0)

    def _ncomp(self, c, name):
        """1 for a scalar cell field (n), 3 for a vector one (n × 3 or 3n flat with ndim 2)."""
        c = np.asarray(c)
        if c.ndim == 1:
            return 1
        if c.ndim == 2 and c.shape[1] == 3:
            return 3
        raise ValueError(f"{name}: expected shape ({self.n_cells},) or ({self.n_cells}, 3), got {c.shape}")

    def boundary_values(self, b_fixed, b_value, c):
        """fixedValue (b_fixed 1: b_value) / zeroGradient (the owner's value) boundary values."""
        nc = self._ncomp(c, "c")
        nb = self.n_boundary
        c = _sized(c, nc * self.n_cells, "c")
        bf = _sized(b_fixed, nb, "b_fixed", np.int32)
        bv = _sized(b_value, nc * nb, "b_value")
        out = np.zeros((max(nb, 1), nc))
        _check(lib().vol_boundary_values(self._h, nc, _p(bf), _p(bv), _p(c), _p(out), VOL_HOST))
        out = out[:nb]
        return out[:, 0] if nc == 1 else out

    def fvc_interpolate(self, c, grad=None, boundary_values=None):
        nc = self._ncomp(c, "c")
        n, nb = self.n_cells, self.n_boundary
        c = _sized(c, nc * n, "c")
        g = None if grad is None else _sized(grad, 3 * nc * n, "grad")
        bv = None if boundary_values is None else _sized(boundary_values, nc * nb, "boundary_values")
        out = np.zeros((max(self._nf_total(), 1), nc))
        _check(lib().vol_fvc_interpolate(self._h, nc, _p(c), _q(g), _q(bv), _p(out), VOL_HOST))
        out = out[: self._nf_total()]
        return out[:, 0] if nc == 1 else out

    def fvc_grad(self, scheme, c, b_fixed, b_value):
        nc = self._ncomp(c, "c")
        nb = self.n_boundary
        c = _sized(c, nc * self.n_cells, "c")
        bf = _sized(b_fixed, nb, "b_fixed", np.int32)
        bv = _sized(b_value, nc * nb, "b_value")
        out = np.zeros((self.n_cells, 3 * nc))
        _check(lib().vol_fvc_grad(self._h, int(scheme), nc, _p(c), _p(bf), _p(bv), _p(out), VOL_HOST))
        return out

    def fvc_div(self, F):
        F = _sized(F, self._nf_total(), "F")
        out = np.zeros(self.n_cells)
        _check(lib().vol_fvc_div(self._h, _p(F), _p(out), VOL_HOST))
        return out

    def fvc_surface_integrate(self, F):
        F = np.asarray(F)
        if F.ndim not in (1, 2) or (F.ndim == 2 and F.shape[1] != 3):
            raise ValueError(f"F: expected shape ({self._nf_total()},) or ({self._nf_total()}, 3), got {F.shape}")
        nc = 1 if F.ndim == 1 else 3
        F = _sized(F, nc * self._nf_total(), "F")
        out = np.zeros((self.n_cells, nc))
        _check(lib().vol_fvc_surface_integrate(self._h, nc, _p(F), _p(out), VOL_HOST))
        return out[:, 0] if nc == 1 else out

    def fvc_flux(self, U, grad_U, b_fixed, Ub):
        n, nb = self.n_cells, self.n_boundary
        U = _sized(U, 3 * n, "U")
        g = None if grad_U is None else _sized(grad_U, 9 * n, "grad_U")
        bf = _sized(b_fixed, nb, "b_fixed", np.int32)
        ub = _sized(Ub, 3 * nb, "Ub")
        out = np.zeros(max(self._nf_total(), 1))
        _check(lib().vol_fvc_flux(self._h, _p(U), _q(g), _p(bf), _p(ub), _p(out), VOL_HOST))
        return out[: self._nf_total()]

    # ---- pressure-corrector steps (IncompressibleFlow.cpp:17-29,242-362)
    def piso_rau_hbya(self, A, H):
        """(rAU, HbyA) = (1/A, rAU·H)."""
        n = self.n_cells
        A, H = _sized(A, n, "A"), _sized(H, 3 * n, "H")
        rau, hbya = np.zeros(n), np.zeros((n, 3))
        _check(lib().vol_piso_rau_hbya(self._h, _p(A), _p(H), _p(rau), _p(hbya), VOL_HOST))
        return rau, hbya

    def piso_hbya_boundary(self, b_fixed, Ub, rAU, gradp):
        """Boundary values of HbyA for its least-squares gradient: U_b + rAU_P ∇p_P on
        fixed-flux faces, 0 elsewhere."""
        n, nb = self.n_cells, self.n_boundary
        bf = _sized(b_fixed, nb, "b_fixed", np.int32)
        ub, ra, gp = _sized(Ub, 3 * nb, "Ub"), _sized(rAU, n, "rAU"), _sized(gradp, 3 * n, "gradp")
        out = np.zeros((max(nb, 1), 3))
        _check(lib().vol_piso_hbya_boundary(self._h, _p(bf), _p(ub), _p(ra), _p(gp), _p(out), VOL_HOST))
        return out[:nb]

    def piso_face_op(self, op, a, b, c=None):
        """Face-wise a + b·c, a − b or a − b/c (FACE_* constants)."""
        nF = self._nf_total()
        a, b = _sized(a, nF, "a"), _sized(b, nF, "b")
        if c is None and op != FACE_A_MINUS_B:
            raise ValueError("c is required for this face operation")
        cc = None if c is None else _sized(c, nF, "c")
        out = np.zeros(max(nF, 1))
        _check(lib().vol_piso_face_op(self._h, int(op), _p(a), _p(b), _q(cc), _p(out), VOL_HOST))
        return out[:nF]

    def piso_ddt_phi_corr(self, coeff, c0, c00, U0, U00, phi_old, phi_old_old, rAUf, phiHbyA):
        """phiHbyA with the ddtPhiCorr coupling added (coeff < 0: automatic coupling)."""
        n, nF = self.n_cells, self._nf_total()
        if phiHbyA is None:
            raise ValueError("phiHbyA is required")
        u0 = _sized(U0, 3 * n, "U0")
        u00 = None if U00 is None else _sized(U00, 3 * n, "U00")
        po = _sized(phi_old, nF, "phi_old")
        poo = None if phi_old_old is None else _sized(phi_old_old, nF, "phi_old_old")
        ra = _sized(rAUf, nF, "rAUf")
        out = _sized(phiHbyA, nF, "phiHbyA").copy()
        _check(lib().vol_piso_ddt_phi_corr(self._h, float(coeff), float(c0), float(c00), _p(u0), _q(u00), _p(po),
                                           _q(poo), _p(ra), _p(out), VOL_HOST))
        return out

    def piso_constrain_pressure(self, start, size, phiHbyA, Ub, rAUf):
        """fixedFluxPressure gradient of boundary faces [start, start + size)."""
        nF, nb = self._nf_total(), self.n_boundary
        ph, ub, ra = _sized(phiHbyA, nF, "phiHbyA"), _sized(Ub, 3 * nb, "Ub"), _sized(rAUf, nF, "rAUf")
        out = np.zeros(max(size, 1))
        _check(lib().vol_piso_constrain_pressure(self._h, int(start), int(size), _p(ph), _p(ub), _p(ra), _p(out),
                                                 VOL_HOST))
        return out[:size]

    def fvc_reconstruct(self, flux, empty_dir=-1):
        """fvc::reconstruct of a face flux: nCells × 3."""
        fl = _sized(flux, self._nf_total(), "flux")
        out = np.zeros((self.n_cells, 3))
        _check(lib().vol_fvc_reconstruct(self._h, _p(fl), int(empty_dir), _p(out), VOL_HOST))
        return out

    def piso_correct_U(self, HbyA, rAU, v, sign, empty_dir=-1):
        """U = HbyA − rAU·v (sign < 0) or HbyA + rAU·v, component empty_dir zeroed."""
        n = self.n_cells
        h, ra, vv = _sized(HbyA, 3 * n, "HbyA"), _sized(rAU, n, "rAU"), _sized(v, 3 * n, "v")
        out = np.zeros((n, 3))
        _check(lib().vol_piso_correct_U(self._h, _p(h), _p(ra), _p(vv), int(sign), int(empty_dir), _p(out), VOL_HOST))
        return out

    def gamg_setup_info(self):
        """(seconds, pairing passes, decision rounds) of the hierarchy build; passes and
        rounds are 0 when the hierarchy was built on the host."""
        s, p, r = ctypes.c_double(), ctypes.c_int(), ctypes.c_int()
        _check(lib().vol_mesh_gamg_setup_info(self._h, ctypes.byref(s), ctypes.byref(p), ctypes.byref(r)))
        return s.value, p.value, r.value

    def gamg_levels(self):
        """[{n, agg, faceMap, faceFlip}] per level, as GamgHierarchy::levels."""
        nl = ctypes.c_int()
        sizes = np.zeros(40, np.int32)
        _check(lib().vol_mesh_gamg_levels(self._h, ctypes.byref(nl), _p(sizes)))
        out = []
        for k in range(nl.value):
            nf = ctypes.c_int()
            _check(lib().vol_mesh_gamg_level(self._h, k, ctypes.byref(nf), None, None, None))
            agg = np.zeros(int(sizes[k]), np.int32)
            fm = np.zeros(max(nf.value, 1), np.int32)
            ff = np.zeros(max(nf.value, 1), np.int32)
            _check(lib().vol_mesh_gamg_level(self._h, k, ctypes.byref(nf), _p(agg), _p(fm), _p(ff)))
            out.append({"n": int(sizes[k]), "agg": agg, "faceMap": fm[: nf.value], "faceFlip": ff[: nf.value]})
        return out


def ddt_coeffs(scheme, dt, dt0=-1.0, have_old_old=False):
    """ddtCoeffs (Schemes.h:71-86): (c, c0, c00) of "steadyState", "Euler" or "backward"
    (variable-step BDF2 when the old-old level exists and dt0 > 0, Euler otherwise)."""
    if scheme == "steadyState":
        return 0.0, 0.0, 0.0
    if scheme not in ("Euler", "backward"):
        raise ValueError(f"unknown ddt scheme: {scheme}")
    rdt = 1.0 / dt
    if scheme == "backward" and have_old_old and dt0 > 0:
        ct = 1.0 + dt / (dt + dt0)
        ct00 = dt * dt / (dt0 * (dt + dt0))
        return ct * rdt, (ct + ct00) * rdt, ct00 * rdt
    return rdt, rdt, 0.0


def device_geometry(ctx: "Context", mesh):
    """Mesh geometry computed on the GPU (vol_geometry) from a voluma_amd.mesh.Mesh's
    points and faces: dict of V, C, Sf, magSf, Cf, d, w, ndc, corr, skew, skewed."""
    nc, nf, ni = mesh.n_cells, mesh.n_faces, mesh.n_internal
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    pts = np.ascontiguousarray(mesh.points, dtype=np.float64)
    g = {"V": np.zeros(nc), "C": np.zeros((nc, 3)), "Sf": np.zeros((nf, 3)), "Cf": np.zeros((nf, 3)),
         "d": np.zeros((nf, 3)), "magSf": np.zeros(nf), "w": np.zeros(nf), "ndc": np.zeros(nf),
         "corr": np.zeros((nf, 3)), "skew": np.zeros((nf, 3))}
    fo, fp, own, nei = i32(mesh.face_off), i32(mesh.face_pts), i32(mesh.owner), i32(mesh.neighbour)
    smax = ctypes.c_double()
    _check(lib().vol_geometry(ctx._h, int(pts.shape[0]), _p(pts), nf, _p(fo), _p(fp), ni, _p(own), _p(nei), nc,
                              *(_p(g[k]) for k in ("V", "C", "Sf", "Cf", "d", "magSf", "w", "ndc", "corr", "skew")),
                              ctypes.byref(smax), VOL_HOST))
    g["skewed"] = bool(smax.value > 1e-8)
    g["max_skewness"] = smax.value
    return g


class LduMatrix:
    """Coefficients over an LduMesh, resident on the GPU (vol_matrix)."""

    def __init__(self, mesh: LduMesh, diag=None, upper=None, lower=None):
        self.mesh = mesh
        h = ctypes.c_void_p()
        _check(lib().vol_matrix_create(mesh._h, ctypes.byref(h)))
        self._h = h
        if diag is not None:
            self.set(diag, upper, lower)

    def set(self, diag, upper, lower=None):
        """diag must already include the boundary contributions; lower None → symmetric."""
        if isinstance(diag, DeviceArray):
            _check(lib().vol_matrix_set(self._h, diag.ptr, upper.ptr, None if lower is None else lower.ptr,
                                        VOL_DEVICE))
            self.symmetric = lower is None
            return
        d = np.ascontiguousarray(diag, dtype=np.float64)
        u = np.ascontiguousarray(upper, dtype=np.float64)
        lo = None if lower is None else np.ascontiguousarray(lower, dtype=np.float64)
        if d.size != self.mesh.n_cells or u.size != self.mesh.n_faces or (lo is not None and lo.size != u.size):
            raise ValueError("coefficient sizes do not match the mesh")
        _check(lib().vol_matrix_set(self._h, _p(d), _p(u), None if lo is None else _p(lo), VOL_HOST))
        self.symmetric = lo is None

    def get(self):
        """(diag, upper, lower) as held on the GPU (lower == upper when symmetric)."""
        m = self.mesh
        d, u, lo = np.zeros(m.n_cells), np.zeros(max(m.n_faces, 1)), np.zeros(max(m.n_faces, 1))
        _check(lib().vol_matrix_get(self._h, _p(d), _p(u), _p(lo), VOL_HOST))
        return d, u[: m.n_faces], lo[: m.n_faces]

    def assemble_laplacian(self, magSf, ndc, bc_fixed, bc_value=None, gamma=1.0, face_gamma=None, b_in=None,
                           return_boundary=False):
        """fvm::laplacian(γ, ψ) + boundary fold on the GPU; returns the folded source b
        (b_in + boundary source) and optionally (internalCoeffs, boundarySource)."""
        m = self.mesh
        n, nb, nF = m.n_cells, m.n_boundary, m.n_faces + m.n_boundary
        opt = lambda a, cnt, name: None if a is None else _sized(a, cnt, name)
        bcf = _sized(bc_fixed, nb, "bc_fixed", np.int32)
        bv, fg, bi = opt(bc_value, nb, "bc_value"), opt(face_gamma, nF, "face_gamma"), opt(b_in, n, "b_in")
        ms, nd = _sized(magSf, nF, "magSf"), _sized(ndc, nF, "ndc")
        b = np.zeros(m.n_cells)
        ic = np.zeros(max(m.n_boundary, 1))
        bs = np.zeros(max(m.n_boundary, 1))
        q = lambda a: None if a is None else _p(a)
        _check(lib().vol_assemble_laplacian(self._h, float(gamma), q(fg), _p(ms), _p(nd), _p(bcf), q(bv), q(bi),
                                            _p(b), _p(ic), _p(bs), VOL_HOST))
        self.symmetric = True
        if return_boundary:
            return b, ic[: m.n_boundary], bs[: m.n_boundary]
        return b

    def assemble_div_upwind(self, flux, bc_fixed, bc_value=None, b_in=None, return_boundary=False):
        """fvm::div(F, ψ, upwind) + boundary fold on the GPU (nonsymmetric)."""
        m = self.mesh
        n, nb, nF = m.n_cells, m.n_boundary, m.n_faces + m.n_boundary
        opt = lambda a, cnt, name: None if a is None else _sized(a, cnt, name)
        bcf = _sized(bc_fixed, nb, "bc_fixed", np.int32)
        F, bv, bi = _sized(flux, nF, "flux"), opt(bc_value, nb, "bc_value"), opt(b_in, n, "b_in")
        b = np.zeros(m.n_cells)
        ic = np.zeros(max(m.n_boundary, 1))
        bs = np.zeros(max(m.n_boundary, 1))
        q = lambda a: None if a is None else _p(a)
        _check(lib().vol_assemble_div_upwind(self._h, _p(F), _p(bcf), q(bv), q(bi), _p(b), _p(ic), _p(bs), VOL_HOST))
        self.symmetric = False
        if return_boundary:
            return b, ic[: m.n_boundary], bs[: m.n_boundary]
        return b

    # ---- FvMatrix algebra on the GPU (host arrays)
    def fvm_reset(self):
        """FvMatrix(psi): zero diag/upper/source/internalCoeffs/boundarySource/ffc."""
        _check(lib().vol_fvm_reset(self._h))
        self._fvm_lower = False

    def fvm_ddt_euler(self, sign, dt, V, psi0):
        n = self.mesh.n_cells
        _check(lib().vol_fvm_ddt_euler(self._h, int(sign), ctypes.c_double(dt), _p(_sized(V, n, "V")),
                                       _p(_sized(psi0, n, "psi0")), VOL_HOST))

    def fvm_ddt(self, sign, coeffs, V, psi0, psi00=None):
        """fvm::ddt with ddtCoeffs (c, c0, c00): diag += c·V, source += V·(c0·ψ⁰ − c00·ψ⁰⁰)."""
        c, c0, c00 = (float(v) for v in coeffs)
        n = self.mesh.n_cells
        p00 = None if psi00 is None else _sized(psi00, n, "psi00")
        _check(lib().vol_fvm_ddt(self._h, int(sign), c, c0, c00, _p(_sized(V, n, "V")), _p(_sized(psi0, n, "psi0")),
                                 _q(p00), VOL_HOST))

    def fvm_div_upwind(self, sign, flux, bc_fixed, bc_value=None):
        m = self.mesh
        nb = m.n_boundary
        bcf = _sized(bc_fixed, nb, "bc_fixed", np.int32)
        bv = None if bc_value is None else _sized(bc_value, nb, "bc_value")
        _check(lib().vol_fvm_div_upwind(self._h, int(sign), _p(_sized(flux, m.n_faces + nb, "flux")), _p(bcf), _q(bv),
                                        VOL_HOST))
        self._fvm_lower = True

    def fvm_laplacian(self, sign, magSf, ndc, bc_fixed, bc_value=None, gamma=1.0, face_gamma=None, nonorth=None):
        """nonorth: dict with psi, w, Sf (nFaces×3), corr (nFaces×3), V, or None."""
        m = self.mesh
        n, nb, nF = m.n_cells, m.n_boundary, m.n_faces + m.n_boundary
        bcf = _sized(bc_fixed, nb, "bc_fixed", np.int32)
        no = None
        keep = []
        if nonorth is not None:
            sizes = {"psi": n, "w": nF, "Sf": 3 * nF, "corr": 3 * nF, "V": n}
            arrs = [_sized(nonorth[k], sizes[k], "nonorth[%s]" % k) for k in ("psi", "w", "Sf", "corr", "V")]
            keep = arrs
            no = _NonOrth(*(a.ctypes.data for a in arrs))
        ms, nd = _sized(magSf, nF, "magSf"), _sized(ndc, nF, "ndc")
        fg = None if face_gamma is None else _sized(face_gamma, nF, "face_gamma")
        bv = None if bc_value is None else _sized(bc_value, nb, "bc_value")
        _check(lib().vol_fvm_laplacian(self._h, int(sign), ctypes.c_double(gamma), _q(fg), _p(ms), _p(nd), _p(bcf),
                                       _q(bv), None if no is None else ctypes.byref(no), VOL_HOST))
        del keep

    def fvm_source(self, su, V):
        """FvMatrix −= su: source += su·V."""
        n = self.mesh.n_cells
        _check(lib().vol_fvm_source(self._h, _p(_sized(su, n, "su")), _p(_sized(V, n, "V")), VOL_HOST))

    def fvm_get(self):
        m = self.mesh
        nb = m.n_boundary
        out = {"diag": np.zeros(m.n_cells), "source": np.zeros(m.n_cells), "ic": np.zeros(max(nb, 1)),
               "bs": np.zeros(max(nb, 1)), "ffc": np.zeros(max(m.n_faces + nb, 1))}
        _check(lib().vol_fvm_get(self._h, *(_p(out[k]) for k in ("diag", "source", "ic", "bs", "ffc")), VOL_HOST))
        out["ic"], out["bs"], out["ffc"] = out["ic"][:nb], out["bs"][:nb], out["ffc"][: m.n_faces + nb]
        return out

    def fvm_finish(self):
        """Fold and make the matrix solvable; returns b (source + folded boundary source)."""
        b = np.zeros(self.mesh.n_cells)
        _check(lib().vol_fvm_finish(self._h, _p(b), VOL_HOST))
        self.symmetric = not getattr(self, "_fvm_lower", False)
        return b

    def fvm_finish_source(self, su, V):
        """Fold a copy of the matrix with −= su (source + su·V); the matrix keeps its own
        source. Returns b."""
        b = np.zeros(self.mesh.n_cells)
        n = self.mesh.n_cells
        _check(lib().vol_fvm_finish_source(self._h, _p(_sized(su, n, "su")), _p(_sized(V, n, "V")), _p(b), VOL_HOST))
        self.symmetric = not getattr(self, "_fvm_lower", False)
        return b

    # ---- pressure-velocity coupling (FvMatrix::relax / setReference / A / H / flux)
    def fvm_relax(self, alpha, psi):
        _check(lib().vol_fvm_relax(self._h, ctypes.c_double(alpha), _p(_sized(psi, self.mesh.n_cells, "psi")),
                                   VOL_HOST))

    def fvm_set_reference(self, cell, value):
        _check(lib().vol_fvm_set_reference(self._h, int(cell), ctypes.c_double(value)))

    def fvm_A(self, V, n_components=1):
        out = np.zeros(self.mesh.n_cells)
        _check(lib().vol_fvm_A(self._h, _p(_sized(V, self.mesh.n_cells, "V")), int(n_components), _p(out), VOL_HOST))
        return out

    def fvm_H(self, psi, V, n_components=1):
        out = np.zeros(self.mesh.n_cells)
        n = self.mesh.n_cells
        _check(lib().vol_fvm_H(self._h, _p(_sized(psi, n, "psi")), _p(_sized(V, n, "V")), int(n_components), _p(out),
                               VOL_HOST))
        return out

    def fvm_flux(self, psi):
        m = self.mesh
        out = np.zeros(max(m.n_faces + m.n_boundary, 1))
        _check(lib().vol_fvm_flux(self._h, _p(_sized(psi, m.n_cells, "psi")), _p(out), VOL_HOST))
        return out[: m.n_faces + m.n_boundary]

    def close(self):
        if getattr(self, "_h", None):
            lib().vol_matrix_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def amul(A: LduMatrix, x, y=None):
    """y = A x (LinearSolver.cpp:121-133). numpy in → numpy out, DeviceArray in → in place."""
    if isinstance(x, DeviceArray):
        _check(lib().vol_amul(A._h, x.ptr, y.ptr, VOL_DEVICE))
        return y
    xx = _sized(x, A.mesh.n_cells, "x")
    yy = np.empty_like(xx)
    _check(lib().vol_amul(A._h, _p(xx), _p(yy), VOL_HOST))
    return yy


class LduBatch:
    """Independent systems solved together (vol_batch): each system keeps its own GAMG
    hierarchy, Krylov scalars and convergence test; one launch per phase serves all.
    systems: sequence of (n_cells, owner, neighbour, gamg_weights) with each system's
    own cell numbers."""

    def __init__(self, ctx: Context, systems):
        self.ctx = ctx
        self.n_cells = np.array([int(s[0]) for s in systems], np.int32)
        self.n_faces = np.array([len(s[1]) for s in systems], np.int32)
        own = np.ascontiguousarray(np.concatenate([np.asarray(s[1], np.int32) for s in systems]), dtype=np.int32)
        nei = np.ascontiguousarray(np.concatenate([np.asarray(s[2], np.int32) for s in systems]), dtype=np.int32)
        w = _f64(np.concatenate([np.asarray(s[3], np.float64) for s in systems]))
        h = ctypes.c_void_p()
        _check(lib().vol_batch_create(ctx._h, len(systems), _p(self.n_cells), _p(self.n_faces), _p(own), _p(nei),
                                      _p(w), ctypes.byref(h)))
        self._h = h
        self.n_systems = len(systems)
        self.cell_off = np.concatenate([[0], np.cumsum(self.n_cells, dtype=np.int64)])
        self.face_off = np.concatenate([[0], np.cumsum(self.n_faces, dtype=np.int64)])
        self.n_total = int(self.cell_off[-1])

    def graph_stats(self, reset=True):
        """(graph replays, eager iterations) of the batched iterations since the last reset."""
        g, e = ctypes.c_int(), ctypes.c_int()
        _check(lib().vol_batch_graph_stats(self._h, ctypes.byref(g), ctypes.byref(e), int(bool(reset))))
        return g.value, e.value

    def levels(self):
        lv = np.zeros(self.n_systems, np.int32)
        _check(lib().vol_batch_info(self._h, None, None, _p(lv)))
        return lv

    def set(self, diag, upper, lower=None):
        """Concatenated diag (boundary-folded) / upper / lower of all systems."""
        if isinstance(diag, DeviceArray):
            _check(lib().vol_batch_set(self._h, diag.ptr, upper.ptr, None if lower is None else lower.ptr, VOL_DEVICE))
            return
        d, u = _f64(diag), _f64(upper)
        lo = None if lower is None else _f64(lower)
        if d.size != self.n_total or u.size != int(self.face_off[-1]):
            raise ValueError("coefficient sizes do not match the batch")
        _check(lib().vol_batch_set(self._h, _p(d), _p(u), None if lo is None else _p(lo), VOL_HOST))

    def solve(self, b, x, controls: SolverControls | None = None):
        """Solve every system; x (concatenated, in/out). Returns one SolverPerformance per system."""
        ctl = controls or SolverControls()
        c = ctl._c()
        pf = (_Perf * self.n_systems)()
        if isinstance(b, DeviceArray):
            _check(lib().vol_batch_solve(self._h, b.ptr, x.ptr, ctypes.byref(c), ctypes.cast(pf, ctypes.c_void_p),
                                         VOL_DEVICE))
        else:
            bb = _f64(b)
            if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
                raise TypeError("x must be a contiguous float64 array")
            _check(lib().vol_batch_solve(self._h, _p(bb), _p(x), ctypes.byref(c), ctypes.cast(pf, ctypes.c_void_p),
                                         VOL_HOST))
        return [SolverPerformance(p.solver.decode(), "", p.initialResidual, p.finalResidual, p.residualSum,
                                  p.normFactor, p.iterations, bool(p.converged), p.outerIterations) for p in pf]

    def close(self):
        if getattr(self, "_h", None):
            lib().vol_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Server:
    """Solve server (vol_server): `slots` workers, each with its own HIP stream, run
    admitted requests concurrently; meshes (and their GAMG hierarchies) are registered
    once and kept resident, a request's matrix is reused by its right-hand sides."""

    def __init__(self, device=0, slots=4, hbm_budget=200e9):
        h = ctypes.c_void_p()
        _check(lib().vol_server_create(int(device), int(slots), ctypes.c_uint64(int(hbm_budget)), ctypes.byref(h)))
        self._h = h
        self._live = {}

    def add_mesh(self, n_cells, owner, neighbour, gamg_weights=None, boundary_owner=None):
        own = np.ascontiguousarray(owner, dtype=np.int32)
        nei = np.ascontiguousarray(neighbour, dtype=np.int32)
        w = None if gamg_weights is None else _f64(gamg_weights)
        bo = None if boundary_owner is None else np.ascontiguousarray(boundary_owner, dtype=np.int32)
        mid = ctypes.c_int()
        _check(lib().vol_server_add_mesh(self._h, int(n_cells), int(own.size), _p(own), _p(nei), _q(w), _q(bo),
                                         0 if bo is None else int(bo.size), ctypes.byref(mid)))
        return mid.value

    def submit(self, mesh_id, diag, upper, lower, bs, x0s=None, controls: SolverControls | None = None):
        """Queue a request; returns a ticket. bs: list of right-hand sides."""
        d, u = _f64(diag), _f64(upper)
        lo = None if lower is None else _f64(lower)
        bs = [_f64(b) for b in bs]
        xs = [np.zeros(b.size) if x0s is None else np.array(x0s[k], dtype=np.float64) for k, b in enumerate(bs)]
        ctl = controls or SolverControls()
        c = ctl._c()
        bp = (ctypes.c_void_p * len(bs))(*[b.ctypes.data for b in bs])
        xp = (ctypes.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
        pf = (_Perf * len(bs))()
        r = _Request(int(mesh_id), d.ctypes.data, u.ctypes.data, None if lo is None else lo.ctypes.data, len(bs),
                     ctypes.cast(bp, ctypes.c_void_p), ctypes.cast(xp, ctypes.c_void_p), c,
                     ctypes.cast(pf, ctypes.c_void_p))
        t = ctypes.c_int64()
        _check(lib().vol_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        self._live[t.value] = (d, u, lo, bs, xs, ctl, bp, xp, pf, r)  # alive until wait()
        return t.value

    def prefetch(self, mesh_id, timeout=600.0):
        """Make a registered mesh resident on every worker (copy, matrix storage, batch of
        copies) so that its first requests pay no set-up; blocks until done."""
        _check(lib().vol_server_prefetch(self._h, int(mesh_id), ctypes.c_double(timeout)))

    def wait(self, ticket, timeout=600.0):
        """(solutions, [SolverPerformance]) of a request; raises VolumaError if it failed."""
        st = ctypes.c_int()
        tm = np.zeros(3)
        try:
            _check(lib().vol_server_wait(self._h, ctypes.c_int64(ticket), ctypes.c_double(timeout), ctypes.byref(st),
                                         _p(tm)))
        except VolumaError as e:
            if e.code != -9:
                self._live.pop(ticket, None)
            raise
        keep = self._live.pop(ticket)
        xs, pf = keep[4], keep[8]
        self.last_times = {"queued": tm[0], "started": tm[1], "finished": tm[2]}
        return xs, [SolverPerformance(p.solver.decode(), "", p.initialResidual, p.finalResidual, p.residualSum,
                                      p.normFactor, p.iterations, bool(p.converged), p.outerIterations) for p in pf]

    def stats(self):
        b, q, c = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_int()
        _check(lib().vol_server_stats(self._h, ctypes.byref(b), ctypes.byref(q), ctypes.byref(c)))
        return {"bytes_in_use": b.value, "queued": q.value, "cached": c.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().vol_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def solve_ldu(A: LduMatrix, b, x, controls: SolverControls = None, field="") -> SolverPerformance:
    """solveLdu (LinearSolver.cpp:135-285): x (numpy float64 or DeviceArray) is updated in place."""
    controls = controls or SolverControls()
    c = controls._c()
    pf = _Perf()
    if isinstance(x, DeviceArray):
        _check(lib().vol_solve(A._h, b.ptr, x.ptr, ctypes.byref(c), ctypes.byref(pf), VOL_DEVICE))
    else:
        if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
            raise TypeError("x must be a contiguous float64 numpy array (updated in place)")
        bb = np.ascontiguousarray(b, dtype=np.float64)
        _check(lib().vol_solve(A._h, _p(bb), _p(x), ctypes.byref(c), ctypes.byref(pf), VOL_HOST))
    return SolverPerformance(pf.solver.decode(), field, pf.initialResidual, pf.finalResidual, pf.residualSum,
                             pf.normFactor, pf.iterations, bool(pf.converged), pf.outerIterations)


solveLdu = solve_ldu
