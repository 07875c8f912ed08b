"""GPU test of vol_solve_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory): called through the C ABI on device buffers, its
results are bit-identical to vol_solve's on the same inputs, for two calls in a row of
which the second has a smaller system, in fast, exact and reproducible sums; sizes
beyond the scratch limit and unsupported preconditioners are refused."""
import ctypes

import numpy as np
import pytest

import voluma_amd as va
from voluma_amd import mesh as M
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _noalloc(A, db, dx, controls):
    c = controls._c()
    p = va._Perf()
    va._check(va.lib().vol_solve_noalloc(A._h, db.ptr, dx.ptr, ctypes.byref(c), ctypes.byref(p)))
    return p


def _solve(A, db, dx, controls):
    c = controls._c()
    p = va._Perf()
    va._check(va.lib().vol_solve(A._h, db.ptr, dx.ptr, ctypes.byref(c), ctypes.byref(p), va.VOL_DEVICE))
    return p


def _perf_tuple(p):
    return (p.iterations, p.converged, p.initialResidual, p.finalResidual, p.residualSum, p.normFactor)


def _system(ctx, m, nonsym):
    if nonsym:
        lm, A, b, inp = W.momentum_system(ctx, m, seed=1, comp=0)
        x0 = np.ascontiguousarray(inp["U0"][:, 0])
    else:
        lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=False)
        x0 = np.zeros(m.n_cells)
    return lm, A, b, x0


CASES = [("PCG", "diagonal", va.SUMS_FAST, False), ("PCG", "none", va.SUMS_FAST, False),
         ("PCG", "diagonal", va.SUMS_EXACT, False), ("PCG", "DIC", va.SUMS_REPRODUCIBLE, False),
         ("PBiCGStab", "diagonal", va.SUMS_FAST, True), ("PBiCGStab", "DILU", va.SUMS_REPRODUCIBLE, True)]


@pytest.mark.parametrize("solver,pre,sums,nonsym", CASES)
def test_noalloc_matches_vol_solve_two_calls(ctx, solver, pre, sums, nonsym):
    """Large call first (box64, the limit's size), then a smaller tet mesh."""
    for m in (W.make_mesh("box64"), M.tet(12, 10, 9, 0.2, 3)):
        print(f"[noalloc] {solver}/{pre} sums {sums}: {m.n_cells} cells", flush=True)
        lm, A, b, x0 = _system(ctx, m, nonsym)
        ctl = va.SolverControls(solver, pre, 1e-8, 0.0, 3000, sums=sums)
        db = va.DeviceArray.from_host(ctx, b)
        dx1 = va.DeviceArray.from_host(ctx, x0)
        dx2 = va.DeviceArray.from_host(ctx, x0)
        p1 = _solve(A, db, dx1, ctl)
        p2 = _noalloc(A, db, dx2, ctl)
        assert p1.converged and _perf_tuple(p1) == _perf_tuple(p2)
        assert np.array_equal(dx1.download(), dx2.download())
        for a in (db, dx1, dx2):
            a.free()
        A.close()
        lm.close()


def test_noalloc_limits(ctx):
    m = M.box(65, 64, 64)  # 266,240 cells > 262,144
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=False)
    db, dx = va.DeviceArray.from_host(ctx, b), va.DeviceArray(ctx, m.n_cells)
    with pytest.raises(va.VolumaError) as e:
        _noalloc(A, db, dx, va.SolverControls("PCG", "diagonal"))
    assert e.value.code == -1 and "262144" in str(e.value)  # VOL_ERR_INVALID_ARGUMENT
    for a in (db, dx):
        a.free()
    A.close()
    lm.close()
    m = M.box(10)
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True)
    db, dx = va.DeviceArray.from_host(ctx, b), va.DeviceArray(ctx, m.n_cells)
    for ctl in (va.SolverControls("PCG", "GAMG"), va.SolverControls("PCG", "DIC"),
                va.SolverControls("GaussSeidel", "none"), va.SolverControls("PCG", "diagonal", precision=va.MIXED)):
        with pytest.raises(va.VolumaError) as e:
            _noalloc(A, db, dx, ctl)
        assert e.value.code == -6  # VOL_ERR_UNSUPPORTED
    for a in (db, dx):
        a.free()
    A.close()
    lm.close()
