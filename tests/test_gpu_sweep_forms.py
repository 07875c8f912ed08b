"""GPU tests of the two forms of the triangular sweeps (Gauss–Seidel solver and GAMG
smoother, DIC/DILU substitutions): level-scheduled (a launch per level or per run of
short levels; the default, VOLUMA_SWEEP=levels) and sync-free (one launch per sweep, each
row waiting for the new values of its neighbours on the sweep's side; VOLUMA_SWEEP=free).
Each row's arithmetic is the same in both, so solutions, iteration counts and residuals
must be bit-identical, in FP64 and mixed precision."""
import os

import numpy as np
import pytest

import voluma_amd as va
from voluma_amd import mesh as M
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu


def _context(sweep):
    old = os.environ.get("VOLUMA_SWEEP")
    try:
        os.environ["VOLUMA_SWEEP"] = sweep
        return va.Context(0)
    finally:
        if old is None:
            os.environ.pop("VOLUMA_SWEEP", None)
        else:
            os.environ["VOLUMA_SWEEP"] = old


def _gs_gamg(prec, maxit):
    return va.SolverControls("PCG", "GAMG", 1e-9, 0.0, maxit, precision=prec, gamgSmoother=va.GAMG_GAUSS_SEIDEL,
                             gamgPreSweeps=2, gamgPostSweeps=2)


PRESSURE = [("gs_gamg_fp64", lambda: _gs_gamg(va.FP64, 12)), ("gs_gamg_mixed", lambda: _gs_gamg(va.MIXED, 12)),
            ("dic_pcg", lambda: va.SolverControls("PCG", "DIC", 1e-8, 0.0, 60)),
            ("gauss_seidel", lambda: va.SolverControls("GaussSeidel", "none", 1e-12, 0.0, 4, nSweeps=2)),
            ("sym_gauss_seidel", lambda: va.SolverControls("symGaussSeidel", "none", 1e-12, 0.0, 3, nSweeps=1))]
MOMENTUM = [("dilu_bicgstab", lambda: va.SolverControls("PBiCGStab", "DILU", 1e-10, 0.0, 50)),
            ("dilu_bicgstab_mixed", lambda: va.SolverControls("PBiCGStab", "DILU", 1e-10, 0.0, 50, precision=va.MIXED))]


def _solves(sweep, m):
    ctx = _context(sweep)
    out = {}
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True, nonorth=False)
    for name, ctl in PRESSURE:
        x = np.zeros(m.n_cells)
        out[name] = (x, va.solve_ldu(A, b, x, ctl()))
        print(f"[sweeps]   {sweep} {name}: {out[name][1].iterations} iterations", flush=True)
    A.close()
    lm.close()
    lm, A, b, inp = W.momentum_system(ctx, m, seed=1, comp=0)
    for name, ctl in MOMENTUM:
        x = np.ascontiguousarray(inp["U0"][:, 0]).copy()
        out[name] = (x, va.solve_ldu(A, b, x, ctl()))
        print(f"[sweeps]   {sweep} {name}: {out[name][1].iterations} iterations", flush=True)
    A.close()
    lm.close()
    ctx.close()
    return out


@pytest.mark.parametrize("mesh", ["tet", "box"])
def test_sweep_forms_bit_identical(mesh):
    m = M.tet(30, 28, 26, 0.2, 3) if mesh == "tet" else M.box(48, 40, 36)
    print(f"[sweeps] {mesh}: {m.n_cells} cells, level-scheduled", flush=True)
    ref = _solves("levels", m)
    print(f"[sweeps] {mesh}: sync-free", flush=True)
    got = _solves("free", m)
    for name in ref:
        (xr, pr), (xg, pg) = ref[name], got[name]
        assert pg.iterations == pr.iterations and pg.finalResidual == pr.finalResidual, name
        assert np.array_equal(xg, xr), name
        print(f"[sweeps] {mesh} {name}: {pr.iterations} iterations, same bits", flush=True)
