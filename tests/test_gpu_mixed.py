"""GPU tests of mixed precision (VOL_MIXED): FP64 defect correction around an FP32
inner Krylov solve. No reference counterpart exists; the bar is the FP64 solution of
the same system: at normalised residual 1e-10 the mixed solution is within 1e-8
relative of the FP64 oracle's tight solution, and every solve's FP64 residual,
recomputed on the host side with the bit-exact FP64 Amul, meets the tolerance.
"""
import numpy as np
import pytest

import voluma_amd as va
from helpers import SMALL_MESHES, load_small
from oracle import oracle as O
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu
NAMES = list(SMALL_MESHES)


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _system(ctx, name, sym=True):
    g = load_small(name)
    m = SMALL_MESHES[name]()
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights())
    if sym:
        A = va.LduMatrix(lm, g["sym_diag"], g["sym_upper"])
        return g, m, lm, A, g["sym_b"]
    A = va.LduMatrix(lm, g["ns_diag"], g["ns_upper"], g["ns_lower"])
    return g, m, lm, A, g["ns_b"]


def _resid(A, b, x, perf):
    """Normalised FP64 residual Σ|b − A·x| / normFactor (LinearSolver.cpp:150-169)."""
    return np.abs(b - va.amul(A, x)).sum() / perf.normFactor


def _mixed(solver, pre, tol, **kw):
    return va.SolverControls(solver, pre, tol, 0.0, 2000, precision=va.MIXED, **kw)


@pytest.mark.parametrize("name", NAMES)
def test_mixed_gamg_pcg_vs_fp64_oracle(ctx, name):
    g, m, lm, A, b = _system(ctx, name)
    om = O.LduMesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    xr, pr = om.solve(g["sym_diag"], g["sym_upper"], None, b, preconditioner="GAMG", tolerance=1e-13, max_iter=500)
    assert pr["converged"]
    for tol, err in ((1e-6, 1e-3), (1e-10, 1e-8)):
        x = np.zeros(b.size)
        p = va.solve_ldu(A, b, x, _mixed("PCG", "GAMG", tol))
        assert p.converged and p.outerIterations >= 1
        res = _resid(A, b, x, p)
        assert res < tol and abs(res - p.finalResidual) <= 1e-6 * tol + 1e-3 * res
        assert np.linalg.norm(x - xr) <= err * np.linalg.norm(xr), (tol, np.linalg.norm(x - xr) / np.linalg.norm(xr))


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("pre", ["diagonal", "GAMG"])
def test_mixed_pbicgstab_nonsymmetric(ctx, name, pre):
    g, m, lm, A, b = _system(ctx, name, sym=False)
    xr = np.zeros(b.size)
    pr = va.solve_ldu(A, b, xr, va.SolverControls("PBiCGStab", pre, 1e-14, 0.0, 2000))
    x = np.zeros(b.size)
    p = va.solve_ldu(A, b, x, _mixed("PBiCGStab", pre, 1e-10))
    assert p.converged and _resid(A, b, x, p) < 1e-10
    assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr), (pr.finalResidual, p.outerIterations)


def test_mixed_diagonal_pcg_and_inner_tolerance(ctx):
    """Any inner reduction gives the FP64 answer; a looser one takes more outer steps."""
    g, m, lm, A, b = _system(ctx, "box12")
    outers = []
    for rel in (1e-2, 1e-4):
        x = np.zeros(b.size)
        p = va.solve_ldu(A, b, x, _mixed("PCG", "diagonal", 1e-10, innerRelTol=rel))
        assert p.converged and _resid(A, b, x, p) < 1e-10
        outers.append(p.outerIterations)
    assert outers[0] > outers[1] >= 2


def test_mixed_errors_and_early_exit(ctx):
    g, m, lm, A, b = _system(ctx, "tet5x4x3")
    x = np.zeros(b.size)
    for ctl, code in ((_mixed("PCG", "GAMG", 1e-6, sums=va.SUMS_REPRODUCIBLE), -6),
                      (_mixed("Jacobi", "none", 1e-6), -6),
                      (_mixed("PCG", "GAMG", 1e-6, innerRelTol=0.0), -1),
                      (_mixed("PCG", "GAMG", 1e-6, innerRelTol=1.0), -1),
                      (va.SolverControls("PCG", "GAMG", precision=7), -1)):
        with pytest.raises(va.VolumaError) as e:
            va.solve_ldu(A, b, x, ctl)
        assert e.value.code == code
    # already converged: no outer step
    p = va.solve_ldu(A, b, x, _mixed("PCG", "GAMG", 1e30))
    assert p.converged and p.iterations == 0 and p.outerIterations == 0
    # iteration budget is shared by the inner solves
    p = va.solve_ldu(A, b, np.zeros(b.size), va.SolverControls("PCG", "diagonal", 1e-14, 0.0, 7, precision=va.MIXED))
    assert p.iterations == 7 and not p.converged


def test_mixed_resident_matrix_fp64_switch(ctx):
    """The FP32 copies follow vol_matrix_set: after new coefficients the mixed solve
    answers the new system."""
    g, m, lm, A, b = _system(ctx, "box12")
    x1 = np.zeros(b.size)
    va.solve_ldu(A, b, x1, _mixed("PCG", "GAMG", 1e-10))
    A.set(2.0 * g["sym_diag"], 2.0 * g["sym_upper"])
    x2 = np.zeros(b.size)
    p = va.solve_ldu(A, b, x2, _mixed("PCG", "GAMG", 1e-10))
    assert p.converged and _resid(A, b, x2, p) < 1e-10
    assert np.linalg.norm(2.0 * x2 - x1) <= 1e-8 * np.linalg.norm(x1)


def test_mixed_full_size_tet70(ctx):
    """2.06M-cell tet pressure system: mixed at 1e-6 and 1e-10 against FP64 GAMG-PCG
    solved to 1e-12."""
    m = W.make_mesh("tet70")
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True)
    xr = np.zeros(m.n_cells)
    pr = va.solve_ldu(A, b, xr, va.SolverControls("PCG", "GAMG", 1e-12, 0.0, 2000))
    assert pr.converged
    for tol, err in ((1e-6, 1e-3), (1e-10, 1e-8)):
        x = np.zeros(m.n_cells)
        p = va.solve_ldu(A, b, x, _mixed("PCG", "GAMG", tol))
        assert p.converged and _resid(A, b, x, p) < tol
        rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        print(f"tet70 mixed tol {tol:g}: iterations {p.iterations}, outer {p.outerIterations}, rel err {rel:.3e}")
        assert rel <= err
