"""GPU tests of the level-scheduled triangular sweeps: the DIC/DILU preconditioner
(LinearSolver.cpp:56-105) and the Gauss–Seidel / symmetric Gauss–Seidel solvers
(LinearSolver.cpp:238-278). Bar: with exact sums, every solve is bit-identical to the
oracle's sequential restatement (which the CPU tests pin to the reference's own
solves); at full size the DILU-preconditioned PBiCGStab of the momentum matrix, the
reference's default momentum solver, converges like the reference's.
"""
import numpy as np
import pytest

import voluma_amd as va
from helpers import SMALL_MESHES, load_json, load_small
from oracle import oracle as O
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu
NAMES = list(SMALL_MESHES)

CASES = [("PCG", "DIC", True), ("PBiCGStab", "DILU", False), ("PBiCGStab", "DIC", False),
         ("PBiCGStab", "DILU", True), ("GaussSeidel", "none", True), ("symGaussSeidel", "none", True),
         ("GaussSeidel", "none", False)]


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", NAMES)
def test_sweeps_exact_mode_bit_exact_vs_oracle(ctx, name):
    g = load_small(name)
    m = SMALL_MESHES[name]()
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
    om = O.LduMesh(m.n_cells, m.owner[:ni], m.neighbour, None)
    mats = {True: (g["sym_diag"], g["sym_upper"], None, g["sym_b"]),
            False: (g["ns_diag"], g["ns_upper"], g["ns_lower"], g["ns_b"])}
    for solver, pre, sym in CASES:
        d, u, lo, b = mats[sym]
        A = va.LduMatrix(lm, d, u, lo)
        for nsw in ((1, 2) if "Gauss" in solver else (1,)):
            tol, mx = (1e-10, 400) if "Gauss" not in solver else (1e-6, 60)
            xo, po = om.solve(d, u, lo, b, solver=solver, preconditioner=pre, tolerance=tol, max_iter=mx,
                              n_sweeps=nsw, mode=O.MODE_EXACT)
            x = np.zeros(b.size)
            p = va.solve_ldu(A, b, x, va.SolverControls(solver, pre, tol, 0.0, mx, nSweeps=nsw, sums=va.SUMS_EXACT))
            assert p.iterations == po["iterations"], (solver, pre, sym, nsw)
            assert p.finalResidual == po["finalResidual"] and p.initialResidual == po["initialResidual"]
            assert np.array_equal(x, xo), (solver, pre, sym, nsw)
        A.close()
    lm.close()


@pytest.mark.parametrize("name", NAMES)
def test_dilu_fast_mode_and_reuse(ctx, name):
    """Fast sums: the same number of iterations as exact sums within ±1 and the same
    solution within 1e-9; a new matrix on the same mesh refactors (no stale factor)."""
    g = load_small(name)
    m = SMALL_MESHES[name]()
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
    A = va.LduMatrix(lm, g["ns_diag"], g["ns_upper"], g["ns_lower"])
    b = g["ns_b"]
    xe, xf = np.zeros(b.size), np.zeros(b.size)
    pe = va.solve_ldu(A, b, xe, va.SolverControls("PBiCGStab", "DILU", 1e-11, 0.0, 400, sums=va.SUMS_EXACT))
    pf = va.solve_ldu(A, b, xf, va.SolverControls("PBiCGStab", "DILU", 1e-11, 0.0, 400))
    assert abs(pe.iterations - pf.iterations) <= 1 and pf.converged
    assert np.linalg.norm(xf - xe) <= 1e-9 * np.linalg.norm(xe)
    # same mesh, another matrix: scaled coefficients give the scaled-back solution
    A.set(2.0 * g["ns_diag"], 2.0 * g["ns_upper"], 2.0 * g["ns_lower"])
    x2 = np.zeros(b.size)
    p2 = va.solve_ldu(A, 2.0 * b, x2, va.SolverControls("PBiCGStab", "DILU", 1e-11, 0.0, 400, sums=va.SUMS_EXACT))
    assert p2.iterations == pe.iterations and np.array_equal(x2, xe)
    # mixed precision with the DILU preconditioner in FP32
    xm = np.zeros(b.size)
    pm = va.solve_ldu(A, 2.0 * b, xm, va.SolverControls("PBiCGStab", "DILU", 1e-10, 0.0, 400, precision=va.MIXED))
    assert pm.converged and np.linalg.norm(xm - xe) <= 1e-8 * np.linalg.norm(xe)
    A.close()
    lm.close()


def test_full_size_momentum_dilu_tet70(ctx):
    """The reference's default momentum solver (PBiCGStab + DILU, plain sums) on the
    full-size momentum matrix: the reference's iteration count (±1) and the solution
    within 1e-9 relative of the reference's at tolerance 1e-10."""
    f = load_json("cfg2_tet70_momentum.json")
    m = W.make_mesh("tet70")
    lm, A, b, inp = W.momentum_system(ctx, m, seed=1, comp=0)
    ref = f["components"][0]["plain_pbicgstab_dilu_1e-10"]
    x = np.ascontiguousarray(inp["U0"][:, 0]).copy()
    p = va.solve_ldu(A, b, x, va.SolverControls("PBiCGStab", "DILU", 1e-10, 0.0, 1000))
    assert p.converged and abs(p.iterations - ref["iterations"]) <= 1, (p.iterations, ref["iterations"])
    assert p.initialResidual == pytest.approx(ref["initialResidual"], rel=1e-10)
    idx = np.array(ref["x"]["sample_idx"])
    xs = np.array(ref["x"]["sample"])
    assert np.abs(x[idx] - xs).max() <= 1e-9 * np.abs(xs).max()
    A.close()
    lm.close()


@pytest.mark.parametrize("name", NAMES)
def test_gamg_gauss_seidel_exact_mode_bit_exact_vs_oracle(ctx, name):
    """GAMG with the reference's smoother (Gauss–Seidel, 2 forward + 2 backward sweeps,
    row-ordered coarsest substitution): with exact sums the PCG and PBiCGStab solves
    are bit-identical to the oracle's restatement of the reference's GAMG."""
    g = load_small(name)
    m = SMALL_MESHES[name]()
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights())
    om = O.LduMesh(m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights())
    d, u, b = g["sym_diag"], g["sym_upper"], g["sym_b"]
    A = va.LduMatrix(lm, d, u)
    for solver in ("PCG", "PBiCGStab"):
        xo, po = om.solve(d, u, None, b, solver=solver, preconditioner="GAMG", tolerance=1e-10, max_iter=200,
                          mode=O.MODE_EXACT, smoother=O.SMOOTHER_GS)
        x = np.zeros(b.size)
        p = va.solve_ldu(A, b, x, va.SolverControls(solver, "GAMG", 1e-10, 0.0, 200, sums=va.SUMS_EXACT,
                                                    gamgSmoother=va.GAMG_GAUSS_SEIDEL))
        assert p.iterations == po["iterations"] and p.finalResidual == po["finalResidual"], solver
        assert np.array_equal(x, xo), solver
    A.close()
    lm.close()


@pytest.mark.parametrize("name", NAMES)
def test_gamg_gauss_seidel_fast_mode_vs_reference(ctx, name):
    """Fast sums with the Gauss–Seidel GAMG smoother against the reference's own GAMG
    solves (plain sums): iterations within ±1, solution within 1e-8; mixed precision
    with the same smoother converges to the same solution."""
    g = load_small(name)
    m = SMALL_MESHES[name]()
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights())
    A = va.LduMatrix(lm, g["sym_diag"], g["sym_upper"])
    b = g["sym_b"]
    from helpers import solve_cases
    seen = 0
    for i, solver, pre, repro, sym in solve_cases(g):
        if pre != "GAMG" or repro or not sym:
            continue
        x = np.zeros(b.size)
        p = va.solve_ldu(A, b, x, va.SolverControls(solver, "GAMG", 1e-10, 0.0, 400,
                                                    gamgSmoother=va.GAMG_GAUSS_SEIDEL))
        ref = g[f"solve{i}_perf"]
        assert abs(p.iterations - int(ref[4])) <= 1, (solver, p.iterations, ref[4])
        xr = g[f"solve{i}_x"]
        assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr), solver
        seen += 1
    assert seen == 2
    xm = np.zeros(b.size)
    pm = va.solve_ldu(A, b, xm, va.SolverControls("PCG", "GAMG", 1e-10, 0.0, 400, precision=va.MIXED,
                                                  gamgSmoother=va.GAMG_GAUSS_SEIDEL))
    assert pm.converged and np.linalg.norm(xm - xr) <= 1e-8 * np.linalg.norm(xr)
    with pytest.raises(va.VolumaError):
        va.solve_ldu(A, b, xm, va.SolverControls("PCG", "GAMG", gamgSmoother=7))
    A.close()
    lm.close()


def test_sweeps_with_split_runs_bit_exact_vs_oracle(ctx):
    """A 48k-cell tet mesh (20³ hexes × 6): its forward levels all hold ≤ 1024 cells, so
    the one-workgroup runs together cover far more cells than one run's 16384 LDS slots
    and are cut at level boundaries. Gauss–Seidel (2 sweeps), DILU-PBiCGStab and
    PCG + GAMG smoothed by Gauss–Seidel, with exact sums: bit-identical to the oracle."""
    from voluma_amd import mesh as M
    m = M.tet(20, 20, 20, 0.2, 1)
    ni, n = m.n_internal, m.n_cells
    w = m.agglomeration_weights()
    rng = np.random.default_rng(4)
    u = -w
    d = np.zeros(n)
    np.add.at(d, m.owner[:ni], w)
    np.add.at(d, m.neighbour, w)
    d = d + 0.05 * w.mean()
    lo = u * rng.uniform(0.8, 1.2, ni)
    b = rng.uniform(-1, 1, n)
    lm = va.LduMesh(ctx, n, m.owner[:ni], m.neighbour, w)
    om = O.LduMesh(n, m.owner[:ni], m.neighbour, w)
    cases = [("GaussSeidel", "none", False, dict(n_sweeps=2), dict(nSweeps=2)),
             ("PBiCGStab", "DILU", False, {}, {}),
             ("PCG", "GAMG", True, dict(smoother=O.SMOOTHER_GS), dict(gamgSmoother=va.GAMG_GAUSS_SEIDEL))]
    for solver, pre, sym, okw, vkw in cases:
        lower = None if sym else lo
        tol, mx = (1e-6, 40) if solver == "GaussSeidel" else (1e-10, 300)
        xo, po = om.solve(d, u, lower, b, solver=solver, preconditioner=pre, tolerance=tol, max_iter=mx,
                          mode=O.MODE_EXACT, **okw)
        A = va.LduMatrix(lm, d, u, lower)
        x = np.zeros(n)
        p = va.solve_ldu(A, b, x, va.SolverControls(solver, pre, tol, 0.0, mx, sums=va.SUMS_EXACT, **vkw))
        assert p.iterations == po["iterations"] and p.finalResidual == po["finalResidual"], solver
        assert np.array_equal(x, xo), solver
        A.close()
    lm.close()
