"""GPU tests of the damped-Jacobi V-cycle's equivalent launch forms. Each form is a
context switch read from the environment when the context is created:

* coarse levels take z = ω·(r·rD) from the restriction (VOLUMA_COARSE_Z);
* FP64 PCG keeps the level-0 z with its residual (VOLUMA_LEVEL0_Z);
* the level-0 post-sweep reads z + xc[agg] from a streaming prolongation pass
  (VOLUMA_PROLONG_SPLIT=1);
* the levels from the first of at most VOLUMA_TAIL_MAX_CELLS cells (up to 4096) down
  to the coarsest run in one workgroup, matrices preloaded into LDS where they fit;
* the last two levels above the coarsest and the coarsest run in one workgroup
  (VOLUMA_LAST2=1, opt-in, where they are small CSR levels);
* the last post-sweep of a coarse level adds its result to the finer level's vector
  (the pushed prolongation, VOLUMA_PUSH_PROLONG, default) instead of the finer sweep
  gathering x + xc[agg];
* SELL coarse levels hold their rows sorted by length within windows of rows
  (VOLUMA_SELL_SORT=<rows>, opt-in), read back in row order by the Gauss–Seidel level
  tables;
* SELL coarse levels are read through their packed rows, int16 column offsets and
  coefficient pairs (VOLUMA_SELL_PAIRS, default);
* the fast-mode A·p finishes its dot product in its last block and runs the scalar
  step there instead of in a finalize launch (VOLUMA_LAST_BLOCK_FIN=1, opt-in);
* levels compute their residual and its restriction (into the coarsest level: and its
  substitution, and the post-sweep of the level) in one launch (VOLUMA_FUSED_RESTRICT,
  default), the packed level 0 reading its rows in restriction order
  (VOLUMA_FUSED_LEVEL0), SELL levels from a copy of their rows in restriction order
  (VOLUMA_FUSED_SELL_MAX: up to that many rows).

Every form computes the same per-row arithmetic in the same order, so solutions,
iteration counts and residuals must equal the plain per-level launches bit for bit, in
fast and exact sums, FP64 and mixed precision, with 1/1 and 2/2 sweeps."""
import os

import numpy as np
import pytest

import voluma_amd as va
from voluma_amd import mesh as M
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu

PLAIN = {"VOLUMA_COARSE_Z": "0", "VOLUMA_LEVEL0_Z": "0", "VOLUMA_PROLONG_SPLIT": "0", "VOLUMA_TAIL_MAX_CELLS": "0",
         "VOLUMA_FUSED_RESTRICT": "0", "VOLUMA_PUSH_PROLONG": "0", "VOLUMA_LAST2": "0", "VOLUMA_SELL_SORT": "0",
         "VOLUMA_SELL_PAIRS": "0"}
FORMS = {
    "default": {},
    "last_two_levels_in_one_workgroup": {"VOLUMA_LAST2": "1"},
    "last_two_levels_in_one_workgroup_gathered": {"VOLUMA_LAST2": "1", "VOLUMA_PUSH_PROLONG": "0"},
    "gathered_prolongation": {"VOLUMA_PUSH_PROLONG": "0"},
    "level0_residual_and_restriction_separate": {"VOLUMA_FUSED_LEVEL0": "0"},
    "sell_levels_fused_direct": {"VOLUMA_FUSED_SELL_DIRECT": "1"},
    "pushed_prolongation_unfused": {"VOLUMA_FUSED_RESTRICT": "0"},
    "fused_sell_levels": {"VOLUMA_FUSED_SELL_MAX": "1048576"},
    "fused_restrict_no_coarse_z": {"VOLUMA_COARSE_Z": "0"},
    "last_block_finalize": {"VOLUMA_LAST_BLOCK_FIN": "1"},
    "prolong_split": {"VOLUMA_PROLONG_SPLIT": "1"},
    "tail": {"VOLUMA_TAIL_MAX_CELLS": "512"},
    "tail_no_coarse_z": {"VOLUMA_TAIL_MAX_CELLS": "128", "VOLUMA_COARSE_Z": "0"},
    "tail_4096": {"VOLUMA_TAIL_MAX_CELLS": "4096"},
}
KEYS = sorted(PLAIN) + ["VOLUMA_FUSED_SELL_MAX", "VOLUMA_FUSED_LEVEL0", "VOLUMA_FUSED_SELL_DIRECT",
                        "VOLUMA_LAST_BLOCK_FIN"]


def _context(env):
    old = {k: os.environ.get(k) for k in KEYS}
    try:
        for k in KEYS:
            os.environ.pop(k, None)
        os.environ.update(env)
        return va.Context(0)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _controls(pre, prec, sums, tol):
    jac = dict(gamgPreSweeps=pre, gamgPostSweeps=pre, gamgOmega=0.9 if pre == 1 else 0.8)
    return va.SolverControls("PCG", "GAMG", tol, 0.0, 2000, sums=sums, precision=prec, **jac)


CASES = [(1, va.FP64, va.SUMS_FAST, 1e-8), (2, va.FP64, va.SUMS_FAST, 1e-8), (1, va.FP64, va.SUMS_EXACT, 1e-7),
         (1, va.MIXED, va.SUMS_FAST, 1e-8)]


def _solves(env, m):
    ctx = _context(env)
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True, nonorth=False)
    out = []
    for pre, prec, sums, tol in CASES:
        x = np.zeros(m.n_cells)
        p = va.solve_ldu(A, b, x, _controls(pre, prec, sums, tol))
        out.append((x, p))
    A.close()
    lm.close()
    ctx.close()
    return out


@pytest.mark.parametrize("mesh", ["box64", "tet"])
def test_vcycle_forms_bit_identical(mesh):
    m = W.make_mesh("box64") if mesh == "box64" else M.tet(34, 30, 28, 0.2, 3)
    print(f"[forms] {mesh}: {m.n_cells} cells, plain per-level launches", flush=True)
    ref = _solves(PLAIN, m)
    for (pre, prec, sums, tol), (x, p) in zip(CASES, ref):
        assert p.converged and p.finalResidual < tol, (pre, prec, sums)
    for name, env in FORMS.items():
        print(f"[forms] {mesh}: {name}", flush=True)
        got = _solves(env, m)
        for case, (xr, pr), (xg, pg) in zip(CASES, ref, got):
            assert pg.iterations == pr.iterations, (name, case)
            assert pg.finalResidual == pr.finalResidual and pg.initialResidual == pr.initialResidual, (name, case)
            assert np.array_equal(xg, xr), (name, case)


def _packed_solves(packed, m):
    """Level-0 gathers with packed rows (the default where the mesh allows them) or
    slot-major (VOLUMA_PACKED=0, read when the mesh is set up): pressure solves of every
    case above, diagonal PCG, A·x, and the nonsymmetric momentum system."""
    old = os.environ.get("VOLUMA_PACKED")
    try:
        if packed:
            os.environ.pop("VOLUMA_PACKED", None)
        else:
            os.environ["VOLUMA_PACKED"] = "0"
        ctx = va.Context(0)
        out = []
        lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True, nonorth=False)
        for pre, prec, sums, tol in CASES:
            x = np.zeros(m.n_cells)
            out.append((x, va.solve_ldu(A, b, x, _controls(pre, prec, sums, tol))))
        x = np.zeros(m.n_cells)
        out.append((x, va.solve_ldu(A, b, x, va.SolverControls("PCG", "diagonal", 1e-8, 0.0, 400))))
        out.append((va.amul(A, b), None))
        A.close()
        lm.close()
        lm, A, b, inp = W.momentum_system(ctx, m, seed=1, comp=0)
        for prec in (va.FP64, va.MIXED):
            x = np.ascontiguousarray(inp["U0"][:, 0]).copy()
            out.append((x, va.solve_ldu(A, b, x, va.SolverControls("PBiCGStab", "diagonal", 1e-10, 0.0, 50,
                                                                      precision=prec))))
        out.append((va.amul(A, b), None))
        A.close()
        lm.close()
        ctx.close()
        return out
    finally:
        if old is None:
            os.environ.pop("VOLUMA_PACKED", None)
        else:
            os.environ["VOLUMA_PACKED"] = old


@pytest.mark.parametrize("mesh", ["box64", "tet"])
def test_packed_rows_bit_identical(mesh):
    m = W.make_mesh("box64") if mesh == "box64" else M.tet(34, 30, 28, 0.2, 3)
    print(f"[packed] {mesh}: {m.n_cells} cells, slot-major level 0", flush=True)
    ref = _packed_solves(False, m)
    print(f"[packed] {mesh}: packed rows", flush=True)
    got = _packed_solves(True, m)
    for k, ((xr, pr), (xg, pg)) in enumerate(zip(ref, got)):
        if pr is not None:
            assert pg.iterations == pr.iterations and pg.finalResidual == pr.finalResidual, k
        assert np.array_equal(xg, xr), k


def _sell_solves(env, m):
    """Every coarse level in SELL form (VOLUMA_COARSE_LAYOUT=sell, read when the mesh is
    set up) with the switches of env: the Jacobi V-cycle cases above and a Gauss–Seidel
    GAMG-PCG solve (its level tables are built from the SELL rows read back)."""
    keys = ["VOLUMA_COARSE_LAYOUT", "VOLUMA_SELL_SORT", "VOLUMA_FUSED_SELL_DIRECT", "VOLUMA_SELL_PAIRS"]
    old = {k: os.environ.get(k) for k in keys}
    try:
        for k in keys:
            os.environ.pop(k, None)
        os.environ["VOLUMA_COARSE_LAYOUT"] = "sell"
        os.environ.update(env)
        ctx = va.Context(0)
        lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True, nonorth=False)
        out = []
        for pre, prec, sums, tol in CASES:
            x = np.zeros(m.n_cells)
            out.append((x, va.solve_ldu(A, b, x, _controls(pre, prec, sums, tol))))
        x = np.zeros(m.n_cells)
        out.append((x, va.solve_ldu(A, b, x, va.SolverControls("PCG", "GAMG", 1e-7, 0.0, 40, sums=va.SUMS_EXACT,
                                                                  gamgSmoother=va.GAMG_GAUSS_SEIDEL))))
        A.close()
        lm.close()
        ctx.close()
        return out
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("mesh", ["box64", "tet"])
def test_sell_row_forms_bit_identical(mesh):
    m = W.make_mesh("box64") if mesh == "box64" else M.tet(34, 30, 28, 0.2, 3)
    print(f"[sell-sort] {mesh}: {m.n_cells} cells, SELL levels in row order, slot-major", flush=True)
    ref = _sell_solves({"VOLUMA_SELL_SORT": "0", "VOLUMA_SELL_PAIRS": "0"}, m)
    for name, env in (("packed_pairs", {}), ("packed_pairs_fused_direct", {"VOLUMA_FUSED_SELL_DIRECT": "1"}),
                      ("sorted", {"VOLUMA_SELL_SORT": "1024"}),
                      ("sorted_fused_direct", {"VOLUMA_SELL_SORT": "1024", "VOLUMA_FUSED_SELL_DIRECT": "1"})):
        print(f"[sell-sort] {mesh}: {name}", flush=True)
        got = _sell_solves(env, m)
        for k, ((xr, pr), (xg, pg)) in enumerate(zip(ref, got)):
            assert pg.iterations == pr.iterations and pg.finalResidual == pr.finalResidual, (name, k)
            assert np.array_equal(xg, xr), (name, k)


def _krylov_solves(env, m):
    """Diagonal PCG and PBiCGStab (FP64 and mixed) on the pressure and momentum systems
    in a context made with env."""
    ctx = _context(env)
    out = []
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True, nonorth=False)
    for prec in (va.FP64, va.MIXED):
        x = np.zeros(m.n_cells)
        out.append((x, va.solve_ldu(A, b, x, va.SolverControls("PCG", "diagonal", 1e-8, 0.0, 400, precision=prec))))
    A.close()
    lm.close()
    lm, A, b, inp = W.momentum_system(ctx, m, seed=1, comp=0)
    for prec in (va.FP64, va.MIXED):
        x = np.ascontiguousarray(inp["U0"][:, 0]).copy()
        out.append((x, va.solve_ldu(A, b, x, va.SolverControls("PBiCGStab", "diagonal", 1e-10, 0.0, 50,
                                                                  precision=prec))))
    A.close()
    lm.close()
    ctx.close()
    return out


def test_last_block_finalize_krylov_bit_identical():
    m = M.tet(34, 30, 28, 0.2, 3)
    ref = _krylov_solves({}, m)
    got = _krylov_solves({"VOLUMA_LAST_BLOCK_FIN": "1"}, m)
    for k, ((xr, pr), (xg, pg)) in enumerate(zip(ref, got)):
        assert pg.iterations == pr.iterations and pg.finalResidual == pr.finalResidual, k
        assert np.array_equal(xg, xr), k
