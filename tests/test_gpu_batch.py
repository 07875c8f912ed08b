"""GPU tests of batched solves (vol_batch_*): many independent pressure systems of
uneven size in one set of launches. No reference counterpart; the bars are
(1) a system's result is bit-identical alone (a batch of one) and inside the batch,
(2) it agrees with the single-system solver vol_solve within the solve tolerance,
(3) mixed precision in a batch matches FP64 within 1e-8 at residual 1e-10.
"""
import numpy as np
import pytest

import voluma_amd as va
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def small_batch(ctx):
    specs = W.batch_specs(n_systems=7, seed=11, n_min=400, n_max=12000, median=2500)
    B, b, meshes = W.batch_system(ctx, specs)
    return specs, B, b, meshes


def _ctl(pre="GAMG", tol=1e-10, maxIter=1000, **kw):
    return va.SolverControls("PCG", pre, tol, 0.0, maxIter, **kw)


def _slice(B, v, s):
    return v[B.cell_off[s]: B.cell_off[s + 1]]


@pytest.mark.parametrize("pre", ["GAMG", "diagonal"])
def test_batch_matches_single_system_solver(ctx, small_batch, pre):
    specs, B, b, meshes = small_batch
    x = np.zeros(B.n_total)
    perf = B.solve(b, x, _ctl(pre))
    assert all(p.converged for p in perf)
    for s, m in enumerate(meshes):
        ni = m.n_internal
        lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights())
        lm.set_boundary(m.owner[ni:])
        A = va.LduMatrix(lm)
        g = m.geometry()
        bc = np.zeros(m.n_boundary, np.int32)
        s0 = int(m.patch_start[specs[s]["patch"]]) - ni
        bc[s0: s0 + int(m.patch_size[specs[s]["patch"]])] = 1
        bs = A.assemble_laplacian(g["magSf"], g["ndc"], bc, b_in=W.pressure_rhs(m.n_cells, 7 + s))
        assert np.array_equal(bs, _slice(B, b, s))
        xs = np.zeros(m.n_cells)
        ps = va.solve_ldu(A, bs, xs, _ctl(pre, gamgPreSweeps=2, gamgPostSweeps=2))
        xb = _slice(B, x, s)
        assert abs(ps.iterations - perf[s].iterations) <= 1, (s, ps.iterations, perf[s].iterations)
        assert np.linalg.norm(xb - xs) <= 1e-8 * np.linalg.norm(xs), s


def test_batch_packed_rows_bit_identical(ctx, small_batch):
    """Batch levels read through their packed SELL rows (the default) or slot-major
    (VOLUMA_SELL_PAIRS=0, read when the context is created): the same bits, FP64 and
    mixed precision."""
    import os
    specs, B, b, meshes = small_batch
    old = os.environ.get("VOLUMA_SELL_PAIRS")
    os.environ["VOLUMA_SELL_PAIRS"] = "0"
    try:
        c0 = va.Context(0)
    finally:
        if old is None:
            os.environ.pop("VOLUMA_SELL_PAIRS", None)
        else:
            os.environ["VOLUMA_SELL_PAIRS"] = old
    B0, b0, _ = W.batch_system(c0, specs)
    assert np.array_equal(b0, b)
    for prec in (va.FP64, va.MIXED):
        x, x0 = np.zeros(B.n_total), np.zeros(B.n_total)
        perf = B.solve(b, x, _ctl(precision=prec))
        perf0 = B0.solve(b0, x0, _ctl(precision=prec))
        for p, p0 in zip(perf, perf0):
            assert p.iterations == p0.iterations and p.finalResidual == p0.finalResidual, prec
        assert np.array_equal(x, x0), prec
    B0.close()
    c0.close()


def test_batch_result_independent_of_batch_mates(ctx, small_batch):
    """System k solved alone (batch of one) is bit-identical to system k in the batch."""
    specs, B, b, meshes = small_batch
    x = np.zeros(B.n_total)
    perf = B.solve(b, x, _ctl())
    for k in (0, 3, len(meshes) - 1):
        sub = [k]
        B1, b1, _ = W.batch_system(ctx, [specs[i] for i in sub], rhs_seed=7 + k)
        assert np.array_equal(b1, _slice(B, b, k))
        x1 = np.zeros(B1.n_total)
        p1 = B1.solve(b1, x1, _ctl())
        assert p1[0].iterations == perf[k].iterations and p1[0].finalResidual == perf[k].finalResidual
        assert np.array_equal(x1, _slice(B, x, k))
        B1.close()


def test_batch_mixed_precision(ctx, small_batch):
    specs, B, b, meshes = small_batch
    xr = np.zeros(B.n_total)
    B.solve(b, xr, _ctl(tol=1e-12))
    for tol, err in ((1e-6, 1e-3), (1e-10, 1e-8)):
        x = np.zeros(B.n_total)
        perf = B.solve(b, x, _ctl(tol=tol, precision=va.MIXED))
        for s in range(B.n_systems):
            assert perf[s].converged and perf[s].outerIterations >= 1
            xs, xrs = _slice(B, x, s), _slice(B, xr, s)
            assert np.linalg.norm(xs - xrs) <= err * np.linalg.norm(xrs), (tol, s)


def test_batch_errors(ctx, small_batch):
    specs, B, b, meshes = small_batch
    x = np.zeros(B.n_total)
    for ctl, code in ((va.SolverControls("PBiCGStab", "GAMG"), -6), (va.SolverControls("PCG", "DIC"), -6),
                      (va.SolverControls("PCG", "ILU"), -3),
                      (va.SolverControls("PCG", "GAMG", sums=va.SUMS_REPRODUCIBLE), -6)):
        with pytest.raises(va.VolumaError) as e:
            B.solve(b, x, ctl)
        assert e.value.code == code
    m = W.batch_mesh({"kind": "box", "nx": 4, "ny": 4, "nz": 4, "seed": 1, "patch": 0})
    with pytest.raises(va.VolumaError) as e:
        va.LduBatch(ctx, [(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())])
    assert e.value.code == -1


# ---------------------------------------------------------------- the benchmark's batch
BENCH_VCYCLE = dict(gamgPreSweeps=1, gamgPostSweeps=1, gamgOmega=0.9)


@pytest.fixture(scope="module")
def bench_batch(ctx):
    """The benchmark's batch: 256 systems, lognormal 5k–200k cells (median 30k)."""
    specs = W.batch_specs(n_systems=256)
    print(f"[batch256] building {len(specs)} systems", flush=True)
    B, b, meshes = W.batch_system(ctx, specs)
    print(f"[batch256] {B.n_total} cells", flush=True)
    yield specs, B, b, meshes
    B.close()


def _true_residuals(B, b, x, meshes, perf):
    """Normalised true residual Σ|b − A·x| / normFactor of every system, A·x by the
    oracle's amul (bit-identical to the GPU's)."""
    from oracle import oracle as O
    out = []
    for s, m in enumerate(meshes):
        c0, c1 = B.cell_off[s], B.cell_off[s + 1]
        f0, f1 = B.face_off[s], B.face_off[s + 1]
        om = O.LduMesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, None)
        r = b[c0:c1] - om.amul(B.diag_h[c0:c1], B.upper_h[f0:f1], None, x[c0:c1])
        out.append(float(np.abs(r).sum() / perf[s].normFactor))
    return np.array(out)


def test_bench_batch_fp64_true_residual_and_batch_mates(ctx, bench_batch):
    """Every system of the 256 converges and its true residual (recomputed with the
    bit-exact A·x) is at the 1e-6 target; the smallest, median and largest systems
    solved alone (a batch of one) are bit-identical to their solution in the batch."""
    specs, B, b, meshes = bench_batch
    assert B.n_systems == 256 and 9_000_000 < B.n_total < 12_000_000
    x = np.zeros(B.n_total)
    perf = B.solve(b, x, _ctl(tol=1e-6, **BENCH_VCYCLE))
    assert all(p.converged and p.finalResidual < 1e-6 for p in perf)
    tr = _true_residuals(B, b, x, meshes, perf)
    assert np.all(tr <= 1e-6 * (1 + 1e-6)), (int(np.argmax(tr)), float(tr.max()))
    order = np.argsort(B.n_cells, kind="stable")
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        print(f"[batch256] system {k} ({B.n_cells[k]} cells) alone", flush=True)
        B1, b1, _ = W.batch_system(ctx, [specs[k]], rhs_seed=7 + k)
        assert np.array_equal(b1, _slice(B, b, k))
        x1 = np.zeros(B1.n_total)
        p1 = B1.solve(b1, x1, _ctl(tol=1e-6, **BENCH_VCYCLE))
        assert p1[0].iterations == perf[k].iterations and p1[0].finalResidual == perf[k].finalResidual
        assert np.array_equal(x1, _slice(B, x, k))
        B1.close()


def test_bench_batch_mixed_precision(ctx, bench_batch):
    """Mixed precision on the 256 systems: every solution at residual 1e-10 within 1e-8
    (relative L2) of the FP64 solution at 1e-12; at 1e-6 every system converges."""
    specs, B, b, meshes = bench_batch
    xr = np.zeros(B.n_total)
    pr = B.solve(b, xr, _ctl(tol=1e-12, maxIter=3000, **BENCH_VCYCLE))
    assert all(p.converged for p in pr)
    x = np.zeros(B.n_total)
    perf = B.solve(b, x, _ctl(tol=1e-10, maxIter=3000, precision=va.MIXED, **BENCH_VCYCLE))
    errs = []
    for s in range(B.n_systems):
        assert perf[s].converged and perf[s].outerIterations >= 1
        xs, xrs = _slice(B, x, s), _slice(B, xr, s)
        errs.append(np.linalg.norm(xs - xrs) / np.linalg.norm(xrs))
    assert max(errs) <= 1e-8, max(errs)
    x6 = np.zeros(B.n_total)
    p6 = B.solve(b, x6, _ctl(tol=1e-6, precision=va.MIXED, **BENCH_VCYCLE))
    assert all(p.converged for p in p6)
    print(f"[batch256] mixed 1e-10: max rel. error {max(errs):.2e}", flush=True)


@pytest.mark.parametrize("prec", [va.FP64, va.MIXED])
def test_batch_graph_replay_matches_eager(ctx, small_batch, prec):
    """Batched iterations replayed from a captured HIP graph give the bits of eagerly
    launched ones, over several solves in a row with eager work between the replays."""
    import os
    specs, B, b, meshes = small_batch
    ctl = _ctl(tol=1e-9, precision=prec, **BENCH_VCYCLE)
    old = os.environ.get("VOLUMA_BATCH_GRAPH")
    try:
        os.environ["VOLUMA_BATCH_GRAPH"] = "0"
        B.graph_stats(reset=True)
        xe = np.zeros(B.n_total)
        pe = B.solve(b, xe, ctl)
        g, e = B.graph_stats(reset=True)
        assert g == 0 and e > 0
        os.environ["VOLUMA_BATCH_GRAPH"] = "1"
        for rep in range(3):
            xg = np.zeros(B.n_total)
            pg = B.solve(b, xg, ctl)
            g, e = B.graph_stats(reset=True)
            assert g > 0, "no graph replays"
            assert np.array_equal(xg, xe), rep
            assert [p.iterations for p in pg] == [p.iterations for p in pe]
            assert [p.finalResidual for p in pg] == [p.finalResidual for p in pe]
            # eager work between the replays
            os.environ["VOLUMA_BATCH_GRAPH"] = "0"
            x2 = np.zeros(B.n_total)
            B.solve(b, x2, _ctl(tol=1e-6, precision=prec, **BENCH_VCYCLE))
            os.environ["VOLUMA_BATCH_GRAPH"] = "1"
    finally:
        if old is None:
            os.environ.pop("VOLUMA_BATCH_GRAPH", None)
        else:
            os.environ["VOLUMA_BATCH_GRAPH"] = old
