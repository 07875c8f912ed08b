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


def _ctl(pre="GAMG", tol=1e-10, **kw):
    return va.SolverControls("PCG", pre, tol, 0.0, 1000, **kw)


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
