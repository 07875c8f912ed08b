"""GPU tests of the solve server (vol_server_*): continuous admission of requests on
concurrent streams, meshes and hierarchies resident, admission by an HBM budget. No
reference counterpart; the bar is that each request's result is the same alone and in
the mix (bit for bit: every request runs on its own worker buffers), and equal to a
plain vol_solve of the same system.
"""
import numpy as np
import pytest

import voluma_amd as va
from voluma_amd import serving as SV

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def pool(ctx):
    return SV.build_pool(ctx, SV.pool_specs(n_meshes=3, seed=5, n_min=2000, n_max=30000))


def _est(e):  # the server's admission estimate per resident mesh copy + matrix
    m = e["mesh"]
    return m.n_cells * 400 + m.n_internal * 96 + (1 << 20)


def _alone(server, pool, stream, precision):
    out = []
    for rq in stream:
        t = server.submit(**SV.request_args(pool, rq, precision))
        out.append(server.wait(t, 120.0))
    return out


@pytest.mark.parametrize("precision", [va.FP64, va.MIXED])
def test_results_same_alone_and_in_the_mix(pool, precision):
    server = va.Server(0, slots=4, hbm_budget=200e9)
    SV.register_pool(server, pool)
    stream = SV.request_stream(pool, rate=1e6, n_requests=12, seed=3)
    alone = _alone(server, pool, stream, precision)
    mixed, stats = SV.run_stream(server, pool, stream, precision, timeout=120.0)
    assert stats["requests"] == 12 and stats["latency_ms"]["p99"] >= stats["latency_ms"]["p50"] > 0
    for (xa, pa), (xm, pm) in zip(alone, mixed):
        assert all(p.converged for p in pa) and all(p.converged for p in pm)
        for a, b in zip(xa, xm):
            assert np.array_equal(a, b)
        assert [p.iterations for p in pa] == [p.iterations for p in pm]
    server.close()


def test_server_result_equals_plain_solve(ctx, pool):
    server = va.Server(0, slots=2)
    SV.register_pool(server, pool)
    rq = {"t": 0.0, "kind": "pressure", "pool": 1, "nrhs": 2, "seed": 42}
    args = SV.request_args(pool, rq, va.FP64)
    xs, perf = server.wait(server.submit(**args))
    e = pool[1]
    m = e["mesh"]
    lm = va.LduMesh(ctx, m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    A = va.LduMatrix(lm, e["p_diag"], e["p_upper"])
    for k, b in enumerate(args["bs"]):
        x = np.zeros(m.n_cells)
        p = va.solve_ldu(A, b, x, args["controls"])
        assert np.array_equal(x, xs[k]) and p.iterations == perf[k].iterations
    server.close()


def test_hbm_budget_admission(pool):
    """A budget that holds the registered meshes plus one resident copy: requests run
    one at a time (evicting), all complete with the same results; a request that can
    never fit is refused at submission."""
    reg = sum(_est(e) // 2 for e in pool)
    big = max(_est(e) for e in pool)
    stream = SV.request_stream(pool, rate=1e6, n_requests=8, seed=4)
    ref = va.Server(0, slots=4)
    SV.register_pool(ref, pool)
    expect = _alone(ref, pool, stream, va.FP64)
    ref.close()
    server = va.Server(0, slots=4, hbm_budget=reg + big)
    SV.register_pool(server, pool)
    got, stats = SV.run_stream(server, pool, stream, va.FP64, timeout=120.0)
    for (xa, _), (xg, _) in zip(expect, got):
        for a, b in zip(xa, xg):
            assert np.array_equal(a, b)
    assert server.stats()["bytes_in_use"] <= reg + big
    server.close()
    tiny = va.Server(0, slots=1, hbm_budget=reg + big - 1)
    SV.register_pool(tiny, pool)
    ibig = max(range(len(pool)), key=lambda i: _est(pool[i]))
    rq = {"t": 0.0, "kind": "pressure", "pool": ibig, "nrhs": 2, "seed": 1}
    with pytest.raises(va.VolumaError) as e:
        tiny.submit(**SV.request_args(pool, rq, va.FP64))
    assert e.value.code == -1
    tiny.close()


# ---------------------------------------------------------------- the benchmark's pool
@pytest.fixture(scope="module")
def bench_pool(ctx):
    """The benchmark's pool: 8 meshes, log-uniform 10k–2M cells (30k … 1.86M)."""
    print("[serve] building the benchmark pool", flush=True)
    pool = SV.build_pool(ctx, SV.pool_specs())
    print("[serve] pool cells " + ", ".join(str(e["n"]) for e in pool), flush=True)
    return pool


@pytest.mark.parametrize("precision", [va.FP64, va.MIXED])
def test_bench_pool_same_alone_and_in_the_mix(bench_pool, precision):
    """Requests of the benchmark's mix (60 % pressure GAMG-PCG / 40 % momentum
    PBiCGStab, 2–3 right-hand sides each, 200 GB budget, 4 workers) arriving faster than
    they are served: a sample that includes pressure and momentum requests on the
    1.86M-cell mesh gives the same bits alone as in the mix."""
    pool = bench_pool
    big = max(range(len(pool)), key=lambda i: pool[i]["n"])
    assert pool[big]["n"] > 1_800_000
    stream = SV.request_stream(pool, rate=200.0, n_requests=24, seed=9)
    t_end = stream[-1]["t"]
    stream += [{"t": t_end + 1e-3, "kind": "pressure", "pool": big, "nrhs": 3, "seed": 900},
               {"t": t_end + 2e-3, "kind": "momentum", "pool": big, "nrhs": 3, "seed": 901}]
    server = va.Server(0, slots=4, hbm_budget=200e9)
    SV.register_pool(server, pool)
    mixed, stats = SV.run_stream(server, pool, stream, precision, timeout=120.0)
    assert stats["requests"] == len(stream)
    sample = sorted(set([k for k, rq in enumerate(stream) if rq["pool"] == big] + [0, 5, 11, 17]))
    print(f"[serve] {len(sample)} requests again alone", flush=True)
    alone = _alone(server, pool, [stream[k] for k in sample], precision)
    for k, (xa, pa) in zip(sample, alone):
        xm, pm = mixed[k]
        assert all(p.converged and p.finalResidual < 1e-6 for p in pm), k
        assert [p.iterations for p in pa] == [p.iterations for p in pm], k
        for a, b in zip(xa, xm):
            assert np.array_equal(a, b), k
    server.close()
