"""GPU tests of the solve server (vol_server_*): continuous admission of requests on
concurrent streams, queued pressure requests batched into shared launches, meshes and
hierarchies resident, admission by an HBM budget. No reference counterpart; the bar is
that each request's result is the same alone and in the mix (bit for bit: a system's
arithmetic in a batch does not depend on its batch mates), and equal to a direct solve
of the same system (a batch of one for pressure, vol_solve for momentum).
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


def _est(e):  # the server's admission estimate per resident mesh copy + matrix (+ batch of 4 copies)
    m = e["mesh"]
    return (m.n_cells * 400 + m.n_internal * 96 + (1 << 20)) * (5 if m.n_cells <= 1 << 18 else 1)


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


def test_server_results_equal_direct_solves(ctx, pool):
    """Pressure requests run as systems of a batch of copies of their mesh: the same
    bits as a batch of one system (vol_batch) of that matrix and right-hand side.
    Momentum requests run with vol_solve: the same bits as a plain vol_solve."""
    server = va.Server(0, slots=2)
    SV.register_pool(server, pool)
    e = pool[1]
    m = e["mesh"]
    rq = {"t": 0.0, "kind": "pressure", "pool": 1, "nrhs": 2, "seed": 42}
    args = SV.request_args(pool, rq, va.FP64)
    xs, perf = server.wait(server.submit(**args))
    B1 = va.LduBatch(ctx, [(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())])
    B1.set(e["p_diag"], e["p_upper"])
    for k, b in enumerate(args["bs"]):
        x = np.zeros(m.n_cells)
        p = B1.solve(b, x, args["controls"])[0]
        assert np.array_equal(x, xs[k]) and p.iterations == perf[k].iterations
    B1.close()
    rq = {"t": 0.0, "kind": "momentum", "pool": 1, "nrhs": 2, "seed": 43}
    args = SV.request_args(pool, rq, va.FP64)
    xs, perf = server.wait(server.submit(**args))
    lm = va.LduMesh(ctx, m.n_cells, m.owner[: m.n_internal], m.neighbour, None)
    A = va.LduMatrix(lm, e["m_diag"], e["m_upper"], e["m_lower"])
    for k, b in enumerate(args["bs"]):
        x = np.array(args["x0s"][k], dtype=np.float64)
        p = va.solve_ldu(A, b, x, args["controls"])
        assert np.array_equal(x, xs[k]) and p.iterations == perf[k].iterations
    server.close()


@pytest.mark.parametrize("precision", [va.FP64, va.MIXED])
def test_queued_pressure_requests_share_a_batch(pool, precision):
    """With one worker busy, pressure requests that queue behind each other on one mesh
    are solved together (the same start and end), and each result is the bits it has
    alone."""
    server = va.Server(0, slots=1)
    SV.register_pool(server, pool)
    busy = {"t": 0.0, "kind": "momentum", "pool": 2, "nrhs": 3, "seed": 50}
    reqs = [{"t": 0.0, "kind": "pressure", "pool": 0, "nrhs": 2, "seed": 60 + k} for k in range(2)]
    alone = _alone(server, pool, reqs, precision)
    t_busy = server.submit(**SV.request_args(pool, busy, precision))
    tickets = [server.submit(**SV.request_args(pool, rq, precision)) for rq in reqs]
    server.wait(t_busy)
    got, times = [], []
    for t in tickets:
        got.append(server.wait(t))
        times.append(dict(server.last_times))
    assert abs(times[0]["started"] - times[1]["started"]) < 1e-3
    assert abs(times[0]["finished"] - times[1]["finished"]) < 1e-3
    for (xa, pa), (xg, pg) in zip(alone, got):
        assert [p.iterations for p in pa] == [p.iterations for p in pg]
        for a, b in zip(xa, xg):
            assert np.array_equal(a, b)
    server.close()


def test_prefetch_makes_every_worker_resident(pool):
    """prefetch builds each worker's copy, matrix storage and batch of copies of a mesh;
    requests after it give the same bits as on a server that built them on first use,
    and the entries are then counted at their measured device bytes, not the estimate."""
    stream = SV.request_stream(pool, rate=1e6, n_requests=10, seed=6)
    ref = va.Server(0, slots=4)
    SV.register_pool(ref, pool)
    expect = _alone(ref, pool, stream, va.FP64)
    ref.close()
    server = va.Server(0, slots=4, hbm_budget=200e9)
    SV.register_pool(server, pool)
    reg = server.stats()["bytes_in_use"]
    for e in pool:
        server.prefetch(e["id"])
    st = server.stats()
    assert st["cached"] == 4 * len(pool) and st["queued"] == 0
    resident = st["bytes_in_use"] - reg
    est = 4 * sum(_est(e) for e in pool)
    print(f"[serve] resident {resident / 1e6:.1f} MB measured, {est / 1e6:.1f} MB estimated", flush=True)
    assert 0 < resident <= est
    got, _ = SV.run_stream(server, pool, stream, va.FP64, timeout=120.0)
    for (xa, pa), (xg, pg) in zip(expect, got):
        assert [p.iterations for p in pa] == [p.iterations for p in pg]
        for a, b in zip(xa, xg):
            assert np.array_equal(a, b)
    with pytest.raises(va.VolumaError):
        server.prefetch(len(pool))
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
