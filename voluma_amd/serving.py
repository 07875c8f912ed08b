"""Serving workload: a Poisson-arrival stream of solve requests against a
:class:`voluma_amd.Server`.

* :func:`pool_specs` / :func:`build_pool` — a pool of meshes with log-uniform sizes
  (box-hex and jittered-tet, seeded), each with its pressure matrix (fvm::laplacian,
  fixedValue 0 on xMax) and momentum matrix (ddt + upwind div − ν laplacian, the
  velocity field of :func:`voluma_amd.workloads.momentum_inputs`), assembled on the GPU;
* :func:`request_stream` — arrivals at a given rate, 60 % pressure (GAMG-PCG) and 40 %
  momentum (PBiCGStab, diagonal), each with 2–3 right-hand sides (PISO correctors or
  velocity components) on one matrix;
* :func:`run_stream` — submits every request at its arrival time, collects the
  results and reports p50/p99 latency (arrival → result) and solves/s.
"""
from __future__ import annotations

import math
import threading
import time

import numpy as np

from . import workloads as W


def pool_specs(n_meshes=8, seed=5, n_min=10_000, n_max=2_000_000):
    """Seeded mesh sizes, log-uniform in [n_min, n_max] (box-hex and tet alternate)."""
    u = W.splitmix_uniform(2 * n_meshes, seed).reshape(n_meshes, 2)
    out = []
    for k in range(n_meshes):
        n = math.exp(math.log(n_min) + float(u[k, 0]) * (math.log(n_max) - math.log(n_min)))
        kind = "tet" if k % 2 == 0 else "box"
        side = W._cbrt(n / (6.0 if kind == "tet" else 1.0))
        nx = max(4, int(side + 0.5))
        out.append({"kind": kind, "nx": nx, "ny": nx, "nz": nx, "seed": 500 + k, "patch": W.XMAX_PATCH})
    return out


def build_pool(ctx, specs):
    """Meshes and their pressure and momentum systems (assembled on the GPU)."""
    import voluma_amd as va

    pool = []
    for sp in specs:
        m = W.batch_mesh(sp)
        ni = m.n_internal
        lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
        lm.set_boundary(m.owner[ni:])
        A = va.LduMatrix(lm)
        g = m.geometry()
        A.assemble_laplacian(g["magSf"], g["ndc"], W.bc_fixed_xmax(m))
        pd, pu, _ = A.get()
        inp = W.momentum_inputs(m, sp["seed"], exact=False)
        mb = [W.momentum_assemble(A, m, inp, c) for c in range(3)]
        md, mu, ml = A.get()
        pool.append({"mesh": m, "n": m.n_cells, "p_diag": pd, "p_upper": pu, "m_diag": md, "m_upper": mu,
                     "m_lower": ml, "m_b": mb, "U0": inp["U0"]})
        A.close()
        lm.close()
    return pool


def register_pool(server, pool, prefetch=False):
    """Register every pool mesh; with prefetch, also make each resident on every worker."""
    for e in pool:
        m = e["mesh"]
        e["id"] = server.add_mesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights(),
                                  m.owner[m.n_internal:])
    if prefetch:
        for e in pool:
            server.prefetch(e["id"])


def request_stream(pool, rate, n_requests, seed=9):
    """Arrival times (s, Poisson at `rate`/s) and request descriptions."""
    u = W.splitmix_uniform(4 * n_requests, seed).reshape(n_requests, 4)
    t = 0.0
    out = []
    for k in range(n_requests):
        t += -math.log(1.0 - float(u[k, 0])) / rate
        out.append({"t": t, "kind": "pressure" if u[k, 1] < 0.6 else "momentum",
                    "pool": int(u[k, 2] * len(pool)), "nrhs": 2 if u[k, 3] < 0.5 else 3, "seed": 100 + k})
    return out


def request_args(pool, rq, precision):
    """Server.submit arguments of one request."""
    import voluma_amd as va

    e = pool[rq["pool"]]
    if rq["kind"] == "pressure":
        bs = [W.pressure_rhs(e["n"], rq["seed"] + j) for j in range(rq["nrhs"])]
        ctl = va.SolverControls("PCG", "GAMG", 1e-6, 0.0, 1000, gamgPreSweeps=1, gamgPostSweeps=1, gamgOmega=0.9,
                                precision=precision)
        return dict(mesh_id=e["id"], diag=e["p_diag"], upper=e["p_upper"], lower=None, bs=bs, x0s=None, controls=ctl)
    bs = [e["m_b"][j] for j in range(rq["nrhs"])]
    x0 = [e["U0"][:, j] for j in range(rq["nrhs"])]
    ctl = va.SolverControls("PBiCGStab", "diagonal", 1e-6, 0.0, 1000, precision=precision)
    return dict(mesh_id=e["id"], diag=e["m_diag"], upper=e["m_upper"], lower=e["m_lower"], bs=bs, x0s=x0,
                controls=ctl)


def run_stream(server, pool, stream, precision, timeout=600.0):
    """Submit each request at its arrival time; returns per-request results and stats."""
    args = [request_args(pool, rq, precision) for rq in stream]
    tickets = [None] * len(stream)
    submitted = [0.0] * len(stream)
    finished = [0.0] * len(stream)
    results = [None] * len(stream)
    ready = threading.Semaphore(0)
    t0 = time.perf_counter()

    def submitter():
        for k, rq in enumerate(stream):
            delay = t0 + rq["t"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            submitted[k] = time.perf_counter()
            tickets[k] = server.submit(**args[k])
            ready.release()

    th = threading.Thread(target=submitter)
    th.start()
    service = [None] * len(stream)
    for k in range(len(stream)):
        ready.acquire()
        results[k] = server.wait(tickets[k], timeout)
        finished[k] = time.perf_counter()
        service[k] = server.last_times
    th.join()
    wall = max(finished) - t0
    # latency: queued → finished on the server's clock, plus any lag of the submit
    # behind the scheduled arrival
    lat = np.array([service[k]["finished"] - service[k]["queued"] + max(0.0, submitted[k] - (t0 + stream[k]["t"]))
                    for k in range(len(stream))])
    wait = np.array([service[k]["started"] - service[k]["queued"] for k in range(len(stream))])
    run = np.array([service[k]["finished"] - service[k]["started"] for k in range(len(stream))])
    n_solves = sum(rq["nrhs"] for rq in stream)

    def pct(v):
        return {"p50": float(np.percentile(v, 50) * 1e3), "p99": float(np.percentile(v, 99) * 1e3),
                "max": float(v.max() * 1e3)}

    # where the time goes: per request kind, and for the largest mesh, queueing and solving
    big = max(range(len(pool)), key=lambda i: pool[i]["n"])
    parts = {}
    for name, sel in (("pressure", [rq["kind"] == "pressure" for rq in stream]),
                      ("momentum", [rq["kind"] == "momentum" for rq in stream]),
                      ("largest_mesh", [rq["pool"] == big for rq in stream])):
        sel = np.array(sel)
        if sel.any():
            parts[name] = {"requests": int(sel.sum()), "latency_ms": pct(lat[sel]), "queued_ms": pct(wait[sel]),
                           "running_ms": pct(run[sel])}
    return results, {"requests": len(stream), "solves": n_solves, "wall_s": wall, "solves_per_s": n_solves / wall,
                     "requests_per_s": len(stream) / wall, "latency_ms": pct(lat), "by_kind": parts}
