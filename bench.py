"""Pressure-solve benchmark on one MI355X.

A step is one pressure solve of the 2.06M-cell jittered-tet Poisson system
(70³ hexes × 6 Kuhn tets, ±20 % interior jitter, seed 1; fvm::laplacian with
fixedValue 0 on xMax and its non-orthogonal correction evaluated on a smooth p⁰) by
GAMG-preconditioned PCG from x = 0 to normalised residual 1e-6, FP64, with the matrix,
hierarchy, b and x resident in HBM. Every step re-zeroes x on the device (device copy)
and solves again. The timed solves run without per-kernel event recording; a
separate run of solves with event recording gives the level-0 Amul duration.

Reported: solves/s (value), p50/p99 solve latency, iterations, the true residual of
the last solution (checked against the tolerance), setup_ms (the Galerkin coarse
matrices and coarsest LU rebuilt after a change of the matrix, which the reference's
solveLdu pays on every call), the level-0 Amul roofline (algorithmic bytes 24n + 16nf
per launch over its HIP-event-timed average duration) and a CPU baseline: the
reference's own solveLdu (serial build) or, without it, the oracle's restatement
(GAMG-PCG with Gauss–Seidel smoothing) on one core, with the host CPU named.

Beside it, "mixed": the same solves in mixed precision (FP64 defect correction around
FP32 GAMG-PCG), timed the same way at 1e-6 and 1e-10, with the speed-up over FP64 and
the relative error of each solution against FP64 GAMG-PCG solved to 1e-12.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "pressure solves/sec + Amul HBM GB/s at 1e-6 resid; p50/p99 latency, mixed vs FP64"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """Model name of the host CPU (from /proc/cpuinfo) and its logical CPU count."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def cpu_baseline(m, b, tol, budget_s=30.0):
    """The same system's GAMG-PCG solve (Gauss–Seidel smoothing) on one host core: the
    reference's own solveLdu (its serial build, oracle/_ref) when present, otherwise the
    oracle's C port of it."""
    from oracle import oracle as O
    from voluma_amd import workloads as W

    ni = m.n_internal
    cpu_name = "%s (%d logical CPUs, 1 used)" % host_cpu()
    d, up, ic, bs = O.laplacian(m, W.bc_fixed_xmax(m))
    d, _ = O.fold(m, d, np.zeros(m.n_cells), ic, bs)
    try:
        from oracle import ref
        have_ref = ref.available()
    except Exception:  # the reference build is optional
        have_ref = False
    if have_ref:
        rm = ref.RefMesh(m)
        t0 = time.perf_counter()
        _, p = rm.solve(d, up, None, b, np.zeros(m.n_cells), solver="PCG", preconditioner="GAMG", tolerance=tol,
                        max_iter=1000)
        dt = time.perf_counter() - t0
        return {"value": 1.0 / dt, "unit": "solves/s", "cores": 1, "kind": "reference", "host_cpu": cpu_name,
                "sample": f"1 full GAMG(GS)-PCG solve of the same system to {tol:g} by the reference's own "
                          f"solveLdu (serial build), {p['iterations']} iterations, {dt:.1f} s"}
    om = O.LduMesh(m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights())
    t0 = time.perf_counter()
    n_solves = 0
    iters = 0
    while True:
        _, p = om.solve(d, up, None, b, preconditioner="GAMG", tolerance=tol, max_iter=1000)
        n_solves += 1
        iters = p["iterations"]
        if time.perf_counter() - t0 > budget_s / 3 or n_solves >= 3:
            break
    dt = time.perf_counter() - t0
    return {"value": n_solves / dt, "unit": "solves/s", "cores": 1, "kind": "port", "host_cpu": cpu_name,
            "sample": f"{n_solves} full GAMG(GS)-PCG solve(s) of the same system to {tol:g}, {iters} iterations, "
                      f"{dt:.1f} s"}


def batch_leg(va, W, ctx, args):
    """256 independent pressure systems (lognormal 5k–200k cells, median 30k, box-hex or
    jittered tets with aspect up to 4:1, one random Dirichlet patch each) solved as one
    batch to 1e-6 by GAMG-PCG, FP64 and mixed; solves/s over whole batches."""
    t0 = time.perf_counter()
    specs = W.batch_specs(n_systems=args.batch_systems)
    B, b, meshes = W.batch_system(ctx, specs)
    setup_s = time.perf_counter() - t0
    db = va.DeviceArray.from_host(ctx, b)
    dx = va.DeviceArray(ctx, B.n_total)
    dz = va.DeviceArray.from_host(ctx, np.zeros(B.n_total))
    out = {"systems": B.n_systems, "cells": B.n_total, "setup_s": setup_s,
           "levels_histogram": np.bincount(B.levels()).tolist()}
    reps = max(3, args.steps // 4)
    for name, prec in (("fp64", va.FP64), ("mixed", va.MIXED)):
        ctl = va.SolverControls("PCG", "GAMG", args.tol, 0.0, 1000, gamgPreSweeps=args.pre, gamgPostSweeps=args.post,
                                gamgOmega=args.omega, precision=prec)
        lat = []
        for k in range(reps + 1):
            dx.copy_from(dz)
            ctx.synchronize()
            t = time.perf_counter()
            perf = B.solve(db, dx, ctl)
            ctx.synchronize()
            if k > 0:  # first batch warms up
                lat.append(time.perf_counter() - t)
            if not all(p.converged for p in perf):
                raise RuntimeError(f"batch {name}: a system did not converge")
        its = [p.iterations for p in perf]
        ms = float(np.median(lat)) * 1e3
        out[name] = {"ms_per_batch": ms, "solves_per_s": B.n_systems / (ms / 1e3),
                     "iterations_min_median_max": [min(its), int(np.median(its)), max(its)]}
    out["mixed_speedup_vs_fp64"] = out["fp64"]["ms_per_batch"] / out["mixed"]["ms_per_batch"]
    log(f"batch: {out['fp64']['solves_per_s']:.0f} solves/s FP64, {out['mixed']['solves_per_s']:.0f} mixed "
        f"({B.n_systems} systems, {B.n_total} cells)")
    B.close()
    return out


def amul_cache_states(va, ctx, lm, A, n):
    """Level-0 A·x (unfused, vol_amul) timed by events around the kernel in two cache
    states: rotating over three copies of the matrix and of x, y, so that each launch
    finds its data evicted from the 256 MB Infinity Cache by the two launches before it
    (DRAM-bound), and the same copy back to back (warm). Returns (dram_us, warm_us)."""
    from voluma_amd import workloads as W

    d_h, u_h, _ = A.get()
    mats = [A] + [va.LduMatrix(lm) for _ in range(2)]
    for M in mats[1:]:
        M.set(d_h, u_h)
    xs = [va.DeviceArray.from_host(ctx, W.splitmix_uniform(n, 50 + k)) for k in range(3)]
    ys = [va.DeviceArray(ctx, n) for _ in range(3)]
    for k in range(3):
        va.amul(mats[k], xs[k], ys[k])
    ctx.timing(True)
    ctx.read_timing(reset=True)
    for _ in range(20):
        for k in range(3):
            va.amul(mats[k], xs[k], ys[k])
    ms_d, cnt_d = ctx.read_timing(reset=True)
    for _ in range(60):
        va.amul(A, xs[0], ys[0])
    ms_w, cnt_w = ctx.read_timing(reset=True)
    ctx.timing(False)
    for a in xs + ys:
        a.free()
    for M in mats[1:]:
        M.close()
    return ms_d / cnt_d * 1e3, ms_w / cnt_w * 1e3


def serve_leg(va, ctx, args):
    """Solve server: 4 workers (own HIP streams), 200 GB HBM budget, a pool of 8 meshes
    with log-uniform sizes 10k–2M cells; Poisson arrivals at --serve-rate of 60 %
    pressure (GAMG-PCG) / 40 % momentum (PBiCGStab) requests with 2–3 right-hand sides
    each, to 1e-6, FP64 and mixed. Also the saturated throughput (all arrivals at once)."""
    from voluma_amd import serving as SV

    t0 = time.perf_counter()
    pool = SV.build_pool(ctx, SV.pool_specs())
    server = va.Server(0, slots=4, hbm_budget=200e9)
    SV.register_pool(server, pool)
    setup_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    for e in pool:  # every worker's copy, matrix storage and batch of copies, before the stream
        server.prefetch(e["id"])
    prefetch_s = time.perf_counter() - t1
    out = {"workers": 4, "hbm_budget_gb": 200, "pool_cells": [e["n"] for e in pool], "setup_s": setup_s,
           "prefetch_s": prefetch_s,
           "rate_per_s": args.serve_rate, "requests": args.serve_requests}
    warm = SV.request_stream(pool, rate=1e6, n_requests=3 * len(pool), seed=1)
    for prec in (va.FP64, va.MIXED):
        SV.run_stream(server, pool, warm, prec)  # kernels and captured iterations warm
    out["bytes_in_use_gb"] = server.stats()["bytes_in_use"] / 1e9
    for name, prec in (("fp64", va.FP64), ("mixed", va.MIXED)):
        stream = SV.request_stream(pool, rate=args.serve_rate, n_requests=args.serve_requests, seed=9)
        res, st = SV.run_stream(server, pool, stream, prec)
        if not all(p.converged for _, perf in res for p in perf):
            raise RuntimeError(f"serve {name}: a solve did not converge")
        sat = SV.request_stream(pool, rate=1e6, n_requests=min(120, args.serve_requests), seed=10)
        _, st_sat = SV.run_stream(server, pool, sat, prec)
        st["saturated_solves_per_s"] = st_sat["solves_per_s"]
        out[name] = st
        log(f"serve {name}: p50 {st['latency_ms']['p50']:.1f} ms, p99 {st['latency_ms']['p99']:.1f} ms at "
            f"{args.serve_rate:g} req/s; {st['solves_per_s']:.0f} solves/s; saturated {st_sat['solves_per_s']:.0f}")
        for kind, v in st["by_kind"].items():
            log(f"  {kind} ({v['requests']}): p50 {v['latency_ms']['p50']:.1f} p99 {v['latency_ms']['p99']:.1f} ms; "
                f"queued p99 {v['queued_ms']['p99']:.1f}, running p50 {v['running_ms']['p50']:.1f} "
                f"p99 {v['running_ms']['p99']:.1f} ms")
    server.close()
    return out


def cpu_baseline_dilu(m, d, u, lo, b, x0):
    """The same momentum solve (PBiCGStab + DILU to 1e-10, plain sums) by the oracle on
    one host core."""
    from oracle import oracle as O
    om = O.LduMesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, None)
    t = time.perf_counter()
    _, po = om.solve(d, u, lo, b, x0, solver="PBiCGStab", preconditioner="DILU", tolerance=1e-10, max_iter=1000)
    return {"ms_per_solve": (time.perf_counter() - t) * 1e3, "iterations": po["iterations"], "cores": 1,
            "kind": "port"}


def corrector_ops_leg(va, W, ctx, m, reps=20):
    """The explicit pressure-corrector operators (Fvc.h) on the benchmark mesh, device
    pointers: milliseconds per call and GB/s of their algorithmic bytes (arrays each
    must read or write once, int32 addressing included)."""
    n, ni, nF = m.n_cells, m.n_internal, m.n_faces
    nb = nF - ni
    lm = va.LduMesh(ctx, n, m.owner[:ni], m.neighbour, None)
    lm.set_boundary(m.owner[ni:])
    g = m.geometry()
    lm.set_geometry(g["V"], g["Sf"], g["magSf"], g["w"], g["d"], g["skew"], g["skewed"])
    L = va.lib()
    dev = lambda a, dt=np.float64: va.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=dt).ravel(), dt)
    p = dev(2.0 * W.splitmix_uniform(n, 81) - 1.0)
    U = dev(2.0 * W.splitmix_uniform(3 * n, 82) - 1.0)
    bfix = dev(np.ones(nb), np.int32)
    pB, UB = dev(np.zeros(nb)), dev(np.zeros(3 * nb))
    gp, gU = va.DeviceArray(ctx, 3 * n), va.DeviceArray(ctx, 9 * n)
    fS, phi, cell = va.DeviceArray(ctx, nF), va.DeviceArray(ctx, nF), va.DeviceArray(ctx, n)
    D = va.VOL_DEVICE
    addr = 8 * ni + 4 * nb
    ops = {
        "grad_ls_scalar": (lambda: L.vol_fvc_grad(lm._h, va.GRAD_LEAST_SQUARES, 1, p.ptr, bfix.ptr, pB.ptr, gp.ptr, D),
                           32 * n + 24 * nF + 48 * nb + addr),
        "grad_ls_vector": (lambda: L.vol_fvc_grad(lm._h, va.GRAD_LEAST_SQUARES, 3, U.ptr, bfix.ptr, UB.ptr, gU.ptr, D),
                           96 * n + 24 * nF + 64 * nb + addr),
        "grad_gauss_skew": (lambda: L.vol_fvc_grad(lm._h, va.GRAD_GAUSS_SKEW, 1, p.ptr, bfix.ptr, pB.ptr, gp.ptr, D),
                            88 * n + 24 * nF + 48 * nb + 32 * ni + 24 * nF + addr),
        "interpolate_skew": (lambda: L.vol_fvc_interpolate(lm._h, 1, p.ptr, gp.ptr, None, fS.ptr, D),
                             32 * n + 40 * ni + 8 * nF + addr),
        "flux_phiHbyA": (lambda: L.vol_fvc_flux(lm._h, U.ptr, gU.ptr, bfix.ptr, UB.ptr, phi.ptr, D),
                         96 * n + 32 * ni + 32 * nF + 28 * nb + addr),
        "div": (lambda: L.vol_fvc_div(lm._h, phi.ptr, cell.ptr, D), 16 * n + 8 * nF + addr),
    }
    out = {}
    for name, (fn, nbytes) in ops.items():
        va._check(fn())
        ctx.synchronize()
        t = time.perf_counter()
        for k in range(reps):
            va._check(fn())
        ctx.synchronize()
        sec = (time.perf_counter() - t) / reps
        out[name] = {"ms": sec * 1e3, "algorithmic_bytes": nbytes, "gbs": nbytes / sec / 1e9}
    for a in (p, U, bfix, pB, UB, gp, gU, fS, phi, cell):
        a.free()
    lm.close()
    log("corrector ops: " + ", ".join(f"{k} {v['ms']:.2f} ms ({v['gbs']:.0f} GB/s)" for k, v in out.items()))
    return out


def piso_step_leg(va, W, ctx, m, args, cpu):
    """Whole incompressible time steps resident on the GPU (vol_flow: the reference's
    IncompressibleFlow::step, upwind convection, Euler, 2 pressure correctors) on the
    benchmark mesh as a lid-driven cavity (zMax moves at (1, 0, 0), p zero-gradient with a
    reference cell), from the synthetic divergence-free velocity of the momentum inputs:
    ms per step with the reference's solvers (momentum PBiCGStab + DILU, pressure PCG +
    GAMG smoothed by Gauss–Seidel) and with the pressure's GAMG smoothed by damped Jacobi.
    With cpu: the reference's own step on the CPU on tet35 beside the GPU's there."""
    from voluma_amd.piso import PisoFlow

    def flow_for(mesh, inp, p_controls=None, p_final=None):
        f = PisoFlow(ctx, mesh, inp["nu"], [(0, 0, 0)] * 5 + [(1, 0, 0)], [None] * 6, ddt="Euler",
                     p_controls=p_controls, p_final_controls=p_final)
        f.U, f.p, f.phi = inp["U0"], np.zeros(mesh.n_cells), inp["phi"]
        return f

    def timed(f, dt, steps):
        f.step(dt)  # warm-up step: builds the solvers' level orders
        ctx.synchronize()
        t = time.perf_counter()
        its = []
        for _ in range(steps):
            f.step(dt)
            its.append([pf.iterations for pf in f.perf])
        ctx.synchronize()
        return (time.perf_counter() - t) / steps * 1e3, its[-1]

    steps = 1 if args.profile_legs else 3
    inp = W.momentum_inputs(m)
    out = {"mesh": args.workload, "cells": m.n_cells, "dt": inp["dt"], "steps_timed": steps}
    f = flow_for(m, inp)
    out["reference_solvers_ms"], out["reference_solvers_iterations"] = timed(f, inp["dt"], steps)
    f.close()
    jac = va.SolverControls("PCG", "GAMG", 1e-8, 0.01, 2000, gamgPreSweeps=args.pre, gamgPostSweeps=args.post,
                            gamgOmega=args.omega)
    jac_final = va.SolverControls("PCG", "GAMG", 1e-8, 0.0, 2000, gamgPreSweeps=args.pre,
                                  gamgPostSweeps=args.post, gamgOmega=args.omega)
    f = flow_for(m, inp, jac, jac_final)
    out["jacobi_gamg_ms"], out["jacobi_gamg_iterations"] = timed(f, inp["dt"], steps)
    f.close()
    log(f"PISO step ({args.workload}): {out['reference_solvers_ms']:.1f} ms with the reference's solvers, "
        f"{out['jacobi_gamg_ms']:.1f} ms with Jacobi-smoothed GAMG")
    if cpu and not args.profile_legs:
        from oracle import ref
        if ref.available():
            ms = W.make_mesh("tet35")
            inp35 = W.momentum_inputs(ms)
            f = flow_for(ms, inp35)
            gpu_ms, _ = timed(f, inp35["dt"], 2)
            f.close()
            f = flow_for(ms, inp35, jac, jac_final)
            gpu_jac_ms, _ = timed(f, inp35["dt"], 2)
            f.close()
            rm = ref.RefMesh(ms)
            t = time.perf_counter()
            rm.piso_run(1, inp35["nu"], inp35["dt"], inp35["U0"], np.zeros(ms.n_cells), inp35["phi"],
                        [(0, 0, 0)] * 5 + [(1, 0, 0)], [None] * 6)
            cpu_ms = (time.perf_counter() - t) * 1e3
            out["cpu_reference"] = {"mesh": "tet35", "cells": ms.n_cells, "cpu_ms_per_step": cpu_ms,
                                    "gpu_ms_per_step": gpu_ms, "gpu_jacobi_gamg_ms_per_step": gpu_jac_ms,
                                    "cores": 1, "kind": "reference",
                                    "note": "the reference's IncompressibleFlow::step (serial build, first step "
                                            "from the same fields) on the GPU host's CPU"}
            log(f"PISO step tet35: reference CPU {cpu_ms:.0f} ms, GPU {gpu_ms:.1f} ms (reference solvers), "
                f"{gpu_jac_ms:.1f} ms (Jacobi-smoothed GAMG)")
    return out


def reference_solvers_leg(va, W, ctx, m, A, db, dx, dzero, args, cpu):
    """The reference's own solver choices on the GPU: GAMG-PCG with its Gauss–Seidel
    smoother (2/2 sweeps) on the pressure system, and its default momentum solver
    (PBiCGStab + DILU) on the momentum matrix of velocity component 0, both as
    level-scheduled sweeps; with cpu, the momentum solve's CPU baseline too."""
    out = {}
    short = args.profile_legs  # a kernel-trace run: one short solve of each kind
    ctl = va.SolverControls("PCG", "GAMG", args.tol, 0.0, 3 if short else 1000, gamgSmoother=va.GAMG_GAUSS_SEIDEL)
    for k in range(0 if short else 2):
        dx.copy_from(dzero)
        va.solve_ldu(A, db, dx, ctl)
    ctx.synchronize()
    steps = 1 if short else max(2, args.steps // 4)
    t = time.perf_counter()
    its = []
    for k in range(steps):
        dx.copy_from(dzero)
        p = va.solve_ldu(A, db, dx, ctl)
        its.append(p.iterations)
    ctx.synchronize()
    el = time.perf_counter() - t
    out["gamg_gauss_seidel"] = {"solves_per_s": steps / el, "ms_per_solve": el / steps * 1e3,
                                "iterations": int(np.median(its)), "converged": bool(p.converged)}
    lm2, A2, b2, inp = W.momentum_system(ctx, m, seed=1, comp=0)
    x0 = np.ascontiguousarray(inp["U0"][:, 0])
    dx0, db2, dxm = va.DeviceArray.from_host(ctx, x0), va.DeviceArray.from_host(ctx, b2), va.DeviceArray(ctx, m.n_cells)
    mctl = va.SolverControls("PBiCGStab", "DILU", 1e-10, 0.0, 1000)
    for k in range(0 if short else 2):
        dxm.copy_from(dx0)
        va.solve_ldu(A2, db2, dxm, mctl)
    ctx.synchronize()
    t = time.perf_counter()
    for k in range(steps):
        dxm.copy_from(dx0)
        p = va.solve_ldu(A2, db2, dxm, mctl)
    ctx.synchronize()
    el = time.perf_counter() - t
    out["momentum_pbicgstab_dilu"] = {"ms_per_solve": el / steps * 1e3, "iterations": p.iterations,
                                      "tolerance": 1e-10, "converged": bool(p.converged)}
    if cpu:
        d, u, lo = A2.get()
        out["momentum_pbicgstab_dilu"]["cpu_baseline"] = cpu_baseline_dilu(m, d, u, lo, b2, x0)
    for a in (dx0, db2, dxm):
        a.free()
    A2.close()
    lm2.close()
    log(f"reference solvers: GS-GAMG {out['gamg_gauss_seidel']['ms_per_solve']:.1f} ms "
        f"({out['gamg_gauss_seidel']['iterations']} it), DILU momentum "
        f"{out['momentum_pbicgstab_dilu']['ms_per_solve']:.1f} ms ({p.iterations} it)")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tet70")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-systems", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-ref-solvers", action="store_true", help="skip the GS-GAMG and DILU legs")
    ap.add_argument("--no-corrector-ops", action="store_true", help="skip the fvc operator timings")
    ap.add_argument("--no-piso", action="store_true", help="skip the whole PISO time-step leg")
    ap.add_argument("--profile-legs", action="store_true",
                    help="kernel-trace aid: one short GS-GAMG (3 iterations) and DILU solve, no warm-up")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="request arrivals per second")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--pre", type=int, default=1)
    ap.add_argument("--post", type=int, default=1)
    ap.add_argument("--omega", type=float, default=0.9)
    ap.add_argument("--inner-rel", type=float, default=1e-2, help="mixed precision: inner residual reduction")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import voluma_amd as va
    from voluma_amd import workloads as W

    t0 = time.perf_counter()
    m = W.make_mesh(args.workload)
    log(f"mesh {args.workload}: {m.n_cells} cells, {m.n_internal} internal faces ({time.perf_counter() - t0:.1f} s)")
    ctx = va.Context(0)
    t0 = time.perf_counter()
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True, nonorth=True)
    levels = [L["n"] for L in lm.gamg_levels()]
    level_faces = [int(L["faceMap"].size) for L in lm.gamg_levels()]
    log(f"GPU setup (assembly + hierarchy) {time.perf_counter() - t0:.1f} s, levels {levels}")
    agg_s, agg_passes, agg_rounds = lm.gamg_setup_info()
    os.environ["VOLUMA_AGGLOMERATION"] = "host"  # the same hierarchy by the host loop, for comparison
    lm_host = va.LduMesh(ctx, m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    del os.environ["VOLUMA_AGGLOMERATION"]
    agglomeration = {"device_s": agg_s, "passes": agg_passes, "rounds": agg_rounds,
                     "host_s": lm_host.gamg_setup_info()[0],
                     "same_maps": all(np.array_equal(a["agg"], b["agg"]) and np.array_equal(a["faceMap"], b["faceMap"])
                                      for a, b in zip(lm.gamg_levels(), lm_host.gamg_levels()))}
    lm_host.close()
    log(f"agglomeration: device {agg_s * 1e3:.1f} ms ({agg_passes} passes, {agg_rounds} rounds), host "
        f"{agglomeration['host_s'] * 1e3:.1f} ms, same maps {agglomeration['same_maps']}")
    n, nf = m.n_cells, m.n_internal
    db = va.DeviceArray.from_host(ctx, b)
    dx = va.DeviceArray(ctx, n)
    dzero = va.DeviceArray.from_host(ctx, np.zeros(n))
    ctl = va.SolverControls("PCG", "GAMG", args.tol, 0.0, 1000, sums=va.SUMS_FAST, gamgPreSweeps=args.pre,
                            gamgPostSweeps=args.post, gamgOmega=args.omega)

    def run(controls, steps, timing=False):
        """steps solves from x = 0 on the device; (elapsed s, latencies s, iterations, perf, x)."""
        if timing:
            ctx.timing(True)
            ctx.read_timing(reset=True)
        lat, its = [], []
        ctx.synchronize()
        t_start = time.perf_counter()
        p = None
        for k in range(steps):
            t1 = time.perf_counter()
            dx.copy_from(dzero)  # x = 0 on the device
            p = va.solve_ldu(A, db, dx, controls)
            lat.append(time.perf_counter() - t1)
            its.append(p.iterations)
            if not p.converged:
                raise RuntimeError(f"solve {k} did not converge: {p}")
        ctx.synchronize()
        elapsed = time.perf_counter() - t_start
        return elapsed, np.array(lat), its, p, dx.download()

    # achievable HBM bandwidth on this GPU: streaming copy of 2^27 doubles (1 GiB each way)
    ncopy = 1 << 27
    src, dst = va.DeviceArray(ctx, ncopy), va.DeviceArray(ctx, ncopy)
    dst.copy_from(src, 2)
    t_copy = dst.copy_from(src, 10)
    stream_gbs = 2 * ncopy * 8 / t_copy / 1e9
    src.free()
    dst.free()
    log(f"stream copy: {stream_gbs:.0f} GB/s")

    run(ctl, args.warmup)
    elapsed, lat, iters, p, x = run(ctl, args.steps)
    lat_ms = lat * 1e3
    # the last solution's true residual: the metric counts only solves that reached it
    r = b - va.amul(A, x)
    true_res = float(np.abs(r).sum() / p.normFactor)
    if not true_res <= 1.5 * args.tol:
        raise RuntimeError(f"true residual {true_res:.3e} above the tolerance {args.tol:g}")
    # level-0 A·p duration: solves with event recording around every A·p launch
    run(ctl, max(3, args.steps // 4), timing=True)
    amul_ms, amul_cnt = ctx.read_timing(reset=True)
    ctx.timing(False)
    # setup after a change of the matrix: set the same coefficients again (device to
    # device), which makes the next solve rebuild the Galerkin levels and coarsest LU
    d_h, u_h, _ = A.get()
    dd, du = va.DeviceArray.from_host(ctx, d_h), va.DeviceArray.from_host(ctx, u_h)
    t_set = []
    for k in range(3):
        ctx.synchronize()
        t1 = time.perf_counter()
        A.set(dd, du)
        dx.copy_from(dzero)
        va.solve_ldu(A, db, dx, ctl)
        ctx.synchronize()
        t_set.append(time.perf_counter() - t1)
    setup_ms = (float(np.median(t_set)) - float(np.median(lat))) * 1e3
    dd.free()
    du.free()
    amul_dram_us, amul_warm_us = amul_cache_states(va, ctx, lm, A, n)
    amul_bytes = 24 * n + 16 * nf
    amul_avg_s = amul_ms / max(amul_cnt, 1) / 1e3
    achieved = amul_bytes / amul_avg_s / 1e9 if amul_cnt else None
    traffic = None
    prof = os.path.join(ROOT, "profiles", f"amul_{args.workload}_traffic.json")
    if os.path.exists(prof):
        with open(prof) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")

    mixed = None
    if not args.no_mixed:
        log("mixed precision ...")
        mctl = dataclasses.replace(ctl, precision=va.MIXED, innerRelTol=args.inner_rel)
        _, _, _, pref, xref = run(dataclasses.replace(ctl, tolerance=1e-12, maxIter=3000), 1)
        nref = np.linalg.norm(xref)
        run(mctl, args.warmup)
        m_el, m_lat, m_its, mp, mx = run(mctl, args.steps, timing=True)
        a32_ms, a32_cnt = ctx.read_timing(reset=True)
        ctx.timing(False)
        k10 = max(2, args.steps // 4)
        f_el10, _, f_its10, _, fx10 = run(dataclasses.replace(ctl, tolerance=1e-10, maxIter=3000), k10)
        m_el10, _, m_its10, mp10, mx10 = run(dataclasses.replace(mctl, tolerance=1e-10, maxIter=3000), k10)
        a32_bytes = 12 * n + 12 * nf
        a32_s = a32_ms / max(a32_cnt, 1) / 1e3
        mixed = {
            "value": args.steps / m_el, "unit": "solves/s", "ms_per_step": m_el / args.steps * 1e3,
            "speedup_vs_fp64": (args.steps / m_el) / (args.steps / elapsed),
            "latency_ms": {"p50": float(np.percentile(m_lat * 1e3, 50)), "p99": float(np.percentile(m_lat * 1e3, 99))},
            "iterations": int(np.median(m_its)), "outer_iterations": mp.outerIterations,
            "final_residual": mp.finalResidual,
            "rel_err_vs_fp64_1e-12": float(np.linalg.norm(mx - xref) / nref),
            "fp64_rel_err_vs_fp64_1e-12": float(np.linalg.norm(x - xref) / nref),
            "tol_1e-10": {"mixed_ms": m_el10 / k10 * 1e3, "fp64_ms": f_el10 / k10 * 1e3,
                          "speedup_vs_fp64": f_el10 / m_el10, "mixed_iterations": int(np.median(m_its10)),
                          "mixed_outer_iterations": mp10.outerIterations, "fp64_iterations": int(np.median(f_its10)),
                          "rel_err_vs_fp64_1e-12": float(np.linalg.norm(mx10 - xref) / nref),
                          "fp64_rel_err_vs_fp64_1e-12": float(np.linalg.norm(fx10 - xref) / nref)},
            "amul32": {"achieved_gbs": a32_bytes / a32_s / 1e9 if a32_cnt else None, "avg_us": a32_s * 1e6,
                       "algorithmic_bytes": a32_bytes, "launches": a32_cnt},
        }
        log(f"mixed: {mixed['value']:.2f} solves/s ({mixed['speedup_vs_fp64']:.2f}x), "
            f"err {mixed['rel_err_vs_fp64_1e-12']:.2e}; 1e-10: {mixed['tol_1e-10']['speedup_vs_fp64']:.2f}x, "
            f"err {mixed['tol_1e-10']['rel_err_vs_fp64_1e-12']:.2e}")
    batch = None
    if not args.no_batch:
        batch = batch_leg(va, W, ctx, args)
    serve = None
    if not args.no_serve:
        serve = serve_leg(va, ctx, args)
    corrector_ops = None
    if not args.no_corrector_ops:
        corrector_ops = corrector_ops_leg(va, W, ctx, m, 2 if args.profile_legs else 20)
    piso = None
    if not args.no_piso:
        piso = piso_step_leg(va, W, ctx, m, args, not args.no_cpu_baseline)
    ref_solvers = None
    if not args.no_ref_solvers:
        ref_solvers = reference_solvers_leg(va, W, ctx, m, A, db, dx, dzero, args, not args.no_cpu_baseline)

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: pressure Poisson (laplacian with non-orthogonal correction), "
                               f"GAMG-PCG (damped-Jacobi V-cycle {args.pre}/{args.post}, omega {args.omega}) to "
                               f"normalised residual {args.tol:g}, FP64",
                   "cells": n, "internal_faces": nf, "gamg_levels": levels, "gamg_level_faces": level_faces,
                   "precision": "fp64"},
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)), "p99": float(np.percentile(lat_ms, 99)),
                       "min": float(lat_ms.min())},
        "mixed": mixed,
        "batch": batch,
        "serve": serve,
        "reference_solvers": ref_solvers,
        "corrector_ops": corrector_ops,
        "piso_step": piso,
        "iterations": int(np.median(iters)),
        "final_residual": p.finalResidual,
        "true_residual": true_res,
        "setup_ms": setup_ms,
        "agglomeration": agglomeration,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sell<OP_AMUL> level-0 A·p", "algorithmic_bytes": amul_bytes,
                     "avg_us": amul_avg_s * 1e6, "launches": amul_cnt, "stream_copy_gbs": stream_gbs,
                     "traffic_kind": "FETCH_SIZE/WRITE_SIZE counter bytes between L2 and the fabric per launch "
                                     "(Infinity Cache hits included)",
                     "dram_us": amul_dram_us, "warm_us": amul_warm_us,
                     "dram_gbs": None if traffic is None else traffic / amul_dram_us / 1e3},
    }
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle, 1 thread) ...")
        out["cpu_baseline"] = cpu_baseline(m, b, args.tol)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
