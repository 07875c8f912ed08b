"""Summaries of a rocprofv3 kernel trace of the solver (``rocprofv3 --kernel-trace -d OUT
-o run -- python3 bench.py ...`` writes ``OUT/run_results.db``, an SQLite database whose
``kernels`` view lists every dispatch with its grid and start/end times in ns).

    python -m voluma_amd.profiling DB stats [N]        per (kernel, grid): calls, total, avg, min, max
    python -m voluma_amd.profiling DB levels BENCH     one PCG iteration of the headline solve,
                                                       per dispatch: level, average time over all
                                                       iterations, bytes, GB/s (BENCH: a file
                                                       holding bench.py's JSON line, for the
                                                       level sizes)
    python -m voluma_amd.profiling DB precisions       the FP64 PCG iteration beside the FP32
                                                       inner iteration of mixed precision,
                                                       dispatch by dispatch
    python -m voluma_amd.profiling FDB traffic WDB BENCH
                                                       counter bytes per dispatch of one PCG
                                                       iteration: FDB and WDB from two
                                                       ``rocprofv3 --pmc FETCH_SIZE`` and
                                                       ``--pmc WRITE_SIZE`` runs of the bench,
                                                       scaled by its copy kernel

The ``levels`` table splits the trace into PCG iterations at the level-0 A·p launches
(``k_sell<0, 0, true, ...>``), keeps the iterations whose dispatch sequence is the most
common one, and walks the V-cycle's fixed order (residual and restriction down the
levels, separate or fused, coarsest solve, prolongation sweeps up) to name each
dispatch's level. Bytes are what the kernel has to move given the stored layout,
without its padding: per row of a level the vectors it reads and writes, per
off-diagonal entry a coefficient and a column (8 + 4 bytes in FP64; 8 + 2 on the packed
level-0 rows, ``k_sell_packed``, and the packed coarse SELL rows, ``k_sell_pk``), the
restriction's member lists and the coarsest LU
factors; a post-sweep that gathers the prolongation (x-load 3) its agg and coarse x, one
with a plain x-load above level 0 the member lists and the finer z it adds its result
to (the pushed prolongation); the x/r update counts the level-0 z it writes for the
V-cycle (FP64 PCG)."""
import json
import re
import sqlite3
import sys
from collections import Counter


def _short(name):
    name = re.sub(r"void vol::(\(anonymous namespace\)::)?", "", name)
    name = re.sub(r"\(.*\)$", "", name)
    return name.replace("vol::", "")


def dispatches(db):
    """[(kernel, grid in workgroups, start ns, duration ns)] in start order."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, grid_x, workgroup_x, start, duration from kernels order by start").fetchall()
    con.close()
    return [(_short(n), g // max(wg, 1), s, d) for n, g, wg, s, d in rows]


def stats(db, top=30):
    rows = dispatches(db)
    busy = sum(r[3] for r in rows)
    agg = {}
    for k, g, _, d in rows:
        a = agg.setdefault((k, g), [0, 0, 1 << 62, 0])
        a[0] += 1
        a[1] += d
        a[2] = min(a[2], d)
        a[3] = max(a[3], d)
    out = [f"busy {busy / 1e6:.3f} ms, {len(rows)} dispatches", "",
           "| kernel | grid (WG) | calls | total ms | avg us | min us | max us | % |", "|---|---|---|---|---|---|---|---|"]
    for (k, g), a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
        out.append(f"| `{k}` | {g} | {a[0]} | {a[1] / 1e6:.3f} | {a[1] / a[0] / 1e3:.2f} | {a[2] / 1e3:.2f} | "
                   f"{a[3] / 1e3:.2f} | {100 * a[1] / busy:.1f} |")
    return "\n".join(out)


def _template_args(kernel):
    m = re.search(r"<(.*)>", kernel)
    return [a.strip() for a in m.group(1).split(",")] if m else []


def _is_sell(kernel):
    """A SELL-64 gather (slot-major or packed rows)."""
    return kernel.startswith(("k_sell<", "k_sell_packed<", "k_sell_pk<", "k_sell_fin<", "k_sell_packed_fin<"))


def _is_amul(kernel):
    a = _template_args(kernel)
    if kernel.startswith(("k_sell_fin<", "k_sell_packed_fin<")):  # always the fused dot
        return len(a) >= 2 and a[0] == "0" and a[1] == "0"
    return _is_sell(kernel) and len(a) >= 3 and a[0] == "0" and a[1] == "0" and a[2] == "true"


def _iterations(rows, elem_type=None):
    """Dispatch sequences between consecutive level-0 A·p launches (of one value type)."""
    starts = [i for i, r in enumerate(rows)
              if _is_amul(r[0]) and (elem_type is None or _template_args(r[0])[-1] == elem_type)]
    return [rows[a:b] for a, b in zip(starts, starts[1:])]


def _common(its):
    """The most common dispatch sequence, the iterations that follow it, and the
    average duration (µs) of each of its dispatches."""
    shape = Counter(tuple((r[0], r[1]) for r in it) for it in its).most_common(1)[0][0]
    keep = [it for it in its if tuple((r[0], r[1]) for r in it) == shape]
    avg = [sum(it[i][3] for it in keep) / len(keep) / 1e3 for i in range(len(shape))]
    return shape, keep, avg


def precisions(db):
    rows = dispatches(db)
    i64, i32 = _iterations(rows, "double"), _iterations(rows, "float")
    if not i64 or not i32:
        raise SystemExit("the trace needs FP64 and FP32 (mixed precision) A·p launches")
    s64, k64, a64 = _common(i64)
    s32, k32, a32 = _common(i32)
    if [g for _, g in s64] != [g for _, g in s32]:
        raise SystemExit("the FP64 and FP32 iterations launch different grids")
    out = [f"FP64: {len(k64)} of {len(i64)} iterations with the common sequence of {len(s64)} dispatches, "
           f"{sum(a64):.1f} us of kernel time per iteration; FP32 inner: {len(k32)} of {len(i32)} iterations, "
           f"{sum(a32):.1f} us ({sum(a64) / sum(a32):.2f}x)", "",
           "| # | FP64 kernel | grid | avg us | FP32 kernel | avg us | FP64/FP32 |", "|---|---|---|---|---|---|---|"]
    for i, ((k, g), (k2, _)) in enumerate(zip(s64, s32)):
        out.append(f"| {i} | `{k}` | {g} | {a64[i]:.2f} | `{k2}` | {a32[i]:.2f} | {a64[i] / a32[i]:.2f} |")
    return "\n".join(out)


def _row_bytes(op_kind, n, nnz, nc, elem, col_bytes=4, n_fine=0):
    """Bytes of one gather sweep over n rows with nnz off-diagonal entries (col_bytes per
    entry for its column: 4, or 2 for the packed level-0 rows)."""
    mat = elem * n + (elem + col_bytes) * nnz  # diag + (coefficient, column) per entry
    if op_kind == "amul":               # x, A·x, dot vector
        return mat + 3 * elem * n
    if op_kind == "resid":              # x (or z), r, residual
        return mat + 3 * elem * n
    if op_kind == "post":               # z, r, result, agg, coarse x (gathered prolongation)
        return mat + 3 * elem * n + 4 * n + elem * nc
    if op_kind == "post_push":          # z, r; pushed into the finer level: lists, z read and written
        return mat + 2 * elem * n + 4 * (n + 1) + 4 * n_fine + 2 * elem * n_fine
    return mat + 3 * elem * n


def _level_sizes(bench_json):
    """Cells and internal faces per GAMG level from the bench's JSON line."""
    cfg = None
    with open(bench_json) as fh:  # the bench's JSON line, wherever it is in the file
        for line in reversed(fh.read().splitlines()):
            try:
                cfg = json.loads(line)["config"]
                break
            except (ValueError, KeyError, TypeError):
                continue
    if cfg is None:
        raise SystemExit(f"no bench.py JSON line in {bench_json}")
    return cfg["gamg_levels"], cfg["gamg_level_faces"]


def _walk(shape, cells, faces):
    """Per dispatch of one iteration: (level, rows, entries, model bytes or None)."""
    nl = len(cells)
    elem = 8
    out = []
    # the V-cycle's order: residual at l, restriction l → l+1 (l = 0 …), coarsest solve,
    # then prolongation sweeps at l = L-2 … 0 (a streaming prolongation, when split,
    # right before the level-0 sweep)
    level = 0
    for k, g in shape:
        lev, nbytes, rws, ent = "", None, "", ""
        cb = 2 if k.startswith(("k_sell_packed<", "k_sell_pk<", "k_sell_packed_fin<")) else 4
        if _is_amul(k):
            lev, nbytes = "0", _row_bytes("amul", cells[0], 2 * faces[0], 0, elem, cb)
        elif _is_sell(k) or k.startswith("k_csrv<"):
            args = _template_args(k)
            if args[0] == "2":  # post-sweep one level up
                level -= 1
                # x-load 3: gathers x + xc[agg]; plain above level 0: pushes into the finer level
                kind = "post" if args[1] == "3" else ("post_push" if level > 0 else "resid")
            else:
                kind = "resid"
            nc = cells[level + 1] if level + 1 < nl else 0
            nfine = cells[level - 1] if level > 0 else 0
            lev, nbytes = str(level), _row_bytes(kind, cells[level], 2 * faces[level], nc, elem, cb, nfine)
        elif k.startswith("k_resid_restrict"):
            # residual of level l (its rows through the restriction lists) summed into l + 1;
            # into the coarsest also the LU substitution, and with _post the level's sweep
            n, nc = cells[level], cells[level + 1]
            lev = f"{level}→{level + 1}"
            rws, ent = n, 2 * faces[level]
            if k.startswith("k_resid_restrict_packed"):  # packed rows, reached through the member list
                mat, rowref = elem * n + (elem + 2) * ent, 4 * n
            else:  # CSR rows (or a SELL level's), reached through row descriptors
                mat, rowref = elem * n + (elem + 4) * ent, 16 * n
            nbytes = mat + 2 * elem * n + rowref + 4 * (nc + 1) + 3 * elem * nc  # r, z; rows; lists; r, z, 1/diag
            if "coarse64" in k:
                nbytes += 8 * nc * nc
                if "_post" in k:  # the sweep re-reads r, z, agg of the rows (and writes x or pushes it)
                    nbytes += 2 * elem * n + 4 * n + elem * n
            level += 1
            if "_post" in k:
                level -= 1  # its post-sweep is done: the next sweep is one level up
        elif k.startswith("k_restrict"):
            n, nc = cells[level], cells[level + 1]
            lev = f"{level}→{level + 1}"
            nbytes = elem * n + 4 * n + 4 * (nc + 1) + 3 * elem * nc  # res, lists; r, z, 1/diag
            level += 1
        elif k.startswith("k_coarse_solve"):
            nc = cells[-1]
            lev = f"{nl - 1}"
            nbytes = 8 * nc * nc + 2 * elem * nc
        elif k.startswith("k_prolong"):
            lev = f"{level}→{level - 1}"
            nbytes = 2 * elem * cells[level - 1] + 4 * cells[level - 1] + elem * cells[level]
        elif k.startswith("k_vec_fast"):
            lev = "0"
            op = _template_args(k)[0]
            if op == "5":    # x += αp, r −= αq (and z = ω·(r·rD)): p, q, x, r, 1/diag in; x, r, z out
                nbytes = 8 * elem * cells[0]
            elif op == "4":  # p = w + βp
                nbytes = 3 * elem * cells[0]
        if lev.isdigit() and not rws:
            rws, ent = cells[int(lev)], 2 * faces[int(lev)]
        out.append((lev, rws, ent, nbytes))
    return out


def levels(db, bench_json):
    cells, faces = _level_sizes(bench_json)
    rows = dispatches(db)
    its = _iterations(rows)
    if not its:
        raise SystemExit("no level-0 A·p launches in the trace")
    shape, keep, avg = _common(its)
    out = [f"{len(keep)} of {len(its)} iterations with the common sequence of {len(shape)} dispatches; "
           f"{sum(avg):.1f} us of kernel time per iteration", "",
           "| # | kernel | grid | level | rows | entries | avg us | MB | GB/s |", "|---|---|---|---|---|---|---|---|---|"]
    for i, ((k, g), us, (lev, rws, ent, nbytes)) in enumerate(zip(shape, avg, _walk(shape, cells, faces))):
        mb = f"{nbytes / 1e6:.1f}" if nbytes else ""
        gbs = f"{nbytes / us / 1e3:.0f}" if nbytes else ""
        out.append(f"| {i} | `{k}` | {g} | {lev} | {rws} | {ent} | {us:.2f} | {mb} | {gbs} |")
    return "\n".join(out)


COPY_KERNEL = "k_vec_fast<13, 0, double>"  # the bench's streaming copy: 2^27 doubles each way
COPY_BYTES = 1 << 30


def _counted(db, counter):
    """Dispatches as in dispatches(db), each with the counter's value (KiB) appended, and
    the bytes per KiB counted, from the largest launch of the bench's copy kernel."""
    con = sqlite3.connect(db)
    vals = dict(con.execute("select dispatch_id, value from counters_collection where counter_name = ?",
                            (counter,)).fetchall())
    rows = con.execute("select name, grid_x, workgroup_x, start, duration, dispatch_id from kernels "
                       "order by start").fetchall()
    con.close()
    out = [(_short(n), g // max(wg, 1), s, d, float(vals.get(di, 0.0))) for n, g, wg, s, d, di in rows]
    copies = [r[4] for r in out if r[0] == COPY_KERNEL]
    if not copies or max(copies) <= 0:
        raise SystemExit(f"{db}: no {COPY_KERNEL} launch with {counter} to calibrate against")
    return out, COPY_BYTES / (max(copies) * 1024.0)


def traffic(fetch_db, write_db, bench_json):
    """Counter bytes per dispatch of one PCG iteration (FETCH_SIZE and WRITE_SIZE from
    two runs, each scaled by the copy kernel's bytes per counted KiB in its run) beside
    the model bytes of the levels table."""
    cells, faces = _level_sizes(bench_json)
    res = []
    for db, counter in ((fetch_db, "FETCH_SIZE"), (write_db, "WRITE_SIZE")):
        rows, scale = _counted(db, counter)
        its = _iterations(rows)
        if not its:
            raise SystemExit(f"{db}: no level-0 A·p launches")
        shape, keep, _ = _common(its)
        per = [sum(it[i][4] for it in keep) / len(keep) * 1024.0 * scale for i in range(len(shape))]
        res.append((shape, keep, scale, per))
    (s_f, k_f, sc_f, fetch), (s_w, k_w, sc_w, write) = res
    if s_f != s_w:
        raise SystemExit("the two runs have different dispatch sequences")
    walk = _walk(s_f, cells, faces)
    tot_model = sum(w[3] or 0 for w in walk)
    tot_ctr = sum(fetch) + sum(write)
    out = [f"{len(k_f)} and {len(k_w)} iterations with the common sequence of {len(s_f)} dispatches; bytes per "
           f"counted KiB from {COPY_KERNEL}: FETCH_SIZE {sc_f:.4f} KiB, WRITE_SIZE {sc_w:.4f} KiB; counted "
           f"{tot_ctr / 1e6:.1f} MB per iteration against {tot_model / 1e6:.1f} MB of model bytes", "",
           "| # | kernel | level | model MB | fetched MB | written MB | counted MB | counted / model |",
           "|---|---|---|---|---|---|---|---|"]
    for i, ((k, _), (lev, _, _, nbytes), f, w) in enumerate(zip(s_f, walk, fetch, write)):
        mb = f"{nbytes / 1e6:.1f}" if nbytes else ""
        ratio = f"{(f + w) / nbytes:.2f}" if nbytes else ""
        out.append(f"| {i} | `{k}` | {lev} | {mb} | {f / 1e6:.1f} | {w / 1e6:.1f} | {(f + w) / 1e6:.1f} | {ratio} |")
    return "\n".join(out)


def main(argv):
    if len(argv) < 3:
        print(__doc__)
        return 2
    db, mode = argv[1], argv[2]
    if mode == "stats":
        print(stats(db, int(argv[3]) if len(argv) > 3 else 30))
    elif mode == "levels":
        print(levels(db, argv[3]))
    elif mode == "precisions":
        print(precisions(db))
    elif mode == "traffic":
        print(traffic(db, argv[3], argv[4]))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
