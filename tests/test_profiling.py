"""CPU tests of the rocprofv3 trace summaries (voluma_amd/profiling.py) on a synthetic
trace: the per-kernel statistics and the per-level walk of one PCG iteration, with the
separate and the fused residual/restriction launches."""
import json
import sqlite3

from voluma_amd import profiling

CELLS = [1000, 250, 60]
FACES = [2000, 600, 150]


def _db(path, iterations):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, grid_x int, workgroup_x int, start int, duration int)")
    t = 0
    for seq in iterations:
        for name, grid, dur in seq:
            con.execute("insert into kernels values (?, ?, ?, ?, ?)",
                        (f"void vol::{name}(vol::SellArgsT<double>)", grid * 256, 256, t, dur))
            t += dur + 100
    con.commit()
    con.close()


def _bench(path):
    line = {"metric": "m", "config": {"gamg_levels": CELLS, "gamg_level_faces": FACES}}
    with open(path, "w") as fh:
        fh.write("[bench] something\n" + json.dumps(line) + "\n")


AMUL = ("k_sell<0, 0, true, 4, double>", 8, 2000)
SEPARATE = [AMUL, ("k_finalize", 1, 500), ("k_sell<1, 0, false, 4, double>", 8, 1000),
            ("k_restrict<double>", 2, 300), ("k_csrv<1, 0, 8, 4, double>", 4, 400),
            ("k_restrict_coarse64<double>", 1, 900), ("k_csrv<2, 3, 8, 4, double>", 4, 450),
            ("k_sell<2, 3, true, 4, double>", 8, 3000)]
FUSED = [AMUL, ("k_finalize", 1, 500), ("k_sell<1, 0, false, 4, double>", 8, 1000),
         ("k_restrict<double>", 2, 300), ("k_resid_restrict_coarse64_post<4, 8, double>", 1, 1200),
         ("k_sell<2, 3, true, 4, double>", 8, 3000)]


def _rows(table):
    return [[c.strip() for c in line.strip("|").split("|")] for line in table.splitlines()
            if line.startswith("| ") and not line.startswith("| #") and not line.startswith("| kernel")]


def test_stats_totals(tmp_path):
    db = str(tmp_path / "run.db")
    _db(db, [SEPARATE] * 3)
    out = profiling.stats(db)
    assert out.startswith(f"busy {3 * sum(d for _, _, d in SEPARATE) / 1e6:.3f} ms, {3 * len(SEPARATE)} dispatches")
    rows = _rows(out)
    top = rows[0]
    assert top[0] == "`k_sell<2, 3, true, 4, double>`" and top[2] == "3" and float(top[4]) == 3.0


def test_levels_separate_launches(tmp_path):
    db, bench = str(tmp_path / "run.db"), str(tmp_path / "bench.log")
    _db(db, [SEPARATE] * 4)
    _bench(bench)
    out = profiling.levels(db, bench)
    assert out.startswith(f"3 of 3 iterations with the common sequence of {len(SEPARATE)} dispatches")
    lev = {r[1]: r[3] for r in _rows(out)}
    assert lev["`k_restrict<double>`"] == "0→1"
    assert lev["`k_csrv<1, 0, 8, 4, double>`"] == "1"
    assert lev["`k_restrict_coarse64<double>`"] == "1→2"
    assert lev["`k_csrv<2, 3, 8, 4, double>`"] == "1"   # the post-sweep one level up
    assert lev["`k_sell<2, 3, true, 4, double>`"] == "0"


def test_levels_fused_launches(tmp_path):
    db, bench = str(tmp_path / "run.db"), str(tmp_path / "bench.log")
    _db(db, [FUSED] * 3)
    _bench(bench)
    rows = {r[1]: r for r in _rows(profiling.levels(db, bench))}
    fused = rows["`k_resid_restrict_coarse64_post<4, 8, double>`"]
    assert fused[3] == "1→2" and fused[4] == str(CELLS[1]) and fused[5] == str(2 * FACES[1])
    # its own post-sweep is inside: the next sweep is level 0's
    assert rows["`k_sell<2, 3, true, 4, double>`"][3] == "0"
    # bytes of the fused launch: level-1 rows and entries, lists, coarse vectors, LU, the sweep
    n, nc, e = CELLS[1], CELLS[2], 2 * FACES[1]
    mb = (8 * n + 12 * e + 16 * n + 16 * n + 4 * (nc + 1) + 24 * nc + 8 * nc * nc + 16 * n + 4 * n + 8 * n) / 1e6
    assert fused[7] == f"{mb:.1f}"


def test_precisions_pairs_dispatches(tmp_path):
    db = str(tmp_path / "run.db")
    f32 = [(k.replace("double", "float"), g, d // 2) for k, g, d in FUSED]
    _db(db, [SEPARATE] * 2 + [FUSED] * 4 + [f32] * 3)
    out = profiling.precisions(db)
    # FP64 iterations run between FP64 A·p launches: 5, three of them the fused sequence;
    # FP32: 2
    assert out.startswith(f"FP64: 3 of 5 iterations with the common sequence of {len(FUSED)} dispatches")
    assert "FP32 inner: 2 of 2 iterations" in out
    rows = _rows(out)
    assert len(rows) == len(FUSED)
    assert all(r[6] == "2.00" for r in rows)


def test_packed_gathers_count_as_level0(tmp_path):
    db, bench = str(tmp_path / "run.db"), str(tmp_path / "bench.log")
    packed = [(k.replace("k_sell<", "k_sell_packed<"), g, d) for k, g, d in FUSED]
    _db(db, [packed] * 3)
    _bench(bench)
    rows = {r[1]: r for r in _rows(profiling.levels(db, bench))}
    assert rows["`k_sell_packed<0, 0, true, 4, double>`"][3] == "0"
    assert rows["`k_sell_packed<1, 0, false, 4, double>`"][3] == "0"
    assert rows["`k_sell_packed<2, 3, true, 4, double>`"][3] == "0"


def test_packed_coarse_rows_keep_their_level(tmp_path):
    db, bench = str(tmp_path / "run.db"), str(tmp_path / "bench.log")
    seq = [AMUL, ("k_finalize", 1, 500), ("k_sell_pk<1, 0, double>", 8, 1000), ("k_restrict<double>", 2, 300),
           ("k_resid_restrict_coarse64_post<4, 8, double>", 1, 1200), ("k_sell_pk<2, 0, double>", 8, 3000)]
    _db(db, [seq] * 3)
    _bench(bench)
    rows = _rows(profiling.levels(db, bench))
    assert [r[3] for r in rows] == ["0", "", "0", "0→1", "1→2", "0"]
    # level-0 residual: diag, offsets at 2 bytes, coefficients; r, z read and the residual written
    n, e = CELLS[0], 2 * FACES[0]
    assert rows[2][7] == f"{(8 * n + 10 * e + 3 * 8 * n) / 1e6:.1f}"


def _pmc_db(path, iterations, counter, per_dispatch, copy_kib):
    """A counter run: the kernels view with dispatch ids, and counters_collection."""
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, grid_x int, workgroup_x int, start int, duration int, "
                "dispatch_id int)")
    con.execute("create table counters_collection (dispatch_id int, counter_name text, value real)")
    t, did = 0, 0
    seq = [("k_vec_fast<13, 0, double>", 4096, 200000, copy_kib)] + [
        (name, grid, dur, per_dispatch[name]) for it in iterations for name, grid, dur in it]
    for name, grid, dur, val in seq:
        did += 1
        con.execute("insert into kernels values (?, ?, ?, ?, ?, ?)",
                    (f"void vol::{name}(vol::SellArgsT<double>)", grid * 256, 256, t, dur, did))
        con.execute("insert into counters_collection values (?, ?, ?)", (did, counter, val))
        t += dur + 100
    con.commit()
    con.close()


def test_traffic_scales_counters_by_the_copy_kernel(tmp_path):
    fdb, wdb, bench = str(tmp_path / "f.db"), str(tmp_path / "w.db"), str(tmp_path / "bench.log")
    names = {k: 0.0 for k, _, _ in FUSED}
    fetch = dict(names, **{AMUL[0]: 1000.0})
    write = dict(names, **{AMUL[0]: 250.0})
    # the copy's FETCH_SIZE counts half its 1 GiB (factor 2), WRITE_SIZE all of it
    _pmc_db(fdb, [FUSED] * 3, "FETCH_SIZE", fetch, 524288.0)
    _pmc_db(wdb, [FUSED] * 3, "WRITE_SIZE", write, 1048576.0)
    _bench(bench)
    out = profiling.traffic(fdb, wdb, bench)
    assert "FETCH_SIZE 2.0000 KiB, WRITE_SIZE 1.0000 KiB" in out
    row = _rows(out)[0]
    assert row[1] == f"`{AMUL[0]}`" and row[2] == "0"
    assert row[4] == f"{1000 * 2 * 1024 / 1e6:.1f}" and row[5] == f"{250 * 1024 / 1e6:.1f}"
