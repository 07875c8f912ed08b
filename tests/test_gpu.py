"""GPU parity tests (MI355X): the HIP path, called through the C ABI, against the
reference fixtures and the oracle.

Bars: integer maps and reproducible-mode (exact-sum) solves are bit-exact; the fast
(tree-sum) mode matches within the stated tolerances; GAMG (damped-Jacobi variant)
in exact-sum mode is bit-exact against the oracle's restatement of that variant.
"""
import numpy as np
import pytest

import voluma_amd as va
from helpers import SMALL_MESHES, load_json, load_small, sha, solve_cases
from oracle import oracle as O
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu
NAMES = list(SMALL_MESHES)


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _system(ctx, name, sym=True, gamg=True):
    g = load_small(name)
    m = SMALL_MESHES[name]()
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights() if gamg else None)
    lm.set_boundary(m.owner[ni:])
    if sym:
        A = va.LduMatrix(lm, g["sym_diag"], g["sym_upper"])
        b = g["sym_b"]
    else:
        A = va.LduMatrix(lm, g["ns_diag"], g["ns_upper"], g["ns_lower"])
        b = g["ns_b"]
    return g, m, lm, A, b


# ---------------------------------------------------------------- Amul / assembly
@pytest.mark.parametrize("name", NAMES)
def test_amul_bit_exact(ctx, name):
    g, m, lm, A, _ = _system(ctx, name)
    assert np.array_equal(va.amul(A, g["amul_x"]), g["amul_y_sym"])
    A2 = va.LduMatrix(lm, g["ns_diag"], g["ns_upper"], g["ns_lower"])
    assert np.array_equal(va.amul(A2, g["amul_x"]), g["amul_y_ns"])


@pytest.mark.parametrize("name", NAMES)
def test_assembly_bit_exact(ctx, name):
    g, m, lm, _, _ = _system(ctx, name)
    geo = m.geometry()
    A = va.LduMatrix(lm)
    bin_ = W.pressure_rhs(m.n_cells, seed=3)
    b, ic, bs = A.assemble_laplacian(geo["magSf"], geo["ndc"], W.bc_fixed_xmax(m), b_in=bin_, return_boundary=True)
    d, u, lo = A.get()
    assert np.array_equal(d, g["sym_diag"]) and np.array_equal(u, g["lap_upper"]) and np.array_equal(lo, u)
    assert np.array_equal(ic, g["lap_ic"]) and np.array_equal(bs, g["lap_bs"])
    assert np.array_equal(b, g["sym_b"])
    D = va.LduMatrix(lm)
    b2, ic2, bs2 = D.assemble_div_upwind(g["div_flux"], W.bc_fixed_xmax(m), return_boundary=True)
    d2, u2, l2 = D.get()
    fd, fb = O.fold(m, g["div_diag"], np.zeros(m.n_cells), g["div_ic"], g["div_bs"])
    assert np.array_equal(u2, g["div_upper"]) and np.array_equal(l2, g["div_lower"])
    assert np.array_equal(ic2, g["div_ic"]) and np.array_equal(bs2, g["div_bs"])
    assert np.array_equal(d2, fd) and np.array_equal(b2, fb)


@pytest.mark.parametrize("name", NAMES)
def test_gamg_hierarchy_bit_exact(ctx, name):
    g, m, lm, _, _ = _system(ctx, name)
    lv = lm.gamg_levels()
    assert [L["n"] for L in lv] == g["gamg_sizes"].tolist()
    nl = len(lv)
    agg = np.concatenate([L["agg"] for L in lv])
    assert np.array_equal(agg, g["gamg_agg"])
    fm = np.concatenate([L["faceMap"] for L in lv[: nl - 1]])
    ff = np.concatenate([L["faceFlip"] for L in lv[: nl - 1]])
    assert np.array_equal(fm, g["gamg_faceMap"][: fm.size])
    assert np.array_equal(ff, g["gamg_faceFlip"][: ff.size])


# ---------------------------------------------------------------- solves vs reference
@pytest.mark.parametrize("name", NAMES)
def test_reproducible_solves_bit_exact(ctx, name):
    """Reference reproducible mode (exact sums + swap) reproduced bit for bit."""
    gsys = {True: _system(ctx, name, True), False: _system(ctx, name, False)}
    g = gsys[True][0]
    checked = 0
    for i, solver, pre, repro, sym in solve_cases(g):
        if not repro:
            continue
        _, _, _, A, b = gsys[sym]
        x = np.zeros(b.size)
        p = va.solve_ldu(A, b, x, va.SolverControls(solver, pre, 1e-10, 0.0, 400, sums=va.SUMS_REPRODUCIBLE))
        ref = g[f"solve{i}_perf"]
        got = np.array([p.initialResidual, p.finalResidual, p.residualSum, p.normFactor, p.iterations, p.converged])
        assert np.array_equal(got, ref), (solver, pre, got, ref)
        assert np.array_equal(x, g[f"solve{i}_x"]), (solver, pre)
        checked += 1
    assert checked == 3


@pytest.mark.parametrize("name", NAMES)
def test_fast_mode_solves_match_reference(ctx, name):
    """Default-mode solves: tree sums instead of sequential sums; same solution within
    1e-8 relative (both at tolerance 1e-10) and iteration counts within ±1 (Krylov) or
    equal (Jacobi, Gauss–Seidel, whose sweeps do not depend on the sums)."""
    gsys = {True: _system(ctx, name, True), False: _system(ctx, name, False)}
    g = gsys[True][0]
    for i, solver, pre, repro, sym in solve_cases(g):
        if repro or pre == "GAMG":
            continue
        _, _, _, A, b = gsys[sym]
        x = np.zeros(b.size)
        p = va.solve_ldu(A, b, x, va.SolverControls(solver, pre, 1e-10, 0.0, 400, sums=va.SUMS_FAST))
        ref = g[f"solve{i}_perf"]
        assert abs(p.iterations - int(ref[4])) <= 1, (solver, pre, p.iterations, ref[4])
        assert p.initialResidual == pytest.approx(ref[0], rel=1e-12)
        xr = g[f"solve{i}_x"]
        if solver in ("GaussSeidel", "symGaussSeidel") and p.iterations == int(ref[4]):
            assert np.array_equal(x, xr), solver  # same sweeps, bit for bit
        elif solver != "Jacobi":
            assert np.linalg.norm(x - xr) <= 1e-8 * np.linalg.norm(xr), (solver, pre)
        else:
            assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("sweeps", [(2, 2), (1, 1), (3, 2)])
def test_gamg_exact_mode_bit_exact_vs_oracle(ctx, name, sweeps):
    g, m, lm, A, b = _system(ctx, name)
    om = O.LduMesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    pre, post = sweeps
    xo, po = om.solve(g["sym_diag"], g["sym_upper"], None, b, preconditioner="GAMG", tolerance=1e-10, max_iter=200,
                      mode=O.MODE_EXACT, smoother=O.SMOOTHER_JACOBI, omega=0.8, n_pre=pre, n_post=post)
    x = np.zeros(b.size)
    p = va.solve_ldu(A, b, x, va.SolverControls("PCG", "GAMG", 1e-10, 0.0, 200, sums=va.SUMS_EXACT,
                                                gamgPreSweeps=pre, gamgPostSweeps=post, gamgOmega=0.8))
    assert p.iterations == po["iterations"] and p.finalResidual == po["finalResidual"]
    assert np.array_equal(x, xo)
    # PBiCGStab with the same preconditioner
    xo2, po2 = om.solve(g["sym_diag"], g["sym_upper"], None, b, solver="PBiCGStab", preconditioner="GAMG",
                        tolerance=1e-10, max_iter=200, mode=O.MODE_EXACT, smoother=O.SMOOTHER_JACOBI, omega=0.8,
                        n_pre=pre, n_post=post)
    x2 = np.zeros(b.size)
    p2 = va.solve_ldu(A, b, x2, va.SolverControls("PBiCGStab", "GAMG", 1e-10, 0.0, 200, sums=va.SUMS_EXACT,
                                                  gamgPreSweeps=pre, gamgPostSweeps=post, gamgOmega=0.8))
    assert p2.iterations == po2["iterations"] and np.array_equal(x2, xo2)


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("sweeps", [(2, 2), (1, 1)])
def test_tail_kernel_bit_exact_vs_oracle(name, sweeps, monkeypatch):
    """The one-workgroup kernel for the small GAMG levels (off by default, enabled by
    VOLUMA_TAIL_MAX_CELLS) gives the same bits as the oracle, FP64 exact mode, and the
    same mixed-precision result as the per-level kernels."""
    monkeypatch.setenv("VOLUMA_TAIL_MAX_CELLS", "512")
    ct = va.Context(0)
    g, m, lm, A, b = _system(ct, name)
    om = O.LduMesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    pre, post = sweeps
    xo, po = om.solve(g["sym_diag"], g["sym_upper"], None, b, preconditioner="GAMG", tolerance=1e-10, max_iter=200,
                      mode=O.MODE_EXACT, smoother=O.SMOOTHER_JACOBI, omega=0.8, n_pre=pre, n_post=post)
    x = np.zeros(b.size)
    p = va.solve_ldu(A, b, x, va.SolverControls("PCG", "GAMG", 1e-10, 0.0, 200, sums=va.SUMS_EXACT,
                                                gamgPreSweeps=pre, gamgPostSweeps=post, gamgOmega=0.8))
    assert p.iterations == po["iterations"] and np.array_equal(x, xo)
    ctl = va.SolverControls("PCG", "GAMG", 1e-10, 0.0, 200, gamgPreSweeps=pre, gamgPostSweeps=post,
                            precision=va.MIXED)
    xm = np.zeros(b.size)
    pm = va.solve_ldu(A, b, xm, ctl)
    monkeypatch.setenv("VOLUMA_TAIL_MAX_CELLS", "0")
    c0 = va.Context(0)
    _, _, _, A0, _ = _system(c0, name)
    xm0 = np.zeros(b.size)
    pm0 = va.solve_ldu(A0, b, xm0, ctl)
    assert pm.iterations == pm0.iterations and np.array_equal(xm, xm0)


@pytest.mark.parametrize("name", NAMES)
def test_gamg_solution_vs_reference_gamg(ctx, name):
    """GPU GAMG-PCG (Jacobi smoother) vs the reference GAMG-PCG (Gauss–Seidel): both at
    normalised residual 1e-13 agree within 1e-10 relative."""
    g, m, lm, A, b = _system(ctx, name)
    om = O.LduMesh(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    xr, pr = om.solve(g["sym_diag"], g["sym_upper"], None, b, preconditioner="GAMG", tolerance=1e-13, max_iter=500)
    x = np.zeros(b.size)
    p = va.solve_ldu(A, b, x, va.SolverControls("PCG", "GAMG", 1e-13, 0.0, 500))
    assert p.converged and pr["converged"]
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_error_codes(ctx):
    g, m, lm, A, b = _system(ctx, "tet5x4x3", sym=False)
    x = np.zeros(b.size)
    with pytest.raises(va.VolumaError) as e:
        va.solve_ldu(A, b, x, va.SolverControls("PCG", "diagonal"))
    assert e.value.code == -4
    for ctl, code in ((va.SolverControls("CG"), -2), (va.SolverControls("PBiCGStab", "ILU0"), -3),
                      (va.SolverControls("GaussSeidel", "none", precision=va.MIXED), -6)):
        with pytest.raises(va.VolumaError) as e:
            va.solve_ldu(A, b, x, ctl)
        assert e.value.code == code
    lm2 = va.LduMesh(ctx, m.n_cells, m.owner[: m.n_internal], m.neighbour, None)
    A2 = va.LduMatrix(lm2, g["sym_diag"], g["sym_upper"])
    with pytest.raises(va.VolumaError) as e:
        va.solve_ldu(A2, g["sym_b"], x, va.SolverControls("PCG", "GAMG"))
    assert e.value.code == -8
    # non-convergence is not an error
    p = va.solve_ldu(A, b, x, va.SolverControls("PBiCGStab", "diagonal", 1e-30, 0.0, 3))
    assert p.iterations == 3 and not p.converged
    # early exit: x already the solution
    p2 = va.solve_ldu(A, b, x, va.SolverControls("PBiCGStab", "diagonal", 1e30, 0.0, 10))
    assert p2.iterations == 0 and p2.converged


def test_device_resident_solve_and_reuse(ctx):
    """Device pointers in and out; the matrix stays resident across several RHS."""
    g, m, lm, A, b = _system(ctx, "box12")
    for seed in (3, 4, 5):
        bb = W.pressure_rhs(m.n_cells, seed)
        db, dx = va.DeviceArray.from_host(ctx, bb), va.DeviceArray.from_host(ctx, np.zeros(m.n_cells))
        p = va.solve_ldu(A, db, dx, va.SolverControls("PCG", "GAMG", 1e-9))
        xh = np.zeros(m.n_cells)
        p2 = va.solve_ldu(A, bb, xh, va.SolverControls("PCG", "GAMG", 1e-9))
        assert np.array_equal(dx.download(), xh) and p.iterations == p2.iterations


# ---------------------------------------------------------------- full-size configurations
@pytest.mark.parametrize("tag,name", [("cfg1", "box64"), ("cfg2", "tet70")])
def test_full_size_reproducible_pcg_bit_exact(ctx, tag, name):
    """Reference reproducible diagonal-PCG at 1e-6 on the full benchmark system: GPU
    solution hash, iterations and residual equal to the reference's."""
    f = load_json(f"{tag}_{name}.json")
    m = W.make_mesh(name)
    lm, A, b = W.pressure_system(ctx, m, seed=1, gamg=True)
    d, u, _ = A.get()
    assert sha(d) == f["diag_sha256"] and sha(u) == f["upper_sha256"] and sha(b) == f["b_sha256"]
    assert [L["n"] for L in lm.gamg_levels()] == f["gamg_sizes"]
    assert [sha(L["agg"]) for L in lm.gamg_levels()] == f["gamg_agg_sha256"]
    x = np.zeros(m.n_cells)
    p = va.solve_ldu(A, b, x, va.SolverControls("PCG", "GAMG", 1e-6, 0.0, 5000, sums=va.SUMS_REPRODUCIBLE))
    r = f["repro_diag_pcg_1e-6"]
    assert p.iterations == r["iterations"] and p.finalResidual == r["finalResidual"]
    assert sha(x) == r["x"]["sha256"]
    # GAMG-PCG to 1e-6 (fast mode) converges; its error against the reference's tight solution
    x2 = np.zeros(m.n_cells)
    p2 = va.solve_ldu(A, b, x2, va.SolverControls("PCG", "GAMG", 1e-6, 0.0, 500))
    assert p2.converged
    idx = np.array(f["gamg_pcg_1e-10"]["x"]["sample_idx"])
    xs = np.array(f["gamg_pcg_1e-10"]["x"]["sample"])
    assert np.abs(x2[idx] - xs).max() <= 2e-3 * np.abs(xs).max()


def test_handle_lifetimes_any_destroy_order():
    """Handles keep their parents alive: closing the context and mesh first leaves a
    working matrix; closing it last frees everything."""
    c = va.Context(0)
    g = load_small("box8x6x5")
    m = SMALL_MESHES["box8x6x5"]()
    lm = va.LduMesh(c, m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights())
    A = va.LduMatrix(lm, g["sym_diag"], g["sym_upper"])
    d = va.DeviceArray.from_host(c, g["amul_x"])
    c.close()
    lm.close()
    assert np.array_equal(va.amul(A, g["amul_x"]), g["amul_y_sym"])
    assert np.array_equal(d.download(), g["amul_x"])
    A.close()
    d.free()


def test_dropin_voluma_program_with_gpu_solver():
    """A Voluma program (reference mesh, fvm::laplacian, FvMatrix::solve) linked with the
    drop-in voluma_amd/integration/LinearSolver_amd.cpp instead of the reference's
    LinearSolver.cpp/GAMG.cpp. Reproducible-mode solves match the reference's own build
    bit for bit; default-mode solves within tolerance."""
    import json
    import os
    import subprocess

    from conftest import ROOT
    exe = os.path.join(ROOT, "oracle", "_ref", "dropin_amd")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/dropin_amd not built (needs the reference sources when building)")
    out = subprocess.run([exe, "24"], check=True, capture_output=True, text=True, timeout=120).stdout
    got = [json.loads(l) for l in out.splitlines() if l.startswith("{")]
    with open(os.path.join(ROOT, "tests", "golden", "dropin_expected.jsonl")) as fh:
        ref = [json.loads(l) for l in fh if l.startswith("{")]
    assert len(got) == len(ref) == 4
    for g, r in zip(got, ref):
        assert (g["solver"], g["preconditioner"], g["reproducible"]) == (r["solver"], r["preconditioner"],
                                                                          r["reproducible"])
        if r["reproducible"]:
            assert g == r
        else:  # GAMG smooths with the reference's Gauss–Seidel: same iteration count ±1
            assert g["final"] < 1e-10 and g["l2sq"] == pytest.approx(r["l2sq"], rel=1e-8)
            assert abs(g["iterations"] - r["iterations"]) <= 1
    # the damped-Jacobi V-cycle is an explicit opt-in; it converges to the same solution
    env = dict(os.environ, VOLUMA_GAMG_SMOOTHER="jacobi")
    out = subprocess.run([exe, "24"], check=True, capture_output=True, text=True, timeout=120, env=env).stdout
    jac = [json.loads(l) for l in out.splitlines() if l.startswith("{")]
    g, r = jac[2], ref[2]
    assert r["preconditioner"] == "GAMG" and g["final"] < 1e-10
    assert g["l2sq"] == pytest.approx(r["l2sq"], rel=1e-8)
    env["VOLUMA_GAMG_SMOOTHER"] = "sor"
    bad = subprocess.run([exe, "24"], capture_output=True, text=True, timeout=120, env=env)
    assert bad.returncode != 0 and "VOLUMA_GAMG_SMOOTHER" in bad.stderr
