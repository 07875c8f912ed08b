"""Write the golden fixtures under tests/golden/ from the reference build.

Runs the reference (Voluma compiled serial by oracle/Makefile into
oracle/_ref/libvolref.so, from the sources in /root/reference) on meshes made by
voluma_amd.mesh, and stores inputs and outputs as data:

* small_<mesh>.npz — full arrays for small meshes: geometry, laplacian / upwind-div
  coefficients, GAMG hierarchy maps, solutions and SolverPerformance of every
  solver/preconditioner pair, residual histories, one GAMG V-cycle, an Amul;
* cfg1_box64.json, cfg2_tet70.json — the two full-size pressure systems: input and
  hierarchy hashes, level sizes, iteration counts, residuals, solution hashes,
  norms and samples for reproducible diagonal-PCG and default GAMG-PCG;
* cfg1_box64_tight.json, cfg2_tet70_tight.json (--tight) — the benchmark's pressure
  systems (non-orthogonal correction included): the reference's GAMG-PCG to 1e-13
  (hash, norm, 4096 samples) and the oracle's exact-sum Jacobi-GAMG-PCG hash;
* exactsum.json — ExactSum outputs for seeded random vectors;
* dropin_expected.jsonl — output of the Voluma drop-in check program with the reference solver.

Usage (in the container that has the reference sources):
    make -C oracle all && python tests/golden/make_golden.py [--small-only]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_ROOT = os.environ.get("VOLUMA_REF", "/root/reference")  # the reference tree (mesh data files)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import ref  # noqa: E402
from voluma_amd import mesh as M  # noqa: E402
from voluma_amd import workloads as W  # noqa: E402

SMALL = {
    "box8x6x5": lambda: M.box(8, 6, 5),
    "tet5x4x3": lambda: M.tet(5, 4, 3, 0.2, 7),
    "box12": lambda: M.box(12),
}

# (solver, preconditioner, reproducible, symmetric system?)
CASES = [
    ("PCG", "diagonal", False, True), ("PCG", "diagonal", True, True), ("PCG", "none", False, True),
    ("PCG", "GAMG", False, True), ("PCG", "DIC", False, True),
    ("PBiCGStab", "diagonal", False, False), ("PBiCGStab", "diagonal", True, False),
    ("PBiCGStab", "none", False, False), ("PBiCGStab", "DILU", False, False), ("PBiCGStab", "GAMG", False, True),
    ("Jacobi", "none", False, True), ("GaussSeidel", "none", False, True), ("symGaussSeidel", "none", False, True),
    ("GaussSeidel", "none", True, True),
]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def mesh_hash(m) -> str:
    h = hashlib.sha256()
    for a in (m.points, m.face_off, m.face_pts, m.owner, m.neighbour, m.patch_start, m.patch_size):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


MOM_DT, MOM_NU, MOM_BC_VALUE = 0.05, 0.01, 0.7


def flux_field(m, seed):
    """A face flux for the nonsymmetric (momentum-like) test system: seeded values
    in [−1, 1)·|S_f| on every face."""
    u = 2.0 * W.splitmix_uniform(m.n_faces, seed) - 1.0
    return u * m.geometry()["magSf"]


def momentum_like(rm, m, seed=11, nu=1e-3, dt=0.05):
    """diag/upper/lower of a ddt + div(upwind) − ν·laplacian style LDU system, composed
    from the reference's own fvm::div and fvm::laplacian coefficients."""
    g = m.geometry()
    d = rm.div_upwind(flux_field(m, seed), [W.XMAX_PATCH])
    lap = rm.laplacian([W.XMAX_PATCH])
    ni, own = m.n_internal, m.owner
    diag = d["diag"] + g["V"] / dt - nu * lap["diag"]
    np.add.at(diag, own[ni:], d["ic"] - nu * lap["ic"])  # boundary fold
    upper = d["upper"] - nu * lap["upper"]
    lower = d["lower"] - nu * lap["upper"]
    return diag, upper, lower


def small_case(name, make):
    m = make()
    rm = ref.RefMesh(mesh=m)
    g = rm.geometry()
    lap = rm.laplacian([W.XMAX_PATCH])
    n, ni = m.n_cells, m.n_internal
    own = m.owner
    dsym = lap["diag"].copy()
    np.add.at(dsym, own[ni:], lap["ic"])
    upper = lap["upper"]
    b = W.pressure_rhs(n, seed=3)
    np.add.at(b, own[ni:], lap["bs"])
    dns, uns, lns = momentum_like(rm, m)
    bns = W.pressure_rhs(n, seed=5)
    out = {"mesh_hash": np.frombuffer(bytes.fromhex(mesh_hash(m)), np.uint8)}
    for k in ("V", "magSf", "ndc", "w"):
        out["geom_" + k] = g[k]
    out["geom_C"] = g["C"]
    out["geom_corr"] = g["corr"]
    for k in ("diag", "upper", "ic", "bs"):
        out["lap_" + k] = lap[k]
    F = flux_field(m, 11)
    dv = rm.div_upwind(F, [W.XMAX_PATCH])
    for k in ("diag", "upper", "lower", "ic", "bs"):
        out["div_" + k] = dv[k]
    out["div_flux"] = F
    # momentum-matrix component: ddt(Euler) + div(F, upwind) − laplacian(nu) with the
    # non-orthogonal correction, fixedValue 0.7 on xMax (IncompressibleFlow.cpp:183-185)
    psi0 = 2.0 * W.splitmix_uniform(n, 31) - 1.0
    vals = np.zeros(len(m.patch_start))
    vals[W.XMAX_PATCH] = MOM_BC_VALUE
    tr = rm.transport(MOM_DT, MOM_NU, psi0, F, [W.XMAX_PATCH], vals, True)
    out["mom_psi0"] = psi0
    for k, v in tr.items():
        out["mom_" + k] = v
    out["sym_diag"], out["sym_upper"], out["sym_b"] = dsym, upper, b
    out["ns_diag"], out["ns_upper"], out["ns_lower"], out["ns_b"] = dns, uns, lns, bns
    # hierarchy
    lv = rm.gamg_hierarchy()
    out["gamg_sizes"] = np.array([L["n"] for L in lv], np.int64)
    out["gamg_nfaces"] = np.array([len(L["l"]) for L in lv], np.int64)
    for key in ("agg", "faceMap", "faceFlip", "l", "u", "weight"):
        out["gamg_" + key] = np.concatenate([L[key] for L in lv])
    # amul and one V-cycle
    x = 2.0 * W.splitmix_uniform(n, 21) - 1.0
    out["amul_x"] = x
    out["amul_y_sym"] = rm.amul(dsym, upper, None, x)
    out["amul_y_ns"] = rm.amul(dns, uns, lns, x)
    out["gamg_apply_r"] = x
    out["gamg_apply_w"] = rm.gamg_apply(dsym, upper, None, x)
    # solves
    meta = []
    for i, (solver, pre, repro, sym) in enumerate(CASES):
        d, u, lo, bb = (dsym, upper, None, b) if sym else (dns, uns, lns, bns)
        kw = dict(solver=solver, preconditioner=pre, tolerance=1e-10, rel_tol=0.0, max_iter=400, reproducible=repro)
        xs, p = rm.solve(d, u, lo, bb, np.zeros(n), **kw)
        out[f"solve{i}_x"] = xs
        out[f"solve{i}_perf"] = np.array([p["initialResidual"], p["finalResidual"], p["residualSum"],
                                          p["normFactor"], p["iterations"], p["converged"]], np.float64)
        meta.append([solver, pre, int(repro), int(sym)])
    out["solve_cases"] = np.array(["|".join(map(str, c)) for c in meta])
    # residual history of two Krylov solves by re-running with maxIter = k
    for tag, (solver, pre, repro) in {"histA": ("PCG", "diagonal", True), "histB": ("PCG", "GAMG", False)}.items():
        _, p = rm.solve(dsym, upper, None, b, np.zeros(n), solver=solver, preconditioner=pre, tolerance=1e-10,
                        max_iter=400, reproducible=repro)
        hist = [p["initialResidual"]]
        for k in range(1, min(p["iterations"], 60) + 1):
            _, pk = rm.solve(dsym, upper, None, b, np.zeros(n), solver=solver, preconditioner=pre,
                             tolerance=1e-10, max_iter=k, reproducible=repro)
            hist.append(pk["finalResidual"])
        out[tag] = np.array(hist)
    np.savez_compressed(os.path.join(HERE, f"small_{name}.npz"), **out)
    print(f"small_{name}: {n} cells, levels {out['gamg_sizes'].tolist()}", flush=True)


POST_ALPHA, POST_DT_VEC, POST_REF_VALUE = 0.7, 1.0, 0.25


def bc_fixed_of(m, patches):
    """Per boundary face: 1 on the faces of the given patches (fixedValue), 0 elsewhere."""
    ni = m.n_internal
    out = np.zeros(m.n_faces - ni, np.int32)
    for p in patches:
        out[m.patch_start[p] - ni: m.patch_start[p] - ni + m.patch_size[p]] = 1
    return out


def post_case(name, make):
    """FvMatrix::relax, A(), H(), flux() and setReference (FvMatrix.h:111-141,179-222)
    from the reference: the momentum matrix of a scalar field (fixedValue on xMax) and of
    a vector field (fixedValue on every patch but xMax, dt = 1 so that the vector form of
    cmptAv changes results), and a pressure matrix (laplacian alone, fixedValue on xMax)."""
    m = make()
    rm = ref.RefMesh(mesh=m)
    n, ni = m.n_cells, m.n_internal
    inp = W.momentum_inputs(m, 2)
    out = {"mesh_hash": np.frombuffer(bytes.fromhex(mesh_hash(m)), np.uint8)}
    xmax = [W.XMAX_PATCH]
    other = [p for p in range(len(m.patch_start)) if p != W.XMAX_PATCH]
    fv_x = np.where(bc_fixed_of(m, xmax) != 0, MOM_BC_VALUE, 0.0)
    cases = {
        "ms": dict(dt=MOM_DT, nu=MOM_NU, psi0=2.0 * W.splitmix_uniform(n, 31) - 1.0, F=flux_field(m, 11),
                   patches=xmax, fv=fv_x, ref_cell=n // 3, pressure=False),
        "mv": dict(dt=POST_DT_VEC, nu=inp["nu"], psi0=np.ascontiguousarray(inp["U0"].T), F=inp["phi"],
                   patches=other, fv=np.ascontiguousarray(inp["Ub"].T), ref_cell=n // 2, pressure=False),
        "p": dict(dt=1.0, nu=1.0, psi0=W.pressure_rhs(n, seed=41), F=np.zeros(m.n_faces), patches=xmax,
                  fv=np.zeros(m.n_faces - ni), ref_cell=0, pressure=True),
    }
    for tag, c in cases.items():
        r = rm.transport_post(c["dt"], c["nu"], c["psi0"], c["F"], c["patches"], c["fv"], POST_ALPHA, c["ref_cell"],
                              POST_REF_VALUE, True, pressure=c["pressure"])
        out[f"{tag}_in_psi0"] = c["psi0"]
        out[f"{tag}_in_F"] = c["F"]
        out[f"{tag}_in_fv"] = c["fv"]
        out[f"{tag}_in_bcfixed"] = bc_fixed_of(m, c["patches"])
        out[f"{tag}_in_scalars"] = np.array([c["dt"], c["nu"], POST_ALPHA, c["ref_cell"], POST_REF_VALUE,
                                             float(c["pressure"])])
        for k, v in r.items():
            out[f"{tag}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, f"post_{name}.npz"), **out)
    print(f"post_{name}: {n} cells", flush=True)


FVC_FIXED = (0,)  # fixedValue patches of the corrector-operator fixtures, with xMax
FVC_MESHES = {"box8x6x5": SMALL["box8x6x5"], "tet5x4x3": SMALL["tet5x4x3"],
              "tet6x5x4": lambda: M.tet(6, 5, 4, 0.24, 11)}


def fvc_case(name, make):
    """The reference's explicit corrector operators (Fvc.h: gradLS of a scalar and a
    vector, interpolate plain and skew-corrected, grad Gauss linear / least squares /
    Gauss skew, div, surfaceIntegrate, the phiHbyA face flux) on seeded fields, and the
    mesh's skew vectors."""
    m = make()
    rm = ref.RefMesh(mesh=m)
    n, nf, ni = m.n_cells, m.n_faces, m.n_internal
    nb = nf - ni
    fixed = sorted(set(FVC_FIXED) | {W.XMAX_PATCH})
    inp = {"p": 2.0 * W.splitmix_uniform(n, 61) - 1.0,
           "U": (2.0 * W.splitmix_uniform(3 * n, 62) - 1.0).reshape(n, 3),
           "pB": 2.0 * W.splitmix_uniform(nb, 63) - 1.0,
           "UB": (2.0 * W.splitmix_uniform(3 * nb, 64) - 1.0).reshape(nb, 3),
           "F": (2.0 * W.splitmix_uniform(nf, 65) - 1.0) * m.geometry()["magSf"],
           "fixed": np.array(fixed, np.int32)}
    out = {"mesh_hash": np.frombuffer(bytes.fromhex(mesh_hash(m)), np.uint8)}
    for k, v in inp.items():
        out["in_" + k] = v
    for k, v in rm.fvc(inp["p"], inp["U"], fixed, inp["pB"], inp["UB"], inp["F"]).items():
        out[k] = v
    sk, flag = rm.skew()
    out["skew"], out["skewed"] = sk, np.array([int(flag)], np.int32)
    np.savez_compressed(os.path.join(HERE, f"fvc_{name}.npz"), **out)
    print(f"fvc_{name}: {n} cells, skewed {flag}", flush=True)


POLYMESH = {"cavity3D_4": "meshes/cavity3D_4/polyMesh",           # 3-D tets, ASCII faceList
            "cavity_binary": "meshes/import/cavity_binary_polyMesh"}  # 2-D hexes, binary faceCompactList
POLYMESH_FILES = ("points", "faces", "owner", "neighbour", "boundary")


def polymesh_case():
    """Two of the reference's own polyMesh test meshes, copied as data (only the five
    files the reader needs), with what the reference's readPolyMesh reads from them and,
    for the 3-D mesh, the geometry of the Mesh it builds."""
    import shutil
    out = {}
    for name, rel in POLYMESH.items():
        src = os.path.join(REF_ROOT, rel)
        dst = os.path.join(HERE, "polymesh", name)
        os.makedirs(dst, exist_ok=True)
        for f in POLYMESH_FILES:
            shutil.copyfile(os.path.join(src, f), os.path.join(dst, f))
        raw, rm = ref.read_polymesh(dst)
        e = {k: sha(np.ascontiguousarray(raw[k], dtype=np.int64 if k != "points" else np.float64))
             for k in ("points", "face_off", "face_pts", "owner", "neighbour")}
        e.update({"n_cells": raw["n_cells"], "patch_names": raw["patch_names"], "patch_types": raw["patch_types"],
                  "patch_start": [int(v) for v in raw["patch_start"]],
                  "patch_size": [int(v) for v in raw["patch_size"]]})
        if name == "cavity3D_4":
            g = rm.geometry()
            sk, flag = rm.skew()
            e["geometry"] = {k: sha(v) for k, v in g.items()}
            e["geometry"]["skew"] = sha(sk)
            e["skewed"] = flag
        out[name] = e
        print(f"polymesh {name}: {raw['n_cells']} cells", flush=True)
    with open(os.path.join(HERE, "polymesh_expected.json"), "w") as fh:
        json.dump(out, fh, indent=1)


GMSH_FILES = {"hybrid3D_v41": "meshes/import/hybrid3D_v41.msh",  # tets, hexes, pyramids (Gmsh 4.1)
              "tet10_v41": "meshes/import/tet10_v41.msh"}        # second-order tets (Gmsh 4.1)


def write_gmsh22(m, path):
    """A Gmsh 2.2 ASCII file of a tet mesh: nodes 1..n, the boundary triangles of each
    patch (physical surface p+1) and the tets (physical volume 100), each tet's nodes
    in first-seen order over its faces."""
    ni = m.n_internal
    cell_nodes = [[] for _ in range(m.n_cells)]
    for f in range(m.n_faces):
        pts = m.face_pts[m.face_off[f]:m.face_off[f + 1]]
        for c in ((m.owner[f], m.neighbour[f]) if f < ni else (m.owner[f],)):
            for q in pts:
                if q not in cell_nodes[c]:
                    cell_nodes[c].append(int(q))
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", str(len(m.patch_start) + 1)]
    for p in range(len(m.patch_start)):
        lines.append(f'2 {p + 1} "side{p}{"Wall" if p % 2 else ""}"')
    lines += ['3 100 "fluid"', "$EndPhysicalNames", "$Nodes", str(len(m.points))]
    lines += [f"{i + 1} {x!r} {y!r} {z!r}" for i, (x, y, z) in enumerate(m.points.tolist())]
    lines += ["$EndNodes", "$Elements"]
    els = []
    for p in range(len(m.patch_start)):
        for f in range(int(m.patch_start[p]), int(m.patch_start[p] + m.patch_size[p])):
            pts = m.face_pts[m.face_off[f]:m.face_off[f + 1]]
            els.append(f"2 2 {p + 1} {p + 1} " + " ".join(str(int(q) + 1) for q in pts))
    for c in range(m.n_cells):
        els.append("4 2 100 1 " + " ".join(str(q + 1) for q in cell_nodes[c]))
    lines.append(str(len(els)))
    lines += [f"{k + 1} {e}" for k, e in enumerate(els)]
    lines += ["$EndElements", ""]
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def gmsh_case():
    """Two of the reference's own Gmsh test files (copied as data) and a Gmsh 2.2 file of
    the jittered tet5x4x3 mesh, with what the reference's readGmsh reads from each and
    the geometry of the Mesh it builds."""
    import shutil
    dst_dir = os.path.join(HERE, "gmsh")
    os.makedirs(dst_dir, exist_ok=True)
    files = {}
    for name, rel in GMSH_FILES.items():
        files[name] = os.path.join(dst_dir, name + ".msh")
        shutil.copyfile(os.path.join(REF_ROOT, rel), files[name])
    files["tet5x4x3_v22"] = os.path.join(dst_dir, "tet5x4x3_v22.msh")
    write_gmsh22(SMALL["tet5x4x3"](), files["tet5x4x3_v22"])
    out = {}
    for name, path in files.items():
        raw, rm = ref.read_gmsh(path)
        e = {k: sha(np.ascontiguousarray(raw[k], dtype=np.int64 if k != "points" else np.float64))
             for k in ("points", "face_off", "face_pts", "owner", "neighbour")}
        e.update({"n_cells": raw["n_cells"], "patch_names": raw["patch_names"], "patch_types": raw["patch_types"],
                  "patch_start": [int(v) for v in raw["patch_start"]],
                  "patch_size": [int(v) for v in raw["patch_size"]]})
        g = rm.geometry()
        sk, flag = rm.skew()
        e["geometry"] = {k: sha(v) for k, v in g.items()}
        e["geometry"]["skew"] = sha(sk)
        e["skewed"] = flag
        out[name] = e
        print(f"gmsh {name}: {raw['n_cells']} cells, patches {raw['patch_names']}", flush=True)
    with open(os.path.join(HERE, "gmsh_expected.json"), "w") as fh:
        json.dump(out, fh, indent=1)


# generateBox cases with cyclic pairs: (nx, ny, nz), periodic directions, 2-D, jittered
PERIODIC_CASES = {
    "box6x5x4_px": ((6, 5, 4), (True, False, False), False, False),
    "box6x5x4_pxyz": ((6, 5, 4), (True, True, True), False, False),
    "box6x5x1_py_2d": ((6, 5, 1), (False, True, False), True, False),
    "channel_periodic": ((8, 5, 4), (True, False, False), False, True),
}


def periodic_case():
    """The reference's generateBox with cyclic patch pairs (and one box with its interior
    points moved, helpers.periodic_channel): the raw mesh, then the Mesh it builds (cyclic
    pairs merged into internal faces, Mesh.cpp:246-300): owner, neighbour and geometry of
    its faces, skew vectors, layout, patches, and a reconstruct of a seeded face flux."""
    from helpers import periodic_channel
    out = {}
    for case, (n, per, two_d, jitter) in PERIODIC_CASES.items():
        pts = periodic_channel().points if jitter else None
        raw, rm = ref.generate_box(*n, per, two_d, points=pts)
        e = {k: sha(np.ascontiguousarray(raw[k], dtype=np.int64 if k != "points" else np.float64))
             for k in ("points", "face_off", "face_pts", "owner", "neighbour")}
        e.update({"n_cells": raw["n_cells"], "patch_names": raw["patch_names"], "patch_types": raw["patch_types"]})
        e["owner_c"], e["neighbour_c"] = sha(rm.owner), sha(rm.neighbour)
        e["n_internal_c"] = rm.n_internal
        g = rm.geometry()
        sk, flag = rm.skew()
        e["geometry"] = {k: sha(v) for k, v in g.items()}
        e["geometry"]["skew"] = sha(sk)
        e["skewed"] = flag
        two, empty_dir, solved, names = rm.layout()
        e["mesh"] = {"two_d": two, "empty_dir": empty_dir, "solved_components": solved, "patch_names": names,
                     "patch_start": [int(v) for v in rm.patch_start], "patch_size": [int(v) for v in rm.patch_size]}
        ssf = np.random.default_rng(11).uniform(-1, 1, rm.n_faces)
        e["reconstruct"] = sha(rm.reconstruct(ssf))
        out[case] = e
        print(f"periodic {case}: {raw['n_cells']} cells, {rm.n_internal} internal faces, patches {names}",
              flush=True)
    with open(os.path.join(HERE, "periodic_expected.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def piso_case(name):
    """Time steps of the reference's IncompressibleFlow (vr_piso_run) for one PISO case
    of helpers.PISO_CASES: U, p and φ after the steps."""
    from helpers import PISO_CASES, piso_inputs
    mk, opt, steps, dt, repro = PISO_CASES[name]
    m = mk()
    if name.startswith("channel_periodic"):
        _, rm = ref.generate_box(8, 5, 4, (True, False, False), points=m.points)
    else:
        rm = ref.RefMesh(m)
    U0, p0, phi0 = piso_inputs(m)
    kw = {k: v for k, v in opt.items() if k not in ("nu", "u_bc", "p_bc")}
    if "ddt_phi_coeff" in kw:
        kw["ddt_phi_coeff"] = kw["ddt_phi_coeff"]
    ref.set_reproducible(repro)
    try:
        U, p, phi = rm.piso_run(steps, opt["nu"], dt, U0, p0, phi0, opt["u_bc"], opt["p_bc"], **kw)
    finally:
        ref.set_reproducible(False)
    np.savez(os.path.join(HERE, f"piso_{name}.npz"), U=U, p=p, phi=phi)
    print(f"piso {name}: |U| {np.abs(U).max():.4g} |p| {np.abs(p).max():.4g} |phi| {np.abs(phi).max():.4g}",
          flush=True)


def summary(x):
    x = np.asarray(x)
    idx = np.linspace(0, x.size - 1, 16).astype(np.int64)
    return {"sha256": sha(x), "l2": float(np.sqrt(np.dot(x, x))), "sample_idx": idx.tolist(),
            "sample": [float(v) for v in x[idx]]}


def full_case(tag, name):
    t0 = time.time()
    m = W.make_mesh(name)
    rm = ref.RefMesh(mesh=m)
    t1 = time.time()
    n, ni = m.n_cells, m.n_internal
    rg = rm.geometry()
    mg = m.geometry()
    geom_equal = all(np.array_equal(rg[k], mg[k]) for k in ("magSf", "ndc", "w", "V", "C"))
    lap = rm.laplacian([W.XMAX_PATCH])
    d = lap["diag"].copy()
    np.add.at(d, m.owner[ni:], lap["ic"])
    b = W.pressure_rhs(n, seed=1)
    np.add.at(b, m.owner[ni:], lap["bs"])
    lv = rm.gamg_hierarchy()
    res = {
        "mesh": name, "n_cells": n, "n_internal_faces": ni, "n_faces": m.n_faces, "mesh_sha256": mesh_hash(m),
        "geometry_bit_equal_to_meshgen": bool(geom_equal),
        "weights_sha256": sha(m.agglomeration_weights()),
        "diag_sha256": sha(d), "upper_sha256": sha(lap["upper"]), "b_sha256": sha(b),
        "gamg_sizes": [int(L["n"]) for L in lv],
        "gamg_agg_sha256": [sha(L["agg"]) for L in lv],
        "gamg_faceMap_sha256": [sha(L["faceMap"]) for L in lv],
        "build_s": t1 - t0,
    }
    for key, (pre, repro, tol) in {"repro_diag_pcg_1e-6": ("diagonal", True, 1e-6),
                                   "gamg_pcg_1e-6": ("GAMG", False, 1e-6),
                                   "gamg_pcg_1e-10": ("GAMG", False, 1e-10)}.items():
        t = time.time()
        x, p = rm.solve(d, lap["upper"], None, b, np.zeros(n), solver="PCG", preconditioner=pre, tolerance=tol,
                        max_iter=5000, reproducible=repro)
        res[key] = {"iterations": p["iterations"], "initialResidual": p["initialResidual"],
                    "finalResidual": p["finalResidual"], "normFactor": p["normFactor"], "converged": p["converged"],
                    "solve_s": time.time() - t, "x": summary(x)}
        print(f"  {tag} {key}: {p['iterations']} iters, final {p['finalResidual']:.3e}, "
              f"{time.time() - t:.1f} s", flush=True)
    with open(os.path.join(HERE, f"{tag}_{name}.json"), "w") as fh:
        json.dump(res, fh, indent=1)
    print(f"{tag}_{name}: {n} cells, levels {res['gamg_sizes']}, geometry equal {geom_equal}", flush=True)


def dense_summary(x, k=4096):
    """summary() plus k evenly spaced entries, for relative-error checks at full size."""
    x = np.asarray(x)
    out = summary(x)
    idx = np.linspace(0, x.size - 1, k).astype(np.int64)
    out["dense_idx_step"] = [int(idx[0]), int(idx[-1]), k]
    out["dense"] = [float(v) for v in x[idx]]
    return out


# Jacobi V-cycle of the benchmark's pressure solve: 1 pre- and 1 post-sweep, ω = 0.9
BENCH_VCYCLE = {"n_pre": 1, "n_post": 1, "omega": 0.9}


def tight_case(tag, name):
    """Full-size pressure system as the benchmark builds it (W.pressure_system with
    nonorth=True: fvm::laplacian(1, p) with its non-orthogonal correction on p⁰ =
    W.pressure_psi0, Gauss-linear gradient, fixedValue 0 on xMax), then
    * the reference's GAMG-PCG (Gauss–Seidel V-cycle, plain sums) to normalised residual
      1e-13 — the tight solution the GPU's solvers are compared with;
    * the oracle's GAMG-PCG at the same tolerance (checked bit-equal to the reference);
    * the oracle's exact-sum GAMG-PCG with the benchmark's Jacobi V-cycle to 1e-6 (its
      bits are what the GPU's exact-sum mode must reproduce)."""
    from oracle import oracle as O
    m = W.make_mesh(name)
    rm = ref.RefMesh(mesh=m)
    n, ni = m.n_cells, m.n_internal
    g = m.geometry()
    psi = W.pressure_psi0(g["C"])
    lap = rm.laplacian([W.XMAX_PATCH], psi=psi, non_orth=True, grad_scheme=0)
    d = lap["diag"].copy()
    np.add.at(d, m.owner[ni:], lap["ic"])
    bf = lap["source"].copy()
    np.add.at(bf, m.owner[ni:], lap["bs"])
    b = W.pressure_rhs(n, seed=1) + bf
    res = {"mesh": name, "n_cells": n, "n_internal_faces": ni, "mesh_sha256": mesh_hash(m),
           "psi0_sha256": sha(psi), "diag_sha256": sha(d), "upper_sha256": sha(lap["upper"]), "b_sha256": sha(b),
           "ffc_sha256": sha(lap["ffc"]), "bench_vcycle": BENCH_VCYCLE}
    t = time.time()
    x, p = rm.solve(d, lap["upper"], None, b, np.zeros(n), solver="PCG", preconditioner="GAMG", tolerance=1e-13,
                    max_iter=5000)
    res["ref_gamg_pcg_1e-13"] = {"iterations": p["iterations"], "initialResidual": p["initialResidual"],
                                 "finalResidual": p["finalResidual"], "normFactor": p["normFactor"],
                                 "converged": p["converged"], "solve_s": time.time() - t, "x": dense_summary(x)}
    print(f"  {tag} ref GS-GAMG-PCG 1e-13: {p['iterations']} iters, {time.time() - t:.1f} s", flush=True)
    om = O.LduMesh(n, m.owner[:ni], m.neighbour, m.agglomeration_weights())
    t = time.time()
    xo, po = om.solve(d, lap["upper"], None, b, preconditioner="GAMG", tolerance=1e-13, max_iter=5000)
    res["oracle_gamg_pcg_1e-13_equal_to_reference"] = bool(np.array_equal(xo, x) and
                                                           po["iterations"] == p["iterations"])
    print(f"  {tag} oracle GS-GAMG-PCG 1e-13: equal {res['oracle_gamg_pcg_1e-13_equal_to_reference']}, "
          f"{time.time() - t:.1f} s", flush=True)
    t = time.time()
    xe, pe = om.solve(d, lap["upper"], None, b, preconditioner="GAMG", tolerance=1e-6, max_iter=5000,
                      mode=O.MODE_EXACT, smoother=O.SMOOTHER_JACOBI, **BENCH_VCYCLE)
    res["oracle_exact_jacobi_gamg_pcg_1e-6"] = {"iterations": pe["iterations"], "finalResidual": pe["finalResidual"],
                                                "normFactor": pe["normFactor"], "x": summary(xe)}
    print(f"  {tag} oracle exact Jacobi-GAMG-PCG 1e-6: {pe['iterations']} iters, {time.time() - t:.1f} s",
          flush=True)
    with open(os.path.join(HERE, f"{tag}_{name}_tight.json"), "w") as fh:
        json.dump(res, fh, indent=1)


def momentum_case(tag, name):
    """Momentum matrix of every velocity component on the full mesh (reference
    fvm::ddt + fvm::div − fvm::laplacian with non-orthogonal correction, fixedValue U
    on every boundary face), and the reference's reproducible PBiCGStab (DILU → diagonal
    swap) from x = U⁰ to 1e-6."""
    m = W.make_mesh(name)
    rm = ref.RefMesh(mesh=m)
    inp = W.momentum_inputs(m, 1)
    allp = list(range(len(m.patch_start)))
    res = {"mesh": name, "dt": inp["dt"], "nu": inp["nu"], "phi_sha256": sha(inp["phi"]),
           "U0_sha256": sha(inp["U0"]), "Ub_sha256": sha(inp["Ub"]), "components": []}
    for comp in range(3):
        tr = rm.transport(inp["dt"], inp["nu"], inp["U0"][:, comp], inp["phi"], allp, None, True,
                          face_values=inp["Ub"][:, comp])
        x0 = np.ascontiguousarray(inp["U0"][:, comp])
        x, p = rm.solve(tr["dF"], tr["upper"], tr["lower"], tr["bF"], x0, solver="PBiCGStab", preconditioner="DILU",
                        tolerance=1e-6, max_iter=1000, reproducible=True)
        res["components"].append({
            "diag_sha256": sha(tr["dF"]), "upper_sha256": sha(tr["upper"]), "lower_sha256": sha(tr["lower"]),
            "b_sha256": sha(tr["bF"]), "ffc_sha256": sha(tr["ffc"]),
            "repro_pbicgstab_1e-6": {"iterations": p["iterations"], "initialResidual": p["initialResidual"],
                                     "finalResidual": p["finalResidual"], "normFactor": p["normFactor"],
                                     "converged": p["converged"], "x": summary(x)}})
        print(f"  {tag} momentum component {comp}: {p['iterations']} iterations", flush=True)
    with open(os.path.join(HERE, f"{tag}_{name}_momentum.json"), "w") as fh:
        json.dump(res, fh, indent=1)


def momentum_dilu_case(tag, name):
    """Adds to the momentum fixture the reference's default momentum solve, PBiCGStab +
    DILU with plain sums (IncompressibleFlow.h:30), of velocity component 0 from x = U⁰
    to 1e-10 (1e-6 takes a single iteration on this diagonally dominant matrix)."""
    path = os.path.join(HERE, f"{tag}_{name}_momentum.json")
    with open(path) as fh:
        res = json.load(fh)
    m = W.make_mesh(name)
    rm = ref.RefMesh(mesh=m)
    inp = W.momentum_inputs(m, 1)
    allp = list(range(len(m.patch_start)))
    tr = rm.transport(inp["dt"], inp["nu"], inp["U0"][:, 0], inp["phi"], allp, None, True,
                      face_values=inp["Ub"][:, 0])
    assert sha(tr["dF"]) == res["components"][0]["diag_sha256"]
    x0 = np.ascontiguousarray(inp["U0"][:, 0])
    t = time.time()
    x, p = rm.solve(tr["dF"], tr["upper"], tr["lower"], tr["bF"], x0, solver="PBiCGStab", preconditioner="DILU",
                    tolerance=1e-10, max_iter=1000, reproducible=False)
    res["components"][0]["plain_pbicgstab_dilu_1e-10"] = {
        "iterations": p["iterations"], "initialResidual": p["initialResidual"], "finalResidual": p["finalResidual"],
        "normFactor": p["normFactor"], "converged": p["converged"], "x": summary(x)}
    print(f"  {tag} momentum DILU: {p['iterations']} iterations, {time.time() - t:.1f} s", flush=True)
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)


def exactsum_case():
    out = {"cases": []}
    for seed, scale in ((1, 1.0), (2, 1e20), (3, 1e-300), (4, 1e300)):
        v = (2.0 * W.splitmix_uniform(10007, seed) - 1.0) * scale
        v[::7] *= 1e-12
        out["cases"].append({"seed": seed, "scale": scale, "n": 10007, "sum": ref.exact_sum(v),
                             "sum_reversed": ref.exact_sum(v[::-1])})
    out["known"] = {"1e20+1-1e20": ref.exact_sum([1e20, 1.0, -1e20]), "-3.5+1.25": ref.exact_sum([-3.5, 1.25])}
    with open(os.path.join(HERE, "exactsum.json"), "w") as fh:
        json.dump(out, fh, indent=1)


def dropin_case():
    """Output of oracle/_ref/dropin_cpu (the Voluma program with the reference's own
    solver) for the drop-in test, which runs the same program linked with the GPU shim."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "dropin_cpu")
    out = subprocess.run([exe, "24"], check=True, capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "dropin_expected.jsonl"), "w") as fh:
        fh.write(out)


if __name__ == "__main__":
    if "--tight" in sys.argv:
        tight_case("cfg1", "box64")
        tight_case("cfg2", "tet70")
        sys.exit(0)
    if "--momentum-dilu" in sys.argv:
        momentum_dilu_case("cfg2", "tet70")
        sys.exit(0)
    if "--post" in sys.argv:
        for nm, mk in SMALL.items():
            post_case(nm, mk)
        sys.exit(0)
    if "--polymesh" in sys.argv:
        polymesh_case()
        gmsh_case()
        sys.exit(0)
    if "--periodic" in sys.argv:
        periodic_case()
        sys.exit(0)
    if "--piso" in sys.argv:
        from helpers import PISO_CASES
        for nm in PISO_CASES:
            piso_case(nm)
        sys.exit(0)
    if "--fvc" in sys.argv:
        for nm, mk in FVC_MESHES.items():
            fvc_case(nm, mk)
        sys.exit(0)
    for nm, mk in SMALL.items():
        post_case(nm, mk)
    for nm, mk in FVC_MESHES.items():
        fvc_case(nm, mk)
    polymesh_case()
    gmsh_case()
    from helpers import PISO_CASES
    for nm in PISO_CASES:
        piso_case(nm)
    dropin_case()
    exactsum_case()
    for nm, mk in SMALL.items():
        small_case(nm, mk)
    if "--small-only" not in sys.argv:
        full_case("cfg1", "box64")
        full_case("cfg2", "tet70")
    if "--small-only" not in sys.argv or "--momentum" in sys.argv:
        momentum_case("cfg2", "tet70")
        momentum_dilu_case("cfg2", "tet70")
