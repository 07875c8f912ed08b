"""Mesh ingest: the polyMesh reader against the reference's readPolyMesh
(RawMesh.cpp:223-336) on two of the reference's own test meshes (3-D tets in ASCII
faceList form; 2-D hexes in binary faceCompactList form), and the Gmsh reader against
the mesh the reference's importer builds (MeshReaders.cpp:131-190,236-482) from its own
hybrid (tet/hex/pyramid) and second-order tet test files (Gmsh 4.1) and a Gmsh 2.2 file
of a jittered tet mesh: points, faces and their order, owner, neighbour, patches and
cell count identical; the meshes' host geometry equal to the geometry of the Mesh the
reference builds from them. Boxes with cyclic patch pairs (one with moved interior
points) against the reference's generateBox and Mesh: raw mesh, merged cyclic faces,
their geometry and a reconstruct on them identical. Malformed input raises."""
import os

import numpy as np
import pytest

from helpers import GOLDEN, load_json, sha
from voluma_amd import mesh as M

CASES = ("cavity3D_4", "cavity_binary")


def _read(name):
    return M.read_polymesh(os.path.join(GOLDEN, "polymesh", name))


@pytest.mark.parametrize("name", CASES)
def test_polymesh_reader_matches_reference(name):
    e = load_json("polymesh_expected.json")[name]
    m = _read(name)
    assert sha(np.ascontiguousarray(m.points)) == e["points"]
    for k in ("face_off", "face_pts", "owner", "neighbour"):
        assert sha(np.ascontiguousarray(getattr(m, k), dtype=np.int64)) == e[k], k
    assert m.n_cells == e["n_cells"]
    assert m.patch_names == e["patch_names"] and m.patch_types == e["patch_types"]
    assert [int(v) for v in m.patch_start] == e["patch_start"] and [int(v) for v in m.patch_size] == e["patch_size"]


def test_polymesh_geometry_matches_reference_mesh():
    e = load_json("polymesh_expected.json")["cavity3D_4"]
    g = _read("cavity3D_4").geometry()
    for k, h in e["geometry"].items():
        assert sha(np.ascontiguousarray(g[k])) == h, k
    assert g["skewed"] == e["skewed"]
    assert abs(g["V"].sum() - 1.0) < 1e-12  # unit cube


def test_polymesh_errors(tmp_path):
    with pytest.raises(RuntimeError):
        M.read_polymesh(tmp_path / "missing")
    d = tmp_path / "bad"
    d.mkdir()
    src = os.path.join(GOLDEN, "polymesh", "cavity3D_4")
    for f in ("points", "faces", "owner", "neighbour", "boundary"):
        data = open(os.path.join(src, f), "rb").read()
        if f == "owner":
            data = data[: len(data) // 2]  # truncated list
        (d / f).write_bytes(data)
    with pytest.raises(RuntimeError):
        M.read_polymesh(d)


GMSH = ("hybrid3D_v41", "tet10_v41", "tet5x4x3_v22")


@pytest.mark.parametrize("name", GMSH)
def test_gmsh_reader_matches_reference(name):
    e = load_json("gmsh_expected.json")[name]
    m = M.read_mesh(os.path.join(GOLDEN, "gmsh", name + ".msh"))
    assert sha(np.ascontiguousarray(m.points)) == e["points"]
    for k in ("face_off", "face_pts", "owner", "neighbour"):
        assert sha(np.ascontiguousarray(getattr(m, k), dtype=np.int64)) == e[k], k
    assert m.n_cells == e["n_cells"]
    assert m.patch_names == e["patch_names"] and m.patch_types == e["patch_types"]
    assert [int(v) for v in m.patch_start] == e["patch_start"] and [int(v) for v in m.patch_size] == e["patch_size"]
    g = m.geometry()
    for k, h in e["geometry"].items():
        assert sha(np.ascontiguousarray(g[k])) == h, k
    assert g["skewed"] == e["skewed"]


def test_gmsh_errors(tmp_path):
    src = open(os.path.join(GOLDEN, "gmsh", "tet5x4x3_v22.msh")).read()
    bad = tmp_path / "bad.msh"
    bad.write_text(src.replace("$EndNodes", "$EndNodez"))  # unterminated section
    with pytest.raises(RuntimeError):
        M.read_gmsh(bad)
    b2 = tmp_path / "b2.msh"
    b2.write_text(src.replace("2.2 0 8", "2.2 1 8"))  # binary flag
    with pytest.raises(RuntimeError):
        M.read_gmsh(b2)
    with pytest.raises(RuntimeError):
        M.read_mesh(tmp_path / "missing.msh")


def test_gmsh_options_and_dispatch(tmp_path):
    """scale multiplies every coordinate; a patch type override replaces the type the
    reader picks; read_mesh tells a polyMesh directory from a Gmsh file and refuses
    anything else; a 2-D mesh (no volume elements) is refused."""
    f = os.path.join(GOLDEN, "gmsh", "hybrid3D_v41.msh")
    m1, m2 = M.read_gmsh(f), M.read_gmsh(f, scale=0.5, patch_types={"defaultFaces": "wall"})
    assert np.array_equal(m2.points, 0.5 * m1.points)
    for k in ("face_off", "face_pts", "owner", "neighbour"):
        assert np.array_equal(getattr(m1, k), getattr(m2, k))
    name = m1.patch_names[0]
    m3 = M.read_gmsh(f, patch_types={name: "symmetry"})
    assert m3.patch_types[0] == "symmetryPlane" and m3.patch_types[1:] == m1.patch_types[1:]
    d = M.read_mesh(os.path.join(GOLDEN, "polymesh", "cavity3D_4"))
    assert d.n_cells == _read("cavity3D_4").n_cells
    assert M.read_mesh(f).n_cells == m1.n_cells
    txt = tmp_path / "mesh.txt"
    txt.write_text("hello\n")
    with pytest.raises(RuntimeError):
        M.read_mesh(txt)
    flat = tmp_path / "flat.msh"
    flat.write_text("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n3\n1 0 0 0\n2 1 0 0\n3 0 1 0\n$EndNodes\n"
                    "$Elements\n1\n1 2 2 1 1 1 2 3\n$EndElements\n")
    with pytest.raises(RuntimeError, match="no volume elements"):
        M.read_gmsh(flat)
    bad_type = tmp_path / "bad_type.msh"
    bad_type.write_text(open(os.path.join(GOLDEN, "gmsh", "tet5x4x3_v22.msh")).read().replace("\n1 2 2 1 1 ", "\n1 99 2 1 1 ", 1))
    with pytest.raises(RuntimeError):
        M.read_gmsh(bad_type)


# ---------------------------------------------------------------- cyclic patch pairs
def _periodic(case):
    from helpers import periodic_channel
    e = load_json("periodic_expected.json")[case]
    if case == "channel_periodic":
        return e, periodic_channel()
    n = {"box6x5x4_px": ((6, 5, 4), (True, False, False), False),
         "box6x5x4_pxyz": ((6, 5, 4), (True, True, True), False),
         "box6x5x1_py_2d": ((6, 5, 1), (False, True, False), True)}[case]
    return e, M.box(*n[0], periodic=n[1], two_d=n[2])


PERIODIC = ("box6x5x4_px", "box6x5x4_pxyz", "box6x5x1_py_2d", "channel_periodic")


@pytest.mark.parametrize("case", PERIODIC)
def test_periodic_mesh_matches_reference(case):
    """generateBox with cyclic pairs (and a box whose interior points are moved, so that
    internal and cyclic faces are skewed): raw mesh identical to the reference's; the
    faces its Mesh computes on (cyclic pairs merged into internal faces after the
    internal ones, Mesh.cpp:246-300) with owner, neighbour, geometry, skew vectors,
    layout and patches identical; reconstruct of a face flux on them identical."""
    from helpers import CompView
    from oracle import oracle as O
    e, m = _periodic(case)
    assert sha(np.ascontiguousarray(m.points)) == e["points"]
    for k in ("face_off", "face_pts", "owner", "neighbour"):
        assert sha(np.ascontiguousarray(getattr(m, k), dtype=np.int64)) == e[k], k
    assert m.n_cells == e["n_cells"]
    assert m.patch_names == e["patch_names"] and m.patch_types == e["patch_types"]
    c = m.computational()
    assert c["n_internal"] == e["n_internal_c"]
    assert sha(c["owner"]) == e["owner_c"] and sha(c["neighbour"]) == e["neighbour_c"]
    for k, h in e["geometry"].items():
        assert sha(np.ascontiguousarray(c["geometry"][k])) == h, k
    assert c["geometry"]["skewed"] == e["skewed"]
    lay = e["mesh"]
    assert (c["two_d"], c["empty_dir"], c["solved_components"]) == (lay["two_d"], lay["empty_dir"],
                                                                    lay["solved_components"])
    assert c["patch_names"] == lay["patch_names"]
    assert c["patch_start"] == lay["patch_start"] and c["patch_size"] == lay["patch_size"]
    cv = CompView(m)
    ssf = np.random.default_rng(11).uniform(-1, 1, cv.n_faces)
    assert sha(O.fvc_reconstruct(cv, ssf, cv.empty_dir)) == e["reconstruct"]


def test_periodic_faces_are_translations():
    """Every merged cyclic face (an xMin face, owned by the first cell layer): d = C_N +
    (C_fa − C_fb) − C_P is the vector to the partner cell's image one period (1 in x)
    away, and w, ndc, corr follow from d as on any internal face (orthogonal box: w = 1/2,
    ndc = 1/|d|, corr = 0)."""
    m = M.box(6, 5, 4, periodic=(True, False, False))
    c = m.computational()
    g = c["geometry"]
    ni_raw, ni = m.n_internal, c["n_internal"]
    f = c["faces"][ni_raw:ni]
    assert len(f) == 5 * 4 and all(m.patch_names[0] == "xMin" for _ in f)
    o, n = c["owner"][ni_raw:ni], c["neighbour"][ni_raw:ni]
    C = g["C"]
    assert np.allclose((C[n] - np.array([1.0, 0, 0])) - C[o], g["d"][ni_raw:ni], atol=1e-15)
    assert np.allclose(g["d"][ni_raw:ni], [-1.0 / 6, 0, 0], atol=1e-15)  # xMin's S_f points to −x
    assert np.allclose(g["w"][ni_raw:ni], 0.5) and np.allclose(g["ndc"][ni_raw:ni], 6.0)
    assert np.abs(g["corr"][ni_raw:ni]).max() < 1e-12


def _foam(path, cls, body):
    path.write_text("FoamFile\n{\n    version 2.0;\n    format ascii;\n    class " + cls + ";\n}\n" + body)


def test_periodic_errors(tmp_path):
    """A periodic direction of one cell is refused by the box generator (as by the
    reference's generateBox, RawMesh.cpp:393), and a read mesh whose cyclic pair couples
    a cell to itself, which the LDU matrices here do not hold, by computational()."""
    with pytest.raises(RuntimeError, match="at least 2 cells"):
        M.box(1, 3, 3, periodic=(True, False, False))
    d = tmp_path / "polyMesh"
    d.mkdir()
    pts = ["(0 0 0)", "(1 0 0)", "(1 1 0)", "(0 1 0)", "(0 0 1)", "(1 0 1)", "(1 1 1)", "(0 1 1)"]
    faces = ["4(0 4 7 3)", "4(1 2 6 5)", "4(0 1 5 4)", "4(3 7 6 2)", "4(0 3 2 1)", "4(4 5 6 7)"]
    _foam(d / "points", "vectorField", "8\n(\n" + "\n".join(pts) + "\n)\n")
    _foam(d / "faces", "faceList", "6\n(\n" + "\n".join(faces) + "\n)\n")
    _foam(d / "owner", "labelList", "6\n(\n0\n0\n0\n0\n0\n0\n)\n")
    _foam(d / "neighbour", "labelList", "0\n(\n)\n")
    _foam(d / "boundary", "polyBoundaryMesh",
          "3\n(\n xMin\n {\n  type cyclic;\n  neighbourPatch xMax;\n  nFaces 1;\n  startFace 0;\n }\n"
          " xMax\n {\n  type cyclic;\n  neighbourPatch xMin;\n  nFaces 1;\n  startFace 1;\n }\n"
          " walls\n {\n  type wall;\n  nFaces 4;\n  startFace 2;\n }\n)\n")
    m = M.read_polymesh(str(d))
    assert m.patch_types == ["cyclic", "cyclic", "wall"] and m.patch_neighbours[:2] == ["xMax", "xMin"]
    with pytest.raises(RuntimeError, match="itself"):
        m.computational()
