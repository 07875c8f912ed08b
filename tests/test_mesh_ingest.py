"""Mesh ingest: the polyMesh reader against the reference's readPolyMesh
(RawMesh.cpp:223-336) on two of the reference's own test meshes (3-D tets in ASCII
faceList form; 2-D hexes in binary faceCompactList form): points, faces, owner,
neighbour, patches and cell count identical; the 3-D mesh's host geometry equal to the
geometry of the Mesh the reference builds from it. Malformed input raises."""
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
