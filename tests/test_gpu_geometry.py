"""GPU mesh geometry (vol_geometry: face area vectors and centres, cell centres and
volumes, d, |S|, w, ndc, corrVecs, skewVecs): byte-identical to the host geometry, which
the CPU tests pin to the reference's Mesh, on small box (quad faces) and tet (triangle
faces) meshes and on both full-size benchmark meshes."""
import numpy as np
import pytest

import voluma_amd as va
from helpers import FVC_MESHES, SMALL_MESHES
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu
KEYS = ("V", "C", "Sf", "Cf", "d", "magSf", "w", "ndc", "corr", "skew")


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _check(ctx, m):
    g = va.device_geometry(ctx, m)
    h = m.geometry()
    for k in KEYS:
        a, b = np.ascontiguousarray(g[k]), np.ascontiguousarray(h[k])
        assert a.shape == b.shape and a.tobytes() == b.tobytes(), k
    assert g["skewed"] == h["skewed"]


@pytest.mark.parametrize("name", list(SMALL_MESHES) + ["tet6x5x4"])
def test_device_geometry_small(ctx, name):
    mk = SMALL_MESHES.get(name) or FVC_MESHES[name]
    _check(ctx, mk())


@pytest.mark.parametrize("name", ["box64", "tet70"])
def test_device_geometry_full_size(ctx, name):
    _check(ctx, W.make_mesh(name))


def test_device_geometry_errors(ctx):
    m = SMALL_MESHES["tet5x4x3"]()
    bad = m.face_pts.copy()
    bad[5] = 10 ** 6

    class Bad:
        pass

    b = Bad()
    for k in ("n_cells", "n_faces", "n_internal", "points", "face_off", "owner", "neighbour"):
        setattr(b, k, getattr(m, k))
    b.face_pts = bad
    with pytest.raises(va.VolumaError):
        va.device_geometry(ctx, b)


def test_read_polymesh_device_geometry_and_solve(ctx):
    """The reference's 3-D cavity mesh read from its polyMesh files: device geometry equal
    to the reference Mesh's, then pressure assembly and GAMG-PCG on it converge."""
    import os
    from helpers import GOLDEN, load_json, sha
    from voluma_amd import mesh as M
    e = load_json("polymesh_expected.json")["cavity3D_4"]
    m = M.read_polymesh(os.path.join(GOLDEN, "polymesh", "cavity3D_4"))
    g = va.device_geometry(ctx, m)
    for k, h in e["geometry"].items():
        assert sha(np.ascontiguousarray(g[k])) == h, k
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, g["magSf"][:ni] * g["ndc"][:ni])
    lm.set_boundary(m.owner[ni:])
    A = va.LduMatrix(lm)
    bc = np.zeros(m.n_faces - ni, np.int32)
    s = int(m.patch_start[1]) - ni
    bc[s: s + int(m.patch_size[1])] = 1  # fixedValue on movingWall
    b = A.assemble_laplacian(g["magSf"], g["ndc"], bc, b_in=W.pressure_rhs(m.n_cells, 3))
    x = np.zeros(m.n_cells)
    p = va.solve_ldu(A, b, x, va.SolverControls("PCG", "GAMG", 1e-8, 0.0, 500))
    assert p.converged
    r = b - va.amul(A, x)
    assert np.abs(r).sum() / p.normFactor < 1e-8


@pytest.mark.parametrize("name", ["hybrid3D_v41", "tet10_v41", "tet5x4x3_v22"])
def test_read_gmsh_device_geometry(ctx, name):
    """Gmsh files read by the library: device geometry equal to the reference Mesh's."""
    import os
    from helpers import GOLDEN, load_json, sha
    from voluma_amd import mesh as M
    e = load_json("gmsh_expected.json")[name]
    m = M.read_gmsh(os.path.join(GOLDEN, "gmsh", name + ".msh"))
    g = va.device_geometry(ctx, m)
    for k, h in e["geometry"].items():
        assert sha(np.ascontiguousarray(g[k])) == h, k
    assert g["skewed"] == e["skewed"]
