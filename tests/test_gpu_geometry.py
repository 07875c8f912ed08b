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
