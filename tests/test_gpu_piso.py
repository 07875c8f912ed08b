"""GPU tests of the pressure-corrector steps (vol_piso_*, IncompressibleFlow.cpp:17-29,
242-362) and vol_fvc_reconstruct (Fvc.h:33-67): byte-identical to the oracle (which the
CPU tests pin to the reference's reconstruct and to closed forms) on an orthogonal box,
two skewed tet meshes and the full-size 2.06M-cell tet mesh; argument errors are
reported."""
import numpy as np
import pytest

import voluma_amd as va
from helpers import FVC_MESHES
from oracle import oracle as O
from voluma_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _same(a, b):
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def _lm(ctx, m):
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
    lm.set_boundary(m.owner[ni:])
    g = m.geometry()
    lm.set_geometry(g["V"], g["Sf"], g["magSf"], g["w"], g["d"], g["skew"], g["skewed"])
    return lm


def _check_all(ctx, m, seed):
    lm = _lm(ctx, m)
    rng = np.random.default_rng(seed)
    n, nF, nb = m.n_cells, m.n_faces, m.n_boundary
    A = rng.uniform(0.5, 2.0, n)
    H = rng.uniform(-1, 1, (n, 3))
    rau, hbya = lm.piso_rau_hbya(A, H)
    orau, ohbya = O.piso_rau_hbya(A, H)
    assert _same(rau, orau) and _same(hbya, ohbya)
    gp = rng.uniform(-1, 1, (n, 3))
    bfix = (np.arange(nb) % 3 == 0).astype(np.int32)
    Ub = rng.uniform(-1, 1, (nb, 3))
    assert _same(lm.piso_hbya_boundary(bfix, Ub, rau, gp), O.piso_hbya_boundary(m, bfix, Ub, rau, gp))
    a, b, c = rng.uniform(-1, 1, (3, nF))
    c = c + 2.0
    for op in (va.FACE_A_PLUS_B_TIMES_C, va.FACE_A_MINUS_B, va.FACE_A_MINUS_B_OVER_C):
        cc = None if op == va.FACE_A_MINUS_B else c
        assert _same(lm.piso_face_op(op, a, b, cc), O.piso_face_op(op, a, b, cc)), op
    U0, U00 = rng.uniform(-1, 1, (2, n, 3))
    phi0, phi00, rAUf, phiH = rng.uniform(-1, 1, (4, nF))
    rAUf = rAUf + 2.0
    for coeff, c0, c00 in ((0.7, 100.0, 0.0), (-1.0, 100.0, 0.0), (-1.0, 150.0, 50.0)):
        got = lm.piso_ddt_phi_corr(coeff, c0, c00, U0, U00 if c00 else None, phi0, phi00 if c00 else None, rAUf,
                                   phiH)
        exp = O.piso_ddt_phi_corr(m, coeff, c0, c00, U0, U00 if c00 else None, phi0, phi00 if c00 else None, rAUf,
                                  phiH)
        assert _same(got, exp), (coeff, c00)
    p = 1
    s, z = int(m.patch_start[p]) - m.n_internal, int(m.patch_size[p])
    assert _same(lm.piso_constrain_pressure(s, z, phiH, Ub, rAUf), O.piso_constrain_pressure(m, s, z, phiH, Ub, rAUf))
    assert _same(lm.fvc_reconstruct(phiH), O.fvc_reconstruct(m, phiH))
    for sign, ed in ((-1, -1), (1, 2)):
        assert _same(lm.piso_correct_U(hbya, rau, gp, sign, ed), O.piso_correct_U(hbya, rau, gp, sign, ed))


@pytest.mark.parametrize("name", list(FVC_MESHES))
def test_piso_steps_small(ctx, name):
    _check_all(ctx, FVC_MESHES[name](), 1)


def test_piso_steps_full_size(ctx):
    _check_all(ctx, W.make_mesh("tet70"), 2)


def test_piso_errors(ctx):
    m = FVC_MESHES["tet5x4x3"]()
    lm = _lm(ctx, m)
    nF = m.n_faces
    with pytest.raises(va.VolumaError):
        lm.piso_face_op(7, np.zeros(nF), np.zeros(nF), np.zeros(nF))
    with pytest.raises(va.VolumaError):
        lm.piso_constrain_pressure(m.n_boundary - 1, 5, np.zeros(nF), np.zeros((m.n_boundary, 3)), np.ones(nF))
    with pytest.raises(va.VolumaError):
        lm.fvc_reconstruct(np.zeros(nF), 3)
    with pytest.raises(va.VolumaError):
        lm.piso_ddt_phi_corr(0.5, 1.0, 1.0, np.zeros((m.n_cells, 3)), None, np.zeros(nF), None, np.ones(nF),
                             np.zeros(nF))
