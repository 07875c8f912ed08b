"""Oracle pins of the explicit pressure-corrector operators (Fvc.h): the oracle's C
restatement against the reference's own fvc functions byte for byte on an orthogonal
box and two skewed tet meshes, and against what the mathematics fixes (linear fields
reproduced exactly by the least-squares and skew-corrected Gauss gradients and by the
skew-corrected interpolate; closed cells integrate a uniform flux to zero)."""
import numpy as np
import pytest

from helpers import FVC_MESHES, fvc_bfix, load_fvc, mesh_hash
from oracle import oracle as O

NAMES = list(FVC_MESHES)


def _same(a, b):
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_fvc_matches_reference(name):
    g = load_fvc(name)
    m = FVC_MESHES[name]()
    assert bytes(g["mesh_hash"]).hex() == mesh_hash(m)
    geo = m.geometry()
    assert _same(geo["skew"], g["skew"]) and geo["skewed"] == bool(g["skewed"][0])
    bfix = fvc_bfix(m, g["in_fixed"])
    p, U, pB, UB, F = g["in_p"], g["in_U"], g["in_pB"], g["in_UB"], g["in_F"]
    gp = O.fvc_grad_ls(m, p, bfix, pB)
    gU = O.fvc_grad_ls(m, U, bfix, UB)
    assert _same(gp, g["gLSp"]) and _same(gU, g["gLSU"])
    assert _same(O.fvc_interpolate(m, p), g["ip"])
    assert _same(O.fvc_interpolate(m, p, gp), g["ipSkew"])
    iU = O.fvc_interpolate(m, U, gU)
    assert _same(iU, g["iU"])
    assert _same(O.fvc_grad_gauss(m, p, bfix, pB), g["gGL"])
    assert _same(gp, g["gLS"])  # grad(p, LeastSquares) of a fixedValue/zeroGradient field is gradLS
    assert _same(O.fvc_grad_gauss_skew(m, p, bfix, pB), g["gGS"])
    assert _same(O.fvc_div(m, F), g["divF"])
    assert _same(O.fvc_surface_integrate(m, iU), g["siU"])
    assert _same(O.fvc_flux(m, U, gU, bfix, UB), g["phi"])


@pytest.mark.parametrize("name", ["tet5x4x3", "tet6x5x4"])
def test_linear_fields_are_exact(name):
    """φ = a·x + b with exact values on every boundary face: the least-squares gradient
    is a, the skew-corrected interpolate is φ(face centre), the skew-corrected Gauss
    gradient is a; the plain Gauss gradient and interpolate are not (skewed mesh)."""
    m = FVC_MESHES[name]()
    geo = m.geometry()
    a, b = np.array([0.7, -1.3, 2.1]), 0.4
    phi = geo["C"] @ a + b
    ni = m.n_internal
    bvals = geo["Cf"][ni:] @ a + b
    allfix = np.ones(m.n_faces - ni, np.int32)
    gls = O.fvc_grad_ls(m, phi, allfix, bvals)
    assert np.abs(gls - a).max() <= 1e-11
    face = O.fvc_interpolate(m, phi, gls)
    exact = geo["Cf"] @ a + b
    assert np.abs(face[:ni] - exact[:ni]).max() <= 1e-12
    assert np.abs(O.fvc_interpolate(m, phi)[:ni] - exact[:ni]).max() > 1e-6
    assert np.abs(O.fvc_grad_gauss_skew(m, phi, allfix, bvals) - a).max() <= 1e-10
    assert np.abs(O.fvc_grad_gauss(m, phi, allfix, bvals) - a).max() > 1e-6
    # vector field U = A·x: the gradient tensor ∇U has (∇U)_ij = ∂U_j/∂x_i = A_ji
    A = np.array([[1.0, 0.5, -0.2], [0.3, -1.1, 0.8], [0.0, 0.4, 0.9]])
    U = geo["C"] @ A.T
    gU = O.fvc_grad_ls(m, U, allfix, geo["Cf"][ni:] @ A.T)
    assert np.abs(gU.reshape(-1, 3, 3) - A.T).max() <= 1e-11


@pytest.mark.parametrize("name", NAMES)
def test_uniform_flux_integrates_to_zero(name):
    """Closed cells: Σ S_f = 0, so a uniform U gives div(U·S) = 0 and the flux of a
    uniform U is U·S exactly."""
    m = FVC_MESHES[name]()
    geo = m.geometry()
    Uc = np.array([0.3, -0.8, 1.1])
    F = geo["Sf"] @ Uc
    scale = np.abs(F).max() / geo["V"].min()
    assert np.abs(O.fvc_div(m, F)).max() <= 1e-12 * scale
    si = O.fvc_surface_integrate(m, np.outer(F, np.ones(3)))
    assert np.abs(si).max() <= 1e-12 * scale
    U = np.tile(Uc, (m.n_cells, 1))
    ni = m.n_internal
    phi = O.fvc_flux(m, U, np.zeros((m.n_cells, 9)), np.zeros(m.n_faces - ni, np.int32), np.zeros((m.n_faces - ni, 3)))
    assert np.abs(phi - F).max() <= 1e-15 * np.abs(F).max() * 10
