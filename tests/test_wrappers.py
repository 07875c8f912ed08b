"""CPU tests of the host-side argument checks of the LduMesh / LduMatrix wrappers: the
C ABI copies a fixed number of elements from each caller array, so a short or wrongly
shaped array must be refused with ValueError before any library call (no GPU needed:
the objects are built without a device handle and never reach the library)."""
import numpy as np
import pytest

import voluma_amd as va


def _mesh(n=10, nf=12, nb=8):
    m = object.__new__(va.LduMesh)
    m._h = None
    m.n_cells, m.n_faces, m.n_boundary = n, nf, nb
    return m


def _matrix(mesh):
    A = object.__new__(va.LduMatrix)
    A._h = None
    A.mesh = mesh
    return A


def test_short_and_misshaped_arrays_are_refused():
    m = _mesh()
    n, nb, nF = 10, 8, 20
    bf = np.zeros(nb, np.int32)
    calls = [
        lambda: m.fvc_div(np.zeros(nF - 1)),
        lambda: m.fvc_surface_integrate(np.zeros((nF, 2))),
        lambda: m.fvc_interpolate(np.zeros(n - 1)),
        lambda: m.fvc_interpolate(np.zeros(n), grad=np.zeros(2 * n)),
        lambda: m.fvc_grad(0, np.zeros((n, 2)), bf, np.zeros(nb)),
        lambda: m.fvc_grad(0, np.zeros(n), bf[:-1], np.zeros(nb)),
        lambda: m.fvc_flux(np.zeros((n, 3)), None, bf, np.zeros(nb)),
        lambda: m.boundary_values(bf, np.zeros(nb), np.zeros(n + 1)),
        lambda: m.piso_rau_hbya(np.zeros(n), np.zeros(n)),
        lambda: m.piso_hbya_boundary(bf, np.zeros((nb, 3)), np.zeros(n), np.zeros(n)),
        lambda: m.piso_face_op(va.FACE_A_PLUS_B_TIMES_C, np.zeros(nF), np.zeros(nF), None),
        lambda: m.piso_face_op(va.FACE_A_MINUS_B, np.zeros(nF), np.zeros(nF - 2)),
        lambda: m.piso_ddt_phi_corr(0.5, 1.0, 0.0, np.zeros((n, 3)), None, np.zeros(nF), None, np.ones(nF), None),
        lambda: m.piso_ddt_phi_corr(0.5, 1.0, 0.0, np.zeros((n, 3)), None, np.zeros(nF), None, np.ones(nF - 1),
                                    np.zeros(nF)),
        lambda: m.piso_constrain_pressure(0, 2, np.zeros(nF), np.zeros(nb), np.ones(nF)),
        lambda: m.fvc_reconstruct(np.zeros(nF + 1)),
        lambda: m.piso_correct_U(np.zeros((n, 3)), np.zeros(n), np.zeros(n), -1),
    ]
    A = _matrix(m)
    calls += [
        lambda: A.fvm_ddt_euler(1, 0.1, np.zeros(n), np.zeros(n - 1)),
        lambda: A.fvm_div_upwind(1, np.zeros(nF - 1), bf),
        lambda: A.fvm_laplacian(1, np.zeros(nF), np.zeros(nF), bf,
                                nonorth={"psi": np.zeros(n), "w": np.zeros(nF), "Sf": np.zeros(nF),
                                         "corr": np.zeros((nF, 3)), "V": np.zeros(n)}),
        lambda: A.fvm_source(np.zeros(n), np.zeros(n - 1)),
        lambda: A.fvm_relax(0.7, np.zeros(n + 1)),
        lambda: A.fvm_A(np.zeros(3)),
        lambda: A.fvm_H(np.zeros(n), np.zeros(2)),
        lambda: A.fvm_flux(np.zeros(n - 1)),
        lambda: A.assemble_laplacian(np.zeros(nF), np.zeros(nF - 1), bf),
        lambda: A.assemble_div_upwind(np.zeros(nF), bf, b_in=np.zeros(n + 1)),
        lambda: va.amul(A, np.zeros(n + 2)),
    ]
    for k, call in enumerate(calls):
        with pytest.raises(ValueError):
            call()
