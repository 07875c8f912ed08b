"""GPU tests of the pressure-velocity coupling functions on the FvMatrix parts
(vol_fvm_relax / set_reference / A / H / flux, FvMatrix.h:111-141,179-222) and of the
fold of a matrix copy with an extra source (vol_fvm_finish_source, as the momentum
predictor's FvMatrix M(UEqn); M −= ngp). Bar: bit-identical to the reference's own
FvMatrix (fixtures written by it) for a scalar and a vector momentum equation and a
pressure equation on three small meshes, with host and device pointers.
"""
import ctypes

import numpy as np
import pytest

import voluma_amd as va
from helpers import POST_CASES, SMALL_MESHES, load_post, post_inputs

pytestmark = pytest.mark.gpu
NAMES = list(SMALL_MESHES)


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _mesh(ctx, m):
    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
    lm.set_boundary(m.owner[ni:])
    return lm


def _assemble(A, m, inp, k):
    """The fixture's matrix for component k (vector) or the scalar field (k None)."""
    geo = m.geometry()
    psi = inp["psi0"] if k is None else np.ascontiguousarray(inp["psi0"][k])
    fv = inp["fv"] if k is None else np.ascontiguousarray(inp["fv"][k])
    bc = inp["bcfixed"]
    nonorth = {"psi": psi, "w": geo["w"], "Sf": geo["Sf"], "corr": geo["corr"], "V": geo["V"]}
    A.fvm_reset()
    if inp["pressure"]:
        A.fvm_laplacian(+1, geo["magSf"], geo["ndc"], bc, fv, gamma=inp["nu"], nonorth=nonorth)
    else:
        A.fvm_ddt_euler(+1, inp["dt"], geo["V"], psi)
        A.fvm_div_upwind(+1, inp["F"], bc, fv)
        A.fvm_laplacian(-1, geo["magSf"], geo["ndc"], bc, fv, gamma=inp["nu"], nonorth=nonorth)
    return psi


def _eq(a, b):
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("tag", POST_CASES)
def test_post_bit_exact_vs_reference(ctx, name, tag):
    g = load_post(name)
    m = SMALL_MESHES[name]()
    inp = post_inputs(g, tag)
    V = m.geometry()["V"]
    lm = _mesh(ctx, m)
    A = va.LduMatrix(lm)
    vec = tag == "mv"
    nc = 3 if vec else 1
    for k in (range(3) if vec else [None]):
        pick = (lambda a: a) if k is None else (lambda a: a[k])
        psi = _assemble(A, m, inp, k)
        A.fvm_relax(inp["alpha"], psi)
        parts = A.fvm_get()
        assert _eq(parts["diag"], g[f"{tag}_diag_relaxed"]), (k, "diag")
        assert _eq(parts["source"], pick(g[f"{tag}_source_relaxed"])), (k, "source")
        assert _eq(A.fvm_A(V, nc), g[f"{tag}_A"]), (k, "A")
        assert _eq(A.fvm_H(psi, V, nc), pick(g[f"{tag}_H"])), (k, "H")
        assert _eq(A.fvm_flux(psi), pick(g[f"{tag}_flux"])), (k, "flux")
        A.fvm_set_reference(inp["ref_cell"], inp["ref_value"])
        parts = A.fvm_get()
        assert _eq(parts["diag"], g[f"{tag}_diag_ref"]), (k, "diag_ref")
        assert _eq(parts["source"], pick(g[f"{tag}_source_ref"])), (k, "source_ref")
    A.close()
    lm.close()


def test_relax_requires_finish_and_finish_source(ctx):
    """relax / setReference leave the solver matrix unset until vol_fvm_finish;
    vol_fvm_finish_source folds source + su·V without touching the matrix's source, so
    H() afterwards is H() of the matrix without su."""
    name = "tet5x4x3"
    g = load_post(name)
    m = SMALL_MESHES[name]()
    inp = post_inputs(g, "ms")
    V = m.geometry()["V"]
    lm = _mesh(ctx, m)
    A = va.LduMatrix(lm)
    psi = _assemble(A, m, inp, None)
    A.fvm_finish()
    A.fvm_relax(inp["alpha"], psi)
    with pytest.raises(va.VolumaError):
        va.amul(A, np.zeros(m.n_cells))
    A.fvm_relax(1.0, psi)  # α ≥ 1: no change
    assert _eq(A.fvm_get()["diag"], g["ms_diag_relaxed"])
    su = 2.0 * va_uniform(m.n_cells, 5) - 1.0
    b = A.fvm_finish_source(su, V)
    parts = A.fvm_get()
    assert _eq(parts["source"], g["ms_source_relaxed"])
    bref = parts["source"] + su * V
    np.add.at(bref, m.owner[m.n_internal:], parts["bs"])
    assert _eq(b, bref)
    assert _eq(A.fvm_H(psi, V), g["ms_H"])
    d, _, _ = A.get()
    dref = parts["diag"].copy()
    np.add.at(dref, m.owner[m.n_internal:], parts["ic"])
    assert _eq(d, dref)
    va.amul(A, np.zeros(m.n_cells))  # finished again: solvable
    A.fvm_set_reference(-1, 3.0)      # no such cell: nothing changes
    A.fvm_set_reference(m.n_cells, 3.0)
    assert _eq(A.fvm_get()["diag"], parts["diag"])
    with pytest.raises(va.VolumaError):
        A.fvm_A(V, 2)
    A.close()
    lm.close()


def va_uniform(n, seed):
    from voluma_amd import workloads as W
    return W.splitmix_uniform(n, seed)


def test_post_device_pointers(ctx):
    """VOL_DEVICE inputs and outputs give the same bytes as host pointers."""
    name = "box12"
    g = load_post(name)
    m = SMALL_MESHES[name]()
    inp = post_inputs(g, "mv")
    V = m.geometry()["V"]
    lm = _mesh(ctx, m)
    A = va.LduMatrix(lm)
    psi = _assemble(A, m, inp, 1)
    dpsi, dV = va.DeviceArray.from_host(ctx, psi), va.DeviceArray.from_host(ctx, V)
    nF = m.n_faces
    dA, dH, dF = va.DeviceArray(ctx, m.n_cells), va.DeviceArray(ctx, m.n_cells), va.DeviceArray(ctx, nF)
    L = va.lib()
    va._check(L.vol_fvm_relax(A._h, ctypes.c_double(inp["alpha"]), dpsi.ptr, va.VOL_DEVICE))
    va._check(L.vol_fvm_A(A._h, dV.ptr, 3, dA.ptr, va.VOL_DEVICE))
    va._check(L.vol_fvm_H(A._h, dpsi.ptr, dV.ptr, 3, dH.ptr, va.VOL_DEVICE))
    va._check(L.vol_fvm_flux(A._h, dpsi.ptr, dF.ptr, va.VOL_DEVICE))
    assert _eq(dA.download(), g["mv_A"])
    assert _eq(dH.download(), g["mv_H"][1])
    assert _eq(dF.download(), g["mv_flux"][1])
    for d in (dpsi, dV, dA, dH, dF):
        d.free()
    A.close()
    lm.close()
