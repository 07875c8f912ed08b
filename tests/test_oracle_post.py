"""Oracle pins of the pressure-velocity coupling functions of FvMatrix (FvMatrix.h:111-141,
179-222): relax, setReference, A(), H() and flux() restated in oracle/oracle.c, checked
byte for byte against fixtures written by the reference's own FvMatrix (the momentum
matrix of a scalar and of a vector field, and a pressure matrix, on three small meshes),
and against identities the mathematics fixes (residual split A·ψ − H, the discrete
divergence of a laplacian's flux, the diagonal dominance relax establishes).
"""
import numpy as np
import pytest

from helpers import POST_CASES, SMALL_MESHES, load_post, mesh_hash, post_inputs
from oracle import oracle as O

NAMES = list(SMALL_MESHES)


def _oracle_post(m, inp):
    return O.transport_post(m, inp["dt"], inp["nu"], inp["psi0"], inp["F"], inp["bcfixed"], inp["fv"], inp["alpha"],
                            inp["ref_cell"], inp["ref_value"], True, pressure=inp["pressure"])


def _same_bytes(a, b):
    a, b = np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("tag", POST_CASES)
def test_oracle_post_matches_reference(name, tag):
    g = load_post(name)
    m = SMALL_MESHES[name]()
    assert bytes(g["mesh_hash"]).hex() == mesh_hash(m)
    inp = post_inputs(g, tag)
    o = _oracle_post(m, inp)
    for k in ("diag_relaxed", "source_relaxed", "A", "H", "flux", "diag_ref", "source_ref"):
        assert _same_bytes(o[k], g[f"{tag}_{k}"]), k


def _matrix(m, inp, k=None):
    psi = inp["psi0"] if k is None else inp["psi0"][k]
    fv = inp["fv"] if k is None else inp["fv"][k]
    if inp["pressure"]:
        return O.pressure_matrix(m, inp["nu"], psi, inp["bcfixed"], fv, True), psi
    return O.transport(m, inp["dt"], inp["nu"], psi, inp["F"], inp["bcfixed"], fv, True), psi


def _offdiag_times(m, M, psi):
    ni = m.n_internal
    y = np.zeros(m.n_cells)
    np.add.at(y, m.owner[:ni], M["upper"] * psi[m.neighbour])
    np.add.at(y, m.neighbour, M["lower"] * psi[m.owner[:ni]])
    return y


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("tag", ("ms", "p"))
def test_A_minus_H_is_the_residual(name, tag):
    """(A ψ − H)·V = (diag + Σ ic) ψ + Σ offdiag ψ_nb − (source + Σ bs): the folded
    system's residual (scalar equations, where ic − cmptAv ic = 0)."""
    g = load_post(name)
    m = SMALL_MESHES[name]()
    inp = post_inputs(g, tag)
    M, psi = _matrix(m, inp)
    d, s = O.relax(m, M, psi, inp["alpha"])
    R = dict(M, diag=d, source=s)
    V = m.geometry()["V"]
    a, h = O.A(m, d, M["ic"]), O.H(m, R, psi)
    dF, bF = O.fold(m, d, s, M["ic"], M["bs"])
    res = dF * psi + _offdiag_times(m, M, psi) - bF
    scale = np.abs(dF * psi).max() + np.abs(bF).max()
    assert np.abs((a * psi - h) * V - res).max() <= 1e-12 * scale


@pytest.mark.parametrize("name", NAMES)
def test_laplacian_flux_divergence(name):
    """For fvm::laplacian alone the discrete divergence of flux() (owner +, neighbour −,
    boundary +) is A_fold ψ − b_fold, the correction flux included."""
    g = load_post(name)
    m = SMALL_MESHES[name]()
    inp = post_inputs(g, "p")
    M, psi = _matrix(m, inp)
    F = O.flux(m, M, psi)
    ni = m.n_internal
    div = np.zeros(m.n_cells)
    np.add.at(div, m.owner[:ni], F[:ni])
    np.add.at(div, m.neighbour, -F[:ni])
    np.add.at(div, m.owner[ni:], F[ni:])
    dF, bF = O.fold(m, M["diag"], M["source"], M["ic"], M["bs"])
    res = dF * psi + _offdiag_times(m, M, psi) - bF
    assert np.abs(div - res).max() <= 1e-12 * (np.abs(dF * psi).max() + np.abs(bF).max())
    assert np.array_equal(F, g["p_flux"])  # flux does not depend on diag, so relax leaves it


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("tag", POST_CASES)
def test_relax_and_reference_properties(name, tag):
    g = load_post(name)
    m = SMALL_MESHES[name]()
    inp = post_inputs(g, tag)
    M, psi = _matrix(m, inp, 0 if tag == "mv" else None)
    ni = m.n_internal
    bD = np.zeros(m.n_cells)
    np.add.at(bD, m.owner[ni:], M["ic"])
    off = np.zeros(m.n_cells)
    np.add.at(off, m.owner[:ni], np.abs(M["upper"]))
    np.add.at(off, m.neighbour, np.abs(M["lower"]))
    d, s = O.relax(m, M, psi, inp["alpha"])
    D = np.maximum(np.abs(M["diag"] + bD), off)
    # relaxed diagonal (with the boundary part) is max(|D|, Σ|offdiag|)/α
    assert np.allclose(d + bD, D / inp["alpha"], rtol=1e-13, atol=0)
    assert np.allclose(s - M["source"], (d - M["diag"]) * psi, rtol=1e-12, atol=1e-300)
    d1, s1 = O.relax(m, M, psi, 1.0)
    assert np.array_equal(d1, M["diag"]) and np.array_equal(s1, M["source"])
    c, v = inp["ref_cell"], inp["ref_value"]
    dr, sr = O.set_reference(d, s, c, v)
    mask = np.arange(m.n_cells) != c
    assert np.array_equal(dr[mask], d[mask]) and np.array_equal(sr[mask], s[mask])
    assert dr[c] == 2.0 * d[c] and sr[c] == s[c] + d[c] * v
    dn, sn = O.set_reference(d, s, m.n_cells, v)  # no such cell: unchanged
    assert np.array_equal(dn, d) and np.array_equal(sn, s)


def test_vector_cmpt_average_is_pinned():
    """The vector fixtures tell cmptAv((v,v,v)) = (v+v+v)/3 from v: the scalar form of
    A() / H() misses them on some cells (a plausible porting slip)."""
    diffs = 0
    for name in NAMES:
        g = load_post(name)
        m = SMALL_MESHES[name]()
        inp = post_inputs(g, "mv")
        for k in range(3):
            M, psi = _matrix(m, inp, k)
            d, s = O.relax(m, M, psi, inp["alpha"])
            R = dict(M, diag=d, source=s)
            diffs += int(np.sum(O.H(m, R, psi, 1) != g["mv_H"][k]))
            if k == 0:
                diffs += int(np.sum(O.A(m, d, M["ic"], 1) != g["mv_A"]))
    assert diffs > 0
