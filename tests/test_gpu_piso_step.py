"""Whole PISO time steps on the GPU (voluma_amd.piso.PisoFlow, every operator through the
library) against the reference's own IncompressibleFlow::step (IncompressibleFlow.cpp:
144-370, upwind convection) run from the same fields: a lid-driven box (Euler), a
skewed tet channel with a fixed-pressure outlet (backward time stepping, automatic
ddtPhiCorr, Gauss-skew ∇p, a non-orthogonal corrector) and a periodic channel. With reproducible sums both sides use the reference's
partition-independent solvers and U, p and φ after 2–3 steps are byte-identical; with
the default solvers (PBiCGStab + DILU, PCG + GAMG with Gauss–Seidel smoothing) they
agree to the solver tolerances."""
import numpy as np
import pytest

import voluma_amd as va
from helpers import GOLDEN, PISO_CASES, piso_inputs
from voluma_amd.piso import PisoFlow

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = va.Context(0)
    yield c
    c.close()


def _run(ctx, name):
    import os
    mk, opt, steps, dt, repro = PISO_CASES[name]
    m = mk()
    ctx.set_reproducible(repro)
    try:
        kw = {k: v for k, v in opt.items() if k not in ("nu", "u_bc", "p_bc")}
        flow = PisoFlow(ctx, m, opt["nu"], opt["u_bc"], opt["p_bc"], **kw)
        flow.U, flow.p, flow.phi = piso_inputs(m)
        for _ in range(steps):
            flow.step(dt)
    finally:
        ctx.set_reproducible(False)
    with np.load(os.path.join(GOLDEN, f"piso_{name}.npz"), allow_pickle=False) as z:
        ref = {k: z[k] for k in z.files}
    return flow, ref


@pytest.mark.parametrize("name", [n for n in PISO_CASES if n.endswith("_repro")])
def test_piso_steps_bit_identical_to_reference(ctx, name):
    flow, ref = _run(ctx, name)
    for k, got in (("U", flow.U), ("p", flow.p), ("phi", flow.phi)):
        got = np.ascontiguousarray(got)
        assert got.shape == ref[k].shape, k
        bad = np.flatnonzero(got.ravel() != ref[k].ravel())
        assert bad.size == 0, (k, bad[:5], got.ravel()[bad[:3]], ref[k].ravel()[bad[:3]])


def test_piso_steps_default_solvers_match_reference(ctx):
    flow, ref = _run(ctx, "box8_lid_default")
    for k, got in (("U", flow.U), ("p", flow.p), ("phi", flow.phi)):
        scale = np.abs(ref[k]).max()
        assert np.abs(got - ref[k]).max() <= 1e-6 * scale, k
    assert all(pf.converged for pf in flow.perf)


def test_boundary_values_and_flow_errors(ctx):
    """vol_boundary_values: fixed faces take their value, the others their owner cell's;
    vol_flow rejects bad options and steps."""
    import ctypes
    from helpers import FVC_MESHES
    m = FVC_MESHES["tet5x4x3"]()
    ni, nb = m.n_internal, m.n_boundary
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
    lm.set_boundary(m.owner[ni:])
    rng = np.random.default_rng(2)
    fix = (np.arange(nb) % 2).astype(np.int32)
    for nc in (1, 3):
        c = rng.uniform(-1, 1, (m.n_cells, nc)).squeeze()
        bv = rng.uniform(-1, 1, (nb, nc)).squeeze()
        got = lm.boundary_values(fix, bv, c)
        exp = np.where((fix == 1) if nc == 1 else (fix == 1)[:, None], bv, c[m.owner[ni:]])
        assert got.tobytes() == np.ascontiguousarray(exp).tobytes()
    from voluma_amd.piso import PisoFlow
    with pytest.raises(ValueError):
        PisoFlow(ctx, m, 0.01, [None] * 5, [None] * 6)  # one entry per patch
    with pytest.raises(va.VolumaError):
        PisoFlow(ctx, m, 0.01, [None] * 6, [None] * 6, n_corr=0)
    f = PisoFlow(ctx, m, 0.01, [(0, 0, 0)] * 6, [None] * 6)
    with pytest.raises(va.VolumaError):
        f.step(0.0)
    f.close()
