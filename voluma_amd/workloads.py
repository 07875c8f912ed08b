"""Synthetic workloads of the pressure/momentum linear-solve benchmarks.

* :func:`splitmix_uniform` — portable seeded uniform stream (splitmix64), so inputs
  are identical on every machine and numpy version;
* :func:`pressure_rhs` — b uniform in [−1, 1) from the seed with its mean removed
  (the mean is computed with math.fsum, i.e. correctly rounded);
* :func:`pressure_system` — the pressure Poisson system of a mesh: fvm::laplacian(1, p)
  with fixedValue 0 on the xMax patch and zeroGradient elsewhere, assembled and
  boundary-folded on the GPU into an :class:`voluma_amd.LduMatrix`.

Named workloads: ``box64`` (64³ hexes, 262,144 cells) and ``tet70`` (70³ hexes × 6
Kuhn tets, 2,058,000 cells, ±20 % interior jitter, seed 1).
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as _mesh

XMAX_PATCH = 1
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix_uniform(n, seed):
    """n doubles in [0, 1) from splitmix64 started at `seed` (same stream as meshgen)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def pressure_rhs(n, seed=1):
    b = 2.0 * splitmix_uniform(n, seed) - 1.0
    return b - math.fsum(b) / n


def make_mesh(name):
    if name == "box64":
        return _mesh.box(64)
    if name == "tet70":
        return _mesh.tet(70, 70, 70, 0.2, 1)
    if name.startswith("box"):
        return _mesh.box(int(name[3:]))
    if name.startswith("tet"):
        return _mesh.tet(int(name[3:]), None, None, 0.2, 1)
    raise ValueError(name)


def bc_fixed_xmax(m):
    bc = np.zeros(m.n_boundary, np.int32)
    s = int(m.patch_start[XMAX_PATCH]) - m.n_internal
    bc[s: s + int(m.patch_size[XMAX_PATCH])] = 1
    return bc


def pressure_system(ctx, m, seed=1, gamg=True):
    """(LduMesh, LduMatrix, b) of the pressure Poisson system on mesh m, assembled on the GPU."""
    import voluma_amd as va

    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights() if gamg else None)
    lm.set_boundary(m.owner[ni:])
    A = va.LduMatrix(lm)
    g = m.geometry()
    b = A.assemble_laplacian(g["magSf"], g["ndc"], bc_fixed_xmax(m), b_in=pressure_rhs(m.n_cells, seed))
    return lm, A, b
