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


def pressure_psi0(C):
    """Previous pressure field of the non-orthogonal correction: p⁰ = cos(2πx)·sin(2πy) + z
    at the cell centres C (n × 3), with the machine-independent sin/cos of _sincos."""
    sx, cx = _sincos(2.0 * math.pi * C[:, 0])
    sy, cy = _sincos(2.0 * math.pi * C[:, 1])
    return cx * sy + C[:, 2]


def pressure_system(ctx, m, seed=1, gamg=True, nonorth=False):
    """(LduMesh, LduMatrix, b) of the pressure Poisson system on mesh m, assembled on the GPU.

    nonorth=False: b = rhs + folded boundary source (the correction's explicit part is
    left out, which is what a solve from p⁰ = 0 sees). nonorth=True: fvm::laplacian(1, p)
    with its non-orthogonal correction evaluated on p⁰ = pressure_psi0 (Fvm.h:155-170:
    Gauss-linear gradient, source[o] −= fc, source[n] += fc) and b = rhs + that folded
    source. The matrix is the same either way."""
    import voluma_amd as va

    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights() if gamg else None)
    lm.set_boundary(m.owner[ni:])
    A = va.LduMatrix(lm)
    g = m.geometry()
    rhs = pressure_rhs(m.n_cells, seed)
    if not nonorth:
        b = A.assemble_laplacian(g["magSf"], g["ndc"], bc_fixed_xmax(m), b_in=rhs)
        return lm, A, b
    psi = pressure_psi0(g["C"])
    A.fvm_reset()
    A.fvm_laplacian(+1, g["magSf"], g["ndc"], bc_fixed_xmax(m), None,
                    nonorth={"psi": psi, "w": g["w"], "Sf": g["Sf"], "corr": g["corr"], "V": g["V"]})
    b = rhs + A.fvm_finish()
    return lm, A, b


# ------------------------------------------------------------------ momentum system
MOMENTUM_NU = 1e-3
MOMENTUM_CFL = 0.5


def _sincos(x):
    """sin and cos by range reduction to [−π, π] and Taylor polynomials in Horner
    form: only +, ×, floor, so the values are the same on every CPU (library sin/cos
    may differ in the last bit between SIMD paths). Error ≲ 1e-14."""
    two_pi = 2.0 * math.pi
    r = x - two_pi * np.floor(x / two_pi + 0.5)
    r2 = r * r
    s = np.zeros_like(r)
    c = np.zeros_like(r)
    for k in range(14, -1, -1):  # Σ (−1)^k r^(2k+1)/(2k+1)!, Σ (−1)^k r^(2k)/(2k)!
        s = s * r2 + ((-1.0) ** k) / math.factorial(2 * k + 1)
        c = c * r2 + ((-1.0) ** k) / math.factorial(2 * k)
    return s * r, c


def _dot3(a, b):
    return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]


class FourierVelocity:
    """U = ∇×A for the seeded vector potential A(x) = Σ_k a_k sin(2π k·x + θ_k), with
    integer wave vectors (|k_i| ≤ 2), scaled so that the RMS of |U| over 4096 seeded
    points of the unit box is 1. The face flux is the circulation of A around each
    face (Stokes), so the flux out of every closed cell sums to zero up to rounding.
    Everything is evaluated with elementwise +, × and fixed-order sums, so the inputs
    are bit-identical on every machine."""

    def __init__(self, seed, n_modes=12, exact=True):
        self.exact = exact
        u = splitmix_uniform(7 * n_modes, seed).reshape(n_modes, 7)
        k = np.floor(u[:, :3] * 5.0) - 2.0
        k[np.all(k == 0, axis=1), 0] = 1.0
        self.k = k
        self.a = 2.0 * u[:, 3:6] - 1.0
        self.theta = 2.0 * math.pi * u[:, 6]
        self.kxa = np.stack([k[:, 1] * self.a[:, 2] - k[:, 2] * self.a[:, 1],
                             k[:, 2] * self.a[:, 0] - k[:, 0] * self.a[:, 2],
                             k[:, 0] * self.a[:, 1] - k[:, 1] * self.a[:, 0]], axis=1)
        q = splitmix_uniform(3 * 4096, seed + 1).reshape(4096, 3)
        self.scale = 1.0
        Uq = self.U(q)
        self.scale = 1.0 / math.sqrt(math.fsum(_dot3(Uq, Uq)) / 4096.0)

    def _phase(self, x):
        x = np.asarray(x, dtype=np.float64)
        ph = np.zeros((x.shape[0], self.k.shape[0]))
        for j in range(self.k.shape[0]):
            kx = x[:, 0] * self.k[j, 0] + x[:, 1] * self.k[j, 1] + x[:, 2] * self.k[j, 2]
            ph[:, j] = 2.0 * math.pi * kx + self.theta[j]
        return ph

    def _modes(self, w, vec):
        out = np.zeros((w.shape[0], 3))
        for j in range(w.shape[1]):
            for d in range(3):
                out[:, d] = out[:, d] + w[:, j] * vec[j, d]
        return out

    def _sc(self, ph):
        # exact: machine-independent bits; otherwise the (faster) library sin/cos
        return _sincos(ph) if self.exact else (np.sin(ph), np.cos(ph))

    def A(self, x):
        s, _ = self._sc(self._phase(x))
        return self.scale * self._modes(s, self.a)

    def U(self, x):  # ∇×(a sin(2πk·x + θ)) = 2π cos(·) k×a
        _, c = self._sc(self._phase(x))
        return self.scale * (2.0 * math.pi) * self._modes(c, self.kxa)

    def face_flux(self, m):
        """Σ over the face's edges (in point order) of ½(A_i + A_j)·(x_j − x_i)."""
        off = np.asarray(m.face_off, dtype=np.int64)
        pts = np.asarray(m.face_pts, dtype=np.int64)
        nf = off.size - 1
        nslot = np.diff(off)
        X = np.asarray(m.points, dtype=np.float64)
        Ap = self.A(X)
        phi = np.zeros(nf)
        for k in range(int(nslot.max())):  # edges in point order, summed left to right
            has = nslot > k
            f = np.nonzero(has)[0]
            p = pts[off[f] + k]
            q = pts[np.where(k + 1 < nslot[f], off[f] + k + 1, off[f])]
            phi[f] = phi[f] + _dot3(0.5 * (Ap[p] + Ap[q]), X[q] - X[p])
        return phi


def momentum_inputs(m, seed=1, exact=True):
    """Inputs of the momentum matrix on mesh m: face flux φ, the cell velocity U⁰ at
    the cell centres (n×3), Dirichlet boundary values (nb×3, U at the boundary face
    centres) and dt for Courant number 0.5 with Co = ½ Σ|φ|/V·dt
    (IncompressibleFlow.cpp:97-107). exact=False evaluates sin/cos with numpy (faster,
    last-bit results may differ between machines)."""
    g = m.geometry()
    fv = FourierVelocity(seed, exact=exact)
    phi = fv.face_flux(m)
    U0 = fv.U(g["C"])
    Ub = fv.U(g["Cf"][m.n_internal:])
    sum_abs = np.zeros(m.n_cells)
    np.add.at(sum_abs, np.asarray(m.owner), np.abs(phi))
    np.add.at(sum_abs, np.asarray(m.neighbour), np.abs(phi[: m.n_internal]))
    co_per_dt = float(np.max(0.5 * sum_abs / g["V"]))
    dt = MOMENTUM_CFL / co_per_dt
    return {"phi": phi, "U0": U0, "Ub": Ub, "dt": float(dt), "nu": MOMENTUM_NU}


def momentum_assemble(A, m, inp, comp):
    """UEqn = ddt(U) + div(φ, U) − laplacian(ν, U) for velocity component comp, with
    fixedValue on every boundary face (IncompressibleFlow.cpp:183-185), assembled on
    the GPU into LduMatrix A; returns b."""
    g = m.geometry()
    bc = np.ones(m.n_boundary, np.int32)
    psi = np.ascontiguousarray(inp["U0"][:, comp])
    bv = np.ascontiguousarray(inp["Ub"][:, comp])
    A.fvm_reset()
    A.fvm_ddt_euler(+1, inp["dt"], g["V"], psi)
    A.fvm_div_upwind(+1, inp["phi"], bc, bv)
    A.fvm_laplacian(-1, g["magSf"], g["ndc"], bc, bv, gamma=inp["nu"],
                    nonorth={"psi": psi, "w": g["w"], "Sf": g["Sf"], "corr": g["corr"], "V": g["V"]})
    return A.fvm_finish()


def momentum_system(ctx, m, seed=1, comp=0, gamg=False):
    """(LduMesh, LduMatrix, b, inputs) of one momentum component on mesh m."""
    import voluma_amd as va

    ni = m.n_internal
    lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, m.agglomeration_weights() if gamg else None)
    lm.set_boundary(m.owner[ni:])
    A = va.LduMatrix(lm)
    inp = momentum_inputs(m, seed)
    b = momentum_assemble(A, m, inp, comp)
    return lm, A, b, inp


# ------------------------------------------------------------------ batches of systems
def _exp(x):
    """e^x by range reduction x = k·ln2 + r and a Horner polynomial (only +, ×, ldexp),
    the same on every CPU."""
    k = math.floor(x / math.log(2.0) + 0.5)
    r = x - k * math.log(2.0)
    s = 0.0
    for j in range(24, -1, -1):
        s = s * r + 1.0 / math.factorial(j)
    return math.ldexp(s, int(k))


def _cbrt(v):
    y = max(1.0, v / 3.0)
    for _ in range(200):  # Newton on y³ = v (division and products only)
        y2 = (2.0 * y + v / (y * y)) / 3.0
        if y2 == y:
            break
        y = y2
    return y


def batch_specs(n_systems=256, seed=4, n_min=5000, n_max=200000, median=30000, sigma=0.8):
    """Seeded descriptions of the systems of a batch: lognormal cell counts (median
    `median`, clipped to [n_min, n_max]; normal deviates from the sum of 12 uniforms),
    box-hex or jittered-tet meshes with per-axis aspect factors in [1, 4) and one random
    Dirichlet patch. Returns dicts (kind, nx, ny, nz, seed, patch)."""
    u = splitmix_uniform(20 * n_systems, seed).reshape(n_systems, 20)
    out = []
    for s in range(n_systems):
        z = float(sum(u[s, :12])) - 6.0
        n = min(max(median * _exp(sigma * z), n_min), n_max)
        kind = "tet" if u[s, 12] < 0.5 else "box"
        fac = [1.0 + 3.0 * float(u[s, 13 + d]) for d in range(3)]
        vol = n / (6.0 if kind == "tet" else 1.0)
        base = _cbrt(vol / (fac[0] * fac[1] * fac[2]))
        dims = [max(2, int(base * f + 0.5)) for f in fac]
        cells = lambda d: d[0] * d[1] * d[2] * (6 if kind == "tet" else 1)
        while cells(dims) < n_min:  # rounding may undershoot the floor
            dims[dims.index(min(dims))] += 1
        out.append({"kind": kind, "nx": dims[0], "ny": dims[1], "nz": dims[2], "seed": 1000 + s,
                    "patch": int(u[s, 16] * 6.0)})
    return out


def batch_mesh(spec):
    if spec["kind"] == "tet":
        return _mesh.tet(spec["nx"], spec["ny"], spec["nz"], 0.2, spec["seed"])
    return _mesh.box(spec["nx"], spec["ny"], spec["nz"])


def batch_system(ctx, specs, rhs_seed=7):
    """(LduBatch, b, meshes): the pressure Poisson system of every spec (fvm::laplacian(1,
    p), fixedValue 0 on the spec's patch, zeroGradient elsewhere), assembled on the GPU
    one system at a time and concatenated."""
    import voluma_amd as va

    meshes = [batch_mesh(sp) for sp in specs]
    diag, upper, bs = [], [], []
    for k, (sp, m) in enumerate(zip(specs, meshes)):
        ni = m.n_internal
        lm = va.LduMesh(ctx, m.n_cells, m.owner[:ni], m.neighbour, None)
        lm.set_boundary(m.owner[ni:])
        A = va.LduMatrix(lm)
        g = m.geometry()
        bc = np.zeros(m.n_boundary, np.int32)
        s0 = int(m.patch_start[sp["patch"]]) - ni
        bc[s0: s0 + int(m.patch_size[sp["patch"]])] = 1
        b = A.assemble_laplacian(g["magSf"], g["ndc"], bc, b_in=pressure_rhs(m.n_cells, rhs_seed + k))
        d, u, _ = A.get()
        diag.append(d)
        upper.append(u)
        bs.append(b)
        A.close()
        lm.close()
    systems = [(m.n_cells, m.owner[: m.n_internal], m.neighbour, m.agglomeration_weights()) for m in meshes]
    B = va.LduBatch(ctx, systems)
    B.diag_h, B.upper_h = np.concatenate(diag), np.concatenate(upper)  # host copies, for checks
    B.set(B.diag_h, B.upper_h)
    return B, np.concatenate(bs), meshes
