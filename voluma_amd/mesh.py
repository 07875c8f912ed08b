"""Benchmark and test meshes: raw polyMesh topology plus finite-volume geometry.

The solve path takes LDU addressing (owner/neighbour over internal faces) and the
GAMG agglomeration weight |S_f|·nonOrthDeltaCoeff_f per internal face (reference
src/fvm/linalg/GAMG.cpp:210-218). This module makes those inputs for two mesh
families with a native generator (``libvoluma_meshgen.so``):

* ``box(nx, ny, nz)`` — the unit-cube hex mesh in generateBox's order
  (reference src/fvm/mesh/RawMesh.cpp:391-502);
* ``tet(nx, ny, nz, jitter, seed)`` — nx·ny·nz hexes split into 6 Kuhn tetrahedra,
  interior vertices jittered by ±jitter·spacing per axis (splitmix64 stream).

Patches are xMin, xMax, yMin, yMax, zMin, zMax in that order.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PATCH_NAMES = ("xMin", "xMax", "yMin", "yMax", "zMin", "zMax")


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libvoluma_meshgen.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.vm_box.restype = P
        lib.vm_box.argtypes = [ctypes.c_int] * 3
        lib.vm_tet.restype = P
        lib.vm_tet.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_uint64]
        lib.vm_free.argtypes = [P]
        lib.vm_sizes.argtypes = [P, P]
        lib.vm_raw.argtypes = [P] * 8
        lib.vm_geometry.argtypes = [P] * 10
        lib.vm_geometry.restype = ctypes.c_int
        lib.vm_skew.argtypes = [P, P, P]
        lib.vm_skew.restype = ctypes.c_int
        lib.vm_read_polymesh.argtypes = [ctypes.c_char_p]
        lib.vm_read_polymesh.restype = P
        for f in (lib.vm_read_gmsh, lib.vm_read_mesh_file):
            f.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_char_p]
            f.restype = P
        lib.vm_patch_info.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int]
        lib.vm_patch_info.restype = ctypes.c_int
        lib.vm_box_spec.argtypes = [ctypes.c_int] * 5
        lib.vm_box_spec.restype = P
        lib.vm_patch_neighbour.argtypes = [P, ctypes.c_int, P, ctypes.c_int]
        lib.vm_patch_neighbour.restype = ctypes.c_int
        lib.vm_set_points.argtypes = [P, P, ctypes.c_int64]
        lib.vm_set_points.restype = ctypes.c_int
        lib.vm_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _mag(v):
    return math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])


def _max_skew(g, ni):
    """Largest |skew|/|d| over the raw internal faces, summed as the reference does."""
    sk, d = g["skew"][:ni], g["d"][:ni]
    if ni == 0:
        return 0.0
    ms = np.sqrt(sk[:, 0] * sk[:, 0] + sk[:, 1] * sk[:, 1] + sk[:, 2] * sk[:, 2])
    md = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    return float(np.max(ms / md))


def _cyclic_geometry(geo, C, Cf, owner, neighbour, ni, n_int, fa, fb):
    """Face quantities of merged cyclic faces (Mesh.cpp:295-300, 598-628): d = C_N +
    (Cf_a − Cf_b) − C_P, then w, skew, ndc and corr from it, in the reference's order."""
    for k in range(n_int - ni):
        f = ni + k
        o, n = owner[f], neighbour[f]
        dd = (C[n] + (Cf[fa[k]] - Cf[fb[k]])) - C[o]
        S, c, ms = geo["Sf"][f], geo["Cf"][f], geo["magSf"][f]
        nv = S / ms
        CP = C[o]
        t = c - CP
        s_own = abs(S[0] * t[0] + S[1] * t[1] + S[2] * t[2])
        t = (CP + dd) - c
        s_nei = abs(S[0] * t[0] + S[1] * t[1] + S[2] * t[2])
        w = s_nei / (s_own + s_nei)
        geo["d"][f] = dd
        geo["w"][f] = w
        geo["skew"][f] = c - (CP + dd * (1.0 - w))
        md = _mag(dd)
        nd = nv[0] * dd[0] + nv[1] * dd[1] + nv[2] * dd[2]
        ndc = 1.0 / max(nd, 0.05 * md)
        geo["ndc"][f] = ndc
        geo["corr"][f] = nv - dd * ndc


class Mesh:
    """Polyhedral mesh with owner/neighbour addressing and reference-order geometry.

    Attributes (numpy): ``points`` (np,3), ``face_off``/``face_pts`` (face CSR),
    ``owner`` (nFaces, int32), ``neighbour`` (nInternalFaces, int32),
    ``patch_start``/``patch_size``; geometry on demand via :meth:`geometry`.
    """

    def __init__(self, handle):
        lib = _lib()
        if not handle:
            raise RuntimeError(lib.vm_last_error().decode())
        self._h = handle
        s = np.zeros(6, np.int64)
        lib.vm_sizes(self._h, _ptr(s))
        n_pts, n_faces, n_int, n_fp, n_cells, n_patch = (int(v) for v in s)
        self.n_cells, self.n_faces, self.n_internal = n_cells, n_faces, n_int
        self.points = np.zeros((n_pts, 3))
        self.face_off = np.zeros(n_faces + 1, np.int64)
        self.face_pts = np.zeros(n_fp, np.int64)
        owner = np.zeros(n_faces, np.int64)
        nei = np.zeros(max(n_int, 1), np.int64)
        self.patch_start = np.zeros(n_patch, np.int64)
        self.patch_size = np.zeros(n_patch, np.int64)
        lib.vm_raw(self._h, _ptr(self.points), _ptr(self.face_off), _ptr(self.face_pts), _ptr(owner), _ptr(nei),
                   _ptr(self.patch_start), _ptr(self.patch_size))
        self.owner = owner.astype(np.int32)
        self.neighbour = nei[:n_int].astype(np.int32)
        self.patch_names, self.patch_types = [], []
        for p in range(n_patch):
            nm, tp = ctypes.create_string_buffer(256), ctypes.create_string_buffer(64)
            lib.vm_patch_info(self._h, p, nm, 256, tp, 64)
            self.patch_names.append(nm.value.decode())
            self.patch_types.append(tp.value.decode())
        self.patch_neighbours = []
        for p in range(n_patch):
            nb = ctypes.create_string_buffer(256)
            lib.vm_patch_neighbour(self._h, p, nb, 256)
            self.patch_neighbours.append(nb.value.decode())
        self._geom = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib().vm_free(h)
            self._h = None

    def set_points(self, points):
        """Replace the point coordinates (same count, same topology); the geometry is
        recomputed on next use."""
        pts = np.ascontiguousarray(points, dtype=np.float64)
        if pts.shape != self.points.shape or _lib().vm_set_points(self._h, _ptr(pts), pts.shape[0]) != 0:
            raise ValueError("points must have the mesh's point count, shape (n, 3)")
        self.points = pts.copy()
        self._geom = None

    @property
    def n_boundary(self):
        return self.n_faces - self.n_internal

    def geometry(self):
        """dict of V, C, Sf, magSf, Cf, d, w, ndc, corr, skew (skewVecs) and skewed
        (reference Mesh accessors)."""
        if self._geom is None:
            nc, nf = self.n_cells, self.n_faces
            g = {"V": np.zeros(nc), "C": np.zeros((nc, 3)), "Sf": np.zeros((nf, 3)), "magSf": np.zeros(nf),
                 "Cf": np.zeros((nf, 3)), "d": np.zeros((nf, 3)), "w": np.zeros(nf), "ndc": np.zeros(nf),
                 "corr": np.zeros((nf, 3))}
            rc = _lib().vm_geometry(self._h, *(_ptr(g[k]) for k in
                                              ("V", "C", "Sf", "magSf", "Cf", "d", "w", "ndc", "corr")))
            if rc != 0:
                raise RuntimeError(_lib().vm_last_error().decode())
            g["skew"] = np.zeros((nf, 3))
            smax = np.zeros(1)
            if _lib().vm_skew(self._h, _ptr(g["skew"]), _ptr(smax)) != 0:
                raise RuntimeError(_lib().vm_last_error().decode())
            g["skewed"] = bool(smax[0] > 1e-8)  # Mesh.cpp:596
            self._geom = g
        return self._geom

    def computational(self):
        """The faces the reference's Mesh computes on (Mesh.cpp:236-330, 598-645): the
        internal faces, then each cyclic patch pair merged into internal faces, then the
        faces of every patch that is neither empty nor cyclic, patches in order. A mesh
        with an empty patch is 2-D: its empty direction is the axis with the largest
        Σ|S_f| component over the empty faces (the first on a tie) and it solves the other
        two vector components.

        Returns a dict: ``faces`` (the raw face of each computational face; a merged
        cyclic face is its first patch's face), ``owner``/``neighbour`` (int32; a merged
        face's neighbour is the partner face's cell), ``n_internal``,
        ``patch_names``/``patch_types``/``patch_start``/``patch_size`` (computational
        numbering), ``two_d``, ``empty_dir`` (−1 in 3-D), ``solved_components`` and
        ``geometry`` (V, C and per computational face Sf, Cf, magSf, d, w, ndc, corr,
        skew, with skewed)."""
        g = self.geometry()
        ni = self.n_internal
        Cf, Sf = g["Cf"], g["Sf"]
        faces = [np.arange(ni)]
        own = [np.asarray(self.owner[:ni])]
        nei = [np.asarray(self.neighbour)]
        cyc_face, cyc_partner = [], []
        done = [False] * len(self.patch_types)
        for a, kind in enumerate(self.patch_types):  # cyclic pairs (Mesh.cpp:246-300)
            if kind != "cyclic" or done[a]:
                continue
            b = len(self.patch_types)
            for k, nm in enumerate(self.patch_names):
                if nm == self.patch_neighbours[a]:
                    b = k
            if b == len(self.patch_types):
                raise RuntimeError(f"periodic patch {self.patch_names[a]!r} names a partner "
                                   f"{self.patch_neighbours[a]!r} that the mesh does not have")
            sa, na = int(self.patch_start[a]), int(self.patch_size[a])
            sb, nbf = int(self.patch_start[b]), int(self.patch_size[b])
            if na != nbf:
                raise RuntimeError(f"periodic pair {self.patch_names[a]!r}/{self.patch_names[b]!r}: {na} faces "
                                   f"against {nbf}")
            done[a] = done[b] = True
            ca, cb = np.zeros(3), np.zeros(3)
            min_a = 1e15  # GREAT
            for k in range(na):
                ca = ca + Cf[sa + k]
                cb = cb + Cf[sb + k]
                min_a = min(min_a, _mag(Sf[sa + k]), _mag(Sf[sb + k]))
            sep = (cb - ca) / float(na)
            h = 0.25 * math.sqrt(min_a)
            key = lambda q: (math.floor(q[0] / h), math.floor(q[1] / h), math.floor(q[2] / h))
            grid = {}
            for k in range(nbf):
                grid.setdefault(key(Cf[sb + k]), []).append(sb + k)
            for k in range(na):
                fa = sa + k
                target = Cf[fa] + sep
                kk = key(target)
                best, best_d = -1, 1e15
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dz in (-1, 0, 1):
                            for fb in grid.get((kk[0] + dx, kk[1] + dy, kk[2] + dz), ()):
                                dd = _mag(Cf[fb] - target)
                                if dd < best_d:
                                    best_d, best = dd, fb
                if best < 0 or best_d > 1e-3 * math.sqrt(_mag(Sf[fa])):
                    raise RuntimeError(f"periodic patch {self.patch_names[a]!r}: face {fa} has no partner face at its "
                                       "translated centre (only translational periodicity is merged)")
                if self.owner[fa] == self.owner[best]:
                    raise RuntimeError(f"cyclic patch {self.patch_names[a]}: a face couples cell "
                                       f"{int(self.owner[fa])} to itself (one cell across the periodic "
                                       "direction), which the LDU matrices here do not hold")
                cyc_face.append(fa)
                cyc_partner.append(best)
        if cyc_face:
            cf_, cp_ = np.asarray(cyc_face), np.asarray(cyc_partner)
            faces.append(cf_)
            own.append(np.asarray(self.owner[cf_]))
            nei.append(np.asarray(self.owner[cp_]))
        n_int = ni + len(cyc_face)
        out = {"patch_names": [], "patch_types": [], "patch_start": [], "patch_size": []}
        n = n_int
        empty_faces = []
        for q, kind in enumerate(self.patch_types):
            f = self.patch_faces(q)
            if kind == "empty":
                empty_faces.append(f)
                continue
            if kind == "cyclic":
                continue
            faces.append(f)
            own.append(np.asarray(self.owner[f]))
            out["patch_names"].append(self.patch_names[q])
            out["patch_types"].append(kind)
            out["patch_start"].append(n)
            out["patch_size"].append(len(f))
            n += len(f)
        two_d = "empty" in self.patch_types
        d = -1
        if two_d:
            # Σ|S_f| over the empty faces, one running sum in patch then face order, as
            # the reference sums (numpy's cumsum adds sequentially)
            ef = np.concatenate(empty_faces)
            empty_n = np.cumsum(np.abs(Sf[ef]), axis=0)[-1] if ef.size else np.zeros(3)
            d = 0
            for a in (1, 2):
                if empty_n[a] > empty_n[d]:
                    d = a
        fl = np.concatenate(faces)
        owner = np.concatenate(own).astype(np.int32)
        neighbour = np.concatenate(nei).astype(np.int32)
        geo = {"V": g["V"], "C": g["C"]}
        for k in ("Sf", "Cf", "magSf", "d", "w", "ndc", "corr", "skew"):
            geo[k] = np.ascontiguousarray(g[k][fl])
        skew_max = _max_skew(g, ni)
        if cyc_face:
            _cyclic_geometry(geo, g["C"], Cf, owner, neighbour, ni, n_int, np.asarray(cyc_face),
                             np.asarray(cyc_partner))
            for f in range(ni, n_int):
                skew_max = max(skew_max, _mag(geo["skew"][f]) / _mag(geo["d"][f]))
        geo["skewed"] = bool(skew_max > 1e-8)  # Mesh.cpp:596
        out.update(faces=fl, owner=owner, neighbour=neighbour, n_internal=n_int, two_d=two_d, empty_dir=d,
                   solved_components=[c for c in range(3) if not two_d or c != d], geometry=geo)
        return out

    def agglomeration_weights(self):
        """|S_f|·nonOrthDeltaCoeff_f over internal faces (reference GAMG.cpp:215)."""
        g = self.geometry()
        ni = self.n_internal
        return g["magSf"][:ni] * g["ndc"][:ni]

    def patch_faces(self, p):
        s = int(self.patch_start[p])
        return np.arange(s, s + int(self.patch_size[p]))


def read_polymesh(directory) -> Mesh:
    """An OpenFOAM polyMesh directory (points, faces, owner, neighbour, boundary; ASCII,
    or binary faceCompactList), as the reference's readPolyMesh reads it
    (RawMesh.cpp:223-336)."""
    return Mesh(_lib().vm_read_polymesh(str(directory).encode()))


def _import_args(path, scale, patch_types):
    over = "".join(f"{name}={kind}\n" for name, kind in (patch_types or {}).items())
    return str(path).encode(), float(scale), over.encode()


def read_gmsh(path, scale=1.0, patch_types=None) -> Mesh:
    """A 3-D mesh from an ASCII Gmsh MSH file (format 2.2 or 4.x; tets, hexes, prisms,
    pyramids, also higher-order ones by their corners). Points, faces, face order and
    patches follow the reference importer's rules (MeshReaders.cpp:131-190,236-482), so
    the mesh is the one the reference builds from the file. ``scale`` multiplies the
    coordinates; ``patch_types`` maps patch names to the type they get (wall, symmetry,
    empty, cyclic, patch)."""
    return Mesh(_lib().vm_read_gmsh(*_import_args(path, scale, patch_types)))


def read_mesh(path, scale=1.0, patch_types=None) -> Mesh:
    """A polyMesh directory or a Gmsh MSH file (options as for :func:`read_gmsh`)."""
    return Mesh(_lib().vm_read_mesh_file(*_import_args(path, scale, patch_types)))


def box(nx, ny=None, nz=None, periodic=(False, False, False), two_d=False) -> Mesh:
    """The unit-cube hex mesh in generateBox's order; periodic directions make their two
    patches a cyclic pair, two_d (nz = 1) one empty frontAndBack patch."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    if not any(periodic) and not two_d:
        return Mesh(_lib().vm_box(nx, ny, nz))
    bits = sum(1 << d for d in range(3) if periodic[d])
    return Mesh(_lib().vm_box_spec(nx, ny, nz, bits, int(bool(two_d))))


def tet(nx, ny=None, nz=None, jitter=0.2, seed=1) -> Mesh:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Mesh(_lib().vm_tet(nx, ny, nz, float(jitter), int(seed)))
