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
        for f in (lib.vm_read_gmsh, lib.vm_read_fluent, lib.vm_read_mesh_file):
            f.argtypes = [ctypes.c_char_p, ctypes.c_double, ctypes.c_double, ctypes.c_char_p]
            f.restype = P
        lib.vm_patch_info.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P, ctypes.c_int]
        lib.vm_patch_info.restype = ctypes.c_int
        lib.vm_last_error.restype = ctypes.c_char_p
        _LIB = lib
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


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
        self._geom = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib().vm_free(h)
            self._h = None

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
        """The faces the reference's Mesh computes on (Mesh.cpp:236-330, 641-645):
        internal faces, then the faces of each patch that is not ``empty``, patches in
        order. A mesh with an empty patch is 2-D: its empty direction is the axis with
        the largest Σ|S_f| component over the empty faces (the first on a tie) and it
        solves the other two vector components. Cyclic patches are not supported.

        Returns a dict: ``faces`` (raw face index of each computational face),
        ``patch_names``/``patch_types``/``patch_start``/``patch_size`` (computational
        numbering), ``two_d``, ``empty_dir`` (−1 in 3-D), ``solved_components``."""
        if "cyclic" in self.patch_types:
            raise RuntimeError("cyclic patches are not supported")
        faces = [np.arange(self.n_internal)]
        out = {"patch_names": [], "patch_types": [], "patch_start": [], "patch_size": []}
        n = self.n_internal
        empty_n = np.zeros(3)
        sf = None
        for p, kind in enumerate(self.patch_types):
            f = self.patch_faces(p)
            if kind == "empty":
                sf = self.geometry()["Sf"] if sf is None else sf
                for v in np.abs(sf[f]):  # in face order, as the reference sums
                    empty_n += v
                continue
            faces.append(f)
            out["patch_names"].append(self.patch_names[p])
            out["patch_types"].append(kind)
            out["patch_start"].append(n)
            out["patch_size"].append(len(f))
            n += len(f)
        two_d = "empty" in self.patch_types
        d = -1
        if two_d:
            d = 0
            for a in (1, 2):
                if empty_n[a] > empty_n[d]:
                    d = a
        out.update(faces=np.concatenate(faces), two_d=two_d, empty_dir=d,
                   solved_components=[c for c in range(3) if not two_d or c != d])
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


def _import_args(path, scale, depth2d, patch_types):
    over = "".join(f"{name}={kind}\n" for name, kind in (patch_types or {}).items())
    return str(path).encode(), float(scale), float(depth2d), over.encode()


def read_gmsh(path, scale=1.0, depth2d=0.1, patch_types=None) -> Mesh:
    """An ASCII Gmsh 2.2 / 4.x file, as the reference's readGmsh reads it
    (MeshReaders.cpp:236-482): same points, faces, face order and patches. A 3-D mesh
    (tets, hexes, prisms, pyramids, also higher-order) is assembled from its cells'
    faces; a 2-D mesh (z = const) is extruded one layer of ``depth2d``. Options follow
    MeshImportOptions: ``scale`` multiplies coordinates, ``patch_types`` maps patch
    names to the type they get (wall, symmetry, empty, cyclic, patch)."""
    return Mesh(_lib().vm_read_gmsh(*_import_args(path, scale, depth2d, patch_types)))


def read_fluent(path, scale=1.0, depth2d=0.1, patch_types=None) -> Mesh:
    """A Fluent mesh file (ASCII or binary sections, 2-D or 3-D), as the reference's
    readFluent reads it (MeshReaders.cpp:484-803); options as for :func:`read_gmsh`."""
    return Mesh(_lib().vm_read_fluent(*_import_args(path, scale, depth2d, patch_types)))


def read_mesh(path, scale=1.0, depth2d=0.1, patch_types=None) -> Mesh:
    """A polyMesh directory, a Gmsh file or a Fluent file, chosen as readMeshFile
    chooses (MeshReaders.cpp:805-816)."""
    return Mesh(_lib().vm_read_mesh_file(*_import_args(path, scale, depth2d, patch_types)))


def box(nx, ny=None, nz=None) -> Mesh:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Mesh(_lib().vm_box(nx, ny, nz))


def tet(nx, ny=None, nz=None, jitter=0.2, seed=1) -> Mesh:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Mesh(_lib().vm_tet(nx, ny, nz, float(jitter), int(seed)))
