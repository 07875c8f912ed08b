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
        lib.vm_read_gmsh.argtypes = [ctypes.c_char_p, ctypes.c_double]
        lib.vm_read_gmsh.restype = P
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


def read_gmsh(path, scale=1.0) -> Mesh:
    """An ASCII Gmsh 2.2 / 4.x file of a 3-D mesh (tets, hexes, prisms, pyramids, also
    higher-order), as the reference's readGmsh reads it (MeshReaders.cpp:236-482):
    same points, faces, face order and patches."""
    return Mesh(_lib().vm_read_gmsh(str(path).encode(), float(scale)))


def read_mesh(path, scale=1.0) -> Mesh:
    """A polyMesh directory or a Gmsh file, chosen as readMeshFile chooses
    (MeshReaders.cpp:805-816); Fluent files are not read here."""
    import os
    if os.path.isdir(path):
        return read_polymesh(path)
    if not os.path.isfile(path):
        raise RuntimeError(f"cannot open {path}")
    with open(path, "rb") as fh:
        first = fh.readline()
    if first.startswith(b"$MeshFormat"):
        return read_gmsh(path, scale)
    raise RuntimeError(f"{path}: not a polyMesh directory or a Gmsh file")


def box(nx, ny=None, nz=None) -> Mesh:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Mesh(_lib().vm_box(nx, ny, nz))


def tet(nx, ny=None, nz=None, jitter=0.2, seed=1) -> Mesh:
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return Mesh(_lib().vm_tet(nx, ny, nz, float(jitter), int(seed)))
