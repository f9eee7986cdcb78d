"""Geometry: STL input, the reference grid rule and GPU ray-cast voxelisation.

Mirrors the reference's ``Geometry(FileName, velSet, spacing, raydirection)`` (ComplexGeometry.h:91,
ComplexGeometry.cpp:8-36): read the ASCII STL, take the bounding box, build the grid and label each
point fluid / not fluid by +z ray parity.  The labelling runs in the gfx950 voxeliser of
libpvlbm.so; boundary detection and classification happen when a :class:`SubDomain` is built.
"""
import ctypes as C

import numpy as np

from . import _native as N
from .lattice import LatticeModel


def read_stl_ascii(path):
    """Triangles [T][12] (v1, v2, v3, file normal) of an ASCII STL (GeometryReader.cpp:33-117)."""
    n = C.c_int64(0)
    N.check(N.lib().pvlbm_read_stl_ascii(str(path).encode(), None, 0, C.byref(n)))
    tri = np.zeros((n.value, 12), dtype=np.float64)
    if n.value:
        N.check(N.lib().pvlbm_read_stl_ascii(str(path).encode(), N.dptr(tri), n.value, C.byref(n)))
    return tri


def read_stl_binary(path):
    """Triangles [T][12] of a binary STL, normals normalised in single precision
    (GeometryReader.cpp:119-176)."""
    n = C.c_int64(0)
    N.check(N.lib().pvlbm_read_stl_binary(str(path).encode(), None, 0, C.byref(n)))
    tri = np.zeros((n.value, 12), dtype=np.float64)
    if n.value:
        N.check(N.lib().pvlbm_read_stl_binary(str(path).encode(), N.dptr(tri), n.value, C.byref(n)))
    return tri


def is_binary_stl(path):
    """A file is binary STL when its size is 84 + 50 * (the uint32 count at byte 80)."""
    import os
    size = os.path.getsize(path)
    if size < 84:
        return False
    with open(path, "rb") as fp:
        fp.seek(80)
        n = int.from_bytes(fp.read(4), "little")
    return size == 84 + 50 * n


def read_stl(path):
    """Binary or ASCII STL by content (the reference reads ASCII, ComplexGeometry.cpp:135)."""
    return read_stl_binary(path) if is_binary_stl(path) else read_stl_ascii(path)


def write_stl_binary(path, tri, name="pvlbm"):
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    if N.gen().pvgen_write_stl_binary(str(path).encode(), N.dptr(tri), len(tri), name.encode()) != 0:
        raise OSError(f"cannot write {path}")


def write_stl_ascii(path, tri, name="pvlbm"):
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    if N.gen().pvgen_write_stl_ascii(str(path).encode(), N.dptr(tri), len(tri), name.encode()) != 0:
        raise OSError(f"cannot write {path}")


def bounding_box(tri):
    """Reference bbox {min_x,max_x,min_y,max_y,min_z,max_z} (ComplexGeometry.cpp:142-154)."""
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    bb = np.zeros(6)
    N.check(N.lib().pvlbm_bbox(N.dptr(tri), len(tri), N.dptr(bb)))
    return bb


def grid_from_bbox(bb, h):
    """Reference grid of a bbox (ComplexGeometry.cpp:163-165): (origin, dims)."""
    g = N.Grid()
    bb = np.ascontiguousarray(bb, dtype=np.float64)
    N.check(N.lib().pvlbm_grid_from_bbox(N.dptr(bb), h, C.byref(g)))
    return np.array(g.origin[:]), tuple(g.dims[:])


def voxelize(tri, h, origin=None, dims=None):
    """GPU ray-cast labels: (fluid uint8 [nz, ny, nx], origin, dims).  origin/dims None = the
    reference grid of the mesh bbox."""
    tri = np.ascontiguousarray(tri, dtype=np.float64)
    gin = None
    if dims is not None:
        gin = N.Grid()
        for a in range(3):
            gin.origin[a] = float(origin[a])
            gin.dims[a] = int(dims[a])
    gout = N.Grid()
    N.check(N.lib().pvlbm_voxelize(N.dptr(tri), len(tri), h, C.byref(gin) if gin else None,
                                   C.byref(gout), None))
    dims = tuple(gout.dims[:])
    fluid = np.zeros((dims[2], dims[1], dims[0]), dtype=np.uint8)
    N.check(N.lib().pvlbm_voxelize(N.dptr(tri), len(tri), h, C.byref(gin) if gin else None,
                                   C.byref(gout), N.u8ptr(fluid)))
    return fluid, np.array(gout.origin[:]), dims


class Geometry:
    """The reference's Geometry: triangles, bbox, grid and fluid labels.

    ``Geometry(FileName, velSet, spacing, raydirection)`` reads an ASCII STL as the reference does;
    ``Geometry.from_triangles`` takes a mesh already in memory and optionally an explicit grid
    (origin, dims) in place of the bbox rule.  ``raydirection`` is stored and, as in the reference,
    does not change the labelling, which always casts +z rays (ComplexGeometry.cpp:994).
    """

    def __init__(self, FileName, velSet=None, spacing=0.05, raydirection='z', *, grid=None):
        tri = read_stl(FileName)
        if len(tri) == 0:
            raise ValueError(f"no facets read from {FileName}")
        self._setup(tri, velSet, spacing, raydirection, grid)
        self.stlFileName = str(FileName)

    @classmethod
    def from_triangles(cls, tri, velSet=None, spacing=1.0, raydirection='z', grid=None):
        self = cls.__new__(cls)
        self._setup(np.ascontiguousarray(tri, dtype=np.float64), velSet, spacing, raydirection, grid)
        self.stlFileName = None
        return self

    def _setup(self, tri, velSet, spacing, raydirection, grid):
        self.m_velSet = velSet if velSet is not None else LatticeModel(19)
        self.spacing = float(spacing)
        self.raydirection = raydirection
        self.globalTriangles = tri
        self.boundingbox = bounding_box(tri)
        origin, dims = (None, None) if grid is None else grid
        self.fluid, self.origin, dims = voxelize(tri, self.spacing, origin, dims)
        self.nx, self.ny, self.nz = dims

    @property
    def dims(self):
        return (self.nx, self.ny, self.nz)

    def coords(self, idx):
        """Coordinates of box indices (x fastest), min + idx*spacing (ComplexGeometry.cpp:337-340)."""
        idx = np.asarray(idx, dtype=np.int64)
        i = idx % self.nx
        j = (idx // self.nx) % self.ny
        k = idx // (self.nx * self.ny)
        return np.stack([self.origin[0] + i * self.spacing, self.origin[1] + j * self.spacing,
                         self.origin[2] + k * self.spacing], axis=-1)
