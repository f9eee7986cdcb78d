"""MI355X-native lattice Boltzmann engine for vessel blood flow.

The reference solver's pipeline LatticeModel -> Geometry -> SubDomain -> Simulation
(LBM.cpp:69-72) over hand-written gfx950 kernels behind the C ABI of include/pvlbm.h.
"""
from .lattice import LatticeModel
from .geometry import Geometry, read_stl_ascii, write_stl_ascii, bounding_box, grid_from_bbox, voxelize
from .simulation import SubDomain, Simulation, face, pv_faces, parabolic_profile, womersley_scale
from . import _native as native

__all__ = ["LatticeModel", "Geometry", "SubDomain", "Simulation", "face", "pv_faces",
           "read_stl_ascii", "write_stl_ascii", "bounding_box", "grid_from_bbox", "voxelize",
           "parabolic_profile", "womersley_scale", "native"]
