"""Small workloads shared by the GPU tests: each builds the GPU simulation and the oracle on the
same voxel flags, faces, triangles and inlet profile."""
import numpy as np

from pv_dec_original_amd import SubDomain, Simulation
from pv_dec_original_amd import _native as N
from pv_dec_original_amd.geometry import Geometry
from pv_dec_original_amd.lattice import LatticeModel
from pv_dec_original_amd.workloads import tube_case, ybif_mesh, xfaces, Case
from tests import oracle_ctypes as O

DT = 8e-6


def small_ybif_case(q=27, box=(96, 48, 48), tau=0.6):
    nx, ny, nz = box
    tri = ybif_mesh(rp=7.0, rd=5.0, angle=30.0, xb=58.0, y0=ny / 2.0, z0=nz / 2.0, xcap0=-2.5,
                    xcap1=nx + 1.5, box=box)
    geo = Geometry.from_triangles(tri, LatticeModel(q), spacing=1.0, grid=((0.0, 0.0, 0.0), box))
    return Case("ybif_small", geo, xfaces(geo), tau, (ny / 2.0, nz / 2.0 - 0.5), 7.0)


def build(case, storage="fp64", u_max=0.05, normals=True, threads=4):
    geo = case.geometry
    sd = SubDomain(geo, case.faces, storage=storage, tau=case.tau, dt=DT)
    sim = Simulation(sd, case.tau, 1.0, DT)
    prof = case.inlet_profile(sd, u_max)
    sim.set_inlet_profile(prof)
    orc = O.OracleSim(geo.fluid, geo.origin, geo.spacing, geo.m_velSet.q, case.faces,
                      geo.globalTriangles, case.tau, DT, normals=normals, threads=threads)
    pbox = np.zeros(geo.fluid.size)
    pbox[sd.list(N.LIST_INLET)] = prof
    orc.set_inlet_profile_box(pbox)
    return sd, sim, orc


def rel_err(a, b):
    d = np.abs(a - b).max()
    s = max(np.abs(b).max(), 1e-300)
    return d / s
