"""Synthetic vessel workloads (meshes, grids, faces, inlet profiles).

tube        straight tube R = 10, L = 80 along x, 48 facets, reference bbox grid (81 x 21 x 20),
            parabolic Zou-He inlet on x-min, pressure outlet on x-max.
ybif        symmetric Y-bifurcation (parent R = 24, daughters R = 17 at +-30 deg) in an explicit
            512 x 256 x 256 box; inlet on x-min, both daughters leave through the x-max outlet.
vessel      seeded tortuous vessel with 1-2 branches in a 96^3 box.
aneurysm    seeded curved tube R = 40 with a spherical bulge R = 70 in a 1024 x 512 x 512 box.
branch      a parent tube with branches leaving through +y and -z faces and a flat cap inside the
            box, with windowed faces (the reference's y-max / z-min outlet closures).

Meshes come from libpvgen.so.  Implicit shapes are contoured on a grid offset from the lattice by
(-6.37, -6.29, -6.13) so that no vertical lattice ray lies in a contour-cell face.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .geometry import Geometry
from .lattice import LatticeModel
from .simulation import face, parabolic_profile

MT_OFFSET = np.array([-6.37, -6.29, -6.13])


def _mesh(fn, *args):
    m = C.c_void_p()
    rc = fn(*args, C.byref(m))
    if rc != 0:
        raise RuntimeError("mesh generation failed")
    n = N.gen().pvgen_mesh_size(m)
    tri = np.zeros((n, 12))
    N.gen().pvgen_mesh_copy(m, N.dptr(tri))
    N.gen().pvgen_mesh_free(m)
    return tri


def tube_mesh(radius=10.0, length=80.0, nseg=48):
    return _mesh(N.gen().pvgen_tube, radius, length, nseg)


def ybif_mesh(rp=24.0, rd=17.0, angle=30.0, xb=330.0, y0=128.0, z0=128.0, xcap0=-2.5,
              xcap1=513.5, box=(512, 256, 256), spacing=1.0):
    lo = np.ascontiguousarray(MT_OFFSET.copy())
    hi = np.array([box[0] + 6.0, box[1] + 6.0, box[2] + 6.0])
    return _mesh(N.gen().pvgen_ybifurcation, rp, rd, angle, xb, y0, z0, xcap0, xcap1,
                 N.dptr(lo), N.dptr(hi), spacing)


def vessel_mesh(seed, n=96, rmin=6.0, rmax=16.0, nbif=None, spacing=1.0):
    if nbif is None:
        nbif = 1 + (seed % 2)
    lo = np.ascontiguousarray(MT_OFFSET.copy())
    hi = np.array([n + 6.0] * 3)
    return _mesh(N.gen().pvgen_vessel, seed, float(n), rmin, rmax, nbif, -2.5, n + 1.5,
                 N.dptr(lo), N.dptr(hi), spacing)


def aneurysm_mesh(seed, box=(1024, 512, 512), rt=40.0, rs=70.0, spacing=2.0):
    lo = np.ascontiguousarray(MT_OFFSET * spacing)
    hi = np.array([box[0] + 8.0, box[1] + 8.0, box[2] + 8.0])
    return _mesh(N.gen().pvgen_aneurysm, seed, float(box[0]), float(box[1]), float(box[2]), rt, rs,
                 -2.5, box[0] + 1.5, N.dptr(lo), N.dptr(hi), spacing)


def tree_mesh(segments, box_lo, box_hi, spacing=1.0):
    """Vessel tree (pvgen_tree): segments = rows of (a, b, ra, rb, end_axis, end_val, end_keep)."""
    seg = np.ascontiguousarray(np.asarray(segments, dtype=np.float64).reshape(-1, 11))
    blo = np.ascontiguousarray(box_lo, dtype=np.float64)
    bhi = np.ascontiguousarray(box_hi, dtype=np.float64)
    lo = np.ascontiguousarray(MT_OFFSET * spacing + np.minimum(blo, 0.0))
    hi = np.ascontiguousarray(bhi + 6.0 * spacing)
    return _mesh(N.gen().pvgen_tree, len(seg), N.dptr(seg), N.dptr(blo), N.dptr(bhi), N.dptr(lo),
                 N.dptr(hi), spacing)


@dataclass
class Case:
    """A workload ready for SubDomain/Simulation: geometry, faces, inlet profile, tau, notes."""
    name: str
    geometry: Geometry
    faces: list
    tau: float
    inlet_center: tuple
    inlet_radius: float
    u_scale: float = 1.0
    extra: dict = field(default_factory=dict)

    def inlet_profile(self, subdomain, u_max):
        coords = self.geometry.coords(subdomain.list(N.LIST_INLET))
        return parabolic_profile(coords, self.inlet_center, self.inlet_radius, u_max)


def xfaces(geo, rho_out=1.0):
    """Inlet on the grid's x-min plane, pressure outlet on its x-max plane (whole faces)."""
    x0 = geo.origin[0]
    x1 = geo.origin[0] + (geo.nx - 1) * geo.spacing
    return [face(N.INLET, N.INLET_XMIN, 0, x0, 1e-2 * geo.spacing),
            face(N.OUTLET1, N.OUTLET_XMAX, 0, x1, 1e-2 * geo.spacing, rho=rho_out)]


def tube_case(q=19, radius=10.0, length=80.0, nseg=48, tau=0.8):
    tri = tube_mesh(radius, length, nseg)
    geo = Geometry.from_triangles(tri, LatticeModel(q), spacing=1.0)
    # nodes are fluid where (x, y, z + h/2) is inside: the lattice sees the axis at z = -h/2
    return Case("tube", geo, xfaces(geo), tau, (0.0, -0.5 * geo.spacing), radius)


def ybif_case(q=27, box=(512, 256, 256), tau=0.6, rp=24.0, rd=17.0, angle=30.0, xb=None):
    nx, ny, nz = box
    if xb is None:
        xb = nx * (330.0 / 512.0)
    tri = ybif_mesh(rp, rd, angle, xb, ny / 2.0, nz / 2.0, -2.5, nx + 1.5, box)
    geo = Geometry.from_triangles(tri, LatticeModel(q), spacing=1.0, grid=((0.0, 0.0, 0.0), box))
    return Case("ybif", geo, xfaces(geo), tau, (ny / 2.0, nz / 2.0 - 0.5), rp,
                extra={"xb": xb, "rd": rd, "angle": angle})


def branch_case(q=19, box=(96, 72, 64), tau=0.8):
    """A vessel with the reference's outlet arrangement (PV.stl: an x-face inlet, outlet 1 on a
    +y face with the y-max closure, outlets 2 and 3 on -z faces with the z-min closure,
    ComplexGeometry.cpp:1376-1408, Simulation.cpp:642-836).  A parent tube (R = 8) enters through
    the x-min face and ends in a rounded tip; branch 1 (R = 6) leaves through the y-max face of the
    box, branch 2 (R = 6) through its z-min face, and branch 3 (R = 5) ends in a flat cap on the plane
    z = 12 inside the box, as PV.stl's outlet 2 does at z = -110 above its outlet 3.  Faces carry
    finite windows around each opening, and branch 2's opening (nodes x = 45..55 on z = 0) is
    split on purpose: outlet 2's window x in (46, 53) is open, so x = 46 and x = 53 on its bounds
    are excluded; a second outlet-3 face on z = 0 with x in (49.5, 60) overlaps it at x = 50..52,
    where outlet 2 must win (role precedence), and takes x = 53..55; x = 45, 46 match nothing and
    are wall nodes."""
    nx, ny, nz = box
    y0, z0 = 24.0, 40.0
    segs = [
        ((-10.0, y0, z0), (70.0, y0, z0), 8.0, 8.0, -1, 0.0, 1),
        ((30.0, y0, z0), (30.0, ny + 10.0, z0), 6.0, 6.0, -1, 0.0, 1),
        ((50.0, y0, z0), (50.0, y0, -10.0), 6.0, 6.0, -1, 0.0, 1),
        ((70.0, y0, z0), (70.0, y0, 12.0), 5.0, 5.0, 2, 12.0, 1),
    ]
    rows = [list(a) + list(b) + [ra, rb, ax, v, keep] for a, b, ra, rb, ax, v, keep in segs]
    tri = tree_mesh(rows, (-2.5, -2.5, -2.5), (nx + 1.5, ny + 1.5, nz + 1.5))
    geo = Geometry.from_triangles(tri, LatticeModel(q), spacing=1.0, grid=((0.0, 0.0, 0.0), box))
    faces = [
        face(N.INLET, N.INLET_XMIN, 0, 0.0, 1e-2, (y0 - 10, z0 - 10), (y0 + 10, z0 + 10)),
        face(N.OUTLET1, N.OUTLET_YMAX, 1, float(ny - 1), 1e-2, (30.0 - 8, z0 - 8), (30.0 + 8, z0 + 8)),
        face(N.OUTLET2, N.OUTLET_ZMIN, 2, 0.0, 1e-2, (46.0, y0 - 8), (53.0, y0 + 8)),
        face(N.OUTLET3, N.OUTLET_ZMIN, 2, 0.0, 1e-2, (49.5, y0 - 8), (60.0, y0 + 8)),
        face(N.OUTLET3, N.OUTLET_ZMIN, 2, 12.0, 1e-2, (70.0 - 7, y0 - 7), (70.0 + 7, y0 + 7)),
    ]
    return Case("branch", geo, faces, tau, (y0, z0 - 0.5), 8.0)


def vessel_case(seed, q=19, n=96, tau=0.7):
    tri = vessel_mesh(seed, n)
    geo = Geometry.from_triangles(tri, LatticeModel(q), spacing=1.0, grid=((0.0, 0.0, 0.0), (n, n, n)))
    return Case(f"vessel{seed}", geo, xfaces(geo), tau, (0.0, 0.0), 1.0)


def aneurysm_case(seed=7, q=19, box=(1024, 512, 512), tau=0.6):
    tri = aneurysm_mesh(seed, box)
    geo = Geometry.from_triangles(tri, LatticeModel(q), spacing=1.0, grid=((0.0, 0.0, 0.0), box))
    return Case("aneurysm", geo, xfaces(geo), tau, (0.0, 0.0), 40.0)


def fluid_box_case(q=19, box=(512, 256, 256), tau=0.6):
    """An all-fluid box (every grid point a node): the step kernel's ceiling, with inlet and outlet
    on the x faces and walls on the other four.  The triangles are the box's own 12 facets."""
    nx, ny, nz = box
    c = np.array([[0, 0, 0], [nx - 1, 0, 0], [nx - 1, ny - 1, 0], [0, ny - 1, 0],
                  [0, 0, nz - 1], [nx - 1, 0, nz - 1], [nx - 1, ny - 1, nz - 1], [0, ny - 1, nz - 1]],
                 dtype=np.float64)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (0, 4, 7, 3), (1, 2, 6, 5)]
    tri = []
    for a, b, cc, d in quads:
        for v1, v2, v3 in ((a, b, cc), (a, cc, d)):
            n = np.cross(c[v2] - c[v1], c[v3] - c[v1])
            tri.append(np.concatenate([c[v1], c[v2], c[v3], n / np.linalg.norm(n)]))
    geo = Geometry.__new__(Geometry)
    geo.m_velSet = LatticeModel(q)
    geo.spacing = 1.0
    geo.raydirection = "z"
    geo.globalTriangles = np.array(tri)
    geo.boundingbox = np.array([0, nx - 1, 0, ny - 1, 0, nz - 1], dtype=np.float64)
    geo.fluid = np.ones((nz, ny, nx), dtype=np.uint8)
    geo.origin = np.zeros(3)
    geo.nx, geo.ny, geo.nz = nx, ny, nz
    geo.stlFileName = None
    return Case("box", geo, xfaces(geo), tau, (ny / 2.0, nz / 2.0), min(ny, nz) / 2.0)


def inlet_disk_profile(geo, subdomain, u_mean):
    """Parabolic profile on an inlet of arbitrary shape: per connected inlet set, fit the
    centre and equivalent radius from the node coordinates and return u(r) with mean u_mean."""
    idx = subdomain.list(N.LIST_INLET)
    c = geo.coords(idx)
    yc, zc = c[:, 1].mean(), c[:, 2].mean()
    area = len(idx) * geo.spacing ** 2
    R = np.sqrt(area / np.pi) + 0.5 * geo.spacing
    return parabolic_profile(c, (yc, zc), R, 2.0 * u_mean)
