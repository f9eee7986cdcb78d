"""SubDomain and Simulation over the device-resident engine of libpvlbm.so.

``SubDomain(geometry, faces)`` replaces the reference's SubDomain(Geometry&) (Subdomain.cpp:3-32):
boundary classification by face specs, node kinds, links, next-to-wall set, normals and the f = w
start.  On one GPU the MPI sub-boxes and halos collapse to a single domain.

``Simulation(subdomain, tau, pressure, del_t)`` replaces Simulation(SubDomain&, tau, pressure,
del_t) (Simulation.h:26).  The reference runs its whole time loop inside the constructor
(Simulation.cpp:13); here that happens with ``iterate=True``, and ``step`` advances any number of
iterations of the loop body (Simulation.cpp:577-586).
"""
import ctypes as C
import math
import os

import numpy as np

from . import _native as N

BIG = 1e300


def split_plane(plane):
    """(hi, lo) doubles with hi + lo equal to the plane coordinate: a float is (plane, 0); a string
    is parsed as an x87 long double (numpy.longdouble on x86-64), as the reference's long double
    literals are (ComplexGeometry.cpp:1376, 1385), and split exactly into two doubles."""
    if isinstance(plane, str):
        ld = np.longdouble(plane)
        hi = float(ld)
        return hi, float(ld - np.longdouble(hi))
    return float(plane), 0.0


def face(role, closure, axis, plane, tol=1e-2, lo=(-BIG, -BIG), hi=(BIG, BIG), rho=1.0):
    """A face spec: boundary nodes with |c[axis] - plane| <= tol inside the open window
    lo < (other two coordinates) < hi (ComplexGeometry.cpp:1380-1407 generalised).  plane may be a
    string holding a long double literal (see split_plane)."""
    f = N.Face()
    f.role, f.closure, f.axis = role, closure, axis
    f.plane, f.plane_lo = split_plane(plane)
    f.tol, f.rho = float(tol), float(rho)
    f.lo[0], f.lo[1] = lo
    f.hi[0], f.hi[1] = hi
    return f


def pv_faces():
    """The reference's hard-coded PV.stl planes (ComplexGeometry.cpp:1376-1408): inlet on
    x = -16.9538L, outlet 1 on y = -170.016L (y-max closure), outlets 2, 3 on z = -110 and
    z = -123 (z-min closure), with their windows and the tolerance 1e-2.  The long double literals
    are carried exactly as double pairs."""
    return [face(N.INLET, N.INLET_XMIN, 0, "-16.9538", 1e-2, (-191.0, -118.0), (-187.0, -114.0)),
            face(N.OUTLET1, N.OUTLET_YMAX, 1, "-170.016", 1e-2, (2.0, -109.0), (5.0, -106.0)),
            face(N.OUTLET2, N.OUTLET_ZMIN, 2, -110.0, 1e-2, (-27.0, -198.0), (-25.0, -195.0)),
            face(N.OUTLET3, N.OUTLET_ZMIN, 2, -123.0, 1e-2, (-33.0, -189.0), (-30.0, -186.0))]


STORAGE = {"fp64": N.FP64, "fp16": N.FP16S, "bf16": N.BF16S}

# reference time step (LBM.cpp:59) used when a SubDomain is built before its Simulation
REFERENCE_DT = 8e-6


class SubDomain:
    """Device-resident lattice built from a Geometry (see module doc)."""

    def __init__(self, geometry, faces=(), storage="fp64", tau=1.0, dt=REFERENCE_DT, triangles=True):
        """triangles=False creates the lattice without the surface: no next-to-wall normals (WSS
        reads zero), no link fractions; setup then skips the normals pass over all triangles."""
        self.Geometry_ = geometry
        self.m_velSet = geometry.m_velSet
        self.q = self.m_velSet.q
        self.storage = storage
        self.faces = list(faces)
        arr = (N.Face * max(1, len(self.faces)))(*self.faces)
        g = N.Grid()
        for a in range(3):
            g.origin[a] = float(geometry.origin[a])
            g.dims[a] = int(geometry.dims[a])
        h = C.c_void_p()
        fluid = np.ascontiguousarray(geometry.fluid, dtype=np.uint8)
        tri = np.ascontiguousarray(geometry.globalTriangles if triangles else np.zeros((0, 12)),
                                   dtype=np.float64)
        N.check(N.lib().pvlbm_create(N.u8ptr(fluid), C.byref(g), geometry.spacing, self.q,
                                     STORAGE[storage], arr, len(self.faces), N.dptr(tri) if len(tri) else None, len(tri),
                                     float(tau), float(dt), C.byref(h)))
        self.handle = h.value
        self.nx, self.ny, self.nz = geometry.dims

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            try:
                N.lib().pvlbm_destroy(h)
            except Exception:
                pass
            self.handle = None

    def _chk(self, rc):
        N.check(rc, self.handle)

    def count(self, which):
        n = C.c_int64(0)
        self._chk(N.lib().pvlbm_count(self.handle, which, C.byref(n)))
        return n.value

    def list(self, which):
        out = np.zeros(self.count(which), dtype=np.int64)
        if len(out):
            self._chk(N.lib().pvlbm_list(self.handle, which, N.i64ptr(out)))
        return out

    # reference list names (Subdomain.h:46-53)
    @property
    def InletBoundaryNode_(self):
        return self.list(N.LIST_INLET)

    @property
    def WallBoundaryNode_(self):
        return self.list(N.LIST_WALL)

    @property
    def nextToTheWall(self):
        return self.list(N.LIST_NTW)

    def link_fractions(self, ndirs=19):
        """q [nwall][Q] of the wall nodes (list order), -1 where a direction has no wall link;
        ndirs = 19 is the reference's loop, ndirs = Q every direction
        (ComplexGeometry.cpp:550-608)."""
        q = np.zeros((self.count(N.LIST_WALL), self.GetVelSet().Get_nDirections()))
        if len(q):
            self._chk(N.lib().pvlbm_link_fractions(self.handle, int(ndirs), N.dptr(q)))
        return q

    def normals(self):
        """[3][nw] unit normals of the next-to-wall nodes in list order (Subdomain.cpp:224-370)."""
        n3 = np.zeros((3, self.count(N.LIST_NTW)))
        if n3.shape[1]:
            self._chk(N.lib().pvlbm_get_normals(self.handle, N.dptr(n3)))
        return n3

    def kinds(self):
        k = np.zeros((self.nz, self.ny, self.nx), dtype=np.uint8)
        self._chk(N.lib().pvlbm_get_kinds(self.handle, N.u8ptr(k)))
        return k

    def sample_kinds(self, ix, iy, iz):
        """Kinds at the box points (ix[i], iy[j], iz[k]) as an array [len(iz), len(iy), len(ix)],
        gathered on the device."""
        ax = [np.ascontiguousarray(a, dtype=np.int32) for a in (ix, iy, iz)]
        out = np.zeros((len(ax[2]), len(ax[1]), len(ax[0])), dtype=np.uint8)
        p = [a.ctypes.data_as(C.POINTER(C.c_int32)) for a in ax]
        self._chk(N.lib().pvlbm_sample_kinds(self.handle, p[0], len(ax[0]), p[1], len(ax[1]), p[2], len(ax[2]),
                                             N.u8ptr(out)))
        return out

    def counts(self):
        return {name: self.count(lid) for name, lid in
                [("fluid", N.LIST_FLUID), ("inlet", N.LIST_INLET), ("outlet1", N.LIST_OUTLET1),
                 ("outlet2", N.LIST_OUTLET2), ("outlet3", N.LIST_OUTLET3), ("wall", N.LIST_WALL),
                 ("ntw", N.LIST_NTW), ("tiles", N.LIST_TILES)]}

    def GetVelSet(self):
        return self.m_velSet


class Simulation:
    """Time stepping, boundary closures, WSS and output of one SubDomain (see module doc)."""

    def __init__(self, SubDomain_, tau, pressure=1.0, del_t=REFERENCE_DT, iterate=False, **loop):
        self.SubDomain_ = SubDomain_
        self.handle = SubDomain_.handle
        self.pressure_ = pressure  # stored; the outlets use their face density (Simulation.cpp:630)
        self.SetRelaxation(tau)
        self.m_dt = float(del_t)
        N.check(N.lib().pvlbm_set_dt(self.handle, self.m_dt), self.handle)
        self.total_runtime = 0.0
        if iterate:
            self.Iterate(**loop)

    def _chk(self, rc):
        N.check(rc, self.handle)

    def SetRelaxation(self, tau):
        self.m_relaxation = float(tau)
        self._chk(N.lib().pvlbm_set_tau(self.handle, self.m_relaxation))

    def set_inlet_profile(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if len(u) != self.SubDomain_.count(N.LIST_INLET):
            raise ValueError("one value per inlet node is required")
        self._chk(N.lib().pvlbm_set_inlet_profile(self.handle, N.dptr(u)))

    def set_wall_closure(self, mode):
        """N.WALL_REFERENCE (the reference's wall rule, default) or N.WALL_INTERPOLATED (linear
        interpolated bounce-back on the wall links' surface fractions)."""
        self._chk(N.lib().pvlbm_set_wall_closure(self.handle, int(mode)))

    def set_store_policy(self, policy):
        """N.STORES_AUTO (default: two lattices, nontemporal population stores when a step moves
        2*Q*S*n > 128 MiB; in place, plain stores and, by the lattice's lines holding nodes,
        nontemporal loads in every launch above 180 MiB, in the even launches above 240 MiB),
        N.STORES_REGULAR or N.STORES_STREAMING (plain loads); results are the same under each."""
        self._chk(N.lib().pvlbm_set_store_policy(self.handle, int(policy)))

    def streaming_stores(self):
        """True when the step writes the populations with nontemporal stores."""
        v = C.c_int(0)
        self._chk(N.lib().pvlbm_get_store_policy(self.handle, C.byref(v)))
        return bool(v.value)

    def streaming_loads(self):
        """0: the step reads the populations with plain loads; 1: nontemporal loads in the even
        in-place launches; 2: in every launch."""
        v = C.c_int(0)
        self._chk(N.lib().pvlbm_get_load_policy(self.handle, C.byref(v)))
        return int(v.value)

    def set_streaming(self, mode):
        """N.STREAMING_AUTO (in place on one lattice while it is at most 512 MiB, the pull from one
        lattice into another beyond, and always the pull under the interpolated wall closure),
        N.STREAMING_TWO_LATTICES or N.STREAMING_IN_PLACE; same results under each."""
        self._chk(N.lib().pvlbm_set_streaming(self.handle, int(mode)))

    @property
    def in_place(self):
        """True when the step streams in place on one lattice."""
        v = C.c_int(0)
        self._chk(N.lib().pvlbm_get_streaming(self.handle, C.byref(v)))
        return bool(v.value)

    def set_wss_every(self, every):
        self._chk(N.lib().pvlbm_set_wss_every(self.handle, int(every)))

    def init_fields(self, rho=None, u=None):
        """Restart from f = f_eq(rho, u) (dense box arrays) or from rest."""
        if rho is None:
            self._chk(N.lib().pvlbm_init_fields(self.handle, None, None))
        else:
            r = np.ascontiguousarray(rho, dtype=np.float64)
            v = np.ascontiguousarray(u, dtype=np.float64)
            self._chk(N.lib().pvlbm_init_fields(self.handle, N.dptr(r), N.dptr(v)))
        self.total_runtime = 0.0

    def init_fields_coarse(self, rho_c, u_c, umax, rho_range, flux_x=0.0, plane_flux=None, plane_drho=None):
        """Restart from a coarse FP32 field held in device memory (torch tensors rho (cz, cy, cx)
        and u (3, cz, cy, cx)), sampled trilinearly at every node, densities clipped into
        rho_range, every x-plane's flux scaled to flux_x when flux_x > 0 (inlet on x-min, outlets
        on x-max) or to plane_flux[x] when that (nx values) is given, plane_drho[x] added to the
        densities of plane x, speeds clipped to umax.  Returns the number of nodes whose sample was
        not finite (started at rest)."""
        if rho_c.dtype.itemsize != 4 or u_c.dtype.itemsize != 4 or not (rho_c.is_contiguous() and u_c.is_contiguous()):
            raise ValueError("rho_c and u_c must be contiguous FP32 device tensors")
        cz, cy, cx = rho_c.shape
        if tuple(u_c.shape) != (3, cz, cy, cx):
            raise ValueError("u_c must be (3,) + rho_c.shape")
        dims = (C.c_int32 * 3)(cx, cy, cz)
        nf = C.c_int64(0)
        nx = self.SubDomain_.nx
        prof = []
        for a in (plane_flux, plane_drho):
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float64)
                if a.shape != (nx,):
                    raise ValueError("plane profiles need one value per x-plane")
            prof.append(a)
        self._chk(N.lib().pvlbm_init_fields_coarse(self.handle, C.c_void_p(rho_c.data_ptr()),
                                                   C.c_void_p(u_c.data_ptr()), dims, float(umax),
                                                   float(rho_range[0]), float(rho_range[1]), float(flux_x),
                                                   N.dptr(prof[0]), N.dptr(prof[1]), C.byref(nf)))
        self.total_runtime = 0.0
        return nf.value

    def step(self, n, inlet_scale=None, timed=False):
        """n iterations of the loop body; inlet_scale: n values (None = 1.0)."""
        sc = None if inlet_scale is None else np.ascontiguousarray(inlet_scale, dtype=np.float64)
        if sc is not None and len(sc) < n:
            raise ValueError("need one inlet scale per iteration")
        if timed:
            ms = C.c_double(0)
            self._chk(N.lib().pvlbm_step_timed(self.handle, int(n), N.dptr(sc), C.byref(ms)))
            return ms.value
        self._chk(N.lib().pvlbm_step(self.handle, int(n), N.dptr(sc)))
        return None

    @property
    def iteration(self):
        it = C.c_int64(0)
        self._chk(N.lib().pvlbm_iteration(self.handle, C.byref(it)))
        return it.value

    @staticmethod
    def getInletVelocity(t, period=0.9):
        """The reference's pulsatile waveform (Simulation.cpp:548-563)."""
        return N.lib().pvlbm_pulsatile_velocity(float(t), float(period))

    def Iterate(self, t_end=2.7, period=0.9, output_every=100, outdir=None, vtu=False, chunk=100):
        """The reference loop (Simulation.cpp:569-601): total_runtime and iter start at 0; while
        total_runtime <= t_end, iteration iter runs with InletVelocity_ =
        getInletVelocity(total_runtime, period), then total_runtime += dt (an accumulated sum, as
        the reference does) and iter += 1.  When outdir is given, WriteOutput(iter) follows every
        iteration with iter % output_every == 0, iteration 0 included (Simulation.cpp:589-591,
        872-876): FlowHemodynamics/<iter>_FlowHemodynamics.csv with that iteration's WSS, OSI and
        TAWSS, and with vtu=True Fluid_Variables/<iter>.vtu with the density and velocity that the
        iteration's Collision computed (Simulation.cpp:52-105).  One launch of the engine finishes
        an iteration and runs the next one's collision, so those fields are read before the launch
        that finishes iteration iter.  Returns the paths written, in order."""
        self.total_runtime = 0.0
        it = 0
        written = []
        while self.total_runtime <= t_end:
            # iterations it .. it + n - 1, ending at the next output iteration or at t_end
            scales = []
            t = self.total_runtime
            while t <= t_end and len(scales) < chunk:
                scales.append(self.getInletVelocity(t, period))
                t += self.m_dt
                if outdir is not None and (it + len(scales) - 1) % output_every == 0:
                    break
            n = len(scales)
            last = it + n - 1
            out = outdir is not None and last % output_every == 0
            fields = None
            if out and vtu:
                if n > 1:
                    self.step(n - 1, scales[:n - 1])
                fields = self.macros()
                self.step(1, scales[n - 1:])
            else:
                self.step(n, scales)
            if out:
                written.append(self.WriteOutput(last, outdir))
                if vtu:
                    written.append(self.WriteVTK(last, outdir, fields))
            it += n
            self.total_runtime = t
        return written

    def macros(self):
        """(rho, u) of the current iteration, dense [nz,ny,nx] and [3,nz,ny,nx]."""
        sd = self.SubDomain_
        rho = np.zeros((sd.nz, sd.ny, sd.nx))
        u = np.zeros((3, sd.nz, sd.ny, sd.nx))
        self._chk(N.lib().pvlbm_get_macros(self.handle, N.dptr(rho), N.dptr(u)))
        return rho, u

    def macros_device(self):
        """(rho, u) of the current iteration as FP32 torch tensors on this handle's device,
        [nz,ny,nx] and [3,nz,ny,nx]; non-node points hold rho = 1, u = 0."""
        import torch
        sd = self.SubDomain_
        dev = torch.device("cuda", torch.cuda.current_device())
        rho = torch.empty((sd.nz, sd.ny, sd.nx), dtype=torch.float32, device=dev)
        u = torch.empty((3, sd.nz, sd.ny, sd.nx), dtype=torch.float32, device=dev)
        torch.cuda.synchronize(dev)
        self._chk(N.lib().pvlbm_get_macros_device(self.handle, C.c_void_p(rho.data_ptr()),
                                                  C.c_void_p(u.data_ptr())))
        return rho, u

    def wss(self):
        """[nw, 6]: x, y, z, WSS, OSI, TAWSS per next-to-wall node (OutputFile.cpp:884-894)."""
        nw = self.SubDomain_.count(N.LIST_NTW)
        out = np.zeros((nw, 6))
        if nw:
            self._chk(N.lib().pvlbm_get_wss(self.handle, N.dptr(out)))
        return out

    def wss_magnitude(self):
        """[nw]: the WSS column of wss() alone (one device copy of nw doubles)."""
        out = np.zeros(self.SubDomain_.count(N.LIST_NTW))
        if len(out):
            self._chk(N.lib().pvlbm_get_wss_magnitude(self.handle, N.dptr(out)))
        return out

    def wss_raw(self):
        nw = self.SubDomain_.count(N.LIST_NTW)
        tau3 = np.zeros((3, nw))
        acc = np.zeros((4, nw))
        if nw:
            self._chk(N.lib().pvlbm_get_wss_raw(self.handle, N.dptr(tau3), N.dptr(acc)))
        return tau3, acc

    def reduce(self):
        """FP64 (sum rho, inlet sum rho*ux, outlet sum rho*ux, non-finite count)."""
        out = np.zeros(4)
        self._chk(N.lib().pvlbm_reduce(self.handle, N.dptr(out)))
        return out

    def gate_events(self):
        n = C.c_int64(0)
        self._chk(N.lib().pvlbm_gate_events(self.handle, C.byref(n)))
        return n.value

    def WriteVTK(self, index, outdir, fields=None):
        """Fluid_Variables/<index>.vtu: velocity, density and speed of every fluid node on the
        lattice's hexahedra (OutputFile::WriteVTK_Collective, OutputFile.cpp:259-429; the reference
        names the file by the iteration, OutputFile.cpp:275).  fields = (rho, u) to write, default
        the current iteration's macros()."""
        from .output import write_fluid_vtu
        folder = os.path.join(outdir, "Fluid_Variables")
        os.makedirs(folder, exist_ok=True)
        path = os.path.join(folder, f"{index}.vtu")
        rho, u = self.macros() if fields is None else fields
        write_fluid_vtu(path, self.SubDomain_.kinds() != N.SOLID, rho, u)
        return path

    def WriteOutput(self, iteration, outdir):
        """FlowHemodynamics/<iter>_FlowHemodynamics.csv (OutputFile.cpp:867-938): the last WSS
        evaluation and the OSI and TAWSS of the accumulators.  These are the reference's
        quantities with WSS evaluated every iteration (set_wss_every(1), the default); with a
        coarser cadence the accumulators hold the sampled iterations only (pvlbm_set_wss_every)."""
        from .output import write_surface_csv
        folder = os.path.join(outdir, "FlowHemodynamics")
        os.makedirs(folder, exist_ok=True)
        path = os.path.join(folder, f"{iteration}_FlowHemodynamics.csv")
        write_surface_csv(path, self.wss())
        return path


def parabolic_profile(coords, center_yz, radius, u_max):
    """u_max (1 - r^2/R^2), clipped at 0, for inlet node coordinates [n, 3]."""
    r2 = (coords[:, 1] - center_yz[0]) ** 2 + (coords[:, 2] - center_yz[1]) ** 2
    return np.clip(u_max * (1.0 - r2 / radius ** 2), 0.0, None)


def womersley_scale(t, period=4000.0, mean=0.03, amp=0.02):
    """Mean inlet velocity of the pulsatile benchmark inlet at iteration t."""
    return mean + amp * math.sin(2.0 * math.pi * t / period)
