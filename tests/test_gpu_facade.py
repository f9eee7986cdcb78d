"""The C++ facade's reference loop on the device, in the test process.

pvlbm::Simulation::Iterate (include/pvlbm.hpp; Simulation::Iterate, Simulation.cpp:569-601) runs
through tests/facade/libfacade_iterate.so, a shim the package build compiles against libpvlbm.so:
total_runtime accumulated by += dt from 0 while it stays <= t_end, the inlet at
getInletVelocity(total_runtime), an output after every iteration whose index is a multiple of the
output period, iteration 0 included (Simulation.cpp:589-591).  The iteration indices of the outputs,
the loop's runtime and iteration count, and the last output's density, velocity and surface rows
(x, y, z, WSS, OSI, TAWSS) must equal the oracle stepped the same way, bit for bit (FP64).
"""
import ctypes as C
import os

import numpy as np
import pytest

from pv_dec_original_amd import _native as N
from tests import oracle_ctypes as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "facade", "libfacade_iterate.so")
DT = 8e-6
T_END = 200 * DT


def _reference_times(t_end, dt):
    t, out = 0.0, []
    while t <= t_end:
        out.append(t)
        t += dt
    return out, t


def test_facade_iterate_matches_the_oracle_loop():
    from pv_dec_original_amd.workloads import tube_case
    N.lib()  # the engine library first (it brings up the HIP runtime the shim shares)
    shim = C.CDLL(SHIM)
    case = tube_case(q=19, tau=0.8)
    geo = case.geometry
    tri = np.ascontiguousarray(geo.globalTriangles, dtype=np.float64)
    faces = (N.Face * len(case.faces))(*case.faces)
    nz, ny, nx = geo.fluid.shape
    nbox = nx * ny * nz
    times, t_after = _reference_times(T_END, DT)
    # the oracle: next-to-wall count for the buffers, then the loop with outputs at 0, 100, 200
    orc = O.OracleSim(geo.fluid, geo.origin, geo.spacing, 19, case.faces, geo.globalTriangles, case.tau, DT)
    orc.set_inlet_profile_box(np.ones(geo.fluid.size))
    vel = np.array([O.load().orc_pulsatile_velocity(t, 0.9) for t in times])
    outs = list(range(0, len(times), 100))
    done = 0
    for k in outs:
        orc.step(k - done, vel[done:k])
        rho_o, u_o = orc.macros()  # iteration k's collision state
        orc.step(1, vel[k:k + 1])
        done = k + 1
    orc.step(len(times) - done, vel[done:])
    wss_o = orc.wss()[0]  # after the last iteration, which is the last output (200)
    assert outs[-1] == len(times) - 1
    nw = wss_o.shape[0]

    iters = np.zeros(16, dtype=np.int64)
    n_out = C.c_int(0)
    rho = np.zeros(nbox)
    u = np.zeros(3 * nbox)
    wss = np.zeros(6 * nw)
    runtime = C.c_double(0.0)
    nit = C.c_int64(0)
    err = C.create_string_buffer(512)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    shim.facade_iterate.restype = C.c_int
    rc = shim.facade_iterate(P(tri), C.c_int64(tri.shape[0]), C.c_double(geo.spacing), C.c_int(19), faces,
                             C.c_int(len(case.faces)), C.c_double(case.tau), C.c_double(DT), C.c_double(T_END),
                             C.c_double(0.9), C.c_int(100), P(iters), C.c_int(16), C.byref(n_out), P(rho), P(u),
                             P(wss), C.c_int64(nw), C.byref(runtime), C.byref(nit), err, C.c_int(512))
    assert rc == 0, err.value.decode()
    assert list(iters[:n_out.value]) == outs
    assert nit.value == len(times) and runtime.value == t_after
    fluid = geo.fluid > 0
    assert np.array_equal(rho.reshape(nz, ny, nx)[fluid], np.asarray(rho_o).reshape(nz, ny, nx)[fluid])
    uo = np.asarray(u_o).reshape(3, nz, ny, nx)
    ud = u.reshape(3, nz, ny, nx)
    for c in range(3):
        assert np.array_equal(ud[c][fluid], uo[c][fluid]), c
    assert np.array_equal(wss.reshape(nw, 6), wss_o)
    assert np.abs(ud).max() > 0 and np.isfinite(wss).all()
