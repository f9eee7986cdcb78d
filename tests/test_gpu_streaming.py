"""The two streaming schemes of the step give the same fields bit for bit.

In place (one lattice, even launches on each node's own slots and odd launches on its
neighbours' slots, links without an upstream node parked where the next launch reads them) against
the pull from one lattice into another (pvlbm_set_streaming), over odd and even numbers of
launches, with WSS and the OSI/TAWSS accumulators every iteration, for both lattices and the three
storages; and across changes of scheme in the middle of a run from either kind of in-place launch,
which move the state between the layouts.  By default a lattice of at most 512 MiB streams in place
and a larger one takes the pull, so the small device parity tests against the oracle (push
streaming, Simulation.cpp:211-237) run in place and the full-size FP64 aneurysm with two lattices.
"""
import numpy as np
import pytest

from pv_dec_original_amd import _native as N
from tests import gpu_cases as G

pytestmark = pytest.mark.gpu


def _state(sim):
    rho, u = sim.macros()
    tau3, acc = sim.wss_raw()
    return rho, u, sim.wss(), tau3, acc


def _same(a, b):
    # OSI is 0/0 (NaN) where a node has seen no shear
    return all(np.array_equal(x, y, equal_nan=True) for x, y in zip(a, b))


@pytest.mark.parametrize("q,storage", [(19, "fp64"), (27, "fp64"), (19, "fp16"), (27, "fp16"), (19, "bf16"),
                                       (27, "bf16")])
def test_in_place_equals_two_lattices(q, storage):
    case = G.small_ybif_case(q=q)
    sd_a, sim_a, _ = G.build(case, storage, normals=False)
    sd_b, sim_b, _ = G.build(case, storage, normals=False)
    assert sim_a.in_place and sim_b.in_place
    sim_b.set_streaming(N.STREAMING_TWO_LATTICES)
    assert not sim_b.in_place
    rng = np.random.default_rng(q)
    sc = rng.uniform(0.6, 1.0, 80)
    done = 0
    for n in (37, 43):  # ends on an odd, then on an even launch
        sim_a.step(n, sc[done:done + n])
        sim_b.step(n, sc[done:done + n])
        done += n
        assert _same(_state(sim_a), _state(sim_b)), (q, storage, done)
    assert np.abs(sim_a.macros()[1]).max() > 1e-3
    if storage == "fp64":
        assert sim_a.gate_events() == 0 and sim_b.gate_events() == 0


def test_scheme_changes_mid_run():
    """Changes of scheme after odd and even launches, and the interpolated closure's switch to the
    pull and back, leave the run identical to one that never changed."""
    case = G.small_ybif_case(q=27)
    _, ref, _ = G.build(case, normals=False)
    _, sim, _ = G.build(case, normals=False)
    sc = np.random.default_rng(5).uniform(0.6, 1.0, 60)
    plan = [(13, N.STREAMING_TWO_LATTICES), (10, N.STREAMING_AUTO), (12, N.STREAMING_TWO_LATTICES),
            (5, N.STREAMING_AUTO), (3, None)]
    done = 0
    for n, mode in plan:
        ref.step(n, sc[done:done + n])
        sim.step(n, sc[done:done + n])
        done += n
        assert _same(_state(ref), _state(sim)), done
        if mode is not None:
            sim.set_streaming(mode)
            assert sim.in_place == (mode == N.STREAMING_AUTO)
    # the interpolated closure takes the pull; switching back returns to the in-place scheme
    sim.set_wall_closure(N.WALL_INTERPOLATED)
    assert not sim.in_place
    sim.set_wall_closure(N.WALL_REFERENCE)
    assert sim.in_place
    ref.step(7, sc[done:done + 7])
    sim.step(7, sc[done:done + 7])
    assert _same(_state(ref), _state(sim))


def test_warm_start_in_both_schemes():
    """A start from given fields parks the rest weights in the links without an upstream node in
    either layout: the runs that follow are identical."""
    from pv_dec_original_amd.workloads import tube_case
    case = tube_case(q=19)
    _, sim_a, _ = G.build(case, normals=False)
    _, sim_b, _ = G.build(case, normals=False)
    sim_a.step(120)
    rho, u = sim_a.macros()
    sim_b.set_streaming(N.STREAMING_TWO_LATTICES)
    for s in (sim_a, sim_b):
        s.init_fields(rho, u)
        s.step(51)
    assert _same(_state(sim_a), _state(sim_b))


def test_automatic_scheme_by_lattice_size_and_forced_modes():
    """A 128 x 128 x 224 all-fluid box, D3Q19: its FP64 lattice (558 MB) exceeds 512 MiB and takes
    the pull by default, its FP16 lattice (139 MB) streams in place; forcing in place on the FP64
    box and running on gives the fields of a run that stayed on two lattices.  The interpolated
    wall closure cannot be combined with forced in-place streaming."""
    from pv_dec_original_amd.simulation import Simulation, SubDomain
    from pv_dec_original_amd.workloads import fluid_box_case
    case = fluid_box_case(q=19, box=(128, 128, 224), tau=0.7)
    n = 128 * 128 * 224
    assert 19 * 8 * n > (512 << 20) >= 19 * 2 * n

    def make(storage):
        sd = SubDomain(case.geometry, case.faces, storage=storage, tau=case.tau, dt=G.DT, triangles=False)
        sim = Simulation(sd, case.tau, 1.0, G.DT)
        sim.set_inlet_profile(case.inlet_profile(sd, 0.05))
        return sd, sim

    _, h16 = make("fp16")
    assert h16.in_place
    del h16
    _, a = make("fp64")
    _, b = make("fp64")
    assert not a.in_place and not b.in_place
    a.step(10)
    b.step(10)
    a.set_streaming(N.STREAMING_IN_PLACE)
    assert a.in_place
    a.step(11)
    b.step(11)
    assert _same(_state(a), _state(b))
    with pytest.raises(N.NativeError, match="two lattices"):
        a.set_wall_closure(N.WALL_INTERPOLATED)
    assert a.in_place
