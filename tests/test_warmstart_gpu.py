"""The learned warm start on the device: a start from given fields reaches the steady state of a
start from rest (the populations the reference never rewrites keep their rest values), and the
trained network shortens the run to converged wall shear stress."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VAL_SEED0 = 1_000_000


@pytest.fixture(scope="module")
def W():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from pv_dec_original_amd import warmstart
    return warmstart


def test_start_from_fields_keeps_the_steady_state(W):
    from pv_dec_original_amd import _native as N
    case, sd, sim, um, ta = W.vessel_setup(VAL_SEED0 + 1)
    fluid = sd.kinds() != N.SOLID
    rho, u, steps, change, status = W.run_to_steady(sim, fluid, tol=1e-7, max_steps=30000)
    assert status == "converged", (status, steps, change)
    sim.set_wss_every(1)
    sim.step(1)
    ref = sim.wss()[:, 3].copy()
    # restart from the converged macroscopic fields (equilibrium populations) and run on
    sim.init_fields(rho, u)
    sim.step(3000)
    w = sim.wss()[:, 3]
    err = np.linalg.norm(w - ref) / np.linalg.norm(ref)
    assert err < 2e-3, err


def test_coarse_start_matches_host_expansion(W):
    """pvlbm_init_fields_coarse against the host restatement of its sampling and clipping."""
    import torch
    from pv_dec_original_amd import _native as N
    case, sd, sim, um, ta = W.vessel_setup(VAL_SEED0 + 2)
    kinds = sd.kinds()
    fl = kinds != N.SOLID
    rng = np.random.default_rng(3)
    cshape = (40, 48, 56)
    rho_c = (1.0 + 0.05 * rng.standard_normal(cshape)).astype(np.float32)
    u_c = (0.05 * rng.standard_normal((3,) + cshape)).astype(np.float32)
    u_c[0, 20, 24, 28] = 2.0  # speeds around it are clipped
    rho_c[10, 30, 30] = 3.0  # densities around it are clipped
    rd, ud = torch.from_numpy(rho_c).cuda(), torch.from_numpy(u_c).cuda()
    assert sim.init_fields_coarse(rd, ud, W.UMAX, W.RHO_CLIP) == 0
    rho, u = sim.macros()
    rr = W._trilinear_up(rho_c.astype(np.float64), kinds.shape)
    uu = np.stack([W._trilinear_up(u_c[k].astype(np.float64), kinds.shape) for k in range(3)])
    rr, uu = W.clip_field(rr, uu)
    assert np.abs(rho[fl] - rr[fl]).max() < 1e-5
    assert np.abs(u[:, fl] - uu[:, fl]).max() < 1e-5
    assert np.sqrt((u[:, fl] ** 2).sum(axis=0)).max() <= W.UMAX * (1 + 1e-9)
    # non-finite samples start at rest and are counted
    rho_c[5, 5, 5] = np.nan
    rd = torch.from_numpy(rho_c).cuda()
    assert sim.init_fields_coarse(rd, ud, W.UMAX, W.RHO_CLIP) > 0
    rho, u = sim.macros()
    assert np.isfinite(rho[fl]).all() and np.isfinite(u[:, fl]).all()


def test_coarse_start_enforces_plane_flux(W):
    """With a target flux every x-plane of the started field carries it (inlet on x-min, outlets
    on x-max: mass conservation), before the solver takes a step."""
    import torch
    from pv_dec_original_amd import _native as N
    case, sd, sim, um, ta = W.vessel_setup(VAL_SEED0 + 3)
    kinds = sd.kinds()
    fl = kinds != N.SOLID
    rng = np.random.default_rng(4)
    cshape = (48, 48, 48)
    rho_c = np.ones(cshape, dtype=np.float32)
    u_c = (0.004 * rng.random((3,) + cshape)).astype(np.float32)
    u_c[0] += 0.01
    q = 0.02 * np.count_nonzero(kinds[:, :, 0] == N.KINLET)
    rd, ud = torch.from_numpy(rho_c).cuda(), torch.from_numpy(u_c).cuda()
    assert sim.init_fields_coarse(rd, ud, W.UMAX, W.RHO_CLIP, q) == 0
    rho, u = sim.macros()
    flux = np.array([u[0][:, :, x][fl[:, :, x]].sum() for x in range(kinds.shape[2])])
    has = np.array([fl[:, :, x].any() for x in range(kinds.shape[2])])
    assert np.allclose(flux[has], q, rtol=1e-9)


def test_trained_network_shortens_the_run_to_converged_wss(W):
    """The committed network against the same solver without it: iterations until the WSS of a
    validation vessel (seeds never trained on) is within 1% (L2) of its converged value, from
    rest and from the warm start, FP64 D3Q19, constant inlet."""
    from pv_dec_original_amd import _native as N
    warm = W.WarmStart(os.path.join(os.path.dirname(W.__file__), "weights", "warmstart_unet.npz"))
    rest_steps, warm_steps = [], []
    for seed in [VAL_SEED0 + i for i in (1, 2, 3, 4, 5, 6, 8)]:
        case, sd, sim, um, ta = W.vessel_setup(seed)
        fluid = sd.kinds() != N.SOLID
        ref, _, status = W.reference_wss(sim, fluid)
        assert status == "converged", seed
        sim.init_fields()
        s_rest, _, _ = W.steps_to_wss(sim, ref, ramp=0, timed=False)
        mode = W.start(sim, sd, warm, um, ta, sim.inlet_flux)
        s_warm, _, _ = W.steps_to_wss(sim, ref, ramp=0, timed=False)
        assert s_rest and s_warm, (seed, s_rest, s_warm)
        rest_steps.append(s_rest)
        warm_steps.append(s_warm if mode == "warm" else s_rest)
    ratio = sum(warm_steps) / sum(rest_steps)
    print(f"steps to 1% WSS: rest {rest_steps} warm {warm_steps} ratio {ratio:.3f}")
    assert ratio <= 0.8, (rest_steps, warm_steps)


def test_phase0_profiles_remove_the_acoustic_transient(W):
    """Pulsatile flow in a wide tube (R = 16, 512 long, tau 0.6, period 2000 steps: Womersley
    number 4.9, 0.44 acoustic wavelengths, the proportions of the large aneurysm).  Both starts
    take the converged steady field at the mean inlet velocity; imposing the phase-0 plane fluxes
    and densities of the one-dimensional model (phase0.phase0_profiles, through
    pvlbm_init_fields_coarse) must leave a several times smaller WSS error against the periodic
    state in the following periods than the steady field alone, whose density misses the
    oscillation's acoustic part and rings (measured: 0.068 / 0.021 against 0.011 / 0.0046 in
    periods 2 and 3)."""
    import torch
    from pv_dec_original_amd import converge as CV
    from pv_dec_original_amd import _native as N
    from pv_dec_original_amd.phase0 import phase0_profiles, plane_areas, plane_areas_from_lists
    from pv_dec_original_amd.simulation import SubDomain, Simulation
    from pv_dec_original_amd.workloads import tube_case
    T, se, mean, amp, tau = 2000, 100, 0.03, 0.02, 0.6
    case = tube_case(q=19, radius=16.0, length=512.0, tau=tau)
    sd = SubDomain(case.geometry, case.faces, storage="fp64", tau=tau, dt=8e-6)
    sim = Simulation(sd, tau, 1.0, 8e-6)
    prof = case.inlet_profile(sd, 2.0)
    prof = prof / prof.mean()
    sim.set_inlet_profile(prof)
    q = mean * float(prof.sum())
    sim.step(24000, np.full(24000, mean))
    rho_s, u_s = sim.macros_device()
    ref = CV.run_periods(sim, "ref", T, se, mean, amp, 30, stop_change=2e-4)
    assert ref.history[-1]["change"] < 2e-4
    rec = ref.records[-1]
    kinds = sd.kinds()
    fl = kinds != 0
    rpl = np.where(fl, rho_s.cpu().numpy(), 0.0).sum(axis=(0, 1)) / np.maximum(fl.sum(axis=(0, 1)), 1)
    area = plane_areas(kinds)
    assert np.array_equal(area, plane_areas_from_lists(kinds.shape[2], sd.list(N.LIST_FLUID), sd.list(N.LIST_WALL)))
    targets, drho = phase0_profiles(area, q, q * amp / mean, T, tau, rpl)
    err = {}
    for name, prof_args in (("steady", (q, None, None)), ("phase0", (0.0, targets, drho))):
        def start(a=prof_args):
            assert sim.init_fields_coarse(rho_s, u_s, W.UMAX, W.RHO_CLIP, *a) == 0
            return name
        r = CV.run_periods(sim, name, T, se, mean, amp, 3, stop_change=0.0, reference=rec, start=start)
        err[name] = [h["error"] for h in r.history]
    print(f"WSS error per period against the periodic state: steady start {np.round(err['steady'], 4)}, "
          f"with phase-0 profiles {np.round(err['phase0'], 4)}")
    assert err["phase0"][0] < err["steady"][0]
    assert err["phase0"][1] < 0.35 * err["steady"][1] and err["phase0"][2] < 0.5 * err["steady"][2]
    torch.cuda.synchronize()


def test_coarse_code_gathered_on_the_device(W):
    """coarse_code_subdomain (device gather of the coarse voxel centres, faces from the lists)
    equals coarse_code over the host copy of the whole box."""
    from pv_dec_original_amd.simulation import SubDomain
    from pv_dec_original_amd.workloads import vessel_case
    case = vessel_case(VAL_SEED0 + 3)
    sd = SubDomain(case.geometry, case.faces, storage="fp16", tau=case.tau)
    kinds = sd.kinds()
    for cshape in ((48, 48, 48), (40, 56, 64), kinds.shape):
        assert np.array_equal(W.coarse_code_subdomain(sd, cshape), W.coarse_code(kinds, cshape)), cshape
