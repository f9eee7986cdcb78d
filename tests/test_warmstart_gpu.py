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
