"""The learned form at the sizes its configurations name.

Training (seeded tortuous vessels in a 96^3 box, 3D conv U-Net of about 2 M parameters, BF16 MFMA
training with FP32 accumulation): one batch of 8 solver-generated vessels at 96^3 through the
full-size network (levels 16, 32, 64, 160; 1.82 M parameters) for 20 Adam steps, and the
16-channel weight gradient at that size against a numpy sum over sampled taps.

Aneurysm (1024 x 512 x 512, D3Q19 FP16 shifted storage, pulsatile inlet, learned warm start, WSS
every 200 steps until the period-to-period change is below 1%): the run of
pv_dec_original_amd/learned_bench.py with its criteria asserted -- within 1% (L2) of the FP64
reference at convergence, at least 5x less wall time than plain FP64 from rest, and fewer periods
than the same start without the network (plug flow with the same phase-0 profiles).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def test_training_at_full_size(T):
    from pv_dec_original_amd import warmstart as W
    from pv_dec_original_amd.warmnet import UNet, UNetSpec
    seeds = list(range(910_000, 910_008))
    samples = []
    for s in seeds:
        smp = W.solve_sample(s, n=96, max_steps=3000)
        samples.append(smp)
        print(f"vessel {s}: {smp.steps} steps, {smp.status}, {smp.seconds:.1f} s", flush=True)
    spec = UNetSpec(levels=(16, 32, 64, 160))
    assert 1.7e6 < spec.nparams < 2.1e6
    net = UNet(spec, spec.init_params(3))
    code = T.from_numpy(np.stack([s.code for s in samples])).to("cuda")
    prm = T.tensor([[s.u_mean, s.tau, s.inlet_area] for s in samples], dtype=T.float32, device="cuda")
    tgt = T.from_numpy(np.stack([s.target for s in samples]).reshape(-1, 4)).to("cuda").to(T.bfloat16)
    assert code.shape == (8, 96, 96, 96)
    losses = []
    for it in range(20):
        net.forward(code, prm)
        losses.append(float(net.loss(tgt, code)[0]))
        net.backward()
        g = net.g
        assert bool(T.isfinite(g).all()), f"non-finite gradient at step {it}"
        if it == 0:
            assert float(g.abs().max()) > 0
        net.adam_step(1e-3)
    T.cuda.synchronize()
    print("losses", [round(v, 4) for v in losses[::4]], flush=True)
    assert np.isfinite(losses).all()
    assert min(losses[-3:]) < 0.8 * losses[0], losses


def test_wgrad16_at_full_size_against_sampled_taps(T):
    """The 16-channel weight gradient (cin = cout = 16, 27 taps) over 8 x 96^3 voxels: selected
    entries dw[n][tap * cin + ci] = sum over voxels m of dy[m][n] x[m + off(tap)][ci] (zero outside
    the box) against numpy in FP64 on the same bf16 operands."""
    from pv_dec_original_amd import _native as N
    NL = N.net()
    shape = (8, 96, 96, 96)
    nb, d, h, w = shape
    cin = cout = 16
    M = nb * d * h * w
    rng = np.random.default_rng(21)
    x = T.randn((M, cin), dtype=T.float32, device="cuda").to(T.bfloat16)
    dy = T.randn((M, cout), dtype=T.float32, device="cuda").to(T.bfloat16)
    ws = T.zeros(int(NL.pvnet_wgrad_ws_floats(C.byref(N.NetShape(*shape)), 27, cin, cout)), dtype=T.float32,
                 device="cuda")
    dw = T.zeros((cout, 27 * cin), dtype=T.float32, device="cuda")
    db = T.zeros(cout, dtype=T.float32, device="cuda")
    g = N.NetWgrad(N.NetShape(*shape), 27, cin, cout, C.c_void_p(x.data_ptr()), cin, cout,
                   C.c_void_p(dy.data_ptr()), C.c_void_p(ws.data_ptr()), C.c_void_p(dw.data_ptr()),
                   C.c_void_p(db.data_ptr()))
    N.net_check(NL.pvnet_conv_wgrad(C.byref(g), C.c_void_p(T.cuda.current_stream().cuda_stream)))
    T.cuda.synchronize()
    X = x.float().cpu().numpy().astype(np.float64).reshape(nb, d, h, w, cin)
    DY = dy.float().cpu().numpy().astype(np.float64).reshape(nb, d, h, w, cout)
    got = dw.cpu().numpy()
    taps = rng.choice(27, 6, replace=False)
    for tap in taps:
        oz, oy, ox = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        # x shifted by the tap's offset, zero where the neighbour is outside the box
        xs = np.zeros_like(X)
        dst = tuple(slice(max(0, -o), L - max(0, o)) for o, L in ((oz, d), (oy, h), (ox, w)))
        src = tuple(slice(max(0, o), L + min(0, o)) for o, L in ((oz, d), (oy, h), (ox, w)))
        xs[(slice(None),) + dst] = X[(slice(None),) + src]
        for n in rng.choice(cout, 2, replace=False):
            ref = np.einsum("bzyx,bzyxc->c", DY[..., n], xs)
            sel = got[n, tap * cin:(tap + 1) * cin]
            sc = np.abs(ref).max()
            assert np.allclose(sel, ref, rtol=1e-5, atol=1e-5 * sc + 1e-3), (tap, n, np.abs(sel - ref).max(), sc)
    assert np.allclose(db.cpu().numpy(), DY.sum(axis=(0, 1, 2, 3)), rtol=1e-5,
                       atol=1e-5 * np.abs(DY.sum(axis=(0, 1, 2, 3))).max() + 1e-3)


def test_fp16_learned_start_on_the_aneurysm(T):
    from pv_dec_original_amd import learned_bench
    out = learned_bench.run(log=lambda m: print(m, flush=True))
    runs = out["runs"]
    warm, plug, plain = runs["fp16-warm"], runs["fp16-plug"], runs["fp64-rest"]
    print({k: v for k, v in out.items() if k != "runs"}, flush=True)
    assert warm["start"] == "warm" and plug["start"] == "warm"
    assert plain["converged_period"] is not None and warm["converged_period"] is not None
    assert warm["error_at_convergence"] <= 0.01
    assert out["speedup_wall"] >= 5.0
    assert plug["converged_period"] is None or warm["converged_period"] < plug["converged_period"]
