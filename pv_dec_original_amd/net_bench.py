"""Throughput of the warm-start network's training step and inference on one device.

    python -m pv_dec_original_amd.net_bench [--batch 8] [--n 96] [--steps 20]

One training step = features, forward, masked loss, backward (input and weight gradients of every
convolution), Adam and weight packing, on a batch of synthetic vessel-like boxes.  Convolution
FLOPs are counted exactly: 2 * voxels * (ntaps * cin) * cout per convolution for the forward, the
same for the input gradient (except the first layer, whose input needs none) and for the weight
gradient.  Prints one JSON line: ms per step, convolution TFLOP/s over the step, and the same
for inference alone, against the 2.5 PFLOP/s dense BF16 peak of the MI355X.
"""
import argparse
import json

import numpy as np

BF16_PEAK_TFLOPS = 2500.0


def conv_flops(spec, shape, training=True):
    nb, d, h, w = shape
    total = 0
    for lay in spec.layers:
        n = lay.name
        lvl = 0 if n == "head" else (len(spec.levels) - 1 if n.startswith("bt") else int(n[1:-1]))
        M = nb * (d >> lvl) * (h >> lvl) * (w >> lvl)
        f = 2 * M * lay.K * lay.cout
        total += f
        if training:
            total += f  # weight gradient
            if n != "e0a":
                total += f  # input gradient
    return total


def synthetic_batch(nb, n, rng):
    """Tubes of random radius along x through the box (codes 0/1/2/3), parameters in range."""
    z, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    codes = np.zeros((nb, n, n, n), dtype=np.uint8)
    prm = np.zeros((nb, 3), dtype=np.float32)
    for b in range(nb):
        r = rng.uniform(6, 16)
        cy, cz = rng.uniform(r + 2, n - r - 2, 2)
        disk = (y - cy) ** 2 + (z - cz) ** 2 < r * r
        codes[b][:, :, :][np.broadcast_to(disk[:, :, None], (n, n, n))] = 1
        codes[b][:, :, 0][disk] = 2
        codes[b][:, :, -1][disk] = 3
        prm[b] = (rng.uniform(0.01, 0.05), rng.uniform(0.55, 0.9), disk.sum())
    return codes, prm


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    a = ap.parse_args(argv)
    import torch
    from .warmnet import UNet, UNetSpec

    rng = np.random.default_rng(0)
    spec = UNetSpec()
    net = UNet(spec, seed=0)
    codes, prm = synthetic_batch(a.batch, a.n, rng)
    cd = torch.from_numpy(codes).cuda()
    pd = torch.from_numpy(prm).cuda()
    tgt = torch.randn((a.batch * a.n ** 3, 4), device="cuda").to(torch.bfloat16)
    shape = (a.batch, a.n, a.n, a.n)

    def step():
        net.forward(cd, pd)
        net.loss(tgt, cd)
        net.backward()
        net.adam_step(1e-4)

    for _ in range(a.warmup):
        step()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(a.steps):
        step()
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    # inference alone (forward of one box per sample)
    for _ in range(3):
        net.forward(cd, pd)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(a.steps):
        net.forward(cd, pd)
    e1.record(st)
    torch.cuda.synchronize()
    ms_inf = e0.elapsed_time(e1) / a.steps
    ft = conv_flops(spec, shape, True)
    fi = conv_flops(spec, shape, False)
    out = {
        "network": {"levels": spec.levels, "params": spec.nparams},
        "batch": list(shape), "train_ms_per_step": round(ms, 3),
        "train_conv_tflop_per_step": round(ft / 1e12, 4),
        "train_conv_tflops": round(ft / (ms * 1e-3) / 1e12, 2),
        "infer_ms": round(ms_inf, 3), "infer_conv_tflops": round(fi / (ms_inf * 1e-3) / 1e12, 2),
        "peak_bf16_tflops": BF16_PEAK_TFLOPS,
        "train_frac_of_peak": round(ft / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS, 4),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
