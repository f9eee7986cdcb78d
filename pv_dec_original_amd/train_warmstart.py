"""Trains the warm-start network on solver-generated vessels (one process, one device).

    python -m pv_dec_original_amd.train_warmstart --samples 4096 --out DIR [--minutes 14]

Worker threads solve seeded 96^3 vessels to steady state with the FP64 D3Q19 engine
(warmstart.solve_sample) while the main thread trains on the samples produced so far; the data
stays resident in device memory (4096 samples: 3.6 GB of codes, 29 GB of bf16 targets).  A fixed
set of validation seeds (disjoint from the training seeds) is solved first.  Writes
DIR/warmstart_unet.npz (levels, FP32 parameters, training metadata) and DIR/train_log.json.
"""
import argparse
import json
import os
import queue
import threading
import time

import numpy as np

VAL_SEED0 = 1_000_000
LOSS_WEIGHTS = (1.0, 1.0, 1.0, 0.25)  # u_x, u_y, u_z, rho: the density channel is heavy-tailed


def augment(code, tgt, k):
    """One of the 8 symmetries of the problem (inlet on x-min, outlet on x-max): bit 0 mirrors y,
    bit 1 mirrors z, bit 2 swaps y and z; velocity components follow.  code (B, z, y, x),
    tgt (B, z, y, x, 4)."""
    if k & 1:
        code, tgt = code.flip(2), tgt.flip(2)
        tgt = tgt * tgt.new_tensor([1.0, -1.0, 1.0, 1.0])
    if k & 2:
        code, tgt = code.flip(1), tgt.flip(1)
        tgt = tgt * tgt.new_tensor([1.0, 1.0, -1.0, 1.0])
    if k & 4:
        code, tgt = code.transpose(1, 2), tgt.transpose(1, 2)[..., [0, 2, 1, 3]]
    return code.contiguous(), tgt.contiguous()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=4096)
    ap.add_argument("--val", type=int, default=32)
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--workers", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=200, help="steps of linear learning-rate warmup")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=2e-5)
    ap.add_argument("--minutes", type=float, default=14.0, help="wall-clock budget of the whole run")
    ap.add_argument("--min-samples", type=int, default=64, help="samples before training starts")
    ap.add_argument("--levels", default="16,32,64,160")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--augment", type=int, default=1, help="random symmetry per minibatch")
    ap.add_argument("--init", default=None, help="continue from these weights (npz of save_weights)")
    ap.add_argument("--seed0", type=int, default=0, help="first training seed (later runs take new vessels)")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    import torch
    from . import warmstart as W
    from .warmnet import UNet, UNetSpec

    t_start = time.perf_counter()
    deadline = t_start + 60.0 * a.minutes
    os.makedirs(a.out, exist_ok=True)

    def log(msg):
        print(f"[{time.perf_counter() - t_start:7.1f}s] {msg}", flush=True)

    n = a.n
    dev = torch.device("cuda")
    codes = torch.zeros((a.samples, n, n, n), dtype=torch.uint8, device=dev)
    targets = torch.zeros((a.samples, n * n * n, 4), dtype=torch.bfloat16, device=dev)
    prms = torch.zeros((a.samples, 3), dtype=torch.float32, device=dev)

    # validation set first (disjoint seeds)
    val = W.generate(list(range(VAL_SEED0, VAL_SEED0 + a.val)), n, workers=a.workers)
    vc = torch.from_numpy(np.stack([s.code for s in val])).to(dev)
    vt = torch.from_numpy(np.stack([s.target for s in val]).reshape(a.val, -1, 4)).to(dev).to(torch.bfloat16)
    vp = torch.tensor([[s.u_mean, s.tau, s.inlet_area] for s in val], dtype=torch.float32, device=dev)
    log(f"validation set: {a.val} samples, statuses {sorted(set(s.status for s in val))}")

    # producer: training samples in seed order, solved by worker threads
    q = queue.Queue(maxsize=4 * a.workers)
    stop = threading.Event()

    def produce():
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=a.workers) as ex:
            futs = []
            for seed in range(a.seed0, a.seed0 + a.samples):
                if stop.is_set():
                    break
                futs.append(ex.submit(W.solve_sample, seed, n))
                # keep a bounded window of work in flight
                while len(futs) >= 2 * a.workers:
                    q.put(futs.pop(0).result())
            for f in futs:
                q.put(f.result())
        q.put(None)

    prod = threading.Thread(target=produce, daemon=True)
    prod.start()

    if a.init:
        levels, params = W.load_weights(a.init)
        spec = UNetSpec(levels=levels)
        net = UNet(spec, params)
    else:
        spec = UNetSpec(levels=tuple(int(v) for v in a.levels.split(",")))
        net = UNet(spec, seed=a.seed)
    log(f"network: {spec.nparams} parameters, levels {spec.levels}")

    have = 0
    statuses = {}
    gen_steps = []
    gen_done_at = None
    rng = np.random.default_rng(a.seed)
    hist = []
    step = 0
    t_train0 = None
    producing = True
    while True:
        # drain finished samples into device memory
        while producing:
            try:
                s = q.get_nowait() if have >= a.min_samples else q.get(timeout=1.0)
            except queue.Empty:
                break
            if s is None:
                producing = False
                gen_done_at = time.perf_counter() - t_start
                log(f"generation done: {have} samples, statuses {statuses}")
                break
            statuses[s.status] = statuses.get(s.status, 0) + 1
            if s.status == "diverged":
                continue
            codes[have].copy_(torch.from_numpy(s.code))
            targets[have].copy_(torch.from_numpy(s.target.reshape(-1, 4)).to(torch.bfloat16))
            prms[have] = torch.tensor([s.u_mean, s.tau, s.inlet_area], dtype=torch.float32)
            gen_steps.append(s.steps)
            have += 1
        if have < a.min_samples:
            continue
        now = time.perf_counter()
        if now > deadline - 30.0:
            break
        if t_train0 is None:
            t_train0 = now
        idx = torch.from_numpy(rng.integers(0, have, a.batch)).to(dev)
        code, tgt = augment(codes.index_select(0, idx), targets.index_select(0, idx).view(a.batch, n, n, n, 4),
                            int(rng.integers(0, 8)) if a.augment else 0)
        tgt = tgt.reshape(-1, 4)
        prm = prms.index_select(0, idx)
        # learning rate: linear warmup, then cosine decay over the remaining wall-clock budget
        frac = min(1.0, (now - t_train0) / max(deadline - 30.0 - t_train0, 1.0))
        lr = a.lr_end + 0.5 * (a.lr - a.lr_end) * (1 + np.cos(np.pi * frac))
        if step < a.warmup:
            lr *= (step + 1) / a.warmup
        net.forward(code, prm)
        loss = net.loss(tgt, code, LOSS_WEIGHTS)
        net.backward()
        net.adam_step(float(lr))
        step += 1
        if step % 100 == 0:
            lv = float(loss[0])
            if not np.isfinite(lv):
                log(f"non-finite loss at step {step}")
                break
            hist.append({"step": step, "loss": lv, "lr": lr, "samples": have,
                         "t": time.perf_counter() - t_start})
            if step % 1000 == 0:
                vm, vr = W.evaluate(net, vc, vt, vp, a.batch)
                k = min(have, 16) // a.batch * a.batch
                tm, tr = W.evaluate(net, codes[:k], targets[:k], prms[:k], a.batch)
                hist[-1].update({"val_mse": vm, "val_rel_u": vr, "train_mse": tm, "train_rel_u": tr})
                log(f"step {step} loss {lv:.4e} val mse {vm:.4e} val rel L2(u) {vr:.3f} "
                    f"train rel L2(u) {tr:.3f} lr {lr:.2e} samples {have}")
    # stop the producer: it finishes the solves in flight and ends the queue with None
    stop.set()
    while producing:
        if q.get() is None:
            producing = False
    prod.join()
    torch.cuda.synchronize()
    vm, vr = W.evaluate(net, vc, vt, vp, a.batch)
    train_s = time.perf_counter() - (t_train0 or time.perf_counter())
    prev = {}
    if a.init:
        with np.load(a.init, allow_pickle=False) as z:
            prev = {k: float(z[k]) for k in ("total_steps", "total_samples") if k in z}
    meta = {"samples": have, "val_samples": a.val, "steps": step, "batch": a.batch, "val_mse": vm,
            "val_rel_u": vr, "train_seconds": train_s, "seed": a.seed, "seed0": a.seed0,
            "total_steps": prev.get("total_steps", 0) + step,
            "total_samples": prev.get("total_samples", 0) + have}
    W.save_weights(os.path.join(a.out, "warmstart_unet.npz"), net, meta)
    rec = {"args": vars(a), "meta": meta, "history": hist, "statuses": statuses,
           "gen_done_at_s": gen_done_at, "gen_steps_mean": float(np.mean(gen_steps)) if gen_steps else None,
           "wall_s": time.perf_counter() - t_start}
    with open(os.path.join(a.out, "train_log.json"), "w") as f:
        json.dump(rec, f, indent=1)
    log(f"done: {step} steps on {have} samples, val mse {vm:.4e}, val rel L2(u) {vr:.3f}")


if __name__ == "__main__":
    main()
