"""Error of the shifted 16-bit storage forms against FP64, per sample, on the Y-bifurcation.

    python -m pv_dec_original_amd.precision_report [--lattice 19 27] [--steps 12000]
                                                   [--every 1000] [--json OUT]

The D3Q19 or D3Q27 Y-bifurcation (512 x 256 x 256, tau 0.6, pulsatile inlet
0.03 + 0.02 sin(2 pi t / 4000) times a parabolic profile of unit mean) runs in lockstep with FP64,
FP16-shifted and BF16-shifted storage (FP32 collision for both 16-bit forms, FP64 global sums).
Every `every` steps the fluid-node fields are compared with the FP64 run: relative L2 and
relative L-infinity (max abs error / max abs FP64 value) of the velocity vector, of rho - 1 and of
the WSS magnitude at the next-to-wall nodes (WSS sampled on the same iterations).  MLUPS per
storage from the launches' HIP-event time.  Prints one JSON line.
"""
import argparse
import json
import sys

import numpy as np


def _err(a, b):
    d = a - b
    nb = max(np.linalg.norm(b), 1e-300)
    return {"l2": float(np.linalg.norm(d) / nb), "linf": float(np.abs(d).max() / max(np.abs(b).max(), 1e-300))}


def run(q, steps=12000, every=1000, box=(512, 256, 256), tau=0.6, log=None):
    from . import _native as N
    from .converge import pulsatile_scales
    from .simulation import SubDomain, Simulation
    from .workloads import ybif_case

    case = ybif_case(q=q, box=box, tau=tau)
    sims = {}
    for st in ("fp64", "fp16", "bf16"):
        sd = SubDomain(case.geometry, case.faces, storage=st, tau=tau, dt=8e-6)
        sim = Simulation(sd, tau, 1.0, 8e-6)
        prof = case.inlet_profile(sd, 2.0)
        sim.set_inlet_profile(prof / prof.mean())
        sims[st] = (sd, sim)
    sd64 = sims["fp64"][0]
    fluid = sd64.kinds() != N.SOLID
    nfl = int(fluid.sum())
    ms = {st: 0.0 for st in sims}
    samples = []
    it = 0
    while it < steps:
        n = min(every, steps - it)
        sc = pulsatile_scales(it, n, 4000, 0.03, 0.02)
        for st, (sd, sim) in sims.items():
            # WSS evaluated on the last iteration of the chunk only
            sim.set_wss_every(1 << 30)
            ms[st] += sim.step(n - 1, sc[:-1], timed=True) if n > 1 else 0.0
            sim.set_wss_every(1)
            ms[st] += sim.step(1, sc[-1:], timed=True)
        it += n
        rho64, u64 = sims["fp64"][1].macros()
        w64 = sims["fp64"][1].wss()[:, 3]
        rec = {"step": it}
        for st in ("fp16", "bf16"):
            rho, u = sims[st][1].macros()
            w = sims[st][1].wss()[:, 3]
            rec[st] = {"u": _err(u[:, fluid].T.ravel(), u64[:, fluid].T.ravel()),
                       "rho": _err(rho[fluid] - 1.0, rho64[fluid] - 1.0),
                       "wss": _err(w, w64)}
        samples.append(rec)
        if log:
            log(f"D3Q{q} step {it}: " + " ".join(
                f"{st} u {rec[st]['u']['l2']:.2e} rho {rec[st]['rho']['l2']:.2e} wss {rec[st]['wss']['l2']:.2e}"
                for st in ("fp16", "bf16")))
    mlups = {st: nfl * steps / (ms[st] / 1e3) / 1e6 for st in sims}
    return {"workload": f"ybif-D3Q{q}-{box[0]}x{box[1]}x{box[2]}", "fluid_nodes": nfl, "steps": steps,
            "every": every, "tau": tau, "mlups": {k: round(v, 1) for k, v in mlups.items()}, "samples": samples}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lattice", type=int, nargs="+", default=[19, 27])
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    out = [run(q, a.steps, a.every, log=lambda m: print(m, file=sys.stderr, flush=True)) for q in a.lattice]
    s = json.dumps(out)
    if a.json:
        with open(a.json, "w") as f:
            f.write(s + "\n")
    print(s)


if __name__ == "__main__":
    main()
