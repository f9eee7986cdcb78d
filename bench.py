"""Benchmark: MLUPS and HBM-roofline fraction of the fused lattice Boltzmann step on one MI355X.

Default workload: the D3Q27 FP64 Y-bifurcation (parent R = 24, daughters R = 17 at +-30 deg) in a
512 x 256 x 256 box, tau = 0.6, pulsatile inlet u(t) = 0.03 + 0.02 sin(2 pi t / 4000) times a
parabolic profile of unit mean, pressure outlet, WSS every 100 iterations.  A "step" is one
iteration of the reference loop body over the whole domain.  The default K = 12000 steps is three
inlet periods.

Output: one JSON line.  value = fluid-node updates per second / 1e6 (MLUPS) over the timed steps,
inputs resident in HBM.  roofline.achieved = fluid nodes * 2 * Q * 8 bytes / average launch time.
reference_cadence = the same handle timed with WSS and its OSI/TAWSS accumulators evaluated every
iteration, as the reference does (Simulation.cpp:585-586); the workload's own cadence (WSS every
100 iterations) is the headline.  roofline.copy_kernel_GBs = a 4-deep 16-byte nontemporal copy,
the attainable HBM rate of a plain stream.
cpu_baseline = the CPU oracle (a restatement of the reference solver, push streaming) on a bounded
number of steps of the same workload, single-threaded like one reference MPI rank;
cpu_baseline_threads = the same over every host core the process may use (OpenMP; capped by an
OMP_NUM_THREADS the environment sets as the host's CPU share).
--workload box / aneurysm run the all-fluid box (the kernel's ceiling) or the aneurysm instead.

"learned" (on by default, --learned 0 to skip): wall time to converged wall shear stress on the
1024 x 512 x 512 aneurysm, the plain FP64 path from rest against FP16 shifted storage started from
the learned warm start with the phase-0 profiles of the pulsatile model
(pv_dec_original_amd/learned_bench.py).
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(workload):
    """HBM bytes per launch of the step kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of the same workload (calibrated against a copy kernel), or None."""
    name = {"ybif-D3Q27-fp64-512x256x256": "k_step_ybif512_q27_fp64.json",
            "ybif-D3Q19-fp16-512x256x256": "k_step_ybif512_q19_fp16.json",
            "aneurysm-D3Q19-fp16-1024x512x512": "k_step_aneurysm1024_q19_fp16.json"}.get(workload)
    if not name:
        return None
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fp:
        return round(json.load(fp)["traffic_bytes"])


def inlet_scales(it0, n, period, mean, amp):
    t = np.arange(it0, it0 + n, dtype=np.float64)
    return mean + amp * np.sin(2.0 * np.pi * t / period)


def cpu_baseline(case, prof_box, q, tau, budget_s, threads):
    """Oracle steps on the host cores: returns (MLUPS, steps, seconds)."""
    from tests import oracle_ctypes as O
    geo = case.geometry
    orc = O.OracleSim(geo.fluid, geo.origin, geo.spacing, q, case.faces, geo.globalTriangles, tau,
                      8e-6, normals=False, threads=threads)
    orc.set_inlet_profile_box(prof_box)
    nfl = orc.counts()["fluid"]
    t0 = time.perf_counter()
    orc.step(2, np.full(2, 0.03), wss_every=100)
    dt2 = time.perf_counter() - t0
    n = max(2, min(200, int(budget_s / max(dt2 / 2, 1e-6))))
    t0 = time.perf_counter()
    orc.step(n, np.full(n, 0.03), wss_every=100)
    el = time.perf_counter() - t0
    return nfl * n / el / 1e6, n, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--lattice", type=int, default=27)
    ap.add_argument("--storage", default="fp64", choices=["fp64", "fp16", "bf16"])
    ap.add_argument("--box", default=None, help="nx,ny,nz (default 512,256,256; 1024,512,512 for aneurysm)")
    ap.add_argument("--workload", default="ybif", choices=["ybif", "box", "aneurysm"],
                    help="ybif: the Y-bifurcation; box: an all-fluid box (kernel ceiling); "
                         "aneurysm: the curved tube with a saccular bulge of the learned-start case")
    ap.add_argument("--wss-every", type=int, default=100)
    ap.add_argument("--streaming", default="auto", choices=["auto", "in_place", "two_lattices"],
                    help="in place on one lattice or the pull from one lattice into another (default: the "
                         "library's choice by lattice size)")
    ap.add_argument("--ref-cadence-steps", type=int, default=2000,
                    help="also time this many steps with WSS every iteration, the reference's cadence "
                         "(0 = skip)")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30,
                    help="size of the calibration copy kernel (0 = skip)")
    ap.add_argument("--learned", type=int, default=1,
                    help="also measure wall time to converged WSS on the large aneurysm: plain FP64 "
                         "from rest against FP16 storage from the learned warm start (1 = on)")
    ap.add_argument("--learned-ablations", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the single-device engine has no multi-GPU path).", file=sys.stderr)
        sys.exit(2)

    from pv_dec_original_amd import SubDomain, Simulation
    from pv_dec_original_amd import _native as N
    from pv_dec_original_amd.workloads import aneurysm_case, fluid_box_case, inlet_disk_profile, ybif_case

    if N.device_count() < 1:
        raise SystemExit("no HIP device visible")
    box = tuple(int(v) for v in (args.box or ("1024,512,512" if args.workload == "aneurysm" else "512,256,256")).split(","))
    q = args.lattice
    tau = 0.6
    t0 = time.perf_counter()
    if args.workload == "ybif":
        case = ybif_case(q=q, box=box, tau=tau)
    elif args.workload == "aneurysm":
        case = aneurysm_case(q=q, box=box, tau=tau)
    else:
        case = fluid_box_case(q, box, tau)
    t_geo = time.perf_counter() - t0
    vox_tests, vox_ms = C.c_int64(0), C.c_double(0.0)
    N.lib().pvlbm_voxelize_stats(C.byref(vox_tests), C.byref(vox_ms))
    log(f"geometry: {len(case.geometry.globalTriangles)} triangles, voxelised in {t_geo:.1f} s "
        f"({vox_tests.value} ray-triangle tests in {vox_ms.value:.3f} ms of kernel time)")
    t0 = time.perf_counter()
    sd = SubDomain(case.geometry, case.faces, storage=args.storage, tau=tau, dt=8e-6)
    sim = Simulation(sd, tau, 1.0, 8e-6)
    if args.streaming != "auto":
        sim.set_streaming(N.STREAMING_IN_PLACE if args.streaming == "in_place" else N.STREAMING_TWO_LATTICES)
    t_setup = time.perf_counter() - t0
    counts = sd.counts()
    nfl = counts["fluid"]
    nbox = box[0] * box[1] * box[2]
    log(f"setup {t_setup:.1f} s: {counts}")
    # parabolic inlet profile of unit mean (peak 2) on the parent tube
    if args.workload == "aneurysm":
        prof = inlet_disk_profile(case.geometry, sd, 1.0)
    else:
        prof = case.inlet_profile(sd, 2.0)
    prof *= 1.0 / prof.mean()
    sim.set_inlet_profile(prof)
    sim.set_wss_every(args.wss_every)
    period, mean, amp = 4000.0, 0.03, 0.02

    sim.step(args.warmup, inlet_scales(0, args.warmup, period, mean, amp))
    log(f"warmup {args.warmup} steps done")
    K = args.steps
    sc = inlet_scales(args.warmup, K, period, mean, amp)
    w0 = time.perf_counter()
    ms = 0.0
    done = 0
    chunk = 2000
    while done < K:
        m = min(chunk, K - done)
        ms += sim.step(m, sc[done:done + m], timed=True)
        done += m
        log(f"timed {done}/{K} steps")
    wall = time.perf_counter() - w0
    red = sim.reduce()
    if red[3] != 0 or not np.isfinite(red[0]):
        raise SystemExit(f"non-finite state after the timed steps: {red}")
    sec = ms / 1e3
    mlups = nfl * K / sec / 1e6
    bytes_per_flup = 2 * q * (8 if args.storage == "fp64" else 2)
    per_launch_s = sec / K
    achieved = nfl * bytes_per_flup / per_launch_s / 1e9

    # the reference's cadence beside the workload's: WSS and its OSI/TAWSS accumulators every
    # iteration (Simulation.cpp:585-586), timed on the same handle over the following steps
    ref_cadence = None
    if args.wss_every != 1 and args.ref_cadence_steps > 0:
        K1 = args.ref_cadence_steps
        sim.set_wss_every(1)
        ms1 = sim.step(K1, inlet_scales(args.warmup + K, K1, period, mean, amp), timed=True)
        sim.set_wss_every(args.wss_every)
        s1 = ms1 / 1e3 / K1
        ref_cadence = {"wss_every": 1, "steps": K1, "ms_per_step": round(ms1 / K1, 5),
                       "value": round(nfl / s1 / 1e6, 2), "unit": "MLUPS",
                       "roofline_frac": round(nfl * bytes_per_flup / s1 / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"WSS every iteration: {ms1 / K1 * 1e3:.2f} us per step")
    dtype = {"fp64": "f64", "fp16": "f16-storage/f32", "bf16": "bf16-storage/f32"}[args.storage]

    copy_gbs = None
    if args.copy_bytes:
        cms = C.c_double(0)
        N.check(N.lib().pvlbm_copy_bench(args.copy_bytes, 10, C.byref(cms)))
        copy_gbs = 2 * (args.copy_bytes // 16 * 16) * 10 / (cms.value / 1e3) / 1e9
        log(f"copy kernel: {copy_gbs:.0f} GB/s (read + write of {args.copy_bytes} bytes)")

    cpu = cpu_threads = None
    if not args.no_cpu:
        k = sd.kinds().reshape(-1)
        pbox = np.zeros(k.size)
        pbox[sd.list(N.LIST_INLET)] = prof
        v, n, el = cpu_baseline(case, pbox, q, tau, args.cpu_budget, 1)
        cpu = {"value": round(v, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
               "sample": f"{n} iterations of the same workload by the CPU oracle (push streaming, "
                         f"-O2 -ffp-contract=off, 1 thread) in {el:.1f} s on {cpu_model()}; "
                         "normals skipped (timing only)"}
        log(f"cpu baseline {v:.2f} MLUPS ({n} steps)")
        # the same oracle over every host core this process may use (a second, threaded baseline);
        # an OMP_NUM_THREADS set by the environment (a host's declared CPU share) caps it
        allowed = len(os.sched_getaffinity(0))
        omp = os.environ.get("OMP_NUM_THREADS", "")
        nthr = max(1, min(allowed, int(omp))) if omp.isdigit() and int(omp) > 0 else allowed
        vt, nt, elt = cpu_baseline(case, pbox, q, tau, args.cpu_budget / 2, nthr)
        cpu_threads = {"value": round(vt, 3), "unit": "MLUPS", "cores": nthr, "kind": "port",
                       "cores_allowed": allowed, "cpu_count": os.cpu_count(),
                       "omp_num_threads": omp or None,
                       "sample": f"{nt} iterations by the CPU oracle with OpenMP over {nthr} threads "
                                 f"({allowed} cores allowed to the process, {os.cpu_count()} in the machine) "
                                 f"in {elt:.1f} s on {cpu_model()}; normals skipped (timing only)"}
        log(f"cpu baseline, {nthr} threads: {vt:.2f} MLUPS ({nt} steps)")

    workload = f"{args.workload}-D3Q{q}-{args.storage}-{box[0]}x{box[1]}x{box[2]}"
    out = {
        "metric": "MLUPS + % HBM roofline",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded generated mesh, voxelised on the GPU)",
        "config": {
            "workload": workload,
            "lattice": q, "storage": args.storage, "box": list(box), "fluid_nodes": nfl,
            "fluid_fraction": round(nfl / nbox, 4),
            # the box's nodes per second: what a dense sweep of the whole box would have to reach
            "box_sweep_mlups_equivalent": round(nbox * K / (ms * 1e-3) / 1e6, 1), "tau": tau,
            "inlet": "0.03+0.02*sin(2*pi*t/4000) x parabolic(mean 1)", "wss_every": args.wss_every,
            "wss_accumulators": ("every iteration" if args.wss_every == 1 else
                                 f"sampled every {args.wss_every} iterations (OSI/TAWSS sums over the samples)"),
            "lists": counts, "population_stores": "nontemporal" if sim.streaming_stores() else "plain",
            "population_loads": {0: "plain", 1: "nontemporal in the even launches",
                                 2: "nontemporal"}[sim.streaming_loads()],
            "streaming": "in place, one lattice" if sim.in_place else "pull, two lattices",
            "wall_s_timed": round(wall, 3), "setup_s": round(t_setup, 2),
            "voxelise_s": round(t_geo, 2),
            "voxelise_kernel_ms": round(vox_ms.value, 3), "ray_triangle_tests": vox_tests.value,
            "ray_triangle_tests_per_s": (round(vox_tests.value / (vox_ms.value * 1e-3), -6)
                                         if vox_ms.value > 0 else None),
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(workload),
            "bytes_per_flup": bytes_per_flup,
            "kernel": f"k_step<{q},{args.storage}>", "launch_avg_us": round(per_launch_s * 1e6, 2),
            "copy_kernel_GBs": None if copy_gbs is None else round(copy_gbs, 1),
        },
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_threads,
        "reference_cadence": ref_cadence,
    }
    if args.learned:
        del sd, sim  # free the Y-bifurcation's device memory first
        from pv_dec_original_amd import learned_bench
        lb = learned_bench.run(ablations=args.learned_ablations, log=log)
        out["metric"] = "MLUPS + % HBM roofline; wall time to converged WSS (<=1% vs FP64) plain vs chosen"
        out["learned"] = {k: v for k, v in lb.items() if k != "runs"}
        out["learned"]["runs"] = {k: {kk: vv for kk, vv in v.items() if kk != "history"}
                                  for k, v in lb["runs"].items()}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
