"""Wall time to converged wall shear stress: the plain FP64 path against a chosen path.

A pulsatile run is sampled every `sample_every` iterations; the samples of one inlet period form
the period's WSS record W_P (samples x next-to-wall nodes, WSS magnitude).  The plain path (FP64
storage, start from rest) runs until the period-to-period change
    |W_P - W_{P-1}| / |W_P|  (L2 over the whole record)
falls below `rel_change`; its last record is the converged reference.  A chosen path (for example
FP16 shifted storage and the learned warm start) runs until its record is within `tol` (L2,
relative) of that reference.  Wall time counts everything the path does from its start: the
network inference and field upload of a warm start, every launch, and the WSS reads.
"""
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class PeriodicRun:
    label: str
    periods: int = 0
    wall_s: float = 0.0
    kernel_ms: float = 0.0
    start_s: float = 0.0
    start_mode: str = "rest"
    converged_at: int = None  # period index (1-based) at which the criterion held
    wall_at_converged_s: float = None
    history: list = field(default_factory=list)  # per period: change and/or error vs reference
    record: np.ndarray = None


def pulsatile_scales(it0, n, period, mean, amp):
    t = np.arange(it0, it0 + n, dtype=np.float64)
    return mean + amp * np.sin(2.0 * np.pi * t / period)


def run_periods(sim, label, period=4000, sample_every=200, mean=0.03, amp=0.02, max_periods=40,
                rel_change=None, reference=None, tol=0.01, start=None, log=None, min_periods=2):
    """Run whole inlet periods from the current state.  `start` (optional) is a callable that
    initialises sim and returns its mode; its time counts.  Stops when the period-to-period change
    is below rel_change (plain path) or when the record is within tol of reference (chosen
    path)."""
    res = PeriodicRun(label)
    t0 = time.perf_counter()
    if start is not None:
        res.start_mode = start()
    res.start_s = time.perf_counter() - t0
    sim.set_wss_every(sample_every)
    per = period // sample_every
    prev = None
    it = 0
    for P in range(1, max_periods + 1):
        rec = []
        for k in range(per):
            ms = sim.step(sample_every, pulsatile_scales(it, sample_every, period, mean, amp), timed=True)
            res.kernel_ms += ms
            it += sample_every
            rec.append(sim.wss()[:, 3].copy())
        rec = np.array(rec)
        res.periods = P
        res.wall_s = time.perf_counter() - t0
        h = {"period": P, "wall_s": res.wall_s, "kernel_ms": res.kernel_ms}
        if not np.isfinite(rec).all():
            h["nonfinite"] = True
            res.history.append(h)
            if log:
                log(f"{label}: non-finite WSS in period {P}")
            break
        if prev is not None:
            h["change"] = float(np.linalg.norm(rec - prev) / max(np.linalg.norm(rec), 1e-300))
        if reference is not None:
            h["error"] = float(np.linalg.norm(rec - reference) / max(np.linalg.norm(reference), 1e-300))
        res.history.append(h)
        if log:
            log(f"{label}: period {P} " + " ".join(f"{k} {v:.4g}" for k, v in h.items() if k != "period"))
        prev = rec
        done = False
        if rel_change is not None and "change" in h and h["change"] < rel_change and P >= min_periods:
            done = True
        if reference is not None and h["error"] <= tol:
            done = True
        if done:
            res.converged_at = P
            res.wall_at_converged_s = res.wall_s
            break
    res.record = prev
    return res
