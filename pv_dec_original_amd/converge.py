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
    history: list = field(default_factory=list)  # per period: wall, kernel time, change, error
    records: list = field(default_factory=list)  # per period: WSS record (samples x nodes)

    def first(self, key, tol, min_period=1):
        """First period (1-based) whose `key` is at most tol, or None."""
        for h in self.history:
            if h["period"] >= min_period and key in h and h[key] <= tol:
                return h["period"]
        return None

    def wall_at(self, period):
        return None if period is None else self.history[period - 1]["wall_s"]


def pulsatile_scales(it0, n, period, mean, amp):
    t = np.arange(it0, it0 + n, dtype=np.float64)
    return mean + amp * np.sin(2.0 * np.pi * t / period)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_periods(sim, label, period=4000, sample_every=200, mean=0.03, amp=0.02, max_periods=40,
                stop_change=0.01, reference=None, start=None, log=None, min_periods=2):
    """Run whole inlet periods from the state set by `start` (a callable returning the start mode;
    its time counts) until the period-to-period change of the WSS record is below stop_change
    (from period min_periods on) or max_periods.  With a reference record, each period's error
    against it is kept too."""
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
            rec.append(sim.wss_magnitude())
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
            h["change"] = rel(rec, prev)
        if reference is not None:
            h["error"] = rel(rec, reference)
        res.history.append(h)
        res.records.append(rec)
        if log:
            log(f"{label}: period {P} " + " ".join(f"{k} {v:.4g}" for k, v in h.items() if k != "period"))
        prev = rec
        if "change" in h and h["change"] < stop_change and P >= min_periods:
            break
    return res
