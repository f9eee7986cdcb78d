"""Learned warm start: solver-generated training data, training, inference and the judged start.

Training data (one sample per seed): a seeded tortuous vessel in a 96^3 box (workloads.vessel_mesh:
cubic-spline centre line, radius 6-16, 1-2 branches, straight flow extensions at the x faces),
u_mean ~ U(0.01, 0.05) and tau ~ U(0.55, 0.9) drawn from the seed, parabolic inlet of mean u_mean
on the x-min face, pressure outlet (rho = 1) on the x-max face.  The FP64 D3Q19 solver runs from
rest, with the inlet ramped over the first RAMP steps, until the relative L2 change of u over
CHECK steps falls below TOL; its (u, rho) is the target, normalised as u / u_mean and
(rho - 1) / u_mean.

Warm start: the network sees the geometry code on a grid of at most NET_MAX voxels per axis (the
solver grid coarsened by an integer factor s when larger), with (u_mean, tau) mapped to the
coarse grid at equal Reynolds number; its output is upsampled trilinearly to the solver grid and
denormalised with the solver's u_mean.  The field is checked (finite, |u| below UMAX, rho within
RHO_RANGE) before pvlbm_init_fields sets f = f_eq(rho, u); a rejected field, or a run that turns
non-finite afterwards, restarts from rest, the reference start.
"""
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import _native as N

RAMP = 500
CHECK = 250
TOL = 2e-4
MAX_STEPS = 12000
UMAX = 0.35
RHO_RANGE = (0.7, 1.4)
NET_MAX = 96


def sample_params(seed):
    """(u_mean, tau) of a training seed."""
    rng = np.random.default_rng([int(seed), 0x5EED])
    return float(rng.uniform(0.01, 0.05)), float(rng.uniform(0.55, 0.9))


def code_from_kinds(kinds):
    """Geometry code of the network input: 0 solid, 1 fluid, 2 inlet face, 3 outlet face."""
    c = np.zeros(kinds.shape, dtype=np.uint8)
    c[kinds != N.SOLID] = 1
    c[kinds == N.KINLET] = 2
    c[kinds == N.KOUTLET] = 3
    return c


def ramp_scales(it0, n, ramp=RAMP):
    if not ramp:
        return np.ones(n)
    return np.minimum(1.0, (it0 + 1 + np.arange(n)) / float(ramp))


def run_to_steady(sim, fluid, ramp=RAMP, check=CHECK, tol=TOL, max_steps=MAX_STEPS):
    """Steps until the relative L2 change of u over `check` steps is below tol (after the ramp).
    Returns (rho, u, steps, last_change, status) with status 'converged', 'max_steps' or
    'diverged'."""
    prev = None
    it = 0
    change = np.inf
    while it < max_steps:
        sim.step(check, ramp_scales(it, check, ramp))
        it += check
        rho, u = sim.macros()
        uu = u[:, fluid]
        if not np.isfinite(uu).all() or not np.isfinite(rho[fluid]).all():
            return rho, u, it, np.inf, "diverged"
        if prev is not None:
            change = float(np.linalg.norm(uu - prev) / max(np.linalg.norm(uu), 1e-30))
            if change < tol and it >= ramp:
                return rho, u, it, change, "converged"
        prev = uu
    return rho, u, it, change, "max_steps"


@dataclass
class Sample:
    seed: int
    u_mean: float
    tau: float
    code: np.ndarray  # (n, n, n) uint8
    target: np.ndarray  # (n, n, n, 4) float32, normalised
    steps: int
    status: str
    seconds: float


def vessel_setup(seed, n=96, q=19, storage="fp64", u_mean=None, tau=None):
    """SubDomain, Simulation and inlet profile of a training/validation vessel."""
    from .simulation import SubDomain, Simulation
    from .workloads import vessel_case, inlet_disk_profile
    um, ta = sample_params(seed)
    um = um if u_mean is None else u_mean
    ta = ta if tau is None else tau
    case = vessel_case(seed, q=q, n=n, tau=ta)
    sd = SubDomain(case.geometry, case.faces, storage=storage, tau=ta, dt=8e-6)
    sim = Simulation(sd, ta, 1.0, 8e-6)
    prof = inlet_disk_profile(case.geometry, sd, um)
    sim.set_inlet_profile(prof)
    sim.set_wss_every(1 << 30)
    return case, sd, sim, um, ta


def solve_sample(seed, n=96, tol=TOL, max_steps=MAX_STEPS):
    t0 = time.perf_counter()
    case, sd, sim, um, ta = vessel_setup(seed, n)
    kinds = sd.kinds()
    fluid = kinds != N.SOLID
    rho, u, steps, change, status = run_to_steady(sim, fluid, tol=tol, max_steps=max_steps)
    target = np.zeros((n, n, n, 4), dtype=np.float32)
    if status != "diverged":
        target[..., 0] = u[0] / um
        target[..., 1] = u[1] / um
        target[..., 2] = u[2] / um
        target[..., 3] = (rho - 1.0) / um
        target[~fluid] = 0.0
    return Sample(int(seed), um, ta, code_from_kinds(kinds), target, steps, status, time.perf_counter() - t0)


def generate(seeds, n=96, workers=4, log=None, **kw):
    """Solve the seeds concurrently (one solver handle and HIP stream per worker thread)."""
    out = []
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for i, s in enumerate(ex.map(lambda sd: solve_sample(sd, n, **kw), seeds)):
            out.append(s)
            if log and (i + 1) % max(1, len(seeds) // 20) == 0:
                log(f"generated {i + 1}/{len(seeds)} (seed {s.seed}: {s.status} after {s.steps} steps)")
    return out


# ---------------------------------------------------------------------------------------------
# weights

def save_weights(path, net, meta=None):
    meta = dict(meta or {})
    np.savez_compressed(path, levels=np.array(net.spec.levels, dtype=np.int32),
                        params=net.params_numpy(), **{k: np.asarray(v) for k, v in meta.items()})


def load_weights(path):
    with np.load(path, allow_pickle=False) as z:
        return tuple(int(v) for v in z["levels"]), z["params"].astype(np.float32)


# ---------------------------------------------------------------------------------------------
# training

def train(net, codes, targets, prms, epochs=1, batch=4, lr=1e-3, lr_end=1e-5, seed=0, log=None,
          time_budget=None, val=None):
    """Adam over minibatches of device-resident samples.
    codes: uint8 (S, d, h, w); targets: bf16 (S, d*h*w, 4); prms: float32 (S, 2) — torch device
    tensors.  Cosine learning-rate decay from lr to lr_end over the planned steps.  Returns the
    per-step training losses."""
    import torch
    S = codes.shape[0]
    steps_per_epoch = S // batch
    total = max(1, steps_per_epoch * epochs)
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.perf_counter()
    step = 0
    for ep in range(epochs):
        order = rng.permutation(S)
        for b in range(steps_per_epoch):
            idx = torch.from_numpy(order[b * batch:(b + 1) * batch].copy()).to(codes.device)
            code = codes.index_select(0, idx).contiguous()
            tgt = targets.index_select(0, idx).reshape(-1, targets.shape[-1]).contiguous()
            prm = prms.index_select(0, idx).contiguous()
            cur = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * step / total))
            net.forward(code, prm)
            loss = net.loss(tgt, code)
            net.backward()
            net.adam_step(float(cur))
            step += 1
            if step % 50 == 0 or step == total:
                losses.append((step, float(loss[0])))
                if log:
                    log(f"step {step}/{total} epoch {ep} loss {losses[-1][1]:.4e} lr {cur:.2e} "
                        f"{time.perf_counter() - t0:.0f}s")
            if time_budget and time.perf_counter() - t0 > time_budget:
                return losses
    return losses


def evaluate(net, codes, targets, prms, batch=4):
    """Mean masked MSE and the relative L2 error of u over the fluid voxels, per sample."""
    import torch
    S = codes.shape[0]
    mse, rel = [], []
    for b in range(0, S - S % batch, batch):
        code = codes[b:b + batch].contiguous()
        tgt = targets[b:b + batch].reshape(-1, targets.shape[-1]).contiguous()
        y = net.forward(code, prms[b:b + batch].contiguous())
        loss = net.loss(tgt, code)
        mse.append(float(loss[0]))
        fl = (code.reshape(-1) > 0)
        yu = y[:, :3][fl].float()
        tu = tgt[:, :3][fl].float()
        rel.append(float(torch.linalg.norm(yu - tu) / torch.linalg.norm(tu)))
    return float(np.mean(mse)), float(np.mean(rel))


# ---------------------------------------------------------------------------------------------
# inference and the judged warm start

def coarse_params(u_mean, tau, s):
    """Network inputs (u_c, tau_c) on a grid coarsened by s with the solver's Reynolds number
    u D / nu, kept inside the training ranges u in [0.01, 0.05], tau in [0.55, 0.9]."""
    nu = (tau - 0.5) / 3.0
    ratio = s * u_mean / nu  # u_c / nu_c
    uc = 0.05
    tc = 0.5 + 3.0 * uc / ratio
    tc = min(max(tc, 0.55), 0.9)
    uc = min(max(ratio * (tc - 0.5) / 3.0, 0.01), 0.05)
    return uc, tc


def coarsen_code(code, s):
    """Code of a grid coarsened by the integer factor s: a coarse voxel is fluid when the fine
    node at its centre is; inlet/outlet faces are the fluid coarse voxels on the x planes that
    hold fine inlet/outlet nodes."""
    if s == 1:
        return code.copy()
    nz, ny, nx = code.shape
    c = code[s // 2::s, s // 2::s, s // 2::s][: nz // s, : ny // s, : nx // s].copy()
    c[c > 1] = 1
    fl = c > 0
    if (code[:, :, 0] == 2).any():
        c[:, :, 0][fl[:, :, 0]] = 2
    if (code[:, :, -1] == 3).any():
        c[:, :, -1][fl[:, :, -1]] = 3
    return c


def _trilinear_up(field, s, shape):
    """Values at fine nodes from values at coarse voxel centres (coarse voxel k covers fine nodes
    k*s .. k*s + s - 1, centre at k*s + (s - 1)/2); clamped at the edges."""
    out = field
    for axis, n in enumerate(shape):
        x = (np.arange(n) - (s - 1) / 2.0) / s
        nc = field.shape[axis]
        x = np.clip(x, 0, nc - 1)
        i0 = np.floor(x).astype(np.int64)
        i1 = np.minimum(i0 + 1, nc - 1)
        wgt = (x - i0).reshape([-1 if a == axis else 1 for a in range(out.ndim)])
        out = np.take(out, i0, axis=axis) * (1 - wgt) + np.take(out, i1, axis=axis) * wgt
    return out


class WarmStart:
    """A trained network and the rule that turns its output into an initial solver field."""

    def __init__(self, path=None, net=None):
        from .warmnet import UNet, UNetSpec
        if net is None:
            levels, params = load_weights(path)
            net = UNet(UNetSpec(levels=levels), params)
        self.net = net

    def predict(self, code, u_mean, tau):
        """(rho, u) on the grid of `code` (d, h, w), u as (3, d, h, w), numpy FP64."""
        import torch
        d, h, w = code.shape
        self.net.spec.check_shape((1, d, h, w))
        cd = torch.from_numpy(np.ascontiguousarray(code)).to("cuda").reshape(1, d, h, w)
        prm = torch.tensor([[u_mean, tau]], dtype=torch.float32, device="cuda")
        y = self.net.forward(cd, prm)
        torch.cuda.synchronize()
        y = y.view(d, h, w, -1).cpu().numpy().astype(np.float64)
        u = np.moveaxis(y[..., 0:3], -1, 0) * u_mean
        rho = 1.0 + y[..., 3] * u_mean
        return rho, u

    def fields(self, kinds, u_mean, tau):
        """Initial (rho, u) on the solver grid of `kinds` (nz, ny, nx)."""
        code = code_from_kinds(kinds)
        div = self.net.spec.divisor
        s = max(1, -(-max(kinds.shape) // NET_MAX))
        cshape = [max(div, (n // s) // div * div) for n in kinds.shape]
        c = coarsen_code(code, s)
        cc = np.zeros(cshape, dtype=np.uint8)
        sl = tuple(slice(0, min(a, b)) for a, b in zip(cshape, c.shape))
        cc[sl] = c[sl]
        uc, tc = coarse_params(u_mean, tau, s) if s > 1 else (u_mean, tau)
        rho_c, u_c = self.predict(cc, uc, tc)
        # at equal Reynolds number u / u_mean is the same field, and rho - 1 (viscous drop
        # nu u L / R^2 or inertial u^2 alike) scales with u_mean^2 / u_c^2
        u_c = u_c * (u_mean / uc)
        rho_c = 1.0 + (rho_c - 1.0) * (u_mean / uc) ** 2
        if s == 1:
            rho, u = rho_c, u_c
        else:
            rho = _trilinear_up(rho_c, s, kinds.shape)
            u = np.stack([_trilinear_up(u_c[k], s, kinds.shape) for k in range(3)])
        fluid = kinds != N.SOLID
        rho = np.where(fluid, rho, 1.0)
        u = np.where(fluid[None], u, 0.0)
        return rho, u


def field_ok(rho, u, fluid):
    """The check a predicted field must pass before the solver starts from it."""
    r = rho[fluid]
    v = u[:, fluid]
    if not (np.isfinite(r).all() and np.isfinite(v).all()):
        return False
    if np.abs(v).max(initial=0.0) > UMAX:
        return False
    return bool(r.min(initial=1.0) >= RHO_RANGE[0] and r.max(initial=1.0) <= RHO_RANGE[1])


def start(sim, sd, warm, u_mean, tau):
    """Initialise sim from the network's field when it passes field_ok, else from rest.
    Returns 'warm' or 'rest'."""
    if warm is not None:
        kinds = sd.kinds()
        try:
            rho, u = warm.fields(kinds, u_mean, tau)
        except (ValueError, N.NativeError):
            rho = u = None
        if rho is not None and field_ok(rho, u, kinds != N.SOLID):
            sim.init_fields(rho, u)
            return "warm"
    sim.init_fields()
    return "rest"
