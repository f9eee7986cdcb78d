"""Learned warm start: solver-generated training data, training, inference and the start.

Training data (one sample per seed): a seeded tortuous vessel in a 96^3 box (workloads.vessel_mesh:
cubic-spline centre line, radius 6-16, 1-2 branches, straight flow extensions at the x faces),
u_mean ~ U(0.01, 0.05) and tau ~ U(0.55, 0.9) drawn from the seed, parabolic inlet of mean u_mean
on the x-min face, pressure outlet (rho = 1) on the x-max face.  The FP64 D3Q19 solver runs from
rest, with the inlet ramped over the first RAMP steps, until the relative L2 change of u over
CHECK steps falls below TOL; its (u, rho) is the target, normalised as u / u_mean and
(rho - 1) / u_mean.

Warm start: the network sees the geometry code on a grid of at most NET_MAX voxels per axis (the
solver grid resampled by a factor s when larger), with (u_mean, tau) mapped to the coarse grid at
equal Reynolds number; its output is upsampled trilinearly to the solver grid and
denormalised with the solver's u_mean.  A field that is not finite, or far out of range (a speed
above REJECT_U, a density outside REJECT_RHO), is rejected and the solver starts from rest, the
reference start; otherwise the field is expanded on the device, every x-plane's flux is scaled to
the inlet flux (mass conservation, for inlets on x-min and outlets on x-max), speeds are clipped to
UMAX and densities into RHO_CLIP, and f = f_eq(rho, u) (pvlbm_init_fields_coarse).
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
UMAX = 0.3  # speeds above are clipped to it (lattice units; stability of BGK)
RHO_CLIP = (0.8, 1.25)  # densities are clipped into this range
REJECT_U = 1.0  # a field with a faster node, or a non-finite value, is rejected
REJECT_RHO = (0.5, 2.0)
NET_MAX = 192


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
    inlet_area: float
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
    sim.inlet_flux = float(prof.sum())
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
    code = code_from_kinds(kinds)
    return Sample(int(seed), um, ta, float(np.count_nonzero(code == 2)), code, target, steps, status,
                  time.perf_counter() - t0)


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
# evaluation (training itself: train_warmstart.py)

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
# inference and the warm start

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


def coarse_code(kinds, cshape):
    """Network input code on the grid cshape from the solver's node kinds: nearest sampling at the
    coarse voxel centres; the fluid voxels of the first and last x planes are the inlet and outlet
    faces when the solver grid has inlet and outlet nodes there.  On the solver grid itself this is
    code_from_kinds."""
    if tuple(cshape) == tuple(kinds.shape):
        return code_from_kinds(kinds)
    idx = [np.minimum(((np.arange(c) + 0.5) * n / c).astype(np.int64), n - 1) for c, n in zip(cshape, kinds.shape)]
    c = (kinds[np.ix_(*idx)] != N.SOLID).astype(np.uint8)
    fl = c > 0
    if (kinds[:, :, 0] == N.KINLET).any():
        c[:, :, 0][fl[:, :, 0]] = 2
    if (kinds[:, :, -1] == N.KOUTLET).any():
        c[:, :, -1][fl[:, :, -1]] = 3
    return c


def coarse_code_subdomain(sd, cshape):
    """coarse_code(sd.kinds(), cshape) without copying the kinds of the whole box: the coarse
    voxel centres are gathered on the device, and the inlet / outlet faces are read from the
    node lists."""
    shape = (sd.nz, sd.ny, sd.nx)
    if tuple(cshape) == shape:
        return code_from_kinds(sd.kinds())
    iz, iy, ix = [np.minimum(((np.arange(c) + 0.5) * n / c).astype(np.int64), n - 1) for c, n in zip(cshape, shape)]
    c = (sd.sample_kinds(ix, iy, iz) != N.SOLID).astype(np.uint8)
    fl = c > 0
    nx = sd.nx
    if (sd.list(N.LIST_INLET) % nx == 0).any():
        c[:, :, 0][fl[:, :, 0]] = 2
    outl = np.concatenate([sd.list(l) for l in (N.LIST_OUTLET1, N.LIST_OUTLET2, N.LIST_OUTLET3)])
    if (outl % nx == nx - 1).any():
        c[:, :, -1][fl[:, :, -1]] = 3
    return c


def _trilinear_up(field, shape):
    """Values at the fine nodes of `shape` from values at the centres of the coarse voxels of
    `field` (coarse voxel k spans fine coordinates [k, k+1) * n/c); clamped at the edges.  The
    host restatement of the device expansion (pvlbm_init_fields_coarse), used by the tests."""
    out = field
    for axis, n in enumerate(shape):
        nc = field.shape[axis]
        x = np.clip((np.arange(n) + 0.5) * nc / n - 0.5, 0, nc - 1)
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

    def net_grid(self, shape):
        """The network's grid for a solver grid: the solver grid itself when it fits (at most
        NET_MAX voxels per axis, sizes divisible by the network's divisor), else a grid with the
        longest axis at NET_MAX and the others scaled alike, each rounded to the divisor.
        Returns (coarse shape, scale s = fine / coarse along the longest axis)."""
        div = self._divisor()
        if max(shape) <= NET_MAX and all(n % div == 0 for n in shape):
            return tuple(shape), 1.0
        s = max(shape) / float(NET_MAX // div * div)
        s = max(s, 1.0)
        return tuple(max(div, int(round(n / s / div)) * div) for n in shape), s

    def _divisor(self):
        return self.net.spec.divisor

    def predict(self, code, u_mean, tau):
        """(rho, u) on the grid of `code` (d, h, w) as FP32 device tensors (d, h, w) and
        (3, d, h, w), denormalised with (u_mean, tau) as given to the network."""
        import torch
        d, h, w = code.shape
        cd = torch.from_numpy(np.ascontiguousarray(code)).to("cuda").reshape(1, d, h, w)
        prm = torch.tensor([[u_mean, tau, float(np.count_nonzero(code == 2))]], dtype=torch.float32,
                           device="cuda")
        y = self.net.forward(cd, prm).view(d, h, w, -1)
        u = y[..., 0:3].permute(3, 0, 1, 2).contiguous() * u_mean
        rho = 1.0 + y[..., 3].contiguous() * u_mean
        return rho, u

    def prepare(self, shape):
        """Load the model for a solver grid of shape (nz, ny, nx) ahead of the first start, like
        reading its weights: the network's buffers, the phase-0 model's module, and one
        inference with its rescaling on a straight tube of the network grid, so that the
        first start does not pay the first launch of those kernels."""
        cshape, s = self.net_grid(shape)
        self.net.buffers((1,) + tuple(cshape))
        from . import phase0  # noqa: F401  (scipy.special)
        cz, cy, cx = cshape
        z, y = np.meshgrid(np.arange(cz), np.arange(cy), indexing="ij")
        disk = (z - 0.5 * cz) ** 2 + (y - 0.5 * cy) ** 2 < (0.25 * min(cz, cy)) ** 2
        cc = np.zeros(cshape, dtype=np.uint8)
        cc[disk, :] = 1
        cc[disk, 0] = 2
        cc[disk, -1] = 3
        self._fields_from_code(cc, 0.03, 0.7, s)

    def coarse_fields(self, kinds, u_mean, tau, sd=None):
        """The network's field for a solver grid, on the network grid: device tensors rho
        (cz, cy, cx) and u (3, cz, cy, cx) for the solver's (u_mean, tau), rest on solid voxels,
        plus the coarse code.  kinds: the grid's node kinds, or None with sd given (the code is
        then gathered on the device, coarse_code_subdomain)."""
        cshape, s = self.net_grid(kinds.shape if kinds is not None else (sd.nz, sd.ny, sd.nx))
        cc = coarse_code(kinds, cshape) if kinds is not None else coarse_code_subdomain(sd, cshape)
        rho, u = self._fields_from_code(cc, u_mean, tau, s)
        return rho, u, cc

    def _fields_from_code(self, cc, u_mean, tau, s):
        uc, tc = coarse_params(u_mean, tau, s) if s > 1 else (u_mean, tau)
        rho, u = self.predict(cc, uc, tc)
        # at equal Reynolds number u / u_mean is the same field, and rho - 1 (viscous drop
        # nu u L / R^2 or inertial u^2 alike) scales with u_mean^2 / u_c^2
        k = u_mean / uc
        import torch
        solid = torch.from_numpy(cc == 0).to("cuda")
        rho = torch.where(solid, torch.ones_like(rho), 1.0 + (rho - 1.0) * (k * k))
        u = torch.where(solid[None], torch.zeros_like(u), u * k)
        return rho, u

    def apply(self, sim, sd, u_mean, tau, flux=None, pulsatile=None):
        """Start sim from the network's field (True) or leave it untouched when the field fails
        field_ok (False).  When the grid's inlet nodes all lie on the x-min plane and its outlet
        nodes on the x-max plane, the start enforces mass conservation: every x-plane carries the
        inlet flux `flux` (default u_mean times the number of inlet nodes).  pulsatile = (amp,
        period): the inlet runs at (u_mean + amp sin(2 pi t / period)) and the start is its
        phase 0; the network's steady field is then taken for the period mean and the
        compressible pulsatile part along x (plane fluxes and a density offset per plane,
        phase0.phase0_profiles) is imposed on it."""
        rho, u, cc = self.coarse_fields(None, u_mean, tau, sd)
        fl = cc > 0
        rho_h = rho.cpu().numpy()
        if not field_ok(rho_h, u.cpu().numpy(), fl):
            return False
        # inlet nodes all on the x-min plane, outlet nodes all on the x-max plane (from the lists)
        nx = sd.nx
        inl = sd.list(N.LIST_INLET) % nx
        outl = np.concatenate([sd.list(l) for l in (N.LIST_OUTLET1, N.LIST_OUTLET2, N.LIST_OUTLET3)]) % nx
        xfaces = len(inl) > 0 and len(outl) > 0 and not inl.any() and (outl == nx - 1).all()
        q = 0.0
        targets = drho = None
        if xfaces:
            q = float(u_mean * len(inl) if flux is None else flux)
            if pulsatile is not None:
                from .phase0 import phase0_profiles, plane_areas_from_lists
                amp, period = pulsatile
                # plane-mean density of the network's field, at the fine planes' coarse coordinates
                cx = rho_h.shape[2]
                cnt = np.maximum(np.count_nonzero(fl, axis=(0, 1)), 1)
                rpl = np.where(fl, rho_h, 0.0).sum(axis=(0, 1)) / cnt
                xc = np.clip((np.arange(nx) + 0.5) * cx / nx - 0.5, 0, cx - 1)
                area = plane_areas_from_lists(nx, sd.list(N.LIST_FLUID), sd.list(N.LIST_WALL))
                targets, drho = phase0_profiles(area, q, q * amp / u_mean, period, tau,
                                                np.interp(xc, np.arange(cx), rpl))
                q = 0.0
        nonfin = sim.init_fields_coarse(rho, u, UMAX, RHO_CLIP, q, targets, drho)
        return nonfin == 0


class PlugStart(WarmStart):
    """The learned start without the network, as its baseline: plug flow, u = (u_mean, 0, 0) and
    rho = 1 on every fluid voxel of the same coarse grid, then everything else WarmStart.apply does
    (each plane scaled to its flux, the phase-0 plane fluxes and densities of the pulsatile model,
    the device expansion).  What a start from WarmStart gains over this one is the network's."""

    def __init__(self, divisor=8):
        self.net = None
        self.divisor = divisor

    def _divisor(self):
        return self.divisor

    def prepare(self, shape):
        from . import phase0  # noqa: F401  (scipy.special), as WarmStart.prepare loads it

    def coarse_fields(self, kinds, u_mean, tau, sd=None):
        import torch
        cshape, _ = self.net_grid(kinds.shape if kinds is not None else (sd.nz, sd.ny, sd.nx))
        cc = coarse_code(kinds, cshape) if kinds is not None else coarse_code_subdomain(sd, cshape)
        fl = torch.from_numpy(cc > 0).to("cuda")
        rho = torch.ones(cshape, dtype=torch.float32, device="cuda")
        u = torch.zeros((3,) + tuple(cshape), dtype=torch.float32, device="cuda")
        u[0] = torch.where(fl, torch.full_like(u[0], float(u_mean)), u[0])
        return rho, u, cc


def field_ok(rho, u, fluid):
    """The check a predicted field must pass before the solver starts from it: finite, and no
    speed or density so far out that clipping would not be a repair."""
    r = rho[fluid]
    v = u[:, fluid]
    if not (np.isfinite(r).all() and np.isfinite(v).all()):
        return False
    if np.sqrt((v * v).sum(axis=0)).max(initial=0.0) > REJECT_U:
        return False
    return bool(r.min(initial=1.0) >= REJECT_RHO[0] and r.max(initial=1.0) <= REJECT_RHO[1])


def clip_field(rho, u):
    """Speeds above UMAX scaled down to UMAX, densities clipped into RHO_CLIP."""
    sp = np.sqrt((u * u).sum(axis=0))
    f = np.where(sp > UMAX, UMAX / np.maximum(sp, 1e-300), 1.0)
    return np.clip(rho, *RHO_CLIP), u * f[None]


def start(sim, sd, warm, u_mean, tau, flux=None, pulsatile=None):
    """Initialise sim from the network's field when it passes field_ok, else from rest.
    Returns 'warm' or 'rest'.  flux: the inlet flux (sum of the inlet velocities); pulsatile:
    (amp, period) for a start at phase 0 of a pulsatile inlet; see WarmStart.apply."""
    if warm is not None:
        try:
            if warm.apply(sim, sd, u_mean, tau, flux, pulsatile):
                return "warm"
        except (ValueError, N.NativeError):
            pass
    sim.init_fields()
    return "rest"


# ---------------------------------------------------------------------------------------------
# convergence of wall shear stress: the measure of the warm start's gain

def reference_wss(sim, fluid, tol=1e-6, max_steps=40000):
    """WSS magnitude per next-to-wall node of the converged steady state (run from rest with the
    inlet ramp until the relative change of u over CHECK steps is below tol)."""
    sim.init_fields()
    rho, u, steps, change, status = run_to_steady(sim, fluid, tol=tol, max_steps=max_steps)
    sim.set_wss_every(1)
    sim.step(1)
    return sim.wss()[:, 3].copy(), steps, status


def steps_to_wss(sim, ref, tol=0.01, check=100, max_steps=40000, ramp=RAMP, timed=True):
    """Iterations (and the kernels' milliseconds) until the relative L2 difference of the WSS
    magnitude from ref is at most tol, sampled every `check` iterations from the current state.
    Returns (steps or None, error, ms)."""
    sim.set_wss_every(check)
    it = 0
    ms = 0.0
    err = np.inf
    nref = max(np.linalg.norm(ref), 1e-300)
    while it < max_steps:
        r = sim.step(check, ramp_scales(it, check, ramp), timed=timed)
        ms += r or 0.0
        it += check
        w = sim.wss()[:, 3]
        if not np.isfinite(w).all():
            return None, np.inf, ms
        err = float(np.linalg.norm(w - ref) / nref)
        if err <= tol:
            return it, err, ms
    return None, err, ms
