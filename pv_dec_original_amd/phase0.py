"""Phase-0 state of a pulsatile run along the vessel: a one-dimensional acoustic model.

With a velocity inlet on the x-min face and a fixed-density outlet on the x-max face, the domain
is an acoustic resonator, and a start whose density field is wrong rings in its lowest mode,
which viscosity damps by only about half per inlet period: on the 1024 x 512 x 512 aneurysm a
start from the exact velocity field with the period-mean density converges no faster than a
start from rest, while the exact phase-0 density (whose oscillatory part is a function of x alone,
to 2% of its norm) converges from the first period.  This module predicts that part.

Per x-plane with fluid cross-section A(x) (lattice units, c_s^2 = 1/3, rho_0 = 1; plane_areas), the pulsatile
part of the flux q and pressure p = c_s^2 rho at angular frequency omega obey the
transmission-line equations
    dp/dx = -Z q,   Z = i omega / (A (1 - F(alpha))),     (momentum, Womersley 1955)
    dq/dx = -Y p,   Y = i omega A / c_s^2,                 (mass, compressibility)
with F(alpha) = 2 J1(L) / (L J0(L)), L = i^(3/2) alpha, alpha = R sqrt(omega / nu), R = sqrt(A/pi):
the Womersley flux-pressure-gradient relation of a tube of the plane's area.  Boundary
conditions: the inlet flux q(0) = Q_amp sin(omega t) (complex amplitude -i Q_amp) and p = 0 at the
outlet plane.  Each cell is integrated exactly with its mean coefficients (cosh / sinh transfer
matrix); two basis solutions meet the two conditions.  The phase-0 state is the real part.
"""
import numpy as np
from scipy.special import jve

from ._native import WALL

CS2 = 1.0 / 3.0


def womersley_f(alpha):
    """F(alpha) = 2 J1(L) / (L J0(L)), L = i^(3/2) alpha (exponentially scaled Bessel functions
    keep the ratio finite for large alpha); F -> 1 - i alpha^2 / 8 as alpha -> 0."""
    alpha = np.asarray(alpha, dtype=np.float64)
    out = np.empty(alpha.shape, dtype=np.complex128)
    small = alpha < 1e-3
    L = alpha[~small] * np.exp(0.75j * np.pi)
    out[~small] = 2.0 * jve(1, L) / (L * jve(0, L))
    out[small] = 1.0 - 1j * alpha[small] ** 2 / 8.0
    return out


def plane_areas(kinds):
    """Effective fluid cross-section per x-plane (kinds: [nz, ny, nx], 0 = solid):
    nodes, with wall nodes counted half, since the bounce-back wall lies halfway between a wall
    node and its solid neighbour (on the 1024 x 512 x 512 aneurysm the measured admittance of the
    tube sections is 2.5% below that of the plain node count, which is what the half cells of
    its perimeter take away)."""
    k = np.asarray(kinds)
    nodes = np.count_nonzero(k != 0, axis=(0, 1)).astype(np.float64)
    return nodes - 0.5 * np.count_nonzero(k == WALL, axis=(0, 1))


def plane_areas_from_lists(nx, nodes, walls):
    """plane_areas from box indices (x fastest) of all nodes and of the wall nodes: the same
    counts without a pass over the whole box."""
    n = np.bincount(np.asarray(nodes) % nx, minlength=nx).astype(np.float64)
    return n - 0.5 * np.bincount(np.asarray(walls) % nx, minlength=nx)


def _expm2(M):
    """exp of 2x2 complex matrices M[..., 2, 2]: with t = tr M / 2 and N = M - t I (N^2 = s^2 I,
    s^2 = -det N), exp M = e^t (cosh s I + sinh(s)/s N)."""
    t = 0.5 * (M[..., 0, 0] + M[..., 1, 1])
    N = M.copy()
    N[..., 0, 0] -= t
    N[..., 1, 1] -= t
    s = np.sqrt(-(N[..., 0, 0] * N[..., 1, 1] - N[..., 0, 1] * N[..., 1, 0]) + 0j)
    small = np.abs(s) < 1e-8
    shs = np.where(small, 1.0 + s * s / 6.0, np.sinh(s) / np.where(small, 1.0, s))
    E = N * shs[..., None, None]
    E[..., 0, 0] += np.cosh(s)
    E[..., 1, 1] += np.cosh(s)
    return E * np.exp(t)[..., None, None]


def oscillation(area, q_amp, period, tau, q_mean=0.0):
    """Complex amplitudes (q_hat, rho_hat) per plane of the pulsatile part, for an inlet flux
    Q_mean + q_amp sin(2 pi t / period) on plane 0 and rho = rho_out on the last plane.

    q_mean > 0 adds the mean flow's convection, linearised about the mean velocity U = q_mean / A:
    the mass flux perturbation is q + U A rho, and the momentum equation's inertia
    i omega q becomes i omega q + U dq/dx.  With a = q_mean / c_s^2 and b = U / (A (1 - F)):
        dp/dx = (b Y p - Z q) / (1 - a b),   dq/dx = -Y p - a dp/dx,
    integrated per cell with the matrix exponential of the mean coefficients.  At Mach 0.05 to
    0.1 this moves the standing wave by about 0.1 rad over a 512-cell tube, which the flux of
    the downstream half feels first."""
    A = np.maximum(np.asarray(area, dtype=np.float64), 1.0)
    nu = (tau - 0.5) / 3.0
    w = 2.0 * np.pi / period
    alpha = np.sqrt(A / np.pi) * np.sqrt(w / nu)
    Z = 1j * w / (A * (1.0 - womersley_f(alpha)))
    Y = 1j * w * A / CS2
    Zm, Ym = 0.5 * (Z[1:] + Z[:-1]), 0.5 * (Y[1:] + Y[:-1])
    n = len(A)
    sol = np.zeros((2, 2, n), dtype=np.complex128)  # basis b: (p, q) along x
    sol[0, :, 0] = (1.0, 0.0)
    sol[1, :, 0] = (0.0, 1.0)
    if q_mean > 0.0:
        a = q_mean / CS2
        bq = (q_mean / A) / (A * (1.0 - womersley_f(alpha)))
        bm = 0.5 * (bq[1:] + bq[:-1])
        d = 1.0 - a * bm
        M = np.empty((n - 1, 2, 2), dtype=np.complex128)
        M[:, 0, 0] = bm * Ym / d
        M[:, 0, 1] = -Zm / d
        M[:, 1, 0] = -Ym - a * M[:, 0, 0]
        M[:, 1, 1] = -a * M[:, 0, 1]
        E = _expm2(M)
        for x in range(n - 1):
            sol[:, :, x + 1] = sol[:, :, x] @ E[x].T
    else:
        g = np.sqrt(Zm * Ym)
        zc = np.sqrt(Zm / Ym)
        ch, sh = np.cosh(g), np.sinh(g)
        for x in range(n - 1):
            p, q = sol[:, 0, x], sol[:, 1, x]
            sol[:, 0, x + 1] = p * ch[x] - zc[x] * q * sh[x]
            sol[:, 1, x + 1] = -p * sh[x] / zc[x] + q * ch[x]
    qin = -1j * q_amp
    a0 = -qin * sol[1, 0, -1] / sol[0, 0, -1]
    p_hat = a0 * sol[0, 0] + qin * sol[1, 0]
    q_hat = a0 * sol[0, 1] + qin * sol[1, 1]
    return q_hat, p_hat / CS2


def phase0_profiles(area, q_mean, q_amp, period, tau, rho_mean_planes=None, convect=False):
    """Per-plane targets at phase 0 (inlet flux rising through its mean) for planes of effective
    cross-section `area` (plane_areas): the flux of u_x over each plane and the density offset
    Re rho_hat to add to a period-mean density field.  The flux is the mean part, which carries
    the inlet's mass flux q_mean rho(0) through planes of period-mean density rho_mean_planes
    (default 1), plus the pulsatile part Re q_hat; on the inlet plane it is q_mean, the flux the
    inlet imposes at phase 0."""
    q_hat, r_hat = oscillation(area, q_amp, period, tau, q_mean if convect else 0.0)
    rpl = np.ones(len(q_hat)) if rho_mean_planes is None else np.asarray(rho_mean_planes, dtype=np.float64)
    return q_mean * rpl[0] / rpl + q_hat.real, r_hat.real
