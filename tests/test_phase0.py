"""The one-dimensional pulsatile model of the phase-0 start (pv_dec_original_amd/phase0.py):
Womersley's function in its limits, the cell-by-cell transmission line against the closed form of
a uniform tube, and the incompressible inertial limit."""
import numpy as np

from pv_dec_original_amd import phase0 as P


def _bessel_series_f(a, terms=60):
    """F from the power series of J0 and J1 (independent of the scaled Bessel routines)."""
    from math import factorial
    z = a * np.exp(0.75j * np.pi)
    j0 = sum((-1) ** k * (z / 2) ** (2 * k) / factorial(k) ** 2 for k in range(terms))
    j1 = sum((-1) ** k * (z / 2) ** (2 * k + 1) / (factorial(k) * factorial(k + 1)) for k in range(terms))
    return 2 * j1 / (z * j0)


def test_womersley_function_limits():
    for a in (0.3, 1.0, 3.0, 6.0, 8.7):
        assert abs(P.womersley_f([a])[0] - _bessel_series_f(a)) < 1e-12
    # low frequency: 1 - F = i a^2 / 8 (1 - i a^2 / 6 + ...), the Poiseuille flux Q = pi R^4 G / (8 nu)
    a = 0.05
    assert abs((1 - P.womersley_f([a])[0]) / (1j * a * a / 8) - (1 - 1j * a * a / 6)) < 1e-6
    # high frequency (inviscid core, Stokes layer): F = (2i / L)(1 - i / (2 L) + O(L^-2))
    for a in (60.0, 400.0):
        L = a * np.exp(0.75j * np.pi)
        assert abs(P.womersley_f([a])[0] / ((2j / L) * (1 - 0.5j / L)) - 1) < 3 / a ** 2


def test_uniform_tube_matches_the_closed_form_line():
    """Constant A: p(x) = q_in Zc sinh(g (L - x)) / cosh(g L), q(x) = q_in cosh(g (L - x)) / cosh(g L)."""
    n, A, T, tau, qa = 300, np.pi * 20.0 ** 2, 3000.0, 0.65, 7.0
    qh, rh = P.oscillation(np.full(n, A), qa, T, tau)
    nu, w = (tau - 0.5) / 3, 2 * np.pi / T
    alpha = np.sqrt(A / np.pi) * np.sqrt(w / nu)
    Z = 1j * w / (A * (1 - P.womersley_f([alpha])[0]))
    Y = 1j * w * A * 3.0
    g, zc = np.sqrt(Z * Y), np.sqrt(Z / Y)
    x = np.arange(n)
    L = n - 1
    qin = -1j * qa
    q_cf = qin * np.cosh(g * (L - x)) / np.cosh(g * L)
    p_cf = qin * zc * np.sinh(g * (L - x)) / np.cosh(g * L)
    assert np.allclose(qh, q_cf, rtol=1e-9, atol=1e-12 * qa)
    assert np.allclose(rh, 3.0 * p_cf, rtol=1e-9, atol=1e-15)
    assert qh[0] == qin and abs(rh[-1]) < 1e-15


def test_short_tube_inertial_limit():
    """Acoustically short, wide tube (alpha >> 1, L << wavelength): the pressure at phase 0 is the
    inviscid acceleration drop rho * omega * Q_amp * (L - x) / A, and the flux is uniform."""
    n, A, T, tau, qa = 64, np.pi * 200.0 ** 2, 2.0e6, 0.500003, 5.0
    qh, rh = P.oscillation(np.full(n, A), qa, T, tau)
    w = 2 * np.pi / T
    x = np.arange(n)
    p0 = (rh / 3.0).real  # phase 0: real part
    expect = w * qa * (n - 1 - x) / A
    assert np.allclose(p0, expect, rtol=2e-2, atol=1e-6 * expect[0])
    assert np.allclose(qh, qh[0], rtol=1e-3)


def test_phase0_profiles_shapes():
    kinds = np.ones((5, 6, 40), dtype=np.uint8)
    kinds[0] = 5  # a wall layer
    area = P.plane_areas(kinds)
    assert np.array_equal(area, np.full(40, 30.0 - 0.5 * 6))
    box = np.flatnonzero(kinds.reshape(-1))
    assert np.array_equal(P.plane_areas_from_lists(40, box, np.flatnonzero(kinds.reshape(-1) == 5)), area)
    t, d = P.phase0_profiles(area, 3.0, 2.0, 500, 0.6)
    assert t.shape == d.shape == (40,) and abs(d[-1]) < 1e-15 and abs(t[0] - 3.0) < 1e-12
    # the mean part keeps the mass flux: q rho(0) / rho(x) through denser or lighter planes
    rpl = np.linspace(1.02, 1.0, 40)
    t2, d2 = P.phase0_profiles(area, 3.0, 2.0, 500, 0.6, rpl)
    assert np.allclose(t2 - t, 3.0 * (rpl[0] / rpl - 1.0), atol=1e-12) and np.array_equal(d2, d)


def test_convective_line_limits():
    """The mean-flow option of the line: without mean flow it is the exact transfer above, and on
    a uniform inviscid line (alpha ~ 5e4) with mean velocity U its solution is the pair of
    Doppler-shifted plane waves p = P+ e^(-i w x / (c + U)) + P- e^(i w x / (c - U)) with
    q = (A / c)(P+ e^(..) - P- e^(..)) (the acoustic impedance holds in the moving fluid), fixed by
    q(0) = q_in and p(L) = 0."""
    A = np.full(240, 800.0)
    A[60:110] = 1500.0
    q0, r0 = P.oscillation(A, 10.0, 2000.0, 0.6)
    q1, r1 = P.oscillation(A, 10.0, 2000.0, 0.6, q_mean=1e-12)
    assert np.allclose(q1, q0, rtol=0, atol=1e-9) and np.allclose(r1, r0, rtol=0, atol=1e-12)
    n, a_, T, qa = 400, np.pi * 16.0 ** 2, 2000.0, 10.0
    for U in (0.0, 0.03, 0.1):
        qh, rh = P.oscillation(np.full(n, a_), qa, T, 0.5 + 1e-9, q_mean=U * a_ if U else 0.0)
        c, w, L = np.sqrt(1.0 / 3.0), 2 * np.pi / T, n - 1
        lp, lm = -1j * w / (c + U), 1j * w / (c - U)
        qin = -1j * qa
        # (A / c)(Pp - Pm) = qin,  Pp e^(lp L) + Pm e^(lm L) = 0
        Pm = -(qin * c / a_) * np.exp(lp * L) / (np.exp(lp * L) + np.exp(lm * L))
        Pp = qin * c / a_ + Pm
        x = np.arange(n)
        p_cf = Pp * np.exp(lp * x) + Pm * np.exp(lm * x)
        q_cf = (a_ / c) * (Pp * np.exp(lp * x) - Pm * np.exp(lm * x))
        assert np.abs(qh - q_cf).max() < 2e-3 * qa, U
        assert np.abs(rh - 3.0 * p_cf).max() < 2e-3 * np.abs(3.0 * p_cf).max(), U
    # the option reaches phase0_profiles
    t0, d0 = P.phase0_profiles(A, 20.0, 10.0, 2000.0, 0.6)
    t1, d1 = P.phase0_profiles(A, 20.0, 10.0, 2000.0, 0.6, convect=True)
    qc, rc = P.oscillation(A, 10.0, 2000.0, 0.6, q_mean=20.0)
    assert np.allclose(t1, 20.0 + qc.real) and np.allclose(d1, rc.real) and not np.allclose(t1, t0)


def test_matrix_exponential_closed_form():
    """_expm2 (e^t (cosh s I + sinh(s)/s N), N = M - t I) against scipy's Pade exponential on
    random complex 2x2 matrices, including nearly nilpotent N (the sinh(s)/s series branch)."""
    from scipy.linalg import expm
    rng = np.random.default_rng(4)
    M = rng.standard_normal((40, 2, 2)) + 1j * rng.standard_normal((40, 2, 2))
    M[:10] *= 1e-3
    N = np.array([[0.0, 1.0], [0.0, 0.0]], dtype=complex)  # nilpotent: exp = I + N
    M[10] = 0.3 * np.eye(2) + 1e-9 * N
    got = P._expm2(M)
    ref = np.array([expm(m) for m in M])
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-14)
