"""Pins of the nAiT special functions (CurlyA, Li_s(-e^x), Ai).

CurlyA is regenerated from the mathematics (quadrature of Ai) at the knots of
the reference's TabFunc_CurlyA.dat and checked against that table; the
polylogarithms (whose reference tables are missing from the reference tree)
are checked against independent series, Taylor and Sommerfeld evaluations."""
import math
import os

import numpy as np
import pytest
from scipy import special as sp

from mpdpft_amd import tables as T

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "curlya_ref_sample.txt")
REF_TABLE = "/root/reference/TabFunc_CurlyA.dat"


@pytest.fixture(scope="module")
def tab():
    T.ensure(verbose=False)
    return T.read()


def test_curlya_knots_match_reference_sample(tab):
    rows = [l.split() for l in open(GOLDEN) if not l.startswith("#")]
    for idx, x, y in rows:
        i, x, y = int(idx), float(x), float(y)
        assert abs((T.C_X0 + T.C_H * i) - x) < 1e-12
        assert abs(tab["c"][i] - y) <= 1e-13 * max(abs(y), 1e-300)


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference tree absent")
def test_curlya_knots_match_full_reference_table(tab):
    v = np.array([float(t) for t in open(REF_TABLE).read().split()])
    y = v[1::2]
    assert len(y) == len(tab["c"])
    assert np.max(np.abs(tab["c"] - y) / np.abs(y)) < 1e-13


def test_curlya_identities(oracle, tab):
    # int_0^inf Ai = 1/3; CurlyA' = -Ai; asymptotic branches meet the spline
    assert abs(oracle.curlyA(0.0) - 1.0 / 3.0) < 1e-15
    for y in (-120.3, -7.77, -0.123, 0.456, 3.21, 11.1):
        h = 1e-4
        d = (oracle.curlyA(y + h) - oracle.curlyA(y - h)) / (2 * h)
        assert abs(d + sp.airy(y)[0]) < 2e-6
    assert abs(oracle.curlyA(-149.99) - oracle.curlyA(-149.99 + 1e-9)) < 1e-6


def test_curlya_spline_error_is_the_reference_interpolation_error(oracle):
    """Between knots the cubic spline (as the reference evaluates CurlyA) is
    within 1e-7 of the exact integral."""
    gx, gw = np.polynomial.legendre.leggauss(30)
    for y in (-140.005, -50.005, -3.005, 0.005, 5.005):
        # exact: CurlyA(knot) - int_knot^y Ai
        k = math.floor((y - T.C_X0) / T.C_H)
        xk = T.C_X0 + T.C_H * k
        a, b = xk, y
        ex = oracle.curlyA(xk) - 0.5 * (b - a) * np.sum(gw * sp.airy(0.5 * (b - a) * gx + 0.5 * (a + b))[0])
        assert abs(oracle.curlyA(y) - ex) < 1e-7


@pytest.mark.parametrize("j", [0, 1, 2, 3])
def test_polylog_against_independent_forms(oracle, j):
    s = T.P_ORDERS[j]
    xs_ser = np.array([-39.9, -20.0, -7.3, -2.2, -1.05])
    for x, ref in zip(xs_ser, T.polylog_series(s, xs_ser)):
        assert abs(oracle.polylog(j, x) - ref) <= 1e-13 * abs(ref)
    xs_tay = np.array([-0.97, -0.4, 0.0, 0.33, 0.9])
    for x, ref in zip(xs_tay, T.polylog_taylor(s, xs_tay)):
        assert abs(oracle.polylog(j, x) - ref) <= 1e-13 * abs(ref) + 1e-16
    xs_asy = np.array([45.0, 59.5, 61.0, 150.0, 650.0])
    for x, ref in zip(xs_asy, T.polylog_asymptotic(s, xs_asy, K=16)):
        assert abs(oracle.polylog(j, x) - ref) <= 1e-12 * abs(ref)
    # the reference's explicit branches
    assert oracle.polylog(j, -700.0) == 0.0
    assert oracle.polylog(j, 800.0) == pytest.approx(-800.0 ** s / math.gamma(s + 1), rel=1e-15)


def test_polylog_chain_is_a_derivative_chain(oracle):
    for j in range(3):
        for x in (-12.0, -1.1, 0.3, 4.7, 33.3):
            h = 1e-5
            d = (oracle.polylog(j, x + h) - oracle.polylog(j, x - h)) / (2 * h)
            assert abs(d - oracle.polylog(j + 1, x)) < 1e-8 * max(1, abs(d))


def test_polylog_tf_limit(oracle):
    """T -> 0: -Li_{3/2}(-e^x) -> x^{3/2}/Gamma(5/2) (the TF density)."""
    x = 2000.0
    assert abs(-oracle.polylog(0, 600.0) - 600.0 ** 1.5 / math.gamma(2.5)) / (600.0 ** 1.5) < 1e-5
    assert oracle.polylog(0, x) == pytest.approx(-x ** 1.5 / math.gamma(2.5), rel=1e-15)


def test_airy_interpolation(oracle):
    ys = np.array([-299.9973, -151.2345, -20.0021, -3.3333, 0.0, 0.0007, 2.5, 17.77, 99.99])
    ai = sp.airy(ys)[0]
    for y, a in zip(ys, ai):
        assert abs(oracle.airy(y) - a) < 1e-11 * max(1.0, abs(y)) ** 0.25 + 1e-15 * abs(a) * 0
    assert oracle.airy(100.5) == 0.0
    y = -400.0
    asym = math.cos(2 * 400 ** 1.5 / 3 - math.pi / 4) / (math.sqrt(math.pi) * 400 ** 0.25)
    assert oracle.airy(y) == pytest.approx(asym, rel=1e-14)


def test_airy_integrates_to_curlya(oracle):
    gx, gw = np.polynomial.legendre.leggauss(40)
    for a in (-80.0, -10.0, -1.0, 0.5):
        b = a + 0.7
        val = 0.5 * (b - a) * sum(w * oracle.airy(0.5 * (b - a) * x + 0.5 * (a + b)) for x, w in zip(gx, gw))
        assert abs(val - (oracle.curlyA(a) - oracle.curlyA(b))) < 1e-9


def _ekin_quad(U, a, lap):
    """Airy average int_0^xmax Ai(x + U/a) ((2ax)^{5/2}/5 - lap sqrt(2ax)/6) dx by adaptive
    quadrature with scipy's Ai (xmax = max(1, 100 - U/a), Ai taken as 0 beyond argument 100)."""
    from scipy import integrate as sint
    y0 = U / a
    xmax = max(1.0, 100.0 - y0)

    def f(x):
        if x + y0 > 100.0:
            return 0.0
        return sp.airy(x + y0)[0] * ((2 * a * x) ** 2.5 / 5 - lap * math.sqrt(2 * a * x) / 6)

    pts = [-y0] if 0 < -y0 < xmax else None
    return sint.quad(f, 0, xmax, points=pts, limit=2000, epsabs=0, epsrel=1e-13)[0]


@pytest.mark.parametrize("U,a,lap,acc,max_it,tol", [
    (-1.0, 0.3, 0.0, 1e-10, 16, 1e-12),
    (-1.0, 0.05, 0.0, 1e-10, 16, 1e-12),
    (0.4, 0.3, 0.0, 1e-10, 16, 1e-12),    # classically forbidden point: only the Airy tail
    (-2.5, 0.6, 0.0, 1e-10, 16, 1e-12),
    (-1.0, 0.3, 0.7, 1e-8, 18, 1e-8),     # sqrt(x) term: trapezoid error ~ h^1.5
    (0.2, 0.5, -0.4, 1e-8, 18, 1e-8),
])
def test_ekin_point_is_the_airy_average(oracle, U, a, lap, acc, max_it, tol):
    """EkinAiT03D's refined trapezoid sum equals the Airy average by independent quadrature."""
    v = oracle.ekin_point(U, a, lap, 0.0, acc, max_it)
    ref = _ekin_quad(U, a, lap)
    assert abs(v - ref) <= tol * abs(ref)


def test_ekin_point_tf_branch_is_the_fermi_gas(oracle):
    """Below the gradient threshold (uniform gas, no Laplacian) the kinetic energy density
    times 1/(4 pi^2) is the T = 0 Fermi gas's (3/5) n mu with n = (2 mu)^{3/2}/(6 pi^2);
    zero outside the classically allowed region."""
    for mu in (0.05, 0.7, 3.0):
        tau = 0.6 * (2 * mu) ** 1.5 / (6 * math.pi ** 2) * mu
        v = oracle.ekin_point(-mu, 0.0, 0.0, 1e-3, 1e-10, 10)
        assert abs(v / (4 * math.pi ** 2) - tau) <= 1e-14 * tau
    assert oracle.ekin_point(0.3, 0.0, 0.0, 1e-3, 1e-10, 10) == 0.0


def test_ekin_point_deep_well_limit(oracle):
    """Deep inside the well (X = -U/a >> 1) the Airy average tends to the TF branch, Laplacian
    term included, and for the x^{5/2} part the gap is the first moment correction
    f'''(X) M3/3! with M3 = int t^3 Ai(t) dt = 2 (the remainder, from the x = 0 endpoint,
    falls off like X^{-3/2} relative to it).  A wrong sign of the Airy shift or a wrong power
    breaks these by orders of magnitude."""
    U, a = -1.0, 0.002                                  # X = 500
    for lap in (0.0, 0.9):
        air = oracle.ekin_point(U, a, lap, 0.0, 1e-10, 18)
        tf = oracle.ekin_point(U, a, lap, 1e9, 1e-10, 18)
        assert abs(air - tf) <= 1e-5 * abs(tf)
    for a, bound in ((0.05, 0.05), (0.02, 0.015)):      # X = 20, 50
        X = -U / a
        c = (2 * a) ** 2.5 / 5
        pred = c * (15 / 8) * X ** -0.5 * (2 / 6) + c * (-225 / 64) * X ** -3.5 * (40 / 720)
        gap = oracle.ekin_point(U, a, 0.0, 0.0, 1e-12, 16) - oracle.ekin_point(U, a, 0.0, 1e9, 1e-12, 16)
        assert abs(gap - pred) <= bound * pred


def _fd(eta, j):
    """Fermi-Dirac integrals by quadrature in u = sqrt(t): F_{1/2} (j = 1) and F_{-3/2} (j = -3)."""
    from scipy import integrate as sint
    up = math.sqrt(max(eta, 0.0) + 40.0)
    if j == 1:
        f, pts = (lambda u: u * u * sp.expit(eta - u * u)), None
        c = 4 / math.sqrt(math.pi)
    else:
        f, pts = (lambda u: sp.expit(u * u - eta) * sp.expit(eta - u * u)), ([math.sqrt(eta)] if eta > 0 else None)
        c = 2 / math.sqrt(math.pi)
    return c * sint.quad(f, 0, up, points=pts, epsabs=0, epsrel=1e-13, limit=200)[0]


def _n_tf(U, lap, T):
    """Finite-T Thomas-Fermi density (T/2pi)^{3/2} F_{1/2}(-U/T) with the Laplacian term in the
    reference's form, -lap sqrt(T)/12 F_{-3/2}(-U/T)/(2pi)^{3/2} (mpDPFT.cpp:3375)."""
    eta = -U / T
    return (T ** 1.5 * _fd(eta, 1) - lap * math.sqrt(T) / 12 * _fd(eta, -3)) / (2 * math.pi) ** 1.5


@pytest.mark.parametrize("U,ar,lap,T", [
    (-1.0, 0.3, 0.0, 0.05), (-1.0, 0.3, 0.8, 0.05), (0.3, 0.4, 0.0, 0.1), (-0.5, 0.1, -0.5, 0.02),
    (-2.0, 0.7, 1.5, 0.2),
])
def test_nait_point_is_the_airy_average(oracle, U, ar, lap, T):
    """The nAiT density (CurlyA form, doubling lower limit, adaptive Boole panels) equals
    int Ai(x) n_TF(U - a_r x) dx by nested quadrature: scipy's Ai outside, Fermi-Dirac
    integrals inside."""
    from scipy import integrate as sint
    n = oracle.nait_point(U, ar, lap, T, 1e-9, 10, 20, 1e-12)
    xlo = (U - 45 * T) / ar
    pts = [U / ar] if xlo < U / ar < 30 else None
    ref = sint.quad(lambda x: sp.airy(x)[0] * _n_tf(U - ar * x, lap, T), xlo, 30, points=pts, epsabs=0,
                    epsrel=1e-12, limit=4000)[0]
    assert abs(n - ref) <= 1e-10 * abs(ref)


def test_nait_tf_branch_is_the_finite_t_fermi_gas(oracle):
    """Below the gradient threshold: (T/2pi)^{3/2} F_{1/2}(mu/T), with F by quadrature, plus the
    Laplacian term; degenerate, intermediate and Boltzmann regimes."""
    for U, T in ((-1.0, 0.05), (-0.2, 0.3), (0.5, 0.1), (2.0, 0.4)):
        for lap in (0.0, 0.9):
            n = oracle.nait_point(U, 0.0, lap, T, 1e-3, 10, 20, 1e-12)
            ref = _n_tf(U, lap, T)
            assert abs(n - ref) <= 1e-13 * abs(ref)


def test_nait_point_boltzmann_closed_form(oracle):
    """Far above mu the Fermi factors expand as Li_s(-z) = sum_k (-1)^k z^k / k^s, and the Airy
    average of each term is exact: int Ai(x) e^{p x} dx = e^{p^3/3}.  So n = beta0 sum_k (-1)^k
    z^k e^{(k a_r/T)^3/3} (T^{3/2} k^{-3/2} - lap sqrt(T)/12 k^{1/2}), z = e^{-U/T}; the k = 1
    factor e^{(a_r/T)^3/3} is the Wigner-Kirkwood beta^3 |grad V|^2/24 (|grad V|^2 = 8 a_r^3).
    T = 4 keeps the integrand far above the absolute cut of the lower-limit search."""
    T = 4.0
    for Ut in (10, 14):
        for bar in (0.3, 0.5, 0.7):
            for lap in (0.0, 3.0):
                U, ar = Ut * T, bar * T
                z = math.exp(-Ut)
                ref = -(2 * math.pi) ** -1.5 * sum(
                    (-1) ** k * z ** k * math.exp((k * bar) ** 3 / 3) * (T ** 1.5 * k ** -1.5 - lap * math.sqrt(T) / 12
                                                                          * k ** 0.5) for k in range(1, 5))
                n = oracle.nait_point(U, ar, lap, T, 1e-12, 10, 20, 1e-12)
                assert abs(n - ref) <= 1e-10 * abs(ref)


def test_nait_laplacian_term_is_the_reference_form(oracle):
    """The Laplacian term keeps the reference's coefficient -lap/(12 sqrt(alpha tau))
    (mpDPFT.cpp:3375): in the Boltzmann regime it shifts n by -lap/(12 T), which is the
    Wigner-Kirkwood -beta^2 lap/12 only at T = 1 (the hbar^2 expansion has 1/T^2)."""
    for T in (1.0, 0.05, 0.3):
        U = 30 * T
        n0 = oracle.nait_point(U, 0.0, 0.0, T, 1e9, 10, 20, 1e-12)
        n1 = oracle.nait_point(U, 0.0, 0.7, T, 1e9, 10, 20, 1e-12)
        assert abs((n1 - n0) / n0 - (-0.7 / (12 * T))) <= 1e-10
