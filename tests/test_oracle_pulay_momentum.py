"""Pins of the oracle's Pulay momentum-space metric (PulayScalarProduct,
mpDPFT.cpp:5212-5227; weight from InitializePulayMixer :7048-7059) by closed
forms of residuals with known spectra.  CPU only.

In the metric, entry (i, j) is
  dk^D [(1 + w A_D sqrt(pi)/2) A0 + (1 + w/q_i^{D-1}) sum_{l != centre} Re(R_i conj R_j)
        - (G - 1) w/q_i^{D-1} e^{-q_i^2} A0],
A0 the product at k = 0, A_D = 2 pi in 2D, and q_i = |k| at grid index i, the
history position of the first residual (the reference's index quirk)."""
import numpy as np

from mpdpft_amd import configs


def params(Mixer2, steps=32, edge=12.0):
    p = configs.c0(steps=steps, edge=edge)
    p.S = 1
    p.Mixer[0], p.Mixer[1], p.Mixer[2] = 1.0, 8.0, Mixer2
    return p


def grid(p):
    EL = p.steps + 1
    dx = p.edge / p.steps
    x = -p.edge / 2 + dx * np.arange(EL)
    k0, dk = -np.pi / dx, 2 * np.pi / p.edge
    return EL, x, k0, dk


def qhist(i, EL, k0, dk):
    """|k| at grid index i (x slowest): (k0 + (i // EL) dk, k0 + (i % EL) dk)."""
    return np.hypot(k0 + (i // EL) * dk, k0 + (i % EL) * dk)


def test_constant_residuals_weight_the_zero_mode(oracle):
    """Constant residuals c_i have the single spectral value c_i L^2 at k = 0,
    so M(i, j) = dk^2 (1 + w 2 pi sqrt(pi)/2 - (G-1) w/q_i e^{-q_i^2}) c_i c_j L^4
    with w = k0^2 + dk^2 (Mixer[2] < 0).  A wrong angular factor or sqrt(pi)/2
    fails this at 1e-12."""
    p = params(-1.0)
    EL, x, k0, dk = grid(p)
    G = EL * EL
    c = np.array([1.0, -0.5, 2.0])
    res = np.array([np.full((1, G), ci) for ci in c])
    M = oracle.pulay_momentum_gram(p, res)
    w = k0 * k0 + dk * dk
    L = p.edge
    for i in range(3):
        q = qhist(i, EL, k0, dk)
        f = 1 + w * 2 * np.pi * np.sqrt(np.pi) / 2 - (G - 1) * (w / q) * np.exp(-q * q)
        for j in range(3):
            ref = dk ** 2 * f * c[i] * c[j] * L ** 4
            assert abs(M[i, j] - ref) <= 1e-12 * abs(ref)


def test_cosine_residuals_use_the_history_index(oracle):
    """Residuals a_i cos(2 pi m x / L) have spectral values a_i L^2 / 2 at
    k = (+-2 pi m / L, 0) and none at k = 0, so M(i, j) = dk^2 (1 + w/q_i)
    a_i a_j L^4 / 2 with q_i = |k| at grid index i (not at the mode): row i
    carries its own factor, the matrix is not symmetric, and a metric that took
    q at the mode would give (1 + w L/(2 pi m)) in every entry.  Explicit weight
    Mixer[2] = 0.7."""
    p = params(0.7)
    EL, x, k0, dk = grid(p)
    X = np.repeat(x, EL)
    m = 3
    a = np.array([1.0, 0.8, -1.3, 0.4])
    res = np.array([(ai * np.cos(2 * np.pi * m * X / p.edge)).reshape(1, -1) for ai in a])
    M = oracle.pulay_momentum_gram(p, res)
    w, L = 0.7, p.edge
    for i in range(len(a)):
        q = qhist(i, EL, k0, dk)
        for j in range(len(a)):
            ref = dk ** 2 * (1 + w / q) * a[i] * a[j] * L ** 4 / 2
            assert abs(M[i, j] - ref) <= 1e-10 * abs(ref)
    assert abs(M[0, 1] - M[1, 0]) > 1e-6 * abs(M[0, 1])


def test_momentum_metric_loop_converges_to_the_euclidean_fixed_point(oracle):
    """The metric changes the mixing coefficients, not the fixed point: a 2D
    two-species TF + contact solve with Mixer (1, 8, -1) converges to the
    density of the Euclidean solve (1e-7 at a 1e-9 stopping criterion)."""
    pe = configs.c0(steps=32, edge=12.0, tol=1e-9)
    pm = configs.c0(steps=32, edge=12.0, tol=1e-9)
    pm.Mixer[2] = -1.0
    rc_e, re, de, _ = oracle.solve(pe)
    rc_m, rm, dm, _ = oracle.solve(pm)
    assert rc_e == 0 and rc_m == 0 and re.converged and rm.converged
    assert rm.PulMix > 0 and rm.SCcount != re.SCcount  # the metric steers the Pulay steps
    assert np.abs(de - dm).max() < 1e-7 * np.abs(de).max()
    assert abs(re.TotalEnergy - rm.TotalEnergy) < 1e-9 * abs(re.TotalEnergy)
