"""Pins of the oracle's loop control that no closed-form solve checks directly:
CheckConvergence (mpDPFT.cpp:4759-4782), the exit rule, the mu bracketing
and bisection of EnforceConstraints (:4569-4743) on a traced search, the
converged nAiT mu as the root of N(mu) = N, and the converged EkinAiT03D
(:5610-5709, with its outer MaxIteration loop) against an independent
quadrature of the whole field.  CPU only."""
import os

import numpy as np
import pytest

from mpdpft_amd import configs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def states(p, dens, Vs, Ns, mus):
    G = (p.steps + 1) ** p.DIM
    return (np.array(dens).reshape(len(dens), p.S, G), np.array(Vs).reshape(len(Vs), p.S, G),
            np.array(Ns, dtype=float).reshape(len(Ns), p.S), np.array(mus, dtype=float).reshape(len(mus), p.S))


# ---------------------------------------------------------- CheckConvergence
def test_overlap_measure_closed_form(oracle):
    """Density mode: with n_new = c_s n_old per species, RelDev_s = (1 - c_s)^2 /
    (1 + c_s^2) exactly, and CC = max_s RelDev_s N_s / sum N; SCcount 0 gives
    1.2345678901234567890.  A dropped factor 2, a swapped weight or a sum in
    place of the maximum fails one of the three species cases."""
    p = configs.c0(steps=16)
    p.ConvergenceMode = 0
    rng = np.random.default_rng(1)
    G = (p.steps + 1) ** 2
    n0 = rng.uniform(0.1, 2.0, size=(2, G))
    for c, Ns in [((1.1, 0.7), (100.0, 50.0)), ((0.7, 1.1), (100.0, 50.0)), ((1.3, 1.3), (20.0, 180.0))]:
        n1 = np.array([c[0] * n0[0], c[1] * n0[1]])
        d, V, N, mu = states(p, [n0, n1], [np.zeros((2, G))] * 2, [Ns, Ns], [[1.0, 1.0]] * 2)
        CC, ex = oracle.convergence_sequence(p, d, V, N, mu)
        rd = [(1 - ci) ** 2 / (1 + ci ** 2) for ci in c]
        expect = max(rd[s] * Ns[s] / sum(Ns) for s in range(2))
        assert CC[0] == 1.2345678901234567890
        assert abs(CC[1] - expect) <= 1e-13 * expect, (c, CC[1], expect)
        assert ex == -1


def test_identical_states_are_steady_and_exit_after_three(oracle):
    """CC(n, n) = 0; with unchanged N and mu the state is 'steady' and counts
    towards the exit even with SCcriterion 0 (CC < 0 never holds); the loop
    exits after the third such iteration (ExitSCcount >= 3).  With mu moving
    by 1e-6 per iteration the same densities are not steady."""
    p = configs.c0(steps=16)
    p.ConvergenceMode = 0
    p.SCcriterion = 0.0
    G = (p.steps + 1) ** 2
    n = np.full((2, G), 0.5)
    d, V, N, mu = states(p, [n] * 6, [np.zeros((2, G))] * 6, [[100.0, 100.0]] * 6, [[3.0, 3.0]] * 6)
    CC, ex = oracle.convergence_sequence(p, d, V, N, mu)
    assert CC[1] == 0.0
    assert ex == 3
    mus = [[3.0 + 1e-6 * k] * 2 for k in range(6)]
    d, V, N, mu = states(p, [n] * 6, [np.zeros((2, G))] * 6, [[100.0, 100.0]] * 6, mus)
    CC, ex = oracle.convergence_sequence(p, d, V, N, mu)
    assert ex == -1


def test_potential_mode_exit_needs_three_consecutive_iterations(oracle):
    """Potential mode (max|V - V_old| < SCcriterion): a step above the criterion
    resets the counter, so with steps 1, 1e-9, 1e-9, 1e-7, 1e-9, 1e-9, 1e-9
    (criterion 1e-8) the loop exits at iteration 7, not 3; CC is the largest
    |dV| over both species."""
    p = configs.c0(steps=16)
    p.ConvergenceMode = 1
    p.SCcriterion = 1e-8
    G = (p.steps + 1) ** 2
    steps = [0.0, 1.0, 1e-9, 1e-9, 1e-7, 1e-9, 1e-9, 1e-9, 1e-9]
    Vs, v = [], np.zeros((2, G))
    for k, dv in enumerate(steps):
        v = v.copy()
        v[k % 2, (37 * k) % G] += dv  # the largest change sits in either species
        Vs.append(v)
    n = np.full((2, G), 0.5)
    d, V, N, mu = states(p, [n] * len(steps), Vs, [[100.0, 100.0]] * len(steps),
                         [[3.0 + k, 3.0] for k in range(len(steps))])
    CC, ex = oracle.convergence_sequence(p, d, V, N, mu)
    assert ex == 7
    assert np.allclose(CC[1:8], steps[1:8], rtol=1e-12, atol=0.)


# ------------------------------------------------------ EnforceConstraints
@pytest.mark.parametrize("mu0", [4.0, 30.0])
def test_mu_search_invariants(oracle, mu0):
    """The traced search of one species (2D TF, target N = 100): the walk moves
    mu monotonically towards the target with steps growing linearly
    (e^DMM (1 + n) |mu_prev|); N(mu) is monotone over every evaluated mu; the
    root lies inside the last walk step; each bisection step is half the
    previous one; the search stops at the first |N - N0| / 2N0 <= RelAcc (or
    on the trap rule) with the final mu within the last half-step of the
    root."""
    from scipy.optimize import brentq
    p = configs.c0(steps=32)
    p.S = 1
    p.Abundances[0] = 100.0
    p.n_interactions = 1
    p.Interactions[0] = 0
    v = oracle.env(p)
    mu, N, nb, tr = oracle.enforce_trace(p, v, [mu0])
    assert nb == 0 and len(tr) > 5
    N0 = 100.0
    root = brentq(lambda m: oracle.density(p, 0, m, v[0])[1] - N0, 0.1, 200.0, xtol=1e-14, rtol=1e-15)
    walk = [(m, n) for ph, m, n in tr if ph in (1, 2)]
    bis = [(m, n) for ph, m, n in tr if ph == 3]
    # N monotone in mu over all evaluations
    pts = sorted((m, n) for _, m, n in tr)
    assert all(b[1] >= a[1] for a, b in zip(pts, pts[1:]))
    # the walk: one direction, step k = e^DMM (1 + k) |mu0| with the DMM update after the first step
    direction = np.sign(walk[0][0] - mu0)
    ms = [mu0] + [m for m, _ in walk]
    d = np.diff(ms)
    assert np.all(np.sign(d) == direction)
    assert np.allclose(d[2:] / d[1:-1], [(k + 2) / (k + 1) for k in range(1, len(d) - 1)], rtol=1e-12)
    lo, hi = (ms[-2], ms[-1]) if direction > 0 else (ms[-1], ms[-2])
    assert lo <= root <= hi
    # bisection: steps halve, the search ends once N is within RelAcc
    bm = [m for m, _ in bis]
    steps = np.abs(np.diff(bm))
    halving = steps[0] * 0.5 ** np.arange(len(steps))
    assert np.all(np.abs(steps - halving) <= 1e-14 * max(abs(m) for m in bm) * np.arange(1, len(steps) + 1))
    rel = [abs(n - N0) / (2 * N0) for _, n in bis]
    assert rel[-1] <= p.RelAcc and all(r > p.RelAcc for r in rel[:-1])
    assert abs(mu[0] - root) <= 2 * steps[-1] + 1e-12 * abs(root)


def test_mu_search_reports_unreachable_targets(oracle):
    """Twenty walk steps cannot reach N = 1e9: mu returns to its start and the
    failure is counted (mpDPFT.cpp:4657-4661)."""
    p = configs.c0(steps=32)
    p.Abundances[0] = 1.0e9
    mu, N, nb, tr = oracle.enforce_trace(p, oracle.env(p), [10.0, 10.0])
    assert nb == 1 and mu[0] == 10.0
    assert sum(1 for ph, _, _ in tr if ph in (1, 2)) == 20


# ----------------------------------------------------------- nAiT at C1
def _c1_fixture():
    path = os.path.join(GOLDEN, "oracle_c1_nait_16_tight.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated")
    return np.load(path)


def test_nait_converged_mu_is_the_root_of_n_of_mu(oracle):
    """C1 at 16^3 converged to max|dV| < 1e-11: N(mu) of the nAiT density on
    the converged potential equals N to RelAcc, and an independent root
    search (brentq on N(mu) - N) returns the loop's mu."""
    from scipy.optimize import brentq
    f = _c1_fixture()
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1, tol=1e-11)
    V = f["V"][0]
    mu = float(f["mu"][0])
    Nmu = oracle.density(p, 0, mu, V)[1]
    assert abs(Nmu - 10.0) <= 2 * p.RelAcc * 10.0
    root = brentq(lambda m: oracle.density(p, 0, m, V)[1] - 10.0, mu - 0.01, mu + 0.01, xtol=1e-15, rtol=1e-14)
    assert abs(mu - root) <= 1e-9 * abs(root)


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_ekin_ait_equals_quadrature_of_the_field(oracle):
    """The converged EkinAiT03D of the C1 16^3 solve (its outer loop raises
    MaxIteration until the total settles) equals prefactor x the trapezoid
    integral of the per-point Airy averages computed independently (scipy
    quadrature, scipy Ai; the TF branch below the gradient threshold).
    Measured agreement 5e-12; without the Laplacian term the total moves by
    3e-4, so the 1e-9 bar also fails a dropped or mis-signed term, a wrong
    Airy scale a = (|grad V|^2 / 8)^(1/3) or a wrong degeneracy/(4 pi^2)."""
    from test_nait_tables import _ekin_quad
    f = _c1_fixture()
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1, tol=1e-11)
    V = f["V"][0]
    mu = float(f["mu"][0])
    d = oracle.delfield(p, V, 5)
    U = V - mu
    a = 0.5 * np.abs(d["gradsq"]) ** (1.0 / 3.0)
    lap = d["laplacian"]
    thr = p.Mpp[2]
    k = np.zeros_like(U)
    for i in range(U.size):
        if a[i] < thr:
            k[i] = (2.0 * -U[i]) ** 2.5 / 5.0 - lap[i] * np.sqrt(2.0 * -U[i]) / 6.0 if -U[i] > 0 else 0.0
        else:
            k[i] = _ekin_quad(U[i], a[i], lap[i])
    E = p.degeneracy / (4.0 * np.pi ** 2) * oracle.integrate(p, k)
    E0 = float(f["DispersalEnergies"][0])
    assert abs(E - E0) <= 1e-9 * abs(E0), (E, E0)
