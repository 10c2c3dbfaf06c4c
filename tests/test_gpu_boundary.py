"""The loop's inner boundary through the C ABI: one iteration at a time
(mpdpft_loop_begin / mpdpft_step / mpdpft_energy), EnforceConstraints and
AdmixDensities as stage calls against the oracle, the recoverable-event
counters, and the forcing hooks of the recovery paths."""
import numpy as np
import pytest

from mpdpft_amd import Context, _abi, configs

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def tf_ion(steps=16):
    p = configs.c1(steps=steps, density_expression=7, hartree_mode=1, tol=1e-9)
    p.n_interactions = 1
    p.Interactions[0] = 8
    return p


@pytest.mark.parametrize("case", ["c0", "tf_ion"])
def test_stepping_the_loop_equals_solve(case):
    print("case", case, flush=True)
    p = configs.c0(steps=32) if case == "c0" else tf_ion()
    ctx = Context(p)
    r0 = ctx.solve()
    n0 = ctx.density().copy()
    ctx.loop_begin()
    steps, done = 0, False
    while not done:
        r, done = ctx.step()
        steps += 1
        assert r.SCcount == steps
    r1 = ctx.energy()
    n1 = ctx.density()
    ctx.close()
    assert r1.SCcount == r0.SCcount and r1.converged == r0.converged
    assert r1.TotalEnergy == r0.TotalEnergy
    assert list(r1.muVec[:p.S]) == list(r0.muVec[:p.S])
    assert np.array_equal(n0, n1)


def test_step_before_loop_begin_is_rejected():
    ctx = Context(configs.c0(steps=16))
    with pytest.raises(Exception) as e:
        ctx.step()
    assert e.value.code == _abi.EINVAL
    with pytest.raises(Exception) as e:
        ctx.mix(0, np.zeros((2, ctx.G)))
    assert e.value.code == _abi.EINVAL
    ctx.close()


@pytest.mark.parametrize("case", ["c0", "nait16"])
def test_enforce_constraints_matches_oracle(oracle, case):
    """GetDensities + EnforceConstraints on a given potential: mu, N and the
    density of the last evaluated mu against the oracle."""
    print("case", case, flush=True)
    if case == "c0":
        p = configs.c0(steps=64)
        v = oracle.env(p)
        v = v + 0.05 * np.sin(np.arange(v.size).reshape(v.shape) * 0.01)
        mu0 = np.array([8.0, 12.0])
    else:
        p = configs.c1(steps=16, density_expression=5, hartree_mode=1)
        v = oracle.env(p) - 0.3
        mu0 = np.array([-1.0])
    ctx = Context(p)
    ctx.set_potential(v)
    mu, N, rc = ctx.enforce_constraints(mu0)
    den = ctx.density()
    ctx.close()
    mo, No, do, nb = oracle.enforce(p, v, mu0)
    print("mu", mu, mo, "N", N, No, "rel n %.2e" % rel(den, do), flush=True)
    assert rc == _abi.OK and nb == 0
    assert np.allclose(mu, mo, rtol=RTOL, atol=0.)
    assert np.allclose(N, No, rtol=RTOL, atol=0.)
    assert rel(den, do) < RTOL


def test_enforce_constraints_reports_a_missing_bracket(oracle):
    """An unreachable particle number: 20 bracketing steps do not find it,
    mu returns to its start (mpDPFT.cpp:4657-4661) and the call reports
    MPDPFT_ENOBRACKET; the oracle counts the same failure."""
    p = configs.c0(steps=32)
    p.Abundances[0] = 1.0e9
    v = oracle.env(p)
    mu0 = np.array([10.0, 10.0])
    ctx = Context(p)
    ctx.set_potential(v)
    mu, N, rc = ctx.enforce_constraints(mu0)
    ctx.close()
    mo, No, do, nb = oracle.enforce(p, v, mu0)
    assert rc == _abi.ENOBRACKET and nb == 1
    assert mu[0] == mu0[0] == mo[0]


def test_mix_sequence_matches_oracle(oracle):
    """AdmixDensities as a stage call: SCcount 0 mixes linearly, later calls
    run the Pulay mixer on the context's history; the mixed density after
    six calls against the oracle's same sequence."""
    p = configs.c0(steps=32)
    ctx = Context(p)
    ctx.loop_begin()
    n_in0 = ctx.density().copy()
    rng = np.random.default_rng(7)
    outs = [n_in0 * (1.0 + 0.05 * rng.standard_normal(n_in0.shape)) for _ in range(6)]
    codes = [ctx.mix(k, outs[k]) for k in range(6)]
    mixed = ctx.density()
    ctx.close()
    ref, rej = oracle.mix_sequence(p, n_in0, np.array(outs))
    print("codes", codes, "oracle rejections", rej, "rel %.2e" % rel(mixed, ref), flush=True)
    assert codes.count(_abi.EPULAY) == rej
    assert rel(mixed, ref) < RTOL


def test_duplicate_pulay_history_falls_back_to_linear_mixing():
    """Forcing hook: the newest Pulay entry pushed twice at iteration 5 makes
    the Gram matrix singular; those solves are rejected (counted in
    pulay_rejections), the iterations mix linearly, and once the duplicate
    leaves the history the loop converges to the unperturbed fixed point.
    The C1 functional at 16^3 runs Pulay on every iteration after the first
    (112 of 113 in the oracle); the converged energies agree to the loop's
    stopping tolerance."""
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1)
    ctx = Context(p)
    r0 = ctx.solve()
    assert r0.pulay_rejections == 0 and r0.LinMix == 1
    ctx.inject_fault(_abi.FAULT_DUPLICATE_PULAY, 5)
    r1 = ctx.solve()
    ctx.inject_fault(0, -1)
    r2 = ctx.solve()
    ctx.close()
    print("rejections", r1.pulay_rejections, "LinMix", r1.LinMix, "SCcount", r0.SCcount, r1.SCcount, flush=True)
    assert r1.pulay_rejections >= 1 and r1.LinMix == 1 + r1.pulay_rejections
    assert r1.converged
    assert abs(r1.TotalEnergy - r0.TotalEnergy) <= 1e-6 * abs(r0.TotalEnergy)
    assert r2.TotalEnergy == r0.TotalEnergy and r2.pulay_rejections == 0


def test_nan_in_the_potential_is_absorbed():
    """Forcing hook: NaN written into V at the grid centre after UpdateV of
    iteration 3.  The TF density there is 0 for that iteration ((mu - NaN) > 0
    is false, as in the reference's GetTFdensity), the next UpdateV rebuilds V
    from finite densities, and the loop converges to the fixed point of the
    unperturbed solve; a later solve without the hook is unaffected.  (C0 at
    32^2 rejects the Pulay solve on every iteration: this loop mixes
    linearly.)"""
    p = configs.c0(steps=32)
    ctx = Context(p)
    r0 = ctx.solve()
    ctx.inject_fault(_abi.FAULT_NAN_POTENTIAL, 3)
    r1 = ctx.solve()
    V1 = ctx.potential()
    ctx.inject_fault(0, -1)
    r2 = ctx.solve()
    ctx.close()
    print("SCcount", r0.SCcount, r1.SCcount, "E", r0.TotalEnergy, r1.TotalEnergy, flush=True)
    assert np.isfinite(V1).all()
    assert r1.converged
    # both stop at max|dV| < 1e-8 on different trajectories: the energies agree to that level
    assert abs(r1.TotalEnergy - r0.TotalEnergy) <= 1e-6 * abs(r0.TotalEnergy)
    assert r2.TotalEnergy == r0.TotalEnergy


def test_c2_fallback_counts_match_oracle():
    """The C2 physics at 16^3 rejects the Pulay solve on nearly every
    iteration: the rejection counter equals the oracle's (fixture)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_c2_nait_16.npz")
    f = np.load(path)
    ctx = Context(configs.c2(steps=16, precision=0))
    r = ctx.solve()
    ctx.close()
    assert r.LinMix == int(f["LinMix"]) and r.PulMix == int(f["PulMix"])
    assert r.pulay_rejections == r.LinMix - 1
