"""The named workloads on the HIP path against the oracle.

C1 (the drop-in functional: nAiT T = 0.05 + isolated Hartree + Dirac +
PW92) and C2 (its Z = N = 18 physics) are compared at 16^3 with converged
oracle solves stored in tests/golden/oracle_*.npz (tests/golden/
make_oracle_fixtures.py wrote them; an oracle nAiT solve takes minutes even at
16^3).  C3 instances (2D two-species TF + contact, grids 64^2/128^2/256^2) are
solved by the oracle here.  C4 (the adjoint of the nAiT cluster) is compared
with the oracle's central differences of whole solves, stored in the same way.

Bars: the FP64 paths match density, potential, mu and total energy to 1e-10
relative with the SC iteration count within +-1 (C1: the converged fixed point
at 1e-10, the count within 15-30 %, see below); the FP32 + FP64-polish paths
match the total energy to 1e-6 relative; adjoint gradients and forces match
the oracle's central differences to 1e-6 relative.
"""
import os

import numpy as np
import pytest

from mpdpft_amd import Batch, Context, configs

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-10


def fixture(name):
    path = os.path.join(GOLDEN, "oracle_%s.npz" % name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated (tests/golden/make_oracle_fixtures.py)" % name)
    return np.load(path)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def check_fp64(r, den, V, f, label, rtol=RTOL, sc_band=1):
    print("%s: SCcount %d vs %d, rel n %.2e, rel V %.2e, rel mu %.2e, rel E %.2e" % (
        label, r.SCcount, int(f["SCcount"]), rel(den, f["den"]), rel(V, f["V"]), rel(r.muVec[0], f["mu"][0]),
        rel(r.TotalEnergy, f["TotalEnergy"])), flush=True)
    assert r.converged
    assert abs(r.SCcount - int(f["SCcount"])) <= sc_band
    assert rel(den, f["den"]) < rtol
    assert rel(V, f["V"]) < rtol
    for s in range(r.S):
        assert abs(r.muVec[s] - f["mu"][s]) <= rtol * abs(f["mu"][s])
        assert abs(r.FinalAbundances[s] - f["N"][s]) <= rtol * abs(f["N"][s])
    assert abs(r.TotalEnergy - float(f["TotalEnergy"])) <= rtol * abs(float(f["TotalEnergy"]))


# Why the C1 iteration count is not held to +-1: the GPU and the oracle run
# the same iteration (identical CC to the printed 8 digits and identical
# bisection counts through iteration 10), but in those first iterations the mu
# bracket is not found (mu held at 0.0034, N between 6.8 and 49) and Pulay
# mixing swings the density strongly, so rounding-level differences of the
# two density evaluations (1e-16..1e-13) grow to 1e-8 in V by iteration 11;
# from there the two trajectories differ visibly and reach the same fixed
# point after 107 and 113 iterations.  The fixed point itself is compared
# with both solves converged far below the 1e-10 bar.
def test_c1_converged_solve_matches_oracle():
    """C1 at 16^3 converged to max|dV| < 1e-11 with mu bisected to RelAcc
    2e-12: density, potential, mu, N and total energy at 1e-10.  (With the
    default RelAcc 1e-11 the bisection itself leaves mu ~2e-11 uncertain,
    1e-10 of |mu|.)  The tail of the two trajectories converges at different
    rates, so the count is only held to within 30 %."""
    f = fixture("c1_nait_16_tight")
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1, tol=1e-11)
    p.RelAcc = 2e-12
    ctx = Context(p)
    r = ctx.solve()
    check_fp64(r, ctx.density(), ctx.potential(), f, "C1 16^3 tol 1e-11",
               sc_band=max(2, int(0.3 * int(f["SCcount"]))))
    ctx.close()


def test_c1_default_tolerance_matches_oracle_to_the_tolerance():
    """The drop-in's own criterion (max|dV| < 1e-8): both loops stop within
    the criterion of the fixed point, so the results agree to ~1e-8 (the
    energy, quadratic in the error, much closer)."""
    f = fixture("c1_nait_16")
    ctx = Context(configs.c1(steps=16, density_expression=5, hartree_mode=1))
    r = ctx.solve()
    den, V = ctx.density(), ctx.potential()
    ctx.close()
    check_fp64(r, den, V, f, "C1 16^3 tol 1e-8", rtol=1e-7, sc_band=max(2, int(0.15 * int(f["SCcount"]))))
    assert abs(r.TotalEnergy - float(f["TotalEnergy"])) <= 1e-9 * abs(float(f["TotalEnergy"]))


def test_c2_physics_fp64_matches_oracle():
    """C2's physics at 16^3 in FP64: the Pulay solve is rejected on nearly
    every iteration here (the loop mixes linearly), so this also pins the
    fallback path."""
    f = fixture("c2_nait_16")
    ctx = Context(configs.c2(steps=16, precision=0))
    r = ctx.solve()
    check_fp64(r, ctx.density(), ctx.potential(), f, "C2 FP64 16^3")
    ctx.close()


def test_c2_precision_variant_matches_oracle():
    """Precision 1 (FP32 transforms and pointwise potentials until the FP32
    floor, then FP64 polish) against the converged FP64 oracle solve."""
    f = fixture("c2_nait_16")
    ctx = Context(configs.c2(steps=16, precision=1))
    r = ctx.solve()
    den = ctx.density()
    ctx.close()
    E0 = float(f["TotalEnergy"])
    dE = abs(r.TotalEnergy - E0) / abs(E0)
    dmu = abs(r.muVec[0] - f["mu"][0]) / abs(f["mu"][0])
    dn = rel(den, f["den"])
    print("C2 precision 1 16^3: FP32 iterations %d, total %d; |dE|/E %.2e, |dmu|/mu %.2e, max|dn|/max n %.2e"
          % (r.fp32_iterations, r.SCcount, dE, dmu, dn), flush=True)
    assert r.converged and r.fp32_iterations > 0
    assert dE < 1e-6
    assert dmu < 1e-6
    assert dn < 1e-5


# --------------------------------------------------------------------- C3 --
def _c3_sample():
    """The first two instances of each grid size in the seeded 512 list."""
    insts = configs.c3(n=512)
    idx = []
    for g in (64, 128, 256):
        idx += [i for i, p in enumerate(insts) if p.steps == g][:2]
    return insts, idx


def test_c3_batch_matches_oracle(oracle):
    insts, idx = _c3_sample()
    sample = [insts[i] for i in idx]
    b64 = Batch(sample)
    r64 = b64.solve()
    f64 = [b64.fields(k) for k in range(len(sample))]
    b64.close()
    b32 = Batch(sample, fp32=True)
    r32 = b32.solve()
    b32.close()
    for k, i in enumerate(idx):
        rc, ro, dO, VO = oracle.solve(insts[i])
        den, V = f64[k]
        r = r64[k]
        print("C3 instance %d (%d^2): SCcount %d vs %d, rel n %.2e, rel V %.2e, rel E %.2e; FP32+polish rel E %.2e"
              % (i, insts[i].steps, r.SCcount, ro.SCcount, rel(den, dO), rel(V, VO),
                 abs(r.TotalEnergy - ro.TotalEnergy) / abs(ro.TotalEnergy),
                 abs(r32[k].TotalEnergy - ro.TotalEnergy) / abs(ro.TotalEnergy)), flush=True)
        assert r.converged and ro.converged
        assert abs(r.SCcount - ro.SCcount) <= 1
        assert rel(den, dO) < RTOL and rel(V, VO) < RTOL
        for s in range(2):
            assert abs(r.muVec[s] - ro.muVec[s]) <= RTOL * abs(ro.muVec[s])
        assert abs(r.TotalEnergy - ro.TotalEnergy) <= RTOL * abs(ro.TotalEnergy)
        assert r32[k].converged
        assert abs(r32[k].TotalEnergy - ro.TotalEnergy) <= 1e-6 * abs(ro.TotalEnergy)


def test_c3_all_512_instances():
    """All 512 seeded instances in one launch each way: every instance
    converges, and FP32 + polish stays within 1e-6 of FP64 in energy."""
    insts = configs.c3(n=512)
    b64 = Batch(insts)
    r64 = b64.solve()
    b64.close()
    b32 = Batch(insts, fp32=True)
    r32 = b32.solve()
    b32.close()
    assert all(r.converged for r in r64)
    assert all(r.converged for r in r32)
    dE = [abs(a.TotalEnergy - b.TotalEnergy) / abs(b.TotalEnergy) for a, b in zip(r32, r64)]
    print("C3 512: max |dE|/E FP32+polish vs FP64 %.2e" % max(dE), flush=True)
    assert max(dE) < 1e-6


# --------------------------------------------------------------------- C4 --
def test_c4_nait_adjoint_matches_oracle_differences():
    """The implicit adjoint of the nAiT + Hartree + Dirac + PW92 cluster
    (16^3, three centres) against the oracle's central differences
    (h = 1e-4) of solves stopped at max|dV| < 1e-12: the loss gradient
    dL/dR and the forces dE/dR.  The central differences carry an absolute
    error (the truncation h^2 f'''/6 and the solves' stopping noise over 2h),
    so each component is held to 1e-6 of the largest component of its
    gradient: dL/dy of centre 1 is 1e-3 of dL/dx of centre 0, and its
    difference quotient is not accurate to 1e-6 of itself."""
    f = fixture("c4_nait_16")
    import sys
    sys.path.insert(0, GOLDEN)
    from make_oracle_fixtures import POS3, c4_params, c4_weights
    p = c4_params(POS3, 1e-12)
    ctx = Context(p)
    r = ctx.solve()
    assert r.converged
    den = ctx.density()
    print("C4 16^3 base solve: rel n %.2e vs oracle, rel E %.2e" % (
        rel(den, f["den"]), abs(r.TotalEnergy - float(f["TotalEnergy"])) / abs(float(f["TotalEnergy"]))), flush=True)
    assert rel(den, f["den"]) < RTOL
    w = c4_weights(p)
    nt = f["n_target"][0]
    lam, info = ctx.adjoint(2.0 * w * (den[0] - nt))
    assert info["converged"], info
    g = ctx.centre_gradient(lam)
    F = ctx.centre_forces()
    ctx.close()
    sL, sE = np.abs(f["fd_loss"]).max(), np.abs(f["fd_energy"]).max()
    for (ci, d), fdL, fdE in zip(f["coords"], f["fd_loss"], f["fd_energy"]):
        eL = abs(g[ci, d] - fdL) / sL
        eE = abs(F[ci, d] - fdE) / sE
        print("C4 centre %d axis %d: dL adjoint %.10e fd %.10e (%.1e); dE adjoint %.10e fd %.10e (%.1e)"
              % (ci, d, g[ci, d], fdL, eL, F[ci, d], fdE, eE), flush=True)
        assert eL < 1e-6
        assert eE < 1e-6
