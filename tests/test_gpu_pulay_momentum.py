"""GPU parity of the Pulay mixer's momentum-space metric (Mixer[2] != 0,
PulayScalarProduct mpDPFT.cpp:5212-5227) against the oracle, and of the
Euclidean Gram rows that now run on the FP64 matrix cores."""
import numpy as np
import pytest

from mpdpft_amd import Context, configs

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("mixer2,steps", [(-1.0, 32), (0.7, 32), (-1.0, 64)])
def test_momentum_metric_solve_matches_oracle(oracle, mixer2, steps):
    """C0 (2D, two species, TF + contact) with the momentum metric, converged
    (max|dV| < 1e-9): density, potential, mu, N and E at 1e-10, the same
    iteration count and the same Pulay/linear mixing and rejection counts."""
    print("case pulay momentum", mixer2, steps, flush=True)
    p = configs.c0(steps=steps, edge=12.0, tol=1e-9)
    p.Mixer[2] = mixer2
    rc, ro, dO, VO = oracle.solve(p)
    assert rc == 0 and ro.converged
    ctx = Context(p)
    rg = ctx.solve()
    dG, VG = ctx.density(), ctx.potential()
    ctx.close()
    assert rg.converged
    assert rg.SCcount == ro.SCcount and rg.PulMix == ro.PulMix and rg.pulay_rejections == ro.pulay_rejections
    assert ro.PulMix > 0
    assert rel(dG, dO) < RTOL and rel(VG, VO) < RTOL
    for s in range(p.S):
        assert abs(rg.muVec[s] - ro.muVec[s]) <= RTOL * abs(ro.muVec[s])
        assert abs(rg.FinalAbundances[s] - ro.FinalAbundances[s]) <= RTOL * abs(ro.FinalAbundances[s])
    assert abs(rg.TotalEnergy - ro.TotalEnergy) <= RTOL * abs(ro.TotalEnergy)


def test_momentum_metric_3d_trajectory_matches_oracle(oracle):
    """3D (C1 physics with the TF functional, periodic Hartree) with the
    automatic weight: 5 iterations on both paths at 1e-10."""
    p = configs.c1(steps=32, density_expression=7, hartree_mode=0, maxsc=5)
    p.Mixer[2] = -1.0
    rc, ro, dO, VO = oracle.solve(p)
    ctx = Context(p)
    rg = ctx.solve()
    dG, VG = ctx.density(), ctx.potential()
    ctx.close()
    assert rg.SCcount == ro.SCcount and rg.PulMix == ro.PulMix
    assert rel(dG, dO) < RTOL and rel(VG, VO) < RTOL
    assert abs(rg.TotalEnergy - ro.TotalEnergy) <= RTOL * abs(ro.TotalEnergy)
