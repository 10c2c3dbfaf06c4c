"""GPU parity of the n3' densities against the oracle: Getn3pFFT
(DensityExpression 6, mpDPFT.cpp:3625-3684: the momentum sum on the device,
then the backward transform) and Getn3p (DensityExpression 10, :3702-3729: the
Bessel-kernel sum on the device).  One density at a given potential and mu at
1e-12 (2D and 3D), then short SC trajectories (the mu search evaluates the
density tens of times per iteration): density, potential, mu and total energy
at 1e-10 with identical iteration counts."""
import numpy as np
import pytest

from mpdpft_amd import Context, configs

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def p2d(de, steps, S=1, maxsc=4):
    p = configs.c0(steps=steps, edge=8.0, N=6.0, maxsc=maxsc)
    p.S = S
    p.DensityExpression = de
    if S == 1:
        p.n_interactions = 1
        p.Interactions[0] = 0
    return p


def p3d(de, steps, maxsc=3):
    p = configs.c1(steps=steps, edge=10.0, Z=4.0, a=0.6, N=4.0, density_expression=7, hartree_mode=1, maxsc=maxsc)
    p.DensityExpression = de
    p.n_interactions = 2
    p.Interactions[0], p.Interactions[1] = 8, 11
    p.degeneracy = 2.0
    return p


def potential(p):
    EL = p.steps + 1
    x = np.linspace(-p.edge / 2, p.edge / 2, EL)
    X = np.stack(np.meshgrid(*([x] * p.DIM), indexing="ij"), axis=0)
    return (0.3 * (X ** 2).sum(0) + 0.25 * X[0] - 0.15 * X[-1] - 1.0).reshape(-1)


@pytest.mark.parametrize("case", [(6, 2, 32), (6, 2, 64), (6, 3, 16), (10, 2, 32), (10, 3, 16)])
def test_n3p_density_matches_oracle(oracle, case):
    de, D, steps = case
    print("case n3p density", case, flush=True)
    p = p2d(de, steps) if D == 2 else p3d(de, steps)
    V = potential(p)
    ctx = Context(p)
    for mu in (0.2, 1.1):
        a, Na = ctx.density_at(0, mu, V)
        b, Nb = oracle.density(p, 0, mu, V)
        assert rel(a, b) < 1e-12, (mu, rel(a, b))
        assert abs(Na - Nb) <= 1e-12 * abs(Nb)
    ctx.close()


@pytest.mark.parametrize("case", ["2d_two_species_n3p_fft", "2d_n3p_bessel", "3d_n3p_fft_hartree_dirac"])
def test_n3p_loops_match_oracle(oracle, case):
    print("case n3p loop", case, flush=True)
    if case == "2d_two_species_n3p_fft":
        p = p2d(6, 32, S=2)
    elif case == "2d_n3p_bessel":
        p = p2d(10, 16, S=2)
    else:
        p = p3d(6, 16)
    rc, ro, dO, VO = oracle.solve(p)
    ctx = Context(p)
    rg = ctx.solve()
    dG, VG = ctx.density(), ctx.potential()
    ctx.close()
    assert rg.SCcount == ro.SCcount
    assert rel(dG, dO) < 1e-10
    assert rel(VG, VO) < 1e-10
    for s in range(p.S):
        assert abs(rg.muVec[s] - ro.muVec[s]) <= 1e-10 * abs(ro.muVec[s])
        assert abs(rg.FinalAbundances[s] - ro.FinalAbundances[s]) <= 1e-10 * abs(ro.FinalAbundances[s])
    assert abs(rg.TotalEnergy - ro.TotalEnergy) <= 1e-10 * abs(ro.TotalEnergy)
