"""Precision 1 (FP32 transforms, multipliers and pointwise potentials, then FP64
polish iterations) against the FP64 loop on the same input."""
import numpy as np
import pytest

from mpdpft_amd import Context, configs

pytestmark = pytest.mark.gpu


def solve(p):
    ctx = Context(p)
    r = ctx.solve()
    den, V = ctx.density().copy(), ctx.potential().copy()
    ctx.close()
    return r, den, V


def case(de, precision):
    if de == 5:  # the C1 functional: nAiT, isolated Hartree, Dirac, PW92
        p = configs.c1(steps=32, density_expression=5, hartree_mode=1, tol=1e-9)
    else:  # TF with isolated Hartree in a Z = 12 well holding 10 electrons (TF + Dirac does not converge)
        p = configs.c1(steps=32, density_expression=7, hartree_mode=1, Z=12.0, tol=1e-9)
        p.n_interactions = 1
        p.Interactions[0] = 8
    p.Precision = precision
    return p


@pytest.mark.parametrize("de", [7, 5])
def test_precision_variant_matches_fp64(de):
    """On 32^3: the FP32 phase ends at its floor and the FP64 polish meets the
    loop's own criterion, so E, mu and n agree with the FP64 solve to the
    loop tolerance (|dE| < 1e-6 relative is the precision target)."""
    r64, n64, V64 = solve(case(de, 0))
    r32, n32, V32 = solve(case(de, 1))
    assert r64.converged and r32.converged
    assert 0 < r32.fp32_iterations < r32.SCcount
    assert r64.fp32_iterations == 0
    dE = abs(r32.TotalEnergy - r64.TotalEnergy) / abs(r64.TotalEnergy)
    dmu = abs(r32.muVec[0] - r64.muVec[0]) / abs(r64.muVec[0])
    dn = np.abs(n32 - n64).max() / np.abs(n64).max()
    assert dE < 1e-6 and dmu < 1e-6 and dn < 1e-5, (dE, dmu, dn, r32.fp32_iterations, r32.SCcount, r64.SCcount)


def test_precision_variant_fp32_stage_is_fp32():
    """The FP32 phase alone (polish cut off by maxSCcount) differs from FP64 at
    the FP32 level, not at rounding level: the stages really run in FP32."""
    p64, p32 = case(7, 0), case(7, 1)
    p64.maxSCcount = p32.maxSCcount = 3
    r64, n64, _ = solve(p64)
    r32, n32, _ = solve(p32)
    dn = np.abs(n32 - n64).max() / np.abs(n64).max()
    assert 1e-12 < dn < 1e-4, dn
