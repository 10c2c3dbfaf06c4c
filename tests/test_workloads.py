"""Seeded workload generators (C2, C3, C4) on the CPU."""
import math

from mpdpft_amd import configs


def test_splitmix64_published_vector():
    """SplitMix64 (Steele, Lea & Flood 2014), seed 1234567: the first outputs of
    the published reference implementation."""
    r = configs.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821]


def test_c3_instances_follow_the_stated_distribution():
    insts = configs.c3(n=512)
    assert len(insts) == 512
    steps = [p.steps for p in insts]
    assert set(steps) == {64, 128, 256}
    frac = [steps.count(g) / 512 for g in (64, 128, 256)]
    assert abs(frac[0] - 0.5) < 0.06 and abs(frac[1] - 0.3) < 0.06 and abs(frac[2] - 0.2) < 0.06
    for p in insts:
        N = p.Abundances[0]
        assert N == p.Abundances[1] and 20 <= N <= 400 and N == int(N)
        assert abs(p.edge - 2.5 * math.sqrt(2 * N)) < 1e-12
        c, e = p.mpp[0], p.mpp[2]
        lam2 = 4 * c - 1  # mpp = ((1 + l^2)/4, ..., (l^2 - 1)/(l^2 + 1))
        assert 1.0 <= lam2 <= 4.0 + 1e-12 and abs(e - (lam2 - 1) / (lam2 + 1)) < 1e-12
        assert 0.2 <= p.beta <= 2.0
    # the same seed gives the same batch
    again = configs.c3(n=8)
    assert [p.Abundances[0] for p in again] == [p.Abundances[0] for p in insts[:8]]


def test_c4_centres_inside_the_sphere():
    pos = configs.centre_positions(32, 424242)
    assert len(pos) == 32 and all(math.sqrt(sum(x * x for x in r)) < 8.0 for r in pos)
    assert pos != configs.centre_positions(32, 434343)
    p = configs.c4()
    assert p.steps == 96 and p.edge == 30.0 and p.n_centres == 32 and p.Environments[0] == 41
    assert p.Abundances[0] == 32.0 and all(p.centres[i][3] == 1.0 and p.centres[i][4] == 0.6 for i in range(32))


def test_c2_is_the_precision_variant_of_c1():
    p = configs.c2()
    assert (p.steps, p.edge, p.mpp[0], p.mpp[1], p.Abundances[0], p.Precision) == (256, 24.0, 18.0, 0.4, 18.0, 1)
    assert p.DensityExpression == 5 and p.HartreeMode == 1 and list(p.Interactions)[:3] == [8, 11, -12]
