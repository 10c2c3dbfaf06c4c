"""Batched solves (mpdpft_batch_*) against single-instance solves on the GPU.

The batch runs each instance's whole SC loop inside one workgroup; the same
instance solved by mpdpft_solve (whose parity with the oracle is tested in
test_gpu_parity.py) must give the same result to the FP64 tolerance."""
import numpy as np
import pytest

from mpdpft_amd import Batch, Context, configs

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def single(p):
    ctx = Context(p)
    r = ctx.solve()
    den, V = ctx.density(), ctx.potential()
    ctx.close()
    return r, den, V


def test_batch_fp64_matches_single_solves():
    print("case batch fp64", flush=True)
    insts = configs.c3(n=6, seed=11, grids=(32, 64, 32))
    b = Batch(insts)
    rb = b.solve()
    for i, p in enumerate(insts):
        r, den, V = single(p)
        bd, bV = b.fields(i)
        print("instance %d: N=%g grid %d SCcount batch %d single %d" % (i, p.Abundances[0], p.steps, rb[i].SCcount,
                                                                       r.SCcount), flush=True)
        assert rb[i].converged == 1 and r.converged == 1
        assert rb[i].SCcount == r.SCcount
        assert rb[i].PulMix == r.PulMix and rb[i].LinMix == r.LinMix
        assert rel(bd, den) < RTOL and rel(bV, V) < RTOL
        for s in range(2):
            assert abs(rb[i].muVec[s] - r.muVec[s]) <= RTOL * abs(r.muVec[s])
            assert abs(rb[i].FinalAbundances[s] - r.FinalAbundances[s]) <= RTOL * abs(r.FinalAbundances[s])
        assert abs(rb[i].TotalEnergy - r.TotalEnergy) <= RTOL * abs(r.TotalEnergy)
    b.close()


def test_batch_queue_refill_is_slot_independent(monkeypatch):
    """Two resident workgroups drain a queue of 6: every instance is retired
    and the next one refilled into the freed slot; results equal those of one
    workgroup per instance bit for bit."""
    print("case batch refill", flush=True)
    insts = configs.c3(n=6, seed=5, grids=(32, 32, 64))
    ref = Batch(insts)
    r0 = ref.solve()
    f0 = [ref.fields(i) for i in range(len(insts))]
    ref.close()
    monkeypatch.setenv("MPDPFT_BATCH_SLOTS", "2")
    b = Batch(insts)
    r1 = b.solve()
    for i in range(len(insts)):
        assert r1[i].SCcount == r0[i].SCcount and r1[i].TotalEnergy == r0[i].TotalEnergy
        d, V = b.fields(i)
        assert np.array_equal(d, f0[i][0]) and np.array_equal(V, f0[i][1])
    b.close()


def test_batch_fp32_with_fp64_polish():
    """FP32 fields with FP64 integrals, mu and Pulay, then an FP64 polish from
    the FP32 result: |dE| < 1e-6 relative to the FP64 solve (the precision
    target), density to 1e-6 of its maximum."""
    print("case batch fp32", flush=True)
    insts = configs.c3(n=4, seed=3, grids=(32, 64, 64))
    b = Batch(insts, fp32=True)
    rb = b.solve()
    for i, p in enumerate(insts):
        r, den, V = single(p)
        bd, bV = b.fields(i)
        print("instance %d: |dE|/E %.2e  max|dn|/max n %.2e" % (i, abs(rb[i].TotalEnergy - r.TotalEnergy) /
                                                                 abs(r.TotalEnergy), rel(bd, den)), flush=True)
        # both loops stop at max|dV| < 1e-8, so they agree to about that level
        assert rb[i].converged == 1
        assert abs(rb[i].TotalEnergy - r.TotalEnergy) <= 1e-6 * abs(r.TotalEnergy)
        assert rel(bd, den) < 1e-6
    b.close()


def test_batch_teams_match_single_workgroup_solves(monkeypatch):
    """Instances of 256^2 points are solved by two-workgroup teams (each member
    sweeps half the points, reductions exchanged through global memory); the
    results equal the one-workgroup solves of the same instances to the FP64
    tolerance (the reductions combine in another order)."""
    print("case batch teams", flush=True)
    insts = configs.c3(n=4, seed=9, grids=(256, 256, 256))
    b = Batch(insts)
    rt = b.solve()
    ft = [b.fields(i) for i in range(len(insts))]
    b.close()
    monkeypatch.setenv("MPDPFT_BATCH_TEAM_POINTS", "0")
    b = Batch(insts)
    r1 = b.solve()
    for i in range(len(insts)):
        d, V = b.fields(i)
        print("instance %d: SCcount team %d single %d, rel E %.2e" % (
            i, rt[i].SCcount, r1[i].SCcount, abs(rt[i].TotalEnergy - r1[i].TotalEnergy) / abs(r1[i].TotalEnergy)),
            flush=True)
        assert rt[i].converged and abs(rt[i].SCcount - r1[i].SCcount) <= 1
        assert abs(rt[i].TotalEnergy - r1[i].TotalEnergy) <= RTOL * abs(r1[i].TotalEnergy)
        assert rel(ft[i][0], d) < RTOL and rel(ft[i][1], V) < RTOL
    b.close()
