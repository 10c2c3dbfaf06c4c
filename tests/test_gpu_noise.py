"""Noise (ImposeNoise mpDPFT.cpp:4490-4504 on the initial densities, the
noise schedules 3, 4, 7 and 11 of InjectSchedule :5252-5306) and TaskType 10
(mpDPFTmain.cpp:179-198, RetrieveTask mpDPFT.cpp:6124-6131) on the GPU.

The generator is seeded through the extension key NoiseSeed (the reference
seeds it from the clock), so the oracle draws the same factors: the GPU's
noisy densities, whole noisy loops and the scheduled injections agree with it
at 1e-13 (one density) and 1e-10 (trajectories, identical iteration counts).
TaskType 10: every repetition is a converged solve of its own noise, the
task file records the running minimum and the sorted energy history, and the
final run (started from the best repetition's potential and mu, no SC loop)
reproduces that repetition's total energy to the loop's tolerance."""
import re

import numpy as np
import pytest

from mpdpft_amd import Context, configs, run_dir

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def noisy(maxsc, schedule=(0,), noise=0.05, seed=24680):
    p = configs.c0(steps=32, edge=8.0, N=6.0, maxsc=maxsc)
    p.Noise = noise
    p.NoiseSeed = seed
    p.n_schedule = len(schedule)
    for i, s in enumerate(schedule):
        p.Schedule[i] = s
    return p


def test_initial_noise_matches_oracle(oracle):
    print("case initial noise, no SC loop", flush=True)
    p = noisy(0)
    rc, ro, dO, VO = oracle.solve(p)
    ctx = Context(p)
    ctx.solve()
    dG = ctx.density()
    ctx.close()
    assert rel(dG, dO) < 1e-13


# Pulay mixing on the rough densities that schedules 7 and 11 keep injecting is
# ill conditioned: in the oracle alone a 1e-14 relative change of the noise
# amplitude moves the density after 6-20 iterations by 3e-12 to 6e-3 (Pulay
# extrapolates from nearly dependent residuals), against < 1e-15 with linear mixing.  Those two
# run with linear mixing; 0, 3 and 4+5 (changes < 4e-14) with Pulay.
@pytest.mark.parametrize("schedule,mixer", [((0,), 1.), ((3,), 1.), ((4, 5), 1.), ((7,), 0.), ((11,), 0.)])
def test_noisy_loops_match_oracle(oracle, schedule, mixer):
    print("case noisy loop, schedule", schedule, "mixer", mixer, flush=True)
    p = noisy(200 if schedule == (0,) else 14, schedule)
    p.Mixer[0] = mixer
    rc, ro, dO, VO = oracle.solve(p)
    ctx = Context(p)
    rg = ctx.solve()
    dG, VG = ctx.density(), ctx.potential()
    ctx.close()
    assert rg.SCcount == ro.SCcount
    if schedule == (0,):
        assert rg.converged
    assert rel(dG, dO) < 1e-10
    assert rel(VG, VO) < 1e-10
    for s in range(p.S):
        assert abs(rg.muVec[s] - ro.muVec[s]) <= 1e-10 * abs(ro.muVec[s])
    assert abs(rg.TotalEnergy - ro.TotalEnergy) <= 1e-10 * abs(ro.TotalEnergy)


def test_task_type_10_repetitions(tmp_path, oracle):
    print("case TaskType 10", flush=True)
    p = noisy(200, noise=0.2, seed=13579)
    p.TaskType = 10
    p.TaskParameterCount = 3
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    r = run_dir(tmp_path)
    text = (tmp_path / "mpDPFT_ControlTask.dat").read_text()
    hist = [[float(v) for v in h.split()] for h in re.findall(r"Etot history \(sorted\) = ([^\n]*)", text)]
    assert [len(h) for h in hist] == [1, 2, 3]
    energies = hist[-1]
    assert energies == sorted(energies)
    assert len(set(energies)) == 3  # the repetitions differ by their noise
    mins = [float(v) for v in re.findall(r"new minEtot = (\S+)", text)]
    assert mins[0] == hist[0][0] and mins[-1] == min(energies)
    # each repetition alone: the solve with that repetition's seed
    best, best_e = None, None
    for k in range(3):
        q = noisy(200, noise=0.2, seed=13579 + k)
        ctx = Context(q)
        rk = ctx.solve()
        ctx.close()
        assert rk.converged
        assert min(abs(e - rk.TotalEnergy) for e in energies) <= 1e-15 * abs(rk.TotalEnergy)  # 16 printed digits
        if best is None or rk.TotalEnergy < best_e:
            best, best_e = k, rk.TotalEnergy
    # the final run: the oracle's densities of the best repetition's potential and mu, no SC loop (its energy
    # differs from the repetition's own by ~1e-7 relative: the loop stopped at max|dV| < 1e-8)
    _, ro, _, VO = oracle.solve(noisy(200, noise=0.2, seed=13579 + best))
    q = noisy(0, noise=0.0)
    for s in range(q.S):
        q.muVec[s] = ro.muVec[s]
    _, rf, dF, _ = oracle.solve(q, VO)
    assert abs(r.TotalEnergy - rf.TotalEnergy) <= 1e-10 * abs(rf.TotalEnergy)
    assert abs(r.TotalEnergy - best_e) <= 1e-6 * abs(best_e)
    assert r.SCcount == 0
    den = np.loadtxt(tmp_path / "mpDPFT_Den.dat")
    assert den.shape[0] == (p.steps + 1) ** 2
