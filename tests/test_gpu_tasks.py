"""Task loops of mpdpft_run_dir (mpDPFTmain.cpp:77-214 with RetrieveTask,
mpDPFT.cpp:6070-6121): successive refinement (type 8) and the rectangular
grid of starting chemical potentials (type 1)."""
import numpy as np
import pytest

from mpdpft_amd import Context, configs, run_dir

pytestmark = pytest.mark.gpu


def read_den(path):
    return np.loadtxt(path, ndmin=2)


def test_task_type8_successive_refinement(tmp_path):
    """32^2 -> 64^2 -> 128^2, each task starting from the previous task's V file
    (LoadV) and mu: the last task reaches the fixed point of a direct 128^2 solve
    and writes 129^2 grid files."""
    p = configs.c0(steps=32, tol=1e-10)
    p.TaskType = 8
    p.TaskParameterCount = 3
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    r = run_dir(tmp_path)
    assert r.converged
    den = read_den(tmp_path / "mpDPFT_Den.dat")
    assert den.shape == (129 * 129, 2)
    lines = (tmp_path / "mpDPFT_ControlTask.dat").read_text().splitlines()
    assert len(lines) == 4 and "steps 128" in lines[-1]
    d = configs.c0(steps=128, tol=1e-10)
    ctx = Context(d)
    rd = ctx.solve()
    nd = ctx.density()
    ctx.close()
    assert abs(r.TotalEnergy - rd.TotalEnergy) <= 1e-8 * abs(rd.TotalEnergy)
    assert abs(r.muVec[0] - rd.muVec[0]) <= 1e-8 * abs(rd.muVec[0])
    assert np.abs(den.T - nd).max() <= 1e-7 * np.abs(nd).max()
    assert r.SCcount < rd.SCcount  # the warm start pays


def test_task_type1_mu_grid(tmp_path):
    """Fixed chemical potentials (RelAcc below MP): four tasks on a 2 x 2 grid of
    mu; the last task runs at mu_s (1 + delta), delta = 1.1^(1/2) - 1, and equals a
    direct solve there."""
    p = configs.c0(steps=32, tol=1e-10)
    p.TaskType = 1
    p.TaskParameterCount = 4
    p.RelAcc = 1e-13
    p.muVec[0] = p.muVec[1] = 14.0
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    r = run_dir(tmp_path)
    lines = (tmp_path / "mpDPFT_ControlTask.dat").read_text().splitlines()
    assert len(lines) == 5
    d = configs.c0(steps=32, tol=1e-10)
    d.RelAcc = 1e-13
    mu = 14.0 * (1.0 + (1.1 ** 0.5 - 1.0))
    d.muVec[0] = d.muVec[1] = mu
    ctx = Context(d)
    rd = ctx.solve()
    ctx.close()
    assert abs(r.muVec[0] - mu) <= 1e-12 * mu
    assert abs(r.TotalEnergy - rd.TotalEnergy) <= 1e-10 * abs(rd.TotalEnergy)
    assert abs(r.FinalAbundances[0] - rd.FinalAbundances[0]) <= 1e-10 * rd.FinalAbundances[0]
