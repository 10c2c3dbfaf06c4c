"""Task loops of mpdpft_run_dir (mpDPFTmain.cpp:77-214 with RetrieveTask,
mpDPFT.cpp:6070-6121): successive refinement (type 8) and the rectangular
grid of starting chemical potentials (type 1)."""
import re

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


def write_reference_table(path, n, edge, f):
    """TabFunc_ReferenceDensity.dat layout: n x n values, one per line, the x index
    fastest, on the lattice -edge/2 + edge i / (n - 1); no newline after the last."""
    x = -edge / 2 + edge * np.arange(n) / (n - 1)
    vals = [f(x[i], x[j]) for j in range(n) for i in range(n)]
    with open(path, "w") as fh:
        fh.write("\n".join("%.17g" % v for v in vals))
    return x


@pytest.mark.parametrize("task_type", [7, 5])
def test_task_parameter_scan_figure_of_merit(tmp_path, task_type):
    """Types 7 (mpp[0]) and 5 (mpp[0], gammaH) over the HyperBox, three values per
    variable: every row of mpDPFT_TASKParameterExploration.dat is (FOM, variables),
    the FOM 1 - 2<n_0, ref>/(|n_0|^2 + |ref|^2) of a direct solve at those
    variables against the reference table interpolated bilinearly by scipy, and
    the last run's files carry it in the Control file."""
    from scipy.interpolate import RegularGridInterpolator
    print("case task scan", task_type, flush=True)
    edge = 8.0
    xt = write_reference_table(tmp_path / "TabFunc_ReferenceDensity.dat", 41, edge,
                               lambda x, y: max(0.0, 3.0 - 0.5 * (x * x + 2.0 * y * y)) / (2 * np.pi))
    p = configs.c0(steps=32, edge=edge, N=6.0, tol=1e-9, maxsc=400)
    p.S = 1
    p.n_interactions = 1
    p.Interactions[0] = 0
    p.TaskType = task_type
    p.TaskParameterCount = 3
    p.n_hyperbox = 4 if task_type == 5 else 2
    for i, v in enumerate([0.3, 0.7, 1.0, 2.0][:p.n_hyperbox]):
        p.TaskHyperBox[i] = v
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    run_dir(tmp_path)
    rows = np.loadtxt(tmp_path / "mpDPFT_TASKParameterExploration.dat", ndmin=2)
    assert rows.shape == ((3, 2) if task_type == 7 else (9, 3))
    table = np.loadtxt(tmp_path / "TabFunc_ReferenceDensity.dat").reshape(41, 41)  # [y][x]
    interp = RegularGridInterpolator((xt, xt), table.T, method="linear")
    x = np.linspace(-edge / 2, edge / 2, 33)
    X, Y = np.meshgrid(x, x, indexing="ij")
    # the reference stores the value at (x_i, y_j) at grid index j EL + i
    ref = interp(np.stack([Y.ravel(), X.ravel()], -1))
    for row in rows:
        q = configs.c0(steps=32, edge=edge, N=6.0, tol=1e-9, maxsc=400)
        q.S = 1
        q.n_interactions = 1
        q.Interactions[0] = 0
        q.mpp[0] = row[1]
        if task_type == 5:
            q.gammaH = row[2]
        ctx = Context(q)
        ctx.solve()
        n0 = ctx.density()[0]
        ctx.close()
        fom = 1 - 2 * np.dot(n0, ref) / (np.dot(n0, n0) + np.dot(ref, ref))
        assert abs(row[0] - fom) <= 1e-12, (row, fom)
    assert np.allclose(rows[:3, 1] if task_type == 7 else rows[::3, 1], [0.3, 0.5, 0.7], rtol=1e-15)
    control = (tmp_path / "mpDPFT_Control.dat").read_text()
    m = re.search(r"FigureOfMerit\s+= ([^\\\\]+)", control)
    assert m and abs(float(m.group(1)) - rows[-1, 0]) <= 1e-15 * abs(rows[-1, 0])
