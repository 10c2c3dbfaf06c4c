"""TaskType 2 and 3: the total energy's minimum over muVec (GetGlobalMinimum_muVec,
mpDPFT.cpp:246-253) or over the abundances (GetGlobalMinimum_Abundances
:256-285), which the reference finds with conjugate gradients on
finite-difference gradients of whole solves (ConjugateGradientDescent
:8365-8522, functionToBeMinimized :8524-8584).

The adjoint gradient (mpdpft_task_gradient: dE/dN with N held, dE/dmu with mu
held) is compared with the oracle's central differences of whole solves
(h = 1e-4), and the task loop's minimum (mpdpft_run_dir, TaskType 2 / 3) with
the oracle's: the root of its central-difference derivative found by secant
steps.  Workload: a 32^3 cluster of three softened Coulomb centres (Z = 1.5,
a = 0.7), TF kinetic energy and isolated Hartree, solves stopped at
max|dV| < 1e-11."""
import re

import pytest

from mpdpft_amd import Context, configs, run_dir

pytestmark = pytest.mark.gpu

POS3 = [[-1.1, 0.3, -0.2], [0.9, -0.6, 0.4], [0.2, 1.0, 0.5]]
H = 1e-4


def cluster(N=4.0, mu=None):
    p = configs.c4(steps=32, edge=14.0, N=N, positions=POS3, Z=1.5, a=0.7, density_expression=7, tol=1e-11,
                   maxsc=3000)
    p.muVec[0] = -1.0
    p.n_interactions = 1
    p.Interactions[0] = 8
    if mu is not None:  # the loop holds mu
        p.RelAcc = 0.0
        p.muVec[0] = mu
    return p


def oracle_energy(oracle, p):
    rc, r, _, _ = oracle.solve(p)
    assert rc == 0 and r.converged
    return r.TotalEnergy


def oracle_fd(oracle, make, x):
    return (oracle_energy(oracle, make(x + H)) - oracle_energy(oracle, make(x - H))) / (2 * H)


def oracle_minimum(oracle, make, x0, x1):
    """Secant steps on the oracle's central-difference derivative."""
    g0, g1 = oracle_fd(oracle, make, x0), oracle_fd(oracle, make, x1)
    for _ in range(30):
        x2 = x1 - g1 * (x1 - x0) / (g1 - g0)
        x0, g0, x1 = x1, g1, x2
        if abs(x1 - x0) < 1e-9:
            break
        g1 = oracle_fd(oracle, make, x1)
    return x1, oracle_energy(oracle, make(x1))


@pytest.mark.parametrize("N", [3.0, 4.0, 6.0])
def test_abundance_gradient_matches_oracle_differences(oracle, N):
    print("case dE/dN", N, flush=True)
    ctx = Context(cluster(N))
    r = ctx.solve()
    assert r.converged
    g = ctx.task_gradient()
    ctx.close()
    fd = oracle_fd(oracle, lambda n: cluster(n), N)
    assert abs(g - fd) <= 1e-6 * abs(fd), (N, g, fd)
    # near the loop's own chemical potential (the discrete energy is not exactly stationary)
    assert abs(g - r.muVec[0]) < 0.05 * max(abs(r.muVec[0]), 0.05)


@pytest.mark.parametrize("mu", [-0.1, 0.05])
def test_mu_gradient_matches_oracle_differences(oracle, mu):
    print("case dE/dmu", mu, flush=True)
    ctx = Context(cluster(mu=mu))
    r = ctx.solve()
    assert r.converged and r.muVec[0] == mu
    g = ctx.task_gradient()
    ctx.close()
    fd = oracle_fd(oracle, lambda m: cluster(mu=m), mu)
    assert abs(g - fd) <= 1e-6 * abs(fd), (mu, g, fd)


def task_dir(tmp_path, p, task_type, box):
    p.TaskType = task_type
    p.System = 1  # EpsF = 0: the loop stops on its scaled-gradient rule
    p.n_hyperbox = 2
    p.TaskHyperBox[0], p.TaskHyperBox[1] = box
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    r = run_dir(tmp_path)
    text = (tmp_path / "mpDPFT_ControlTask.dat").read_text()
    m = re.search(r"Emin = (\S+) @ (\w+) = (\S+)", text)
    assert m, text
    assert "#Gradient = adjoint" in text and "#TerminationType = 4" in text, text
    assert (tmp_path / "mpDPFT_Den.dat").exists()
    return r, float(m.group(1)), float(m.group(3)), text


@pytest.mark.parametrize("task_type", [3, 2])
def test_task_loop_minimum_matches_the_oracle(oracle, tmp_path, task_type):
    print("case task", task_type, flush=True)
    if task_type == 3:
        make = lambda n: cluster(n)  # noqa: E731
        r, E, x, text = task_dir(tmp_path, cluster(), 3, (3.0, 6.0))
        xo, Eo = oracle_minimum(oracle, make, 4.0, 4.5)
        assert abs(r.FinalAbundances[0] - x) < 1e-9 * x
    else:
        make = lambda m: cluster(mu=m)  # noqa: E731
        r, E, x, text = task_dir(tmp_path, cluster(mu=-0.06), 2, (-0.2, 0.1))
        xo, Eo = oracle_minimum(oracle, make, -0.06, 0.0)
        assert abs(r.muVec[0] - x) <= 1e-15 * abs(x)  # x as printed (16 digits)
    print(task_type, "minimum", x, E, "oracle", xo, Eo, flush=True)
    assert r.converged and E == r.TotalEnergy
    assert abs(x - xo) < 1e-5 * max(abs(xo), 0.1), (x, xo)
    assert abs(E - Eo) < 1e-10 * abs(Eo), (E, Eo)


def test_task_type_9_annealed_descent(oracle, tmp_path):
    """TaskType 9 (GradientDescent mpDPFT.cpp:8586-8710) on the same cluster,
    abundance box [3, 6], one descent (TaskParameterCount 1), seed 8642: the
    start is 3 + 3 u with u the generator's first draw (the oracle's MT19937-64
    gives it), the energies the log prints at its first iterates are the
    oracle's at those abundances, and the lowest printed energy reaches the
    oracle's minimum over N (secant on central differences) to 1e-8."""
    print("case task 9", flush=True)
    p = cluster()
    p.TaskType = 9
    p.TaskParameterCount = 1
    p.NoiseSeed = 8642
    p.n_hyperbox = 2
    p.TaskHyperBox[0], p.TaskHyperBox[1] = 3.0, 6.0
    configs.write_input(p, str(tmp_path / "mpDPFT.input"))
    run_dir(tmp_path)
    text = (tmp_path / "mpDPFT_ControlTask.dat").read_text()
    start = float(re.search(r"\[xStart = (\S+)\]", text).group(1))
    u = oracle.noise_draws(8642, 1, raw=True)[0] / 2.0 ** 64
    assert abs(start - (3.0 + u * 3.0)) <= 1e-15 * start
    its = re.findall(r"iter = (\d+)/200: f = (\S+)\n gamma = .*?---> xNew = (\S+)", text)
    assert len(its) >= 3, text[-3000:]
    xs = [start] + [float(t[2]) for t in its]
    fs = [float(t[1]) for t in its]
    for k in (0, 1, 2, len(its) - 1):
        Eo = oracle_energy(oracle, cluster(xs[k]))
        assert abs(fs[k] - Eo) <= 1e-10 * abs(Eo), (k, xs[k], fs[k], Eo)
    xo, Eo = oracle_minimum(oracle, lambda n: cluster(n), 4.0, 4.5)
    print("task 9: %d iterations, lowest f %.16g at %.10g; oracle minimum %.16g at %.10g" %
          (len(its), min(fs), xs[fs.index(min(fs))], Eo, xo), flush=True)
    assert Eo - 1e-10 * abs(Eo) <= min(fs) <= Eo + 1e-8 * abs(Eo)
    assert "Local minimum f = " in text and (tmp_path / "mpDPFT_Den.dat").exists()
