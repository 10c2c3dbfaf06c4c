"""HIP path vs the oracle on the same inputs (the parity gate).

Integer/count results must agree exactly (iteration counts within +-1 for a
converged solve); floating point within 1e-10 relative (the FP64 tolerance of
the drop-in: density, potential, mu and total energy).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import mpdpft_amd
from mpdpft_amd import Context, configs

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rand_field(G, seed, cplx=False):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=G)
    if cplx:
        f = f + 1j * rng.normal(size=G)
    return f


# ------------------------------------------------------------------ stages
@pytest.mark.parametrize("D,steps", [(2, 16), (2, 64), (2, 128), (3, 16), (3, 32), (3, 96), (3, 128)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_centred(oracle, D, steps, sign):
    print("case fft D=%d steps=%d sign=%d" % (D, steps, sign), flush=True)
    p = configs.c0(steps=steps) if D == 2 else configs.c1(steps=steps)
    p.HartreeMode = 0
    p.Interactions[0] = 0
    p.n_interactions = 1
    G = (steps + 1) ** D
    f = rand_field(G, steps, cplx=True)
    # a centred field must be periodic for the backward test to be meaningful
    ctx = Context(p)
    a = ctx.fft(f, sign)
    b = oracle.fft_centered(p, f, sign)
    assert rel(a, b) < 1e-12
    ctx.close()


@pytest.mark.parametrize("D,steps,method", [(3, 32, 5), (3, 64, 5), (3, 32, 4), (2, 64, 5), (3, 96, 5)])
def test_delfield(oracle, D, steps, method):
    print("case delfield D=%d steps=%d method=%d" % (D, steps, method), flush=True)
    p = configs.c1(steps=steps) if D == 3 else configs.c0(steps=steps, edge=16.0)
    p.DelFieldMethod = method
    ctx = Context(p)
    env = oracle.env(p)[0]
    f = env + 0.3 * np.sin(np.arange(env.size) * 0.001)
    a = ctx.delfield(f)
    b = oracle.delfield(p, f, method=method)
    for k in ("dx", "dy", "gradsq", "laplacian") + (("dz",) if D == 3 else ()):
        assert rel(a[k], b[k]) < 1e-10, k
    ctx.close()


@pytest.mark.parametrize("steps", [64, 128])
def test_fft_block_shape_does_not_change_results(monkeypatch, steps):
    """The FFT autotuner picks the lines per block of each transform length by
    timing, which varies from run to run; DelField and the isolated Hartree
    potential must not depend on that choice, bit for bit (C1 sizes: 128^3
    and its 256^3 padding)."""
    print("case fft block shapes", steps, flush=True)
    p = configs.c1(steps=steps, density_expression=5, hartree_mode=1)
    x = np.linspace(-1.0, 1.0, p.steps + 1)
    r2 = (x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2).reshape(-1)
    den = 3.0 * np.exp(-8.0 * r2)
    out = []
    for sel in ("0", "1", "2"):
        monkeypatch.setenv("MPDPFT_FFT_B", sel)
        ctx = Context(p)
        f = ctx.environment()[0] + 0.3 * np.sin(np.arange(ctx.G) * 0.001)
        d = ctx.delfield(f)
        out.append([d[k] for k in ("dx", "dy", "dz", "gradsq", "laplacian")] + [ctx.interaction_potential(8, 0, den)])
        ctx.close()
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("case", ["contact2d", "contact3", "dirac", "ldax", "pw92", "hartree_per", "hartree_iso",
                                  "hartree_iso2d", "hartree_per2d", "single_contact", "gombas", "vwn", "renorm"])
def test_interaction_potentials(oracle, case):
    print("case vint", case, flush=True)
    if case in ("contact2d", "hartree_iso2d", "hartree_per2d", "renorm"):
        p = configs.c0(steps=64, edge=12.0)
    elif case == "contact3":
        p = configs.c0(steps=32, edge=12.0)
        p.S = 3
        p.Abundances[2] = 50.0
        p.Environments[2] = 1
        p.mpp[2] = 0.5
        p.n_mpp = 3
    else:
        p = configs.c1(steps=32, edge=12.0, hartree_mode=1 if case.startswith("hartree_iso") else 0)
    typ = {"contact2d": 1, "contact3": 1, "dirac": 11, "ldax": -1, "pw92": -12, "hartree_per": 8,
           "hartree_iso": 8, "hartree_iso2d": 8, "hartree_per2d": 8, "single_contact": 10, "gombas": 12,
           "vwn": 13, "renorm": 9}[case]
    if case == "renorm":
        p.beta = 0.05  # 2D scattering length: 4 pi a^2 n stays below 1 (repulsive branch)
    if case in ("gombas", "vwn"):
        p.Units = 2  # the reference evaluates these potentials in 3D with Units 2 only
    p.n_interactions = 1
    p.Interactions[0] = typ
    if case == "hartree_iso2d":
        p.HartreeMode = 1
    G = (p.steps + 1) ** p.DIM
    den = np.abs(rand_field(p.S * G, 11)).reshape(p.S, G)
    ctx = Context(p)
    for s in range(p.S):
        a = ctx.interaction_potential(typ, s, den)
        b = oracle.vint(p, typ, s, den)
        assert rel(a, b) < 1e-12, s
    ctx.close()


@pytest.mark.parametrize("D,steps", [(2, 64), (3, 32)])
def test_tf_density_and_integral(oracle, D, steps):
    p = configs.c0(steps=steps) if D == 2 else configs.c1(steps=steps)
    ctx = Context(p)
    v = oracle.env(p)[0]
    mu = 4.0 if D == 2 else -0.3
    a, Na = ctx.density_at(0, mu, v)
    b, Nb = oracle.density(p, 0, mu, v)
    assert rel(a, b) < 1e-14
    assert abs(Na - Nb) < 1e-13 * abs(Nb)
    ctx.close()


def test_environments(oracle):
    p = configs.c0(steps=32)
    p.Environments[1] = 3
    p.n_mpp = 4
    p.mpp[2], p.mpp[3] = 0.2, -0.3
    ctx = Context(p)
    assert rel(ctx.environment(), oracle.env(p)) < 1e-15
    ctx.close()
    q = configs.c1(steps=32)
    q.Environments[0] = 41
    q.n_centres = 3
    for c, row in enumerate([(1.0, 2.0, -1.0, 1.0, 0.6), (0.0, 0.0, 0.0, 2.0, 0.4), (-3.0, 1.0, 2.0, 1.0, 0.6)]):
        for k, v in enumerate(row):
            q.centres[c][k] = v
    ctx = Context(q)
    assert rel(ctx.environment(), oracle.env(q)) < 1e-15
    ctx.close()


# ------------------------------------------------------------------ full loop
def solve_both(oracle, p):
    rc, ro, dO, VO = oracle.solve(p)
    ctx = Context(p)
    rg = ctx.solve()
    dG, VG = ctx.density(), ctx.potential()
    ctx.close()
    return ro, dO, VO, rg, dG, VG


def check_loop(ro, dO, VO, rg, dG, VG, S, exact_count=True):
    if exact_count:
        assert rg.SCcount == ro.SCcount
    else:
        assert abs(rg.SCcount - ro.SCcount) <= 1
    assert rel(dG, dO) < RTOL
    assert rel(VG, VO) < RTOL
    for s in range(S):
        assert abs(rg.muVec[s] - ro.muVec[s]) <= RTOL * abs(ro.muVec[s])
        assert abs(rg.FinalAbundances[s] - ro.FinalAbundances[s]) <= RTOL * abs(ro.FinalAbundances[s])
    assert abs(rg.TotalEnergy - ro.TotalEnergy) <= RTOL * abs(ro.TotalEnergy)


@pytest.mark.parametrize("steps,mode", [(32, 1), (64, 1), (128, 1), (128, 0)])
def test_c0_converged_solve(oracle, steps, mode):
    print("case c0 steps=%d mode=%d" % (steps, mode), flush=True)
    # the overlap rule CC ~ (dn/n)^2 has a rounding floor near 1e-16, where the
    # exit iteration is decided by noise: test it well above the floor
    p = configs.c0(steps=steps, convergence_mode=mode, tol=1e-8 if mode == 1 else 1e-13)
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    assert ro.converged and rg.converged
    check_loop(ro, dO, VO, rg, dG, VG, p.S, exact_count=False)
    assert (rg.LinMix, rg.PulMix) == (ro.LinMix, ro.PulMix)


@pytest.mark.parametrize("hmode,iters", [(1, 3), (0, 3)])
def test_3d_tf_hartree_xc_trajectory(oracle, hmode, iters):
    """C1 physics with the TF density: the same K iterations on both paths.  The
    TF-Dirac neutral atom does not converge (its density edge is discontinuous)
    and its mixed iteration amplifies rounding differences ~7x per step, so
    the trajectory is compared over a few steps only."""
    print("case 3d traj hmode=%d iters=%d" % (hmode, iters), flush=True)
    p = configs.c1(steps=32, density_expression=7, hartree_mode=hmode, maxsc=iters)
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    assert ro.SCcount == iters + 1 and rg.SCcount == iters + 1
    check_loop(ro, dO, VO, rg, dG, VG, 1)
    assert rg.bisection_steps == ro.bisection_steps


@pytest.mark.parametrize("maxsc", [4, 400])
def test_2d_periodic_hartree_loop(oracle, maxsc):
    """2D two-species TF gas with contact and the periodic Hartree term (Bessel-I0
    compensation, mpDPFT.cpp:2627-2659): a 4-iteration trajectory and a converged
    solve on both paths."""
    print("case 2d periodic hartree", maxsc, flush=True)
    p = configs.c0(steps=64, edge=16.0, maxsc=maxsc)
    p.n_interactions = 2
    p.Interactions[1] = 8
    p.HartreeMode = 0
    p.gammaH = 0.05
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    if maxsc == 4:
        assert ro.SCcount == 5 and rg.SCcount == 5
        check_loop(ro, dO, VO, rg, dG, VG, p.S)
    else:
        assert ro.converged and rg.converged
        check_loop(ro, dO, VO, rg, dG, VG, p.S, exact_count=False)


def test_2d_renormalised_contact_trajectory(oracle):
    """2D two-species TF gas with the renormalised contact interaction (mpDPFT.cpp:2891-2904)
    instead of the plain contact: 4 iterations on both paths, with its energy."""
    print("case renormalised contact trajectory", flush=True)
    p = configs.c0(steps=64, edge=16.0, maxsc=4)
    p.Interactions[0] = 9
    p.beta = 0.05
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    assert ro.SCcount == 5 and rg.SCcount == 5
    check_loop(ro, dO, VO, rg, dG, VG, p.S)
    assert ro.InteractionEnergies[0] != 0.0
    assert abs(rg.InteractionEnergies[0] - ro.InteractionEnergies[0]) <= 1e-10 * abs(ro.InteractionEnergies[0])


def test_3d_gombas_vwn_trajectory(oracle):
    """3D TF in Units 2 with isolated Hartree, Dirac exchange and the Gombas and VWN
    correlation terms (potentials and energies): 3 iterations on both paths."""
    print("case gombas vwn trajectory", flush=True)
    p = configs.c1(steps=16, density_expression=7, hartree_mode=1, maxsc=3)
    p.Units = 2
    p.n_interactions = 4
    for k, t in enumerate((8, 11, 12, 13)):
        p.Interactions[k] = t
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    assert ro.SCcount == 4 and rg.SCcount == 4
    check_loop(ro, dO, VO, rg, dG, VG, 1)
    for k in range(4):
        assert abs(rg.InteractionEnergies[k] - ro.InteractionEnergies[k]) <= 1e-10 * abs(ro.InteractionEnergies[k]) + 1e-14
        # GetAuxEint (S = 1, 3D, Units 2): Dirac, its potential form, Gombas, VWN
        assert ro.AuxEint[0][k] != 0.0
        assert abs(rg.AuxEint[0][k] - ro.AuxEint[0][k]) <= 1e-10 * abs(ro.AuxEint[0][k])


def test_3d_tf_ion_converges(oracle):
    """A positive ion (N < Z) converges in both implementations."""
    p = configs.c1(steps=32, density_expression=7, hartree_mode=1, N=6.0, maxsc=600)
    p.Interactions[1] = 0
    p.Interactions[2] = 0
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    assert ro.converged and rg.converged
    check_loop(ro, dO, VO, rg, dG, VG, 1, exact_count=False)


# ------------------------------------------------------------------ drop-in
def test_drop_in_directory_run(oracle, tmp_path):
    p = configs.c0(steps=32)
    configs.write_input(p, tmp_path / "mpDPFT.input")
    cli = os.path.join(os.path.dirname(mpdpft_amd.__file__), "bin", "mpdpft_hip")
    out = subprocess.run([cli, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    G = 33 * 33
    den_lines = (tmp_path / "mpDPFT_Den.dat").read_text().splitlines()
    assert len(den_lines) == G
    assert all(len(l.split(" ")) == 2 and not l.endswith(" ") for l in den_lines)
    v_lines = (tmp_path / "mpDPFT_V.dat").read_text().splitlines()
    assert len(v_lines) == G and all(l.endswith(" ") and len(l.split()) == 4 for l in v_lines)
    assert v_lines[0].split()[:2] == ["-16", "-16"]
    ctl = (tmp_path / "mpDPFT_Control.dat").read_text()
    for key in ("FinalAbundances", "muVec", "TotalEnergy", "SCiterations", "ConvergenceCheck", "AuxilliaryEint[s=1]",
                "AlternativeDispersalEnergies", "FigureOfMerit"):
        assert key in ctl
    rc, ro, dO, VO = oracle.solve(p)
    den = np.array([[float(x) for x in l.split()] for l in den_lines]).T
    assert rel(den, dO) < RTOL
    V = np.array([[float(x) for x in l.split()[2:]] for l in v_lines]).T
    assert rel(V, VO) < RTOL
    it = [l for l in ctl.split("\\\\") if l.startswith("SCiterations")][0]
    assert it.split("=")[1].strip() == "%d/%d" % (ro.SCcount - 1, p.maxSCcount)
    check_plot_files(tmp_path, p, den, V)


def check_plot_files(d, p, den, V):
    """mpDPFT_ContourData.dat and mpDPFT_CutData.dat (ExportData, mpDPFT.cpp:12527-12660)
    against the run's own Den/V files and the environment, recomputed here: the
    x-y plane (3D: z index steps/2) with a blank line before each new x row, and
    the default cut along the x axis, where the interpolated fields are the grid
    values of row y = 0 (3D: plus the lattice coordinate and the column density
    sum_k n dx)."""
    EL, S, D = p.steps + 1, p.S, p.DIM
    ctx = Context(p)
    env = ctx.environment().reshape(S, -1)
    ctx.close()
    x = -p.edge / 2 + np.arange(EL) * (p.edge / p.steps)
    plane = lambda f, i, j: f[..., i * EL + j] if D == 2 else f[..., (i * EL + j) * EL + p.steps // 2]
    lines = (d / "mpDPFT_ContourData.dat").read_text().split("\n")
    assert lines[-1] == "" and lines[-2] == ""  # the closing blank line
    blocks = "\n".join(lines).strip("\n").split("\n\n")
    assert len(blocks) == EL
    for i, blk in enumerate(blocks):
        rows = np.array([[float(t) for t in l.split(" ")] for l in blk.split("\n")])
        assert rows.shape == (EL, 3 + 3 * S)
        zc = x[p.steps // 2] if D == 3 else 0.0
        assert np.allclose(rows[:, 0], x[i], rtol=0, atol=1e-13) and np.allclose(rows[:, 1], x, rtol=0, atol=1e-13)
        assert np.allclose(rows[:, 2], np.sqrt(x[i] ** 2 + x ** 2 + zc ** 2), rtol=1e-14, atol=1e-13)
        for s in range(S):
            j = np.arange(EL)
            assert np.array_equal(rows[:, 3 + 3 * s], plane(den[s], i, j))
            assert np.allclose(rows[:, 4 + 3 * s], plane(env[s], i, j), rtol=1e-15, atol=0)
            assert np.array_equal(rows[:, 5 + 3 * s], plane(V[s], i, j))
    cut = np.array([[float(t) for t in l.split(" ")] for l in (d / "mpDPFT_CutData.dat").read_text().splitlines()])
    assert cut.shape == (EL, 1 + 3 * S + (1 + S if D == 3 else 0))
    assert np.allclose(cut[:, 0], x, rtol=0, atol=1e-13)
    i = np.arange(EL)
    for s in range(S):
        for k, f in enumerate((den[s], env[s], V[s])):
            ref = plane(f, i, p.steps // 2)
            assert np.allclose(cut[:, 1 + 3 * s + k], ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    if D == 3:
        assert np.allclose(cut[:, 1 + 3 * S], x, rtol=0, atol=1e-13)
        cube = den.reshape(S, EL, EL, EL)
        for s in range(S):
            col = cube[s, :, p.steps // 2, :].sum(axis=1) * (p.edge / p.steps)
            assert np.allclose(cut[:, 2 + 3 * S + s], col, rtol=1e-13, atol=0)


def test_drop_in_plot_files_3d(tmp_path):
    """The 3D form of the plot files: contour plane at z index steps/2, cut with
    the lattice and column-density columns (TF + Hartree ion, 16^3, 3 iterations)."""
    print("case drop-in plot files 3d", flush=True)
    p = configs.c1(steps=16, density_expression=7, hartree_mode=1, maxsc=3)
    p.n_interactions = 1
    configs.write_input(p, tmp_path / "mpDPFT.input")
    cli = os.path.join(os.path.dirname(mpdpft_amd.__file__), "bin", "mpdpft_hip")
    out = subprocess.run([cli, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode in (0, 3), out.stderr  # 3: not converged after maxSCcount (files written)
    den = np.array([[float(t) for t in l.split()] for l in (tmp_path / "mpDPFT_Den.dat").read_text().splitlines()]).T
    V = np.array([[float(t) for t in l.split()[3:]] for l in (tmp_path / "mpDPFT_V.dat").read_text().splitlines()]).T
    check_plot_files(tmp_path, p, den, V)


def test_warm_start_from_v_file(oracle, tmp_path):
    """InterpolVQ = 1: the CLI reads mpDPFT_V.dat of a coarser run, interpolates
    it onto its own lattice and starts the loop from it (LoadV, mpDPFT.cpp:184,
    :6445-6502); the oracle runs the same loop from the same imported field."""
    print("case warm start", flush=True)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    cli = os.path.join(os.path.dirname(mpdpft_amd.__file__), "bin", "mpdpft_hip")
    p = configs.c0(steps=32)
    configs.write_input(p, a / "mpDPFT.input")
    assert subprocess.run([cli, str(a)], capture_output=True, timeout=120).returncode == 0
    q = configs.c0(steps=64)
    q.InterpolVQ = 1
    configs.write_input(q, b / "mpDPFT.input")
    shutil.copy(a / "mpDPFT_V.dat", b / "mpDPFT_V.dat")
    v0, warn = mpdpft_amd.import_v(b / "mpDPFT_V.dat", q)
    out = subprocess.run([cli, str(b)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    ctl = (b / "mpDPFT_Control.dat").read_text()
    assert "LoadV: Interpolate V from mpDPFT_V.dat" in ctl
    rc, ro, dO, VO = oracle.solve(q, v0=v0)
    den = np.array([[float(x) for x in l.split()] for l in (b / "mpDPFT_Den.dat").read_text().splitlines()]).T
    assert rel(den, dO) < RTOL
    it = [l for l in ctl.split("\\\\") if l.startswith("SCiterations")][0]
    assert it.split("=")[1].strip() == "%d/%d" % (ro.SCcount - 1, q.maxSCcount)


def test_manual_break_file(tmp_path):
    """ManualSCloopBreakQ (mpDPFT.cpp:11959-11976): a 1 in the break file ends
    the loop after the current iteration and the file is reset to 0."""
    print("case break file", flush=True)
    p = configs.c0(steps=32)
    configs.write_input(p, tmp_path / "mpDPFT.input")
    (tmp_path / "mpDPFTmanualSCloopBreakQ.dat").write_text("1\n")
    cli = os.path.join(os.path.dirname(mpdpft_amd.__file__), "bin", "mpdpft_hip")
    out = subprocess.run([cli, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 3, out.stderr  # not converged
    ctl = (tmp_path / "mpDPFT_Control.dat").read_text()
    assert "SCloop--Manually--Terminated" in ctl
    it = [l for l in ctl.split("\\\\") if l.startswith("SCiterations")][0]
    assert it.split("=")[1].strip() == "0/%d" % p.maxSCcount
    assert (tmp_path / "mpDPFTmanualSCloopBreakQ.dat").read_text().split()[0] == "0"


@pytest.mark.parametrize("schedule", [5, 9])
def test_schedule_retunes_theta(oracle, schedule):
    """InjectSchedule 5 and 9 (mpDPFT.cpp:5261-5285) on both paths."""
    print("case schedule", schedule, flush=True)
    p = configs.c0(steps=32, maxsc=40)
    p.n_schedule = 1
    p.Schedule[0] = schedule
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    check_loop(ro, dO, VO, rg, dG, VG, 2)


def test_not_converged_is_reported(oracle):
    p = configs.c0(steps=32, maxsc=5)
    ctx = Context(p)
    r = ctx.solve()
    assert r.converged == 0 and r.SCcount == 6
    ctx.close()


# ------------------------------------------------------------------ nAiT (C1 functional)
@pytest.mark.parametrize("kernel,steps,bp", [("lanes", 16, 10), ("lanes", 16, 15), ("lanes", 16, 3),
                                              ("reuse", 16, 10), ("plain", 16, 10), ("reuse", 16, 17),
                                              ("lanes", 32, 10)])
def test_nait_density_stage(oracle, monkeypatch, kernel, steps, bp):
    """nAiT density at fixed V and mu against the oracle: the fixed-lane-layout
    kernel (default for BP <= 15), the panel-reuse kernel for any BP (one or two
    samples per lane and half; BP = 17 also takes it by default) and the
    per-node kernel."""
    print("case nait density stage", kernel, steps, bp, flush=True)
    monkeypatch.setenv("MPDPFT_NAIT_KERNEL", kernel)
    p = configs.c1(steps=steps, density_expression=5, hartree_mode=1)
    p.Mpp[0] = bp
    ctx = Context(p)
    v = oracle.env(p)[0] + 0.05 * np.cos(np.arange((p.steps + 1) ** 3) * 0.01)
    for mu in (-0.6, -0.2):
        a, Na = ctx.density_at(0, mu, v)
        b, Nb = oracle.density(p, 0, mu, v)
        assert rel(a, b) < RTOL
        assert abs(Na - Nb) <= RTOL * abs(Nb)
    ctx.close()


@pytest.mark.parametrize("iters", [3])
def test_nait_trajectory(oracle, iters):
    """C1 (nAiT + isolated Hartree + Dirac + PW92) at 16^3: K iterations on both paths."""
    print("case nait trajectory", iters, flush=True)
    p = configs.c1(steps=16, density_expression=5, hartree_mode=1, maxsc=iters)
    ro, dO, VO, rg, dG, VG = solve_both(oracle, p)
    assert ro.SCcount == iters + 1 and rg.SCcount == iters + 1
    check_loop(ro, dO, VO, rg, dG, VG, 1)
    assert rg.bisection_steps == ro.bisection_steps


@pytest.mark.gpu
def test_airy_energy_tail_is_exact(monkeypatch):
    """The Airy kinetic energy leaves out the trapezoid nodes where Ai is below
    1e-37 of its scale; summing every node (MPDPFT_EKIN_FULL=1, the reference's
    loop) gives the same dispersal energy to the bit."""
    print("case airy energy tail", flush=True)
    p = configs.c1(steps=32, density_expression=5, hartree_mode=1, maxsc=3)
    out = []
    for full in ("0", "1"):
        monkeypatch.setenv("MPDPFT_EKIN_FULL", full)
        ctx = Context(p)
        r = ctx.solve()
        ctx.close()
        out.append((r.DispersalEnergies[0], r.TotalEnergy))
        print("full=%s  E_disp %.17g  E %.17g" % (full, out[-1][0], out[-1][1]), flush=True)
    assert out[0] == out[1]
