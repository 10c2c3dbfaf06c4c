"""GPU parity of the 2D-3D crossover (14), the quasi-2D contact (24), and the
dipolar interactions in position (17) and momentum space (15) against the
oracle: each potential alone at 1e-12, and short SC trajectories whose
interaction energies enter the total energy (density, potential, mu, N and E
at 1e-10, identical iteration counts)."""
import os
import shutil

import numpy as np
import pytest

from mpdpft_amd import Context, configs, set_data_dir

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ETA = os.path.join(HERE, "golden", "TabFunc_2D3D_eta_l5.dat")
RTOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture
def eta_dir(tmp_path, oracle):
    """The reference's eta table (levels l1 <= 5) as TabFunc_2D3D_eta.dat in a
    data directory, for the library and the oracle."""
    shutil.copy(ETA, tmp_path / "TabFunc_2D3D_eta.dat")
    set_data_dir(str(tmp_path))
    oracle.load_eta(str(tmp_path / "TabFunc_2D3D_eta.dat"))
    yield tmp_path
    set_data_dir("")


def p2d(interactions, steps=64, S=2, alpha=1.3, beta=0.4):
    p = configs.c0(steps=steps, edge=12.0, maxsc=4)
    p.S = S
    p.alpha, p.beta = alpha, beta
    p.n_interactions = len(interactions)
    for a, t in enumerate(interactions):
        p.Interactions[a] = t
    return p


def p3d(interactions, steps, alpha):
    p = configs.c1(steps=steps, edge=12.0, hartree_mode=0, maxsc=3)
    p.Units = 1
    p.alpha = alpha
    p.Abundances[0] = 3.0
    p.n_interactions = len(interactions)
    for a, t in enumerate(interactions):
        p.Interactions[a] = t
    return p


@pytest.mark.parametrize("case", ["crossover", "quasi2d", "dipolar_pos", "dipolar_k"])
def test_interaction_potentials_ext(oracle, eta_dir, case):
    print("case vint ext", case, flush=True)
    rng = np.random.default_rng(17)
    if case in ("crossover", "quasi2d"):
        typ = 14 if case == "crossover" else 24
        p = p2d([typ])
        G = (p.steps + 1) ** 2
        # densities over the first five bands of transverse levels (alpha = 1.3: edges 0.21, 0.62, ...)
        den = rng.uniform(0., 3.0, size=(p.S, G))
    else:
        typ = 17 if case == "dipolar_pos" else 15
        p = p3d([typ], 32 if typ == 17 else 16, 0.7)
        G = (p.steps + 1) ** 3
        x = np.linspace(-6, 6, p.steps + 1)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        r2 = (X ** 2 + Y ** 2 + 1.5 * Z ** 2).ravel()
        den = (np.exp(-r2 / 6) * (1 + 0.1 * rng.normal(size=G))).reshape(1, G)
    ctx = Context(p)
    for s in range(p.S):
        a = ctx.interaction_potential(typ, s, den)
        b = oracle.vint(p, typ, s, den)
        assert rel(a, b) < 1e-12, s
    ctx.close()


@pytest.mark.parametrize("case", ["2d_contact_crossover_quasi2d", "3d_tf_dipolar_pos", "3d_tf_dipolar_k"])
def test_loops_with_new_interactions(oracle, eta_dir, case):
    """A few SC iterations on both paths; the total energy carries the
    crossover and quasi-2D contact energies (the dipolar ones are 0 in the
    reference's GetInteractionEnergy).  The 3D dipolar cases mix linearly."""
    print("case loop ext", case, flush=True)
    if case.startswith("2d"):
        p = p2d([1, 14, 24], steps=64)
    else:
        # linear mixing: with the weak dipolar term the first Pulay Gram matrices are
        # nearly singular and turn rounding-level differences of V (3e-15) into 1e-9
        # differences of the mixed density by the fourth iteration
        p = p3d([17], 32, 0.05) if case == "3d_tf_dipolar_pos" else p3d([15], 16, 0.05)
        p.Mixer[0] = 0.0
    rc, ro, dO, VO = oracle.solve(p)
    ctx = Context(p)
    rg = ctx.solve()
    dG, VG = ctx.density(), ctx.potential()
    ctx.close()
    assert rg.SCcount == ro.SCcount
    assert rel(dG, dO) < RTOL
    assert rel(VG, VO) < RTOL
    for s in range(p.S):
        assert abs(rg.muVec[s] - ro.muVec[s]) <= RTOL * abs(ro.muVec[s])
    assert abs(rg.TotalEnergy - ro.TotalEnergy) <= RTOL * abs(ro.TotalEnergy)
    for a in range(p.n_interactions):
        assert abs(rg.InteractionEnergies[a] - ro.InteractionEnergies[a]) <= RTOL * max(
            abs(ro.InteractionEnergies[a]), 1e-30)
    if case.startswith("2d"):
        assert ro.InteractionEnergies[1] != 0. and ro.InteractionEnergies[2] != 0.


def test_quasi2d_needs_its_table(tmp_path):
    """Interaction 24 without TabFunc_2D3D_eta.dat in the data directory is an
    I/O error at create, not a silent zero."""
    set_data_dir(str(tmp_path))
    try:
        with pytest.raises(Exception):
            Context(p2d([24]))
    finally:
        set_data_dir("")
