"""Implicit adjoint of the converged fixed point (mpdpft_adjoint,
mpdpft_centre_forces) against central finite differences of whole solves,
the oracle's FP64 finite differences, and the torch.autograd binding."""
import numpy as np
import pytest

from mpdpft_amd import Context, configs

pytestmark = pytest.mark.gpu

POS3 = [[-1.1, 0.3, -0.2], [0.9, -0.6, 0.4], [0.2, 1.0, 0.5]]


def weights(p):
    EL = p.steps + 1
    w1 = np.ones(EL)
    w1[0] = w1[-1] = 0.5
    dx = p.edge / p.steps
    return (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]).reshape(-1) * dx ** 3


def cluster(pos, de=7, edge=14.0, steps=32, tol=1e-11, inter=(8,)):
    p = configs.c4(steps=steps, edge=edge, N=4.0, positions=pos, Z=1.5, a=0.7, density_expression=de, tol=tol,
                   maxsc=3000)
    p.muVec[0] = -1.0
    p.n_interactions = len(inter)
    for k, t in enumerate(inter):
        p.Interactions[k] = t
    return p


def centres(pos):
    return np.array([[r[0], r[1], r[2], 1.5, 0.7] for r in pos], dtype=np.float64)


def fd_loss(ctx, pos, nt, w, ci, d, h, V0=None):
    out = []
    for sgn in (1, -1):
        c = centres(pos)
        c[ci, d] += sgn * h
        ctx.set_centres(c)
        if V0 is not None:
            ctx.set_initial_potential(V0)
        r = ctx.solve()
        assert r.converged
        n = ctx.density()[0]
        out.append((float(np.sum(w * (n - nt) ** 2)), r.TotalEnergy))
    ctx.set_centres(centres(pos))
    ctx.set_initial_potential(None)
    return (out[0][0] - out[1][0]) / (2 * h), (out[0][1] - out[1][1]) / (2 * h)


def target_density(p_of_pos):
    shifted = [[r[0] + 0.3, r[1] - 0.1, r[2]] for r in POS3]
    ct = Context(p_of_pos(shifted))
    r = ct.solve()
    assert r.converged
    nt = ct.density()[0].copy()
    ct.close()
    return nt


def test_adjoint_tf_hartree_matches_finite_differences():
    mk = lambda pos: cluster(pos)  # noqa: E731
    nt = target_density(mk)
    p = mk(POS3)
    w = weights(p)
    ctx = Context(p)
    r = ctx.solve()
    assert r.converged
    n = ctx.density()[0].copy()
    lam, info = ctx.adjoint(2.0 * w * (n - nt))
    assert info["converged"] and info["relative_residual"] < 1e-11, info
    g = ctx.centre_gradient(lam)
    for ci, d in [(0, 0), (1, 1), (2, 2), (0, 2)]:
        fd, _ = fd_loss(ctx, POS3, nt, w, ci, d, 1e-4)
        assert abs(g[ci, d] - fd) <= 1e-6 * abs(fd), (ci, d, g[ci, d], fd)
    ctx.close()


def test_adjoint_matches_oracle_fp64_gradient():
    """The oracle (plain C, FP64) solves at theta +- h: its central difference
    is the FP64 reference gradient."""
    from oracle import oracle as O
    mk = lambda pos: cluster(pos)  # noqa: E731
    shifted = [[r[0] + 0.3, r[1] - 0.1, r[2]] for r in POS3]
    _, rt, nt, _ = O.solve(mk(shifted))
    nt = nt[0]
    p = mk(POS3)
    w = weights(p)
    ctx = Context(p)
    ctx.solve()
    n = ctx.density()[0].copy()
    lam, info = ctx.adjoint(2.0 * w * (n - nt))
    g = ctx.centre_gradient(lam)
    ctx.close()
    h = 1e-4
    for ci, d in [(1, 0), (2, 1)]:
        Ls = []
        for sgn in (1, -1):
            pos = [list(r) for r in POS3]
            pos[ci][d] += sgn * h
            rc, r, den, _ = O.solve(mk(pos))
            assert r.converged
            Ls.append(float(np.sum(w * (den[0] - nt) ** 2)))
        ref = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(g[ci, d] - ref) <= 1e-6 * abs(ref), (ci, d, g[ci, d], ref)


def test_centre_forces_match_energy_differences():
    """dE/dR of the loop's total energy through the fixed point (TF kinetic
    energy, isolated Hartree with the density reaching the box faces, where a
    Hellmann-Feynman sum int n dEnv/dR is off by 4e-4)."""
    p = cluster(POS3)
    ctx = Context(p)
    r = ctx.solve()
    assert r.converged
    f = ctx.centre_forces()
    w = weights(p)
    for ci, d in [(0, 0), (1, 1), (2, 2)]:
        _, fdE = fd_loss(ctx, POS3, np.zeros_like(w), w, ci, d, 1e-4)
        assert abs(f[ci, d] - fdE) <= 1e-6 * abs(fdE), (ci, d, f[ci, d], fdE)
    ctx.close()


def test_adjoint_nait_matches_finite_differences():
    """nAiT density (adaptive Airy quadrature) with Hartree, Dirac and PW92:
    the adjoint differentiates the computed density with its quadrature tree
    held fixed.  Central differences (h = 1e-4) of whole solves stopped at
    max|dV| < 1e-10 carry about 1e-10 / 2h of noise, 1e-7..4e-6 of these
    derivatives; the bound is 1e-5."""
    mk = lambda pos: cluster(pos, de=5, edge=20.0, tol=1e-10, inter=(8, 11, -12))  # noqa: E731
    nt = target_density(mk)
    p = mk(POS3)
    w = weights(p)
    ctx = Context(p)
    r = ctx.solve()
    assert r.converged
    n = ctx.density()[0].copy()
    V = ctx.potential().copy()
    lam, info = ctx.adjoint(2.0 * w * (n - nt))
    assert info["converged"], info
    g = ctx.centre_gradient(lam)
    f = ctx.centre_forces()
    for ci, d in [(0, 0), (1, 1), (2, 2)]:
        fd, fdE = fd_loss(ctx, POS3, nt, w, ci, d, 1e-4, V0=V)
        assert abs(g[ci, d] - fd) <= 1e-5 * abs(fd), (ci, d, g[ci, d], fd)
        assert abs(f[ci, d] - fdE) <= 1e-5 * abs(fdE), (ci, d, f[ci, d], fdE)
    ctx.close()


def test_autograd_function_backward_is_the_adjoint():
    import torch
    from mpdpft_amd.autograd import CentreDensity, CentreEnergy, CentreSolver
    mk = lambda pos: cluster(pos)  # noqa: E731
    nt = torch.from_numpy(target_density(mk))
    p = mk(POS3)
    w = torch.from_numpy(weights(p))
    solver = CentreSolver(p)
    theta = torch.tensor(POS3, dtype=torch.float64, requires_grad=True)
    n = CentreDensity.apply(theta, solver)
    loss = (w * (n - nt) ** 2).sum()
    loss.backward()
    lam, _ = solver.ctx.adjoint((2.0 * w * (n.detach() - nt)).numpy())
    ref = solver.ctx.centre_gradient(lam)
    assert np.allclose(theta.grad.numpy(), ref, rtol=1e-12, atol=0.)
    theta2 = torch.tensor(POS3, dtype=torch.float64, requires_grad=True)
    E = CentreEnergy.apply(theta2, solver)
    E.backward()
    assert np.allclose(theta2.grad.numpy(), solver.ctx.centre_forces(), rtol=1e-12, atol=0.)
    # a second solve invalidates the first density's backward
    n1 = CentreDensity.apply(torch.tensor(POS3, dtype=torch.float64, requires_grad=True), solver)
    CentreEnergy.apply(torch.tensor(POS3, dtype=torch.float64, requires_grad=True), solver)
    with pytest.raises(RuntimeError):
        n1.sum().backward()
    solver.close()
