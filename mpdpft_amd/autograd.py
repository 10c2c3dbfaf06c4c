"""torch.autograd bindings of the fixed-point adjoint (mpdpft_adjoint).

CentreDensity maps the centre coordinates of an environment-41 cluster
((n, 3) float64 tensor) to the converged density ([G] tensor); its backward
solves the adjoint of the converged self-consistent fixed point once per call
(implicit differentiation: no iterations are unrolled).  CentreEnergy maps
the coordinates to the loop's total energy; its backward returns dE/dR
through the fixed point (mpdpft_centre_forces).

    solver = CentreSolver(params)            # environment 41, one species, 3D
    theta = torch.tensor(xyz, dtype=torch.float64, requires_grad=True)
    n = CentreDensity.apply(theta, solver)
    loss = (w * (n - n_target) ** 2).sum()
    loss.backward()                          # theta.grad = dL/dtheta

The backward of a CentreDensity output uses the solver's state, so it must run
before the solver solves again (the solver checks this).
"""
import numpy as np
import torch

from . import Context


class CentreSolver:
    """A context whose environment-41 centre positions are the variables."""

    def __init__(self, params, device=-1):
        self.ctx = Context(params, device)
        p = self.ctx.params
        self.centres = np.array([[p.centres[i][k] for k in range(5)] for i in range(p.n_centres)])
        self.solves = 0
        self.last = None

    def solve_at(self, xyz):
        c = self.centres.copy()
        c[:, :3] = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.ctx.set_centres(c)
        self.last = self.ctx.solve()
        if not self.last.converged:
            raise RuntimeError("SC loop did not converge (SCcount %d)" % self.last.SCcount)
        self.solves += 1
        return self.last

    def close(self):
        self.ctx.close()


class CentreDensity(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, solver):
        solver.solve_at(theta.detach().cpu().numpy())
        fctx.solver = solver
        fctx.stamp = solver.solves
        return torch.from_numpy(solver.ctx.density()[0].copy())

    @staticmethod
    def backward(fctx, grad_n):
        solver = fctx.solver
        if solver.solves != fctx.stamp:
            raise RuntimeError("the solver has solved again since this density was computed")
        lam, info = solver.ctx.adjoint(grad_n.detach().cpu().numpy().astype(np.float64))
        if not info["converged"]:
            raise RuntimeError("adjoint solve did not converge: %r" % info)
        return torch.from_numpy(solver.ctx.centre_gradient(lam)), None


class CentreEnergy(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, solver):
        r = solver.solve_at(theta.detach().cpu().numpy())
        fctx.save_for_backward(torch.from_numpy(solver.ctx.centre_forces()))
        return torch.tensor(r.TotalEnergy, dtype=torch.float64)

    @staticmethod
    def backward(fctx, grad_e):
        (forces,) = fctx.saved_tensors
        return grad_e * forces, None
