"""Differentiable steady state of the per-cell chemistry.

Observables of the coupled solve, the [CII] 158 um cooling (solver_3d.py:
956-965) and the CO column along +x (the exit-face value of the columns of
fuv_rt_3d.py:386-388), are differentiated with respect to log10 of every rate
coefficient, ln G0, ln zeta_CR and ln of the density scale, with the
temperature, radiation field (A_V) and shielding of a converged fixed point
held fixed.

The state differentiated is the charge-neutral steady state of every cell
(p3d_chem_polish): the network conserves nuclei but not charge, so its raw
rate equations have no fixed point, and the upstream solve closes charge with
e- = sum of the cations after integrating (chemistry/solver.py:382-386).  The
forward map integrates each cell from its initial abundances with the GPU BDF
(tight tolerances) and polishes the result onto that state (pseudo-transient
continuation into Newton) with the element totals of the initial abundances,
so the map is smooth in the parameters; the
gradient is one transposed solve per cell at the state (p3d_chem_adjoint,
implicit function theorem), so the backward pass costs a small fraction of
the forward integration.

Parameters are offsets theta (n_reactions + 3): theta[r] adds to log10 k_r,
theta[-3] to ln G0, theta[-2] to ln zeta, theta[-1] to ln of the density
scale (n_H, and with it the CO column and the collision rates).
"""

import numpy as np
import torch

from .chemistry import CII_158, ChemistrySolver, solve_opts
from .device import device
from .network import DeviceNetwork, Network


class SteadyStateProblem:
    """Per-cell steady-state chemistry with frozen fields and the objective
    J = sum over cells of (w_cii Lambda_CII + w_co n_H x_CO)."""

    def __init__(self, n_H, T, G0, A_V, zeta, fsh_H2, fsh_CO, y0, w_cii=1.0, w_co=1.0,
                 network=None, rtol=1e-9, atol=1e-20, variant=0, polish_iter=600):
        dev = device()
        f = lambda v: torch.as_tensor(v, dtype=torch.float64, device=dev).reshape(-1).contiguous()
        self.n_H, self.T, self.G0, self.A_V = f(n_H), f(T), f(G0), f(A_V)
        self.zeta, self.fsh_H2, self.fsh_CO = f(zeta), f(fsh_H2), f(fsh_CO)
        self.n = self.n_H.numel()
        self.y0 = torch.as_tensor(y0, dtype=torch.float64, device=dev)
        if self.y0.ndim == 1:
            self.y0 = self.y0[:, None].expand(-1, self.n)
        self.y0 = self.y0.contiguous()
        self.w_cii, self.w_co = float(w_cii), float(w_co)
        self.network = network or Network()
        self.base = ChemistrySolver(DeviceNetwork(self.network))
        self.nr = self.network.n_reactions
        self.n_params = self.nr + 3
        self.opts = dict(t_max=None, t_start=1e4, n_checkpoints=2, early_stop=False, rtol=rtol,
                         atol=atol, post_mode=-1, variant=variant)
        self.polish_iter = int(polish_iter)
        self.last_iters = None

    def _solver(self, dlog10k):
        if not np.any(dlog10k):
            return self.base
        net = Network(self.network.species, self.network.reactions)
        net.alpha = net.alpha * 10.0 ** dlog10k
        return ChemistrySolver(DeviceNetwork(net))

    def inputs(self, theta):
        th = np.zeros(self.n_params) if theta is None else \
            np.asarray(torch.as_tensor(theta).detach().cpu(), dtype=np.float64)
        e = np.exp(th[-3:])
        return th, (self.n_H * e[2], self.T, self.G0 * e[0], self.A_V, self.zeta * e[1],
                    self.fsh_H2, self.fsh_CO)

    def state(self, theta=None):
        """Charge-neutral steady state (n_species, n) at parameters theta."""
        th, ins = self.inputs(theta)
        cs = self._solver(th[:self.nr])
        y = self.y0.clone()
        cs.solve_steady_batch(y, *ins, opts=solve_opts(**self.opts), sync=False)
        self.last_iters = cs.polish_steady_batch(y, *ins, max_iter=self.polish_iter,
                                                 y_totals=self.y0)
        return y, cs, ins

    def evaluate(self, theta=None, want_grad=True):
        """(J, per-cell objective (n, 2), dJ/dtheta or None, state)."""
        y, cs, ins = self.state(theta)
        obj, grad = cs.steady_adjoint_batch(y, *ins, w_cii=self.w_cii, w_co=self.w_co,
                                            line=CII_158)
        J = self.w_cii * obj[:, 0].sum() + self.w_co * obj[:, 1].sum()
        g = grad.sum(dim=0) if want_grad else None
        return J, obj, g, y


class SteadyObjective(torch.autograd.Function):
    """J(theta) of a SteadyStateProblem with its adjoint gradient:
    ``J = SteadyObjective.apply(theta, problem)``; ``J.backward()`` fills
    theta.grad with dJ/dtheta."""

    @staticmethod
    def forward(ctx, theta, problem):
        J, _, g, _ = problem.evaluate(theta, want_grad=True)
        ctx.save_for_backward(g.to(theta.device, theta.dtype))
        return J.to(theta.device, theta.dtype)

    @staticmethod
    def backward(ctx, gJ):
        (g,) = ctx.saved_tensors
        return gJ * g, None


def steady_objective(theta, problem):
    return SteadyObjective.apply(theta, problem)
