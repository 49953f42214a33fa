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
        self.last = None

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

    def objective(self, theta=None):
        """(J, per-cell objective (n, 2)); the gradient is formed with the
        forward (one transposed solve per cell) and kept for gradient()."""
        J, obj, g, _ = self.evaluate(theta, want_grad=True)
        self.last = dict(g=g, obj=obj)
        return J, obj

    def gradient(self, last=None):
        """(dJ/dtheta, per-cell objective, None, None) of a forward objective()."""
        L = last if last is not None else self.last
        if L is None:
            raise RuntimeError("gradient() needs a forward objective() first")
        return L["g"], L["obj"], None, None


class ChemistryObjective(torch.autograd.Function):
    """J(theta) of a chemistry problem with its adjoint gradient:
    ``J = ChemistryObjective.apply(theta, problem)``; ``J.backward()`` fills
    theta.grad with dJ/dtheta.  The problem chooses the mode:
    SteadyStateProblem (the implicit-function adjoint at the charge-neutral
    steady state, computed with the forward) or TransientProblem (the
    discrete adjoint of the implicit steps, run in the backward pass over
    the tape of this forward)."""

    @staticmethod
    def forward(ctx, theta, problem):
        J, _ = problem.objective(theta)
        ctx.problem, ctx.last = problem, problem.last
        ctx.meta = (theta.device, theta.dtype)
        return J.to(theta.device, theta.dtype)

    @staticmethod
    def backward(ctx, gJ):
        g = ctx.problem.gradient(ctx.last)[0]
        dev, dt = ctx.meta
        return gJ * g.to(dev, dt), None


SteadyObjective = TransientObjective = ChemistryObjective


def chemistry_objective(theta, problem):
    return ChemistryObjective.apply(theta, problem)


steady_objective = transient_objective = chemistry_objective


class TransientProblem(SteadyStateProblem):
    """The same objective at the end of the BDF integration itself: every
    cell integrated from y0 over one span to t_end (the reference's
    integration, chemistry/solver.py:185-217; t_end None = its steady-state
    horizon min(3e17, max(1e16, 3e17 / n_H)) s) with frozen fields.  The
    gradient is the discrete adjoint of the implicit steps with the step
    sequence held fixed (p3d_chem_tape_adjoint): the taped forward records
    each step's c, order and step-size change and the differences every 16
    steps; the adjoint recomputes each 16-step segment from its checkpoint,
    stores its steps' factors of I - c J and reuses them in the transposed
    solves.  cap bounds the accepted steps per cell of the tape (a cell past
    it gets no gradient: status -2).  Cells without a valid gradient (status
    != 0) are left out of the summed gradient and reported by gradient().

    Memory: the tape takes ~(24 cap + 2048 (cap / 16 + 1)) bytes per cell.
    When the cells' tapes exceed tape_budget bytes the forward keeps none:
    it runs the taped integration in chunks of cells that fit, and the
    gradient recomputes each chunk's taped forward before its adjoint walk
    (one more forward of work, the same results)."""

    def __init__(self, n_H, T, G0, A_V, zeta, fsh_H2, fsh_CO, y0, w_cii=1.0, w_co=1.0,
                 network=None, t_end=None, rtol=1e-9, atol=1e-20, cap=4096, max_steps=1000000,
                 tape_budget=48 * 2 ** 30):
        super().__init__(n_H, T, G0, A_V, zeta, fsh_H2, fsh_CO, y0, w_cii=w_cii, w_co=w_co,
                         network=network, rtol=rtol, atol=atol)
        self.opts = dict(t_max=t_end, t_start=1e4, n_checkpoints=2, early_stop=False,
                         rtol=rtol, atol=atol, post_mode=-1, max_steps=max_steps)
        self.cap = int(cap)
        per_cell = 24 * self.cap + 2048 * (self.cap // 16 + 1) + 16
        self.chunk = max(1, min(self.n, int(tape_budget // per_cell)))
        self.last = None

    def _chunks(self):
        return [(a, min(a + self.chunk, self.n)) for a in range(0, self.n, self.chunk)]

    def state(self, theta=None):
        """State at t_end (n_species, n) at parameters theta (its tape kept
        when all cells' tapes fit the budget)."""
        th, ins = self.inputs(theta)
        cs = self._solver(th[:self.nr])
        y = self.y0.clone()
        o = solve_opts(**self.opts)
        tape = None
        if self.chunk >= self.n:
            _, tape = cs.solve_tape(y, *ins, opts=o, cap=self.cap, sync=False)
        else:
            for a, b in self._chunks():
                yc = y[:, a:b].contiguous()
                cs.solve_tape(yc, *[t[a:b] for t in ins], opts=o, cap=self.cap, sync=False)
                y[:, a:b] = yc
        self.last = dict(theta=th, cs=cs, ins=ins, y=y, tape=tape)
        return y, cs, ins

    def objective(self, theta=None):
        """(J, per-cell objective (n, 2)) at the end of the integration."""
        y, cs, ins = self.state(theta)
        obj = _objective_values(y, ins, self.network, CII_158)
        J = self.w_cii * obj[:, 0].sum() + self.w_co * obj[:, 1].sum()
        return J, obj

    def gradient(self, last=None, w_cii=None, w_co=None):
        """dJ/dtheta (n_params,), per-cell objective (n, 2), status (n,) and
        per-cell gradient (n, n_params) of a forward state() (the last one by
        default): the adjoint walk over its tape (weights: the problem's
        unless given)."""
        L = last if last is not None else self.last
        if L is None:
            raise RuntimeError("gradient() needs a forward state() first")
        wc = self.w_cii if w_cii is None else float(w_cii)
        wo = self.w_co if w_co is None else float(w_co)
        cs, ins = L["cs"], L["ins"]
        if L["tape"] is not None:
            obj, grad, status = cs.tape_adjoint(self.y0, L["y"], *ins, tape=L["tape"],
                                                w_cii=wc, w_co=wo, line=CII_158, sync=False)
        else:
            o = solve_opts(**self.opts)
            parts = []
            for a, b in self._chunks():
                sub = [t[a:b] for t in ins]
                y0c = self.y0[:, a:b].contiguous()
                yc = y0c.clone()
                _, tape = cs.solve_tape(yc, *sub, opts=o, cap=self.cap, sync=False)
                parts.append(cs.tape_adjoint(y0c, yc, *sub, tape=tape, w_cii=wc, w_co=wo,
                                             line=CII_158, sync=False))
                del tape
            obj, grad, status = (torch.cat([p_[k] for p_ in parts]) for k in range(3))
        # a cell without a valid gradient (status != 0: failed integration, tape
        # too short, or a recomputed segment that missed the forward's state)
        # contributes nothing to the sum; status reports it
        grad = torch.where((status == 0)[:, None], grad, torch.zeros_like(grad))
        return grad.sum(dim=0), obj, status, grad

    def evaluate(self, theta=None, want_grad=True):
        """(J, per-cell objective (n, 2), dJ/dtheta or None, state)."""
        y, cs, ins = self.state(theta)
        g, obj, status, _ = self.gradient()
        J = self.w_cii * obj[:, 0].sum() + self.w_co * obj[:, 1].sum()
        return J, obj, (g if want_grad else None), y


def _objective_values(y, ins, network, line):
    """{Lambda_CII, n_H x_CO} per cell at abundances y (the objective of the
    adjoints, solver_3d.py:956-965), on the device with torch arithmetic."""
    n_H, T = ins[0], ins[1]
    idx = network.idx
    x = lambda s: y[idx[s]]
    freq, A, Tul, gu, gl = [float(v) for v in line]
    ge = 8.63e-6 / (gl * torch.sqrt(T)) * 2.15 * torch.exp(-Tul / T)
    gh = 8.0e-10 * (T / 100.0) ** 0.07
    gh2 = 4.9e-10 * (T / 100.0) ** 0.12
    q = ge * (x('e-') * n_H) + gh * (x('H') * n_H) + gh2 * (x('H2') * n_H)
    ar = (gu / gl) * torch.exp(-Tul / T)
    frac = ar * q / (ar * q + A + q)
    lam = x('C+') * n_H * frac * (A * 6.62607015e-27 * freq)
    return torch.stack([lam, n_H * x('CO')], dim=1)
