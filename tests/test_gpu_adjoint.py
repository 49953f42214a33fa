"""GPU parity of the charge-neutral steady-state polish and the steady-state
adjoint (p3d_chem_polish, p3d_chem_adjoint) against the oracle's Newton
polish and forward sensitivities, and of the autograd objective against
central finite differences of the whole forward map."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cells(n, seed):
    from oracle import oracle as O
    rng = np.random.RandomState(seed)
    par = np.stack([10 ** rng.uniform(2, 6, n), 10 ** rng.uniform(1, 2.7, n),
                    10 ** rng.uniform(0, 4, n), rng.uniform(0, 8, n),
                    10 ** rng.uniform(-17, -15.5, n), 10 ** rng.uniform(-4, 0, n),
                    10 ** rng.uniform(-4, 0, n)])
    par[:, :3] = np.array([[1e4, 3e2, 2e5], [40.0, 120.0, 15.0], [1e3, 30.0, 1e3],
                           [2.0, 0.4, 6.0], [2e-16, 1e-16, 2e-16], [0.05, 0.6, 1e-4],
                           [0.3, 0.9, 1e-3]])
    y0 = np.full(O.NSP, 1e-30)
    for s, v in (('H', 1.0), ('C+', 1.4e-4), ('O', 3e-4), ('e-', 1.4e-4), ('He', 0.1),
                 ('S+', 1e-6), ('Si+', 1e-7), ('Fe+', 1e-8)):
        y0[O.IDX[s]] = v
    return par, y0


def test_polish_and_adjoint_match_oracle(gpu):
    """64 cells.  Polish: from the same integrated state, the GPU polish equals
    the oracle's to rounding.  Whole pipeline (each side's own integration,
    then its polish): the two integrations conserve the element totals to
    ~1e-10, which moves species carrying a tiny share of an element by that
    absolute amount, so species are compared where they hold more than 1e-7
    of their element.  Gradient: at the GPU state, against the oracle's
    forward sensitivities (one solve per parameter), relative to the largest
    component (a contraction of large terms of both signs through matrices
    of condition 1e13-1e16)."""
    import torch
    from oracle import oracle as O
    from prism3d_amd.chemistry import ChemistrySolver, solve_opts
    n = 64
    par, y0 = _cells(n, 3)
    cs = ChemistrySolver()
    y = torch.tensor(np.repeat(y0[:, None], n, axis=1), device=gpu)
    ins = [torch.tensor(par[k], device=gpu) for k in range(7)]
    # the integration that precedes the polish runs in the one-cell-per-
    # workgroup form (the batched forms' integrations are compared with the
    # oracle in test_gpu_lane.py)
    cs.solve_steady_batch(y, *ins, opts=solve_opts(t_max=None, n_checkpoints=2, rtol=1e-9,
                                                   atol=1e-20, early_stop=False, post_mode=-1,
                                                   variant=2))
    y_int = y.cpu().numpy().copy()
    it = cs.polish_steady_batch(y, *ins)
    assert (it.cpu().numpy() > 0).all()
    x_gpu = y.cpu().numpy()
    w_cii, w_co = 1e24, 1e-6
    obj, grad = cs.steady_adjoint_batch(y, *ins, w_cii=w_cii, w_co=w_co)
    obj, grad = obj.cpu().numpy(), grad.cpu().numpy()
    yi, _, _ = O.solve_steady_batch(np.repeat(y0[:, None], n, axis=1), *par, t_max=0.0,
                                    n_checkpoints=2, rtol=1e-9, atol=1e-20, early_stop=False,
                                    post_mode=-1)
    _, C = O.element_matrix()
    worst_same = worst_pipe = worst_g = 0.0
    for c in range(n):
        xs, its = O.polish_steady(y_int[:, c], *par[:, c])
        assert its == int(it[c].item())
        m = xs > 1e-20
        worst_same = max(worst_same, np.max(np.abs(x_gpu[m, c] - xs[m]) / xs[m]))
        xo, ito = O.polish_steady(yi[:, c], *par[:, c])
        assert ito > 0
        share = np.max(C * xo[None, :] / (C @ xo)[:, None], axis=0)  # of its element
        m = share > 1e-7
        worst_pipe = max(worst_pipe, np.max(np.abs(x_gpu[m, c] - xo[m]) / xo[m]))
        ob, ga, gf = O.steady_sensitivities(x_gpu[:, c], *par[:, c], w_cii=w_cii, w_co=w_co)
        assert np.allclose(obj[c], ob, rtol=1e-13, atol=0)
        worst_g = max(worst_g, np.max(np.abs(grad[c] - gf)) / np.max(np.abs(gf)))
    print(f"polish (same input) {worst_same:.2e}, pipeline {worst_pipe:.2e}, "
          f"gradient {worst_g:.2e}", flush=True)
    assert worst_same < 1e-10  # same iteration path; the last solves' rounding
    assert worst_pipe < 1e-6
    assert worst_g < 1e-6


def test_steady_objective_gradient_against_finite_differences(gpu):
    """J(theta) through the autograd function on 128 cells; central
    differences (step 1e-4 in log10) of the forward map (GPU integration +
    polish) on 20 reactions drawn with RandomState(7) and on ln G0, ln zeta
    and ln of the density scale, to 1e-4 relative above the differences'
    noise floor (1e-12 of J per unit step: the map takes its element totals
    from the initial abundances, so it is smooth to rounding)."""
    import torch
    from prism3d_amd.adjoint import SteadyStateProblem, steady_objective
    par, y0 = _cells(128, 5)
    prob = SteadyStateProblem(*par, y0=y0, w_cii=1e24, w_co=1e-6)
    theta = torch.zeros(prob.n_params, dtype=torch.float64, requires_grad=True)
    J = steady_objective(theta, prob)
    J.backward()
    g = theta.grad.numpy()
    assert (prob.last_iters.cpu().numpy() > 0).all()
    J0 = float(J.detach())
    h = 1e-4
    idx = list(np.random.RandomState(7).choice(prob.nr, 20, replace=False)) + \
        [prob.nr, prob.nr + 1, prob.nr + 2]
    worst = 0.0
    for j in idx:
        step = h if j < prob.nr else h * np.log(10.0)
        tp = np.zeros(prob.n_params)
        tp[j] = step
        Jp = float(prob.evaluate(tp, want_grad=False)[0])
        Jm = float(prob.evaluate(-tp, want_grad=False)[0])
        fd = (Jp - Jm) / (2 * step)
        err = abs(fd - g[j])
        worst = max(worst, err / max(abs(fd), 1e-300))
        print(f"param {j}: fd {fd:.6e} adjoint {g[j]:.6e}", flush=True)
        assert err <= 1e-4 * abs(fd) + 1e-12 * abs(J0) / h, (j, fd, g[j])


def test_polish_rejects_bad_arguments(gpu):
    import torch
    from prism3d_amd.chemistry import ChemistrySolver
    cs = ChemistrySolver()
    y = torch.full((31, 2), 1e-5, dtype=torch.float64, device=gpu)
    with pytest.raises(ValueError):
        cs.polish_steady_batch(y, 1e3, 50.0, 1.0, 0.0, 1e-16, max_iter=0)
