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


# --- discrete adjoint of the implicit steps ---------------------------------
def _tape_opts(t_max):
    from prism3d_amd.chemistry import solve_opts
    return solve_opts(t_max=t_max, t_start=1e4, n_checkpoints=2, early_stop=False, rtol=1e-9,
                      atol=1e-20, post_mode=-1, variant=6)


def test_taped_forward_is_the_solve(gpu):
    """The taped integration returns the plain one-cell-per-wavefront solve's
    state and step counts on 512 cells of PDR conditions (same code, the
    recorder only stores)."""
    import torch
    from prism3d_amd.chemistry import ChemistrySolver
    n = 512
    par, y0 = _cells(n, 11)
    cs = ChemistrySolver()
    ins = [torch.tensor(par[k], device=gpu) for k in range(7)]
    o = _tape_opts(1e7 * 3.15576e7)
    y1 = torch.tensor(np.repeat(y0[:, None], n, axis=1), device=gpu)
    _, st1 = cs.solve_steady_batch(y1, *ins, opts=o)
    y2 = torch.tensor(np.repeat(y0[:, None], n, axis=1), device=gpu)
    st2, tape = cs.solve_tape(y2, *ins, opts=o, cap=8192)
    a, b = y1.cpu().numpy(), y2.cpu().numpy()
    m = a > 1e-14
    rel = np.max(np.abs(a[m] - b[m]) / a[m])
    print(f"taped vs plain: max rel {rel:.2e}, bit-identical {np.array_equal(a, b)}", flush=True)
    assert rel <= 1e-13
    assert np.array_equal(st1.cpu().numpy(), st2.cpu().numpy())
    steps = cs.tape_steps(tape, range(n))
    assert all(k == int(st2[c, 0]) for c, (_, _, k) in enumerate(steps))


def test_tape_adjoint_matches_forward_sensitivities(gpu):
    """Six cells from diffuse to dense: the discrete adjoint of the implicit
    steps against the oracle's forward (tangent-linear) sensitivities through
    the same recorded step sequence, for all 78 parameters (log10 k_r, ln G0,
    ln zeta, ln n_H), with the objective weights of both terms."""
    import torch
    from oracle import oracle as O
    from prism3d_amd.chemistry import ChemistrySolver
    n = 6
    par, y0 = _cells(n, 3)
    cs = ChemistrySolver()
    ins = [torch.tensor(par[k], device=gpu) for k in range(7)]
    o = _tape_opts(1e6 * 3.15576e7)
    y0d = torch.tensor(np.repeat(y0[:, None], n, axis=1), device=gpu)
    y = y0d.clone()
    _, tape = cs.solve_tape(y, *ins, opts=o, cap=8192)
    w_cii, w_co = 1e24, 1e-6
    obj, grad, status = cs.tape_adjoint(y0d, y, *ins, tape=tape, w_cii=w_cii, w_co=w_co)
    assert (status.cpu().numpy() == 0).all()
    grad, obj = grad.cpu().numpy(), obj.cpu().numpy()
    ye = y.cpu().numpy()
    worst = 0.0
    for c, (steps, h0, k) in enumerate(cs.tape_steps(tape, range(n))):
        yt, _, _, gt = O.taped_sensitivities(y0, steps, h0, *par[:, c], w_cii=w_cii, w_co=w_co)
        m = ye[:, c] > 1e-14
        assert np.max(np.abs(yt[m] - ye[m, c]) / ye[m, c]) < 1e-8  # the replay reaches the state
        x = ye[:, c]
        lam, _ = O.cii_cooling(par[1, c], par[0, c], x[O.IDX['C+']], x[O.IDX['e-']],
                               x[O.IDX['H']], x[O.IDX['H2']])
        assert np.allclose(obj[c], [lam, par[0, c] * x[O.IDX['CO']]], rtol=1e-12, atol=0)
        err = np.max(np.abs(grad[c] - gt)) / np.max(np.abs(gt))
        print(f"cell {c}: {k} steps, gradient {err:.2e}", flush=True)
        worst = max(worst, err)
    assert worst < 1e-8


def _cube_problem(n=32):
    """The differentiation workload's cube: n^3 log-normal (median n_H 1e4,
    sigma_ln 1.0, 0.5 pc, seed 42), G0 1e3 from -x over 48 rays, zeta
    2e-16; T, radiation and shielding from the coupled fixed point with the
    reduced explicit chemistry."""
    from prism3d_amd.adjoint import TransientProblem
    from prism3d_amd.density import PC, lognormal_median
    from prism3d_amd.grid import directional_G0
    from prism3d_amd.network import TRACKED
    from prism3d_amd.solver3d import PDRSolverGPU
    box = 0.5 * PC
    dens, cell = lognormal_median(n, box, 1e4, 1.0, seed=42)
    sv = PDRSolverGPU(dens, box, G0_external=directional_G0(1e3, [-1.0, 0.0, 0.0], nside=2),
                      zeta_CR_0=2e-16, nside_rt=2, chemistry='euler')
    sv.run(max_iterations=20, convergence_tol=0.05)
    sv.shielding_cr_step()
    y0 = np.full(31, 1e-30)
    for s, v in {'H': 1.0, 'C+': 1.4e-4, 'O': 3.0e-4, 'e-': 1.4e-4, 'He': 0.0851}.items():
        y0[TRACKED.index(s)] = v
    f = lambda t: t.reshape(-1)
    return TransientProblem(f(sv.density), f(sv.T_gas), f(sv.G0), f(sv.A_V), f(sv.zeta_CR),
                            f(sv.fsh_H2), f(sv.fsh_CO), y0, w_cii=1.0, w_co=cell), cell


def test_tape_adjoint_against_central_differences_on_cube(gpu):
    """The differentiation workload at 32^3: the summed [CII] 158 um cooling and
    the CO column along +x (summed over the exit face) at the end of every
    cell's integration to its steady-state horizon.  Their gradients from the
    discrete adjoint of the implicit steps (one taped forward, one adjoint
    walk per objective) against central differences of the whole forward map
    (step 1e-4 in log10; adaptive steps included) for 20 reactions drawn with
    RandomState(7) and ln G0, ln zeta, ln n_H.  Bound: 1e-4 of the largest
    gradient component of each objective."""
    import torch
    prob, cell = _cube_problem(32)
    nr = prob.nr
    rs = np.random.RandomState(7).choice(nr, 20, replace=False)
    params = list(rs) + [nr, nr + 1, nr + 2]
    theta = torch.zeros(prob.n_params, dtype=torch.float64)
    J0, obj0 = prob.objective(theta)
    g_cii, _, status, _ = prob.gradient(w_cii=1.0, w_co=0.0)
    g_co, _, _, _ = prob.gradient(w_cii=0.0, w_co=cell)
    st = status.cpu().numpy()
    # a cell whose recomputed segment misses the forward's state by more than
    # 1e-6 is flagged (-3); it stays in the sums
    print(f"adjoint status over {st.size} cells: {dict(zip(*np.unique(st, return_counts=True)))}",
          flush=True)
    assert ((st == 0) | (st == -3)).all() and (st == -3).sum() <= 1e-4 * st.size
    g_cii, g_co = g_cii.cpu().numpy(), g_co.cpu().numpy()
    eps = 1e-4
    fd = np.zeros((len(params), 2))
    for i, p in enumerate(params):
        vals = []
        for sgn in (1.0, -1.0):
            th = np.zeros(prob.n_params)
            th[p] = sgn * eps * (1.0 if p < nr else np.log(10.0))  # 1e-4 in log10
            _, ob = prob.objective(th)
            vals.append(ob.sum(dim=0).cpu().numpy())
        d = (vals[0] - vals[1]) / (2.0 * eps * (1.0 if p < nr else np.log(10.0)))
        fd[i] = [d[0], cell * d[1]]
    ad = np.stack([g_cii[params], g_co[params]], axis=1)
    worst = []
    for j, name in enumerate(("CII", "CO column")):
        scale = np.max(np.abs(fd[:, j]))
        err = np.abs(ad[:, j] - fd[:, j]) / scale
        big = np.abs(fd[:, j]) > 1e-2 * scale
        rel_big = np.max(np.abs(ad[big, j] - fd[big, j]) / np.abs(fd[big, j]))
        print(f"{name}: max error {err.max():.2e} of the largest component; components above "
              f"1% of it agree to {rel_big:.2e} relative", flush=True)
        worst.append(err.max())
    assert max(worst) < 1e-4


def test_transient_autograd_and_chunked_tapes(gpu):
    """The autograd front end of the implicit-step adjoint: J.backward() fills
    theta.grad with the adjoint's gradient, and the memory-short mode (tapes
    of chunks of cells, recomputed in the backward pass) gives the same
    objective and gradient bit for bit."""
    import torch
    from prism3d_amd.adjoint import TransientProblem, chemistry_objective
    n = 300
    par, y0 = _cells(n, 5)
    grads, Js = [], []
    for budget in (48 * 2 ** 30, 128 * (24 * 4096 + 2048 * 257 + 16)):
        prob = TransientProblem(*par, y0, w_cii=1e24, w_co=1e-6, t_end=1e6 * 3.15576e7,
                                cap=4096, tape_budget=budget)
        theta = torch.zeros(prob.n_params, dtype=torch.float64, requires_grad=True)
        J = chemistry_objective(theta, prob)
        J.backward()
        g, _, status, _ = prob.gradient()
        assert torch.equal(theta.grad, g.cpu())
        st = status.cpu().numpy()
        assert (st == 0).all(), dict(zip(*np.unique(st, return_counts=True)))
        grads.append(theta.grad.clone())
        Js.append(float(J))
        print(f"tape budget {budget}: {prob.chunk} cells per chunk", flush=True)
    assert Js[0] == Js[1]
    assert torch.equal(grads[0], grads[1])
