"""GPU parity of the whole-grid kernels (columns, shielding, cosmic rays, dust
state, rate precompute, thermal context, diagnostics), the post-solve
corrections and the device-resident driver, against the upstream golden
vectors and the oracle.  Every call goes through the C ABI."""

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

THERMAL_KEYS = ['gamma_fixed', 'gg_coeff', 't_dust', 'n_e', 'n_hi', 'n_h2', 'n_co', 'n_cp_nh',
                'n_o_nh', 'n_c_nh']


def T(a, dev):
    import torch
    return torch.tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)


def close(a, b, rtol):
    a = a.cpu().numpy().ravel() if hasattr(a, "cpu") else np.ravel(a)
    b = np.ravel(b)
    rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
    return rel.max()


@pytest.mark.parametrize("case", ["c1", "ln", "ln_dir"])
def test_fuv_columns_match_upstream(gpu, case):
    from oracle import oracle as O
    from prism3d_amd import grid
    g = golden("rt_golden.npz")
    if case == "c1":
        dens, cell = O.uniform_slab(16, 1.0 * O.PC, 1e3)
        xh2, xco, g0, nside = g['c1_xH2'], g['c1_xCO'], g['g0rays_c1'], 1
    else:
        dens, cell = g['ln_density'], float(g['ln_cell'])
        xh2, xco, nside = g['ln_xH2'], g['ln_xCO'], 2
        g0 = 1e3 if case == "ln" else g['g0rays_c2']
    out = grid.fuv_columns(T(dens, gpu), T(xh2, gpu), T(xco, gpu), cell, nside, g0)
    for k, name in enumerate(['G0', 'AV', 'NH', 'NH2', 'NCO']):
        ref = g[f'{case}_{name}']
        assert close(out[k], ref, 0) < 1e-12, name


def test_shielding_and_cosmic_rays_match_upstream(gpu):
    import torch
    from prism3d_amd import grid
    g = golden("rt_golden.npz")
    tm = torch.tensor([float(g['sh_T'])], dtype=torch.float64, device=gpu)
    fh2, fco, zeta = grid.shield_cr(T(g['sh_NH2'], gpu), T(g['sh_NCO'], gpu), T(g['cr_NH'], gpu),
                                    tm, 2e-16)
    assert close(fh2, g['sh_fH2'], 0) < 1e-13
    assert close(fco, g['sh_fCO'], 0) < 1e-12
    assert close(zeta, g['cr_zeta'], 0) < 1e-13
    _, _, z5 = grid.shield_cr(T(g['sh_NH2'], gpu), T(g['sh_NCO'], gpu), T(g['cr_NH'], gpu), tm,
                              5e-17)
    assert close(z5, g['cr_zeta_5e17'], 0) < 1e-13


@pytest.mark.parametrize("n", [1, 5, 8192, 8193, 100003, 9000011])
def test_mean_kernel(gpu, n):
    """Single- and multi-block partitions (up to the 1024-partial cap), against
    numpy, and bit-identical on a second call."""
    from prism3d_amd import grid
    x = np.random.RandomState(n).uniform(10, 1000, n)
    xt = T(x, gpu)
    m = grid.mean(xt).item()
    assert abs(m - np.mean(x)) / np.mean(x) < 1e-13
    assert grid.mean(xt).item() == m


def test_dust_state_matches_upstream(gpu):
    from prism3d_amd import grid
    from prism3d_amd.dust_tables import PETable
    g = golden("pe_golden.npz")
    pe = PETable()
    Td, _ = grid.dust_state(T(g['g0'], gpu), T(g['a_v'], gpu), T(g['n_h'], gpu),
                            T(g['t_gas'], gpu), T(g['f_nano'], gpu), pe.sigma_nano,
                            pe.sigma_largeC, pe.sigma_Sil)
    assert close(Td, g['t_dust'], 0) < 1e-13
    # R_H2 at the given dust temperature: feed the fixture's T_dust as the
    # dust temperature by using a density below the gas-grain coupling limit
    from oracle import oracle as O
    Td_in = g['t_dust_in']
    # find G0 giving exactly T_dust_in is not possible in general; compare the
    # kernel's R_H2 with the oracle at the kernel's own T_dust instead
    _, R = grid.dust_state(T(g['g0'], gpu), T(g['a_v'], gpu), T(g['n_h'], gpu),
                           T(g['t_gas'], gpu), T(g['f_nano'], gpu), pe.sigma_nano,
                           pe.sigma_largeC, pe.sigma_Sil)
    Rref = O.h2_formation_rate(g['t_gas'], g['t_dust'], g['f_nano'], pe.sigma_nano,
                               pe.sigma_largeC, pe.sigma_Sil)
    assert close(R, Rref, 0) < 1e-12
    assert Td_in.size == g['t_gas'].size


def test_euler_rates_match_upstream(gpu):
    from prism3d_amd import grid
    e = golden("euler_golden.npz")
    out = grid.euler_rates(T(e['n_h'], gpu), T(e['G0'], gpu), T(e['A_V'], gpu),
                           T(e['T_gas'], gpu), T(e['zeta_CR'], gpu), T(e['fsh_H2'], gpu),
                           T(e['fsh_CO'], gpu), T(e['R_H2'], gpu))
    for k, name in enumerate(['k_pd_base', 'k_ci', 'alpha_cp', 'k_co_total', 'r_h2_half_nh',
                              'r_cr_base']):
        assert close(out[k], e[name], 0) < 1e-13, name


def test_thermal_context_matches_upstream(gpu):
    from prism3d_amd import grid
    s = golden("state8_golden.npz")
    t = golden("thermal_golden.npz")
    ctx = grid.thermal_ctx(*[T(s[k], gpu) for k in
                             ['density', 'G0', 'A_V', 'fsh_H2', 'zeta_CR', 'T_dust', 'Gamma_PE',
                              '_R_H2', 'x_HI', 'x_H2', 'x_e', 'x_CO', 'x_Cp', 'x_O', 'x_C']])
    names = ['gamma_fixed', 'gg_coeff', 'n_e', 'n_hi', 'n_h2', 'n_co', 'n_cp_nh', 'n_o_nh',
             'n_c_nh']
    for k, name in enumerate(names):
        assert close(ctx[k], t[name], 0) < 1e-13, name


def test_heating_cooling_matches_upstream(gpu):
    from prism3d_amd import grid
    s = golden("state8_golden.npz")
    d = golden("diag_golden.npz")
    G, L = grid.heating_cooling(T(d['T'], gpu), *[T(s[k], gpu) for k in
                                                  ['density', 'G0', 'A_V', 'zeta_CR', 'fsh_H2',
                                                   'x_e', 'x_HI', 'x_H2', 'x_C', 'x_Cp', 'x_O',
                                                   'x_CO', 'T_dust', 'Gamma_PE', '_R_H2']])
    assert close(G, d['Gamma'], 0) < 1e-12
    assert close(L, d['Lambda'], 0) < 1e-12


def test_post_corrections_match_upstream(gpu):
    from prism3d_amd.chemistry import ChemistrySolver
    p = golden("post_golden.npz")
    cs = ChemistrySolver()
    y = T(p['y'].T, gpu)
    cs.post_batch(y, p['n_H'], p['T'], p['zeta'], mode=0, x_c_total=float(p['x_C_total']),
                  x_o_total=float(p['x_O_total']))
    got = y.cpu().numpy().T
    # the upstream corrects H+/e- first, then conserves
    from oracle import oracle as O
    ref = np.array([O.enforce_conservation(p['hp_e'][i], float(p['x_C_total']),
                                           float(p['x_O_total']), 0)
                    for i in range(p['y'].shape[0])])
    rel = np.abs(got - ref) / ref
    assert rel.max() < 1e-12
    # exact projection: H nuclei and carbon on the totals, e- = sum of cations
    net = cs.network
    hw = np.array([1.0 if s == 'H' else 2.0 if s == 'H2' else 0.0 for s in net.species])
    # feasible states only: H nuclei in the minor species below the total
    ok = (net.post_w[0] @ p['y'].T) < 0.9
    y = T(p['y'][ok].T, gpu)
    cs.post_batch(y, p['n_H'][ok], p['T'][ok], p['zeta'][ok], mode=1)
    got = y.cpu().numpy()
    hsum = (hw + net.post_w[0]) @ got
    assert ok.sum() > 10
    assert np.all(np.abs(hsum - 1.0) < 1e-12)


def _rel(a, b):
    return np.abs(a - b) / np.abs(b)


@pytest.mark.parametrize("iters", [1, 4])
def test_driver_euler_mode_matches_oracle_loop(gpu, iters):
    """The device-resident loop (native-core chemistry semantics) against the
    restated loop on the upstream 8^3 parity fixture.

    The loop is sensitive at the last bit: the explicit chemistry's step
    limiter and the thermal bisection's sign tests turn an ulp into a
    different branch, and the oracle itself, run on densities scaled by
    (1 + 2e-16), moves a few cells after one iteration and about a third of
    the x_H2 values by more than 1e-3 after four.  One iteration is held to a
    tight median with at most 1% outliers; four iterations are held to the
    oracle's own spread under that perturbation."""
    from oracle import oracle as O
    from prism3d_amd.solver3d import PDRSolverGPU
    r = golden("run_golden.npz")
    dens = r['density'].reshape(8, 8, 8)
    sv = PDRSolverGPU(dens, float(r['box']), G0_external=100.0, zeta_CR_0=2e-16, nside_rt=1,
                      chemistry='euler')
    sv.run(max_iterations=iters, convergence_tol=1e-9)
    st, hist = O.run_loop(dens, float(r['box']), 100.0, 2e-16, 1, iters)
    sp, _ = O.run_loop(dens * (1.0 + 2e-16), float(r['box']), 100.0, 2e-16, 1, iters)
    for k in ('T_gas', 'x_H2', 'x_CO', 'x_e', 'G0', 'A_V'):
        rel = _rel(getattr(sv, k).cpu().numpy().ravel(), st[k].ravel())
        own = _rel(sp[k].ravel(), st[k].ravel())
        print(k, "median", np.median(rel), "(oracle", np.median(own), ") n>1e-3",
              int(np.sum(rel > 1e-3)), "(oracle", int(np.sum(own > 1e-3)), ")")
        if iters == 1:
            assert np.median(rel) < 1e-12, k
            assert np.mean(rel > 1e-6) <= 0.01, k
        else:
            assert np.median(rel) <= 10.0 * np.median(own) + 1e-12, k
            assert np.mean(rel > 1e-3) <= 1.5 * np.mean(own > 1e-3) + 0.02, k
    assert sv.convergence_history[0]['dT'] == pytest.approx(hist[0]['dT'], rel=1e-9)


def test_driver_network_mode_small_grid(gpu):
    """Network chemistry inside the loop on a 8^3 cube: every iteration's
    chemistry equals the oracle BDF on the same inputs."""
    from oracle import oracle as O
    from prism3d_amd.solver3d import PDRSolverGPU
    r = golden("run_golden.npz")
    dens = r['density'].reshape(8, 8, 8)
    sv = PDRSolverGPU(dens, float(r['box']), G0_external=100.0, zeta_CR_0=2e-16, nside_rt=1,
                      chemistry='network', chem_rtol=1e-9, chem_atol=1e-20)
    sv.record_inputs = np.arange(0, 512, 3)
    sv.recorded = []
    sv.run(max_iterations=3, convergence_tol=1e-9)
    rec = sv.recorded
    assert len(rec) == 3
    last = rec[-1]
    ref, _, _ = O.solve_steady_batch(last['y'], last['n_H'], last['T'], last['G0'], last['A_V'],
                                     last['zeta'], last['fsh_H2'], last['fsh_CO'],
                                     t_max=last['t_max'], n_checkpoints=2, rtol=1e-9, atol=1e-20,
                                     early_stop=False, post_mode=last['post_mode'],
                                     x_c_total=last['x_c_total'], x_o_total=last['x_o_total'])
    got = last['y_out']
    m = ref > 1e-14
    assert np.max(np.abs(got[m] - ref[m]) / ref[m]) < 1e-6


@pytest.mark.parametrize("chem", ["euler", "network"])
def test_graph_replay_matches_eager_loop(gpu, chem):
    """run(graph=True) replays one captured HIP graph of the inner iteration
    from the third iteration on, with the host convergence test (and its
    synchronisations) between replays.  Every replay's convergence record and
    the final state equal the eager loop's (same kernels, same order)."""
    from prism3d_amd.solver3d import PDRSolverGPU
    r = golden("run_golden.npz")
    dens = r['density'].reshape(8, 8, 8)
    runs = []
    for graph in (False, True):
        print(f"case: {chem} graph={graph}", flush=True)
        sv = PDRSolverGPU(dens, float(r['box']), G0_external=100.0, zeta_CR_0=2e-16,
                          nside_rt=1, chemistry=chem)
        sv.run(max_iterations=6, convergence_tol=0.0, graph=graph)
        runs.append(sv)
    a, b = runs
    assert b._graph is not None
    assert len(a.convergence_history) == len(b.convergence_history) == 6
    for ra, rb in zip(a.convergence_history, b.convergence_history):
        for k in ('dT', 'dH2', 'dCO'):
            assert ra[k] == rb[k], (ra['iteration'], k)
    for k in ('T_gas', 'x_H2', 'x_CO', 'x_e', 'G0', 'Gamma', 'Lambda'):
        assert np.array_equal(getattr(a, k).cpu().numpy(), getattr(b, k).cpu().numpy()), k


def test_refine_matches_oracle_steady_solve(gpu):
    """refine() after three Euler iterations on the 8^3 fixture: every cell's
    network solve at the upstream solve_steady_state settings from the ten
    grid species (others at the floor), shielding from the columns and the
    mean T, against the oracle on the same inputs; then the thermal balance
    at the refined abundances against the oracle thermal solve."""
    from oracle import oracle as O
    from prism3d_amd.solver3d import PDRSolverGPU
    r = golden("run_golden.npz")
    dens = r['density'].reshape(8, 8, 8)
    sv = PDRSolverGPU(dens, float(r['box']), G0_external=100.0, zeta_CR_0=2e-16, nside_rt=1)
    sv.run(max_iterations=3, convergence_tol=1e-9)
    f = lambda a: a.cpu().numpy().ravel()
    grid_sp = ['H', 'H2', 'C+', 'C', 'CO', 'O', 'e-', 'OH', 'H2O', 'HCO+']
    y0 = np.full((O.NSP, 512), 1e-30)
    ycur = sv.y.cpu().numpy()
    for s in grid_sp:
        y0[O.IDX[s]] = ycur[sv.sidx[s]]
    T0 = f(sv.T_gas)
    fH2 = O.f_shield_H2(f(sv.N_H2), T=np.mean(T0))
    fCO = O.f_shield_CO(f(sv.N_CO), f(sv.N_H2))
    ref, conv_ref, _ = O.solve_steady_batch(y0, f(sv.density), T0, f(sv.G0), f(sv.A_V),
                                            f(sv.zeta_CR), fH2, fCO)
    n_conv = sv.refine(verbose=False)
    got = sv.y.cpu().numpy()
    m = ref > 1e-14
    rel = np.abs(got[m] - ref[m]) / ref[m]
    print(f"refine: frac < 1e-3 {np.mean(rel < 1e-3):.4f}, median {np.median(rel):.2e}, "
          f"converged {n_conv} (oracle {int(conv_ref.sum())})", flush=True)
    assert np.mean(rel < 1e-3) > 0.99
    assert abs(n_conv - int(conv_ref.sum())) <= 0.05 * 512
    assert np.allclose(f(sv.fsh_H2), fH2, rtol=1e-13)
    # thermal at the refined state
    Td, gpe, rh2, zeta = f(sv.T_dust), f(sv.Gamma_PE), f(sv._R_H2), f(sv.zeta_CR)
    xs = {k: got[sv.sidx[s]] for k, s in (('x_HI', 'H'), ('x_H2', 'H2'), ('x_e', 'e-'),
                                           ('x_CO', 'CO'), ('x_Cp', 'C+'), ('x_O', 'O'),
                                           ('x_C', 'C'))}
    ctx = O.thermal_context(f(sv.density), f(sv.G0), f(sv.A_V), fH2, zeta, Td, gpe, rh2,
                            xs['x_HI'], xs['x_H2'], xs['x_e'], xs['x_CO'], xs['x_Cp'], xs['x_O'],
                            xs['x_C'])
    Tref = O.solve_thermal_vec(*ctx, *O.line_params())
    relT = np.abs(f(sv.T_gas) - Tref) / Tref
    assert np.median(relT) < 1e-12 and np.mean(relT > 1e-4) <= 0.02
