"""The workload configurations, downscaled to fit the test budget, through the
GPU driver against the oracle:

  * the headline grid (log-normal, median n_H 1e4, sigma_ln 1.0, 0.5 pc, G0
    1e4 from -x over 48 rays, zeta 2e-16): the whole coupled loop to the
    upstream convergence test on a 16^3 field, against the oracle's loop;
    and the in-loop chemistry of a 32^3 field through every form against the
    oracle on 4096 recorded cells at two absolute tolerances;
  * the batching grid (sigma_ln 1.5, n_H clipped to 1e2-1e6): a 16^3 field's
    first iterations against the oracle's loop, and a 4-cloud ensemble
    (G0, median n_H and sigma_ln drawn per cloud) through one persistent
    kernel launch per iteration against each cloud's oracle loop.
"""

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

Y0 = {'H': 1.0, 'C+': 1.4e-4, 'O': 3.0e-4, 'e-': 1.4e-4}
XC, XO = 1.4e-4, 3.0e-4


def _solver(dens, box, g0, nside=2, **kw):
    from prism3d_amd.grid import directional_G0
    from prism3d_amd.solver3d import PDRSolverGPU
    return PDRSolverGPU(dens, box, G0_external=directional_G0(g0, [-1.0, 0.0, 0.0], nside=nside),
                        zeta_CR_0=2e-16, nside_rt=nside, chemistry='network', x_c_total=XC,
                        x_o_total=XO, initial_abundances=Y0, T_init=50.0, chem_atol=1e-20,
                        chem_mode='integrate', abundance_damping='reference', chem_post_mode=0,
                        **kw)


def _oracle_loop(dens, box, g0, iterations, nside=2):
    from prism3d_amd.grid import directional_G0
    y0 = np.full(O.NSP, 1e-30)
    for s, v in Y0.items():
        y0[O.IDX[s]] = v
    return O.run_loop_network(dens, box, directional_G0(g0, [-1.0, 0.0, 0.0], nside=nside),
                              2e-16, nside, iterations, y0, x_c_total=XC, x_o_total=XO)


def _compare(sv, st, what):
    """Relative differences of T and the abundances above 1e-14."""
    T = sv.T_gas.cpu().numpy().ravel()
    y = sv.y.cpu().numpy()
    rT = np.abs(T - st['T_gas'].ravel()) / st['T_gas'].ravel()
    m = st['y'] > 1e-14
    ry = np.abs(y[m] - st['y'][m]) / st['y'][m]
    print(f"{what}: T median {np.median(rT):.1e} p99 {np.percentile(rT, 99):.1e} max "
          f"{rT.max():.1e}; abundances median {np.median(ry):.1e} p99 "
          f"{np.percentile(ry, 99):.1e} max {ry.max():.1e}", flush=True)
    return rT, ry


def test_headline_loop_converges_like_the_oracle_loop(gpu):
    """16^3 headline field to convergence (tol 0.05): the same number of
    iterations as the oracle's loop, the same convergence figures per
    iteration, and the same converged state.  The loop carries rounding
    differences of the chemistry (~1e-8) through the thermal root-find and
    the columns, so the state is compared in distribution: the median cell
    to 1e-9 and 99% of values to 1e-6 (measured: 1e-11 and 3e-9)."""
    from prism3d_amd.density import PC, lognormal_median
    print("case: headline loop 16^3 vs oracle loop", flush=True)
    dens, _ = lognormal_median(16, 0.5 * PC, 1e4, 1.0, seed=42)
    sv = _solver(dens, 0.5 * PC, 1e4)
    conv = sv.run(max_iterations=40, convergence_tol=0.05)
    st, hist = _oracle_loop(dens, 0.5 * PC, 1e4, 40)
    print(f"GPU {len(sv.convergence_history)} iterations (converged {conv}), oracle "
          f"{len(hist)}", flush=True)
    assert conv and hist[-1]['delta_max'] < 0.05
    assert abs(len(sv.convergence_history) - len(hist)) <= 1
    for a, b in zip(sv.convergence_history[:10], hist[:10]):
        for k in ('dT', 'dH2', 'dCO'):
            assert a[k] == pytest.approx(b[k], rel=1e-3, abs=1e-6), (a['iteration'], k)
    rT, ry = _compare(sv, st, "converged 16^3 state")
    assert np.median(rT) < 1e-9 and np.percentile(rT, 99) < 1e-6
    assert np.median(ry) < 1e-9 and np.percentile(ry, 99) < 1e-6


@pytest.mark.parametrize("atol", [1e-20, 1e-24])
def test_headline_inloop_chemistry_all_forms_vs_oracle(gpu, atol):
    """32^3 headline field after two coupled iterations: the in-loop chemistry
    of 4096 strided cells at those conditions through the drop-in (2) and
    both batched forms (4, 6), against the oracle at the same tolerances;
    every abundance above 1e-14 within 1e-6."""
    import torch
    from prism3d_amd.chemistry import solve_opts
    from prism3d_amd.density import PC, lognormal_median
    print(f"case: headline 32^3 in-loop chemistry, atol {atol:g}", flush=True)
    dens, _ = lognormal_median(32, 0.5 * PC, 1e4, 1.0, seed=42)
    sv = _solver(dens, 0.5 * PC, 1e4)
    sv.record_inputs = np.linspace(0, sv.n_cells - 1, 4096).astype(np.int64)
    sv.recorded = []
    sv.run(max_iterations=3, convergence_tol=0.0)
    r = sv.recorded[-1]
    keys = ('n_H', 'T', 'G0', 'A_V', 'zeta', 'fsh_H2', 'fsh_CO')
    ref = O.solve_steady_batch(r['y'], *[r[k] for k in keys], t_max=r['t_max'], n_checkpoints=2,
                               rtol=1e-9, atol=atol, early_stop=False, post_mode=r['post_mode'],
                               x_c_total=r['x_c_total'], x_o_total=r['x_o_total'])[0]
    ins = [torch.tensor(r[k], device=gpu) for k in keys]
    m = ref > 1e-14
    for var in (2, 4, 6):
        y = torch.tensor(r['y'], device=gpu)
        o = solve_opts(t_max=r['t_max'], t_start=1e4, n_checkpoints=2, early_stop=False,
                       rtol=1e-9, atol=atol, post_mode=r['post_mode'], x_c_total=r['x_c_total'],
                       x_o_total=r['x_o_total'], variant=var, waves_per_cu=1 if var == 4 else 0)
        sv.chem.solve_steady_batch(y, *ins, opts=o)
        got = y.cpu().numpy()
        rel = np.abs(got[m] - ref[m]) / ref[m]
        print(f"variant {var}: max {rel.max():.2e}, p99.9 {np.percentile(rel, 99.9):.2e}",
              flush=True)
        assert rel.max() < 1e-6, var


def test_batching_field_first_iterations_vs_oracle_loop(gpu):
    """16^3 batching field (sigma_ln 1.5, n_H clipped to 1e2-1e6): three
    coupled iterations through the default batched form against the
    oracle's loop."""
    from prism3d_amd.density import PC, lognormal_median
    print("case: batching field 16^3, 3 iterations vs oracle loop", flush=True)
    dens, _ = lognormal_median(16, 0.5 * PC, 1e4, 1.5, seed=42, clip=(1e2, 1e6))
    sv = _solver(dens, 0.5 * PC, 1e4)
    sv.run(max_iterations=3, convergence_tol=0.0)
    st, hist = _oracle_loop(dens, 0.5 * PC, 1e4, 3)
    for a, b in zip(sv.convergence_history, hist):
        for k in ('dT', 'dH2', 'dCO'):
            assert a[k] == pytest.approx(b[k], rel=1e-3, abs=1e-6), (a['iteration'], k)
    rT, ry = _compare(sv, st, "batching field after 3 iterations")
    assert np.median(rT) < 1e-9 and np.percentile(rT, 99) < 1e-6
    assert np.median(ry) < 1e-9 and np.percentile(ry, 99) < 1e-6


@pytest.mark.parametrize("variant", [4, 6])
def test_cloud_ensemble_in_one_launch_vs_oracle(gpu, variant):
    """Four 8^3 clouds of the batching ensemble (cloud_ensemble: G0, median
    n_H and sigma_ln drawn per cloud), two iterations with the chemistry of
    all clouds in one persistent-kernel launch per iteration (CloudBatch),
    each cloud against its own oracle loop."""
    from prism3d_amd.batch import CloudBatch
    from prism3d_amd.density import cloud_ensemble
    print(f"case: 4-cloud ensemble, one launch per iteration, variant {variant}", flush=True)
    ens = cloud_ensemble(4, 8)
    clouds = [_solver(e['density'], e['box_size'], e['G0']) for e in ens]
    b = CloudBatch(clouds, chem_variant=variant, chem_waves_per_cu=1 if variant == 4 else 0)
    b.run(max_iterations=2, convergence_tol=0.0)
    assert b.launch_cells == [4 * 512, 4 * 512]
    for k, (c, e) in enumerate(zip(clouds, ens)):
        st, hist = _oracle_loop(e['density'], e['box_size'], e['G0'], 2)
        rT, ry = _compare(c, st, f"cloud {k} (G0 {e['G0']:.3g})")
        assert np.median(rT) < 1e-9 and np.percentile(rT, 99) < 1e-6
        assert np.median(ry) < 1e-9 and np.percentile(ry, 99) < 1e-6
