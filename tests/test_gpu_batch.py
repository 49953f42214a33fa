"""Batched clouds (prism3d_amd.batch.CloudBatch): several independent grids
advanced together, their chemistry in one p3d_chem_solve launch, against
the same clouds run one by one."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE = ('T_gas', 'x_HI', 'x_H2', 'x_Cp', 'x_C', 'x_CO', 'x_O', 'x_e', 'x_OH', 'x_HCOp')


def _clouds(n_clouds, n, variant=0, seed=3):
    from prism3d_amd.density import cloud_ensemble
    from prism3d_amd.grid import directional_G0
    from prism3d_amd.solver3d import PDRSolverGPU
    out = []
    for e in cloud_ensemble(n_clouds, n, param_seed=seed):
        out.append(PDRSolverGPU(e['density'], e['box_size'],
                                G0_external=directional_G0(e['G0'], [-1.0, 0.0, 0.0], nside=1),
                                zeta_CR_0=2e-16, nside_rt=1, chemistry='network',
                                x_c_total=1.4e-4, x_o_total=3.0e-4,
                                initial_abundances={'H': 1.0, 'C+': 1.4e-4, 'O': 3.0e-4,
                                                    'e-': 1.4e-4},
                                chem_variant=variant))
    return out


def test_batch_equals_clouds_run_alone(gpu):
    """Three 8^3 clouds, up to 6 iterations at a tolerance some clouds meet
    before others: with the one-cell-per-workgroup kernel (deterministic per
    cell) every cloud's history and state equal its own run bit for bit, and
    converged clouds leave the shared launch."""
    import torch
    from prism3d_amd.batch import CloudBatch
    print("case: batch of 3 8^3 clouds vs single runs, workgroup kernel", flush=True)
    solo = _clouds(3, 8, variant=2)
    flags = [c.run(max_iterations=6, convergence_tol=0.3) for c in solo]
    batch_clouds = _clouds(3, 8, variant=2)
    b = CloudBatch(batch_clouds, chem_variant=2)
    got = b.run(max_iterations=6, convergence_tol=0.3)
    assert got == flags
    assert b.launch_cells[0] == 3 * 512
    assert all(x >= y for x, y in zip(b.launch_cells, b.launch_cells[1:]))
    for a, c in zip(solo, batch_clouds):
        assert len(a.convergence_history) == len(c.convergence_history)
        for ra, rc in zip(a.convergence_history, c.convergence_history):
            assert ra == rc
        for k in STATE:
            assert torch.equal(getattr(a, k), getattr(c, k)), k
    print("launch cells per iteration:", b.launch_cells, "converged:", got, flush=True)


def test_batch_through_the_lane_kernel(gpu):
    """Four 8^3 clouds of very different irradiation and density, one
    iteration, chemistry of all clouds in one launch of the persistent
    one-cell-per-lane kernel (with its hand-over): the same states as the
    clouds solved alone by the workgroup kernel, to the solver tolerance.
    (Over further iterations the coupled loop amplifies the two kernels'
    ~1e-8 rounding differences through the thermal root-find and the
    columns; each iteration's chemistry is checked against the oracle on
    its own inputs in test_gpu_lane.py.)"""
    from prism3d_amd.batch import CloudBatch
    print("case: batch of 4 8^3 clouds, lane kernel vs single runs", flush=True)
    solo = _clouds(4, 8, variant=2, seed=11)
    for c in solo:
        c.run(max_iterations=1, convergence_tol=0.0)
    batch_clouds = _clouds(4, 8, seed=11)
    b = CloudBatch(batch_clouds, chem_variant=4, chem_waves_per_cu=1)
    b.run(max_iterations=1, convergence_tol=0.0)
    assert (b.chem_stats[:, 0].cpu().numpy() > 0).all()
    worst = {}
    for a, c in zip(solo, batch_clouds):
        for k in STATE:
            x, y = getattr(a, k).cpu().numpy().ravel(), getattr(c, k).cpu().numpy().ravel()
            m = np.abs(x) > 1e-14
            if not m.any():
                continue
            rel = np.abs(y[m] - x[m]) / np.abs(x[m])
            worst[k] = max(worst.get(k, 0.0), float(np.percentile(rel, 99.9)))
            assert rel.max() < 1e-4, k
    print("p99.9 rel:", {k: f"{v:.1e}" for k, v in worst.items()}, flush=True)
    assert max(worst.values()) < 1e-6


def test_batch_rejects_mixed_settings(gpu):
    from prism3d_amd.batch import CloudBatch
    a, b = _clouds(2, 4)
    b.chem_rtol = 1e-6
    with pytest.raises(ValueError):
        CloudBatch([a, b])
    with pytest.raises(ValueError):
        CloudBatch([])
