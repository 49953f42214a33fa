"""The tail of an inner iteration on the device (p3d_loop_snapshot,
p3d_damp_abundances, p3d_damp_temperature, p3d_convergence) against the
reference's own numpy expressions (solver_3d.py:330-436): np.percentile and
np.max on the masked relative changes, and the damping formulas, bit for bit."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ref_convergence(old_T, T, old_H2, H2, old_CO, CO, tol):
    """solver_3d.py:381-432 restated with numpy, as the reference runs it."""
    mask_T, mask_H2, mask_CO = old_T > 10, old_H2 > 1e-6, old_CO > 1e-8
    rel_T = (np.abs(T[mask_T] - old_T[mask_T]) / old_T[mask_T]) if np.any(mask_T) \
        else np.array([0.0])
    rel_H2 = (np.abs(H2[mask_H2] - old_H2[mask_H2]) / np.maximum(old_H2[mask_H2], 0.01)) \
        if np.any(mask_H2) else np.array([0.0])
    rel_CO = (np.abs(CO[mask_CO] - old_CO[mask_CO]) / np.maximum(old_CO[mask_CO], 1e-6)) \
        if np.any(mask_CO) else np.array([0.0])
    dT_p99 = float(np.percentile(rel_T, 99))
    dT = float(np.percentile(rel_T, 95))
    dH2 = float(np.percentile(rel_H2, 99))
    dCO = float(np.percentile(rel_CO, 99))
    return {'dT': dT, 'dT_p99': dT_p99, 'dH2': dH2, 'dCO': dCO,
            'dT_max': float(np.max(rel_T)), 'dH2_max': float(np.max(rel_H2)),
            'n_oscillating': int(np.sum(rel_T > tol)), 'delta_max': max(dT, dH2, dCO)}


def state(rng, n, frac_masked=0.3):
    old_T = 10 ** rng.uniform(0.5, 3.5, n)
    T = old_T * (1 + rng.normal(0, 0.05, n) * (rng.uniform(size=n) < 0.9))
    old_H2 = np.where(rng.uniform(size=n) < frac_masked, 1e-8, 10 ** rng.uniform(-7, -0.3, n))
    H2 = old_H2 * np.exp(rng.normal(0, 0.1, n))
    old_CO = np.where(rng.uniform(size=n) < frac_masked, 1e-12, 10 ** rng.uniform(-10, -3.9, n))
    CO = old_CO * np.exp(rng.normal(0, 0.2, n))
    # exact repeats and unchanged cells: ties at the selected ranks
    T[: n // 10] = old_T[: n // 10]
    return [old_T, T, old_H2, H2, old_CO, CO]


def device_convergence(arrs, tol):
    import torch
    from prism3d_amd import grid
    d = [torch.tensor(a, device="cuda") for a in arrs]
    rec = torch.full((12,), -7.0, dtype=torch.float64, device="cuda")
    ws = grid.conv_workspace()
    ws.fill_(np.nan)  # the kernels must not rely on a cleared workspace
    grid.convergence(*d, tol, rec, ws)
    return grid.convergence_record(rec), rec.cpu().numpy()


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 4096, 262144, 300001])
def test_convergence_figures_equal_numpy(gpu, n):
    rng = np.random.RandomState(n)
    arrs = state(rng, n)
    print(f"case: convergence n={n}", flush=True)
    got, raw = device_convergence(arrs, 0.05)
    want = ref_convergence(*arrs, 0.05)
    for k, v in want.items():
        assert got[k] == v, (k, got[k], v)
    assert raw[8] == np.sum(arrs[0] > 10) and raw[9] == np.sum(arrs[2] > 1e-6)
    assert raw[11] == (1.0 if want['delta_max'] < 0.05 else 0.0)


def test_convergence_empty_masks_and_repeat_calls(gpu):
    n = 5000
    cold = [np.full(n, 5.0), np.full(n, 7.0), np.full(n, 1e-7), np.full(n, 2e-7),
            np.full(n, 1e-9), np.full(n, 3e-9)]
    got, _ = device_convergence(cold, 0.05)
    assert got == ref_convergence(*cold, 0.05)
    assert got['delta_max'] == 0.0 and got['n_oscillating'] == 0
    # the same workspace and record over calls of different sizes
    import torch
    from prism3d_amd import grid
    ws = grid.conv_workspace()
    rec = torch.zeros(12, dtype=torch.float64, device="cuda")
    for n2, seed in ((70000, 1), (33, 2), (70000, 3)):
        arrs = state(np.random.RandomState(seed), n2)
        grid.convergence(*[torch.tensor(a, device="cuda") for a in arrs], 0.02, rec, ws)
        assert grid.convergence_record(rec) == ref_convergence(*arrs, 0.02)


def test_convergence_nan_propagates_like_numpy(gpu):
    rng = np.random.RandomState(11)
    arrs = state(rng, 1000)
    arrs[0][17] = 100.0  # a masked-in cell (old T > 10) ...
    arrs[1][17] = np.nan  # ... whose new temperature is NaN
    got, _ = device_convergence(arrs, 0.05)
    want = ref_convergence(*arrs, 0.05)
    assert np.isnan(got['dT']) and np.isnan(want['dT'])
    assert np.isnan(got['dT_max']) and np.isnan(got['dT_p99'])
    assert got['dH2'] == want['dH2'] and got['dCO'] == want['dCO']
    assert got['n_oscillating'] == want['n_oscillating']
    assert (np.isnan(got['delta_max']) and np.isnan(want['delta_max'])) or \
        got['delta_max'] == want['delta_max']


@pytest.mark.parametrize("mode", ["reference", "uniform", "oscillating"])
def test_abundance_damping_matches_reference_formulas(gpu, mode):
    import torch
    from prism3d_amd import grid
    rng = np.random.RandomState(3)
    nsp, n = 31, 20000
    y_old = 10 ** rng.uniform(-12, -0.3, (nsp, n))
    y = y_old * np.exp(rng.normal(0, 0.3, (nsp, n)))
    w = rng.uniform(0.5, 0.98, n)
    rows = [0, 1, 4, 5, 6]  # H, H2, C+, C, CO rows of this synthetic state
    xc = 1.4e-4
    dy, dyo, dw = (torch.tensor(a, device="cuda") for a in (y, y_old, w))
    grid.damp_abundances(dy, dyo, mode, rows, xc, dw)
    got = dy.cpu().numpy()
    want = y.copy()
    H, H2, Cp, C, CO = rows
    if mode == "uniform":
        want = 0.5 * y_old + 0.5 * y
    elif mode == "reference":  # solver_3d.py:345-354
        want[H2] = 0.5 * y_old[H2] + 0.5 * y[H2]
        want[H] = np.maximum(1.0 - 2.0 * want[H2], 1e-30)
        want[CO] = 0.5 * y_old[CO] + 0.5 * y[CO]
        want[Cp] = np.maximum(xc - want[CO] - y[C], 1e-30)
    else:  # solver_3d.py:330-345
        osc = w > 0.55
        want[H2][osc] = w[osc] * y_old[H2][osc] + (1 - w[osc]) * y[H2][osc]
        want[H][osc] = np.maximum(1.0 - 2.0 * want[H2][osc], 1e-30)
        want[CO][osc] = w[osc] * y_old[CO][osc] + (1 - w[osc]) * y[CO][osc]
        want[Cp][osc] = np.maximum(xc - want[CO][osc] - y[C][osc], 1e-30)
    assert np.array_equal(got, want)


def test_temperature_damping_and_snapshot(gpu):
    import torch
    from prism3d_amd import grid
    rng = np.random.RandomState(4)
    n = 50000
    old_T = 10 ** rng.uniform(0.5, 3.5, n)
    T_new = old_T * np.exp(rng.normal(0, 0.5, n))
    w = rng.uniform(0.5, 0.98, n)
    dT, dTn, dw = (torch.tensor(a, device="cuda") for a in (old_T, T_new, w))
    grid.damp_temperature(dT, dTn, dw)
    rel = np.abs(T_new - old_T) / np.maximum(old_T, 10.0)  # solver_3d.py:362-375
    w2 = np.clip(np.where(rel > 0.5, w + 0.15, np.where(rel < 0.1, w - 0.02, w)), 0.5, 0.98)
    assert np.array_equal(dw.cpu().numpy(), w2)
    assert np.array_equal(dTn.cpu().numpy(), w2 * old_T + (1.0 - w2) * T_new)
    y = torch.tensor(rng.uniform(size=(31, n)), device="cuda")
    t_old, y_old = torch.empty_like(dTn), torch.empty_like(y)
    grid.loop_snapshot(dTn, y, t_old, y_old)
    assert torch.equal(t_old, dTn) and torch.equal(y_old, y)
