"""Host-side product logic that runs without a GPU: the perfect-tree layout
of the surrogate, the THEMIS photoelectric table builder and the synthetic
input generators, against the upstream golden vectors."""

import numpy as np
import pytest

from conftest import golden

PACKED = ('left', 'right', 'feature', 'threshold', 'value', 'tree_offset', 'base', 'scale',
          'divisor', 'n_trees')


def _walk(D, feat, thr, leaf, t, X32):
    i = np.zeros(len(X32), dtype=np.int64)
    for _ in range(D):
        i = 2 * i + np.where(X32[np.arange(len(X32)), feat[t][i]] <= thr[t][i], 1, 2)
    return leaf[t][i - (2 ** D - 1)]


def _features(g):
    ins = [g['in_' + k] for k in ('n_H', 'T', 'G0', 'A_V', 'zeta', 'fsh_H2', 'fsh_CO')]
    X = np.column_stack([np.log10(np.maximum(ins[0], 1)), np.log10(np.maximum(ins[1], 5)),
                         np.log10(np.maximum(ins[2], 1e-2)), ins[3],
                         np.log10(np.maximum(ins[4], 1e-18)), np.log10(np.maximum(ins[5], 1e-10)),
                         np.log10(np.maximum(ins[6], 1e-10))])
    return ((X - g['x_mean']) / g['x_scale']).astype(np.float32)


@pytest.mark.parametrize("kind", ["gbrt", "rf"])
def test_perfect_tree_layout_reproduces_the_tree_library(kind):
    """The device layout (perfect trees, float32 round-down thresholds), walked
    on the host exactly as the kernel walks it, gives the library's own
    predictions bit for bit."""
    from prism3d_amd.surrogate import perfect_trees
    g = golden("surrogate_golden.npz")
    m = {k: g[f'{kind}_{k}'] for k in PACKED}
    D, feat, thr, leaf = perfect_trees(m)
    assert thr.dtype == np.float32 and leaf.shape == (feat.shape[0], 2 ** D)
    X32 = _features(g)
    t0 = 0
    for k in range(9):
        raw = np.full(len(X32), m['base'][k])
        for t in range(t0, t0 + int(m['n_trees'][k])):
            raw = raw + m['scale'][k] * _walk(D, feat, thr, leaf, t, X32)
        t0 += int(m['n_trees'][k])
        assert np.array_equal(raw / m['divisor'][k], g[f'{kind}_raw'][k])


def test_perfect_tree_layout_edge_cases():
    """An unbalanced tree (leaf at depth 1 beside a depth-3 branch), a
    single-leaf tree, thresholds between float32 neighbours and NaN input
    (goes right, as the library)."""
    from prism3d_amd.surrogate import _round_down_f32, perfect_trees
    t_mid = float(np.float32(0.1)) + 1e-12  # just above a float32 value
    left = np.array([1, -1, 3, -1, 5, -1, -1, -1], dtype=np.int32)
    right = np.array([2, -1, 4, -1, 6, -1, -1, -1], dtype=np.int32)
    feat = np.array([0, -2, 1, -2, 2, -2, -2, -2], dtype=np.int32)
    thr = np.array([0.0, -2, t_mid, -2, -1.0, -2, -2, -2])
    val = np.array([0, 10.0, 0, 20.0, 0, 30.0, 40.0, 50.0])
    m = dict(left=left, right=right, feature=feat, threshold=thr, value=val,
             tree_offset=np.array([0, 7, 8]), base=np.zeros(9), scale=np.ones(9),
             divisor=np.ones(9), n_trees=np.array([2] + [0] * 8, dtype=np.int32))
    D, f, t, lf = perfect_trees(m)
    assert D == 3 and f.shape == (2, 7)
    X = np.array([[-1, 0, 0], [1, 0.1, 0], [1, 0.2, -2], [1, 0.2, 0], [np.nan, 0, 0]],
                 dtype=np.float32)
    X = np.pad(X, ((0, 0), (0, 4)))
    # row 1: float32(0.1) <= t_mid goes left; row 4: NaN at the root goes right,
    # then 0 <= t_mid goes left
    assert list(_walk(D, f, t, lf, 0, X)) == [10.0, 20.0, 30.0, 40.0, 20.0]
    assert list(_walk(D, f, t, lf, 1, X)) == [50.0] * 5
    r = _round_down_f32(np.array([t_mid, 0.5, np.inf, 1e39, -1e39]))
    assert r[0] == np.float32(0.1) and r[1] == np.float32(0.5) and np.isinf(r[2])
    assert r[3] == np.finfo(np.float32).max and np.isneginf(r[4])


def test_pe_table_builder_matches_upstream():
    """The product's one-time THEMIS table build (dust_tables.PETable) is the
    upstream table bit for bit (themis_tables.py:57-151, themis.py:138-171)."""
    from prism3d_amd.dust_tables import PETable
    g = golden("pe_golden.npz")
    pe = PETable()
    assert np.array_equal(pe.table, g['pe_table'])
    assert np.array_equal(pe.sigma_abs, g['bin_sigma_abs'])
    assert np.array_equal(pe.dn_da_da, g['bin_dn_da_da'])
    assert np.array_equal(pe.is_nano, g['bin_is_nano'])
    for k in ('sigma_nano', 'sigma_largeC', 'sigma_Sil', 'log_psi_min', 'log_psi_step',
              'eg_min', 'eg_step'):
        assert getattr(pe, k) == float(g[k]), k


def test_embedded_star_cloud_and_blowout_preset_match_upstream():
    from prism3d_amd.density import PC, blowout, embedded_star_cloud
    g = golden("preset_golden.npz")
    for n, seed in ((16, 42), (12, 7)):
        d, cs, pos = embedded_star_cloud(n, 3.0 * PC, n_cloud=3000, sigma_ln=1.2, seed=seed)
        assert np.array_equal(d, g[f'd{n}']) and cs == float(g[f'cell{n}'])
        assert np.array_equal(pos, g[f'star{n}'])
    d, box, kw = blowout(16)
    assert np.array_equal(d, g['d16']) and box == 3.0 * PC
    assert kw['star']['L_FUV'] == float(g['L_FUV']) and kw['G0_external'] == 1.0


def test_cloud_ensemble_parameters_and_fields():
    """Per-cloud G0, median and sigma from one seeded draw, within the stated
    ranges; each field is the log-normal generator at its own seed, so its
    median and log-width match the drawn values; the default field seeds are
    0..n-1."""
    from prism3d_amd.density import PC, cloud_ensemble, lognormal_median
    ens = cloud_ensemble(6, 16, param_seed=5, field_seeds=[9, 8, 7, 6, 5, 4])
    again = cloud_ensemble(6, 16, param_seed=5, field_seeds=[9, 8, 7, 6, 5, 4])
    rng = np.random.RandomState(5)
    for i, e in enumerate(ens):
        g0 = 10.0 ** rng.uniform(1, 5)
        med = 10.0 ** rng.uniform(2, 6)
        sig = rng.uniform(0.5, 1.5)
        assert (e['G0'], e['median'], e['sigma_ln']) == (g0, med, sig)
        assert 1e1 <= e['G0'] <= 1e5 and 1e2 <= e['median'] <= 1e6
        assert 0.5 <= e['sigma_ln'] <= 1.5
        d, _ = lognormal_median(16, 0.5 * PC, med, sig, seed=9 - i)
        assert np.array_equal(e['density'], d)
        assert np.array_equal(again[i]['density'], d)
        ln = np.log(d)
        assert abs(np.std(ln) - sig) < 1e-9 * sig + 1e-6 or np.any(d == 1.0)
        assert abs(np.median(ln) - np.log(med)) < 0.5 * sig
    assert [e['seed'] for e in cloud_ensemble(3, 8)] == [0, 1, 2]


def test_mlp_oracle_equals_sklearn_predict():
    """The oracle's MLP forward pass (the checker of the MFMA MLP surrogate)
    against MLPRegressor.predict of a fitted network of the accelerator's
    architecture, and the host-side shape checks of MLPChemistrySurrogate."""
    import warnings
    import pytest
    from sklearn.neural_network import MLPRegressor
    from oracle import oracle as O
    from prism3d_amd.surrogate import MLPChemistrySurrogate
    rng = np.random.RandomState(0)
    X = rng.normal(size=(300, 7))
    y = np.sin(X[:, 0]) + X[:, 1] * X[:, 2]
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        m = MLPRegressor(hidden_layer_sizes=(128, 64, 32), activation='relu', max_iter=30,
                         random_state=42).fit(X, y)
    Xt = rng.normal(size=(500, 7))
    assert np.array_equal(m.predict(Xt), O.mlp_predict(m.coefs_, m.intercepts_, Xt))
    bad = [(m.coefs_, m.intercepts_)] * 8  # one short
    with pytest.raises(ValueError):
        MLPChemistrySurrogate(bad, np.zeros(7), np.ones(7), np.zeros(9), np.ones(9))
    wrong = [([c.T for c in m.coefs_], m.intercepts_)] * 9
    with pytest.raises(ValueError):
        MLPChemistrySurrogate(wrong, np.zeros(7), np.ones(7), np.zeros(9), np.ones(9))
