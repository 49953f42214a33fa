"""GPU parity of the tree-ensemble chemistry surrogate (the upstream ML
accelerator's inference) and of the driver with it, plus the blowout preset
with its embedded star, through the C ABI."""

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

PACKED = ('left', 'right', 'feature', 'threshold', 'value', 'tree_offset', 'base', 'scale',
          'divisor', 'n_trees')
INPUTS = ['n_H', 'T', 'G0', 'A_V', 'zeta', 'fsh_H2', 'fsh_CO']


def _surrogate(g, kind):
    from prism3d_amd.surrogate import ChemistrySurrogate
    return ChemistrySurrogate({k: g[f'{kind}_{k}'] for k in PACKED}, g['x_mean'], g['x_scale'],
                              g['y_mean'], g['y_scale'])


def _rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


@pytest.mark.parametrize("kind", ["gbrt", "rf"])
def test_surrogate_matches_upstream_predict_batch(gpu, kind):
    """Tree walks select the library's leaves exactly; the abundances differ
    from the upstream ones by the rounding of log10 / 10^y only."""
    from prism3d_amd.surrogate import SPECIES
    g = golden("surrogate_golden.npz")
    print(f"case: surrogate {kind}", flush=True)
    s = _surrogate(g, kind)
    out = s.predict_batch(*[g['in_' + k] for k in INPUTS])
    ref = g[f'{kind}_out']
    for k, sp in enumerate(SPECIES):
        got = out[sp].cpu().numpy()
        assert _rel(got, ref[k]).max() < 1e-13, sp


def _random_ensemble(rng, n_trees, depth, p_split=0.85):
    """Nine outputs of n_trees random unbalanced trees of depth <= `depth`
    (node ids in creation order, as the tree library stores them)."""
    left, right, feat, thr, val, off = [], [], [], [], [], [0]
    for _ in range(9 * n_trees):
        L, R, F, TH, V = [], [], [], [], []

        def grow(d):
            k = len(L)
            L.append(-1); R.append(-1); F.append(-2); TH.append(-2.0)
            V.append(rng.normal())
            if d < depth and (d == 0 or rng.rand() < p_split):
                F[k] = rng.randint(7)
                TH[k] = rng.normal() * 1.2
                L[k] = grow(d + 1)
                R[k] = grow(d + 1)
            return k
        grow(0)
        left += L; right += R; feat += F; thr += TH; val += V
        off.append(off[-1] + len(L))
    return dict(left=np.array(left), right=np.array(right), feature=np.array(feat),
                threshold=np.array(thr), value=np.array(val), tree_offset=np.array(off),
                base=rng.normal(size=9), scale=np.full(9, 0.1), divisor=np.ones(9),
                n_trees=np.full(9, n_trees, dtype=np.int32))


@pytest.mark.parametrize("n_trees,depth", [(200, 6), (7, 2), (13, 3), (9, 8), (5, 10)])
def test_surrogate_ensembles_match_oracle(gpu, n_trees, depth):
    """Random unbalanced ensembles: 200 stages of depth 6 (the upstream
    training settings) and the depths either side of the LDS-staged kernel's
    range (4-8) that run the direct kernel, on 20k random cells."""
    import torch
    from oracle import oracle as O
    from prism3d_amd.surrogate import SPECIES, ChemistrySurrogate
    print(f"case: ensemble {n_trees} trees, depth {depth}", flush=True)
    rng = np.random.RandomState(11 + depth)
    packed = _random_ensemble(rng, n_trees, depth)
    xm, xs = rng.normal(size=7), rng.uniform(0.5, 2, 7)
    ym, ys = rng.uniform(-8, -2, 9), rng.uniform(0.2, 1.0, 9)
    s = ChemistrySurrogate(packed, xm, xs, ym, ys)
    assert s.depth == depth
    n = 20000
    ins = [10 ** rng.uniform(0, 7, n), 10 ** rng.uniform(0.5, 4, n), 10 ** rng.uniform(-3, 5, n),
           rng.uniform(0, 20, n), 10 ** rng.uniform(-19, -15, n), 10 ** rng.uniform(-11, 0, n),
           10 ** rng.uniform(-11, 0, n)]
    out = s.predict_batch(*ins)
    ref = O.surrogate_predict_batch(packed, xm, xs, ym, ys, *ins)
    for k, sp in enumerate(SPECIES):
        rel = _rel(out[sp].cpu().numpy(), ref[k])
        assert np.median(rel) < 1e-15 and rel.max() < 1e-12, sp
    # into preallocated rows of a state array
    y = torch.zeros((9, n), dtype=torch.float64, device=gpu)
    dev_in = [torch.tensor(v, device=gpu) for v in ins]
    s.predict_into([y[k] for k in range(9)], *dev_in)
    assert torch.equal(y[2], out['C+'])
    s.close()


def test_surrogate_save_load_and_errors(gpu, tmp_path):
    from prism3d_amd.surrogate import ChemistrySurrogate
    g = golden("surrogate_golden.npz")
    s = _surrogate(g, 'gbrt')
    p = tmp_path / "model.npz"
    s.save(str(p))
    s2 = ChemistrySurrogate.load(str(p))
    ins = [g['in_' + k] for k in INPUTS]
    a, b = s.predict_batch(*ins), s2.predict_batch(*ins)
    for k in a:
        assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy())
    with pytest.raises(ValueError):
        ChemistrySurrogate({k: g[f'gbrt_{k}'] for k in PACKED}, g['x_mean'][:6], g['x_scale'],
                           g['y_mean'], g['y_scale'])
    bad = {k: g[f'gbrt_{k}'].copy() for k in PACKED}
    bad['feature'][0] = 9  # out of range
    with pytest.raises(ValueError):
        ChemistrySurrogate(bad, g['x_mean'], g['x_scale'], g['y_mean'], g['y_scale'])
    s.close()
    s2.close()


def test_driver_with_surrogate_matches_upstream_run(gpu):
    """run() with the accelerator on the 8^3 fixture (3 iterations) against the
    upstream run: convergence figures and state to rounding."""
    from prism3d_amd.solver3d import PDRSolverGPU
    g = golden("surrogate_golden.npz")
    r = golden("run_golden.npz")
    sv = PDRSolverGPU(r['density'].reshape(8, 8, 8), float(r['box']), G0_external=100.0,
                      zeta_CR_0=2e-16, nside_rt=1)
    sv.accelerator = _surrogate(g, 'gbrt')
    sv.run(max_iterations=3, convergence_tol=1e-9)
    h = sv.convergence_history
    for k in ('dT', 'dH2', 'dCO'):
        assert [x[k] for x in h] == pytest.approx(list(g['run_' + k]), rel=1e-9), k
    for k in ('T_gas', 'x_HI', 'x_H2', 'x_Cp', 'x_C', 'x_CO', 'x_O', 'x_e', 'x_OH', 'x_HCOp'):
        rel = _rel(getattr(sv, k).cpu().numpy().ravel(), g['run_' + k])
        print(k, "max rel", rel.max(), flush=True)
        assert rel.max() < 1e-9, k


def test_blowout_preset_first_iteration_matches_oracle(gpu):
    """The blowout preset (embedded O5V star in a 3 pc cloud, point-source RT,
    G0 = 1 background) at 16^3: one inner iteration against the restated loop,
    columns exactly and the state within the oracle's own last-bit spread."""
    from oracle import oracle as O
    from prism3d_amd.density import blowout
    from prism3d_amd.solver3d import PDRSolverGPU
    dens, box, kw = blowout(16)
    sv = PDRSolverGPU(dens, box, **kw)
    sv.run(max_iterations=1, convergence_tol=1e-9)
    st, hist = O.run_loop(dens, box, 1.0, 2e-16, 1, 1, star=kw['star'])
    sp, _ = O.run_loop(dens * (1.0 + 2e-16), box, 1.0, 2e-16, 1, 1, star=kw['star'])
    assert np.array_equal(sv.N_H.cpu().numpy(), st['N_H'])
    assert _rel(sv.G0.cpu().numpy(), st['G0']).max() < 1e-14
    for k in ('T_gas', 'x_H2', 'x_CO', 'x_e'):
        rel = _rel(getattr(sv, k).cpu().numpy().ravel(), st[k].ravel())
        own = _rel(sp[k].ravel(), st[k].ravel())
        assert np.median(rel) < 1e-12, k
        assert np.mean(rel > 1e-6) <= 1.5 * np.mean(own > 1e-6) + 0.01, k


def _random_mlps(rng, widths):
    dims = [7] + list(widths) + [1]
    out = []
    for _ in range(9):
        coefs = [rng.normal(size=(a, b)) * np.sqrt(2.0 / a) for a, b in zip(dims[:-1], dims[1:])]
        inter = [rng.normal(size=b) * 0.1 for b in dims[1:]]
        out.append((coefs, inter))
    return out


@pytest.mark.parametrize("widths", [(128, 64, 32), (20, 7), (128,), (16, 16, 16, 16)])
def test_mlp_surrogate_matches_oracle(gpu, widths):
    """The accelerator's MLP model type on FP64 MFMA tiles against the
    oracle's forward pass (pinned to MLPRegressor.predict), 20k random cells:
    the upstream architecture (128, 64, 32) and widths that need padding."""
    import torch
    from oracle import oracle as O
    from prism3d_amd.surrogate import SPECIES, ChemistrySurrogate, MLPChemistrySurrogate
    print(f"case: MLP surrogate {widths}", flush=True)
    rng = np.random.RandomState(len(widths) * 7 + widths[0])
    mlps = _random_mlps(rng, widths)
    xm, xs = rng.normal(size=7), rng.uniform(0.5, 2, 7)
    ym, ys = rng.uniform(-8, -2, 9), rng.uniform(0.2, 1.0, 9)
    s = MLPChemistrySurrogate(mlps, xm, xs, ym, ys)
    n = 20000 + 5  # not a multiple of the 16-cell tile
    ins = [10 ** rng.uniform(0, 7, n), 10 ** rng.uniform(0.5, 4, n), 10 ** rng.uniform(-3, 5, n),
           rng.uniform(0, 20, n), 10 ** rng.uniform(-19, -15, n), 10 ** rng.uniform(-11, 0, n),
           10 ** rng.uniform(-11, 0, n)]
    out = s.predict_batch(*ins)
    ref = O.surrogate_predict_batch(mlps, xm, xs, ym, ys, *ins)
    from prism3d_amd.chemistry import GAS_O
    for k, sp in enumerate(SPECIES):
        got = out[sp].cpu().numpy()
        if sp == 'O':  # x_O_total - CO - OH: rounding relative to the total
            assert np.max(np.abs(got - ref[k])) < 1e-12 * GAS_O
        else:
            rel = _rel(got, ref[k])
            assert rel.max() < 1e-12, (sp, rel.max())
    if widths == (128, 64, 32):
        import tempfile, os
        with tempfile.TemporaryDirectory() as d:
            p = os.path.join(d, "mlp.npz")
            s.save(p)
            s2 = ChemistrySurrogate.load(p)
            assert isinstance(s2, MLPChemistrySurrogate)
            b = s2.predict_batch(*ins)
            for sp in SPECIES:
                assert torch.equal(b[sp], out[sp])
            s2.close()
    s.close()


def test_mlp_surrogate_rejects_bad_models(gpu):
    from prism3d_amd.surrogate import MLPChemistrySurrogate
    rng = np.random.RandomState(0)
    mlps = _random_mlps(rng, (256,))
    with pytest.raises(ValueError):  # wider than the 128-unit tile budget
        MLPChemistrySurrogate(mlps, np.zeros(7), np.ones(7), np.zeros(9), np.ones(9))
