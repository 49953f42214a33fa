"""The synthetic 64-species network and the oracle's generic-network
restatement (CPU only)."""

import numpy as np

from conftest import golden


def test_synthetic_network_is_balanced_and_reproducible():
    from prism3d_amd.network import Network
    from prism3d_amd.network_synth import (EXTRA, balanced, element_totals_check, formula,
                                           synthetic_network)
    sp, rx = synthetic_network()
    sp2, rx2 = synthetic_network()
    assert sp == sp2 and rx == rx2
    assert len(sp) == 64 and len(set(sp)) == 64
    assert len(rx) >= 1000
    assert element_totals_check(sp, rx) == []
    for s in EXTRA:
        assert any(s in r[3] for r in rx), s    # formed
        assert any(s in r[2] for r in rx), s    # destroyed
    assert formula('C2H2+') == ({'C': 2, 'H': 2}, 1)
    assert formula('HCNH+') == ({'H': 2, 'C': 1, 'N': 1}, 1)
    assert not balanced(['C', 'O'], ['CO+'])
    net = Network(sp, rx)
    assert net.n_species == 64 and net.n_reactions == len(rx)
    assert (net.roles >= 0).all()


def test_wide_synthetic_network_is_balanced_and_reproducible():
    """The 103-species network: reproducible, every reaction balancing
    nuclei and charge, every added species formed and destroyed, and its
    reaction list conserving each element (C S = 0)."""
    from prism3d_amd.network import Network
    from prism3d_amd.network_synth import (EXTRA, EXTRA_WIDE, element_totals_check, formula,
                                           synthetic_network_wide)
    sp, rx = synthetic_network_wide()
    assert (sp, rx) == synthetic_network_wide()
    assert len(sp) == 103 and len(set(sp)) == 103 and len(rx) == 1349
    assert sp[64:] == EXTRA_WIDE and sp[31:64] == EXTRA
    assert element_totals_check(sp, rx) == []
    for s in EXTRA_WIDE:
        assert any(s in r[3] for r in rx), s
        assert any(s in r[2] for r in rx), s
    net = Network(sp, rx)
    elems = sorted({e for x in sp for e in formula(x)[0]})
    C = np.array([[formula(x)[0].get(e, 0) for x in sp] for e in elems], dtype=np.float64)
    assert np.abs(C @ net.stoichiometry()).max() == 0.0


def test_generic_oracle_equals_the_bundled_oracle():
    """The reaction-list restatement against the 31-species C oracle (pinned
    by the upstream golden vectors) on random states: rates to one rounding,
    RHS and analytic Jacobian to rounding of their largest entries."""
    from oracle import oracle as O
    from oracle.generic_network import GenericNetwork
    from prism3d_amd.network import REACTIONS, TRACKED
    net = GenericNetwork(TRACKED, REACTIONS)
    rng = np.random.RandomState(2)
    for _ in range(20):
        par = (10 ** rng.uniform(2, 6), 10 ** rng.uniform(1, 3.5), 10 ** rng.uniform(0, 5),
               rng.uniform(0, 10), 10 ** rng.uniform(-17, -15), 10 ** rng.uniform(-4, 0),
               10 ** rng.uniform(-4, 0))
        y = 10 ** rng.uniform(-25, -1, 31)
        y[rng.randint(31)] = 1e-32  # one species below the floor
        k = net.rate_coeffs(*par)
        ko = O.rate_coeffs(*par)
        assert np.max(np.abs(k - ko) / np.maximum(np.abs(ko), 1e-300)) < 5e-16
        d, do = net.rhs(y, k, *par[:3]), O.rhs(y, *par)
        assert np.max(np.abs(d - do)) <= 1e-15 * np.max(np.abs(do))
        J, Jo = net.jacobian(y, k, *par[:3]), O.jacobian(y, *par)
        assert np.max(np.abs(J - Jo)) <= 1e-15 * np.max(np.abs(Jo))


def test_generic_jacobian_directional_derivatives_on_the_synthetic_network():
    """J v against central differences of the RHS along v = y * u (u random
    signs), at states where every species is well above the floor, so the
    difference quotient resolves every term."""
    from oracle.generic_network import GenericNetwork
    from prism3d_amd.network_synth import synthetic_network
    sp, rx = synthetic_network()
    net = GenericNetwork(sp, rx)
    rng = np.random.RandomState(3)
    for _ in range(5):
        par = (10 ** rng.uniform(2, 5), 10 ** rng.uniform(1, 3), 10 ** rng.uniform(0, 3),
               rng.uniform(0, 5), 1e-16, 0.1, 0.1)
        y = 10 ** rng.uniform(-6, -2, len(sp))
        k = net.rate_coeffs(*par)
        J = net.jacobian(y, k, *par[:3])
        v = y * rng.choice([-1.0, 1.0], len(sp))
        h = 1e-6
        fd = (net.rhs(y + h * v, k, *par[:3]) - net.rhs(y - h * v, k, *par[:3])) / (2 * h)
        scale = np.abs(J) @ np.abs(v)
        assert np.max(np.abs(J @ v - fd) / (scale + 1e-300)) < 1e-6


def test_synthetic_fixture_conserves_nuclei():
    """The scipy reference states of both synthetic networks keep every
    element's nuclei (the reactions all balance; PAH recombinations move
    only charge), to the integration tolerance."""
    from prism3d_amd.network_synth import formula, synthetic_network, synthetic_network_wide
    for name, make in (("synth_golden.npz", synthetic_network),
                       ("synth_wide_golden.npz", synthetic_network_wide)):
        g = golden(name)
        sp, _ = make()
        elems = sorted({e for s in sp for e in formula(s)[0]})
        C = np.array([[formula(s)[0].get(e, 0) for s in sp] for e in elems], dtype=np.float64)
        t0 = C @ g['y0']
        assert len(g['y_end']) >= 5
        for y in g['y_end']:
            assert np.max(np.abs(C @ y - t0) / t0) < 1e-8
