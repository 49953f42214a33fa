"""The one-cell-per-lane solver's code (prism3d_amd/csrc/lane_bdf.h and the
generated lane_net.h) built for the host and checked against the oracle:
generated RHS, Jacobian and no-pivot LU, then the whole trip-stepped solve
against the oracle BDF and the upstream golden vectors.  The gfx950 kernel
compiles the same source; tests/test_gpu_lane.py checks it on the GPU."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "lane_host.cpp")


@pytest.fixture(scope="module")
def lanelib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("lane") / "lane_host.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-o", out, SRC], check=True)
    lib = ctypes.CDLL(out)
    lib.lane_solve_host.restype = ctypes.c_int64
    return lib


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def net_tables():
    from prism3d_amd.network import Network
    net = Network()
    return net, [np.ascontiguousarray(net.rtype, dtype=np.int32), np.ascontiguousarray(net.alpha),
                 np.ascontiguousarray(net.beta), np.ascontiguousarray(net.gamma),
                 np.ascontiguousarray(net.post_w.reshape(-1))]


def lane_solve(lib, y, p, t_max=0.0, t_start=1e4, nk=20, early_stop=1, rtol=1e-4, atol=1e-22,
               stop_rtol=1e-3, post_mode=0, max_steps=100000):
    from oracle import oracle as O
    net, tabs = net_tables()
    y = np.ascontiguousarray(y, dtype=np.float64).copy()
    n = y.shape[1]
    cells = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))
             for v in p]
    conv = np.zeros(n, dtype=np.int32)
    stats = np.zeros((n, 4), dtype=np.int32)
    d = ctypes.c_double
    lib.lane_solve_host(ctypes.c_int64(n), P(y), *[P(c) for c in cells], *[P(t) for t in tabs],
                        d(t_max), d(t_start), nk, early_stop, d(rtol), d(atol), d(stop_rtol),
                        post_mode, max_steps, d(O.GAS_C), d(O.GAS_O), P(conv), P(stats))
    return y, conv, stats


def test_generated_header_is_current():
    from prism3d_amd import gen_lane
    text, info = gen_lane.generate()
    assert open(gen_lane.OUT).read() == text
    assert info["NJ"] == 232 and info["NLU"] == 260


def _random_states(n, seed):
    from oracle import oracle as O
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n):
        y = 10 ** rng.uniform(-14, -3, O.NSP)
        y[O.IDX['H']] = rng.uniform(0.0, 1.0)
        y[O.IDX['H2']] = rng.uniform(0.0, 0.5)
        if rng.rand() < 0.2:
            y[rng.randint(O.NSP)] = 1e-35  # below the floor: inactive partials
        cell = [10 ** rng.uniform(1, 6), 10 ** rng.uniform(1, 4), 10 ** rng.uniform(-1, 5),
                rng.uniform(0, 20), 10 ** rng.uniform(-17, -15), rng.uniform(0, 1),
                rng.uniform(0, 1)]
        out.append((y, cell))
    return out


def test_generated_rhs_matches_oracle(lanelib):
    from oracle import oracle as O
    for y, cell in _random_states(200, 1):
        k = O.rate_coeffs(*cell)
        f = np.zeros(O.NSP)
        c7 = np.array(cell)
        lanelib.lane_rhs_host(P(c7), P(k), P(y), P(f))
        ref = O.rhs(y, *cell)
        sc = np.abs(ref) + 1e-300 + np.max(np.abs(ref)) * 1e-14
        assert np.max(np.abs(f - ref) / sc) < 1e-12


def test_generated_jacobian_matches_oracle(lanelib):
    from oracle import oracle as O
    from prism3d_amd import gen_lane
    _, info = gen_lane.generate()
    rows = np.array([c[0] for c in info["jcells"]])
    cols = np.array([c[1] for c in info["jcells"]])
    for y, cell in _random_states(200, 2):
        k = O.rate_coeffs(*cell)
        jz = np.zeros(info["NJ"])
        c7 = np.array(cell)
        lanelib.lane_jac_host(P(c7), P(k), P(y), P(jz))
        J = O.jacobian(y, *cell)
        dense = np.zeros_like(J)
        dense[rows, cols] = jz
        # identical term order: bit-exact against the oracle's analytic Jacobian
        sc = np.abs(J) + 1e-300
        assert np.max(np.abs(dense - J) / sc) < 1e-13
        assert np.all((J != 0) <= (dense != 0) | (J == 0))


def test_no_pivot_lu_solves(lanelib):
    """(I - c J) x = b by the generated statically ordered LU, over step
    scales c = 1e-2 .. 1e17 s, against an iteratively refined solution (long
    double residuals), in the integrator's error norm (atol 1e-20, rtol
    1e-9).  Random states include badly conditioned matrices, so the bound
    is the forward-error bound of a stable solve, err <= 1e-15 cond(M) +
    1e-12, with rounding-level error in the typical case."""
    from oracle import oracle as O
    from prism3d_amd import gen_lane
    _, info = gen_lane.generate()
    rows = np.array([c[0] for c in info["jcells"]])
    cols = np.array([c[1] for c in info["jcells"]])
    errs = []
    for y, cell in _random_states(40, 3):
        k = O.rate_coeffs(*cell)
        jz = np.zeros(info["NJ"])
        lanelib.lane_jac_host(P(np.array(cell)), P(k), P(y), P(jz))
        J = np.zeros((O.NSP, O.NSP))
        J[rows, cols] = jz
        for c in 10.0 ** np.arange(-2, 18, 2.0):
            b = np.random.RandomState(0).standard_normal(O.NSP) * (1e-20 + 1e-9 * y)
            x = b.copy()
            lanelib.lane_lu_solve_host(P(jz), ctypes.c_double(c), P(x))
            M = np.eye(O.NSP) - c * J
            xt = np.linalg.solve(M, b)
            Ml = M.astype(np.longdouble)
            for _ in range(5):
                r = (b.astype(np.longdouble) - Ml @ xt.astype(np.longdouble)).astype(np.float64)
                xt = xt + np.linalg.solve(M, r)
            sc = 1e-20 + 1e-9 * np.abs(y)
            e = np.sqrt(np.mean(((x - xt) / sc) ** 2)) / max(np.sqrt(np.mean((xt / sc) ** 2)),
                                                              1e-300)
            errs.append(e)
            assert e <= 1e-15 * np.linalg.cond(M) + 1e-12
    assert np.median(errs) < 1e-13


def test_lane_solver_tight_matches_oracle_and_upstream(lanelib):
    """Integration to 1e7 yr at rtol 1e-9 / atol 1e-24, one span, no
    corrections: the trip-stepped lane solve against the oracle BDF on the
    same inputs and against scipy on the upstream RHS (tight fixture)."""
    from oracle import oracle as O
    g = golden("tight_golden.npz")
    n = g['n_H'].size
    y0 = np.repeat(g['y0'][:, None], n, axis=1)
    p = [g[k] for k in ('n_H', 'T', 'G0', 'A_V', 'zeta', 'fsh_H2', 'fsh_CO')]
    t_end = float(g['t_end'])
    y, conv, st = lane_solve(lanelib, y0, p, t_max=t_end, nk=2, early_stop=0, rtol=1e-9,
                             atol=1e-24, post_mode=-1)
    ref, _, _ = O.solve_steady_batch(y0, *p, t_max=t_end, n_checkpoints=2, rtol=1e-9,
                                     atol=1e-24, early_stop=False, post_mode=-1)
    m = ref > 1e-14
    e_or = np.max(np.abs(y[m] - ref[m]) / ref[m])
    up = g['y_end'].T
    m = up > 1e-14
    e_up = np.max(np.abs(y[m] - up[m]) / up[m])
    print(f"lane tight: vs oracle {e_or:.2e}, vs upstream-scipy {e_up:.2e}, "
          f"steps {st[:, 0].mean():.0f}")
    assert e_or < 1e-6
    assert e_up < 1e-6
    assert np.all(st[:, 0] > 0)


def test_lane_solver_upstream_mode_matches_golden(lanelib):
    """Upstream defaults (20 checkpoints, early stop, clipped corrections) on
    the steady fixture: within the fixture's own tolerance."""
    from oracle import oracle as O
    g = golden("steady_golden.npz")
    n = g['n_H'].size
    y0 = np.full((O.NSP, n), 1e-30)
    for q, k in enumerate(g['init_keys']):
        y0[O.IDX[str(k)]] = g['x_init'][:, q]
    p = [g[k] for k in ('n_H', 'T', 'G0', 'A_V', 'zeta', 'fsh_H2', 'fsh_CO')]
    y, conv, _ = lane_solve(lanelib, y0, p)
    ref = g['x_out'].T
    m = ref > 1e-12
    rel = np.abs(y[m] - ref[m]) / ref[m]
    assert np.mean(rel < 1e-3) > 0.99
    assert np.array_equal(conv.astype(bool), g['converged'].astype(bool)) or \
        np.mean(conv.astype(bool) == g['converged'].astype(bool)) > 0.9
