"""The static elimination schedule of the sparse no-pivot LU (variant 7),
read from the library's host-only p3d_sparse_lu_schedule, checked on the CPU:

- the order is the greedy minimum-Markowitz order of the generator that
  writes the bundled network's register LU (gen_lane.markowitz, an
  independent implementation);
- the factorisation, emulated pass by pass as the kernel runs it (a lane
  reads other rows as they stood when the pass began, so a pass that read
  what the same pass writes would give a wrong answer), and the dense
  substitutions in position order solve M x = b to rounding against LAPACK
  for the bundled, the 64-species and the 103-species networks, with the
  factors zero outside the filled structure the schedule names."""

import ctypes

import numpy as np
import pytest


def _structure(which):
    from prism3d_amd.gen_lane import structure
    from prism3d_amd.network import Network
    if which == "bundled":
        net = Network()
    else:
        from prism3d_amd import network_synth as S
        sp, rx = S.synthetic_network() if which == "synth64" else S.synthetic_network_wide()
        net = Network(sp, rx)
    _, _, terms = structure(net)
    ij = np.array(sorted(terms), dtype=np.int32)
    return net.n_species, ij[:, 0].copy(), ij[:, 1].copy()


def _schedule(ns, jrow, jcol):
    from prism3d_amd import _lib
    L = _lib.lib()
    cap = ns * ns
    out = {k: np.zeros(ns, np.int32) for k in ("pos", "lev", "lvp")}
    out.update(lvoff=np.zeros(ns + 1, np.int32), uoff=np.zeros(ns + 1, np.int32),
               uj=np.zeros(cap, np.int32))
    counts = np.zeros(2, np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    rc = L.p3d_sparse_lu_schedule(ns, jrow.size, P(jrow), P(jcol), cap, P(out["pos"]),
                                  P(out["lev"]), P(out["lvoff"]), P(out["lvp"]), P(out["uoff"]),
                                  P(out["uj"]), P(counts))
    assert rc == 0, L.p3d_last_error()
    out["nlev"] = int(counts[0])
    out["uj"] = out["uj"][:counts[1]]
    out["lvoff"] = out["lvoff"][:out["nlev"] + 1]
    return out


def _factor_solve(s, M0, b):
    """The kernel's passes on a host copy (M0 and b in species order): in
    pass v every pivot k of level v, in position order, is applied to each
    row p > k with a nonzero entry at k; other rows are read as they stood
    when the pass began.  Then the dense substitutions in position order."""
    ns = M0.shape[0]
    perm = np.argsort(s["pos"])             # position -> species
    M = M0[np.ix_(perm, perm)].copy()       # position coordinates
    lev = s["lev"]
    for p in np.nonzero(lev == 0)[0]:
        M[p, p] = 1.0 / M[p, p]
    for v in range(s["nlev"]):
        snap = M.copy()
        for k in s["lvp"][s["lvoff"][v]:s["lvoff"][v + 1]]:
            js = s["uj"][s["uoff"][k]:s["uoff"][k + 1]]
            for p in range(k + 1, ns):
                if M[p, k] != 0.0:
                    l = M[p, k] * snap[k, k]
                    M[p, k] = l
                    M[p, js] -= l * snap[k, js]
        for p in np.nonzero(lev == v + 1)[0]:
            M[p, p] = 1.0 / M[p, p]
    x = b[perm].copy()
    for k in range(ns - 1):
        x[k + 1:] -= M[k + 1:, k] * x[k]
    for k in range(ns - 1, -1, -1):
        x[k] *= M[k, k]
        x[:k] -= M[:k, k] * x[k]
    out = np.empty(ns)
    out[perm] = x
    return out, M


def test_order_is_the_generators_markowitz_order():
    from prism3d_amd.gen_lane import markowitz
    ns, jr, jc = _structure("bundled")
    s = _schedule(ns, jr, jc)
    A = np.eye(ns, dtype=bool)
    A[jr, jc] = True
    order = markowitz(A, ns)
    assert list(np.argsort(s["pos"])) == list(order)


@pytest.mark.parametrize("which", ["bundled", "synth64", "wide103"])
def test_level_schedule_solves_against_lapack(which):
    ns, jr, jc = _structure(which)
    s = _schedule(ns, jr, jc)
    assert s["nlev"] <= ns and s["lvoff"][-1] == ns
    assert sorted(s["lvp"]) == list(range(ns))
    # the upper entries are those of the generator's fill (gen_lane.fill) of
    # the structure in the schedule's order
    from prism3d_amd.gen_lane import fill
    perm = np.argsort(s["pos"])
    A = np.eye(ns, dtype=bool)
    A[jr, jc] = True
    filled = fill(A[np.ix_(perm, perm)], ns)
    for p in range(ns):
        assert list(s["uj"][s["uoff"][p]:s["uoff"][p + 1]]) == list(np.nonzero(filled[p, p + 1:])[0] + p + 1)
    rng = np.random.RandomState(5)
    for c in (1e-3, 1e2, 1e8):
        # M = I - c J with a chemistry-like J: non-negative formation terms
        # off the diagonal, destruction (column losses) on it
        J = np.zeros((ns, ns))
        J[jr, jc] = rng.uniform(0.0, 1.0, jr.size) * 10.0 ** rng.uniform(-3, 0, jr.size)
        np.fill_diagonal(J, 0.0)
        J[np.arange(ns), np.arange(ns)] = -(J.sum(axis=0) + rng.uniform(0.1, 1.0, ns))
        M = np.eye(ns) - c * J
        b = rng.standard_normal(ns)
        x, F = _factor_solve(s, M, b)
        ref = np.linalg.solve(M, b)
        # the factors vanish outside the filled structure
        assert not np.any(F[~filled])
        assert np.max(np.abs(x - ref)) <= 1e-10 * np.max(np.abs(ref)), (which, c)


def test_schedule_rejects_bad_structure():
    from prism3d_amd import _lib
    L = _lib.lib()
    z = np.zeros(16, np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    bad = np.array([0, 5], np.int32)
    assert L.p3d_sparse_lu_schedule(3, 2, P(bad), P(bad), 8, *[P(z)] * 7) != 0
    assert b"out of range" in L.p3d_last_error()


def _packed(ns, jrow, jcol):
    from prism3d_amd import _lib
    L = _lib.lib()
    cap = ns * ns + jrow.size + 64
    desc = np.zeros(ns, np.uint32)
    ub, zb, pb = np.zeros(cap, np.uint8), np.zeros(cap, np.uint8), np.zeros(128, np.uint8)
    sizes = np.zeros(2, np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    V = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = L.p3d_sparse_lu_packed(ns, jrow.size, P(jrow), P(jcol), cap, V(desc), V(ub), V(zb),
                                V(pb), P(sizes))
    assert rc == 0, L.p3d_last_error()
    assert sizes[0] % 4 == 0 and sizes[1] % 4 == 0 and sizes[1] >= jrow.size
    return desc, ub[:sizes[0]], zb[:sizes[1]], pb


def _factor_solve_packed(desc, ub, zb, pb, jrow, jcol, Jv, c, b):
    """LinSparseG as the kernel runs it, from the packed arrays: each row
    zeroed, the Jacobian entries scattered by zb, the descriptors walked in
    order (k, len, ubeg, last-of-level decoded by the kernel's shifts and
    masks), rows read as they stood when the level began, a row inverting its
    diagonal at the end of the level before its own; then the dense
    substitutions in position order."""
    ns = desc.size
    p_of = pb[:ns].astype(int)
    M = np.zeros((ns, ns))
    for s_ in range(ns):
        M[p_of[s_], p_of[s_]] = 1.0
    for z in range(jrow.size):
        p, j = p_of[jrow[z]], int(zb[z])
        M[p, j] = (1.0 if j == p else 0.0) - c * Jv[z]
    # level of each row: the pass at whose end its diagonal inverts
    lev = np.zeros(ns, int)
    v = 0
    for q in range(ns):
        lev[int(desc[q]) & 127] = v
        if desc[q] >> 31:
            v += 1
    assert v == int(np.sum(desc >> 31))
    for p in np.nonzero(lev == 0)[0]:
        M[p, p] = 1.0 / M[p, p]
    v = 0
    snap = M.copy()
    for q in range(ns):
        d = int(desc[q])
        k, ln, u0 = d & 127, (d >> 7) & 127, (d >> 14) & 0x1FFFF
        assert u0 % 4 == 0
        js = ub[u0:u0 + ln].astype(int)
        for p in range(k + 1, ns):
            if M[p, k] != 0.0:
                l = M[p, k] * snap[k, k]
                M[p, k] = l
                M[p, js] -= l * snap[k, js]
        if d >> 31:
            for p in np.nonzero(lev == v + 1)[0]:
                M[p, p] = 1.0 / M[p, p]
            v += 1
            snap = M.copy()
    x = np.zeros(ns)
    x[p_of] = b
    for k in range(ns - 1):
        x[k + 1:] -= M[k + 1:, k] * x[k]
    for k in range(ns - 1, -1, -1):
        x[k] *= M[k, k]
        x[:k] -= M[:k, k] * x[k]
    return x[p_of]


@pytest.mark.parametrize("which", ["bundled", "wide103"])
def test_packed_schedule_as_the_kernel_reads_it(which):
    """The packed arrays the kernel stages in LDS, decoded with the kernel's
    field layout, factor and solve M = I - c J to 1e-10 of LAPACK; the
    descriptors list every pivot once, in the schedule's level order."""
    ns, jr, jc = _structure(which)
    s = _schedule(ns, jr, jc)
    desc, ub, zb, pb = _packed(ns, jr, jc)
    assert list(desc & 127) == list(s["lvp"])
    assert int(np.sum(desc >> 31)) == s["nlev"]
    assert list(pb[:ns]) == list(s["pos"])
    rng = np.random.RandomState(8)
    for c in (1e-2, 1e6):
        Jv = rng.uniform(0.0, 1.0, jr.size) * 10.0 ** rng.uniform(-3, 0, jr.size)
        diag = jr == jc
        J = np.zeros((ns, ns))
        J[jr[~diag], jc[~diag]] = Jv[~diag]
        Jv[diag] = -(J.sum(axis=0)[jc[diag]] + 0.5)
        J[jr[diag], jc[diag]] = Jv[diag]
        b = rng.standard_normal(ns)
        x = _factor_solve_packed(desc, ub, zb, pb, jr, jc, Jv, c, b)
        ref = np.linalg.solve(np.eye(ns) - c * J, b)
        assert np.max(np.abs(x - ref)) <= 1e-10 * np.max(np.abs(ref)), (which, c)
