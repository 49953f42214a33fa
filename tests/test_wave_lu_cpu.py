"""The generated register LU of the one-cell-per-wavefront form
(prism3d_amd/csrc/wave_net.h) executed on the CPU: every statement of the
generated factorisation and solves runs for all 64 lanes at once, as a
wavefront runs it (a broadcast bc(v, P) reads lane P's value at that point of
the program), and the solution of M x = b is checked against LAPACK for
random matrices M = I - c J with the network's structural Jacobian pattern.
This pins the schedule (elimination order, dependency levels, lane masks) of
the generated code, independently of the GPU."""

import os
import re

import numpy as np
import pytest

from conftest import ROOT

HDR = os.path.join(ROOT, "prism3d_amd", "csrc", "wave_net.h")


def _parse():
    txt = open(HDR).read()
    perm = [int(x) for x in re.search(r"kPerm\[kNS\] = \{([^}]*)\}", txt).group(1).split(",")]
    lpos = [int(x) for x in re.search(r"P3D_WAVE_POS \{([^}]*)\}", txt).group(1).split(",")]
    llev = [int(x) for x in re.search(r"P3D_WAVE_LEVELS \{([^}]*)\}", txt).group(1).split(",")]
    lmask = [int(x.strip().rstrip("u"), 16) for x in
             re.search(r"P3D_WAVE_LMASK \{([^}]*)\}", txt).group(1).split(",")]
    umask = [int(x.strip().rstrip("u"), 16) for x in
             re.search(r"P3D_WAVE_UMASK \{([^}]*)\}", txt).group(1).split(",")]
    def body(start):
        i = txt.index(start)
        i = txt.index("{\n", i) + 2
        return [l.strip() for l in txt[i:txt.index("\n}\n", i)].split("\n") if l.strip()]
    tum = [int(x.strip().rstrip("u"), 16) for x in
           re.search(r"P3D_WAVE_TUMASK \{([^}]*)\}", txt).group(1).split(",")]
    tlm = [int(x.strip().rstrip("u"), 16) for x in
           re.search(r"P3D_WAVE_TLMASK \{([^}]*)\}", txt).group(1).split(",")]
    tlev = [int(x) for x in re.search(r"P3D_WAVE_TLEVELS \{([^}]*)\}", txt).group(1).split(",")]
    _parse.transposed = (tum, tlm, np.array(tlev), body("double lu_solve_t"))
    return (perm, lmask, umask, np.array(lpos), np.array(llev), body("void lu_factor"),
            body("double lu_solve("))


def _run(lines, env):
    """Interpret the generated statements, vectorised over 64 lanes."""
    lane = np.arange(64)
    stack = []  # active-lane masks of the open if-blocks
    act = lambda: np.logical_and.reduce(stack) if stack else np.ones(64, bool)
    for l in lines:
        if l.startswith("//") or l in ("{", "}") or l.startswith("{  //") or l == "return b;":
            if l == "}" and stack:
                stack.pop()
            continue
        m = re.match(r"if \(\(lm >> (\d+)\) & 1u\) \{$", l)
        if m:
            stack.append(((env["lm"] >> int(m.group(1))) & 1).astype(bool))
            continue
        if l in ("const int lvf = lev & 0xff, lvb = lev >> 8;",
                 "const int lvf = lev & 0xff, lvb = (lev >> 8) & 0xff;"):
            env["lvf"], env["lvb"] = env["lev"] & 0xFF, (env["lev"] >> 8) & 0xFF
            continue
        if l == "const int lvk = lev >> 16;":
            env["lvk"] = env["lev"] >> 16
            continue
        m = re.match(r"if \(lvk == (\d+)\) a\[pos\] = 1\.0 / a\[pos\];$", l)
        if m:
            # the pivots of this factorisation level invert their own diagonals
            sel = (env["lvk"] == int(m.group(1))) & act()
            lanes = np.nonzero(sel)[0]
            env["a"][env["pos"][lanes], lanes] = 1.0 / env["a"][env["pos"][lanes], lanes]
            continue
        m = re.match(r"if \((lvf|lvb) == (\d+)\) xs\[pos\] = (b|bx);$", l)
        if m:
            # the lanes of this level record their component (LDS: one slot per position)
            sel = (env[m.group(1)] == int(m.group(2))) & act()
            env["xs"][env["pos"][sel]] = env[m.group(3)][sel]
            continue
        if l == "b = xs[pos];":
            env["b"] = env["xs"][env["pos"]].copy()
            continue
        if l == "return xs[pos];":
            env["ret"] = env["xs"][env["pos"]].copy()
            continue
        m = re.match(r"(?:\{ )?const double (\w+) = (.*);(?: \})?$", l)
        if m:
            name, e = m.groups()
            env[name] = _eval(e, env)
            continue
        m = re.match(r"if \(\((lm|um) >> (\d+)\) & 1u\) (.*);$", l)
        if m:
            msk = ((env[m.group(1)] >> int(m.group(2))) & 1).astype(bool)
            _assign(m.group(3), env, msk & act())
            continue
        m = re.match(r"if \(lane == (\d+)\) (.*);$", l)
        if m:
            _assign(m.group(2), env, (lane == int(m.group(1))) & act())
            continue
        if l.endswith(";"):
            _assign(l[:-1], env, act())
            continue
        raise AssertionError(f"unparsed generated line: {l}")


def _eval(e, env):
    e = re.sub(r"bc\(a\[(\d+)\], (\d+)\)", r"A[\1][\2]", e)
    e = re.sub(r"bc\(b, (\d+)\)", r"B[\1]", e)
    e = re.sub(r"bc\(bx, (\d+)\)", r"bx[\1]", e)
    e = re.sub(r"a\[(\d+)\]", r"A[\1]", e)
    e = re.sub(r"\bb\b", "B", e)
    v = eval(e, {}, {**env, "A": env["a"], "B": env["b"]})
    return np.broadcast_to(np.asarray(v, dtype=np.float64), (64,)).copy()


def _assign(stmt, env, mask):
    m = re.match(r"(a\[(\d+)\]|b) ([-*]?=) (.*)$", stmt)
    assert m, stmt
    tgt, idx, op, e = m.groups()
    v = _eval(e, env)
    cur = env["a"][int(idx)] if idx else env["b"]
    new = v if op == "=" else (cur - v if op == "-=" else cur * v)
    res = np.where(mask, new, cur)
    if idx:
        env["a"][int(idx)] = res
    else:
        env["b"] = res


def _pattern():
    from prism3d_amd.gen_lane import structure
    from prism3d_amd.network import Network
    net = Network()
    _, _, terms = structure(net)
    pat = np.zeros((net.n_species, net.n_species), bool)
    for (i, j) in terms:
        pat[i, j] = True
    return pat


@pytest.mark.parametrize("seed", range(6))
def test_generated_register_lu_solves_like_lapack(seed):
    perm, lmask, umask, lpos, llev, fac, sol = _parse()
    pat = _pattern()
    ns = pat.shape[0]
    rng = np.random.RandomState(seed)
    J = np.where(pat, rng.normal(size=pat.shape) * 10 ** rng.uniform(-3, 3, pat.shape), 0.0)
    c = 10 ** rng.uniform(-3, 1)
    M = np.eye(ns) - c * J
    # a strongly diagonal matrix keeps the no-pivot LU well defined
    M[np.diag_indices(ns)] += np.sign(np.diag(M)) * (np.abs(M).sum(axis=1) + 1.0)
    J = (np.eye(ns) - M) / c
    lanes = np.arange(64)
    a = np.zeros((32, 64))
    for q in range(ns):
        for s in range(ns):
            a[q, s] = (1.0 if perm[q] == s else 0.0) - c * J[s, perm[q]]
    lm = np.array([lmask[s] if s < ns else 0 for s in lanes], dtype=np.int64)
    um = np.array([umask[s] if s < ns else 0 for s in lanes], dtype=np.int64)
    env = {"a": a, "lm": lm, "um": um, "lane": lanes, "b": np.zeros(64), "pos": lpos,
           "lev": llev}
    _run(fac, env)
    rhs = rng.normal(size=ns)
    # the solve's inputs as LinStatic::solve passes them: the lane's 1 / u_ii
    # (a[pos]), its position and levels, and a 32-slot record in LDS whose
    # spare slot 31 holds garbage
    dinv = np.array([env["a"][lpos[s], s] if s < ns else 0.0 for s in lanes])
    xs = np.full(32, np.nan)
    env.update({"b": np.concatenate([rhs, np.zeros(64 - ns)]), "pos": lpos, "lev": llev,
                "dinv": dinv, "xs": xs})
    _run(sol, env)
    x = env["ret"][:ns]
    ref = np.linalg.solve(M, rhs)
    assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref)) * np.linalg.cond(M)
    # transposed solve M^T x = g in the column layout: lane s holds entry
    # (i, pos(s)) of the factors at a[i], gathered from the row layout
    tum, tlm, tlev, solt = _parse.transposed
    fac_rows = env["a"].copy()
    at = np.zeros((32, 64))
    for s in range(ns):
        for i in range(ns):
            at[i, s] = fac_rows[lpos[s], perm[i]]
    g = rng.normal(size=ns)
    envt = {"a": at, "lane": lanes, "pos": lpos, "lev": tlev, "dinv": dinv,
            "xs": np.full(32, np.nan), "b": np.concatenate([g, np.zeros(64 - ns)]),
            "lm": np.array([tlm[s] if s < ns else 0 for s in lanes], dtype=np.int64),
            "um": np.array([tum[s] if s < ns else 0 for s in lanes], dtype=np.int64)}
    _run(solt, envt)
    xt = envt["ret"][:ns]
    reft = np.linalg.solve(M.T, g)
    assert np.max(np.abs(xt - reft)) <= 1e-12 * np.max(np.abs(reft)) * np.linalg.cond(M)
