"""Generate the straight-line network code of the one-cell-per-lane solver.

    python -m prism3d_amd.gen_lane   # writes prism3d_amd/csrc/lane_net.h and wave_net.h

The lane solver keeps one cell per GPU lane, so every per-species array of a
cell must be indexed with compile-time constants (registers cannot be
indexed per lane).  This script turns the network structure (species,
reactants, products; prism3d_amd/network.py) into plain C++ for that:

  * rhs:       reaction rates and dy/dt, accumulated per species in the
               upstream order (reaction by reaction, reactant occurrences then
               product occurrences: chemistry/solver.py:104-143);
  * jacobian:  the analytic Jacobian, one row at a time (row entries in
               registers, then stored), every entry summed over its reaction
               terms in reaction order;
  * lu_factor: M = I - c J factorised without pivoting in a fill-reducing
               order (greedy minimum Markowitz product on the structural
               pattern), row by row (left-looking), storing the unit-lower
               multipliers, the upper factor and the reciprocal diagonal;
  * lu_solve:  the two triangular solves in that order.

Rate coefficients stay table driven (any alpha/beta/gamma of the same
structure), so the code depends only on which species each reaction reads and
writes.  The emitted header is included by csrc/lane_bdf.h.
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "csrc", "lane_net.h")
OUT_WAVE = os.path.join(HERE, "csrc", "wave_net.h")

# reaction types (include/prism3d_amd.h)
TWO_BODY, PHOTO, PHOTO_FH2, PHOTO_FCO, CR, GRAIN = 0, 1, 2, 3, 4, 5
PAH_MINUS, NONE = -2, -1

# compiler-fence spacing in the emitted code (0: none): bounds how many slot
# loads the compiler may keep in flight at once, i.e. register pressure
# against memory-latency overlap.  The RHS reads the abundances once into
# registers (measured ~3% faster on the 64^3 solves than re-reading them per
# reaction group), so its rate-coefficient loads run unfenced.
RHS_FENCE_EVERY = 0
JAC_ROWS_PER_CHUNK = 2
JAC_TERMS_PER_GROUP = 8
LU_ROWS_PER_GROUP = 5
SOLVE_ROWS_PER_GROUP = 8


def structure(net):
    """Dependency species per rate slot, net stoichiometry per reaction and
    the Jacobian term lists, built the way the library builds them for a
    network handle."""
    ns, nr = net.n_species, net.n_reactions
    rt, re, pr, e = net.rtype, net.reactants, net.products, net.e_index
    dep = np.full(2 * nr, -1, dtype=np.int64)
    for r in range(nr):
        a, b = re[2 * r], re[2 * r + 1]
        dep[2 * r] = a
        if rt[r] == TWO_BODY:
            if b == PAH_MINUS:
                dep[2 * r + 1] = e
            elif b != a:
                dep[2 * r + 1] = b
    nets = []
    for r in range(nr):
        slots = 2 if rt[r] in (TWO_BODY, GRAIN) else 1
        m = {}
        for k in range(slots):
            s = re[2 * r + k]
            if s >= 0:
                m[s] = m.get(s, 0.0) - 1.0
        for k in range(3):
            s = pr[3 * r + k]
            if s >= 0:
                m[s] = m.get(s, 0.0) + 1.0
        nets.append(sorted((s, v) for s, v in m.items()))
    terms = {}  # (i, j) -> [(r, slot, coef)] in reaction order
    for r in range(nr):
        for sl in range(2):
            j = dep[2 * r + sl]
            if j < 0:
                continue
            for i, v in nets[r]:
                if v != 0.0:
                    terms.setdefault((int(i), int(j)), []).append((r, sl, v))
    return dep, nets, terms


def markowitz(pattern, n):
    """Greedy minimum Markowitz product ordering (lowest index on ties)."""
    A = pattern.copy()
    left = list(range(n))
    order = []
    while left:
        best = None
        for k in left:
            rc = sum(1 for j in left if j != k and A[k, j])
            cc = sum(1 for i in left if i != k and A[i, k])
            if best is None or rc * cc < best[0]:
                best = (rc * cc, k)
        k = best[1]
        order.append(k)
        left.remove(k)
        rows = [i for i in left if A[i, k]]
        cols = [j for j in left if A[k, j]]
        for i in rows:
            for j in cols:
                A[i, j] = True
    return order


def fill(pattern_pos, n):
    """Filled pattern of the no-pivot LU in position order."""
    A = pattern_pos.copy()
    for k in range(n):
        rows = [i for i in range(k + 1, n) if A[i, k]]
        cols = [j for j in range(k + 1, n) if A[k, j]]
        for i in rows:
            for j in cols:
                A[i, j] = True
    return A


def generate(net=None):
    from .network import Network
    net = net or Network()
    ns, nr, e = net.n_species, net.n_reactions, net.e_index
    rt, re, pr = net.rtype, net.reactants, net.products
    dep, nets, terms = structure(net)
    pat = np.zeros((ns, ns), bool)
    for (i, j) in terms:
        pat[i, j] = True
    # every row needs its diagonal for the no-pivot factorisation
    for i in range(ns):
        if not pat[i, i]:
            raise ValueError(f"species {net.species[i]} has no diagonal Jacobian entry")
    perm = markowitz(pat, ns)               # position -> species
    pos = {s: p for p, s in enumerate(perm)}
    ppat = pat[np.ix_(perm, perm)]
    F = fill(ppat, ns)
    # storage: J entries in position-major order of the structural pattern,
    # LU entries in position-major order of the filled pattern
    jz = {}
    for pi in range(ns):
        for pj in range(ns):
            if ppat[pi, pj]:
                jz[(pi, pj)] = len(jz)
    lz = {}
    for pi in range(ns):
        for pj in range(ns):
            if F[pi, pj]:
                lz[(pi, pj)] = len(lz)
    NJ, NLU = len(jz), len(lz)
    L = []
    w = L.append
    w("// Generated by prism3d_amd/gen_lane.py from the network structure of")
    w("// prism3d_amd/network.py.  Do not edit; rerun the generator.")
    w("//")
    w(f"// {ns} species, {nr} reactions, {NJ} structural Jacobian entries,")
    w(f"// {NLU} entries in the filled no-pivot LU factors.")
    w("#pragma once")
    w("")
    w("namespace p3d {")
    w("namespace lane {")
    w("")
    w(f"constexpr int kNS = {ns}, kNR = {nr}, kNJ = {NJ}, kNLU = {NLU}, kE = {e};")
    w("constexpr int kRoles[7] = {" + ", ".join(str(int(x)) for x in net.roles) + "};")
    w("constexpr int kRtype[kNR] = {" + ", ".join(str(int(x)) for x in rt) + "};")
    w("constexpr int kReact[2 * kNR] = {" + ", ".join(str(int(x)) for x in re) + "};")
    w("constexpr int kProd[3 * kNR] = {" + ", ".join(str(int(x)) for x in pr) + "};")
    w("// elimination order: position -> species")
    w("constexpr int kPerm[kNS] = {" + ", ".join(str(int(x)) for x in perm) + "};")
    w("")
    # ------------------------------------------------------------------ rhs
    w("// dy/dt (chemistry/solver.py:69-143) at the abundances held in slot")
    w("// entries yoff..yoff+kNS, floored at 1e-30 as the rates see them;")
    w("// accumulation per species follows the reaction order.  The abundances")
    w("// are read once into registers; rate coefficients are read per reaction.")
    w("template <class Slot>")
    w("P3D_HD P3D_INLINE void rhs(const Cell &p, const Slot &m, int yoff, double (&f)[kNS]) {")
    w("#pragma unroll")
    w("  for (int i = 0; i < kNS; ++i) f[i] = 0.0;")
    w("  double yv[kNS];")
    w("#pragma unroll")
    w("  for (int i = 0; i < kNS; ++i) yv[i] = fmax((double)m.at(yoff + i), 1e-30);")
    w("  double pm, dpm;")
    w("  pah_minus(yv[kE], p, pm, dpm);")
    w("  (void)dpm;")
    for r in range(nr):
        a, b = int(re[2 * r]), int(re[2 * r + 1])
        t = int(rt[r])
        ya = f"yv[{a}]"
        if t == TWO_BODY:
            if b == PAH_MINUS:
                expr = f"m.k({r}) * ya * pm * p.nH"
                loads = f"const double ya = {ya}; "
            elif b == a:
                expr = f"m.k({r}) * ya * ya * p.nH"
                loads = f"const double ya = {ya}; "
            else:
                expr = f"m.k({r}) * ya * yb * p.nH"
                loads = f"const double ya = {ya}; const double yb = yv[{b}]; "
            slots = 2
        elif t == GRAIN:
            expr = f"m.k({r}) * ya * p.nH * 0.5"
            loads = f"const double ya = {ya}; "
            slots = 2
        else:
            expr = f"m.k({r}) * ya"
            loads = f"const double ya = {ya}; "
            slots = 1
        upd = []
        for k in range(slots):
            s_ = int(re[2 * r + k])
            if s_ >= 0:
                upd.append(f"f[{s_}] -= q;")
        for k in range(3):
            s_ = int(pr[3 * r + k])
            if s_ >= 0:
                upd.append(f"f[{s_}] += q;")
        w(f"  {{ {loads}const double q = {expr}; {' '.join(upd)} }}  // {net.names[r]}")
        if RHS_FENCE_EVERY and r % RHS_FENCE_EVERY == RHS_FENCE_EVERY - 1:
            w("  P3D_FENCE();")
    w("}")
    w("")
    # ------------------------------------------------------------- jacobian
    def partial(r, sl):
        a, b = int(re[2 * r]), int(re[2 * r + 1])
        t = int(rt[r])
        act_a = "(ya > 1e-30 ? 1.0 : 0.0)"
        if t == TWO_BODY:
            if b == PAH_MINUS:
                return ([a], f"k * pm * p.nH * {act_a}") if sl == 0 else \
                    ([a], "k * ya * p.nH * dpm * (ye > 1e-30 ? 1.0 : 0.0)")
            if a == b:
                return [a], f"2.0 * k * ya * p.nH * {act_a}"
            return ([a, b], f"k * yb * p.nH * {act_a}") if sl == 0 else \
                ([a, b], "k * ya * p.nH * (yb > 1e-30 ? 1.0 : 0.0)")
        if t == GRAIN:
            return [a], f"k * p.nH * 0.5 * {act_a}"
        return [a], f"k * {act_a}"

    w("// Analytic Jacobian at the abundances held in slot entries yoff..yoff+kNS")
    w("// (floored at 1e-30 as the rates see them) into the slot's J entries, one")
    w("// row at a time; abundances are read per term so that only the current")
    w("// row's entries stay in registers.")
    w("template <class Slot>")
    w("P3D_HD P3D_INLINE void jacobian(const Cell &p, const Slot &m, int yoff) {")
    w("  const double ye = fmax((double)m.at(yoff + kE), 1e-30);")
    w("  double pm, dpm;")
    w("  pah_minus(ye, p, pm, dpm);")
    for c0 in range(0, ns, JAC_ROWS_PER_CHUNK):
        rows = list(range(c0, min(ns, c0 + JAC_ROWS_PER_CHUNK)))
        w(f"  {{  // rows {', '.join(net.species[perm[pi]] for pi in rows)}")
        for pi in rows:
            for pj in [pj for pj in range(ns) if ppat[pi, pj]]:
                w(f"    double j{pi}_{pj} = 0.0;")
        nterm = 0
        for pi in rows:
            i = perm[pi]
            cols = [pj for pj in range(ns) if ppat[pi, pj]]
            by_r = {}
            for pj in cols:
                for (r, sl, v) in terms[(i, perm[pj])]:
                    by_r.setdefault((r, sl), []).append((pj, v))
            for q, (r, sl) in enumerate(sorted(by_r)):
                upd = " ".join(f"j{pi}_{pj} += {v!r} * d;" for pj, v in by_r[(r, sl)])
                used, expr = partial(r, sl)
                loads = f"const double ya = fmax((double)m.at(yoff + {used[0]}), 1e-30); "
                if len(used) > 1:
                    loads += f"const double yb = fmax((double)m.at(yoff + {used[1]}), 1e-30); "
                w(f"    {{ const double k = m.k({r}); {loads}const double d = {expr}; {upd} }}")
                nterm += 1
                if nterm % JAC_TERMS_PER_GROUP == 0:
                    w("    P3D_FENCE();")
        for pi in rows:
            for pj in [pj for pj in range(ns) if ppat[pi, pj]]:
                w(f"    m.J({jz[(pi, pj)]}) = j{pi}_{pj};")
        w("  }")
        w("  P3D_FENCE();")
    w("}")
    w("")
    # ------------------------------------------------------------ lu_factor
    w("// M = I - c J factorised without pivoting in kPerm order (left-looking,")
    w("// row by row).  Stores the unit-lower multipliers, the upper factor and,")
    w("// on the diagonal, 1 / u_ii.")
    w("template <class Slot>")
    w("P3D_HD P3D_INLINE void lu_factor(const Slot &m, double c) {")
    # rows grouped by dependency level: a level's rows read only rows of
    # earlier levels, so each level loads, computes and stores as one group
    lev = [0] * ns
    for pi in range(ns):
        lev[pi] = 1 + max([lev[pk] for pk in range(pi) if F[pi, pk]], default=0)
    groups = []
    for L_ in range(1, max(lev) + 1):
        rows_l = [pi for pi in range(ns) if lev[pi] == L_]
        for g0 in range(0, len(rows_l), LU_ROWS_PER_GROUP):
            groups.append((L_, rows_l[g0:g0 + LU_ROWS_PER_GROUP]))
    for L_, rows in groups:
        w(f"  {{  // level {L_}: {', '.join(net.species[perm[pi]] for pi in rows)}")
        for pi in rows:
            cols = [pj for pj in range(ns) if F[pi, pj]]
            for pj in cols:
                if (pi, pj) in jz:
                    one = "1.0" if pi == pj else "0.0"
                    w(f"    double a{pi}_{pj} = {one} - c * m.J({jz[(pi, pj)]});")
                else:
                    w(f"    double a{pi}_{pj} = 0.0;")
        for pi in rows:
            cols = [pj for pj in range(ns) if F[pi, pj]]
            for pk in [pj for pj in cols if pj < pi]:
                w(f"    a{pi}_{pk} *= m.LU({lz[(pk, pk)]});")
                for pj in [q for q in range(pk + 1, ns) if F[pk, q]]:
                    w(f"    a{pi}_{pj} -= a{pi}_{pk} * m.LU({lz[(pk, pj)]});")
        for pi in rows:
            cols = [pj for pj in range(ns) if F[pi, pj]]
            for pj in cols:
                if pj == pi:
                    w(f"    m.LU({lz[(pi, pj)]}) = 1.0 / a{pi}_{pj};")
                else:
                    w(f"    m.LU({lz[(pi, pj)]}) = a{pi}_{pj};")
        w("  }")
        w("  P3D_FENCE();")
    w("}")
    w("")
    # ------------------------------------------------------------- lu_solve
    w("// Solve M x = b in place (b in species order).")
    w("template <class Slot>")
    w("P3D_HD P3D_INLINE void lu_solve(const Slot &m, double (&b)[kNS]) {")
    for pi in range(ns):
        for pk in [q for q in range(pi) if F[pi, q]]:
            w(f"  b[{perm[pi]}] -= m.LU({lz[(pi, pk)]}) * b[{perm[pk]}];")
        if pi % SOLVE_ROWS_PER_GROUP == SOLVE_ROWS_PER_GROUP - 1:
            w("  P3D_FENCE();")
    w("  P3D_FENCE();")
    for pi in range(ns - 1, -1, -1):
        for pj in [q for q in range(ns - 1, pi, -1) if F[pi, q]]:
            w(f"  b[{perm[pi]}] -= m.LU({lz[(pi, pj)]}) * b[{perm[pj]}];")
        w(f"  b[{perm[pi]}] *= m.LU({lz[(pi, pi)]});")
        if pi % SOLVE_ROWS_PER_GROUP == 0:
            w("  P3D_FENCE();")
    w("}")
    w("")
    w("}  // namespace lane")
    w("}  // namespace p3d")
    jcells = [None] * NJ
    for (pi, pj), z in jz.items():
        jcells[z] = (perm[pi], perm[pj])
    return "\n".join(L) + "\n", dict(NJ=NJ, NLU=NLU, perm=perm, jcells=jcells)


def op_counts(net=None):
    """FP64 operations of one evaluation of each block of the sparse solve
    the generated code runs (a multiply-add counts 2): the RHS (rate products
    and stoichiometric updates), the analytic Jacobian (partials and their
    row updates), the no-pivot LU of M = I - c J in the minimum-Markowitz
    order (matrix build, multipliers, row updates, reciprocal pivots) and
    the two triangular solves.  These are the work the algorithm needs,
    independent of how a kernel stores or pivots the matrix."""
    from .network import Network
    net = net or Network()
    ns, nr = net.n_species, net.n_reactions
    rt, re, pr = net.rtype, net.reactants, net.products
    dep, nets, terms = structure(net)
    pat = np.zeros((ns, ns), bool)
    for (i, j) in terms:
        pat[i, j] = True
    for i in range(ns):
        pat[i, i] = True
    perm = markowitz(pat, ns)
    F = fill(pat[np.ix_(perm, perm)], ns)
    rhs = 0
    for r in range(nr):
        t = int(rt[r])
        rhs += 3 if t in (TWO_BODY, GRAIN) else 1
        slots = 2 if t in (TWO_BODY, GRAIN) else 1
        rhs += sum(1 for k in range(slots) if re[2 * r + k] >= 0)
        rhs += sum(1 for k in range(3) if pr[3 * r + k] >= 0)
    jac = sum(3 for _ in {(r, sl) for v in terms.values() for (r, sl, _c) in v})
    jac += 2 * sum(len(v) for v in terms.values())
    nj = int(pat.sum())
    lu = 2 * nj + ns  # M entries, reciprocal pivots
    solve = ns
    for pi in range(ns):
        for pk in range(pi):
            if F[pi, pk]:
                lu += 1 + 2 * int(F[pk, pk + 1:].sum())
                solve += 2
        solve += 2 * int(F[pi, pi + 1:].sum())
    return dict(rhs=rhs, jacobian=jac, lu=lu, solve=solve, n_species=ns)


def generate_wave(net=None):
    """Straight-line register code of the same no-pivot LU (minimum-Markowitz
    order, filled pattern) for the one-cell-per-wavefront form, lane =
    species: each lane holds its species' row of M in a[0..kNS) indexed by
    elimination position.  Step k of the factorisation broadcasts (readlane)
    the pivot lane's reciprocal diagonal and only the structural entries of
    its U row; lanes whose row has a structural multiplier at k update (a
    per-lane mask).  The solves broadcast one component per position.  Every
    entry receives its updates in the order of the one-cell-per-lane code."""
    from .network import Network
    net = net or Network()
    ns = net.n_species
    if ns > 32:
        raise ValueError("the register form holds at most 32 species")
    dep, nets, terms = structure(net)
    pat = np.zeros((ns, ns), bool)
    for (i, j) in terms:
        pat[i, j] = True
    perm = markowitz(pat, ns)
    pos = {s_: p_ for p_, s_ in enumerate(perm)}
    F = fill(pat[np.ix_(perm, perm)], ns)
    lmask = [0] * ns
    umask = [0] * ns
    for s_ in range(ns):
        i = pos[s_]
        for k in range(ns):
            if F[i, k] and k < i:
                lmask[s_] |= 1 << k
            if F[i, k] and k > i:
                umask[s_] |= 1 << k
    L = []
    w = L.append
    w("// Generated by prism3d_amd/gen_lane.py from the network structure of")
    w("// prism3d_amd/network.py.  Do not edit; rerun the generator.")
    w("//")
    w(f"// Register LU of M = I - c J for {ns} species, one cell per wavefront, lane =")
    w("// species: the no-pivot factorisation in the minimum-Markowitz order of")
    w("// lane_net.h, right-looking, broadcasting only structural U entries.")
    w("#pragma once")
    w("")
    w("namespace p3d {")
    w("namespace wave {")
    w("")
    w(f"constexpr int kNS = {ns};")
    w(f"constexpr int kNR = {net.n_reactions};  // reactions")
    w(f"constexpr int kNNZ = {int(pat.sum())};  // structural entries of the Jacobian")
    w("constexpr int kPerm[kNS] = {" + ", ".join(str(int(x)) for x in perm) + "};")
    w("// per species (lane): bit k set where its row has a structural multiplier")
    w("// (kLMask) or upper entry (kUMask) at elimination position k")
    w("#define P3D_WAVE_LMASK {" + ", ".join(f"0x{m:08x}u" for m in lmask) + "}")
    w("#define P3D_WAVE_UMASK {" + ", ".join(f"0x{m:08x}u" for m in umask) + "}")
    w("")
    # dependency levels.  Forward solve: component k is final once every
    # earlier component its row multiplies is (row dependencies).  Factor:
    # step k needs its pivot row final (row dependencies) and its column of
    # multipliers final, i.e. every earlier step k' whose pivot row has an
    # upper entry in column k (column dependencies); steps of one level touch
    # disjoint data, so their broadcasts are issued together.
    lev_r = [0] * ns
    for k in range(ns):
        lev_r[k] = 1 + max([lev_r[q] for q in range(k) if F[k, q]], default=0)
    lev_f = [0] * ns
    for k in range(ns):
        lev_f[k] = 1 + max([lev_f[q] for q in range(k) if F[k, q] or F[q, k]], default=0)
    lev_b = [0] * ns
    for k in range(ns - 1, -1, -1):
        lev_b[k] = 1 + max([lev_b[q] for q in range(k + 1, ns) if F[k, q]], default=0)
    groups = lambda lev: [[k for k in range(ns) if lev[k] == v] for v in range(1, max(lev) + 1)]
    has_l = [any((lmask[s_] >> k) & 1 for s_ in range(ns)) for k in range(ns)]
    has_u = [any((umask[s_] >> k) & 1 for s_ in range(ns)) for k in range(ns)]
    # RHS pieces: every species row (reaction order, reactant occurrences then
    # product occurrences, as the library's CSR lists) cut into pieces of at
    # most K entries, K the smallest that gives <= 64 pieces; lane l sums piece l
    rt, re, pr = net.rtype, net.reactants, net.products
    rows = [[] for _ in range(ns)]
    for r in range(net.n_reactions):
        slots = 2 if rt[r] in (TWO_BODY, GRAIN) else 1
        for k in range(slots):
            if re[2 * r + k] >= 0:
                rows[re[2 * r + k]].append(r << 1)
        for k in range(3):
            if pr[3 * r + k] >= 0:
                rows[pr[3 * r + k]].append((r << 1) | 1)
    if max(e for row in rows for e in row) >= 255:
        raise ValueError("RHS pieces hold reaction indices below 127")
    K = next(k for k in range(1, 9) if sum(-(-len(x) // k) for x in rows) <= 64)
    pieces, row_pc = [], [0]
    for row in rows:
        for c0 in range(0, len(row), K):
            pieces.append(row[c0:c0 + K])
        row_pc.append(len(pieces))
    words = []
    for pc in pieces + [[]] * (64 - len(pieces)):
        b = pc + [0xFF] * (8 - len(pc))
        words.append((b[0] | b[1] << 8 | b[2] << 16 | b[3] << 24, b[4] | b[5] << 8 | b[6] << 16 | b[7] << 24))
    w(f"// RHS pieces: lane l sums piece l, at most kPieceLen entries of one species'")
    w("// row (byte q of the two words: reaction << 1 | product, 0xff unused); the")
    w("// pieces of species s are lanes kRowPiece[s] .. kRowPiece[s + 1] - 1")
    w(f"constexpr int kPieceLen = {K};")
    w("#define P3D_WAVE_PIECE_W0 {" + ", ".join(f"0x{a:08x}u" for a, _ in words) + "}")
    w("#define P3D_WAVE_PIECE_W1 {" + ", ".join(f"0x{b_:08x}u" for _, b_ in words) + "}")
    w("#define P3D_WAVE_ROW_PIECE {" + ", ".join(str(x) for x in row_pc) + "}")
    w("constexpr int kRowPieceMax = %d;  // pieces of the longest row" %
      max(b_ - a_ for a_, b_ in zip(row_pc, row_pc[1:])))
    w("")
    w("// a: the lane's row of M (positions); lm: the lane's kLMask entry; pos, lev:")
    w("// its elimination position and P3D_WAVE_LEVELS entry.  Afterwards a[k] holds")
    w("// the multiplier l_ik below the diagonal, 1 / u_ii on it and u_ij above it")
    w("// (structural entries).  At each dependency level the level's pivots invert")
    w("// their diagonals in place at once, then the pivot rows are broadcast")
    w("// together and applied in position order.  A: an accessor of the lane's row")
    w("// in LDS (a[pos] indexes it at run time; bc reads the pivot row there).")
    w("template <class A, class BC>")
    w("__device__ __forceinline__ void lu_factor(A &a, unsigned lm, int lane, const BC &bc,")
    w("                                          int pos, int lev) {")
    w("  const int lvk = lev >> 16;")
    for g in groups(lev_f):
        w(f"  {{  // level {lev_f[g[0]]}: {', '.join(net.species[perm[k]] for k in g)}")
        # the pivots of this level invert their own diagonals in one vector
        # division; the other lanes read the stored reciprocals
        w(f"    if (lvk == {lev_f[g[0]]}) a[pos] = 1.0 / a[pos];")
        for k in g:
            P = perm[k]
            w(f"    const double inv{k} = bc(a[{k}], {P});")
            if has_l[k]:
                for j in range(k + 1, ns):
                    if F[k, j]:
                        w(f"    const double u{k}_{j} = bc(a[{j}], {P});")
        for k in g:
            P = perm[k]
            if has_l[k]:
                w(f"    if ((lm >> {k}) & 1u) {{")
                w(f"      const double m = a[{k}] * inv{k};")
                for j in range(k + 1, ns):
                    if F[k, j]:
                        w(f"      a[{j}] = a[{j}] - m * u{k}_{j};")
                w(f"      a[{k}] = m;")
                w("    }")
        w("  }")
    w("}")
    w("")
    # per lane (species): elimination position and its forward / backward
    # solve levels (inactive lanes: a spare slot and level 0, never written)
    lpos = [pos[s_] if s_ < ns else 31 for s_ in range(64)]
    llev = [(lev_r[pos[s_]] | (lev_b[pos[s_]] << 8) | (lev_f[pos[s_]] << 16))
            if s_ < ns else 0 for s_ in range(64)]
    # a component whose broadcast reaches a lane that is not final yet has to
    # be masked: forward, lane i with an upper entry at k > pos(i) in a later
    # level than k; backward, lane i with a multiplier at k < pos(i) in a
    # later backward level than k
    fw_mask = [any(F[i, k] and k > i and lev_r[i] > lev_r[k] for i in range(ns))
               for k in range(ns)]
    bw_mask = [any(F[i, k] and k < i and lev_b[i] > lev_b[k] for i in range(ns))
               for k in range(ns)]
    w("// per lane: elimination position (31: no row) and the forward | backward << 8")
    w("// dependency levels of the solves | the factorisation's level << 16")
    w("#define P3D_WAVE_POS {" + ", ".join(str(x) for x in lpos) + "}")
    w("#define P3D_WAVE_LEVELS {" + ", ".join(str(x) for x in llev) + "}")
    w("")
    w("// Solve M x = b with the factors; b and x: the lane's species component.")
    w("// Level by level: the components of one level are broadcast, each lane of")
    w("// the level records its component in xs[pos], then every lane subtracts")
    w("// a[k] x_k with no mask.  A lane reached by a component it has no entry")
    w("// for in this triangle (an entry of the other triangle, or its own")
    w("// diagonal) is always final by then, so only its register is spoilt and")
    w("// the recorded xs stays exact; the few components for which that is not")
    w("// so keep their mask.  dinv: the lane's 1 / u_ii.")
    w("template <class A, class BC>")
    w("__device__ __forceinline__ double lu_solve(const A &a, unsigned lm, unsigned um, int lane,")
    w("                                          double b, const BC &bc, double *xs, int pos,")
    w("                                          int lev, double dinv) {")
    w("  const int lvf = lev & 0xff, lvb = (lev >> 8) & 0xff;")
    for g in groups(lev_r):
        w("  {")
        gl = [k for k in g if has_l[k]]
        for k in gl:
            w(f"    const double x{k} = bc(b, {perm[k]});")
        w(f"    if (lvf == {lev_r[g[0]]}) xs[pos] = b;")
        for k in gl:
            if fw_mask[k]:
                w(f"    if ((lm >> {k}) & 1u) b -= a[{k}] * x{k};")
            else:
                w(f"    b -= a[{k}] * x{k};")
        w("  }")
    w("  b = xs[pos];")
    for g in groups(lev_b):
        w("  {")
        w("    const double bx = b * dinv;")
        w(f"    if (lvb == {lev_b[g[0]]}) xs[pos] = bx;")
        gu = [k for k in sorted(g, reverse=True) if has_u[k]]
        for k in gu:
            w(f"    const double x{k} = bc(bx, {perm[k]});")
        for k in gu:
            if bw_mask[k]:
                w(f"    if ((um >> {k}) & 1u) b -= a[{k}] * x{k};")
            else:
                w(f"    b -= a[{k}] * x{k};")
        w("  }")
    w("  return xs[pos];")
    w("}")
    w("")
    # transposed solve M^T x = b in the column layout (lane s holds column
    # pos(s) of the factors): U^T z = b forward, then L^T x = z backward
    lev_ft = [0] * ns
    for q in range(ns):
        lev_ft[q] = 1 + max([lev_ft[i] for i in range(q) if F[i, q]], default=0)
    lev_bt = [0] * ns
    for q in range(ns - 1, -1, -1):
        lev_bt[q] = 1 + max([lev_bt[i] for i in range(q + 1, ns) if F[i, q]], default=0)
    ft_mask = [any(F[k, q] and k > q and lev_ft[q] > lev_ft[k] for q in range(ns))
               for k in range(ns)]
    bt_mask = [any(F[k, q] and k < q and lev_bt[q] > lev_bt[k] for q in range(ns))
               for k in range(ns)]
    tum = [0] * ns  # per species: bit i where its column has a U^T entry (F[i, pos], i < pos)
    tlm = [0] * ns  # bit i where its column has an L^T entry (F[i, pos], i > pos)
    for s_ in range(ns):
        q = pos[s_]
        for i in range(ns):
            if F[i, q] and i < q:
                tum[s_] |= 1 << i
            if F[i, q] and i > q:
                tlm[s_] |= 1 << i
    tlev = [(lev_ft[pos[s_]] | (lev_bt[pos[s_]] << 8)) if s_ < ns else 0 for s_ in range(64)]
    w("// Transposed solve: per species (lane), bit i where its column of the")
    w("// factors holds a U^T (upper of the lane's row of U^T) or L^T entry, and")
    w("// the forward | backward << 8 levels of the transposed solves")
    w("#define P3D_WAVE_TUMASK {" + ", ".join(f"0x{m:08x}u" for m in tum) + "}")
    w("#define P3D_WAVE_TLMASK {" + ", ".join(f"0x{m:08x}u" for m in tlm) + "}")
    w("#define P3D_WAVE_TLEVELS {" + ", ".join(str(x) for x in tlev) + "}")
    w("")
    w("// Solve M^T x = b with the factors of lu_factor in the column layout: a[i]")
    w("// is entry (i, pos) of the factored matrix in elimination positions, i.e.")
    w("// u_(i,pos) above the diagonal, l_(i,pos) below it and 1 / u_(pos,pos) on")
    w("// it (dinv).  U^T z = b runs level by level with the division, then")
    w("// L^T x = z; components are recorded in xs as in lu_solve, and the")
    w("// components that could reach a lane that is not final keep a mask (um:")
    w("// the lane's U^T entries, lm: its L^T entries).")
    w("template <class BC>")
    w("__device__ __forceinline__ double lu_solve_t(const double (&a)[32], unsigned lm, unsigned um,")
    w("                                            int lane, double b, const BC &bc, double *xs,")
    w("                                            int pos, int lev, double dinv) {")
    w("  const int lvf = lev & 0xff, lvb = lev >> 8;")
    has_ut = [any(F[k, q] for q in range(k + 1, ns)) for k in range(ns)]  # row k of U has entries
    has_lt = [any(F[k, q] for q in range(k)) for k in range(ns)]          # row k of L has entries
    for g in groups(lev_ft):
        w("  {")
        w("    const double bx = b * dinv;")
        w(f"    if (lvf == {lev_ft[g[0]]}) xs[pos] = bx;")
        gu = [k for k in g if has_ut[k]]
        for k in gu:
            w(f"    const double x{k} = bc(bx, {perm[k]});")
        for k in gu:
            if ft_mask[k]:
                w(f"    if ((um >> {k}) & 1u) b -= a[{k}] * x{k};")
            else:
                w(f"    b -= a[{k}] * x{k};")
        w("  }")
    w("  b = xs[pos];")
    for g in groups(lev_bt):
        w("  {")
        w(f"    if (lvb == {lev_bt[g[0]]}) xs[pos] = b;")
        gl = [k for k in sorted(g, reverse=True) if has_lt[k]]
        for k in gl:
            w(f"    const double x{k} = bc(b, {perm[k]});")
        for k in gl:
            if bt_mask[k]:
                w(f"    if ((lm >> {k}) & 1u) b -= a[{k}] * x{k};")
            else:
                w(f"    b -= a[{k}] * x{k};")
        w("  }")
    w("  return xs[pos];")
    w("}")
    w("")
    w("}  // namespace wave")
    w("}  // namespace p3d")
    return "\n".join(L) + "\n"


def write_wave(path=OUT_WAVE):
    text = generate_wave()
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


def write(path=OUT):
    text, info = generate()
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path, info


if __name__ == "__main__":
    write_wave()
    p, info = write()
    print(p, "J entries", info["NJ"], "LU entries", info["NLU"], file=sys.stderr)
