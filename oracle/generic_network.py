"""CPU oracle for a chemistry network given as a reaction list.

TEST INFRASTRUCTURE ONLY (the tests use it as the checker; the product
never imports it).

The upstream RHS is written over whatever reaction list the network holds
(prism3d/chemistry/solver.py:43-143: the per-reaction rate forms of
:100-133, the PAH- closed form of :74-87, the floor max(y, 1e-30) and the
accumulation reactant by reactant, product by product of :135-141), and the
upstream solve integrates it with scipy's BDF (solver.py:195-201).  This
module restates both for an arbitrary list of
(name, type, reactants, products, alpha, beta, gamma) tuples, the format
of the bundled table, so networks larger than the bundled one have a
checker.

Pinned by: the 31-species oracle (p3d_oracle.c, itself pinned against the
upstream golden vectors) on the bundled network, to one rounding in the rates
and to rounding in the RHS (tests/test_oracle_golden.py); nuclei
conservation of every reaction that balances.
"""

import numpy as np

TWO_BODY = ('two_body', 'recomb', 'charge_exchange', 'grain_recomb')
PHOTO = ('photodiss', 'photoion')
CR = ('cosmic_ray', 'cr_photo')
SHIELDED = ('H2_photodiss', 'CO_photodiss')  # solver.py:116-121


class GenericNetwork:
    """Index tables of a (species, reactions) network."""

    def __init__(self, species, reactions):
        self.species = list(species)
        self.idx = {s: i for i, s in enumerate(self.species)}
        self.ns, self.nr = len(self.species), len(reactions)
        self.kind = []
        self.a = np.array([r[4] for r in reactions], dtype=np.float64)
        self.b = np.array([r[5] for r in reactions], dtype=np.float64)
        self.g = np.array([r[6] for r in reactions], dtype=np.float64)
        self.reac = np.full((self.nr, 2), -1, dtype=np.int64)   # -2: PAH-
        self.prod = []
        for r, (name, typ, reac, prod, *_x) in enumerate(reactions):
            if typ in TWO_BODY:
                k = 0
            elif typ in PHOTO:
                k = 2 if name == SHIELDED[0] else (3 if name == SHIELDED[1] else 1)
            elif typ in CR:
                k = 4
            elif typ == 'grain_surface':
                k = 5
            else:
                raise ValueError(f"unknown reaction type {typ!r}")
            self.kind.append(k)
            for m, s in enumerate(reac[:2]):
                self.reac[r, m] = -2 if s == 'PAH-' else self.idx.get(s, -1)
            self.prod.append([self.idx[s] for s in prod if s in self.idx])
        self.kind = np.array(self.kind)
        self.e = self.idx['e-']
        self.h = self.idx['H']
        # accumulation list in the upstream order: per reaction its tracked
        # reactants (-rate), then its products (+rate)
        acc_i, acc_r, acc_s = [], [], []
        for r in range(self.nr):
            for i in self.reac[r]:
                if i >= 0:
                    acc_i.append(i); acc_r.append(r); acc_s.append(-1.0)
            for i in self.prod[r]:
                acc_i.append(i); acc_r.append(r); acc_s.append(1.0)
        self.acc_i, self.acc_r = np.array(acc_i), np.array(acc_r)
        self.acc_s = np.array(acc_s)

    def rate_coeffs(self, nH, T, G0, AV, zeta, fH2=1.0, fCO=1.0):
        """k_r (solver.py:100-133); the grain form keeps n_H / 2 outside."""
        Tc = min(max(T, 10.0), 41000.0)
        k = np.empty(self.nr)
        for r in range(self.nr):
            a, b, g, t = self.a[r], self.b[r], self.g[r], self.kind[r]
            if t == 0:
                k[r] = a * (Tc / 300.0) ** b * np.exp(-g / Tc)
            elif t in (1, 2, 3):
                k[r] = a * G0 * np.exp(-g * AV) * (fH2 if t == 2 else fCO if t == 3 else 1.0)
            elif t == 4:
                k[r] = a * zeta
            else:
                k[r] = a * (T / 300.0) ** b
        return k

    def pah_minus(self, ye, nH, T, G0):
        """PAH- abundance from the floored electron abundance (solver.py:74-87)."""
        ye = max(ye, 1e-10)
        ne = ye * nH
        if G0 > 0.0 and ne > 0.0:
            psi = G0 * np.sqrt(T) / ne
            return 1e-7 * max(500.0 / (psi + 500.0) * 0.5, 0.0)
        return 1e-7 * 0.4

    def rates(self, y, k, nH, T, G0):
        """Per-reaction rates at max(y, 1e-30) (solver.py:89-133)."""
        yf = np.maximum(y, 1e-30)
        pm = self.pah_minus(yf[self.e], nH, T, G0)
        i0, i1 = self.reac[:, 0], self.reac[:, 1]
        y0 = yf[np.maximum(i0, 0)]
        y1 = np.where(i1 == -2, pm, yf[np.maximum(i1, 0)])
        return np.where(self.kind == 0, k * y0 * y1 * nH,
                        np.where(self.kind == 5, k * yf[self.h] * nH * 0.5, k * y0))

    def jacobian(self, y, k, nH, T, G0):
        """Analytic d(dy/dt)/dy: the floor max(y, 1e-30) is the identity
        above 1e-30 and constant below it (as the 31-species oracle's)."""
        yf = np.maximum(y, 1e-30)
        live = (y > 1e-30).astype(np.float64)
        ye = max(yf[self.e], 1e-10)
        pm = self.pah_minus(yf[self.e], nH, T, G0)
        dpm = 0.0
        if G0 > 0.0 and ye * nH > 0.0 and yf[self.e] > 1e-10:
            psi = G0 * np.sqrt(T) / (ye * nH)
            dpm = 1e-7 * 250.0 * psi / ((psi + 500.0) ** 2 * ye)
        # d rate_r / d y_j as (r, j, value) triples
        R, Jc, V = [], [], []
        for r in range(self.nr):
            i0, i1 = self.reac[r]
            t = self.kind[r]
            if t == 0:
                if i1 == -2:
                    R += [r, r]; Jc += [i0, self.e]
                    V += [k[r] * pm * nH * live[i0], k[r] * yf[i0] * dpm * nH * live[self.e]]
                elif i0 == i1:
                    R.append(r); Jc.append(i0); V.append(2.0 * k[r] * yf[i0] * nH * live[i0])
                else:
                    R += [r, r]; Jc += [i0, i1]
                    V += [k[r] * yf[i1] * nH * live[i0], k[r] * yf[i0] * nH * live[i1]]
            elif t == 5:
                R.append(r); Jc.append(self.h); V.append(k[r] * nH * 0.5 * live[self.h])
            else:
                R.append(r); Jc.append(i0); V.append(k[r] * live[i0])
        dr = np.zeros((self.nr, self.ns))
        np.add.at(dr, (np.array(R), np.array(Jc)), np.array(V))
        S = np.zeros((self.ns, self.nr))
        np.add.at(S, (self.acc_i, self.acc_r), self.acc_s)
        return S @ dr

    def rhs(self, y, k, nH, T, G0):
        """dy/dt at y with rate coefficients k (solver.py:89-141), summed
        reaction by reaction in the upstream order."""
        rate = self.rates(y, k, nH, T, G0)
        d = np.zeros(self.ns)
        np.add.at(d, self.acc_i, self.acc_s * rate[self.acc_r])
        return d


def integrate(net, y0, nH, T, G0, AV, zeta, fH2, fCO, t_end, t_start=1e4, rtol=1e-10,
              atol=1e-26, analytic_jacobian=True):
    """y(t_end) of one cell from y0 at t_start with scipy's BDF on the
    restated RHS, with the analytic Jacobian (or scipy's finite-difference
    one, as upstream, when analytic_jacobian is False)."""
    from scipy.integrate import solve_ivp
    k = net.rate_coeffs(nH, T, G0, AV, zeta, fH2, fCO)
    jac = (lambda t, y: net.jacobian(y, k, nH, T, G0)) if analytic_jacobian else None
    sol = solve_ivp(lambda t, y: net.rhs(y, k, nH, T, G0), (t_start, t_end),
                    np.asarray(y0, dtype=np.float64), method='BDF', rtol=rtol, atol=atol,
                    jac=jac)
    if not sol.success:
        raise RuntimeError(sol.message)
    return sol.y[:, -1]
