"""THEMIS grain populations and the photoelectric-heating lookup table.

One-time host set-up of the table the PE kernel interpolates
(grains/themis_tables.py:41-151, populations of grains/themis.py:106-289):
per grain size bin, the charge-averaged yield x photoelectron energy on a
(log10 psi, E_g) grid, plus the per-bin cross sections and the per-population
surface areas used by the H2 formation rate.
"""

import numpy as np

M_H = 1.6735575e-24
EV = 1.602176634e-12
E_CHARGE = 4.80320451e-10


def _population(name, a_min, a_max, n_bins, rho, e_g, mass_fraction):
    edges = np.logspace(np.log10(a_min), np.log10(a_max), n_bins + 1)
    a = np.sqrt(edges[:-1] * edges[1:])
    da = edges[1:] - edges[:-1]
    raw = a ** -3.5
    mass = np.sum((4. / 3.) * np.pi * a ** 3 * rho * raw * da)
    dn_da = raw * (mass_fraction * 0.01 * M_H / mass)
    x = 2 * np.pi * a / 0.1e-4  # size parameter at 1000 A
    if name.startswith('aC'):
        q = np.where(x > 1, 1.0, x) * np.clip(1.0 + 0.5 * (0.1 - min(e_g, 0.5)), 0.5, 1.5)
    else:
        q = np.where(x > 1, 0.7, 0.7 * x)
    return {'name': name, 'a': a, 'da': da, 'dn_da': dn_da, 'q_fuv': q,
            'carbon': name.startswith('aC')}


def themis_populations(f_nano=1.0, E_g=0.1):
    """Nano carbons (12 bins), large carbons (10), silicates (10)."""
    nano = _population('aC_nano', 4e-8, 2e-7, 12, 1.6, E_g, 0.005 * f_nano)
    nano['dn_da'] = nano['dn_da'] * (1e-7 * f_nano / np.sum(nano['dn_da'] * nano['da']))
    return [nano,
            _population('aC_large', 2e-7, 2e-5, 10, 1.6, 0.1, 0.30),
            _population('aSil', 2e-7, 3e-5, 10, 3.5, 0.1, 0.695)]


class PETable:
    """Lookup table (n_bins, n_psi, n_eg) and the per-bin factors."""

    def __init__(self, f_nano=1.0, E_g=0.1, n_psi=200, n_eg=20):
        pops = themis_populations(f_nano, E_g)
        lp = np.linspace(-1, 7, n_psi)
        egs = np.linspace(0.01, 2.5, n_eg)
        self.n_psi, self.n_eg = n_psi, n_eg
        self.log_psi_min, self.log_psi_step = float(lp[0]), float(lp[1] - lp[0])
        self.eg_min, self.eg_step = float(egs[0]), float(egs[1] - egs[0])
        psi = 10.0 ** lp[:, None]
        tables, sig, dnda, nano = [], [], [], []
        for p in pops:
            for i, a in enumerate(p['a']):
                sig.append(np.pi * a ** 2 * p['q_fuv'][i])
                dnda.append(p['dn_da'][i] * p['da'][i])
                nano.append(p['name'] == 'aC_nano')
                tables.append(self._bin_table(a, p['carbon'], psi, egs[None, :]))
        self.table = np.ascontiguousarray(np.array(tables))
        self.n_bins = self.table.shape[0]
        self.sigma_abs = np.array(sig)
        self.dn_da_da = np.array(dnda)
        self.is_nano = np.array(nano, dtype=np.float64)
        area = {p['name']: np.sum(np.pi * p['a'] ** 2 * p['dn_da'] * p['da']) for p in pops}
        self.sigma_nano = area['aC_nano']
        self.sigma_largeC = area['aC_large']
        self.sigma_Sil = area['aSil']

    @staticmethod
    def _bin_table(a, carbon, psi, eg):
        """Gaussian charge distribution around Z_mean = psi a / 5e-7 (capped at
        Z_max), yield Y(Z) and photoelectron energy E(Z), averaged over Z."""
        y0 = 0.14 if carbon else 0.05
        work = 4.0 + eg if carbon else np.full((1, 1), 8.0)
        zmax = np.maximum((10.0 - work) * a * EV / E_CHARGE ** 2, 0.5)
        z = np.arange(-1, int(np.max(zmax)) + 2)[None, None, :]
        zmean = np.minimum(psi * a / 5e-7, zmax)
        sz = np.maximum(1.0, np.sqrt(zmean + 1))
        p = np.exp(-0.5 * ((z - zmean[:, :, None]) / sz[:, :, None]) ** 2)
        p /= np.sum(p, axis=2, keepdims=True)
        zpos = np.maximum(z, 0)
        yz = y0 * np.maximum(1.0 - zpos / (zmax[:, :, None] + 1), 0.01)
        if a < 1e-7:
            yz = yz * min(3.0, 1e-7 / a)
        ip = work[:, :, None] * EV + zpos * E_CHARGE ** 2 / max(a, 1e-8)
        e_pe = np.maximum(10.0 * EV - ip, 0.0)
        t = np.broadcast_to(np.sum(p * yz * e_pe, axis=2), (psi.shape[0], eg.shape[1])).copy()
        if not carbon:
            t[:, :] = t[:, 0:1]
        return t
