"""Per-cell chemistry on the GPU: the ChemistrySolver call surface.

``ChemistrySolver`` mirrors prism3d/chemistry/solver.py:22-388: the same
tracked species, ``solve_steady_state`` signature, return value and
post-solve corrections.  Every solve runs the gfx950 BDF kernel
(p3d_chem_solve); the batched entry points take (n_species, n_cells)
structure-of-arrays device tensors and solve all cells in one launch.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from .device import device, stream_handle
from .network import DeviceNetwork, Network, default_device_network

X_FLOOR = 1e-30
CII_158 = (1.9005e12, 2.321e-6, 91.21, 4.0, 2.0)  # freq, A, T_ul, g_u, g_l (constants.py:103)
ELEMENTS = ('H', 'He', 'C', 'O', 'S', 'Si', 'Fe')


def element_counts(species):
    """(len(ELEMENTS), len(species)) nuclei per species, from the formulas."""
    import re
    C = np.zeros((len(ELEMENTS), len(species)), dtype=np.int64)
    for s, name in enumerate(species):
        for el, cnt in re.findall(r'([A-Z][a-z]?)(\d*)', name):
            C[ELEMENTS.index(el), s] += int(cnt) if cnt else 1
    return C


def steady_constraints(network):
    """p3d_steady_cons of a network: element nuclei and the closure cations."""
    from .network import CATIONS
    if network.n_species > 64:
        raise ValueError("at most 64 species")
    C = element_counts(network.species)
    k = _lib.SteadyCons()
    k.n_elem = len(ELEMENTS)
    for e in range(len(ELEMENTS)):
        for s in range(network.n_species):
            k.nuclei[e][s] = int(C[e, s])
    for s, name in enumerate(network.species):
        k.cation[s] = 1 if name in CATIONS else 0
    return k
GAS_C = 2.69e-4 * 0.4  # gas-phase C/H (prism3d/utils/constants.py:90-99)
GAS_O = 4.90e-4 * 0.6
GAS_S = 1.32e-5 * 0.4
GAS_SI = 3.24e-5 * 0.03
GAS_FE = 3.16e-5 * 0.006
GAS_HE = 8.51e-2


def solve_opts(t_max=None, t_start=1e4, n_checkpoints=20, early_stop=True, rtol=1e-4,
               atol=1e-22, stop_rtol=1e-3, post_mode=0, max_steps=1000000, x_c_total=GAS_C,
               x_o_total=GAS_O, variant=0, waves_per_cu=0, rescue_frac=0.0, rate_fp32=False,
               fast=False):
    """p3d_solve_opts; the defaults are the upstream refine-path settings
    (solver.py:150-223).  variant: 0 auto, 1-3 one cell per 64-lane
    workgroup (LU in LDS / registers / unrolled LDS), 4 persistent
    one-cell-per-lane kernel, 5 the FP64-MFMA panel LU, 6 the persistent
    one-cell-per-wavefront kernel (the bundled network's automatic form), 7
    the static sparse LU by levels on two wavefronts per cell (the automatic
    form from 64 species); waves_per_cu sizes variant 4's resident grid and
    rescue_frac sets when its long-tail cells move to the workgroup kernel;
    rate_fp32 evaluates the rate coefficients' exp/log in FP32 (the solve
    itself stays FP64).  max_steps caps the BDF steps of one span (a span
    that reaches it fails and leaves the state unchanged, as a failed
    upstream span does); the upstream integrator has no cap, and 1e6 is far
    above the ~1e5 steps the slowest bundled-network cells take at rtol 1e-9,
    atol 1e-24.  fast: the upstream fast mode (ten explicit steps,
    chemistry/solver.py:164-183) instead of the BDF."""
    o = _lib.SolveOpts()
    o.t_max = -1.0 if t_max is None else float(t_max)
    o.t_start = float(t_start)
    o.n_checkpoints = int(n_checkpoints)
    o.early_stop = int(bool(early_stop))
    o.rtol, o.atol, o.stop_rtol = float(rtol), float(atol), float(stop_rtol)
    o.post_mode = int(post_mode)
    o.max_steps = int(max_steps)
    o.x_c_total, o.x_o_total = float(x_c_total), float(x_o_total)
    o.variant = int(variant)
    o.waves_per_cu = int(waves_per_cu)
    o.rescue_frac = float(rescue_frac)
    o.rate_fp32 = int(bool(rate_fp32))
    o.fast = int(bool(fast))
    return o


def _cell_arrays(n, *vals):
    dev = device()
    out = []
    for v in vals:
        t = torch.as_tensor(v, dtype=torch.float64, device=dev).reshape(-1)
        if t.numel() == 1 and n > 1:
            t = t.expand(n).contiguous()
        if t.numel() != n:
            raise ValueError(f"parameter array of length {t.numel()} for {n} cells")
        out.append(t.contiguous())
    return out


class ChemistrySolver:
    """GPU twin of prism3d.chemistry.solver.ChemistrySolver."""

    def __init__(self, network=None):
        if network is None:
            self.dnet = default_device_network()
        elif isinstance(network, DeviceNetwork):
            self.dnet = network
        elif isinstance(network, Network):
            self.dnet = DeviceNetwork(network)
        else:
            raise TypeError("network must be a prism3d_amd Network or DeviceNetwork")
        self.network = self.dnet.net
        self.tracked_species = list(self.network.species)
        self.n_tracked = len(self.tracked_species)
        self.idx = {s: i for i, s in enumerate(self.tracked_species)}

    # --- host conversions (solver.py:62-67) --------------------------------
    def _abundances_to_dict(self, x_arr):
        return {self.tracked_species[i]: max(float(x_arr[i]), X_FLOOR)
                for i in range(self.n_tracked)}

    def _dict_to_array(self, x_dict):
        return np.array([x_dict.get(s, X_FLOOR) for s in self.tracked_species], dtype=np.float64)

    def _default_initial(self, G0, A_V, gas=None):
        """tanh initial abundances (solver.py:300-343)."""
        g = gas or {'C': GAS_C, 'O': GAS_O, 'He': GAS_HE, 'S': GAS_S, 'Si': GAS_SI,
                    'Fe': GAS_FE}
        x = {}
        f_H2 = 0.5 * (1.0 + np.tanh((A_V - 1.5) / 0.5))
        f_CO = 0.5 * (1.0 + np.tanh((A_V - 3.0) / 0.8))
        x['H'] = max(1.0 - f_H2, 0.003)
        x['H2'] = f_H2 * 0.5
        x['C+'] = g['C'] * (1.0 - f_CO)
        x['CO'] = g['C'] * f_CO * 0.95
        x['C'] = g['C'] * f_CO * 0.05
        x['O'] = max(g['O'] - x['CO'], 1e-10)
        for ion, neu, tot in (('S+', 'S', g.get('S', 0)), ('Si+', 'Si', g.get('Si', 0)),
                              ('Fe+', 'Fe', g.get('Fe', 0))):
            f = max(0.01, 1.0 - A_V / 3.0) if A_V < 3 else 0.01
            x[ion] = tot * f
            x[neu] = tot * (1 - f)
        x['H+'] = 1e-6 if A_V < 2 else 1e-7
        x['e-'] = max(x.get('C+', 0) + x.get('S+', 0) + x.get('Si+', 0) + x.get('Fe+', 0)
                      + x.get('H+', 0), 1e-7)
        x['H-'] = 1e-14
        x['H2+'] = 1e-13
        x['H3+'] = 5e-10 if A_V > 2 else 1e-12
        x['OH'] = 1e-7 if A_V > 2 else 1e-9
        x['OH+'] = 1e-13
        x['H2O'] = 1e-7 if A_V > 3 else 1e-9
        x['H2O+'] = 1e-14
        x['H3O+'] = 1e-13
        x['O2'] = 1e-8 if A_V > 3 else 1e-12
        x['CH'] = 1e-10
        x['CH+'] = 1e-13
        x['CH2+'] = 1e-14
        x['CO+'] = 1e-15
        x['HCO+'] = 1e-9 if A_V > 2 else 1e-12
        x['O+'] = 1e-13
        x['He'] = g.get('He', GAS_HE)
        x['He+'] = 1e-13
        return x

    # --- single-cell API (solver.py:145-225) --------------------------------
    def solve_steady_state(self, n_H, T, G0, A_V, zeta_CR, x_init=None, f_shield_H2=1.0,
                           f_shield_CO=1.0, t_max=None, rtol=1e-3, atol=1e-20, fast=False):
        """Steady-state abundances of one cell -> (dict, converged).

        ``rtol`` is the checkpoint stop criterion and ``atol`` is unused, as
        upstream; the BDF itself runs at rtol 1e-4, atol 1e-22.  ``fast``
        runs the upstream fast mode (ten explicit steps, then the
        corrections; converged is False), as chemistry/solver.py:164-183."""
        if x_init is None:
            x_init = self._default_initial(G0, A_V)
        y = torch.tensor(self._dict_to_array(x_init), dtype=torch.float64,
                         device=device()).reshape(-1, 1)
        conv = self.solve_steady_batch(y, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO,
                                       opts=solve_opts(t_max=t_max, stop_rtol=rtol,
                                                       fast=fast))[0]
        return self._abundances_to_dict(y[:, 0].cpu().numpy()), bool(conv[0].item())

    # --- batched SoA API ----------------------------------------------------
    def solve_steady_batch(self, y, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0, f_shield_CO=1.0,
                           opts=None, sync=True):
        """Solve every cell of y (n_species, n_cells), float64 device tensor,
        in place.  Returns (converged int32[n], stats int32[n, 4]) with stats
        columns accepted steps, LU factorisations, RHS and Jacobian
        evaluations."""
        self._check_state(y)
        n = y.shape[1]
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        # every cell's flag and counters are written by the solve (no fill, so
        # the call can be captured in a graph)
        conv = torch.empty(n, dtype=torch.int32, device=y.device)
        stats = torch.empty((n, 4), dtype=torch.int32, device=y.device)
        o = opts if opts is not None else solve_opts()
        rc = _lib.lib().p3d_chem_solve(self.dnet.handle, n, y.data_ptr(),
                                       *[t.data_ptr() for t in p], o, conv.data_ptr(),
                                       stats.data_ptr(), stream_handle())
        _lib.check(rc, "p3d_chem_solve")
        if sync:
            torch.cuda.synchronize(y.device)
        return conv, stats

    def post_batch(self, y, n_H, T, zeta_CR, mode=0, x_c_total=GAS_C, x_o_total=GAS_O):
        """H+/e- correction and conservation on every cell of y, in place
        (solver.py:227-388); mode 1 projects exactly onto the totals."""
        self._check_state(y)
        n = y.shape[1]
        p = _cell_arrays(n, n_H, T, zeta_CR)
        rc = _lib.lib().p3d_chem_post(self.dnet.handle, n, y.data_ptr(),
                                      *[t.data_ptr() for t in p], int(mode), float(x_c_total),
                                      float(x_o_total), stream_handle())
        _lib.check(rc, "p3d_chem_post")
        torch.cuda.synchronize(y.device)
        return y

    # --- charge-neutral steady state and its sensitivities -------------------
    def steady_constraints(self):
        """p3d_steady_cons of this network: nuclei of each element per species
        (from the species formulas) and the cations of the electron closure."""
        if getattr(self, "_cons", None) is None:
            self._cons = steady_constraints(self.network)
        return self._cons

    def polish_steady_batch(self, y, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0, f_shield_CO=1.0,
                            max_iter=600, sync=True, y_totals=None):
        """Polish every cell of y in place onto its charge-neutral steady
        state (p3d_chem_polish), with the element totals of y_totals (same
        shape, e.g. the initial abundances) or of y.  Returns the iterations
        per cell (-1: not converged in max_iter)."""
        self._check_state(y)
        if y_totals is not None:
            self._check_state(y_totals)
            if y_totals.shape != y.shape:
                raise ValueError("y_totals must have the shape of y")
        n = y.shape[1]
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        iters = torch.empty(n, dtype=torch.int32, device=y.device)
        cons = self.steady_constraints()
        rc = _lib.lib().p3d_chem_polish(self.dnet.handle, n, y.data_ptr(),
                                        *[t.data_ptr() for t in p], ctypes.byref(cons),
                                        None if y_totals is None else y_totals.data_ptr(),
                                        int(max_iter), iters.data_ptr(), stream_handle())
        _lib.check(rc, "p3d_chem_polish")
        if sync:
            torch.cuda.synchronize(y.device)
        return iters

    def steady_adjoint_batch(self, y, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0,
                             f_shield_CO=1.0, w_cii=1.0, w_co=1.0, line=CII_158):
        """Per-cell objective {Lambda_CII, n_H x_CO} (n, 2) and gradient (n,
        n_reactions + 3) of w_cii Lambda_CII + w_co n_H x_CO at the
        charge-neutral steady state y, with respect to log10 k_r, ln G0,
        ln zeta, ln n_H (p3d_chem_adjoint)."""
        self._check_state(y)
        n = y.shape[1]
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        obj = torch.empty((n, 2), dtype=torch.float64, device=y.device)
        grad = torch.empty((n, self.network.n_reactions + 3), dtype=torch.float64,
                           device=y.device)
        cons = self.steady_constraints()
        lp = np.ascontiguousarray(line, dtype=np.float64)
        rc = _lib.lib().p3d_chem_adjoint(self.dnet.handle, n, y.contiguous().data_ptr(),
                                         *[t.data_ptr() for t in p], ctypes.byref(cons),
                                         _lib.dptr(lp), float(w_cii), float(w_co),
                                         obj.data_ptr(), grad.data_ptr(), stream_handle())
        _lib.check(rc, "p3d_chem_adjoint")
        return obj, grad

    # --- discrete adjoint of the implicit steps ------------------------------
    def solve_tape(self, y, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0, f_shield_CO=1.0,
                   opts=None, cap=4096, sync=True):
        """The integration of solve_steady_batch (one span, no post-solve
        corrections: opts must have n_checkpoints 2, early_stop False,
        post_mode -1) with its step sequence recorded for tape_adjoint.
        y (n_species, n) is updated in place.  Returns (stats int32[n, 4],
        tape) where tape is a dict holding the device buffer and its cap."""
        self._check_state(y)
        n = y.shape[1]
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        nbytes = ctypes.c_int64(0)
        _lib.check(_lib.lib().p3d_chem_tape_bytes(n, int(cap), ctypes.byref(nbytes)),
                   "p3d_chem_tape_bytes")
        buf = torch.empty(max(1, (nbytes.value + 7) // 8), dtype=torch.float64, device=y.device)
        stats = torch.empty((n, 4), dtype=torch.int32, device=y.device)
        o = opts if opts is not None else solve_opts(n_checkpoints=2, early_stop=False,
                                                     post_mode=-1)
        rc = _lib.lib().p3d_chem_solve_tape(self.dnet.handle, n, y.data_ptr(),
                                            *[t.data_ptr() for t in p], o, buf.data_ptr(),
                                            nbytes.value, int(cap), stats.data_ptr(),
                                            stream_handle())
        _lib.check(rc, "p3d_chem_solve_tape")
        if sync:
            torch.cuda.synchronize(y.device)
        return stats, {"buffer": buf, "bytes": nbytes.value, "cap": int(cap), "opts": o, "n": n}

    @staticmethod
    def tape_steps(tape, cells):
        """Host copy of the recorded step sequence of the given cells: a list
        of (steps (n_steps, 3) rows {c, F, k}, h0, n_steps); the layout of
        p3d_chem_tape_bytes: step records n x cap x 3, differences every 16
        steps n x (cap / 16 + 1) x 8 x 32, initial step sizes n, step counts n
        (int32)."""
        n, cap = tape["n"], tape["cap"]
        buf = tape["buffer"]
        rec_d = n * cap * 3
        ck_d = n * (cap // 16 + 1) * 8 * 32
        h0 = buf[rec_d + ck_d:rec_d + ck_d + n].cpu().numpy()
        ns = buf[rec_d + ck_d + n:rec_d + ck_d + n + (n + 1) // 2].cpu().numpy()
        nsteps = ns.view(np.int32)[:n]
        out = []
        for c in cells:
            k = int(nsteps[c])
            rec = buf[c * cap * 3:c * cap * 3 + 3 * max(k, 0)].cpu().numpy().reshape(-1, 3)
            out.append((rec, float(h0[c]), k))
        return out

    def tape_adjoint(self, y0, y_end, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0,
                     f_shield_CO=1.0, tape=None, w_cii=1.0, w_co=1.0, line=CII_158,
                     sync=True):
        """Per-cell objective {Lambda_CII, n_H x_CO} (n, 2), gradient (n,
        n_reactions + 3) of w_cii Lambda_CII + w_co n_H x_CO at y_end with
        respect to log10 k_r, ln G0, ln zeta, ln n_H through the taped
        integration from y0 (p3d_chem_tape_adjoint), and status int32[n]
        (0 valid)."""
        self._check_state(y_end)
        self._check_state(y0)
        n = y_end.shape[1]
        if tape is None or tape["n"] != n:
            raise ValueError("tape_adjoint needs the tape of solve_tape on the same cells")
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        obj = torch.empty((n, 2), dtype=torch.float64, device=y_end.device)
        grad = torch.empty((n, self.network.n_reactions + 3), dtype=torch.float64,
                           device=y_end.device)
        status = torch.empty(n, dtype=torch.int32, device=y_end.device)
        lp = np.ascontiguousarray(line, dtype=np.float64)
        rc = _lib.lib().p3d_chem_tape_adjoint(
            self.dnet.handle, n, y0.contiguous().data_ptr(), y_end.contiguous().data_ptr(),
            *[t.data_ptr() for t in p], tape["opts"], tape["buffer"].data_ptr(), tape["bytes"],
            tape["cap"], _lib.dptr(lp), float(w_cii), float(w_co), obj.data_ptr(),
            grad.data_ptr(), status.data_ptr(), stream_handle())
        _lib.check(rc, "p3d_chem_tape_adjoint")
        if sync:
            torch.cuda.synchronize(y_end.device)
        return obj, grad, status

    def rhs_batch(self, y, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0, f_shield_CO=1.0):
        self._check_state(y)
        n = y.shape[1]
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        out = torch.empty_like(y)
        rc = _lib.lib().p3d_chem_rhs(self.dnet.handle, n, y.data_ptr(), *[t.data_ptr() for t in p],
                                     out.data_ptr(), stream_handle())
        _lib.check(rc, "p3d_chem_rhs")
        torch.cuda.synchronize(y.device)
        return out

    def jacobian_batch(self, y, n_H, T, G0, A_V, zeta_CR, f_shield_H2=1.0, f_shield_CO=1.0):
        self._check_state(y)
        n = y.shape[1]
        ns = self.n_tracked
        p = _cell_arrays(n, n_H, T, G0, A_V, zeta_CR, f_shield_H2, f_shield_CO)
        out = torch.empty((n, ns, ns), dtype=torch.float64, device=y.device)
        rc = _lib.lib().p3d_chem_jacobian(self.dnet.handle, n, y.data_ptr(),
                                          *[t.data_ptr() for t in p], out.data_ptr(),
                                          stream_handle())
        _lib.check(rc, "p3d_chem_jacobian")
        torch.cuda.synchronize(y.device)
        return out

    def _check_state(self, y):
        if not isinstance(y, torch.Tensor) or y.dtype != torch.float64 or not y.is_cuda:
            raise TypeError("y must be a float64 GPU tensor of shape (n_species, n_cells)")
        if y.ndim != 2 or y.shape[0] != self.n_tracked:
            raise ValueError(f"y must have shape ({self.n_tracked}, n_cells), got {tuple(y.shape)}")
        if not y.is_contiguous():
            raise TypeError("y must be contiguous")
