"""Device-resident whole-grid operations around the per-cell step.

Thin wrappers over the grid kernels of libp3d_hip.so; every argument is a
float64 CUDA tensor (the grid in C order) and every result is allocated on
the device.  The ray set is host data (a few hundred doubles).
"""

import numpy as np
import torch

from . import _lib
from .device import device, stream_handle

AV_PER_NH = 5.34e-22  # prism3d/utils/constants.py:52
G0_TO_FLUX = 1.6e-3   # Habing flux [erg/cm^2/s], prism3d/utils/constants.py:45
LINES = {  # (freq, A, T_ul, g_u, g_l) of prism3d/utils/constants.py:103-109
    "CII_158": (1.9005e12, 2.321e-6, 91.21, 4.0, 2.0),
    "OI_63": (4.7448e12, 8.910e-5, 227.71, 3.0, 5.0),
    "OI_145": (2.0604e12, 1.750e-5, 326.58, 1.0, 3.0),
    "CI_609": (4.9231e11, 7.880e-8, 23.62, 3.0, 1.0),
    "CI_370": (8.0929e11, 2.650e-7, 62.46, 5.0, 3.0),
}


def line_params():
    """The four parameter vectors of the thermal call (solver_3d.py:787-790)."""
    return (np.array(LINES["CII_158"]), np.array(LINES["OI_63"]), np.array(LINES["OI_145"]),
            np.array(LINES["CI_609"] + LINES["CI_370"]))


def healpix_directions(nside=1):
    """Fibonacci-spiral ray directions and solid-angle weights
    (radiative_transfer/fuv_rt.py:19-57)."""
    n = 12 * nside ** 2
    i = np.arange(n, dtype=np.float64)
    ct = 1.0 - 2.0 * (i + 0.5) / n
    st = np.sqrt(1.0 - ct ** 2)
    phi = np.pi * (3.0 - np.sqrt(5.0)) * i
    return np.stack([st * np.cos(phi), st * np.sin(phi), ct], axis=1), np.full(n, 4.0 * np.pi / n)


def directional_G0(G0, source_direction, nside=1, background=0.0):
    """Per-ray external field for a distant source (fuv_rt_3d.py:21-71)."""
    d, w = healpix_directions(nside)
    s = np.asarray(source_direction, dtype=np.float64)
    s = s / np.linalg.norm(s)
    cth = d @ s
    lit = cth > 0
    ws = np.sum(np.where(lit, cth * w, 0.0))
    g = np.where(lit, G0 * cth * (4.0 * np.pi) / ws, 0.0) if ws > 0 else np.full(len(d), G0)
    return g + background


def _t(x):
    if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64):
        raise TypeError("expected a float64 CUDA tensor")
    return x.contiguous()


def _empty(like, k=1):
    return [torch.empty_like(like) for _ in range(k)]


def fuv_columns(density, x_h2, x_co, cell_size, nside=1, G0_external=1.0,
                av_per_nh=AV_PER_NH, workspace=None):
    """G0, A_V, N_H, N_H2, N_CO on the grid (fuv_rt_3d.py:215-294)."""
    density, x_h2, x_co = _t(density), _t(x_h2), _t(x_co)
    if density.ndim != 3 or x_h2.shape != density.shape or x_co.shape != density.shape:
        raise ValueError("density, x_h2 and x_co must share one 3-D shape")
    nx, ny, nz = density.shape
    dirs, _ = healpix_directions(nside)
    nr = len(dirs)
    g0r = np.asarray(G0_external, dtype=np.float64).reshape(-1)
    if g0r.size == 1:
        g0r = np.full(nr, float(g0r[0]))
    elif g0r.size < nr:  # upstream: rays beyond the array reuse element 0
        g0r = np.concatenate([g0r, np.full(nr - g0r.size, g0r[0])])
    dirs = np.ascontiguousarray(dirs.reshape(-1))
    g0r = np.ascontiguousarray(g0r[:nr])
    if workspace is None:
        workspace = torch.empty(18 * density.numel(), dtype=torch.float64, device=density.device)
    outs = _empty(density, 5)
    rc = _lib.lib().p3d_fuv_columns(nx, ny, nz, density.data_ptr(), x_h2.data_ptr(),
                                    x_co.data_ptr(), float(cell_size), nr, _lib.dptr(dirs),
                                    _lib.dptr(g0r), float(av_per_nh),
                                    *[o.data_ptr() for o in outs], workspace.data_ptr(),
                                    stream_handle())
    _lib.check(rc, "p3d_fuv_columns")
    return tuple(outs)


def fuv_point_source(density, star_pos_cm, L_FUV, cell_size, x_h2=None, x_co=None,
                     av_per_nh=AV_PER_NH, g0_to_flux=G0_TO_FLUX, out=None):
    """G0, A_V, N_H, N_H2, N_CO of an embedded point source of FUV luminosity
    L_FUV [erg/s] at star_pos_cm (fuv_rt_3d.py:74-212); x_h2 / x_co default
    to zero abundances.  `out` may hold five preallocated result tensors."""
    density = _t(density)
    if density.ndim != 3:
        raise ValueError("density must be a 3-D grid")
    opt = []
    for x in (x_h2, x_co):
        if x is not None:
            x = _t(x)
            if x.shape != density.shape:
                raise ValueError("x_h2 and x_co must match the density grid")
        opt.append(x)
    pos = np.ascontiguousarray(np.asarray(star_pos_cm, dtype=np.float64).reshape(-1))
    if pos.size != 3:
        raise ValueError("star position must have 3 components")
    nx, ny, nz = density.shape
    outs = list(out) if out is not None else _empty(density, 5)
    rc = _lib.lib().p3d_fuv_point_source(
        nx, ny, nz, density.data_ptr(), *[x.data_ptr() if x is not None else None for x in opt],
        _lib.dptr(pos), float(L_FUV), float(cell_size), float(av_per_nh), float(g0_to_flux),
        *[o.data_ptr() for o in outs], stream_handle())
    _lib.check(rc, "p3d_fuv_point_source")
    return tuple(outs)


def dust_evolve(g0, n_h, f_nano, e_g):
    """One outer dust step of f_nano and E_g, in place (solver_3d.py:465-507)."""
    g0, n_h = _t(g0), _t(n_h)
    for x in (f_nano, e_g):
        if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous()):
            raise TypeError("f_nano and E_g must be contiguous float64 CUDA tensors")
        if x.numel() != g0.numel():
            raise ValueError("f_nano and E_g must match the grid")
    rc = _lib.lib().p3d_dust_evolve(g0.numel(), g0.data_ptr(), n_h.data_ptr(), f_nano.data_ptr(),
                                    e_g.data_ptr(), stream_handle())
    _lib.check(rc, "p3d_dust_evolve")
    return f_nano, e_g


MEAN_WORKSPACE = 1024  # doubles of partial sums p3d_mean needs


def mean(x, out=None, workspace=None):
    """Deterministic device mean of x into out[0]."""
    x = _t(x)
    if out is None:
        out = torch.empty(1, dtype=torch.float64, device=x.device)
    if workspace is None:
        workspace = torch.empty(MEAN_WORKSPACE, dtype=torch.float64, device=x.device)
    _lib.check(_lib.lib().p3d_mean(x.numel(), x.data_ptr(), out.data_ptr(),
                                   workspace.data_ptr(), stream_handle()), "p3d_mean")
    return out


def shield_cr(n_h2_col, n_co_col, n_h_col, t_mean, zeta0):
    """f_shield_H2(N_H2, mean T), f_shield_CO(N_CO, N_H2), zeta_CR(N_H)."""
    a, b, c = _t(n_h2_col), _t(n_co_col), _t(n_h_col)
    outs = _empty(a, 3)
    rc = _lib.lib().p3d_shield_cr(a.numel(), a.data_ptr(), b.data_ptr(), c.data_ptr(),
                                  _t(t_mean).data_ptr(), float(zeta0),
                                  *[o.data_ptr() for o in outs], stream_handle())
    _lib.check(rc, "p3d_shield_cr")
    return tuple(outs)


def dust_state(g0, a_v, n_h, t_gas, f_nano, sigma_nano, sigma_large_c, sigma_sil):
    """T_dust and R_H2 (grains/themis_tables.py:278-354)."""
    ins = [_t(v) for v in (g0, a_v, n_h, t_gas, f_nano)]
    outs = _empty(ins[0], 2)
    rc = _lib.lib().p3d_dust_state(ins[0].numel(), *[v.data_ptr() for v in ins],
                                   float(sigma_nano), float(sigma_large_c), float(sigma_sil),
                                   *[o.data_ptr() for o in outs], stream_handle())
    _lib.check(rc, "p3d_dust_state")
    return tuple(outs)


def euler_rates(n_h, g0, a_v, t_gas, zeta, fsh_h2, fsh_co, r_h2):
    """k_pd_base, k_ci, alpha_cp, k_co_total, r_h2_half_nh, r_cr_base."""
    ins = [_t(v) for v in (n_h, g0, a_v, t_gas, zeta, fsh_h2, fsh_co, r_h2)]
    outs = _empty(ins[0], 6)
    rc = _lib.lib().p3d_euler_rates(ins[0].numel(), *[v.data_ptr() for v in ins],
                                    *[o.data_ptr() for o in outs], stream_handle())
    _lib.check(rc, "p3d_euler_rates")
    return tuple(outs)


def thermal_ctx(n_h, g0, a_v, fsh_h2, zeta, t_dust, gamma_pe, r_h2, x_hi, x_h2, x_e, x_co,
                x_cp, x_o, x_c):
    """The 9 computed context arrays of solver_3d.py:693-731 (t_dust passes through)."""
    ins = [_t(v) for v in (n_h, g0, a_v, fsh_h2, zeta, t_dust, gamma_pe, r_h2, x_hi, x_h2, x_e,
                           x_co, x_cp, x_o, x_c)]
    outs = _empty(ins[0], 9)
    rc = _lib.lib().p3d_thermal_ctx(ins[0].numel(), *[v.data_ptr() for v in ins],
                                    *[o.data_ptr() for o in outs], stream_handle())
    _lib.check(rc, "p3d_thermal_ctx")
    return tuple(outs)


def heating_cooling(t_gas, n_h, g0, a_v, zeta, fsh_h2, x_e, x_hi, x_h2, x_c, x_cp, x_o, x_co,
                    t_dust, gamma_pe, r_h2):
    """Gamma and Lambda at t_gas (solver_3d.py:898-1050)."""
    ins = [_t(v) for v in (t_gas, n_h, g0, a_v, zeta, fsh_h2, x_e, x_hi, x_h2, x_c, x_cp, x_o,
                           x_co, t_dust, gamma_pe, r_h2)]
    outs = _empty(ins[0], 2)
    lp = [np.ascontiguousarray(p) for p in line_params()]
    rc = _lib.lib().p3d_heating_cooling(ins[0].numel(), *[v.data_ptr() for v in ins],
                                        *[_lib.dptr(p) for p in lp],
                                        *[o.data_ptr() for o in outs], stream_handle())
    _lib.check(rc, "p3d_heating_cooling")
    return tuple(outs)


def solve_thermal(gamma_fixed, gg_coeff, t_dust, n_e, n_hi, n_h2, n_co, n_cp, n_o, n_c,
                  t_gas, n_bisect=10, n_newton=3):
    """Thermal root-find on device tensors, t_gas overwritten."""
    ins = [_t(v) for v in (gamma_fixed, gg_coeff, t_dust, n_e, n_hi, n_h2, n_co, n_cp, n_o,
                           n_c)]
    lp = [np.ascontiguousarray(p) for p in line_params()]
    rc = _lib.lib().p3d_solve_thermal_vec(ins[0].numel(), *[v.data_ptr() for v in ins],
                                          *[_lib.dptr(p) for p in lp], int(n_bisect),
                                          int(n_newton), _t(t_gas).data_ptr(), stream_handle())
    _lib.check(rc, "p3d_solve_thermal_vec")
    return t_gas


# --- the tail of an inner iteration (solver_3d.py:262-436) --------------------
CONV_KEYS = ('dT', 'dT_p99', 'dH2', 'dCO', 'dT_max', 'dH2_max', 'n_oscillating', 'delta_max')
DAMP_MODES = {'reference': _lib.P3D_DAMP_REFERENCE, 'uniform': _lib.P3D_DAMP_UNIFORM,
              'oscillating': _lib.P3D_DAMP_OSCILLATING}


def conv_workspace(dev=None):
    """Device workspace of one convergence() call sequence."""
    return torch.empty(_lib.P3D_CONV_WORKSPACE_BYTES // 8, dtype=torch.float64,
                       device=dev if dev is not None else device())


def loop_snapshot(t_gas, y, t_old, y_old):
    """t_old <- t_gas, y_old <- y (the state at the start of an iteration)."""
    t_gas, t_old = _t(t_gas), _t(t_old)
    y, y_old = _t(y), _t(y_old)
    n = t_gas.numel()
    if t_old.numel() != n or y.shape != y_old.shape or y.numel() != y.shape[0] * n:
        raise ValueError("snapshot buffers do not match the state")
    _lib.check(_lib.lib().p3d_loop_snapshot(n, y.shape[0], t_gas.data_ptr(), y.data_ptr(),
                                            t_old.data_ptr(), y_old.data_ptr(), stream_handle()),
               "p3d_loop_snapshot")


def damp_abundances(y, y_old, mode, rows5, x_c_total, w_damp=None):
    """Abundance damping of solver_3d.py:330-354 in place on y (n_species, n):
    mode 'reference', 'uniform' or 'oscillating'; rows5 = rows of H, H2, C+,
    C, CO."""
    y, y_old = _t(y), _t(y_old)
    if y.shape != y_old.shape or y.ndim != 2:
        raise ValueError("y and y_old must share one (n_species, n) shape")
    n = y.shape[1]
    w = None
    if w_damp is not None:
        w = _t(w_damp)
        if w.numel() != n:
            raise ValueError("w_damp must hold one weight per cell")
    r = np.ascontiguousarray(rows5, dtype=np.int32)
    _lib.check(_lib.lib().p3d_damp_abundances(n, DAMP_MODES[mode], y.shape[0], y.data_ptr(),
                                              y_old.data_ptr(),
                                              w.data_ptr() if w is not None else None,
                                              _lib.i32ptr(r), float(x_c_total), stream_handle()),
               "p3d_damp_abundances")


def damp_temperature(t_old, t_gas, w_damp):
    """Adaptive T damping of solver_3d.py:357-375, in place on t_gas and w_damp."""
    t_old, t_gas, w = _t(t_old), _t(t_gas), _t(w_damp)
    n = t_gas.numel()
    if t_old.numel() != n or w.numel() != n:
        raise ValueError("t_old, t_gas and w_damp must have one value per cell")
    _lib.check(_lib.lib().p3d_damp_temperature(n, t_old.data_ptr(), t_gas.data_ptr(),
                                               w.data_ptr(), stream_handle()),
               "p3d_damp_temperature")


def convergence(t_old, t_gas, h2_old, x_h2, co_old, x_co, tol, record, workspace):
    """Convergence figures of solver_3d.py:381-432 into the device record
    (P3D_CONV_FIELDS doubles), enqueued on the current stream."""
    a = [_t(v) for v in (t_old, t_gas, h2_old, x_h2, co_old, x_co)]
    n = a[0].numel()
    if any(v.numel() != n for v in a):
        raise ValueError("the six state arrays must have one value per cell")
    if record.numel() < _lib.P3D_CONV_FIELDS or workspace.numel() * workspace.element_size() \
            < _lib.P3D_CONV_WORKSPACE_BYTES:
        raise ValueError("record or workspace too small")
    _lib.check(_lib.lib().p3d_convergence(n, *[v.data_ptr() for v in a], float(tol),
                                          record.data_ptr(), workspace.data_ptr(),
                                          stream_handle()), "p3d_convergence")


def convergence_record(record):
    """The device record as the reference's convergence-history fields."""
    r = record.cpu().numpy()
    rec = {k: float(r[i]) for i, k in enumerate(CONV_KEYS)}
    rec['n_oscillating'] = int(r[6])
    return rec
