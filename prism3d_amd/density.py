"""Synthetic density fields (input generators of prism3d/density_fields.py).

``fractal_turbulent`` reproduces the upstream log-normal field bit for bit
(same legacy RandomState stream and numpy FFT, density_fields.py:80-141);
``uniform_slab`` is density_fields.py:14-36.
"""

import numpy as np

PC = 3.0856776e18


def uniform_slab(n_cells, box_size, n_H=1e3):
    return np.full((n_cells, n_cells, n_cells), float(n_H), dtype=np.float64), box_size / n_cells


def fractal_turbulent(n_cells, box_size, n_mean=1e3, sigma_ln=1.5, spectral_index=-3.7, seed=42):
    rng = np.random.RandomState(seed)
    cell = box_size / n_cells
    k = np.fft.fftfreq(n_cells, d=cell)
    kx, ky, kz = np.meshgrid(k, k, k, indexing='ij')
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    kk[0, 0, 0] = 1.0
    amp = kk ** (spectral_index / 2.0)
    amp[0, 0, 0] = 0.0
    phases = rng.uniform(0, 2 * np.pi, (n_cells, n_cells, n_cells))
    field = np.real(np.fft.ifftn(amp * np.exp(1j * phases)))
    field -= np.mean(field)
    if np.std(field) > 0:
        field *= sigma_ln / np.std(field)
    return np.maximum(n_mean * np.exp(field - sigma_ln ** 2 / 2.0), 1.0), cell


def lognormal_median(n_cells, box_size, median, sigma_ln, seed=42, clip=None):
    """Log-normal field with the given median (n_mean = median e^{sigma^2/2})."""
    d, cell = fractal_turbulent(n_cells, box_size, n_mean=median * np.exp(sigma_ln ** 2 / 2.0),
                                sigma_ln=sigma_ln, spectral_index=-11.0 / 3.0, seed=seed)
    if clip is not None:
        d = np.clip(d, clip[0], clip[1])
    return d, cell


LSUN_ERG = 3.828e33  # prism3d/utils/constants.py:34


def _gaussian_field(rng, n_cells, cell, sigma_ln, spectral_index=-3.7):
    """Zero-mean Gaussian random field with P(k) ~ k^spectral_index and the
    given standard deviation (the draw order of density_fields.py:224-238)."""
    k = np.fft.fftfreq(n_cells, d=cell)
    kx, ky, kz = np.meshgrid(k, k, k, indexing='ij')
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    kk[0, 0, 0] = 1.0
    amp = kk ** (spectral_index / 2.0)
    amp[0, 0, 0] = 0.0
    g = np.real(np.fft.ifftn(amp * np.exp(1j * rng.uniform(0, 2 * np.pi, (n_cells,) * 3))))
    g -= np.mean(g)
    if np.std(g) > 0:
        g *= sigma_ln / np.std(g)
    return g


def embedded_star_cloud(n_cells, box_size, n_cloud=3000, n_cavity=10, R_cavity_frac=0.15,
                        n_shell_factor=10, shell_width_frac=0.02, sigma_ln=1.2,
                        star_offset=(0.0, 0.0, -0.05), blowout_axis=2, blowout_thin=0.4,
                        seed=42):
    """Turbulent cloud around an embedded O star (density_fields.py:144-296):
    a cavity of density n_cavity whose radius grows by up to 30% toward the
    blowout axis, a Gaussian swept-up shell (thinned on the blowout side), the
    log-normal ambient cloud outside, and four dense pillars pointing at the
    star.  Returns (density, cell_size, star position [cm])."""
    rng = np.random.RandomState(seed)
    cell = box_size / n_cells
    star = 0.5 * box_size + np.asarray(star_offset, dtype=np.float64) * box_size
    c = (np.arange(n_cells) + 0.5) * cell
    X, Y, Z = np.meshgrid(c, c, c, indexing='ij')
    r = np.sqrt((X - star[0]) ** 2 + (Y - star[1]) ** 2 + (Z - star[2]) ** 2)
    R_cav = R_cavity_frac * box_size
    w = shell_width_frac * box_size
    along = (X - star[0], Y - star[1], Z - star[2])[blowout_axis]
    f_blow = np.clip(along / (R_cav + 1e-10), 0, 2)
    R_eff = R_cav * (1.0 + 0.3 * f_blow)
    cloud = n_cloud * np.exp(_gaussian_field(rng, n_cells, cell, sigma_ln) - sigma_ln ** 2 / 2)
    shell = (n_cloud * n_shell_factor * np.exp(-0.5 * ((r - R_eff) / w) ** 2)
             * (1.0 - blowout_thin * np.clip(f_blow, 0, 1)))
    outside = 0.5 * (1.0 + np.tanh((r - R_eff) / (0.5 * w)))
    density = n_cavity + outside * cloud + shell
    for _ in range(4):  # pillars: random direction, length, width (in this draw order)
        theta = np.arccos(2 * rng.rand() - 1)
        phi = 2 * np.pi * rng.rand()
        u = np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
        length = (0.3 + 0.4 * rng.rand()) * R_cav
        width = (0.02 + 0.02 * rng.rand()) * box_size
        base = star + R_cav * u
        spine = (star + (R_cav - length) * u) - base
        s_len = np.linalg.norm(spine)
        s_hat = spine / s_len
        t = np.clip(np.sum(np.stack([X - base[0], Y - base[1], Z - base[2]], axis=-1) * s_hat,
                           axis=-1), 0, s_len)
        near = base + t[..., np.newaxis] * s_hat
        d = np.sqrt((X - near[..., 0]) ** 2 + (Y - near[..., 1]) ** 2 + (Z - near[..., 2]) ** 2)
        pillar = 5 * n_cloud * np.exp(-0.5 * (d / width) ** 2) * np.clip(1.0 - t / s_len, 0, 1)
        density = np.maximum(density, pillar)
    return np.maximum(density, 1.0), cell, star


def blowout(n_cells=64, seed=42, n_mean=3000.0, sigma_ln=1.2, box_pc=3.0):
    """The upstream 'blowout' preset (prism3d/run.py:110-117,158-169,184-185):
    an O5V star (L_FUV = 0.3 x 2e5 L_sun) in a 3 pc cloud, point-source RT,
    a weak background field G0 = 1.  Returns (density, box_size, solver
    keyword arguments)."""
    box = box_pc * PC
    density, _, star_pos = embedded_star_cloud(n_cells, box, n_cloud=n_mean, sigma_ln=sigma_ln,
                                               seed=seed)
    star = {'position': star_pos, 'L_FUV': 0.3 * 2e5 * LSUN_ERG}
    return density, box, {'star': star, 'G0_external': 1.0, 'nside_rt': 1}


def cloud_ensemble(n_clouds=64, n_cells=32, box_size=0.5 * PC, param_seed=2026, field_seeds=None,
                   g0_range=(1e1, 1e5), median_range=(1e2, 1e6), sigma_range=(0.5, 1.5)):
    """Independent log-normal clouds for batched runs: per cloud a field
    G0 drawn log-uniform in g0_range, a median n_H log-uniform in
    median_range and sigma_ln uniform in sigma_range (one
    RandomState(param_seed) draw of the three per cloud, in that order); cloud
    i's field uses field_seeds[i] (default i).  Returns a list of dicts
    (density, box_size, G0, median, sigma_ln, seed)."""
    rng = np.random.RandomState(param_seed)
    seeds = list(range(int(n_clouds))) if field_seeds is None else list(field_seeds)
    if len(seeds) != int(n_clouds):
        raise ValueError("one field seed per cloud")
    lg, lm = np.log10(g0_range), np.log10(median_range)
    out = []
    for i in range(int(n_clouds)):
        g0 = 10.0 ** rng.uniform(*lg)
        med = 10.0 ** rng.uniform(*lm)
        sig = rng.uniform(*sigma_range)
        d, _ = lognormal_median(n_cells, box_size, med, sig, seed=int(seeds[i]))
        out.append(dict(density=d, box_size=float(box_size), G0=float(g0), median=float(med),
                        sigma_ln=float(sig), seed=int(seeds[i])))
    return out
