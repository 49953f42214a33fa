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
