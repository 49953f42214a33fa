"""The allocation-free convergence entry point (p3d_convergence_noalloc,
whose workspace is the library's own per-GPU device array) against the
entry point it varies (p3d_convergence with a caller workspace), through the
C ABI: the same record bit for bit on two calls in a row, the second on a
smaller grid, and sizes beyond its limit rejected."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(rng, n):
    old_T = 10 ** rng.uniform(0.5, 3.5, n)
    T = old_T * (1 + rng.normal(0, 0.05, n))
    old_H2 = np.where(rng.uniform(size=n) < 0.3, 1e-8, 10 ** rng.uniform(-7, -0.3, n))
    H2 = old_H2 * np.exp(rng.normal(0, 0.1, n))
    old_CO = np.where(rng.uniform(size=n) < 0.3, 1e-12, 10 ** rng.uniform(-10, -3.9, n))
    CO = old_CO * np.exp(rng.normal(0, 0.2, n))
    return [old_T, T, old_H2, H2, old_CO, CO]


def test_noalloc_convergence_equals_workspace_entry_point(gpu):
    import torch
    from prism3d_amd import _lib
    from prism3d_amd.device import stream_handle
    L = _lib.lib()
    ws = torch.empty(_lib.P3D_CONV_WORKSPACE_BYTES // 8, dtype=torch.float64, device=gpu)
    for n, seed in ((300001, 1), (1000, 2)):
        print(f"case: noalloc convergence, n={n}", flush=True)
        d = [torch.tensor(a, device=gpu) for a in _state(np.random.RandomState(seed), n)]
        ptrs = [t.data_ptr() for t in d]
        rec_ws = torch.full((_lib.P3D_CONV_FIELDS,), -1.0, dtype=torch.float64, device=gpu)
        rec_na = torch.full((_lib.P3D_CONV_FIELDS,), -2.0, dtype=torch.float64, device=gpu)
        _lib.check(L.p3d_convergence(n, *ptrs, 0.05, rec_ws.data_ptr(), ws.data_ptr(),
                                     stream_handle()), "p3d_convergence")
        _lib.check(L.p3d_convergence_noalloc(n, *ptrs, 0.05, rec_na.data_ptr(),
                                             stream_handle()), "p3d_convergence_noalloc")
        torch.cuda.synchronize()
        a, b = rec_ws.cpu().numpy(), rec_na.cpu().numpy()
        assert np.array_equal(a.view(np.int64), b.view(np.int64)), (a, b)
        assert b[8] == np.sum(d[0].cpu().numpy() > 10)
    rec = torch.zeros(_lib.P3D_CONV_FIELDS, dtype=torch.float64, device=gpu)
    rc = L.p3d_convergence_noalloc((1 << 40) + 1, *ptrs, 0.05, rec.data_ptr(), stream_handle())
    assert rc == -1  # P3D_EINVAL: beyond the 2^40-cell limit
