"""Device-buffer plumbing for the C ABI: numpy or torch in, device pointers out.

PyTorch-ROCm is used only as the allocator and for host<->device copies.  Every
entry point of the package goes through :class:`Call`, which validates and
stages the arguments the way the upstream native core does (1-D, float64,
C-contiguous; outputs writable), plus length checks that the upstream skips.
"""

import numpy as np
import torch

_DEV = None


def device():
    """The HIP device used by the package (cuda:0); raises without a GPU."""
    global _DEV
    if _DEV is None:
        if not torch.cuda.is_available():
            raise RuntimeError("prism3d_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        _DEV = torch.device("cuda", 0)
    return _DEV


def stream_handle():
    """Raw hipStream_t of torch's current stream on the package device."""
    return torch.cuda.current_stream(device()).cuda_stream


def _check_array(a, name, writable=False):
    if isinstance(a, torch.Tensor):
        if a.dtype != torch.float64:
            raise TypeError(f"{name}: expected float64, got {a.dtype}")
        if not a.is_contiguous():
            raise TypeError(f"{name}: array is not contiguous")
        return a
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name}: expected a numpy array or torch tensor, got {type(a).__name__}")
    if a.dtype != np.float64:
        raise TypeError(f"{name}: expected float64, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise TypeError(f"{name}: array is not contiguous")
    if writable and not a.flags.writeable:
        raise TypeError(f"{name}: array is not writeable")
    return a


class Call:
    """Stages one native call.

    ``inp(a)`` / ``out(a)`` return device pointers; host arrays are copied to
    the device (outputs too, for in-place semantics) and ``finish()`` copies
    outputs back into the host arrays after the launch completed.
    """

    def __init__(self):
        self.dev = device()
        self.keep = []
        self.back = []

    def tensor(self, a, name, writable=False, ndim=None):
        a = _check_array(a, name, writable)
        if ndim is not None and a.ndim != ndim:
            raise TypeError(f"{name}: expected a {ndim}-D array, got {a.ndim}-D")
        if isinstance(a, torch.Tensor):
            if a.device != self.dev:
                raise ValueError(f"{name}: tensor is on {a.device}, expected {self.dev}")
            return a
        t = torch.from_numpy(a).to(self.dev, non_blocking=False)
        if writable:
            self.back.append((a, t))
        self.keep.append(t)
        return t

    def inp(self, a, name, ndim=1):
        return self.tensor(a, name, ndim=ndim).data_ptr()

    def out(self, a, name, ndim=1):
        return self.tensor(a, name, writable=True, ndim=ndim).data_ptr()

    def finish(self):
        torch.cuda.synchronize(self.dev)
        for host, t in self.back:
            host[...] = t.cpu().numpy().reshape(host.shape)


def lengths(named, expect=None):
    """All arrays must have the same length; returns it."""
    n = expect
    for name, a in named:
        m = a.shape[0] if a.ndim else 0
        if n is None:
            n = m
        elif m != n:
            raise ValueError(f"{name}: length {m} does not match the cell count {n}")
    return n
