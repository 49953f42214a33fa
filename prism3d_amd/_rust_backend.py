"""Loader twin of the upstream prism3d/_rust_backend.py:11-33.

``has_rust()`` reports whether the native core can be used (here: the HIP
library loads and a GPU is present); ``get_rust()`` returns the drop-in
module.  Upstream semantics: failures to import give False / ImportError.
"""

_available = None
_module = None


def has_rust():
    """True when the gfx950 core is usable."""
    global _available
    if _available is None:
        try:
            from . import _lib
            _lib.lib()
            import torch
            _available = bool(torch.cuda.is_available())
        except (ImportError, OSError):
            _available = False
    return _available


def get_rust():
    """The drop-in core module; raises ImportError if it is not usable."""
    global _module
    if _module is None:
        if not has_rust():
            raise ImportError("prism3d_amd native core unavailable (library or GPU missing)")
        from . import _prism3d_core as mod
        _module = mod
    return _module
