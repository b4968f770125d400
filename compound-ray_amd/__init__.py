"""MI355X compound-eye renderer: host-side package.

The renderer itself is the C-ABI shared library ``lib/libEyeRenderer3.so``
(HIP for gfx950, built by ``build.py``).  This package locates and loads it
(``load_library``), exposes ctypes bindings that mirror the reference's
``eyeRendererHelperFunctions`` (``eyerenderer``), and generates the seeded
synthetic scenes and eyes of the benchmark configurations (``scenes``).
"""
from __future__ import annotations

import ctypes
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libEyeRenderer3.so"

_LIB = None


class LibraryMissing(RuntimeError):
    """Raised when the HIP library has not been built."""


def load_library(path: str | pathlib.Path | None = None) -> ctypes.CDLL:
    """Loads libEyeRenderer3.so once and configures every signature.

    Fails loudly when the library is missing: there is no CPU fallback.
    """
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise LibraryMissing(f"{p} not found: build it with compound-ray_amd/build.py (hipcc, gfx950)")
    from . import eyerenderer

    lib = ctypes.CDLL(str(p))
    eyerenderer.configure_functions(lib)
    eyerenderer.configure_extensions(lib)
    if path is None:
        _LIB = lib
    return lib
