"""Loads the native extension (pybind11 module over libgpu_mm_amd.so).

torch is imported first on purpose: torch ships its own HIP runtime, and loading it first makes
the extension bind to that same runtime (one HIP runtime per process). There is no fallback:
if the extension is missing, importing gpu_mm_amd fails.
"""

import importlib
import os

import torch  # noqa: F401  (must precede the extension, see above)

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
_mod = None


def ext():
    global _mod
    if _mod is None:
        try:
            _mod = importlib.import_module("gpu_mm_amd._gpu_mm_amd_py")
        except ImportError as e:
            raise ImportError(
                "gpu_mm_amd: native extension not built (run `python gpu_mm_amd/_build.py` "
                f"or __graft_entry__.build()): {e}") from e
    return _mod


def lib_path():
    return os.path.join(_PKG_DIR, "libgpu_mm_amd.so")
