"""gpu_fem -- MI355X (gfx950) implementation of the guarded-floor GMG-FGMRES
solver path, with the module layout and call signatures of the reference's
``src/gpu_fem`` package.  Device arrays are torch tensors on the ROCm device;
the arithmetic runs in libgfhip.so (HIP kernels + C ABI)."""

from . import _lib  # noqa: F401

__all__ = ["_lib"]
