"""MI355X-native backward-LPDM step engine.

Hand-written gfx950 HIP kernels (``glide_amd/csrc``) behind a C ABI
(``include/glide.h``), loaded with ctypes (:mod:`glide_amd._lib`); Python host
code mirrors the reference's plugin interfaces for the per-step path.
"""

__all__ = ["__version__"]
__version__ = "0.1.0"
