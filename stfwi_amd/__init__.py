"""stfwi_amd — MI355X-native elastic velocity-strain FWI gradient path.

Hand-written HIP kernels for gfx950 behind a C-ABI (include/stfwi.h,
libstfwi.so), driven from Python through ctypes on PyTorch-ROCm tensors.
``stfwi_amd.fwi`` mirrors the reference package's call signatures
(fwi.forward, fwi.adjoint_jax_kernel, fwi.gradient, ...).
"""
from .plan import Plan, padded_shape  # noqa: F401
from . import config  # noqa: F401

__all__ = ["Plan", "padded_shape", "config"]
