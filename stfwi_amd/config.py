"""Global precision switch, the counterpart of jax_enable_x64 in the
reference's scripts (e.g. Main_strain_fwi.py:48): float32 unless enabled."""
import torch

_X64 = False


def enable_x64(flag: bool = True):
    global _X64
    _X64 = bool(flag)


def default_dtype():
    return torch.float64 if _X64 else torch.float32
