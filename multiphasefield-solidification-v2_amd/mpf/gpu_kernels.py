"""The 2D kernel set under the reference's names (src/gpu_kernels.py:19-20,
:338-525; x periodic, y mirror), launched as ``kernel[blocks, threads](*args)``
on mpf.cuda arrays."""
from ._kernel_objects import make_kernels

KMAX = 50
LIQ = 0

_k = make_kernels(2, KMAX)
kernel_update_nfmf = _k["kernel_update_nfmf"]
kernel_update_phasefield_active = _k["kernel_update_phasefield_active"]
kernel_update_temp = _k["kernel_update_temp"]

__all__ = ["KMAX", "LIQ", "kernel_update_nfmf", "kernel_update_phasefield_active", "kernel_update_temp"]
