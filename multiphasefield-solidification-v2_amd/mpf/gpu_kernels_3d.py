"""The 3D kernel set under the reference's names (src/gpu_kernels_3d.py:20-21,
:462-838), launched as ``kernel[blocks, threads](*args)`` on mpf.cuda arrays.

    from mpf import cuda
    from mpf.gpu_kernels_3d import (KMAX, kernel_update_nfmf_3d_checked,
                                    kernel_update_phasefield_active_3d,
                                    kernel_update_temp_3d)
"""
from ._kernel_objects import make_kernels

KMAX = 50
LIQ = 0

_k = make_kernels(3, KMAX)
kernel_update_nfmf_3d = _k["kernel_update_nfmf_3d"]
kernel_update_nfmf_3d_checked = _k["kernel_update_nfmf_3d_checked"]
kernel_update_phasefield_active_3d = _k["kernel_update_phasefield_active_3d"]
kernel_update_phasefield_active_3d_switchable = _k["kernel_update_phasefield_active_3d_switchable"]
kernel_update_temp_3d = _k["kernel_update_temp_3d"]

__all__ = ["KMAX", "LIQ", "kernel_update_nfmf_3d", "kernel_update_nfmf_3d_checked",
           "kernel_update_phasefield_active_3d", "kernel_update_phasefield_active_3d_switchable",
           "kernel_update_temp_3d"]
