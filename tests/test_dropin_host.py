"""Host-side checks of the reference-idiom kernel objects (no GPU needed):
names, launch syntax, argument counts and rejected launches."""
import numpy as np
import pytest

from mpf import gpu_kernels as K2
from mpf import gpu_kernels_3d as K3
from mpf._kernel_objects import _check_cover, blocks_for


def test_names_and_constants_match_the_reference_modules():
    assert K3.KMAX == 50 and K3.LIQ == 0 and K2.KMAX == 50 and K2.LIQ == 0
    for n in ("kernel_update_nfmf_3d", "kernel_update_nfmf_3d_checked", "kernel_update_phasefield_active_3d",
              "kernel_update_phasefield_active_3d_switchable", "kernel_update_temp_3d"):
        assert getattr(K3, n).__name__ == n
    for n in ("kernel_update_nfmf", "kernel_update_phasefield_active", "kernel_update_temp"):
        assert getattr(K2, n).__name__ == n


def test_argument_counts_follow_the_reference_signatures():
    # src/gpu_kernels_3d.py:463, :490, :531-537, :684-692, :834; src/gpu_kernels.py:338, :362-368, :521
    expect3 = {"kernel_update_nfmf_3d": 7, "kernel_update_nfmf_3d_checked": 8,
               "kernel_update_phasefield_active_3d": 26, "kernel_update_phasefield_active_3d_switchable": 28,
               "kernel_update_temp_3d": 5}
    for n, k in expect3.items():
        assert getattr(K3, n)._nargs == k
        with pytest.raises(TypeError):
            getattr(K3, n)[(1, 1, 1), (1, 1, 1)](*([0] * (k - 1)))
    expect2 = {"kernel_update_nfmf": 6, "kernel_update_phasefield_active": 25, "kernel_update_temp": 4}
    for n, k in expect2.items():
        assert getattr(K2, n)._nargs == k


def test_host_arrays_and_bad_launches_are_rejected():
    with pytest.raises(TypeError):
        K3.kernel_update_temp_3d[(1, 1, 1), (4, 4, 4)](np.zeros((4, 4, 4)), 0.1, 4, 4, 4)
    with pytest.raises(TypeError):
        K3.kernel_update_temp_3d[4]
    with pytest.raises(ValueError):
        _check_cover("k", (1, 1, 1), (4, 4, 4), (5, 4, 4))
    with pytest.raises(ValueError):
        _check_cover("k", (2, 2), (4, 4, 4), (5, 4, 4))
    _check_cover("k", (2, 1, 1), (4, 4, 4), (5, 4, 4))
    assert blocks_for((65, 64, 1), (8, 8, 4)) == (9, 8, 1)
