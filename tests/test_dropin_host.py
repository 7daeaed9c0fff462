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


def _fake(shape, dtype):
    """A DeviceArray header without device memory (the checks below must
    reject the call before any pointer is used)."""
    from mpf.dense import DeviceArray
    a = object.__new__(DeviceArray)
    a.shape = tuple(shape)
    a.dtype = np.dtype(dtype)
    a.size = int(np.prod(a.shape))
    a.nbytes = a.size * a.dtype.itemsize
    a._ptr = None
    return a


def test_dense_dropins_reject_mismatched_shapes_before_launch():
    from mpf import dense
    N, g = 5, (6, 4, 3)
    phi, new = _fake((N,) + g, np.float64), _fake((N,) + g, np.float64)
    mf, nf, st = _fake((50,) + g, np.int32), _fake(g, np.int32), _fake((2,), np.int32)
    T = _fake(g, np.float64)
    tabs = [_fake((N, N), np.float64) for _ in range(3)] + [_fake((N, 8, 3), np.float64)]
    sc = {k: 1.0 for k in ("dx", "dt", "T_melt", "Sf", "eps0_sl", "w0_sl", "a0", "delta_a", "mu_a",
                           "p_round", "g2_floor", "ksi", "theta_c_rad")}
    bad_apt = [
        dict(phi=phi, mf=mf, nf=nf, status=st, n_phase=N + 1),                       # more phases than phi
        dict(phi=phi, mf=_fake((50, 6, 4, 2), np.int32), nf=nf, status=st, n_phase=N),  # mf grid
        dict(phi=phi, mf=_fake((50,) + g, np.int64), nf=nf, status=st, n_phase=N),       # mf dtype
        dict(phi=phi, mf=mf, nf=_fake((6, 4), np.int32), status=st, n_phase=N),          # nf grid
        dict(phi=phi, mf=mf, nf=nf, status=_fake((1,), np.int32), n_phase=N),            # status size
        dict(phi=phi, mf=mf, nf=nf, status=None, n_phase=N),                             # checked, no status
    ]
    for kw in bad_apt:
        with pytest.raises(ValueError):
            dense.apt(**kw)
    good = dict(phi=phi, phi_new=new, temp=T, mf=mf, nf=nf, wij=tabs[0], aij=tabs[1], mij=tabs[2],
                n111=tabs[3], scalars=sc, n_phase=N)
    for key, val in (("phi_new", _fake((N, 6, 4, 2), np.float64)), ("temp", _fake((6, 4), np.float64)),
                     ("wij", _fake((N + 1, N + 1), np.float64)), ("mij", _fake((N, N - 1), np.float64)),
                     ("n111", _fake((N, 4, 3), np.float64)), ("nf", _fake((6, 4, 2), np.int32))):
        kw = dict(good)
        kw[key] = val
        with pytest.raises(ValueError):
            dense.update(**kw)
