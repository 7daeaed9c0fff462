"""The reference-idiom drop-in (mpf.cuda + mpf.gpu_kernels_3d / mpf.gpu_kernels):
the reference's own time loop (run_randommode_3d.py:267-287 with the checked
APT of tests/verification_3d.py:830-905) written against the kernel objects,
compared with the oracle.  FP64, tolerance: bit-identical phi, mf, nf."""
import math

import numpy as np
import pytest

import cases
from mpf import cuda
from mpf import gpu_kernels as K2
from mpf import gpu_kernels_3d as K3
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SCALAR_ORDER = ("dx", "dt", "T_melt", "Sf", "eps0_sl", "w0_sl", "a0", "delta_a", "mu_a", "p_round",
                "g2_floor", "ksi", "theta_c_rad")


def _loop(c, steps, cooling, dim, switchable=True):
    phi = c["phi"]
    N = phi.shape[0]
    grid = phi.shape[1:]
    tpb = (8, 8, 4) if dim == 3 else (16, 16)
    blocks = tuple(math.ceil(n / t) for n, t in zip(grid, tpb))
    d_phi = cuda.to_device(phi)
    d_new = cuda.to_device(np.zeros_like(phi))
    d_T = cuda.to_device(c["temp"])
    d_mf = cuda.to_device(np.zeros((K3.KMAX,) + grid, np.int32))
    d_nf = cuda.to_device(np.zeros(grid, np.int32))
    d_st = cuda.to_device(np.zeros(2, np.int32))
    w, a, m, n111 = (cuda.to_device(np.ascontiguousarray(t, np.float64)) for t in c["tables"])
    sc = [c["scalars"][k] for k in SCALAR_ORDER]
    for _ in range(steps):
        if dim == 3:
            K3.kernel_update_temp_3d[blocks, tpb](d_T, cooling, *grid)
            K3.kernel_update_nfmf_3d_checked[blocks, tpb](d_phi, d_mf, d_nf, d_st, *grid, N)
            if switchable:
                K3.kernel_update_phasefield_active_3d_switchable[blocks, tpb](
                    d_phi, d_new, d_T, d_mf, d_nf, w, a, m, n111, *grid, N, *sc, 1, 1)
            else:
                K3.kernel_update_phasefield_active_3d[blocks, tpb](
                    d_phi, d_new, d_T, d_mf, d_nf, w, a, m, n111, *grid, N, *sc)
        else:
            K2.kernel_update_temp[blocks, tpb](d_T, cooling, *grid)
            K2.kernel_update_nfmf[blocks, tpb](d_phi, d_mf, d_nf, *grid, N)
            K2.kernel_update_phasefield_active[blocks, tpb](
                d_phi, d_new, d_T, d_mf, d_nf, w, a, m, n111, *grid, N, *sc)
        d_phi, d_new = d_new, d_phi
    cuda.synchronize()
    return d_phi.copy_to_host(), d_T.copy_to_host(), d_st.copy_to_host()


@pytest.mark.parametrize("switchable", [True, False])
def test_reference_loop_3d_bit_exact(switchable):
    c = cases.voronoi_case()
    steps, cool = 5, 0.25
    ref, T_ref, done, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50,
                                 cooling_rate=cool)
    got, T, st_got = _loop(c, steps, cool, 3, switchable)
    print("max |diff| =", np.abs(got - ref).max())
    assert np.array_equal(got, ref)
    assert np.array_equal(T, T_ref)
    assert st_got[0] == 0 and st_got[1] == st[1]


def test_reference_loop_2d_bit_exact():
    c = cases.disk2d_case(n=32, radius=6.0)
    steps, cool = 5, 0.001
    ref, T_ref, done, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50,
                                 cooling_rate=cool)
    got, T, _ = _loop(c, steps, cool, 2)
    print("max |diff| =", np.abs(got - ref).max())
    assert np.array_equal(got, ref)
    assert np.array_equal(T, T_ref)
