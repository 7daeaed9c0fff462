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


def _production_loop(c, steps, cooling, dtype):
    """run_randommode_3d.py:244-287 as written: phi in `dtype`, the
    temperature in float64, tables in `dtype`, mf of depth MAX_GRAINS = N,
    the unchecked APT kernel_update_nfmf_3d and the plain update kernel."""
    phi = c["phi"].astype(dtype)
    N = phi.shape[0]
    grid = phi.shape[1:]
    tpb = (8, 8, 4)
    blocks = tuple(math.ceil(n / t) for n, t in zip(grid, tpb))
    d_phi = cuda.to_device(phi)
    d_new = cuda.to_device(phi.copy())
    d_T = cuda.to_device(np.asarray(c["temp"], np.float64))
    d_mf = cuda.to_device(np.zeros((N,) + grid, np.int32))
    d_nf = cuda.to_device(np.zeros(grid, np.int32))
    w, a, m, n111 = (cuda.to_device(np.ascontiguousarray(t, dtype)) for t in c["tables"])
    sc = [dtype(c["scalars"][k]) for k in SCALAR_ORDER]
    for _ in range(steps):
        K3.kernel_update_temp_3d[blocks, tpb](d_T, cooling, *grid)
        K3.kernel_update_nfmf_3d[blocks, tpb](d_phi, d_mf, d_nf, *grid, N)
        K3.kernel_update_phasefield_active_3d[blocks, tpb](
            d_phi, d_new, d_T, d_mf, d_nf, w, a, m, n111, *grid, N, *sc)
        d_phi, d_new = d_new, d_phi
    cuda.synchronize()
    return d_phi.copy_to_host(), d_mf.copy_to_host(), d_nf.copy_to_host(), d_T.copy_to_host()


def test_production_loop_unchecked_apt_3d():
    """The production driver's loop with the unchecked APT on a grid of
    several 8x8x4 launch tiles (40x36x24, 9 grains, seed layer 10, cooling):
    FP64 phi, mf and nf bit-identical to the oracle (mf over the nf active
    entries; the clamp of the checked APT never acts at depth N); FP32 phi
    within 1e-4."""
    dc = cases.directional_case(40, 36, 24, 9, seed_height=10, G=2.0e4)
    p, d = dc["params"], dc["derived"]
    phi = cases.inputs.init_phi_from_grain_map_3d(dc["gmap"], 9, 40, 36, 24, p["dx"], d["delta"], 10)
    c = {"phi": phi.astype(np.float64), "temp": dc["temp"], "tables": dc["tables"], "scalars": dc["scalars"]}
    N = 10
    steps, cool = 8, dc["cooling"] * 100
    ref, T_ref, done, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=N,
                                 cooling_rate=cool)
    # the APT the last step used: that of the state before it
    prev, _, _, _ = O.run(c["phi"], c["temp"], steps - 1, c["tables"], c["scalars"], kmax=N, cooling_rate=cool)
    mf_o, nf_o, _ = O.apt(prev, N, depth=N)
    got, mf, nf, T = _production_loop(c, steps, cool, np.float64)
    print("max |diff| =", np.abs(got - ref).max(), "max nf", nf.max())
    assert np.array_equal(got, ref)
    assert np.array_equal(nf, nf_o)
    mask = np.arange(N)[:, None, None, None] < nf[None]
    assert np.array_equal(np.where(mask, mf, -1), np.where(mask, mf_o, -1))
    assert np.array_equal(T, T_ref)
    got32, _, nf32, _ = _production_loop(c, steps, cool, np.float32)
    err = float(np.abs(got32 - ref).max())
    print("fp32 max |diff| =", err)
    assert err <= 1e-4
