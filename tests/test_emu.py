"""The solver's per-cell device code compiled for the host (tests/emu) against
the oracle, step by step: FP64 bit-exact on every kernel path (single-phase
stencils, the 4-grain register path, the scratch path), FP32 within 1e-4."""
import numpy as np
import pytest

import cases
from emu import emu
from oracle import oracle as O


def _run(phi, temp, tabs, scal, steps, mode=0, precision="fp64", aniso=True, torque=True, use_class=False,
         kmax=50):
    cur_ref = phi.copy()
    cur = phi.copy()
    used = np.zeros(4, int)
    for _ in range(steps):
        mf, nf, _ = O.apt(cur_ref, kmax)
        ref = O.update(cur_ref, temp, mf, nf, tabs, scal, aniso=aniso, torque=torque)
        out, paths, _ = emu.step(cur if precision == "fp64" else cur_ref, temp, tabs, scal, mode=mode,
                                 precision=precision, aniso=aniso, torque=torque, use_class=use_class, kmax=kmax)
        used += paths
        yield ref, out, used
        cur_ref, cur = ref, out


@pytest.mark.parametrize("seed,n_solid", [(17, 4), (3, 7), (11, 12)])
@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("use_class", [False, True])
def test_fp64_voronoi_paths_bit_exact(seed, n_solid, mode, use_class):
    c = cases.voronoi_case(nx=10, ny=9, nz=12, n_solid=n_solid, seed=seed, qseed=seed + 1)
    for ref, out, used in _run(c["phi"], c["temp"], c["tables"], c["scalars"], 3, mode=mode,
                               use_class=use_class):
        assert np.array_equal(out, ref)
    if mode == 0:
        assert used[0] > 0 and used[1] > 0


def test_fp64_quad_junction_and_many_phase_cells():
    """A four-grain junction meeting the front (liquid + 4 solids, the last
    register slot) and a blob of up to 8 grains (scratch path)."""
    p = cases.ref_params()
    d, _ = cases.tables_for(p, np.tile([0, 0, 0, 1.0], (5, 1)))
    nx, ny, nz = 8, 8, 12
    gm = np.ones((nx, ny), np.int32)
    gm[4:, :4], gm[:4, 4:], gm[4:, 4:] = 2, 3, 4
    phi = cases.inputs.init_phi_from_grain_map_3d(gm, 4, nx, ny, nz, p["dx"], d["delta"], 5).astype(np.float64)
    rng = np.random.default_rng(2)
    q = np.vstack([[0, 0, 0, 1.0], rng.standard_normal((4, 4))])
    q[1:] /= np.linalg.norm(q[1:], axis=1, keepdims=True)
    d, tabs = cases.tables_for(p, q)
    scal = cases.inputs.kernel_scalars(p, d)
    temp = np.full((nx, ny, nz), p["T_melt"] - 2.0)
    for ref, out, used in _run(phi, temp, tabs, scal, 4):
        assert np.array_equal(out, ref)
    _, nf, st = O.apt(ref, 50)
    assert st[1] >= 5

    N = 9
    phi = np.zeros((N, 8, 8, 8))
    phi[0] = 0.3
    phi[1:, 2:6, 2:6, 2:6] = rng.uniform(0.0, 1.0, (N - 1, 4, 4, 4))
    phi[1:] = np.where(phi[1:] < 0.35, 0.0, phi[1:])
    phi /= phi.sum(0)
    q = np.vstack([[0, 0, 0, 1.0], rng.standard_normal((N - 1, 4))])
    q[1:] /= np.linalg.norm(q[1:], axis=1, keepdims=True)
    tabs = cases.tables_for(p, q)[1]
    temp = np.full((8, 8, 8), 1682.0)
    for ref, out, used in _run(phi, temp, tabs, scal, 3):
        assert np.array_equal(out, ref)
    assert used[3] > 0


@pytest.mark.parametrize("kmax", [1, 2, 3])
def test_fp64_kmax_clamp_bit_exact(kmax):
    """KMAX below the active count: cells keep the liquid and the lowest
    grain ids (the two-grain gather and the k-way merge alike)."""
    c = cases.voronoi_case(nx=10, ny=9, nz=12, n_solid=7, seed=3, qseed=4)
    for ref, out, used in _run(c["phi"], c["temp"], c["tables"], c["scalars"], 3, kmax=kmax):
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("aniso,torque", [(True, False), (False, False)])
def test_fp64_switchable_bit_exact(aniso, torque):
    c = cases.voronoi_case()
    for ref, out, _ in _run(c["phi"], c["temp"], c["tables"], c["scalars"], 3, aniso=aniso, torque=torque):
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("use_class", [False, True])
def test_fp64_2d_bit_exact(use_class):
    c = cases.disk2d_case(n=32, radius=6.0)
    for ref, out, _ in _run(c["phi"], c["temp"], c["tables"], c["scalars"], 5, use_class=use_class):
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_paired_facet_scan_identical(precision):
    """The n111 rows that compute_rotated_n111 builds come in opposite pairs;
    scanning four rows then gives the same facet, so the step is identical
    bit for bit to the eight-row scan (FP64 and FP32), and FP64 equals the
    oracle.  A table whose pairs are broken is refused for the four-row scan."""
    c = cases.voronoi_case(nx=10, ny=9, nz=12, n_solid=7, seed=3, qseed=4)
    cur8 = cur4 = c["phi"]
    for _ in range(3):
        out8, _, _ = emu.step(cur8, c["temp"], c["tables"], c["scalars"], precision=precision)
        out4, _, _ = emu.step(cur4, c["temp"], c["tables"], c["scalars"], precision=precision, pairs=True)
        assert np.array_equal(out4, out8)
        if precision == "fp64":
            mf, nf, _ = O.apt(cur8, 50)
            assert np.array_equal(out4, O.update(cur8, c["temp"], mf, nf, c["tables"], c["scalars"]))
        cur8, cur4 = out8, out4
    w, a, m, n111 = c["tables"]
    broken = n111.copy()
    broken[2, 6, 1] += 1e-3
    with pytest.raises(ValueError):
        emu.step(c["phi"], c["temp"], (w, a, m, broken), c["scalars"], pairs=True)


def test_class_form_detection():
    """Tables from build_interaction_matrices have the class form; changing one
    off-diagonal entry, or a non-zero diagonal, removes it (the general lookup
    path then runs, and still matches the oracle)."""
    c = cases.voronoi_case(n_solid=5)
    w, a, m, n111 = c["tables"]
    emu.step(c["phi"], c["temp"], c["tables"], c["scalars"], use_class=True)
    for t, (i, j) in ((0, (2, 3)), (1, (0, 4)), (2, (3, 3))):
        tabs = [x.copy() for x in (w, a, m)] + [n111]
        tabs[t][i, j] *= 1.5 if i != j else 0.0
        if i == j:
            tabs[t][i, j] = 1e-3
        with pytest.raises(ValueError):
            emu.step(c["phi"], c["temp"], tabs, c["scalars"], use_class=True)
        mf, nf, _ = O.apt(c["phi"], 50)
        ref = O.update(c["phi"], c["temp"], mf, nf, tabs, c["scalars"])
        out, _, _ = emu.step(c["phi"], c["temp"], tabs, c["scalars"])
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("use_class", [False, True])
def test_fp32_step_error_small(use_class):
    c = cases.voronoi_case()
    for ref, out, _ in _run(c["phi"], c["temp"], c["tables"], c["scalars"], 3, precision="fp32",
                            use_class=use_class):
        assert np.abs(out - ref).max() <= 1e-4
