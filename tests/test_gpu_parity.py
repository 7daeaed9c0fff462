"""GPU parity of the HIP solver against the CPU oracle.

FP64: the solver and the oracle (include/mpf_math.h transcendentals, no
contraction) must agree bit for bit on phi, and the APT sets mf/nf must be
identical (tolerance written below: exact).  FP32: max |phi - phi_oracle_fp64|
<= 1e-4 over short runs.
"""
import numpy as np
import pytest

import cases
from mpf import Solver
from mpf import dense as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10   # the FP64 phi bound; the checks below also demand bit-equality
FP32_TOL = 1e-4


def _solver_for(c, precision, **kw):
    phi = c["phi"]
    s = Solver(phi.shape[1:], phi.shape[0], c["scalars"], precision=precision, **kw)
    s.set_tables(*c["tables"])
    s.upload_phi(phi)
    s.upload_temp(c["temp"])
    return s


def _assert_apt_equal(solver, phi_ref, kmax=50):
    mf, nf, st = solver.apt(depth=kmax)
    mf_o, nf_o, st_o = O.apt(phi_ref, kmax)
    assert np.array_equal(nf, nf_o)
    depth_mask = np.arange(kmax)[:, None, None, None] < nf[None] if nf.ndim == 3 else \
        np.arange(kmax)[:, None, None] < nf[None]
    assert np.array_equal(np.where(depth_mask, mf, -1), np.where(depth_mask, mf_o, -1))
    assert np.array_equal(st, st_o)


@pytest.mark.parametrize("aniso,torque", [(True, True), (True, False), (False, False)])
def test_fp64_voronoi_bit_exact(aniso, torque):
    c = cases.voronoi_case()
    steps = 12
    ref, T_ref, done, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50,
                                 aniso=aniso, torque=torque, temp_f32=True)
    with _solver_for(c, "fp64", temp_f32=True, anisotropy=aniso, torque=torque) as s:
        assert s.step(steps) == steps
        got = s.phi()
        print("max |diff| =", np.abs(got - ref).max())
        assert np.abs(got - ref).max() <= FP64_TOL
        assert np.array_equal(got, ref), "FP64 solver must be bit-identical to the oracle"
        _assert_apt_equal(s, ref)
        assert s.stats()["max_active"] == st[1]


def test_fp64_directional_bit_exact():
    """Seed layer, linear temperature, uniform cooling (run_randommode_3d.py)."""
    dc = cases.directional_case(24, 20, 32, 6, seed_height=8, G=2.0e4)
    gm, pi_, sol, liq = cases.inputs.columns_from_grain_map_3d(dc["gmap"], 32, dc["params"]["dx"],
                                                               dc["derived"]["delta"], 8)
    phi0 = cases.inputs.init_phi_from_grain_map_3d(dc["gmap"], 6, 24, 20, 32, dc["params"]["dx"],
                                                   dc["derived"]["delta"], 8).astype(np.float64)
    steps = 40
    cool = dc["cooling"] * 50
    ref, T_ref, done, st = O.run(phi0, dc["temp"], steps, dc["tables"], dc["scalars"], kmax=50,
                                 cooling_rate=cool, temp_f32=False)
    s = Solver((24, 20, 32), 7, dc["scalars"], precision="fp64", cooling_rate=cool)
    s.set_tables(*dc["tables"])
    s.init_columns(gm, pi_, sol, liq)
    assert np.array_equal(s.phi(), phi0)     # columnar init == dense init
    s.upload_temp(dc["temp"])
    assert s.step(steps) == steps
    got = s.phi()
    assert np.array_equal(got, ref)
    np.testing.assert_array_equal(s.temp(), T_ref[0, 0, :])
    _assert_apt_equal(s, ref)
    s.close()


def test_fp64_2d_disk_bit_exact():
    """Config-1 style 2D half disc (x periodic, y mirror), 200 steps."""
    c = cases.disk2d_case()
    steps = 200
    ref, _, done, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50)
    with _solver_for(c, "fp64") as s:
        assert s.step(steps) == steps
        got = s.phi()
        assert np.array_equal(got, ref)
        _assert_apt_equal(s, ref)


def test_fp64_2d_disk_linear_cooling_bit_exact():
    """The 2D single-grain case with linear cooling (T -= rate every step),
    200 steps, FP64 bit-exact."""
    c = cases.disk2d_case()
    steps, cool = 200, 5.0e-4
    ref, T_ref, done, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50,
                                 cooling_rate=cool)
    with _solver_for(c, "fp64", temp_dense=True, cooling_rate=cool) as s:
        assert s.step(steps) == steps
        got = s.phi()
        assert np.array_equal(got, ref)
        assert np.array_equal(s.temp(), T_ref)
        _assert_apt_equal(s, ref)
    # the cooling changed the run: more solid than without it
    still, _, _, _ = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50)
    assert ref[1].sum() > still[1].sum()


def test_fp32_within_tolerance():
    c = cases.voronoi_case()
    steps = 12
    ref, _, _, _ = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50, temp_f32=True)
    with _solver_for(c, "fp32", temp_f32=True) as s:
        assert s.step(steps) == steps
        err = np.abs(s.phi() - ref).max()
        print("fp32 max |diff| =", err)
        assert err <= FP32_TOL


def test_dense_dropin_kernels_bit_exact():
    """The three drop-in kernels on the reference's dense layout, FP64."""
    c = cases.voronoi_case()
    phi = c["phi"]
    N = phi.shape[0]
    grid = phi.shape[1:]
    temp = c["temp"].copy()
    cool = 0.125
    T_o = temp.copy()
    O.lib().orc_temp(T_o.reshape(-1), T_o.size, cool, 0)
    mf_o, nf_o, st_o = O.apt(phi, 50)
    new_o = O.update(phi, T_o, mf_o, nf_o, c["tables"], c["scalars"])

    d_phi = D.to_device(phi)
    d_new = D.to_device(np.zeros_like(phi))
    d_T = D.to_device(temp)
    d_mf = D.to_device(np.zeros((50,) + grid, np.int32))
    d_nf = D.to_device(np.zeros(grid, np.int32))
    d_st = D.to_device(np.zeros(2, np.int32))
    tabs = [D.to_device(np.ascontiguousarray(t, np.float64)) for t in c["tables"]]
    D.temp(d_T, cool)
    D.apt(d_phi, d_mf, d_nf, d_st, N, checked=True, kmax=50)
    D.update(d_phi, d_new, d_T, d_mf, d_nf, *tabs, c["scalars"], N, kmax=50)
    np.testing.assert_array_equal(d_T.copy_to_host(), T_o)
    assert np.array_equal(d_nf.copy_to_host(), nf_o)
    assert np.array_equal(d_st.copy_to_host(), st_o)
    assert np.array_equal(d_new.copy_to_host(), new_o)


def test_kmax_overflow_probe():
    """tests/verification_3d.py:1459-1480: 52 phases at 1/52 in one cell of a
    4^3 grid; the checked APT reports status [1, 52] (report :156)."""
    N = 52
    phi = np.zeros((N, 4, 4, 4), np.float64)
    phi[0] = 1.0
    phi[:, 1, 1, 1] = 1.0 / N
    d_phi = D.to_device(phi)
    d_mf = D.to_device(np.zeros((N, 4, 4, 4), np.int32))
    d_nf = D.to_device(np.zeros((4, 4, 4), np.int32))
    d_st = D.to_device(np.zeros(2, np.int32))
    D.apt(d_phi, d_mf, d_nf, d_st, N, checked=True, kmax=50)
    st = d_st.copy_to_host()
    assert list(st) == [1, 52]
    mf_o, nf_o, st_o = O.apt(phi, 50)
    assert np.array_equal(d_nf.copy_to_host(), nf_o)
    # the resident solver halts before updating (KMaxOverflowError semantics)
    c = cases.voronoi_case()
    p = dict(c["scalars"])
    tabs = cases.tables_for(c["params"], np.tile([0, 0, 0, 1.0], (N, 1)))[1]
    s = Solver((4, 4, 4), N, p, precision="fp64", halt_on_kmax=True, slot_capacity=N - 1)
    s.set_tables(*tabs)
    s.upload_phi(phi)
    s.upload_temp(np.full(4, 1680.0))
    assert s.step(3) == 0
    st = s.stats()
    assert st["kmax_overflow"] == 1 and st["max_active"] == 52 and st["halt_step"] == 0
    assert np.array_equal(s.phi(), phi)
    s.close()


@pytest.mark.parametrize("N", [9, 13])
def test_general_path_many_phases(N):
    """Cells with more active solids than the register fast path holds take
    the general path and still match the oracle bit for bit: up to 8 solids
    in unrolled register arrays, one lane per solid (N = 9: hundreds of cells
    hold the liquid and all 8 solids active at every step, so the register
    path runs at its full width), more in the scratch arrays (N = 13)."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 8, 8, 8
    phi = np.zeros((N, nx, ny, nz))
    phi[0] = 0.3
    blob = rng.uniform(0.0, 1.0, (N - 1, 4, 4, 4))
    phi[1:, 2:6, 2:6, 2:6] = blob
    phi[1:] = np.where(phi[1:] < 0.35, 0.0, phi[1:])
    s_ = phi.sum(0)
    phi /= s_
    c = cases.voronoi_case()
    quats = np.vstack([[0, 0, 0, 1.0], rng.standard_normal((N - 1, 4))])
    quats[1:] /= np.linalg.norm(quats[1:], axis=1, keepdims=True)
    tabs = cases.tables_for(c["params"], quats)[1]
    temp = np.full((nx, ny, nz), 1682.0)
    ref, _, _, st = O.run(phi, temp, 5, tabs, c["scalars"], kmax=50)
    s = Solver((nx, ny, nz), N, c["scalars"], precision="fp64")
    s.set_tables(*tabs)
    s.upload_phi(phi)
    s.upload_temp(temp)
    assert s.step(5) == 5
    assert s.stats()["general_cells"] > 0
    if N == 13:
        assert s.stats()["max_active"] >= 10   # liquid + 9 or more solids somewhere
    assert np.array_equal(s.phi(), ref)
    full = np.count_nonzero(O.apt(ref, 50)[1] == N)
    print(f"N={N}: {full} cells with every phase active after 5 steps")
    if N == 9:
        assert full >= 100
    assert np.array_equal(s.apt()[1], O.apt(ref, 50)[1])
    s.close()


def test_observables_match_numpy():
    c = cases.voronoi_case()
    with _solver_for(c, "fp64", temp_f32=True) as s:
        s.step(4)
        phi = s.phi()
        ob = s.observables()
        np.testing.assert_allclose(ob["phase_sums"], phi.reshape(phi.shape[0], -1).sum(1), rtol=1e-12)
        assert ob["phi_sum_error_max"] == pytest.approx(np.abs(phi.sum(0) - 1).max(), abs=1e-15)
        solid = 1.0 - phi[0]
        mask = solid > 0.5
        nz = phi.shape[3]
        pos = np.where(mask.any(2), nz - 1 - np.argmax(mask[:, :, ::-1], axis=2), np.nan)
        np.testing.assert_array_equal(ob["front"], pos)


@pytest.mark.parametrize("seed,n_solid", [(3, 7), (11, 12)])
def test_fp64_voronoi_seeds_bit_exact(seed, n_solid):
    c = cases.voronoi_case(nx=10, ny=9, nz=12, n_solid=n_solid, seed=seed, qseed=seed + 1)
    steps = 8
    ref, _, _, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50, temp_f32=True)
    with _solver_for(c, "fp64", temp_f32=True) as s:
        assert s.step(steps) == steps
        assert np.array_equal(s.phi(), ref)
        _assert_apt_equal(s, ref)


def test_fp64_quad_junction_bit_exact():
    """Four grains meeting at a column that crosses the front: cells with the
    liquid and four solids active fill the last register slot."""
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
    ref, _, _, st = O.run(phi, temp, 6, tabs, scal, kmax=50)
    assert st[1] >= 5
    s = Solver((nx, ny, nz), 5, scal, precision="fp64")
    s.set_tables(*tabs)
    s.upload_phi(phi)
    s.upload_temp(temp)
    assert s.step(6) == 6
    assert np.array_equal(s.phi(), ref)
    s.close()


def test_fp64_general_and_class_tables_bit_exact():
    """Reference tables take the class-form kernels and the four-row facet
    scan (opposite n111 row pairs); a table with one entry changed takes the
    general lookup kernels, and an n111 row with one value changed the
    eight-row scan.  All bit-identical to the oracle."""
    c = cases.voronoi_case(n_solid=5)
    steps = 6
    with _solver_for(c, "fp64", temp_f32=True) as s:
        assert s.stats()["class_tables"] and s.stats()["n111_pairs"]
        s.step(steps)
        ref, _, _, _ = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50, temp_f32=True)
        assert np.array_equal(s.phi(), ref)
    w, a, m, n111 = (t.copy() for t in c["tables"])
    w[2, 3] *= 1.25
    a[0, 4] *= 0.9
    n111[3, 5, 2] *= 1.0 + 1e-6
    tabs = (w, a, m, n111)
    c2 = dict(c, tables=tabs)
    with _solver_for(c2, "fp64", temp_f32=True) as s:
        assert not s.stats()["class_tables"] and not s.stats()["n111_pairs"]
        s.step(steps)
        ref, _, _, _ = O.run(c["phi"], c["temp"], steps, tabs, c["scalars"], kmax=50, temp_f32=True)
        assert np.array_equal(s.phi(), ref)


@pytest.mark.parametrize("shape", [(10, 8, 14), (12, 17, 14), (68, 18, 20),
                                   (1, 9, 14), (9, 1, 14), (1, 1, 12), (10, 8, 3)])
def test_fp64_classification_kernels_bit_exact(shape):
    """Grids whose classification tiles (64 x 8 cells) are partial in x and y,
    including the 3D path with 32-bit offsets on odd sizes, and degenerate
    grids: one cell across x, y or both (the periodic neighbour is the cell
    itself) and three planes, all inside the seed layer."""
    nx, ny, nz = shape
    c = cases.voronoi_case(nx=nx, ny=ny, nz=nz, n_solid=5, seed=5, qseed=6, seed_height=5)
    steps = 6
    ref, _, _, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50, temp_f32=True)
    with _solver_for(c, "fp64", temp_f32=True) as s:
        assert s.step(steps) == steps
        got = s.phi()
        print(shape, "max |diff| =", np.abs(got - ref).max())
        assert np.array_equal(got, ref)
        _assert_apt_equal(s, ref)


@pytest.mark.parametrize("kmax", [2, 3])
def test_fp64_small_kmax_clamp_bit_exact(kmax):
    """KMAX below the number of active phases: the checked APT keeps the
    lowest ids (liquid first) and the update runs on that set.  The
    classification pass sends an interface cell straight to its bucket only
    when no clamp applies, so clamped cells take the merge pass.  Bit-identical
    to the oracle, overflow flag included."""
    c = cases.voronoi_case(n_solid=6, seed=5, qseed=6)
    steps = 6
    ref, _, _, st = O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=kmax, temp_f32=True)
    assert st[0] == 1  # the case does overflow
    with _solver_for(c, "fp64", temp_f32=True, kmax=kmax, slot_capacity=6) as s:
        assert s.step(steps) == steps
        assert np.array_equal(s.phi(), ref)
        assert s.stats()["kmax_overflow"] == 1


def test_fp64_negative_upload_values_bit_exact():
    """An uploaded phi may hold small negative solid values (the slot layout
    stores every non-zero value).  A negative value at a cell makes that grain
    inactive there although it is stored, so the classification pass does not
    route such cells directly; they take the merge pass.  Bit-identical to the
    oracle."""
    c = cases.voronoi_case(n_solid=4)
    phi = c["phi"].copy()
    # grains absent at a cell but present at an x neighbour: store -1e-4
    pos = phi[1:] > 0
    nb = np.roll(pos, 1, axis=1) | np.roll(pos, -1, axis=1)
    cand = np.argwhere(nb & ~pos & (phi[0][None] > 0.2))
    assert len(cand) > 10
    for g, l, m, k in cand[::3]:
        phi[g + 1, l, m, k] = -1e-4
    c2 = dict(c, phi=phi)
    steps = 4
    ref, _, _, _ = O.run(phi, c["temp"], steps, c["tables"], c["scalars"], kmax=50, temp_f32=True)
    with _solver_for(c2, "fp64", temp_f32=True) as s:
        assert s.step(steps) == steps
        assert np.array_equal(s.phi(), ref)


def test_fp64_large_grid_equals_tiled_small_grid():
    """Size-independent check of the large-grid code paths.  A 512x512x128
    grid (4096 classification tiles: 16-plane work items, interface passes
    on side streams) made of 8 x 8 copies of a 64x64x128 pattern evolves,
    with x and y periodic, exactly like the pattern alone (64 tiles: 4-plane
    items, one interface launch).  FP64, bit for bit."""
    c = cases.voronoi_case(nx=64, ny=64, nz=128, n_solid=4, seed=9, qseed=10, seed_height=40)
    steps = 12
    with _solver_for(c, "fp64", temp_f32=True) as s:
        assert s.launch()["side_stream"] == 2
        s.step(steps)
        small = s.phi()
    big = dict(c, phi=np.tile(c["phi"], (1, 8, 8, 1)), temp=np.tile(c["temp"], (8, 8, 1)))
    with _solver_for(big, "fp64", temp_f32=True) as s:
        assert s.launch()["side_stream"] == 1
        s.step(steps)
        st = s.stats()
        got = s.phi()
    assert st["front_cells"] > 0 and st["merge_cells"] > 0
    assert np.array_equal(got, np.tile(small, (1, 8, 8, 1)))
