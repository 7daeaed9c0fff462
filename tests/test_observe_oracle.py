"""CPU checks of the sweep measurements and the sweep driver's host logic.

The probe oracle (oracle/observe.py) is pinned by closed forms: along a
linear field, trilinear interpolation and the crossing interpolation are
exact, so the probe returns the analytic crossing; the tanh sphere of
init_singlemode_sphere_3d has its 0.5 level at the radius.  The host
arithmetic of mpf.observe (growth-rate fit, anisotropy index, winner ratio)
is checked against exact answers."""
import math

import numpy as np
import pytest

from mpf import inputs, observe, sweep
from oracle import observe as OO


def _linear_field(shape, f0, grad):
    x, y, z = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    return f0 + grad[0] * x + grad[1] * y + grad[2] * z


@pytest.mark.parametrize("direction,crosses", [((1, 0, 0), True), ((0, -1, 0), True), ((0, 0, 1), True),
                                               ((1, 1, 1), True), ((-2, 1, 0.5), False), ((0, 1, 0), False)])
def test_probe_linear_field_exact_crossing(direction, crosses):
    shape = (21, 19, 23)
    grad = np.array([-0.09, 0.07, -0.05])
    c = np.array([10.25, 9.5, 11.0])
    f = _linear_field(shape, 0.0, grad)
    f += 0.9 - OO.trilinear_sample(f, *c)          # value 0.9 at the origin
    u = OO.unit(direction)
    r = OO.radius_one_side(f, c, u, 0.5, 0.5)
    if crosses:
        s_star = (0.9 - 0.5) / -float(grad @ u)
        assert s_star <= OO.max_distance_to_boundary(c, u, shape)
        assert r == pytest.approx(s_star, abs=1e-9)
    else:
        assert math.isnan(r)


def test_probe_sphere_level_set_at_radius():
    n, R = 40, 9.0
    d = inputs.derive_constants(sweep.nondimensional_params())
    phi = inputs.init_singlemode_sphere_3d(n, n, n, 1.0, 1.0, 1.0, d["delta"], R)
    c = np.array([n / 2, n / 2, n / 2])
    for name, vec in OO.DIRECTIONS.items():
        r = OO.radius_symmetric(phi[1].astype(np.float64), c, vec, 0.5, 0.5)
        assert abs(r - R) < 0.05, (name, r)
    row = OO.measure_observables(phi, 7, 2.0, 1.0, c)
    assert row["time"] == 14.0
    assert row["radius_x"] == pytest.approx(R, abs=0.05)


def test_probe_special_cases():
    f = np.zeros((8, 8, 8))
    c = np.array([3.0, 3.0, 3.0])
    u = OO.unit((1, 0, 0))
    assert math.isnan(OO.radius_one_side(f, c, u, 0.5, 0.5))            # origin below threshold
    f[:] = 1.0
    assert math.isnan(OO.radius_one_side(f, c, u, 0.5, 0.5))            # no crossing before the boundary
    assert math.isnan(OO.radius_one_side(f, np.array([-1.0, 3, 3]), u, 0.5, 0.5))   # origin outside
    assert OO.unit((0, 0, 0)) is None
    assert OO.max_distance_to_boundary(c, u, f.shape) == 4.0


def test_equivalent_radius_closed_form():
    phi = np.zeros((2, 10, 10, 10))
    phi[1, 2:5, 2:6, 3:8] = 1.0      # 60 cells
    row = OO.measure_observables(phi, 0, 1.0, 0.5, np.array([3.0, 3.0, 5.0]))
    assert row["volume"] == 60 * 0.125
    assert row["equivalent_radius"] == pytest.approx((3 * 7.5 / (4 * math.pi)) ** (1 / 3), rel=1e-15)
    host = observe.sphere_row(0, 1.0, 0.5, 60.0, {"x": float("nan")})
    assert host["volume"] == row["volume"] and host["equivalent_radius"] == row["equivalent_radius"]
    assert math.isnan(host["radius_x"])


def test_fit_growth_rate_exact_slopes():
    rows = [{"time": float(t), "equivalent_radius": 2.0 + 0.125 * t} for t in range(0, 11)]
    rate, i0, i1, n = observe.fit_growth_rate(rows, 0.5)
    assert rate == pytest.approx(0.125, rel=1e-12)
    assert (i0, i1, n) == (5, 10, 6)                       # trailing ceil(11 * 0.5) = 6 rows
    rows2 = [dict(r) for r in rows]
    for r in rows2[6:]:
        r["equivalent_radius"] = 100.0                      # beyond the cut-off: ignored
    rate, i0, i1, n = observe.fit_growth_rate(rows2, 0.5, fit_end_time_s=5.0)
    assert rate == pytest.approx(0.125, rel=1e-12) and (i0, i1, n) == (0, 5, 6)
    rows2[3]["equivalent_radius"] = float("nan")
    rate, _, _, n = observe.fit_growth_rate(rows2[:6], 1.0)
    assert rate == pytest.approx(0.125, rel=1e-12) and n == 5
    assert math.isnan(observe.fit_growth_rate(rows[:1], 0.5)[0])


def test_anisotropy_index_and_winner():
    assert observe.anisotropy_index({"radius_x": 1.0, "radius_y": 2.0, "radius_z": 3.0,
                                     "radius_diag111": float("nan")}) == pytest.approx(1.0)
    assert math.isnan(observe.anisotropy_index({"radius_x": 1.0}))
    m = observe.bicrystal_group_metrics([5.0, 3.0, 6.0], [1], [2], 0.5)
    assert m["winner_family"] == "B" and m["winner_loser_volume_ratio"] == 2.0
    assert m["volume_a"] == 3.0 * 0.125


def test_bicrystal_cases_deterministic_and_in_range():
    a = sweep.bicrystal_cases(40, seed=3)
    b = sweep.bicrystal_cases(40, seed=3)
    assert [c.shape for c in a] == [c.shape for c in b]
    assert all(np.array_equal(x.quats, y.quats) for x, y in zip(a, b))
    for c in a:
        assert c.shape[0] == c.shape[1] and c.shape[0] in (64, 96, 128) and c.shape[2] in (96, 128, 192)
        assert 0.05 <= c.undercooling < 0.3
        assert np.allclose(np.linalg.norm(c.quats, axis=1), 1.0)
        assert c.split_index == c.shape[0] // 2 and c.n_phase == 3
    assert len({c.shape for c in a}) > 3
    # the default sweep: 256 grids, about 3.5e8 cells (1.05e12 cell-updates over 3000 steps)
    full = sweep.bicrystal_cases()
    assert len(full) == 256
    assert 2.5e8 < sum(c.cells for c in full) < 4.5e8


def test_pack_covers_every_case_once():
    cs = sweep.bicrystal_cases(50, seed=1)
    cap = 6_000_000
    bins = sweep.pack(cs, cap)
    flat = sorted(i for b in bins for i in b)
    assert flat == list(range(50))
    for b in bins:
        cells = sum(cs[i].cells for i in b)
        assert cells <= cap or len(b) == 1
    mixed = cs[:3] + sweep.orientation_cases(shape=(16, 16, 16), radius_cells=4)
    for k, c in enumerate(mixed):
        c.index = k
    for b in sweep.pack(mixed, 1 << 30):
        assert len({mixed[i].kind for i in b}) == 1


def test_orientation_cases_reference_defaults():
    from scipy.spatial.transform import Rotation
    cs = sweep.orientation_cases(shape=(32, 30, 28), radius_cells=5.0)
    assert [c.label for c in cs] == ["{100}", "{110}", "{111}", "mid1", "mid2"]
    assert cs[0].center == (16.0, 15.0, 14.0)
    q111 = cs[2].quats[1]
    v = Rotation.from_quat(q111).apply(np.ones(3) / math.sqrt(3.0))
    assert np.allclose(v, [0, 0, 1], atol=1e-12)
    assert all(c.n_phase == 2 for c in cs)


def test_nondimensional_set_is_stable():
    p = sweep.nondimensional_params()
    d = inputs.derive_constants(p)
    D = d["m_sl_phi"] * d["eps0_sl"] ** 2
    assert D * p["dt"] / p["dx"] ** 2 == pytest.approx(0.05, rel=1e-12)
    assert d["delta"] == 6.0
