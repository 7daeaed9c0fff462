"""Host side of the verification suite: the measurements against closed
forms, the host APT scan on the reference's overflow field, the builders,
the case table and the records."""
import math
import os

import numpy as np
import pytest

from mpf import verification_3d as V


def test_front_heights_and_profile():
    phi = np.zeros((2, 3, 2, 6), np.float32)
    phi[0] = 1.0
    phi[1, 0, :, :3] = 1.0
    phi[0, 0, :, :3] = 0.0
    phi[1, 2, :, :5] = 0.8
    phi[0, 2, :, :5] = 0.2
    h = V.front_heights(phi)
    assert h[0, 0] == 2 and math.isnan(h[1, 0]) and h[2, 1] == 4
    prof = V.front_profile(phi)
    assert prof[0] == 2 and math.isnan(prof[1]) and prof[2] == 4


def test_kink_angle_closed_form():
    prof = np.array([5.0, 4.0, 5.0])
    assert V.kink_angle(prof, 1.0, 1) == pytest.approx(45.0)
    assert math.isnan(V.kink_angle(prof, 1.0, 0))


def test_crossing_heights_linear_interpolation():
    col = np.array([[1.0, 0.9, 0.7, 0.1, 0.0],      # 0.5 between k=2 (0.7) and k=3 (0.1): 2 + 0.2/0.6
                    [0.2, 0.1, 0.0, 0.0, 0.0],      # never reaches 0.5
                    [1.0, 1.0, 1.0, 1.0, 1.0]])     # solid to the top
    z = V.crossing_heights(col)
    assert z[0] == pytest.approx(2.0 + 0.2 / 0.6) and math.isnan(z[1]) and z[2] == 4.0


def _v_groove(nx=21, nz=24, root=10.0, slope=0.5, base=12.0):
    """Two grains (split at root+0.5) under a V-shaped front z = base -
    slope*(r - |x - root|) capped at base: known flank slopes and root."""
    x = np.arange(nx)
    height = base - slope * np.maximum(4.0 - np.abs(x - root), 0.0)
    z = np.arange(nz)
    solid = np.clip(height[:, None] + 0.5 - z[None, :], 0.0, 1.0).astype(np.float32)
    phi = np.zeros((3, nx, 3, nz), np.float32)
    left = (x <= root)[:, None, None]
    phi[1] = np.where(left, solid[:, None, :], 0.0)
    phi[2] = np.where(~left, solid[:, None, :], 0.0)
    phi[0] = 1.0 - phi[1] - phi[2]
    return phi, height


def test_boundary_root_and_groove_fit_on_a_v():
    phi, height = _v_groove()
    rx, rz = V.boundary_root(phi, 10.0)
    assert 10.0 <= rx <= 11.0 and np.isfinite(rz)
    g = V.groove_fit(phi, 10.0)
    # flanks: slope -0.5 left of the root, +0.5 right of it (front height
    # crossings sit at height exactly), so the angle is 2 atan(0.5)
    assert g["left_angle_deg"] == pytest.approx(-math.degrees(math.atan(0.5)), abs=1.0)
    assert g["right_angle_deg"] == pytest.approx(math.degrees(math.atan(0.5)), abs=1.0)
    assert g["dihedral_angle_like_deg"] == pytest.approx(2 * math.degrees(math.atan(0.5)), abs=2.0)
    assert g["groove_depth_cells"] > 1.0
    # no boundary: the guess and the front height there
    one = phi.copy()
    one[1] += one[2]
    one[2] = 0.0
    g1 = V.groove_fit(one, 7.0)
    assert g1["root_x_cells"] == 7.0


def test_host_apt_on_the_overflow_field():
    N = 52
    phi = V.field_single_cell_crowd(N)
    info = V.first_overflow(phi, 50)
    assert info["location"] == [0, 1, 1] and info["count"] == 52       # the reference's report
    act = V.host_active(phi)
    assert list(np.flatnonzero(act[:, 3, 3, 3])) == [0]
    assert act[:, 1, 1, 1].all() and act[:, 1, 1, 0].all() and act[:, 1, 1, 2].all()


def test_host_apt_boundaries():
    act = V.host_active(V.field_edge_seeds())
    assert act[1, 5, 2, 2]            # x periodic
    assert act[2, 3, 4, 2]            # y periodic
    assert act[3, 3, 2, 1]            # z neighbour
    assert not act[3, 3, 2, 4]        # z does not wrap


def test_motion_and_family_metrics():
    dx = 1e-4
    phi0 = V.field_flat(4, 3, 12, dx, 6e-4, 4)
    phi1 = V.field_flat(4, 3, 12, dx, 6e-4, 6)
    m = V.motion_metrics(V.field_metrics(phi1, np.ones((4, 3, 12)), dx, 0.0), phi0, phi1, dx, 2e-3)
    assert m["front_velocity"] == pytest.approx(2 * dx / 2e-3)
    assert m["solid_volume_growth"] == pytest.approx(2 * 4 * 3 * dx ** 3, rel=1e-3)
    two = V.field_bicrystal(8, 2, 10, dx, 6e-4, 5.0, 3, 0.0)
    f = V.family_volumes(two, [1], [2], dx)
    assert f["winner_family"] == "B" and f["winner_loser_volume_ratio"] == pytest.approx(5 / 3, rel=1e-5)
    assert V.signal_level(1.0, 0.5, 2.0, True) == "L1" and V.signal_level(3.0, 0.5, 2.0, False) == "L0"


def test_builders():
    b = V.params()
    assert b["KMAX"] == 50 and b["dx"] == 1e-4
    phi = V.field_flat(4, 3, 10, 1e-4, 6e-4, 5)
    assert phi.dtype == np.float32 and np.allclose(phi.sum(0), 1.0)
    w = V.field_wavy(8, 2, 12, 1e-4, 6e-4, 4.8, 1.2)
    assert w.shape == (2, 8, 2, 12) and np.allclose(w.sum(0), 1.0, atol=1e-6)
    assert np.array_equal(w[:, :, 0, :], V.field_wavy_2d(8, 12, 1e-4, 6e-4, 4.8, 1.2))
    assert np.array_equal(V.stripes(9, 2, 3)[:, 0], [1, 1, 1, 2, 2, 2, 3, 3, 3])
    q = V.quaternions(["identity", "aligned_111_to_z"])
    assert q.shape == (3, 4) and np.allclose(np.linalg.norm(q, axis=1), 1.0)
    z = V.orientation("aligned_111_to_z").apply(np.array([1.0, 1.0, 1.0]) / math.sqrt(3.0))
    assert np.allclose(z, [0.0, 0.0, 1.0])


def test_case_table():
    assert len(V.CASES) == 20
    kinds = [c.kind for c in V.CASES.values()]
    assert kinds.count("quick") == 11 and kinds.count("main") == 9
    # every case's runs build for both toggle settings, and the scans share
    # the threshold sweep's delta_a = 0.5 runs
    for c in V.CASES.values():
        for a in (False, True):
            runs = c.runs(a, a)
            assert runs and all(isinstance(r, V.Run) for r in runs.values())
    scan = set(V.CASES["single_grain_preferred_growth_benchmark"].runs(True, False).values())
    thr = set(V.CASES["anisotropy_threshold_test"].runs(True, False).values())
    assert len(scan & thr) == 2


def test_records(tmp_path):
    out = V.Outcome("x", "quick", True, "L0", "ok", "", {"a": np.float32(1.5)}, {"b": [1, 2]}, str(tmp_path / "x"))
    V.write_record(out, {"t.csv": [{"u": 1, "v": 2.0}]})
    for f in ("case_metrics.csv", "case_result.json", "case_report.md", "t.csv"):
        assert os.path.exists(tmp_path / "x" / f)
    rep = V.write_report(str(tmp_path), [out])
    assert "| x | quick | PASS | L0 | ok |" in open(rep).read()
