"""Host helpers of the verification counterpart: interface positions,
groove angles, the host APT brute force, field builders and the records."""
import math
import os

import numpy as np
import pytest

from mpf import verification_3d as V


def test_interface_positions_and_profile():
    phi = np.zeros((2, 3, 2, 6), np.float32)
    phi[0] = 1.0
    phi[1, 0, :, :3] = 1.0
    phi[0, 0, :, :3] = 0.0
    phi[1, 2, :, :5] = 0.8
    phi[0, 2, :, :5] = 0.2
    pos = V.interface_positions_3d(phi)
    assert pos[0, 0] == 2 and math.isnan(pos[1, 0]) and pos[2, 1] == 4
    prof = V.interface_profile_xz(phi)
    assert prof[0] == 2 and math.isnan(prof[1]) and prof[2] == 4


def test_groove_angle_closed_form():
    prof = np.array([5.0, 4.0, 5.0])
    assert V.groove_angle_from_profile(prof, 1.0, 1) == pytest.approx(45.0)
    assert math.isnan(V.groove_angle_from_profile(prof, 1.0, 0))


def test_host_apt_on_the_overflow_field():
    N = 52
    phi = np.zeros((N, 4, 4, 4), np.float32)
    phi[:, 1, 1, 1] = np.float32(1.0 / N)
    phi[0] = np.clip(1.0 - phi[1:].sum(axis=0), 0.0, 1.0)
    info = V.find_kmax_exceedance_3d(phi, N, 50)
    assert info["location"] == [0, 1, 1] and info["count"] == 52       # the reference's report
    assert V.active_phase_ids_at(phi, 3, 3, 3, N) == [0]


def test_builders_and_records(tmp_path):
    b = V.base_params()
    assert b["KMAX"] == 50 and b["dx"] == 1e-4
    phi = V.flat_single_phi_3d(4, 3, 10, 1e-4, 6e-4, 5)
    assert phi.dtype == np.float32 and np.allclose(phi.sum(0), 1.0)
    w = V.wavy_single_phi_3d(8, 2, 12, 1e-4, 6e-4, 4.8, 1.2)
    assert w.shape == (2, 8, 2, 12) and np.allclose(w.sum(0), 1.0, atol=1e-6)
    assert np.array_equal(np.unique(V.stripe_grain_map(9, 2, 3)), [1, 2, 3])
    q = V.quaternions_from_rotations([V._rotation("identity"), V._rotation("aligned_111")])
    assert q.shape == (3, 4) and np.allclose(np.linalg.norm(q, axis=1), 1.0)
    r = V.make_case_result("x", True, "ok", "", {"a": np.float32(1.5)}, {"b": [1, 2]}, str(tmp_path / "x"))
    assert r.passed and os.path.exists(tmp_path / "x" / "case_report.md")
    rep = V.write_report(str(tmp_path), [r])
    assert "| x | quick | PASS |" in open(rep).read()
