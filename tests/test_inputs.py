"""Host input builders (mpf/inputs.py) against the reference's own outputs
(tests/golden/inputs.npz, written from src/seed_modes_3d.py,
src/seed_modes.py and src/orientation_utils.py)."""
import math
import os

import numpy as np

from mpf import inputs
from mpf.config import flat_params, load_config

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inputs.npz")


def _p():
    p = flat_params(load_config())
    return p, inputs.derive_constants(p)


def test_voronoi_map_and_seed_layer():
    g = np.load(GOLD)
    p, d = _p()
    gm = inputs.generate_random_grain_map_3d(12, 8, 4, random_seed=17)
    assert np.array_equal(gm, g["gmap"])
    assert np.array_equal(inputs.generate_random_grain_map_3d(64, 48, 40, random_seed=42), g["big_map"])
    phi = inputs.init_phi_from_grain_map_3d(gm, 4, 12, 8, 14, p["dx"], d["delta"], 4)
    assert phi.dtype == np.float32 and np.array_equal(phi, g["phi0"])
    gmc, pidx, sol, liq = inputs.columns_from_grain_map_3d(gm, 14, p["dx"], d["delta"], 4)
    dense = np.zeros_like(phi)
    for l in range(12):
        for m in range(8):
            dense[gmc[l, m], l, m] = sol[pidx[l, m]]
            dense[0, l, m] = liq[pidx[l, m]]
    assert np.array_equal(dense, phi)


def test_orientations_and_tables():
    g = np.load(GOLD)
    p, d = _p()
    q = inputs.assign_quaternions_to_grains(4, "random", 23)
    assert np.array_equal(q, g["quats"])
    assert np.array_equal(inputs.compute_rotated_n111(q), g["n111"])
    q200 = inputs.assign_quaternions_to_grains(200, "random", 42)
    assert np.array_equal(q200, g["q200"])
    assert np.array_equal(inputs.compute_rotated_n111(q200), g["n200"])
    w, a, m = inputs.build_interaction_matrices(5, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"],
                                                d["w_gb"], d["m_gb_phi"])
    for mine, ref in ((w, "wij"), (a, "aij"), (m, "mij")):
        assert mine.dtype == np.float32 and np.array_equal(mine, g[ref])


def test_other_initial_conditions():
    g = np.load(GOLD)
    p, d = _p()
    assert np.array_equal(inputs.init_temperature_field_3d(6, 5, 40, p["T_melt"], p["G"], p["dx"], 24),
                          g["tfield"])
    assert np.array_equal(inputs.init_twomode_phi_3d(10, 4, 20, p["dx"], d["delta"], 6, 5, 1, -1), g["twomode"])
    assert np.array_equal(inputs.init_singlemode_sphere_3d(10, 9, 8, p["dx"], p["dx"], p["dx"], d["delta"],
                                                           3 * p["dx"]), g["sphere"])
    assert np.array_equal(inputs.init_singlemode_disk_2d(24, 24, 1.0, 1.0, 6.0, 6.0, center_x=12.0,
                                                         center_y=0.0), g["disk"])
    gm, pidx, sol, liq = inputs.columns_twomode_3d(10, 4, 20, p["dx"], d["delta"], 6, 5, 1, -1)
    two = g["twomode"]
    for l in range(10):
        assert np.array_equal(two[gm[l, 0], l, 0], sol[pidx[l, 0]].astype(np.float32))
        assert np.array_equal(two[0, l, 0], liq[pidx[l, 0]].astype(np.float32))


def test_derived_constants_closed_form():
    """run_randommode_3d.py:56-65 and tests/verification_3d.py:177-204 at the
    config_3d.yaml defaults, against hand-evaluated closed forms."""
    p, d = _p()
    assert d["delta"] == pytest_approx(6e-4)
    assert d["eps0_sl"] ** 2 == pytest_approx(2.13990e-4)
    assert d["w0_sl"] == pytest_approx(2933.33)
    assert d["m_sl_phi"] == pytest_approx(0.102808)
    assert d["m_gb_phi"] == pytest_approx(7.19659e-3)
    assert d["g2_floor"] == pytest_approx(1e6)
    assert d["cooling_rate"] == pytest_approx(2.5e-4)
    assert math.tan(d["a0"]) == pytest_approx(1.41235)


def pytest_approx(v):
    import pytest
    return pytest.approx(v, rel=1e-5)
