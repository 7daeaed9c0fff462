"""Shared builders of parity cases (seeded synthetic inputs in the
reference's dense layout) used by the CPU and GPU tests."""
from __future__ import annotations

import math

import numpy as np

from mpf import inputs
from mpf.config import flat_params, load_config


def ref_params(**over) -> dict:
    """config_3d.yaml values (the reference's verification defaults)."""
    p = flat_params(load_config())
    p.update(over)
    return p


def tables_for(p: dict, quats: np.ndarray):
    d = inputs.derive_constants(p)
    N = quats.shape[0]
    w, a, m = inputs.build_interaction_matrices(N, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"],
                                                d["eps_gb"], d["w_gb"], d["m_gb_phi"])
    n111 = inputs.compute_rotated_n111(quats)
    return d, (w.astype(np.float64), a.astype(np.float64), m.astype(np.float64),
               n111.astype(np.float64))


def voronoi_case(nx=12, ny=8, nz=14, n_solid=4, seed=17, qseed=23, seed_height=4, undercool=5.0,
                 **over):
    """tests/verification_3d.py:1433-1456 (basic_constraints) generalised."""
    p = ref_params(**over)
    d = inputs.derive_constants(p)
    gmap = inputs.generate_random_grain_map_3d(nx, ny, n_solid, random_seed=seed)
    phi = inputs.init_phi_from_grain_map_3d(gmap, n_solid, nx, ny, nz, p["dx"], d["delta"], seed_height)
    quats = inputs.assign_quaternions_to_grains(n_solid, "random", qseed)
    d, tabs = tables_for(p, quats)
    temp = np.full((nx, ny, nz), np.float32(p["T_melt"] - undercool), np.float32)
    scal = inputs.kernel_scalars(p, d, as_float32=True)
    return {"phi": phi.astype(np.float64), "temp": temp.astype(np.float64), "tables": tabs,
            "scalars": scal, "params": p, "derived": d, "quats": quats}


def directional_case(nx, ny, nz, n_solid, seed_height, seed=42, **over):
    """run_randommode_3d.py setup: Voronoi seed layer, linear temperature,
    uniform cooling."""
    p = ref_params(**over)
    d = inputs.derive_constants(p)
    gmap = inputs.generate_random_grain_map_3d(nx, ny, n_solid, random_seed=seed)
    quats = inputs.assign_quaternions_to_grains(n_solid, "random", seed)
    d, tabs = tables_for(p, quats)
    temp = inputs.init_temperature_field_3d(nx, ny, nz, p["T_melt"], p["G"], p["dx"], seed_height)
    scal = inputs.kernel_scalars(p, d, as_float32=True)
    cooling = float(np.float32(d["cooling_rate"]))
    return {"gmap": gmap, "temp": temp.astype(np.float64), "tables": tabs, "scalars": scal,
            "params": p, "derived": d, "quats": quats, "cooling": cooling,
            "seed_height": seed_height, "n_solid": n_solid, "shape": (nx, ny, nz)}


def disk2d_case(n=64, radius=8.0, seed=0, undercool=0.5, steps_dt=0.05):
    """Config-1 style 2D case: half disc of radius 8 cells on the mirror edge,
    nondimensional units (dx = 1)."""
    rng = np.random.default_rng(seed)
    theta = float(rng.uniform(0.0, math.pi / 2))
    p = ref_params(dx=1.0, dy=1.0, dz=1.0, dt=steps_dt, T_melt=1.0, Sf=1.0, G=0.0, V_pulling=0.0)
    d = inputs.derive_constants(p)
    phi = inputs.init_singlemode_disk_2d(n, n, 1.0, 1.0, d["delta"], radius, center_x=n / 2, center_y=0.0)
    quats = np.array([[0, 0, 0, 1.0], [0, 0, math.sin(theta / 2), math.cos(theta / 2)]])
    d, tabs = tables_for(p, quats)
    temp = np.full((n, n), 1.0 - undercool)
    scal = inputs.kernel_scalars(p, d, as_float32=False)
    return {"phi": phi.astype(np.float64), "temp": temp, "tables": tabs, "scalars": scal,
            "params": p, "derived": d, "theta": theta}
