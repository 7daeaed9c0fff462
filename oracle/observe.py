"""CPU oracle of the sweep measurements on a dense phi.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
observables (mpf_batch_observables_all, mpf_probe_rays) and of the sweep
driver's host arithmetic.

It restates the reference's host-side measurement functions in plain
Python/numpy on the reference's dense layout phi[N, nx, ny, nz]:
- trilinear_sample, max_distance_to_boundary, directional_radius_one_side,
  directional_radius_symmetric and measure_observables
  (run_orientation_scan_3d.py:236-394);
- bicrystal_group_metrics (tests/verification_3d.py:1366-1383).

Pins (tests/test_observe_oracle.py):
- probes: exact crossings of linear ramps (trilinear interpolation of a
  linear field is exact, so is the linear crossing interpolation); the 0.5
  level of the tanh sphere of init_singlemode_sphere_3d at radius R; brute
  force on a fine grid of samples;
- equivalent radius: closed form of a known volume.
"""
from __future__ import annotations

import math

import numpy as np


def trilinear_sample(field: np.ndarray, x: float, y: float, z: float) -> float:
    """run_orientation_scan_3d.py:236-267"""
    nx, ny, nz = field.shape
    if x < 0.0 or y < 0.0 or z < 0.0 or x > (nx - 1) or y > (ny - 1) or z > (nz - 1):
        return float("nan")
    x0, y0, z0 = int(math.floor(x)), int(math.floor(y)), int(math.floor(z))
    x1, y1, z1 = min(x0 + 1, nx - 1), min(y0 + 1, ny - 1), min(z0 + 1, nz - 1)
    tx, ty, tz = x - x0, y - y0, z - z0
    c000, c100 = field[x0, y0, z0], field[x1, y0, z0]
    c010, c110 = field[x0, y1, z0], field[x1, y1, z0]
    c001, c101 = field[x0, y0, z1], field[x1, y0, z1]
    c011, c111 = field[x0, y1, z1], field[x1, y1, z1]
    c00 = c000 * (1.0 - tx) + c100 * tx
    c01 = c001 * (1.0 - tx) + c101 * tx
    c10 = c010 * (1.0 - tx) + c110 * tx
    c11 = c011 * (1.0 - tx) + c111 * tx
    c0 = c00 * (1.0 - ty) + c10 * ty
    c1 = c01 * (1.0 - ty) + c11 * ty
    return float(c0 * (1.0 - tz) + c1 * tz)


def max_distance_to_boundary(center, direction, shape) -> float:
    """run_orientation_scan_3d.py:270-287"""
    upper = (shape[0] - 1.0, shape[1] - 1.0, shape[2] - 1.0)
    lims = []
    for c, d, u in zip(center, direction, upper):
        if abs(d) < 1.0e-12:
            continue
        lim = (u - c) / d if d > 0.0 else (0.0 - c) / d
        if lim > 0.0:
            lims.append(float(lim))
    return min(lims) if lims else 0.0


def unit(direction) -> np.ndarray | None:
    d = np.asarray(direction, dtype=np.float64).copy()
    n = np.linalg.norm(d)
    if n < 1.0e-12:
        return None
    d /= n
    return d


def radius_one_side(field, center, direction, threshold: float, step_cells: float) -> float:
    """run_orientation_scan_3d.py:290-331 (direction already unit)."""
    center = np.asarray(center, np.float64)
    d = np.asarray(direction, np.float64)
    max_s = max_distance_to_boundary(center, d, field.shape)
    if max_s <= 0.0:
        return float("nan")
    v0 = trilinear_sample(field, center[0], center[1], center[2])
    if not np.isfinite(v0) or v0 < threshold:
        return float("nan")
    prev_s, prev_v = 0.0, v0
    n_steps = max(2, int(math.ceil(max_s / step_cells)) + 1)
    for i in range(1, n_steps + 1):
        s = min(i * step_cells, max_s)
        pos = center + d * s
        v = trilinear_sample(field, pos[0], pos[1], pos[2])
        if not np.isfinite(v):
            break
        if prev_v >= threshold and v < threshold:
            dv = prev_v - v
            if abs(dv) < 1.0e-12:
                return float(s)
            frac = float(np.clip((prev_v - threshold) / dv, 0.0, 1.0))
            return float(prev_s + frac * (s - prev_s))
        prev_s, prev_v = s, v
    return float("nan")


def radius_symmetric(field, center, direction, threshold=0.5, step_cells=0.5) -> float:
    """run_orientation_scan_3d.py:334-348"""
    dp, dm = unit(direction), unit(-np.asarray(direction, np.float64))
    rp = radius_one_side(field, center, dp, threshold, step_cells) if dp is not None else float("nan")
    rm = radius_one_side(field, center, dm, threshold, step_cells) if dm is not None else float("nan")
    vals = [v for v in (rp, rm) if np.isfinite(v)]
    return float(np.mean(vals)) if vals else float("nan")


DIRECTIONS = {"x": (1.0, 0.0, 0.0), "y": (0.0, 1.0, 0.0), "z": (0.0, 0.0, 1.0), "diag111": (1.0, 1.0, 1.0)}


def measure_observables(phi, step, dt, dx, center_cells, probe_step_cells=0.5) -> dict:
    """run_orientation_scan_3d.py:351-394"""
    solid = np.asarray(phi[1], np.float64)
    volume = float(np.sum(solid) * (dx ** 3))
    r_eq = float((3.0 * volume / (4.0 * math.pi)) ** (1.0 / 3.0)) if volume > 0.0 else 0.0
    row = {"step": int(step), "time": float(step * dt), "volume": volume, "equivalent_radius": r_eq}
    for name, vec in DIRECTIONS.items():
        r = radius_symmetric(solid, center_cells, vec, 0.5, probe_step_cells)
        row[f"radius_{name}"] = float(r * dx) if np.isfinite(r) else float("nan")
    return row


def group_volumes(phi, dx) -> np.ndarray:
    """Grain volumes phi[g].sum() * dx^3 (tests/verification_3d.py:1367)."""
    return np.array([float(phi[g].sum() * (dx ** 3)) for g in range(1, phi.shape[0])], np.float64)
