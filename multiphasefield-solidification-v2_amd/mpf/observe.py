"""Sweep observables on device data.

The reference measures its sweeps on the host from full D2H copies of phi
(run_orientation_scan_3d.py:236-432, tests/verification_3d.py:1366-1383).
Here the grid-sized work runs on the GPU: phase sums and fronts
(mpf_batch_observables_all) and the directional ray probes
(mpf_probe_rays).  What is left on the host is the reference's arithmetic on
a handful of numbers per grid: the symmetric radius mean, the equivalent
radius, the growth-rate fit, the anisotropy index and the winner/loser ratio.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, ptr

# measure_observables' probe directions (run_orientation_scan_3d.py:341-346)
PROBE_DIRECTIONS = {
    "x": (1.0, 0.0, 0.0),
    "y": (0.0, 1.0, 0.0),
    "z": (0.0, 0.0, 1.0),
    "diag111": (1.0, 1.0, 1.0),
}


def observables_all(h, shapes, n_phase: int, with_front: bool = True):
    """Phase sums [n, N], scalars [n, 7] and per-grid fronts [nx, ny] (or
    None) of every grid of a handle, from one device pass."""
    n = len(shapes)
    gv = np.zeros((n, n_phase), np.float64)
    out = np.zeros((n, 7), np.float64)
    cols = sum(int(s[0]) * int(s[1]) for s in shapes)
    fr = np.zeros(cols, np.int32) if with_front else None
    check(_lib.lib().mpf_batch_observables_all(h, ptr(gv), ptr(fr, C.c_int32) if with_front else None, ptr(out)))
    fronts = None
    if with_front:
        fronts, o = [], 0
        for nx, ny, _ in shapes:
            fronts.append(fr[o:o + nx * ny].reshape(nx, ny))
            o += nx * ny
    return gv, out, fronts


def _unit(v) -> np.ndarray | None:
    """directional_radius_one_side's normalisation (run_orientation_scan_3d.py:296-300)."""
    d = np.asarray(v, dtype=np.float64).copy()
    d_norm = np.linalg.norm(d)
    if d_norm < 1.0e-12:
        return None
    d /= d_norm
    return d


def probe_rays(h, inst, phase, origin, direction, threshold: float = 0.5, step_cells: float = 0.5) -> np.ndarray:
    """Raw device ray probes: distance (cells) to the first downward threshold
    crossing along each (already unit) direction, NaN where there is none."""
    inst = np.ascontiguousarray(inst, np.int32).reshape(-1)
    phase = np.ascontiguousarray(phase, np.int32).reshape(-1)
    origin = np.ascontiguousarray(origin, np.float64).reshape(-1, 3)
    direction = np.ascontiguousarray(direction, np.float64).reshape(-1, 3)
    n = inst.size
    if not (phase.size == n and origin.shape[0] == n and direction.shape[0] == n):
        raise ValueError("inst, phase, origin and direction must describe the same number of rays")
    out = np.empty(n, np.float64)
    check(_lib.lib().mpf_probe_rays(h, n, ptr(inst, C.c_int32), ptr(phase, C.c_int32), ptr(origin),
                                    ptr(direction), float(threshold), float(step_cells), ptr(out)))
    return out


def directional_radii(h, insts, centers, phase: int = 1, threshold: float = 0.5, step_cells: float = 0.5,
                      directions=PROBE_DIRECTIONS) -> list[dict]:
    """directional_radius_symmetric for every named direction and every grid
    (run_orientation_scan_3d.py:334-348): the mean of the finite one-sided
    radii along +d and -d, in cells."""
    insts = list(insts)
    centers = np.asarray(centers, np.float64).reshape(len(insts), 3)
    ri, ro, rd, slot = [], [], [], []
    for q, ii in enumerate(insts):
        for name, vec in directions.items():
            for sgn, v in ((0, vec), (1, -np.asarray(vec))):
                d = _unit(v)
                if d is None:
                    continue
                ri.append(ii)
                ro.append(centers[q])
                rd.append(d)
                slot.append((q, name, sgn))
    res = [{name: [math.nan, math.nan] for name in directions} for _ in insts]
    if ri:
        vals = probe_rays(h, ri, np.full(len(ri), phase), ro, rd, threshold, step_cells)
        for (q, name, sgn), v in zip(slot, vals):
            res[q][name][sgn] = float(v)
    out = []
    for r in res:
        row = {}
        for name, (rp, rm) in r.items():
            vals = [v for v in (rp, rm) if np.isfinite(v)]
            row[name] = float(np.mean(vals)) if vals else float("nan")
        out.append(row)
    return out


def sphere_row(step: int, dt: float, dx: float, phase_sum: float, radii_cells: dict) -> dict:
    """One row of measure_observables (run_orientation_scan_3d.py:351-394)
    from the device phase sum and probe radii."""
    volume = float(phase_sum * (dx ** 3))
    equivalent_radius = float((3.0 * volume / (4.0 * math.pi)) ** (1.0 / 3.0)) if volume > 0.0 else 0.0
    row = {"step": int(step), "time": float(step * dt), "volume": volume, "equivalent_radius": equivalent_radius}
    for name, r in radii_cells.items():
        row[f"radius_{name}"] = float(r * dx) if np.isfinite(r) else float("nan")
    return row


def fit_growth_rate(timeseries: list[dict], fit_fraction: float, fit_end_time_s: float | None = None):
    """Slope of r_eq(t) (run_orientation_scan_3d.py:397-432): over the rows up
    to fit_end_time_s when at least two qualify, else over the trailing
    fit_fraction of the finite rows.  Returns (slope, first, last, count)."""
    if len(timeseries) < 2:
        return float("nan"), 0, 0, 0
    times = np.array([float(r["time"]) for r in timeseries], dtype=np.float64)
    req = np.array([float(r["equivalent_radius"]) for r in timeseries], dtype=np.float64)
    idx = np.arange(len(timeseries), dtype=np.int32)
    valid = np.isfinite(times) & np.isfinite(req)
    if fit_end_time_s is not None and fit_end_time_s > 0.0:
        cut = valid & (times <= fit_end_time_s + 1.0e-15)
        if np.count_nonzero(cut) >= 2:
            sel = idx[cut]
            slope = float(np.polyfit(times[cut], req[cut], 1)[0])
            return slope, int(sel[0]), int(sel[-1]), int(sel.size)
    vt, vr, vi = times[valid], req[valid], idx[valid]
    if vt.size < 2:
        return float("nan"), 0, 0, 0
    n_use = max(2, int(math.ceil(vt.size * fit_fraction)))
    x, y, sel = vt[-n_use:], vr[-n_use:], vi[-n_use:]
    slope = float(np.polyfit(x, y, 1)[0])
    return slope, int(sel[0]), int(sel[-1]), int(x.size)


def anisotropy_index(row: dict) -> float:
    """(max - min) / mean of the finite directional radii (run_orientation_scan_3d.py:435-444)."""
    vals = np.array([float(row.get(k, float("nan"))) for k in ("radius_x", "radius_y", "radius_z",
                                                                "radius_diag111")], dtype=np.float64)
    vals = vals[np.isfinite(vals)]
    if vals.size < 2:
        return float("nan")
    mean_v = float(np.mean(vals))
    if abs(mean_v) < 1.0e-30:
        return float("nan")
    return float((np.max(vals) - np.min(vals)) / mean_v)


def bicrystal_group_metrics(phase_sums, family_a, family_b, dx: float) -> dict:
    """Winner and winner/loser volume ratio of two grain families
    (tests/verification_3d.py:1366-1383) from the device phase sums."""
    vols = np.asarray(phase_sums, np.float64)[1:] * (dx ** 3)
    vol_a = float(np.sum([vols[g - 1] for g in family_a]))
    vol_b = float(np.sum([vols[g - 1] for g in family_b]))
    if vol_a >= vol_b:
        winner, ratio = "A", vol_a / max(vol_b, 1.0e-30)
    else:
        winner, ratio = "B", vol_b / max(vol_a, 1.0e-30)
    return {"volume_a": vol_a, "volume_b": vol_b, "winner_loser_volume_ratio": float(ratio),
            "winner_family": winner}
