"""Batched sweeps of small independent grids.

The reference runs its sweeps one grid at a time, one launch per kernel per
step, and measures each grid on the host from a full copy of phi:
- the bicrystal runs (run_twomode_3d.py:200-322; the verification's
  run_bicrystal_competition, tests/verification_3d.py:1951-2013);
- the isothermal orientation scan (run_orientation_scan_3d.py:495-871).

Here a sweep is a list of cases, packed into batches (mpf.Batch).  Every
kernel launch of a step advances every grid of a batch, and the grid-sized
measurements run on the GPU (mpf.observe).  Optionally, the launch shape is
autotuned and steps are replayed from HIP graphs.

Bicrystal sweep: n grids drawn from default_rng(seed).  Each has
- nx = ny from sizes_xy and nz from sizes_z;
- N = 3, two grains split at nx/2 over a flat seed layer (init_twomode_phi_3d);
- two random unit quaternions (normalised standard_normal(4));
- a uniform undercooling from U[undercooling), isothermal.

Orientation scan: one spherical grain per case (init_singlemode_sphere_3d),
N = 2, isothermal, with the reference's five default orientations.

Both use the nondimensional parameter set below unless given another.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import inputs
from .batch import Batch
from .config import flat_params, load_config
from . import observe

# Nondimensional units: dx = dt = 1 and T_melt = Sf = 1.  The interface
# energy is 1 and the grain-boundary energy keeps the configuration's ratio
# (0.60 / 0.44).  M_SL = 0.05 makes m_phi * eps^2 * dt / dx^2 = M_SL * gamma = 0.05
# (explicit stability needs it well below 1/6).  The flat-front velocity
# M * Sf * dT is then 0.0025-0.015 cells/step for dT in [0.05, 0.3].
NONDIMENSIONAL = {
    "dx": 1.0, "dy": 1.0, "dz": 1.0, "dt": 1.0, "T_melt": 1.0, "Sf": 1.0, "G": 0.0, "V_pulling": 0.0,
    "gamma_100": 1.0, "gamma_GB": 0.60 / 0.44, "M_SL": 0.05,
}


def nondimensional_params(**over) -> dict:
    """config_3d.yaml (anisotropy, KMAX, interface width in cells) with the
    nondimensional units above."""
    p = flat_params(load_config())
    p.update(NONDIMENSIONAL)
    p.update(over)
    return p


@dataclass
class Case:
    index: int
    kind: str                       # "bicrystal" or "sphere"
    shape: tuple
    quats: np.ndarray               # [N, 4], liquid first
    undercooling: float
    label: str = ""
    seed_height: float = 0.0        # bicrystal
    split_index: int = 0
    radius: float = 0.0             # sphere, cells
    center: tuple = (0.0, 0.0, 0.0)  # sphere centre, cells
    meta: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return int(np.prod(self.shape))

    @property
    def n_phase(self) -> int:
        return int(self.quats.shape[0])


def bicrystal_cases(n: int = 256, seed: int = 0, sizes_xy=(64, 96, 128), sizes_z=(96, 128, 192),
                    undercooling=(0.05, 0.3), seed_height: float = 24.0) -> list[Case]:
    """The bicrystal sweep's cases, drawn in order from default_rng(seed):
    nx (= ny), nz, two quaternions, undercooling."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nxy = int(rng.choice(np.asarray(sizes_xy)))
        nz = int(rng.choice(np.asarray(sizes_z)))
        q = rng.standard_normal((2, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        dT = float(rng.uniform(undercooling[0], undercooling[1]))
        quats = np.vstack([[0.0, 0.0, 0.0, 1.0], q])
        out.append(Case(i, "bicrystal", (nxy, nxy, nz), quats, dT, label=f"bicrystal_{i:04d}",
                        seed_height=float(seed_height), split_index=nxy // 2))
    return out


def align_direction_to_z_quaternion(direction_xyz) -> list:
    """run_orientation_scan_3d.py:91-96"""
    from scipy.spatial.transform import Rotation
    src = np.array([direction_xyz], dtype=np.float64)
    src /= np.linalg.norm(src, axis=1, keepdims=True)
    tgt = np.array([[0.0, 0.0, 1.0]], dtype=np.float64)
    rot, _ = Rotation.align_vectors(tgt, src)
    return rot.as_quat().astype(np.float64).tolist()


def default_orientation_cases() -> list[dict]:
    """run_orientation_scan_3d.py:99-127"""
    return [
        {"label": "{100}", "orientation_type": "euler", "euler_deg": [0.0, 0.0, 0.0]},
        {"label": "{110}", "orientation_type": "euler", "euler_deg": [0.0, 45.0, 0.0]},
        {"label": "{111}", "orientation_type": "quaternion",
         "quaternion": align_direction_to_z_quaternion([1.0, 1.0, 1.0])},
        {"label": "mid1", "orientation_type": "euler", "euler_deg": [0.0, 15.0, 0.0]},
        {"label": "mid2", "orientation_type": "euler", "euler_deg": [20.0, 35.0, 10.0]},
    ]


def orientation_cases(raw_cases=None, shape=(96, 96, 96), radius_cells: float = 10.0,
                      undercooling: float = 0.1, center_frac=(0.5, 0.5, 0.5), dx: float = 1.0) -> list[Case]:
    """resolve_orientation_cases (run_orientation_scan_3d.py:130-160) and the
    scan's sphere geometry (:553-556, :698-712)."""
    import warnings

    from scipy.spatial.transform import Rotation
    raw_cases = raw_cases or default_orientation_cases()
    nx, ny, nz = (int(v) for v in shape)
    center_m = (center_frac[0] * nx * dx, center_frac[1] * ny * dx, center_frac[2] * nz * dx)
    center = (center_m[0] / dx, center_m[1] / dx, center_m[2] / dx)
    out = []
    for idx, raw in enumerate(raw_cases, start=1):
        if not isinstance(raw, dict):
            raise ValueError(f"orientation_cases[{idx - 1}] must be a dict.")
        q = inputs.build_quaternion_from_config(raw)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            euler = Rotation.from_quat(q).as_euler("xyz", degrees=True)
        quats = np.vstack([[0.0, 0.0, 0.0, 1.0], q])
        out.append(Case(idx - 1, "sphere", (nx, ny, nz), quats, float(undercooling),
                        label=str(raw.get("label", f"case_{idx:02d}")), radius=float(radius_cells),
                        center=center,
                        meta={"center_m": center_m, "orientation_type": str(raw.get("orientation_type", "euler")),
                              "euler_deg": [float(v) for v in euler], "quaternion": [float(v) for v in q]}))
    return out


def pack(cases: list[Case], max_cells: int) -> list[list[int]]:
    """First-fit decreasing by cells: batches of at most max_cells cells (a
    case larger than that gets a batch of its own), one phase count per
    batch, at most 4096 grids per batch."""
    order = sorted(range(len(cases)), key=lambda i: (-cases[i].cells, i))
    bins: list[list[int]] = []
    load: list[int] = []
    for i in order:
        c = cases[i]
        for b, members in enumerate(bins):
            if (load[b] + c.cells <= max_cells and len(members) < 4096
                    and cases[members[0]].n_phase == c.n_phase and cases[members[0]].kind == c.kind):
                members.append(i)
                load[b] += c.cells
                break
        else:
            bins.append([i])
            load.append(c.cells)
    return [sorted(b) for b in bins]


ANISO_MODES = {  # run_orientation_scan_3d.py:186-196
    "full": lambda da, ks: (1, 1, da, ks),
    "isotropic": lambda da, ks: (0, 0, 0.0, 1.0),
    "energetic_only": lambda da, ks: (1, 1, da, 1.0),
    "kinetic_only": lambda da, ks: (1, 0, 0.0, ks),
}


def _percentiles(v):
    v = np.asarray(v, np.float64)
    if v.size == 0:
        return {"p50": float("nan"), "p99": float("nan")}
    return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99))}


def physics(params: dict | None = None, aniso_mode: str = "full"):
    """(params, derived constants, kernel scalars, anisotropy on, torque on)
    of a sweep; the anisotropy mode adjusts delta_a and ksi as the
    orientation scan does."""
    p = dict(params) if params is not None else nondimensional_params()
    if p["dx"] != p["dy"] or p["dx"] != p["dz"]:
        raise ValueError("3D kernels assume dx == dy == dz.")
    en_aniso, en_torque, delta_a, ksi = ANISO_MODES[aniso_mode](p["delta_a"], p["ksi"])
    p["delta_a"], p["ksi"] = delta_a, ksi
    d = inputs.derive_constants(p)
    return p, d, inputs.kernel_scalars(p, d, as_float32=True), bool(en_aniso), bool(en_torque)


def case_tables(c: Case, d: dict):
    """(wij, aij, mij, n111) of a case, float64."""
    w, a, m = inputs.build_interaction_matrices(c.n_phase, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"],
                                                d["w_gb"], d["m_gb_phi"])
    return (np.asarray(w, np.float64), np.asarray(a, np.float64), np.asarray(m, np.float64),
            np.asarray(inputs.compute_rotated_n111(c.quats), np.float64))


def case_initial_phi(c: Case, p: dict, d: dict) -> np.ndarray:
    """Dense initial phi of a case (float32, the reference's init functions)."""
    nx, ny, nz = c.shape
    dx = p["dx"]
    if c.kind == "sphere":
        cm = c.meta.get("center_m", tuple(v * dx for v in c.center))
        return inputs.init_singlemode_sphere_3d(nx, ny, nz, dx, dx, dx, d["delta"], c.radius * dx,
                                                center_x=cm[0], center_y=cm[1], center_z=cm[2])
    return inputs.init_twomode_phi_3d(nx, ny, nz, dx, d["delta"], c.seed_height, c.split_index)


def case_temperature(c: Case, p: dict) -> np.ndarray:
    """Uniform temperature column: float32-rounded for the orientation scan
    (its temp_iso is float32, run_orientation_scan_3d.py:678), double for
    the bicrystal runs."""
    t = p["T_melt"] - c.undercooling
    return np.full(c.shape[2], float(np.float32(t)) if c.kind == "sphere" else t)


def run_sweep(cases: list[Case], steps: int, params: dict | None = None, precision: str = "fp32",
              max_batch_cells: int = 1 << 30, sample_every: int = 0, graphs: bool = False, autotune: int = 0,
              aniso_mode: str = "full", fit_fraction: float = 0.5, fit_end_time: float | None = None,
              probe_step_cells: float = 0.5, log=None) -> dict:
    """Run every case for `steps` steps in batches; returns per-case results
    and the sweep's timing (instances/s, MLUPS, per-grid completion latency
    and per-chunk step time percentiles)."""
    if steps < 1:
        raise ValueError("steps must be positive")
    p, d, scal, en_aniso, en_torque = physics(params, aniso_mode)
    dx, dt = p["dx"], p["dt"]
    chunk = sample_every if sample_every and sample_every > 0 else steps
    results: list[dict | None] = [None] * len(cases)
    batches = pack(cases, max_batch_cells)
    say = log or (lambda *_: None)
    t0 = time.perf_counter()
    batch_rows = []
    step_ms = []
    for bi, members in enumerate(batches):
        bc = [cases[i] for i in members]
        N = bc[0].n_phase
        sphere = bc[0].kind == "sphere"
        tb = time.perf_counter()
        with Batch([c.shape for c in bc], N, scal, precision=precision, anisotropy=en_aniso,
                   torque=en_torque, temp_f32=sphere) as b:
            for q, c in enumerate(bc):
                b.set_tables(q, *case_tables(c, d))
                nx, ny, nz = c.shape
                if sphere:
                    b.upload_phi(q, case_initial_phi(c, p, d))
                else:
                    # the columnar form of init_twomode_phi_3d: no dense array on the host
                    b.init_columns(q, *inputs.columns_twomode_3d(nx, ny, nz, dx, d["delta"], c.seed_height,
                                                                c.split_index))
                b.upload_temp(q, case_temperature(c, p))
            t_tune = time.perf_counter()
            tuned = b.autotune(autotune) if autotune else None
            t_run = time.perf_counter()
            if graphs:
                b.set_graphs(True)
            series = [[] for _ in bc]
            centers = [c.center for c in bc]

            def sample(step):
                gv, out, fronts = observe.observables_all(b._h, b.shapes, N, with_front=not sphere)
                radii = observe.directional_radii(b._h, range(len(bc)), centers, phase=1,
                                                  step_cells=probe_step_cells) if sphere else None
                for q in range(len(bc)):
                    if sphere:
                        series[q].append(observe.sphere_row(step, dt, dx, gv[q, 1], radii[q]))
                    else:
                        f = fronts[q].astype(np.float64)
                        f = f[fronts[q] >= 0]
                        series[q].append({"step": int(step), "time": float(step * dt),
                                          "phase_sums": gv[q].tolist(),
                                          "solid_fraction": float(out[q, 4] / out[q, 5]),
                                          "front_mean": float(f.mean()) if f.size else float("nan"),
                                          "front_std": float(f.std()) if f.size else float("nan"),
                                          "phi_sum_error_max": float(out[q, 0])})

            sample(0)
            done = 0
            gpu_ms = 0.0
            while done < steps:
                n = min(chunk, steps - done)
                b.step(n)
                ms = b.stats()["last_step_ms"]
                gpu_ms += ms
                step_ms.append(ms / n)
                done += n
                sample(done)
            st = b.stats()
            launch = b.launch()
        tb1 = time.perf_counter()
        for q, c in enumerate(bc):
            results[members[q]] = _case_result(c, series[q], dx, fit_fraction, fit_end_time, latency=tb1 - t0)
        cells = sum(c.cells for c in bc)
        batch_rows.append({"instances": len(bc), "cells": cells, "gpu_ms_per_step": gpu_ms / steps,
                           "wall_s": tb1 - tb, "setup_s": t_tune - tb, "autotune_s": t_run - t_tune,
                           "run_s": tb1 - t_run, "mlups_gpu": cells * steps / (gpu_ms * 1e-3) / 1e6,
                           "launch": launch, "autotune": tuned, "max_active": st["max_active"],
                           "kmax_overflow": st["kmax_overflow"]})
        say(f"batch {bi + 1}/{len(batches)}: {len(bc)} grids, {cells} cells, {gpu_ms / steps:.3f} ms/step, "
            f"wall {tb1 - tb:.2f} s (setup {t_tune - tb:.2f}, autotune {t_run - t_tune:.2f}, "
            f"steps+samples {tb1 - t_run:.2f}), launch {launch}")
    wall = time.perf_counter() - t0
    cell_updates = sum(c.cells for c in cases) * steps
    lat = [r["latency_s"] for r in results]
    gpu_total = sum(r["gpu_ms_per_step"] * steps for r in batch_rows) * 1e-3
    timing = {
        "instances": len(cases), "steps": steps, "cell_updates": cell_updates, "wall_s": wall,
        "instances_per_s": len(cases) / wall, "mlups_wall": cell_updates / wall / 1e6,
        "gpu_s": gpu_total, "mlups_gpu": cell_updates / gpu_total / 1e6,
        "latency_s": _percentiles(lat), "step_ms": _percentiles(step_ms), "batches": batch_rows,
        "precision": precision, "graphs": bool(graphs), "sample_every": chunk,
    }
    return {"cases": results, "timing": timing}


def _case_result(c: Case, series, dx, fit_fraction, fit_end_time, latency) -> dict:
    base = {"index": c.index, "label": c.label, "kind": c.kind, "shape": list(c.shape),
            "undercooling": c.undercooling, "quaternions": c.quats[1:].tolist(), "latency_s": latency,
            "timeseries": series}
    if c.kind == "sphere":
        rate, i0, i1, npts = observe.fit_growth_rate(series, fit_fraction, fit_end_time)
        fin = series[-1]
        base.update({
            "orientation_label": c.label, **{k: v for k, v in c.meta.items() if k != "center_m"},
            "final_volume": fin["volume"], "final_equivalent_radius": fin["equivalent_radius"],
            "fitted_growth_rate": rate, "anisotropy_index": observe.anisotropy_index(fin),
            "fit_start_step": int(series[i0]["step"] if npts > 0 else 0),
            "fit_end_step": int(series[i1]["step"] if npts > 0 else 0),
            "fit_start_time": float(series[i0]["time"] if npts > 0 else 0.0),
            "fit_end_time": float(series[i1]["time"] if npts > 0 else 0.0),
            "fit_points": int(npts),
        })
    else:
        fin = series[-1]
        fam = observe.bicrystal_group_metrics(fin["phase_sums"], [1], [2], dx)
        base.update({
            "seed_height": c.seed_height, "split_index": c.split_index, **fam,
            "winner_gid": 1 if fam["winner_family"] == "A" else 2,
            "solid_fraction": fin["solid_fraction"], "front_mean": fin["front_mean"],
            "front_std": fin["front_std"], "phi_sum_error_max": fin["phi_sum_error_max"],
            "front_velocity_cells_per_step": (fin["front_mean"] - series[0]["front_mean"]) / max(fin["step"], 1),
        })
    return base


# ----------------------------------------------------------------- outputs
ORIENTATION_SUMMARY_FIELDS = [  # run_orientation_scan_3d.py:846-866
    "case_name", "orientation_label", "orientation_type", "euler_x_deg", "euler_y_deg", "euler_z_deg",
    "quat_x", "quat_y", "quat_z", "quat_w", "final_volume", "final_equivalent_radius", "fitted_growth_rate",
    "anisotropy_index", "fit_start_step", "fit_end_step", "fit_start_time", "fit_end_time", "fit_points",
]
TIMESERIES_FIELDS = ["step", "time", "volume", "equivalent_radius", "radius_x", "radius_y", "radius_z",
                     "radius_diag111"]
BICRYSTAL_SUMMARY_FIELDS = [
    "index", "label", "nx", "ny", "nz", "undercooling", "volume_a", "volume_b", "winner_gid",
    "winner_loser_volume_ratio", "solid_fraction", "front_mean", "front_std", "front_velocity_cells_per_step",
    "phi_sum_error_max", "latency_s",
]


def write_outputs(res: dict, out_dir: str) -> list[str]:
    """summary.csv (and, for the orientation scan, one timeseries.csv per
    case under NN_label/) in the reference's columns, plus sweep.json."""
    import re
    os.makedirs(out_dir, exist_ok=True)
    written = []
    cases = res["cases"]
    path = os.path.join(out_dir, "summary.csv")
    with open(path, "w", encoding="utf-8", newline="") as f:
        if cases and cases[0]["kind"] == "sphere":
            wr = csv.DictWriter(f, fieldnames=ORIENTATION_SUMMARY_FIELDS)
            wr.writeheader()
            for r in cases:
                name = f"{r['index'] + 1:02d}_" + (re.sub(r"[^A-Za-z0-9._-]+", "_", r["label"].strip())
                                                  .strip("_") or "case")
                e, q = r["euler_deg"], r["quaternion"]
                wr.writerow({"case_name": name, "orientation_label": r["label"],
                             "orientation_type": r["orientation_type"],
                             "euler_x_deg": e[0], "euler_y_deg": e[1], "euler_z_deg": e[2],
                             "quat_x": q[0], "quat_y": q[1], "quat_z": q[2], "quat_w": q[3],
                             **{k: r[k] for k in ORIENTATION_SUMMARY_FIELDS[10:]}})
                cdir = os.path.join(out_dir, name)
                os.makedirs(cdir, exist_ok=True)
                tp = os.path.join(cdir, "timeseries.csv")
                with open(tp, "w", encoding="utf-8", newline="") as g:
                    tw = csv.DictWriter(g, fieldnames=TIMESERIES_FIELDS)
                    tw.writeheader()
                    tw.writerows(r["timeseries"])
                written.append(tp)
        else:
            wr = csv.DictWriter(f, fieldnames=BICRYSTAL_SUMMARY_FIELDS)
            wr.writeheader()
            for r in cases:
                nx, ny, nz = r["shape"]
                wr.writerow({"nx": nx, "ny": ny, "nz": nz,
                             **{k: r[k] for k in BICRYSTAL_SUMMARY_FIELDS if k not in ("nx", "ny", "nz")}})
    written.append(path)
    jp = os.path.join(out_dir, "sweep.json")
    with open(jp, "w", encoding="utf-8") as f:
        json.dump(res, f, default=float)
    written.append(jp)
    return written


def main(argv=None):
    ap = argparse.ArgumentParser(description="Batched bicrystal or orientation-scan sweep on one GPU.")
    ap.add_argument("kind", choices=["bicrystal", "orientation"])
    ap.add_argument("--instances", type=int, default=256, help="bicrystal grids")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--sample-every", type=int, default=500)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--batch-cells", type=int, default=1 << 30,
                    help="cells per batch (default: the whole bicrystal sweep in one batch)")
    ap.add_argument("--graphs", type=int, default=0)
    ap.add_argument("--autotune", type=int, default=4)
    ap.add_argument("--nx", type=int, default=96)
    ap.add_argument("--ny", type=int, default=96)
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--initial-radius-cells", type=float, default=10.0)
    ap.add_argument("--undercooling", type=float, default=0.1)
    ap.add_argument("--aniso-mode", choices=sorted(ANISO_MODES), default="full")
    ap.add_argument("--fit-fraction", type=float, default=0.5)
    ap.add_argument("--fit-end-time", type=float, default=None)
    ap.add_argument("--probe-step-cells", type=float, default=0.5)
    ap.add_argument("--out-dir", default=None, help="write summary.csv / timeseries.csv / sweep.json here")
    args = ap.parse_args(argv)
    if args.kind == "bicrystal":
        cases = bicrystal_cases(args.instances, args.seed)
    else:
        cases = orientation_cases(shape=(args.nx, args.ny, args.nz), radius_cells=args.initial_radius_cells,
                                  undercooling=args.undercooling)
    res = run_sweep(cases, args.steps, precision=args.precision, max_batch_cells=args.batch_cells,
                    sample_every=args.sample_every, graphs=bool(args.graphs), autotune=args.autotune,
                    aniso_mode=args.aniso_mode, fit_fraction=min(1.0, max(0.1, args.fit_fraction)),
                    fit_end_time=args.fit_end_time, probe_step_cells=max(0.1, args.probe_step_cells),
                    log=lambda s: print(s, file=sys.stderr, flush=True))
    if args.out_dir:
        write_outputs(res, args.out_dir)
    t = res["timing"]
    print(json.dumps({"kind": args.kind, **{k: v for k, v in t.items() if k != "batches"},
                      "batches": len(t["batches"])}, default=float), flush=True)


if __name__ == "__main__":
    main()
