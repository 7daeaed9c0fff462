"""Counterpart of the reference's verification suite (tests/verification_3d.py):
the same cases, setups, metrics, pass rules and per-case records
(case_metrics.csv, case_result.json, case_report.md) and a summary
report, with the step run by the HIP solver (FP32 storage, float32
temperature, as the reference's verification loop, :771-936).

    python -m mpf.verification_3d --out-dir result/verification [--cases name ...]

Cases (reference builders): basic_constraints (:1433),
kmax_overflow_detection (:1459), static_flat_interface (:1483),
two_d_limit_consistency (:1509), boundary_conditions (:1538),
isotropic_orientation_independence (:1568), anisotropic_preferred_growth
(:1597), torque_term_contribution (:1637), grain_competition (:1663),
grain_boundary_groove_trijunction (:1696), convergence_grid_dt_delta (:1726),
single_grain_preferred_growth_benchmark (:2016), directional_preference_map
(:2070), anisotropy_threshold_test (:2148),
bicrystal_competition_low_high_driving (:2201),
bicrystal_interfacial_energy_dominated_regime_test (:2260),
bicrystal_kinetic_dominated_regime_test (:2326),
multigrain_competition_extension (:2388), groove_depth_only (:2453),
robust_groove_angle_estimation (:2491).

The reference measures on host copies of small fields; so does this module
(these grids are at most 24x12x32), with the device doing the steps and the
active-phase tables.
"""
from __future__ import annotations

import argparse
import copy
import csv
import datetime
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import inputs
from .config import DEFAULT_CONFIG, flat_params, load_config
from .solver import KMaxOverflowError, Solver

KMAX_3D = 50


# ------------------------------------------------------------ parameters
def base_params(config_path: str | None = None) -> dict:
    """SimulationParams from config_3d.yaml (tests/verification_3d.py:110-173)."""
    p = flat_params(load_config(config_path or DEFAULT_CONFIG))
    p["KMAX"] = KMAX_3D
    return p


def replace(p: dict, **kw) -> dict:
    q = dict(p)
    q.update(kw)
    return q


def _rotation(kind):
    from scipy.spatial.transform import Rotation
    if kind == "identity":
        return Rotation.identity()
    if kind == "aligned_111":
        target = np.array([[0.0, 0.0, 1.0]])
        source = np.array([[1.0, 1.0, 1.0]]) / math.sqrt(3.0)
        rot, _ = Rotation.align_vectors(target, source)
        return rot
    return Rotation.from_euler("xyz", kind, degrees=True)


def quaternions_from_rotations(rotations) -> np.ndarray:
    """Liquid identity first, then one quaternion per grain (:306-311)."""
    q = np.zeros((len(rotations) + 1, 4))
    q[0] = (0.0, 0.0, 0.0, 1.0)
    for g, r in enumerate(rotations, start=1):
        q[g] = r.as_quat()
    return q


# ------------------------------------------------------------ initial fields
def constant_temp(shape, value) -> np.ndarray:
    return np.full(shape, np.float32(value), np.float32)


def flat_single_phi_3d(nx, ny, nz, dx, delta, seed_height) -> np.ndarray:
    """:243-252"""
    phi = np.zeros((2, nx, ny, nz), np.float32)
    factor = np.float32(2.2 / delta)
    dist = (np.arange(nz, dtype=np.float64) - seed_height) * dx
    phi[1] = (0.5 * (1.0 - np.tanh(factor * dist))).astype(np.float32)[None, None, :]
    phi[0] = 1.0 - phi[1]
    return phi


def wavy_single_phi_3d(nx, ny, nz, dx, delta, base_height, amplitude_cells, phase_shift=0.0) -> np.ndarray:
    """:255-274"""
    phi = np.zeros((2, nx, ny, nz), np.float32)
    factor = np.float32(2.2 / delta)
    x = np.arange(nx, dtype=np.float64).reshape(nx, 1, 1)
    z = np.arange(nz, dtype=np.float64).reshape(1, 1, nz)
    height = base_height + amplitude_cells * np.cos(2.0 * math.pi * x / max(nx, 1) + phase_shift)
    solid = (0.5 * (1.0 - np.tanh(factor * ((z - height) * dx)))).astype(np.float32)
    phi[1] = np.broadcast_to(solid, (nx, ny, nz))
    phi[0] = 1.0 - phi[1]
    return phi


def wavy_single_phi_2d(nx, ny, dx, delta, base_height, amplitude_cells, phase_shift=0.0) -> np.ndarray:
    """:277-295"""
    phi = np.zeros((2, nx, ny), np.float32)
    factor = np.float32(2.2 / delta)
    x = np.arange(nx, dtype=np.float64).reshape(nx, 1)
    y = np.arange(ny, dtype=np.float64).reshape(1, ny)
    height = base_height + amplitude_cells * np.cos(2.0 * math.pi * x / max(nx, 1) + phase_shift)
    phi[1] = (0.5 * (1.0 - np.tanh(factor * ((y - height) * dx)))).astype(np.float32)
    phi[0] = 1.0 - phi[1]
    return phi


def stripe_grain_map(nx, ny, n_solid) -> np.ndarray:
    """:298-303"""
    gm = np.zeros((nx, ny), np.int32)
    edges = np.linspace(0, nx, n_solid + 1, dtype=int)
    for g in range(1, n_solid + 1):
        gm[edges[g - 1]:edges[g], :] = g
    return gm


# ------------------------------------------------------------ measurements
def active_phase_ids_at(phi, l, m, k, n_phase):
    """Host APT of one cell (:360-379): x, y periodic, z mirror."""
    nx, ny, nz = phi.shape[1:]
    lp, lm = (l + 1) % nx, (l - 1) % nx
    mp, mm = (m + 1) % ny, (m - 1) % ny
    kp, km = min(k + 1, nz - 1), max(k - 1, 0)
    out = []
    for g in range(n_phase):
        c = phi[g]
        if c[l, m, k] > 0.0 or (c[l, m, k] == 0.0 and (c[lp, m, k] > 0.0 or c[lm, m, k] > 0.0 or c[l, mp, k] > 0.0
                                                        or c[l, mm, k] > 0.0 or c[l, m, kp] > 0.0
                                                        or c[l, m, km] > 0.0)):
            out.append(g)
    return out


def find_kmax_exceedance_3d(phi, n_phase, kmax):
    """:391-406"""
    nx, ny, nz = phi.shape[1:]
    for l in range(nx):
        for m in range(ny):
            for k in range(nz):
                a = active_phase_ids_at(phi, l, m, k, n_phase)
                if len(a) > kmax:
                    return {"location": [l, m, k], "count": len(a), "active_ids_preview": a[:16]}
    return None


def interface_positions_3d(phi, threshold=0.5) -> np.ndarray:
    """:409-416"""
    mask = (1.0 - phi[0]) > threshold
    nz = mask.shape[2]
    has = mask.any(axis=2)
    rev = np.argmax(mask[:, :, ::-1], axis=2)
    return np.where(has, nz - 1 - rev, np.nan).astype(np.float64)


def interface_profile_xz(phi, y_index=None, threshold=0.5) -> np.ndarray:
    """:419-430"""
    solid = 1.0 - phi[0]
    ny = solid.shape[1]
    y = ny // 2 if y_index is None else int(np.clip(y_index, 0, ny - 1))
    mask = solid[:, y, :] > threshold
    nz = mask.shape[1]
    has = mask.any(axis=1)
    rev = np.argmax(mask[:, ::-1], axis=1)
    return np.where(has, nz - 1 - rev, np.nan).astype(np.float64)


def groove_angle_from_profile(profile, dx, center_idx) -> float:
    """:433-442"""
    if center_idx <= 0 or center_idx >= profile.size - 1:
        return float("nan")
    left = profile[center_idx] - profile[center_idx - 1]
    right = profile[center_idx + 1] - profile[center_idx]
    if not (np.isfinite(left) and np.isfinite(right)):
        return float("nan")
    la = math.degrees(math.atan(abs(left) / max(dx, 1e-30)))
    ra = math.degrees(math.atan(abs(right) / max(dx, 1e-30)))
    return 0.5 * (la + ra)


def compute_metrics_3d(phi, nf, dx, runtime_s, interface_angle_deg=float("nan")) -> dict:
    """:612-654 on the float32 field."""
    phi_sum = phi.sum(axis=0)
    grain_volumes = [float(phi[g].sum() * (dx ** 3)) for g in range(1, phi.shape[0])]
    interface = interface_positions_3d(phi)
    finite = interface[np.isfinite(interface)]
    mean_z = float(np.mean(finite) * dx) if finite.size else float("nan")
    std_z = float(np.std(finite) * dx) if finite.size else float("nan")
    nan = float("nan")
    return {
        "phi_sum_error_max": float(np.max(np.abs(phi_sum - 1.0))),
        "phi_min": float(np.min(phi)), "phi_max": float(np.max(phi)),
        "nf_max": int(np.max(nf)) if nf.size else 0,
        "nan_count": int(np.isnan(phi).sum()),
        "solid_fraction": float(np.mean(1.0 - phi[0])),
        "solid_volume": float(np.sum(1.0 - phi[0]) * (dx ** 3)),
        "grain_volumes": grain_volumes,
        "surviving_grains": int(sum(v > dx ** 3 for v in grain_volumes)),
        "interface_position_mean": mean_z,
        "interface_position_max": float(np.nanmax(finite) * dx) if finite.size else nan,
        "interface_position_std": std_z, "interface_roughness": std_z,
        "interface_angle_deg": float(interface_angle_deg),
        "solid_volume_growth": nan, "winner_loser_volume_ratio": nan, "grain_boundary_root_shift": nan,
        "trijunction_x": nan, "trijunction_z": nan, "dihedral_angle_like_metric": nan,
        "angle_deviation_from_final_state": nan, "front_velocity": nan, "runtime_s": float(runtime_s),
    }


# ------------------------------------------------------------ runs
def _scalars(p, d):
    return inputs.kernel_scalars(p, d, as_float32=True)


def _tables(p, d, quats):
    N = quats.shape[0]
    w, a, m = inputs.build_interaction_matrices(N, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"], d["w_gb"],
                                                d["m_gb_phi"])
    return w, a, m, inputs.compute_rotated_n111(quats)


def run_3d_simulation(label, p, phi0, temp0, quats, anisotropy, torque, sample_every=None, observer=None,
                      precision="fp32"):
    """The verification loop (:771-936) on the device: float32 state and
    temperature, the checked APT every step (a KMAX overflow raises),
    optional observer samples; metrics of the initial and final fields."""
    N = phi0.shape[0]
    nx, ny, nz = phi0.shape[1:]
    if N != quats.shape[0]:
        raise ValueError("phi0 phase count and quaternion count do not match.")
    d = inputs.derive_constants(p)
    cooling = float(np.float32(d["cooling_rate"]))
    with Solver((nx, ny, nz), N, _scalars(p, d), kmax=p["KMAX"], precision=precision, temp_dense=True,
                temp_f32=True, anisotropy=bool(anisotropy), torque=bool(anisotropy and torque), halt_on_kmax=True,
                cooling_rate=cooling) as s:
        s.set_tables(*_tables(p, d, quats))
        s.upload_phi(phi0.astype(np.float32))
        s.upload_temp(np.asarray(temp0, np.float64))
        nf0 = s.apt()[1]
        initial = compute_metrics_3d(phi0, nf0, p["dx"], 0.0)
        samples = []
        t0 = time.perf_counter()
        done = 0
        chunk = sample_every if (observer is not None and sample_every) else p["nsteps"]
        nf_max = int(nf0.max())
        while done < p["nsteps"]:
            n = min(chunk, p["nsteps"] - done)
            if s.step(n) != n:
                st = s.stats()
                phi_h = s.phi(np.float32)
                raise KMaxOverflowError(f"KMAX overflow in {label}: max active phases {st['max_active']}, "
                                        f"details={find_kmax_exceedance_3d(phi_h, N, p['KMAX'])}")
            done += n
            if observer is not None and sample_every:
                rec = observer(s.phi(np.float32), done)
                if rec is not None:
                    samples.append({"step": done, **rec})
        runtime = time.perf_counter() - t0
        st = s.stats()
        if st["kmax_overflow"]:
            raise KMaxOverflowError(f"KMAX overflow in final APT of {label}: max active {st['max_active']}")
        phi = s.phi(np.float32)
        nf = s.apt()[1]
        temp = s.temp()
    metrics = compute_metrics_3d(phi, nf, p["dx"], runtime)
    metrics["nf_max"] = max(metrics["nf_max"], int(st["max_active"]), nf_max)
    return {"phi": phi, "temp": temp, "nf": nf, "metrics": metrics, "initial_metrics": initial,
            "number_of_grain": N, "samples": samples}


def run_2d_simulation(label, p, phi0, temp0, quats, precision="fp32"):
    """:939-1012: the 2D kernels, float32 state and temperature."""
    N = phi0.shape[0]
    nx, ny = phi0.shape[1:]
    d = inputs.derive_constants(p)
    with Solver((nx, ny), N, _scalars(p, d), kmax=p["KMAX"], precision=precision, temp_dense=True, temp_f32=True,
                cooling_rate=float(np.float32(d["cooling_rate"]))) as s:
        s.set_tables(*_tables(p, d, quats))
        s.upload_phi(phi0.astype(np.float32))
        s.upload_temp(np.asarray(temp0, np.float64))
        t0 = time.perf_counter()
        s.step(p["nsteps"])
        rt = time.perf_counter() - t0
        phi = s.phi(np.float32)
    return {"phi": phi, "runtime_s": rt}


def run_checked_apt_only(p, phi):
    """The checked APT alone on a dense field (:737-768), through the
    reference-layout drop-in kernel."""
    from . import dense
    N = phi.shape[0]
    nx, ny, nz = phi.shape[1:]
    depth = max(p["MAX_GRAINS"], p["KMAX"], N)
    d_phi = dense.to_device(phi.astype(np.float32))
    d_mf = dense.to_device(np.zeros((depth, nx, ny, nz), np.int32))
    d_nf = dense.to_device(np.zeros((nx, ny, nz), np.int32))
    d_st = dense.to_device(np.zeros(2, np.int32))
    t0 = time.perf_counter()
    dense.apt(d_phi, d_mf, d_nf, d_st, N, checked=True, kmax=p["KMAX"])
    dense.synchronize()
    rt = time.perf_counter() - t0
    mf, nf, status = d_mf.copy_to_host(), d_nf.copy_to_host(), d_st.copy_to_host()
    metrics = compute_metrics_3d(phi, nf, p["dx"], rt)
    metrics["nf_max"] = max(metrics["nf_max"], int(status[1]))
    return {"mf": mf, "nf": nf, "status": status, "metrics": metrics}


def resolve_toggles(default_anisotropy, default_torque, overrides):
    """:1015-1020"""
    a = default_anisotropy if overrides.get("anisotropy") is None else bool(overrides["anisotropy"])
    t = default_torque if overrides.get("torque") is None else bool(overrides["torque"])
    return a, (t if a else False)


# ------------------------------------------------------------ records
@dataclass
class CaseOutcome:
    name: str
    passed: bool
    summary: str
    failure_reason: str
    metrics: dict
    extra: dict
    case_dir: str
    category: str = "quick"
    level: str = "L0"


def _ser(o):
    if isinstance(o, np.ndarray):
        return o.tolist()
    if isinstance(o, (np.bool_,)):
        return bool(o)
    if isinstance(o, np.integer):
        return int(o)
    if isinstance(o, np.floating):
        return float(o)
    if isinstance(o, dict):
        return {k: _ser(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_ser(v) for v in o]
    return o


def make_case_result(name, passed, summary, failure_reason, metrics, extra, case_dir, category="quick",
                     level="L0") -> CaseOutcome:
    """case_metrics.csv, case_result.json and case_report.md (:1040-1082, :1384-1430)."""
    os.makedirs(case_dir, exist_ok=True)
    with open(os.path.join(case_dir, "case_metrics.csv"), "w", encoding="utf-8", newline="") as f:
        w = csv.writer(f)
        w.writerow(["key", "value"])
        rows = {"name": name, "category": category, "passed": passed, "level": level, **metrics,
                **{f"extra_{k}": v for k, v in extra.items()}}
        for k, v in rows.items():
            w.writerow([k, json.dumps(_ser(v), ensure_ascii=False)])
    with open(os.path.join(case_dir, "case_result.json"), "w", encoding="utf-8") as f:
        json.dump(_ser({"name": name, "passed": passed, "category": category, "level": level, "summary": summary,
                        "failure_reason": failure_reason, "metrics": metrics, "extra": extra}), f,
                  ensure_ascii=False, indent=2)
    lines = [f"# {name}", "", f"- Category: {category}", f"- Status: {'PASS' if passed else 'FAIL'}",
             f"- Physical level: {level}", f"- Summary: {summary}"]
    if failure_reason:
        lines.append(f"- Failure reason: {failure_reason}")
    lines += ["", "## Metrics", ""] + [f"- {k}: `{json.dumps(_ser(v), ensure_ascii=False)}`"
                                       for k, v in metrics.items()]
    if extra:
        lines += ["", "## Extra", ""] + [f"- {k}: `{json.dumps(_ser(v), ensure_ascii=False)}`"
                                         for k, v in extra.items()]
    with open(os.path.join(case_dir, "case_report.md"), "w", encoding="utf-8") as f:
        f.write("\n".join(lines))
    return CaseOutcome(name, bool(passed), summary, failure_reason, metrics, extra, case_dir, category, level)


# ------------------------------------------------------------ cases
def case_basic_constraints(case_dir, base, ov):
    """:1433-1456"""
    p = replace(base, nx=12, ny=8, nz=14, nsteps=6)
    d = inputs.derive_constants(p)
    gm = inputs.generate_random_grain_map_3d(p["nx"], p["ny"], 4, random_seed=17)
    phi0 = inputs.init_phi_from_grain_map_3d(gm, 4, p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 4)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - 5.0)
    quats = inputs.assign_quaternions_to_grains(4, "random", 23)
    a, t = resolve_toggles(True, True, ov)
    m = run_3d_simulation("basic_constraints", p, phi0, temp0, quats, a, t)["metrics"]
    passed = (m["nan_count"] == 0 and m["phi_min"] >= -1e-5 and m["phi_max"] <= 1.0 + 1e-5
              and m["phi_sum_error_max"] <= 5e-4 and m["nf_max"] <= p["KMAX"])
    return make_case_result("basic_constraints", passed,
                            f"Short multi-grain run stayed bounded with phi_sum_error_max="
                            f"{m['phi_sum_error_max']:.2e} and nf_max={m['nf_max']}.",
                            "" if passed else "Constraint violation detected in short random-mode run.", m,
                            {"enable_anisotropy": a, "enable_torque": t}, case_dir)


def case_kmax_overflow_detection(case_dir, base, ov):
    """:1459-1480"""
    p = replace(base, nx=4, ny=4, nz=4, nsteps=0)
    N = p["KMAX"] + 2
    phi = np.zeros((N, 4, 4, 4), np.float32)
    phi[:, 1, 1, 1] = np.float32(1.0 / N)
    phi[0] = np.clip(1.0 - phi[1:].sum(axis=0), 0.0, 1.0)
    apt = run_checked_apt_only(p, phi)
    info = find_kmax_exceedance_3d(phi, N, p["KMAX"])
    det = int(apt["status"][0]) == 1 and info is not None
    return make_case_result("kmax_overflow_detection", det,
                            f"KMAX guard {'detected' if det else 'missed'} an overflow request of "
                            f"{info['count'] if info else 'unknown'} active phases.",
                            "" if det else "Checked APT kernel did not report the crafted KMAX overflow.",
                            apt["metrics"], {"status": apt["status"].tolist(), "overflow_info": info,
                                             "requested_number_of_grain": N}, case_dir)


def case_static_flat_interface(case_dir, base, ov):
    """:1483-1506"""
    p = replace(base, nx=12, ny=6, nz=14, nsteps=6)
    d = inputs.derive_constants(p)
    phi0 = flat_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 5)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"])
    a, t = resolve_toggles(False, False, ov)
    r = run_3d_simulation("static_flat", p, phi0, temp0, quaternions_from_rotations([_rotation("identity")]), a, t)
    m, i = r["metrics"], r["initial_metrics"]
    shift = abs(m["interface_position_mean"] - i["interface_position_mean"])
    rough = m["interface_roughness"]
    passed = shift <= 1.05 * p["dx"] and rough <= 0.15 * p["dx"]
    return make_case_result("static_flat_interface", passed,
                            f"Flat interface mean shift={shift:.2e} m, roughness={rough:.2e} m "
                            f"with anisotropy={a}, torque={t}.",
                            "" if passed else "Planar interface drifted or roughened more than the static threshold.",
                            m, {"initial_interface_mean": i["interface_position_mean"],
                                "final_interface_mean": m["interface_position_mean"], "mean_shift": shift}, case_dir)


def case_two_d_limit_consistency(case_dir, base, ov):
    """:1509-1535"""
    p = replace(base, nx=12, ny=4, nz=14, nsteps=6, delta_a=0.0, ksi=1.0)
    d = inputs.derive_constants(p)
    phi3 = wavy_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 4.8, 1.2)
    q = quaternions_from_rotations([_rotation("identity")])
    a, t = resolve_toggles(False, False, ov)
    r3 = run_3d_simulation("two_d_limit_3d", p, phi3, constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - 4.0),
                           q, a, t)
    phi2 = wavy_single_phi_2d(p["nx"], p["nz"], p["dx"], d["delta"], 4.8, 1.2)
    r2 = run_2d_simulation("two_d_limit_2d", p, phi2, constant_temp((p["nx"], p["nz"]), p["T_melt"] - 4.0), q)
    diff = r3["phi"][1, :, p["ny"] // 2, :].astype(np.float64) - r2["phi"][1].astype(np.float64)
    linf = float(np.max(np.abs(diff)))
    l2 = float(np.sqrt(np.mean(diff * diff)))
    passed = linf <= 2e-3 and l2 <= 8e-4
    return make_case_result("two_d_limit_consistency", passed,
                            f"3D mid-y slice matched 2D reference with L_inf={linf:.2e}, L2={l2:.2e}.",
                            "" if passed else "3D y-uniform evolution diverged from the 2D reference beyond tolerance.",
                            r3["metrics"], {"phi_linf_vs_2d": linf, "phi_l2_vs_2d": l2}, case_dir)


def case_boundary_conditions(case_dir, base, ov):
    """:1538-1565"""
    p = replace(base, nx=6, ny=5, nz=5, nsteps=0)
    phi = np.zeros((4, 6, 5, 5), np.float32)
    phi[1, 0, 2, 2] = 1.0
    phi[2, 3, 0, 2] = 1.0
    phi[3, 3, 2, 0] = 1.0
    phi[0] = np.clip(1.0 - phi[1:].sum(axis=0), 0.0, 1.0)
    apt = run_checked_apt_only(p, phi)
    mf, nf = apt["mf"], apt["nf"]

    def has(cell, g):
        l, m, k = cell
        return g in mf[:nf[l, m, k], l, m, k].tolist()

    checks = {"x_periodic_wrap": has((p["nx"] - 1, 2, 2), 1), "y_periodic_wrap": has((3, p["ny"] - 1, 2), 2),
              "z_mirror_adjacent": has((3, 2, 1), 3), "z_no_wrap": not has((3, 2, p["nz"] - 1), 3)}
    passed = all(checks.values())
    return make_case_result("boundary_conditions", passed,
                            "Boundary-condition probes: " + ", ".join(f"{k}={v}" for k, v in checks.items()),
                            "" if passed else "At least one x/y periodic or z-Neumann boundary probe failed.",
                            apt["metrics"], {"checks": checks}, case_dir)


def case_isotropic_orientation_independence(case_dir, base, ov):
    """:1568-1594"""
    p = replace(base, nx=12, ny=6, nz=14, nsteps=6, delta_a=0.0, ksi=1.0)
    d = inputs.derive_constants(p)
    phi0 = flat_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 5)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - 6.0)
    a, t = resolve_toggles(False, False, ov)
    ra = run_3d_simulation("iso_a", p, phi0, temp0, quaternions_from_rotations([_rotation("identity")]), a, t)
    rb = run_3d_simulation("iso_b", p, phi0, temp0, quaternions_from_rotations([_rotation("aligned_111")]), a, t)
    mean_diff = abs(ra["metrics"]["interface_position_mean"] - rb["metrics"]["interface_position_mean"])
    solid_diff = abs(ra["metrics"]["solid_fraction"] - rb["metrics"]["solid_fraction"])
    dphi = ra["phi"][1].astype(np.float64) - rb["phi"][1].astype(np.float64)
    l2 = float(np.sqrt(np.mean(dphi * dphi)))
    passed = mean_diff <= 0.15 * p["dx"] and solid_diff <= 2e-4 and l2 <= 2e-3
    return make_case_result("isotropic_orientation_independence", passed,
                            f"Isotropic orientation sensitivity stayed small: mean_diff={mean_diff:.2e} m, "
                            f"solid_diff={solid_diff:.2e}, L2={l2:.2e}.",
                            "" if passed else "Orientation changed the isotropic solution more than the allowed tolerance.",
                            ra["metrics"], {"interface_mean_diff": mean_diff, "solid_fraction_diff": solid_diff,
                                            "phi_l2_diff": l2}, case_dir)


def case_anisotropic_preferred_growth(case_dir, base, ov):
    """:1597-1634"""
    p = replace(base, nx=16, ny=6, nz=18, nsteps=12, delta_a=0.55, p_round=0.02)
    d = inputs.derive_constants(p)
    phi0 = wavy_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 6.0, 2.0, phase_shift=0.2)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"])
    a, t = resolve_toggles(True, False, ov)
    ri = run_3d_simulation("aniso_identity", p, phi0, temp0, quaternions_from_rotations([_rotation("identity")]), a, t)
    rl = run_3d_simulation("aniso_aligned", p, phi0, temp0, quaternions_from_rotations([_rotation("aligned_111")]),
                           a, t)
    mi, ml = ri["metrics"]["interface_position_mean"], rl["metrics"]["interface_position_mean"]
    dmean = abs(mi - ml)
    rdiff = abs(ri["metrics"]["interface_roughness"] - rl["metrics"]["interface_roughness"])
    dphi = ri["phi"][1].astype(np.float64) - rl["phi"][1].astype(np.float64)
    l2 = float(np.sqrt(np.mean(dphi * dphi)))
    pref = "identity" if mi > ml else "aligned111"
    passed = dmean >= 0.10 * p["dx"] or rdiff >= 0.05 * p["dx"] or l2 >= 5e-5
    return make_case_result("anisotropic_preferred_growth", passed,
                            f"Anisotropy created orientation-dependent evolution: |delta mean|={dmean:.2e} m, "
                            f"roughness_diff={rdiff:.2e} m, L2={l2:.2e}; {pref} advanced further.",
                            "" if passed else "Orientation dependence was too weak to confirm anisotropic preference.",
                            ri["metrics"], {"interface_mean_identity": mi, "interface_mean_aligned111": ml,
                                            "interface_mean_abs_diff": dmean, "roughness_diff": rdiff,
                                            "phi_l2_diff": l2, "preferred_orientation": pref,
                                            "enable_torque": t}, case_dir)


def case_torque_term_contribution(case_dir, base, ov):
    """:1637-1660"""
    p = replace(base, nx=16, ny=6, nz=18, nsteps=12, delta_a=0.55, p_round=0.02)
    d = inputs.derive_constants(p)
    phi0 = wavy_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 6.0, 2.0, phase_shift=0.3)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"])
    q = quaternions_from_rotations([_rotation([20.0, 35.0, 10.0])])
    a, _ = resolve_toggles(True, False, ov)
    t_on = True if ov.get("torque") is None else bool(ov["torque"])
    r_off = run_3d_simulation("torque_off", p, phi0, temp0, q, a, False)
    r_on = run_3d_simulation("torque_on", p, phi0, temp0, q, a, t_on)
    dphi = r_on["phi"][1].astype(np.float64) - r_off["phi"][1].astype(np.float64)
    l2 = float(np.sqrt(np.mean(dphi * dphi)))
    rdiff = abs(r_on["metrics"]["interface_roughness"] - r_off["metrics"]["interface_roughness"])
    passed = l2 >= 5e-5 or rdiff >= 0.05 * p["dx"]
    return make_case_result("torque_term_contribution", passed,
                            f"Torque changed the solution by L2={l2:.2e}, roughness_diff={rdiff:.2e} m.",
                            "" if passed else "Torque on/off produced no measurable difference in the chosen anisotropic case.",
                            r_on["metrics"], {"phi_l2_diff": l2, "roughness_diff": rdiff, "anisotropy_enabled": a},
                            case_dir)


def case_grain_competition(case_dir, base, ov):
    """:1663-1693"""
    p = replace(base, nx=12, ny=12, nz=16, nsteps=8)
    d = inputs.derive_constants(p)
    gm = inputs.generate_random_grain_map_3d(p["nx"], p["ny"], 6, random_seed=31)
    phi0 = inputs.init_phi_from_grain_map_3d(gm, 6, p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 5)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - 5.0)
    quats = inputs.assign_quaternions_to_grains(6, "random", 11)
    a, t = resolve_toggles(True, True, ov)
    r = run_3d_simulation("grain_competition", p, phi0, temp0, quats, a, t)
    v0 = np.array(r["initial_metrics"]["grain_volumes"])
    v1 = np.array(r["metrics"]["grain_volumes"])
    dv = float(np.max(np.abs(v1 - v0)))
    top0 = np.unique(np.argmax(phi0, axis=0)[:, :, -1])
    top1 = np.unique(np.argmax(r["phi"], axis=0)[:, :, -1])
    s0, s1 = r["initial_metrics"]["surviving_grains"], r["metrics"]["surviving_grains"]
    passed = s1 <= s0 and dv >= 2e-14
    return make_case_result("grain_competition", passed,
                            f"Competition changed grain volumes by max {dv:.2e} m^3 and surviving grains {s0}->{s1}.",
                            "" if passed else "Short multi-grain run did not show enough measurable competition.",
                            r["metrics"], {"max_volume_change": dv,
                                           "top_surface_phase_count_initial": int(np.sum(top0 > 0)),
                                           "top_surface_phase_count_final": int(np.sum(top1 > 0))}, case_dir)


def case_grain_boundary_groove(case_dir, base, ov):
    """:1696-1723"""
    p = replace(base, nx=18, ny=6, nz=18, nsteps=10, delta_a=0.0, ksi=1.0)
    d = inputs.derive_constants(p)
    gm = stripe_grain_map(p["nx"], p["ny"], 3)
    phi0 = inputs.init_phi_from_grain_map_3d(gm, 3, p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 6)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"])
    q = quaternions_from_rotations([_rotation("identity")] * 3)
    a, t = resolve_toggles(False, False, ov)
    r = run_3d_simulation("grain_boundary_groove", p, phi0, temp0, q, a, t)
    prof = interface_profile_xz(r["phi"])
    s1, s2 = p["nx"] // 3, 2 * p["nx"] // 3
    far = np.nanmean(np.concatenate([prof[1:s1 - 1], prof[s1 + 1:s2 - 1], prof[s2 + 1:-1]]))
    l1 = np.nanmin(prof[max(s1 - 1, 0):min(s1 + 2, prof.size)])
    l2_ = np.nanmin(prof[max(s2 - 1, 0):min(s2 + 2, prof.size)])
    depth = float(max(far - l1, far - l2_, 0.0) * p["dx"])
    a1 = groove_angle_from_profile(prof, p["dx"], s1)
    a2 = groove_angle_from_profile(prof, p["dx"], s2)
    angle = float(np.nanmean([a1, a2]))
    m = compute_metrics_3d(r["phi"], r["nf"], p["dx"], r["metrics"]["runtime_s"], interface_angle_deg=angle)
    passed = depth >= 0.15 * p["dx"] and np.isfinite(angle)
    return make_case_result("grain_boundary_groove_trijunction", passed,
                            f"GB groove depth={depth:.2e} m, groove angle={angle:.2f} deg.",
                            "" if passed else "Capillary relaxation did not produce a measurable groove at the trijunctions.",
                            m, {"groove_depth": depth, "angle_split1_deg": a1, "angle_split2_deg": a2}, case_dir)


def case_convergence_sweep(case_dir, base, ov):
    """:1726-1780"""
    coarse = replace(base, nx=12, ny=4, nz=16, dx=1e-4, dt=5e-5, nsteps=6, delta_a=0.0, ksi=1.0)
    dt_half = replace(coarse, dt=2.5e-5, nsteps=12)
    delta_wide = replace(coarse, delta_factor=8.0)
    fine = replace(base, nx=24, ny=4, nz=32, dx=5e-5, dt=1.25e-5, nsteps=24, delta_a=0.0, ksi=1.0)

    def build(p):
        d = inputs.derive_constants(p)
        phi = wavy_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 0.35 * p["nz"], 0.08 * p["nx"],
                                 phase_shift=0.25)
        return phi, constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"])

    q = quaternions_from_rotations([_rotation("identity")])
    rc = run_3d_simulation("convergence_coarse", coarse, *build(coarse), q, False, False)
    rd = run_3d_simulation("convergence_dt_half", dt_half, *build(dt_half), q, False, False)
    rw = run_3d_simulation("convergence_delta_wide", delta_wide, *build(delta_wide), q, False, False)
    rf = run_3d_simulation("convergence_fine", fine, *build(fine), q, False, False)
    yc, yf = coarse["ny"] // 2, fine["ny"] // 2
    cs = rc["phi"][1, :, yc, :].astype(np.float64)
    ds = rd["phi"][1, :, yc, :].astype(np.float64)
    ws = rw["phi"][1, :, yc, :].astype(np.float64)
    fs = rf["phi"][1, ::2, yf, ::2].astype(np.float64)
    grid_error = float(np.sqrt(np.mean((cs - fs) ** 2)))
    dt_error = float(np.sqrt(np.mean((ds - fs) ** 2)))
    delta_sens = float(np.sqrt(np.mean((ws - cs) ** 2)))
    im = lambda r: r["metrics"]["interface_position_mean"]   # noqa: E731
    gdiff, ddiff, wdiff = abs(im(rc) - im(rf)), abs(im(rc) - im(rd)), abs(im(rc) - im(rw))
    passed = gdiff <= 2.5 * coarse["dx"] and ddiff <= 1.5 * coarse["dx"] and wdiff <= 1.5 * coarse["dx"]
    return make_case_result("convergence_grid_dt_delta", passed,
                            f"Convergence sanity check: grid_error={grid_error:.2e}, dt_error={dt_error:.2e}, "
                            f"delta_sensitivity={delta_sens:.2e}, interface_diffs=({gdiff:.2e}, {ddiff:.2e}, "
                            f"{wdiff:.2e}).",
                            "" if passed else "Grid, dt, or interface-width sensitivity exceeded the small-run sanity thresholds.",
                            rc["metrics"], {"grid_error": grid_error, "dt_error": dt_error,
                                            "delta_sensitivity": delta_sens, "grid_interface_diff": gdiff,
                                            "dt_interface_diff": ddiff, "delta_interface_diff": wdiff,
                                            "fine_runtime_s": rf["metrics"]["runtime_s"]}, case_dir)



# ------------------------------------------------------------ groove and bicrystal measurements
def interface_curve_from_field_xz(field_xz, threshold=0.5):
    """Linear threshold crossing per column (:445-467)."""
    nx, nz = field_xz.shape
    xs = np.arange(nx, dtype=np.float64)
    zc = np.full(nx, np.nan)
    for x in range(nx):
        col = field_xz[x]
        above = np.where(col >= threshold)[0]
        if above.size == 0:
            continue
        k0 = int(above[-1])
        if k0 >= nz - 1:
            zc[x] = float(nz - 1)
            continue
        v0, v1 = float(col[k0]), float(col[k0 + 1])
        zc[x] = float(k0) if abs(v1 - v0) < 1e-12 else float(k0) + (threshold - v0) / (v1 - v0)
    return xs, zc


def solid_interface_curve_xz(phi, y_index=None, threshold=0.5):
    """:470-474"""
    solid = 1.0 - phi[0]
    ny = solid.shape[1]
    y = ny // 2 if y_index is None else int(np.clip(y_index, 0, ny - 1))
    return interface_curve_from_field_xz(solid[:, y, :], threshold)


def sample_column_at_fractional_z(col, z):
    """:477-486"""
    if not np.isfinite(z):
        return float("nan")
    nz = col.shape[0]
    z0 = int(np.clip(int(np.floor(z)), 0, nz - 1))
    z1 = min(z0 + 1, nz - 1)
    f = float(z - z0)
    return float((1.0 - f) * col[z0] + f * col[z1])


def boundary_delta_profile_x(phi, gl, gr, y_index=None):
    """:489-499"""
    xs, zc = solid_interface_curve_xz(phi, y_index)
    ny = phi.shape[2]
    y = ny // 2 if y_index is None else int(np.clip(y_index, 0, ny - 1))
    delta = np.full_like(xs, np.nan)
    for x in range(phi.shape[1]):
        delta[x] = sample_column_at_fractional_z(phi[gl, x, y, :] - phi[gr, x, y, :], zc[x])
    return xs, zc, delta


def interpolate_root_from_sign_change(xs, zc, delta, x_guess=None):
    """:502-528"""
    pairs = []
    for i in range(len(xs) - 1):
        d0, d1 = delta[i], delta[i + 1]
        if not (np.isfinite(d0) and np.isfinite(d1)):
            continue
        if d0 == 0.0:
            pairs.append((float(xs[i]), float(zc[i])))
            continue
        if d0 * d1 > 0.0:
            continue
        f = abs(d0) / max(abs(d0) + abs(d1), 1e-12)
        xr = float(xs[i] + f * (xs[i + 1] - xs[i]))
        z0, z1 = zc[i], zc[i + 1]
        zr = float((1.0 - f) * z0 + f * z1) if (np.isfinite(z0) and np.isfinite(z1)) else float("nan")
        pairs.append((xr, zr))
    if not pairs:
        return float("nan"), float("nan")
    if x_guess is None:
        return pairs[len(pairs) // 2]
    return min(pairs, key=lambda q: abs(q[0] - x_guess))


def grain_boundary_root_and_trijunction(phi, gl=1, gr=2, y_index=None, x_guess=None):
    """:531-545"""
    xs, zc, delta = boundary_delta_profile_x(phi, gl, gr, y_index)
    rx, rz = interpolate_root_from_sign_change(xs, zc, delta, x_guess)
    return {"root_x_cells": float(rx), "root_z_cells": float(rz), "trijunction_x_cells": float(rx),
            "trijunction_z_cells": float(rz)}


def fit_interface_line(xs, zc, x_min, x_max):
    """:548-568"""
    mask = np.isfinite(zc) & (xs >= x_min) & (xs <= x_max)
    x, z = xs[mask], zc[mask]
    if x.size < 2:
        return {"slope": float("nan"), "intercept": float("nan"), "angle_deg": float("nan"), "rmse": float("nan")}
    slope, intercept = np.polyfit(x, z, deg=1)
    fit = slope * x + intercept
    return {"slope": float(slope), "intercept": float(intercept), "angle_deg": float(math.degrees(math.atan(slope))),
            "rmse": float(np.sqrt(np.mean((fit - z) ** 2)))}


def robust_groove_metrics(phi, root_x_guess, y_index=None, fit_half_width=4.0, exclude_half_width=1.0):
    """:571-609"""
    xs, zc = solid_interface_curve_xz(phi, y_index)
    root = grain_boundary_root_and_trijunction(phi, 1, 2, y_index, root_x_guess)
    rx, rz = root["root_x_cells"], root["root_z_cells"]
    if not np.isfinite(rx):
        rx = float(root_x_guess)
        i = int(np.clip(round(rx), 0, len(zc) - 1))
        rz = float(zc[i])
    left = fit_interface_line(xs, zc, rx - fit_half_width, rx - exclude_half_width)
    right = fit_interface_line(xs, zc, rx + exclude_half_width, rx + fit_half_width)
    far_mask = np.isfinite(zc) & ((xs <= rx - fit_half_width) | (xs >= rx + fit_half_width))
    far = float(np.nanmean(zc[far_mask])) if np.any(far_mask) else float("nan")
    depth = float(max(far - rz, 0.0)) if (np.isfinite(far) and np.isfinite(rz)) else float("nan")
    dih = (float(abs(right["angle_deg"] - left["angle_deg"]))
           if np.isfinite(left["angle_deg"]) and np.isfinite(right["angle_deg"]) else float("nan"))
    return {"root_x_cells": float(rx), "root_z_cells": float(rz), "trijunction_x_cells": float(rx),
            "trijunction_z_cells": float(rz), "left_angle_deg": left["angle_deg"],
            "right_angle_deg": right["angle_deg"], "left_fit_rmse": left["rmse"], "right_fit_rmse": right["rmse"],
            "groove_depth_cells": depth, "dihedral_angle_like_deg": dih}


def finalize_metrics(metrics, phi_initial, phi_final, dx, total_time, root_initial_cells=None,
                     trijunction_final=None, groove_metrics=None, angle_history_deg=None):
    """:657-705"""
    ini = interface_positions_3d(phi_initial)
    fin = ini[np.isfinite(ini)]
    ini_mean = float(np.mean(fin) * dx) if fin.size else float("nan")
    metrics["solid_volume_growth"] = float(np.sum((1.0 - phi_final[0]) - (1.0 - phi_initial[0])) * (dx ** 3))
    if np.isfinite(metrics.get("interface_position_mean", float("nan"))) and np.isfinite(ini_mean):
        metrics["front_velocity"] = float((metrics["interface_position_mean"] - ini_mean) / max(total_time, 1e-30))
    pos = [float(v) for v in metrics.get("grain_volumes", []) if float(v) > 0.5 * (dx ** 3)]
    if len(pos) >= 2:
        metrics["winner_loser_volume_ratio"] = float(max(pos) / max(min(pos), 1e-30))
    if trijunction_final is not None:
        metrics["trijunction_x"] = float(trijunction_final.get("trijunction_x_cells", float("nan")) * dx)
        metrics["trijunction_z"] = float(trijunction_final.get("trijunction_z_cells", float("nan")) * dx)
        if root_initial_cells is not None and np.isfinite(root_initial_cells):
            metrics["grain_boundary_root_shift"] = float((trijunction_final.get("root_x_cells", float("nan"))
                                                          - root_initial_cells) * dx)
    if groove_metrics is not None:
        a = float(groove_metrics.get("dihedral_angle_like_deg", float("nan")))
        metrics["interface_angle_deg"] = a
        metrics["dihedral_angle_like_metric"] = a
        metrics["trijunction_x"] = float(groove_metrics.get("trijunction_x_cells", float("nan")) * dx)
        metrics["trijunction_z"] = float(groove_metrics.get("trijunction_z_cells", float("nan")) * dx)
        if root_initial_cells is not None and np.isfinite(root_initial_cells):
            metrics["grain_boundary_root_shift"] = float((groove_metrics.get("root_x_cells", float("nan"))
                                                          - root_initial_cells) * dx)
    if angle_history_deg:
        fa = [float(a) for a in angle_history_deg if np.isfinite(a)]
        if fa and np.isfinite(metrics.get("interface_angle_deg", float("nan"))):
            metrics["angle_deviation_from_final_state"] = float(np.mean(np.abs(np.array(fa)
                                                                               - float(metrics["interface_angle_deg"]))))
    return metrics


def blank_metrics():
    """:708-733"""
    nan = float("nan")
    return {"phi_sum_error_max": nan, "phi_min": nan, "phi_max": nan, "nf_max": -1, "nan_count": -1,
            "solid_fraction": nan, "solid_volume": nan, "grain_volumes": [], "surviving_grains": -1,
            "interface_position_mean": nan, "interface_position_max": nan, "interface_position_std": nan,
            "interface_roughness": nan, "interface_angle_deg": nan, "solid_volume_growth": nan,
            "winner_loser_volume_ratio": nan, "grain_boundary_root_shift": nan, "trijunction_x": nan,
            "trijunction_z": nan, "dihedral_angle_like_metric": nan, "angle_deviation_from_final_state": nan,
            "front_velocity": nan, "runtime_s": nan}


def classify_signal_level(signal, l1, l2, passed):
    """:1030-1037"""
    if not passed or not np.isfinite(signal):
        return "L0"
    return "L2" if signal >= l2 else ("L1" if signal >= l1 else "L0")


def bicrystal_phi_3d(nx, ny, nz, dx, delta, seed_height, split_index, left_gid=1, right_gid=2, left_offset=0.0,
                     right_offset=0.0, waviness=0.0):
    """:1085-1111"""
    phi = np.zeros((3, nx, ny, nz), np.float32)
    factor = np.float32(2.2 / delta)
    x = np.arange(nx, dtype=np.float64).reshape(nx, 1, 1)
    z = np.arange(nz, dtype=np.float64).reshape(1, 1, nz)
    base = seed_height + waviness * np.cos(2.0 * math.pi * x / max(nx, 1))
    ls = (0.5 * (1.0 - np.tanh(factor * ((z - (base + left_offset)) * dx)))).astype(np.float32)
    rs = (0.5 * (1.0 - np.tanh(factor * ((z - (base + right_offset)) * dx)))).astype(np.float32)
    phi[left_gid, :split_index] = np.broadcast_to(ls[:split_index], (split_index, ny, nz))
    phi[right_gid, split_index:] = np.broadcast_to(rs[split_index:], (nx - split_index, ny, nz))
    phi[0] = np.clip(1.0 - phi[1] - phi[2], 0.0, 1.0)
    return phi


def alternating_multigrain_phi_3d(nx, ny, nz, dx, delta, seed_height, ngrains, waviness=0.0):
    """:1114-1135"""
    gm = stripe_grain_map(nx, ny, ngrains)
    phi = inputs.init_phi_from_grain_map_3d(gm, ngrains, nx, ny, nz, dx, delta, int(seed_height))
    if waviness == 0.0:
        return phi
    factor = np.float32(2.2 / delta)
    x = np.arange(nx, dtype=np.float64).reshape(nx, 1, 1)
    z = np.arange(nz, dtype=np.float64).reshape(1, 1, nz)
    base = seed_height + waviness * np.cos(2.0 * math.pi * x / max(nx, 1))
    prof = (0.5 * (1.0 - np.tanh(factor * ((z - base) * dx)))).astype(np.float32)
    for g in range(1, ngrains + 1):
        phi[g] = np.where((gm == g)[:, :, None], prof, 0.0)
    phi[0] = np.clip(1.0 - phi[1:].sum(axis=0), 0.0, 1.0)
    return phi


def orientation_library():
    """:1138-1149"""
    from scipy.spatial.transform import Rotation
    return {"identity": Rotation.identity(), "y_15": Rotation.from_euler("y", 15.0, degrees=True),
            "y_30": Rotation.from_euler("y", 30.0, degrees=True), "y_45": Rotation.from_euler("y", 45.0, degrees=True),
            "x_30": Rotation.from_euler("x", 30.0, degrees=True),
            "tilt_xyz": Rotation.from_euler("xyz", [20.0, 35.0, 10.0], degrees=True),
            "aligned_111_to_z": _rotation("aligned_111")}


def bicrystal_group_metrics(phi, fa, fb, dx):
    """:1366-1381"""
    vols = np.array([float(phi[g].sum() * (dx ** 3)) for g in range(1, phi.shape[0])])
    va = float(np.sum([vols[g - 1] for g in fa]))
    vb = float(np.sum([vols[g - 1] for g in fb]))
    if va >= vb:
        return {"volume_a": va, "volume_b": vb, "winner_loser_volume_ratio": float(va / max(vb, 1e-30)),
                "winner_family": "A"}
    return {"volume_a": va, "volume_b": vb, "winner_loser_volume_ratio": float(vb / max(va, 1e-30)),
            "winner_family": "B"}


# ------------------------------------------------------------ directional scans
_SCAN_CACHE: dict = {}


def build_directional_scan_case(base, delta_a, ksi, nsteps):
    """:1157-1163"""
    return replace(base, nx=16, ny=6, nz=20, nsteps=nsteps, delta_a=delta_a, ksi=ksi, p_round=0.02, G=0.0,
                   V_pulling=0.0)


def main_directional_params(base):
    """main_directional_low/high_params (:1787-1794)"""
    return replace(build_directional_scan_case(base, 0.5, 0.05, 8), nx=16, ny=6, nz=24)


def single_grain_benchmark_phi(p):
    """:1166-1168"""
    d = inputs.derive_constants(p)
    return wavy_single_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 6.5, 1.8, phase_shift=0.3)


def run_directional_scan(case_dir, p, undercoolings, anisotropy, torque, include_iso=False, prefix=None):
    """:1171-1288: every library orientation at every undercooling (plus the
    isotropic baseline), front velocity per run; rows cached per signature."""
    key = (tuple(sorted(p.items())), tuple(sorted(undercoolings.items())), anisotropy, torque, include_iso)
    if key not in _SCAN_CACHE:
        rows = []
        phi0 = single_grain_benchmark_phi(p)
        lib = orientation_library()
        for lab, dT in undercoolings.items():
            temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - dT)
            runs = [(lab, p, anisotropy, torque)]
            if include_iso:
                runs.append((f"{lab}_iso", replace(p, delta_a=0.0, ksi=1.0), False, False))
            for rlab, rp, a, t in runs:
                for name, rot in lib.items():
                    r = run_3d_simulation(f"scan_{rlab}_{name}", rp, phi0, temp0, quaternions_from_rotations([rot]), a, t)
                    m = finalize_metrics(dict(r["metrics"]), phi0, r["phi"], p["dx"], p["dt"] * p["nsteps"])
                    rows.append({"orientation": name, "driving_label": rlab, "undercooling": dT,
                                 "interface_mean": m["interface_position_mean"],
                                 "interface_max": m["interface_position_max"],
                                 "solid_volume_growth": m["solid_volume_growth"], "front_velocity": m["front_velocity"],
                                 "runtime_s": m["runtime_s"]})
        _SCAN_CACHE[key] = {"rows": rows}
    rows = _SCAN_CACHE[key]["rows"]
    pre = f"{prefix}_" if prefix else ""
    with open(os.path.join(case_dir, f"{pre}directional_scan.csv"), "w", encoding="utf-8", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    return _SCAN_CACHE[key]


def directional_stats(rows, label):
    """:1291-1303"""
    sub = [r for r in rows if r["driving_label"] == label]
    if not sub:
        raise RuntimeError(f"No directional-scan rows were found for driving label '{label}'.")
    best = max(sub, key=lambda r: r["front_velocity"])
    worst = min(sub, key=lambda r: r["front_velocity"])
    return {"rows": sub, "best": best, "worst": worst, "spread": float(best["front_velocity"] - worst["front_velocity"])}


def directional_row(rows, label, name):
    """:1306-1316"""
    for r in rows:
        if r["driving_label"] == label and r["orientation"] == name:
            return r
    raise RuntimeError(f"No directional-scan row for '{label}' and '{name}'.")


def groove_observer(split_index):
    """:1877-1887"""
    def obs(phi, step):
        g = robust_groove_metrics(phi, float(split_index))
        return {"step": int(step), "dihedral_angle_like_deg": g["dihedral_angle_like_deg"],
                "trijunction_x_cells": g["trijunction_x_cells"], "trijunction_z_cells": g["trijunction_z_cells"]}
    return obs


def run_bicrystal_competition(label, p, left, right, undercooling, anisotropy, torque):
    """:1951-2013"""
    d = inputs.derive_constants(p)
    split = p["nx"] // 2
    phi0 = bicrystal_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 7.0, split, waviness=0.8)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - undercooling)
    lib = orientation_library()
    quats = quaternions_from_rotations([lib[left], lib[right]])
    root0 = grain_boundary_root_and_trijunction(phi0, 1, 2, x_guess=split)
    r = run_3d_simulation(label, p, phi0, temp0, quats, anisotropy, torque, sample_every=max(2, p["nsteps"] // 4),
                          observer=groove_observer(split))
    groove = robust_groove_metrics(r["phi"], float(split))
    hist = [smp.get("dihedral_angle_like_deg", float("nan")) for smp in r["samples"]]
    m = finalize_metrics(dict(r["metrics"]), phi0, r["phi"], p["dx"], p["dt"] * p["nsteps"],
                         root_initial_cells=root0["root_x_cells"],
                         trijunction_final=grain_boundary_root_and_trijunction(r["phi"], 1, 2, x_guess=split),
                         groove_metrics=groove, angle_history_deg=hist)
    fam = bicrystal_group_metrics(r["phi"], [1], [2], p["dx"])
    m["winner_loser_volume_ratio"] = fam["winner_loser_volume_ratio"]
    win_a = fam["winner_family"] == "A"
    return {"phi0": phi0, "run": r, "metrics": m, "winner_gid": 1 if win_a else 2,
            "winner_orientation": left if win_a else right, "loser_orientation": right if win_a else left,
            "family_metrics": fam, "root_initial_cells": root0["root_x_cells"], "split_index": split}


def run_multigrain_family_competition(label, p, fa_name, fb_name, undercooling, ngrains, waviness, anisotropy,
                                      torque):
    """:1890-1948"""
    d = inputs.derive_constants(p)
    phi0 = alternating_multigrain_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 7.0, ngrains, waviness)
    fa = list(range(1, ngrains + 1, 2))
    fb = list(range(2, ngrains + 1, 2))
    lib = orientation_library()
    quats = np.zeros((ngrains + 1, 4))
    quats[0] = (0.0, 0.0, 0.0, 1.0)
    for g in fa:
        quats[g] = lib[fa_name].as_quat()
    for g in fb:
        quats[g] = lib[fb_name].as_quat()
    r = run_3d_simulation(label, p, phi0, constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - undercooling),
                          quats, anisotropy, torque)
    m = finalize_metrics(dict(r["metrics"]), phi0, r["phi"], p["dx"], p["dt"] * p["nsteps"])
    g = bicrystal_group_metrics(r["phi"], fa, fb, p["dx"])
    m["winner_loser_volume_ratio"] = g["winner_loser_volume_ratio"]
    return {"phi0": phi0, "run": r, "metrics": m, "group": g,
            "winner_orientation": fa_name if g["winner_family"] == "A" else fb_name, "ngrains": ngrains}


SCHEDULE = {"low": 0.25, "high": 3.0}          # main_directional_schedule (:1869-1870)
PAIR = ("aligned_111_to_z", "identity")       # main_bicrystal_pair (:1873-1874)


def _bicrystal_params(base, nsteps, delta_a):
    """main_bicrystal_switch/energy/kinetic_params (:1797-1835)"""
    return replace(base, nx=18, ny=6, nz=28, nsteps=nsteps, delta_a=delta_a, ksi=0.05 if delta_a == 0.5 else 0.01,
                   p_round=0.02, G=0.0, V_pulling=0.0)


def case_single_grain_preferred_growth_benchmark(case_dir, base, ov):
    """:2016-2067"""
    a, t = resolve_toggles(True, False, ov)
    rows = run_directional_scan(case_dir, main_directional_params(base), {"high": SCHEDULE["high"]}, a, t, True,
                                "benchmark")["rows"]
    on, off = directional_stats(rows, "high"), directional_stats(rows, "high_iso")
    al = directional_row(rows, "high", "aligned_111_to_z")
    pen = float(on["best"]["front_velocity"] - al["front_velocity"])
    passed = on["spread"] >= 5.0 * max(off["spread"], 1e-12) and pen >= 5e-2
    m = blank_metrics()
    m.update(interface_position_mean=float(on["best"]["interface_mean"]),
             interface_position_max=float(on["best"]["interface_max"]),
             solid_volume_growth=float(on["best"]["solid_volume_growth"]),
             front_velocity=float(on["best"]["front_velocity"]))
    return make_case_result(
        "single_grain_preferred_growth_benchmark", passed,
        f"Single-grain preferred growth spread_on={on['spread']:.2e} m/s vs isotropic spread_off={off['spread']:.2e} m/s; "
        f"best={on['best']['orientation']}, worst={on['worst']['orientation']}, aligned penalty={pen:.2e} m/s.",
        "" if passed else "Preferred growth did not separate enough from the isotropic baseline.", m,
        {"spread_on": on["spread"], "spread_off": off["spread"], "best_orientation": on["best"]["orientation"],
         "worst_orientation": on["worst"]["orientation"], "best_front_velocity": on["best"]["front_velocity"],
         "worst_front_velocity": on["worst"]["front_velocity"], "aligned_front_velocity": al["front_velocity"],
         "aligned_penalty": pen}, case_dir, "main", classify_signal_level(on["spread"], 3e-2, 8e-2, passed))


def case_directional_preference_map(case_dir, base, ov):
    """:2070-2145"""
    a, t = resolve_toggles(True, False, ov)
    lo = run_directional_scan(case_dir, main_directional_params(base), {"low": SCHEDULE["low"]}, a, t, True, "low")["rows"]
    hi = run_directional_scan(case_dir, main_directional_params(base), {"high": SCHEDULE["high"]}, a, t, True,
                              "high")["rows"]
    ls, hs = directional_stats(lo, "low"), directional_stats(hi, "high")
    lis, his = directional_stats(lo, "low_iso"), directional_stats(hi, "high_iso")
    pl = float(ls["best"]["front_velocity"] - directional_row(lo, "low", "aligned_111_to_z")["front_velocity"])
    ph = float(hs["best"]["front_velocity"] - directional_row(hi, "high", "aligned_111_to_z")["front_velocity"])
    passed = (ls["spread"] >= 1.5e-2 and hs["spread"] >= 5e-2 and hs["spread"] > ls["spread"]
              and lis["spread"] <= 1e-10 and his["spread"] <= 1e-10 and ph > pl + 2e-2)
    m = blank_metrics()
    m.update(interface_position_mean=float(hs["best"]["interface_mean"]),
             interface_position_max=float(hs["best"]["interface_max"]),
             solid_volume_growth=float(hs["best"]["solid_volume_growth"]),
             front_velocity=float(hs["best"]["front_velocity"]))

    def part(st, iso, pen):
        return {"best_orientation": st["best"]["orientation"], "worst_orientation": st["worst"]["orientation"],
                "spread": st["spread"], "iso_spread": iso["spread"], "aligned_penalty": pen}

    return make_case_result(
        "directional_preference_map", passed,
        f"Directional map spreads: low={ls['spread']:.2e}, high={hs['spread']:.2e}; aligned penalty grew "
        f"{pl:.2e} -> {ph:.2e} m/s.",
        "" if passed else "Directional preference map did not show a clear anisotropic spread.", m,
        {"low": part(ls, lis, pl), "high": part(hs, his, ph)}, case_dir, "main",
        classify_signal_level(hs["spread"], 4e-2, 8e-2, passed))


def case_anisotropy_threshold_test(case_dir, base, ov):
    """:2148-2198"""
    a, _ = resolve_toggles(True, False, ov)
    rows = []
    for da in (0.0, 0.1, 0.2, 0.35, 0.5):
        p = replace(build_directional_scan_case(base, da, 0.05, 8), nx=16, ny=6, nz=24)
        phi0 = single_grain_benchmark_phi(p)
        temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"] - 3.0)
        ra = run_3d_simulation(f"thr_id_{da:.2f}", p, phi0, temp0,
                               quaternions_from_rotations([_rotation("identity")]), a and da > 0.0, False)
        rb = run_3d_simulation(f"thr_al_{da:.2f}", p, phi0, temp0,
                               quaternions_from_rotations([_rotation("aligned_111")]), a and da > 0.0, False)
        ma = finalize_metrics(dict(ra["metrics"]), phi0, ra["phi"], p["dx"], p["dt"] * p["nsteps"])
        mb = finalize_metrics(dict(rb["metrics"]), phi0, rb["phi"], p["dx"], p["dt"] * p["nsteps"])
        rows.append({"delta_a": da, "front_velocity_diff": abs(ma["front_velocity"] - mb["front_velocity"])})
    with open(os.path.join(case_dir, "anisotropy_threshold.csv"), "w", encoding="utf-8", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
        w.writeheader()
        w.writerows(rows)
    mono = all(rows[i + 1]["front_velocity_diff"] >= rows[i]["front_velocity_diff"] - 2e-2 for i in range(len(rows) - 1))
    passed = rows[0]["front_velocity_diff"] <= 5e-4 and rows[-1]["front_velocity_diff"] >= 7e-2 and mono
    m = blank_metrics()
    m["front_velocity"] = float(rows[-1]["front_velocity_diff"])
    return make_case_result(
        "anisotropy_threshold_test", passed,
        f"Directional contrast grew from {rows[0]['front_velocity_diff']:.2e} to {rows[-1]['front_velocity_diff']:.2e} "
        "as delta_a increased.",
        "" if passed else "Directional contrast did not grow cleanly with anisotropy strength.", m,
        {"rows": rows, "monotonic_non_decreasing": mono}, case_dir, "main",
        classify_signal_level(rows[-1]["front_velocity_diff"], 7e-2, 9e-2, passed))


def case_bicrystal_competition_low_high_driving(case_dir, base, ov):
    """:2201-2257"""
    p = _bicrystal_params(base, 8, 0.5)
    a, t = resolve_toggles(True, False, ov)
    lo_o, hi_o = PAIR
    lo = run_bicrystal_competition("bicrystal_low", p, lo_o, hi_o, SCHEDULE["low"], a, t)
    hi = run_bicrystal_competition("bicrystal_high", p, lo_o, hi_o, SCHEDULE["high"], a, t)
    switched = lo["winner_orientation"] == lo_o and hi["winner_orientation"] == hi_o
    sig = min(lo["metrics"]["winner_loser_volume_ratio"], hi["metrics"]["winner_loser_volume_ratio"])
    passed = switched and sig >= 1.01
    return make_case_result(
        "bicrystal_competition_low_high_driving", passed,
        f"Bicrystal winner switched {lo['winner_orientation']} -> {hi['winner_orientation']} between low/high driving.",
        "" if passed else "Low/high driving did not produce the expected bicrystal winner switch.", dict(hi["metrics"]),
        {"pair": {"low_orientation": lo_o, "high_orientation": hi_o}, "low_winner": lo["winner_orientation"],
         "high_winner": hi["winner_orientation"], "low_ratio": lo["metrics"]["winner_loser_volume_ratio"],
         "high_ratio": hi["metrics"]["winner_loser_volume_ratio"],
         "low_root_shift": lo["metrics"]["grain_boundary_root_shift"],
         "high_root_shift": hi["metrics"]["grain_boundary_root_shift"]}, case_dir, "main",
        classify_signal_level(sig, 1.01, 1.015, passed))


def case_bicrystal_interfacial_energy_dominated_regime(case_dir, base, ov):
    """:2260-2323"""
    p = _bicrystal_params(base, 14, 0.2)
    a, t = resolve_toggles(True, False, ov)
    al, ident = PAIR
    on = run_bicrystal_competition("bicrystal_low_on", p, al, ident, 0.01, a, t)
    off = run_bicrystal_competition("bicrystal_low_off", replace(p, delta_a=0.0, ksi=1.0), al, ident, 0.01, False, False)
    rg = on["metrics"]["winner_loser_volume_ratio"] - off["metrics"]["winner_loser_volume_ratio"]
    vg = on["metrics"]["front_velocity"] - off["metrics"]["front_velocity"]
    sg = abs(on["metrics"]["grain_boundary_root_shift"]) - abs(off["metrics"]["grain_boundary_root_shift"])
    passed = on["winner_orientation"] == al and vg > 5e-3 and rg > 3e-3 and sg > 5e-7
    return make_case_result(
        "bicrystal_interfacial_energy_dominated_regime_test", passed,
        f"Low-driving bicrystal favored {on['winner_orientation']} with front-velocity gain {vg:.2e} m/s and ratio gain "
        f"{rg:.2e} over anisotropy-off.",
        "" if passed else "Low-driving bicrystal did not show a clean aligned-facet preference over the anisotropy-off baseline.",
        dict(on["metrics"]),
        {"expected_low_winner": al, "winner_on": on["winner_orientation"], "winner_off": off["winner_orientation"],
         "ratio_on": on["metrics"]["winner_loser_volume_ratio"], "ratio_off": off["metrics"]["winner_loser_volume_ratio"],
         "front_velocity_on": on["metrics"]["front_velocity"], "front_velocity_off": off["metrics"]["front_velocity"],
         "front_velocity_gain": vg, "root_shift_on": on["metrics"]["grain_boundary_root_shift"],
         "root_shift_off": off["metrics"]["grain_boundary_root_shift"], "root_shift_gain": sg}, case_dir, "main",
        classify_signal_level(vg, 5e-3, 1.2e-2, passed))


def case_bicrystal_kinetic_dominated_regime(case_dir, base, ov):
    """:2326-2385"""
    p = _bicrystal_params(base, 10, 0.35)
    a, t = resolve_toggles(True, False, ov)
    al, ident = PAIR
    on = run_bicrystal_competition("bicrystal_high_on", p, al, ident, 6.0, a, t)
    off = run_bicrystal_competition("bicrystal_high_off", replace(p, delta_a=0.0, ksi=1.0), al, ident, 6.0, False, False)
    rg = on["metrics"]["winner_loser_volume_ratio"] - off["metrics"]["winner_loser_volume_ratio"]
    sw = on["metrics"]["grain_boundary_root_shift"] * off["metrics"]["grain_boundary_root_shift"] < 0.0
    passed = on["winner_orientation"] == ident and off["winner_orientation"] == al and rg > 0.02 and sw
    return make_case_result(
        "bicrystal_kinetic_dominated_regime_test", passed,
        f"High-driving bicrystal favored {on['winner_orientation']} with ratio gain {rg:.2e} over the anisotropy-off "
        "baseline.",
        "" if passed else "High-driving bicrystal did not show the expected kinetic-dominated winner.",
        dict(on["metrics"]),
        {"expected_high_winner": ident, "winner_on": on["winner_orientation"], "winner_off": off["winner_orientation"],
         "ratio_on": on["metrics"]["winner_loser_volume_ratio"], "ratio_off": off["metrics"]["winner_loser_volume_ratio"],
         "root_shift_on": on["metrics"]["grain_boundary_root_shift"],
         "root_shift_off": off["metrics"]["grain_boundary_root_shift"], "root_switch": sw}, case_dir, "main",
        classify_signal_level(rg, 0.02, 0.05, passed))


def case_multigrain_competition_extension(case_dir, base, ov):
    """:2388-2450"""
    a, t = resolve_toggles(True, False, ov)
    al, ident = PAIR
    lo = run_multigrain_family_competition(
        "multigrain_low", replace(base, nx=24, ny=10, nz=30, nsteps=22, delta_a=0.2, ksi=0.01, p_round=0.02, G=0.0,
                                  V_pulling=0.0), al, ident, 0.01, 6, 1.5, a, t)
    hi = run_multigrain_family_competition(
        "multigrain_high", replace(base, nx=24, ny=10, nz=28, nsteps=10, delta_a=0.35, ksi=0.01, p_round=0.02, G=0.0,
                                   V_pulling=0.0), al, ident, 6.0, 6, 1.0, a, t)
    lr, hr = lo["group"]["winner_loser_volume_ratio"], hi["group"]["winner_loser_volume_ratio"]
    passed = (lo["winner_orientation"] == al and hi["winner_orientation"] == ident and hr > 1.05
              and (hr - 1.0) > 10.0 * max(lr - 1.0, 1e-12))
    return make_case_result(
        "multigrain_competition_extension", passed,
        f"Multigrain extension switched dominant family {lo['winner_orientation']} -> {hi['winner_orientation']}; "
        f"winner ratios were {lr:.4f} and {hr:.4f}.",
        "" if passed else "Multigrain extension did not preserve the bicrystal family switch trend strongly enough.",
        dict(hi["metrics"]),
        {"low_family_name": lo["winner_orientation"], "high_family_name": hi["winner_orientation"],
         "low_group": lo["group"], "high_group": hi["group"],
         "low_surviving_grains": lo["metrics"]["surviving_grains"],
         "high_surviving_grains": hi["metrics"]["surviving_grains"]}, case_dir, "main",
        classify_signal_level(min(lr, hr), 1.005, 1.02, passed))


def _groove_setup(base):
    p = replace(base, nx=20, ny=6, nz=24, nsteps=20, delta_a=0.0, ksi=1.0, G=0.0, V_pulling=0.0)
    d = inputs.derive_constants(p)
    split = p["nx"] // 2
    phi0 = bicrystal_phi_3d(p["nx"], p["ny"], p["nz"], p["dx"], d["delta"], 7.0, split, waviness=0.6)
    temp0 = constant_temp((p["nx"], p["ny"], p["nz"]), p["T_melt"])
    q = quaternions_from_rotations([_rotation("identity"), _rotation("identity")])
    return p, split, phi0, temp0, q, grain_boundary_root_and_trijunction(phi0, 1, 2, x_guess=split)


def case_groove_depth_only(case_dir, base, ov):
    """:2453-2488"""
    p, split, phi0, temp0, q, root0 = _groove_setup(base)
    r = run_3d_simulation("groove_depth_only", p, phi0, temp0, q, False, False)
    g = robust_groove_metrics(r["phi"], float(split))
    m = finalize_metrics(dict(r["metrics"]), phi0, r["phi"], p["dx"], p["dt"] * p["nsteps"],
                         root_initial_cells=root0["root_x_cells"], groove_metrics=g)
    depth = g["groove_depth_cells"] * p["dx"]
    passed = bool(np.isfinite(depth) and depth >= 0.10 * p["dx"])
    return make_case_result(
        "groove_depth_only", passed, f"Groove depth benchmark gave depth={depth:.2e} m.",
        "" if passed else "Groove depth remained too small for a reliable capillarity benchmark.", m,
        {"groove_depth": depth, "groove_metrics": g}, case_dir, "main",
        classify_signal_level(depth / p["dx"] if np.isfinite(depth) else float("nan"), 0.10, 0.25, passed))


def case_robust_groove_angle_estimation(case_dir, base, ov):
    """:2491-2542"""
    p, split, phi0, temp0, q, root0 = _groove_setup(base)
    r = run_3d_simulation("robust_groove_angle", p, phi0, temp0, q, False, False, sample_every=2,
                          observer=groove_observer(split))
    g = robust_groove_metrics(r["phi"], float(split))
    hist = [smp.get("dihedral_angle_like_deg", float("nan")) for smp in r["samples"]]
    m = finalize_metrics(dict(r["metrics"]), phi0, r["phi"], p["dx"], p["dt"] * p["nsteps"],
                         root_initial_cells=root0["root_x_cells"], groove_metrics=g, angle_history_deg=hist)
    rmse = max(g["left_fit_rmse"], g["right_fit_rmse"])
    passed = bool(np.isfinite(m["interface_angle_deg"]) and rmse <= 0.6)
    return make_case_result(
        "robust_groove_angle_estimation", passed,
        f"Robust groove fit produced dihedral-like angle={m['interface_angle_deg']:.2f} deg with fit_rmse={rmse:.2e} cells.",
        "" if passed else "Robust groove-angle fit remained numerically unstable.", m,
        {"groove_metrics": g, "angle_history_deg": hist, "fit_rmse": rmse}, case_dir, "main",
        classify_signal_level(m["interface_angle_deg"], 2.0, 5.0, passed))

CASES = {
    "basic_constraints": case_basic_constraints,
    "kmax_overflow_detection": case_kmax_overflow_detection,
    "static_flat_interface": case_static_flat_interface,
    "two_d_limit_consistency": case_two_d_limit_consistency,
    "boundary_conditions": case_boundary_conditions,
    "isotropic_orientation_independence": case_isotropic_orientation_independence,
    "anisotropic_preferred_growth": case_anisotropic_preferred_growth,
    "torque_term_contribution": case_torque_term_contribution,
    "grain_competition": case_grain_competition,
    "grain_boundary_groove_trijunction": case_grain_boundary_groove,
    "convergence_grid_dt_delta": case_convergence_sweep,
    "single_grain_preferred_growth_benchmark": case_single_grain_preferred_growth_benchmark,
    "directional_preference_map": case_directional_preference_map,
    "anisotropy_threshold_test": case_anisotropy_threshold_test,
    "bicrystal_competition_low_high_driving": case_bicrystal_competition_low_high_driving,
    "bicrystal_interfacial_energy_dominated_regime_test": case_bicrystal_interfacial_energy_dominated_regime,
    "bicrystal_kinetic_dominated_regime_test": case_bicrystal_kinetic_dominated_regime,
    "multigrain_competition_extension": case_multigrain_competition_extension,
    "groove_depth_only": case_groove_depth_only,
    "robust_groove_angle_estimation": case_robust_groove_angle_estimation,
}


def run_cases(out_dir, names=None, overrides=None, config_path=None, log=None) -> list[CaseOutcome]:
    """Run the named cases (all by default); a case that raises becomes a
    failed case (:2605-2630)."""
    base = base_params(config_path)
    ov = overrides or {}
    out = []
    for name in (names or list(CASES)):
        cd = os.path.join(out_dir, name)
        os.makedirs(cd, exist_ok=True)
        try:
            res = CASES[name](cd, base, ov)
        except Exception as e:   # the reference records crashes as failed cases
            res = make_case_result(name, False, f"{name} raised {type(e).__name__}.", str(e), {}, {}, cd)
        if log:
            log(f"{'PASS' if res.passed else 'FAIL'} {name}: {res.summary}")
        out.append(res)
    return out


def write_report(out_dir, results, notes=()) -> str:
    """verification_report.md: a summary table and one section per case."""
    path = os.path.join(out_dir, "verification_report.md")
    n_pass = sum(r.passed for r in results)
    lines = ["# Verification report", "", f"- Generated: {datetime.datetime.now():%Y-%m-%d %H:%M:%S}",
             f"- Passed: {n_pass} / {len(results)}", ""] + [f"- Note: {n}" for n in notes]
    lines += ["", "| case | category | status | level | summary |", "|---|---|---|---|---|"]
    lines += [f"| {r.name} | {r.category} | {'PASS' if r.passed else 'FAIL'} | {r.level} | {r.summary} |"
              for r in results]
    for r in results:
        lines += ["", f"## {r.name}", "", f"- Status: {'PASS' if r.passed else 'FAIL'}", f"- Summary: {r.summary}"]
        if r.failure_reason:
            lines.append(f"- Failure reason: {r.failure_reason}")
        for k, v in r.extra.items():
            lines.append(f"- {k}: `{json.dumps(_ser(v), ensure_ascii=False)}`")
    with open(path, "w", encoding="utf-8") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(out_dir, "verification_summary.json"), "w", encoding="utf-8") as f:
        json.dump(_ser([{"name": r.name, "passed": r.passed, "category": r.category, "level": r.level,
                         "summary": r.summary, "failure_reason": r.failure_reason} for r in results]), f, indent=2)
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description="Verification cases of the 3D solver.")
    ap.add_argument("--out-dir", default="result/verification_3d")
    ap.add_argument("--cases", nargs="*", default=None, choices=sorted(CASES))
    ap.add_argument("--config", default=None)
    ap.add_argument("--anisotropy", type=int, default=None, help="override the cases' anisotropy toggle (0/1)")
    ap.add_argument("--torque", type=int, default=None, help="override the cases' torque toggle (0/1)")
    args = ap.parse_args(argv)
    res = run_cases(args.out_dir, args.cases, {"anisotropy": args.anisotropy, "torque": args.torque}, args.config,
                    log=lambda s: print(s, flush=True))
    print(write_report(args.out_dir, res))
    return 0 if all(r.passed for r in res) else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
