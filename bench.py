#!/usr/bin/env python3
"""Headline benchmark: 3D random-mode plain method at 512 x 512 x 256 with 200
Voronoi grains in a 24-cell seed layer, random orientations, uniform cooling,
KMAX = 50 (config_3d.yaml physics), a 5000-step run.  One "step" is one full
time step (temperature, APT, RHS and update) over the whole grid.

The run is advanced to its middle (step 2500, timed in segments), then W
warm-up steps, the launch-shape autotuner (state restored) and the K timed
steps (value), then the rest of the run.  The FP64 run does the same, and at
the end of both runs the FP32 state is compared with the FP64 state on the
device.

Prints ONE JSON line: MLUPS of the FP32 solver (value) at the mid-run state,
the mean over the whole run (trajectory), the FP64 solver's rates, FP32 vs
FP64 at the end of the run (and of the 128^3 2000-step case), the HBM
roofline of the step, its vector-ALU issue, and a CPU baseline (the CPU
oracle in the reference's dense layout on a bounded slab of the workload).

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --summarize-trace <rocprofv3 csv> --steps K
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "multiphasefield-solidification-v2_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

from mpf import Solver, inputs  # noqa: E402
from mpf.config import flat_params, load_config  # noqa: E402

METRIC = "MLUPS at 512²×256 (FP32 vs FP64 plain) + surrogate speedup at ≤1-cell front err"
HBM_PEAK_GBS = 8000.0
NX, NY, NZ, N_SOLID, SEED_H = 512, 512, 256, 200, 24
RUN_STEPS = 5000


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(nx=NX, ny=NY, nz=NZ, n_solid=N_SOLID, seed_height=SEED_H):
    p = flat_params(load_config())
    d = inputs.derive_constants(p)
    gmap = inputs.generate_random_grain_map_3d(nx, ny, n_solid, random_seed=42)
    quats = inputs.assign_quaternions_to_grains(n_solid, "random", 42)
    n111 = inputs.compute_rotated_n111(quats)
    w, a, m = inputs.build_interaction_matrices(n_solid + 1, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"],
                                                d["eps_gb"], d["w_gb"], d["m_gb_phi"])
    cols = inputs.columns_from_grain_map_3d(gmap, nz, p["dz"], d["delta"], seed_height)
    temp = inputs.init_temperature_field_3d(1, 1, nz, p["T_melt"], p["G"], p["dz"], seed_height)[0, 0]
    return {"p": p, "d": d, "gmap": gmap, "tables": (w, a, m, n111), "cols": cols,
            "temp": temp.astype(np.float64), "scalars": inputs.kernel_scalars(p, d, as_float32=True),
            "cooling": float(np.float32(d["cooling_rate"])), "N": n_solid + 1}


def new_solver(W, precision):
    s = Solver((NX, NY, NZ), W["N"], W["scalars"], precision=precision, kmax=50, cooling_rate=W["cooling"])
    s.set_tables(*W["tables"])
    s.init_columns(*W["cols"])
    s.upload_temp(W["temp"])
    return s


def segment(s, n, log_name=None):
    """Advance n steps; the GPU time (HIP events) and the cell counts of
    the passes over them."""
    s.reset_stats()
    done = s.step(n)
    st = s.stats()
    rec = {"steps": done, "ms": st["last_step_ms"], "light_cells": st["light_cells"],
           "merge_cells": st["merge_cells"], "front_cells": st["front_cells"],
           "general_cells": st["general_cells"], "max_active": st["max_active"],
           "kmax_overflow": st["kmax_overflow"]}
    if log_name:
        log(f"{log_name}: {n} steps, {st['last_step_ms'] / max(done, 1):.3f} ms/step")
    return rec


def run_trajectory(W, precision, mid, warmup, steps, total, kernel_steps=0, graphs=False, autotune=0,
                   seg=500):
    """One run of the workload: `mid` steps in segments of `seg`, `warmup`
    steps, the launch-shape autotuner (state restored), `steps` timed steps
    (HIP events on the solver's stream, inputs resident), optionally
    `kernel_steps` steps with every kernel bracketed by HIP events, then the
    rest of the `total`-step run in segments.  Every step of the run is
    timed; the solver is returned open at step `total`."""
    s = new_solver(W, precision)
    segs = []
    done = 0
    tuned0 = None
    if autotune and mid > 0:
        # the first steps at the default launch shape, then a shape tuned on
        # the early state for the first half of the run
        r = segment(s, min(5, mid))
        r["from"] = 0
        segs.append(r)
        done = r["steps"]
        tuned0 = s.autotune(autotune)
    while done < mid:
        n = min(seg - done % seg, mid - done)
        r = segment(s, n)
        r["from"] = done
        segs.append(r)
        done += n
        ob = s.observables(dx=W["p"]["dx"])
        r["front_mean"] = ob["front_mean"]
        r["stored_slots"] = ob["stored_slots"]
        if done % seg == 0 or done == mid:
            log(f"{precision}: step {done}: {r['ms'] / n:.3f} ms/step, front {ob['front_mean']:.1f}, "
                f"merge {r['merge_cells'] // n}, light {r['light_cells'] // n}")
    if warmup:
        r = segment(s, warmup)
        r["from"] = done
        segs.append(r)
        done += warmup
    tuned = s.autotune(autotune) if autotune else None
    if graphs:
        s.set_graphs(True)
        s.step(2)   # capture both state parities outside the timed region
        done += 2
    ob0 = s.observables(dx=W["p"]["dx"])
    t0 = time.perf_counter()
    r = segment(s, steps)
    wall = time.perf_counter() - t0
    r["from"] = done
    segs.append(r)
    timed = dict(r, wall=wall, start_step=done)
    done += steps
    ob1 = s.observables(dx=W["p"]["dx"])
    kern = None
    if kernel_steps:
        s.set_kernel_timing(True)
        s.reset_stats()
        s.step(kernel_steps)
        ks = s.stats()
        kern = {k: v / ks["kernel_steps"] for k, v in ks["kernel_ms"].items()}
        kern["step"] = ks["last_step_ms"] / ks["kernel_steps"]
        s.set_kernel_timing(False)
        segs.append({"from": done, "steps": kernel_steps, "ms": ks["last_step_ms"], "kernel_events": True})
        done += kernel_steps
    while done < total:
        n = min(seg - done % seg, total - done)
        r = segment(s, n)
        r["from"] = done
        segs.append(r)
        done += n
        if done % seg == 0 or done == total:
            ob = s.observables(dx=W["p"]["dx"])
            r["front_mean"] = ob["front_mean"]
            r["stored_slots"] = ob["stored_slots"]
            log(f"{precision}: step {done}: {r['ms'] / n:.3f} ms/step, front {ob['front_mean']:.1f}")
    return {"solver": s, "segments": segs, "timed": timed, "ob0": ob0, "ob1": ob1, "kernel_ms": kern,
            "tuned": tuned, "tuned_early": tuned0, "launch": s.launch(), "bytes": s.device_bytes,
            "steps_done": done}


def trajectory_summary(run, ncell):
    segs = run["segments"]
    steps = sum(r["steps"] for r in segs)
    ms = sum(r["ms"] for r in segs)
    rows = []
    for r in segs:
        if r.get("kernel_events"):
            continue
        rows.append({"from": r["from"], "steps": r["steps"], "ms_per_step": round(r["ms"] / r["steps"], 4),
                     "front_mean": round(r["front_mean"], 2) if "front_mean" in r else None,
                     "light_cells_per_step": r["light_cells"] // r["steps"],
                     "merge_cells_per_step": r["merge_cells"] // r["steps"],
                     "front_cells_per_step": r["front_cells"] // r["steps"],
                     "general_cells_per_step": r["general_cells"] // r["steps"]})
    return {"steps": steps, "mean_ms_per_step": round(ms / steps, 4),
            "mlups": round(ncell * steps / (ms * 1e-3) / 1e6, 1), "segments": rows}


def fp32_vs_fp64(s32, s64, dx, ncell):
    """The FP32 state against the FP64 state after the same steps, reduced
    on the device (grain slots matched by id), with the observables the
    reference reports: front heights and grain volume fractions."""
    d = s32.compare(s64, tol=1e-4)
    o32, o64 = s32.observables(dx=dx), s64.observables(dx=dx)
    v32, v64 = o32["grain_volumes"], o64["grain_volumes"]
    f32, f64 = v32 / v32.sum(), v64 / v64.sum()
    sel = f64 > 0
    return {"phi_max_abs": d["phi_max_abs"], "phi_mean_abs_per_cell": d["phi_sum_abs"] / ncell,
            "cells_above_1e-4": d["cells_above_tol"], "cells_above_1e-4_share": d["cells_above_tol"] / ncell,
            "grain_set_diff_cells": d["grain_set_diff_cells"],
            "front_mean_fp32": o32["front_mean"], "front_mean_fp64": o64["front_mean"],
            "front_mean_diff_cells": abs(o32["front_mean"] - o64["front_mean"]),
            "front_max_abs_diff_cells": float(np.nanmax(np.abs(o32["front"] - o64["front"]))),
            "grain_fraction_max_rel_diff": float(np.max(np.abs(f32[sel] - f64[sel]) / f64[sel])),
            "grain_fraction_max_abs_diff": float(np.max(np.abs(f32 - f64))),
            "grain_fraction_max_rel_diff_above_1e-3": float(np.max(np.where(f64 >= 1e-3, np.abs(f32 - f64), 0.0) /
                                                                   np.maximum(f64, 1e-3))),
            "solid_fraction_diff": abs(o32["solid_fraction"] - o64["solid_fraction"])}


def c2_precision(steps=2000):
    """The 128^3 16-grain case (seed layer 16, the config_3d.yaml physics
    with its cooling) run `steps` steps in FP32 and in FP64 on the device."""
    p = flat_params(load_config())
    d = inputs.derive_constants(p)
    n, ns, sh = 128, 16, 16
    gmap = inputs.generate_random_grain_map_3d(n, n, ns, random_seed=42)
    quats = inputs.assign_quaternions_to_grains(ns, "random", 42)
    w, a, m = inputs.build_interaction_matrices(ns + 1, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"],
                                                d["eps_gb"], d["w_gb"], d["m_gb_phi"])
    n111 = inputs.compute_rotated_n111(quats)
    cols = inputs.columns_from_grain_map_3d(gmap, n, p["dz"], d["delta"], sh)
    temp = inputs.init_temperature_field_3d(1, 1, n, p["T_melt"], p["G"], p["dz"], sh)[0, 0]
    sc = inputs.kernel_scalars(p, d, as_float32=True)
    ss, ms = [], {}
    try:
        for prec in ("fp32", "fp64"):
            s = Solver((n, n, n), ns + 1, sc, precision=prec, kmax=50, cooling_rate=float(np.float32(d["cooling_rate"])))
            ss.append(s)
            s.set_tables(w, a, m, n111)
            s.init_columns(*cols)
            s.upload_temp(temp.astype(np.float64))
            s.step(steps)
            ms[prec] = s.stats()["last_step_ms"] / steps
        out = fp32_vs_fp64(ss[0], ss[1], p["dx"], n ** 3)
    finally:
        for s in ss:
            s.close()
    out.update({"workload": f"128^3, 16 grains, seed layer 16, {steps} steps", "at_step": steps,
                "fp32_ms_per_step": round(ms["fp32"], 4), "fp64_ms_per_step": round(ms["fp64"], 4)})
    return out


def alg_bytes_per_step(ob0, ob1, bv, ncell, nz):
    """Compulsory HBM bytes of one step in the slot layout: every cell's liquid
    value and slot count read once and written once, every stored grain slot
    (id + value) read once and written once, the per-plane temperature."""
    slots = 0.5 * (ob0["stored_slots"] + ob1["stored_slots"])
    return 2.0 * ncell * (bv + 1) + 2.0 * slots * (bv + 2) + 2.0 * nz * 8, slots / ncell


def touched_bytes_per_step(light_cells, slots, bv, nz):
    """Compulsory bytes of the cells a step actually computes: the cells of
    the tile planes the classification pass computes (the other planes hold
    liquid exactly 1 with liquid 1 around them and are neither read nor
    written) read and written once, every stored grain slot (all of them lie
    in those planes) read and written once, the per-plane temperature."""
    return 2.0 * light_cells * (bv + 1) + 2.0 * slots * (bv + 2) + 2.0 * nz * 8


def measured_valu(path):
    """Vector-ALU instructions per step from the committed rocprofv3 PMC
    summary (SQ_INSTS_VALU per dispatch)."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return {"valu_wave_instr_per_step": t["valu_wave_instr_per_step"],
            "valu_time_at_full_issue_us": t["valu_time_at_full_issue_us_per_step"],
            "model": t["model"], "source": os.path.relpath(path, ROOT)}


def measured_traffic(path):
    """HBM bytes per step of the step's kernels from the committed rocprofv3
    PMC summary (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by what the
    same counters read for a device copy of known size)."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return {"bytes_per_step": t["bytes_per_step"], "source": os.path.relpath(path, ROOT)}


def batched_sweep(steps, graphs):
    """The batched bicrystal sweep: 256 grids (nx = ny in {64, 96, 128},
    nz in {96, 128, 192}, two grains split at nx/2, random quaternion pairs,
    undercooling U[0.05, 0.3), nondimensional units), FP32, all grids in one
    batch (every kernel launch covers all of them), launch shape autotuned,
    observables every 500 steps on the GPU."""
    from mpf import sweep
    cases = sweep.bicrystal_cases(256, seed=0)
    res = sweep.run_sweep(cases, steps, precision="fp32", max_batch_cells=1 << 30, sample_every=500,
                          graphs=graphs, autotune=4, log=log)
    t = res["timing"]
    b = t["batches"][0]
    wins = sum(1 for r in res["cases"] if r["winner_gid"] == 1)
    return {"workload": "256 bicrystal grids, nx=ny in {64,96,128}, nz in {96,128,192}, N=3, FP32, "
                        f"{steps} steps, one batch", "instances": t["instances"], "steps": steps,
            "cell_updates": t["cell_updates"], "instances_per_s": round(t["instances_per_s"], 3),
            "wall_s": round(t["wall_s"], 3), "gpu_s": round(t["gpu_s"], 3),
            "mlups_gpu": round(t["mlups_gpu"], 1), "mlups_wall": round(t["mlups_wall"], 1),
            "per_instance_time_s": {k: round(v, 3) for k, v in t["latency_s"].items()},
            "step_ms": {k: round(v, 4) for k, v in t["step_ms"].items()},
            "setup_s": round(b["setup_s"], 3), "autotune_s": round(b["autotune_s"], 3),
            "launch": b["launch"], "graphs": graphs, "grain1_wins": wins,
            "kmax_overflow": b["kmax_overflow"]}


BF16_DENSE_PEAK_TFLOPS = 2500.0


def surrogate_block():
    """The learned surrogate: trained here on 1920 trajectory pairs from the
    batched solver (48 grids of 64^3, 2000 steps), then 2000 steps of the
    128^3 16-grain random-mode case plain and surrogate-assisted.  The
    speedup counted against the target is the best over the tolerances
    whose run stays within 1 cell of front error and 1% grain fraction."""
    from mpf import surrogate_run as SR
    summ, _ = SR.run(grids=48, traj_steps=2000, every=50, train_steps=2500, eval_steps=2000, log=log)
    c2 = summ["c2"]
    ok = [v for k, v in c2.items() if k.startswith("surrogate") and v["front_max_err_cells"] <= 1.0
          and v["grain_fraction_max_err"] <= 0.01]
    best = max(ok, key=lambda v: v["speedup_vs_plain"]) if ok else None
    pr = summ["proposal"]
    return {
        "workload": "128^3, 16 grains, 2000 steps, FP32; proposals of 16 steps over a 128x128x32 band",
        "speedup_at_1cell_front_err": round(best["speedup_vs_plain"], 3) if best else None,
        "accepted_at_that_tolerance": best["accepted"] if best else None,
        "plain_wall_s": round(c2["plain"]["wall_s"], 4),
        "runs": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()
                     if kk in ("wall_s", "speedup_vs_plain", "accepted", "rejected", "acceptance",
                               "front_mean_err_cells", "front_max_err_cells", "grain_fraction_max_err",
                               "rate_tol", "residual_ratio_p50")}
                 for k, v in c2.items() if k != "plain"},
        "held_out_skill": round(summ["skill_trained"], 4), "train_s": round(summ["train_s"], 2),
        "train_ms_per_step": round(summ["train_ms_per_step"], 2), "pairs": summ["pairs"], "params": summ["params"],
        "roofline": {"bound": "mfma", "kernel": "network forward of one proposal (k_conv_fwd dominant)",
                     "achieved": round(pr["tflops"], 1), "peak": BF16_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(pr["tflops"] / BF16_DENSE_PEAK_TFLOPS, 4), "ms": round(pr["ms"], 3),
                     "gflop": round(pr["gflop"], 1), "dtype": "bf16 x bf16 -> fp32"},
    }


def pcie_inclusive(steps=2000):
    """The 128^3 16-grain case driven the way a caller holding the
    reference's dense host arrays drives it: phi [17][128][128][128] float32
    uploaded (host-to-device copy and conversion to the slot layout), `steps`
    FP32 steps, phi downloaded again.  The cell-update rate with the
    transfers included is reported beside the headline value, never as it."""
    from mpf import surrogate_run as SR
    make, _, _, _ = SR.c2_case()
    n = 128
    with make() as s:
        s.step(16)
        phi = s.phi(np.float32)
        s.upload_phi(phi)
        s.step(1)
        t0 = time.perf_counter()
        s.upload_phi(phi)
        t1 = time.perf_counter()
        s.step(steps)
        t2 = time.perf_counter()
        s.phi(np.float32)
        t3 = time.perf_counter()
    cells = float(n ** 3) * steps
    return {"workload": f"128^3, 16 grains, FP32, {steps} steps between one dense upload and one dense download",
            "dense_phi_bytes": int(phi.nbytes), "upload_s": round(t1 - t0, 4), "steps_s": round(t2 - t1, 4),
            "download_s": round(t3 - t2, 4), "upload_GBps": round(phi.nbytes / (t1 - t0) / 1e9, 2),
            "download_GBps": round(phi.nbytes / (t3 - t2) / 1e9, 2),
            "mlups_steps_only": round(cells / (t2 - t1) / 1e6, 1),
            "mlups_with_transfers": round(cells / (t3 - t0) / 1e6, 1)}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(W, steps=3, sz=32):
    """The CPU oracle (the reference's algorithm in its dense layout, OpenMP
    over cells) on a full-xy 512 x 512 x sz slab of the workload holding the
    24-cell seed layer and the front, `steps` steps after one untimed step,
    in pure FP64 and in the reference's own precision (float32 stores
    between FP64 arithmetic)."""
    from oracle import oracle as O

    sx, sy = NX, NY
    p, d = W["p"], W["d"]
    N = W["N"]
    phi = np.zeros((N, sx, sy, sz), np.float64)
    prof = inputs.tanh_profile(sz, p["dz"], d["delta"], SEED_H)
    gm = W["gmap"][:sx, :sy]
    ll, mm = np.meshgrid(np.arange(sx), np.arange(sy), indexing="ij")
    phi[gm, ll, mm, :] = prof
    phi[0] = inputs.liquid_from_solid(prof)[None, None, :]
    temp = np.broadcast_to(W["temp"][:sz], (sx, sy, sz)).copy()
    tabs = tuple(np.asarray(t, np.float64) for t in W["tables"])
    out = {}
    for name, faithful in (("fp64", False), ("reference_precision", True)):
        ph, tp, _, _ = O.run(phi, temp, 1, tabs, W["scalars"], kmax=50, cooling_rate=W["cooling"], faithful=faithful)
        t0 = time.perf_counter()
        O.run(ph, tp, steps, tabs, W["scalars"], kmax=50, cooling_rate=W["cooling"], faithful=faithful)
        dt = time.perf_counter() - t0
        out[name] = {"mlups": round(sx * sy * sz * steps / dt / 1e6, 4), "seconds": round(dt, 2)}
        del ph, tp
        log(f"cpu baseline {name}: {out[name]}")
    return {"value": out["fp64"]["mlups"], "unit": "MLUPS", "cores": O.threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "reference_precision_mlups": out["reference_precision"]["mlups"],
            "sample": f"full-xy {sx}x{sy}x{sz} slab of the workload's initial state (24-cell seed layer and the "
                      f"front), N={N}, the oracle in the reference's dense layout, FP64 (value) and the "
                      f"reference's float32-store precision, {steps} steps each after 1 untimed step "
                      f"({out['fp64']['seconds']} s and {out['reference_precision']['seconds']} s); "
                      f"fewer than 20 steps to keep the sample within a few tens of seconds"}


def summarize_trace(path, window):
    """Per-kernel averages over the last `window` dispatches of each step
    kernel in a rocprofv3 CSV: durations (ms) from a kernel trace, counter
    values from a counter collection."""
    import csv
    import re

    def short(name):
        m = re.search(r"(k_[a-z0-9_]+)", name)
        return m.group(1) if m else name.split("(")[0][:60]

    rows = list(csv.DictReader(open(path)))
    if not rows:
        return {}
    out = {}
    if "Counter_Name" in rows[0]:
        per = {}
        for r in rows:
            key = (short(r["Kernel_Name"]), r["Counter_Name"])
            d = per.setdefault(key, {})
            i = int(r["Dispatch_Id"])
            d[i] = d.get(i, 0.0) + float(r["Counter_Value"])
        for (k, cn), d in per.items():
            last = [d[i] for i in sorted(d)[-window:]]
            out.setdefault(k, {})[cn] = sum(last) / len(last)
        return out
    per = {}
    for r in rows:
        t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        per.setdefault(short(r["Kernel_Name"]), []).append((t0, (t1 - t0) * 1e-6))
    for k, v in per.items():
        v.sort()
        last = [d for _, d in v[-window:]]
        out[k] = {"dispatches": len(v), "avg_ms_last": sum(last) / len(last), "n_last": len(last)}
    return out


STEP_KERNELS = ("k_tiles", "k_light", "k_merge", "k_interface", "k_front1", "k_front_multi", "k_step_general")
CLOCK_GHZ, N_SIMD = 2.4, 1024


def profile_summary(kind, window, trace, counters, copy_kib=262144):
    """Per-step PMC summary of the step kernels (the last `window`
    dispatches of each) from rocprofv3 CSVs, as committed under profiles/:
    kind 'traffic': counters = (FETCH_SIZE csv, WRITE_SIZE csv), each scaled
    to bytes by what it reads for the 256 MiB device copy of --copy-probe
    (the largest __amd_rocclr_copyBuffer dispatch); kind 'valu': counters =
    (csv with SQ_INSTS_VALU, SQ_INSTS_SALU, SQ_INSTS_VMEM_RD, SQ_INSTS_LDS,
    SQ_WAVES).  Durations come from the kernel trace."""
    import csv
    dur = summarize_trace(trace, window)
    # the kernels of the timed steps: those dispatched after the first of the
    # last `window` classification passes (autotuner trials of other launch
    # modes come before it)
    rows = list(csv.DictReader(open(trace)))
    light = sorted(int(r["Start_Timestamp"]) for r in rows if "k_light" in r["Kernel_Name"])
    t0 = light[-window]
    timed = {k for k in STEP_KERNELS
             if any(k in r["Kernel_Name"] and int(r["Start_Timestamp"]) >= t0 for r in rows)}
    out = {"window_dispatches": window, "kernels": {}}
    if kind == "traffic":
        scales, per = [], []
        for path, name in zip(counters, ("FETCH_SIZE", "WRITE_SIZE")):
            probe = {}
            for r in csv.DictReader(open(path)):
                if r["Counter_Name"] == name and "copyBuffer" in r["Kernel_Name"]:
                    i = int(r["Dispatch_Id"])
                    probe[i] = probe.get(i, 0.0) + float(r["Counter_Value"])
            # the probe is the first device copy of the run
            scales.append(copy_kib / probe[min(probe)])
            per.append(summarize_trace(path, window))
        out["fetch_scale"], out["write_scale"] = round(scales[0], 4), round(scales[1], 4)
        tot = 0
        for k in STEP_KERNELS:
            if k not in timed or k not in per[0]:
                continue
            rd = per[0][k]["FETCH_SIZE"] * 1024 * out["fetch_scale"]
            wr = per[1][k]["WRITE_SIZE"] * 1024 * out["write_scale"]
            out["kernels"][k] = {"read_bytes": round(rd), "write_bytes": round(wr),
                                 "duration_ms": round(dur[k]["avg_ms_last"], 4)}
            tot += rd + wr
        out["bytes_per_step"] = round(tot)
        return out
    c = summarize_trace(counters[0], window)
    out["model"] = (f"a wavefront-64 VALU instruction occupies a 16-lane SIMD for 4 cycles; {N_SIMD} SIMDs "
                    f"(256 CUs x 4) at {CLOCK_GHZ} GHz")
    tv = tt = 0.0
    for k in STEP_KERNELS:
        if k not in timed or k not in c:
            continue
        v = c[k]["SQ_INSTS_VALU"]
        t_us = v * 4 / (N_SIMD * CLOCK_GHZ * 1e3)
        d_us = dur[k]["avg_ms_last"] * 1e3
        out["kernels"][k] = {"valu_wave_instr": round(v), "salu_wave_instr": round(c[k]["SQ_INSTS_SALU"]),
                             "vmem_rd_wave_instr": round(c[k]["SQ_INSTS_VMEM_RD"]),
                             "lds_wave_instr": round(c[k]["SQ_INSTS_LDS"]), "waves": round(c[k]["SQ_WAVES"]),
                             "valu_time_at_full_issue_us": round(t_us, 1), "duration_us": round(d_us, 1),
                             "valu_issue_frac": round(t_us / d_us, 3)}
        tv += v
        tt += t_us
    out["valu_wave_instr_per_step"] = round(tv)
    out["valu_time_at_full_issue_us_per_step"] = round(tt, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mid", type=int, default=RUN_STEPS // 2,
                    help="steps of the run advanced (timed, in segments) before the warm-up: the timed steps "
                         "start at step mid + warmup of the run")
    ap.add_argument("--run-steps", type=int, default=RUN_STEPS,
                    help="length of the workload's run; with --traj the FP32 and FP64 runs are completed to it")
    ap.add_argument("--traj", type=int, default=1,
                    help="complete the runs to --run-steps and report the trajectory and FP32 vs FP64 there")
    ap.add_argument("--kernel-steps", type=int, default=20,
                    help="steps after the timed ones with every kernel bracketed by HIP events")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy (reference for HBM counters)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "pmc_traffic_512x512x256_fp32.json"),
                    help="rocprofv3 PMC traffic summary to report as roofline.traffic")
    ap.add_argument("--valu-json", default=os.path.join(ROOT, "profiles", "valu_512x512x256_fp32.json"),
                    help="rocprofv3 PMC vector-ALU summary to report as roofline.valu")
    ap.add_argument("--graphs", type=int, default=0, help="replay steps from HIP graphs (1) or launch eagerly (0)")
    ap.add_argument("--autotune", type=int, default=8,
                    help="time launch shapes for this many steps each after the warm-up (0: default shape)")
    ap.add_argument("--sweep-steps", type=int, default=3000,
                    help="steps of the batched bicrystal sweep (256 grids); 0 skips it")
    ap.add_argument("--sweep-graphs", type=int, default=0)
    ap.add_argument("--surrogate", type=int, default=1,
                    help="train the surrogate and measure assisted stepping at 128^3 (0 skips it)")
    ap.add_argument("--c2", type=int, default=1, help="FP32 vs FP64 on the 128^3 16-grain case (0 skips it)")
    ap.add_argument("--pcie", type=int, default=1, help="the host-buffer (PCIe-inclusive) rate at 128^3 (0 skips it)")
    ap.add_argument("--summarize-trace", default=None,
                    help="print per-kernel averages over the last --steps dispatches of a rocprofv3 CSV and exit")
    ap.add_argument("--profile-summary", nargs="+", default=None, metavar="ARG",
                    help="traffic KERNEL_TRACE FETCH_CSV WRITE_CSV | valu KERNEL_TRACE PMC_CSV: print the "
                         "per-step PMC summary of the step kernels over the last --steps dispatches and exit")
    args = ap.parse_args()
    if args.summarize_trace:
        print(json.dumps(summarize_trace(args.summarize_trace, args.steps), indent=1))
        return
    if args.profile_summary:
        kind, trace, *csvs = args.profile_summary
        print(json.dumps(profile_summary(kind, args.steps, trace, csvs), indent=1))
        return
    # torch's HIP runtime is loaded before the solver's (the surrogate needs
    # it, and every run then binds the same runtime, profiled or not)
    import mpf.surrogate  # noqa: F401
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the solver has no "
              "multi-GPU path in this build.", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        from mpf import dense
        n = 256 << 20
        src, dst = dense.DeviceArray((n,), np.uint8), dense.DeviceArray((n,), np.uint8)
        from mpf import _lib
        _lib.check(_lib.lib().mpf_memcpy_dtod(dst.ptr, src.ptr, n))
        dense.synchronize()
        log("copy probe: 256 MiB device-to-device copy done")
    W = workload()
    ncell = NX * NY * NZ
    mid = max(0, min(args.mid, args.run_steps))
    total = args.run_steps if args.traj else mid + args.warmup + args.steps + args.kernel_steps
    total = max(total, mid + args.warmup + args.steps + args.kernel_steps)
    log(f"workload ready: {NX}x{NY}x{NZ}, N={W['N']}; timed steps start at step {mid + args.warmup}")
    r32 = run_trajectory(W, "fp32", mid, args.warmup, args.steps, total, kernel_steps=args.kernel_steps,
                         graphs=bool(args.graphs), autotune=args.autotune)
    t32 = r32["timed"]
    ms32 = t32["ms"] / args.steps
    log(f"fp32: {ms32:.3f} ms/step at steps {t32['start_step']}..{t32['start_step'] + args.steps}")
    log(f"fp32 per-kernel ms (live HIP events): {r32['kernel_ms']}")
    log(f"fp32 launch {r32['launch']} graphs={bool(args.graphs)} autotune={r32['tuned']}")
    B32, slot32 = alg_bytes_per_step(r32["ob0"], r32["ob1"], 4, ncell, NZ)
    achieved = B32 / (ms32 * 1e-3) / 1e9
    light_per_step = t32["light_cells"] / args.steps
    T32 = touched_bytes_per_step(light_per_step, slot32 * ncell, 4, NZ)
    traffic = measured_traffic(args.traffic_json)
    valu = measured_valu(args.valu_json)

    out = {
        "metric": METRIC, "value": round(ncell * args.steps / (t32["ms"] * 1e-3) / 1e6, 2), "unit": "MLUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms32, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"3D random mode, plain method, 512x512x256, 200 Voronoi grains, 24-cell seed "
                               f"layer, random quaternions, uniform cooling, KMAX=50, {args.run_steps}-step run; "
                               f"timed steps start at step {t32['start_step']} (mid-run state)",
                   "nx": NX, "ny": NY, "nz": NZ, "n_solid": N_SOLID, "seed_height": SEED_H, "kmax": 50,
                   "precision": "fp32", "parallelism": "single GPU", "graphs": bool(args.graphs),
                   "timed_from_step": t32["start_step"], "run_steps": args.run_steps,
                   "launch": r32["launch"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes_per_step"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "alg_bytes_per_step": round(B32), "mean_stored_slots_per_cell": round(slot32, 4),
                     "alg_bytes_note": "every cell's state read and written once per step (a dense sweep's "
                                       "compulsory bytes); the step skips the liquid planes that cannot change",
                     "touched_cells_per_step": round(light_per_step),
                     "touched_bytes_per_step": round(T32),
                     "touched_achieved": round(T32 / (ms32 * 1e-3) / 1e9, 1),
                     "valu": valu,
                     "kernel": "one step: k_tiles, k_light, k_merge, then the interface and many-phase passes "
                               "(launch mode in config.launch)",
                     "kernel_ms_live": {k: round(v, 4) for k, v in (r32["kernel_ms"] or {}).items()}},
        "wall_s_timed": round(t32["wall"], 3),
        "front_mean_cells": r32["ob1"]["front_mean"],
        "cells_per_step": {"light": t32["light_cells"] // args.steps, "merge": t32["merge_cells"] // args.steps,
                           "front": t32["front_cells"] // args.steps,
                           "general": t32["general_cells"] // args.steps},
        "max_active": t32["max_active"],
    }
    out["trajectory"] = {"fp32": trajectory_summary(r32, ncell)}
    log(f"fp32 run: {out['trajectory']['fp32']['mean_ms_per_step']} ms/step over "
        f"{out['trajectory']['fp32']['steps']} steps")
    if not args.no_fp64:
        r64 = run_trajectory(W, "fp64", mid, args.warmup, args.steps, total, autotune=args.autotune)
        ms64 = r64["timed"]["ms"] / args.steps
        log(f"fp64: {ms64:.3f} ms/step")
        out["fp64_mlups"] = round(ncell * args.steps / (r64["timed"]["ms"] * 1e-3) / 1e6, 2)
        out["fp64_ms_per_step"] = round(ms64, 4)
        out["trajectory"]["fp64"] = trajectory_summary(r64, ncell)
        if r64["steps_done"] == r32["steps_done"]:
            cmp = fp32_vs_fp64(r32["solver"], r64["solver"], W["p"]["dx"], ncell)
            cmp["at_step"] = r32["steps_done"]
            cmp["note"] = ("max |phi_32 - phi_64| jumps to ~1e-2..1e-1 at the few cells whose set of non-zero "
                           "grains differs between the precisions (a value one rounds to 0); the reference's own "
                           "float32-store precision shows the same jumps against FP64 (tests/test_gpu_workloads.py)")
            out["fp32_vs_fp64"] = cmp
            log(f"fp32 vs fp64 at step {cmp['at_step']}: {cmp}")
        r64["solver"].close()
    r32["solver"].close()
    if args.c2:
        out["c2_fp32_vs_fp64"] = c2_precision()
        log(f"c2: {out['c2_fp32_vs_fp64']}")
    if args.pcie:
        out["pcie_inclusive"] = pcie_inclusive()
        log(f"pcie inclusive: {out['pcie_inclusive']}")
    if args.sweep_steps > 0:
        out["batched_sweep"] = batched_sweep(args.sweep_steps, bool(args.sweep_graphs))
    if args.surrogate:
        out["surrogate"] = surrogate_block()
    if not args.no_cpu:
        cb = cpu_baseline(W)
        log(f"cpu baseline: {cb}")
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
