"""3D two-grain (bicrystal) solidification: the counterpart of
run_twomode_3d.py.

Same config_3d.yaml keys (grid, physical, interface, anisotropy, mobility,
output and the twomode block: seed_height, split_ratio, grain seed offsets,
grain1/grain2 orientations with the reference's defaults,
run_twomode_3d.py:128-178), the same bicrystal seed (init_twomode_phi_3d, in
columnar form on the device), linear temperature and uniform cooling, and
the same artefacts.

    python -m mpf.run_twomode_3d --config config_3d.yaml
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import inputs, plot
from .config import DEFAULT_CONFIG, flat_params, load_config
from .solver import Solver

# run_twomode_3d.py:167-168
DEFAULT_GRAIN1 = {"orientation_type": "euler", "euler_deg": [0.0, 0.0, 0.0]}
DEFAULT_GRAIN2 = {"orientation_type": "euler", "euler_deg": [0.0, 54.736, 45.0]}


def setup(c: dict):
    p = flat_params(c)
    if p["dx"] != p["dy"] or p["dx"] != p["dz"]:
        raise ValueError("3D kernels assume dx == dy == dz.")
    tm = c.get("twomode", {})
    seed_height = int(tm.get("seed_height", 32))
    split_ratio = float(tm.get("split_ratio", 0.5))
    split_index = int(p["nx"] * split_ratio)
    off1, off2 = int(tm.get("grain1_seed_offset", 0)), int(tm.get("grain2_seed_offset", 0))
    q = np.zeros((3, 4))
    q[0] = [0.0, 0.0, 0.0, 1.0]
    q[1] = inputs.build_quaternion_from_config(tm.get("grain1", DEFAULT_GRAIN1))
    q[2] = inputs.build_quaternion_from_config(tm.get("grain2", DEFAULT_GRAIN2))
    d = inputs.derive_constants(p)
    w, a, m = inputs.build_interaction_matrices(3, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"], d["w_gb"],
                                                d["m_gb_phi"])
    temp = inputs.init_temperature_field_3d(1, 1, p["nz"], p["T_melt"], p["G"], p["dz"], seed_height)[0, 0]
    return {"p": p, "d": d, "seed_height": seed_height, "split_ratio": split_ratio, "split_index": split_index,
            "offsets": (off1, off2), "quats": q, "tables": (w, a, m, inputs.compute_rotated_n111(q)),
            "temp": temp.astype(np.float64), "cooling": float(np.float32(p["G"] * p["V_pulling"] * p["dt"])),
            "scalars": inputs.kernel_scalars(p, d, as_float32=True)}


def run(c: dict, out_dir: str, plots: bool = True, keep_phi: bool = False, log=print) -> dict:
    S = setup(c)
    p = S["p"]
    nx, ny, nz, nsteps = p["nx"], p["ny"], p["nz"], p["nsteps"]
    save_every = int(c["output"]["save_every"])
    save_slices = list(c["output"].get("save_slices", ["xy", "xz"]))
    slice_index = c["output"].get("slice_index")
    log("Mode        : twomode")
    log(f"Grid        : {nx}x{ny}x{nz}, dx={p['dx']:.1e}, dt={p['dt']:.1e}, nsteps={nsteps}")
    log(f"Seed height : {S['seed_height']}  (split_index={S['split_index']}, split_ratio={S['split_ratio']:.2f})")
    log(f"Output dir  : {out_dir}")
    with Solver((nx, ny, nz), 3, S["scalars"], kmax=p["KMAX"], precision=p["precision"],
                cooling_rate=S["cooling"]) as s:
        s.set_tables(*S["tables"])
        s.init_columns(*inputs.columns_twomode_3d(nx, ny, nz, p["dz"], S["d"]["delta"], S["seed_height"],
                                                  S["split_index"], *S["offsets"]))
        s.upload_temp(S["temp"])
        files = []
        if plots:
            files += plot.save_requested_slices(s, out_dir, 3, save_slices, slice_index, "step_0",
                                                "twomode3d -- step 0")
        files += plot.save_run_config(out_dir, c, {
            "mode": "twomode", "number_of_grain": 3, "seed_height": S["seed_height"],
            "split_ratio": S["split_ratio"], "split_index": S["split_index"],
            "grain1_seed_offset": S["offsets"][0], "grain2_seed_offset": S["offsets"][1],
            "grain_quaternions": S["quats"], "out_dir": out_dir})
        done, gpu_ms = 0, 0.0
        t0 = time.perf_counter()
        while done < nsteps:
            n = min(save_every - done % save_every, nsteps - done)
            s.step(n)
            gpu_ms += s.stats()["last_step_ms"]
            done += n
            if done % save_every == 0:
                if plots:
                    files += plot.save_requested_slices(s, out_dir, 3, save_slices, slice_index, f"step_{done}",
                                                        f"twomode3d -- step {done}")
                log(f"  saved step_{done}")
        ob = s.observables(dx=p["dx"])
        out = {"out_dir": out_dir, "steps": done, "files": files, "gpu_s": gpu_ms * 1e-3,
               "wall_s": time.perf_counter() - t0, "front_mean": ob["front_mean"],
               "grain_volumes": ob["grain_volumes"].tolist()}
        if keep_phi:
            out["phi"] = s.phi()
    log("\nDone.")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run 3D two-grain solidification.")
    ap.add_argument("--config", default=DEFAULT_CONFIG)
    ap.add_argument("--out-dir", default=None)
    ap.add_argument("--no-plots", action="store_true")
    args = ap.parse_args(argv)
    c = load_config(args.config)
    ph = c["physical"]
    out_dir = args.out_dir or f"result/twomode_3d/newgrain/{ph['V_pulling'] * ph['G']}"
    return run(c, out_dir, plots=not args.no_plots, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main(sys.argv[1:])
