"""3D random-mode solidification: the counterpart of run_randommode_3d.py.

Same command-line flags (run_randommode_3d.py:42-53), the same
config_3d.yaml keys and overrides (:88-177), the same inputs (Voronoi seed
layer, quaternions, linear temperature, uniform cooling G*V*dt as float32)
and the same artefacts:
- step_N phase-map slices and interface maps every save_every steps;
- config_snapshot.yaml / run_params.yaml.
The step runs on the GPU through mpf.Solver in the slot layout; the seed
layer is built on the device from its columns (no dense phi on the host).

    python -m mpf.run_randommode_3d --config config_3d.yaml --nx 128 --nsteps 2000
"""
from __future__ import annotations

import argparse
import copy
import sys
import time

import numpy as np

from . import inputs, plot
from .config import DEFAULT_CONFIG, load_config
from .solver import Solver


def parse_args(argv=None):
    """run_randommode_3d.py:42-53, plus --no-plots (skip the PNGs)."""
    ap = argparse.ArgumentParser(description="Run 3D random-mode solidification.")
    ap.add_argument("--config", default=DEFAULT_CONFIG, help="Path to config YAML.")
    ap.add_argument("--out-dir", default=None, help="Override output directory.")
    ap.add_argument("--nx", type=int, default=None, help="Override grid nx.")
    ap.add_argument("--ny", type=int, default=None, help="Override grid ny.")
    ap.add_argument("--nz", type=int, default=None, help="Override grid nz.")
    ap.add_argument("--nsteps", type=int, default=None, help="Override total time steps.")
    ap.add_argument("--save-every", type=int, default=None, help="Override save interval.")
    ap.add_argument("--seed-height", type=int, default=None, help="Override initial seed height.")
    ap.add_argument("--n-solid", type=int, default=None, help="Override number of solid grains.")
    ap.add_argument("--no-plots", action="store_true", help="Skip the PNG artefacts.")
    return ap.parse_args(argv)


def runtime_config(cfg: dict, args) -> dict:
    """The reference's command-line overrides (run_randommode_3d.py:90-106)."""
    c = copy.deepcopy(cfg)
    if args.nx is not None:
        c["grid"]["nx"] = args.nx
    if args.ny is not None:
        c["grid"]["ny"] = args.ny
    if args.nz is not None:
        c["grid"]["nz"] = args.nz
    if args.nsteps is not None:
        c["grid"]["nsteps"] = args.nsteps
    if args.save_every is not None:
        c["output"]["save_every"] = args.save_every
    if args.seed_height is not None:
        c.setdefault("randommode", {})["seed_height"] = args.seed_height
    if args.n_solid is not None:
        c.setdefault("randommode", {})["n_solid"] = args.n_solid
    return c


def setup(c: dict):
    """Resolved parameters, tables and inputs (run_randommode_3d.py:108-226).
    Raises the reference's configuration errors."""
    from .config import flat_params
    p = flat_params(c)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    if p["dx"] != p["dy"] or p["dx"] != p["dz"]:
        raise ValueError("3D kernels assume dx == dy == dz.")
    tpb = c["gpu"].get("threads_per_block", [4, 4, 4])
    if len(tpb) != 3:
        raise ValueError("gpu.threads_per_block must have 3 entries for 3D.")
    if not 1 <= p["KMAX"] <= 64:
        raise ValueError(f"config_3d.yaml gpu.KMAX={p['KMAX']} is outside the supported 1..64.")
    rm = c.get("randommode", {})
    seed_height = int(rm.get("seed_height", 32))
    n_solid = int(rm.get("n_solid", 10))
    N = n_solid + 1
    if N > p["MAX_GRAINS"]:
        raise RuntimeError(f"number_of_grain={N} exceeds MAX_GRAINS={p['MAX_GRAINS']}. "
                           "Reduce randommode.n_solid or increase gpu.MAX_GRAINS.")
    d = inputs.derive_constants(p)
    quats = inputs.assign_quaternions_to_grains(n_solid, str(rm.get("orientation_mode", "random")),
                                                int(rm.get("orientation_seed", 42)),
                                                rm.get("orientation_csv", None) or None)
    w, a, m = inputs.build_interaction_matrices(N, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"], d["w_gb"],
                                                d["m_gb_phi"])
    n111 = inputs.compute_rotated_n111(quats)
    gmap = inputs.generate_random_grain_map_3d(nx, ny, n_solid, random_seed=int(rm.get("random_seed", 42)))
    temp = inputs.init_temperature_field_3d(1, 1, nz, p["T_melt"], p["G"], p["dz"], seed_height)[0, 0]
    return {"p": p, "d": d, "N": N, "n_solid": n_solid, "seed_height": seed_height, "quats": quats,
            "tables": (w, a, m, n111), "grain_map": gmap, "temp": temp.astype(np.float64),
            "cooling": float(np.float32(p["G"] * p["V_pulling"] * p["dt"])),
            "scalars": inputs.kernel_scalars(p, d, as_float32=True)}


def run(c: dict, out_dir: str, plots: bool = True, keep_phi: bool = False, log=print) -> dict:
    """Create the solver, build the seed layer, step nsteps saving every
    save_every steps; returns a summary (and the final dense phi with keep_phi)."""
    S = setup(c)
    p = S["p"]
    nx, ny, nz, nsteps = p["nx"], p["ny"], p["nz"], p["nsteps"]
    save_every = int(c["output"]["save_every"])
    save_slices = list(c["output"].get("save_slices", ["xy", "xz"]))
    slice_index = c["output"].get("slice_index")
    log("Mode             : randommode_3d")
    log(f"Grid             : {nx}x{ny}x{nz}, dx={p['dx']:.1e}, dt={p['dt']:.1e}, nsteps={nsteps}")
    log(f"Seed height      : {S['seed_height']}")
    log(f"n_solid          : {S['n_solid']}  (number_of_phases={S['N']})")
    log(f"Precision        : {p['precision']}")
    log(f"Output dir       : {out_dir}")
    with Solver((nx, ny, nz), S["N"], S["scalars"], kmax=p["KMAX"], precision=p["precision"],
                cooling_rate=S["cooling"]) as s:
        s.set_tables(*S["tables"])
        s.init_columns(*inputs.columns_from_grain_map_3d(S["grain_map"], nz, p["dz"], S["d"]["delta"],
                                                         S["seed_height"]))
        s.upload_temp(S["temp"])
        files = []
        if plots:
            files += plot.save_requested_slices(s, out_dir, S["N"], save_slices, slice_index, "step_0",
                                                "randommode3d -- step 0")
        files += plot.save_run_config(out_dir, c, {
            "mode": "randommode_3d", "number_of_grain": S["N"], "n_solid": S["n_solid"],
            "seed_height": S["seed_height"], "grain_quaternions": S["quats"], "out_dir": out_dir,
            "precision": p["precision"]})
        log(f"\nStarting time evolution ({nsteps} steps, saving every {save_every}) ...")
        done, gpu_ms = 0, 0.0
        t0 = time.perf_counter()
        while done < nsteps:
            n = min(save_every - done % save_every, nsteps - done)
            s.step(n)
            gpu_ms += s.stats()["last_step_ms"]
            done += n
            if done % save_every == 0:
                if plots:
                    files += plot.save_requested_slices(s, out_dir, S["N"], save_slices, slice_index,
                                                        f"step_{done}", f"randommode3d -- step {done}")
                log(f"saved step_{done}")
        wall = time.perf_counter() - t0
        ob = s.observables(dx=p["dx"])
        st = s.stats()
        out = {"out_dir": out_dir, "steps": done, "files": files, "gpu_s": gpu_ms * 1e-3, "wall_s": wall,
               "mlups_gpu": nx * ny * nz * done / max(gpu_ms * 1e-3, 1e-12) / 1e6,
               "front_mean": ob["front_mean"], "solid_fraction": ob["solid_fraction"],
               "grain_volumes": ob["grain_volumes"].tolist(), "kmax_overflow": st["kmax_overflow"],
               "max_active": st["max_active"]}
        if keep_phi:
            out["phi"] = s.phi()
    log(f"\nDone. {out['mlups_gpu']:.0f} MLUPS on the GPU, front at {out['front_mean']:.2f} cells.")
    return out


def main(argv=None):
    args = parse_args(argv)
    c = runtime_config(load_config(args.config), args)
    n_solid = int(c.get("randommode", {}).get("n_solid", 10))
    ph = c["physical"]
    out_dir = args.out_dir or (c["output"]["outdir"] + f"/{n_solid} grains/{ph['V_pulling'] * ph['G']:.1f}K")
    return run(c, out_dir, plots=not args.no_plots, log=lambda s: print(s, flush=True))


if __name__ == "__main__":
    main(sys.argv[1:])
