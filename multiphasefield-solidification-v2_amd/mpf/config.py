"""Loading of the config_3d.yaml schema (reference: config_3d.yaml,
run_randommode_3d.py:88-177)."""
from __future__ import annotations

import copy
import os

import yaml

DEFAULT_CONFIG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                              "configs", "config_3d.yaml")


def load_config(path: str | None = None) -> dict:
    with open(path or DEFAULT_CONFIG, "r", encoding="utf-8") as f:
        return yaml.safe_load(f)


def flat_params(cfg: dict) -> dict:
    """The parameter set of tests/verification_3d.py:SimulationParams (:39-64)."""
    g, ph, itf, an, mo, gpu = (cfg["grid"], cfg["physical"], cfg["interface"], cfg["anisotropy"],
                               cfg["mobility"], cfg.get("gpu", {}))
    return {
        "nx": int(g["nx"]), "ny": int(g["ny"]), "nz": int(g["nz"]),
        "dx": float(g["dx"]), "dy": float(g.get("dy", g["dx"])), "dz": float(g.get("dz", g["dx"])),
        "dt": float(g["dt"]), "nsteps": int(g["nsteps"]),
        "T_melt": float(ph["T_melt"]), "G": float(ph["G"]), "V_pulling": float(ph["V_pulling"]),
        "Sf": float(ph["Sf"]),
        "delta_factor": float(itf["delta_factor"]), "gamma_100": float(itf["gamma_100"]),
        "gamma_GB": float(itf["gamma_GB"]),
        "a0_deg": float(an["a0_deg"]), "delta_a": float(an["delta_a"]), "mu_a": float(an["mu_a"]),
        "p_round": float(an["p_round"]), "ksi": float(an["ksi"]), "omg_deg": float(an["omg_deg"]),
        "M_SL": float(mo["M_SL"]), "M_GB_ratio": float(mo["M_GB_ratio"]),
        "MAX_GRAINS": int(gpu.get("MAX_GRAINS", 50)), "KMAX": int(gpu.get("KMAX", 50)),
        "precision": str(gpu.get("precision", "fp32")),
    }


def with_overrides(cfg: dict, **kw) -> dict:
    out = copy.deepcopy(cfg)
    for k, v in kw.items():
        if v is None:
            continue
        sec, key = k.split(".")
        out.setdefault(sec, {})[key] = v
    return out
