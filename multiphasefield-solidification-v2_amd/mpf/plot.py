"""Run artefacts in the reference's formats: phase-map slices, interface
height maps (PNG) and the two YAML run records.

The reference draws from a full host copy of phi
(src/plot_utils_3d.py:36-96, src/plot_utils.py:67-100).  Here the slice's
dominant-phase map (mpf_phase_map) and the front heights
(mpf_observables) come from the device, so a 512x512x256 run never copies
its 50+ GB dense phi to the host.  File names and directory layout match.
"""
from __future__ import annotations

import datetime
import os

import numpy as np
import yaml


def resolve_slice_index(axis: str, shape, index) -> int:
    """src/plot_utils_3d.py:17-33"""
    nx, ny, nz = shape
    if axis == "xy":
        upper, default = nz - 1, nz // 2
    elif axis == "xz":
        upper, default = ny - 1, ny // 2
    elif axis == "yz":
        upper, default = nx - 1, nx // 2
    else:
        raise ValueError(f"Unsupported slice axis: {axis}")
    if index is None:
        return default
    return int(np.clip(index, 0, upper))


AXES = {"xy": 0, "xz": 1, "yz": 2}


def _plt():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    return plt


def save_phase_map_slice(slc: np.ndarray, out_dir: str, filename: str, number_of_grain: int, axis: str,
                         idx: int, title: str = "", dpi: int = 150) -> str:
    """One phase-map slice (save_phase_map_slice_3d, src/plot_utils_3d.py:36-72)."""
    plt = _plt()
    fig, ax = plt.subplots(figsize=(10, 10))
    im = ax.imshow(slc.T, cmap="tab20", origin="lower", vmin=0, vmax=max(number_of_grain - 1, 1))
    cbar = fig.colorbar(im, ax=ax, ticks=np.arange(number_of_grain), shrink=0.7)
    cbar.set_label("Phase ID", fontsize=24)
    cbar.ax.tick_params(labelsize=14)
    ax.set_title(title or f"{axis} slice @ {idx}", fontsize=24)
    ax.axis("off")
    plt.tight_layout()
    d = os.path.join(out_dir + "/3dphase", "3dphase")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, filename)
    plt.savefig(path, dpi=dpi)
    plt.close(fig)
    return path


def interface_z(front: np.ndarray) -> np.ndarray:
    """Topmost solid cell per column, 0 where a column has none
    (save_interface_position_3d, src/plot_utils_3d.py:75-85)."""
    f = np.asarray(front)
    return np.where(f >= 0, f, 0)


def save_interface_position(front: np.ndarray, out_dir: str, filename: str, title: str = "",
                            dpi: int = 150) -> str:
    """src/plot_utils_3d.py:75-96"""
    plt = _plt()
    fig, ax = plt.subplots(figsize=(8, 8))
    im = ax.imshow(interface_z(front).T, cmap="viridis", origin="lower")
    cbar = fig.colorbar(im, ax=ax)
    cbar.set_label("Interface z position [cells]")
    ax.set_title(title or "Interface z position")
    ax.axis("off")
    plt.tight_layout()
    d = os.path.join(out_dir + "/interface", "interface")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, filename)
    plt.savefig(path, dpi=dpi)
    plt.close(fig)
    return path


def _serializable(obj):
    if isinstance(obj, np.ndarray):
        return obj.tolist()
    if isinstance(obj, np.generic):
        return obj.item()
    if isinstance(obj, dict):
        return {k: _serializable(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_serializable(v) for v in obj]
    return obj


def save_run_config(out_dir: str, cfg: dict, runtime_info: dict) -> list[str]:
    """config_snapshot.yaml (the configuration as loaded) and run_params.yaml
    (resolved parameters with a timestamp), src/plot_utils.py:67-100."""
    os.makedirs(out_dir, exist_ok=True)
    info = _serializable(dict(runtime_info))
    info["saved_at"] = datetime.datetime.now().strftime("%Y-%m-%d %H:%M:%S")
    a = os.path.join(out_dir, "config_snapshot.yaml")
    with open(a, "w", encoding="utf-8") as f:
        yaml.dump(cfg, f, allow_unicode=True, default_flow_style=False, sort_keys=False)
    b = os.path.join(out_dir, "run_params.yaml")
    with open(b, "w", encoding="utf-8") as f:
        yaml.dump(info, f, allow_unicode=True, default_flow_style=False, sort_keys=False)
    return [a, b]


def save_requested_slices(solver, out_dir: str, number_of_grain: int, save_slices, slice_index, prefix: str,
                          title_suffix: str, grid: int | None = None) -> list[str]:
    """The phase-map slices named in save_slices plus the interface height map
    of a Solver (grid None) or of grid `grid` of a Batch
    (run_randommode_3d.py save_requested_slices)."""
    shape = solver.shape if grid is None else solver.shapes[grid]
    paths = []
    for axis in save_slices:
        idx = resolve_slice_index(axis, shape, slice_index)
        slc = solver.phase_map(axis, idx) if grid is None else solver.phase_map(grid, axis, idx)
        paths.append(save_phase_map_slice(slc, out_dir, f"{prefix}_{axis}.png", number_of_grain, axis, idx,
                                          title=f"{title_suffix} ({axis})"))
    ob = solver.observables() if grid is None else solver.observables(grid)
    front = np.where(np.isfinite(ob["front"]), ob["front"], -1).astype(np.int64)
    paths.append(save_interface_position(front, out_dir, f"{prefix}_interface.png",
                                         title=f"Interface z position -- {title_suffix}"))
    return paths
