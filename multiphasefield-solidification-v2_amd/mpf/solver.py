"""Resident solver: the reference's host time loop (run_randommode_3d.py:244-300,
tests/verification_3d.py:771-936) over the library's device state."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FP32, FP64, check, ptr


class KMaxOverflowError(RuntimeError):
    """A cell needed more active phases than KMAX (tests/verification_3d.py:86-87)."""


class Solver:
    """One grid advanced on the GPU.

    Arrays cross the boundary in the reference's dense layouts:
    phi[N, nx, ny, nz] (2D: [N, nx, ny]), temp[nx, ny, nz].
    """

    def __init__(self, shape, n_phase: int, scalars: dict, *, kmax: int = 50,
                 precision: str = "fp32", temp_dense: bool = False, temp_f32: bool = False,
                 anisotropy: bool = True, torque: bool = True, halt_on_kmax: bool = False,
                 cooling_rate: float = 0.0, slot_capacity: int = 0):
        shape = tuple(int(v) for v in shape)
        if len(shape) not in (2, 3):
            raise ValueError("shape must be (nx, ny, nz) or (nx, ny)")
        self.dim = len(shape)
        self.shape = shape
        self.nx, self.ny = shape[0], shape[1]
        self.nz = shape[2] if self.dim == 3 else 1
        self.n_phase = int(n_phase)
        self.kmax = int(kmax)
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self.temp_dense = bool(temp_dense)
        self.temp_f32 = bool(temp_f32)
        cfg = _lib.Config()
        cfg.nx, cfg.ny, cfg.nz, cfg.dim = self.nx, self.ny, self.nz, self.dim
        cfg.n_phase, cfg.kmax = self.n_phase, self.kmax
        cfg.precision = FP64 if precision == "fp64" else FP32
        cfg.temp_dense, cfg.temp_f32 = int(self.temp_dense), int(self.temp_f32)
        cfg.enable_anisotropy, cfg.enable_torque = int(bool(anisotropy)), int(bool(torque))
        cfg.halt_on_kmax = int(bool(halt_on_kmax))
        cfg.slot_capacity = int(slot_capacity)
        cfg.s = _lib.scalars_struct(scalars)
        cfg.cooling_rate = float(cooling_rate)
        self._cfg = cfg
        h = C.c_void_p()
        check(_lib.lib().mpf_create(C.byref(cfg), C.byref(h)))
        self._h = h

    # ------------------------------------------------------------ lifetime
    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().mpf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_bytes(self) -> int:
        return int(_lib.lib().mpf_device_bytes(self._h))

    # -------------------------------------------------------------- inputs
    def set_tables(self, wij, aij, mij, n111):
        N = self.n_phase
        w = np.ascontiguousarray(wij, np.float64).reshape(N, N)
        a = np.ascontiguousarray(aij, np.float64).reshape(N, N)
        m = np.ascontiguousarray(mij, np.float64).reshape(N, N)
        n = np.ascontiguousarray(n111, np.float64).reshape(N, 8, 3)
        check(_lib.lib().mpf_set_tables(self._h, ptr(w), ptr(a), ptr(m), ptr(n)))

    def upload_phi(self, phi):
        phi = np.asarray(phi)
        expect = (self.n_phase,) + self.shape
        if phi.shape != expect:
            raise ValueError(f"phi has shape {phi.shape}, expected {expect}")
        if phi.dtype == np.float64:
            a, dt = np.ascontiguousarray(phi), FP64
        else:
            a, dt = np.ascontiguousarray(phi, np.float32), FP32
        check(_lib.lib().mpf_upload_phi(self._h, a.ctypes.data_as(C.c_void_p), dt))

    def init_columns(self, grain_map, profile_index, solid_profiles, liquid_profiles):
        """Columnar state: grain grain_map[l,m] with z profile solid_profiles[profile_index[l,m]]."""
        g = np.ascontiguousarray(grain_map, np.int32).reshape(self.nx, self.ny)
        p = np.ascontiguousarray(profile_index, np.int32).reshape(self.nx, self.ny)
        s = np.ascontiguousarray(solid_profiles, np.float64).reshape(-1, self.nz)
        l = np.ascontiguousarray(liquid_profiles, np.float64).reshape(-1, self.nz)
        if s.shape != l.shape:
            raise ValueError("solid and liquid profiles must have the same shape")
        check(_lib.lib().mpf_init_columns(self._h, ptr(g, C.c_int32), ptr(p, C.c_int32),
                                          s.shape[0], ptr(s), ptr(l)))

    def upload_temp(self, temp):
        """temp: dense [nx, ny(, nz)] or, without temp_dense, a [nz] column."""
        t = np.asarray(temp, np.float64)
        if self.temp_dense:
            t = np.ascontiguousarray(np.broadcast_to(t, self.shape))
        else:
            if t.shape == self.shape:
                col = t.reshape(-1, self.nz)
                if not np.all(col == col[0]):
                    raise ValueError("temperature varies in x/y: create the solver with temp_dense=True")
                t = col[0]
            t = np.ascontiguousarray(np.broadcast_to(t, (self.nz,)))
        check(_lib.lib().mpf_upload_temp(self._h, ptr(t)))

    # ---------------------------------------------------------------- step
    def step(self, nsteps: int = 1) -> int:
        """Advance; returns the completed steps (fewer only on a KMAX halt)."""
        return check(_lib.lib().mpf_step(self._h, int(nsteps)))

    # ------------------------------------------------------------- outputs
    def phi(self, dtype=np.float64) -> np.ndarray:
        dt = FP64 if np.dtype(dtype) == np.float64 else FP32
        out = np.empty((self.n_phase,) + self.shape, np.float64 if dt == FP64 else np.float32)
        check(_lib.lib().mpf_download_phi(self._h, out.ctypes.data_as(C.c_void_p), dt))
        return out

    def temp(self) -> np.ndarray:
        n = int(np.prod(self.shape)) if self.temp_dense else self.nz
        out = np.empty(n, np.float64)
        check(_lib.lib().mpf_download_temp(self._h, ptr(out)))
        return out.reshape(self.shape) if self.temp_dense else out

    def apt(self, depth: int | None = None):
        """Checked APT of the current state: (mf[depth,...], nf, status[2])."""
        depth = int(depth or self.kmax)
        mf = np.zeros((depth,) + self.shape, np.int32)
        nf = np.zeros(self.shape, np.int32)
        st = np.zeros(2, np.int32)
        check(_lib.lib().mpf_download_apt(self._h, ptr(mf, C.c_int32), depth, ptr(nf, C.c_int32),
                                          ptr(st, C.c_int32)))
        return mf, nf, st

    def stats(self) -> dict:
        s = _lib.Stats()
        check(_lib.lib().mpf_get_stats(self._h, C.byref(s)))
        return stats_dict(s)

    def compare(self, other: "Solver", tol: float = 1e-4) -> dict:
        """Difference between this solver's phi and another's on the same grid
        (any precisions), reduced on the device in FP64: max and sum of
        |phi_a - phi_b| over cells and phases, cells whose sets of non-zero
        grains differ, cells above tol."""
        out = np.zeros(4, np.float64)
        check(_lib.lib().mpf_compare_phi(self._h, other._h, float(tol), ptr(out)))
        return {"phi_max_abs": float(out[0]), "phi_sum_abs": float(out[1]),
                "grain_set_diff_cells": int(out[2]), "cells_above_tol": int(out[3]), "tol": float(tol)}

    def set_launch(self, merge_blocks: int = 0, front_blocks: int = 0, general_blocks: int = 0,
                   side_stream: int = 2):
        """Block counts (0: default) and the mode of the passes after the
        merge pass: 0 in sequence, 1 on side streams, 2 one launch."""
        set_launch(self._h, merge_blocks, front_blocks, general_blocks, side_stream)

    def launch(self) -> dict:
        return get_launch(self._h)

    def set_graphs(self, on: bool = True):
        check(_lib.lib().mpf_set_graphs(self._h, int(bool(on))))

    def set_tiling(self, planes_per_item: int = 0, light_blocks: int = 0):
        """Classification work items of planes_per_item z-planes (4, 8, 16;
        0: by grid size) and the classification grid (0: 1024 blocks)."""
        set_tiling(self._h, planes_per_item, light_blocks)

    def autotune(self, trial_steps: int = 4) -> dict:
        return autotune(self._h, trial_steps)

    def phase_map(self, axis: str = "xy", index: int | None = None) -> np.ndarray:
        """Dominant phase id over one plane (np.argmax over phases, lowest id
        on ties): axis 'xy' at z = index -> [nx, ny]; 'xz' at y -> [nx, nz];
        'yz' at x -> [ny, nz].  index None: the middle plane."""
        from .plot import AXES, resolve_slice_index
        nx, ny, nz = self.shape
        idx = resolve_slice_index(axis, self.shape, index)
        shp = {"xy": (nx, ny), "xz": (nx, nz), "yz": (ny, nz)}[axis]
        out = np.empty(shp, np.int32)
        check(_lib.lib().mpf_phase_map(self._h, AXES[axis], idx, ptr(out, C.c_int32)))
        return out

    def directional_radii(self, center, phase: int = 1, threshold: float = 0.5, step_cells: float = 0.5) -> dict:
        """Directional radii (cells) of phase `phase` about `center` (cells),
        from device ray probes (mpf.observe.directional_radii); 3D only."""
        from .observe import directional_radii
        return directional_radii(self._h, [0], [center], phase, threshold, step_cells)[0]

    def set_kernel_timing(self, on: bool = True):
        """Time each kernel of every step with HIP events (summed in stats()["kernel_ms"])."""
        check(_lib.lib().mpf_set_kernel_timing(self._h, int(bool(on))))

    def reset_stats(self):
        check(_lib.lib().mpf_reset_stats(self._h))

    def observables(self, dx: float | None = None) -> dict:
        """FP64 reductions on the device (tests/verification_3d.py:612-654)."""
        N = self.n_phase
        gv = np.zeros(N, np.float64)
        front = np.zeros((self.nx, self.ny), np.int32)
        out = np.zeros(7, np.float64)
        check(_lib.lib().mpf_observables(self._h, ptr(gv), ptr(front, C.c_int32), ptr(out)))
        return summarize_observables(gv, front, out, dx)


def set_launch(h, merge_blocks, front_blocks, general_blocks, side_stream):
    check(_lib.lib().mpf_set_launch(h, int(merge_blocks), int(front_blocks), int(general_blocks),
                                    int(side_stream)))


def get_launch(h) -> dict:
    out = np.zeros(4, np.int32)
    check(_lib.lib().mpf_get_launch(h, ptr(out, C.c_int32)))
    til = np.zeros(2, np.int32)
    check(_lib.lib().mpf_get_tiling(h, ptr(til, C.c_int32)))
    return {"merge_blocks": int(out[0]), "front_blocks": int(out[1]), "general_blocks": int(out[2]),
            "side_stream": int(out[3]), "light_planes": int(til[0]), "light_blocks": int(til[1])}


def set_tiling(h, planes_per_item, light_blocks):
    check(_lib.lib().mpf_set_tiling(h, int(planes_per_item), int(light_blocks)))


def autotune(h, trial_steps) -> dict:
    out = np.zeros(5, np.int32)
    check(_lib.lib().mpf_autotune(h, int(trial_steps), ptr(out, C.c_int32)))
    r = get_launch(h)
    r["us_per_step"] = int(out[4])
    return r


def stats_dict(s) -> dict:
    return {"max_active": s.max_active, "kmax_overflow": s.kmax_overflow,
            "overflow_cells": s.overflow_cells, "general_cells": s.general_cells,
            "halt_step": s.halt_step, "steps_done": s.steps_done,
            "last_step_ms": s.last_step_ms, "merge_cells": s.merge_cells,
            "front_cells": s.front_cells, "class_tables": bool(s.class_tables),
            "kernel_steps": s.kernel_steps, "light_cells": s.light_cells, "n111_pairs": bool(s.n111_pairs),
            "kernel_ms": dict(zip(("light", "merge", "front", "general"), list(s.kernel_ms)))}


def summarize_observables(gv, front, out, dx) -> dict:
    """Observables from the device reductions: per-phase sums, the topmost
    solid cell per column (interface_positions_3d, tests/verification_3d.py:409-416)
    and the scalar partials."""
    scale = float(dx) ** 3 if dx is not None else 1.0
    f = front.astype(np.float64)
    f[front < 0] = np.nan
    finite = f[np.isfinite(f)]
    return {
        "phase_sums": gv,
        "grain_volumes": gv[1:] * scale,
        "phi_sum_error_max": out[0], "phi_min": out[1], "phi_max": out[2],
        "nan_count": int(out[3]), "solid_sum": out[4], "cells": int(out[5]),
        "stored_slots": int(out[6]),
        "solid_fraction": out[4] / out[5],
        "front": f,
        "front_mean": float(np.mean(finite)) if finite.size else float("nan"),
        "front_max": float(np.max(finite)) if finite.size else float("nan"),
        "front_std": float(np.std(finite)) if finite.size else float("nan"),
    }
