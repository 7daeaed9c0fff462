"""A batch of independent 3D grids stepped together on the GPU.

The reference runs its bicrystal and orientation sweeps one grid at a time
(run_twomode_3d.py:200-322, run_orientation_scan_3d.py:495-841).  Here all
grids of a sweep live in one set of device buffers and every kernel launch
of a step covers all of them.  The grids share the physics scalars, N, KMAX
and precision; each has its own size, tables (orientations) and temperature.
Arrays cross the boundary per grid in the reference's dense layouts."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FP32, FP64, check, ptr


class Batch:
    def __init__(self, shapes, n_phase: int, scalars: dict, *, kmax: int = 50, precision: str = "fp32",
                 anisotropy: bool = True, torque: bool = True, cooling_rate: float = 0.0,
                 temp_f32: bool = False, slot_capacity: int = 0):
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        if not self.shapes or any(len(s) != 3 for s in self.shapes):
            raise ValueError("shapes must be a non-empty list of (nx, ny, nz)")
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.n_phase = int(n_phase)
        self.kmax = int(kmax)
        self.precision = precision
        cfg = _lib.Config()
        cfg.nx, cfg.ny, cfg.nz, cfg.dim = self.shapes[0][0], self.shapes[0][1], self.shapes[0][2], 3
        cfg.n_phase, cfg.kmax = self.n_phase, self.kmax
        cfg.precision = FP64 if precision == "fp64" else FP32
        cfg.temp_dense, cfg.temp_f32 = 0, int(bool(temp_f32))
        cfg.enable_anisotropy, cfg.enable_torque = int(bool(anisotropy)), int(bool(torque))
        cfg.halt_on_kmax = 0
        cfg.slot_capacity = int(slot_capacity)
        cfg.s = _lib.scalars_struct(scalars)
        cfg.cooling_rate = float(cooling_rate)
        dims = np.ascontiguousarray(self.shapes, np.int32).reshape(-1)
        h = C.c_void_p()
        check(_lib.lib().mpf_batch_create(C.byref(cfg), len(self.shapes), ptr(dims, C.c_int32), C.byref(h)))
        self._h = h

    def __len__(self):
        return len(self.shapes)

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

    @property
    def cells(self) -> int:
        return int(sum(nx * ny * nz for nx, ny, nz in self.shapes))

    def set_tables(self, i: int, wij, aij, mij, n111):
        N = self.n_phase
        w = np.ascontiguousarray(wij, np.float64).reshape(N, N)
        a = np.ascontiguousarray(aij, np.float64).reshape(N, N)
        m = np.ascontiguousarray(mij, np.float64).reshape(N, N)
        n = np.ascontiguousarray(n111, np.float64).reshape(N, 8, 3)
        check(_lib.lib().mpf_batch_set_tables(self._h, int(i), ptr(w), ptr(a), ptr(m), ptr(n)))

    def upload_phi(self, i: int, phi):
        phi = np.asarray(phi)
        expect = (self.n_phase,) + self.shapes[i]
        if phi.shape != expect:
            raise ValueError(f"phi has shape {phi.shape}, expected {expect}")
        a, dt = (np.ascontiguousarray(phi), FP64) if phi.dtype == np.float64 else \
            (np.ascontiguousarray(phi, np.float32), FP32)
        check(_lib.lib().mpf_batch_upload_phi(self._h, int(i), a.ctypes.data_as(C.c_void_p), dt))

    def init_columns(self, i: int, grain_map, profile_index, solid_profiles, liquid_profiles):
        nx, ny, nz = self.shapes[i]
        g = np.ascontiguousarray(grain_map, np.int32).reshape(nx, ny)
        p = np.ascontiguousarray(profile_index, np.int32).reshape(nx, ny)
        s = np.ascontiguousarray(solid_profiles, np.float64).reshape(-1, nz)
        l = np.ascontiguousarray(liquid_profiles, np.float64).reshape(-1, nz)
        check(_lib.lib().mpf_batch_init_columns(self._h, int(i), ptr(g, C.c_int32), ptr(p, C.c_int32),
                                                s.shape[0], ptr(s), ptr(l)))

    def upload_temp(self, i: int, temp):
        """Temperature of grid i: a [nz] column, or a field uniform in x and y."""
        nz = self.shapes[i][2]
        t = np.asarray(temp, np.float64)
        if t.shape == self.shapes[i]:
            col = t.reshape(-1, nz)
            if not np.all(col == col[0]):
                raise ValueError("a batch keeps one temperature per z-plane")
            t = col[0]
        t = np.ascontiguousarray(np.broadcast_to(t, (nz,)))
        check(_lib.lib().mpf_batch_upload_temp(self._h, int(i), ptr(t)))

    def step(self, nsteps: int = 1) -> int:
        return check(_lib.lib().mpf_step(self._h, int(nsteps)))

    def phi(self, i: int, dtype=np.float64) -> np.ndarray:
        dt = FP64 if np.dtype(dtype) == np.float64 else FP32
        out = np.empty((self.n_phase,) + self.shapes[i], np.float64 if dt == FP64 else np.float32)
        check(_lib.lib().mpf_batch_download_phi(self._h, int(i), out.ctypes.data_as(C.c_void_p), dt))
        return out

    def temp(self, i: int) -> np.ndarray:
        out = np.empty(self.shapes[i][2], np.float64)
        check(_lib.lib().mpf_batch_download_temp(self._h, int(i), ptr(out)))
        return out

    def apt(self, i: int, depth: int | None = None):
        depth = int(depth or self.kmax)
        shape = self.shapes[i]
        mf = np.zeros((depth,) + shape, np.int32)
        nf = np.zeros(shape, np.int32)
        st = np.zeros(2, np.int32)
        check(_lib.lib().mpf_batch_download_apt(self._h, int(i), ptr(mf, C.c_int32), depth, ptr(nf, C.c_int32),
                                                ptr(st, C.c_int32)))
        return mf, nf, st

    def observables(self, i: int, dx: float | None = None) -> dict:
        from .solver import summarize_observables
        nx, ny, _ = self.shapes[i]
        gv = np.zeros(self.n_phase, np.float64)
        front = np.zeros((nx, ny), np.int32)
        out = np.zeros(7, np.float64)
        check(_lib.lib().mpf_batch_observables(self._h, int(i), ptr(gv), ptr(front, C.c_int32), ptr(out)))
        return summarize_observables(gv, front, out, dx)

    def observables_all(self, with_front: bool = True):
        """(phase sums [n, N], scalars [n, 7], fronts) of every grid at once
        (mpf.observe.observables_all)."""
        from .observe import observables_all
        return observables_all(self._h, self.shapes, self.n_phase, with_front)

    def phase_map(self, i: int, axis: str = "xy", index: int | None = None) -> np.ndarray:
        """Dominant phase id over one plane of grid i (Solver.phase_map)."""
        from .plot import AXES, resolve_slice_index
        nx, ny, nz = self.shapes[i]
        idx = resolve_slice_index(axis, self.shapes[i], index)
        shp = {"xy": (nx, ny), "xz": (nx, nz), "yz": (ny, nz)}[axis]
        out = np.empty(shp, np.int32)
        check(_lib.lib().mpf_batch_phase_map(self._h, int(i), AXES[axis], idx, ptr(out, C.c_int32)))
        return out

    def directional_radii(self, centers, phase: int = 1, threshold: float = 0.5, step_cells: float = 0.5,
                          grids=None) -> list[dict]:
        """Directional radii (cells) of phase `phase` about each grid's centre
        (mpf.observe.directional_radii)."""
        from .observe import directional_radii
        grids = range(len(self.shapes)) if grids is None else grids
        return directional_radii(self._h, grids, centers, phase, threshold, step_cells)

    def stats(self) -> dict:
        from .solver import stats_dict
        s = _lib.Stats()
        check(_lib.lib().mpf_get_stats(self._h, C.byref(s)))
        return stats_dict(s)

    def reset_stats(self):
        check(_lib.lib().mpf_reset_stats(self._h))

    def set_launch(self, merge_blocks: int = 0, front_blocks: int = 0, general_blocks: int = 0,
                   side_stream: int = 2):
        from .solver import set_launch
        set_launch(self._h, merge_blocks, front_blocks, general_blocks, side_stream)

    def launch(self) -> dict:
        from .solver import get_launch
        return get_launch(self._h)

    def set_graphs(self, on: bool = True):
        check(_lib.lib().mpf_set_graphs(self._h, int(bool(on))))

    def set_tiling(self, planes_per_item: int = 0, light_blocks: int = 0):
        from .solver import set_tiling
        set_tiling(self._h, planes_per_item, light_blocks)

    def autotune(self, trial_steps: int = 4) -> dict:
        from .solver import autotune
        return autotune(self._h, trial_steps)

    def set_kernel_timing(self, on: bool = True):
        check(_lib.lib().mpf_set_kernel_timing(self._h, int(bool(on))))
