"""Device arrays and the drop-in kernels on the reference's dense layout.

DeviceArray mirrors the part of Numba's DeviceNDArray the reference uses
(cuda.to_device, copy_to_host, copy_to_device; run_randommode_3d.py:244-290)
on memory from the library's own allocator.  apt/update/temp call the
mpf_dense_* entry points, one per reference kernel.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FP32, FP64, check


class DeviceArray:
    """A C-contiguous array in device memory."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.nbytes = self.size * self.dtype.itemsize
        p = C.c_void_p()
        check(_lib.lib().mpf_malloc(C.byref(p), max(self.nbytes, 1)))
        self._ptr = p

    @property
    def ptr(self) -> C.c_void_p:
        return self._ptr

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def copy_to_device(self, ary):
        a = np.ascontiguousarray(ary, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} does not match device array {self.shape}")
        check(_lib.lib().mpf_memcpy_htod(self._ptr, a.ctypes.data_as(C.c_void_p), self.nbytes))

    def copy_to_host(self, ary=None) -> np.ndarray:
        out = np.empty(self.shape, self.dtype) if ary is None else ary
        check(_lib.lib().mpf_memcpy_dtoh(out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes))
        return out

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p is not None and p.value:
            try:
                _lib.lib().mpf_free(p)
            except Exception:
                pass
            self._ptr = None


def to_device(ary) -> DeviceArray:
    a = np.ascontiguousarray(ary)
    d = DeviceArray(a.shape, a.dtype)
    d.copy_to_device(a)
    return d


def device_array(shape, dtype=np.float32) -> DeviceArray:
    return DeviceArray(shape, dtype)


def synchronize():
    check(_lib.lib().mpf_device_synchronize())


def _dt(a: DeviceArray) -> int:
    if a.dtype == np.float64:
        return FP64
    if a.dtype == np.float32:
        return FP32
    raise TypeError(f"unsupported dtype {a.dtype}")


def _grid(phi: DeviceArray):
    if phi.ndim == 4:
        return 3, phi.shape[1], phi.shape[2], phi.shape[3]
    if phi.ndim == 3:
        return 2, phi.shape[1], phi.shape[2], 1
    raise ValueError("phi must be [N, nx, ny, nz] or [N, nx, ny]")


def _check_shapes(name, phi, n_phase, mf, nf, status=None, phi_new=None, temp=None, tables=None):
    """The entry points see only pointers and sizes, so every array the
    kernels index is checked against the grid and phase count here, before a
    launch: a mismatch would otherwise read or write out of bounds."""
    n_phase = int(n_phase)
    grid = phi.shape[1:]
    if n_phase < 1 or phi.shape[0] < n_phase:
        raise ValueError(f"{name}: phi holds {phi.shape[0]} phases, number_of_grain is {n_phase}")
    if phi_new is not None and phi_new.shape != phi.shape:
        raise ValueError(f"{name}: phi_new {phi_new.shape} does not match phi {phi.shape}")
    if mf.ndim != phi.ndim or tuple(mf.shape[1:]) != grid or mf.shape[0] < 1 or mf.dtype != np.int32:
        raise ValueError(f"{name}: mf must be int32 [depth, {', '.join(map(str, grid))}], got {mf.dtype} {mf.shape}")
    if tuple(nf.shape) != grid or nf.dtype != np.int32:
        raise ValueError(f"{name}: nf must be int32 {grid}, got {nf.dtype} {nf.shape}")
    if status is not None and (status.size < 2 or status.dtype != np.int32):
        raise ValueError(f"{name}: status must be int32 with 2 entries, got {status.dtype} {status.shape}")
    if temp is not None and temp.size != int(np.prod(grid)):
        raise ValueError(f"{name}: temp {temp.shape} does not match the grid {grid}")
    if tables is not None:
        wij, aij, mij, n111 = tables
        for tn, t in (("wij", wij), ("aij", aij), ("mij", mij)):
            if tuple(t.shape) != (n_phase, n_phase):
                raise ValueError(f"{name}: {tn} must be [{n_phase}, {n_phase}], got {t.shape}")
        if tuple(n111.shape) != (n_phase, 8, 3):
            raise ValueError(f"{name}: n111 must be [{n_phase}, 8, 3], got {n111.shape}")


def apt(phi: DeviceArray, mf: DeviceArray, nf: DeviceArray, status: DeviceArray | None, n_phase: int,
        checked: bool = True, kmax: int = 50):
    """kernel_update_nfmf_3d(_checked) / kernel_update_nfmf."""
    dim, nx, ny, nz = _grid(phi)
    _check_shapes("apt", phi, n_phase, mf, nf, status=status if checked else None)
    if checked and status is None:
        raise ValueError("apt: the checked APT needs a status array")
    check(_lib.lib().mpf_dense_apt(phi.ptr, _dt(phi), mf.ptr, nf.ptr, status.ptr if status else None,
                                   int(n_phase), nx, ny, nz, dim, mf.shape[0], int(kmax),
                                   int(bool(checked)), None))


def update(phi: DeviceArray, phi_new: DeviceArray, temp: DeviceArray, mf: DeviceArray, nf: DeviceArray,
           wij: DeviceArray, aij: DeviceArray, mij: DeviceArray, n111: DeviceArray, scalars: dict,
           n_phase: int, kmax: int = 50, anisotropy: bool = True, torque: bool = True):
    """kernel_update_phasefield_active_3d(_switchable) / kernel_update_phasefield_active."""
    dim, nx, ny, nz = _grid(phi)
    for t in (wij, aij, mij, n111, phi_new):
        if t.dtype != phi.dtype:
            raise TypeError("phi, phi_new and the tables must share one dtype")
    _check_shapes("update", phi, n_phase, mf, nf, phi_new=phi_new, temp=temp, tables=(wij, aij, mij, n111))
    s = _lib.scalars_struct(scalars)
    check(_lib.lib().mpf_dense_update(phi.ptr, phi_new.ptr, _dt(phi), temp.ptr, _dt(temp), mf.ptr, nf.ptr,
                                      wij.ptr, aij.ptr, mij.ptr, n111.ptr, C.byref(s), int(n_phase),
                                      nx, ny, nz, dim, mf.shape[0], int(kmax), int(bool(anisotropy)),
                                      int(bool(torque)), None))


def temp(t: DeviceArray, cooling_rate: float):
    """kernel_update_temp_3d / kernel_update_temp."""
    check(_lib.lib().mpf_dense_temp(t.ptr, _dt(t), float(cooling_rate), t.size, None))
