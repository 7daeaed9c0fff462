"""Kernel objects in the reference's launch idiom, over the dense drop-ins.

The reference launches Numba kernels as ``kernel[blocks, threads](*args)``
with positional argument lists (src/gpu_kernels_3d.py:463, :490, :531-537,
:684-692, :834; src/gpu_kernels.py:338, :362-368, :521).  The objects here
accept the same launch syntax and the same argument lists and run the
library's mpf_dense_* entry points on arrays from ``mpf.cuda``.  A launch must
cover the whole grid (the reference's callers always size it so); one that
does not is rejected instead of silently leaving cells untouched.
"""
from __future__ import annotations

import math

import numpy as np

from . import dense
from .dense import DeviceArray

SCALAR_NAMES = ("dx", "dt", "T_melt", "Sf", "eps0_sl", "w0_sl", "a0", "delta_a", "mu_a", "p_round",
                "g2_floor", "ksi", "theta_c_rad")


def _as_tuple(v):
    if isinstance(v, (tuple, list)):
        return tuple(int(x) for x in v)
    return (int(v),)


def _check_cover(name, blocks, threads, extents):
    b, t = _as_tuple(blocks), _as_tuple(threads)
    if len(b) != len(extents) or len(t) != len(extents):
        raise ValueError(f"{name}: launch needs {len(extents)}-D blocks and threads, got {blocks} x {threads}")
    for bi, ti, n in zip(b, t, extents):
        if bi < 1 or ti < 1:
            raise ValueError(f"{name}: empty launch {blocks} x {threads}")
        if bi * ti < n:
            raise ValueError(f"{name}: launch {blocks} x {threads} does not cover the grid {extents}")


def _dev(name, a):
    if not isinstance(a, DeviceArray):
        raise TypeError(f"{name}: arrays must be device arrays from mpf.cuda (got {type(a).__name__})")
    return a


class _Kernel:
    """``kernel[blocks, threads](*args)``"""

    def __init__(self, name, dim, nargs, run):
        self.__name__ = name
        self._dim, self._nargs, self._run = dim, nargs, run

    def __getitem__(self, cfg):
        if not isinstance(cfg, tuple) or len(cfg) < 2:
            raise TypeError(f"{self.__name__}: launch as kernel[blocks, threads](...)")
        blocks, threads = cfg[0], cfg[1]

        def launch(*args):
            if len(args) != self._nargs:
                raise TypeError(f"{self.__name__} takes {self._nargs} arguments ({len(args)} given)")
            self._run(self.__name__, blocks, threads, *args)

        return launch

    def __repr__(self):
        return f"<mpf kernel {self.__name__}>"


def _grid_args(dim, args):
    return tuple(int(v) for v in args[:dim])


def make_kernels(dim: int, kmax: int):
    """The reference kernel set of one dimension (3: src/gpu_kernels_3d.py, 2: src/gpu_kernels.py)."""

    def nfmf(name, blocks, threads, phi, mf, nf, *rest, status=None):
        ext = _grid_args(dim, rest)
        n_phase = int(rest[dim])
        _check_cover(name, blocks, threads, ext)
        dense.apt(_dev(name, phi), _dev(name, mf), _dev(name, nf), status, n_phase,
                  checked=status is not None, kmax=kmax)

    def nfmf_checked(name, blocks, threads, phi, mf, nf, status, *rest):
        nfmf(name, blocks, threads, phi, mf, nf, *rest, status=_dev(name, status))

    def update(name, blocks, threads, *args, switchable=False):
        phi, phi_new, temp, mf, nf, wij, aij, mij, n111 = (_dev(name, a) for a in args[:9])
        ext = _grid_args(dim, args[9:])
        n_phase = int(args[9 + dim])
        vals = args[10 + dim:10 + dim + len(SCALAR_NAMES)]
        scalars = {k: float(v) for k, v in zip(SCALAR_NAMES, vals)}
        aniso = torque = True
        if switchable:
            aniso, torque = bool(int(args[-2])), bool(int(args[-1]))
        _check_cover(name, blocks, threads, ext)
        dense.update(phi, phi_new, temp, mf, nf, wij, aij, mij, n111, scalars, n_phase, kmax=kmax,
                     anisotropy=aniso, torque=torque)

    def update_sw(name, blocks, threads, *args):
        update(name, blocks, threads, *args, switchable=True)

    def temp(name, blocks, threads, t, cooling_rate, *ext):
        ext = tuple(int(v) for v in ext)
        _check_cover(name, blocks, threads, ext)
        t = _dev(name, t)
        if int(np.prod(ext)) != t.size:
            raise ValueError(f"{name}: grid {ext} does not match the temperature array {t.shape}")
        dense.temp(t, float(cooling_rate))

    nupd = 9 + dim + 1 + len(SCALAR_NAMES)
    if dim == 3:
        return {
            "kernel_update_nfmf_3d": _Kernel("kernel_update_nfmf_3d", 3, 7, nfmf),
            "kernel_update_nfmf_3d_checked": _Kernel("kernel_update_nfmf_3d_checked", 3, 8, nfmf_checked),
            "kernel_update_phasefield_active_3d": _Kernel("kernel_update_phasefield_active_3d", 3, nupd, update),
            "kernel_update_phasefield_active_3d_switchable": _Kernel(
                "kernel_update_phasefield_active_3d_switchable", 3, nupd + 2, update_sw),
            "kernel_update_temp_3d": _Kernel("kernel_update_temp_3d", 3, 5, temp),
        }
    return {
        "kernel_update_nfmf": _Kernel("kernel_update_nfmf", 2, 6, nfmf),
        "kernel_update_phasefield_active": _Kernel("kernel_update_phasefield_active", 2, nupd, update),
        "kernel_update_temp": _Kernel("kernel_update_temp", 2, 4, temp),
    }


def blocks_for(extents, threads):
    """ceil(n / tpb) per dimension, as the reference's drivers compute it."""
    return tuple(math.ceil(int(n) / int(t)) for n, t in zip(extents, _as_tuple(threads)))
