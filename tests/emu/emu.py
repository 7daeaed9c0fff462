"""Python driver of the CPU emulation build (tests/emu/mpf_emu.cpp).

TEST INFRASTRUCTURE: packs the reference's dense phi into the solver's slot
layout, runs one step of the device per-cell code on the CPU and unpacks."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmpf_emu.so")
_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B" if force else "all", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(LIB)
        _lib.emu_step.restype = C.c_int
    return _lib


def pack(phi):
    """Dense [N, nx, ny, nz] -> (liq, cnt, sid[scap, ncell], sval[scap, ncell]) in
    cell order c = (k*ny + m)*nx + l."""
    N = phi.shape[0]
    grid = phi.shape[1:]
    a = np.ascontiguousarray(np.transpose(phi, (0,) + tuple(range(len(grid), 0, -1)))).reshape(N, -1)
    liq = a[0].copy()
    sol = a[1:]
    nz_mask = sol != 0
    cnt = nz_mask.sum(0).astype(np.uint8)
    scap = max(int(cnt.max()) if cnt.size else 1, 1)
    pos = np.cumsum(nz_mask, axis=0) - 1
    sid = np.zeros((scap, liq.size), np.uint16)
    sval = np.zeros((scap, liq.size), phi.dtype)
    g, c = np.nonzero(nz_mask)
    sid[pos[g, c], c] = g + 1
    sval[pos[g, c], c] = sol[g, c]
    return liq, cnt, sid, sval


def unpack(liq, cnt, sid, sval, N, grid):
    ncell = liq.size
    out = np.zeros((N, ncell), liq.dtype)
    out[0] = liq
    for s in range(sid.shape[0]):
        sel = cnt > s
        out[sid[s, sel].astype(np.int64), np.nonzero(sel)[0]] = sval[s, sel]
    rev = tuple(reversed(grid))
    out = out.reshape((N,) + rev)
    return np.ascontiguousarray(np.transpose(out, (0,) + tuple(range(len(grid), 0, -1))))


def step(phi, T_cooled, tables, scalars, kmax=50, aniso=True, torque=True, precision="fp64", mode=0,
         use_class=False, pairs=False):
    """One step of the device code on the CPU.  T_cooled is the dense [nx, ny, nz]
    temperature after this step's cooling.  use_class selects the class-form
    table code (the tables must have that form); pairs the four-row facet scan
    (the n111 rows must come in opposite pairs).  Returns (phi_new, path
    counts, per-cell path)."""
    from mpf import _lib as L

    dt = np.float64 if precision == "fp64" else np.float32
    phi = np.asarray(phi, dt)
    N = phi.shape[0]
    grid = phi.shape[1:]
    dim = len(grid)
    nx, ny = grid[0], grid[1]
    nz = grid[2] if dim == 3 else 1
    liq, cnt, sid, sval = pack(phi)
    scap = max(sid.shape[0], min(kmax, N - 1), 2)   # the device code reads slots 0 and 1 unconditionally
    sid = np.vstack([sid, np.zeros((scap - sid.shape[0], liq.size), np.uint16)])
    sval = np.vstack([sval, np.zeros((scap - sval.shape[0], liq.size), dt)])
    tab = np.concatenate([np.asarray(t, np.float64).ravel() for t in tables]).astype(dt)
    T = np.ascontiguousarray(np.transpose(np.broadcast_to(np.asarray(T_cooled, np.float64), grid),
                                          tuple(range(dim - 1, -1, -1)))).ravel()
    lo, co = np.zeros_like(liq), np.zeros_like(cnt)
    so, vo = np.zeros_like(sid), np.zeros_like(sval)
    paths = np.zeros(4, np.int32)
    path_of = np.zeros(liq.size, np.int8)
    s = L.scalars_struct(scalars)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = lib().emu_step(C.c_int(1 if precision == "fp64" else 0), C.c_int(dim), C.byref(s), nx, ny, nz, N, kmax,
                        int(aniso), int(torque), vp(tab), vp(liq), vp(cnt), vp(sid), vp(sval), vp(lo), vp(co),
                        vp(so), vp(vo), T.ctypes.data_as(C.POINTER(C.c_double)), mode,
                        paths.ctypes.data_as(C.POINTER(C.c_int)), vp(path_of),
                        int(bool(use_class)) | (2 if pairs else 0))
    if rc == -2:
        raise ValueError("the n111 rows are not in opposite pairs")
    if rc != 0:
        raise ValueError("the tables do not have the class form")
    po = np.ascontiguousarray(np.transpose(path_of.reshape(tuple(reversed(grid))), tuple(range(dim - 1, -1, -1))))
    return unpack(lo, co, so, vo, N, grid), paths, po
