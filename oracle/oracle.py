"""ctypes front end of the CPU oracle (oracle/mpf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP solver.

Arrays use the reference's dense layout: phi[N, nx, ny, nz] (3D) or
phi[N, nx, ny] (2D), float64, C order; temp is dense [nx, ny(, nz)].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("N", C.c_int),
        ("dx", C.c_double), ("dt", C.c_double), ("T_melt", C.c_double), ("Sf", C.c_double),
        ("eps0_sl", C.c_double), ("w0_sl", C.c_double),
        ("a0", C.c_double), ("delta_a", C.c_double), ("mu_a", C.c_double),
        ("p_round", C.c_double), ("g2_floor", C.c_double), ("ksi", C.c_double),
        ("theta_c_rad", C.c_double),
        ("use_libm", C.c_int), ("faithful", C.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liborc.so with oracle/Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        P = C.POINTER(OrcParams)
        L.orc_apt3.argtypes = [dp, ip, ip, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_apt2.argtypes = [dp, ip, ip, ip, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_update3.argtypes = [dp, dp, dp, ip, ip, C.c_int, dp, dp, dp, dp, P, C.c_int, C.c_int]
        L.orc_update2.argtypes = [dp, dp, dp, ip, ip, C.c_int, dp, dp, dp, dp, P]
        L.orc_temp.argtypes = [dp, C.c_size_t, C.c_double, C.c_int]
        L.orc_probe3.argtypes = [dp, dp, P, C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.orc_calc_a.argtypes = [P, C.c_double]
        L.orc_calc_a.restype = C.c_double
        L.orc_calc_b.argtypes = [P, C.c_double]
        L.orc_calc_b.restype = C.c_double
        L.orc_run.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp, P,
                              C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, ip]
        L.orc_run.restype = C.c_int
        L.orc_threads.restype = C.c_int
        for fn in ("orc_mpf_tan", "orc_mpf_acos", "orc_mpf_tanh"):
            getattr(L, fn).argtypes = [C.c_double]
            getattr(L, fn).restype = C.c_double
        _lib = L
    return _lib


PARAM_KEYS = ("dx", "dt", "T_melt", "Sf", "eps0_sl", "w0_sl", "a0", "delta_a", "mu_a",
              "p_round", "g2_floor", "ksi", "theta_c_rad")


def make_params(shape, N: int, scalars: dict, use_libm: bool = False, faithful: bool = False) -> OrcParams:
    p = OrcParams()
    p.nx = int(shape[0])
    p.ny = int(shape[1])
    p.nz = int(shape[2]) if len(shape) > 2 else 1
    p.N = int(N)
    for k in PARAM_KEYS:
        setattr(p, k, float(scalars[k]))
    p.use_libm = 1 if use_libm else 0
    p.faithful = 1 if faithful else 0
    return p


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def apt(phi, kmax: int, depth: int | None = None):
    """Checked APT (src/gpu_kernels_3d.py:489-527) -> (mf[depth,...], nf, status[2])."""
    phi = _f64(phi)
    N = phi.shape[0]
    grid = phi.shape[1:]
    depth = int(depth or kmax)
    mf = np.zeros((depth,) + grid, np.int32)
    nf = np.zeros(grid, np.int32)
    st = np.zeros(2, np.int32)
    if len(grid) == 3:
        lib().orc_apt3(phi, mf, nf, st, depth, kmax, grid[0], grid[1], grid[2], N)
    else:
        lib().orc_apt2(phi, mf, nf, st, depth, kmax, grid[0], grid[1], N)
    return mf, nf, st


def update(phi, temp, mf, nf, tables, scalars, aniso=True, torque=True, use_libm=False):
    """One update over dense phi given the APT result (returns phi_new)."""
    phi = _f64(phi)
    N = phi.shape[0]
    grid = phi.shape[1:]
    p = make_params(grid, N, scalars, use_libm)
    out = np.empty_like(phi)
    wij, aij, mij, n111 = (_f64(t) for t in tables)
    mf = np.ascontiguousarray(mf, np.int32)
    nf = np.ascontiguousarray(nf, np.int32)
    temp = _f64(np.broadcast_to(temp, grid))
    if len(grid) == 3:
        lib().orc_update3(phi, out, temp, mf, nf, mf.shape[0], wij, aij, mij, n111, C.byref(p),
                          int(bool(aniso)), int(bool(torque)))
    else:
        lib().orc_update2(phi, out, temp, mf, nf, mf.shape[0], wij, aij, mij, n111, C.byref(p))
    return out


def run(phi, temp, nsteps, tables, scalars, kmax=50, aniso=True, torque=True,
        cooling_rate=0.0, temp_f32=False, halt_on_kmax=False, use_libm=False, faithful=False):
    """Verification-style loop on the CPU.  Returns (phi, temp, steps_done, status).
    faithful (3D): float32 rounding at the reference's truncation points,
    including the float32 phi stores; the initial phi is rounded to float32
    as the reference's to_device of a float32 array would hold it."""
    phi = _f64(phi).copy()
    N = phi.shape[0]
    grid = phi.shape[1:]
    if faithful:
        if len(grid) != 3:
            raise ValueError("the faithful mode covers the 3D update")
        phi = phi.astype(np.float32).astype(np.float64)
    temp = _f64(np.broadcast_to(temp, grid)).copy()
    p = make_params(grid, N, scalars, use_libm, faithful)
    wij, aij, mij, n111 = (_f64(t) for t in tables)
    st = np.zeros(2, np.int32)
    done = lib().orc_run(len(grid), phi, temp, int(nsteps), int(kmax), wij, aij, mij, n111,
                         C.byref(p), int(bool(aniso)), int(bool(torque)), float(cooling_rate),
                         int(bool(temp_f32)), int(bool(halt_on_kmax)), st)
    if done < 0:
        raise MemoryError("oracle run could not allocate its buffers")
    return phi, temp, done, st


def threads() -> int:
    """OpenMP threads the oracle's loops use."""
    return int(lib().orc_threads())


def probe3(phi, n111, scalars, l, m, k, s, use_libm=False):
    """Per-solid terms at one cell: dict of aniso_term1, torque, b, A13, A12, best_cos."""
    phi = _f64(phi)
    p = make_params(phi.shape[1:], phi.shape[0], scalars, use_libm)
    out = np.zeros(11)
    lib().orc_probe3(phi, _f64(n111), C.byref(p), l, m, k, s, out)
    return {"aniso_term1": out[0], "torque": out[1], "b_sl": out[2], "A13": out[3:6].copy(),
            "A12": out[6:10].copy(), "best_cos": out[10]}


def calc_a(c, scalars, use_libm=False):
    p = make_params((1, 1, 1), 1, scalars, use_libm)
    return lib().orc_calc_a(C.byref(p), float(c))


def calc_b(c, scalars, use_libm=False):
    p = make_params((1, 1, 1), 1, scalars, use_libm)
    return lib().orc_calc_b(C.byref(p), float(c))
