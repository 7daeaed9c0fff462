"""Host-side inputs of the time step: derived constants, pairwise tables,
rotated {111} normals and initial conditions.

Each builder reproduces the reference's values bit for bit (same NumPy and
SciPy calls in the same dtypes) so that the GPU solver and the CPU oracle can
be fed exactly what the reference feeds its kernels.  Function names follow
the reference so callers read the same.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.spatial.transform import Rotation

LIQ = 0

# ----------------------------------------------------------------- constants


def eps_from_gamma(delta: float, gamma: float) -> float:
    """run_randommode_3d.py:56-57"""
    return math.sqrt(8.0 * delta * gamma / (math.pi * math.pi))


def w_from_gamma(delta: float, gamma: float) -> float:
    """run_randommode_3d.py:60-61"""
    return 4.0 * gamma / delta


def mij_from_M(delta: float, mobility: float) -> float:
    """run_randommode_3d.py:64-65"""
    return (math.pi * math.pi / (8.0 * delta)) * mobility


def derive_constants(p: dict) -> dict:
    """tests/verification_3d.py:177-204 from a flat parameter dict with keys
    dx, dt, G, V_pulling, delta_factor, gamma_100, gamma_GB, a0_deg, omg_deg,
    M_SL, M_GB_ratio."""
    delta = p["delta_factor"] * p["dx"]
    m_gb = p["M_SL"] * p["M_GB_ratio"]
    return {
        "delta": delta,
        "a0": math.radians(p["a0_deg"]),
        "theta_c_rad": math.radians(p["omg_deg"]),
        "eps0_sl": eps_from_gamma(delta, p["gamma_100"]),
        "w0_sl": w_from_gamma(delta, p["gamma_100"]),
        "m_sl_phi": mij_from_M(delta, p["M_SL"]),
        "eps_gb": eps_from_gamma(delta, p["gamma_GB"]),
        "w_gb": w_from_gamma(delta, p["gamma_GB"]),
        "m_gb_phi": mij_from_M(delta, m_gb),
        "g2_floor": (0.1 / p["dx"]) ** 2,
        "cooling_rate": p["G"] * p["V_pulling"] * p["dt"],
    }


def kernel_scalars(p: dict, d: dict, as_float32: bool = True) -> dict:
    """The 13 scalars of the update kernel.  The reference passes them as
    np.float32 (run_randommode_3d.py:280-285, tests/verification_3d.py:869-881);
    as_float32 reproduces that rounding."""
    raw = {
        "dx": p["dx"], "dt": p["dt"], "T_melt": p["T_melt"], "Sf": p["Sf"],
        "eps0_sl": d["eps0_sl"], "w0_sl": d["w0_sl"], "a0": d["a0"],
        "delta_a": p["delta_a"], "mu_a": p["mu_a"], "p_round": p["p_round"],
        "g2_floor": d["g2_floor"], "ksi": p["ksi"], "theta_c_rad": d["theta_c_rad"],
    }
    if as_float32:
        return {k: float(np.float32(v)) for k, v in raw.items()}
    return {k: float(v) for k, v in raw.items()}


def build_interaction_matrices(N: int, eps0_sl, w0_sl, m_sl_phi, eps_GB, w_GB, m_GB_phi):
    """wij, aij, mij [N, N] float32 (src/seed_modes.py:333-380): solid-solid
    pairs get the grain-boundary values, solid-liquid pairs the SL baseline,
    the diagonal stays 0."""
    wij = np.zeros((N, N), np.float32)
    aij = np.zeros((N, N), np.float32)
    mij = np.zeros((N, N), np.float32)
    if N > 1:
        solid = np.ones((N - 1, N - 1), bool) & ~np.eye(N - 1, dtype=bool)
        wij[1:, 1:][solid] = w_GB
        aij[1:, 1:][solid] = eps_GB
        mij[1:, 1:][solid] = m_GB_phi
        for t, v in ((wij, w0_sl), (aij, eps0_sl), (mij, m_sl_phi)):
            t[0, 1:] = v
            t[1:, 0] = v
    return wij, aij, mij


# -------------------------------------------------------------- orientations

# The 8 {111} normals in the order of src/orientation_utils.py:202-205 (facet
# ties resolve to the first in this order).
_N111_SIGNS = ((1, 1, 1), (1, 1, -1), (1, -1, 1), (-1, 1, 1),
               (1, -1, -1), (-1, 1, -1), (-1, -1, 1), (-1, -1, -1))


def compute_rotated_n111(grain_quaternions: np.ndarray) -> np.ndarray:
    """[N, 8, 3] float32 normals rotated by each grain's (x, y, z, w)
    quaternion (src/orientation_utils.py:182-212)."""
    q = np.asarray(grain_quaternions, np.float64)
    base = np.array(_N111_SIGNS, dtype=np.float32) / np.sqrt(3.0)
    out = np.zeros((q.shape[0], 8, 3), np.float32)
    for g in range(q.shape[0]):
        R = Rotation.from_quat(q[g]).as_matrix().astype(np.float32)
        out[g] = (R @ base.T).T
    return out


def assign_quaternions_to_grains(n_solid: int, mode: str = "random", orientation_seed: int = 42,
                                 csv_path: str | None = None) -> np.ndarray:
    """[n_solid + 1, 4] unit quaternions; row 0 is the liquid dummy
    (src/orientation_utils.py:123-179; modes 'random' and 'file')."""
    q = np.zeros((n_solid + 1, 4), np.float64)
    q[0] = (0.0, 0.0, 0.0, 1.0)
    if mode == "random":
        rng = np.random.default_rng(orientation_seed)
        for g in range(1, n_solid + 1):
            v = rng.standard_normal(4)
            q[g] = v / np.linalg.norm(v)
    elif mode == "file":
        if not csv_path:
            raise ValueError("orientation_mode='file' requires orientation_csv")
        data = np.atleast_2d(np.loadtxt(csv_path, delimiter=",", dtype=np.float64))
        if data.shape[0] < n_solid or data.shape[1] != 4:
            raise ValueError(f"{csv_path}: need {n_solid} rows of x,y,z,w")
        for g in range(1, n_solid + 1):
            v = data[g - 1]
            n = np.linalg.norm(v)
            q[g] = v / n if n > 1e-12 else (0.0, 0.0, 0.0, 1.0)
    else:
        raise ValueError(f"unknown orientation_mode {mode!r} (expected 'random' or 'file')")
    return q


def build_quaternion_from_config(cfg: dict) -> np.ndarray:
    """Euler 'xyz' degrees or a direct quaternion (src/orientation_utils.py:19-61)."""
    kind = cfg.get("orientation_type", "euler")
    if kind == "euler":
        return Rotation.from_euler("xyz", cfg.get("euler_deg", [0.0, 0.0, 0.0]), degrees=True).as_quat()
    if kind == "quaternion":
        v = np.asarray(cfg["quaternion"], np.float64)
        n = np.linalg.norm(v)
        return v / n if n >= 1e-12 else np.array([0.0, 0.0, 0.0, 1.0])
    raise ValueError(f"unknown orientation_type {kind!r}")


# ------------------------------------------------------ initial conditions


def generate_random_grain_map_3d(nx: int, ny: int, n_solid: int, random_seed: int = 42) -> np.ndarray:
    """Voronoi map [nx, ny] of grain ids 1..n_solid (src/seed_modes_3d.py:15-24):
    seeds from default_rng(seed).uniform for x then y, nearest seed wins, ties
    to the lowest id."""
    rng = np.random.default_rng(random_seed)
    sx = rng.uniform(0, nx, n_solid)
    sy = rng.uniform(0, ny, n_solid)
    lx = np.arange(nx, dtype=np.float32).reshape(nx, 1)
    ly = np.arange(ny, dtype=np.float32).reshape(1, ny)
    best = None
    gmap = np.zeros((nx, ny), np.int32)
    for g in range(n_solid):
        d2 = (lx - sx[g]) ** 2 + (ly - sy[g]) ** 2
        if best is None:
            best = d2
            gmap[:] = 1
        else:
            closer = d2 < best
            best = np.where(closer, d2, best)
            gmap[closer] = g + 1
    return gmap


def tanh_profile(nz: int, dz: float, delta: float, seed_height: float) -> np.ndarray:
    """0.5(1 - tanh(2.2/delta * (k - h) dz)) as float32 (src/seed_modes_3d.py:46-49)."""
    factor = np.float32(2.2 / delta)
    dist = (np.arange(nz, dtype=np.float64) - seed_height) * dz
    return (0.5 * (1.0 - np.tanh(factor * dist))).astype(np.float32)


def liquid_from_solid(solid: np.ndarray) -> np.ndarray:
    """clip(1 - phi_s, 0, 1) in float32 for a single-grain column."""
    return np.clip(np.float32(1.0) - solid.astype(np.float32), 0.0, 1.0).astype(np.float32)


def init_phi_from_grain_map_3d(grain_map, n_solid, nx, ny, nz, dz, delta, seed_height) -> np.ndarray:
    """Dense phi [N, nx, ny, nz] float32 (src/seed_modes_3d.py:38-57)."""
    N = n_solid + 1
    phi = np.zeros((N, nx, ny, nz), np.float32)
    prof = tanh_profile(nz, dz, delta, seed_height)
    for g in range(1, N):
        mask = grain_map == g
        phi[g][mask] = prof
    phi[0] = np.clip(1.0 - phi[1:].sum(axis=0), 0.0, 1.0)
    return phi


def columns_from_grain_map_3d(grain_map, nz, dz, delta, seed_height):
    """The same state as init_phi_from_grain_map_3d in columnar form for
    Solver.init_columns: (grain_map, profile_index, solid[1, nz], liquid[1, nz])."""
    prof = tanh_profile(nz, dz, delta, seed_height)
    return (np.asarray(grain_map, np.int32), np.zeros(grain_map.shape, np.int32),
            prof.astype(np.float64)[None, :], liquid_from_solid(prof).astype(np.float64)[None, :])


def init_temperature_field_3d(nx, ny, nz, T_melt, G, dz, seed_height) -> np.ndarray:
    """T(k) = T_melt + G (k - h) dz as float32, uniform in x, y (src/seed_modes_3d.py:27-35)."""
    t1 = (T_melt + G * (np.arange(nz, dtype=np.float64) - seed_height) * dz).astype(np.float32)
    return np.broadcast_to(t1[None, None, :], (nx, ny, nz)).copy()


def init_twomode_phi_3d(nx, ny, nz, dz, delta, seed_height, split_index,
                        grain1_seed_offset=0, grain2_seed_offset=0) -> np.ndarray:
    """Bicrystal: grain 1 for l < split, grain 2 beyond (src/seed_modes_3d.py:117-166)."""
    phi = np.zeros((3, nx, ny, nz), np.float32)
    phi[1, :split_index] = tanh_profile(nz, dz, delta, seed_height + grain1_seed_offset)
    phi[2, split_index:] = tanh_profile(nz, dz, delta, seed_height + grain2_seed_offset)
    phi[0] = 1.0 - phi[1] - phi[2]
    np.clip(phi[0], 0.0, 1.0, out=phi[0])
    return phi


def columns_twomode_3d(nx, ny, nz, dz, delta, seed_height, split_index,
                       grain1_seed_offset=0, grain2_seed_offset=0):
    """Columnar form of init_twomode_phi_3d."""
    p1 = tanh_profile(nz, dz, delta, seed_height + grain1_seed_offset)
    p2 = tanh_profile(nz, dz, delta, seed_height + grain2_seed_offset)
    gmap = np.full((nx, ny), 2, np.int32)
    gmap[:split_index] = 1
    pidx = np.where(gmap == 1, 0, 1).astype(np.int32)
    solid = np.stack([p1, p2]).astype(np.float64)
    liquid = np.stack([liquid_from_solid(p1), liquid_from_solid(p2)]).astype(np.float64)
    return gmap, pidx, solid, liquid


def init_singlemode_sphere_3d(nx, ny, nz, dx, dy, dz, delta, radius, center_x=None, center_y=None,
                              center_z=None, solid_gid=1) -> np.ndarray:
    """One spherical grain in liquid, [2, nx, ny, nz] float32 (src/seed_modes_3d.py:59-114)."""
    phi = np.zeros((2, nx, ny, nz), np.float32)
    factor = np.float32(2.2 / delta)
    cx = nx * dx / 2.0 if center_x is None else center_x
    cy = ny * dy / 2.0 if center_y is None else center_y
    cz = nz * dz / 2.0 if center_z is None else center_z
    X, Y, Z = np.meshgrid(np.arange(nx, dtype=np.float64) * dx, np.arange(ny, dtype=np.float64) * dy,
                          np.arange(nz, dtype=np.float64) * dz, indexing="ij")
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2)
    s = (0.5 * (1.0 - np.tanh(factor * (r - radius)))).astype(np.float32)
    phi[solid_gid] = s
    phi[0] = 1.0 - s
    return phi


def init_singlemode_disk_2d(nx, ny, dx, dy, delta, radius, center_x=None, center_y=None,
                            solid_gid=1) -> np.ndarray:
    """One disc-shaped grain in liquid, [2, nx, ny] float32 (src/seed_modes.py:135-182)."""
    phi = np.zeros((2, nx, ny), np.float32)
    factor = np.float32(2.2 / delta)
    cx = nx * dx / 2.0 if center_x is None else center_x
    cy = ny * dy / 2.0 if center_y is None else center_y
    X, Y = np.meshgrid(np.arange(nx, dtype=np.float64) * dx, np.arange(ny, dtype=np.float64) * dy,
                       indexing="ij")
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
    s = (0.5 * (1.0 - np.tanh(factor * (r - radius)))).astype(np.float32)
    phi[solid_gid] = s
    phi[0] = 1.0 - s
    return phi


def init_temperature_field(nx, ny, T_melt, G, dy, seed_height) -> np.ndarray:
    """2D linear temperature (src/seed_modes.py:185-210)."""
    t1 = (T_melt + G * (np.arange(ny, dtype=np.float64) - seed_height) * dy).astype(np.float32)
    return np.broadcast_to(t1[None, :], (nx, ny)).copy()
