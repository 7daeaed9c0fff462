"""The driver counterparts on the GPU: a small random-mode and two-grain run
(FP64 through the configuration's gpu.precision) equal the oracle run of the
reference's dense inputs bit for bit, and the device phase map equals
np.argmax over the oracle's phases on every slice axis."""
import os

import numpy as np
import pytest

from mpf import inputs
from mpf import run_randommode_3d as R
from mpf import run_twomode_3d as T
from mpf.config import load_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _small(c, nx=20, ny=18, nz=26, nsteps=6, save_every=3):
    c["grid"].update(nx=nx, ny=ny, nz=nz, nsteps=nsteps)
    c["output"]["save_every"] = save_every
    c["gpu"]["precision"] = "fp64"
    return c


def test_randommode_run_matches_oracle(tmp_path):
    c = _small(load_config())
    c["randommode"].update(n_solid=4, seed_height=7)
    res = R.run(c, str(tmp_path / "rm"), plots=True, keep_phi=True, log=lambda *_: None)
    S = R.setup(c)
    p = S["p"]
    phi0 = inputs.init_phi_from_grain_map_3d(S["grain_map"], 4, p["nx"], p["ny"], p["nz"], p["dz"],
                                             S["d"]["delta"], S["seed_height"])
    T0 = inputs.init_temperature_field_3d(p["nx"], p["ny"], p["nz"], p["T_melt"], p["G"], p["dz"], S["seed_height"])
    ref, _, _, _ = O.run(phi0, T0, 6, S["tables"], S["scalars"], kmax=p["KMAX"], cooling_rate=S["cooling"])
    assert res["steps"] == 6
    assert np.array_equal(res["phi"], ref)
    names = sorted(os.path.basename(f) for f in res["files"])
    for k in (0, 3, 6):
        assert f"step_{k}_xy.png" in names and f"step_{k}_xz.png" in names and f"step_{k}_interface.png" in names
    assert "config_snapshot.yaml" in names and "run_params.yaml" in names


def test_twomode_run_and_phase_map(tmp_path):
    c = _small(load_config(), nx=24, ny=12, nz=22, nsteps=4, save_every=4)
    c["twomode"].update(seed_height=6, split_ratio=0.5)
    res = T.run(c, str(tmp_path / "tm"), plots=False, keep_phi=True, log=lambda *_: None)
    S = T.setup(c)
    p = S["p"]
    phi0 = inputs.init_twomode_phi_3d(p["nx"], p["ny"], p["nz"], p["dz"], S["d"]["delta"], S["seed_height"],
                                      S["split_index"])
    T0 = inputs.init_temperature_field_3d(p["nx"], p["ny"], p["nz"], p["T_melt"], p["G"], p["dz"], S["seed_height"])
    ref, _, _, _ = O.run(phi0, T0, 4, S["tables"], S["scalars"], kmax=p["KMAX"], cooling_rate=S["cooling"])
    assert np.array_equal(res["phi"], ref)


def test_phase_map_equals_argmax():
    import cases
    from mpf import Solver
    c = cases.voronoi_case(nx=14, ny=10, nz=16, n_solid=6)
    s = Solver(c["phi"].shape[1:], c["phi"].shape[0], c["scalars"], precision="fp64")
    with s:
        s.set_tables(*c["tables"])
        s.upload_phi(c["phi"])
        s.upload_temp(c["temp"])
        s.step(5)
        phi = s.phi()
        pm = np.argmax(phi, axis=0)
        for axis, sl in (("xy", lambda i: pm[:, :, i]), ("xz", lambda i: pm[:, i, :]), ("yz", lambda i: pm[i])):
            for idx in (0, 3, None):
                from mpf.plot import resolve_slice_index
                i = resolve_slice_index(axis, phi.shape[1:], idx)
                assert np.array_equal(s.phase_map(axis, idx), sl(i)), (axis, idx)
        with pytest.raises(Exception):
            from mpf import _lib
            out = np.zeros((14, 10), np.int32)
            _lib.check(_lib.lib().mpf_phase_map(s._h, 0, 16, _lib.ptr(out, _lib.C.c_int32)))
