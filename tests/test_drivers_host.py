"""Host logic of the driver counterparts (mpf.run_randommode_3d,
mpf.run_twomode_3d) and of the artefact writers (mpf.plot): command-line
overrides, configuration errors, slice indices, YAML records and PNGs."""
import os

import numpy as np
import pytest
import yaml

from mpf import plot
from mpf import run_randommode_3d as R
from mpf import run_twomode_3d as T
from mpf.config import load_config


def test_cli_overrides_match_reference_keys():
    a = R.parse_args(["--nx", "40", "--ny", "36", "--nz", "50", "--nsteps", "7", "--save-every", "3",
                      "--seed-height", "9", "--n-solid", "5"])
    c = R.runtime_config(load_config(a.config), a)
    assert (c["grid"]["nx"], c["grid"]["ny"], c["grid"]["nz"], c["grid"]["nsteps"]) == (40, 36, 50, 7)
    assert c["output"]["save_every"] == 3
    assert c["randommode"]["seed_height"] == 9 and c["randommode"]["n_solid"] == 5
    base = load_config(a.config)
    assert base["grid"]["nx"] != 40 or base["grid"]["nsteps"] != 7     # the loaded dict is untouched


def test_setup_inputs_and_errors():
    c = load_config()
    c["grid"].update(nx=20, ny=18, nz=30)
    c["randommode"].update(n_solid=4, seed_height=6)
    S = R.setup(c)
    assert S["N"] == 5 and S["grain_map"].shape == (20, 18) and S["temp"].shape == (30,)
    assert S["cooling"] == float(np.float32(c["physical"]["G"] * c["physical"]["V_pulling"] * c["grid"]["dt"]))
    assert set(np.unique(S["grain_map"])) <= set(range(1, 5))
    bad = load_config()
    bad["randommode"]["n_solid"] = bad["gpu"]["MAX_GRAINS"]
    with pytest.raises(RuntimeError):
        R.setup(bad)
    bad = load_config()
    bad["grid"]["dz"] = 2 * bad["grid"]["dx"]
    with pytest.raises(ValueError):
        R.setup(bad)
    bad = load_config()
    bad["gpu"]["threads_per_block"] = [8, 8]
    with pytest.raises(ValueError):
        R.setup(bad)
    t = T.setup(load_config())
    assert t["split_index"] == int(load_config()["grid"]["nx"] * 0.5)
    assert np.allclose(np.linalg.norm(t["quats"], axis=1), 1.0)


def test_slice_index_and_interface_map():
    assert plot.resolve_slice_index("xy", (10, 12, 14), None) == 7
    assert plot.resolve_slice_index("xz", (10, 12, 14), 99) == 11
    assert plot.resolve_slice_index("yz", (10, 12, 14), -3) == 0
    with pytest.raises(ValueError):
        plot.resolve_slice_index("zz", (10, 12, 14), None)
    assert np.array_equal(plot.interface_z(np.array([[-1, 3], [5, -1]])), [[0, 3], [5, 0]])


def test_artefacts_written(tmp_path):
    out = str(tmp_path / "run")
    files = plot.save_run_config(out, {"grid": {"nx": 4}}, {"q": np.eye(2), "n": np.int64(3)})
    rp = yaml.safe_load(open(files[1]))
    assert rp["q"] == [[1.0, 0.0], [0.0, 1.0]] and rp["n"] == 3 and "saved_at" in rp
    assert yaml.safe_load(open(files[0])) == {"grid": {"nx": 4}}
    p1 = plot.save_phase_map_slice(np.arange(12).reshape(3, 4) % 3, out, "step_0_xy.png", 3, "xy", 1)
    p2 = plot.save_interface_position(np.array([[1, 2], [-1, 4]]), out, "step_0_interface.png")
    assert p1.endswith(os.path.join("3dphase", "3dphase", "step_0_xy.png")) and os.path.getsize(p1) > 0
    assert p2.endswith(os.path.join("interface", "interface", "step_0_interface.png")) and os.path.getsize(p2) > 0
