"""Batches of independent grids (mpf.Batch, mpf_batch_* in include/mpf.h):
every grid of a batch, stepped together with the others, must equal the
oracle run of that grid alone.  FP64, tolerance: bit-identical phi, mf/nf."""
import numpy as np
import pytest

import cases
from mpf import Batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bicrystal(nx, ny, nz, seed, undercool):
    """Two grains split along x under a uniform undercooling (the bicrystal
    sweep's cases: init_twomode_phi_3d, src/seed_modes_3d.py:117-166)."""
    p = cases.ref_params()
    d = cases.inputs.derive_constants(p)
    rng = np.random.default_rng(seed)
    q = np.vstack([[0, 0, 0, 1.0], rng.standard_normal((2, 4))])
    q[1:] /= np.linalg.norm(q[1:], axis=1, keepdims=True)
    d, tabs = cases.tables_for(p, q)
    phi = cases.inputs.init_twomode_phi_3d(nx, ny, nz, p["dz"], d["delta"], 4, nx // 2, 0, 0)
    temp = np.full(nz, p["T_melt"] - undercool)
    return {"phi": phi.astype(np.float64), "temp": temp, "tables": tabs,
            "scalars": cases.inputs.kernel_scalars(p, d, as_float32=True)}


def _cases():
    out = []
    for (nx, ny, nz), seed, du in (((16, 12, 10), 3, 2.0), ((10, 14, 12), 5, 4.0), ((20, 8, 9), 7, 3.0)):
        out.append(_bicrystal(nx, ny, nz, seed, du))
    return out


@pytest.mark.parametrize("cooling", [0.0, 0.25])
def test_batch_matches_each_grid_alone(cooling):
    cs = _cases()
    steps = 6
    shapes = [c["phi"].shape[1:] for c in cs]
    with Batch(shapes, 3, cs[0]["scalars"], precision="fp64", cooling_rate=cooling) as b:
        for i, c in enumerate(cs):
            b.set_tables(i, *c["tables"])
            b.upload_phi(i, c["phi"])
            b.upload_temp(i, c["temp"])
        assert b.step(steps) == steps
        for i, c in enumerate(cs):
            T = np.broadcast_to(c["temp"], c["phi"].shape[1:]).copy()
            ref, T_ref, _, st = O.run(c["phi"], T, steps, c["tables"], c["scalars"], kmax=50, cooling_rate=cooling)
            got = b.phi(i)
            print(i, shapes[i], "max |diff| =", np.abs(got - ref).max())
            assert np.array_equal(got, ref)
            mf, nf, s = b.apt(i)
            mf_o, nf_o, _ = O.apt(ref, 50)
            assert np.array_equal(nf, nf_o)
            np.testing.assert_array_equal(b.temp(i), T_ref[0, 0])
            ob = b.observables(i)
            assert ob["cells"] == int(np.prod(shapes[i]))
            np.testing.assert_allclose(ob["phase_sums"], ref.reshape(3, -1).sum(1), rtol=1e-12)


def test_batch_fp32_and_errors():
    cs = _cases()
    shapes = [c["phi"].shape[1:] for c in cs]
    with Batch(shapes, 3, cs[0]["scalars"], precision="fp32") as b:
        with pytest.raises(Exception):
            b.step(1)                     # tables not set
        for i, c in enumerate(cs):
            b.set_tables(i, *c["tables"])
            b.upload_phi(i, c["phi"])
            b.upload_temp(i, c["temp"])
        b.step(4)
        for i, c in enumerate(cs):
            T = np.broadcast_to(c["temp"], c["phi"].shape[1:]).copy()
            ref, _, _, _ = O.run(c["phi"], T, 4, c["tables"], c["scalars"], kmax=50)
            assert np.abs(b.phi(i) - ref).max() <= 1e-4
        assert b.stats()["class_tables"]
        with pytest.raises(Exception):
            b.upload_phi(len(cs), cs[0]["phi"])
