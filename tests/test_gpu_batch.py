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


@pytest.fixture(scope="module")
def c4_sized():
    """One grid of each size class of the batched bicrystal sweep (nx = ny in
    {64, 96, 128}, nz in {96, 128, 192}), two grains split along x, random
    quaternion pair, nondimensional undercooling from the sweep's range, and
    the oracle's FP64 run of each grid alone over 100 steps."""
    from mpf import sweep
    p, d, scal, _, _ = sweep.physics()
    cs = sweep.bicrystal_cases(3, seed=41)
    for c, (nxy, nz) in zip(cs, ((64, 96), (96, 128), (128, 192))):
        c.shape, c.split_index = (nxy, nxy, nz), nxy // 2
    steps = 100
    runs = []
    for c in cs:
        phi0 = sweep.case_initial_phi(c, p, d).astype(np.float64)
        T = sweep.case_temperature(c, p)
        tabs = sweep.case_tables(c, d)
        ref, _, _, st = O.run(phi0, np.broadcast_to(T, c.shape).copy(), steps, tabs, scal, kmax=50)
        runs.append({"case": c, "phi0": phi0, "temp": T, "tables": tabs, "ref": ref})
    return runs, scal, steps


def _c4_batch(runs, scal, steps, precision, graphs):
    shapes = [r["case"].shape for r in runs]
    out = []
    with Batch(shapes, 3, scal, precision=precision) as b:
        for i, r in enumerate(runs):
            b.set_tables(i, *r["tables"])
            b.upload_phi(i, r["phi0"])
            b.upload_temp(i, r["temp"])
        if graphs:
            b.set_graphs(True)
        assert b.step(steps) == steps
        for i in range(len(runs)):
            out.append((b.phi(i), b.apt(i)[1], b.observables(i, dx=1.0)))
    return out


def test_batch_c4_sizes_fp64_bit_exact(c4_sized):
    """The three size classes in one batch, 100 FP64 steps: phi and nf
    identical to the oracle's run of each grid alone."""
    runs, scal, steps = c4_sized
    for r, (phi, nf, ob) in zip(runs, _c4_batch(runs, scal, steps, "fp64", graphs=False)):
        print(r["case"].shape, "max |phi - oracle| =", float(np.abs(phi - r["ref"]).max()))
        assert np.array_equal(phi, r["ref"])
        assert np.array_equal(nf, O.apt(r["ref"], 50)[1])


def test_batch_c4_sizes_fp32_graphs(c4_sized):
    """The same batch in FP32 with the steps captured in HIP graphs, against
    the FP64 oracle: max |phi32 - phi64| <= 1e-4, grain volumes within 1e-5
    relative, front heights (topmost cell with 1 - phi_L > 0.5) identical."""
    runs, scal, steps = c4_sized
    for r, (phi, nf, ob) in zip(runs, _c4_batch(runs, scal, steps, "fp32", graphs=True)):
        ref = r["ref"]
        err = float(np.abs(phi - ref).max())
        vol = ref.reshape(3, -1).sum(1)
        vrel = np.abs(np.asarray(ob["phase_sums"]) - vol) / vol
        solid = 1.0 - ref[0]
        k = np.arange(solid.shape[2])
        top = np.max(np.where(solid > 0.5, k, -1), axis=2)
        front_ref = np.where(top >= 0, top, np.nan)
        front = ob["front"]
        differ = int((np.nan_to_num(front, nan=-1) != np.nan_to_num(front_ref, nan=-1)).sum())
        print(r["case"].shape, f"max |phi32 - phi64| = {err:.3e}, grain volume rel = {vrel.max():.3e}, "
              f"front columns differing = {differ}")
        assert err <= 1e-4
        assert vrel.max() <= 1e-5
        assert differ == 0
