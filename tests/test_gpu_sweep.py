"""Device sweep measurements and the batched sweep driver against the oracle.

- mpf_probe_rays computes the orientation scan's ray probes in double with
  contraction off: identical to the oracle's restatement of the reference on
  the same phi (FP64 state, bit-identical to the oracle's).
- mpf_batch_observables_all equals the per-grid observables (phase sums up to
  the order of their atomic accumulation).
- mpf.sweep.run_sweep on small bicrystal and orientation cases (FP64):
  - phase sums within 1e-12 relative of the oracle's sums at every sample;
  - orientation rows (volume, equivalent radius, directional radii)
    identical to measure_observables on the oracle's phi, up to the summation
    order of the volume.
"""
import math

import numpy as np
import pytest

from mpf import Batch, Solver, inputs, observe, sweep
from oracle import observe as OO
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _sphere_solver(n=24, R=6.0, steps=4, dT=0.2, q=None):
    p, d, scal, _, _ = sweep.physics()
    c = sweep.orientation_cases([{"label": "t", "orientation_type": "euler",
                                  "euler_deg": q or [10.0, 20.0, 30.0]}], shape=(n, n, n), radius_cells=R,
                                undercooling=dT)[0]
    phi0 = sweep.case_initial_phi(c, p, d)
    tabs = sweep.case_tables(c, d)
    T = sweep.case_temperature(c, p)
    s = Solver((n, n, n), 2, scal, precision="fp64", temp_f32=True)
    s.set_tables(*tabs)
    s.upload_phi(phi0)
    s.upload_temp(T)
    s.step(steps)
    ref, _, _, _ = O.run(phi0, np.broadcast_to(T, (n, n, n)), steps, tabs, scal, kmax=50, temp_f32=True)
    return s, c, ref


def test_probe_rays_match_oracle():
    s, c, ref = _sphere_solver()
    with s:
        got_phi = s.phi()
        assert np.array_equal(got_phi, ref)
        field = ref[1]
        rng = np.random.default_rng(4)
        origins = np.vstack([np.array(c.center), np.array(c.center) + rng.uniform(-3, 3, (20, 3)),
                             [[-1.0, 5.0, 5.0], [2.0, 2.0, 2.0]]])
        dirs = [OO.unit(v) for v in rng.standard_normal((origins.shape[0], 3))]
        dirs[0] = OO.unit((1, 1, 1))
        got = observe.probe_rays(s._h, np.zeros(len(dirs), np.int32), np.ones(len(dirs), np.int32), origins,
                                 np.array(dirs), 0.5, 0.5)
        exp = np.array([OO.radius_one_side(field, o, u, 0.5, 0.5) for o, u in zip(origins, dirs)])
        print("finite rays:", int(np.isfinite(exp).sum()), "of", exp.size)
        assert np.isfinite(exp).sum() >= 15
        assert np.array_equal(np.isnan(got), np.isnan(exp))
        assert np.array_equal(got[np.isfinite(exp)], exp[np.isfinite(exp)])
        # liquid phase, coarser step: same rule
        got0 = observe.probe_rays(s._h, [0], [0], [[0.5, 0.5, 0.5]], [OO.unit((1, 1, 1))], 0.5, 0.7)
        assert got0[0] == OO.radius_one_side(ref[0], np.array([0.5, 0.5, 0.5]), OO.unit((1, 1, 1)), 0.5, 0.7)
        radii = s.directional_radii(c.center)
        for name, vec in OO.DIRECTIONS.items():
            assert radii[name] == OO.radius_symmetric(field, np.array(c.center), vec, 0.5, 0.5), name
        with pytest.raises(Exception):
            observe.probe_rays(s._h, [1], [1], [c.center], [OO.unit((1, 0, 0))])   # no grid 1
        with pytest.raises(Exception):
            observe.probe_rays(s._h, [0], [2], [c.center], [OO.unit((1, 0, 0))])   # no phase 2


def test_observables_all_matches_per_grid():
    p, d, scal, _, _ = sweep.physics()
    cs = sweep.bicrystal_cases(3, seed=9, sizes_xy=(16, 20), sizes_z=(18, 24))
    with Batch([c.shape for c in cs], 3, scal, precision="fp64") as b:
        for q, c in enumerate(cs):
            b.set_tables(q, *sweep.case_tables(c, d))
            b.upload_phi(q, sweep.case_initial_phi(c, p, d))
            b.upload_temp(q, sweep.case_temperature(c, p))
        b.step(3)
        gv, out, fronts = b.observables_all()
        for q in range(len(cs)):
            ob = b.observables(q)
            # phase sums accumulate with atomics: equal up to summation order
            np.testing.assert_allclose(gv[q], ob["phase_sums"], rtol=1e-13)
            assert out[q, 0] == ob["phi_sum_error_max"] and out[q, 4] == ob["solid_sum"]
            assert int(out[q, 5]) == ob["cells"] and int(out[q, 6]) == ob["stored_slots"]
            f = fronts[q].astype(np.float64)
            f[fronts[q] < 0] = np.nan
            np.testing.assert_array_equal(f, ob["front"])


def test_bicrystal_sweep_matches_oracle():
    cs = sweep.bicrystal_cases(5, seed=2, sizes_xy=(16, 20), sizes_z=(20, 26), seed_height=8.0)
    steps, every = 8, 4
    res = sweep.run_sweep(cs, steps, precision="fp64", max_batch_cells=12_000, sample_every=every)
    assert len(res["timing"]["batches"]) >= 2
    p, d, scal, _, _ = sweep.physics()
    for c, r in zip(cs, res["cases"]):
        assert r["index"] == c.index and r["shape"] == list(c.shape)
        phi0 = sweep.case_initial_phi(c, p, d)
        T = np.broadcast_to(sweep.case_temperature(c, p), c.shape)
        tabs = sweep.case_tables(c, d)
        for row in r["timeseries"]:
            k = row["step"]
            ref = phi0.astype(np.float64) if k == 0 else \
                O.run(phi0, T, k, tabs, scal, kmax=50)[0]
            np.testing.assert_allclose(row["phase_sums"], ref.reshape(3, -1).sum(1), rtol=1e-12, atol=1e-12)
        vols = OO.group_volumes(ref, 1.0)
        assert r["winner_gid"] == (1 if vols[0] >= vols[1] else 2)
        assert r["volume_a"] == pytest.approx(vols[0], rel=1e-12)
    t = res["timing"]
    assert t["instances"] == 5 and t["cell_updates"] == sum(c.cells for c in cs) * steps
    assert t["latency_s"]["p50"] <= t["latency_s"]["p99"] <= t["wall_s"] + 1e-9


def test_orientation_sweep_matches_oracle():
    cs = sweep.orientation_cases(shape=(22, 22, 22), radius_cells=5.0, undercooling=0.25)
    steps, every = 10, 5
    res = sweep.run_sweep(cs, steps, precision="fp64", sample_every=every, graphs=True)
    p, d, scal, _, _ = sweep.physics()
    for c, r in zip(cs, res["cases"]):
        phi0 = sweep.case_initial_phi(c, p, d)
        T = np.broadcast_to(sweep.case_temperature(c, p), c.shape)
        tabs = sweep.case_tables(c, d)
        exp_rows = []
        for row in r["timeseries"]:
            k = row["step"]
            ref = phi0.astype(np.float64) if k == 0 else \
                O.run(phi0, T, k, tabs, scal, kmax=50, temp_f32=True)[0]
            e = OO.measure_observables(ref, k, p["dt"], p["dx"], np.array(c.center))
            exp_rows.append(e)
            assert row["volume"] == pytest.approx(e["volume"], rel=1e-12)
            assert row["equivalent_radius"] == pytest.approx(e["equivalent_radius"], rel=1e-12)
            for name in ("x", "y", "z", "diag111"):
                a, b = row[f"radius_{name}"], e[f"radius_{name}"]
                assert (math.isnan(a) and math.isnan(b)) or a == b, (c.label, k, name, a, b)
        rate, _, _, npts = observe.fit_growth_rate(exp_rows, 0.5)
        assert r["fitted_growth_rate"] == pytest.approx(rate, rel=1e-9)
        assert r["fit_points"] == npts
    assert all(np.isfinite(r["fitted_growth_rate"]) for r in res["cases"])
