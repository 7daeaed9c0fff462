"""Launch shapes, HIP-graph replay and the autotuner (mpf_set_launch,
mpf_set_graphs, mpf_autotune in include/mpf.h).  None of these may change a
result: every variant is compared bit for bit (FP64) with the oracle and
with eager stepping at the default shape."""
import numpy as np
import pytest

import cases
from mpf import Batch, Solver
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _many_phase_case():
    """A blob of 8 grains: cells with more solids than the interface kernel's
    buckets take the many-phase kernel (side stream)."""
    rng = np.random.default_rng(5)
    N, nx, ny, nz = 9, 12, 10, 9
    phi = np.zeros((N, nx, ny, nz))
    phi[0] = 0.3
    phi[1:, 3:8, 2:7, 2:6] = rng.uniform(0.0, 1.0, (N - 1, 5, 5, 4))
    phi[1:] = np.where(phi[1:] < 0.35, 0.0, phi[1:])
    phi /= phi.sum(0)
    c = cases.voronoi_case()
    quats = np.vstack([[0, 0, 0, 1.0], rng.standard_normal((N - 1, 4))])
    quats[1:] /= np.linalg.norm(quats[1:], axis=1, keepdims=True)
    tabs = cases.tables_for(c["params"], quats)[1]
    return {"phi": phi, "temp": np.full((nx, ny, nz), 1682.0), "tables": tabs, "scalars": c["scalars"]}


def _solver(c, **kw):
    phi = c["phi"]
    s = Solver(phi.shape[1:], phi.shape[0], c["scalars"], precision="fp64", **kw)
    s.set_tables(*c["tables"])
    s.upload_phi(phi)
    s.upload_temp(c["temp"])
    return s


def _oracle(c, steps, cooling=0.0):
    return O.run(c["phi"], c["temp"], steps, c["tables"], c["scalars"], kmax=50, cooling_rate=cooling)


@pytest.mark.parametrize("shape", [(1, 1, 1, 0), (3, 5, 2, 1), (7, 1, 1, 2), (4096, 4096, 256, 1),
                                   (2, 8, 4, 2), (4096, 4096, 256, 2)])
def test_launch_shapes_bit_exact(shape):
    """Any block count (one block walking every list included) and every
    mode of the passes after the merge pass (in sequence, side streams, one
    launch) gives the same state."""
    c = _many_phase_case()
    steps = 5
    ref, _, _, _ = _oracle(c, steps)
    with _solver(c) as s:
        s.set_launch(*shape)
        got = s.launch()
        assert (got["merge_blocks"], got["front_blocks"], got["general_blocks"], got["side_stream"]) == shape
        assert s.step(steps) == steps
        assert s.stats()["general_cells"] > 0
        assert np.array_equal(s.phi(), ref)


@pytest.mark.parametrize("tiling", [(4, 1), (8, 3), (16, 0), (4, 2048), (16, 1)])
def test_tilings_bit_exact(tiling):
    """Every classification tiling (planes per work item, persistent grid
    down to one block) gives the oracle's state, on a grid of several tiles
    in x, y and z with a partial tile in each."""
    c = cases.voronoi_case(nx=70, ny=13, nz=20, n_solid=6, seed_height=8)
    steps = 6
    ref, _, _, _ = _oracle(c, steps)
    with _solver(c) as s:
        s.set_tiling(*tiling)
        got = s.launch()
        assert got["light_planes"] == tiling[0]
        assert got["light_blocks"] == (tiling[1] or 1024)
        assert s.step(steps) == steps
        assert np.array_equal(s.phi(), ref)


def test_tiling_rejects_bad_arguments():
    """Planes per work item other than 0, 4, 8 or 16 and negative block
    counts are refused before any launch; the tiling in use is unchanged."""
    from mpf._lib import MPFError
    c = cases.voronoi_case()
    with _solver(c) as s:
        before = s.launch()
        for bad in ((5, 0), (32, 0), (4, -1)):
            with pytest.raises(MPFError):
                s.set_tiling(*bad)
        assert s.launch() == before


@pytest.mark.parametrize("which", ["voronoi", "many_phase"])
def test_graph_replay_matches_eager(which):
    """Steps replayed from captured graphs, in several calls with downloads,
    observables and statistics between them, equal eager stepping and the
    oracle after every call."""
    c = cases.voronoi_case(n_solid=6) if which == "voronoi" else _many_phase_case()
    cooling = 0.25 if which == "voronoi" else 0.0
    with _solver(c, cooling_rate=cooling) as g, _solver(c, cooling_rate=cooling) as e:
        g.set_graphs(True)
        done = 0
        for n in (1, 3, 2, 4):
            assert g.step(n) == n
            assert e.step(n) == n
            done += n
            ref, T_ref, _, _ = _oracle(c, done, cooling)
            pg = g.phi()
            assert np.array_equal(pg, e.phi())
            assert np.array_equal(pg, ref), f"graph replay differs after {done} steps"
            np.testing.assert_array_equal(g.temp(), T_ref[0, 0])
            g.observables()
            assert g.stats()["max_active"] == e.stats()["max_active"]
        # a new launch shape drops the captured graphs and captures again
        g.set_launch(64, 64, 8, False)
        e.step(3)
        g.step(3)
        assert np.array_equal(g.phi(), e.phi())


def test_graphs_refused_with_halt():
    c = cases.voronoi_case()
    with _solver(c, halt_on_kmax=True) as s:
        with pytest.raises(Exception):
            s.set_graphs(True)


def test_autotune_restores_state():
    c = cases.voronoi_case(n_solid=6)
    with _solver(c) as s:
        s.step(2)
        phi0 = s.phi()
        st0 = s.stats()
        r = s.autotune(trial_steps=2)
        assert r["merge_blocks"] > 0 and r["front_blocks"] > 0 and r["general_blocks"] > 0
        assert r["us_per_step"] > 0
        assert s.launch() == {k: v for k, v in r.items() if k != "us_per_step"}
        assert r["light_planes"] in (4, 8, 16) and r["light_blocks"] in (512, 1024, 2048)
        assert np.array_equal(s.phi(), phi0)
        st1 = s.stats()
        for k in ("steps_done", "max_active", "merge_cells", "front_cells", "general_cells"):
            assert st1[k] == st0[k], k
        s.set_graphs(True)
        s.step(4)
        ref, _, _, _ = _oracle(c, 6)
        assert np.array_equal(s.phi(), ref)


def test_batch_graph_replay():
    """A batch of grids replayed from graphs equals the same batch stepped
    eagerly."""
    shapes = [(16, 12, 10), (10, 14, 12), (20, 8, 9)]
    cs = [cases.voronoi_case(nx=nx, ny=ny, nz=nz, n_solid=4 + i, seed=3 + i, qseed=9 + i)
          for i, (nx, ny, nz) in enumerate(shapes)]
    # one set of tables for the batch: same number of phases everywhere
    N = 7
    c0 = cases.voronoi_case(n_solid=N - 1)

    def pad(phi):
        out = np.zeros((N,) + phi.shape[1:])
        out[:phi.shape[0]] = phi
        return out

    def make():
        b = Batch(shapes, N, c0["scalars"], precision="fp64")
        for i, c in enumerate(cs):
            b.set_tables(i, *c0["tables"])
            b.upload_phi(i, pad(c["phi"]))
            b.upload_temp(i, c["temp"][0, 0])
        return b

    with make() as g, make() as e:
        g.set_graphs(True)
        for n in (2, 3):
            g.step(n)
            e.step(n)
            for i in range(len(shapes)):
                assert np.array_equal(g.phi(i), e.phi(i))
        for i, c in enumerate(cs):
            T = np.broadcast_to(c["temp"][0, 0], shapes[i]).copy()
            ref, _, _, _ = O.run(pad(c["phi"]), T, 5, c0["tables"], c0["scalars"], kmax=50)
            assert np.array_equal(g.phi(i), ref)

