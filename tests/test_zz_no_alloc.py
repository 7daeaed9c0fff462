"""mpf_dense_update_noalloc (the update on the reference's dense layout with
its tables in the library's file-scope device array) against
mpf_dense_update, both through the C ABI: two calls in a row, the second on a
smaller grid with fewer phases, each bit-identical to the existing entry
point in FP64 and FP32; and n_phase above the array's capacity is rejected
before any device work."""
import ctypes as C

import numpy as np
import pytest

import cases
from mpf import _lib
from mpf import dense as D

pytestmark = pytest.mark.gpu


def _inputs(c, dt):
    phi = c["phi"].astype(dt)
    N = phi.shape[0]
    grid = phi.shape[1:]
    d = {"N": N, "grid": grid, "phi": D.to_device(phi), "temp": D.to_device(np.asarray(c["temp"], np.float64)),
         "mf": D.to_device(np.zeros((N,) + grid, np.int32)), "nf": D.to_device(np.zeros(grid, np.int32)),
         "st": D.to_device(np.zeros(2, np.int32)),
         "tabs": [D.to_device(np.ascontiguousarray(t, dt)) for t in c["tables"]],
         "scalars": _lib.scalars_struct(c["scalars"]), "dtype": _lib.FP64 if dt == np.float64 else _lib.FP32}
    D.apt(d["phi"], d["mf"], d["nf"], d["st"], N, checked=True, kmax=50)
    return d


def _update(fn, d):
    out = D.to_device(np.full((d["N"],) + d["grid"], -1.0, d["phi"].dtype))
    nx, ny, nz = d["grid"]
    w, a, m, n111 = d["tabs"]
    rc = getattr(_lib.lib(), fn)(d["phi"].ptr, out.ptr, d["dtype"], d["temp"].ptr, _lib.FP64, d["mf"].ptr,
                                 d["nf"].ptr, w.ptr, a.ptr, m.ptr, n111.ptr, C.byref(d["scalars"]), d["N"], nx, ny,
                                 nz, 3, d["mf"].shape[0], 50, 1, 1, None)
    _lib.check(rc)
    return out


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_noalloc_matches_dense_update_on_two_calls(dt):
    big = cases.voronoi_case(nx=24, ny=20, nz=16, n_solid=9, seed=3, qseed=4, seed_height=6)
    small = cases.voronoi_case()   # 12 x 8 x 14, 4 grains
    di = [_inputs(c, dt) for c in (big, small)]
    refs = [_update("mpf_dense_update", d).copy_to_host() for d in di]
    outs = [_update("mpf_dense_update_noalloc", d) for d in di]   # back to back, no synchronisation between
    D.synchronize()
    got = [o.copy_to_host() for o in outs]
    for r, g, c in zip(refs, got, (big, small)):
        assert g.shape == c["phi"].shape
        assert np.array_equal(g, r)
        assert np.all(g >= 0.0)   # every phase written (the -1 fill is gone)


def test_noalloc_rejects_too_many_phases():
    d = _inputs(cases.voronoi_case(), np.float64)
    w, a, m, n111 = d["tabs"]
    rc = _lib.lib().mpf_dense_update_noalloc(d["phi"].ptr, d["phi"].ptr, _lib.FP64, d["temp"].ptr, _lib.FP64,
                                             d["mf"].ptr, d["nf"].ptr, w.ptr, a.ptr, m.ptr, n111.ptr,
                                             C.byref(d["scalars"]), 257, 12, 8, 14, 3, 50, 50, 1, 1, None)
    assert rc == _lib.MPF_EINVAL
    assert b"257" in _lib.lib().mpf_last_error()
