"""The surrogate's HIP kernels against the FP64 oracle (oracle/nn_oracle.py).

Single layers take bf16-rounded inputs and weights on both sides, so the
only differences are fp32 against fp64 accumulation and the final bf16
rounding: outputs within 1 bf16 ulp (relative 2^-7), weight gradients (fp32)
within 1e-3 relative to their norm.  The whole model is compared with the oracle
model that stores activations and their gradients in bf16 where the device
does (FP64 arithmetic) and takes the device's ReLU decisions (outputs within
rounding of 0 would otherwise flip and dominate the difference):
predictions within 1e-2 relative L2 and every layer's weight and bias
gradients within 1.5e-2 (cosine >= 0.9999); against the pure FP64 oracle
the predictions stay within 5e-2."""
import numpy as np
import pytest

from oracle import nn_oracle as NO

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _dev(a, dtype):
    torch = _torch()
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def _host(t):
    return t.detach().float().cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("dims,cin,cout,ntap,periodic", [
    ((1, 8, 6, 10), 32, 32, 27, (1, 1, 0)),
    ((2, 5, 4, 7), 64, 96, 27, (0, 0, 0)),
    ((1, 4, 4, 4), 96, 32, 27, (1, 1, 1)),
    ((1, 3, 9, 21), 32, 64, 27, (1, 1, 0)),
    ((2, 5, 3, 16), 128, 64, 27, (0, 1, 0)),
    ((3, 6, 5, 9), 64, 32, 1, (1, 1, 0)),
    # sizes at which the forward kernel takes its larger wave tiles (4x4,
    # 2x4 with 2x2 for the backward data, 4x2 with 2x4 for the backward data)
    ((1, 64, 64, 16), 64, 64, 27, (1, 1, 0)),
    ((1, 32, 32, 32), 32, 64, 27, (1, 1, 0)),
    ((1, 128, 128, 32), 64, 32, 27, (1, 1, 0)),
])
def test_conv_forward_and_backward(dims, cin, cout, ntap, periodic):
    torch = _torch()
    from mpf import _nnlib
    from mpf.surrogate import Conv, _stream
    rng = np.random.default_rng(7)
    x = NO.bf16(rng.standard_normal(dims + (cin,)))
    layer = Conv(cin, cout, ntap, True, rng)
    layer.b.copy_(torch.tensor(rng.standard_normal(cout) * 0.1, dtype=torch.float32, device="cuda"))
    layer.refresh()
    w = NO.bf16(_host(layer.w))
    b = _host(layer.b)
    xd = _dev(x.reshape(-1, cin), torch.bfloat16)
    y = layer.forward(xd, dims, periodic)
    ref = NO.bf16(NO.conv(x, w, b, periodic, relu=True)).reshape(-1, cout)
    got = _host(y)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= scale * 2 ** -7, np.abs(got - ref).max()
    # backward through the ReLU: weight, bias and input gradients
    dy = NO.bf16(rng.standard_normal(ref.shape))
    dyd = _dev(dy, torch.bfloat16)
    dx = layer.backward(dyd)
    torch.cuda.synchronize()
    mask = (got > 0)
    dym = dy * mask
    dx_ref, dw_ref = NO.conv_bwd(x, w, dym.reshape(dims + (cout,)), periodic)
    gw = _host(layer.gw)
    assert np.linalg.norm(gw - dw_ref) <= 1e-3 * np.linalg.norm(dw_ref)
    np.testing.assert_allclose(_host(layer.gb), dym.sum(axis=0), rtol=1e-3, atol=1e-3)
    dxh = _host(dx)
    dxr = dx_ref.reshape(-1, cin)
    assert np.abs(dxh - dxr).max() <= np.abs(dxr).max() * 2 ** -7 + 1e-6


@pytest.mark.parametrize("C", [32, 12])
def test_pool_upsample_copy_kernels(C):
    """C = 32 takes the 8-channel kernels, C = 12 the one-channel kernels."""
    torch = _torch()
    from mpf import _nnlib
    from mpf.surrogate import _stream, _p
    rng = np.random.default_rng(8)
    dims = (2, 4, 6, 8)
    x = NO.bf16(rng.standard_normal(dims + (C,)))
    xd = _dev(x.reshape(-1, C), torch.bfloat16)
    n_c = 2 * 2 * 3 * 4
    p = torch.empty((n_c, C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_pool", xd.data_ptr(), p.data_ptr(), _nnlib.ivec(*dims), C, 0, _stream())
    np.testing.assert_allclose(_host(p), NO.bf16(NO.pool(x)).reshape(-1, C), rtol=2 ** -8, atol=1e-6)
    g = NO.bf16(rng.standard_normal((2, 2, 3, 4, C)))
    gd = _dev(g.reshape(-1, C), torch.bfloat16)
    pb = torch.empty((x.size // C, C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_pool", gd.data_ptr(), pb.data_ptr(), _nnlib.ivec(*dims), C, 1, _stream())
    np.testing.assert_allclose(_host(pb), NO.bf16(NO.pool_bwd(g)).reshape(-1, C), rtol=0, atol=0)
    u = torch.empty((x.size // C, C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_upsample", gd.data_ptr(), u.data_ptr(), _nnlib.ivec(2, 2, 3, 4), C, 0, _stream())
    np.testing.assert_array_equal(_host(u), NO.up(g).reshape(-1, C))
    ub = torch.empty((n_c, C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_upsample", xd.data_ptr(), ub.data_ptr(), _nnlib.ivec(2, 2, 3, 4), C, 1, _stream())
    np.testing.assert_allclose(_host(ub), NO.bf16(NO.up_bwd(x)).reshape(-1, C), rtol=2 ** -7, atol=1e-6)
    # channel copies: place x into channels [16, 48) of a 64-channel tensor, then accumulate
    nv = x.size // C
    big = torch.zeros((nv, 64), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_copy_channels", xd.data_ptr(), C, 0, big.data_ptr(), 64, 16, C, nv, 0, _stream())
    _nnlib.call("mpf_nn_copy_channels", xd.data_ptr(), C, 0, big.data_ptr(), 64, 16, C, nv, 1, _stream())
    hb = _host(big)
    np.testing.assert_array_equal(hb[:, 16:16 + C], NO.bf16(2 * x.reshape(-1, C)))
    assert not hb[:, :16].any() and not hb[:, 16 + C:].any()
    # an offset that is not a multiple of 8 channels takes the one-channel kernel
    odd = torch.zeros((nv, 40), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_copy_channels", xd.data_ptr(), C, 0, odd.data_ptr(), 40, 4, C, nv, 0, _stream())
    ho = _host(odd)
    np.testing.assert_array_equal(ho[:, 4:4 + C], x.reshape(-1, C))
    assert not ho[:, :4].any() and not ho[:, 4 + C:].any()


@pytest.mark.parametrize("C", [32, 12])
def test_slot_pool_and_head_input_kernels(C):
    """Sum pooling over valid slots, the head's input and their adjoints
    against numpy, for channel counts that take the 8-channel kernels (32)
    and the one-channel kernels (12)."""
    torch = _torch()
    from mpf import _nnlib
    from mpf.surrogate import _stream
    rng = np.random.default_rng(11)
    K, nv = 3, 200
    sid = rng.integers(-1, 5, size=(K, nv)).astype(np.int32)
    e = NO.bf16(rng.standard_normal((K, nv, C)))
    ed, sd = _dev(e, torch.bfloat16), _dev(sid, torch.int32)
    h = torch.empty((nv, C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_slot_pool", ed.data_ptr(), sd.data_ptr(), h.data_ptr(), nv, K, C, 0, 0, _stream())
    ref = NO.bf16(((sid >= 0)[..., None] * e).sum(axis=0))
    np.testing.assert_allclose(_host(h), ref, rtol=2 ** -7, atol=1e-6)
    dh = NO.bf16(rng.standard_normal((nv, C)))
    de = _dev(np.ones((K, nv, C)), torch.bfloat16)
    _nnlib.call("mpf_nn_slot_pool", _dev(dh, torch.bfloat16).data_ptr(), sd.data_ptr(), de.data_ptr(), nv, K, C, 1,
                1, _stream())
    np.testing.assert_allclose(_host(de), NO.bf16(1.0 + (sid >= 0)[..., None] * dh[None]), rtol=2 ** -7, atol=1e-6)
    F = NO.bf16(rng.standard_normal((nv, C)))
    Fd = _dev(F, torch.bfloat16)
    hin = torch.empty((K, nv, 2 * C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_head_in", Fd.data_ptr(), C, ed.data_ptr(), C, hin.data_ptr(), nv, K, 0, _stream())
    np.testing.assert_array_equal(_host(hin), np.concatenate([np.broadcast_to(F, (K, nv, C)), e], axis=2))
    dhin = NO.bf16(rng.standard_normal((K, nv, 2 * C)))
    dF = torch.empty((nv, C), dtype=torch.bfloat16, device="cuda")
    de2 = torch.empty((K, nv, C), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_head_in", dF.data_ptr(), C, de2.data_ptr(), C, _dev(dhin, torch.bfloat16).data_ptr(), nv, K,
                1, _stream())
    np.testing.assert_allclose(_host(dF), NO.bf16(dhin[..., :C].sum(axis=0)), rtol=2 ** -7, atol=1e-6)
    np.testing.assert_array_equal(_host(de2), dhin[..., C:])


def test_band_reductions():
    """Band maxima of the assisted stepper's acceptance test against numpy."""
    torch = _torch()
    from mpf import _nnlib
    from mpf.surrogate import _p, _stream
    rng = np.random.default_rng(12)
    N, nx, ny, nz, z0, zb = 3, 8, 6, 20, 5, 9
    a, b, c = (rng.standard_normal((N, nx, ny, nz)).astype(np.float32) for _ in range(3))
    band = (slice(None), slice(None), slice(None), slice(z0, z0 + zb))
    # outside the band the arrays differ far more, which must not count
    b[..., :z0] += 100.0
    ad, bd, cd = (_dev(t, torch.float32) for t in (a, b, c))

    def reduce(name, *args):   # the maximum comes back as float bits in an int32
        out = torch.zeros(1, dtype=torch.int32, device="cuda")
        _nnlib.call(name, *args, _p(out), _stream())
        return np.frombuffer(np.int32(out.item()).tobytes(), np.float32)[0]

    got = reduce("mpf_nn_band_maxdiff", _p(ad), _p(bd), N, nx, ny, nz, z0, zb)
    assert got == np.abs(a[band] - b[band]).max()
    inv_h = np.float32(1.0 / 16)
    got = reduce("mpf_nn_band_residual", _p(ad), _p(bd), _p(cd), N, nx, ny, nz, z0, zb, float(inv_h))
    ref = np.abs((c[band] - b[band]) - (b[band] - a[band]) * inv_h).max()
    assert abs(got - ref) <= 1e-6 * abs(ref)


@pytest.mark.parametrize("n", [1, 1000, 3 * 128 * 128 * 32 + 17])
def test_fill_and_sum_kernels(n):
    """The fill and sum kernels that replace torch's fills and reductions in
    the training loop: a 0/1 mask's count exactly (integer sums below 2^24),
    a float sum within rounding, on a buffer the fill kernel set first;
    the sum overwrites whatever the output held."""
    torch = _torch()
    from mpf import _nnlib
    from mpf.surrogate import _p, _stream, _zeros
    rng = np.random.default_rng(n)
    mask = (rng.uniform(size=n) < 0.3).astype(np.float32)
    md = _dev(mask, torch.float32)
    out = _dev(np.array([123.0], np.float32), torch.float32)
    _nnlib.call("mpf_nn_sum", _p(md), n, _p(out), _stream())
    assert out.item() == mask.sum()
    z = _zeros(n)
    assert torch.count_nonzero(z).item() == 0
    _nnlib.call("mpf_nn_fill", _p(z), 0.5, n, _stream())
    _nnlib.call("mpf_nn_sum", _p(z), n, _p(out), _stream())
    assert out.item() == 0.5 * n
    x = rng.standard_normal(n).astype(np.float32)
    _nnlib.call("mpf_nn_sum", _p(_dev(x, torch.float32)), n, _p(out), _stream())
    assert abs(out.item() - x.astype(np.float64).sum()) <= 1e-5 * np.abs(x).sum() + 1e-6


@pytest.mark.parametrize("nvox,c", [(1, 8), (37, 3), (1000, 32), (4097, 96), (70001, 64), (2051, 512),
                                    (3 * 64 * 64 * 16 + 5, 256)])
def test_bias_gradient_kernel(nvox, c):
    """db += column sums of a bf16 [nvox, c] gradient: the 16-byte kernel
    (c a multiple of 8: one channel group per thread, several rows per
    block, also more than 256 channels) and the one-channel-per-thread
    kernel (c = 3), against FP64 sums; db accumulates onto what it held."""
    torch = _torch()
    from mpf import _nnlib
    from mpf.surrogate import _p, _stream
    rng = np.random.default_rng(nvox + c)
    dy = NO.bf16(rng.standard_normal((nvox, c)))
    db0 = rng.standard_normal(c).astype(np.float32)
    dbd = _dev(db0, torch.float32)
    _nnlib.call("mpf_nn_bias_grad", _p(_dev(dy, torch.bfloat16)), _p(dbd), nvox, c, _stream())
    ref = db0.astype(np.float64) + dy.sum(axis=0)
    got = _host(dbd)
    assert np.abs(got - ref).max() <= 1e-5 * (np.abs(dy).sum(axis=0).max() + 1.0)


def test_slot_features_match_oracle():
    torch = _torch()
    import cases
    from mpf import Solver
    from mpf.surrogate import SlotModel, orientation_features
    c = cases.voronoi_case(nx=12, ny=10, nz=16, n_solid=6, seed_height=6)
    with Solver(c["phi"].shape[1:], c["phi"].shape[0], c["scalars"], precision="fp32") as s:
        s.set_tables(*c["tables"])
        s.upload_phi(c["phi"])
        s.upload_temp(c["temp"])
        s.step(3)
        phi = s.phi(np.float32).astype(np.float64)
    N, nx, ny, nz = phi.shape
    orient = orientation_features(c["tables"][3])
    tfeat = np.linspace(-1, 1, nz).astype(np.float32)
    m = SlotModel.__new__(SlotModel)
    m.k = 3
    sid, feat, aux, over = m.features(_dev(phi.reshape(N, -1), torch.float32), N, (nx, ny, nz), 2, 8,
                                      _dev(orient, torch.float32), _dev(tfeat, torch.float32))
    rs, rf, ra, ro = NO.slot_features(phi, 2, 8, 3, orient.astype(np.float64), tfeat.astype(np.float64))
    np.testing.assert_array_equal(sid.cpu().numpy(), rs)
    np.testing.assert_array_equal(_host(feat)[..., :4], NO.bf16(rf))
    assert not _host(feat)[..., 4:].any()
    np.testing.assert_array_equal(_host(aux)[:, :3], NO.bf16(ra))
    assert int(over.item()) == ro


def _ref_from(model, bf16_storage):
    ws = [_host(c.w) for c in model.layers()]
    bs = [_host(c.b) for c in model.layers()]
    return NO.RefModel([NO.bf16(w) for w in ws], bs, 32, bf16_storage)


def test_whole_model_forward_and_gradients():
    torch = _torch()
    from mpf.surrogate import SlotModel
    rng = np.random.default_rng(11)
    m = SlotModel(seed=3, k=2)
    # a non-zero head so that every layer receives gradient
    m.h2.w.copy_(torch.tensor(rng.standard_normal(m.h2.w.shape) * 0.2, dtype=torch.float32, device="cuda"))
    m.h2.refresh()
    nx = ny = zb = 16
    k, nv = 2, nx * ny * zb
    sid = np.where(rng.random((k, nv)) < 0.7, 1 + np.arange(k)[:, None], -1).astype(np.int32)
    feat = np.zeros((k, nv, 32))
    feat[..., :4] = rng.random((k, nv, 4))
    feat = NO.bf16(feat * (sid >= 0)[..., None])
    aux = np.zeros((nv, 32))
    aux[:, :3] = rng.random((nv, 3))
    aux = NO.bf16(aux)
    out = m.forward(_dev(sid, torch.int32), _dev(feat, torch.bfloat16), _dev(aux, torch.bfloat16), (nx, ny, 16), zb)
    ref = _ref_from(m, bf16_storage=True)
    rout = ref.forward(sid, feat, aux, (nx, ny, zb))
    got = _host(out).reshape(k, nv, 32)
    err = np.linalg.norm(got - rout) / np.linalg.norm(rout)
    plain = _ref_from(m, bf16_storage=False).forward(sid, feat, aux, (nx, ny, zb))
    err64 = np.linalg.norm(got - plain) / np.linalg.norm(plain)
    print("forward rel L2 vs bf16-storage oracle", err, "vs FP64 oracle", err64)
    assert err <= 1e-2 and err64 <= 5e-2
    dout = NO.bf16(rng.standard_normal((k, nv, 32)) * 1e-2)
    m.zero_grad()
    m.backward(_dev(dout.reshape(k * nv, 32), torch.bfloat16))
    torch.cuda.synchronize()
    # the device's ReLU decisions (values within rounding of 0 would
    # otherwise flip between the two sides and dominate the difference)
    masks = {i: (_host(c.y) > 0).reshape(ref.cache[i][1].shape) for i, c in enumerate(m.layers()) if c.relu}
    gw, gb = ref.backward(dout, masks)
    res = []
    for i, c in enumerate(m.layers()):
        g, r = _host(c.gw).ravel(), gw[i].ravel()
        cos = g @ r / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-30)
        rel = np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-30)
        gbh = _host(c.gb)
        relb = np.linalg.norm(gbh - gb[i]) / (np.linalg.norm(gb[i]) + 1e-30)
        print(i, c.ntap, c.cin, c.cout, "cos", round(float(cos), 5), "rel", round(float(rel), 4),
              "bias rel", round(float(relb), 4))
        res.append((cos, rel, relb))
    for i, (cos, rel, relb) in enumerate(res):
        assert cos >= 0.9999 and rel <= 1.5e-2 and relb <= 1.5e-2, i


def test_training_reduces_loss_and_stepper_restores_state():
    torch = _torch()
    from mpf import Solver, inputs
    from mpf import surrogate as S
    samples, meta = S.generate_samples(n_grids=4, shape=(32, 32, 32), steps=96, every=32, seed=1, zb=16)
    assert len(samples) == 4 * 3
    m = S.SlotModel(seed=0)
    before = np.mean([S.sample_loss(m, smp, backward=False) for smp in samples])
    hist = S.train(m, samples, 60, lr=1e-3)
    after = np.mean([S.sample_loss(m, smp, backward=False) for smp in samples])
    print("loss before", before, "after", after)
    assert np.isfinite(hist).all() and after < before
    # stepper without the network: every proposal rejected, the state restored,
    # so the run equals plain stepping bit for bit
    from mpf.config import flat_params, load_config
    p = flat_params(load_config())
    d = inputs.derive_constants(p)
    scal = inputs.kernel_scalars(p, d, as_float32=True)
    cool = meta["cooling"]
    n_solid, shape, h = 4, (32, 32, 32), 12
    q = inputs.assign_quaternions_to_grains(n_solid, "random", 5)
    w, a, mm = inputs.build_interaction_matrices(n_solid + 1, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"],
                                                 d["w_gb"], d["m_gb_phi"])
    n111 = inputs.compute_rotated_n111(q)
    gmap = inputs.generate_random_grain_map_3d(32, 32, n_solid, random_seed=3)
    T0 = inputs.init_temperature_field_3d(1, 1, 32, p["T_melt"], p["G"], p["dz"], h)[0, 0]

    def solver():
        s = Solver(shape, n_solid + 1, scal, precision="fp32", cooling_rate=cool)
        s.set_tables(w, a, mm, n111)
        s.init_columns(*inputs.columns_from_grain_map_3d(gmap, 32, p["dz"], d["delta"], h))
        s.upload_temp(T0)
        return s

    with solver() as a_, solver() as b_:
        st = S.SurrogateStepper(a_, m, n111, p["T_melt"], meta["t_scale"], cool, zb=16)
        for _ in range(3):
            n, acc = st.advance(model=False)
            assert not acc and n == S.HORIZON
            assert st.last_residual == st.last_change > 0
        b_.step(3 * S.HORIZON)
        assert np.array_equal(a_.phi(np.float32), b_.phi(np.float32))
        np.testing.assert_array_equal(a_.temp(), b_.temp())
        assert st.rejected == 3 and st.accepted == 0
        # with the (briefly trained) network the cycle runs and accounts its steps
        n, acc = st.advance(model=True)
        assert n == (S.HORIZON + 1 if acc else S.HORIZON) and np.isfinite(st.last_residual)
