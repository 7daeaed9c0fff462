"""CPU pins of the surrogate's layer oracle (oracle/nn_oracle.py): the
vectorised convolution equals the six-fold loop definition; its backward
satisfies the adjoint identities; pooling and upsampling are adjoint; the
whole-model gradient equals finite differences; slot features on a
hand-made field."""
import numpy as np
import pytest

from oracle import nn_oracle as NO


@pytest.mark.parametrize("ntap,periodic", [(27, (1, 1, 0)), (27, (0, 0, 0)), (27, (1, 1, 1)), (1, (1, 1, 0))])
def test_conv_matches_loop_definition(ntap, periodic):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 4, 5, 3))
    w = rng.standard_normal((ntap, 4, 3))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(NO.conv(x, w, b, periodic), NO.conv_direct(x, w, b, periodic), rtol=1e-12,
                               atol=1e-12)


@pytest.mark.parametrize("periodic", [(1, 1, 0), (0, 1, 0), (1, 1, 1)])
def test_conv_backward_adjoint_identities(periodic):
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 4, 3, 6, 5))
    w = rng.standard_normal((27, 7, 5))
    dy = rng.standard_normal((1, 4, 3, 6, 7))
    dx, dw = NO.conv_bwd(x, w, dy, periodic)
    assert np.sum(NO.conv(x, w, None, periodic) * dy) == pytest.approx(np.sum(x * dx), rel=1e-12)
    w2 = rng.standard_normal(w.shape)
    assert np.sum(dw * w2) == pytest.approx(np.sum(NO.conv(x, w2, None, periodic) * dy), rel=1e-12)


def test_pool_upsample_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 6, 8, 3))
    y = rng.standard_normal((1, 2, 3, 4, 3))
    assert np.sum(NO.pool(x) * y) == pytest.approx(np.sum(x * NO.pool_bwd(y)), rel=1e-12)
    assert np.sum(NO.up(y) * x) == pytest.approx(np.sum(y * NO.up_bwd(x)), rel=1e-12)
    assert np.allclose(NO.pool(np.full((1, 4, 4, 4, 2), 0.75)), 0.75)


def test_bf16_rounding():
    v = NO.bf16(np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38]))
    assert v[0] == 1.0 and v[1] == 1.0            # tie to even
    assert v[2] == 1.0 + 2 ** -7                  # rounds up
    assert v[3] == -2.5


def _tiny_model(rng, base=32, k=2, shape=(8, 8, 8)):
    from mpf.surrogate import BASE  # noqa: F401  (layer order documented there)
    specs = [(1, 32, 32)]
    c = [base, 2 * base, 4 * base, 8 * base]
    cin = 64
    for lvl in range(4):
        specs += [(27, c[lvl], cin), (27, c[lvl], c[lvl])]
        cin = c[lvl]
    for lvl in (2, 1, 0):
        specs += [(27, c[lvl], c[lvl + 1] + c[lvl]), (27, c[lvl], c[lvl])]
    specs += [(1, 32, base + 32), (1, 32, 32)]
    ws = [rng.standard_normal(s) * np.sqrt(2.0 / (s[0] * s[2])) for s in specs]
    bs = [rng.standard_normal(s[1]) * 0.1 for s in specs]
    return NO.RefModel(ws, bs, base)


def test_whole_model_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    m = _tiny_model(rng)
    nx = ny = zb = 8
    k, nv = 2, 8 * 8 * 8
    sid = np.where(rng.random((k, nv)) < 0.7, 1, -1)
    feat = rng.random((k, nv, 32)) * (sid >= 0)[..., None]
    aux = rng.random((nv, 32))
    R = rng.standard_normal((k, nv, 32))
    out = m.forward(sid, feat, aux, (nx, ny, zb))
    gw, gb = m.backward(R)
    base_w = [w.copy() for w in m.w]
    for layer, idx in ((0, (0, 3, 1)), (3, (13, 5, 7)), (9, (4, 2, 100)), (16, (0, 0, 20)), (15, (0, 3, 40)), (8, (26, 255, 200))):
        eps = 1e-6
        m.w = [w.copy() for w in base_w]
        m.w[layer][idx] += eps
        lp = np.sum(m.forward(sid, feat, aux, (nx, ny, zb)) * R)
        m.w[layer][idx] -= 2 * eps
        lm = np.sum(m.forward(sid, feat, aux, (nx, ny, zb)) * R)
        fd = (lp - lm) / (2 * eps)
        assert gw[layer][idx] == pytest.approx(fd, rel=1e-5, abs=1e-7), (layer, idx)
    m.w = base_w
    assert out.shape == (k, nv, 32)


def test_slot_features_on_hand_made_field():
    phi = np.zeros((4, 3, 3, 6))
    phi[0] = 1.0
    phi[1, 0, 0, 1] = 0.6
    phi[0, 0, 0, 1] = 0.4
    phi[3, 1, 1, 4] = 0.2
    phi[0, 1, 1, 4] = 0.8
    orient = np.arange(12, dtype=float).reshape(4, 3)
    sid, feat, aux, over = NO.slot_features(phi, 1, 4, 2, orient, np.arange(6) * 0.5)
    ny, zb = 3, 4
    v = (0 * ny + 0) * zb + 0          # cell (0, 0, 1)
    assert list(sid[:, v]) == [1, -1] and feat[0, v, 0] == 0.6 and list(feat[0, v, 1:]) == [3, 4, 5]
    v2 = (0 * ny + 0) * zb + 1         # cell (0, 0, 2): grain 1 is a face neighbour below
    assert list(sid[:, v2]) == [1, -1] and feat[0, v2, 0] == 0.0
    v3 = (2 * ny + 0) * zb + 0         # cell (2, 0, 1): periodic x neighbour of (0, 0, 1)
    assert sid[0, v3] == 1
    v4 = (1 * ny + 1) * zb + 3         # cell (1, 1, 4)
    assert list(sid[:, v4]) == [3, -1] and aux[v4, 0] == 0.8 and aux[v4, 1] == 2.0
    assert over == 0
