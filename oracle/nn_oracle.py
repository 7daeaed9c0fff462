"""CPU FP64 oracle of the surrogate's layers (include/mpf_nn.h).

TEST INFRASTRUCTURE ONLY: imported by tests/ to check the device kernels of
the learned surrogate and their gradients.  There is no reference
counterpart (the reference has no learned model); these are the textbook
definitions on channels-last arrays x[B, X, Y, Z, C]:
- 3x3x3 or 1x1x1 convolution with per-axis periodic or zero padding;
- its adjoints;
- 2x average pooling, 2x nearest upsampling;
- the slot features of a dense phi.

Pins (tests/test_nn_oracle.py):
- convolution: a direct six-fold loop on tiny inputs; the adjoint
  identities <conv(x), y> = <x, conv_T(y)> and <dW, W'> = <conv(x; W'), dy>;
- pooling and upsampling: adjointness; pooling of a constant is the
  constant;
- finite differences of a scalar loss through the layers.
"""
from __future__ import annotations

import numpy as np

TAPS27 = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def _shift(x, d, axis, periodic):
    """y[..i..] = x[..i+d..] along a spatial axis (1..3), wrap or zero."""
    if d == 0:
        return x
    y = np.roll(x, -d, axis=axis)
    if not periodic:
        idx = [slice(None)] * x.ndim
        idx[axis] = slice(-d, None) if d > 0 else slice(None, -d)
        y[tuple(idx)] = 0.0
    return y


def conv(x, w, b=None, periodic=(1, 1, 0), relu=False):
    """x [B,X,Y,Z,Cin], w [ntap, Cout, Cin] -> [B,X,Y,Z,Cout]."""
    ntap = w.shape[0]
    taps = TAPS27 if ntap == 27 else [(0, 0, 0)]
    y = np.zeros(x.shape[:4] + (w.shape[1],))
    for t, (dx, dy, dz) in enumerate(taps):
        xs = _shift(_shift(_shift(x, dx, 1, periodic[0]), dy, 2, periodic[1]), dz, 3, periodic[2])
        y += xs @ w[t].T
    if b is not None:
        y += b
    return np.maximum(y, 0.0) if relu else y


def conv_bwd(x, w, dy, periodic=(1, 1, 0)):
    """(dx, dw) of y = conv(x, w) for the output gradient dy."""
    ntap = w.shape[0]
    taps = TAPS27 if ntap == 27 else [(0, 0, 0)]
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, (ax, ay, az) in enumerate(taps):
        xs = _shift(_shift(_shift(x, ax, 1, periodic[0]), ay, 2, periodic[1]), az, 3, periodic[2])
        dw[t] = np.einsum("bxyzo,bxyzi->oi", dy, xs)
        g = dy @ w[t]
        dx += _shift(_shift(_shift(g, -ax, 1, periodic[0]), -ay, 2, periodic[1]), -az, 3, periodic[2])
    return dx, dw


def conv_direct(x, w, b=None, periodic=(1, 1, 0)):
    """Six-fold loop definition, for tiny inputs."""
    B, X, Y, Z, _ = x.shape
    ntap, co, _ = w.shape
    taps = TAPS27 if ntap == 27 else [(0, 0, 0)]
    y = np.zeros((B, X, Y, Z, co))
    dims = (X, Y, Z)
    for bb in range(B):
        for i in range(X):
            for j in range(Y):
                for k in range(Z):
                    for t, d in enumerate(taps):
                        p = [i + d[0], j + d[1], k + d[2]]
                        ok = True
                        for a in range(3):
                            if not 0 <= p[a] < dims[a]:
                                if periodic[a]:
                                    p[a] %= dims[a]
                                else:
                                    ok = False
                        if ok:
                            y[bb, i, j, k] += w[t] @ x[bb, p[0], p[1], p[2]]
    if b is not None:
        y += b
    return y


def pool(x):
    B, X, Y, Z, C = x.shape
    return x.reshape(B, X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(dy):
    B, X, Y, Z, C = dy.shape
    return dy.reshape(B, X // 2, 2, Y // 2, 2, Z // 2, 2, C).sum(axis=(2, 4, 6))


def slot_features(phi, z0, zb, k, orient, tfeat):
    """Slots of the band of a dense phi[N, nx, ny, nz]: grains with phi > 0
    at the cell or a face neighbour (x, y periodic, z clamped), ascending,
    at most k.  Returns sid[k, nv], feat[k, nv, 4], aux[nv, 3], overflow."""
    N, nx, ny, nz = phi.shape
    nv = nx * ny * zb
    sid = -np.ones((k, nv), np.int64)
    feat = np.zeros((k, nv, 4))
    aux = np.zeros((nv, 3))
    over = 0
    pos = phi[1:] > 0
    act = pos | np.roll(pos, 1, 1) | np.roll(pos, -1, 1) | np.roll(pos, 1, 2) | np.roll(pos, -1, 2)
    up_ = np.concatenate([pos[..., 1:], pos[..., -1:]], axis=3)
    dn_ = np.concatenate([pos[..., :1], pos[..., :-1]], axis=3)
    act |= up_ | dn_
    for l in range(nx):
        for m in range(ny):
            for kz in range(zb):
                z = z0 + kz
                v = (l * ny + m) * zb + kz
                gs = [g + 1 for g in range(N - 1) if act[g, l, m, z]]
                if len(gs) > k:
                    over += 1
                for s, g in enumerate(gs[:k]):
                    sid[s, v] = g
                    feat[s, v] = (phi[g, l, m, z], *orient[g])
                aux[v] = (phi[0, l, m, z], tfeat[z], 1.0)
    return sid, feat, aux, over


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) and back to float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


# ------------------------------------------------------------- whole model
class RefModel:
    """FP64 restatement of mpf.surrogate.SlotModel (same layer order and
    shapes) from its fp32 master weights: per-slot encoder, slot sum,
    4-level U-Net, per-slot head.  forward() returns the per-slot outputs
    [k, nv, 32]; backward(dout) returns the weight and bias gradients in
    layer order."""

    def __init__(self, weights, biases, base=32, bf16_storage=False):
        """bf16_storage: round every stored activation and activation
        gradient to bf16 where the device kernels store them (conv outputs,
        pooled and summed tensors and their gradients); arithmetic stays
        FP64."""
        self.w = [np.asarray(w, np.float64) for w in weights]
        self.b = [np.asarray(b, np.float64) for b in biases]
        self.base = base
        self.r = bf16 if bf16_storage else (lambda a: a)

    def _conv(self, i, x, relu, per, store=True):
        y = conv(x, self.w[i], self.b[i], per, relu)
        if store:
            y = self.r(y)
        self.cache[i] = (x, y, relu, per)
        return y

    def _conv_bwd(self, i, dy):
        x, y, relu, per = self.cache[i]
        if relu:
            m = self.masks[i] if self.masks is not None and i in self.masks else (y > 0)
            dy = dy * m
        dx, dw = conv_bwd(x, self.w[i], dy, per)
        self.gw[i] += dw
        self.gb[i] += dy.sum(axis=(0, 1, 2, 3))
        return self.r(dx)

    def forward(self, sid, feat, aux, shape):
        """sid [k, nv], feat [k, nv, 32], aux [nv, 32]; shape (nx, ny, zb)."""
        nx, ny, zb = shape
        k = sid.shape[0]
        self.cache = {}
        self.k, self.shape = k, shape
        per1 = (1, 1, 0)
        valid = (sid >= 0)[..., None]
        e = self._conv(0, feat.reshape(k, nx, ny, zb, 32), True, per1).reshape(k, -1, 32)
        h = self.r((e * valid).sum(axis=0))
        x = np.concatenate([h, aux], axis=1).reshape(1, nx, ny, zb, 64)
        skips = []
        i = 1
        for lvl in range(4):
            x = self._conv(i, x, True, per1)
            x = self._conv(i + 1, x, True, per1)
            i += 2
            if lvl < 3:
                skips.append(x)
                x = self.r(pool(x))
        for lvl in (2, 1, 0):
            x = np.concatenate([up(x), skips[lvl]], axis=4)
            x = self._conv(i, x, True, per1)
            x = self._conv(i + 1, x, True, per1)
            i += 2
        F = x.reshape(-1, self.base)
        self.cf = F.shape[1]
        hin = np.concatenate([np.broadcast_to(F, (k,) + F.shape), e], axis=2).reshape(k, nx, ny, zb, -1)
        y1 = self._conv(i, hin, True, per1)
        out = self._conv(i + 1, y1, False, per1, store=False)
        self.valid = valid
        self.nl = i + 2
        return out.reshape(k, -1, 32)

    def backward(self, dout, masks=None):
        """masks: optional {layer: boolean array shaped like its output}
        replacing the ReLU decisions (y > 0), e.g. by another
        implementation's, so that outputs sitting at zero within rounding
        do not flip the comparison."""
        self.masks = masks
        nx, ny, zb = self.shape
        k = self.k
        self.gw = [np.zeros_like(w) for w in self.w]
        self.gb = [np.zeros_like(b) for b in self.b]
        i = self.nl - 2
        d = self._conv_bwd(i + 1, dout.reshape(k, nx, ny, zb, 32))
        dhin = self._conv_bwd(i, d).reshape(k, -1, self.cf + 32)
        dF = self.r(dhin[..., :self.cf].sum(axis=0)).reshape(1, nx, ny, zb, self.cf)
        de = dhin[..., self.cf:]
        # U-Net backward: decoder levels 0, 1, 2 in reverse, then encoder
        x = dF
        dskip = {}
        j = i - 2
        for lvl in (0, 1, 2):
            x = self._conv_bwd(j + 1, x)
            x = self._conv_bwd(j, x)
            skip_c = self.base * (1 << lvl)
            dskip[lvl] = x[..., -skip_c:]
            x = self.r(up_bwd(x[..., :-skip_c]))
            j -= 2
        for lvl in (3, 2, 1, 0):
            if lvl < 3:
                x = self.r(self.r(pool_bwd(x)) + dskip[lvl])
            x = self._conv_bwd(j + 1, x)
            x = self._conv_bwd(j, x)
            j -= 2
        dh = x.reshape(-1, 64)[:, :32]
        de = self.r(de + self.valid * dh[None])
        self._conv_bwd(0, de.reshape(k, nx, ny, zb, 32))
        return self.gw, self.gb
