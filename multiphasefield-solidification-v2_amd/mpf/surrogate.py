"""Learned surrogate: a slot-equivariant 3D U-Net that proposes 16-step
advances of the phase field.  The reference has no learned model.

Model (per cell, over a z-band of the grid that holds the front):
- slots: the cell's candidate grains (phi > 0 at the cell or a face
  neighbour), at most K.  Each slot's features are its phi value and three
  orientation features of its grain: max over the 8 {111} normals of
  |n.x|, |n.y| and |n.z|.
- a shared per-slot encoder (1x1x1 conv, 32 channels) whose outputs are
  summed over the cell's slots: a permutation-invariant 32-channel cell
  code.  With the liquid fraction, the temperature and a constant, that code
  is the U-Net input (64 channels).
- a 4-level U-Net with 32 base channels.  Two 3x3x3 convs per level, 2x
  average pooling, nearest upsampling and skip concatenation; x and y
  periodic, z zero-padded.
- a shared per-slot head (1x1x1 convs 64 -> 32 -> 1) on (U-Net features,
  slot code) predicting that slot's 16-step increment of phi.
Because the encoder and the head are shared and the pooling is a sum,
permuting the grains permutes the predictions: the model is
slot-equivariant.

Everything runs in the HIP kernels of include/mpf_nn.h: bf16 MFMA
convolutions with fp32 accumulation, their backward passes and Adam.  torch
only provides device buffers, so importing this module imports torch; do it
before the first solver of the process (see _torch).  The slot features are built on the device
from the solver's dense phi (mpf_download_phi_device); a proposal is written
back with mpf_upload_phi_device.

Surrogate-assisted stepping (SurrogateStepper):
- propose a 16-step advance P;
- take one exact step from P;
- accept when that step's change agrees with the proposal's mean rate,
  max |(step(P) - P) - (P - phi)/16| <= rate_tol x max |step(P) - P|;
- otherwise restore the state and take 16 exact steps instead.
"""
from __future__ import annotations

import math

import numpy as np

from . import _nnlib
from ._lib import FP32, check
from . import _lib

try:
    import torch as _torch_module      # before libmpf.so: the two must share torch's HIP runtime
except ImportError:
    _torch_module = None

SCALE = 20.0        # targets are 20 x the 16-step increment
HORIZON = 16
K_SLOTS = 6
BASE = 32


def _torch():
    """torch, checked to see the GPU.

    PyTorch-ROCm carries its own libamdhip64 under the soname of the system
    runtime that libmpf.so links against, and the dynamic linker binds every
    later load to the copy loaded first.  With torch imported first, libmpf.so
    shares torch's runtime and both see the GPU; with libmpf.so loaded first,
    torch is bound to the system runtime and finds no GPU."""
    if _torch_module is None:
        raise RuntimeError("the surrogate needs PyTorch (ROCm) for its device buffers")
    if not _torch_module.cuda.is_available():
        hint = (" libmpf.so was loaded before torch in this process: import mpf.surrogate (or torch) before "
                "creating the first solver" if _lib.loaded() else "")
        raise RuntimeError("the surrogate runs on the GPU (HIP) only and torch sees no GPU." + hint)
    return _torch_module


def _stream():
    return _torch().cuda.current_stream().cuda_stream


def _p(t):
    return t.data_ptr()


def _zeros(*shape, dtype=None):
    """A device buffer cleared by the project's fill kernel (4-byte elements;
    +0.0f has the bits of integer 0)."""
    torch = _torch()
    t = torch.empty(shape, dtype=dtype or torch.float32, device="cuda")
    assert t.element_size() == 4
    _nnlib.call("mpf_nn_fill", _p(t), 0.0, t.numel(), _stream())
    return t


def _zero(t):
    assert t.element_size() == 4 and t.is_contiguous()
    _nnlib.call("mpf_nn_fill", _p(t), 0.0, t.numel(), _stream())


class Conv:
    """One convolution layer: fp32 master weights [ntap][cout][cin] (He
    initialisation), bias, Adam moments, bf16 copies for the forward and
    backward-data passes."""

    def __init__(self, cin, cout, ntap, relu, rng, zero=False):
        torch = _torch()
        self.cin, self.cout, self.ntap, self.relu = cin, cout, ntap, relu
        std = 0.0 if zero else math.sqrt(2.0 / (cin * ntap))
        w = rng.standard_normal((ntap, cout, cin)).astype(np.float32) * std
        dev = "cuda"
        self.w = torch.tensor(w, device=dev)
        self.b = _zeros(cout)
        self.gw, self.gb = _zeros(*self.w.shape), _zeros(cout)
        self.mw, self.vw = _zeros(*self.w.shape), _zeros(*self.w.shape)
        self.mb, self.vb = _zeros(cout), _zeros(cout)
        self.wb = torch.empty(self.w.shape, dtype=torch.bfloat16, device=dev)
        self.wf = torch.empty((ntap, cin, cout), dtype=torch.bfloat16, device=dev)
        self.refresh()

    def params(self):
        return [(self.w, self.gw, self.mw, self.vw), (self.b, self.gb, self.mb, self.vb)]

    def refresh(self):
        s = _stream()
        _nnlib.call("mpf_nn_weights_bf16", _p(self.w), _p(self.wb), self.ntap, self.cout, self.cin, 0, s)
        _nnlib.call("mpf_nn_weights_bf16", _p(self.w), _p(self.wf), self.ntap, self.cout, self.cin, 1, s)

    def forward(self, x, dims, periodic, out_f32=False):
        torch = _torch()
        nv = dims[0] * dims[1] * dims[2] * dims[3]
        if x.numel() != nv * self.cin or x.dtype != torch.bfloat16 or not x.is_contiguous():
            raise ValueError(f"conv input: {tuple(x.shape)} {x.dtype}, expected {nv} x {self.cin} contiguous bf16")
        x = x.view(nv, self.cin)
        y = torch.empty((nv, self.cout), dtype=torch.bfloat16, device="cuda")
        yf = torch.empty((nv, self.cout), dtype=torch.float32, device="cuda") if out_f32 else None
        _nnlib.call("mpf_nn_conv_fwd", _p(x), _p(self.wb), _p(self.b), _p(y), _p(yf) if yf is not None else None,
                    _nnlib.ivec(*dims), self.cin, self.cout, self.ntap, _nnlib.ivec(*periodic), int(self.relu),
                    _stream())
        self.x, self.y, self.dims, self.periodic = x, y, dims, periodic
        return (y, yf) if out_f32 else y

    def backward(self, dy, need_dx=True):
        """dy: bf16 [nv, cout] (modified in place by the ReLU mask).
        Accumulates gw, gb; returns dx or None."""
        torch = _torch()
        s = _stream()
        nv = self.dims[0] * self.dims[1] * self.dims[2] * self.dims[3]
        if dy.numel() != nv * self.cout or dy.dtype != torch.bfloat16 or not dy.is_contiguous():
            raise ValueError(f"conv output gradient: {tuple(dy.shape)} {dy.dtype}, expected {nv} x {self.cout} "
                             "contiguous bf16")
        dy = dy.view(nv, self.cout)
        if self.relu:
            _nnlib.call("mpf_nn_relu_bwd", _p(dy), _p(self.y), dy.numel(), s)
        _nnlib.call("mpf_nn_bias_grad", _p(dy), _p(self.gb), nv, self.cout, s)
        _nnlib.call("mpf_nn_conv_bwd_w", _p(self.x), _p(dy), _p(self.gw), _nnlib.ivec(*self.dims), self.cin,
                    self.cout, self.ntap, _nnlib.ivec(*self.periodic), s)
        if not need_dx:
            return None
        dx = torch.empty((nv, self.cin), dtype=torch.bfloat16, device="cuda")
        _nnlib.call("mpf_nn_conv_fwd", _p(dy), _p(self.wf), None, _p(dx), None, _nnlib.ivec(*self.dims),
                    self.cout, self.cin, self.ntap, _nnlib.ivec(*self.periodic), 0, s)
        return dx


def _need(t, n, dtype, what, at_least=False):
    """Host-side check of a device operand before any kernel reads it."""
    ok = (t.numel() >= n if at_least else t.numel() == n) and t.dtype == dtype and t.is_contiguous() and t.is_cuda
    if not ok:
        raise ValueError(f"{what}: {tuple(t.shape)} {t.dtype}, expected {n} contiguous {dtype} on the GPU")


def _new(nv, c):
    torch = _torch()
    return torch.empty((nv, c), dtype=torch.bfloat16, device="cuda")


def _cat(a, ca, b, cb, nv):
    out = _new(nv, ca + cb)
    s = _stream()
    _nnlib.call("mpf_nn_copy_channels", _p(a), ca, 0, _p(out), ca + cb, 0, ca, nv, 0, s)
    _nnlib.call("mpf_nn_copy_channels", _p(b), cb, 0, _p(out), ca + cb, ca, cb, nv, 0, s)
    return out


def _split(d, ca, cb, nv):
    a, b = _new(nv, ca), _new(nv, cb)
    s = _stream()
    _nnlib.call("mpf_nn_copy_channels", _p(d), ca + cb, 0, _p(a), ca, 0, ca, nv, 0, s)
    _nnlib.call("mpf_nn_copy_channels", _p(d), ca + cb, ca, _p(b), cb, 0, cb, nv, 0, s)
    return a, b


def _add_into(dst, src, c, nv):
    _nnlib.call("mpf_nn_copy_channels", _p(src), c, 0, _p(dst), c, 0, c, nv, 1, _stream())


class UNet:
    """4 levels, channels BASE * (1, 2, 4, 8); input cin channels, output
    BASE channels (ReLU)."""

    def __init__(self, cin, rng, base=BASE):
        c1, c2, c3, c4 = base, 2 * base, 4 * base, 8 * base
        self.ch = (c1, c2, c3, c4)
        self.enc = [(Conv(cin, c1, 27, True, rng), Conv(c1, c1, 27, True, rng)),
                    (Conv(c1, c2, 27, True, rng), Conv(c2, c2, 27, True, rng)),
                    (Conv(c2, c3, 27, True, rng), Conv(c3, c3, 27, True, rng)),
                    (Conv(c3, c4, 27, True, rng), Conv(c4, c4, 27, True, rng))]
        self.dec = [(Conv(c4 + c3, c3, 27, True, rng), Conv(c3, c3, 27, True, rng)),
                    (Conv(c3 + c2, c2, 27, True, rng), Conv(c2, c2, 27, True, rng)),
                    (Conv(c2 + c1, c1, 27, True, rng), Conv(c1, c1, 27, True, rng))]

    def layers(self):
        return [c for pair in self.enc + self.dec for c in pair]

    def forward(self, x, dims, periodic):
        s = _stream()
        B, X, Y, Z = dims
        self.dims_l = [(B, X >> i, Y >> i, Z >> i) for i in range(4)]
        self.skips = []
        h = x
        for lvl, (ca, cb) in enumerate(self.enc):
            d = self.dims_l[lvl]
            h = cb.forward(ca.forward(h, d, periodic), d, periodic)
            if lvl < 3:
                self.skips.append(h)
                dn = self.dims_l[lvl + 1]
                p = _new(dn[0] * dn[1] * dn[2] * dn[3], cb.cout)
                _nnlib.call("mpf_nn_pool", _p(h), _p(p), _nnlib.ivec(*d), cb.cout, 0, s)
                h = p
        self.up_src_c = []
        for j, (ca, cb) in enumerate(self.dec):
            lvl = 2 - j
            d, dc = self.dims_l[lvl], self.dims_l[lvl + 1]
            nv = d[0] * d[1] * d[2] * d[3]
            cc = self.enc[lvl + 1][1].cout if j == 0 else self.dec[j - 1][1].cout
            u = _new(nv, cc)
            _nnlib.call("mpf_nn_upsample", _p(h), _p(u), _nnlib.ivec(*dc), cc, 0, s)
            skip = self.skips[lvl]
            cat = _cat(u, cc, skip, self.enc[lvl][1].cout, nv)
            self.up_src_c.append(cc)
            h = cb.forward(ca.forward(cat, d, periodic), d, periodic)
        return h

    def backward(self, dh):
        s = _stream()
        dskip = [None, None, None]
        for j in reversed(range(3)):
            ca, cb = self.dec[j]
            lvl = 2 - j
            d, dc = self.dims_l[lvl], self.dims_l[lvl + 1]
            nv = d[0] * d[1] * d[2] * d[3]
            dcat = ca.backward(cb.backward(dh))
            cc = self.up_src_c[j]
            du, dskip[lvl] = _split(dcat, cc, self.enc[lvl][1].cout, nv)
            nvc = dc[0] * dc[1] * dc[2] * dc[3]
            dh = _new(nvc, cc)
            _nnlib.call("mpf_nn_upsample", _p(du), _p(dh), _nnlib.ivec(*dc), cc, 1, s)
        for lvl in reversed(range(4)):
            ca, cb = self.enc[lvl]
            d = self.dims_l[lvl]
            nv = d[0] * d[1] * d[2] * d[3]
            if lvl < 3:
                # dh is the gradient of the pooled output of this level
                g = _new(nv, cb.cout)
                _nnlib.call("mpf_nn_pool", _p(dh), _p(g), _nnlib.ivec(*d), cb.cout, 1, s)
                _add_into(g, dskip[lvl], cb.cout, nv)
                dh = g
            dh = ca.backward(cb.backward(dh))
        return dh


def orientation_features(n111: np.ndarray) -> np.ndarray:
    """[N, 3]: max over the 8 normals of |n.x|, |n.y|, |n.z| per grain."""
    n = np.asarray(n111, np.float64).reshape(-1, 8, 3)
    return np.abs(n).max(axis=1).astype(np.float32)


class SlotModel:
    """Slot encoder + U-Net + slot head; see the module docstring."""

    def __init__(self, seed=0, k=K_SLOTS, base=BASE):
        rng = np.random.default_rng(seed)
        self.k = k
        self.enc = Conv(32, 32, 1, True, rng)
        self.unet = UNet(64, rng, base)
        self.h1 = Conv(base + 32, 32, 1, True, rng)
        self.h2 = Conv(32, 32, 1, False, rng, zero=True)
        self.t = 0

    def layers(self):
        return [self.enc] + self.unet.layers() + [self.h1, self.h2]

    def n_params(self):
        return sum(c.w.numel() + c.b.numel() for c in self.layers())

    def features(self, phi, n_phase, shape, z0, zb, orient, tfeat):
        """Device slot features of the band of a dense float32 phi."""
        torch = _torch()
        nx, ny, nz = shape
        nv = nx * ny * zb
        _need(phi, n_phase * nx * ny * nz, torch.float32, "phi")
        _need(orient, n_phase * 3, torch.float32, "orientation features", at_least=True)
        _need(tfeat, nz, torch.float32, "temperature feature", at_least=True)
        if not (0 <= z0 and zb >= 8 and z0 + zb <= nz):
            raise ValueError(f"band [{z0}, {z0 + zb}) outside the grid of nz = {nz}")
        sid = torch.empty((self.k, nv), dtype=torch.int32, device="cuda")
        feat = torch.empty((self.k, nv, 32), dtype=torch.bfloat16, device="cuda")
        aux = torch.empty((nv, 32), dtype=torch.bfloat16, device="cuda")
        over = _zeros(1, dtype=torch.int32)
        _nnlib.call("mpf_nn_slot_features", _p(phi), n_phase, nx, ny, nz, z0, zb, self.k, _p(orient), _p(tfeat),
                    _p(sid), _p(feat), _p(aux), _p(over), _stream())
        return sid, feat, aux, over

    def forward(self, sid, feat, aux, shape, zb):
        """Per-slot predictions: fp32 [k, nv, 32], channel 0 = SCALE x the
        16-step increment."""
        torch = _torch()
        s = _stream()
        nx, ny, _ = shape
        nv = nx * ny * zb
        if nx % 8 or ny % 8 or zb % 8:
            raise ValueError("the U-Net needs nx, ny and the band height divisible by 8")
        _need(sid, self.k * nv, torch.int32, "sid")
        _need(feat, self.k * nv * 32, torch.bfloat16, "slot features")
        _need(aux, nv * 32, torch.bfloat16, "cell features")
        per = (1, 1, 0)
        self.sid, self.nv = sid, nv
        self.dims = (1, nx, ny, zb)
        self.sdims = (self.k, nx, ny, zb)
        e = self.enc.forward(feat.view(-1, 32), self.sdims, per)
        h = _new(nv, 32)
        _nnlib.call("mpf_nn_slot_pool", _p(e), _p(sid), _p(h), nv, self.k, 32, 0, 0, s)
        x = _cat(h, 32, aux, 32, nv)
        F = self.unet.forward(x, self.dims, per)
        self.cf = F.shape[1]
        hin = _new(self.k * nv, self.cf + 32)
        _nnlib.call("mpf_nn_head_in", _p(F), self.cf, _p(e), 32, _p(hin), nv, self.k, 0, s)
        y1 = self.h1.forward(hin, self.sdims, per)
        _, out = self.h2.forward(y1, self.sdims, per, out_f32=True)
        return out

    def backward(self, dout):
        s = _stream()
        nv = self.nv
        dhin = self.h1.backward(self.h2.backward(dout))
        dF = _new(nv, self.cf)
        de = _new(self.k * nv, 32)
        _nnlib.call("mpf_nn_head_in", _p(dF), self.cf, _p(de), 32, _p(dhin), nv, self.k, 1, s)
        dx = self.unet.backward(dF)
        dh, _ = _split(dx, 32, 32, nv)
        _nnlib.call("mpf_nn_slot_pool", _p(dh), _p(self.sid), _p(de), nv, self.k, 32, 1, 1, s)
        self.enc.backward(de, need_dx=False)

    def zero_grad(self):
        for c in self.layers():
            _zero(c.gw)
            _zero(c.gb)

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        s = _stream()
        for c in self.layers():
            for p, g, m, v in c.params():
                _nnlib.call("mpf_nn_adam", _p(p), _p(g), _p(m), _p(v), p.numel(), lr, b1, b2, eps, self.t, s)
            c.refresh()

    def state_dict(self):
        return {f"l{i}_{n}": t.detach().cpu().numpy() for i, c in enumerate(self.layers())
                for n, t in (("w", c.w), ("b", c.b))}

    def load_state_dict(self, sd):
        torch = _torch()
        for i, c in enumerate(self.layers()):
            c.w.copy_(torch.tensor(sd[f"l{i}_w"], device="cuda"))
            c.b.copy_(torch.tensor(sd[f"l{i}_b"], device="cuda"))
            c.refresh()


# ------------------------------------------------------------------ training
class Sample:
    """A training pair on the device: the band of a dense phi at t and
    t + HORIZON, the grid shape, the orientation features and the
    temperature feature column."""

    def __init__(self, phi0, phi1, n_phase, shape, z0, zb, orient, tfeat):
        self.phi0, self.phi1 = phi0, phi1
        self.n_phase, self.shape, self.z0, self.zb = n_phase, shape, z0, zb
        self.orient, self.tfeat = orient, tfeat


def sample_loss(model: SlotModel, smp: Sample, backward=True):
    """Masked MSE of the scaled 16-step increments (and its gradients)."""
    torch = _torch()
    s = _stream()
    sid, feat, aux, _ = model.features(smp.phi0, smp.n_phase, smp.shape, smp.z0, smp.zb, smp.orient, smp.tfeat)
    out = model.forward(sid, feat, aux, smp.shape, smp.zb)
    nv = model.nv
    n = model.k * nv
    _need(smp.phi1, smp.phi0.numel(), torch.float32, "phi at t + HORIZON")
    tgt = torch.empty(n, dtype=torch.float32, device="cuda")
    mask = torch.empty(n, dtype=torch.float32, device="cuda")
    nx, ny, nz = smp.shape
    _nnlib.call("mpf_nn_slot_targets", _p(smp.phi0), _p(smp.phi1), nx, ny, nz, smp.z0, smp.zb, model.k, _p(sid),
                SCALE, _p(tgt), _p(mask), s)
    msum = torch.empty(1, dtype=torch.float32, device="cuda")
    _nnlib.call("mpf_nn_sum", _p(mask), n, _p(msum), s)
    cnt = float(msum.item())
    loss = _zeros(1)
    dout = torch.empty((n, 32), dtype=torch.bfloat16, device="cuda")
    _nnlib.call("mpf_nn_mse", _p(out), 32, _p(tgt), _p(mask), n, 1.0 / max(cnt, 1.0), _p(dout), _p(loss), s)
    if backward:
        model.backward(dout)
    return float(loss.item()) / max(cnt, 1.0)


def train(model: SlotModel, samples, steps, lr=1e-3, log=None, seed=0):
    """Adam over randomly drawn samples; returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    for it in range(steps):
        smp = samples[int(rng.integers(len(samples)))]
        model.zero_grad()
        lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * it / max(steps, 1)))
        hist.append(sample_loss(model, smp))
        model.adam(lr_t)
        if log and (it % 50 == 0 or it == steps - 1):
            log(f"train step {it}: loss {np.mean(hist[-50:]):.5f}")
    return hist


# ------------------------------------------------------------- trajectories
def generate_samples(n_grids=32, shape=(64, 64, 64), steps=2000, every=100, seed=0, n_phase=9, zb=32,
                     t_scale=0.5, log=None):
    """Training pairs from the batched solver.  Each grid is a random-mode
    seed layer in the configuration's physics (config_3d.yaml, as
    run_randommode_3d.py): 1 to n_phase-1 Voronoi grains, seed height 10-20,
    random orientations, linear temperature, uniform cooling.  At every
    `every` steps the band around each grid's front is recorded at t and
    t + HORIZON.  Returns (samples, meta)."""
    torch = _torch()
    from . import inputs
    from .batch import Batch
    from .config import flat_params, load_config
    p = flat_params(load_config())
    d = inputs.derive_constants(p)
    scal = inputs.kernel_scalars(p, d, as_float32=True)
    cooling = float(np.float32(p["G"] * p["V_pulling"] * p["dt"]))
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    w, a, m = inputs.build_interaction_matrices(n_phase, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"],
                                                d["w_gb"], d["m_gb_phi"])
    orients = []
    samples = []
    with Batch([shape] * n_grids, n_phase, scal, precision="fp32", cooling_rate=cooling) as b:
        for i in range(n_grids):
            n_solid = int(rng.integers(1, n_phase))
            h = int(rng.integers(10, 21))
            q = inputs.assign_quaternions_to_grains(n_phase - 1, "random", int(rng.integers(1 << 30)))
            n111 = inputs.compute_rotated_n111(q)
            b.set_tables(i, w, a, m, n111)
            gmap = inputs.generate_random_grain_map_3d(nx, ny, n_solid, random_seed=int(rng.integers(1 << 30)))
            b.init_columns(i, *inputs.columns_from_grain_map_3d(gmap, nz, p["dz"], d["delta"], h))
            b.upload_temp(i, inputs.init_temperature_field_3d(1, 1, nz, p["T_melt"], p["G"], p["dz"], h)[0, 0])
            orients.append(torch.tensor(orientation_features(n111), device="cuda"))
        ncell = nx * ny * nz
        t = 0
        while t + HORIZON <= steps:
            gv, out, fronts = b.observables_all()
            phi0 = [torch.empty((n_phase, ncell), dtype=torch.float32, device="cuda") for _ in range(n_grids)]
            tf = []
            for i in range(n_grids):
                check(_lib.lib().mpf_batch_download_phi_device(b._h, i, _p(phi0[i]), FP32))
                T = b.temp(i)
                tf.append(torch.tensor(((T - p["T_melt"]) / t_scale).astype(np.float32), device="cuda"))
            b.step(HORIZON)
            for i in range(n_grids):
                phi1 = torch.empty_like(phi0[i])
                check(_lib.lib().mpf_batch_download_phi_device(b._h, i, _p(phi1), FP32))
                f = fronts[i]
                lo = int(f[f >= 0].min()) if (f >= 0).any() else 0
                z0 = max(0, min(nz - zb, lo - zb // 2))
                samples.append(Sample(phi0[i], phi1, n_phase, shape, z0, zb, orients[i], tf[i]))
            t += HORIZON
            gap = every - HORIZON
            if gap > 0:
                if t + gap + HORIZON > steps:
                    break
                b.step(gap)
                t += gap
            if log:
                log(f"trajectories: step {t}, {len(samples)} samples")
    meta = {"grids": n_grids, "shape": list(shape), "steps": steps, "every": every, "n_phase": n_phase, "zb": zb,
            "t_scale": t_scale, "T_melt": p["T_melt"], "cooling": cooling}
    return samples, meta


# ---------------------------------------------------------- assisted stepping
class SurrogateStepper:
    """Surrogate-assisted stepping of a Solver (3D, FP32 state, column
    temperature).  A cycle proposes HORIZON steps, takes one exact step from
    the proposal P and accepts when that step's change agrees with the
    proposal's own mean rate:
        max |(step(P) - P) - (P - phi) / HORIZON| <= rate_tol * max |step(P) - P|
    (maxima over the band).  A proposal that leaves the state unchanged has
    a residual equal to the exact change and is rejected wherever the front
    moves."""

    def __init__(self, solver, model: SlotModel, n111, T_melt, t_scale, cooling, zb=32, rate_tol=0.5):
        torch = _torch()
        self.s, self.m = solver, model
        self.shape = solver.shape
        self.N = solver.n_phase
        nx, ny, nz = self.shape
        self.zb = min(zb, nz - nz % 8)
        self.orient = torch.tensor(orientation_features(n111), device="cuda")
        self.T_melt, self.t_scale, self.cooling = T_melt, t_scale, cooling
        ncell = nx * ny * nz
        self.phi = torch.empty((self.N, ncell), dtype=torch.float32, device="cuda")
        self.prop = torch.empty_like(self.phi)
        self.chk = torch.empty_like(self.phi)
        self.rate_tol = rate_tol
        self.accepted = self.rejected = self.overflow = 0
        self.steps = 0
        self.last_residual = self.last_change = float("nan")
        self.ratios = []                     # residual / exact change of every cycle

    def _download(self, dst):
        check(_lib.lib().mpf_download_phi_device(self.s._h, _p(dst), FP32))

    def _upload(self, src):
        check(_lib.lib().mpf_upload_phi_device(self.s._h, _p(src), FP32))

    def _band(self):
        f = self.s.observables()["front"]
        nz = self.shape[2]
        lo = int(np.nanmin(f)) if np.isfinite(f).any() else 0
        return max(0, min(nz - self.zb, lo - self.zb // 2))

    def _reduce(self, name, *args):
        torch = _torch()
        out = _zeros(1, dtype=torch.int32)
        _nnlib.call(name, *args, _p(out), _stream())
        return float(np.frombuffer(np.int32(out.item()).tobytes(), np.float32)[0])

    def propose(self, z0, model=True):
        """The proposal (dense, in self.prop) from the current state; with
        model False the state itself (the method without the network)."""
        torch = _torch()
        nx, ny, nz = self.shape
        self._download(self.phi)
        self.prop.copy_(self.phi)
        if not model:
            return 0
        T = self.s.temp()
        tf = torch.tensor(((T - self.T_melt) / self.t_scale).astype(np.float32), device="cuda")
        sid, feat, aux, over = self.m.features(self.phi, self.N, self.shape, z0, self.zb, self.orient, tf)
        out = self.m.forward(sid, feat, aux, self.shape, self.zb)
        _nnlib.call("mpf_nn_slot_apply", _p(self.prop), self.N, nx, ny, nz, z0, self.zb, self.m.k, _p(sid), _p(out),
                    32, 1.0 / SCALE, _stream())
        return int(over.item())

    def advance(self, model=True):
        """One cycle: HORIZON + 1 steps by an accepted proposal, or HORIZON
        exact steps after a rejected one.  Returns (steps, accepted)."""
        nx, ny, nz = self.shape
        z0 = self._band()
        T0 = self.s.temp()
        # cells with more candidate grains than slots keep the extra grains'
        # values; the residual check judges the proposal as a whole
        over = self.propose(z0, model)
        self.overflow += int(over > 0)
        T = T0.copy()
        for _ in range(HORIZON):          # the proposal's temperature: HORIZON cooling steps
            T -= self.cooling
        try:
            self._upload(self.prop)
        except _lib.CapacityError:          # more solids in a cell than the state's slots hold
            self._upload(self.phi)
            self.rejected += 1
            self.ratios.append(float("inf"))
            self.s.step(HORIZON)
            self.steps += HORIZON
            return HORIZON, False
        self.s.upload_temp(T)
        self.s.step(1)
        self._download(self.chk)
        r = self._reduce("mpf_nn_band_residual", _p(self.phi), _p(self.prop), _p(self.chk), self.N, nx, ny, nz,
                         z0, self.zb, 1.0 / HORIZON)
        c = self._reduce("mpf_nn_band_maxdiff", _p(self.prop), _p(self.chk), self.N, nx, ny, nz, z0, self.zb)
        self.last_residual, self.last_change = r, c
        self.ratios.append(r / c if c > 0 else float("inf"))
        if r <= self.rate_tol * c:
            self.accepted += 1
            self.steps += HORIZON + 1
            return HORIZON + 1, True
        self._upload(self.phi)             # restore the state and its temperature
        self.s.upload_temp(T0)
        self.rejected += 1
        self.s.step(HORIZON)
        self.steps += HORIZON
        return HORIZON, False
