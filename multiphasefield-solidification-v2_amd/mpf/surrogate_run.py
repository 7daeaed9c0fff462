"""Train the surrogate on the solver's own trajectories and measure
surrogate-assisted stepping against plain FP32 stepping.

    python -m mpf.surrogate_run --grids 48 --train-steps 2500 --out result/surrogate

Steps:
1. Trajectories: random-mode seed layers on 64^3 grids in the
   configuration's physics (config_3d.yaml), all grids of a batch stepped
   together; pairs (t, t + 16) every `every` steps (mpf.surrogate.generate_samples).
2. Training: Adam on the GPU (the HIP forward and backward kernels), a
   cosine learning-rate schedule, held-out grids for validation.
3. Skill on held-out pairs: 1 - |P - phi16| / |phi - phi16| over the
   predicted slots, with P the proposal and phi16 the exact state 16 steps
   on.  The method without the network (P = phi) scores 0.
4. The 128^3, 16-grain random-mode case (seed layer of 16 cells), 2000
   steps, three ways: plain FP32; surrogate-assisted; and the stepper with
   the do-nothing proposal (the method without the network).  Each reports
   wall time, front height, grain fractions and the acceptance rate.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import inputs
from . import surrogate as S
from .config import flat_params, load_config
from .solver import Solver


def skill(model, samples):
    """Mean skill over samples (see the module docstring)."""
    torch = S._torch()
    out = []
    for smp in samples:
        nx, ny, nz = smp.shape
        sid, feat, aux, over = model.features(smp.phi0, smp.n_phase, smp.shape, smp.z0, smp.zb, smp.orient, smp.tfeat)
        pred = model.forward(sid, feat, aux, smp.shape, smp.zb)
        prop = smp.phi0.clone()
        S._nnlib.call("mpf_nn_slot_apply", S._p(prop), smp.n_phase, nx, ny, nz, smp.z0, smp.zb, model.k, S._p(sid),
                      S._p(pred), 32, 1.0 / S.SCALE, S._stream())
        band = slice(None)
        p0 = smp.phi0.view(smp.n_phase, nx, ny, nz)[..., smp.z0:smp.z0 + smp.zb]
        p1 = smp.phi1.view(smp.n_phase, nx, ny, nz)[..., smp.z0:smp.z0 + smp.zb]
        pp = prop.view(smp.n_phase, nx, ny, nz)[..., smp.z0:smp.z0 + smp.zb]
        base = float(torch.linalg.vector_norm((p0 - p1)[band]).item())
        err = float(torch.linalg.vector_norm((pp - p1)[band]).item())
        if base > 0:
            out.append(1.0 - err / base)
    return float(np.mean(out)) if out else float("nan")


def unet_flops(model, nx, ny, zb):
    """Multiply-add FLOPs (2 per MAC) of one proposal: every convolution
    at its level's voxel count, for the U-Net and the per-slot layers."""
    f = 0
    for lvl, (ca, cb) in enumerate(model.unet.enc):
        nv = (nx >> lvl) * (ny >> lvl) * (zb >> lvl)
        f += sum(2 * c.ntap * c.cin * c.cout * nv for c in (ca, cb))
    for j, (ca, cb) in enumerate(model.unet.dec):
        lvl = 2 - j
        nv = (nx >> lvl) * (ny >> lvl) * (zb >> lvl)
        f += sum(2 * c.ntap * c.cin * c.cout * nv for c in (ca, cb))
    nv = nx * ny * zb * model.k
    f += sum(2 * c.ntap * c.cin * c.cout * nv for c in (model.enc, model.h1, model.h2))
    return f


def proposal_rate(model, nx=128, ny=128, zb=32, n_phase=17, reps=10):
    """Time the network's forward pass of one proposal over an nx x ny x zb
    band with HIP events on the stream its kernels run on; returns ms and
    TFLOP/s (bf16 MFMA, fp32 accumulate)."""
    torch = S._torch()
    rng = np.random.default_rng(0)
    k = model.k
    nv = nx * ny * zb
    sid = torch.tensor(np.where(rng.random((k, nv)) < 0.5, 1, -1).astype(np.int32), device="cuda")
    feat = torch.tensor(rng.random((k, nv, 32)).astype(np.float32), device="cuda").to(torch.bfloat16)
    aux = torch.tensor(rng.random((nv, 32)).astype(np.float32), device="cuda").to(torch.bfloat16)
    model.forward(sid, feat, aux, (nx, ny, zb), zb)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        model.forward(sid, feat, aux, (nx, ny, zb), zb)
    e1.record(st)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    fl = unet_flops(model, nx, ny, zb)
    return {"band": [nx, ny, zb], "ms": ms, "gflop": fl / 1e9, "tflops": fl / (ms * 1e-3) / 1e12}


def c2_case(n=128, n_solid=16, seed_height=16, seed=42):
    """The 128^3, 16-grain random-mode case in the configuration's physics."""
    p = flat_params(load_config())
    p.update(nx=n, ny=n, nz=n)
    d = inputs.derive_constants(p)
    N = n_solid + 1
    q = inputs.assign_quaternions_to_grains(n_solid, "random", seed)
    w, a, m = inputs.build_interaction_matrices(N, d["eps0_sl"], d["w0_sl"], d["m_sl_phi"], d["eps_gb"], d["w_gb"],
                                                d["m_gb_phi"])
    n111 = inputs.compute_rotated_n111(q)
    gmap = inputs.generate_random_grain_map_3d(n, n, n_solid, random_seed=seed)
    cols = inputs.columns_from_grain_map_3d(gmap, n, p["dz"], d["delta"], seed_height)
    temp = inputs.init_temperature_field_3d(1, 1, n, p["T_melt"], p["G"], p["dz"], seed_height)[0, 0]
    cooling = float(np.float32(p["G"] * p["V_pulling"] * p["dt"]))

    def make():
        s = Solver((n, n, n), N, inputs.kernel_scalars(p, d, True), precision="fp32", cooling_rate=cooling)
        s.set_tables(w, a, m, n111)
        s.init_columns(*cols)
        s.upload_temp(temp.astype(np.float64))
        return s

    return make, n111, p, cooling


def _summary(s, dx):
    ob = s.observables(dx=dx)
    gv = ob["grain_volumes"]
    return {"front_mean": ob["front_mean"], "front": ob["front"], "fractions": gv / max(gv.sum(), 1e-300)}


def compare_c2(model, steps=2000, zb=32, rate_tol=(0.5, 1.0, 2.0), t_scale=0.5, log=print):
    torch = S._torch()
    make, n111, p, cooling = c2_case()
    res = {}
    # plain
    with make() as s:
        s.step(S.HORIZON)                          # warm the kernels outside the timing
        torch.cuda.synchronize()
    with make() as s:
        t0 = time.perf_counter()
        s.step(steps)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
        ref = _summary(s, p["dx"])
    res["plain"] = {"wall_s": t_plain, "steps": steps, "front_mean": ref["front_mean"]}
    log(f"plain FP32: {steps} steps in {t_plain:.3f} s, front {ref['front_mean']:.3f}")
    runs = [(f"surrogate_tol{t:g}", True, t) for t in rate_tol] + [("no_network", False, rate_tol[0])]
    for name, use_model, tol in runs:
        with make() as s:
            st = S.SurrogateStepper(s, model, n111, p["T_melt"], t_scale, cooling, zb=zb, rate_tol=tol)
            t0 = time.perf_counter()
            while st.steps + S.HORIZON + 1 <= steps:
                st.advance(model=use_model)
            if st.steps < steps:
                s.step(steps - st.steps)
                st.steps = steps
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            got = _summary(s, p["dx"])
        fe = np.abs(got["front"] - ref["front"])
        fe = fe[np.isfinite(fe)]
        r = {"wall_s": t, "steps": st.steps, "speedup_vs_plain": t_plain / t,
             "accepted": st.accepted, "rejected": st.rejected, "overflow": st.overflow,
             "acceptance": st.accepted / max(st.accepted + st.rejected, 1),
             "front_mean": got["front_mean"], "front_mean_err_cells": abs(got["front_mean"] - ref["front_mean"]),
             "front_max_err_cells": float(fe.max()) if fe.size else float("nan"),
             "grain_fraction_max_err": float(np.max(np.abs(got["fractions"] - ref["fractions"]))),
             "rate_tol": tol, "residual_ratio_p50": float(np.median(st.ratios)) if st.ratios else float("nan"),
             "residual_ratio_min": float(np.min(st.ratios)) if st.ratios else float("nan")}
        res[name] = r
        log(f"{name}: {r}")
    return res


def run(grids=48, traj_steps=2000, every=50, train_steps=2500, lr=1e-3, eval_steps=2000, rate_tol=(0.5, 1.0, 2.0),
        seed=0, log=None):
    """Trajectories, training, held-out skill, the network's proposal rate
    and the 128^3 comparison; returns the summary dict and the model."""
    torch = S._torch()
    log = log or (lambda m: None)
    t0 = time.perf_counter()
    samples, meta = S.generate_samples(n_grids=grids, steps=traj_steps, every=every, seed=seed)
    t_data = time.perf_counter() - t0
    n_hold = max(1, len(samples) // 8)
    order = np.random.default_rng(seed).permutation(len(samples))
    held = [samples[i] for i in order[:n_hold]]
    train_set = [samples[i] for i in order[n_hold:]]
    log(f"{len(samples)} pairs from {grids} grids in {t_data:.1f} s; {len(held)} held out")
    model = S.SlotModel(seed=seed)
    sk0 = skill(model, held)
    t1 = time.perf_counter()
    hist = S.train(model, train_set, train_steps, lr=lr, log=log, seed=seed)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    sk1 = skill(model, held)
    log(f"trained {train_steps} steps in {t_train:.1f} s; held-out skill {sk0:.3f} -> {sk1:.3f}")
    del samples, held, train_set
    rate = proposal_rate(model)
    log(f"proposal forward: {rate}")
    res = compare_c2(model, steps=eval_steps, rate_tol=tuple(rate_tol), t_scale=meta["t_scale"], log=log)
    summary = {"pairs": int(len(order)), "grids": grids, "data_s": t_data, "train_steps": train_steps,
               "train_s": t_train, "train_ms_per_step": 1e3 * t_train / max(train_steps, 1),
               "loss_first": float(np.mean(hist[:50])), "loss_last": float(np.mean(hist[-50:])),
               "skill_untrained": sk0, "skill_trained": sk1, "params": model.n_params(), "proposal": rate,
               "c2": res, "meta": dict(meta)}
    return summary, model


def main(argv=None):
    ap = argparse.ArgumentParser(description="Train the surrogate and measure assisted stepping.")
    ap.add_argument("--grids", type=int, default=48)
    ap.add_argument("--traj-steps", type=int, default=2000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--train-steps", type=int, default=2500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-steps", type=int, default=2000)
    ap.add_argument("--rate-tol", type=float, nargs="+", default=[0.5, 1.0, 2.0])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None, help="write weights.npz and summary.json here")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)   # noqa: E731
    summary, model = run(args.grids, args.traj_steps, args.every, args.train_steps, args.lr, args.eval_steps,
                         args.rate_tol, args.seed, log)
    if args.out:
        os.makedirs(args.out, exist_ok=True)
        np.savez(os.path.join(args.out, "weights.npz"), **model.state_dict())
        with open(os.path.join(args.out, "summary.json"), "w") as f:
            json.dump(summary, f, indent=1, default=float)
    print(json.dumps(summary, default=float), flush=True)
    return summary


if __name__ == "__main__":
    main(sys.argv[1:])
