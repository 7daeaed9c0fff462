"""bench.py's host-side summaries of rocprofv3 CSVs (no GPU): the per-step
traffic and VALU summaries keep only the kernels of the timed steps, scale the
HBM counters by what they read for the first device copy (the copy probe),
and average the last dispatches of each kernel."""
import csv
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _trace(path, kernels):
    """kernels: (name, start_ns, duration_ns) in dispatch order."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kind", "Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for i, (name, t0, d) in enumerate(kernels):
            w.writerow(["KERNEL_DISPATCH", i + 1, name, t0, t0 + d])


def _counters(path, rows):
    """rows: (dispatch_id, kernel, counter, value); two rows per dispatch so
    that the per-dispatch sum is exercised."""
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, k, c, v in rows:
            w.writerow([i, k, c, v / 2])
            w.writerow([i, k, c, v / 2])


@pytest.fixture
def run(tmp_path):
    """An autotuner trial of the side-stream mode (k_front1), then two timed
    steps in one-launch mode (k_interface); a 256 MiB copy probe first and a
    larger state copy of the autotuner later."""
    k = []
    t = 0
    seq = [("__amd_rocclr_copyBuffer", 100), ("k_tiles", 10), ("k_light", 300), ("k_merge", 400), ("k_front1", 70),
           ("__amd_rocclr_copyBuffer", 5000)]
    for _ in range(2):
        seq += [("k_tiles", 10), ("k_light", 280), ("k_merge", 410), ("k_interface", 180)]
    for name, d in seq:
        k.append((f"void mpf::impl_f32::{name}<3, 1, 0>(args)" if name.startswith("k_") else name, t, d * 1000))
        t += d * 1000 + 50
    tr = tmp_path / "kt.csv"
    _trace(tr, k)
    fetch, write = [], []
    for i, (name, _, _) in enumerate(k, start=1):
        if "copyBuffer" in name:
            big = i > 1
            fetch.append((i, name, "FETCH_SIZE", 131072.0 * (20 if big else 1)))
            write.append((i, name, "WRITE_SIZE", 262144.0 * (20 if big else 1)))
        else:
            fetch.append((i, name, "FETCH_SIZE", 1000.0))
            write.append((i, name, "WRITE_SIZE", 500.0))
    _counters(tmp_path / "f.csv", fetch)
    _counters(tmp_path / "w.csv", write)
    valu = []
    for i, (name, _, _) in enumerate(k, start=1):
        for c, v in (("SQ_INSTS_VALU", 6e6), ("SQ_INSTS_SALU", 1e6), ("SQ_INSTS_VMEM_RD", 2e5), ("SQ_INSTS_LDS", 1e5),
                     ("SQ_WAVES", 4096.0)):
            valu.append((i, name, c, v))
    _counters(tmp_path / "v.csv", valu)
    return tmp_path


def test_traffic_summary_timed_kernels_and_probe_scale(run):
    b = _bench()
    s = b.profile_summary("traffic", 2, str(run / "kt.csv"), [str(run / "f.csv"), str(run / "w.csv")])
    assert set(s["kernels"]) == {"k_tiles", "k_light", "k_merge", "k_interface"}   # no autotuner-only kernel
    assert s["fetch_scale"] == 2.0 and s["write_scale"] == 1.0                    # from the first copy, not the big one
    kl = s["kernels"]["k_light"]
    assert kl["read_bytes"] == 1000 * 1024 * 2 and kl["write_bytes"] == 500 * 1024
    assert kl["duration_ms"] == pytest.approx(0.28)                                # last 2 dispatches, not the trial's
    assert s["bytes_per_step"] == 4 * (1000 * 1024 * 2 + 500 * 1024)


def test_valu_summary(run):
    b = _bench()
    s = b.profile_summary("valu", 2, str(run / "kt.csv"), [str(run / "v.csv")])
    assert set(s["kernels"]) == {"k_tiles", "k_light", "k_merge", "k_interface"}
    m = s["kernels"]["k_merge"]
    t_us = 6e6 * 4 / (1024 * 2.4e3)                      # 4 cycles per instruction, 1024 SIMDs at 2.4 GHz
    assert m["valu_time_at_full_issue_us"] == pytest.approx(t_us, abs=0.05)
    assert m["valu_issue_frac"] == pytest.approx(t_us / 410.0, abs=1e-3)
    assert s["valu_wave_instr_per_step"] == 4 * 6e6
