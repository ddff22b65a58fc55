"""Benchmark of the zldsp::fft hot path on one MI355X.

Workload (one "step"): one forward pass of 16384 complex FP32 CFFTs of
N = 4096 points, AoS in and AoS out (1 GiB of algorithmic HBM traffic:
16 B per point, one read of the input and one write of the output), with the
inputs resident in HBM.  The step is one launch through the C-ABI
(zlfft_cfft_exec) on torch's current stream: the persistent streaming
Stockham kernel for N = 4096.

Printed: ONE JSON line with whole-job throughput in GB/s of algorithmic
traffic over exactly --steps back-to-back steps (one event pair around the
timed region), the kernel's roofline fraction against the 8 TB/s HBM peak
(average launch duration from HIP events bracketing every launch of a second
window right after it),
the accuracy of the timed output against the FP64 oracle, and a CPU baseline:
the oracle's FP32 restatement of the reference algorithm, timed on this host.

Timing rules, for every line:
* warm-up by time: back-to-back launches for at least WARM_S seconds of GPU
  work (and at least --warmup launches for the headline), with no
  synchronisation between the warm-up and the timed launches, so every line
  starts at the same clock state;
* the MFMA variants and the plain FP32 kernel are timed interleaved (plain,
  variant 1, ..., variant 5, plain, ...) in short windows with an event pair
  around every launch, at least 100 launches each; their medians are
  reported, so the comparison does not depend on the order of measurement;
* the secondary lines call the C-ABI directly (no Python wrapper per call),
  rotate over buffer sets larger than the 256 MB MALL, and report both the
  wall rate of back-to-back launches and the mean of per-launch event times.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ORDER, BATCH = 12, 16384
N = 1 << ORDER
BYTES_PER_STEP = 16 * N * BATCH  # read N complex64 + write N complex64, per transform
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cfft4096_b16384_pmc.md")
WARM_S = 0.3          # seconds of back-to-back launches before a line is timed
MIN_TIMED = 100       # launches per secondary line / MFMA variant, whatever --steps is
ROTATE = 5            # buffer sets per secondary line: > 256 MiB, so the MALL cannot hold the working set


def parse():
    global WARM_S, MIN_TIMED
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="CPU baseline sampling budget (T = all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mfma", action="store_true", help="skip the MFMA split-precision variants")
    ap.add_argument("--no-configs", action="store_true", help="skip the analyzer and large-order lines")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged-batch line")
    ap.add_argument("--only", choices=("main", "mfma", "configs", "ragged", "large", "analyzer", "orders"),
                    default=None,
                    help="run one section alone (for profiler runs): the headline batch, the MFMA variants, "
                         "the secondary lines, the ragged batch, the large-order lines, the analyzer lines, or "
                         "the plain batched path by size (orders: only this way, not in the default run)")
    ap.add_argument("--mfma-precisions", default=None,
                    help="comma-separated subset of the MFMA variant names")
    ap.add_argument("--no-copy-ref", action="store_true",
                    help="skip the plain-copy reference rates beside the lines")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 512 MiB device copy (reference kernel for profiler byte counters)")
    ap.add_argument("--warm-seconds", type=float, default=WARM_S,
                    help="time-based warm-up per line (profiler runs use 0)")
    ap.add_argument("--min-timed", type=int, default=MIN_TIMED,
                    help="timed launches per secondary line and MFMA variant (profiler runs use a few)")
    args = ap.parse_args()
    WARM_S, MIN_TIMED = args.warm_seconds, args.min_timed
    return args


# ---------------------------------------------------------------- timing helpers

def warm_up(torch, fns, seconds=None, min_calls=0):
    """Back-to-back calls of fns (in turn) for about `seconds` of GPU work and
    at least min_calls calls.  Ends without a synchronisation, so the timed
    launches queued next follow the warm-up with no idle gap."""
    seconds = WARM_S if seconds is None else min(seconds, WARM_S)
    probe = max(3, len(fns))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(probe):
        fns[i % len(fns)]()
    torch.cuda.synchronize()
    per = max((time.perf_counter() - t0) / probe, 1e-6)
    calls = max(min_calls, min(50000, int(math.ceil(seconds / per))))
    for i in range(calls):
        fns[i % len(fns)]()


def time_window(torch, stream, fns, reps):
    """ms per call over `reps` back-to-back calls (fns in turn), one event pair
    around the window on `stream`."""
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for i in range(reps):
        fns[i % len(fns)]()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def time_each(torch, stream, fns, reps):
    """Per-call ms of `reps` calls (fns in turn), an event pair around each."""
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for i in range(reps):
        ev[i][0].record(stream)
        fns[i % len(fns)]()
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in ev]


def _line(nbytes, flops, count, ms, rel=None, kern_ms=None):
    gbs = nbytes / (ms * 1e-3) / 1e9
    d = {"ms": round(ms, 5), "bytes": int(nbytes), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
         "transforms_per_s": round(count / (ms * 1e-3), 1), "gflops": round(flops / (ms * 1e-3) / 1e9, 1)}
    if kern_ms is not None:
        kg = nbytes / (kern_ms * 1e-3) / 1e9
        d["kernel_ms_mean"] = round(kern_ms, 5)
        d["kernel_GB/s"] = round(kg, 1)
        d["kernel_frac_hbm"] = round(kg / HBM_PEAK_GBS, 4)
    if rel is not None:
        d["rel_l2_vs_fp64"] = rel
    return d


def measure_line(torch, stream, fns, reps=None):
    """Warm-up by time, then (wall ms per call of back-to-back launches,
    mean of per-launch event ms)."""
    reps = MIN_TIMED if reps is None else min(reps, MIN_TIMED)
    warm_up(torch, fns)
    wall = time_window(torch, stream, fns, reps)
    warm_up(torch, fns, seconds=0.05)
    each = time_each(torch, stream, fns, reps)
    return wall, float(np.mean(each))


def _call(lib, fn, *args):
    rc = fn(*args)
    if rc:
        raise RuntimeError(lib.zlfft_last_error().decode())


# ---------------------------------------------------------------- CPU baseline

def host_cpus() -> dict:
    """CPU model, the CPUs this process may run on, and the cgroup CPU quota."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    usable = affinity if quota is None else max(1, min(affinity, int(math.floor(quota + 1e-9))))
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity": affinity, "cgroup_quota": quota,
            "usable": usable}


def cpu_baseline(seconds: float) -> dict:
    """The oracle's FP32 restatement of the reference CFFT<float> (a scalar C
    port: the reference itself needs Google Highway, which is not installed)
    on this host: a contiguous slice of the batch per thread, one workspace per
    thread (reference instances are not shareable,
    src/common/zldsp_fft_common_init.hpp:24-43).  T = every usable host CPU,
    and T = 1."""
    import oracle
    cpus = host_cpus()

    def run(threads, sample, secs):
        x = np.random.default_rng(0).standard_normal((sample, 2 * N)).astype(np.float32)
        y = np.empty_like(x)
        oracle.cfft_batch_raw(ORDER, x, y, True, "f32", threads)  # warm-up
        times = []
        t_end = time.perf_counter() + secs
        while time.perf_counter() < t_end or len(times) < 3:
            t0 = time.perf_counter()
            oracle.cfft_batch_raw(ORDER, x, y, True, "f32", threads)
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        return t, len(times)

    t_all = cpus["usable"]
    sample = max(256, 128 * t_all)
    t, nrun = run(t_all, sample, seconds)
    t1, nrun1 = run(1, 128, max(1.0, seconds / 2))
    l1, sw = oracle.switch_order("f32")
    path = "stockham" if ORDER < sw else "dif-macro"
    return {
        "value": round(16 * N * sample / t / 1e9, 3),
        "unit": "GB/s",
        "cores": t_all,
        "kind": "port",
        "sample": f"{sample} x 4096-point CFFT<float> forward, AoS, median of {nrun} runs on {t_all} threads; "
                  f"scalar C restatement of the reference ({path} plan, switch order {sw}), -O3 -march=x86-64-v3",
        "transforms_per_s": round(sample / t, 1),
        "gflops_5nlog2n": round(5 * N * ORDER * sample / t / 1e9, 2),
        "t1": {"value": round(16 * N * 128 / t1 / 1e9, 3), "unit": "GB/s", "cores": 1,
               "sample": f"128 transforms, median of {nrun1} runs", "transforms_per_s": round(128 / t1, 1),
               "gflops_5nlog2n": round(5 * N * ORDER * 128 / t1 / 1e9, 2)},
        "cpu_model": cpus["model"],
        "host_cpus": cpus,
    }


# ---------------------------------------------------------------- MFMA variants

MFMA_VARIANTS = (("f16_split", 2), ("bf16_split3", 5), ("f16", 1), ("bf16", 3), ("bf16_split", 4))


def mfma_variants(x, y, stream, lib, args, plain_step) -> dict:
    """The headline workload on the MFMA split-precision kernels (each radix-16
    pass as a 16x16 DFT-matrix product on FP16 / BF16 matrix cores with FP32
    accumulate), timed interleaved with the plain FP32 kernel: rounds of
    (plain, variant 1, ..., variant 5) windows with an event pair around every
    launch, medians over at least MIN_TIMED launches each.  Accuracy vs the
    FP64 oracle on two input sets: A, i.i.d. N(0, 1) (the timed batch, 16
    sampled transforms), and B, the dynamic-range set (unit sinusoid, -60 dB
    sinusoid, -100 dB white noise), 64 transforms through the same kernels."""
    import torch
    import oracle
    import zldsp_fft_amd as zf
    from zldsp_fft_amd import signals
    wanted = args.mfma_precisions.split(",") if args.mfma_precisions else None
    variants = [(name, prec) for name, prec in MFMA_VARIANTS if wanted is None or name in wanted]
    sp = stream.cuda_stream
    plans, steps = {}, {"plain": plain_step}
    for name, prec in variants:
        plans[name] = zf.CFFT(ORDER, prec)
        h = plans[name]._h

        def step(h=h):
            _call(lib, lib.zlfft_cfft_exec, h, zf.FORWARD, zf.AOS, x.data_ptr(), None, zf.AOS, y.data_ptr(), None,
                  BATCH, sp)
        steps[name] = step
    order = ["plain"] + [n for n, _ in variants]
    window = 10
    rounds = max(1, -(-MIN_TIMED // window))
    for name in order:
        warm_up(torch, [steps[name]], seconds=0.1)
    warm_up(torch, [steps[n] for n in order])
    times = {n: [] for n in order}
    for _ in range(rounds):
        for name in order:
            times[name] += time_each(torch, stream, [steps[name]], window)
    # accuracy: set A from the timed batch, set B from a separate batch
    idx = list(range(0, BATCH, BATCH // 16))
    xa = x[idx].cpu().numpy().astype(np.complex128)
    ref_a = np.stack([oracle.cfft(r, True, "f64") for r in xa])
    nb = 64
    xb = signals.dynamic_range_set(5, nb, N).astype(np.complex64)
    ref_b = np.stack([oracle.cfft(r, True, "f64") for r in xb.astype(np.complex128)])
    xb_d = torch.from_numpy(xb).to(x.device)
    yb_d = torch.empty_like(xb_d)

    def errs(got, ref):
        rel = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
        mbe = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
        return float(rel.max()), float(rel.mean()), float(mbe.max())

    res = {}
    plain_ms = float(np.median(times["plain"]))
    for name in order:
        ms = float(np.median(times[name]))
        gbs = BYTES_PER_STEP / (ms * 1e-3) / 1e9
        d = {"GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4), "ms_median": round(ms, 5),
             "ms_mean": round(float(np.mean(times[name])), 5), "launches": len(times[name])}
        if name != "plain":
            steps[name]()
            torch.cuda.synchronize()
            ra, rma, ma = errs(y[idx].cpu().numpy().astype(np.complex128), ref_a)
            _call(lib, lib.zlfft_cfft_exec, plans[name]._h, zf.FORWARD, zf.AOS, xb_d.data_ptr(), None, zf.AOS,
                  yb_d.data_ptr(), None, nb, sp)
            torch.cuda.synchronize()
            rb, rmb, mb = errs(yb_d.cpu().numpy().astype(np.complex128), ref_b)
            d.update({"rel_l2_max": ra, "rel_l2_mean": rma, "max_bin_err": ma,
                      "setB_rel_l2_max": rb, "setB_rel_l2_mean": rmb, "setB_max_bin_err": mb,
                      "meets_1e-6": max(ra, rb) <= 1e-6, "vs_plain": round(plain_ms / ms, 4),
                      "beats_plain": ms < plain_ms})
            plans[name].close()
        res["plain_f32" if name == "plain" else name] = d
    res["timing"] = (f"interleaved: {rounds} rounds x {window} launches per kernel, event pair per launch, "
                     f"median; warm-up {WARM_S} s")
    return res


# ---------------------------------------------------------------- secondary lines

def _ragged_line(orders) -> dict:
    """10,000 transforms with orders uniform in `orders`, CFFT / RFFT,
    forward / backward, AoS / SoA (50/50 each), in one graph-launched
    persistent kernel.  Throughput in algorithmic bytes (one read of each
    input, one write of each output) per graph launch."""
    import torch
    import zldsp_fft_amd as zf
    from zldsp_fft_amd import workloads
    descs = workloads.ragged_mix(10000, 6, orders)
    nbytes = sum(workloads.algorithmic_bytes(d) for d in descs)
    rb = zf.RaggedBatch(descs)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    lib = zf.load_library()
    h = rb._h

    def run():
        _call(lib, lib.zlfft_ragged_exec, h, sp)
    wall, kern = measure_line(torch, stream, [run])
    rb.close()
    gbs = nbytes / (wall * 1e-3) / 1e9
    kg = nbytes / (kern * 1e-3) / 1e9
    return {"workload": f"ragged_10000_orders{orders[0]}-{orders[1]}_mixed", "bytes_per_launch": nbytes,
            "ms_per_launch": round(wall, 5), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
            "kernel_ms_mean": round(kern, 5), "kernel_GB/s": round(kg, 1),
            "transforms_per_s": round(len(descs) / (wall * 1e-3), 1),
            "check": "every output vs the per-transform plain path <= 1e-6 rel-L2 in tests/test_gpu_ragged.py"}


def ragged_section(args) -> dict:
    """The ragged batch of orders 6..14 (512-thread kernel); beside it the
    same mix without order 14, which runs in the 256-thread kernel."""
    out = _ragged_line((6, 14))
    out["without_order_14"] = _ragged_line((6, 13))
    return out


def analyzer_lines(args) -> dict:
    """The analyzer batch (4096 RFFTs of 2048 Hann-windowed reals: AoS bins and
    squared magnitudes), rocFFT's rfft beside it, and the STFT front-end."""
    import torch
    import oracle
    import zldsp_fft_amd as zf
    from zldsp_fft_amd import signals
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    lib = zf.load_library()
    out = {}
    n, frames = 2048, 4096
    nb = n // 2 + 1
    plan = zf.RFFT(11)
    h = plan._h
    x0 = torch.from_numpy(signals.analyzer_frames(2, frames, n).astype(np.float32)).to(dev)
    xs = [x0] + [x0.clone() for _ in range(ROTATE - 1)]  # distinct addresses, same frames
    ys = [torch.empty(frames * nb, dtype=torch.complex64, device=dev) for _ in range(ROTATE)]
    ps = [torch.empty(frames * nb, dtype=torch.float32, device=dev) for _ in range(ROTATE)]
    flops = 2.5 * n * 11 * frames
    fwd = [lambda i=i: _call(lib, lib.zlfft_rfft_forward_exec, h, xs[i].data_ptr(), zf.AOS, ys[i].data_ptr(), None,
                             frames, sp) for i in range(ROTATE)]
    wall, kern = measure_line(torch, stream, fwd)
    ref = np.stack([oracle.rfft_forward(xs[0][f].cpu().numpy().astype(np.float64), "f64") for f in (7, 2048, 4095)])
    got = ys[0].view(frames, nb)[[7, 2048, 4095]].cpu().numpy()
    rel = float((np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)).max())
    out["rfft2048_b4096_aos"] = _line(frames * (4 * n + 8 * nb), flops, frames, wall, rel, kern)
    sq = [lambda i=i: _call(lib, lib.zlfft_rfft_forward_sqr_mag_exec, h, xs[i].data_ptr(), ps[i].data_ptr(), frames,
                            sp) for i in range(ROTATE)]
    wall, kern = measure_line(torch, stream, sq)
    got = ps[0].view(frames, nb)[[7, 2048, 4095]].cpu().numpy()
    rsq = np.abs(ref) ** 2
    rel = float((np.linalg.norm(got - rsq, axis=1) / np.linalg.norm(rsq, axis=1)).max())
    out["rfft2048_b4096_sqr_mag"] = _line(frames * (4 * n + 4 * nb), flops, frames, wall, rel, kern)
    # RFFT<float>::backward over the batch: the bins just computed back to
    # frames (N/2 x the frames), AoS bins in
    rs = [torch.empty(frames * n, dtype=torch.float32, device=dev) for _ in range(ROTATE)]
    bw = [lambda i=i: _call(lib, lib.zlfft_rfft_backward_exec, h, zf.AOS, ys[i].data_ptr(), None, rs[i].data_ptr(),
                            frames, sp) for i in range(ROTATE)]
    for f in fwd:
        f()
    wall, kern = measure_line(torch, stream, bw)
    binsd = ys[0].view(frames, nb)[[7, 2048, 4095]].cpu().numpy().astype(np.complex128)
    refb = np.stack([oracle.rfft_backward(b, n, "f64") for b in binsd])
    gotb = rs[0].view(frames, n)[[7, 2048, 4095]].cpu().numpy().astype(np.float64)
    relb = float((np.linalg.norm(gotb - refb, axis=1) / np.linalg.norm(refb, axis=1)).max())
    out["rfft2048_b4096_backward"] = _line(frames * (8 * nb + 4 * n), flops, frames, wall, relb, kern)
    wall, kern = measure_line(torch, stream, [lambda i=i: torch.fft.rfft(xs[i], dim=-1) for i in range(ROTATE)])
    out["rocfft_rfft2048_b4096_aos"] = _line(frames * (4 * n + 8 * nb), flops, frames, wall, None, kern)
    # STFT front-end: 4096 Hann-windowed frames of 2048 samples at hop 512 of
    # one signal, dB levels (frames cut and windowed at load time); bytes count
    # the frame samples read and the levels written
    hop = 512
    sig = [torch.rand((frames - 1) * hop + n, device=dev) * 2 - 1 for _ in range(ROTATE)]
    win = torch.from_numpy(signals.hann(n).astype(np.float32)).to(dev)
    lv = [torch.empty(frames * nb, dtype=torch.float32, device=dev) for _ in range(ROTATE)]
    st = [lambda i=i: _call(lib, lib.zlfft_rfft_stft_exec, h, sig[i].data_ptr(), hop, win.data_ptr(), frames,
                            zf.STFT_DB, zf.AOS, lv[i].data_ptr(), None, 1e-12, sp) for i in range(ROTATE)]
    wall, kern = measure_line(torch, stream, st)
    out["stft2048_hop512_b4096_hann_db"] = _line(frames * (4 * n + 4 * nb), flops, frames, wall, None, kern)
    plan.close()
    return out


def hbm_passes(order: int) -> int:
    """HBM passes of the FP32 CFFT of an order above 14 (2 up to order 20,
    then 3): the order's multi-pass plan in zldsp_fft_amd/csrc/zlfft_registry.hpp."""
    return 2 if order <= 20 else 3


def large_lines(args) -> dict:
    """The large SoA CFFTs: orders 16/18/20/22 at batch 64/16/4/1 (4,194,304
    points per call), forward and backward, through the C-ABI; then the same
    shapes as interleaved complex64, forward, through the C-ABI and through
    rocFFT (torch.fft) for comparison."""
    import torch
    import oracle
    import zldsp_fft_amd as zf
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    lib = zf.load_library()
    out = {}
    for order, batch in ((16, 64), (18, 16), (20, 4), (22, 1)):
        n = 1 << order
        plan = zf.CFFT(order)
        h = plan._h
        g = torch.Generator(device=dev).manual_seed(3)
        ins = [(torch.randn(batch * n, device=dev, generator=g), torch.randn(batch * n, device=dev, generator=g))
               for _ in range(ROTATE)]
        outs = [(torch.empty(batch * n, device=dev), torch.empty(batch * n, device=dev)) for _ in range(ROTATE)]
        for name, d in (("forward", zf.FORWARD), ("backward", zf.BACKWARD)):
            fns = [lambda i=i, d=d: _call(lib, lib.zlfft_cfft_exec, h, d, zf.SOA, ins[i][0].data_ptr(),
                                          ins[i][1].data_ptr(), zf.SOA, outs[i][0].data_ptr(), outs[i][1].data_ptr(),
                                          batch, sp) for i in range(ROTATE)]
            wall, kern = measure_line(torch, stream, fns)
            x = ins[0][0][:n].cpu().numpy().astype(np.float64) + 1j * ins[0][1][:n].cpu().numpy()
            ref = oracle.cfft(x, name == "forward", "f64")
            got = outs[0][0][:n].cpu().numpy() + 1j * outs[0][1][:n].cpu().numpy().astype(np.float64)
            out[f"cfft_o{order}_b{batch}_soa_{name}"] = _line(16 * n * batch, 5.0 * n * order * batch, batch, wall,
                                                              float(np.linalg.norm(got - ref) / np.linalg.norm(ref)),
                                                              kern)
            out[f"cfft_o{order}_b{batch}_soa_{name}"]["hbm_passes"] = hbm_passes(order)
        # the same shape as interleaved complex64, here and through rocFFT
        # (torch.fft, which takes no split layout): an external reference point
        xa = [torch.complex(a, b).view(batch, n) for a, b in ins]
        ya = [torch.empty_like(t) for t in xa]
        fns = [lambda i=i: _call(lib, lib.zlfft_cfft_exec, h, zf.FORWARD, zf.AOS, xa[i].data_ptr(), None, zf.AOS,
                                 ya[i].data_ptr(), None, batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        out[f"cfft_o{order}_b{batch}_aos_forward"] = _line(16 * n * batch, 5.0 * n * order * batch, batch, wall, None,
                                                           kern)
        out[f"cfft_o{order}_b{batch}_aos_forward"]["hbm_passes"] = hbm_passes(order)
        fns = [lambda i=i: torch.fft.fft(xa[i], dim=-1) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        out[f"rocfft_o{order}_b{batch}_aos_forward"] = _line(16 * n * batch, 5.0 * n * order * batch, batch, wall,
                                                             None, kern)
        plan.close()
        del ins, outs, xa, ya
        torch.cuda.empty_cache()
    out.update(large_rfft_lines(torch, lib, stream))
    return out


def large_rfft_lines(torch, lib, stream) -> dict:
    """Large RFFTs, 8,388,608 reals per call: orders 17 x 64 (inner CFFT of 2
    HBM passes) and 23 x 1 (3 passes), forward to AoS bins, squared
    magnitudes and backward, through the C-ABI, with rocFFT's rfft
    (torch.fft) on the same shapes.  The post-twiddle runs in the last HBM
    pass and the pre-twiddle in the first."""
    import oracle
    import zldsp_fft_amd as zf
    dev = torch.device("cuda:0")
    sp = stream.cuda_stream
    out = {}
    for order, batch in ((17, 64), (23, 1)):
        rn, nb = 1 << order, (1 << order) // 2 + 1
        plan = zf.RFFT(order)
        h = plan._h
        g = torch.Generator(device=dev).manual_seed(5)
        xs = [torch.randn(batch * rn, device=dev, generator=g) for _ in range(ROTATE)]
        ys = [torch.empty(batch * nb, dtype=torch.complex64, device=dev) for _ in range(ROTATE)]
        ps = [torch.empty(batch * nb, device=dev) for _ in range(ROTATE)]
        rs = [torch.empty(batch * rn, device=dev) for _ in range(ROTATE)]
        flops = 2.5 * rn * order * batch
        passes = hbm_passes(order - 1)
        fwd = [lambda i=i: _call(lib, lib.zlfft_rfft_forward_exec, h, xs[i].data_ptr(), zf.AOS, ys[i].data_ptr(),
                                 None, batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fwd)
        ref = oracle.rfft_forward(xs[0][:rn].cpu().numpy().astype(np.float64), "f64")
        got = ys[0][:nb].cpu().numpy().astype(np.complex128)
        key = f"rfft_o{order}_b{batch}"
        out[key + "_bins"] = _line((4 * rn + 8 * nb) * batch, flops, batch, wall,
                                   float(np.linalg.norm(got - ref) / np.linalg.norm(ref)), kern)
        sq = [lambda i=i: _call(lib, lib.zlfft_rfft_forward_sqr_mag_exec, h, xs[i].data_ptr(), ps[i].data_ptr(),
                                batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, sq)
        out[key + "_sqr_mag"] = _line((4 * rn + 4 * nb) * batch, flops, batch, wall, None, kern)
        for f in fwd:
            f()
        bw = [lambda i=i: _call(lib, lib.zlfft_rfft_backward_exec, h, zf.AOS, ys[i].data_ptr(), None,
                                rs[i].data_ptr(), batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, bw)
        x0 = xs[0][:rn].cpu().numpy().astype(np.float64)
        r0 = rs[0][:rn].cpu().numpy().astype(np.float64)
        out[key + "_backward"] = _line((8 * nb + 4 * rn) * batch, flops, batch, wall,
                                       float(np.linalg.norm(r0 - (rn / 2) * x0) / np.linalg.norm((rn / 2) * x0)), kern)
        for k in ("_bins", "_sqr_mag", "_backward"):
            out[key + k]["hbm_passes"] = passes
        xv = [t.view(batch, rn) for t in xs]
        fns = [lambda i=i: torch.fft.rfft(xv[i], dim=-1) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        out[f"rocfft_rfft_o{order}_b{batch}"] = _line((4 * rn + 8 * nb) * batch, flops, batch, wall, None, kern)
        plan.close()
        del xs, ys, ps, rs, xv
        torch.cuda.empty_cache()
    return out


def order_lines(args) -> dict:
    """The plain batched path by size (not part of the default run): CFFT
    orders 3-14 (AoS forward) and RFFT orders 4-15 (AoS bins, squared
    magnitudes, and backward from AoS bins), 2^26 points per call, through the C-ABI over ROTATE buffer
    sets; the persistent streaming kernel runs wherever the order has one."""
    import torch
    import zldsp_fft_amd as zf
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    lib = zf.load_library()
    out = {}
    pts = 1 << 26
    for order in range(3, 15):
        n = 1 << order
        batch = pts // n
        plan = zf.CFFT(order)
        h = plan._h
        ins = [torch.rand((batch, n), dtype=torch.complex64, device=dev) for _ in range(ROTATE)]
        outs = [torch.empty_like(t) for t in ins]
        fns = [lambda i=i: _call(lib, lib.zlfft_cfft_exec, h, zf.FORWARD, zf.AOS, ins[i].data_ptr(), None, zf.AOS,
                                 outs[i].data_ptr(), None, batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        out[f"cfft_o{order}_aos"] = _line(16 * n * batch, 5.0 * n * order * batch, batch, wall, None, kern)
        plan.close()
        del ins, outs
        rn, nb = 2 * n, n + 1  # RFFT of real order order + 1
        plan = zf.RFFT(order + 1)
        h = plan._h
        ins = [torch.rand((batch, rn), device=dev) for _ in range(ROTATE)]
        bins = [torch.empty((batch, nb), dtype=torch.complex64, device=dev) for _ in range(ROTATE)]
        sq = [torch.empty((batch, nb), device=dev) for _ in range(ROTATE)]
        fns = [lambda i=i: _call(lib, lib.zlfft_rfft_forward_exec, h, ins[i].data_ptr(), zf.AOS, bins[i].data_ptr(),
                                 None, batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        flops = 2.5 * rn * (order + 1) * batch
        out[f"rfft_o{order + 1}_bins"] = _line((4 * rn + 8 * nb) * batch, flops, batch, wall, None, kern)
        fns = [lambda i=i: _call(lib, lib.zlfft_rfft_forward_sqr_mag_exec, h, ins[i].data_ptr(), sq[i].data_ptr(),
                                 batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        out[f"rfft_o{order + 1}_sqr_mag"] = _line((4 * rn + 4 * nb) * batch, flops, batch, wall, None, kern)
        fns = [lambda i=i: _call(lib, lib.zlfft_rfft_backward_exec, h, zf.AOS, bins[i].data_ptr(), None, ins[i].data_ptr(),
                                 batch, sp) for i in range(ROTATE)]
        wall, kern = measure_line(torch, stream, fns)
        out[f"rfft_o{order + 1}_backward"] = _line((8 * nb + 4 * rn) * batch, flops, batch, wall, None, kern)
        plan.close()
        del ins, bins, sq
        torch.cuda.empty_cache()
    return out


def config_lines(args) -> dict:
    """Secondary lines: the analyzer batch and STFT, CFFT<double>
    (8192 x 4096 points), the headline batch from host memory (PCIe-inclusive,
    never `value`), and the large SoA CFFTs.  GB/s counts one read of the
    inputs and one write of the outputs; GFLOP/s uses 5 N log2 N per CFFT and
    2.5 N log2 N per RFFT.  rocFFT (through torch.fft) is timed beside the
    analyzer batch as an external point of reference."""
    import torch
    import oracle
    import zldsp_fft_amd as zf
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream
    lib = zf.load_library()
    out = analyzer_lines(args)
    torch.cuda.empty_cache()
    # CFFT<double>: 8192 x 4096 points AoS (1 GiB of traffic per batch, like the headline batch)
    n64, b64 = 4096, 8192
    plan = zf.CFFT(12, zf.PREC_F64)
    h = plan._h
    g = torch.Generator(device=dev).manual_seed(7)
    ins = [torch.randn((b64, n64), dtype=torch.complex128, device=dev, generator=g) for _ in range(2)]
    outs = [torch.empty_like(ins[0]) for _ in range(2)]
    fns = [lambda i=i: _call(lib, lib.zlfft_cfft_exec, h, zf.FORWARD, zf.AOS, ins[i].data_ptr(), None, zf.AOS,
                             outs[i].data_ptr(), None, b64, sp) for i in range(2)]
    wall, kern = measure_line(torch, stream, fns, reps=40)
    ref = oracle.cfft(ins[0][5].cpu().numpy(), True, "f64")
    got = outs[0][5].cpu().numpy()
    out["cfft64_o12_b8192_aos_forward"] = _line(32 * n64 * b64, 5.0 * n64 * 12 * b64, b64, wall,
                                                 float(np.linalg.norm(got - ref) / np.linalg.norm(ref)), kern)
    plan.close()
    del ins, outs
    torch.cuda.empty_cache()
    # host buffers (PCIe-inclusive): the headline batch from host memory through
    # the pipelined host path, pageable and pinned, wall-clock per call
    xh = np.random.default_rng(1).standard_normal((BATCH, 2 * N)).astype(np.float32).view(np.complex64)
    yh = np.empty_like(xh)
    plan = zf.CFFT(ORDER)
    for name, (src, dst) in (("pageable", (xh, yh)),
                             ("pinned", (torch.from_numpy(xh).pin_memory(),
                                         torch.empty(xh.shape, dtype=torch.complex64).pin_memory()))):
        plan.forward(src, dst)  # warm-up: staging allocation
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            plan.forward(src, dst)
            t.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(t))
        out[f"host_{name}_cfft4096_b16384"] = _line(BYTES_PER_STEP, 5.0 * N * ORDER * BATCH, BATCH, ms)
    plan.close()
    del xh, yh
    out.update(large_lines(args))
    return out


def rocfft_c4(x, stream, args) -> dict:
    """rocFFT (through torch.fft) on the same 16384 x 4096 AoS batch: an
    external reference point, timed the same way."""
    import torch
    wall, kern = measure_line(torch, stream, [lambda: torch.fft.fft(x, dim=-1)], reps=40)
    return _line(BYTES_PER_STEP, 5.0 * N * ORDER * BATCH, BATCH, wall, None, kern)


def copy_reference(nbytes: int, structure: int = 0, per_cu: int = 2) -> dict:
    """A plain device copy of the same algorithmic traffic (nbytes / 2 read,
    nbytes / 2 written), timed like the lines (warm-up by time, back-to-back
    launches over ROTATE buffer sets): the practical ceiling of a memory-bound
    kernel that moves those bytes in one launch.  structure 0: one 32 KiB chunk
    per workgroup; 1: the persistent streaming kernels' form, per_cu resident
    256-thread workgroups per CU (libzlprobe.so)."""
    import ctypes
    import torch
    lib = _probe_lib()
    chunk = lib.zlprobe_chunk_bytes()
    half = max(chunk, (nbytes // 2) // chunk * chunk)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    src = [torch.empty(half // 4, device="cuda") for _ in range(ROTATE)]
    dst = [torch.empty(half // 4, device="cuda") for _ in range(ROTATE)]
    for t in src:
        t.uniform_()

    def run(i):
        if lib.zlprobe_copy(ctypes.c_void_p(src[i].data_ptr()), ctypes.c_void_p(dst[i].data_ptr()),
                            ctypes.c_size_t(half), structure, per_cu, ctypes.c_void_p(sp)):
            raise RuntimeError("zlprobe_copy failed")
    wall, _ = measure_line(torch, stream, [lambda i=i: run(i) for i in range(ROTATE)])
    if not torch.equal(src[0], dst[0]):
        raise RuntimeError("reference copy is wrong")
    del src, dst
    torch.cuda.empty_cache()
    gbs = 2 * half / (wall * 1e-3) / 1e9
    return {"bytes": 2 * half, "ms": round(wall, 5), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}


_PROBE = None


def _probe_lib():
    global _PROBE
    if _PROBE is None:
        import ctypes
        path = os.path.join(ROOT, "zldsp_fft_amd", "libzlprobe.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() first")
        _PROBE = ctypes.CDLL(path)
        _PROBE.zlprobe_chunk_bytes.restype = ctypes.c_size_t
    return _PROBE


def annotate_with_copies(lines: dict, cache: dict) -> None:
    """Adds to every line with a byte count the rate of a copy of the same
    bytes (copy_GB/s) and the line's fraction of it (of_copy)."""
    for name, d in lines.items():
        if not isinstance(d, dict) or "GB/s" not in d or name.startswith(("host_", "rocfft")):
            continue
        nbytes = d.get("bytes")
        if not nbytes:
            continue
        if nbytes not in cache:
            cache[nbytes] = copy_reference(nbytes)
        c = cache[nbytes]
        d["copy_GB/s"] = c["GB/s"]
        d["of_copy"] = round(d["GB/s"] / c["GB/s"], 4)
        if d.get("hbm_passes", 1) > 1:
            # a k-pass transform moves its bytes k times: k copies are its ceiling
            d["of_pass_copies"] = round(d["GB/s"] * d["hbm_passes"] / c["GB/s"], 4)


def traffic_bytes():
    """HBM bytes per launch measured with rocprofv3 PMC counters (FETCH_SIZE +
    WRITE_SIZE, scaled by a copy kernel of known size), if profiled."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except (OSError, ValueError):
        pass
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the batch of independent transforms "
              "would shard with no communication, see the sharding paragraph of the design notes.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import oracle
    import zldsp_fft_amd as zf

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    sections = {"configs": config_lines, "ragged": ragged_section, "large": large_lines, "analyzer": analyzer_lines,
                "orders": order_lines}
    if args.copy_probe:
        # known-size copy: reads and writes 512 MiB, for scaling HBM counters
        a = torch.empty(1 << 27, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.only in sections:
        res = sections[args.only](args)
        if args.only != "ragged" and not args.no_copy_ref:
            annotate_with_copies(res, {})
        print(json.dumps({"section": args.only, args.only: res}), flush=True)
        return
    dev = torch.device("cuda:0")
    plan = zf.CFFT(ORDER)
    g = torch.Generator(device=dev).manual_seed(4)
    x = torch.randn((BATCH, N), dtype=torch.complex64, device=dev, generator=g)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)
    lib = zf.load_library()
    xp, yp, h, sp = x.data_ptr(), y.data_ptr(), plan._h, stream.cuda_stream

    def step():
        _call(lib, lib.zlfft_cfft_exec, h, zf.FORWARD, zf.AOS, xp, None, zf.AOS, yp, None, BATCH, sp)

    if args.only == "mfma":
        print(json.dumps({"section": "mfma", "mfma": mfma_variants(x, y, stream, lib, args, step)}), flush=True)
        return

    # warm-up: --warmup steps, extended by time so the clocks settle; no
    # synchronisation before the timed region
    warm_up(torch, [step], min_calls=args.warmup)
    # timed region: exactly --steps back-to-back steps, one event pair on the
    # launch stream around the region (an event pair around every launch adds
    # its marker packets between the kernels: measured about 5 us per step)
    t_start = torch.cuda.Event(enable_timing=True)
    t_stop = torch.cuda.Event(enable_timing=True)
    t_start.record(stream)
    for i in range(args.steps):
        step()
    t_stop.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_stop)
    ms_per_step = total_ms / args.steps
    # the kernel's launch duration for roofline.achieved: an event pair around
    # each launch of a second window (at least MIN_TIMED launches), same stream
    launch_ms = time_each(torch, stream, [step], max(args.steps, MIN_TIMED))
    kern_ms = float(np.mean(launch_ms))

    # accuracy of the timed output against the FP64 oracle (sampled transforms)
    idx = list(range(0, BATCH, BATCH // 16))
    xs = x[idx].cpu().numpy().astype(np.complex128)
    ys = y[idx].cpu().numpy().astype(np.complex128)
    ref = np.stack([oracle.cfft(r, True, "f64") for r in xs])
    rel = float((np.linalg.norm(ys - ref, axis=1) / np.linalg.norm(ref, axis=1)).max())

    gbs = BYTES_PER_STEP / (ms_per_step * 1e-3) / 1e9
    achieved = BYTES_PER_STEP / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": "batched FFT GB/s (% HBM peak) at rel-L2 <=1e-6; MFMA-split vs plain FP32 GPU",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "cfft_f32_aos_n4096_batch16384_forward", "order": ORDER, "batch": BATCH,
                   "layout": "aos->aos", "path": "plain-f32-streaming-stockham"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_bytes()},
        "rel_l2_max_vs_fp64": rel,
        "transforms_per_s": round(BATCH / (ms_per_step * 1e-3), 1),
        "gflops_5nlog2n": round(5 * N * ORDER * BATCH / (ms_per_step * 1e-3) / 1e9, 1),
        "kernel_ms_avg": round(kern_ms, 5),
        "kernel_ms_median": round(float(np.median(launch_ms)), 5),
        "warmup_rule": f"{args.warmup} steps, extended to >= {WARM_S} s of back-to-back launches, no sync gap",
    }
    if args.only == "main":
        args.no_mfma = args.no_configs = args.no_ragged = args.no_cpu = True
    if not args.no_mfma:
        out["mfma"] = mfma_variants(x, y, stream, lib, args, step)
    if args.only != "main":
        out["rocfft"] = rocfft_c4(x, stream, args)
    copies = {}
    if not args.no_copy_ref:
        # the same 1 GiB moved by a plain copy in one launch, in the fastest
        # form measured and in the streaming kernel's form
        grid = copy_reference(BYTES_PER_STEP, 0)
        copies[BYTES_PER_STEP] = grid
        out["copy_reference"] = {"one_chunk_per_workgroup": grid,
                                 "streaming_form_2_per_cu": copy_reference(BYTES_PER_STEP, 1, 2),
                                 "value_of_copy": round(gbs / grid["GB/s"], 4)}
    if not args.no_configs:
        del x, y
        torch.cuda.empty_cache()
        out["configs"] = config_lines(args)
        if not args.no_copy_ref:
            annotate_with_copies(out["configs"], copies)
    if not args.no_ragged:
        torch.cuda.empty_cache()
        out["ragged"] = ragged_section(args)
        if not args.no_copy_ref:
            for d in (out["ragged"], out["ragged"]["without_order_14"]):
                c = copy_reference(d["bytes_per_launch"])
                d["copy_GB/s"] = c["GB/s"]
                d["of_copy"] = round(d["GB/s"] / c["GB/s"], 4)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    if rel > 1e-6:
        out["value_invalid_reason"] = "accuracy above rel-L2 1e-6"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
