"""Benchmark of the zldsp::fft hot path on one MI355X.

Workload (one "step"): one forward pass of 16384 complex FP32 CFFTs of
N = 4096 points, AoS in and AoS out (1 GiB of algorithmic HBM traffic:
16 B per point, one read of the input and one write of the output), with the
inputs resident in HBM.  The step is one launch of the LDS-resident kernel
through the C-ABI (zlfft_cfft_exec) on torch's current stream: the persistent
streaming Stockham kernel for N = 4096.

Printed: ONE JSON line with whole-job throughput in GB/s of algorithmic
traffic, the kernel's roofline fraction against the 8 TB/s HBM peak (average
launch duration from HIP events bracketing every launch in the timed region),
the accuracy of the timed output against the FP64 oracle, and a CPU baseline:
the oracle's FP32 restatement of the reference algorithm, timed on this host.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="CPU baseline sampling budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mfma", action="store_true", help="skip the MFMA split-precision variants")
    ap.add_argument("--no-configs", action="store_true", help="skip the analyzer and large-order lines")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged-batch line")
    ap.add_argument("--only", choices=("main", "mfma", "configs", "ragged"), default=None,
                    help="run one section alone (for profiler runs): the headline batch, the MFMA variants, "
                         "the analyzer / large-order lines or the ragged batch")
    ap.add_argument("--mfma-precisions", default=None,
                    help="comma-separated subset of the MFMA variant names")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 512 MiB device copy (reference kernel for profiler byte counters)")
    return ap.parse_args()


def cpu_baseline(seconds: float) -> dict:
    """The oracle's FP32 restatement of the reference CFFT<float> on this host:
    a contiguous slice of the batch per thread, one workspace per thread."""
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    sample = 2048  # transforms per timed run (1/8 of the batch)
    x = np.random.default_rng(0).standard_normal((sample, 2 * N)).astype(np.float32)
    y = np.empty_like(x)
    oracle.cfft_batch_raw(ORDER, x, y, True, "f32", threads)  # warm-up
    times = []
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        oracle.cfft_batch_raw(ORDER, x, y, True, "f32", threads)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    l1, sw = oracle.switch_order("f32")
    path = "stockham" if ORDER < sw else "dif-macro"
    return {
        "value": round(16 * N * sample / t / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} x 4096-point CFFT<float> forward, AoS, median of {len(times)} runs; "
                  f"oracle FP32 restatement of the reference ({path} plan, switch order {sw})",
        "transforms_per_s": round(sample / t, 1),
    }


MFMA_VARIANTS = (("f16_split", 2), ("bf16_split3", 5), ("f16", 1), ("bf16", 3), ("bf16_split", 4))


def mfma_variants(x, y, xs, ref, stream, lib, args) -> dict:
    """The same workload on the MFMA split-precision kernels: each radix-16
    pass as a 16x16 DFT-matrix product on FP16 / BF16 matrix cores with FP32
    accumulate.  Throughput (event-timed like the main line) and accuracy vs
    the FP64 oracle on the same sampled transforms."""
    import torch
    import zldsp_fft_amd as zf
    res = {}
    idx = list(range(0, BATCH, BATCH // 16))
    steps = args.steps
    wanted = args.mfma_precisions.split(",") if args.mfma_precisions else None
    for name, prec in MFMA_VARIANTS:
        if wanted is not None and name not in wanted:
            continue
        plan = zf.CFFT(ORDER, prec)
        h, sp = plan._h, stream.cuda_stream

        def step():
            rc = lib.zlfft_cfft_exec(h, zf.FORWARD, zf.AOS, x.data_ptr(), None, zf.AOS, y.data_ptr(), None, BATCH, sp)
            if rc:
                raise RuntimeError(lib.zlfft_last_error().decode())

        for _ in range(max(20, args.warmup)):
            step()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for _ in range(steps):
            step()
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        ys = y[idx].cpu().numpy().astype(np.complex128)
        rel = float((np.linalg.norm(ys - ref, axis=1) / np.linalg.norm(ref, axis=1)).max())
        mbe = float((np.abs(ys - ref).max(axis=1) / np.abs(ref).max(axis=1)).max())
        gbs = BYTES_PER_STEP / (ms * 1e-3) / 1e9
        res[name] = {"GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4), "rel_l2_max": rel,
                     "max_bin_err": mbe, "meets_1e-6": rel <= 1e-6}
        plan.close()
    return res


def ragged_section(args) -> dict:
    """The ragged batch: 10,000 transforms with orders uniform in 6..14,
    CFFT / RFFT, forward / backward, AoS / SoA (50/50 each), in one
    graph-launched persistent kernel.  Throughput in algorithmic bytes (one
    read of each input, one write of each output) per graph launch."""
    import torch
    import zldsp_fft_amd as zf
    from zldsp_fft_amd import workloads
    descs = workloads.ragged_mix(10000, 6)
    nbytes = sum(workloads.algorithmic_bytes(d) for d in descs)
    rb = zf.RaggedBatch(descs)
    ms = workloads.time_ms(torch, rb.run, reps=max(20, args.steps // 4), warmup=10)
    rb.close()
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"workload": "ragged_10000_orders6-14_mixed", "bytes_per_launch": nbytes, "ms_per_launch": round(ms, 5),
            "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
            "transforms_per_s": round(len(descs) / (ms * 1e-3), 1)}


ROTATE = 5  # buffer sets per line: > 256 MiB in total, so the MALL cannot hold the working set


def _timed(torch, stream, fns, reps, warmup=5):
    """Mean ms per call over `reps` calls cycling through `fns`, HIP events on `stream`."""
    for i in range(warmup):
        fns[i % len(fns)]()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for i in range(reps):
        fns[i % len(fns)]()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def _line(nbytes, flops, count, ms, rel=None):
    gbs = nbytes / (ms * 1e-3) / 1e9
    d = {"ms": round(ms, 5), "GB/s": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
         "transforms_per_s": round(count / (ms * 1e-3), 1), "gflops": round(flops / (ms * 1e-3) / 1e9, 1)}
    if rel is not None:
        d["rel_l2_vs_fp64"] = rel
    return d


def config_lines(args) -> dict:
    """Throughput of the analyzer batch (4096 RFFTs of 2048 reals: AoS bins and
    squared magnitudes), of CFFT<double> (8192 x 4096 points) and of the
    large SoA CFFTs (orders 16/18/20/22 at batch 64/16/4/1, forward and
    backward).  Each line cycles through
    ROTATE buffer sets; GB/s counts one read of the inputs and one write of
    the outputs; GFLOP/s uses 5 N log2 N per CFFT and 2.5 N log2 N per RFFT.
    rocFFT (through torch.fft) is timed beside it as an external reference
    point on the AoS layouts it supports."""
    import torch
    import oracle
    import zldsp_fft_amd as zf
    from zldsp_fft_amd import signals
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    reps = max(20, args.steps // 4)
    out = {}
    # analyzer batch: 4096 Hann-windowed frames of 2048 reals
    n, frames = 2048, 4096
    nb = n // 2 + 1
    plan = zf.RFFT(11)
    x0 = torch.from_numpy(signals.analyzer_frames(2, frames, n).astype(np.float32)).to(dev)
    xs = [x0] + [x0.clone() for _ in range(ROTATE - 1)]  # distinct addresses, same frames
    ys = [torch.empty(frames * nb, dtype=torch.complex64, device=dev) for _ in range(ROTATE)]
    ps = [torch.empty(frames * nb, dtype=torch.float32, device=dev) for _ in range(ROTATE)]
    flops = 2.5 * n * 11 * frames
    ms = _timed(torch, stream, [lambda i=i: plan.forward(xs[i].view(-1), ys[i]) for i in range(ROTATE)], reps)
    ref = oracle.rfft_forward(xs[0][7].cpu().numpy().astype(np.float64), "f64")
    got = ys[0].view(frames, nb)[7].cpu().numpy()
    out["rfft2048_b4096_aos"] = _line(frames * (4 * n + 8 * nb), flops, frames, ms,
                                       float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    ms = _timed(torch, stream, [lambda i=i: plan.forward_sqr_mag(xs[i].view(-1), ps[i]) for i in range(ROTATE)], reps)
    got = ps[0].view(frames, nb)[7].cpu().numpy()
    out["rfft2048_b4096_sqr_mag"] = _line(frames * (4 * n + 4 * nb), flops, frames, ms,
                                           float(np.linalg.norm(got - np.abs(ref) ** 2) / np.linalg.norm(np.abs(ref) ** 2)))
    ms = _timed(torch, stream, [lambda i=i: torch.fft.rfft(xs[i], dim=-1) for i in range(ROTATE)], reps)
    out["rocfft_rfft2048_b4096_aos"] = _line(frames * (4 * n + 8 * nb), flops, frames, ms)
    # STFT front-end: 4096 Hann-windowed frames of 2048 samples at hop 512 of
    # one signal, dB levels (frames cut and windowed at load time); bytes count
    # the frame samples read and the levels written
    hop = 512
    sig = [torch.rand((frames - 1) * hop + n, device=dev) * 2 - 1 for _ in range(ROTATE)]
    win = torch.from_numpy(signals.hann(n).astype(np.float32)).to(dev)
    lv = [torch.empty(frames * nb, dtype=torch.float32, device=dev) for _ in range(ROTATE)]
    ms = _timed(torch, stream, [lambda i=i: plan.stft(sig[i], hop, lv[i], window=win, output=zf.STFT_DB,
                                                       db_floor=1e-12, nframes=frames) for i in range(ROTATE)], reps)
    out["stft2048_hop512_b4096_hann_db"] = _line(frames * (4 * n + 4 * nb), flops, frames, ms)
    del sig, lv
    plan.close()
    del xs, ys, ps
    torch.cuda.empty_cache()
    # CFFT<double>: 8192 x 4096 points AoS (1 GiB of traffic per batch, like the headline batch)
    n64, b64 = 4096, 8192
    plan = zf.CFFT(12, zf.PREC_F64)
    g = torch.Generator(device=dev).manual_seed(7)
    ins = [torch.randn((b64, n64), dtype=torch.complex128, device=dev, generator=g) for _ in range(2)]
    outs = [torch.empty_like(ins[0]) for _ in range(2)]
    ms = _timed(torch, stream, [lambda i=i: plan.forward(ins[i], outs[i]) for i in range(2)], reps)
    ref = oracle.cfft(ins[0][5].cpu().numpy(), True, "f64")
    got = outs[0][5].cpu().numpy()
    out["cfft64_o12_b8192_aos_forward"] = _line(32 * n64 * b64, 5.0 * n64 * 12 * b64, b64, ms,
                                                 float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    plan.close()
    del ins, outs
    torch.cuda.empty_cache()
    # host buffers (PCIe-inclusive): the headline batch from host memory through
    # the pipelined host path, pageable and pinned, wall-clock per call
    import time as _time
    xh = np.random.default_rng(1).standard_normal((BATCH, 2 * N)).astype(np.float32).view(np.complex64)
    yh = np.empty_like(xh)
    plan = zf.CFFT(ORDER)
    for name, (src, dst) in (("pageable", (xh, yh)),
                             ("pinned", (torch.from_numpy(xh).pin_memory(), torch.empty(xh.shape, dtype=torch.complex64).pin_memory()))):
        plan.forward(src, dst)  # warm-up: staging allocation
        t = []
        for _ in range(3):
            t0 = _time.perf_counter()
            plan.forward(src, dst)
            t.append(_time.perf_counter() - t0)
        ms = 1e3 * float(np.median(t))
        out[f"host_{name}_cfft4096_b16384"] = _line(BYTES_PER_STEP, 5.0 * N * ORDER * BATCH, BATCH, ms)
    plan.close()
    del xh, yh
    # large SoA CFFTs: 4,194,304 points per call
    for order, batch in ((16, 64), (18, 16), (20, 4), (22, 1)):
        n = 1 << order
        plan = zf.CFFT(order)
        g = torch.Generator(device=dev).manual_seed(3)
        ins = [(torch.randn(batch * n, device=dev, generator=g), torch.randn(batch * n, device=dev, generator=g))
               for _ in range(ROTATE)]
        outs = [(torch.empty(batch * n, device=dev), torch.empty(batch * n, device=dev)) for _ in range(ROTATE)]
        for name, fn in (("forward", plan.forward), ("backward", plan.backward)):
            ms = _timed(torch, stream, [lambda i=i, fn=fn: fn(ins[i], outs[i]) for i in range(ROTATE)], reps)
            x = ins[0][0][:n].cpu().numpy().astype(np.float64) + 1j * ins[0][1][:n].cpu().numpy()
            ref = oracle.cfft(x, name == "forward", "f64")
            got = outs[0][0][:n].cpu().numpy() + 1j * outs[0][1][:n].cpu().numpy().astype(np.float64)
            out[f"cfft_o{order}_b{batch}_soa_{name}"] = _line(16 * n * batch, 5.0 * n * order * batch, batch, ms,
                                                              float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
        plan.close()
        del ins, outs
        torch.cuda.empty_cache()
    return out


def rocfft_c4(x, stream, args) -> dict:
    """rocFFT (through torch.fft) on the same 16384 x 4096 AoS batch: an
    external reference point, timed the same way."""
    import torch
    ms = _timed(torch, stream, [lambda: torch.fft.fft(x, dim=-1)], max(20, args.steps // 4), warmup=10)
    return _line(BYTES_PER_STEP, 5.0 * N * ORDER * BATCH, BATCH, ms)


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
    if args.only in ("configs", "ragged"):
        out = {"section": args.only}
        out[args.only] = config_lines(args) if args.only == "configs" else ragged_section(args)
        print(json.dumps(out), flush=True)
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
        rc = lib.zlfft_cfft_exec(h, zf.FORWARD, zf.AOS, xp, None, zf.AOS, yp, None, BATCH, sp)
        if rc:
            raise RuntimeError(lib.zlfft_last_error().decode())

    if args.copy_probe:
        # known-size copy: reads and writes BATCH * N * 8 bytes, for scaling HBM counters
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
    if args.only == "mfma":
        idx = list(range(0, BATCH, BATCH // 16))
        xs = x[idx].cpu().numpy().astype(np.complex128)
        ref = np.stack([oracle.cfft(r, True, "f64") for r in xs])
        print(json.dumps({"section": "mfma", "mfma": mfma_variants(x, y, xs, ref, stream, lib, args)}), flush=True)
        return
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # timed region: K steps, each launch bracketed by events on the launch stream
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for i in range(args.steps):
        ev[i][0].record(stream)
        step()
        ev[i][1].record(stream)
    t_stop.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_stop)
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    ms_per_step = total_ms / args.steps
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
    }
    if args.only == "main":
        args.no_mfma = args.no_configs = args.no_ragged = args.no_cpu = True
    if not args.no_mfma:
        out["mfma"] = mfma_variants(x, y, xs, ref, stream, lib, args)
    if args.only != "main":
        out["rocfft"] = rocfft_c4(x, stream, args)
    if not args.no_configs:
        out["configs"] = config_lines(args)
    if not args.no_ragged:
        del x, y
        torch.cuda.empty_cache()
        out["ragged"] = ragged_section(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    if rel > 1e-6:
        out["value_invalid_reason"] = "accuracy above rel-L2 1e-6"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
