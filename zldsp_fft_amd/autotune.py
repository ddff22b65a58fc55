"""Autotuner: per CFFT order, the ragged kernel's points per thread (8, 16 or
32, i.e. the threads per transform, the transforms per workgroup and, at 32,
whether the next item's input is prefetched), measured on this GPU and written
to zlfft_tune.txt next to libzlfft.so, where plans read it.

    python -m zldsp_fft_amd.autotune [--out PATH] [--points 2**24]

For every order, a batch of one size that cycles through every variant
(CFFT / RFFT x forward / backward x AoS / SoA, sqr-mag) is timed with each
legal choice; the choices' outputs must agree (rel-L2 <= 1e-6) before one is
kept.
"""
from __future__ import annotations

import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import zldsp_fft_amd as zf  # noqa: E402
from zldsp_fft_amd import workloads  # noqa: E402


LOGRS = (3, 4, 5)      # log2 points per thread the ragged kernel supports


def legal(order: int, logr: int) -> bool:
    """Whether the ragged kernel runs CFFT order `order` at 2^logr points per
    thread: 8 for orders 3-7, 16 for 4-13, 32 for 8-14."""
    return (logr == 3 and 3 <= order <= 7) or (logr == 4 and 4 <= order <= 13) or (logr == 5 and 8 <= order <= 14)


def default_logr(order: int) -> int:
    """The library's choice without a tuning table."""
    return 3 if order <= 3 else 4 if legal(order, 4) else 5


def _rel(a, b):
    """Largest per-transform rel-L2 between two (transforms, floats) output buffers."""
    import torch
    num = torch.linalg.vector_norm((a - b).double(), dim=1)
    den = torch.linalg.vector_norm(b.double(), dim=1).clamp_min(1e-300)
    return float((num / den).max())


def tune(points: int, out_path: str, verbose: bool = True) -> dict:
    import torch
    results = {}
    base = [default_logr(n) for n in range(15)]
    for order in range(3, 15):
        count = max(len(workloads.VARIANTS), points >> order)
        ut = workloads.UniformTable(order, count, seed=order)
        times = {}
        outs = {}
        for logr in LOGRS:
            if not legal(order, logr):
                continue
            table = list(base)
            table[order] = logr
            rb = zf.RaggedBatch.from_table(ut.table, keep=ut, logr=table)
            times[logr] = workloads.time_ms(torch, rb.run)
            torch.cuda.synchronize()
            outs[logr] = [o.clone() for o in ut.outputs()]
            rb.close()
        first = min(outs)
        for lg in outs:
            worst = max(_rel(a, b) for a, b in zip(outs[lg], outs[first]))
            if worst > 1e-6:
                raise RuntimeError(f"order {order}: configurations {first} and {lg} disagree (rel-L2 {worst:.3e})")
        del ut, outs
        torch.cuda.empty_cache()
        best = min(times, key=times.get)
        results[order] = (best, times)
        if verbose:
            print(f"order {order:2d}: " + ", ".join(f"{1 << k:2d} pts/thread {v:.4f} ms" for k, v in sorted(times.items()))
                  + f" -> {1 << best}", flush=True)
    props = torch.cuda.get_device_properties(0)
    with open(out_path, "w") as f:
        f.write("# zlfft tuning table, written by python -m zldsp_fft_amd.autotune\n")
        f.write(f"# device: {props.name} ({getattr(props, 'gcnArchName', '?')}), {props.multi_processor_count} CUs\n")
        f.write(f"# batch per order: {points} points cycling through every transform variant\n")
        f.write("# key order value   (ragged_logr: log2 points per thread of the ragged kernel)\n")
        for order, (best, times) in sorted(results.items()):
            detail = " ".join(f"{1 << k}pts={v:.4f}ms" for k, v in sorted(times.items()))
            f.write(f"ragged_logr {order} {best}   # {detail}\n")
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "zlfft_tune.txt"))
    ap.add_argument("--points", type=int, default=1 << 24)
    args = ap.parse_args()
    tune(args.points, args.out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
