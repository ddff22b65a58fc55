"""Autotuner: per CFFT order, the ragged kernel's points per thread (8, 16 or
32, i.e. the threads per transform, the transforms per workgroup and, at 32,
whether the next item's input is prefetched), measured on this GPU and written
to zlfft_tune.txt next to libzlfft.so, where plans read it.

    python -m zldsp_fft_amd.autotune [--out PATH] [--points 2**24]

For every order, a batch of one size that cycles through every variant
(CFFT / RFFT x forward / backward x AoS / SoA, sqr-mag) is timed with each
legal choice; the choices' outputs must agree (rel-L2 <= 1e-6) before one is
kept.  Then the table is refined on a mixed batch (10,000 transforms, orders
uniform in 6..14, every variant): largest order first, each other legal
choice is timed in the mix and kept when it is more than 1% faster.  In one
launch, the orders share the workgroups' runs (items at 32 points per thread
run without the prefetch, the others with it), so an order's best choice
alone is not always its best choice in a mix: at order 13, 16 points per
thread were faster alone and 32 points faster in the mix by 6-9%.
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
    """Whether the ragged kernels run CFFT order `order` at 2^logr points per
    thread: 8 for orders 3-7, 16 for 4-12, 32 for 8-14 (a transform's
    2^(order - logr) threads fit a workgroup: 256 threads up to order 13, 512
    at order 14)."""
    ok = (logr == 3 and 3 <= order <= 7) or (logr == 4 and 4 <= order <= 13) or (logr == 5 and 8 <= order <= 14)
    return ok and (1 << (order - logr)) <= (512 if order >= 14 else 256)


def default_logr(order: int) -> int:
    """The library's choice without a tuning table."""
    return 3 if order <= 3 else 4 if legal(order, 4) else 5


def _rel(a, b):
    """Largest per-transform rel-L2 between two (transforms, floats) output buffers."""
    import torch
    num = torch.linalg.vector_norm((a - b).double(), dim=1)
    den = torch.linalg.vector_norm(b.double(), dim=1).clamp_min(1e-300)
    return float((num / den).max())


def tune(points: int, out_path: str, verbose: bool = True, mix_count: int = 10000) -> dict:
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
    table = [default_logr(n) for n in range(15)]
    for order, (best, _) in results.items():
        table[order] = best
    mixed = refine_on_mix(table, verbose, count=mix_count) if mix_count else {}
    props = torch.cuda.get_device_properties(0)
    with open(out_path, "w") as f:
        f.write("# zlfft tuning table, written by python -m zldsp_fft_amd.autotune\n")
        f.write(f"# device: {props.name} ({getattr(props, 'gcnArchName', '?')}), {props.multi_processor_count} CUs\n")
        f.write(f"# batch per order: {points} points cycling through every transform variant\n")
        if mixed:
            f.write(f"# refined on a mix of {mix_count} transforms, orders 6..14 (mix: ms per launch)\n")
        f.write("# key order value   (ragged_logr: log2 points per thread of the ragged kernel)\n")
        for order, (best, times) in sorted(results.items()):
            detail = " ".join(f"{1 << k}pts={v:.4f}ms" for k, v in sorted(times.items()))
            if order in mixed:
                detail += "  mix: " + " ".join(f"{1 << k}pts={v:.4f}ms" for k, v in sorted(mixed[order].items()))
            f.write(f"ragged_logr {order} {table[order]}   # {detail}\n")
    return results


def refine_on_mix(table: list, verbose: bool = True, count: int = 10000, seed: int = 99) -> dict:
    """Coordinate pass over orders 14..6 on one mixed batch: each other legal
    choice of an order is timed with the rest of `table` fixed, and kept
    (in `table`) when more than 1% faster.  Returns {order: {logr: ms}}."""
    import torch
    descs = workloads.ragged_mix(count, seed)

    def run_ms(t):
        rb = zf.RaggedBatch(descs, logr=t)
        ms = min(workloads.time_ms(torch, rb.run, reps=20, warmup=5) for _ in range(2))
        rb.close()
        return ms

    cur = run_ms(table)
    out = {}
    for order in range(14, 5, -1):
        times = {table[order]: cur}
        for logr in LOGRS:
            if logr == table[order] or not legal(order, logr):
                continue
            t = list(table)
            t[order] = logr
            times[logr] = run_ms(t)
        best = min(times, key=times.get)
        if best != table[order] and times[best] < 0.99 * cur:
            table[order] = best
            cur = times[best]
        out[order] = times
        if verbose:
            print(f"mix, order {order:2d}: " + ", ".join(f"{1 << k:2d} pts/thread {v:.4f} ms" for k, v in sorted(times.items()))
                  + f" -> {1 << table[order]}", flush=True)
    del descs
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "zlfft_tune.txt"))
    ap.add_argument("--points", type=int, default=1 << 24)
    ap.add_argument("--mix", type=int, default=10000, help="transforms in the refinement mix (0: no refinement)")
    args = ap.parse_args()
    tune(args.points, args.out, mix_count=args.mix)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
