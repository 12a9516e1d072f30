#!/usr/bin/env python3
"""Suffix-array build throughput on one MI355X.

Workload (default): the repeat-rich 268,435,456-symbol u8 text over {0,1,2,3}
(seeded order-3 Markov chain with Dirichlet(1) rows, 40% overwritten by
planted repeats of mean length 5,000 with 1% substitutions), u32 index, every
suffix sorted.  The text is generated on the host, copied to HBM, and each
timed step is one full GPU build from the HBM-resident text to the
HBM-resident suffix array.  The result of the last step is checked on the
host (permutation of 0..n-1 with strictly increasing adjacent suffixes, which
pins the unique suffix array).

Prints one JSON line (see the keys below).  `roofline` is computed for the
kernel with the largest share of the step, from event pairs around its
launches on the build's stream; `cpu_baseline` times the CPU restatement of
the reference's in-memory sample sort on a bounded prefix of the same text.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "caps-sa_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (generator, n, description)
    "repeat-rich-268M": ("markov", 1 << 28, "order-3 Markov {0..3} + 40% planted repeats (mean 5000, 1% subst)"),
    "uniform-1M": ("uniform", 1 << 20, "i.i.d. uniform {0..3}"),
}


def make_text(kind: str, n: int, seed: int):
    from caps_sa import gen

    if kind == "markov":
        return gen.markov_repeats(n, seed=seed)
    return gen.uniform_u8(n, 4, seed=seed)


# Algorithmic HBM bytes per suffix of each kernel (u8 text, index of I bytes).
def algorithmic_bytes_per_suffix(kernel: str, I: int) -> float:
    return {
        "k_classify": 1 + 4,            # text once + bucket id write
        "k_scatter": 1 + 4 + 8 + I,     # text once + bucket id read + (key, idx) write
        "k_tile_sort": 2 * (8 + I),     # (key, idx) read + write
        "k_detect": 8,                  # keys read once
        "k_boundary_lcp": 8 + I + I,    # keys, idx, lcp write
    }.get(kernel, 0.0)


def cpu_baseline(text, seconds_budget: float, threads: int):
    """The reference's in-memory sample sort (restated in C, oracle/) on a
    prefix of the text, p = 1024 subproblems as in the reference run."""
    import numpy as np

    import oracle as O

    n = 1 << 20
    # grow the prefix until one run takes about a tenth of the budget
    while True:
        ctx = O.Ctx(np.ascontiguousarray(text[:n]))
        t0 = time.perf_counter()
        ctx.sample_sort(p=1024, threads=threads, emit=lambda a: 0)
        dt = time.perf_counter() - t0
        if dt * 8 > seconds_budget or n * 2 > len(text):
            break
        n *= 2
    runs = []
    t_end = time.perf_counter() + seconds_budget
    while time.perf_counter() < t_end or len(runs) < 1:
        t0 = time.perf_counter()
        ctx.sample_sort(p=1024, threads=threads, emit=lambda a: 0)
        runs.append(time.perf_counter() - t0)
        if len(runs) >= 3:
            break
    runs.sort()
    med = runs[len(runs) // 2]
    return {"value": n / med / 1e6, "unit": "M suffixes/s", "cores": threads, "kind": "port",
            "sample": f"first {n} symbols of the same text, build_in_memory_sample_sort p=1024, "
                      f"median of {len(runs)} runs ({med:.2f} s each)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="repeat-rich-268M", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the text length")
    ap.add_argument("--seed", type=int, default=0xC1)
    ap.add_argument("--router", default="bsearch", choices=["bsearch", "pwl", "mlp"])
    ap.add_argument("--refine", default="plan", choices=["plan", "queue"],
                    help="groups of 2..64: the default per-class plan, or persistent work-queue slots for all")
    ap.add_argument("--tile-sort", default="merge", choices=["merge", "radix"])
    ap.add_argument("--bucket-keys", type=int, default=0, help="mean records per bucket (0: library default)")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device three times (calibrates the HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        return 2

    import numpy as np
    import torch

    import caps_sa
    from caps_sa import device as dev

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU"}))
        return 1
    if args.copy_probe:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 2 * src.numel() * 4}))
        return 0
    kind, n_default, desc = WORKLOADS[args.workload]
    n = args.n or n_default
    print(f"[bench] generating {args.workload} n={n}", file=sys.stderr, flush=True)
    text = make_text(kind, n, args.seed)
    d_text = dev.text_to_device(text)
    d_out = torch.empty(n, dtype=torch.int32, device="cuda:0")
    router = {"bsearch": 0, "pwl": 1, "mlp": 2}[args.router]

    def step():
        dev.build_device(d_text, n, 1, idx_bytes=4, router=router, d_out=d_out, tile_keys=args.bucket_keys,
                         flags=(caps_sa.FLAG_PERSISTENT_REFINE if args.refine == "queue" else 0)
                         | (caps_sa.FLAG_TILE_RADIX if args.tile_sort == "radix" else 0))
        return caps_sa.last_stats()

    for i in range(args.warmup):
        st = step()
        print(f"[bench] warmup {i}: {st['ms_total']:.2f} ms", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(step())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / max(args.steps, 1)
    ev_ms = sum(s["ms_total"] for s in stats) / max(len(stats), 1)
    value = n / (ms_step / 1e3) / 1e6
    print(f"[bench] {args.steps} steps: {ms_step:.2f} ms/step (device events {ev_ms:.2f} ms)", file=sys.stderr,
          flush=True)

    # roofline of the dominant kernel
    kernels = ["k_classify", "k_scatter", "k_tile_sort", "k_detect", "k_pairs", "k_groups_small", "k_groups_mid",
               "k_boundary_lcp"]
    avg = {k: sum(s[k] for s in stats) / len(stats) for k in kernels}
    dom = max(kernels, key=lambda k: avg[k] if algorithmic_bytes_per_suffix(k, 4) else -1)
    bpsuf = algorithmic_bytes_per_suffix(dom, 4)
    achieved = n * bpsuf / (avg[dom] / 1e3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_suffix": bpsuf, "kernel_ms": round(avg[dom], 4)}

    verified = None
    if not args.no_verify:
        import oracle as O

        sa = d_out.cpu().numpy().view(np.uint32).astype(np.uint64)
        ctx = O.Ctx(text)
        bad = ctx.verify(sa, strict=True, canon=True, threads=args.cpu_threads)
        verified = bool(bad == -1)
        print(f"[bench] verify: {'ok' if verified else f'FAILED at {bad}'}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(text, args.cpu_seconds, args.cpu_threads)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    last = stats[-1]
    line = {
        "metric": "M suffixes/s (bit-exact SA), learned-router vs plain GPU; HBM GB/s % peak",
        "value": round(value, 2),
        "unit": "M suffixes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "text": desc, "index": "u32", "router": args.router,
                   "refine": args.refine, "tile_sort": args.tile_sort, "seed": args.seed, "buckets": last["buckets"]},
        "verified_bit_exact": verified,
        "phases_ms": {k: round(sum(s[k] for s in stats) / len(stats), 3)
                      for k in ("ms_alphabet", "ms_sample", "ms_classify", "ms_scatter", "ms_tile_sort", "ms_merge",
                                "ms_detect", "ms_refine", "ms_giant")},
        "kernels_ms": {k: round(v, 3) for k, v in avg.items()},
        "groups": {k: last[k] for k in ("pairs", "small_groups", "mid_groups", "giant_groups", "oversized")},
        "buckets": {"count": last["buckets"], "max": last["bucket_max"],
                    "imbalance": round(last["bucket_max"] * last["buckets"] / n, 3)},
        "router": None if args.router == "bsearch" else {
            "kind": args.router,
            "mean_mispredict": round(last["router_err_sum"] / max(last["router_count"], 1), 4),
            "max_mispredict": last["router_err_max"],
            "fallback_frac": round(last["router_fallbacks"] / max(last["router_count"], 1), 6),
            "train_and_sample_ms": round(sum(s["ms_sample"] for s in stats) / len(stats), 3)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0 if verified in (None, True) else 3


if __name__ == "__main__":
    sys.exit(main())
