#!/usr/bin/env python3
"""Suffix-array build throughput on one MI355X.

Workload (default): the repeat-rich 268,435,456-symbol u8 text over {0,1,2,3}
(seeded order-3 Markov chain with Dirichlet(1) rows, 40% overwritten by
planted repeats of mean length 5,000 with 1% substitutions), u32 index, every
suffix sorted.  The text is generated on the host, copied to HBM, and each
timed step is one full GPU build from the HBM-resident text to the
HBM-resident suffix array.  `value` is the default build (binary search over
the splitters, 1,024 classification buckets).  `learned_vs_plain` times the
same text with 4,096 classification buckets decided by the learned router
(piecewise-linear CDF by default) and by binary search; all suffix arrays
must be equal, and the default build's is checked on the host (permutation
of 0..n-1 with strictly increasing adjacent suffixes, which pins the unique
suffix array).

Prints one JSON line (see the keys below).  `roofline` is the single-launch
kernel with the longest average launch, from event pairs around its launch on
the build's stream, against its algorithmic bytes; `cpu_baseline` times the
CPU restatement of the reference's in-memory sample sort on the host cores.
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
    "genome-share-820M": ("genome", 820_000_000,
                          "one GPU's eighth of the 6.56e9-symbol genome: 175,000 lognormal segments (median 2,500, "
                          "sigma 1.6) of i.i.d. {0..3} ending in spacer 4, ~5.8% N runs (symbol 5, geometric mean "
                          "10,000); positions {p : text[p] < 4}, SegmentedText ShorterFirst, u64 index"),
    "genome-6.56G": ("genome-full", 6_560_000_000,
                     "the whole 6.56e9-symbol genome on one GPU: 1,400,000 lognormal segments (median 2,500, sigma "
                     "1.6) of i.i.d. {0..3} ending in spacer 4, ~5.8% N runs (symbol 5, geometric mean 10,000); kept "
                     "positions {p : text[p] < 4} as a bitmap (build_ext_mem_for_filter), SegmentedText "
                     "ShorterFirst, u64 index, bucket batches"),
    "batch-4096": ("batch", 0, "4,096 texts, lognormal lengths (median 65,536, sigma 1.5, clipped to [1,024, "
                              "16,777,216]); even texts u8 i.i.d. {0..3}, odd texts u16 Zipf(1.1) over 50,000 symbols"),
}


def make_text(kind: str, n: int, seed: int):
    from caps_sa import gen

    if kind == "markov":
        return gen.markov_repeats(n, seed=seed)
    return gen.uniform_u8(n, 4, seed=seed)


def launch_bytes(st: dict, n: int, w: int, I: int) -> dict:
    """Algorithmic HBM bytes of each single-launch kernel of one build (what
    the kernel must read and write at the least), from the build's own
    counts.  Text bytes are w per symbol; records are R = 8 + I bytes (key and
    index).  The pair refinement reads the packed code text (fields of pb
    bits, the power of two at or above the seed field width): its text bytes
    are the symbols it actually read (a CAPS_FLAG_COUNT_WORK build's device
    counters) times pb / 8."""
    R = 8 + I
    out = {
        "k_classify": n * (w + 2),            # text once + coarse interval id (u16) written
        "k_scatter1": n * (w + 2 + R),        # text once + interval id read + record written
        "k_scatter2": n * 2 * R,              # record read + record written to its bucket
        # records of the class read once, their positions written once (a
        # build without an LCP array writes no sorted keys back for tiles
        # that hold a whole bucket, and here every tile does)
        "k_tile_1024": st["tile_records"][0] * (R + I),
        "k_tile_2048": st["tile_records"][1] * (R + I),
        "k_tile_4096": st["tile_records"][2] * (R + I),
    }
    if st.get("cmp_symbols") and any(st["cmp_symbols"]):
        # pair records (start + two positions) read, two positions written, packed text compared
        pb = 1
        while pb < st["key_bits"]:
            pb *= 2
        out["k_pairs"] = st["pairs"] * (8 + 4 * I) + st["cmp_symbols"][0] * pb / 8
    return out


def launch_ms(st: dict) -> dict:
    return {"k_classify": st["k_classify"], "k_scatter1": st["k_scatter1"], "k_scatter2": st["k_scatter2"],
            "k_tile_1024": st["k_tile"][0], "k_tile_2048": st["k_tile"][1], "k_tile_4096": st["k_tile"][2],
            "k_pairs": st["k_pairs"]}


# rocprofv3 names of those kernels (u8 text, u32 index, merge-path tile sort)
ROCPROF_NAMES = {"k_classify": "k_classify_seq<u8, u32, {rk}, false, 8>", "k_scatter1": "k_scatter1_staged<u8, u32, false>",
                 "k_scatter2": "k_scatter2<u32, 1024, 8192, true>", "k_tile_1024": "k_tile_merge<u32, 1024, true>",
                 "k_tile_2048": "k_tile_merge<u32, 2048, true>", "k_tile_4096": "k_tile_merge<u32, 4096, true>",
                 "k_pairs": "k_pairs_packed<u8, u32>"}


def measured_traffic(workload: str, kernel: str, rk: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    runs of this workload (profiles/hbm_traffic.json: FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by what the counter reads for a
    plain copy of known size); None when not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        k = d["workloads"][workload][ROCPROF_NAMES[kernel].format(rk=rk)]
        cal = d["calibration"]
        return round(k["fetch_kib"] * 1024 * cal["fetch"] + k["write_kib"] * 1024 * cal["write"])
    except (OSError, KeyError, ValueError):
        return None


def rooflines_of(stats: list, n: int, w: int, I: int, counted: dict, workload: str, rk: int) -> dict:
    """Per-launch roofline of every kernel with algorithmic bytes: bytes of
    one launch / its average duration (HIP events around the launch on the
    build's stream, averaged over the timed builds)."""
    bytes_ = launch_bytes(counted, n, w, I)
    out = {}
    for k, b in bytes_.items():
        ms = sum(launch_ms(s)[k] for s in stats) / len(stats)
        if ms <= 0 or b <= 0:
            continue
        achieved = b / (ms / 1e3) / 1e9
        out[k] = {"bound": "hbm", "kernel": k, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(workload, k, rk),
                  "algorithmic_bytes": int(b), "kernel_ms": round(ms, 4)}
    return out


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(text, seconds_budget: float, threads: int, p: int):
    """The reference's in-memory sample sort (the oracle's C restatement of
    ext_mem.rs, p subproblems) on the host cores, pinned to them as taskset
    would (sched_setaffinity), median of 3 runs: on the whole text when one
    run fits the budget, else on the largest power-of-two prefix that does."""
    import numpy as np

    import oracle as O

    allowed = sorted(os.sched_getaffinity(0))
    threads = min(threads, len(allowed))
    cpus = allowed[:threads]
    old = os.sched_getaffinity(0)
    os.sched_setaffinity(0, cpus)
    try:
        m = min(len(text), 1 << 24)
        ctx = O.Ctx(np.ascontiguousarray(text[:m]))
        t0 = time.perf_counter()
        ctx.sample_sort(p=p, threads=threads, emit=lambda a: 0)
        rate = m / (time.perf_counter() - t0)  # suffixes per second on the probe
        n = len(text)
        while n > m and n / rate > seconds_budget / 3:
            n //= 2
        ctx = O.Ctx(np.ascontiguousarray(text[:n]))
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            ctx.sample_sort(p=p, threads=threads, emit=lambda a: 0)
            runs.append(time.perf_counter() - t0)
    finally:
        os.sched_setaffinity(0, old)
    runs.sort()
    med = runs[1]
    whole = n == len(text)
    return {"value": round(n / med / 1e6, 3), "unit": "M suffixes/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_info(), "host_cpus_allowed": len(allowed), "pinned_cpus": f"{cpus[0]}-{cpus[-1]}",
            "sample": (f"{'the whole text' if whole else f'first {n} symbols of the same text'} "
                       f"({n} suffixes), the oracle's build_in_memory_sample_sort with p={p} on {threads} pinned "
                       f"threads, median of 3 runs ({med:.2f} s; runs {', '.join(f'{r:.2f}' for r in runs)} s)")}


def run_batch(args, count: int) -> dict:
    """Batched builds: every text's suffix array from one pipeline per symbol
    width (caps_build_batch_device), texts resident in HBM back to back."""
    import numpy as np
    import torch

    import caps_sa
    from caps_sa import device as dev
    from caps_sa import gen

    lengths = gen.lognormal_lengths(count, 65536.0, 1.5, 1024, 1 << 24, seed=args.seed).astype(np.int64)
    texts = []
    for i, m in enumerate(lengths.tolist()):
        sd = args.seed * 1_000_003 + i
        texts.append(gen.uniform_u8(m, 4, seed=sd) if i % 2 == 0 else gen.zipf_u16(m, 50000, 1.1, seed=sd))
    groups = []
    for w, sel in ((1, [i for i in range(count) if i % 2 == 0]), (2, [i for i in range(count) if i % 2 == 1])):
        if not sel:
            continue
        cat = np.concatenate([texts[i] for i in sel])
        ends = np.cumsum([len(texts[i]) for i in sel]).astype(np.uint64)
        d_text = dev.text_to_device(cat)
        d_out = torch.empty(len(cat), dtype=torch.int32, device="cuda:0")
        groups.append((w, sel, ends, d_text, d_out, len(cat)))
    total = sum(g[5] for g in groups)
    print(f"[bench] batch: {count} texts, {total} symbols", file=sys.stderr, flush=True)
    knobs = [{"tile_keys": 0, "flags": 0} for _ in groups]
    tuned = None
    if args.autotune:
        tuned = []
        for gi, (w, sel, ends, d_text, d_out, m) in enumerate(groups):
            choice, report = dev.autotune_batch(d_text, m, w, ends, 4)
            knobs[gi] = choice
            tuned.append({"sym_bytes": w, "choice": choice, "report": report})
        print(f"[bench] autotune: {knobs}", file=sys.stderr, flush=True)

    def step():
        out = []
        for gi, (w, sel, ends, d_text, d_out, m) in enumerate(groups):
            dev.build_batch_device(d_text, m, w, ends, 4, d_out, tile_keys=knobs[gi]["tile_keys"],
                                   flags=knobs[gi]["flags"] | args.flags)
            out.append(caps_sa.last_stats())
        return out

    for i in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / max(args.steps, 1)
    verified = None
    sample = []
    if not args.no_verify:
        import oracle as O

        rng = np.random.default_rng(args.seed)
        sample = sorted(rng.choice(count, size=min(count, 48), replace=False).tolist())
        ok = True
        for w, sel, ends, d_text, d_out, m in groups:
            got = d_out.cpu().numpy().view(np.uint32).astype(np.uint64)
            starts = np.concatenate([[0], ends[:-1]]).astype(np.int64)
            for k, i in enumerate(sel):
                if i not in sample:
                    continue
                a, b = int(starts[k]), int(ends[k])
                want = O.Ctx(texts[i]).build_in_memory(threads=args.cpu_threads)
                ok = ok and bool(np.array_equal(got[a:b] - a, want))
        verified = ok
        print(f"[bench] verify {len(sample)} texts: {'ok' if ok else 'FAILED'}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        import oracle as O

        t_end = time.perf_counter() + args.cpu_seconds
        done, t_cpu, k = 0, 0.0, 0
        order = np.random.default_rng(args.seed + 1).permutation(count)
        while time.perf_counter() < t_end and k < count:
            tx = texts[int(order[k])]
            t1 = time.perf_counter()
            O.Ctx(tx).build_in_memory(threads=args.cpu_threads)
            t_cpu += time.perf_counter() - t1
            done += len(tx)
            k += 1
        cpu = {"value": round(done / t_cpu / 1e6, 3), "unit": "M suffixes/s", "cores": args.cpu_threads,
               "kind": "port", "sample": f"build_in_memory on {k} texts of the same batch ({done} symbols)"}
    kern = ["k_classify", "k_scatter1", "k_scatter2", "k_scatter", "k_tile_sort", "k_pairs", "k_groups_small",
            "k_groups_mid", "k_boundary_lcp"]
    avg = {kk: sum(sum(x[kk] for x in st) for st in stats) / len(stats) for kk in kern}
    phases = ("ms_alphabet", "ms_sample", "ms_classify", "ms_scatter", "ms_tile_sort", "ms_merge", "ms_detect",
              "ms_refine", "ms_giant")
    per_width = []
    for gi, (w, sel, ends, d_text, d_out, m) in enumerate(groups):
        last = stats[-1][gi]
        per_width.append({"sym_bytes": w, "texts": len(sel), "symbols": m, "key_k": last["key_k"],
                          "phases_ms": {ph: round(sum(st[gi][ph] for st in stats) / len(stats), 3) for ph in phases},
                          "groups": {k: last[k] for k in ("pairs", "small_groups", "mid_groups", "giant_groups",
                                                           "oversized", "buckets", "bucket_max")}})
    return {"ms_step": ms_step, "n": total, "verified": verified, "sample_texts": len(sample), "cpu": cpu,
            "kernels": avg, "last": stats[-1], "per_width": per_width, "autotune": tuned}


def run_genome(args, n: int, filtered: bool = False) -> dict:
    """Segmented + filtered build at genome scale on one GPU: suffix array in
    HBM (u64), SegmentedText limits.  The kept positions are a u64 position
    list in HBM, or (filtered) a bitmap walked beside the text, as the
    crate's build_ext_mem_for_filter does."""
    import numpy as np
    import torch

    import caps_sa
    from caps_sa import device as dev
    from caps_sa import gen

    n_seg = max(1, round(n * 1.40e6 / 6.56e9))
    t_gen = time.perf_counter()
    print(f"[bench] generating genome n={n} segments={n_seg}", file=sys.stderr, flush=True)
    text, ends, pos = gen.genome(n, n_seg, seed=args.seed, positions=not filtered)
    lp = caps_sa.SegmentedText(n, ends)
    d_text = dev.text_to_device(text)
    flags = (caps_sa.FLAG_POW2_SEED if args.pow2_seed else 0) | args.flags
    if filtered:
        keep = gen.keep_below(text, 4)
        n_pos = int(np.bitwise_count(keep).sum())
        d_keep = torch.from_numpy(keep.view(np.int64)).to("cuda:0")
        d_out = torch.empty(n_pos, dtype=torch.int64, device="cuda:0")

        def step():
            dev.build_device_filtered(d_text, n, 1, d_keep, n_pos, idx_bytes=8, lp=lp, d_out=d_out, flags=flags,
                                      gpu_buckets=args.genome_buckets, tile_keys=args.bucket_keys,
                                      batch_records=args.batch_records)
            return caps_sa.last_stats()
    else:
        n_pos = len(pos)
        d_pos = torch.from_numpy(pos.view(np.int64)).to("cuda:0")
        d_out = torch.empty(n_pos, dtype=torch.int64, device="cuda:0")

        def step():
            dev.build_device(d_text, n, 1, d_positions=d_pos, n_pos=n_pos, idx_bytes=8, lp=lp, d_out=d_out,
                             gpu_buckets=args.genome_buckets, tile_keys=args.bucket_keys,
                             flags=flags, batch_records=args.batch_records)
            return caps_sa.last_stats()
    print(f"[bench] genome: n={n} segments={n_seg} positions={n_pos} "
          f"(inputs ready in {time.perf_counter() - t_gen:.1f} s)", file=sys.stderr, flush=True)

    for i in range(args.warmup):
        st = step()
        print(f"[bench] warmup {i}: {st['ms_total']:.1f} ms, {st['batches']} batch(es)", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(step())
        print(f"[bench] step {i}: {stats[-1]['ms_total']:.1f} ms (scatter {stats[-1]['ms_scatter']:.1f}, giant {stats[-1]['ms_giant']:.1f})", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / max(args.steps, 1)
    verified = None
    if not args.no_verify:
        import oracle as O

        t_v = time.perf_counter()
        sa = d_out.cpu().numpy().view(np.uint64)
        ctx = O.Ctx(text, ends, O.ORDER_SHORTER_FIRST)
        if filtered:
            bad = ctx.verify_keep(sa, keep, strict=True, canon=True, threads=args.cpu_threads)
        else:
            bad = ctx.verify(sa, positions=pos, strict=True, canon=True, threads=args.cpu_threads)
        verified = bool(bad == -1)
        print(f"[bench] verify: {'ok' if verified else f'FAILED at {bad}'} ({time.perf_counter() - t_v:.1f} s)",
              file=sys.stderr, flush=True)
    kern = ["k_classify", "k_scatter1", "k_scatter2", "k_scatter", "k_tile_sort", "k_pairs", "k_groups_small",
            "k_groups_mid", "k_boundary_lcp"]
    avg = {k: sum(st[k] for st in stats) / len(stats) for k in kern}
    phases = ("ms_alphabet", "ms_sample", "ms_classify", "ms_scatter", "ms_tile_sort", "ms_merge", "ms_detect",
              "ms_refine", "ms_giant")
    last = stats[-1]
    return {"ms_step": ms_step, "n_pos": n_pos, "n": n, "segments": n_seg, "verified": verified,
            "batches": last["batches"], "runs": last["runs"], "classification_buckets": last["coarse_buckets"],
            "kernels": avg, "phases": {ph: round(sum(st[ph] for st in stats) / len(stats), 3) for ph in phases},
            "groups": {k: last[k] for k in ("pairs", "small_groups", "mid_groups", "giant_groups", "oversized",
                                             "buckets", "bucket_max")}, "key_k": last["key_k"]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="repeat-rich-268M", choices=sorted(WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the text length")
    ap.add_argument("--seed", type=int, default=0xC1)
    ap.add_argument("--router", default="spline", choices=["bsearch", "pwl", "mlp", "spline"],
                    help="the learned router timed for `value` (repeat-rich workload); binary search is the plain "
                         "build it is compared with")
    ap.add_argument("--buckets", type=int, default=4096,
                    help="classification buckets the classify pass decides among (1..4096; 0: library default)")
    ap.add_argument("--refine", default="plan", choices=["plan", "queue"],
                    help="groups of 2..64: the default per-class plan, or persistent work-queue slots for all")
    ap.add_argument("--tile-sort", default="merge", choices=["merge", "radix"])
    ap.add_argument("--bucket-keys", type=int, default=0, help="mean records per bucket (0: library default)")
    ap.add_argument("--texts", type=int, default=4096, help="batch workload: number of texts")
    ap.add_argument("--genome-buckets", type=int, default=0,
                    help="genome workloads: classification buckets (0: library default)")
    ap.add_argument("--batch-records", type=int, default=0,
                    help="genome workloads: records per bucket batch (0: one batch while it fits in HBM)")
    ap.add_argument("--flags", type=int, default=0, help="extra CAPS_FLAG_* bits")
    ap.add_argument("--pow2-seed", action="store_true",
                    help="genome workload: seed keys with the packed text's power-of-two field width")
    ap.add_argument("--autotune", action="store_true",
                    help="pick bucket size / tile sort / refinement style / merge width by timing builds first")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="repeat-rich workload: skip the plain comparison builds")
    ap.add_argument("--cpu-seconds", type=float, default=40.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: every allowed host CPU)")
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
        # 1 GiB read + 1 GiB written per launch: a device copy (runtime copy
        # kernel) and an elementwise kernel (src + 0 into dst)
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
            torch.add(src, 0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 4, "copy_probe_bytes_written": src.numel() * 4}))
        return 0
    # the host CPUs this job may use: its affinity, within the thread count
    # the environment grants it (OMP_NUM_THREADS)
    share = len(os.sched_getaffinity(0))
    try:
        share = min(share, int(os.environ.get("OMP_NUM_THREADS", "0")) or share)
    except ValueError:
        pass
    cpu_threads = args.cpu_threads or share
    args.cpu_threads = cpu_threads
    kind, n_default, desc = WORKLOADS[args.workload]
    if kind == "batch":
        r = run_batch(args, args.texts)
        n = r["n"]
        value = n / (r["ms_step"] / 1e3) / 1e6
        avg = r["kernels"]
        dom = max(("k_classify", "k_scatter1", "k_scatter2"), key=lambda k: avg[k])
        per = {"k_classify": 1 + 2, "k_scatter1": 1 + 2 + 12, "k_scatter2": 24}[dom]
        achieved = n * per / (avg[dom] / 1e3) / 1e9 if avg[dom] else 0.0
        line = {
            "metric": "M suffixes/s (bit-exact SA), learned-router vs plain GPU; HBM GB/s % peak",
            "value": round(value, 2), "unit": "M suffixes/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(r["ms_step"], 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8+u16", "data": "synthetic",
            "config": {"workload": args.workload, "texts": args.texts, "symbols": n, "text": desc, "index": "u32", "flags": args.flags,
                       "seed": args.seed},
            "verified_bit_exact": r["verified"], "verified_texts": r["sample_texts"],
            "kernels_ms": {k: round(v, 3) for k, v in avg.items()},
            "builds": r["per_width"],
            "autotune": r["autotune"],
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                         "bytes_per_suffix": per, "kernel_ms": round(avg[dom], 4)},
            "cpu_baseline": r["cpu"],
        }
        print(json.dumps(line))
        return 0 if r["verified"] in (None, True) else 3
    if kind in ("genome", "genome-full"):
        r = run_genome(args, args.n or n_default, filtered=kind == "genome-full")
        n_pos = r["n_pos"]
        value = n_pos / (r["ms_step"] / 1e3) / 1e6
        avg = r["kernels"]
        rooflines = {}
        for k in ("k_classify", "k_scatter1", "k_scatter2"):
            bpsuf = {"k_classify": 8 + 1 + 2, "k_scatter1": 8 + 2 + 16, "k_scatter2": 2 * 16}[k]
            achieved = n_pos * bpsuf / (avg[k] / 1e3) / 1e9 if avg[k] else 0.0
            rooflines[k] = {"bound": "hbm", "kernel": k, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                            "bytes_per_suffix": bpsuf, "kernel_ms": round(avg[k], 4)}
        dom = max(rooflines, key=lambda k: avg[k])
        line = {
            "metric": "M suffixes/s (bit-exact SA), learned-router vs plain GPU; HBM GB/s % peak",
            "value": round(value, 2), "unit": "M suffixes/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(r["ms_step"], 3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": args.workload, "n": r["n"], "segments": r["segments"], "positions": n_pos,
                       "text": desc, "index": "u64", "key_symbols": r["key_k"], "bucket_batches": r["batches"],
                       "classification_buckets": r["classification_buckets"], "run_memo_runs": r["runs"],
                       "seed": args.seed},
            "verified_bit_exact": r["verified"], "phases_ms": r["phases"],
            "kernels_ms": {k: round(v, 3) for k, v in avg.items()}, "groups": r["groups"],
            "roofline": rooflines[dom], "rooflines": rooflines, "cpu_baseline": None,
        }
        print(json.dumps(line))
        return 0 if r["verified"] in (None, True) else 3
    return run_single(args, kind, desc, n_default)


def run_single(args, kind: str, desc: str, n_default: int) -> int:
    """One text, every suffix.  `value` is the library's default build
    (binary search over 1,023 coarse splitters: 1,024 classification
    buckets, the reference run's 1,024 splitters).  Beside it, on the same
    text: the build with `--buckets` classification buckets decided by the
    learned router, and the same build with binary search (the plain build
    the router is compared with).  All three suffix arrays must be equal."""
    import numpy as np
    import torch

    import caps_sa
    from caps_sa import device as dev

    n = args.n or n_default
    print(f"[bench] generating {args.workload} n={n}", file=sys.stderr, flush=True)
    text = make_text(kind, n, args.seed)
    d_text = dev.text_to_device(text)
    d_out = torch.empty(n, dtype=torch.int32, device="cuda:0")
    d_other = torch.empty(n, dtype=torch.int32, device="cuda:0")
    rk = {"bsearch": 0, "pwl": 1, "mlp": 2, "spline": 3}[args.router]
    tile_keys = args.bucket_keys
    flags = ((caps_sa.FLAG_PERSISTENT_REFINE if args.refine == "queue" else 0)
             | (caps_sa.FLAG_TILE_RADIX if args.tile_sort == "radix" else 0) | args.flags)
    tuned = None
    if args.autotune:
        choice, report = dev.autotune(d_text, n, 1, idx_bytes=4)
        tile_keys, flags = choice["tile_keys"], choice["flags"]
        tuned = {"choice": choice, "report": report}
        print(f"[bench] autotune: {choice}", file=sys.stderr, flush=True)

    def build(router, buckets, out, extra=0):
        dev.build_device(d_text, n, 1, idx_bytes=4, router=router, d_out=out, tile_keys=tile_keys,
                         gpu_buckets=buckets, flags=flags | extra)
        return caps_sa.last_stats()

    def timed(router, buckets, out, label):
        for i in range(args.warmup):
            st = build(router, buckets, out)
            print(f"[bench] {label} warmup {i}: {st['ms_total']:.2f} ms", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = [build(router, buckets, out) for _ in range(args.steps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / max(args.steps, 1)
        ev = sum(s["ms_total"] for s in stats) / max(len(stats), 1)
        print(f"[bench] {label}: {args.steps} steps, {ms:.2f} ms/step (device events {ev:.2f} ms)",
              file=sys.stderr, flush=True)
        return ms, stats

    ms_step, stats = timed(0, 0, d_out, "default build (binary search, 1,024 classification buckets)")
    value = n / (ms_step / 1e3) / 1e6
    same = True
    cmp = None
    if not args.no_plain and args.buckets:
        ms_l, st_l = timed(rk, args.buckets, d_other, f"router={args.router} buckets={args.buckets}")
        same = same and bool(torch.equal(d_out, d_other))
        ms_p, st_p = timed(0, args.buckets, d_other, f"router=bsearch buckets={args.buckets}")
        same = same and bool(torch.equal(d_out, d_other))
        ll, lp = st_l[-1], st_p[-1]
        C = ll["coarse_buckets"]
        avg = lambda sts, k: round(sum(s[k] for s in sts) / len(sts), 3)
        cmp = {"workload": f"the same text, {C} classification buckets",
               "learned": {"router": args.router, "ms_per_step": round(ms_l, 3),
                           "value": round(n / (ms_l / 1e3) / 1e6, 2), "k_classify_ms": avg(st_l, "k_classify"),
                           "train_and_sample_ms": avg(st_l, "ms_sample"),
                           "mean_mispredict": round(ll["router_err_sum"] / max(ll["router_count"], 1), 4),
                           "max_mispredict": ll["router_err_max"],
                           "fallback_frac": round(ll["router_fallbacks"] / max(ll["router_count"], 1), 6),
                           "error_bound": ([ll["router_bound_lo"], ll["router_bound_hi"]]
                                           if args.router == "spline" else None)},
               "plain": {"router": "bsearch", "ms_per_step": round(ms_p, 3),
                         "value": round(n / (ms_p / 1e3) / 1e6, 2), "k_classify_ms": avg(st_p, "k_classify"),
                         "sample_ms": avg(st_p, "ms_sample")},
               "kernels_ms": {w: {k: round(sum(launch_ms(x)[k] for x in sts) / len(sts), 3) for k in launch_ms(sts[0])}
                              for w, sts in (("learned", st_l), ("plain", st_p))},
               "speedup": round(ms_p / ms_l, 4),
               "imbalance": {"classification": round(lp["coarse_max"] * C / n, 4),
                             "bucket": round(lp["bucket_max"] * lp["buckets"] / n, 4)},
               "bit_equal": same}
    # a measurement build: the refinement's device work counters (not timed)
    counted = build(0, 0, d_other, caps_sa.FLAG_COUNT_WORK)
    torch.cuda.synchronize()
    same = same and bool(torch.equal(d_out, d_other))

    rooflines = rooflines_of(stats, n, 1, 4, counted, args.workload, 0)
    dom = max(rooflines, key=lambda k: rooflines[k]["kernel_ms"])
    avg_k = {k: round(sum(launch_ms(s)[k] for s in stats) / len(stats), 3) for k in launch_ms(stats[0])}
    avg_k["k_groups_small (5 launches)"] = round(sum(s["k_groups_small"] for s in stats) / len(stats), 3)
    avg_k["k_groups_mid"] = round(sum(s["k_groups_mid"] for s in stats) / len(stats), 3)

    verified = None
    if not args.no_verify:
        import oracle as O

        sa = d_out.cpu().numpy().view(np.uint32).astype(np.uint64)
        bad = O.Ctx(text).verify(sa, strict=True, canon=True, threads=args.cpu_threads)
        verified = bool(bad == -1) and same
        print(f"[bench] verify: {'ok' if bad == -1 else f'FAILED at {bad}'}, all builds equal: {same}",
              file=sys.stderr, flush=True)

    # host pointers in and out: text upload + build + SA download (never `value`)
    pcie = None
    if not args.no_plain:
        t0 = time.perf_counter()
        caps_sa.build_in_memory_with(text, None, None, np.uint32)
        pcie_ms = (time.perf_counter() - t0) * 1e3
        pcie = {"ms": round(pcie_ms, 2), "value": round(n / (pcie_ms / 1e3) / 1e6, 2),
                "what": f"caps_build_in_memory: {n} B text host->HBM, build, {4 * n} B SA HBM->host (pageable)"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(text, args.cpu_seconds, args.cpu_threads, p=1024)
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
        "config": {"workload": f"{args.workload}: build_in_memory_sample_sort, 1,024 splitters", "n": n,
                   "text": desc, "index": "u32", "router": "bsearch",
                   "classification_buckets": last["coarse_buckets"], "buckets": last["buckets"],
                   "refine": args.refine, "tile_sort": args.tile_sort, "seed": args.seed},
        "verified_bit_exact": verified,
        "learned_vs_plain": cmp,
        "imbalance": {"classification": round(last["coarse_max"] * last["coarse_buckets"] / n, 4),
                      "bucket": round(last["bucket_max"] * last["buckets"] / n, 4)},
        "phases_ms": {k: round(sum(s[k] for s in stats) / len(stats), 3)
                      for k in ("ms_alphabet", "ms_sample", "ms_classify", "ms_scatter", "ms_tile_sort", "ms_merge",
                                "ms_detect", "ms_refine", "ms_giant")},
        "kernels_ms": avg_k,
        "groups": {k: last[k] for k in ("pairs", "small_groups", "mid_groups", "giant_groups", "oversized")},
        "refine_symbols_read": {"pairs": counted["cmp_symbols"][0], "groups_3_64": counted["cmp_symbols"][1],
                                "groups_65_plus": counted["cmp_symbols"][2]},
        "autotune": tuned,
        "roofline": rooflines[dom],
        "rooflines": rooflines,
        "pcie_inclusive": pcie,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0 if verified in (None, True) else 3


if __name__ == "__main__":
    sys.exit(main())
