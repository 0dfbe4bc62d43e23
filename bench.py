"""Benchmark of the TNF contig-graph path on one MI355X.

A step is one full pass of the pair stage over the 200,000-contig synthetic
metagenome (1,000 order-3 Markov genomes, log-normal lengths median 8 kb,
sigma 1.0, clipped to [1.5 kb, 500 kb], 0.5% N): the thresholded, per-row
capped edge list (cutoff 0.95, maxEdges 200) of all n(n-1)/2 pairs, from
HBM-resident FP32 TNF rows and FP64 log sizes.  The timed mode is the BF16
MFMA Gram prefilter + FP32 refinement; the exact FP32 mode and the FP16 mode
run in the same process, and their edge lists must be identical.

value = n(n-1)/2 / (seconds per step): effective pair evaluations per second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import mdet_amd as M  # noqa: E402
from mdet_amd import workload as W  # noqa: E402

PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16/FP16 MFMA
PEAK_FP32_TFLOPS = 157.3   # MI355X FP32
FLOPS_PER_PAIR = 2 * 136   # algorithmic: one K=136 dot product per unordered pair
CUTOFF, MAX_EDGES = 0.95, 200


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_mode(ctx, w, mode, steps, warmup):
    for _ in range(warmup):
        ctx.tnf_edges(w.tnf, w.logsize, w.n, CUTOFF, MAX_EDGES, mode, fetch=False)
    ctx.sync()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        _, st = ctx.tnf_edges(w.tnf, w.logsize, w.n, CUTOFF, MAX_EDGES, mode, fetch=False)
        stats.append(st.as_dict())
    ctx.sync()
    dt = time.perf_counter() - t0
    return dt / steps, stats


def fetch_edges(ctx, w, mode):
    f, t, s, st = ctx.tnf_edges(w.tnf, w.logsize, w.n, CUTOFF, MAX_EDGES, mode)
    return f, t, s, st


def cpu_baseline(tnf, ls, budget_s=12.0):
    """The oracle (C restatement of the reference pair loop, gen_tnf_graph's
    per-row scan over all columns) on the host cores, on a row sample."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    n = tnf.shape[0]
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    # calibrate on one thread
    t0 = time.perf_counter()
    O.edges(tnf, ls, CUTOFF, MAX_EDGES, rows=(0, 8))
    per_row = (time.perf_counter() - t0) / 8
    rows_per_thread = max(4, int(budget_s / per_row))
    total_rows = min(n, rows_per_thread * threads)
    chunks = np.array_split(np.arange(total_rows), threads)
    out = [None] * threads

    def work(k):
        c = chunks[k]
        if len(c):
            out[k] = O.edges(tnf, ls, CUTOFF, MAX_EDGES, rows=(int(c[0]), int(c[-1]) + 1))

    th = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    dt = time.perf_counter() - t0
    pairs = total_rows * (n - 1)
    edges = (np.concatenate([o[0] for o in out if o is not None]), np.concatenate([o[1] for o in out if o is not None]))
    return {"value": pairs / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"rows 0..{total_rows - 1} of {n} against all {n} columns ({pairs:.3g} ordered pairs, "
                      f"{dt:.1f} s), oracle/tnf_oracle.c per-row scan of gen_tnf_graph semantics"}, edges, total_rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3_200k", choices=sorted(W.CONFIGS))
    ap.add_argument("--n", type=int, default=None, help="override the contig count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="exact_f32,f16_refine")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    ctx = M.Ctx(0)
    log(f"generating {args.workload}")
    w = W.make(ctx, args.workload, n=args.n, seed=1)
    log(f"{w.n} contigs, {w.n_bases} bases, generated in {w.gen_s:.1f} s, counted in {w.count_ms:.1f} ms")
    P = w.n * (w.n - 1) // 2

    # headline: BF16 MFMA prefilter + FP32 refine
    log("timing bf16_refine")
    sec, stats = run_mode(ctx, w, M.BF16_REFINE, args.steps, args.warmup)
    pf_ms = float(np.mean([s["ms_prefilter"] for s in stats]))
    modes = {"bf16_refine": {"pairs_per_s": P / sec, "ms_per_step": sec * 1e3,
                             "stage_ms": {k: float(np.mean([s[k] for s in stats]))
                                          for k in ("ms_prefilter", "ms_refine", "ms_select", "ms_total")},
                             "survivors": stats[-1]["survivors"], "candidates": stats[-1]["candidates"],
                             "adjudicated_fp64": stats[-1]["adjudicated"]}}
    fb, tb, sb, stb = fetch_edges(ctx, w, M.BF16_REFINE)
    modes["bf16_refine"]["edges"] = int(len(fb))

    identical = {}
    for name in [m for m in args.modes.split(",") if m]:
        mode = M.PRECISIONS[name]
        steps = max(1, args.steps // 3)
        log(f"timing {name} ({steps} steps)")
        sec_m, st_m = run_mode(ctx, w, mode, steps, 1)
        f, t, s, _ = fetch_edges(ctx, w, mode)
        same = bool(np.array_equal(f, fb) and np.array_equal(t, tb))
        identical[name] = same
        modes[name] = {"pairs_per_s": P / sec_m, "ms_per_step": sec_m * 1e3,
                       "stage_ms": {k: float(np.mean([x[k] for x in st_m]))
                                    for k in ("ms_prefilter", "ms_refine", "ms_select", "ms_total")},
                       "edges": int(len(f)), "edge_set_identical_to_bf16": same,
                       "max_abs_s_diff_vs_bf16": float(np.max(np.abs(s - sb))) if same and len(s) else None}
    if "exact_f32" in modes:
        modes["bf16_refine"]["speedup_vs_exact_f32"] = modes["bf16_refine"]["pairs_per_s"] / modes["exact_f32"]["pairs_per_s"]

    achieved = FLOPS_PER_PAIR * P / (pf_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "kernel": "k_prefilter_mfma<bf16>", "achieved": achieved, "peak": PEAK_BF16_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None,
                "flops_per_launch": FLOPS_PER_PAIR * P, "avg_launch_ms": pf_ms}

    cpu = None
    parity = {}
    if not args.no_cpu:
        log("cpu baseline (oracle, row sample)")
        tnf, ls = w.tnf.numpy(), w.logsize.numpy()
        cpu, (fo, to), rows = cpu_baseline(tnf, ls)
        m = fb < rows
        parity = {"oracle_rows_checked": rows,
                  "edges_in_rows_equal_oracle": bool(set(zip(fo.tolist(), to.tolist())) ==
                                                     set(zip(fb[m].tolist(), tb[m].tolist())))}

    line = {
        "metric": "TNF pair evals/s: exact vs BF16-refine vs learned filter at >=99.9% edge recall",
        "value": P / sec,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.n} contigs, {W.CONFIGS[args.workload]['n_genomes']} genomes, "
                               f"{w.n_bases} bases; cutoff {CUTOFF}, maxEdges {MAX_EDGES}; step = full edge list "
                               f"of {P} pairs (bf16_refine)",
                   "n_contigs": w.n, "pairs": P, "cutoff": CUTOFF, "max_edges": MAX_EDGES,
                   "count_ms": w.count_ms, "count_bases_per_s": w.n_bases / (w.count_ms * 1e-3)},
        "modes": modes,
        "edge_sets_identical": identical,
        "parity": parity,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    ctx.close()


if __name__ == "__main__":
    main()
