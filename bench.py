#!/usr/bin/env python3
"""bench.py -- the nGIA2 clustering hot path on one MI355X.

One step = one pass of the hot path over the resident database: minhash
sketch of every sequence, band bucketing (64 bands), then the graph build at
the given identity (candidates, alignment of every candidate pair, sorted
unique edge list).  The database (packed records, offsets, lengths) is
uploaded before the timed region.  Workload: config 2 of the benchmark suite,
1,000,000 synthetic proteins in 50,000 families, at -i 90.

Prints one JSON line.  `value` is sequences through the whole hot path per
second in the batched form; the plain one-launch-per-band form is measured
on the same GPU and reported beside it.  `cpu_baseline` times the oracle (the
reference algorithm's CPU restatement, OpenMP) on a bounded sample of the
same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, public figure


def valu_peak_tops(info):
    # int32 VALU: 64 lanes per CU per clock
    return info["cu_count"] * 64 * info["clock_khz"] * 1e3 / 1e12


def sketch_work(lengths):
    """Algorithmic work of the sketch: 17 int32 ops per (window, band), 64 bands,
    max(L-7, 1) windows for L >= 1; bytes = packed record read + 256 B written."""
    import numpy as np
    L = lengths.astype(np.int64)
    windows = np.where(L >= 8, L - 7, np.where(L >= 1, 1, 0)).sum()
    ops = 17 * 64 * int(windows)
    nbytes = int((4 * (1 + 4 * ((L + 31) // 32))).sum()) + 256 * len(L)
    return ops, nbytes


def cpu_baseline(config, scale, seed, identity):
    """The oracle's sketch + bucket + graph over a config sample, all host threads."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ngia2_amd as ng
    import oracle_lib as orc
    threads = len(os.sched_getaffinity(0))
    d = ng.gen_packed(ng.gen_config(config, scale, seed))
    t0 = time.perf_counter()
    sig = orc.sketch_all(d.packed, d.word_off, threads)
    rep = orc.bucket(sig, threads)
    edges, tot = orc.build_graph(d.packed, d.word_off, d.lengths, rep, 64, identity, 1, threads)
    dt = time.perf_counter() - t0
    return {"value": d.n / dt, "unit": "seqs/s", "cores": threads, "kind": "port",
            "sample": "config %d scaled to %d sequences (%g of the families), seed %d, -i %d: "
                      "sketch + bucket + 64-band graph, %.1f s, %d candidate pairs, %d edges"
                      % (config, d.n, scale, seed, identity, dt, tot, len(edges))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--identity", type=int, default=90)
    ap.add_argument("--cpu-scale", type=float, default=1.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the hot path runs on a single MI355X in this "
              "build (--gpus %d / WORLD_SIZE %d requested)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ngia2_amd as ng

    t0 = time.perf_counter()
    d = ng.gen_packed(ng.gen_config(args.config, args.scale, args.seed))
    gen_s = time.perf_counter() - t0
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    log("generated %d sequences in %.1f s" % (d.n, gen_s))

    ctx = ng.Context(0)
    info = ctx.info()
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)

    def step(mode):
        db.sketch()
        t_s = ctx.timing()["sketch_ms"]
        db.bucket()
        t_b = ctx.timing()["bucket_ms"]
        st = db.graph(args.identity, 64, mode)
        tm = ctx.timing()
        return {"sketch_ms": t_s, "bucket_ms": t_b, "graph_ms": tm["graph_ms"],
                "finalize_ms": tm["finalize_ms"], "stats": st}

    def run(mode, steps, warmup):
        for _ in range(warmup):
            step(mode)
        ctx.sync()
        rec = []
        t0 = time.perf_counter()
        for _ in range(steps):
            rec.append(step(mode))
        ctx.sync()
        dt = time.perf_counter() - t0
        return dt, rec

    dt, rec = run(ng.GRAPH_BATCHED, args.steps, args.warmup)
    log("batched: %.2f ms/step" % (1e3 * dt / args.steps))
    mean = lambda k: float(np.mean([r[k] for r in rec]))
    st = rec[-1]["stats"]
    sk_ms, bk_ms, gr_ms, fin_ms = mean("sketch_ms"), mean("bucket_ms"), mean("graph_ms"), \
        mean("finalize_ms")
    ms_step = 1e3 * dt / args.steps
    value = d.n * args.steps / dt

    # dominant kernel: the sketch (one kernel, timed with HIP events on the context stream)
    ops, nbytes = sketch_work(d.lengths)
    peak = valu_peak_tops(info)
    achieved = ops / (sk_ms * 1e-3) / 1e12
    roofline = {"kernel": "k_sketch", "bound": "valu", "achieved": round(achieved, 3),
                "peak": round(peak, 3), "unit": "TOP/s (int32)",
                "frac": round(achieved / peak, 4), "traffic": None,
                "hbm_gbps": round(nbytes / (sk_ms * 1e-3) / 1e9, 2),
                "hbm_frac": round(nbytes / (sk_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5),
                "algorithmic_ops": ops, "algorithmic_bytes": nbytes,
                "kernel_ms": round(sk_ms, 4)}

    out = {"metric": "seqs clustered/s + aligned pairs/s at -i %d (1 GPU); p50/p99 stream latency"
                     % args.identity,
           "value": round(value, 1), "unit": "seqs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic (seeded generator, config %d)" % args.config,
           "config": {"workload": "config %d: %d proteins in %d families, median length %g, -i %d, "
                                  "64 bands, batched form" % (args.config, d.n,
                                                              ng.gen_config(args.config,
                                                                            args.scale).n_families,
                                                              ng.gen_config(args.config).len_a,
                                                              args.identity),
                      "sequences": d.n, "identity": args.identity, "bands": 64,
                      "residues": int(d.lengths.astype(np.int64).sum()), "seed": args.seed},
           "aligned_pairs_per_s": round(st["total_pairs"] / ((gr_ms + fin_ms) * 1e-3), 1),
           "distinct_pairs_per_s": round(st["distinct_pairs"] / ((gr_ms + fin_ms) * 1e-3), 1),
           "pairs": {"candidates_with_duplicates": st["total_pairs"],
                     "distinct": st["distinct_pairs"], "edges": st["n_edges"],
                     "ref_tiles": st["ref_tiles"]},
           "stage_ms": {"sketch": round(sk_ms, 3), "bucket": round(bk_ms, 3),
                        "graph": round(gr_ms, 3), "finalize": round(fin_ms, 3)},
           "roofline": roofline,
           "device": {"name": info["name"], "arch": info["arch"], "cu": info["cu_count"],
                      "clock_mhz": info["clock_khz"] / 1e3}}

    if not args.no_plain:
        pdt, prec = run(ng.GRAPH_PLAIN, max(1, min(args.steps, 3)), 1)
        psteps = max(1, min(args.steps, 3))
        pv = d.n * psteps / pdt
        out["plain_form"] = {"value": round(pv, 1), "ms_per_step": round(1e3 * pdt / psteps, 3),
                             "graph_ms": round(float(np.mean([r["graph_ms"] for r in prec])), 3),
                             "host_syncs": prec[-1]["stats"]["host_syncs"],
                             "align_launches": prec[-1]["stats"]["align_launches"]}
        out["batched_vs_plain"] = round(value / pv, 3)
        log("plain: %.2f ms/step" % (1e3 * pdt / psteps))

    db.close()
    ctx.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.config, args.cpu_scale, args.seed, args.identity)
        log("cpu baseline: %.1f seqs/s" % out["cpu_baseline"]["value"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
