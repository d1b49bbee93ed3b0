#!/usr/bin/env python3
"""bench.py -- the nGIA2 clustering hot path on one MI355X.

One step = one pass of the hot path over the resident database: minhash
sketch of every sequence, band bucketing (64 bands), then the graph build at
-i 90 (candidates, alignment of every candidate pair, sorted unique edge
list).  The database (packed records, offsets, lengths) is uploaded before
the timed region.

Workload: 1,000,000 synthetic proteins in 50,000 families with heavy-tailed
lengths (log-normal body, median 300, plus a Pareto tail with alpha 1.5 up to
35,000 residues), 0-40% divergence, at -i 90 -- the set on which the batched
form is compared with the plain one-launch-per-band form.  The 1M-protein
set with a log-normal length body only (median 350, at most 5,000) is
measured the same way and reported beside it.

Prints one JSON line.  `value` is sequences through the whole hot path per
second in the batched form (captured hipGraph).  The plain form is measured
on the same GPU and reported beside it.  A second pass of the same steps in
the uncaptured batched form times every alignment kernel on its own stream
with HIP events; from it come the alignment rooflines (reference 32x32 tiles
x 128 int32 ops per tile / kernel time, against 64 lanes x CUs x clock).
`cpu_baseline` times the oracle (the reference algorithm's CPU restatement,
OpenMP over all host threads) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, public figure
OPS_PER_TILE = 128     # 32 rep-word x 32 target-char updates, 4 int32 ops each


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def valu_peak_tops(info):
    """int32 VALU peak: 128 lane-operations per CU per clock (two 64-lane VALU
    instructions), the rate of the published 157.3 TFLOP/s vector FP32 figure
    counted per operation.  One 64-lane instruction per CU-clock is not a
    ceiling on this device: k_sketch executes 43 T VALU lane-ops/s
    (SQ_INSTS_VALU) and v_xor_b32 alone reaches 68 T."""
    return info["cu_count"] * 128 * info["clock_khz"] * 1e3 / 1e12


def sketch_work(lengths):
    """Algorithmic work of the sketch: 17 int32 ops per (window, band), 64 bands,
    max(L-7, 1) windows for L >= 1; bytes = packed record read + 256 B written."""
    import numpy as np
    L = lengths.astype(np.int64)
    windows = np.where(L >= 8, L - 7, np.where(L >= 1, 1, 0)).sum()
    ops = 17 * 64 * int(windows)
    nbytes = int((4 * (1 + 4 * ((L + 31) // 32))).sum()) + 256 * len(L)
    return ops, nbytes


def kernel_name(b):
    if b["kind"] == "window":
        return "k_align_win<%d>" % b["width"]
    if b["kind"] == "wave":
        return "k_align_wave<%d>" % b["width"]
    if b["kind"] == "seg":
        return "k_align_seg<%d>" % b["width"]
    if b["kind"] in ("seg2", "seg3", "seg4"):
        return "k_align_seg<%d, %s>" % (b["width"], b["kind"][3])
    return "k_align_plain"


TRAFFIC_SOURCE = ("not measured in this run: HBM bytes per launch from the committed "
                  "profiles/traffic.json, two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of "
                  "the same workload scaled by a 1 GiB copy; null when the kernel or workload "
                  "has no entry there")


def load_traffic():
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by a plain copy of known size), keyed by kernel and workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def copy_probe(nbytes):
    """One device-to-device copy of nbytes: a kernel of known HBM traffic (nbytes
    read + nbytes written), for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy probe: %d bytes device to device" % nbytes)


def measure(ng, np, ctx, info, d, identity, steps, warmup, plain_steps, label, split=True):
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)

    def step(mode):
        if mode == "step":  # sketch + bucketing + graph in one call, head round split off
            st = db.step(identity, 64, split=split)
            tm = ctx.timing()
            return {"step_ms": tm["step_ms"], "head": tm["step_head"], "stats": st}
        db.sketch()
        t_s = ctx.timing()["sketch_ms"]
        db.bucket()
        t_b = ctx.timing()["bucket_ms"]
        st = db.graph(identity, 64, mode)
        tm = ctx.timing()
        r = {"sketch_ms": t_s, "bucket_ms": t_b, "graph_ms": tm["graph_ms"],
             "finalize_ms": tm["finalize_ms"], "align_ms": tm["align_ms"], "stats": st}
        if mode == ng.GRAPH_BATCHED_EAGER:
            r["bins"] = db.bin_stats()
        return r

    def run(mode, k, w):
        for _ in range(w):
            step(mode)
        ctx.sync()
        rec = []
        t0 = time.perf_counter()
        for _ in range(k):
            rec.append(step(mode))
        ctx.sync()
        return time.perf_counter() - t0, rec

    out = {}
    mean = lambda rs, k: float(np.mean([r[k] for r in rs]))
    # the timed step: ngia_db_step, the head's round beside the rest's sketch
    sdt, srec = run("step", steps, warmup)
    log("%s step: %.2f ms/step, head %d records" % (label, 1e3 * sdt / steps, srec[-1]["head"]))
    step_edges = db.edges()
    out["dt"], out["steps"] = sdt, steps
    out["value"] = d.n * steps / sdt
    out["ms_per_step"] = 1e3 * sdt / steps
    out["stats"] = srec[-1]["stats"]
    out["step"] = {"head_records": srec[-1]["head"],
                   "device_ms_per_step": round(mean(srec, "step_ms"), 3)}
    # the same work as three calls (sketch, bucket, captured graph), per stage
    dt, rec = run(ng.GRAPH_BATCHED, steps, warmup)
    log("%s batched: %.2f ms/step" % (label, 1e3 * dt / steps))
    edges = db.edges()  # the captured form's sorted unique edge list of the last step
    out["step"]["unsplit_ms_per_step"] = round(1e3 * dt / steps, 3)
    out["step"]["edges_identical"] = bool(np.array_equal(step_edges, edges))
    for k in ("sketch_ms", "bucket_ms", "graph_ms", "finalize_ms"):
        out[k] = mean(rec, k)

    # the same steps, uncaptured, with every alignment kernel timed on its stream
    edt, erec = run(ng.GRAPH_BATCHED_EAGER, steps, 1)
    log("%s eager batched: %.2f ms/step" % (label, 1e3 * edt / steps))
    out["eager_ms_per_step"] = 1e3 * edt / steps
    out["align_ms"] = mean(erec, "align_ms")
    bins = erec[-1]["bins"]
    for i, b in enumerate(bins):
        b["ms"] = float(np.mean([r["bins"][i]["ms"] for r in erec]))
    out["eager_edges_identical"] = bool(np.array_equal(db.edges(), edges))
    # one more round with the tally on: the reference tiles of the blocks each
    # bin's kernel ran (a pair that can no longer pass ends early), untimed
    db.graph(identity, 64, ng.GRAPH_BATCHED_TALLY)
    for b, t in zip(bins, db.bin_stats()):
        b["run_tiles"] = t["run_tiles"]
    out["tally_edges_identical"] = bool(np.array_equal(db.edges(), edges))
    out["bins"] = bins
    out["edges"] = edges

    if plain_steps:
        pdt, prec = run(ng.GRAPH_PLAIN, plain_steps, 1)
        log("%s plain: %.2f ms/step" % (label, 1e3 * pdt / plain_steps))
        out["plain"] = {"value": round(d.n * plain_steps / pdt, 1),
                        "ms_per_step": round(1e3 * pdt / plain_steps, 3),
                        "steps": plain_steps,
                        "graph_ms": round(mean(prec, "graph_ms"), 3),
                        "host_syncs": prec[-1]["stats"]["host_syncs"],
                        "align_launches": prec[-1]["stats"]["align_launches"],
                        "edges_identical": bool(np.array_equal(db.edges(), edges))}
    db.close()
    return out


def offline_recluster(ng, params, q, identity):
    """Offline re-clustering of residents + queries, measured: the residents'
    FASTA (the same generator) with the queries appended, through the command
    line (makeDB, then cluster with the graph on the GPU, Leiden and the result
    file on the host)."""
    import shutil
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="ngia2_recluster_")
    try:
        fa, db, res = (os.path.join(tmp, x) for x in ("all.fa", "all.packed", "out.txt"))
        n_res = ng.gen_fasta(params, fa)
        with open(fa, "ab") as f:
            for k in range(q.n):
                f.write(b">query%d\n" % k + q.seq(k) + b"\n")
        exe = os.path.join(ng.BIN_DIR, "ngia2")
        t0 = time.perf_counter()
        r1 = subprocess.run([exe, "-m", "makedb", "-f", fa, "-p", db], capture_output=True,
                            text=True, timeout=600)
        t1 = time.perf_counter()
        r2 = subprocess.run([exe, "-m", "clustering", "-p", db, "-r", res, "-i", str(identity)],
                            capture_output=True, text=True, timeout=600)
        t2 = time.perf_counter()
        if r1.returncode or r2.returncode:
            raise RuntimeError("re-clustering failed: %s %s" % (r1.stdout[-400:], r2.stdout[-400:]))
        with open(res, "rb") as f:
            clusters = sum(line.startswith(b">") for line in f)
        return {"sequences": n_res + q.n, "makedb_s": round(t1 - t0, 3),
                "cluster_s": round(t2 - t1, 3), "ms": round(1e3 * (t2 - t0), 1),
                "clusters": clusters,
                "basis": "measured: makeDB + cluster command lines over residents + queries "
                         "(two processes, files on local disk)"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def serving_bench(ng, np, ctx, d, params, identity, n_queries, seed):
    """Serving over the resident set: cluster it offline (batched graph + host
    Leiden), then assign a stream of query batches (70% mutated residents at
    0-30% divergence, 30% novel; batch sizes log-uniform over 1..4096) with
    Poisson arrivals at several offered loads.  Every served query joins the
    snapshot, so each run starts a new session (reset) over the residents."""
    t0 = time.perf_counter()
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    db.bucket()
    db.graph(identity, 64, ng.GRAPH_BATCHED)
    edges = db.edges()
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    part = ng.leiden(edges, d.n, ng.resolution(identity))
    t_leiden = time.perf_counter() - t0
    new_base = int(part.max()) + 1
    log("serving: resident clustering %.2f s (graph) + %.2f s (Leiden), %d clusters"
        % (t_graph, t_leiden, new_base))
    q = ng.gen_queries(d.packed, d.word_off, n_queries, mutated=0.7, sub=(0.0, 0.3), indel=0.01,
                       len_median=300.0, len_sigma=0.5, len_min=30, len_max=5000, seed=seed)
    rng = np.random.default_rng(seed)
    sizes = []
    left = q.n
    while left > 0:
        k = int(min(max(np.floor(np.exp(rng.uniform(0.0, np.log(4097.0)))), 1), 4096, left))
        sizes.append(k)
        left -= k
    count = np.array(sizes, dtype=np.int64)
    first = np.concatenate([[0], np.cumsum(count)[:-1]])
    sv = ng.Serve(db, part, identity, new_base)
    sv.assign(q)  # warm-up
    sv.reset()
    t0 = time.perf_counter()
    offline, st_off = sv.assign(q)
    t_off = time.perf_counter() - t0
    sat_qps = q.n / t_off
    log("serving: offline assignment of %d queries in %.1f ms (%.0f queries/s)"
        % (q.n, 1e3 * t_off, sat_qps))
    loads = []
    identical = True
    for rho in (0.25, 0.5, 0.8):
        rate = rho * sat_qps / float(count.mean())  # requests per second
        arrival = np.cumsum(rng.exponential(1.0 / rate, len(count)))
        arrival -= arrival[0]
        sv.reset()
        got, lat, st = sv.replay(q, arrival, first, count)
        same = bool(np.array_equal(got, offline))
        identical &= same
        loads.append({"offered_load": rho, "request_rate_per_s": round(rate, 1),
                      "queries_per_s": round(q.n / st["wall_s"], 1),
                      "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
                      "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
                      "max_ms": round(1e3 * float(lat.max()), 3),
                      "batches": st["batches"], "device_s": round(st["device_s"], 4),
                      "wall_s": round(st["wall_s"], 4), "identical_to_offline": same})
        log("serving: load %.2f: p50 %.2f ms, p99 %.2f ms, %d batches"
            % (rho, loads[-1]["p50_ms"], loads[-1]["p99_ms"], st["batches"]))
    sv.close()
    db.close()
    kinds = q.kind.astype(bool)
    out = {"workload": "resident: the benchmark set clustered at -i %d (batched graph + Leiden, "
                       "%d clusters); %d queries, 70%% mutated residents (0-30%% substitutions, 1%% "
                       "indels), 30%% novel (log-normal lengths, median 300); %d requests, "
                       "log-uniform sizes 1-4096, Poisson arrivals" % (identity, new_base, q.n,
                                                                         len(count)),
           "residents": d.n, "queries": q.n, "requests": len(count),
           "lanes": st_off["lanes"], "max_queries_per_batch": st_off["max_queries"],
           "max_residues_per_batch": st_off["max_residues"],
           "saturated_queries_per_s": round(sat_qps, 1),
           "offline_assign_ms": round(1e3 * t_off, 3),
           "loads": loads, "assignments_identical_to_offline": identical,
           "new_clusters": int((offline["resident"] == ng.NO_RESIDENT).sum()),
           "joined_earlier_queries": int(((offline["resident"] != ng.NO_RESIDENT) &
                                          (offline["resident"] >= d.n)).sum()),
           "mutated_joined": round(float((offline["resident"][kinds] != ng.NO_RESIDENT).mean()), 4)
           if kinds.any() else None,
           "resident_clustering_s": {"graph": round(t_graph, 3), "leiden": round(t_leiden, 3)},
           "offline_recluster": offline_recluster(ng, params, q, identity)}
    out["offline_recluster"]["serving_speedup"] = round(
        out["offline_recluster"]["ms"] / (1e3 * t_off), 2)
    return out


def align_rooflines(m, peak):
    """Per alignment kernel and for the whole phase: `achieved` counts the
    reference tiles of the blocks the kernels ran (run_tiles, from the tally
    round; an early exit skips the rest of a failing pair), so `frac` is the
    share of the VALU peak doing that work.  `ref_rate` counts every
    reference tile of the pairs, the work the reference's loop does for them:
    with early exits it can exceed the peak, as it measures skipped work too."""
    rows = []
    for b in m["bins"]:
        if b["kind"] not in ("window", "wave", "seg", "seg2", "seg3", "seg4", "plain") or not b["pairs"] or b["ms"] <= 0:
            continue
        sec = b["ms"] * 1e-3
        ach = b["run_tiles"] * OPS_PER_TILE / sec / 1e12
        rows.append({"kernel": kernel_name(b), "pairs": b["pairs"], "ref_tiles": b["ref_tiles"],
                     "run_tiles": b["run_tiles"], "kernel_ms": round(b["ms"], 4),
                     "achieved": round(ach, 3), "frac": round(ach / peak, 4),
                     "ref_rate": round(b["ref_tiles"] * OPS_PER_TILE / sec / 1e12, 3)})
    rows.sort(key=lambda r: -r["kernel_ms"])
    tiles = sum(b["ref_tiles"] for b in m["bins"])
    run = sum(b.get("run_tiles", 0) for b in m["bins"])
    sec = m["align_ms"] * 1e-3
    phase = run * OPS_PER_TILE / sec / 1e12 if sec > 0 else 0.0
    ref_rate = tiles * OPS_PER_TILE / sec / 1e12 if sec > 0 else 0.0
    return rows, {"ref_tiles": tiles, "run_tiles": run, "ms": round(m["align_ms"], 4),
                  "achieved": round(phase, 3), "frac": round(phase / peak, 4),
                  "ref_rate": round(ref_rate, 3)}


def chunked_graph(ng, np, ctx, d, identity, blocks):
    """The graph of the whole set computed as if it did not fit the device:
    blocks of 1/blocks of the packed area, each (rep block, target block)
    chunk uploaded and aligned alone, edges merged on the host; compared with
    the resident batched graph's edges."""
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    db.bucket()
    rep = db.reps()
    db.graph(identity, 64, ng.GRAPH_BATCHED)
    want = db.edges()
    db.close()
    words = int(len(d.packed))
    bw = (words + blocks - 1) // blocks
    t0 = time.perf_counter()
    got, st, chunks = ctx.build_graph_chunked(d.packed, d.word_off, d.lengths, rep, 64, identity,
                                              block_words=bw)
    wall = time.perf_counter() - t0
    return {"blocks": int(len(ng.plan_blocks(d.lengths, bw)) - 1), "block_words": bw,
            "chunks": chunks, "wall_s": round(wall, 3), "edges": int(len(got)),
            "edges_identical": bool(np.array_equal(got, want)),
            "note": "host-resident database, packed area cut into blocks; each chunk is "
                    "uploaded, aligned (batched form) and its edges merged on the host"}


def end_to_end(ng, config, scale, seed, identity):
    """The command line as a user runs it: FASTA -> makeDB -> packed DB ->
    cluster (graph on the GPU, Leiden and the result file on the host)."""
    import shutil
    import subprocess
    import tempfile
    tmp = tempfile.mkdtemp(prefix="ngia2_e2e_")
    try:
        fa, db, res = (os.path.join(tmp, x) for x in ("in.fa", "in.packed", "out.txt"))
        n = ng.gen_fasta(ng.gen_config(config, scale, seed), fa)
        exe = os.path.join(ng.BIN_DIR, "ngia2")
        t0 = time.perf_counter()
        r1 = subprocess.run([exe, "-m", "makedb", "-f", fa, "-p", db], capture_output=True,
                            text=True, timeout=600)
        t1 = time.perf_counter()
        r2 = subprocess.run([exe, "-m", "clustering", "-p", db, "-r", res, "-i", str(identity)],
                            capture_output=True, text=True, timeout=600)
        t2 = time.perf_counter()
        if r1.returncode or r2.returncode:
            raise RuntimeError("command line failed: %s %s" % (r1.stdout[-400:], r2.stdout[-400:]))
        # one cluster per representative line (">name" at the line start;
        # members are written indented, saveResult.hpp:31-83)
        clusters = 0
        with open(res, "rb") as f:
            for line in f:
                clusters += line.startswith(b">")
        return {"sequences": n, "fasta_bytes": os.path.getsize(fa),
                "makedb_s": round(t1 - t0, 3), "cluster_s": round(t2 - t1, 3),
                "seqs_per_s": round(n / (t2 - t0), 1),
                "cluster_seqs_per_s": round(n / (t2 - t1), 1), "clusters": clusters,
                "note": "wall time of two processes (each opens the GPU), files on local disk"}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(ng, orc, config, scale, seed, identity):
    """The oracle's sketch + bucket + graph over a sample of the workload, all host threads."""
    threads = len(os.sched_getaffinity(0))
    d = ng.gen_packed(ng.gen_config(config, scale, seed))
    t0 = time.perf_counter()
    sig = orc.sketch_all(d.packed, d.word_off, threads)
    rep = orc.bucket(sig, threads)
    edges, tot = orc.build_graph(d.packed, d.word_off, d.lengths, rep, 64, identity, 1, threads)
    dt = time.perf_counter() - t0
    return {"value": round(d.n / dt, 1), "unit": "seqs/s", "cores": threads, "kind": "port",
            "cpu_model": host_cpu(),
            "sample": "the same generator scaled to %d sequences (%g of the families), seed %d, "
                      "-i %d: sketch + bucket + 64-band graph in %.1f s, %d candidate pairs, "
                      "%d edges" % (d.n, scale, seed, identity, dt, tot, len(edges))}


def host_cpu():
    """Model name of the host CPU (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def config1_cpu_build(ng, np, orc, c1_gpu):
    """Config 1 at full size: the reference algorithm's CPU build (the oracle,
    all host threads; Leiden single-threaded, seed 42) beside the GPU path
    measured above, with both edge lists and both partitions compared."""
    threads = len(os.sched_getaffinity(0))
    d = c1_gpu["data"]
    t0 = time.perf_counter()
    sig = orc.sketch_all(d.packed, d.word_off, threads)
    rep = orc.bucket(sig, threads)
    t_sb = time.perf_counter() - t0
    res = {"cores": threads, "cpu_model": host_cpu(), "kind": "port",
           "sketch_bucket_s": round(t_sb, 4), "by_identity": {}}
    for ident in sorted(k for k in c1_gpu if k != "data"):
        edges, part, gpu_ms = c1_gpu[ident]
        t0 = time.perf_counter()
        want, tot = orc.build_graph(d.packed, d.word_off, d.lengths, rep, 64, ident, 1, threads)
        t_g = time.perf_counter() - t0
        t0 = time.perf_counter()
        cl = orc.leiden(want, d.n, orc.resolution(ident))
        t_l = time.perf_counter() - t0
        res["by_identity"][str(ident)] = {
            "graph_s": round(t_g, 4), "leiden_s": round(t_l, 4),
            "hot_path_seqs_per_s": round(d.n / (t_sb + t_g), 1),
            "gpu_hot_path_seqs_per_s": round(d.n / (gpu_ms * 1e-3), 1),
            "edges": int(len(want)), "candidates_with_duplicates": int(tot),
            "edges_identical": bool(np.array_equal(edges, want)),
            "clusters": int(cl.max()) + 1 if d.n else 0,
            "clusters_identical": bool(np.array_equal(part, cl))}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4, help="generator config of the workload")
    ap.add_argument("--also-configs", default="1:50/70/90,2:50/70/90,3:97",
                    help="further workloads reported beside the first, "
                         "config:identity[/identity...],... ('' for none)")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--identity", type=int, default=90)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--cpu-scale", type=float, default=0.1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-queries", type=int, default=300000,
                    help="queries of the serving stream over the resident set (0: skip)")
    ap.add_argument("--no-tune", action="store_true", help="skip the window-bin autotuner")
    ap.add_argument("--no-split", action="store_true",
                    help="time the step without the head round (every dispatch of a kernel then "
                         "does the same work: for counter profiles)")
    ap.add_argument("--no-e2e", action="store_true", help="skip the command-line end-to-end run")
    ap.add_argument("--chunks", type=int, default=4,
                    help="blocks of the out-of-HBM graph run (0 skips it)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="bytes of one device-to-device copy run first (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the hot path runs on a single MI355X in this "
              "build (--gpus %d / WORLD_SIZE %d requested)" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ngia2_amd as ng

    if args.copy_probe:
        copy_probe(args.copy_probe)
    ctx = ng.Context(0)
    info = ctx.info()
    peak = valu_peak_tops(info)
    tune = None
    if not args.no_tune:
        t0 = time.perf_counter()
        pts = ctx.autotune(args.identity)
        t = ctx.tuning()
        sizes = [c for c in t["class_size"] if c]

        def shapes(table):
            # per window width: the shape chosen for each bin-size class
            return {str(W): {str(sizes[k]): ng.SHAPES[table[W][k]] for k in range(len(sizes))}
                    for W in range(2, 17)}
        tune = {"identity": args.identity, "seconds": round(time.perf_counter() - t0, 3),
                "class_size": sizes, "shapes": shapes(t["shape"]),
                "shapes_compact": shapes(t["shape_compact"]),
                "points": [[p["width"], p["dna"], p["pairs"]] +
                           [round(p["ms"].get(x, 0.0), 4) for x in ng.SHAPES] for p in pts
                           if p["width"] <= 16],
                "points_columns": ["width", "dna", "pairs"] + ["%s_ms" % x for x in ng.SHAPES],
                # windows of 17..60 words, by group: several words per lane or one
                "long_windows": {"17-21": "8 lanes x 3 words, or 32 lanes x 1",
                                 "22-30": "16 lanes x 2 words, or 32 lanes x 1",
                                 "32-45": "16 lanes x 3 words, or a wavefront per pair",
                                 "46-60": "16 lanes x 4 words, or a wavefront per pair"},
                "long_shapes": {g: {str(sizes[k]): ng.LONG_SHAPES[t["long"][i][k]]
                                    for k in range(len(sizes))}
                                for i, g in enumerate(("17-21", "22-30", "32-45", "46-60"))},
                "long_points": [[p["width"], p["pairs"], round(p["ms"]["multi"], 4),
                                 round(p["ms"]["single"], 4)] for p in pts if p["width"] > 16],
                "long_points_columns": ["tuned_at_window", "pairs", "multi_ms", "single_ms"],
                "fixed": "K per wavefront bin is the smallest that holds the band window "
                         "(64K >= W + K - 1); window 31 takes 32 lanes of one word; persistent "
                         "grids are as many blocks as the CUs hold"}
        log("autotune: %.2f s, shapes %s" % (tune["seconds"], {
            W: sorted(set(v.values())) for W, v in tune["shapes"].items()}))
    traffic = load_traffic()

    t0 = time.perf_counter()
    params = ng.gen_config(args.config, args.scale, args.seed)
    d = ng.gen_packed(params)
    log("generated %d sequences (config %d) in %.1f s" % (d.n, args.config,
                                                        time.perf_counter() - t0))
    plain_steps = 0 if args.no_plain else args.plain_steps
    m = measure(ng, np, ctx, info, d, args.identity, args.steps, args.warmup, plain_steps,
                "config %d" % args.config, split=not args.no_split)
    st = m["stats"]
    workload = ("%d proteins in %d families, %s lengths (median %g, max %d), -i %d, 64 bands"
                % (d.n, params.n_families,
                   "log-normal + Pareto(%.1f) tail" % params.tail_alpha if params.tail_frac > 0
                   else "log-normal", params.len_a, params.tail_max if params.tail_frac > 0
                   else params.len_max, args.identity))

    # sketch: one kernel on the context stream
    ops, nbytes = sketch_work(d.lengths)
    sk_ms = m["sketch_ms"]
    sk_ach = ops / (sk_ms * 1e-3) / 1e12
    sketch_roof = {"kernel": "k_sketch", "bound": "valu", "achieved": round(sk_ach, 3),
                   "peak": round(peak, 3), "unit": "TOP/s int32", "frac": round(sk_ach / peak, 4),
                   "traffic": traffic.get("k_sketch", {}).get(workload),
                   "traffic_source": TRAFFIC_SOURCE,
                   "kernel_ms": round(sk_ms, 4), "algorithmic_ops": ops,
                   "algorithmic_bytes": nbytes,
                   "hbm_gbps": round(nbytes / (sk_ms * 1e-3) / 1e9, 2),
                   "hbm_frac": round(nbytes / (sk_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 5)}
    align_rows, phase = align_rooflines(m, peak)
    # the dominant kernel: the one with the most device time per step
    roofline = sketch_roof
    if align_rows and align_rows[0]["kernel_ms"] > sk_ms:
        a = align_rows[0]
        roofline = {"kernel": a["kernel"], "bound": "valu", "achieved": a["achieved"],
                    "peak": round(peak, 3), "unit": "TOP/s int32", "frac": a["frac"],
                    "traffic": traffic.get(a["kernel"], {}).get(workload),
                    "traffic_source": TRAFFIC_SOURCE,
                    "kernel_ms": a["kernel_ms"], "algorithmic_ops": a["run_tiles"] * OPS_PER_TILE,
                    "ref_tiles": a["ref_tiles"], "run_tiles": a["run_tiles"],
                    "ref_rate": a["ref_rate"], "pairs": a["pairs"]}

    out = {"metric": "seqs clustered/s + aligned pairs/s at -i %d (1 GPU); p50/p99 stream latency"
                     % args.identity,
           "value": round(m["value"], 1), "unit": "seqs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(m["ms_per_step"], 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic (seeded generator, config %d, seed %d)" % (args.config, args.seed),
           "config": {"workload": workload, "sequences": d.n, "families": params.n_families,
                      "identity": args.identity, "bands": 64,
                      "residues": int(d.lengths.astype(np.int64).sum()),
                      "max_length": int(d.lengths.max()), "seed": args.seed,
                      "length_model": {"body": "log-normal", "median": params.len_a,
                                       "sigma": params.len_b, "min": params.len_min,
                                       "max": params.len_max,
                                       "tail_fraction_of_families": params.tail_frac,
                                       "tail_x_min": params.tail_xmin,
                                       "tail_alpha": params.tail_alpha,
                                       "tail_max": params.tail_max},
                      "divergence": {"substitutions": [params.sub_lo, params.sub_hi],
                                     "indels": [params.indel_lo, params.indel_hi]},
                      "form": "one ngia_db_step per step: sketch, bucketing and the batched "
                              "graph (length-binned kernels, captured hipGraph), with the head "
                              "round (records whose own band window spans 64 words or more) beside "
                              "the rest's sketch"},
           "value_covers": "sequences per second through the device hot path: sketch, band "
                           "bucketing and the graph up to the sorted unique edge list on the "
                           "device, database resident in HBM; host Leiden and the files are "
                           "outside it (end_to_end times the command line with them)",
           "aligned_pairs_per_s": round(st["total_pairs"] / ((m["graph_ms"] + m["finalize_ms"])
                                                              * 1e-3), 1),
           "distinct_pairs_per_s": round(st["distinct_pairs"] / ((m["graph_ms"] + m["finalize_ms"])
                                                                  * 1e-3), 1),
           "pairs": {"candidates_with_duplicates": st["total_pairs"],
                     "distinct": st["distinct_pairs"], "edges": st["n_edges"],
                     "ref_tiles": st["ref_tiles"]},
           "stage_ms": {"sketch": round(m["sketch_ms"], 3), "bucket": round(m["bucket_ms"], 3),
                        "graph": round(m["graph_ms"], 3), "finalize": round(m["finalize_ms"], 3),
                        "note": "the three stages called one after another (unsplit)"},
           "step_split": m["step"],
           "roofline": roofline,
           "sketch_roofline": sketch_roof,
           "align_phase": dict(phase, peak=round(peak, 3), unit="TOP/s int32",
                               eager_ms_per_step=round(m["eager_ms_per_step"], 3),
                               eager_edges_identical=m["eager_edges_identical"],
                               tally_edges_identical=m["tally_edges_identical"]),
           "align_kernels": align_rows,
           "autotune": tune,
           "device": {"name": info["name"], "arch": info["arch"], "cu": info["cu_count"],
                      "clock_mhz": info["clock_khz"] / 1e3}}
    if "plain" in m:
        out["plain_form"] = m["plain"]
        out["batched_vs_plain"] = round(m["value"] / m["plain"]["value"], 3)

    if args.chunks > 1:
        out["out_of_hbm"] = chunked_graph(ng, np, ctx, d, args.identity, args.chunks)
        log("out of HBM: %d blocks, %d chunks, %.2f s, edges identical %s" % (
            out["out_of_hbm"]["blocks"], out["out_of_hbm"]["chunks"], out["out_of_hbm"]["wall_s"],
            out["out_of_hbm"]["edges_identical"]))

    if args.serve_queries:
        sv = serving_bench(ng, np, ctx, d, params, args.identity, args.serve_queries,
                           args.seed + 100)
        out["serving"] = sv
        mid = sv["loads"][1]
        out["latency_ms"] = {"p50": mid["p50_ms"], "p99": mid["p99_ms"],
                             "offered_load": mid["offered_load"],
                             "queries_per_s": mid["queries_per_s"]}

    c1_gpu = {}
    tuned_at = None if args.no_tune else args.identity
    for spec in [x for x in args.also_configs.split(",") if x]:
        cfg, idents = spec.split(":")
        cfg = int(cfg)
        p2 = ng.gen_config(cfg, args.scale, args.seed)
        d2 = ng.gen_packed(p2)
        kind = "DNA (ACGT)" if p2.dna else "proteins"
        lens = ("normal lengths (mean %g, sd %g, %d-%d)" % (p2.len_a, p2.len_b, p2.len_min, p2.len_max)
                if p2.len_normal else "log-normal lengths (median %g, max %d)" % (p2.len_a, p2.len_max))
        o2 = {"workload": "%d %s in %d families, %s" % (d2.n, kind, p2.n_families, lens),
              "residues": int(d2.lengths.astype(np.int64).sum()), "by_identity": {}}
        for ident in (int(v) for v in idents.split("/")):
            if tuned_at is not None and tuned_at != ident:
                # the window-bin routing is tuned at the identity measured
                ctx.autotune(ident)
                tuned_at = ident
            plain = 0 if args.no_plain else (3 if cfg != 3 else 1)
            m2 = measure(ng, np, ctx, info, d2, ident, args.steps, args.warmup, plain,
                         "config %d -i %d" % (cfg, ident), split=not args.no_split)
            rows2, phase2 = align_rooflines(m2, peak)
            r = {"identity": ident, "routing_tuned_at_identity": tuned_at,
                 "value": round(m2["value"], 1),
                 "ms_per_step": round(m2["ms_per_step"], 3),
                 "stage_ms": {"sketch": round(m2["sketch_ms"], 3),
                              "bucket": round(m2["bucket_ms"], 3),
                              "graph": round(m2["graph_ms"], 3),
                              "finalize": round(m2["finalize_ms"], 3)},
                 "pairs": {"candidates_with_duplicates": m2["stats"]["total_pairs"],
                           "distinct": m2["stats"]["distinct_pairs"],
                           "edges": m2["stats"]["n_edges"], "ref_tiles": m2["stats"]["ref_tiles"]},
                 "aligned_pairs_per_s": round(m2["stats"]["total_pairs"] /
                                              ((m2["graph_ms"] + m2["finalize_ms"]) * 1e-3), 1),
                 "align_phase": phase2, "align_kernels": rows2[:4],
                 "step_split": m2["step"],
                 "eager_edges_identical": m2["eager_edges_identical"],
                 "tally_edges_identical": m2["tally_edges_identical"]}
            ops2, nbytes2 = sketch_work(d2.lengths)
            r["sketch_hbm_gbps"] = round(nbytes2 / (m2["sketch_ms"] * 1e-3) / 1e9, 2)
            if "plain" in m2:
                r["plain_form"] = m2["plain"]
                r["batched_vs_plain"] = round(m2["value"] / m2["plain"]["value"], 3)
            if cfg == 1:
                t0 = time.perf_counter()
                part = ng.leiden(m2["edges"], d2.n, ng.resolution(ident))
                r["leiden_s"] = round(time.perf_counter() - t0, 4)
                r["clusters"] = int(part.max()) + 1 if d2.n else 0
                c1_gpu[ident] = (m2["edges"], part, m2["ms_per_step"])
            o2["by_identity"][str(ident)] = r
        out["config_%d" % cfg] = o2
        if cfg == 1:
            c1_gpu["data"] = d2
        else:
            del d2

    ctx.close()
    if not args.no_e2e:
        out["end_to_end"] = end_to_end(ng, args.config, args.scale, args.seed, args.identity)
        log("end to end: makedb %.2f s, cluster %.2f s" % (out["end_to_end"]["makedb_s"],
                                                          out["end_to_end"]["cluster_s"]))
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as orc
        out["cpu_baseline"] = cpu_baseline(ng, orc, args.config, args.cpu_scale, args.seed,
                                           args.identity)
        log("cpu baseline: %.1f seqs/s" % out["cpu_baseline"]["value"])
        if "data" in c1_gpu:
            out["config_1"]["cpu_build"] = config1_cpu_build(ng, np, orc, c1_gpu)
            log("config 1 cpu build: %s" % {k: (v["edges_identical"], v["clusters_identical"])
                                            for k, v in out["config_1"]["cpu_build"]["by_identity"].items()})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
