"""The benchmark workloads under the GPU parity gate, at sizes the oracle
finishes in seconds.

- Config 1 at full size (10,000 proteins in 200 families): makeDB bytes, the
  result file of `cluster` at -i 50/70/90, and both graph forms' edges.
- Config 2 (log-normal proteins, median 350, at most 5,000) at 1% of its
  families: signatures, rows, every candidate pair, and both forms' edges at
  -i 50, 70 and 90.
- Config 4 (heavy-tailed proteins, Pareto tail to 35,000) at 2% of its
  families, a sample whose tail reaches past 10,000 residues: the same checks
  at -i 90 (and 70), with the wavefront bins in use.
- Config 5: serving over a config-4 resident set at 2% of its families,
  clustered on the device; offline, replayed and one-query batches against the
  oracle's assignment.
Each test prints the case before touching the GPU.
"""
import os

import numpy as np
import pytest

import ngia2_amd as ng
import oracle_lib as orc

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def check_graph(ctx, d, rep, identity, bands=64):
    """Device sketch/bucket against the oracle, then the batched form's
    distinct pairs against the union of the oracle's candidates, and the edges
    of both forms against orc.build_graph."""
    want, tot = orc.build_graph(d.packed, d.word_off, d.lengths, rep, bands, identity)
    per_band = [orc.candidates(rep[b], d.lengths, identity) for b in range(bands)]
    union = np.unique(np.concatenate(per_band))
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.set_reps(rep, bands)
    st = db.graph(identity, bands, ng.GRAPH_BATCHED)
    keys, sc = db.pair_scores()
    np.testing.assert_array_equal(keys, union)
    assert st["total_pairs"] == tot == sum(len(c) for c in per_band)
    np.testing.assert_array_equal(db.edges(), want)
    jobs = np.zeros((len(keys), 3), dtype=np.uint32)
    jobs[:, 0] = keys >> np.uint64(32)
    jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
    thr = float(np.float32(identity) / np.float32(100))
    np.testing.assert_array_equal(sc, orc.align_jobs(d.packed, d.word_off, jobs, thr)[:, 2])
    bins = {(b["kind"], b["width"]) for b in db.bin_stats() if b["pairs"]}
    st_p = db.graph(identity, bands, ng.GRAPH_PLAIN)
    np.testing.assert_array_equal(db.edges(), want)
    assert st_p["total_pairs"] == tot
    db.close()
    return want, bins


@pytest.fixture(scope="module")
def config1(tmp_path_factory):
    d = tmp_path_factory.mktemp("config1")
    fa, db = str(d / "c1.fa"), str(d / "c1.packed")
    p = ng.gen_config(1, scale=1.0, seed=1)
    assert ng.gen_fasta(p, fa) == 10000
    assert orc.makedb(fa, db) == 0
    return fa, db, orc.read_db_numpy(db)


def test_config1_makedb(ctx, config1, tmp_path):
    fa, db, _ = config1
    out = str(tmp_path / "gpu.packed")
    say("config 1 full: makeDB on the device, 10,000 sequences")
    ctx.makedb(fa, out)
    assert open(out, "rb").read() == open(db, "rb").read()


@pytest.mark.parametrize("identity", [50, 70, 90])
def test_config1_cluster_result_file(ctx, config1, tmp_path, identity):
    """cluster -i on the full config-1 DB: the result file equals the oracle's
    pipeline (build_graph, single-thread Leiden with seed 42, save)."""
    _, db, d = config1
    res = str(tmp_path / "gpu.txt")
    say("config 1 full: cluster -i", identity)
    st = ctx.cluster(db, res, identity)
    edges, tot = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
    assert st["n_edges"] == len(edges) and st["total_pairs"] == tot
    cl = orc.leiden(edges, d["n"], orc.resolution(identity))
    want = str(tmp_path / "oracle.txt")
    assert orc.save_result(cl, db, want) == 0
    assert open(res, "rb").read() == open(want, "rb").read()
    if identity == 50:
        assert len(edges) > 10000 and len(np.unique(cl)) < d["n"] // 2


@pytest.mark.parametrize("identity", [50, 90])
def test_config1_graph_both_forms(ctx, config1, identity):
    _, _, d = config1
    p = ng.PackedDB()
    p.packed, p.word_off, p.lengths, p.n = d["packed"], d["word_off"], d["lengths"], d["n"]
    say("config 1 full: graph, both forms, -i", identity)
    check_graph(ctx, p, d["rep"], identity)


@pytest.fixture(scope="module")
def config2():
    d = ng.gen_packed(ng.gen_config(2, scale=0.01, seed=2))
    sig = orc.sketch_all(d.packed, d.word_off)
    return d, sig, orc.bucket(sig)


def test_config2_sketch_bucket(ctx, config2):
    d, sig, rep = config2
    say("config 2 at 1%: sketch + bucket", d.n)
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    db.bucket()
    np.testing.assert_array_equal(db.sigs(), sig)
    np.testing.assert_array_equal(db.reps(), rep)
    db.close()


@pytest.mark.parametrize("identity", [50, 70, 90])
def test_config2_graph(ctx, config2, identity):
    d, _, rep = config2
    say("config 2 at 1%: graph -i", identity, "records", d.n)
    want, _ = check_graph(ctx, d, rep, identity)
    assert len(want) > (1000 if identity < 90 else 50)


@pytest.fixture(scope="module")
def config4():
    d = ng.gen_packed(ng.gen_config(4, scale=0.02, seed=1))
    assert int(d.lengths.max()) > 10000 and int((d.lengths > 5000).sum()) >= 20  # the tail
    sig = orc.sketch_all(d.packed, d.word_off)
    return d, sig, orc.bucket(sig)


def test_config4_sketch_bucket(ctx, config4):
    d, sig, rep = config4
    say("config 4 at 2%: sketch + bucket", d.n, "longest", int(d.lengths.max()))
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    db.bucket()
    np.testing.assert_array_equal(db.sigs(), sig)
    np.testing.assert_array_equal(db.reps(), rep)
    db.close()


@pytest.mark.parametrize("identity", [70, 90])
def test_config4_graph(ctx, config4, identity):
    d, _, rep = config4
    say("config 4 at 2%: graph -i", identity, "records", d.n)
    want, bins = check_graph(ctx, d, rep, identity)
    assert len(want) > 500
    assert any(k == "wave" for k, _ in bins) and any(k == "window" for k, _ in bins), bins


@pytest.fixture(scope="module")
def split_ctx():
    """A context that splits heads of 16 records or more, whose own band
    windows span 32 words or more (the samples here hold a few dozen to a few
    hundred such records), and aligns windows of 17-60 words with two to four
    words per lane whatever a bin's size: the head round's pairs of 32-60
    words then go to the segmented bins too."""
    os.environ["NGIA_HEAD_MIN"] = "16"
    os.environ["NGIA_HEAD_WINDOW"] = "32"
    os.environ["NGIA_SEG2_MIN"] = "0"
    try:
        c = ng.Context(0)
    finally:
        del os.environ["NGIA_HEAD_MIN"]
        del os.environ["NGIA_HEAD_WINDOW"]
        del os.environ["NGIA_SEG2_MIN"]
    yield c
    c.close()


@pytest.mark.parametrize("cfg,identity", [(4, 90), (4, 70), (2, 50), (1, 90)])
def test_step_split_equals_stages(split_ctx, config2, config4, config1, cfg, identity):
    """ngia_db_step: the head's round (records whose own window exceeds 31
    words) beside the rest's sketch, then the main round from the head's end.
    Signatures, rows and edges equal the oracle's and those of the three
    stages run one after another; a second step (replayed main round) agrees,
    and a plain graph() afterwards still covers every target."""
    if cfg == 1:
        d = config1[2]
        p = ng.PackedDB()
        p.packed, p.word_off, p.lengths, p.n = d["packed"], d["word_off"], d["lengths"], d["n"]
        d, sig, rep = p, orc.sketch_all(p.packed, p.word_off), d["rep"]
    else:
        d, sig, rep = config4 if cfg == 4 else config2
    want, tot = orc.build_graph(d.packed, d.word_off, d.lengths, rep, 64, identity)
    say("step, config", cfg, "-i", identity, "records", d.n)
    ctx = split_ctx
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    for split in (True, False, True):
        st = db.step(identity, 64, split=split)
        head = ctx.timing()["step_head"]
        if split and cfg != 1:
            assert head > 0
        if not split or cfg == 1:
            assert head == 0
        np.testing.assert_array_equal(db.edges(), want)
        assert st["total_pairs"] == tot
        np.testing.assert_array_equal(db.sigs(), sig)
        np.testing.assert_array_equal(db.reps(), rep)
    db.graph(identity, 64, ng.GRAPH_BATCHED)
    np.testing.assert_array_equal(db.edges(), want)
    db.close()


def test_step_rows_after_partition_overflow(ctx):
    """A step whose partitioned bucketing overflows: 9,000 copies of one
    record among 20,000 put 9,000 equal signatures into one partition of
    every band, past its 8,192 capacity, and the step's bucketing redoes the
    rows with the full sort: signatures, rows, candidate count and edges
    equal the oracle's, in two steps in a row."""
    rng = np.random.default_rng(23)
    aa = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", dtype=np.uint8)
    common = aa[rng.integers(0, 20, 120)].tobytes()
    seqs = [common] * 9000 + [aa[rng.integers(0, 20, int(rng.integers(80, 200)))].tobytes()
                              for _ in range(11000)]
    seqs = [seqs[i] for i in rng.permutation(len(seqs))]
    packed, word_off = orc.pack_many(seqs)
    lengths = np.array([len(x) for x in seqs], dtype=np.uint32)
    sig = orc.sketch_all(packed, word_off)
    rep = orc.bucket(sig)
    want, tot = orc.build_graph(packed, word_off, lengths, rep, 64, 90)
    assert len(want) >= 8999
    say("step with a partition overflow, records", len(seqs))
    db = ng.Database(ctx, packed, word_off, lengths)
    for _ in range(2):
        st = db.step(90, 64)
        np.testing.assert_array_equal(db.sigs(), sig)
        np.testing.assert_array_equal(db.reps(), rep)
        np.testing.assert_array_equal(db.edges(), want)
        assert st["total_pairs"] == tot
    db.close()


def test_config5_serving(ctx):
    """Serving over a config-4 resident set (2% of the families, with its
    Pareto tail) clustered on the device (batched graph + Leiden): 3,000
    queries, 70% mutated residents (0-30% substitutions) and 30% novel, in
    requests of log-uniform size 1..4096 with Poisson arrivals.  The replayed
    stream, the stream cut into batches of one query, and the offline call
    all equal the oracle's assignment at -i 90."""
    d = ng.gen_packed(ng.gen_config(5, scale=0.02, seed=1))
    assert int(d.lengths.max()) > 10000
    say("config 5: resident clustering", d.n)
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    db.bucket()
    db.graph(90, 64, ng.GRAPH_BATCHED)
    part = ng.leiden(db.edges(), d.n, ng.resolution(90))
    new_base = int(part.max()) + 1
    q = ng.gen_queries(d.packed, d.word_off, 3000, mutated=0.7, sub=(0.0, 0.3), indel=0.01,
                       len_median=300.0, len_sigma=0.5, len_min=30, len_max=5000, seed=5)
    strs = [q.seq(k) for k in range(q.n)]
    want = orc.serve_assign(d.packed, d.word_off, db.sigs(), part, strs, 90, new_base)
    rng = np.random.default_rng(5)
    sizes, left = [], q.n
    while left:
        k = int(min(max(np.floor(np.exp(rng.uniform(0.0, np.log(4097.0)))), 1), 4096, left))
        sizes.append(k)
        left -= k
    count = np.array(sizes)
    first = np.concatenate([[0], np.cumsum(count)[:-1]])
    arrival = np.cumsum(rng.exponential(0.002, len(count)))
    rows = lambda a: np.stack([a["cluster"], a["resident"], a["score"], a["candidates"]], axis=1)
    sv = ng.Serve(db, part, 90, new_base)
    say("config 5: offline", q.n)
    got, _ = sv.assign(q)
    np.testing.assert_array_equal(rows(got), want)
    sv.reset()
    say("config 5: replay", len(count), "requests")
    got, lat, st = sv.replay(q, arrival, first, count)
    np.testing.assert_array_equal(rows(got), want)
    assert (lat > 0).all()
    sv.close()
    sv = ng.Serve(db, part, 90, new_base, lanes=2, max_queries=1)
    say("config 5: batches of one query")
    got, _, st = sv.replay(q, np.zeros(q.n), np.arange(q.n), np.ones(q.n, dtype=np.int64))
    np.testing.assert_array_equal(rows(got), want)
    assert st["batches"] == q.n
    sv.close()
    kinds = q.kind.astype(bool)
    assert (want[kinds, 1] != 0xFFFFFFFF).mean() > 0.1  # mutated members find clusters
    assert (want[~kinds, 1] == 0xFFFFFFFF).mean() > 0.9  # novel queries mostly open them
    db.close()
