"""Parity of the HIP path (through the C-ABI) with the oracle, bit for bit.

Each test prints the case it runs (flushed) before touching the GPU.
"""
import os

import numpy as np
import pytest

import ngia2_amd as ng
import oracle_lib as orc
from conftest import AMINO, family_fasta, mutate

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def c1db(tmp_path_factory):
    """A config-1 style set (scaled to 40 families) as a packed DB built by the oracle."""
    d = tmp_path_factory.mktemp("c1")
    fa, db = str(d / "c1.fa"), str(d / "c1.bin")
    p = ng.gen_config(1, scale=0.2, seed=7)
    ng.gen_fasta(p, fa)
    assert orc.makedb(fa, db) == 0
    return fa, db, orc.read_db_numpy(db)


@pytest.fixture(scope="module")
def longdb(tmp_path_factory):
    """Long, close families (window widths far beyond the per-lane kernel)."""
    d = tmp_path_factory.mktemp("long")
    fa, db = str(d / "long.fa"), str(d / "long.bin")
    rng = np.random.default_rng(11)
    lengths = [900, 1500, 2600, 4100, 6000, 9000, 700, 300, 64, 31]
    family_fasta(fa, rng, lengths, members=4, sub=0.03, indel=0.01)
    assert orc.makedb(fa, db) == 0
    return fa, db, orc.read_db_numpy(db)


def edge_case_records():
    """Packed records of awkward sequences: empty, shorter than 8, 32-boundaries,
    unknown letters, lowercase."""
    rng = np.random.default_rng(3)
    seqs = [b"", b"A", b"ACDEFG", b"ACDEFGH", b"ACDEFGHI", b"XXXXXXXXXX", b"acdefghiklmn",
            b"BJOZ*-AC"]
    for L in (31, 32, 33, 63, 64, 65, 95, 96, 97, 500, 2049):
        seqs.append(bytes(AMINO[i] for i in rng.integers(20, size=L)))
    recs = [orc.pack(s) for s in seqs]
    off = np.cumsum([0] + [len(r) for r in recs[:-1]]).astype(np.uint64)
    return np.concatenate(recs).astype(np.uint32), off, seqs


def test_sketch_edge_cases(ctx):
    packed, off, _ = edge_case_records()
    say("sketch edge cases", len(off))
    got = ctx.sketch(packed, off)
    want = orc.sketch_all(packed, off)
    np.testing.assert_array_equal(got, want)


def test_sketch_c1(ctx, c1db):
    _, _, d = c1db
    say("sketch c1", d["n"])
    got = ctx.sketch(d["packed"], d["word_off"])
    want = orc.sketch_all(d["packed"], d["word_off"])
    np.testing.assert_array_equal(got, want)


def test_bucket_c1(ctx, c1db):
    _, _, d = c1db
    sig = orc.sketch_all(d["packed"], d["word_off"])
    say("bucket c1", d["n"])
    np.testing.assert_array_equal(ctx.bucket(sig), orc.bucket(sig))
    # the rep rows the oracle wrote into the DB file are the same rows
    np.testing.assert_array_equal(ctx.bucket(sig), d["rep"])


def test_bucket_wide_keys(c1db):
    """The 38-bit (band, sig) key form that n >= 2^26 needs gives the same rows."""
    _, _, d = c1db
    sig = orc.sketch_all(d["packed"], d["word_off"])
    os.environ["NGIA_BUCKET_WIDE"] = "1"
    try:
        say("bucket wide keys", d["n"])
        with ng.Context(0) as c2:
            got = c2.bucket(sig)
    finally:
        del os.environ["NGIA_BUCKET_WIDE"]
    np.testing.assert_array_equal(got, d["rep"])


def test_bucket_small_blocks():
    """The small bucketing (one block per band, the sort in LDS: a step's
    head and databases of at most 16,384 records), through NGIA_BUCKET_SMALL:
    the three block shapes and their capacities, 1 to 16,384 records, with
    heavy collisions and all-ones signatures (the sort's padding key)."""
    rng = np.random.default_rng(17)
    os.environ["NGIA_BUCKET_SMALL"] = "1"
    try:
        c2 = ng.Context(0)
    finally:
        del os.environ["NGIA_BUCKET_SMALL"]
    with c2:
        for n in (1, 2, 100, 2048, 2049, 5000, 8192, 8193, 12000, 16384):
            say("bucket small", n)
            sig = rng.integers(0, 1 << 32, size=(n, 64), dtype=np.uint64).astype(np.uint32)
            sig[:, 1] = rng.integers(0, 7, size=n)
            sig[:, 2] = 0xFFFFFFFF
            sig[::3, 3] = 0xFFFFFFFF
            np.testing.assert_array_equal(c2.bucket(sig), orc.bucket(sig))


@pytest.mark.parametrize("n", [16384, 16385])
def test_bucket_small_database_boundary(ctx, n):
    """A database at the one-block-per-band bucketing's limit (16,384
    records) and one past it (the partitioned form), in families of eight
    close copies: signatures and rows of the stage calls, and the rows,
    candidate count and edges of a step, equal the oracle's."""
    rng = np.random.default_rng(n)
    aa = np.frombuffer(AMINO, dtype=np.uint8)
    seqs = []
    while len(seqs) < n:
        anc = aa[rng.integers(0, 20, int(rng.integers(60, 160)))]
        for _ in range(min(8, n - len(seqs))):
            x = anc.copy()
            sub = rng.random(len(x)) < 0.05
            x[sub] = aa[rng.integers(0, 20, int(sub.sum()))]
            seqs.append(x.tobytes())
    packed, word_off = orc.pack_many(seqs)
    lengths = np.array([len(x) for x in seqs], dtype=np.uint32)
    sig = orc.sketch_all(packed, word_off)
    rep = orc.bucket(sig)
    want, tot = orc.build_graph(packed, word_off, lengths, rep, 64, 90)
    assert len(want) > n // 8
    say("bucket and step at the small-database boundary", n)
    db = ng.Database(ctx, packed, word_off, lengths)
    db.sketch()
    db.bucket()
    np.testing.assert_array_equal(db.sigs(), sig)
    np.testing.assert_array_equal(db.reps(), rep)
    st = db.step(90, 64)
    np.testing.assert_array_equal(db.reps(), rep)
    np.testing.assert_array_equal(db.edges(), want)
    assert st["total_pairs"] == tot
    db.close()


def test_bucket_collisions(ctx):
    # heavy signature collisions: few distinct values per band
    rng = np.random.default_rng(5)
    sig = rng.integers(0, 7, size=(3000, 64)).astype(np.uint32)
    sig[:, 5] = 0xFFFFFFFF
    say("bucket collisions", sig.shape)
    np.testing.assert_array_equal(ctx.bucket(sig), orc.bucket(sig))


def test_bucket_partition_collisions(ctx):
    """Many equal signatures inside the partitions' capacity (the sort in LDS
    orders up to ~40 copies of each value per band); a band where all 5,000
    records share one signature (one partition of 5,000 sorted in LDS, the
    largest capacity tier); and 9,000 such records, beyond the largest
    capacity (8,192), so that the partition overflows and the bucketing is
    redone with the full sort."""
    rng = np.random.default_rng(8)
    sig = rng.integers(0, 50, size=(2000, 64)).astype(np.uint32) * np.uint32(0x01000193)
    say("bucket partition collisions", sig.shape)
    np.testing.assert_array_equal(ctx.bucket(sig), orc.bucket(sig))
    for n in (5000, 9000):
        big = rng.integers(0, 1 << 32, size=(n, 64), dtype=np.uint64).astype(np.uint32)
        big[:, 3] = 12345
        say("bucket one band of equal signatures", big.shape)
        np.testing.assert_array_equal(ctx.bucket(big), orc.bucket(big))


def family_sigs(rng, n, fam=20, keep=0.7):
    """Signature rows of n records in families of `fam`: each entry is the
    family's value with probability `keep`, else random."""
    base = rng.integers(0, 1 << 32, size=((n + fam - 1) // fam, 64), dtype=np.uint64).astype(np.uint32)
    sig = base[np.arange(n) // fam]
    noise = rng.integers(0, 1 << 32, size=(n, 64), dtype=np.uint64).astype(np.uint32)
    return np.where(rng.random((n, 64)) < keep, sig, noise)


@pytest.mark.parametrize("n", [1000000, 1000001, 1500000])
def test_bucket_partitioned_large(ctx, n):
    """At a million records (8-bit partitions, one radix pass) and beyond
    (10-bit partitions, both block capacities) the rows equal those of the
    38-bit full sort."""
    sig = family_sigs(np.random.default_rng(n), n)
    say("bucket partitioned vs full sort", n)
    got = ctx.bucket(sig)
    os.environ["NGIA_BUCKET_WIDE"] = "1"
    try:
        with ng.Context(0) as c2:
            want = c2.bucket(sig)
    finally:
        del os.environ["NGIA_BUCKET_WIDE"]
    np.testing.assert_array_equal(got, want)
    assert (got != np.arange(n, dtype=np.uint32)[None, :]).mean() > 0.5


@pytest.mark.parametrize("identity", [1, 50, 70, 90, 97, 99])
def test_align_random_pairs(ctx, c1db, identity):
    _, _, d = c1db
    n = d["n"]
    rng = np.random.default_rng(identity)
    jobs = np.zeros((4000, 3), dtype=np.uint32)
    jobs[:, 0] = rng.integers(0, n, size=4000)
    jobs[:, 1] = rng.integers(0, n, size=4000)
    # include same-family neighbours: use the DB's band-0 chains
    r0 = d["rep"][0]
    t = np.nonzero(r0 != np.arange(n))[0][:2000]
    jobs[:len(t), 0] = r0[t]
    jobs[:len(t), 1] = t
    thr = np.float32(identity) / np.float32(100)
    say("align plain kernel, identity", identity)
    got = ctx.align(d["packed"], d["word_off"], jobs, float(thr))
    want = orc.align_jobs(d["packed"], d["word_off"], jobs, float(thr))
    np.testing.assert_array_equal(got, want)
    assert (want[:, 2] > 0).any() or identity >= 90


@pytest.mark.parametrize("identity", [50, 70, 90])
@pytest.mark.parametrize("mode", [ng.GRAPH_PLAIN, ng.GRAPH_BATCHED])
def test_graph_c1(ctx, c1db, identity, mode):
    _, _, d = c1db
    say("graph c1 identity", identity, "mode", mode)
    got, st = ctx.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity, mode)
    want, tot = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
    assert st["total_pairs"] == tot
    np.testing.assert_array_equal(got, want)


def candidate_union(rep, lengths, bands, identity):
    """Every band's isJob candidates from the oracle (buildGraph.hpp:262-300):
    the deduplicated union as sorted rep<<32|tar keys, and the count with
    duplicates."""
    per_band = [orc.candidates(rep[b], lengths, identity) for b in range(bands)]
    allc = np.concatenate(per_band) if per_band else np.zeros(0, np.uint64)
    return np.unique(allc), len(allc)


@pytest.mark.parametrize("identity", [50, 70, 90])
def test_plain_band_jobs_elementwise(ctx, c1db, identity):
    """The plain form's job list of every band, element by element and in the
    reference's order (ascending target), equals the oracle's candidates, and
    every job's score equals the oracle's."""
    _, _, d = c1db
    say("plain band jobs elementwise, identity", identity)
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.set_reps(d["rep"], 64)
    thr = float(np.float32(identity) / np.float32(100))
    seen = 0
    for b in range(64):
        jobs = db.band_jobs(identity, b)
        want = orc.candidates(d["rep"][b], d["lengths"], identity)
        keys = (jobs[:, 0].astype(np.uint64) << np.uint64(32)) | jobs[:, 1].astype(np.uint64)
        np.testing.assert_array_equal(keys, want)
        wj = np.zeros((len(want), 3), dtype=np.uint32)
        wj[:, 0] = want >> np.uint64(32)
        wj[:, 1] = want & np.uint64(0xFFFFFFFF)
        np.testing.assert_array_equal(jobs[:, 2], orc.align_jobs(d["packed"], d["word_off"], wj, thr)[:, 2])
        seen += len(jobs)
    db.close()
    assert seen > 1000


@pytest.mark.parametrize("bands", [64, 20])
@pytest.mark.parametrize("identity", [50, 90])
def test_batched_candidates_elementwise(ctx, c1db, longdb, identity, bands):
    """The batched form's distinct pairs are exactly the deduplicated union of
    the oracle's per-band candidates (no pair dropped, none added), and its
    counts with and without duplicates are the union's."""
    for name, (_, _, d) in (("c1", c1db), ("long", longdb)):
        say("batched candidates elementwise", name, "identity", identity, "bands", bands)
        db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
        db.set_reps(np.ascontiguousarray(d["rep"][:bands]), bands)
        st = db.graph(identity, bands, ng.GRAPH_BATCHED)
        keys, _ = db.pair_scores()
        want, tot = candidate_union(d["rep"], d["lengths"], bands, identity)
        np.testing.assert_array_equal(keys, want)
        assert st["total_pairs"] == tot and st["distinct_pairs"] == len(want)
        db.close()


@pytest.mark.parametrize("identity", [50, 90])
def test_batched_pair_scores(ctx, c1db, identity):
    """Every distinct pair's score from the length-binned kernels equals the oracle's."""
    _, _, d = c1db
    say("batched pair scores identity", identity)
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.set_reps(d["rep"], 64)
    db.graph(identity, 64, ng.GRAPH_BATCHED)
    keys, sc = db.pair_scores()
    jobs = np.zeros((len(keys), 3), dtype=np.uint32)
    jobs[:, 0] = keys >> np.uint64(32)
    jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
    thr = float(np.float32(identity) / np.float32(100))
    want = orc.align_jobs(d["packed"], d["word_off"], jobs, thr)
    np.testing.assert_array_equal(sc, want[:, 2])


@pytest.mark.parametrize("identity", [30, 50, 80, 95])
def test_long_pairs(ctx, longdb, identity):
    """Pairs routed to the wavefront-per-pair kernels (K = 1..8) and the plain bin."""
    _, _, d = longdb
    say("long pairs identity", identity)
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.set_reps(d["rep"], 64)
    db.graph(identity, 64, ng.GRAPH_BATCHED)
    keys, sc = db.pair_scores()
    jobs = np.zeros((len(keys), 3), dtype=np.uint32)
    jobs[:, 0] = keys >> np.uint64(32)
    jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
    thr = float(np.float32(identity) / np.float32(100))
    want = orc.align_jobs(d["packed"], d["word_off"], jobs, thr)
    np.testing.assert_array_equal(sc, want[:, 2])
    got, _ = ctx.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity,
                             ng.GRAPH_BATCHED)
    ref, _ = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
    np.testing.assert_array_equal(got, ref)


def test_all_pairs_long(ctx, longdb):
    """All ordered pairs of the long set through the batched kernels: every window bin."""
    _, _, d = longdb
    n = d["n"]
    # rep rows that chain every sequence to each of the previous ones, band by band
    rows = []
    for b in range(64):
        r = np.arange(n, dtype=np.int64) - 1 - b
        r[r < 0] = np.arange(n)[r < 0]
        rows.append(r.astype(np.uint32))
    rep = np.stack(rows)
    say("all pairs long", n)
    for identity in (20, 60, 90):
        db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
        db.set_reps(rep, 64)
        db.graph(identity, 64, ng.GRAPH_BATCHED)
        keys, sc = db.pair_scores()
        jobs = np.zeros((len(keys), 3), dtype=np.uint32)
        jobs[:, 0] = keys >> np.uint64(32)
        jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
        thr = float(np.float32(identity) / np.float32(100))
        want = orc.align_jobs(d["packed"], d["word_off"], jobs, thr)
        np.testing.assert_array_equal(sc, want[:, 2])


def band_edge_groups(rng):
    """Groups of sequences whose pairs make the band edges matter.

    Per base length: the ancestor, same-length copies shifted left and right
    by up to a quarter of the length (the alignment runs along or outside a
    band edge), copies with a junk prefix, substituted copies, unrelated
    sequences, and low-complexity sequences (one letter with a few others at
    random places), whose band-edge words decide single matches.
    """
    groups = []
    for L in (90, 400, 1100, 2300, 5000, 9000):
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        g = [anc]
        for _ in range(3):
            sh = int(rng.integers(1, L // 4 + 2))
            junk = bytes(AMINO[i] for i in rng.integers(20, size=sh))
            g.append(anc[sh:] + junk)       # path offset +sh, same length
            g.append(junk + anc[:L - sh])   # path offset -sh, same length
            g.append(anc[sh:])              # shorter suffix
        g.append(mutate(rng, anc, 0.05, 0.01))
        g.append(mutate(rng, anc, 0.3, 0.02))
        for _ in range(2):
            g.append(bytes(AMINO[i] for i in rng.integers(20, size=int(L * rng.uniform(0.8, 1.0)))))
        for _ in range(4):
            n = int(L * rng.uniform(0.85, 1.0))
            s = bytearray(b"A" * n)
            for p in rng.integers(0, n, size=max(1, n // 200)):
                s[p] = AMINO[int(rng.integers(1, 20))]
            g.append(bytes(s))
        groups.append(g)
    # widths that need K = 32 at low identity
    anc = bytes(AMINO[i] for i in rng.integers(20, size=20000))
    groups.append([anc, anc[3000:] + anc[:3000], mutate(rng, anc, 0.4, 0.02)])
    return groups


def test_band_edge_pairs(ctx):
    """Every ordered pair within each group through the batched bins (K = 1..32)."""
    rng = np.random.default_rng(29)
    groups = band_edge_groups(rng)
    seqs, gid = [], []
    for k, g in enumerate(groups):
        g = sorted(g, key=lambda s: -len(s))  # the earlier record is the longer one
        seqs += g
        gid += [k] * len(g)
    recs = [orc.pack(s) for s in seqs]
    off = np.cumsum([0] + [len(r) for r in recs[:-1]]).astype(np.uint64)
    packed = np.concatenate(recs).astype(np.uint32)
    lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
    n = len(seqs)
    gid = np.array(gid)
    rows = []
    for b in range(64):
        r = np.arange(n, dtype=np.int64) - 1 - b
        ok = (r >= 0) & (gid[np.maximum(r, 0)] == gid)
        r[~ok] = np.arange(n)[~ok]
        rows.append(r.astype(np.uint32))
    rep = np.stack(rows)
    # A second context stores LCS lengths instead of scores: the band edges
    # decide the LCS of pairs that fail the identity test, whose score is 0
    # whatever the LCS, so only the raw value checks the edges exactly.
    # The raw context keeps windows of 17-30 words on 8 lanes of three words
    # or 16 of two, and those of 32-60 on 16 lanes of three or four
    # (NGIA_SEG2_MIN=0), whatever the bin's size; the default context sends
    # this small set's bins of them to 32 lanes, one word per lane, and to a
    # wavefront per pair.
    os.environ["NGIA_RAW_LCS"] = "1"
    os.environ["NGIA_SEG2_MIN"] = "0"
    try:
        raw_ctx = ng.Context(0)
    finally:
        del os.environ["NGIA_RAW_LCS"]
        del os.environ["NGIA_SEG2_MIN"]
    kinds = set()
    with raw_ctx:
        for identity in (5, 30, 60, 90, 99):
            thr = float(np.float32(identity) / np.float32(100))
            for c, raw in ((ctx, False), (raw_ctx, True)):
                say("band edge pairs, identity", identity, "records", n, "raw lcs", raw)
                db = ng.Database(c, packed, off, lengths)
                db.set_reps(rep, 64)
                db.graph(identity, 64, ng.GRAPH_BATCHED)
                keys, sc = db.pair_scores()
                kinds |= {(b["kind"], b["width"]) for b in db.bin_stats() if b["pairs"]}
                jobs = np.zeros((len(keys), 3), dtype=np.uint32)
                jobs[:, 0] = keys >> np.uint64(32)
                jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
                want = orc.align_jobs(packed, off, jobs, thr, raw_lcs=raw)
                np.testing.assert_array_equal(sc, want[:, 2])
                db.close()
    waves = {w for k, w in kinds if k == "wave"}
    assert {1, 2, 4, 8, 16, 32} <= waves, kinds
    assert {("seg3", 8), ("seg2", 16), ("seg", 32), ("seg3", 16), ("seg4", 16)} <= kinds, kinds


def test_makedb_bytes(ctx, c1db, tmp_path):
    fa, db, _ = c1db
    out = str(tmp_path / "mine.bin")
    say("makedb c1")
    ctx.makedb(fa, out)
    assert open(out, "rb").read() == open(db, "rb").read()


@pytest.mark.parametrize("identity", [50, 90])
def test_cluster_result_file(ctx, c1db, tmp_path, identity):
    fa, db, d = c1db
    res = str(tmp_path / "r.txt")
    say("cluster c1 identity", identity)
    ctx.cluster(db, res, identity)
    edges, _ = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
    cl = orc.leiden(edges, d["n"], orc.resolution(identity))
    want = str(tmp_path / "w.txt")
    orc.save_result(cl, db, want)
    assert open(res, "rb").read() == open(want, "rb").read()
    # the same command through the out-of-HBM path: blocks of a sixth of the packed area
    os.environ["NGIA_BLOCK_WORDS"] = str(len(d["packed"]) // 6)
    try:
        say("cluster c1 in blocks, identity", identity)
        st = ctx.cluster(db, str(tmp_path / "rb.txt"), identity)
    finally:
        del os.environ["NGIA_BLOCK_WORDS"]
    assert st["n_edges"] == len(edges)
    assert open(str(tmp_path / "rb.txt"), "rb").read() == open(want, "rb").read()


@pytest.mark.parametrize("identity", [50, 90])
def test_fast_mode(ctx, c1db, identity):
    _, _, d = c1db
    say("fast c1 identity", identity)
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.set_reps(d["rep"], 64)
    got, tot = db.fast(identity, 64)
    want, wtot = orc.fast_cluster(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, identity)
    assert tot == wtot
    np.testing.assert_array_equal(got, want)


def test_db_pipeline_matches_oracle_db(ctx, c1db):
    """Resident DB: sketch + bucket on the device give the rep rows of the oracle DB."""
    _, _, d = c1db
    say("db sketch+bucket c1")
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.sketch()
    db.bucket()
    np.testing.assert_array_equal(db.reps(), d["rep"])
    np.testing.assert_array_equal(db.sigs(), orc.sketch_all(d["packed"], d["word_off"]))


def test_graph_replay(ctx, c1db):
    """The captured batched round replays to the same edges, interleaved with eager work."""
    _, _, d = c1db
    say("graph replay")
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.set_reps(d["rep"], 64)
    want, _ = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, 70)
    for k in range(4):
        db.graph(70, 64, ng.GRAPH_BATCHED)
        np.testing.assert_array_equal(db.edges(), want)
        db.graph(70, 64, ng.GRAPH_PLAIN)  # eager work between replays
        np.testing.assert_array_equal(db.edges(), want)


@pytest.mark.parametrize("identity", [50, 90])
def test_eager_batched_and_bin_accounting(ctx, longdb, identity):
    """The uncaptured batched form scores like the captured one; every distinct pair
    sits in exactly one bin, and each bin's tile count is the oracle's reference
    tile count over that bin's pairs."""
    _, _, d = longdb
    n = d["n"]
    rows = []
    for b in range(64):
        r = np.arange(n, dtype=np.int64) - 1 - b
        r[r < 0] = np.arange(n)[r < 0]
        rows.append(r.astype(np.uint32))
    rep = np.stack(rows)
    say("eager batched identity", identity)
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    db.set_reps(rep, 64)
    st_g = db.graph(identity, 64, ng.GRAPH_BATCHED)
    keys_g, sc_g = db.pair_scores()
    edges_g = db.edges()
    st_e = db.graph(identity, 64, ng.GRAPH_BATCHED_EAGER)
    keys_e, sc_e = db.pair_scores()
    np.testing.assert_array_equal(keys_e, keys_g)
    np.testing.assert_array_equal(sc_e, sc_g)
    np.testing.assert_array_equal(db.edges(), edges_g)
    assert st_e["ref_tiles"] == st_g["ref_tiles"]
    assert ctx.timing()["align_ms"] > 0
    bins = db.bin_stats()
    assert sum(b["pairs"] for b in bins) == st_e["distinct_pairs"] == len(keys_e)
    assert sum(b["ref_tiles"] for b in bins) == st_e["ref_tiles"]
    thr = float(np.float32(identity) / np.float32(100))
    lens = d["lengths"].astype(np.int64)
    R = lens[(keys_e >> np.uint64(32)).astype(np.int64)]
    T = lens[(keys_e & np.uint64(0xFFFFFFFF)).astype(np.int64)]
    tiles = np.array([orc.align_tiles(int(r), int(t), thr) for r, t in zip(R, T)])
    assert int(tiles.sum()) == st_e["ref_tiles"]
    used = [b for b in bins if b["pairs"]]
    assert any(b["kind"] == "wave" for b in used) and any(b["kind"] == "window" for b in used)
    for b in used:
        if b["kind"] in ("window", "wave", "seg", "plain"):
            assert b["ms"] > 0, b


@pytest.mark.parametrize("shape", [ng.SHAPE_SEG16, ng.SHAPE_SEG32, ng.SHAPE_WAVE, ng.SHAPE_SEG8])
def test_routed_window_bins_score_alike(ctx, c1db, longdb, shape):
    """Window bins routed to another shape (16 or 32 lanes per pair, 8 lanes
    of two words, or a wavefront per pair) score every pair as the oracle does and as one thread
    per pair does, and the bin accounting follows the pairs: the shape's bins
    hold them and their tiles."""
    say("routed window bins, shape", ng.SHAPES[shape])
    saved = ctx.tuning()
    try:
        for name, (_, _, d) in (("c1", c1db), ("long", longdb)):
            db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
            db.set_reps(d["rep"], 64)
            for identity in (50, 90):
                ctx.route_all(ng.SHAPE_WINDOW)
                db.graph(identity, 64, ng.GRAPH_BATCHED_EAGER)
                keys0, sc0 = db.pair_scores()
                edges0 = db.edges()
                win0 = sum(b["pairs"] for b in db.bin_stats() if b["kind"] == "window")
                ctx.route_all(shape)
                st = db.graph(identity, 64, ng.GRAPH_BATCHED_EAGER)
                keys, sc = db.pair_scores()
                np.testing.assert_array_equal(keys, keys0)
                np.testing.assert_array_equal(sc, sc0)
                np.testing.assert_array_equal(db.edges(), edges0)
                bins = db.bin_stats()
                assert all(b["pairs"] == 0 for b in bins if b["kind"] == "window")
                assert sum(b["ref_tiles"] for b in bins) == st["ref_tiles"]
                if shape != ng.SHAPE_WAVE and win0:
                    assert sum(b["pairs"] for b in bins if b["kind"].startswith("seg")) >= win0
                jobs = np.zeros((len(keys), 3), dtype=np.uint32)
                jobs[:, 0] = keys >> np.uint64(32)
                jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
                thr = float(np.float32(identity) / np.float32(100))
                np.testing.assert_array_equal(sc, orc.align_jobs(d["packed"], d["word_off"], jobs, thr)[:, 2])
            db.close()
    finally:
        ctx.set_tuning(saved)


def expected_run_tiles(R, T, thr, lcs, tiles, period):
    """Reference tiles of the blocks a kernel that tests every `period` blocks
    runs: it stops after block k (k a multiple of the period, k < the block
    count) when the LCS so far plus the characters left fails the pass test
    (buildGraph.hpp:248), else runs every block.  period 0: no test."""
    nb = len(lcs)
    thr32 = np.float32(thr)
    if period:
        for k in range(period, nb, period):
            best = int(lcs[k - 1]) + max(T - 32 * k, 0)
            if not np.float32(best) > np.float32(R + T - best) * thr32:
                return int(tiles[k - 1])
    return int(tiles[-1])


def test_early_exit_tally_matches_oracle():
    """The early exit, block for block: a tally round
    (GRAPH_BATCHED_TALLY) scores every pair as the oracle does, and each bin's
    kernel ran exactly the reference tiles the oracle's per-block LCS predicts
    -- one thread per pair tests after every W blocks, a wavefront per pair
    after every 8, the segmented kernels not at all -- on the band-edge groups
    (shifted copies, unrelated and low-complexity sequences fail part way)."""
    rng = np.random.default_rng(71)
    groups = band_edge_groups(rng)[:5]  # base lengths 90 .. 5000
    seqs, gid = [], []
    for k, g in enumerate(groups):
        g = sorted(g, key=lambda x: -len(x))
        seqs += g
        gid += [k] * len(g)
    recs = [orc.pack(x) for x in seqs]
    off = np.cumsum([0] + [len(r) for r in recs[:-1]]).astype(np.uint64)
    packed = np.concatenate(recs).astype(np.uint32)
    lengths = np.array([len(x) for x in seqs], dtype=np.uint32)
    n = len(seqs)
    gid = np.array(gid)
    rows = []
    for b in range(64):
        r = np.arange(n, dtype=np.int64) - 1 - b
        ok = (r >= 0) & (gid[np.maximum(r, 0)] == gid)
        r[~ok] = np.arange(n)[~ok]
        rows.append(r.astype(np.uint32))
    rep = np.stack(rows)
    # several words per lane for every window of 17-30 and 32-60 words,
    # whatever the bin's size
    os.environ["NGIA_SEG2_MIN"] = "0"
    try:
        c2 = ng.Context(0)
    finally:
        del os.environ["NGIA_SEG2_MIN"]
    kinds = set()
    with c2:
        c2.route_all(ng.SHAPE_WINDOW)
        db = ng.Database(c2, packed, off, lengths)
        db.set_reps(rep, 64)
        for identity in (50, 70, 90):
            say("early-exit tally, identity", identity, "records", n)
            thr = float(np.float32(identity) / np.float32(100))
            st = db.graph(identity, 64, ng.GRAPH_BATCHED_TALLY)
            keys, sc = db.pair_scores()
            jobs = np.zeros((len(keys), 3), dtype=np.uint32)
            jobs[:, 0] = keys >> np.uint64(32)
            jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
            np.testing.assert_array_equal(sc, orc.align_jobs(packed, off, jobs, thr)[:, 2])
            want = {}
            for u, v in jobs[:, :2]:
                R, T = int(lengths[u]), int(lengths[v])
                lcs, tiles = orc.align_blocks(recs[u], recs[v], thr)
                if len(lcs) == 0:
                    continue  # length-rejected: no kernel
                W = orc.band_window(R, T, thr)
                if W <= 16:
                    key, period = ("window", W), W
                elif W <= 21:
                    key, period = ("seg3", 8), 0
                elif W <= 30:
                    key, period = ("seg2", 16), 0
                elif W == 31:
                    key, period = ("seg", 32), 0
                elif W <= 45:
                    key, period = ("seg3", 16), 0
                elif W <= 60:
                    key, period = ("seg4", 16), 0
                else:
                    K = next(K for K in (1, 2, 4, 8, 16, 32) if W + K - 1 <= 64 * K)
                    key, period = ("wave", K), 8
                want[key] = want.get(key, 0) + expected_run_tiles(R, T, thr, lcs, tiles, period)
            got = {(b["kind"], b["width"]): b["run_tiles"] for b in db.bin_stats() if b["pairs"]}
            assert got == want, (identity, got, want)
            bins = db.bin_stats()
            assert sum(b["ref_tiles"] for b in bins) == st["ref_tiles"]
            assert sum(b["run_tiles"] for b in bins) < st["ref_tiles"]  # some pairs ended early
            kinds |= set(k for k, _ in want)
        db.close()
    assert {"window", "seg2", "seg3", "seg4", "wave"} <= kinds, kinds
    assert kinds <= {"window", "seg", "seg2", "seg3", "seg4", "wave"}, kinds


def test_segmented_kernels_raw_lcs(ctx):
    """The segmented kernels at the LCS level: every window bin and the 32-lane
    bin's own windows (17..31 words) of the band-edge groups, with raw LCS
    lengths compared to the oracle's banded LCS (a band edge off by one word
    changes only the LCS of pairs that fail)."""
    rng = np.random.default_rng(57)
    groups = band_edge_groups(rng)[:4]  # base lengths 90 .. 2300
    seqs, gid = [], []
    for k, g in enumerate(groups):
        g = sorted(g, key=lambda x: -len(x))
        seqs += g
        gid += [k] * len(g)
    recs = [orc.pack(x) for x in seqs]
    off = np.cumsum([0] + [len(r) for r in recs[:-1]]).astype(np.uint64)
    packed = np.concatenate(recs).astype(np.uint32)
    lengths = np.array([len(x) for x in seqs], dtype=np.uint32)
    n = len(seqs)
    gid = np.array(gid)
    rows = []
    for b in range(64):
        r = np.arange(n, dtype=np.int64) - 1 - b
        ok = (r >= 0) & (gid[np.maximum(r, 0)] == gid)
        r[~ok] = np.arange(n)[~ok]
        rows.append(r.astype(np.uint32))
    rep = np.stack(rows)
    os.environ["NGIA_RAW_LCS"] = "1"
    try:
        raw_ctx = ng.Context(0)
    finally:
        del os.environ["NGIA_RAW_LCS"]
    seg_pairs = 0
    with raw_ctx:
        for shape in (ng.SHAPE_SEG16, ng.SHAPE_SEG32, ng.SHAPE_SEG8):
            raw_ctx.route_all(shape)
            for identity in (30, 60, 90, 99):
                say("segmented raw lcs, shape", ng.SHAPES[shape], "identity", identity)
                thr = float(np.float32(identity) / np.float32(100))
                db = ng.Database(raw_ctx, packed, off, lengths)
                db.set_reps(rep, 64)
                db.graph(identity, 64, ng.GRAPH_BATCHED)
                keys, sc = db.pair_scores()
                seg_pairs += sum(b["pairs"] for b in db.bin_stats() if b["kind"].startswith("seg"))
                jobs = np.zeros((len(keys), 3), dtype=np.uint32)
                jobs[:, 0] = keys >> np.uint64(32)
                jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
                np.testing.assert_array_equal(sc, orc.align_jobs(packed, off, jobs, thr, raw_lcs=True)[:, 2])
                db.close()
    assert seg_pairs > 1000


def test_captured_round_follows_tuning_and_rebucket(ctx, c1db):
    """The captured batched round is replayed while everything it was captured
    with stands (a fresh bucketing into the same rows included) and captured
    again when the routing changes: the bins follow the new routing, and the
    edges stay the oracle's throughout, with another database's rounds in
    between."""
    _, _, d = c1db
    say("captured round: tuning change, rebucket, two databases")
    want, _ = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, 90)
    db = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    other = ng.Database(ctx, d["packed"], d["word_off"], d["lengths"])
    other.set_reps(d["rep"], 64)
    saved = ctx.tuning()
    try:
        ctx.route_all(ng.SHAPE_WINDOW)
        db.sketch()
        db.bucket()
        db.graph(90, 64, ng.GRAPH_BATCHED)
        np.testing.assert_array_equal(db.edges(), want)
        win = sum(b["pairs"] for b in db.bin_stats() if b["kind"] == "window")
        assert win > 0
        ctx.route_all(ng.SHAPE_SEG16)  # every window bin -> four pairs per wavefront
        db.graph(90, 64, ng.GRAPH_BATCHED)
        np.testing.assert_array_equal(db.edges(), want)
        assert all(b["pairs"] == 0 for b in db.bin_stats() if b["kind"] == "window")
        other.graph(90, 64, ng.GRAPH_BATCHED)
        np.testing.assert_array_equal(other.edges(), want)
        ctx.route_all(ng.SHAPE_WINDOW)
        for _ in range(3):
            db.sketch()
            db.bucket()
            db.graph(90, 64, ng.GRAPH_BATCHED)
            np.testing.assert_array_equal(db.edges(), want)
            assert sum(b["pairs"] for b in db.bin_stats() if b["kind"] == "window") == win
            other.graph(90, 64, ng.GRAPH_BATCHED)
            np.testing.assert_array_equal(other.edges(), want)
    finally:
        ctx.set_tuning(saved)
        other.close()
        db.close()


def test_autotune_installs_fastest_shapes(ctx):
    """The tuner times every applicable shape per window width and bin size
    (proteins, then DNA) and installs the fastest one per size class."""
    say("autotune")
    pts = ctx.autotune(90)
    t = ctx.tuning()
    sizes = [16, 64, 256, 1024, 4096, 16384, 65536, 262144]
    assert t["class_size"] == sizes
    widths = sorted({p["width"] for p in pts if not p["dna"] and p["width"] <= 16})
    assert widths == list(range(3, 17))  # W = 2 cannot pass the length test
    # the groups of 17..60-word windows, on proteins
    assert sorted({p["width"] for p in pts if p["width"] > 16}) == list(ng.LONG_WINDOWS)
    assert not any(p["dna"] for p in pts if p["width"] > 16)
    for p in pts:
        assert all(v > 0 for v in p["ms"].values())
        k = sizes.index(p["pairs"])
        if p["width"] > 16:
            assert set(p["ms"]) == {"multi", "single"}
            g = ng.LONG_WINDOWS.index(p["width"])
            want = ng.LONG_SINGLE if p["ms"]["single"] < p["ms"]["multi"] else ng.LONG_MULTI
            assert t["long"][g][k] == want, p
            continue
        assert set(p["ms"]) == ({"window", "seg32", "wave"} | ({"seg16"} if p["width"] <= 15 else set())
                                | ({"seg8"} if p["width"] <= 14 else set()))
        table = t["shape_compact"] if p["dna"] else t["shape"]
        best = min(p["ms"], key=lambda s_: (p["ms"][s_], ng.SHAPES.index(s_)))
        assert table[p["width"]][k] == ng.SHAPES.index(best), p
    ctx.route_all(ng.SHAPE_WINDOW)
    assert ctx.tuning()["long"] == [[ng.LONG_DEFAULT] * 8] * 4


def test_set_tuning_rejects_what_no_kernel_holds(ctx):
    """ngia_set_tuning refuses shapes that cannot hold a window (16 lanes at
    W = 16, 8 lanes of two words above 14 words), unknown shapes and unknown
    long-window choices, and leaves the installed table as it was."""
    base = ctx.tuning()
    one = [1, 0, 0, 0, 0, 0, 0, 0]
    bad = []
    for W, sh in ((16, ng.SHAPE_SEG16), (15, ng.SHAPE_SEG8), (16, ng.SHAPE_SEG8), (5, 5)):
        t = {"class_size": one, "shape": [[0] * 8 for _ in range(17)]}
        t["shape"][W] = [sh] * 8
        bad.append(t)
    bad.append({"class_size": one, "long": [[3] * 8] * 4})
    bad.append({"class_size": [16, 8, 0, 0, 0, 0, 0, 0]})  # sizes must ascend
    for t in bad:
        with pytest.raises(ng.NgiaError):
            ctx.set_tuning(t)
        assert ctx.tuning() == base
    ok = {"class_size": one, "shape": [[0] * 8 for _ in range(17)], "long": [[2] * 8] * 4}
    ok["shape"][14] = [ng.SHAPE_SEG8] * 8
    ctx.set_tuning(ok)
    t = ctx.tuning()
    assert t["shape"][14] == [ng.SHAPE_SEG8] * 8 and t["long"] == [[2] * 8] * 4
    ctx.set_tuning(base)


def test_long_window_routing_table(ctx):
    """The tuned shapes of 17..60-word windows route their bins: with every
    group on several words per lane, then on one word per lane (or a
    wavefront per pair), whatever the bin's size, the band-edge groups'
    pair scores equal the oracle's and the bins in use are those the table
    names."""
    rng = np.random.default_rng(29)
    groups = band_edge_groups(rng)
    seqs, gid = [], []
    for k, g in enumerate(groups):
        g = sorted(g, key=lambda s: -len(s))
        seqs += g
        gid += [k] * len(g)
    recs = [orc.pack(s) for s in seqs]
    off = np.cumsum([0] + [len(r) for r in recs[:-1]]).astype(np.uint64)
    packed = np.concatenate(recs).astype(np.uint32)
    lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
    n = len(seqs)
    gid = np.array(gid)
    rows = []
    for b in range(64):
        r = np.arange(n, dtype=np.int64) - 1 - b
        ok = (r >= 0) & (gid[np.maximum(r, 0)] == gid)
        r[~ok] = np.arange(n)[~ok]
        rows.append(r.astype(np.uint32))
    rep = np.stack(rows)
    base = ctx.tuning()
    several = {("seg3", 8), ("seg2", 16), ("seg3", 16), ("seg4", 16)}
    try:
        for choice in (ng.LONG_MULTI, ng.LONG_SINGLE):
            ctx.set_tuning({"class_size": [1, 0, 0, 0, 0, 0, 0, 0],
                            "long": [[choice] * 8] * 4})
            kinds = set()
            for identity in (30, 60, 90):
                say("long-window routing", ng.LONG_SHAPES[choice], "identity", identity)
                thr = float(np.float32(identity) / np.float32(100))
                db = ng.Database(ctx, packed, off, lengths)
                db.set_reps(rep, 64)
                db.graph(identity, 64, ng.GRAPH_BATCHED)
                keys, sc = db.pair_scores()
                kinds |= {(b["kind"], b["width"]) for b in db.bin_stats() if b["pairs"]}
                jobs = np.zeros((len(keys), 3), dtype=np.uint32)
                jobs[:, 0] = keys >> np.uint64(32)
                jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
                np.testing.assert_array_equal(sc, orc.align_jobs(packed, off, jobs, thr)[:, 2])
                db.close()
            if choice == ng.LONG_MULTI:
                assert several <= kinds, kinds
            else:
                assert not (several & kinds) and ("seg", 32) in kinds and ("wave", 1) in kinds, kinds
    finally:
        ctx.set_tuning(base)


@pytest.fixture(scope="module")
def dnadb():
    """The DNA families of the benchmark suite (ACGT, ~1400 nt, 0-5% substitutions),
    scaled to 60 families of 100."""
    p = ng.gen_config(3, scale=0.003, seed=4)
    d = ng.gen_packed(p)
    sig = orc.sketch_all(d.packed, d.word_off)
    rep = orc.bucket(sig)
    want = {i: orc.build_graph(d.packed, d.word_off, d.lengths, rep, 64, i)[0] for i in (90, 97)}
    return d, sig, rep, want


def test_dna_sketch_bucket(ctx, dnadb):
    d, sig, rep, _ = dnadb
    say("dna sketch/bucket", d.n)
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    db.bucket()
    np.testing.assert_array_equal(db.sigs(), sig)
    np.testing.assert_array_equal(db.reps(), rep)


@pytest.mark.parametrize("identity", [90, 97])
@pytest.mark.parametrize("mode", [ng.GRAPH_PLAIN, ng.GRAPH_BATCHED])
def test_dna_graph(ctx, dnadb, identity, mode):
    d, sig, rep, want = dnadb
    say("dna graph", identity, mode)
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.set_reps(rep, 64)
    db.graph(identity, 64, mode)
    np.testing.assert_array_equal(db.edges(), want[identity])
    assert len(want[identity]) > 500


@pytest.mark.parametrize("alpha", [b"ACGT", b"ACGTN", b"ACGTX", b"ACD"])
def test_compact_alphabet_window_kernels(ctx, tmp_path, alpha):
    """A database whose residues use at most five nonzero codes (DNA; DNA with
    N, code 11; with X, code 0; a three-letter protein alphabet) is aligned by the
    window kernels with a table indexed by slot: every pair scores as with the
    16-code table (a context with NGIA_FULL_ALPHABET=1) and as the oracle."""
    fa, dbp = str(tmp_path / "a.fa"), str(tmp_path / "a.bin")
    rng = np.random.default_rng(21)
    lengths = [40, 90, 200, 333, 480, 700, 1000, 1500, 2100, 3000]
    family_fasta(fa, rng, lengths, members=6, sub=0.04, indel=0.01, alpha=alpha)
    assert orc.makedb(fa, dbp) == 0
    d = orc.read_db_numpy(dbp)
    os.environ["NGIA_FULL_ALPHABET"] = "1"
    try:
        full_ctx = ng.Context(0)
    finally:
        del os.environ["NGIA_FULL_ALPHABET"]
    with full_ctx:
        for identity in (50, 80, 90, 97):
            say("compact alphabet", alpha, "identity", identity)
            got = []
            for c in (ctx, full_ctx):
                db = ng.Database(c, d["packed"], d["word_off"], d["lengths"])
                db.set_reps(d["rep"], 64)
                db.graph(identity, 64, ng.GRAPH_BATCHED)
                got.append((db.pair_scores(), db.edges(),
                            sum(b["pairs"] for b in db.bin_stats() if b["kind"] == "window")))
                db.close()
            (keys, sc), edges, win = got[0]
            np.testing.assert_array_equal(keys, got[1][0][0])
            np.testing.assert_array_equal(sc, got[1][0][1])
            np.testing.assert_array_equal(edges, got[1][1])
            assert win > 0
            jobs = np.zeros((len(keys), 3), dtype=np.uint32)
            jobs[:, 0] = keys >> np.uint64(32)
            jobs[:, 1] = keys & np.uint64(0xFFFFFFFF)
            thr = float(np.float32(identity) / np.float32(100))
            np.testing.assert_array_equal(sc, orc.align_jobs(d["packed"], d["word_off"], jobs, thr)[:, 2])


def test_sketch_past_2gib_of_codes(ctx):
    """Sequences whose byte codes start beyond 2^31 bytes of the code area
    (1.6M DNA sequences of ~1400 nt): their signatures match the oracle."""
    p = ng.gen_config(3, scale=0.8, seed=6)
    d = ng.gen_packed(p)
    codes_bytes = int((32 * ((d.lengths.astype(np.int64) + 31) // 32)).sum())
    assert codes_bytes > (1 << 31) + (1 << 24)
    say("sketch past 2 GiB", d.n, codes_bytes)
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    got = db.sigs()
    db.close()
    tail = np.arange(d.n - 2000, d.n)
    head = np.arange(0, 2000)
    for idx in (head, tail):
        want = np.stack([orc.sketch_record(d.packed[int(d.word_off[i]):]) for i in idx])
        np.testing.assert_array_equal(got[idx], want)


@pytest.mark.parametrize("identity,bands", [(50, 64), (90, 64), (70, 20)])
def test_chunked_graph_c1(ctx, c1db, identity, bands):
    """The out-of-HBM form: blocks of about a fifth of the packed area, every
    (rep block, target block) chunk uploaded and aligned alone; the merged
    edges and the candidate count equal the oracle's (64 bands, and the
    first 20 rows as with -b 20)."""
    _, _, d = c1db
    words = int(d["word_off"][-1]) + 1
    rep = np.ascontiguousarray(d["rep"][:bands])
    say("chunked graph c1 identity", identity, "bands", bands, "block words", words // 5)
    got, st, chunks = ctx.build_graph_chunked(d["packed"], d["word_off"], d["lengths"], rep,
                                              bands, identity, block_words=words // 5)
    want, tot = orc.build_graph(d["packed"], d["word_off"], d["lengths"], rep, bands, identity)
    assert chunks > 5
    assert st["total_pairs"] == tot
    np.testing.assert_array_equal(got, want)


def test_chunked_graph_single_records(ctx, c1db):
    """Every record a block of its own (block_words = 1): one chunk per
    (representative, target) pair; edges equal the oracle's."""
    _, _, d = c1db
    n = 400
    rep = np.ascontiguousarray(d["rep"][:, :n])
    say("chunked graph, one record per block, records", n)
    got, st, chunks = ctx.build_graph_chunked(d["packed"], d["word_off"][:n], d["lengths"][:n], rep,
                                              64, 70, block_words=1)
    want, tot = orc.build_graph(d["packed"], d["word_off"][:n], d["lengths"][:n], rep, 64, 70)
    assert st["total_pairs"] == tot and chunks == st["distinct_pairs"]
    np.testing.assert_array_equal(got, want)


def test_chunked_graph_long_and_whole(ctx, longdb):
    """Long records, each longer than the block limit (blocks of one record),
    and the automatic size (the whole set fits: one chunk, the same edges)."""
    _, _, d = longdb
    say("chunked graph long records")
    want, tot = orc.build_graph(d["packed"], d["word_off"], d["lengths"], d["rep"], 64, 90)
    got, st, chunks = ctx.build_graph_chunked(d["packed"], d["word_off"], d["lengths"], d["rep"],
                                              64, 90, block_words=64)
    assert chunks > 1 and st["total_pairs"] == tot
    np.testing.assert_array_equal(got, want)
    got, st, chunks = ctx.build_graph_chunked(d["packed"], d["word_off"], d["lengths"], d["rep"],
                                              64, 90)
    assert chunks == 1
    np.testing.assert_array_equal(got, want)
