"""Serving: assignment of streamed queries against a resident clustered snapshot.

CPU tests pin the oracle's assignment against a brute restatement that finds
bucket heads by following the groupSignatures chains (rep rows from the
bucketing oracle) instead of the oracle's sorted signature arrays.  GPU tests
check the HIP serving loop against the oracle: offline, alone, mixed into
continuous batches, and with lanes small enough to split requests.
"""
import numpy as np
import pytest

import ngia2_amd as ng
import oracle_lib as orc
from conftest import AMINO, mutate


def resident_set(seed, n_fam=30, members=6):
    """Families of mutated ancestors in makeDB order (stable length-descending)."""
    rng = np.random.default_rng(seed)
    seqs = []
    for f in range(n_fam):
        L = int(rng.integers(40, 400))
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        for m in range(members):
            seqs.append(mutate(rng, anc, float(rng.uniform(0, 0.15)), 0.01))
    seqs.append(b"")  # an empty record: its signature is all ones
    order = sorted(range(len(seqs)), key=lambda i: -len(seqs[i]))
    seqs = [seqs[i] for i in order]
    packed, off = orc.pack_many(seqs)
    lengths = np.array([len(s) for s in seqs], dtype=np.uint32)
    sig = orc.sketch_all(packed, off)
    return seqs, packed, off, lengths, sig


def queries_for(seqs, rng, nq):
    out = []
    for k in range(nq):
        u = rng.random()
        if u < 0.15:
            out.append(seqs[int(rng.integers(len(seqs)))])  # a resident, unchanged
        elif u < 0.75:
            out.append(mutate(rng, seqs[int(rng.integers(len(seqs)))], float(rng.uniform(0, 0.3)),
                              0.01))
        else:
            out.append(bytes(AMINO[i] for i in rng.integers(20, size=int(rng.integers(0, 500)))))
    return out + [b"", b"A", b"ACDEFGH", b"ACDEFGHI", b"XXXXXXXXXXXX"]


def brute_assign(packed, off, lengths, sig, cluster_of, qs, identity, new_base, first_qid):
    """Every query by the rule, in arrival order, with heads found through the
    bucketing oracle's chains over the grown snapshot (residents, then the
    queries in order): the head of query k in band b is the end of the
    groupSignatures chain through its own index n + k, and is a candidate when
    it lies before it."""
    n = len(lengths)
    thr = float(np.float32(identity) / np.float32(100))
    qrecs = [orc.pack(q) for q in qs]
    qsig = np.stack([orc.sketch_record(r) for r in qrecs])
    rep = orc.bucket(np.concatenate([sig, qsig]))
    rows = []
    for k, q in enumerate(qs):
        i = n + k
        heads = []
        for b in range(64):
            h = i
            while rep[b, h] != h:  # walk the chain to its head
                h = int(rep[b, h])
            if h != i and h not in heads:
                heads.append(h)
        qrec = qrecs[k]
        best, best_r, cand = 0, None, 0
        for h in heads:
            hrec = packed[int(off[h]):] if h < n else qrecs[h - n]
            Lr, Lq = int(hrec[0]), len(q)
            r_rec, t_rec = (hrec, qrec) if Lr >= Lq else (qrec, hrec)
            R, T = (Lr, Lq) if Lr >= Lq else (Lq, Lr)
            if np.float32(T) < np.float32(R) * np.float32(thr):
                continue
            cand += 1
            sc = orc.align_record(r_rec, t_rec, thr)
            if sc > best or (sc == best and sc > 0 and h < best_r):
                best, best_r = sc, h
        if best == 0:
            rows.append((new_base + first_qid + k, 0xFFFFFFFF, 0, cand))
        else:
            cl = int(cluster_of[best_r]) if best_r < n else rows[best_r - n][0]
            rows.append((cl, best_r, best, cand))
    return rows


@pytest.mark.parametrize("identity", [50, 90])
def test_oracle_serve_assign_matches_brute_force(identity):
    seqs, packed, off, lengths, sig = resident_set(3)
    rng = np.random.default_rng(identity)
    cluster_of = rng.integers(0, 1000, len(seqs)).astype(np.uint32)
    qs = queries_for(seqs, rng, 120)
    got = orc.serve_assign(packed, off, sig, cluster_of, qs, identity, 5000, first_qid=7)
    want = brute_assign(packed, off, lengths, sig, cluster_of, qs, identity, 5000, 7)
    for k in range(len(qs)):
        assert tuple(int(x) for x in got[k]) == want[k], (k, len(qs[k]))
    # exact copies of residents pass with score 100 unless the rule's length test bars them
    assert (got[:, 2] == 100).sum() >= 10
    assert (got[:, 1] == 0xFFFFFFFF).sum() >= 5  # novel queries open clusters


def test_oracle_serve_snapshot_grows():
    """Queries join the snapshot: a novel sequence opens a cluster, and a
    later copy or close variant of it joins that cluster through it (its best
    index is the first one's snapshot index), while the same variant served
    before it opens its own."""
    seqs, packed, off, lengths, sig = resident_set(5)
    n = len(seqs)
    rng = np.random.default_rng(12)
    cluster_of = np.arange(n, dtype=np.uint32)
    novel = bytes(AMINO[i] for i in rng.integers(20, size=300))
    close = mutate(rng, novel, 0.02, 0.0)
    junk = bytes(AMINO[i] for i in rng.integers(20, size=200))
    got = orc.serve_assign(packed, off, sig, cluster_of, [novel, junk, close, novel], 90, 10 ** 6)
    assert got[0][1] == 0xFFFFFFFF and got[0][0] == 10 ** 6
    assert got[2][1] == n and got[2][0] == 10 ** 6 and got[2][2] >= 90
    assert got[3][1] == n and got[3][2] == 100 and got[3][0] == 10 ** 6
    rev = orc.serve_assign(packed, off, sig, cluster_of, [close, novel], 90, 10 ** 6)
    assert rev[0][1] == 0xFFFFFFFF and rev[0][0] == 10 ** 6
    assert rev[1][1] == n and rev[1][0] == 10 ** 6
    # a chain: the third query joins through the second, which joined the first
    assert tuple(int(x) for x in got[3]) == tuple(brute_assign(
        packed, off, lengths, sig, cluster_of, [novel, junk, close, novel], 90, 10 ** 6, 0)[3])


def test_oracle_serve_resident_queries_point_at_their_chain_heads():
    """A resident sent as a query finds, band by band, the head of its own chain."""
    seqs, packed, off, lengths, sig = resident_set(8)
    rep = orc.bucket(sig)
    cluster_of = np.arange(len(seqs), dtype=np.uint32)
    for i in range(0, len(seqs), 7):
        got = orc.serve_assign(packed, off, sig, cluster_of, [seqs[i]], 90, 10 ** 6)[0]
        heads = set()
        for b in range(64):
            h = i
            while rep[b, h] != h:
                h = int(rep[b, h])
            heads.add(h)
        if got[1] != 0xFFFFFFFF:
            assert int(got[1]) in heads
        assert got[3] <= len(heads)


# ------------------------------------------------------------------ GPU
@pytest.fixture(scope="module")
def served(ctx):
    d = ng.gen_packed(ng.gen_config(1, scale=0.1, seed=5))
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    sig = db.sigs()
    rng = np.random.default_rng(17)
    cluster_of = (np.arange(d.n) % 37 + 1000).astype(np.uint32)
    q1 = ng.gen_queries(d.packed, d.word_off, 300, mutated=0.7, sub=(0.0, 0.3), seed=9)
    q2 = ng.gen_queries(d.packed, d.word_off, 40, mutated=1.0, sub=(0.0, 0.0), indel=0.0, seed=10)
    extra = [b"", b"A", b"ACDEFGH", b"XXXXXXXXXXXXXXX", bytes(AMINO[i] for i in rng.integers(20, size=3000))]
    strs = [q1.seq(k) for k in range(q1.n)] + [q2.seq(k) for k in range(q2.n)] + extra
    queries = ng.Queries.from_strings(strs)
    want = orc.serve_assign(d.packed, d.word_off, sig, cluster_of, strs, 90, 10 ** 6)
    return d, db, cluster_of, queries, want


def as_rows(a):
    return np.stack([a["cluster"], a["resident"], a["score"], a["candidates"]], axis=1)


@pytest.mark.gpu
def test_serve_offline_matches_oracle(served):
    d, db, cluster_of, queries, want = served
    print("[case] serve offline", queries.n, flush=True)
    sv = ng.Serve(db, cluster_of, 90, 10 ** 6)
    got, st = sv.assign(queries)
    np.testing.assert_array_equal(as_rows(got), want)
    assert st["queries"] == queries.n and st["batches"] >= 1 and st["grown"] == queries.n
    assert (want[:, 2] == 100).sum() >= 30  # exact copies found themselves or a twin
    assert (want[:, 1] >= d.n).any() and (want[:, 1] == 0xFFFFFFFF).any()
    sv.close()


@pytest.mark.gpu
def test_serve_alone_and_in_the_mix(served):
    """Each query's result is the same alone (a batch of its own), in a mixed
    continuous stream and offline: it depends on the queries before it, not
    on how the stream was cut."""
    d, db, cluster_of, queries, want = served
    # alone: batches of one query, two lanes alternating
    sv = ng.Serve(db, cluster_of, 90, 10 ** 6, lanes=2, max_queries=1)
    print("[case] serve alone", queries.n, flush=True)
    n1 = np.ones(queries.n, dtype=np.int64)
    got, lat, st = sv.replay(queries, np.zeros(queries.n), np.arange(queries.n), n1)
    np.testing.assert_array_equal(as_rows(got), want)
    assert st["batches"] == queries.n and (lat > 0).all()
    sv.close()
    # mix: requests of random sizes arriving within 30 ms, batched as lanes free up
    sv = ng.Serve(db, cluster_of, 90, 10 ** 6, lanes=2, max_queries=64)
    rng = np.random.default_rng(23)
    for trial in range(2):
        print("[case] serve mix", trial, flush=True)
        bounds = np.sort(rng.choice(np.arange(1, queries.n), 40, replace=False))
        first = np.concatenate([[0], bounds])
        count = np.diff(np.concatenate([first, [queries.n]]))
        arrival = np.sort(rng.uniform(0, 0.03, len(first)))
        got, lat, st = sv.replay(queries, arrival, first, count)
        np.testing.assert_array_equal(as_rows(got), want)
        assert len(lat) == len(first) and (lat > 0).all()
        assert st["requests"] == len(first)
        sv.reset()
    sv.close()


@pytest.mark.gpu
def test_serve_small_lanes_split_requests(served):
    """Lanes smaller than the requests: requests are split across batches."""
    d, db, cluster_of, queries, want = served
    print("[case] serve small lanes", flush=True)
    sv = ng.Serve(db, cluster_of, 90, 10 ** 6, lanes=3, max_queries=7, max_residues=70000)
    got, lat, st = sv.replay(queries, [0.0, 0.0, 0.001], [0, 100, 200], [100, 100, queries.n - 200])
    np.testing.assert_array_equal(as_rows(got), want)
    assert st["batches"] >= queries.n // 7
    sv.reset()
    got2, _ = sv.assign(queries, first_qid=0)
    np.testing.assert_array_equal(as_rows(got2), want)
    sv.close()


@pytest.mark.gpu
def test_serve_session_continues_across_calls(served):
    """Without a reset the session continues: a second call's queries see the
    first call's.  Serving the set in two calls (ids continued) equals one
    offline call; a reset starts afresh; a session past grow_queries fails."""
    d, db, cluster_of, queries, want = served
    print("[case] serve session across calls", flush=True)
    sv = ng.Serve(db, cluster_of, 90, 10 ** 6, grow_queries=queries.n + 5)
    h = queries.n // 2
    a = ng.Queries(queries.res, queries.off[:h], queries.length[:h])
    b = ng.Queries(queries.res, queries.off[h:], queries.length[h:])
    got_a, _ = sv.assign(a, first_qid=0)
    got_b, st = sv.assign(b, first_qid=h)
    np.testing.assert_array_equal(as_rows(np.concatenate([got_a, got_b])), want)
    assert st["grown"] == queries.n
    with pytest.raises(ng.NgiaError):  # 6 more queries exceed grow_queries
        sv.assign(ng.Queries(queries.res, queries.off[:6], queries.length[:6]), first_qid=10 ** 5)
    sv.reset()
    got, st = sv.assign(queries)
    np.testing.assert_array_equal(as_rows(got), want)
    sv.close()


@pytest.mark.gpu
def test_serve_grown_queries_join_each_other(ctx):
    """Novel families arriving over the stream: their later members join the
    cluster their first member opened (through snapshot indices >= n), alone,
    mixed and offline alike, as the oracle assigns them."""
    d = ng.gen_packed(ng.gen_config(1, scale=0.05, seed=8))
    db = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    db.sketch()
    sig = db.sigs()
    cluster_of = np.arange(d.n, dtype=np.uint32)
    rng = np.random.default_rng(31)
    strs = []
    for f in range(12):  # 12 new families of 6, interleaved over the stream
        anc = bytes(AMINO[i] for i in rng.integers(20, size=int(rng.integers(80, 900))))
        strs += [mutate(rng, anc, float(rng.uniform(0, 0.04)), 0.005) for _ in range(6)]
    order = rng.permutation(len(strs))
    strs = [strs[k] for k in order]
    queries = ng.Queries.from_strings(strs)
    want = orc.serve_assign(d.packed, d.word_off, sig, cluster_of, strs, 90, 10 ** 6)
    assert (want[:, 1] >= d.n).sum() >= 30 and (want[:, 1] != 0xFFFFFFFF).sum() < len(strs)
    for lanes, mq in ((1, 0), (2, 1), (3, 5)):
        print("[case] grown families lanes", lanes, "max_queries", mq, flush=True)
        sv = ng.Serve(db, cluster_of, 90, 10 ** 6, lanes=lanes, max_queries=mq)
        if mq == 0:
            got, _ = sv.assign(queries)
        else:
            first = np.arange(0, queries.n, 4)
            count = np.minimum(4, queries.n - first)
            got, _, _ = sv.replay(queries, np.linspace(0, 0.01, len(first)), first, count)
        np.testing.assert_array_equal(as_rows(got), want)
        sv.close()
    db.close()


@pytest.mark.gpu
def test_serve_rejects_bad_input(served, ctx):
    d, db, cluster_of, queries, want = served
    with pytest.raises(ng.NgiaError):
        ng.Serve(db, cluster_of, 0, 0)
    fresh = ng.Database(ctx, d.packed, d.word_off, d.lengths)
    with pytest.raises(ng.NgiaError):  # no signatures yet
        ng.Serve(fresh, cluster_of, 90, 0)
    sv = ng.Serve(db, cluster_of, 90, 0)
    with pytest.raises(ng.NgiaError):  # request outside the query set
        sv.replay(queries, [0.0], [queries.n - 1], [5])
    with pytest.raises(ng.NgiaError):  # decreasing arrivals
        sv.replay(queries, [0.01, 0.0], [0, 1], [1, 1])
    sv.close()
    fresh.close()


@pytest.mark.gpu
def test_handles_outlive_their_context(served):
    """A database or serving handle destroyed after its context only frees its memory."""
    d, _, cluster_of, queries, _ = served
    print("[case] handles outlive context", flush=True)
    c2 = ng.Context(0)
    db = ng.Database(c2, d.packed, d.word_off, d.lengths)
    db.sketch()
    sv = ng.Serve(db, cluster_of, 90, 0)
    ng._lib.lib().ngia_close(c2.h)  # the native context goes first
    c2.h = None
    with pytest.raises(ng.NgiaError):
        sv.assign(queries)
    sv.close()
    db.close()
