"""Pins of the oracle against what the mathematics fixes: closed forms,
brute force, special cases, an independent textbook LCS.  CPU only."""
import os

import numpy as np
import pytest

import oracle_lib as orc
from conftest import AMINO, mutate

# makedb.hpp:164-169, written out independently as a dictionary
TABLE = {"A": 1, "C": 2, "D": 3, "E": 4, "Q": 4, "F": 5, "W": 5, "Y": 5, "G": 6, "H": 7,
         "I": 8, "V": 8, "K": 9, "L": 10, "M": 10, "N": 11, "P": 12, "R": 13, "S": 14,
         "T": 15, "U": 15}


def py_code(b):
    ch = chr(b).upper() if 0 <= b < 128 else ""
    return TABLE.get(ch, 0) if chr(b).isalpha() and b < 128 else 0


def test_code_table_all_bytes():
    for b in range(256):
        assert orc.code(b) == py_code(b), b


def test_pack_layout_bit_by_bit():
    rng = np.random.default_rng(1)
    for L in (0, 1, 31, 32, 33, 100):
        s = bytes(AMINO[i] for i in rng.integers(20, size=L))
        rec = orc.pack(s)
        assert rec[0] == L and len(rec) == 1 + (L + 31) // 32 * 4
        for j, ch in enumerate(s):
            c = 0
            for e in range(4):
                c |= ((int(rec[1 + j // 32 * 4 + e]) >> (j % 32)) & 1) << e
            assert c == TABLE[chr(ch)]
        # padding bits are code 0
        if L % 32:
            for e in range(4):
                assert int(rec[1 + (L // 32) * 4 + e]) >> (L % 32) == 0


def py_hash(kmer, k):
    M = 0xFFFFFFFF
    h = (kmer ^ ((k * 0x9e3779b1 + 0x85ebca6b) & M)) & M
    h = (h ^ 61) ^ (h >> 16)
    h = (h + (h << 3)) & M
    h = h ^ (h >> 4)
    h = (h * 0x27d4eb2d) & M
    h = h ^ (h >> 15)
    h = ((h ^ (h >> 8)) * 0x9e3779b1) & M
    h = h ^ (h >> 14)
    return h


def py_sketch(codes):
    """Brute force over explicit windows: positions 7..L-1 (8-mer of codes
    j-7..j), or only L-1 when L < 8 (a partial k-mer)."""
    L = len(codes)
    sig = [0xFFFFFFFF] * 64
    positions = range(7, L) if L >= 8 else ([L - 1] if L else [])
    for j in positions:
        kmer = 0
        for t in range(max(0, j - 7), j + 1):
            kmer = ((kmer << 4) + codes[t]) & 0xFFFFFFFF
        for k in range(64):
            sig[k] = min(sig[k], py_hash(kmer, k))
    return np.array(sig, dtype=np.uint32)


@pytest.mark.parametrize("L", [0, 1, 2, 7, 8, 9, 33, 120])
def test_sketch_brute_force(L):
    rng = np.random.default_rng(L)
    s = bytes(AMINO[i] for i in rng.integers(20, size=L))
    codes = [TABLE[chr(c)] for c in s]
    np.testing.assert_array_equal(orc.sketch_record(orc.pack(s)), py_sketch(codes))


def test_sketch_empty_is_all_ones():
    assert (orc.sketch_record(orc.pack(b"")) == 0xFFFFFFFF).all()


def brute_bucket(sig):
    n = sig.shape[0]
    rep = np.zeros((64, n), dtype=np.uint32)
    for b in range(64):
        order = np.lexsort((np.arange(n), sig[:, b]))
        s = sig[order, b]
        row = order.copy()
        same = np.concatenate([[False], s[1:] == s[:-1]])
        row[same] = order[np.nonzero(same)[0] - 1]
        rep[b, order] = row
    return rep


def test_bucket_brute_force_and_invariants():
    rng = np.random.default_rng(4)
    sig = rng.integers(0, 40, size=(700, 64)).astype(np.uint32)
    rep = orc.bucket(sig)
    np.testing.assert_array_equal(rep, brute_bucket(sig))
    idx = np.arange(700)
    assert (rep <= idx).all()
    for b in range(64):
        heads = rep[b] == idx
        # a head is the first index of its signature in the band
        first = {}
        for i in range(700):
            first.setdefault(int(sig[i, b]), i)
        assert all(heads[i] == (first[int(sig[i, b])] == i) for i in range(700))


def test_candidates_closed_form():
    rng = np.random.default_rng(6)
    n = 500
    lens = np.sort(rng.integers(1, 900, size=n))[::-1].astype(np.uint32)
    rep = np.array([rng.integers(0, i + 1) for i in range(n)], dtype=np.uint32)
    for identity in (1, 50, 90, 99):
        thr = np.float32(identity) / np.float32(100)
        want = [(int(rep[i]) << 32) | i for i in range(n)
                if rep[i] != i and np.float32(lens[i]) >= np.float32(lens[rep[i]]) * thr]
        np.testing.assert_array_equal(orc.candidates(rep, lens, identity), np.array(want, np.uint64))


def codes_of(s):
    return np.array([orc.code(c) for c in s], dtype=np.uint8)


def full_score(a, b, thr):
    """Score from the textbook LCS through the reference's pass test."""
    R, T = len(a), len(b)
    if np.float32(T) < np.float32(R) * thr:
        return 0
    lcs = orc.lcs_full(codes_of(a), codes_of(b))
    d = R + T - lcs
    return lcs * 100 // d if np.float32(lcs) > np.float32(d) * thr else 0


def test_align_special_cases():
    rng = np.random.default_rng(8)
    thr = float(np.float32(0.9))
    for L in (1, 5, 31, 32, 33, 64, 300, 1000):
        s = bytes(AMINO[i] for i in rng.integers(20, size=L))
        assert orc.align_record(orc.pack(s), orc.pack(s), thr) == 100  # identical
    assert orc.align_record(orc.pack(b""), orc.pack(b""), thr) == 0
    x = b"X" * 50  # code 0 never matches, not even itself
    assert orc.align_record(orc.pack(x), orc.pack(x), 0.01) == 0
    a = bytes(AMINO[i] for i in rng.integers(20, size=100))
    assert orc.align_record(orc.pack(a), orc.pack(a[:80]), thr) == 0  # 80 < 100*0.9


def test_align_tile_loop_equals_textbook_lcs_score():
    """Property of the reference's banded tile loop: its score equals the
    textbook-LCS score on every pair (it can only fall below the full LCS on
    pairs that fail the identity test either way)."""
    rng = np.random.default_rng(12)
    checked = 0
    for trial in range(1500):
        L = int(rng.integers(1, 260))
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        div = float(rng.uniform(0, 0.5))
        b = mutate(rng, anc, div, float(rng.uniform(0, 0.05)))
        a = mutate(rng, anc, float(rng.uniform(0, 0.2)), 0.01)
        if len(b) > len(a):
            a, b = b, a
        identity = int(rng.integers(1, 100))
        thr = float(np.float32(identity) / np.float32(100))
        got = orc.align_record(orc.pack(a), orc.pack(b), thr)
        assert got == full_score(a, b, np.float32(thr)), (trial, len(a), len(b), identity)
        checked += got > 0
    assert checked > 200  # the property was exercised on passing pairs


def banded_lcs_bigint(a, b, thr32):
    """The reference's banded recurrence (buildGraph.hpp:195-246) on whole-band
    integers: per target character the band's rep bits form one Python
    integer, the addition carries across word boundaries by itself, and the
    carry out of the band's top word is masked off.  Band edges as the fused
    float32 expressions: T - T*thr and R - T*thr rounded once."""
    R, T = len(a), len(b)
    if np.float32(T) < np.float32(R) * thr32:
        return 0
    l32 = np.float32(T - T * np.float64(thr32))
    r32 = np.float32(R - T * np.float64(thr32))
    lsft = (max(int(np.ceil(l32)), 0) + 31) & ~31
    rsft = max(int(np.ceil(r32)), 0) + 32
    nw = (R + 31) // 32
    masks = {}
    for j, ch in enumerate(a):
        c = py_code(ch)
        if c:
            masks[c] = masks.get(c, 0) | (1 << j)
    V = (1 << (32 * nw)) - 1
    for i in range(T):
        blk = i // 32
        js = max(32 * blk - lsft, 0)
        je = min(32 * blk + rsft, R)
        if je <= js:
            continue
        width = 32 * ((je + 31) // 32) - js
        mask = (1 << width) - 1
        v = (V >> js) & mask
        m = (masks.get(py_code(b[i]), 0) >> js) & mask if py_code(b[i]) else 0
        nv = ((v + (v & m)) | (v & ~m)) & mask
        V = (V & ~(mask << js)) | (nv << js)
    return 32 * nw - bin(V).count("1")


def test_align_lcs_whole_band_restatement():
    """orc_align_lcs on pairs whose alignment runs along or outside the band
    edges (shifted, unrelated, low-complexity), where it differs from the
    textbook LCS."""
    rng = np.random.default_rng(31)
    differs = 0
    for trial in range(500):
        L = int(rng.integers(1, 300))
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        kind = trial % 5
        if kind == 0:
            b = bytes(AMINO[i] for i in rng.integers(20, size=int(rng.integers(1, L + 1))))
        elif kind == 1:
            sh = int(rng.integers(0, L // 2 + 1))
            b = anc[sh:] + bytes(AMINO[i] for i in rng.integers(20, size=sh))
        elif kind == 2:
            sh = int(rng.integers(0, L // 2 + 1))
            b = bytes(AMINO[i] for i in rng.integers(20, size=sh)) + anc[:L - sh]
        elif kind == 3:
            anc = b"".join(b"A" if rng.random() < 0.9 else AMINO[i:i + 1] for i in rng.integers(20, size=L))
            b = mutate(rng, anc, 0.1, 0.05)
        else:
            b = mutate(rng, anc, float(rng.uniform(0, 0.5)), float(rng.uniform(0, 0.1)))
        a = anc
        if len(b) > len(a):
            a, b = b, a
        thr = np.float32(int(rng.integers(1, 100))) / np.float32(100)
        want = banded_lcs_bigint(a, b, thr)
        got = orc.align_lcs(orc.pack(a), orc.pack(b), float(thr))
        assert got == want, (trial, len(a), len(b), float(thr))
        if len(a) and np.float32(len(b)) >= np.float32(len(a)) * thr:
            differs += got != orc.lcs_full(codes_of(a), codes_of(b))
    assert differs > 10  # the band bound the recurrence on many pairs


def masked_chain_by_block(a, b, thr32):
    """Longest chain of matches (strictly increasing in both sequences) in the
    band-masked match matrix, over the target's first 32(k+1) characters for
    each block k: the textbook LCS recurrence with match(j, t) true only where
    the residues' codes agree (code 0 never matches) and rep residue j lies in
    a word the reference updates for t's block (buildGraph.hpp:208-216)."""
    R, T = len(a), len(b)
    l32 = np.float32(T - T * np.float64(thr32))
    r32 = np.float32(R - T * np.float64(thr32))
    lsft = (max(int(np.ceil(l32)), 0) + 31) & ~31
    rsft = max(int(np.ceil(r32)), 0) + 32
    ca, cb_ = codes_of(a).astype(np.int64), codes_of(b).astype(np.int64)
    row = np.zeros(R + 1, dtype=np.int64)  # row[j]: chain length over rep[:j]
    out = []
    pos = np.arange(R)
    for t in range(T):
        blk = t // 32
        js, je = max(32 * blk - lsft, 0), min(32 * blk + rsft, R)
        inband = (pos >= js) & (pos < 32 * ((je + 31) // 32)) if je > js else np.zeros(R, bool)
        m = (ca == cb_[t]) & (ca != 0) & inband
        cand = np.maximum(row[1:], row[:-1] + m)
        row = np.concatenate([[0], np.maximum.accumulate(cand)])
        if t % 32 == 31 or t == T - 1:
            out.append(int(row[-1]))
    return out, lsft // 32 + (rsft + 31) // 32


def test_align_blocks_prefix_lcs_is_masked_chain():
    """orc_align_blocks: after each block, the LCS so far equals the longest
    chain of band-masked matches over the target read so far (a DP that shares
    nothing with the bit-parallel loop), so it rises by at most one per target
    character and the final LCS is at most the LCS so far plus the characters
    left, the bound an early exit rests on; the last entry is orc_align_lcs,
    the tiles' last entry orc_align_tiles, and orc_band_window the window
    width of the same band edges."""
    rng = np.random.default_rng(41)
    exits = 0
    for trial in range(160):
        L = int(rng.integers(1, 200))
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        kind = trial % 4
        if kind == 0:
            b = bytes(AMINO[i] for i in rng.integers(20, size=int(rng.integers(1, L + 1))))
        elif kind == 1:
            sh = int(rng.integers(0, L // 2 + 1))
            b = anc[sh:] + bytes(AMINO[i] for i in rng.integers(20, size=sh))
        elif kind == 2:
            b = b"".join(b"X" if rng.random() < 0.05 else anc[i:i + 1] for i in range(L))
        else:
            b = mutate(rng, anc, float(rng.uniform(0, 0.5)), float(rng.uniform(0, 0.1)))
        a = anc
        if len(b) > len(a):
            a, b = b, a
        thr = np.float32(int(rng.integers(1, 100))) / np.float32(100)
        lcs, tiles = orc.align_blocks(orc.pack(a), orc.pack(b), float(thr))
        R, T = len(a), len(b)
        if np.float32(T) < np.float32(R) * thr:
            assert len(lcs) == 0
            continue
        want, W = masked_chain_by_block(a, b, thr)
        assert lcs.tolist() == want, (trial, R, T, float(thr))
        assert orc.band_window(R, T, float(thr)) == W
        assert lcs[-1] == orc.align_lcs(orc.pack(a), orc.pack(b), float(thr))
        assert tiles[-1] == orc.align_tiles(R, T, float(thr))
        done = np.minimum(32 * np.arange(1, len(lcs) + 1), T)
        assert (np.diff(np.concatenate([[0], lcs])) <= np.diff(np.concatenate([[0], done]))).all()
        assert (int(lcs[-1]) <= lcs.astype(np.int64) + (T - done)).all()
        d = R + T - int(lcs[-1])
        if len(lcs) > 1 and not np.float32(lcs[-1]) > np.float32(d) * thr:
            exits += 1
    assert exits > 20  # failing pairs (where an early exit can apply) were covered


def test_fused_and_unfused_band_edges_score_alike():
    rng = np.random.default_rng(13)
    for trial in range(800):
        L = int(rng.integers(30, 700))
        anc = bytes(AMINO[i] for i in rng.integers(20, size=L))
        a = mutate(rng, anc, 0.05, 0.01)
        b = mutate(rng, anc, float(rng.uniform(0, 0.3)), 0.02)
        if len(b) > len(a):
            a, b = b, a
        thr = float(np.float32(int(rng.integers(50, 100))) / np.float32(100))
        assert orc.align_record(orc.pack(a), orc.pack(b), thr, 1) == \
            orc.align_record(orc.pack(a), orc.pack(b), thr, 0)


def test_tile_counts_equal_lengths():
    # tiles of the reference loop for rep = target = L
    thr = lambda i: float(np.float32(i) / np.float32(100))
    assert orc.align_tiles(300, 300, thr(90)) == 28
    assert orc.align_tiles(350, 350, thr(90)) == 49
    assert orc.align_tiles(1400, 1400, thr(90)) == 454
    assert orc.align_tiles(5000, 5000, thr(90)) == 4909
    assert orc.align_tiles(35000, 35000, thr(90)) == 229564


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937_64
    assert orc.mt19937_64_nth(5489, 10000) == 9981545732273789042


def test_resolution_closed_form():
    for i in range(1, 100):
        want = np.float32((np.float32(100.0) - np.float32(abs(np.float32(i) - np.float32(50.0)))
                           * np.float32(2.0)) / np.float32(100.0))
        assert orc.resolution(i) == min(max(want, np.float32(0)), np.float32(1))


def test_fast_cluster_brute_force():
    rng = np.random.default_rng(21)
    fams = [bytes(AMINO[i] for i in rng.integers(20, size=int(rng.integers(40, 200))))
            for _ in range(6)]
    seqs = [mutate(rng, fams[k % 6], 0.04, 0.01) for k in range(60)]
    seqs.sort(key=len, reverse=True)
    recs = [orc.pack(s) for s in seqs]
    off = np.cumsum([0] + [len(r) for r in recs[:-1]]).astype(np.uint64)
    packed = np.concatenate(recs)
    lens = np.array([len(s) for s in seqs], dtype=np.uint32)
    n = len(seqs)
    rep = np.stack([np.array([rng.integers(0, i + 1) for i in range(n)], dtype=np.uint32)
                    for _ in range(8)])
    got, _ = orc.fast_cluster(packed, off, lens, rep, 8, 80)
    thr = float(np.float32(0.8))
    parent = [None] * n
    for b in range(8):
        todo = [(int(rep[b, i]), i) for i in range(n) if rep[b, i] != i and parent[i] is None]
        for r, t in todo:
            if orc.align_record(recs[r], recs[t], thr) > 0:
                parent[t] = r
    cl = []
    for i in range(n):
        cl.append(i if parent[i] is None else cl[parent[i]])
    np.testing.assert_array_equal(got, np.array(cl, dtype=np.uint32))


def graph_from_parts(packed, word_off, lengths, rep, bands, identity):
    """The edge list of buildGraph.hpp:256-362 assembled in numpy from its
    pinned pieces: per band the isJob candidates in closed form (buildGraph.hpp:
    262-266, the loop of test_candidates_closed_form written with numpy), each
    scored by orc.align_jobs (pinned by the textbook LCS); the passing ones
    appended as (rep, tar, (float)score / 100.0f) (:339-350); then the
    lexicographic sort of (u, v, w) and std::unique (:355-362).  Returns the
    edges and the candidate count with duplicates (:353)."""
    n = len(lengths)
    thr = np.float32(identity) / np.float32(100)
    idx = np.arange(n, dtype=np.uint32)
    lens = np.asarray(lengths, dtype=np.uint32)
    u, v, w = [], [], []
    total = 0
    for b in range(bands):
        r = rep[b]
        ok = (r != idx) & (lens.astype(np.float32) >= lens[r].astype(np.float32) * thr)
        tar = idx[ok]
        total += len(tar)
        if not len(tar):
            continue
        jobs = np.zeros((len(tar), 3), dtype=np.uint32)
        jobs[:, 0] = r[tar]
        jobs[:, 1] = tar
        sc = orc.align_jobs(packed, word_off, jobs, float(thr))[:, 2]
        keep = sc > 0
        u.append(jobs[keep, 0].astype(np.int32))
        v.append(jobs[keep, 1].astype(np.int32))
        w.append(sc[keep].astype(np.float32) / np.float32(100))
    e = np.zeros(sum(len(x) for x in u), dtype=orc.EDGE_DTYPE)
    if len(e):
        e["u"], e["v"], e["w"] = np.concatenate(u), np.concatenate(v), np.concatenate(w)
        e = e[np.lexsort((e["w"], e["v"], e["u"]))]
        first = np.ones(len(e), dtype=bool)
        first[1:] = (e["u"][1:] != e["u"][:-1]) | (e["v"][1:] != e["v"][:-1]) | \
                    (e["w"][1:] != e["w"][:-1])
        e = e[first]
    return e, total


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden_db(name):
    return orc.read_db_numpy(os.path.join(GOLDEN, name))


@pytest.mark.parametrize("source", ["edge_cases", "c1_small", "config1_full"])
def test_build_graph_equals_assembly_from_pinned_parts(tmp_path, source):
    """orc_build_graph against its parts.  Sets: the reference-built edge-case
    DB (empty, short, 65,534-residue and unknown-letter records) with its own
    rows, the reference-built 500-record C1 sample with its rows, and the full
    10,000-sequence C1 set (200 families of 50) bucketed by the oracle."""
    import ngia2_amd as ng
    if source == "config1_full":
        d = ng.gen_packed(ng.gen_config(1, scale=1.0, seed=1))
        packed, off, lens = d.packed, d.word_off, d.lengths
        rep = orc.bucket(orc.sketch_all(packed, off))
    else:
        if source == "edge_cases":
            d = _golden_db("edge_cases.packed")
        else:  # the oracle's DB of c1_small.fa is byte-equal to the reference's (test_golden)
            db = str(tmp_path / "c1.packed")
            assert orc.makedb(os.path.join(GOLDEN, "c1_small.fa"), db) == 0
            d = orc.read_db_numpy(db)
        packed, off, lens, rep = d["packed"], d["word_off"], d["lengths"], d["rep"]
    edges_seen = 0
    for identity in (30, 50, 70, 90, 97):
        for bands in (64, 7):
            want, wtot = graph_from_parts(packed, off, lens, rep, bands, identity)
            got, tot = orc.build_graph(packed, off, lens, rep, bands, identity)
            assert tot == wtot
            np.testing.assert_array_equal(got, want)
            # weights are exactly score/100 of a score in 1..100, and u < v
            sc = np.rint(got["w"].astype(np.float64) * 100)
            assert ((sc >= 1) & (sc <= 100)).all()
            np.testing.assert_array_equal(sc.astype(np.float32) / np.float32(100), got["w"])
            assert (got["u"] < got["v"]).all()
            edges_seen += len(got)
    assert edges_seen > 0
