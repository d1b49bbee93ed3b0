"""ctypes access to the oracle (oracle/liboracle.so) -- the checker.

Test infrastructure only: the parity tests compare the HIP product path with
these CPU restatements of the reference algorithm.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libngia2_ref.so")

P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64

EDGE_DTYPE = np.dtype([("u", np.int32), ("v", np.int32), ("w", np.float32)])

_orc = None
_ref = None


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def orc():
    global _orc
    if _orc is None:
        L = C.CDLL(ORACLE_SO)
        L.orc_code.restype = U32
        L.orc_code.argtypes = [C.c_uint8]
        L.orc_pack.argtypes = [P, U32, P]
        L.orc_sketch_record.argtypes = [P, P]
        L.orc_sketch_all.argtypes = [P, P, U32, P, C.c_int]
        L.orc_bucket.argtypes = [P, U32, P, C.c_int]
        L.orc_candidates.restype = U64
        L.orc_candidates.argtypes = [P, P, U32, U32, P]
        L.orc_align_record.restype = U32
        L.orc_align_record.argtypes = [P, P, C.c_float, C.c_int]
        L.orc_align_tiles.restype = U64
        L.orc_align_tiles.argtypes = [U32, U32, C.c_float, C.c_int]
        L.orc_align_jobs.argtypes = [P, P, P, U32, C.c_float, C.c_int, C.c_int]
        L.orc_align_lcs.restype = U32
        L.orc_align_lcs.argtypes = [P, P, C.c_float, C.c_int]
        L.orc_align_jobs_lcs.argtypes = [P, P, P, U32, C.c_float, C.c_int, C.c_int]
        L.orc_align_blocks.restype = U32
        L.orc_align_blocks.argtypes = [P, P, C.c_float, C.c_int, P, P]
        L.orc_band_window.restype = U32
        L.orc_band_window.argtypes = [U32, U32, C.c_float, C.c_int]
        L.orc_lcs_full.restype = U32
        L.orc_lcs_full.argtypes = [P, U32, P, U32]
        L.orc_build_graph.restype = C.c_int
        L.orc_build_graph.argtypes = [P, P, P, U32, P, U32, U32, C.c_int, C.c_int,
                                      C.POINTER(C.c_void_p), C.POINTER(U64), C.POINTER(U64)]
        L.orc_resolution.restype = C.c_float
        L.orc_resolution.argtypes = [U32]
        L.orc_leiden.restype = C.c_int
        L.orc_leiden.argtypes = [P, U64, C.c_int32, C.c_float, U32, C.c_int32, P]
        L.orc_fast_cluster.restype = C.c_int
        L.orc_fast_cluster.argtypes = [P, P, P, U32, P, U32, U32, C.c_int, C.c_int, P,
                                       C.POINTER(U64)]
        L.orc_makedb.restype = C.c_int
        L.orc_serve_assign.argtypes = [P, P, U32, P, P, P, P, U32, U32, U32, U64, P, C.c_int]
        L.orc_serve_assign.restype = None
        L.orc_makedb.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        L.orc_save_result.restype = C.c_int
        L.orc_save_result.argtypes = [P, U32, C.c_char_p, C.c_char_p]
        L.orc_mt19937_64_nth.restype = U64
        L.orc_mt19937_64_nth.argtypes = [U64, U64]
        L.orc_free.argtypes = [P]
        _orc = L
    return _orc


def ref():
    """The reference's own host headers built by oracle/Makefile (ref target); None if absent."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        L = C.CDLL(REF_SO)
        L.ref_makedb.restype = C.c_int
        L.ref_makedb.argtypes = [C.c_char_p, C.c_char_p]
        L.ref_leiden.restype = C.c_int
        L.ref_leiden.argtypes = [P, P, P, U64, C.c_int32, C.c_float, U32, C.c_int32, P]
        L.ref_save.argtypes = [P, U32, C.c_char_p, C.c_char_p]
        _ref = L
    return _ref


# ------------------------------------------------------------------ wrappers
def code(c):
    return orc().orc_code(c)


def pack(residues: bytes):
    L = len(residues)
    rec = np.zeros(1 + (L + 31) // 32 * 4, dtype=np.uint32)
    buf = np.frombuffer(residues, dtype=np.uint8) if L else np.zeros(1, np.uint8)
    orc().orc_pack(_p(buf), L, _p(rec))
    return rec


def sketch_record(rec):
    rec = np.ascontiguousarray(rec, dtype=np.uint32)
    sig = np.zeros(64, dtype=np.uint32)
    orc().orc_sketch_record(_p(rec), _p(sig))
    return sig


def sketch_all(packed, word_off, threads=0):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    word_off = np.ascontiguousarray(word_off, dtype=np.uint64)
    n = len(word_off)
    sig = np.zeros((n, 64), dtype=np.uint32)
    orc().orc_sketch_all(_p(packed), _p(word_off), n, _p(sig), threads)
    return sig


def bucket(sig, threads=0):
    sig = np.ascontiguousarray(sig, dtype=np.uint32)
    n = sig.shape[0]
    rep = np.zeros((64, n), dtype=np.uint32)
    orc().orc_bucket(_p(sig), n, _p(rep), threads)
    return rep


def candidates(rep_row, lengths, identity):
    rep_row = np.ascontiguousarray(rep_row, dtype=np.uint32)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    n = len(rep_row)
    out = np.zeros(max(n, 1), dtype=np.uint64)
    c = orc().orc_candidates(_p(rep_row), _p(lengths), n, identity, _p(out))
    return out[:c]


def align_record(rep_rec, tar_rec, thr, fused=1):
    a = np.ascontiguousarray(rep_rec, dtype=np.uint32)
    b = np.ascontiguousarray(tar_rec, dtype=np.uint32)
    return orc().orc_align_record(_p(a), _p(b), C.c_float(thr), fused)


def align_lcs(rep_rec, tar_rec, thr, fused=1):
    a = np.ascontiguousarray(rep_rec, dtype=np.uint32)
    b = np.ascontiguousarray(tar_rec, dtype=np.uint32)
    return orc().orc_align_lcs(_p(a), _p(b), C.c_float(thr), fused)


def align_tiles(R, T, thr, fused=1):
    return orc().orc_align_tiles(R, T, C.c_float(thr), fused)


def align_blocks(rep_rec, tar_rec, thr, fused=1):
    """Per 32-character block of the target: the LCS so far and the tiles
    visited so far (two arrays of ceil(T/32); empty when the length filter
    rejects the pair)."""
    a = np.ascontiguousarray(rep_rec, dtype=np.uint32)
    b = np.ascontiguousarray(tar_rec, dtype=np.uint32)
    nb = (int(b[0]) + 31) // 32
    lcs = np.zeros(max(nb, 1), dtype=np.uint32)
    tiles = np.zeros(max(nb, 1), dtype=np.uint64)
    k = orc().orc_align_blocks(_p(a), _p(b), C.c_float(thr), fused, _p(lcs), _p(tiles))
    return lcs[:k], tiles[:k]


def band_window(R, T, thr, fused=1):
    return orc().orc_band_window(R, T, C.c_float(thr), fused)


def align_jobs(packed, word_off, jobs, thr, fused=1, threads=0, raw_lcs=False):
    """Score (or, with raw_lcs, the banded LCS length) of every job pair."""
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    word_off = np.ascontiguousarray(word_off, dtype=np.uint64)
    jobs = np.array(jobs, dtype=np.uint32, order="C", copy=True)
    fn = orc().orc_align_jobs_lcs if raw_lcs else orc().orc_align_jobs
    fn(_p(packed), _p(word_off), _p(jobs), len(jobs), C.c_float(thr), fused, threads)
    return jobs


def lcs_full(a_codes, b_codes):
    a = np.ascontiguousarray(a_codes, dtype=np.uint8)
    b = np.ascontiguousarray(b_codes, dtype=np.uint8)
    aa = a if len(a) else np.zeros(1, np.uint8)
    bb = b if len(b) else np.zeros(1, np.uint8)
    return orc().orc_lcs_full(_p(aa), len(a), _p(bb), len(b))


def build_graph(packed, word_off, lengths, rep_rows, bands, identity, fused=1, threads=0):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    word_off = np.ascontiguousarray(word_off, dtype=np.uint64)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    rep_rows = np.ascontiguousarray(rep_rows, dtype=np.uint32)
    ep = C.c_void_p()
    m = U64()
    tot = U64()
    orc().orc_build_graph(_p(packed), _p(word_off), _p(lengths), len(lengths), _p(rep_rows),
                          bands, identity, fused, threads, C.byref(ep), C.byref(m),
                          C.byref(tot))
    out = np.frombuffer(C.string_at(ep, m.value * 12), dtype=EDGE_DTYPE).copy() if m.value else \
        np.zeros(0, EDGE_DTYPE)
    orc().orc_free(ep)
    return out, tot.value


def resolution(identity):
    return orc().orc_resolution(identity)


def leiden(edges, n, res, seed=42, iters=1):
    edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
    out = np.zeros(n, dtype=np.uint32)
    e = edges if len(edges) else np.zeros(1, EDGE_DTYPE)
    orc().orc_leiden(_p(e), len(edges), n, C.c_float(res), seed, iters, _p(out))
    return out


def ref_leiden(edges, n, res, seed=42, iters=1):
    L = ref()
    if L is None:
        return None
    edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
    u = np.ascontiguousarray(edges["u"]) if len(edges) else np.zeros(1, np.int32)
    v = np.ascontiguousarray(edges["v"]) if len(edges) else np.zeros(1, np.int32)
    w = np.ascontiguousarray(edges["w"]) if len(edges) else np.zeros(1, np.float32)
    out = np.zeros(n, dtype=np.uint32)
    L.ref_leiden(_p(u), _p(v), _p(w), len(edges), n, C.c_float(res), seed, iters, _p(out))
    return out


def fast_cluster(packed, word_off, lengths, rep_rows, bands, identity, fused=1, threads=0):
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    word_off = np.ascontiguousarray(word_off, dtype=np.uint64)
    lengths = np.ascontiguousarray(lengths, dtype=np.uint32)
    rep_rows = np.ascontiguousarray(rep_rows, dtype=np.uint32)
    n = len(lengths)
    out = np.zeros(n, dtype=np.uint32)
    tot = U64()
    orc().orc_fast_cluster(_p(packed), _p(word_off), _p(lengths), n, _p(rep_rows), bands,
                           identity, fused, threads, _p(out), C.byref(tot))
    return out, tot.value


def pack_many(seqs):
    """Packed records of residue strings: (packed area, word offsets)."""
    recs = [pack(x) for x in seqs]
    off = np.zeros(len(recs), dtype=np.uint64)
    if len(recs) > 1:
        off[1:] = np.cumsum([len(r) for r in recs[:-1]])
    packed = np.concatenate(recs).astype(np.uint32) if recs else np.zeros(1, np.uint32)
    return packed, off


def serve_assign(packed, word_off, sig, cluster_of, qseqs, identity, new_base, first_qid=0,
                 threads=0):
    """Serving assignment of residue strings qseqs: (nq, 4) uint32 rows of
    (cluster, resident, score, candidates)."""
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    word_off = np.ascontiguousarray(word_off, dtype=np.uint64)
    sig = np.ascontiguousarray(sig, dtype=np.uint32)
    cluster_of = np.ascontiguousarray(cluster_of, dtype=np.uint32)
    qp, qo = pack_many(qseqs)
    out = np.zeros((len(qseqs), 4), dtype=np.uint32)
    orc().orc_serve_assign(_p(packed), _p(word_off), len(word_off), _p(sig), _p(cluster_of),
                           _p(qp), _p(qo), len(qseqs), identity, new_base, first_qid, _p(out),
                           threads)
    return out


def makedb(fasta, packed, threads=0):
    return orc().orc_makedb(fasta.encode(), packed.encode(), threads)


def save_result(cluster, packed, result):
    cluster = np.ascontiguousarray(cluster, dtype=np.uint32)
    return orc().orc_save_result(_p(cluster), len(cluster), packed.encode(), result.encode())


def mt19937_64_nth(seed, nth):
    return orc().orc_mt19937_64_nth(seed, nth)


def read_db_numpy(path, bands=64):
    """Parse a packed DB file with numpy (independent of the product reader)."""
    d = np.fromfile(path, dtype=np.uint8)
    n = int(d[:4].view(np.uint32)[0])
    lens = d[4 + 4 * n:4 + 8 * n].view(np.uint32).copy()
    po = d[4 + 8 * n:4 + 16 * n].copy().view(np.uint64)
    to = d[4 + 16 * n:4 + 24 * n].copy().view(np.uint64)
    rep = d[4 + 24 * n:4 + 24 * n + 4 * 64 * n].view(np.uint32).reshape(64, n)[:bands].copy()
    packed = d[int(po[0]):int(to[0])].view(np.uint32).copy()
    woff = ((po - po[0]) // 4).astype(np.uint64)
    return {"n": n, "lengths": lens, "pack_off": po, "text_off": to, "rep": rep,
            "packed": packed, "word_off": woff, "raw": d}
