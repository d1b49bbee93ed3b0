"""MI355X-native nGIA2 clustering hot path -- Python plumbing over the C-ABI.

The product is the C-ABI library ``lib/libngia2.so`` (include/ngia2.h) and the
``bin/ngia2`` command line (with ``makeDB`` / ``cluster`` aliases).  This
module wraps the library with numpy arrays for tests and benchmarks; every
call goes to the HIP implementation and raises ``NgiaError`` on failure.
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import (BANDS, GRAPH_BATCHED, GRAPH_BATCHED_EAGER, GRAPH_BATCHED_TALLY, GRAPH_PLAIN,
                   NGIA_ERR_ARG, NGIA_ERR_IO, NGIA_ERR_NODEV, NGIA_ERR_STATE, NGIA_OK, BIN_DIR,
                   LIB_PATH, SHAPE_WINDOW, SHAPE_SEG16, SHAPE_SEG32, SHAPE_WAVE, SHAPE_SEG8, SHAPES,
                   LONG_DEFAULT, LONG_MULTI, LONG_SINGLE, LONG_SHAPES, LONG_WINDOWS)

__all__ = ["NgiaError", "Context", "Database", "PackedDB", "Serve", "Queries", "gen_config",
           "gen_fasta", "gen_packed", "gen_queries", "read_db", "leiden", "resolution",
           "save_result", "fasta_check", "edges_to_array", "BANDS", "GRAPH_PLAIN",
           "GRAPH_BATCHED", "BIN_DIR", "LIB_PATH", "ASSIGN_DTYPE", "NO_RESIDENT"]

NO_RESIDENT = _lib.NO_RESIDENT
ASSIGN_DTYPE = np.dtype([("cluster", np.uint32), ("resident", np.uint32), ("score", np.uint32),
                         ("candidates", np.uint32)])


class NgiaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (status %d)" % (msg, code))
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _check(rc, ctx=None):
    if rc != NGIA_OK:
        L = _lib.lib()
        msg = L.ngia_last_error(ctx) if ctx else None
        if not msg:
            msg = L.ngia_last_error(None)
        raise NgiaError(rc, (msg or b"").decode(errors="replace"))


EDGE_DTYPE = np.dtype([("u", np.int32), ("v", np.int32), ("w", np.float32)])


def edges_to_array(ptr, m):
    if m == 0:
        return np.zeros(0, dtype=EDGE_DTYPE)
    buf = C.string_at(ptr, m * 12)
    return np.frombuffer(buf, dtype=EDGE_DTYPE).copy()


class Context:
    """One HIP device and stream (ngia_open / ngia_close)."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = C.c_void_p()
        _check(L.ngia_open(device, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # databases and serving handles on this context

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            _lib.lib().ngia_close(self.h)
            self.h = None

    def tuning(self):
        """Shapes of the window bins by size (ngia_get_tuning): a dict with
        class_size (8 ascending bin sizes, 0 ends the list), shape /
        shape_compact, 17 x 8 lists of SHAPE_* for window width W and size
        class k (the 16-code table and compact alphabets), and long, 4 x 8
        lists of LONG_* for the groups of 17..60-word windows."""
        t = _lib.Tuning()
        _check(_lib.lib().ngia_get_tuning(self.h, C.byref(t)), self.h)
        return {"class_size": list(t.class_size),
                "shape": [list(t.shape[W]) for W in range(17)],
                "shape_compact": [list(t.shape_compact[W]) for W in range(17)],
                "long": [list(t.long_shape[g]) for g in range(_lib.LONG_GROUPS)]}

    def set_tuning(self, tuning):
        t = _lib.Tuning()
        for k, v in enumerate(tuning.get("class_size", [])):
            t.class_size[k] = int(v)
        for name in ("shape", "shape_compact"):
            for W, row in enumerate(tuning.get(name, [])):
                for k, v in enumerate(row):
                    getattr(t, name)[W][k] = int(v)
        for g, row in enumerate(tuning.get("long", [])):
            for k, v in enumerate(row):
                t.long_shape[g][k] = int(v)
        _check(_lib.lib().ngia_set_tuning(self.h, C.byref(t)), self.h)

    def route_all(self, shape):
        """Send every window bin, whatever its size, to one shape (SHAPE_SEG16
        falls back to 32 lanes at W = 16, SHAPE_SEG8 to 16 and 32 lanes at W =
        15 and 16); SHAPE_WINDOW restores the default."""
        def row(W):
            sh = shape
            if shape == _lib.SHAPE_SEG8 and W > 14:
                sh = _lib.SHAPE_SEG16 if W == 15 else _lib.SHAPE_SEG32
            if sh == _lib.SHAPE_SEG16 and W == 16:
                sh = _lib.SHAPE_SEG32
            return [sh] * 8
        self.set_tuning({"class_size": [1, 0, 0, 0, 0, 0, 0, 0],
                         "shape": [row(W) for W in range(17)],
                         "shape_compact": [row(W) for W in range(17)]})

    def autotune(self, identity):
        """Time every window-bin shape per window width and bin size, proteins
        then DNA, and install the fastest (ngia_autotune); returns the points."""
        cap = 512
        pts = (_lib.TunePoint * cap)()
        n = C.c_uint32()
        _check(_lib.lib().ngia_autotune(self.h, identity, pts, cap, C.byref(n)), self.h)
        # points of the 17..60-word groups (width > 16) time two shapes:
        # several words per lane (ms[0]) and one (ms[1])
        return [{"width": p.width, "dna": p.dna, "length": p.length, "pairs": p.pairs,
                 "ms": ({_lib.SHAPES[k]: p.ms[k] for k in range(5) if p.ms[k] > 0} if p.width <= 16
                        else {"multi": p.ms[0], "single": p.ms[1]})}
                for p in pts[:min(n.value, cap)]]

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        d = _lib.DeviceInfo()
        _check(_lib.lib().ngia_device_query(self.h, C.byref(d)), self.h)
        return {"device": d.device, "cu_count": d.cu_count, "clock_khz": d.clock_khz,
                "warp_size": d.warp_size, "hbm_total": d.hbm_total, "hbm_free": d.hbm_free,
                "name": d.name.decode(), "arch": d.arch.decode()}

    def timing(self):
        t = _lib.Timing()
        _check(_lib.lib().ngia_get_timing(self.h, C.byref(t)), self.h)
        return {k: getattr(t, k) for k, _ in t._fields_}

    def sync(self):
        _check(_lib.lib().ngia_sync(self.h), self.h)

    # ---- the three reference seams, host pointers
    def sketch(self, packed, word_off):
        packed, word_off = _u32(packed), _u64(word_off)
        n = len(word_off)
        sig = np.zeros((n, BANDS), dtype=np.uint32)
        _check(_lib.lib().ngia_sketch(self.h, _ptr(packed), len(packed), _ptr(word_off), n,
                                      _ptr(sig)), self.h)
        return sig

    def bucket(self, sig):
        sig = _u32(sig)
        n = sig.shape[0]
        rep = np.zeros((BANDS, n), dtype=np.uint32)
        _check(_lib.lib().ngia_bucket(self.h, _ptr(sig), n, _ptr(rep)), self.h)
        return rep

    def align(self, packed, word_off, jobs, threshold):
        """kernel_dynamic contract: jobs (count,3) uint32, column 2 overwritten."""
        packed, word_off = _u32(packed), _u64(word_off)
        jobs = np.array(jobs, dtype=np.uint32, order="C", copy=True)
        _check(_lib.lib().ngia_align(self.h, _ptr(packed), len(packed), _ptr(word_off),
                                     len(word_off), _ptr(jobs), len(jobs),
                                     C.c_float(threshold)), self.h)
        return jobs

    def build_graph(self, packed, word_off, lengths, rep_rows, bands, identity,
                    mode=GRAPH_BATCHED):
        packed, word_off, lengths = _u32(packed), _u64(word_off), _u32(lengths)
        rep_rows = _u32(rep_rows)
        ep = C.POINTER(_lib.Edge)()
        m = C.c_uint64()
        st = _lib.GraphStats()
        _check(_lib.lib().ngia_build_graph(self.h, _ptr(packed), len(packed), _ptr(word_off),
                                           _ptr(lengths), len(lengths), _ptr(rep_rows), bands,
                                           identity, mode, C.byref(ep), C.byref(m),
                                           C.byref(st)), self.h)
        try:
            arr = edges_to_array(ep, m.value)
        finally:
            _lib.lib().ngia_free(ep)
        return arr, st.as_dict()

    def build_graph_chunked(self, packed, word_off, lengths, rep_rows, bands, identity,
                            block_words=0):
        """Edges of the whole database, aligned chunk by chunk (pairs of blocks
        of at most block_words packed words; 0 sizes them from free device
        memory).  Returns (edges, stats, chunks)."""
        packed, word_off, lengths = _u32(packed), _u64(word_off), _u32(lengths)
        rep_rows = _u32(rep_rows)
        ep = C.POINTER(_lib.Edge)()
        m = C.c_uint64()
        st = _lib.GraphStats()
        nc = C.c_uint32()
        _check(_lib.lib().ngia_build_graph_chunked(
            self.h, _ptr(packed), len(packed), _ptr(word_off), _ptr(lengths), len(lengths),
            _ptr(rep_rows), bands, identity, block_words, C.byref(ep), C.byref(m), C.byref(st),
            C.byref(nc)), self.h)
        try:
            arr = edges_to_array(ep, m.value)
        finally:
            _lib.lib().ngia_free(ep)
        return arr, st.as_dict(), nc.value

    def makedb(self, fasta, packed):
        _check(_lib.lib().ngia_makedb(self.h, fasta.encode(), packed.encode()), self.h)

    def cluster(self, packed, result, identity, bands=BANDS, mode=GRAPH_BATCHED):
        st = _lib.GraphStats()
        _check(_lib.lib().ngia_cluster(self.h, packed.encode(), result.encode(), identity, bands,
                                       mode, C.byref(st)), self.h)
        return st.as_dict()


class Database:
    """A device-resident packed database (ngia_db_*)."""

    def __init__(self, ctx, packed, word_off, lengths=None):
        self.ctx = ctx
        self._packed, self._word_off = _u32(packed), _u64(word_off)
        self._lengths = None if lengths is None else _u32(lengths)
        self.n = len(self._word_off)
        h = C.c_void_p()
        _check(_lib.lib().ngia_db_create(ctx.h, _ptr(self._packed), len(self._packed),
                                         _ptr(self._word_off), _ptr(self._lengths), self.n,
                                         C.byref(h)), ctx.h)
        self.h = h
        ctx._children.add(self)

    def close(self):
        if self.h:
            _lib.lib().ngia_db_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sketch(self):
        _check(_lib.lib().ngia_db_sketch(self.ctx.h, self.h), self.ctx.h)

    def bucket(self):
        _check(_lib.lib().ngia_db_bucket(self.ctx.h, self.h), self.ctx.h)

    def set_reps(self, rep_rows, bands):
        rep_rows = _u32(rep_rows)
        _check(_lib.lib().ngia_db_set_reps(self.ctx.h, self.h, _ptr(rep_rows), bands), self.ctx.h)

    def sigs(self):
        s = np.zeros((self.n, BANDS), dtype=np.uint32)
        _check(_lib.lib().ngia_db_get_sigs(self.ctx.h, self.h, _ptr(s)), self.ctx.h)
        return s

    def reps(self):
        r = np.zeros((BANDS, self.n), dtype=np.uint32)
        _check(_lib.lib().ngia_db_get_reps(self.ctx.h, self.h, _ptr(r)), self.ctx.h)
        return r

    def graph(self, identity, bands=BANDS, mode=GRAPH_BATCHED):
        st = _lib.GraphStats()
        _check(_lib.lib().ngia_db_graph(self.ctx.h, self.h, identity, bands, mode, C.byref(st)),
               self.ctx.h)
        return st.as_dict()

    def step(self, identity, bands=BANDS, split=True):
        """Sketch, bucketing and the batched graph in one call (ngia_db_step):
        the same signatures, rows and edges as sketch(), bucket() and
        graph(identity, bands); with split, the head's graph (records whose
        own band window spans 64 words or more) runs beside the rest's sketch."""
        st = _lib.GraphStats()
        _check(_lib.lib().ngia_db_step(self.ctx.h, self.h, identity, bands,
                                       0 if split else _lib.STEP_NO_SPLIT, C.byref(st)),
               self.ctx.h)
        return st.as_dict()

    def edges(self):
        m = C.c_uint64()
        L = _lib.lib()
        _check(L.ngia_db_get_edges(self.ctx.h, self.h, None, 0, C.byref(m)), self.ctx.h)
        out = np.zeros(m.value, dtype=EDGE_DTYPE)
        if m.value:
            _check(L.ngia_db_get_edges(self.ctx.h, self.h, _ptr(out), m.value, C.byref(m)),
                   self.ctx.h)
        return out

    def pair_scores(self):
        m = C.c_uint64()
        L = _lib.lib()
        _check(L.ngia_db_get_pair_scores(self.ctx.h, self.h, None, None, 0, C.byref(m)),
               self.ctx.h)
        keys = np.zeros(m.value, dtype=np.uint64)
        sc = np.zeros(m.value, dtype=np.uint32)
        if m.value:
            _check(L.ngia_db_get_pair_scores(self.ctx.h, self.h, _ptr(keys), _ptr(sc), m.value,
                                             C.byref(m)), self.ctx.h)
        return keys, sc

    def band_jobs(self, identity, band):
        """The plain form's scored job triplets of one band (ngia_db_band_jobs):
        (count, 3) uint32 rows (rep, tar, score) in the reference's job order."""
        jobs = np.zeros((max(self.n, 1), 3), dtype=np.uint32)
        c = C.c_uint64()
        _check(_lib.lib().ngia_db_band_jobs(self.ctx.h, self.h, identity, band, _ptr(jobs),
                                            self.n, C.byref(c)), self.ctx.h)
        return jobs[:c.value].copy()

    def bin_stats(self):
        """Length bins of the last batched graph: kind, width, pairs, ref_tiles, ms and
        run_tiles (the reference tiles of the blocks the kernels ran; counted by
        GRAPH_BATCHED_TALLY rounds only, else 0)."""
        L = _lib.lib()
        nb = C.c_uint32()
        _check(L.ngia_db_bin_stats(self.ctx.h, self.h, None, 0, C.byref(nb)), self.ctx.h)
        arr = (_lib.BinStat * nb.value)()
        _check(L.ngia_db_bin_stats(self.ctx.h, self.h, arr, nb.value, C.byref(nb)), self.ctx.h)
        return [{"kind": _lib.BIN_KINDS[b.kind], "width": b.width, "pairs": b.pairs,
                 "ref_tiles": b.ref_tiles, "ms": b.ms, "run_tiles": b.run_tiles} for b in arr]

    def fast(self, identity, bands=BANDS):
        cl = np.zeros(self.n, dtype=np.uint32)
        tot = C.c_uint64()
        _check(_lib.lib().ngia_db_fast(self.ctx.h, self.h, identity, bands, _ptr(cl),
                                       C.byref(tot)), self.ctx.h)
        return cl, tot.value


class Queries:
    """Residue strings for serving: query q is res[off[q]:off[q] + length[q]]."""

    def __init__(self, res, off, length, kind=None):
        self.res = np.ascontiguousarray(res, dtype=np.uint8)
        self.off = _u64(off)
        self.length = _u32(length)
        self.kind = None if kind is None else np.ascontiguousarray(kind, dtype=np.uint8)
        self.n = len(self.length)

    @classmethod
    def from_strings(cls, seqs):
        seqs = [bytes(x) for x in seqs]
        length = np.array([len(x) for x in seqs], dtype=np.uint32)
        off = np.zeros(len(seqs), dtype=np.uint64)
        if len(seqs) > 1:
            off[1:] = np.cumsum(length[:-1], dtype=np.uint64)
        res = np.frombuffer(b"".join(seqs) or b"\0", dtype=np.uint8)
        return cls(res, off, length)

    def seq(self, q):
        o = int(self.off[q])
        return self.res[o:o + int(self.length[q])].tobytes()


class Serve:
    """A resident clustered snapshot serving query batches (ngia_serve_*)."""

    def __init__(self, db, cluster_of, identity, new_base, lanes=0, max_queries=0,
                 max_residues=0, hbm_fraction=0.0, grow_queries=0, grow_residues=0):
        self.ctx = db.ctx
        self._cluster_of = _u32(cluster_of)
        assert len(self._cluster_of) == db.n
        cfg = _lib.ServeConfig(lanes, max_queries, max_residues, hbm_fraction, grow_queries,
                               grow_residues)
        h = C.c_void_p()
        _check(_lib.lib().ngia_serve_create(self.ctx.h, db.h, _ptr(self._cluster_of), identity,
                                            new_base, C.byref(cfg), C.byref(h)), self.ctx.h)
        self.h = h
        self.ctx._children.add(self)

    def close(self):
        if self.h:
            _lib.lib().ngia_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        """A new session: the snapshot returns to the residents (ngia_serve_reset)."""
        _check(_lib.lib().ngia_serve_reset(self.h), self.ctx.h)

    def assign(self, queries, first_qid=0):
        """All queries at once (the offline run): structured array of ASSIGN_DTYPE."""
        out = np.zeros(queries.n, dtype=ASSIGN_DTYPE)
        st = _lib.ServeStats()
        _check(_lib.lib().ngia_serve_assign(self.h, _ptr(queries.res), _ptr(queries.off),
                                            _ptr(queries.length), queries.n, first_qid,
                                            _ptr(out), C.byref(st)), self.ctx.h)
        return out, st.as_dict()

    def replay(self, queries, arrival, first, count):
        """Real-time replay of requests (arrival seconds, query ranges): assignments,
        per-request latency in seconds and the run's statistics."""
        nr = len(arrival)
        reqs = (_lib.Request * max(nr, 1))()
        for r in range(nr):
            reqs[r].arrival = float(arrival[r])
            reqs[r].first = int(first[r])
            reqs[r].count = int(count[r])
        out = np.zeros(queries.n, dtype=ASSIGN_DTYPE)
        lat = np.zeros(max(nr, 1), dtype=np.float64)
        st = _lib.ServeStats()
        _check(_lib.lib().ngia_serve_replay(self.h, _ptr(queries.res), _ptr(queries.off),
                                            _ptr(queries.length), queries.n, reqs, nr, _ptr(out),
                                            _ptr(lat), C.byref(st)), self.ctx.h)
        return out, lat[:nr], st.as_dict()


def gen_queries(packed, word_off, n_queries, mutated=0.7, sub=(0.0, 0.3), indel=0.01,
                len_median=300.0, len_sigma=0.5, len_min=50, len_max=1000, seed=1):
    """Serving queries from a resident packed area (ngia_gen_queries)."""
    packed, word_off = _u32(packed), _u64(word_off)
    p = _lib.QueryParams(n_queries, mutated, sub[0], sub[1], indel, len_median, len_sigma,
                         len_min, len_max, seed)
    res, off, ln, kind = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    L = _lib.lib()
    _check(L.ngia_gen_queries(_ptr(packed), _ptr(word_off), len(word_off), C.byref(p),
                              C.byref(res), C.byref(off), C.byref(ln), C.byref(kind)))
    try:
        length = np.ctypeslib.as_array(C.cast(ln, C.POINTER(C.c_uint32)), (max(n_queries, 1),))[:n_queries].copy()
        offs = np.ctypeslib.as_array(C.cast(off, C.POINTER(C.c_uint64)), (max(n_queries, 1),))[:n_queries].copy()
        kd = np.ctypeslib.as_array(C.cast(kind, C.POINTER(C.c_uint8)), (max(n_queries, 1),))[:n_queries].copy()
        total = int(length.astype(np.int64).sum())
        r = np.ctypeslib.as_array(C.cast(res, C.POINTER(C.c_uint8)), (max(total, 1),))[:max(total, 1)].copy()
    finally:
        for h in (res, off, ln, kind):
            L.ngia_free(h)
    return Queries(r, offs, length, kd)


class PackedDB:
    """A packed DB in host memory (ngia_packed_db), copied into numpy arrays."""

    @classmethod
    def _from_raw(cls, raw, bands):
        db = cls()
        n = raw.n

        def arr(p, count):
            if count == 0 or not p:
                return None
            return np.ctypeslib.as_array(p, shape=(count,)).copy()

        db.n = n
        db.name_len = arr(raw.name_len, n)
        db.lengths = arr(raw.lengths, n)
        db.pack_off = arr(raw.pack_off, n)
        db.text_off = arr(raw.text_off, n)
        db.word_off = arr(raw.word_off, n)
        db.packed = arr(raw.packed, raw.n_words)
        rep = arr(raw.rep_rows, bands * n) if bands else None
        db.rep_rows = None if rep is None else rep.reshape(bands, n)
        _lib.lib().ngia_release_db(C.byref(raw))
        return db


def read_db(path, bands=BANDS):
    raw = _lib.PackedDB()
    _check(_lib.lib().ngia_read_db(path.encode(), bands, C.byref(raw)))
    return PackedDB._from_raw(raw, bands)


def plan_blocks(lengths, block_words, block_records=0):
    """Block cut of build_graph_chunked (host only): the list of block bounds."""
    lengths = _u32(lengths)
    bounds = np.zeros(len(lengths) + 1, dtype=np.uint32)
    nb = C.c_uint32()
    _check(_lib.lib().ngia_plan_blocks(_ptr(lengths), len(lengths), block_words, block_records,
                                       _ptr(bounds), C.byref(nb)))
    return bounds[:nb.value + 1].copy()


def gen_config(config, scale=1.0, seed=1):
    p = _lib.GenParams()
    _check(_lib.lib().ngia_gen_config(config, scale, seed, C.byref(p)))
    return p


def gen_fasta(params, path):
    n = _lib.lib().ngia_gen_fasta(C.byref(params), path.encode())
    if n < 0:
        _check(int(n))
    return int(n)


def gen_packed(params):
    raw = _lib.PackedDB()
    _check(_lib.lib().ngia_gen_packed(C.byref(params), C.byref(raw)))
    return PackedDB._from_raw(raw, 0)


def leiden(edges, n, resolution_, seed=42, iters=1):
    edges = np.ascontiguousarray(edges, dtype=EDGE_DTYPE)
    out = np.zeros(n, dtype=np.uint32)
    _check(_lib.lib().ngia_leiden(_ptr(edges), len(edges), n, C.c_float(resolution_), seed, iters,
                                  _ptr(out)))
    return out


def resolution(identity):
    return _lib.lib().ngia_resolution(identity)


def save_result(cluster, packed_path, result_path):
    cluster = _u32(cluster)
    _check(_lib.lib().ngia_save_result(_ptr(cluster), len(cluster), packed_path.encode(),
                                       result_path.encode()))


def fasta_check(path):
    rc = _lib.lib().ngia_fasta_check(path.encode())
    if rc != NGIA_OK:
        return _lib.lib().ngia_last_error(None).decode()
    return None
