"""ctypes binding of the C-ABI library ``ngia2_amd/lib/libngia2.so`` (include/ngia2.h).

The library is built in-tree by ``ngia2_amd/csrc/Makefile``.  Loading fails
loudly when it is missing: there is no fallback implementation.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libngia2.so")
BIN_DIR = os.path.join(HERE, "bin")

NGIA_OK = 0
NGIA_ERR_ARG = -1
NGIA_ERR_HIP = -2
NGIA_ERR_NOMEM = -3
NGIA_ERR_IO = -4
NGIA_ERR_STATE = -5
NGIA_ERR_NODEV = -6
GRAPH_PLAIN = 0
GRAPH_BATCHED = 1
GRAPH_BATCHED_EAGER = 2
GRAPH_BATCHED_TALLY = 3
STEP_NO_SPLIT = 1
BANDS = 64


class Edge(C.Structure):
    _fields_ = [("u", C.c_int32), ("v", C.c_int32), ("w", C.c_float)]


class GraphStats(C.Structure):
    _fields_ = [
        ("total_pairs", C.c_uint64),
        ("distinct_pairs", C.c_uint64),
        ("n_edges", C.c_uint64),
        ("ref_tiles", C.c_uint64),
        ("ms_candidates", C.c_double),
        ("ms_align", C.c_double),
        ("ms_edges", C.c_double),
        ("ms_total", C.c_double),
        ("align_launches", C.c_uint32),
        ("host_syncs", C.c_uint32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BinStat(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("width", C.c_uint32), ("pairs", C.c_uint64),
                ("ref_tiles", C.c_uint64), ("ms", C.c_double), ("run_tiles", C.c_uint64)]


BIN_KINDS = {0: "zero", 1: "window", 2: "wave", 3: "plain", 4: "seg", 5: "seg2", 6: "seg3",
             7: "seg4"}

SHAPE_WINDOW, SHAPE_SEG16, SHAPE_SEG32, SHAPE_WAVE, SHAPE_SEG8 = 0, 1, 2, 3, 4
SHAPES = ("window", "seg16", "seg32", "wave", "seg8")
TUNE_CLASSES = 8
# groups of 17..60-word windows (17-21, 22-30, 32-45, 46-60) and their shapes
LONG_GROUPS = 4
LONG_WINDOWS = (19, 26, 38, 53)  # the window each group is tuned at
LONG_DEFAULT, LONG_MULTI, LONG_SINGLE = 0, 1, 2
LONG_SHAPES = ("default", "multi", "single")


class Tuning(C.Structure):
    _fields_ = [("class_size", C.c_uint32 * TUNE_CLASSES),
                ("shape", (C.c_uint32 * TUNE_CLASSES) * 17),
                ("shape_compact", (C.c_uint32 * TUNE_CLASSES) * 17),
                ("long_shape", (C.c_uint32 * TUNE_CLASSES) * LONG_GROUPS)]


class TunePoint(C.Structure):
    _fields_ = [("width", C.c_uint32), ("dna", C.c_uint32), ("length", C.c_uint32),
                ("pairs", C.c_uint32), ("ms", C.c_double * 5)]


class Timing(C.Structure):
    _fields_ = [("sketch_ms", C.c_double), ("bucket_ms", C.c_double), ("graph_ms", C.c_double),
                ("finalize_ms", C.c_double), ("align_ms", C.c_double), ("step_ms", C.c_double),
                ("step_head", C.c_uint32), ("reserved", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("cu_count", C.c_int32),
        ("clock_khz", C.c_int32),
        ("warp_size", C.c_int32),
        ("hbm_total", C.c_uint64),
        ("hbm_free", C.c_uint64),
        ("name", C.c_char * 128),
        ("arch", C.c_char * 64),
    ]


class PackedDB(C.Structure):
    _fields_ = [
        ("n", C.c_uint32),
        ("name_len", C.POINTER(C.c_uint32)),
        ("lengths", C.POINTER(C.c_uint32)),
        ("pack_off", C.POINTER(C.c_uint64)),
        ("text_off", C.POINTER(C.c_uint64)),
        ("rep_rows", C.POINTER(C.c_uint32)),
        ("packed", C.POINTER(C.c_uint32)),
        ("n_words", C.c_uint64),
        ("word_off", C.POINTER(C.c_uint64)),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("n_families", C.c_uint32),
        ("members", C.c_uint32),
        ("dna", C.c_uint32),
        ("len_normal", C.c_uint32),
        ("len_a", C.c_double),
        ("len_b", C.c_double),
        ("len_min", C.c_uint32),
        ("len_max", C.c_uint32),
        ("tail_frac", C.c_double),
        ("tail_xmin", C.c_double),
        ("tail_alpha", C.c_double),
        ("tail_max", C.c_uint32),
        ("sub_lo", C.c_double),
        ("sub_hi", C.c_double),
        ("indel_lo", C.c_double),
        ("indel_hi", C.c_double),
        ("seed", C.c_uint64),
    ]


class Assignment(C.Structure):
    _fields_ = [("cluster", C.c_uint32), ("resident", C.c_uint32), ("score", C.c_uint32),
                ("candidates", C.c_uint32)]


class ServeConfig(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("max_queries", C.c_uint32),
                ("max_residues", C.c_uint64), ("hbm_fraction", C.c_double),
                ("grow_queries", C.c_uint32), ("grow_residues", C.c_uint64)]


class Request(C.Structure):
    _fields_ = [("arrival", C.c_double), ("first", C.c_uint32), ("count", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("queries", C.c_uint64), ("requests", C.c_uint64), ("batches", C.c_uint64),
                ("pairs", C.c_uint64), ("new_clusters", C.c_uint64), ("lanes", C.c_uint32),
                ("max_queries", C.c_uint32), ("max_residues", C.c_uint64),
                ("wall_s", C.c_double), ("device_s", C.c_double), ("grown", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class QueryParams(C.Structure):
    _fields_ = [("n_queries", C.c_uint32), ("mutated", C.c_double), ("sub_lo", C.c_double),
                ("sub_hi", C.c_double), ("indel", C.c_double), ("len_median", C.c_double),
                ("len_sigma", C.c_double), ("len_min", C.c_uint32), ("len_max", C.c_uint32),
                ("seed", C.c_uint64)]


NO_RESIDENT = 0xFFFFFFFF

P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
I32 = C.c_int32

# name -> (restype, argtypes); every entry point declared in include/ngia2.h
PROTOTYPES = {
    "ngia_open": (C.c_int, [C.c_int, C.POINTER(P)]),
    "ngia_close": (None, [P]),
    "ngia_last_error": (C.c_char_p, [P]),
    "ngia_device_query": (C.c_int, [P, C.POINTER(DeviceInfo)]),
    "ngia_sync": (C.c_int, [P]),
    "ngia_get_timing": (C.c_int, [P, C.POINTER(Timing)]),
    "ngia_free": (None, [P]),
    "ngia_sketch": (C.c_int, [P, P, U64, P, U32, P]),
    "ngia_bucket": (C.c_int, [P, P, U32, P]),
    "ngia_align": (C.c_int, [P, P, U64, P, U32, P, U32, C.c_float]),
    "ngia_align_noalloc": (C.c_int, [P, P, P, P, U32, C.c_float, U32, P]),
    "ngia_build_graph": (C.c_int, [P, P, U64, P, P, U32, P, U32, U32, U32,
                                   C.POINTER(C.POINTER(Edge)), C.POINTER(U64),
                                   C.POINTER(GraphStats)]),
    "ngia_build_graph_chunked": (C.c_int, [P, P, U64, P, P, U32, P, U32, U32, U64,
                                           C.POINTER(C.POINTER(Edge)), C.POINTER(U64),
                                           C.POINTER(GraphStats), C.POINTER(U32)]),
    "ngia_plan_blocks": (C.c_int, [P, U32, U64, U32, P, C.POINTER(U32)]),
    "ngia_db_create": (C.c_int, [P, P, U64, P, P, U32, C.POINTER(P)]),
    "ngia_db_destroy": (None, [P]),
    "ngia_db_sketch": (C.c_int, [P, P]),
    "ngia_db_bucket": (C.c_int, [P, P]),
    "ngia_db_set_reps": (C.c_int, [P, P, P, U32]),
    "ngia_db_get_sigs": (C.c_int, [P, P, P]),
    "ngia_db_get_reps": (C.c_int, [P, P, P]),
    "ngia_db_graph": (C.c_int, [P, P, U32, U32, U32, C.POINTER(GraphStats)]),
    "ngia_db_step": (C.c_int, [P, P, U32, U32, U32, C.POINTER(GraphStats)]),
    "ngia_db_get_edges": (C.c_int, [P, P, P, U64, C.POINTER(U64)]),
    "ngia_db_get_pair_scores": (C.c_int, [P, P, P, P, U64, C.POINTER(U64)]),
    "ngia_db_band_jobs": (C.c_int, [P, P, U32, U32, P, U64, C.POINTER(U64)]),
    "ngia_db_bin_stats": (C.c_int, [P, P, P, U32, C.POINTER(U32)]),
    "ngia_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "ngia_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "ngia_autotune": (C.c_int, [P, U32, P, U32, C.POINTER(U32)]),
    "ngia_db_fast": (C.c_int, [P, P, U32, U32, P, C.POINTER(U64)]),
    "ngia_serve_create": (C.c_int, [P, P, P, U32, U32, C.POINTER(ServeConfig), C.POINTER(P)]),
    "ngia_serve_destroy": (None, [P]),
    "ngia_serve_reset": (C.c_int, [P]),
    "ngia_serve_assign": (C.c_int, [P, P, P, P, U32, U64, P, C.POINTER(ServeStats)]),
    "ngia_serve_replay": (C.c_int, [P, P, P, P, U32, P, U32, P, P, C.POINTER(ServeStats)]),
    "ngia_leiden": (C.c_int, [P, U64, I32, C.c_float, U32, I32, P]),
    "ngia_resolution": (C.c_float, [U32]),
    "ngia_save_result": (C.c_int, [P, U32, C.c_char_p, C.c_char_p]),
    "ngia_fasta_check": (C.c_int, [C.c_char_p]),
    "ngia_makedb": (C.c_int, [P, C.c_char_p, C.c_char_p]),
    "ngia_cluster": (C.c_int, [P, C.c_char_p, C.c_char_p, U32, U32, U32,
                               C.POINTER(GraphStats)]),
    "ngia_read_db": (C.c_int, [C.c_char_p, U32, C.POINTER(PackedDB)]),
    "ngia_release_db": (None, [C.POINTER(PackedDB)]),
    "ngia_gen_config": (C.c_int, [C.c_int, C.c_double, U64, C.POINTER(GenParams)]),
    "ngia_gen_fasta": (C.c_int64, [C.POINTER(GenParams), C.c_char_p]),
    "ngia_gen_packed": (C.c_int, [C.POINTER(GenParams), C.POINTER(PackedDB)]),
    "ngia_gen_queries": (C.c_int, [P, P, U32, C.POINTER(QueryParams), C.POINTER(P), C.POINTER(P),
                                   C.POINTER(P), C.POINTER(P)]),
}

_lib = None


def lib():
    """Load libngia2.so once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libngia2.so is not built (%s); run __graft_entry__.build() or "
                "make -C ngia2_amd/csrc" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib
