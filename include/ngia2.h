/*
 * ngia2.h -- C-ABI of the MI355X-native nGIA2 clustering hot path.
 *
 * The drop-in boundary.  Plain pointers and sizes, int status codes (0 = ok,
 * negative = error, message via ngia_last_error), no C++ exceptions and no
 * torch types cross it.  Every entry point that has a counterpart in the
 * reference names it (paths relative to the nGIA2 source tree).
 *
 * Ownership: the caller owns every host buffer it passes; a context owns the
 * device memory and the HIP stream it creates; arrays returned through a
 * pointer-to-pointer are released with ngia_free.  One context per host
 * thread; calls on one context are serialised by the caller.
 *
 * Integer results (signatures, representative rows, candidate pairs, scores,
 * edges) are bit-identical to the reference algorithm on the same inputs.
 */
#ifndef NGIA2_H
#define NGIA2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGIA_BANDS 64        /* GROUP_BANDS, makedb.hpp:45 */
#define NGIA_MAX_LEN 0xFFFEu /* reads >= 0xFFFF are dropped, makedb.hpp:118 */

/* status codes */
#define NGIA_OK 0
#define NGIA_ERR_ARG (-1)    /* bad argument (size, range, null pointer) */
#define NGIA_ERR_HIP (-2)    /* HIP runtime / kernel failure */
#define NGIA_ERR_NOMEM (-3)  /* device or host allocation failed */
#define NGIA_ERR_IO (-4)     /* file read/write failure */
#define NGIA_ERR_STATE (-5)  /* call out of order (e.g. graph before reps) */
#define NGIA_ERR_NODEV (-6)  /* no HIP device */

typedef struct ngia_ctx ngia_ctx;
typedef struct ngia_db ngia_db;
typedef struct ngia_serve ngia_serve;

/* One graph edge, the reference's std::tuple<int,int,float> (buildGraph.hpp:256). */
typedef struct {
  int32_t u;   /* representative index (u < v) */
  int32_t v;   /* target index */
  float w;     /* (float)score / 100.0f, score in 1..100 */
} ngia_edge;

/* Graph-build modes (flags of ngia_db_graph / ngia_build_graph). */
#define NGIA_GRAPH_PLAIN 0u    /* one align launch + host sync per band (buildGraph.hpp:274-350) */
#define NGIA_GRAPH_BATCHED 1u  /* cross-band dedupe, length-binned kernels, captured hipGraphs */
#define NGIA_GRAPH_BATCHED_EAGER 2u /* the batched kernels launched without the graph, with
                                     * the alignment phase timed (ngia_timing.align_ms) */
#define NGIA_GRAPH_BATCHED_TALLY 3u /* as EAGER, and each alignment kernel also counts the
                                     * reference tiles of the blocks it ran (ngia_bin_stat.
                                     * run_tiles); the counting slows that call */

typedef struct {
  uint64_t total_pairs;     /* candidates summed over bands, duplicates included ("compare pairs", buildGraph.hpp:353) */
  uint64_t distinct_pairs;  /* distinct (rep,tar) pairs actually aligned (batched); = total_pairs in plain mode */
  uint64_t n_edges;         /* sorted unique edges */
  uint64_t ref_tiles;       /* reference 32x32 tiles over the aligned pairs */
  double ms_candidates;     /* device time per stage, milliseconds (HIP events) */
  double ms_align;
  double ms_edges;
  double ms_total;
  uint32_t align_launches;  /* alignment kernel launches */
  uint32_t host_syncs;      /* host<->device synchronisations inside the call */
} ngia_graph_stats;

typedef struct {
  int32_t device;
  int32_t cu_count;
  int32_t clock_khz;
  int32_t warp_size;
  uint64_t hbm_total;
  uint64_t hbm_free;
  char name[128];
  char arch[64];
} ngia_device_info;

/* ---------------------------------------------------------------- context */
/* Replaces the device set-up in BuildGraph::check (buildGraph.hpp:100-116). */
int ngia_open(int device, ngia_ctx** ctx);
void ngia_close(ngia_ctx* ctx);
/* Last error message of ctx (or of the calling thread when ctx is NULL). */
const char* ngia_last_error(const ngia_ctx* ctx);
int ngia_device_query(ngia_ctx* ctx, ngia_device_info* info);
int ngia_sync(ngia_ctx* ctx);
/* Device time (HIP events on the context stream) of the most recent stages.
 * No reference counterpart: replaces the reference's wall-clock phase timers
 * (buildGraph.hpp:272, 352-361). */
typedef struct {
  double sketch_ms;   /* k_sketch of the last ngia_db_sketch */
  double bucket_ms;   /* transpose + radix sort + predecessor of the last ngia_db_bucket */
  double graph_ms;    /* candidates .. raw edges of the last ngia_db_graph */
  double finalize_ms; /* sort + unique of the last ngia_db_graph */
  double align_ms;    /* alignment bins alone (NGIA_GRAPH_BATCHED_EAGER only, else 0) */
  double step_ms;     /* the whole of the last ngia_db_step */
  uint32_t step_head; /* records in the head round of the last ngia_db_step (0: none) */
  uint32_t reserved;
} ngia_timing;
int ngia_get_timing(ngia_ctx* ctx, ngia_timing* out);
void ngia_free(void* p);

/* ------------------------------------------- the three reference seams */
/* (1a) minhash sketch: the packData inner loop (makedb.hpp:221-247) over all
 * records.  packed: the packed area (records "L, ceil(L/32) x 4 planes");
 * word_off[i]: uint32 word offset of record i inside it (buildGraph.hpp:173-177).
 * sig: n x 64 uint32, sequence-major (the layout packData writes at signBase). */
int ngia_sketch(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                const uint64_t* word_off, uint32_t n, uint32_t* sig);

/* (1b) band bucketing: MakeDB::groupSignatures (makedb.hpp:289-364).
 * sig as above; rep: 64 x n uint32, band-major, rep[b][i] = previous index with
 * the same band-b signature, or i for a bucket head. */
int ngia_bucket(ngia_ctx* ctx, const uint32_t* sig, uint32_t n, uint32_t* rep);

/* (2) banded bit-parallel LCS: kernel_dynamic (buildGraph.hpp:192-253), same
 * contract: jobs holds count triplets (rep, tar, score); score is written with
 * lcs*100/(R+T-lcs) when lcs > (R+T-lcs)*threshold, else 0.  n_seqs bounds the
 * rep/tar indices (checked). */
int ngia_align(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
               const uint64_t* word_off, uint32_t n_seqs, uint32_t* jobs,
               uint32_t count, float threshold);

/* (2') The same contract on device memory, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc).
 * packed, word_off and jobs are device pointers (the packed area, the uint32
 * word offset of each record inside it, count job triplets); the scores are
 * written into jobs in place, asynchronously on `stream` (a hipStream_t; NULL
 * = the context's stream), and nothing is synchronised.  max_len (at most
 * NGIA_MAX_LEN, else NGIA_ERR_ARG) is the longest rep among the jobs: it sizes
 * the per-thread LCS rows; a longer rep is still aligned exactly, on fewer
 * threads, and a rep of 65,535 residues or more is left unaligned (its score
 * keeps the caller's value).  The rows live in a static device array of
 * NGIA_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and shared
 * by every call on that GPU: calls that overlap on one GPU (other streams,
 * contexts or threads) race and must be serialised by the caller. */
#define NGIA_NOALLOC_SCRATCH_BYTES (32ull << 20)
int ngia_align_noalloc(ngia_ctx* ctx, const uint32_t* packed, const uint64_t* word_off,
                       uint32_t* jobs, uint32_t count, float threshold, uint32_t max_len,
                       void* stream);

/* (3) edge emission: BuildGraph::buildGraph (buildGraph.hpp:256-363).
 * rep_rows: the first `bands` band-major rows; identity 1..99; bands 1..64.
 * *edges receives the sorted unique edge list (release with ngia_free). */
int ngia_build_graph(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                     const uint64_t* word_off, const uint32_t* lengths, uint32_t n,
                     const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                     uint32_t flags, ngia_edge** edges, uint64_t* n_edges,
                     ngia_graph_stats* stats);

/* (3') edge emission for a database larger than the device holds.  No kernel
 * counterpart: the reference keeps the packed area in managed memory and
 * lets the driver page it (README.md:8, buildGraph.hpp:143, 164).  The
 * records are cut into blocks of consecutive indices of at most block_words
 * packed words each (a longer record is a block by itself); every candidate
 * pair (rep r, target t) is aligned in the chunk of r's block and t's block,
 * which is all that is uploaded for it, in the batched form.  The edge list
 * equals ngia_build_graph's.  block_words = 0 sizes blocks from the free
 * device memory (two blocks within half of it), which is one block when the
 * whole database fits.  *n_chunks (may be NULL) receives the chunks run. */
int ngia_build_graph_chunked(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                             const uint64_t* word_off, const uint32_t* lengths, uint32_t n,
                             const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                             uint64_t block_words, ngia_edge** edges, uint64_t* n_edges,
                             ngia_graph_stats* stats, uint32_t* n_chunks);
/* The block cut of ngia_build_graph_chunked: records in index order, a new
 * block when the next record would take the block past block_words packed
 * words (1 + 4 ceil(L/32) per record) or past block_records records
 * (0 = no limit).  bounds holds n + 1 entries; block k is
 * [bounds[k], bounds[k+1]).  Host only. */
int ngia_plan_blocks(const uint32_t* lengths, uint32_t n, uint64_t block_words,
                     uint32_t block_records, uint32_t* bounds, uint32_t* n_blocks);

/* ------------------------------------------- device-resident database */
/* Upload the packed area, word offsets and lengths once (BuildGraph::readData,
 * buildGraph.hpp:125-183, with explicit device buffers instead of managed
 * memory).  Everything below works on the resident copy. */
int ngia_db_create(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                   const uint64_t* word_off, const uint32_t* lengths, uint32_t n,
                   ngia_db** db);
void ngia_db_destroy(ngia_db* db);
/* Sketch + bucket on the device (the makeDB half of the path). */
int ngia_db_sketch(ngia_ctx* ctx, ngia_db* db);
int ngia_db_bucket(ngia_ctx* ctx, ngia_db* db);
/* Install representative rows read from a packed DB file instead. */
int ngia_db_set_reps(ngia_ctx* ctx, ngia_db* db, const uint32_t* rep_rows, uint32_t bands);
int ngia_db_get_sigs(ngia_ctx* ctx, ngia_db* db, uint32_t* sig);   /* n x 64 */
int ngia_db_get_reps(ngia_ctx* ctx, ngia_db* db, uint32_t* rep);   /* 64 x n */
/* Build the edge list on the device; it stays resident until the next call. */
int ngia_db_graph(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                  uint32_t flags, ngia_graph_stats* stats);

/* Sketch, bucketing and the batched graph of the whole database in one call:
 * the signatures, rows and sorted unique edges of ngia_db_sketch,
 * ngia_db_bucket and ngia_db_graph(NGIA_GRAPH_BATCHED) in sequence.  No
 * reference counterpart (the reference runs makeDB and clustering as two
 * programs, main.cu:62-117).  The records whose own band window spans at
 * least 64 words at this identity (the leading run: makeDB orders records
 * longest first) form a head; its graph, whose pairs depend on the head alone,
 * runs beside the sketch of the rest, and the main round takes the other
 * targets.  NGIA_STEP_NO_SPLIT runs the three stages one after another.
 * Bin figures (ngia_db_bin_stats) and pair scores then cover the main
 * round only. */
#define NGIA_STEP_NO_SPLIT 1u
int ngia_db_step(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands, uint32_t flags,
                 ngia_graph_stats* st);
int ngia_db_get_edges(ngia_ctx* ctx, ngia_db* db, ngia_edge* out, uint64_t cap,
                      uint64_t* n_edges);
/* Scores of the distinct pairs of the last batched graph, sorted by (rep,tar):
 * keys[i] = rep<<32 | tar.  Either array may be NULL. */
int ngia_db_get_pair_scores(ngia_ctx* ctx, ngia_db* db, uint64_t* keys, uint32_t* scores,
                            uint64_t cap, uint64_t* n_pairs);

/* One band of the plain form (buildGraph.hpp:274-333): the band's job
 * triplets (rep[i], i, score) for ascending i with rep[i] != i and the isJob
 * length test (buildGraph.hpp:262-300, 305-321), scored by the kernel_dynamic
 * contract (buildGraph.hpp:192-253), in the order the reference's jobs array
 * holds them after the launch.  *count receives the band's job count (at most
 * n); jobs (3 * count uint32, may be NULL) receives the triplets when cap >=
 * count.  Needs representative rows for the band (ngia_db_bucket or
 * ngia_db_set_reps). */
int ngia_db_band_jobs(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t band,
                      uint32_t* jobs, uint64_t cap, uint64_t* count);

/* Length bins of the last batched graph.  No reference counterpart (the
 * reference aligns every band's jobs in one kernel_dynamic launch).  A pair's
 * bin follows from its band window W = lsft/32 + ceil(rsft/32) words:
 * W <= 16 -> a window bin, aligned by the shape the tuning picks for its size
 * (one thread per pair by default); W <= 21 -> 8 lanes per pair with three
 * words per lane, eight pairs per wavefront; W <= 30 -> 16 lanes per pair with
 * two words per lane, four pairs per wavefront; W = 31 ->
 * 32 lanes per pair, two pairs per wavefront; W <= 45 and W <= 60 -> 16 lanes
 * per pair with three and four words per lane; W <= 64*K - K + 1 -> a
 * wavefront per pair with K words per lane; wider -> the plain kernel;
 * length-rejected or empty targets score 0 without a kernel.  A bin of 17..30
 * words holding fewer than 32,768 pairs goes to 32 lanes per pair, one of
 * 32..60 words to a wavefront per pair (NGIA_SEG2_MIN in the environment of
 * ngia_open overrides the limit). */
#define NGIA_BIN_ZERO 0u
#define NGIA_BIN_WINDOW 1u   /* k_align_win<W>: one thread per pair; width = W */
#define NGIA_BIN_WAVE 2u     /* k_align_wave<K>: one wavefront per pair; width = K */
#define NGIA_BIN_PLAIN 3u
#define NGIA_BIN_SEG 4u      /* k_align_seg<G>: 64/G pairs per wavefront; width = G */
#define NGIA_BIN_SEG2 5u     /* k_align_seg<G, 2>: 64/G pairs per wavefront, two words per
                              * lane; width = G */
#define NGIA_BIN_SEG3 6u     /* k_align_seg<G, 3>: three words per lane; width = G */
#define NGIA_BIN_SEG4 7u     /* k_align_seg<G, 4>: four words per lane; width = G */
typedef struct {
  uint32_t kind;       /* NGIA_BIN_* */
  uint32_t width;      /* W (window), K (wave), G (seg); 0 otherwise */
  uint64_t pairs;      /* distinct pairs in the bin */
  uint64_t ref_tiles;  /* reference 32x32 tiles of those pairs */
  double ms;           /* the bin's kernel time on its stream (NGIA_GRAPH_BATCHED_EAGER), else 0 */
  uint64_t run_tiles;  /* NGIA_GRAPH_BATCHED_TALLY: the reference tiles of the column blocks
                        * the kernel ran (a pair that can no longer pass ends early), else 0 */
} ngia_bin_stat;
int ngia_db_bin_stats(ngia_ctx* ctx, ngia_db* db, ngia_bin_stat* out, uint32_t cap,
                      uint32_t* n_bins);

/* ------------------------------------------- autotuner */
/* No reference counterpart; the reference has one fixed shape, a thread per
 * job in blocks of 64 (buildGraph.hpp:192, 323).  A window bin (band window
 * W = 2..16 words) can be aligned in five shapes: one thread per pair (the
 * most work per instruction, but a lone pair's T x W word-characters run
 * serially on one lane, and its table of W x 4 KiB per wavefront limits the
 * wavefronts per CU), 16 or 32 lanes per pair (several pairs per wavefront,
 * one word per lane), 8 lanes of two words per pair, or a wavefront per pair.
 * Which is fastest depends on W, the bin's size, the identity (pair length)
 * and the device. */
#define NGIA_SHAPE_WINDOW 0u  /* k_align_win<W> */
#define NGIA_SHAPE_SEG16 1u   /* k_align_seg<16>: four pairs per wavefront (W <= 15) */
#define NGIA_SHAPE_SEG32 2u   /* k_align_seg<32>: two pairs per wavefront */
#define NGIA_SHAPE_WAVE 3u    /* k_align_wave<1> */
#define NGIA_SHAPE_SEG8 4u    /* k_align_seg<8, 2>: eight pairs per wavefront (W <= 14) */
#define NGIA_TUNE_CLASSES 8
/* Windows of 17..60 words come in four groups, each aligned with several
 * words per lane or one: 17..21 (eight 8-lane segments of three words, or 32
 * lanes of one), 22..30 (four 16-lane segments of two words, or 32 lanes),
 * 32..45 and 46..60 (four 16-lane segments of three and four words, or a
 * wavefront per pair). */
#define NGIA_LONG_GROUPS 4
#define NGIA_LONG_DEFAULT 0u /* several words per lane from 32,768 pairs (NGIA_SEG2_MIN) */
#define NGIA_LONG_MULTI 1u   /* several words per lane */
#define NGIA_LONG_SINGLE 2u  /* one word per lane, or a wavefront per pair */
typedef struct {
  /* window bin W holding c pairs takes shape[W][k], k the last size class with
   * class_size[k] <= c (class 0 below class_size[0]); class sizes ascend, 0
   * ends the list.  A fresh context's table is all zero: every window bin on
   * one thread per pair. */
  uint32_t class_size[NGIA_TUNE_CLASSES];
  uint32_t shape[17][NGIA_TUNE_CLASSES];
  /* the same for databases whose residues use at most five codes (DNA), whose
   * window kernels index a compact table */
  uint32_t shape_compact[17][NGIA_TUNE_CLASSES];
  /* group g of windows of 17..60 words holding c pairs: NGIA_LONG_* by the
   * same size classes (both alphabets) */
  uint32_t long_shape[NGIA_LONG_GROUPS][NGIA_TUNE_CLASSES];
} ngia_tuning;
int ngia_get_tuning(ngia_ctx* ctx, ngia_tuning* out);
/* NGIA_ERR_ARG for an unknown shape, NGIA_SHAPE_SEG16 at W = 16 or
 * NGIA_SHAPE_SEG8 at W > 14. */
int ngia_set_tuning(ngia_ctx* ctx, const ngia_tuning* t);
typedef struct {
  uint32_t width;    /* band window W */
  uint32_t dna;      /* 1: the DNA pass (compact table), 0: proteins */
  uint32_t length;   /* rep length of the synthetic pairs (targets are up to 64 shorter) */
  uint32_t pairs;    /* pairs in the timed bin */
  double ms[5];      /* device time (HIP events) per NGIA_SHAPE_*, 0 where not applicable;
                      * for W > 16 (a window of a group of 17..60 words): ms[0] several
                      * words per lane, ms[1] one word per lane or a wavefront per pair */
} ngia_tune_point;
/* Time every shape on synthetic pairs of every window width W at `identity`
 * (random reps with targets of nearly the same length, chosen so that the
 * band window is exactly W; proteins, then DNA), at bins of 16..262144 pairs,
 * and install the fastest shape per (W, size class) in the context's tuning;
 * likewise both shapes of each group of 17..60-word windows, on proteins at
 * the windows 19, 26, 38 and 53.  points (may be NULL) receive the
 * measurements. */
int ngia_autotune(ngia_ctx* ctx, uint32_t identity, ngia_tune_point* points, uint32_t cap,
                  uint32_t* n_points);

/* ------------------------------------------- fast mode (no graph, no Leiden) */
/* FastClustering::cluster (fastClustering.hpp:249-352): per band, align the
 * still-unclustered targets, parent[tar] = rep on a pass; cluster[i] is the
 * representative index of i. */
int ngia_db_fast(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                 uint32_t* cluster, uint64_t* total_pairs);

/* ------------------------------------------- serving */
/* No reference counterpart: the reference only clusters offline.  A serving
 * handle keeps a clustered database resident -- packed records, byte codes,
 * the 64 signatures of every sequence and its cluster id -- and assigns
 * streamed query sequences to clusters.  The snapshot grows with the stream:
 * the session's p-th query (counted in arrival order over every assign and
 * replay call since creation or the last ngia_serve_reset) becomes snapshot
 * index n + p, so later queries can join it.
 *
 * Assignment of one query (fixed by the residents and the session's earlier
 * queries alone, so it is the same whether the query is served alone, in a
 * mixed batch or offline):
 *  1. sketch the query as makeDB sketches a record (makedb.hpp:224-247);
 *  2. per band, the candidate is the bucket head of the query's signature
 *     among the residents and the earlier queries: the lowest snapshot index
 *     with that band signature (the row that groupSignatures makes every
 *     member point back to, makedb.hpp:340-347), when there is one;
 *  3. each distinct candidate forms a pair with the longer sequence as rep
 *     (ties: the candidate), kept if (float)len[tar] >= (float)len[rep]*thr
 *     (the isJob test, buildGraph.hpp:262-266), and is scored by the
 *     kernel_dynamic recurrence (buildGraph.hpp:192-253);
 *  4. the query joins the cluster of its best candidate (highest score, then
 *     lowest index): cluster_of[r] for a resident r, the cluster assigned to
 *     an earlier query otherwise; with no passing pair it opens a new cluster
 *     numbered new_base + its query id.
 * Requests wait in arrival order and are admitted into batches as lanes free
 * up; a batch takes at most max_queries queries and max_residues residues,
 * and a lane's device buffers grow toward max_queries at admission only
 * while its share (hbm_fraction / lanes) of the free HBM at that moment
 * allows -- else the batch is cut to what the lane holds.  Lanes run on their
 * own streams, so one batch's host packing and copies overlap another's
 * kernels; batches join the snapshot in arrival order. */
#define NGIA_NO_RESIDENT 0xFFFFFFFFu
typedef struct {
  uint32_t cluster;     /* the best candidate's cluster, or new_base + query id */
  uint32_t resident;    /* snapshot index of the best candidate (>= n: the session's query
                         * resident - n); NGIA_NO_RESIDENT when a new cluster opens */
  uint32_t score;       /* its score 1..100, 0 for a new cluster */
  uint32_t candidates;  /* distinct candidates aligned (after the length test) */
} ngia_assignment;
typedef struct {
  uint32_t lanes;          /* batches in flight at once (streams); 0 -> 2 */
  uint32_t max_queries;    /* queries per batch; 0 -> 8192 */
  uint64_t max_residues;   /* residues per batch; 0 -> 2^24 */
  double hbm_fraction;     /* share of the free HBM the lanes may grow into; 0 -> 0.25 */
  uint32_t grow_queries;   /* queries a session can add to the snapshot; 0 -> 2^20 */
  uint64_t grow_residues;  /* residues a session can add; 0 -> 2^28 */
} ngia_serve_config;
typedef struct {
  double arrival;          /* seconds after the start of the replay (non-decreasing) */
  uint32_t first, count;   /* queries [first, first + count) of the query set */
} ngia_request;
typedef struct {
  uint64_t queries, requests, batches, pairs, new_clusters;
  uint32_t lanes, max_queries;   /* lane shape in effect */
  uint64_t max_residues;
  double wall_s;           /* start of the replay to the last completion */
  double device_s;         /* summed device time of the batches (HIP events) */
  uint64_t grown;          /* queries in the snapshot's grown part after the call */
} ngia_serve_stats;
/* Resident snapshot from db (which must be sketched: ngia_db_sketch) and
 * cluster_of[db n]; identity 1..99.  The snapshot and room for the grown part
 * are allocated here (NGIA_ERR_NOMEM when they exceed the free HBM). */
int ngia_serve_create(ngia_ctx* ctx, ngia_db* db, const uint32_t* cluster_of, uint32_t identity,
                      uint32_t new_base, const ngia_serve_config* cfg, ngia_serve** out);
void ngia_serve_destroy(ngia_serve* sv);
/* Start a new session: the snapshot returns to the residents alone.  A call
 * that would take the grown part past grow_queries or grow_residues fails
 * with NGIA_ERR_NOMEM (the batches admitted before stay in the snapshot). */
int ngia_serve_reset(ngia_serve* sv);
/* Queries are residue strings (FASTA letters): query q is
 * res[res_off[q] .. res_off[q] + res_len[q]), res_len[q] < 65535; its query id
 * is first_qid + q (ids name new clusters: keep them distinct within a
 * session). */
int ngia_serve_assign(ngia_serve* sv, const uint8_t* res, const uint64_t* res_off,
                      const uint32_t* res_len, uint32_t n_queries, uint64_t first_qid,
                      ngia_assignment* out, ngia_serve_stats* stats);
/* Real-time replay of a request stream (continuous batching): request r
 * arrives at reqs[r].arrival; latency_s[r] = its completion (all its queries'
 * results in host memory) minus its arrival.  Query ids are the query
 * indices. */
int ngia_serve_replay(ngia_serve* sv, const uint8_t* res, const uint64_t* res_off,
                      const uint32_t* res_len, uint32_t n_queries, const ngia_request* reqs,
                      uint32_t n_reqs, ngia_assignment* out, double* latency_s,
                      ngia_serve_stats* stats);

/* ------------------------------------------- host side (no device work) */
/* Leiden over the edge list (LeidenOmp::cluster, leidenParallel.hpp:265-322). */
int ngia_leiden(const ngia_edge* edges, uint64_t m, int32_t n, float resolution,
                uint32_t seed, int32_t iters, uint32_t* result);
/* Leiden resolution from identity (main.cu:114-116). */
float ngia_resolution(uint32_t identity);
/* Result file (SaveResult::save, saveResult.hpp:31-83). */
int ngia_save_result(const uint32_t* cluster, uint32_t n, const char* packed_path,
                     const char* result_path);
/* FASTA checks of MakeDB::check (makedb.hpp:59-91): the file exists, its first
 * line has no trailing '\r', a non-empty file ends with '\n'.  NGIA_OK or
 * NGIA_ERR_ARG with the reason in ngia_last_error(NULL).  No device work. */
int ngia_fasta_check(const char* fasta_path);
/* makeDB end to end: FASTA -> packed DB, sketch and bucket on the device
 * (main.cu:54-83 -> makedb.hpp:59-364).  0 ok; NGIA_ERR_ARG when the FASTA
 * fails the reference's checks (missing, CRLF first line, no final newline). */
int ngia_makedb(ngia_ctx* ctx, const char* fasta_path, const char* packed_path);
/* cluster end to end: packed DB -> edges -> Leiden -> result file
 * (main.cu:86-127).  flags as for ngia_db_graph.  The batched form goes
 * through ngia_build_graph_chunked: the whole database when it fits the
 * device, else blocks of it; env NGIA_BLOCK_WORDS=N forces blocks of N
 * packed words. */
int ngia_cluster(ngia_ctx* ctx, const char* packed_path, const char* result_path,
                 uint32_t identity, uint32_t bands, uint32_t flags, ngia_graph_stats* stats);

/* Packed DB reader (Appendix layout of makedb.hpp:1-25): arrays are malloc'd,
 * release with ngia_free.  word_off is relative to the packed area. */
typedef struct {
  uint32_t n;
  uint32_t* name_len;
  uint32_t* lengths;
  uint64_t* pack_off;    /* absolute byte offsets as stored */
  uint64_t* text_off;
  uint32_t* rep_rows;    /* 64 x n */
  uint32_t* packed;      /* the packed area */
  uint64_t n_words;
  uint64_t* word_off;    /* uint32 word offsets inside packed */
} ngia_packed_db;
int ngia_read_db(const char* path, uint32_t bands, ngia_packed_db* out);
void ngia_release_db(ngia_packed_db* db);

/* ------------------------------------------- synthetic families */
/* No reference counterpart: seeded generator of the benchmark inputs. */
typedef struct {
  uint32_t n_families;
  uint32_t members;
  uint32_t dna;           /* 0: 20 amino acids, 1: ACGT */
  uint32_t len_normal;    /* 0: log-normal(len_a = median, len_b = sigma); 1: normal(mean, sd) */
  double len_a, len_b;
  uint32_t len_min, len_max;
  double tail_frac;       /* fraction of families drawn from the Pareto tail */
  double tail_xmin, tail_alpha;
  uint32_t tail_max;
  double sub_lo, sub_hi;  /* per-member substitution rate ~ U[lo, hi] */
  double indel_lo, indel_hi;
  uint64_t seed;
} ngia_gen_params;
/* config 1..5 of the benchmark suite; scale multiplies the family count. */
int ngia_gen_config(int config, double scale, uint64_t seed, ngia_gen_params* out);
/* Write the families as FASTA (shuffled, one sequence line per record). */
int64_t ngia_gen_fasta(const ngia_gen_params* p, const char* path);
/* Same sequences, packed in memory in index order (length-descending stable
 * sort of the FASTA order, exactly as makeDB would order them). */
int ngia_gen_packed(const ngia_gen_params* p, ngia_packed_db* out);
/* Serving queries: a share `mutated` of them are resident sequences (chosen
 * uniformly; residues spelled back from their codes) with substitutions at a
 * rate ~ U[sub_lo, sub_hi] and single-residue indels at rate `indel`; the rest
 * are uniform random proteins with log-normal lengths (median len_median,
 * sigma len_sigma, clipped to [len_min, len_max]).  Arrays are malloc'd
 * (release with ngia_free); kind[q] = 1 for a mutated resident, 0 for novel. */
typedef struct {
  uint32_t n_queries;
  double mutated;
  double sub_lo, sub_hi, indel;
  double len_median, len_sigma;
  uint32_t len_min, len_max;
  uint64_t seed;
} ngia_query_params;
int ngia_gen_queries(const uint32_t* packed, const uint64_t* word_off, uint32_t n,
                     const ngia_query_params* p, uint8_t** res, uint64_t** res_off,
                     uint32_t** res_len, uint8_t** kind);

#ifdef __cplusplus
}
#endif
#endif
