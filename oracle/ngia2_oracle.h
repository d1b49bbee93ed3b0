/*
 * ngia2_oracle.h -- CPU restatement of the nGIA2 clustering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark link it to compare
 * against the HIP product path.  Nothing in the product (the C-ABI library,
 * the CLI) links, calls or falls back to it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the nGIA2 source tree).  The restatement is plain C with
 * OpenMP; float32 expressions are written out operation by operation and the
 * file is compiled with -ffp-contract=off so no contraction happens behind the
 * code's back.
 *
 * How each function is pinned (besides the reference build in oracle/_ref):
 *   orc_code / orc_pack        closed-form table and bit layout, file-size
 *                              invariant, byte-identical packed DB vs the
 *                              reference makeDB build
 *   orc_sketch_record          closed-form hash recomputed from residues in
 *                              the tests (different data path: codes, not
 *                              planes), L=0 / L<8 special cases, reference DB
 *   orc_bucket                 brute-force qsort on (sig,i) + predecessor;
 *                              reference DB rep rows
 *   orc_candidates             closed form (brute loop in tests)
 *   orc_align_record           textbook O(RT) LCS DP + closed-form score on
 *                              passing pairs; identical / empty / filtered
 *                              special cases; fused vs unfused band edges give
 *                              identical scores on the fixtures
 *   orc_align_lcs              the LCS length orc_align_record scores; a
 *                              per-bit restatement of the banded recurrence
 *                              in the tests (band-binding pairs included)
 *   orc_build_graph            sort/unique brute force over per-band scores
 *   orc_leiden                 reference LeidenOmp build (oracle/_ref) on the
 *                              same edge lists; mt19937_64 known-answer value
 *   orc_save_result            reference SaveResult build (oracle/_ref)
 *   orc_fast_cluster           brute restatement in tests (parent chains)
 *   orc_serve_assign           heads equal the groupSignatures chain heads of
 *                              resident queries; scores equal orc_align_record
 *                              of the chosen pair; brute force in tests
 * Parity unpinned: nvcc's choice of FMA contraction for the band edges in the
 * reference kernel (buildGraph.hpp:208-210) cannot be observed without nvcc;
 * both variants are provided (fused = nvcc -fmad=true default).
 */
#ifndef NGIA2_ORACLE_H
#define NGIA2_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BANDS 64

typedef struct { int32_t u; int32_t v; float w; } orc_edge;

/* makedb.hpp:161-171 */
uint32_t orc_code(uint8_t c);
/* makedb.hpp:222-231; rec must hold 1 + 4*ceil(L/32) words */
void orc_pack(const uint8_t* residues, uint32_t L, uint32_t* rec);
/* makedb.hpp:224-247, from a packed record */
void orc_sketch_record(const uint32_t* rec, uint32_t sig[ORC_BANDS]);
/* makedb.hpp:202-248 over n records; sig is seq-major n x 64 */
void orc_sketch_all(const uint32_t* packed, const uint64_t* word_off, uint32_t n,
                    uint32_t* sig, int nthreads);
/* makedb.hpp:289-364; rep is band-major 64 x n */
void orc_bucket(const uint32_t* sig, uint32_t n, uint32_t* rep, int nthreads);
/* buildGraph.hpp:259-300 for one band; returns count, fills (rep<<32|i) */
uint64_t orc_candidates(const uint32_t* rep_row, const uint32_t* len, uint32_t n,
                        uint32_t identity, uint64_t* out);
/* buildGraph.hpp:195-252 on two packed records; returns the job score */
uint32_t orc_align_record(const uint32_t* rep, const uint32_t* tar, float thr, int fused);
/* The LCS length of the banded recurrence (0 when the length filter rejects). */
uint32_t orc_align_lcs(const uint32_t* rep, const uint32_t* tar, float thr, int fused);
/* The same loop, recording after each 32-character block k (k < ceil(T/32))
 * the LCS so far, lcs_out[k], and the tiles visited through block k,
 * tiles_out[k]; returns the block count (0 when the length filter rejects the
 * pair, which is not aligned).  The LCS so far is that of the band-masked
 * match matrix over the target's first 32(k+1) characters. */
uint32_t orc_align_blocks(const uint32_t* rep, const uint32_t* tar, float thr, int fused,
                          uint32_t* lcs_out, uint64_t* tiles_out);
/* band window in words, lsft/32 + ceil(rsft/32) (buildGraph.hpp:208-216) */
uint32_t orc_band_window(uint32_t R, uint32_t T, float thr, int fused);
/* number of 32x32 tiles the reference loop visits for (R,T,thr) */
uint64_t orc_align_tiles(uint32_t R, uint32_t T, float thr, int fused);
/* kernel_dynamic contract: jobs[3i+2] = score of (jobs[3i], jobs[3i+1]) */
void orc_align_jobs(const uint32_t* packed, const uint64_t* word_off, uint32_t* jobs,
                    uint32_t count, float thr, int fused, int nthreads);
/* the same with jobs[3i+2] = orc_align_lcs of the pair */
void orc_align_jobs_lcs(const uint32_t* packed, const uint64_t* word_off, uint32_t* jobs,
                        uint32_t count, float thr, int fused, int nthreads);
/* textbook LCS on residue codes; code 0 never matches (buildGraph.hpp:220-225) */
uint32_t orc_lcs_full(const uint8_t* a, uint32_t R, const uint8_t* b, uint32_t T);
/* buildGraph.hpp:256-363: full band loop, sorted unique edges (malloc'd) */
int orc_build_graph(const uint32_t* packed, const uint64_t* word_off, const uint32_t* len,
                    uint32_t n, const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                    int fused, int nthreads, orc_edge** edges, uint64_t* n_edges,
                    uint64_t* total_pairs);
/* main.cu:114-116 */
float orc_resolution(uint32_t identity);
/* leidenParallel.hpp:265-322 */
int orc_leiden(const orc_edge* edges, uint64_t m, int32_t n, float resolution,
               uint32_t seed, int32_t iters, uint32_t* result);
/* fastClustering.hpp:249-352 (scores via orc_align_record) */
int orc_fast_cluster(const uint32_t* packed, const uint64_t* word_off, const uint32_t* len,
                     uint32_t n, const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                     int fused, int nthreads, uint32_t* cluster, uint64_t* total_pairs);
/* Serving assignment (no reference counterpart; the rule is stated with
 * ngia_serve_create in include/ngia2.h and built from the pieces above).
 * The snapshot grows: query q is snapshot index n + q and sees the residents
 * and the queries before it.  Per query, sketch (makedb.hpp:224-247); per
 * band the lowest snapshot index before q with the query's signature (the
 * bucket head of makedb.hpp:340-347 over residents + earlier queries); the
 * distinct heads as pairs with the longer sequence as rep (ties: the earlier
 * index), the isJob length test (buildGraph.hpp:262-266) and the
 * kernel_dynamic score; best = highest score, then lowest index; a query
 * joins its best's cluster (a resident's cluster_of, an earlier query's
 * assigned cluster) or opens new_base + first_qid + q.  out[4q..4q+3] =
 * cluster, best index (0xFFFFFFFF for a new cluster), score, candidates. */
void orc_serve_assign(const uint32_t* packed, const uint64_t* word_off, uint32_t n,
                      const uint32_t* sig, const uint32_t* cluster_of, const uint32_t* qpacked,
                      const uint64_t* qword_off, uint32_t nq, uint32_t identity,
                      uint32_t new_base, uint64_t first_qid, uint32_t* out, int nthreads);
/* makedb.hpp:94-364 end to end; returns 0 ok, 1 runtime failure, 2 check failure */
int orc_makedb(const char* fasta, const char* packed, int nthreads);
/* saveResult.hpp:31-83 */
int orc_save_result(const uint32_t* cluster, uint32_t n, const char* packed, const char* result);
/* mt19937_64 (C++ [rand.predef]) -- exposed for its known-answer test */
uint64_t orc_mt19937_64_nth(uint64_t seed, uint64_t nth);
void orc_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
