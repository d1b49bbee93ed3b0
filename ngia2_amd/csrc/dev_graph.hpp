// dev_graph.hpp -- launchers of dev_graph.hip (device pointers, one stream).
#pragma once

#include "dev_common.hpp"

namespace ngia {

// Pair bins: 0 = scored 0 without alignment, 2..kWinMax = k_align_win<W>,
// kBinWave0 + log2(K) = k_align_wave<K>, kBinPlain = the plain kernel.
constexpr int kBinZero = 0;
constexpr int kBinWave0 = kWinMax + 1;  // kWinMax+1 .. +6 for K = 1..32
constexpr int kBinPlain = kBinWave0 + 6;
constexpr int kBins = kBinPlain + 1;

// Routing of small window bins, set by the autotuner: window bin W with
// fewer than route_below[W] pairs is aligned by k_align_wave<1> instead (a
// wavefront per pair finishes a few pairs sooner than one thread per pair).
struct RouteTable {
  uint32_t route_below[kWinMax + 1];
};

void fill_u32(uint32_t* p, uint32_t v, uint64_t count, hipStream_t s);
uint64_t scan_scratch_words(uint32_t n);
// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total (n >= 1).
void scan_u32(const uint32_t* in, uint32_t n, uint64_t* out, uint64_t* sums, hipStream_t s);

// Plain form, one band: flags/pos sized n (+1), jobs sized 3n.  parent != null
// selects the fast-mode predicate.
void band_jobs(const uint32_t* rep, const uint32_t* len, uint32_t n, float thr,
               const uint32_t* parent, uint32_t* flags, uint64_t* pos, uint64_t* sums,
               uint32_t* jobs, hipStream_t s);
void jobs_to_edges(const uint32_t* jobs, uint32_t count, uint64_t* ekeys, uint32_t* escores,
                   unsigned long long* ecount, hipStream_t s);
void fast_apply(const uint32_t* jobs, uint32_t count, uint32_t* parent, hipStream_t s);

// Batched form.
// Candidates: the count pass writes cnt[i] (distinct pairs of target i),
// keep[i] (the bands whose representatives are kept) and adds the count with
// duplicates to *total; after cnt is scanned into off, the emit pass writes
// the pairs grouped by target, in target order.
void pair_count(const uint32_t* rep, const uint32_t* len, uint32_t n, uint32_t bands, float thr,
                uint32_t* cnt, uint64_t* keep, unsigned long long* total, hipStream_t s);
void pair_emit(const uint32_t* rep, uint32_t n, uint32_t bands, const uint64_t* keep,
               const uint64_t* off, uint2* pairs, hipStream_t s);
void bin_pairs(const uint2* pairs, const uint64_t* npairs, const uint32_t* len, float thr,
               uint32_t* bin_of, uint32_t* bin_cnt, uint32_t* bin_off, uint32_t* cursor,
               uint32_t* list, unsigned long long* bin_tiles, uint32_t* score, uint32_t kmax,
               uint32_t* remap, const RouteTable& route, uint32_t grid, hipStream_t s);
void plain_bin_jobs(const uint2* pairs, const uint32_t* list, const uint32_t* bin_off,
                    const uint32_t* bin_cnt, uint32_t* jobs, hipStream_t s);
void plain_bin_scores(const uint32_t* list, const uint32_t* bin_off, const uint32_t* bin_cnt,
                      const uint32_t* jobs, uint32_t* score, hipStream_t s);
void pairs_to_edges(const uint2* pairs, const uint64_t* npairs, const uint32_t* score,
                    uint64_t* ekeys, uint32_t* escores, unsigned long long* ecount, uint32_t grid,
                    hipStream_t s);

// Sort + unique of m raw edges in place (keys/scores); *d_unique = count.
size_t finalize_scratch_bytes(uint64_t m);
void finalize_edges(uint64_t* keys, uint32_t* scores, uint64_t m, uint64_t* keys_tmp,
                    uint32_t* scores_tmp, void* temp, size_t temp_bytes,
                    unsigned long long* d_unique, hipStream_t s);

// Sketch over byte codes (dev_sketch.hip).
void launch_sketch(const uint8_t* codes, const uint64_t* code_off, const uint32_t* lengths,
                   uint32_t n, uint32_t* sig, hipStream_t stream);

}  // namespace ngia
