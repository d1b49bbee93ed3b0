// dev_graph.hpp -- launchers of dev_graph.hip (device pointers, one stream).
#pragma once

#include "dev_common.hpp"

namespace ngia {

// Pair bins: 0 = scored 0 without alignment, 2..kWinMax = k_align_win<W>,
// kBinWave0 + log2(K) = k_align_wave<K>, kBinPlain = the plain kernel,
// kBinSeg16 / kBinSeg32 = k_align_seg<16, 1> / <32, 1> (several pairs per
// wavefront, one word per lane), kBinSeg16x2..x4 = k_align_seg<16, 2..4>
// (four pairs per wavefront, 2-4 words per lane), kBinSeg8x3 / x2 =
// k_align_seg<8, 3 / 2> (eight pairs per wavefront; x2 only for window bins
// the tuning routes there): band windows of
// kWinMax+1 .. 21 words go to kBinSeg8x3, of 22..30 to kBinSeg16x2, of 31
// words to kBinSeg32, of 32..45 to kBinSeg16x3 and of 46..60 to
// kBinSeg16x4.
constexpr int kBinZero = 0;
constexpr int kBinWave0 = kWinMax + 1;  // kWinMax+1 .. +6 for K = 1..32
constexpr int kBinPlain = kBinWave0 + 6;
constexpr int kBinSeg16 = kBinPlain + 1;
constexpr int kBinSeg32 = kBinSeg16 + 1;
constexpr int kBinSeg16x2 = kBinSeg32 + 1;
constexpr int kBinSeg16x3 = kBinSeg16x2 + 1;
constexpr int kBinSeg16x4 = kBinSeg16x3 + 1;
constexpr int kBinSeg8x3 = kBinSeg16x4 + 1;
constexpr int kBinSeg8x2 = kBinSeg8x3 + 1;
constexpr int kBins = kBinSeg8x2 + 1;
static_assert(kBins <= 32, "bin masks are 32-bit");
// A kBinSeg16x2 or kBinSeg8x3 bin of fewer pairs goes to kBinSeg32, a
// kBinSeg16x3 or x4 bin to k_align_wave<1> (k_bin_offsets; the seg2_min
// argument of bin_pairs, NGIA_SEG2_MIN): on the heavy-tailed set at -i 90 its
// 14-28k long pairs made the step 0.3 ms longer with two words per lane; the
// log-normal set's 0.37M and 2.3M at -i 70 and 50 took 22% and 20% less
// time.
constexpr uint32_t kSeg2MinPairs = 32768;
// The segmented bins and their (lanes per pair, words per lane).
constexpr bool is_seg_bin(int b) { return b >= kBinSeg16 && b <= kBinSeg8x2; }
constexpr int seg_lanes(int b) {
  return b == kBinSeg32 ? 32 : (b == kBinSeg8x3 || b == kBinSeg8x2) ? 8 : 16;
}
constexpr int seg_words(int b) {
  return b == kBinSeg8x3 ? 3 : b == kBinSeg8x2 ? 2 : b >= kBinSeg16x2 ? b - kBinSeg16x2 + 2 : 1;
}

// Shapes a window bin (W = 2..kWinMax) can be aligned with, chosen per bin
// size by the autotuner: one thread per pair (k_align_win<W>), 16 or 32 lanes
// per pair (k_align_seg), a wavefront per pair (k_align_wave<1>), or 8 lanes
// of two words per pair (k_align_seg<8, 2>).
constexpr uint32_t kShapeWindow = NGIA_SHAPE_WINDOW;
constexpr uint32_t kShapeSeg16 = NGIA_SHAPE_SEG16;
constexpr uint32_t kShapeSeg32 = NGIA_SHAPE_SEG32;
constexpr uint32_t kShapeWave = NGIA_SHAPE_WAVE;
constexpr uint32_t kShapeSeg8 = NGIA_SHAPE_SEG8;
constexpr uint32_t kShapes = kShapeSeg8 + 1;
constexpr int kTuneClasses = NGIA_TUNE_CLASSES;
// Window bin W holding c pairs is aligned with shape[W][k], k the last size
// class with class_size[k] <= c (class 0 below class_size[0]).  A zeroed table
// keeps every window bin on k_align_win<W>.
// Groups of 17..60-word windows: lng[g][k] = NGIA_LONG_* for group g
// (17..21, 22..30, 32..45, 46..60 words) in size class k.
constexpr int kLongGroups = NGIA_LONG_GROUPS;
struct RouteTable {
  uint32_t class_size[kTuneClasses];
  uint8_t shape[kWinMax + 1][kTuneClasses];
  uint8_t lng[kLongGroups][kTuneClasses];
};
// Size class of a bin of c pairs.
__host__ __device__ inline uint32_t route_class(const RouteTable& rt, uint32_t c) {
  uint32_t k = 0;
  for (uint32_t j = 1; j < (uint32_t)kTuneClasses; j++)
    if (rt.class_size[j] && rt.class_size[j] <= c) k = j;
  return k;
}

// The bin that aligns window bin W's c pairs under the routing table.
__host__ __device__ inline uint32_t route_window_bin(const RouteTable& rt, uint32_t W, uint32_t c) {
  const uint32_t k = route_class(rt, c);
  switch (rt.shape[W][k]) {
    case kShapeSeg16: return W <= 15 ? (uint32_t)kBinSeg16 : (uint32_t)kBinSeg32;
    case kShapeSeg32: return (uint32_t)kBinSeg32;
    case kShapeWave: return (uint32_t)kBinWave0;
    case kShapeSeg8:
      return W <= 14 ? (uint32_t)kBinSeg8x2 : W <= 15 ? (uint32_t)kBinSeg16 : (uint32_t)kBinSeg32;
    default: return W;
  }
}

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
// the pairs grouped by target, in target order.  Targets below t0 get none.
void pair_count(const uint32_t* rep, const uint32_t* len, uint32_t n, uint32_t t0, uint32_t bands,
                float thr, uint32_t* cnt, uint64_t* keep, unsigned long long* total, hipStream_t s);
void pair_emit(const uint32_t* rep, uint32_t n, uint32_t bands, const uint64_t* keep,
               const uint64_t* off, uint2* pairs, hipStream_t s);
void bin_pairs(const uint2* pairs, const uint64_t* npairs, const uint32_t* len, float thr,
               uint32_t* bin_of, uint32_t* bin_cnt, uint32_t* bin_off, uint32_t* cursor,
               uint32_t* list, unsigned long long* bin_tiles, uint32_t* score, uint32_t kmax,
               uint32_t* remap, const RouteTable& route, uint32_t seg2_min, uint32_t grid,
               hipStream_t s);
void plain_bin_jobs(const uint2* pairs, const uint32_t* list, const uint32_t* bin_off,
                    const uint32_t* bin_cnt, uint32_t* jobs, hipStream_t s);
void plain_bin_scores(const uint32_t* list, const uint32_t* bin_off, const uint32_t* bin_cnt,
                      const uint32_t* jobs, uint32_t* score, hipStream_t s);
void pairs_to_edges(const uint2* pairs, const uint64_t* npairs, const uint32_t* score,
                    uint64_t* ekeys, uint32_t* escores, unsigned long long* ecount, uint32_t grid,
                    hipStream_t s);

// Sort + unique of m raw edges in place (keys/scores); *d_unique = count.
// The keys are u << 32 | v with u, v < n: the sort skips the bits above
// 32 + bits(n - 1).
size_t finalize_scratch_bytes(uint64_t m);
void finalize_edges(uint64_t* keys, uint32_t* scores, uint64_t m, uint32_t n, uint64_t* keys_tmp,
                    uint32_t* scores_tmp, void* temp, size_t temp_bytes,
                    unsigned long long* d_unique, hipStream_t s);

// Sketch over byte codes (dev_sketch.hip).
void launch_sketch(const uint8_t* codes, const uint64_t* code_off, const uint32_t* lengths,
                   uint32_t n, uint32_t* sig, hipStream_t stream);

}  // namespace ngia
