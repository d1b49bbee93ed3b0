// dev_common.hpp -- shared declarations of the HIP side of the C-ABI library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "ngia2.h"

namespace ngia {

// Error carried from a HIP call site to the C-ABI boundary.
struct HipError {
  int code;
  std::string msg;
};

#define NGIA_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::ngia::HipError{NGIA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

inline unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  return (unsigned)(g == 0 ? 1 : g);
}

// Band edges of the reference kernel (buildGraph.hpp:208-210), computed with
// an explicit fused multiply-add: nvcc contracts `T - T*thr` into
// fma(-T, thr, T) under its default -fmad=true.  Float -> uint32 conversion
// saturates negative values to 0 on both GPU ISAs.
struct Band {
  uint32_t lsft, rsft;
};
__host__ __device__ inline uint32_t f2u_sat(float x) {
  return x > 0.0f ? (x >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)x) : 0u;
}
__host__ __device__ inline Band band_edges(uint32_t R, uint32_t T, float thr) {
  const float fT = (float)T, fR = (float)R;
  Band b;
  b.lsft = (f2u_sat(ceilf(fmaf(-fT, thr, fT))) + 31u) & ~31u;
  b.rsft = f2u_sat(ceilf(fmaf(-fT, thr, fR))) + 32u;
  return b;
}
// The reference kernel's length filter (buildGraph.hpp:201): the job is
// skipped (score stays 0) when T < R*thr.
__host__ __device__ inline bool length_rejects(uint32_t R, uint32_t T, float thr) {
  return (float)T < (float)R * thr;
}
// Reference tiles (one rep word by one 32-character block, 1,024 cells) of
// the first `blocks` column blocks of a pair: the word range the reference
// updates for block cb (buildGraph.hpp:213-216), [max(32cb - lsft, 0),
// min(32cb + rsft, R)), in words.
__host__ __device__ inline uint64_t ref_tiles(uint32_t R, uint32_t T, float thr,
                                              uint32_t blocks) {
  const Band bd = band_edges(R, T, thr);
  uint64_t t = 0;
  for (uint32_t cb = 0; cb < blocks; cb++) {
    int32_t js = (int32_t)(32 * cb) - (int32_t)bd.lsft;
    if (js < 0) js = 0;
    int32_t je = (int32_t)(32 * cb) + (int32_t)bd.rsft;
    if (je > (int32_t)R) je = (int32_t)R;
    if (je > js) t += ((uint32_t)(je - js) + 31) / 32;
  }
  return t;
}
// Pass test and integer percent score (buildGraph.hpp:248-252).
__host__ __device__ inline uint32_t lcs_score(uint32_t lcs, uint32_t R, uint32_t T, float thr) {
  const uint32_t d = R + T - lcs;
  return ((float)lcs > (float)d * thr) ? lcs * 100u / d : 0u;
}

// ---------------------------------------------------------------- launchers
// All take device pointers and run on `stream`.

// Band-major bucketing of n x 64 signatures into 64 x n representative rows.
// scratch holds bucket_scratch_bytes(n, force_wide) bytes; force_wide selects
// the 38-bit-key full sort that large n needs anyway.  Below that, n up to
// 2.8M takes the partitioned form, which synchronises the stream once (to
// learn whether a partition overflowed and the full sort must run instead).
// few_values skips the partitioned form: signatures that take few distinct
// values per band (8-mers of an alphabet of at most five letters, DNA among
// them, have at most 5^8 values and their minima crowd near the smallest
// hashes) make long runs of equal signatures that overflow any partition.
size_t bucket_scratch_bytes(uint32_t n, bool force_wide = false);
// pinned_flag (optional): 4 bytes of pinned host memory that receive the
// overflow flag of the partitioned form.
void launch_bucket(const uint32_t* sig, uint32_t n, uint32_t* rep, void* scratch,
                   hipStream_t stream, bool force_wide = false, bool few_values = false,
                   uint32_t* pinned_flag = nullptr);
// The same rows for n <= kBucketSmallMax records: one block per band, no
// scratch and no host synchronisation.
constexpr uint32_t kBucketSmallMax = 16384;
void launch_bucket_small(const uint32_t* sig, uint32_t n, uint32_t* rep, hipStream_t stream);
// Signatures of long records in chunks of `len` window positions:
// chunks[c] = (record, first position); sig rows set to all ones beforehand.
void launch_sketch_chunks(const uint8_t* codes, const uint64_t* code_off, const uint32_t* lengths,
                          const uint2* chunks, uint32_t nchunks, uint32_t len, uint32_t* sig,
                          hipStream_t stream);

// Byte-per-residue codes (code of residue j at byte code_off[i] + j, padding
// to a multiple of 32 is code 0) built from the packed planes.
// Bit c of *present |= 1 when code c occurs among the first `bytes` codes.
void launch_code_set(const uint8_t* codes, uint64_t bytes, uint32_t* present, hipStream_t stream);
// slots[k] = slot of codes[k] under lut (slot of code c in bits 4c..4c+3).
void launch_code_slots(const uint8_t* codes, uint64_t bytes, uint64_t lut, uint8_t* slots,
                       hipStream_t stream);
void launch_codes(const uint32_t* packed, const uint64_t* word_off, const uint64_t* code_off,
                  uint32_t n, uint8_t* codes, hipStream_t stream);

// Geometry of one aligned pair.
struct PairGeom {
  uint32_t R, T;   // lengths
  uint32_t a, c;   // window below / above: band words = [cb-a, cb+c)
};

// Alignment of a pair list: score[p] for pairs rep[p], tar[p], p < count.
struct AlignArgs {
  const uint32_t* packed;
  const uint64_t* word_off;
  const uint8_t* codes;
  const uint64_t* code_off;
  // Record lengths, as in each record's header word.  The binned kernels
  // read a target's length here: its packed record is otherwise untouched,
  // and its header would cost a line of memory traffic per pair.
  const uint32_t* lengths = nullptr;
  float thr;
  uint32_t raw_lcs = 0;  // nonzero: store the LCS length instead of the score
  // Compact alphabet (a database whose residues use at most kSlotCodes
  // nonzero codes, DNA among them): slots[] holds each residue's slot
  // (0 for code 0, 1..kSlotCodes for the codes present), laid out as codes[];
  // slot_code holds the code of slot s in bits 4(s-1)..4s-1.  nullptr: the
  // full 16-code alphabet.
  const uint8_t* slots = nullptr;
  uint32_t slot_code = 0;
  // Non-null: the kernel adds the reference tiles of the column blocks it
  // ran (an early exit skips the rest of a pair's) here, one atomic per lane
  // at its end.  Measurement only; nullptr in the product's rounds.
  unsigned long long* tally = nullptr;
};
// The tally of a lane's alignments, added once when the lane is done.
__device__ inline void tally_add(const AlignArgs& A, uint64_t ran) {
  if (A.tally && ran) atomicAdd(A.tally, (unsigned long long)ran);
}
constexpr int kSlotCodes = 5;
// What an alignment kernel stores for a pair it aligned: the score, or the
// LCS length itself when raw_lcs is set (a check of the recurrence that does
// not go through the score's integer rounding and threshold).
__host__ __device__ inline uint32_t pair_value(const AlignArgs& A, uint32_t lcs, uint32_t R,
                                               uint32_t T) {
  return A.raw_lcs ? lcs : lcs_score(lcs, R, T, A.thr);
}
// Plain per-band form: one thread per job triplet, any length.  The rep of
// every job has at most nw_max words (ceil(R/32)).  With count_ptr set the
// job count is read from device memory (count is then an upper bound).  The
// grid has at most max_blocks blocks of 64 threads, which stride over the
// jobs; the row scratch holds align_plain_scratch_words(nw_max, max_blocks)
// words.
constexpr uint32_t kPlainBlocks = 4096;    // 262,144 threads: the plain form
constexpr uint32_t kPlainBinBlocks = 256;  // the batched form's widest-window bin
void launch_align_plain(const AlignArgs& a, uint32_t* jobs, uint32_t count, uint32_t nw_max,
                        uint32_t* scratch, uint64_t scratch_words, hipStream_t stream,
                        const uint32_t* count_ptr = nullptr, uint32_t max_blocks = kPlainBlocks);
uint64_t align_plain_scratch_words(uint32_t nw_max, uint32_t max_blocks = kPlainBlocks);
// The plain kernel over a scratch of fixed size (scratch_words words) and
// the packed area alone (a.codes unused: target codes come from the planes):
// jobs whose rep has at most nw_cap words run on threads with nw_cap-word
// rows, longer ones (up to kPlainMaxWords) on 1,024 threads with full rows.
constexpr uint32_t kPlainMaxWords = 2048;  // reps of at most 65,534 residues
void launch_align_plain_tiers(const AlignArgs& a, uint32_t* jobs, uint32_t count,
                              uint32_t* scratch, uint64_t scratch_words, uint32_t nw_cap,
                              hipStream_t stream);
// Work queue of a bin: items [0, count) in kQueueSlices contiguous slices,
// one claim counter per slice, each on a line of its own
// (queue[s * kQueueStride]); a queue is kQueueWords zeroed words.  A
// wavefront claims from the slice of its XCD (blockIdx.x mod 8 under the
// round-robin dispatch of workgroups to XCDs) and, once that is exhausted,
// from the others: one counter for every wavefront of a bin serialised the
// claims of bins with millions of short pairs.
constexpr uint32_t kQueueSlices = 8;
constexpr uint32_t kQueueStride = 64;  // words: 256 bytes
constexpr uint32_t kQueueWords = kQueueSlices * kQueueStride;
// Length-binned forms over a pair list (pairs[list[*off_ptr + i]], i <
// *count_ptr); work is a zeroed work queue (kQueueWords words); blocks is the
// persistent grid.
size_t win_lds_bytes(int W, bool compact = false);
size_t wave_lds_bytes(int K);
void launch_align_win(int W, const AlignArgs& a, const uint2* pairs, const uint32_t* list,
                      const uint32_t* off_ptr, const uint32_t* count_ptr, uint32_t* work, uint32_t* score,
                      uint32_t blocks, hipStream_t stream);
void launch_align_wave(int K, const AlignArgs& a, const uint2* pairs, const uint32_t* list,
                       const uint32_t* off_ptr, const uint32_t* count_ptr, uint32_t* work, uint32_t* score,
                       uint32_t blocks, hipStream_t stream);
// 64 / G pairs per wavefront, one per G-lane segment, K words per lane
// ((G, K) = (16, 1..4), (32, 1) or (8, 2..3)), for band windows of at most
// K (G - 1) words; any alphabet (the 16-code table).
size_t seg_lds_bytes();
// Persistent blocks per CU for k_align_seg<G, K> (from the kernel's occupancy).
uint32_t seg_blocks_per_cu(int G, int K = 1);
void launch_align_seg(int G, int K, const AlignArgs& a, const uint2* pairs, const uint32_t* list,
                      const uint32_t* off_ptr, const uint32_t* count_ptr, uint32_t* work,
                      uint32_t* score, uint32_t blocks, hipStream_t stream);
// Raise the dynamic-LDS limit of the kernels that need more than 64 KiB;
// returns the largest K of k_align_wave<K> the device accepts.
int align_kernels_setup(size_t lds_per_block);
constexpr int kWinMax = 16;    // window widths 2..16 use k_align_win
constexpr int kWaveKMax = 32;  // k_align_wave handles W + K - 1 <= 64*K, K <= 32

}  // namespace ngia
