// dev_align.hip -- banded bit-parallel LCS of candidate pairs on CDNA4.
//
// Semantics: kernel_dynamic (buildGraph.hpp:192-253).  For column block cb
// (target characters 32cb..32cb+31) the reference updates rep words
// [ws, we) = [max(cb - lsft/32, 0), ceil(min(32cb + rsft, R)/32)) in order,
// with carry-in 0 into word ws and the carry out of word we-1 dropped; words
// outside the range are frozen.  Because lsft is a multiple of 32, that range
// is the window [cb - a, cb + c) with a = lsft/32, c = ceil(rsft/32), clipped
// to [0, ceil(R/32)).  The window slides up one word per block, so a word
// enters at the top as all ones (never touched before) and leaves at the
// bottom frozen for good.  Words of the window outside [0, ceil(R/32)) are
// "virtual": all-ones rows whose match mask is 0 -- such a row is a fixed
// point of the update for either carry-in, so it absorbs the dropped carry at
// the top and passes carry 0 at the bottom.  Target characters beyond T have
// code 0 (match mask 0) and with carry-in 0 leave every word unchanged.
// The kernels below all compute that recurrence exactly; they differ in how
// pairs and words map onto lanes.
//
//   k_align_plain   one thread per job, one launch per band (the plain form)
//   k_align_win<W>  one thread per pair, W window words in registers, the
//                   per-lane match masks (Peq) in LDS, lane-interleaved
//   k_align_wave<K> one wavefront per pair, K window words per lane, carry
//                   propagated across lanes by a ballot carry-lookahead
#include <algorithm>

#include "dev_common.hpp"

namespace ngia {

// ------------------------------------------------------------- byte codes
// codes[code_off[i] + j] = code of residue j of sequence i, 0 past the end.
__global__ void k_codes(const uint32_t* __restrict__ packed, const uint64_t* __restrict__ word_off,
                        const uint64_t* __restrict__ code_off, uint32_t n, uint8_t* __restrict__ codes) {
  const uint32_t i = blockIdx.x;
  if (i >= n) return;
  const uint32_t* rec = packed + word_off[i];
  const uint32_t nb = (rec[0] + 31) / 32;
  uint32_t* out = (uint32_t*)(codes + code_off[i]);
  for (uint32_t t = threadIdx.x; t < nb * 8; t += blockDim.x) {
    const uint32_t b = t >> 3, q = t & 7;  // block b, chars 4q..4q+3
    const uint32_t* pl = rec + 1 + 4 * b;
    uint32_t w = 0;
    #pragma unroll
    for (int m = 0; m < 4; m++) {
      const uint32_t k = 4 * q + m;
      uint32_t c = ((pl[0] >> k) & 1u) | (((pl[1] >> k) & 1u) << 1) | (((pl[2] >> k) & 1u) << 2) |
                   (((pl[3] >> k) & 1u) << 3);
      w |= c << (8 * m);
    }
    out[t] = w;
  }
}

void launch_codes(const uint32_t* packed, const uint64_t* word_off, const uint64_t* code_off,
                  uint32_t n, uint8_t* codes, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_codes, dim3(n), dim3(64), 0, stream, packed, word_off, code_off, n, codes);
  NGIA_HIP(hipGetLastError());
}

// Which codes occur: one bit per code, OR-reduced per block, then one atomic.
__global__ __launch_bounds__(256) void k_code_set(const uint32_t* __restrict__ codes, uint64_t words,
                                                  uint32_t* __restrict__ present) {
  uint32_t bits = 0;
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < words;
       k += (uint64_t)gridDim.x * 256) {
    const uint32_t w = codes[k];
    bits |= (1u << (w & 0xFu)) | (1u << ((w >> 8) & 0xFu)) | (1u << ((w >> 16) & 0xFu)) |
            (1u << ((w >> 24) & 0xFu));
  }
  __shared__ uint32_t acc;
  if (threadIdx.x == 0) acc = 0;
  __syncthreads();
  if (bits) atomicOr(&acc, bits);
  __syncthreads();
  if (threadIdx.x == 0 && acc) atomicOr(present, acc);
}

void launch_code_set(const uint8_t* codes, uint64_t bytes, uint32_t* present, hipStream_t stream) {
  const uint64_t words = bytes / 4;  // the code area is a whole number of 32-byte blocks
  if (words == 0) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((words + 255) / 256, 4096);
  hipLaunchKernelGGL(k_code_set, dim3(grid), dim3(256), 0, stream, (const uint32_t*)codes, words,
                     present);
  NGIA_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_code_slots(const uint32_t* __restrict__ codes, uint64_t words,
                                                    uint64_t lut, uint32_t* __restrict__ slots) {
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < words;
       k += (uint64_t)gridDim.x * 256) {
    const uint32_t w = codes[k];
    uint32_t o = 0;
    #pragma unroll
    for (int b = 0; b < 4; b++)
      o |= (uint32_t)((lut >> (4 * ((w >> (8 * b)) & 0xFu))) & 0xFu) << (8 * b);
    slots[k] = o;
  }
}

void launch_code_slots(const uint8_t* codes, uint64_t bytes, uint64_t lut, uint8_t* slots,
                       hipStream_t stream) {
  const uint64_t words = bytes / 4;
  if (words == 0) return;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((words + 255) / 256, 16384);
  hipLaunchKernelGGL(k_code_slots, dim3(grid), dim3(256), 0, stream, (const uint32_t*)codes, words,
                     lut, (uint32_t*)slots);
  NGIA_HIP(hipGetLastError());
}

// The 15 match masks of one rep word from its 4 bit planes: mask[k] has bit
// r set iff residue r of the word has code k (buildGraph.hpp:220-225).
__device__ inline void peq_masks(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3,
                                 uint32_t m[16]) {
  const uint32_t a0 = ~p0 & ~p1, a1 = p0 & ~p1, a2 = ~p0 & p1, a3 = p0 & p1;
  const uint32_t b0 = ~p2 & ~p3, b1 = p2 & ~p3, b2 = ~p2 & p3, b3 = p2 & p3;
  m[0] = 0;
  m[1] = a1 & b0;  m[2] = a2 & b0;  m[3] = a3 & b0;
  m[4] = a0 & b1;  m[5] = a1 & b1;  m[6] = a2 & b1;  m[7] = a3 & b1;
  m[8] = a0 & b2;  m[9] = a1 & b2;  m[10] = a2 & b2; m[11] = a3 & b2;
  m[12] = a0 & b3; m[13] = a1 & b3; m[14] = a2 & b3; m[15] = a3 & b3;
}

// One word-character step of the LCS row update (buildGraph.hpp:231-239):
// row' = (row + (row & M) + cin) | (row & ~M), carry out returned in *cout.
__device__ inline uint32_t lcs_step(uint32_t row, uint32_t M, uint32_t cin, uint32_t* cout) {
  const uint32_t t0 = row & M;
  const uint32_t s = __builtin_addc(row, t0, cin, cout);
  uint32_t r = s | (row & ~M);  // one v_bitop3
  // Keep the new row opaque: otherwise the next step's row & M is rebuilt
  // from s and row & ~M, which costs an extra bitop3 per word-character.
  asm("" : "+v"(r));
  return r;
}

// ------------------------------------------------------------- plain form
// One thread per job triplet, the whole band of any width; the row vector
// lives in a lane-interleaved global scratch (row w of the thread at
// rows[w * stride]), the 16 match masks of the current rep word in a
// lane-interleaved LDS table.  The target's codes come from the byte codes
// (kPlanes = false) or straight from its four bit planes (kPlanes = true, for
// callers that hand over the packed area alone).
template <bool kPlanes>
__device__ inline void plain_job(const AlignArgs& A, uint32_t* __restrict__ jobs, uint32_t job,
                                 uint32_t R, uint32_t* __restrict__ rows, uint32_t stride,
                                 uint32_t* peq, uint32_t lane) {
  const uint32_t rep = jobs[3 * job], tar = jobs[3 * job + 1];
  const uint32_t* rr = A.packed + A.word_off[rep];
  const uint32_t* tr = A.packed + A.word_off[tar];
  const uint32_t T = tr[0];
  if (length_rejects(R, T, A.thr)) return;  // score keeps its initial 0
  const uint32_t nw = (R + 31) / 32;
  for (uint32_t w = 0; w < nw; w++) rows[(uint64_t)w * stride] = 0xFFFFFFFFu;
  const Band bd = band_edges(R, T, A.thr);
  const uint8_t* tc = kPlanes ? nullptr : A.codes + A.code_off[tar];
  for (uint32_t i = 0; i < T; i += 32) {
    int32_t js = (int32_t)i - (int32_t)bd.lsft;
    if (js < 0) js = 0;
    int32_t je = (int32_t)i + (int32_t)bd.rsft;
    if (je > (int32_t)R) je = (int32_t)R;
    uint32_t cw[8];  // codes of characters 4q..4q+3 of the block, one per byte
    if (kPlanes) {
      const uint32_t* pl = tr + 1 + (i / 32) * 4;
      const uint32_t p0 = pl[0], p1 = pl[1], p2 = pl[2], p3 = pl[3];
      #pragma unroll
      for (int q = 0; q < 8; q++) {
        uint32_t w = 0;
        #pragma unroll
        for (int m = 0; m < 4; m++) {
          const uint32_t k = 4 * q + m;
          w |= (((p0 >> k) & 1u) | (((p1 >> k) & 1u) << 1) | (((p2 >> k) & 1u) << 2) |
                (((p3 >> k) & 1u) << 3)) << (8 * m);
        }
        cw[q] = w;
      }
    } else {
      const uint4 c0 = *(const uint4*)(tc + i), c1 = *(const uint4*)(tc + i + 16);
      cw[0] = c0.x; cw[1] = c0.y; cw[2] = c0.z; cw[3] = c0.w;
      cw[4] = c1.x; cw[5] = c1.y; cw[6] = c1.z; cw[7] = c1.w;
    }
    uint32_t carrys = 0;  // carry out of the previous word, bit k for char k
    for (uint32_t j = (uint32_t)js; j < (uint32_t)je; j += 32) {
      const uint32_t* pl = rr + 1 + (j / 32) * 4;
      uint32_t m[16];
      peq_masks(pl[0], pl[1], pl[2], pl[3], m);
      #pragma unroll
      for (int k = 0; k < 16; k++) peq[k * 64 + lane] = m[k];
      uint32_t row = rows[(uint64_t)(j / 32) * stride];
      uint32_t cout_bits = 0;
      #pragma unroll
      for (int k = 0; k < 32; k++) {
        const uint32_t code = (cw[k >> 2] >> (8 * (k & 3))) & 0xFu;
        const uint32_t M = peq[code * 64 + lane];
        uint32_t co;
        row = lcs_step(row, M, (carrys >> k) & 1u, &co);
        cout_bits |= co << k;
      }
      carrys = cout_bits;
      rows[(uint64_t)(j / 32) * stride] = row;
    }
  }
  uint32_t lcs = 0;
  for (uint32_t w = 0; w < nw; w++) lcs += 32u - __popc(rows[(uint64_t)w * stride]);
  jobs[3 * job + 2] = pair_value(A, lcs, R, T);
}

// Grid-stride over the jobs; `stride` threads in all.
__global__ __launch_bounds__(64) void k_align_plain(AlignArgs A, uint32_t* __restrict__ jobs,
                                                    uint32_t count, const uint32_t* count_ptr,
                                                    uint32_t* __restrict__ scratch,
                                                    uint32_t stride) {
  __shared__ uint32_t peq[16 * 64];
  const uint32_t lane = threadIdx.x;
  const uint32_t tid = blockIdx.x * 64 + lane;
  if (count_ptr) count = *count_ptr;
  uint64_t ran = 0;  // tally: reference tiles of the jobs run
  for (uint32_t job = tid; job < count; job += stride) {
    const uint32_t R = A.packed[A.word_off[jobs[3 * job]]];
    plain_job<false>(A, jobs, job, R, scratch + tid, stride, peq, lane);
    if (A.tally) {
      const uint32_t T = A.packed[A.word_off[jobs[3 * job + 1]]];
      if (!length_rejects(R, T, A.thr)) ran += ref_tiles(R, T, A.thr, (T + 31) / 32);
    }
  }
  tally_add(A, ran);
}

// The same over a scratch of fixed size, in two tiers of threads.  Threads
// [0, stride_a) take the jobs whose rep has at most nw_cap words, with rows
// of nw_cap words; threads [stride_a, stride_a + stride_b) take the longer
// ones, with rows of kPlainMaxWords words (a rep of at most 65,534 residues,
// the reference's lines[2048]).  Every job is taken by exactly one tier and
// every row index stays inside the tier's rows, whatever the records hold: a
// rep longer than kPlainMaxWords words is left unaligned (its score keeps
// the caller's value, as a length-rejected job's does).
__global__ __launch_bounds__(64) void k_align_plain_tiers(AlignArgs A, uint32_t* __restrict__ jobs,
                                                          uint32_t count,
                                                          uint32_t* __restrict__ scratch,
                                                          uint32_t nw_cap, uint32_t stride_a,
                                                          uint32_t stride_b) {
  __shared__ uint32_t peq[16 * 64];
  const uint32_t lane = threadIdx.x;
  const uint32_t tid = blockIdx.x * 64 + lane;
  const bool tier_a = tid < stride_a;
  const uint32_t t = tier_a ? tid : tid - stride_a;
  const uint32_t stride = tier_a ? stride_a : stride_b;
  uint32_t* rows = tier_a ? scratch + t : scratch + (uint64_t)nw_cap * stride_a + t;
  for (uint32_t job = t; job < count; job += stride) {
    const uint32_t R = A.packed[A.word_off[jobs[3 * job]]];
    const uint32_t nw = (R + 31) / 32;
    if (tier_a ? nw > nw_cap : (nw <= nw_cap || nw > kPlainMaxWords)) continue;
    plain_job<true>(A, jobs, job, R, rows, stride, peq, lane);
  }
}

uint64_t align_plain_tiers_words(uint32_t nw_cap, uint32_t stride_a, uint32_t stride_b) {
  return (uint64_t)nw_cap * stride_a + (uint64_t)kPlainMaxWords * stride_b;
}

void launch_align_plain_tiers(const AlignArgs& a, uint32_t* jobs, uint32_t count,
                              uint32_t* scratch, uint64_t scratch_words, uint32_t nw_cap,
                              hipStream_t stream) {
  if (count == 0) return;
  if (nw_cap < 1 || nw_cap > kPlainMaxWords)
    throw HipError{NGIA_ERR_ARG, "row cap outside 1..2048 words"};
  // Tier b: 1,024 threads with full-length rows (8 MiB); tier a: what the
  // rest of the scratch holds at nw_cap words per thread, in whole blocks,
  // at most the plain form's grid and no more than the jobs need.
  const uint32_t stride_b = 1024;
  const uint64_t rest = scratch_words - (uint64_t)kPlainMaxWords * stride_b;
  if (scratch_words < (uint64_t)kPlainMaxWords * stride_b + 64ull * nw_cap)
    throw HipError{NGIA_ERR_ARG, "scratch too small for two tiers"};
  uint64_t blocks_a = rest / nw_cap / 64;
  blocks_a = std::min<uint64_t>(blocks_a, kPlainBlocks);
  blocks_a = std::min<uint64_t>(blocks_a, ((uint64_t)count + 63) / 64);
  blocks_a = std::max<uint64_t>(blocks_a, 1);
  const uint32_t stride_a = (uint32_t)blocks_a * 64;
  if (align_plain_tiers_words(nw_cap, stride_a, stride_b) > scratch_words)
    throw HipError{NGIA_ERR_ARG, "plain alignment scratch too small"};
  const uint32_t blocks = (uint32_t)blocks_a + stride_b / 64;
  hipLaunchKernelGGL(k_align_plain_tiers, dim3(blocks), dim3(64), 0, stream, a, jobs, count,
                     scratch, nw_cap, stride_a, stride_b);
  NGIA_HIP(hipGetLastError());
}

uint64_t align_plain_scratch_words(uint32_t nw_max, uint32_t max_blocks) {
  return (uint64_t)max_blocks * 64 * (nw_max ? nw_max : 1);
}

void launch_align_plain(const AlignArgs& a, uint32_t* jobs, uint32_t count, uint32_t nw_max,
                        uint32_t* scratch, uint64_t scratch_words, hipStream_t stream,
                        const uint32_t* count_ptr, uint32_t max_blocks) {
  if (count == 0) return;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)count + 63) / 64, max_blocks);
  const uint32_t stride = blocks * 64;
  if ((uint64_t)stride * nw_max > scratch_words)
    throw HipError{NGIA_ERR_ARG, "plain alignment scratch too small"};
  hipLaunchKernelGGL(k_align_plain, dim3(blocks), dim3(64), 0, stream, a, jobs, count, count_ptr,
                     scratch, stride);
  NGIA_HIP(hipGetLastError());
}

// Early exit.  The kernels compute the LCS of the band-masked match matrix
// (the bit-parallel update with the masks of words outside the band set to
// 0 is exactly the banded recurrence), and an LCS grows by at most one per
// target character; the pass test (buildGraph.hpp:248) is monotone in the
// LCS.  So a pair whose LCS after `done` characters, plus the characters left,
// fails the test scores 0 whatever follows.
__device__ inline bool cannot_pass(uint32_t lcs_now, uint32_t done, uint32_t R, uint32_t T,
                                   float thr) {
  const uint32_t left = T > done ? T - done : 0u;
  return lcs_score(lcs_now + left, R, T, thr) == 0u;
}

// A wavefront's claim on its bin's work queue.  Every pair taken through a
// per-pair atomic on one counter serialises a bin of many short pairs on that
// address, so one atomic claims `chunk` consecutive items of a slice and the
// wave hands them out in order.  The chunk follows the bin's size, so each
// wave still gets many claims and the queue drains evenly: count / (16
// waves' worth), between `lo` and `hi` (lo >= the most items handed out at
// once).  The wave starts on the slice of its XCD (kQueueSlices slices, one
// counter each); when a claim comes back empty it reads every slice's
// counter and moves to the next slice with items left, and with none left
// the queue is drained.  All members are wave-uniform; take() is called by
// the whole wave at a convergent point.
struct WaveClaim {
  uint32_t next = 0, end = 0, chunk = 1, count, slice;
  bool drained = false;
  __device__ WaveClaim(uint32_t count_, uint32_t lo, uint32_t hi) : count(count_) {
    const uint32_t per = count / (gridDim.x * (blockDim.x / 64u) * 16u);
    chunk = per < lo ? lo : per > hi ? hi : per;
    slice = blockIdx.x % kQueueSlices;
  }
  __device__ uint32_t slice_lo(uint32_t s) const {
    return (uint32_t)((uint64_t)count * s / kQueueSlices);
  }
  // n (wave-uniform, n <= lo) lanes take an item each; `rank` (< n) is this
  // lane's place among them.  A lane that gets no item (the queue is
  // drained) gets 0xFFFFFFFF, which is >= count.  Every item below count is
  // handed out exactly once: each slice's counter hands out its items once,
  // and a wave hands out its claims in order.
  __device__ uint32_t take(uint32_t n, uint32_t rank, uint32_t* queue) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t item = 0xFFFFFFFFu, given = 0;
    while (given < n && !drained) {
      if (next == end) {
        const uint32_t lo = slice_lo(slice), len = slice_lo(slice + 1) - lo;
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(queue + slice * kQueueStride, chunk);
        b = __builtin_amdgcn_readfirstlane(b);
        if (b < len) {
          next = lo + b;
          end = lo + (len - b < chunk ? len : b + chunk);
        } else {
          // this slice is exhausted: find one with items left, looking from
          // the next slice on (a counter only grows, so an exhausted slice
          // stays so)
          bool left = false;
          if (lane < kQueueSlices) {
            const uint32_t c = __hip_atomic_load(queue + lane * kQueueStride, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            left = c < slice_lo(lane + 1) - slice_lo(lane);
          }
          const uint32_t m = (uint32_t)__builtin_amdgcn_ballot_w64(left);
          if (m == 0) {
            drained = true;
          } else {
            const uint32_t rot = ((m >> slice) | (m << (kQueueSlices - slice))) & 0xFFu;
            slice = (slice + (uint32_t)__builtin_ctz(rot)) % kQueueSlices;
          }
        }
        continue;
      }
      const uint32_t k = n - given < end - next ? n - given : end - next;
      if (rank >= given && rank - given < k) item = next + (rank - given);
      next += k;
      given += k;
    }
    return item;
  }
};

// ------------------------------------------------ short window, thread/pair
// Pair list entry: pairs[p] = (rep, tar); geometry computed by the binning
// pass so that a + c == W for every pair of the list.
//
// Block cb of a W-block round uses register and Peq slot (cb + j) mod W for
// relative word j; the round is instantiated W times (Q = cb mod W) so that
// every index is a compile-time constant.
template <int W>
struct WinPair {
  uint32_t* rows;        // W row words (registers after inlining)
  uint32_t lcs;          // zero bits of the words that left the window
  const uint32_t* rr;    // rep record
  const uint32_t* tc;    // target byte codes, 8 dwords per block
  int32_t c, nw;
  uint32_t nb;
  uint32_t* lds;         // [W][NC][64] Peq slots
  uint32_t lane, lane4;
  uint4 nc0, nc1;        // codes of the next block, loaded one block ahead
  uint32_t slot_code;    // NC < 16: the code of slot s (1..NC-1) in bits 4(s-1)..
};

// The masks of one entering rep word into its Peq slot: the 15 nonzero codes
// (NC = 16), or the codes of slots 1..NC-1 of a compact alphabet.
template <int NC>
__device__ __forceinline__ void win_store_masks(uint32_t* slot, uint32_t lane, uint32_t p0,
                                                uint32_t p1, uint32_t p2, uint32_t p3,
                                                uint32_t slot_code) {
  if constexpr (NC == 16) {
    uint32_t m[16];
    peq_masks(p0, p1, p2, p3, m);
    #pragma unroll
    for (int k = 1; k < 16; k++) slot[k * 64 + lane] = m[k];
  } else {
    #pragma unroll
    for (int k = 1; k < NC; k++) {
      const uint32_t code = (slot_code >> (4 * (k - 1))) & 0xFu;
      const uint32_t x0 = (code & 1u) ? p0 : ~p0, x1 = (code & 2u) ? p1 : ~p1;
      const uint32_t x2 = (code & 4u) ? p2 : ~p2, x3 = (code & 8u) ? p3 : ~p3;
      slot[k * 64 + lane] = x0 & x1 & x2 & x3;
    }
  }
}

// The masks of one group of four characters (code bytes of cw) for every
// window register: M[mm][r] = Peq slot r's mask of character mm's code.
template <int W, int NC>
__device__ __forceinline__ void win_load_group(const uint32_t* lds, uint32_t cw, uint32_t lane4,
                                               uint32_t (&M)[4][W]) {
  constexpr int kSlot = NC * 64;
  #pragma unroll
  for (int mm = 0; mm < 4; mm++) {
    // byte address of Peq[code][lane] = code*256 + lane*4: v_perm takes byte
    // 0 from lane4 (selector 0) and byte 1 from code byte mm of cw (selector
    // 4+mm); bytes 2,3 are zero (selector 0x0C).  With a compact alphabet the
    // byte is the residue's slot.
    const uint32_t addr =
        __builtin_amdgcn_perm(cw, lane4, 0x0C0C0000u | ((4u + (uint32_t)mm) << 8));
    const uint32_t* pq = (const uint32_t*)((const char*)lds + addr);
    #pragma unroll
    for (int r = 0; r < W; r++) M[mm][r] = pq[r * kSlot];
  }
}

// Four characters against the window: relative word j lives in register
// (Q + j) mod W, carries run from relative word 0 upward.
template <int W, int Q>
__device__ __forceinline__ void win_compute_group(uint32_t* rows, const uint32_t (&M)[4][W]) {
  #pragma unroll
  for (int mm = 0; mm < 4; mm++) {
    uint32_t cin = 0;
    #pragma unroll
    for (int j = 0; j < W; j++) {
      constexpr int kW = W;
      const int r = (Q + j) % kW;
      rows[r] = lcs_step(rows[r], M[mm][r], cin, &cin);
    }
  }
}

// Window widths up to this one read the next group's masks ahead.
constexpr int kWinPrefetchMax = 8;

template <int W, int NC, int Q>
__device__ __forceinline__ bool win_block(WinPair<W>& s, uint32_t cb0) {
  constexpr int kSlot = NC * 64;  // dwords per Peq slot
  const uint32_t cb = cb0 + Q;
  if (cb >= s.nb) return false;
  uint32_t* const lds = s.lds;
  // This block's codes arrived during the previous block; fetch the next
  // block's now (past the last block the read stays inside the buffer's
  // 64-byte tail padding), and the planes of the word entering at the end.
  uint32_t q[8] = {s.nc0.x, s.nc0.y, s.nc0.z, s.nc0.w, s.nc1.x, s.nc1.y, s.nc1.z, s.nc1.w};
  s.nc0 = *(const uint4*)(s.tc + 8 * (cb + 1));
  s.nc1 = *(const uint4*)(s.tc + 8 * (cb + 1) + 4);
  const int32_t w = (int32_t)cb + s.c;
  uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
  if (w < s.nw) {
    const uint32_t* pl = s.rr + 1 + 4 * w;
    p0 = pl[0]; p1 = pl[1]; p2 = pl[2]; p3 = pl[3];
  }
  if constexpr (W <= kWinPrefetchMax) {
    // Eight groups of four characters.  The masks of the next group are read
    // while the current group is computed (two buffers), so only the block's
    // first group waits for LDS; the code words stay in registers, consumed
    // two groups per step.
    uint32_t MA[4][W], MB[4][W];
    win_load_group<W, NC>(lds, q[0], s.lane4, MA);
    #pragma unroll 1
    for (int h = 0; h < 4; h++) {
      win_load_group<W, NC>(lds, q[1], s.lane4, MB);
      win_compute_group<W, Q>(s.rows, MA);
      if (h < 3) win_load_group<W, NC>(lds, q[2], s.lane4, MA);
      win_compute_group<W, Q>(s.rows, MB);
      #pragma unroll
      for (int k = 0; k < 6; k++) q[k] = q[k + 2];
    }
  } else {
    // Wide windows: one group per step, its code word staged in LDS, so that
    // the round (unrolled W times) stays small enough for the instruction
    // cache; two buffers of 4W masks would double it.
    uint32_t* const cst = lds + W * kSlot;
    #pragma unroll
    for (int k = 0; k < 8; k++) cst[k * 64 + s.lane] = q[k];
    for (int g = 0; g < 8; g++) {
      uint32_t M[4][W];
      win_load_group<W, NC>(lds, cst[g * 64 + s.lane], s.lane4, M);
      win_compute_group<W, Q>(s.rows, M);
    }
  }
  // Slide: relative word 0 (register Q) leaves, word cb + c enters there.
  s.lcs += 32u - __popc(s.rows[Q]);
  s.rows[Q] = 0xFFFFFFFFu;
  // Past the rep's last word the planes are 0, and so is every mask of a
  // nonzero code (each needs a set plane bit or a clear one with another set).
  win_store_masks<NC>(lds + Q * kSlot, s.lane, p0, p1, p2, p3, s.slot_code);
  return true;
}

template <int W, int NC, int Q>
__device__ __forceinline__ void win_round(WinPair<W>& s, uint32_t cb0) {
  if constexpr (Q < W) {
    if (win_block<W, NC, Q>(s, cb0)) win_round<W, NC, Q + 1>(s, cb0);
  }
}

// NC = 16 indexes the Peq table by residue code; NC < 16 by the slot of a
// compact alphabet (A.slots), which shrinks the table and raises occupancy.
// Each lane holds one pair and takes the next from the wave's claim on the
// work queue (WaveClaim: chunks of 64 to 256 items) when its pair ends --
// at the end of a round of W blocks, where the registers are back in their
// starting order -- or when the pair can no longer pass (cannot_pass).
template <int W, int NC>
__global__ __launch_bounds__(64) void k_align_win(AlignArgs A, const uint2* __restrict__ pairs,
                                                  const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ off_ptr,
                                                  const uint32_t* __restrict__ count_ptr,
                                                  uint32_t* __restrict__ work,
                                                  uint32_t* __restrict__ score) {
  constexpr int kSlot = NC * 64;
  // Peq slots [W][NC codes][64 lanes] (then, for wide windows, the block's
  // staged codes [8][64]).  Static so that its LDS address is a compile-time
  // constant.
  __shared__ uint32_t lds[W * kSlot + (W <= kWinPrefetchMax ? 0 : 8 * 64)];
  const uint32_t lane = threadIdx.x;
  const uint32_t lane4 = lane * 4;
  const uint8_t* const rcodes = NC == 16 ? A.codes : A.slots;
  #pragma unroll
  for (int s = 0; s < W; s++) lds[s * kSlot + lane] = 0;  // code (slot) 0 never matches
  const uint32_t count = *count_ptr;
  const uint32_t* const lst = list + (off_ptr ? *off_ptr : 0u);
  uint32_t rows[W];
  #pragma unroll
  for (int j = 0; j < W; j++) rows[j] = 0xFFFFFFFFu;
  const uint4 zero4 = make_uint4(0u, 0u, 0u, 0u);
  WinPair<W> st{rows, 0u, A.packed, (const uint32_t*)rcodes, 0, 0, 0u, lds, lane, lane4,
                zero4, zero4, A.slot_code};
  bool active = false, drained = false;
  uint32_t p = 0, R = 0, T = 0, cb0 = 0;
  uint64_t ran = 0;  // tally: reference tiles of the blocks run
  WaveClaim q(count, 64u, 256u);
  for (;;) {
    const uint64_t want = __builtin_amdgcn_ballot_w64(!active && !drained);
    if (want) {
      const uint32_t idx = q.take((uint32_t)__popcll(want),
                                  (uint32_t)__popcll(want & ((1ull << lane) - 1ull)), work);
      if (!active && !drained) {
        if (idx >= count) {
          drained = true;
        } else {
          p = lst[idx];
          const uint2 pr = pairs[p];
          const uint32_t* rr = A.packed + A.word_off[pr.x];
          R = rr[0];
          T = A.lengths[pr.y];
          const Band bd = band_edges(R, T, A.thr);
          const int32_t a = (int32_t)(bd.lsft / 32);
          const int32_t nw = (int32_t)((R + 31) / 32);
          const uint32_t* tc = (const uint32_t*)(rcodes + A.code_off[pr.y]);
          // The first block's codes, and the planes of the W window words
          // at block 0, all in one round trip: relative word j <-> absolute
          // j - a, register j, slot j.  A word outside the record reads the
          // packed area's first four words (the buffer holds at least 64) and
          // is zeroed (the masks of all-zero planes: 0 for every nonzero
          // code).  Loaded word by word, before their masks were stored, the
          // planes took W round trips.
          st.nc0 = *(const uint4*)tc;
          st.nc1 = *(const uint4*)(tc + 4);
          uint4 pw[W];
          #pragma unroll
          for (int j = 0; j < W; j++) {
            const int32_t w = j - a;
            const uint32_t* pl = (w >= 0 && w < nw) ? rr + 1 + 4 * w : A.packed;
            pw[j] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
          }
          #pragma unroll
          for (int j = 0; j < W; j++) {
            rows[j] = 0xFFFFFFFFu;
            const int32_t w = j - a;
            const bool in = w >= 0 && w < nw;
            win_store_masks<NC>(lds + j * kSlot, lane, in ? pw[j].x : 0u, in ? pw[j].y : 0u,
                                in ? pw[j].z : 0u, in ? pw[j].w : 0u, A.slot_code);
          }
          st.lcs = 0u;
          st.rr = rr;
          st.tc = tc;
          st.c = (int32_t)((bd.rsft + 31) / 32);
          st.nw = nw;
          st.nb = (T + 31) / 32;
          cb0 = 0;
          active = true;
        }
      }
    }
    if (__builtin_amdgcn_ballot_w64(active) == 0) break;
    if (active) {
      win_round<W, NC, 0>(st, cb0);
      cb0 += W;
      uint32_t lcs = st.lcs;
      #pragma unroll
      for (int j = 0; j < W; j++) lcs += 32u - __popc(rows[j]);
      const bool end = cb0 >= st.nb;
      const bool fail = !end && !A.raw_lcs && cannot_pass(lcs, 32u * cb0, R, T, A.thr);
      if (end || fail) {
        score[p] = fail ? 0u : pair_value(A, lcs, R, T);
        if (A.tally) ran += ref_tiles(R, T, A.thr, end ? st.nb : cb0);
        active = false;
      }
    }
  }
  tally_add(A, ran);
}

// ------------------------------------------------ one wavefront per pair
// The wave holds the 64K consecutive rep words [B, B + 64K): lane l keeps
// words B + lK + m in rows[m].  B = K * floor((cb - a) / K) trails the band
// bottom: when the bottom crosses a multiple of K the registers shift down
// one lane (lane 0's K words leave, frozen for good), so the band
// [cb - a, cb + c) lies inside the held range whenever W + K - 1 <= 64K.
// Held words outside the band see match mask 0: below the band they are
// frozen words (mask 0 and carry-in 0 leave a word unchanged and pass carry
// 0 up), above it untouched all-ones words (a fixed point for either
// carry-in, absorbing the carry the reference drops at the band top).
//
// Match masks: peq[code * S + (w mod S)], S = 64K slots per code.  A word's
// 15 masks are written by lanes 1..15 when it enters the band and cleared
// when it leaves; a lane's K slots are contiguous and K-aligned, so they load
// as one ds_read_b32 / b64 / b128 (or K/4 of the latter) per character.
// Target codes of the next block and the bit planes of the next entering
// word are loaded one block ahead.  The block is one wavefront, whose LDS
// accesses execute in program order, so no workgroup barrier is needed; the
// empty asm with a memory clobber keeps the compiler from reordering them.
__device__ inline void lds_order() { asm volatile("" ::: "memory"); }

// Match mask of code k (1..15) over one word from its four bit planes.
__device__ inline uint32_t code_mask(uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3,
                                     uint32_t k) {
  const uint32_t x0 = (k & 1u) ? p0 : ~p0, x1 = (k & 2u) ? p1 : ~p1;
  const uint32_t x2 = (k & 4u) ? p2 : ~p2, x3 = (k & 8u) ? p3 : ~p3;
  return x0 & x1 & x2 & x3;
}

template <int K>
__device__ inline void load_masks(const uint32_t* p, uint32_t (&M)[K]) {
  if constexpr (K == 1) {
    M[0] = p[0];
  } else if constexpr (K == 2) {
    const uint2 v = *(const uint2*)p;
    M[0] = v.x;
    M[1] = v.y;
  } else {
    #pragma unroll
    for (int q = 0; q < K / 4; q++) {
      const uint4 v = ((const uint4*)p)[q];
      M[4 * q] = v.x;
      M[4 * q + 1] = v.y;
      M[4 * q + 2] = v.z;
      M[4 * q + 3] = v.w;
    }
  }
}

template <int K>
__global__ __launch_bounds__(64) void k_align_wave(AlignArgs A, const uint2* __restrict__ pairs,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ off_ptr,
                                                   const uint32_t* __restrict__ count_ptr,
                                                   uint32_t* __restrict__ work,
                                                   uint32_t* __restrict__ score) {
  constexpr uint32_t S = 64 * K;  // slots per code
  constexpr int LG = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : K == 8 ? 3 : K == 16 ? 4 : 5;
  extern __shared__ __align__(16) uint32_t peq[];  // [16][S]
  const uint32_t lane = threadIdx.x;
  const uint32_t count = *count_ptr;
  const uint32_t* const lst = list + (off_ptr ? *off_ptr : 0u);
  // The wider the window, the longer the pair: such waves are the critical
  // path of the concurrent bins, so the SIMD issues them first.
  __builtin_amdgcn_s_setprio(K >= 4 ? 3 : K == 2 ? 2 : 1);
  uint64_t ran = 0;  // tally (lane 0): reference tiles of the blocks run
  WaveClaim q(count, 1u, K == 1 ? 4u : 1u);
  for (;;) {
    const uint32_t it = q.take(1u, 0u, work);
    if (it >= count) break;
    const uint32_t p = lst[it];
    const uint2 pr = pairs[p];
    const uint32_t* rr = A.packed + A.word_off[pr.x];
    const uint32_t R = rr[0];
    const uint32_t T = A.lengths[pr.y];
    const Band bd = band_edges(R, T, A.thr);
    const int32_t a = (int32_t)(bd.lsft / 32);
    const int32_t c = (int32_t)((bd.rsft + 31) / 32);
    const int32_t nw = (int32_t)((R + 31) / 32);
    const uint32_t nb = (T + 31) / 32;
    const uint32_t* tcw = (const uint32_t*)(A.codes + A.code_off[pr.y]);
    // Clear the table, then write the words in the band at block 0: [0, min(c, nw)).
    for (uint32_t x = lane; x < 16 * S / 4; x += 64) ((uint4*)peq)[x] = make_uint4(0u, 0u, 0u, 0u);
    lds_order();
    const int32_t top0 = c < nw ? c : nw;
    for (int32_t w = (int32_t)lane; w < top0; w += 64) {
      const uint32_t* pl = rr + 1 + 4 * w;
      uint32_t m[16];
      peq_masks(pl[0], pl[1], pl[2], pl[3], m);
      #pragma unroll
      for (int k = 1; k < 16; k++) peq[k * S + (uint32_t)w % S] = m[k];
    }
    lds_order();
    uint32_t rows[K];
    #pragma unroll
    for (int m = 0; m < K; m++) rows[m] = 0xFFFFFFFFu;
    int32_t B = ((-a) >> LG) * K;  // K * floor(-a / K)
    uint32_t vs = (uint32_t)(B + (int32_t)(lane * K) + (int32_t)(64 * S)) % S;
    uint32_t lcs = 0;
    uint32_t cur = (lane < 8 && nb > 0) ? tcw[lane] : 0u;
    // Wait for `cur` here, so that the loop below never has to wait on the
    // next block's codes while they are in flight.
    asm volatile("" : "+v"(cur));
    // bit planes of the word entering at the end of the current block
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    if (c < nw) {
      const uint32_t* pl = rr + 1 + 4 * c;
      q0 = pl[0]; q1 = pl[1]; q2 = pl[2]; q3 = pl[3];
    }
    bool failed = false;
    uint32_t ran_blocks = nb;
    for (uint32_t cb = 0; cb < nb; cb++) {
      const uint32_t nxt = (lane < 8 && cb + 1 < nb) ? tcw[8 * (cb + 1) + lane] : 0u;
      const uint32_t* pqb = peq + vs;
      // The masks of group g's four characters; for K <= 8 those of group
      // g + 1 are loaded while group g is computed (two buffers, g mod 2), so
      // that only the block's first group waits for its loads.
      constexpr bool kAhead = K <= 8;
      uint32_t Mb[kAhead ? 2 : 1][4][K];
      auto load_group = [&](int g, uint32_t (&Mx)[4][K]) {
        const uint32_t cw = __builtin_amdgcn_readlane(cur, g);
        #pragma unroll
        for (int mm = 0; mm < 4; mm++) load_masks<K>(pqb + ((cw >> (8 * mm)) & 0xFu) * S, Mx[mm]);
      };
      if constexpr (kAhead) load_group(0, Mb[0]);
      #pragma unroll
      for (int g = 0; g < 8; g++) {
        uint32_t(&Mq)[4][K] = Mb[kAhead ? (g & 1) : 0];
        if constexpr (kAhead) {
          if (g < 7) load_group(g + 1, Mb[(g + 1) & 1]);
        } else {
          load_group(g, Mq);
        }
        #pragma unroll
        for (int mm = 0; mm < 4; mm++) {
          uint32_t(&M)[K] = Mq[mm];
          // Both K-word add chains carry through a lane mask (the carry
          // operand of v_add_co / v_addc_co): the first chain's carry-out is
          // the lane's generate bit G as it stands, and the lookahead's
          // carries C enter the second chain as its carry-in mask, so no
          // carry is kept per lane in a VGPR and tested around the ballots.
          // The asm also keeps the compiler from folding the two chains.
          uint32_t s0[K];
          uint64_t cy;
          asm("v_add_co_u32 %0, %1, %2, %3" : "=v"(s0[0]), "=s"(cy) : "v"(rows[0]), "v"(rows[0] & M[0]));
          if constexpr (K == 1) {
            const uint64_t G = cy;  // lanes whose word carries out
            const uint64_t P = __builtin_amdgcn_ballot_w64(s0[0] == 0xFFFFFFFFu);
            uint64_t ci = ((G | P) + G) ^ P;  // carry into each lane
            uint32_t s;
            asm("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(s), "=s"(ci) : "v"(s0[0]), "s"(ci));
            rows[0] = s | (rows[0] & ~M[0]);
            asm("" : "+v"(rows[0]));  // as in lcs_step: one bitop3, not two
          } else {
            #pragma unroll
            for (int m = 1; m < K; m++)
              asm("v_addc_co_u32 %0, %1, %2, %3, %4"
                  : "=v"(s0[m]), "=s"(cy) : "v"(rows[m]), "v"(rows[m] & M[m]), "s"(cy));
            uint32_t all = s0[0];
            #pragma unroll
            for (int m = 1; m < K; m++) all &= s0[m];
            const uint64_t G = cy;
            const uint64_t P = __builtin_amdgcn_ballot_w64(all == 0xFFFFFFFFu);
            uint64_t ci = ((G | P) + G) ^ P;
            #pragma unroll
            for (int m = 0; m < K; m++) {
              uint32_t s;
              asm("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(s), "=s"(ci) : "v"(s0[m]), "s"(ci));
              rows[m] = s | (rows[m] & ~M[m]);
              asm("" : "+v"(rows[m]));
            }
          }
        }
      }
      // Word cb - a leaves the band, word cb + c enters it for block cb + 1.
      const int32_t wl = (int32_t)cb - a, we = (int32_t)cb + c;
      if (lane >= 1 && lane < 16) {
        if (wl >= 0 && wl < nw) peq[lane * S + (uint32_t)wl % S] = 0u;
        if (we < nw) peq[lane * S + (uint32_t)we % S] = code_mask(q0, q1, q2, q3, lane);
      }
      lds_order();
      if (we + 1 < nw) {
        const uint32_t* pl = rr + 1 + 4 * (we + 1);
        q0 = pl[0]; q1 = pl[1]; q2 = pl[2]; q3 = pl[3];
      }
      // Trail the band bottom of block cb + 1.
      const int32_t Bn = (((int32_t)cb + 1 - a) >> LG) * K;
      if (Bn != B) {  // Bn == B + K: lane 0's words leave the held range
        #pragma unroll
        for (int m = 0; m < K; m++) {
          lcs += 32u - __popc(__builtin_amdgcn_readfirstlane(rows[m]));
          const uint32_t up = __shfl_down(rows[m], 1);
          rows[m] = lane == 63 ? 0xFFFFFFFFu : up;
        }
        B = Bn;
        vs = (vs + K) % S;
      }
      cur = nxt;
      asm volatile("" : "+v"(cur));
      // Every 8 blocks: a pair that cannot pass any more ends with score 0.
      if (!A.raw_lcs && (cb & 7u) == 7u && cb + 1 < nb) {
        uint32_t part = 0;
        #pragma unroll
        for (int m = 0; m < K; m++) part += 32u - __popc(rows[m]);
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
        if (cannot_pass(lcs + part, 32u * (cb + 1), R, T, A.thr)) {
          failed = true;
          ran_blocks = cb + 1;
          break;
        }
      }
    }
    uint32_t part = 0;
    #pragma unroll
    for (int m = 0; m < K; m++) part += 32u - __popc(rows[m]);
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
    lcs += part;
    if (lane == 0) score[p] = failed ? 0u : pair_value(A, lcs, R, T);
    if (A.tally && lane == 0) ran += ref_tiles(R, T, A.thr, ran_blocks);
    lds_order();
  }
  tally_add(A, ran);
}

// ------------------------------------------------ several pairs per wavefront
// The codes a segment without work reads: 32 zero bytes (code 0 matches
// nothing, so its rows stay as they are).
__device__ const uint4 kNoCodes[2] = {};

// Lane (id xor OFF) of the same 32-lane half (ds_swizzle's bit-mask mode:
// and-mask 31, or-mask 0, xor-mask OFF).
template <int OFF>
__device__ inline uint32_t swizzle_xor(uint32_t v) {
  static_assert(OFF >= 1 && OFF <= 16, "within 32 lanes");
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (OFF << 10) | 0x1F);
}

// One pair per G-lane segment (64 / G pairs per wavefront), K words per lane
// ((G, K) = (32, 1), (16, 1..4) or (8, 2..3)), for band windows of at most
// K (G - 1) words.  In the segment
// of lanes [base, base + G), lane base + l holds the words B + K l + m,
// m < K, of its pair's rep; B = K floor((cb - a) / K) trails the band bottom
// cb - a, so the segment's registers shift down one lane every K blocks
// (as in the wavefront kernel), and the band [cb - a, cb + c) lies inside the
// held words [B, B + K G - 1).  The top lane's last word is a sentinel whose
// row restarts at 0 each block: within a block of 32 characters that row
// stays below 2^31 (a row at most doubles plus one per character), so it
// neither generates nor propagates a carry.  One 64-lane carry lookahead per
// character thus serves every segment, and no carry crosses from one segment
// into the next.
// Match masks: K tables of [16 codes][64 slots]; word w of a segment is in
// table w mod K at slot base + (floor(w / K) mod G), so a lane's K words sit
// at one slot of the K tables, one v_perm address per character.  The
// sentinel's slot is the one the word that left last held (cleared), so its
// mask is 0.
// With K = 2 a segment of 16 lanes takes windows of up to 30 words, which
// K = 1 needs 32 lanes for: four pairs per wavefront share each character's
// lookahead instead of two, and a lane's two words share its mask address.
// K = 3 and 4 (windows of up to 45 and 60 words) keep four pairs per
// wavefront where a wavefront per pair would hold 64 words for each; with
// 8-lane segments of three words (windows of up to 21 words) eight pairs
// share each lookahead, and a pair holds fewer words beyond its band than in
// 16 lanes of two words.
// Each segment takes the next pair from the wave's claim on the work queue
// (WaveClaim) when its pair ends;
// the segments run their blocks in lockstep, one without work holding code 0
// (no change), and the wave leaves when every segment found the queue empty.
template <int G, int K>
__global__ __launch_bounds__(64) void k_align_seg(AlignArgs A, const uint2* __restrict__ pairs,
                                                  const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ off_ptr,
                                                  const uint32_t* __restrict__ count_ptr,
                                                  uint32_t* __restrict__ work,
                                                  uint32_t* __restrict__ score) {
  static_assert((G == 32 && K == 1) || (G == 16 && K >= 1 && K <= 4) || (G == 8 && (K == 2 || K == 3)),
                "segments of 32 lanes with one word, 16 with 1-4, 8 with 2-3");
  // floor(x / K) for either sign (x / K rounds toward zero)
  auto fdiv = [](int32_t x) { return (x >= 0 ? x : x - (K - 1)) / K; };
  // [K][16 codes][64 slots]; code 0 stays 0
  __shared__ __align__(16) uint32_t peq[K * 16 * 64];
  const uint32_t lane = threadIdx.x;
  const uint32_t l = lane & (uint32_t)(G - 1), base = lane - l;
  const uint32_t count = *count_ptr;
  const uint32_t* const lst = list + (off_ptr ? *off_ptr : 0u);
  __builtin_amdgcn_s_setprio(1);
  for (uint32_t x = lane; x < K * 16 * 64; x += 64) peq[x] = 0u;
  lds_order();
  // The state of this lane's segment (equal in its G lanes; lcs in lane base).
  bool active = false, need = true;
  uint32_t p = 0, R = 0, T = 0, nb = 0, cb = 0, lcs = 0;
  int32_t a = 0, c = 0, nw = 0, B = 0;
  const uint32_t* rr = A.packed;
  const uint4* tcq = (const uint4*)A.codes;  // the target's codes, 32 bytes per block
  uint32_t row[K], vs = base, q0 = 0, q1 = 0, q2 = 0, q3 = 0;
  #pragma unroll
  for (int m = 0; m < K; m++) row[m] = 0u;
  uint4 c0 = make_uint4(0u, 0u, 0u, 0u), c1 = c0;  // this block's 32 codes, one byte each
  uint64_t ran = 0;  // tally (lane base): reference tiles of the blocks run
  // the sentinel: the top lane's last word
  auto sentinel = [&](int m) { return l == (uint32_t)(G - 1) && m == K - 1; };
  // the slot of word w in table w mod K
  auto slot_of = [&](int32_t w) { return base + ((uint32_t)fdiv(w) & (uint32_t)(G - 1)); };
  WaveClaim q(count, 64u / G, 4u * (64u / G));
  for (;;) {
    // Segments whose pair ended take the next ones; one that finds the queue
    // empty stays idle, and one whose pair has no characters scores it and
    // takes another in the next pass.
    const uint64_t want = __builtin_amdgcn_ballot_w64(need && l == 0);
    if (want) {
      const uint32_t it = __shfl(q.take((uint32_t)__popcll(want),
                                        (uint32_t)__popcll(want & ((1ull << lane) - 1ull)), work),
                                 (int)base);
      if (need) {
        need = false;
        active = false;
        if (it >= count) {
          tcq = kNoCodes;
          c0 = c1 = make_uint4(0u, 0u, 0u, 0u);
        } else {
          p = lst[it];
          const uint2 pr = pairs[p];
          rr = A.packed + A.word_off[pr.x];
          R = rr[0];
          T = A.lengths[pr.y];
          const Band bd = band_edges(R, T, A.thr);
          a = (int32_t)(bd.lsft / 32);
          c = (int32_t)((bd.rsft + 31) / 32);
          nw = (int32_t)((R + 31) / 32);
          nb = (T + 31) / 32;
          tcq = (const uint4*)(A.codes + A.code_off[pr.y]);
          if (nb == 0) {
            if (l == 0) score[p] = pair_value(A, 0u, R, T);
            need = true;
          } else {
            // The segment's slots: cleared, then the words in the band at
            // block 0, [0, min(c, nw)) with c <= K (G - 1): word w = K l + m
            // at slot base + l of table m.
            #pragma unroll
            for (int m = 0; m < K; m++) {
              #pragma unroll
              for (int k = 1; k < 16; k++) peq[(m * 16 + k) * 64 + lane] = 0u;
            }
            lds_order();
            const int32_t top0 = c < nw ? c : nw;
            #pragma unroll
            for (int m = 0; m < K; m++) {
              const int32_t w = (int32_t)(K * l) + m;
              if (w < top0) {
                const uint32_t* pl = rr + 1 + 4 * w;
                uint32_t mk[16];
                peq_masks(pl[0], pl[1], pl[2], pl[3], mk);
                #pragma unroll
                for (int k = 1; k < 16; k++) peq[(m * 16 + k) * 64 + lane] = mk[k];
              }
            }
            lds_order();
            B = fdiv(-a) * K;
            #pragma unroll
            for (int m = 0; m < K; m++) row[m] = sentinel(m) ? 0u : 0xFFFFFFFFu;
            vs = base + ((uint32_t)(fdiv(B) + (int32_t)l) & (uint32_t)(G - 1));
            lcs = 0;
            cb = 0;
            c0 = tcq[0];
            c1 = tcq[1];
            q0 = q1 = q2 = q3 = 0;
            if (c < nw) {
              const uint32_t* pl = rr + 1 + 4 * c;
              q0 = pl[0]; q1 = pl[1]; q2 = pl[2]; q3 = pl[3];
            }
            active = true;
          }
        }
      }
      if (__builtin_amdgcn_ballot_w64(need) != 0) continue;
    }
    if (__builtin_amdgcn_ballot_w64(active) == 0) break;
    // This block's codes (held in registers by every lane of the segment);
    // the next block's are loaded now, to arrive during this one.
    const uint32_t cw[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    // Clamped to the pair's last block: a segment whose pair ends takes its
    // next pair's codes at the claim, and one left without work reads the
    // zero codes tcq then points at.
    {
      const uint32_t nx = active ? (cb + 1 < nb ? cb + 1 : nb - 1) : 0u;
      c0 = tcq[2 * nx];
      c1 = tcq[2 * nx + 1];
    }
    const uint32_t vs4 = vs * 4;
    // The masks of group g's four characters; those of group g + 1 are
    // loaded while group g is computed (two buffers, g mod 2).
    uint32_t Mb[2][4][K];
    auto load_group = [&](int g, uint32_t (&Mx)[4][K]) {
      #pragma unroll
      for (int mm = 0; mm < 4; mm++) {
        // byte address code * 256 + slot * 4 (slot < 64); table m 4 KiB on
        const uint32_t addr =
            __builtin_amdgcn_perm(cw[g], vs4, 0x0C0C0000u | ((4u + (uint32_t)mm) << 8));
        #pragma unroll
        for (int m = 0; m < K; m++)
          Mx[mm][m] = *(const uint32_t*)((const char*)peq + m * 4096 + addr);
      }
    };
    load_group(0, Mb[0]);
    #pragma unroll
    for (int g = 0; g < 8; g++) {
      uint32_t(&Mq)[4][K] = Mb[g & 1];
      if (g < 7) load_group(g + 1, Mb[(g + 1) & 1]);
      #pragma unroll
      for (int mm = 0; mm < 4; mm++) {
        const uint32_t(&M)[K] = Mq[mm];
        // As in k_align_wave: the carry chain through a lane mask (the carry
        // operand of v_add_co / v_addc_co) and the ballot lookahead.
        uint32_t s0[K];
        uint64_t cy;
        asm("v_add_co_u32 %0, %1, %2, %3" : "=v"(s0[0]), "=s"(cy) : "v"(row[0]), "v"(row[0] & M[0]));
        uint32_t all = s0[0];
        #pragma unroll
        for (int m = 1; m < K; m++) {
          asm("v_addc_co_u32 %0, %1, %2, %3, %4"
              : "=v"(s0[m]), "=s"(cy) : "v"(row[m]), "v"(row[m] & M[m]), "s"(cy));
          all &= s0[m];
        }
        const uint64_t G1 = cy;
        const uint64_t P = __builtin_amdgcn_ballot_w64(all == 0xFFFFFFFFu);
        uint64_t ci = ((G1 | P) + G1) ^ P;
        #pragma unroll
        for (int m = 0; m < K; m++) {
          uint32_t sum;
          asm("v_addc_co_u32 %0, %1, %2, 0, %3" : "=v"(sum), "=s"(ci) : "v"(s0[m]), "s"(ci));
          row[m] = sum | (row[m] & ~M[m]);
          asm("" : "+v"(row[m]));
        }
      }
    }
    // Word cb - a leaves the band, word cb + c enters it for block cb + 1.
    if (active) {
      const int32_t wl = (int32_t)cb - a, we = (int32_t)cb + c;
      // lane l writes code l's mask (and, in 8-lane segments, code l + 8's)
      #pragma unroll
      for (int h = 0; h < (G < 16 ? 2 : 1); h++) {
        const uint32_t k = l + (uint32_t)(h * G);
        if (k >= 1 && k < 16) {
          if (wl >= 0 && wl < nw)
            peq[((wl % K) * 16 + (int32_t)k) * 64 + slot_of(wl)] = 0u;
          if (we < nw)
            peq[((we % K) * 16 + (int32_t)k) * 64 + slot_of(we)] = code_mask(q0, q1, q2, q3, k);
        }
      }
      lds_order();
      if (we + 1 < nw) {
        const uint32_t* pl = rr + 1 + 4 * (we + 1);
        q0 = pl[0]; q1 = pl[1]; q2 = pl[2]; q3 = pl[3];
      }
    }
    // B trails the band bottom of block cb + 1: when it moves up K words, the
    // words in lane base leave (counted there), the rows shift down one lane
    // and the top lane's words are new (all-ones) but for the sentinel; the
    // sentinel restarts at 0 every block.
    const int32_t Bn = fdiv((int32_t)cb + 1 - a) * K;
    if (K == 1 || Bn != B) {
      #pragma unroll
      for (int m = 0; m < K; m++) {
        if (l == 0) lcs += 32u - __popc(row[m]);
        // lane l + 1's word: within a 16-lane DPP row (row_shl:1) for
        // segments of up to 16 lanes, where the row's last lane is a
        // segment's top and takes a fresh word anyway
        uint32_t up;
        if constexpr (G <= 16)
          up = (uint32_t)__builtin_amdgcn_mov_dpp((int)row[m], 0x101, 0xF, 0xF, true);
        else
          up = __shfl_down(row[m], 1);
        // lane G - 2 took the top lane's words: its last is a new word where
        // the sentinel was
        const bool fresh = l == (uint32_t)(G - 1) || (l == (uint32_t)(G - 2) && m == K - 1);
        row[m] = sentinel(m) ? 0u : (fresh ? 0xFFFFFFFFu : up);
      }
      vs = base + ((vs + 1u) & (uint32_t)(G - 1));
      B = Bn;
    } else {
      #pragma unroll
      for (int m = 0; m < K; m++)
        if (sentinel(m)) row[m] = 0u;
    }
    cb++;
    // At the pair's end: the LCS summed over the segment (the words that left
    // are counted in lane base).  No early exit here: a test every 8 blocks
    // cost more than the pairs it ended saved (these pairs are short).
    if (active && cb == nb) {
      uint32_t part = l == 0 ? lcs : 0u;
      #pragma unroll
      for (int m = 0; m < K; m++)
        if (!sentinel(m)) part += 32u - __popc(row[m]);
      // butterfly sum over the segment: swizzles within 32 lanes (lane id
      // xor off), no address arithmetic
      if constexpr (G >= 32) part += swizzle_xor<16>(part);
      if constexpr (G >= 16) part += swizzle_xor<8>(part);
      part += swizzle_xor<4>(part);
      part += swizzle_xor<2>(part);
      part += swizzle_xor<1>(part);
      if (l == 0) score[p] = pair_value(A, part, R, T);
      if (A.tally && l == 0) ran += ref_tiles(R, T, A.thr, nb);
      active = false;
      need = true;
    }
  }
  lds_order();
  tally_add(A, ran);
}

// ------------------------------------------------ launchers for the bins
size_t win_lds_bytes(int W, bool compact) {
  return (size_t)W * (compact ? kSlotCodes + 1 : 16) * 256 + (W <= kWinPrefetchMax ? 0 : 8 * 256);
}
size_t wave_lds_bytes(int K) { return (size_t)16 * 64 * K * 4; }
size_t seg_lds_bytes() { return (size_t)16 * 64 * 4; }

// The binned kernels read target lengths from a.lengths.
static void require_lengths(const AlignArgs& a) {
  if (!a.lengths) throw HipError{NGIA_ERR_ARG, "binned alignment without record lengths"};
}

#define NGIA_WIN_CASE(Wv)                                                                  \
  case Wv:                                                                                 \
    if (a.slots)                                                                           \
      hipLaunchKernelGGL((k_align_win<Wv, kSlotCodes + 1>), dim3(blocks), dim3(64), 0,     \
                         stream, a, pairs, list, off_ptr, count_ptr, work, score);         \
    else                                                                                   \
      hipLaunchKernelGGL((k_align_win<Wv, 16>), dim3(blocks), dim3(64), 0, stream, a,      \
                         pairs, list, off_ptr, count_ptr, work, score);                    \
    break;

void launch_align_win(int W, const AlignArgs& a, const uint2* pairs, const uint32_t* list,
                      const uint32_t* off_ptr, const uint32_t* count_ptr, uint32_t* work, uint32_t* score,
                      uint32_t blocks, hipStream_t stream) {
  require_lengths(a);
  switch (W) {
    NGIA_WIN_CASE(2)
    NGIA_WIN_CASE(3)
    NGIA_WIN_CASE(4)
    NGIA_WIN_CASE(5)
    NGIA_WIN_CASE(6)
    NGIA_WIN_CASE(7)
    NGIA_WIN_CASE(8)
    NGIA_WIN_CASE(9)
    NGIA_WIN_CASE(10)
    NGIA_WIN_CASE(11)
    NGIA_WIN_CASE(12)
    NGIA_WIN_CASE(13)
    NGIA_WIN_CASE(14)
    NGIA_WIN_CASE(15)
    NGIA_WIN_CASE(16)
    default:
      throw HipError{NGIA_ERR_ARG, "no window kernel for W=" + std::to_string(W)};
  }
  NGIA_HIP(hipGetLastError());
}

int align_kernels_setup(size_t lds_per_block) {
  // K = 16, 32 need more than 64 KiB of LDS per workgroup; when the device
  // refuses, windows wider than 64*8 words go to the plain kernel instead.
  if (lds_per_block < wave_lds_bytes(32)) return 8;
  if (hipFuncSetAttribute((const void*)k_align_wave<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)wave_lds_bytes(16)) != hipSuccess ||
      hipFuncSetAttribute((const void*)k_align_wave<32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)wave_lds_bytes(32)) != hipSuccess) {
    (void)hipGetLastError();
    return 8;
  }
  return kWaveKMax;
}

// The segmented kernel of (G, K): (16, 1..4), (32, 1) or (8, 2..3).
using SegKernel = void (*)(AlignArgs, const uint2*, const uint32_t*, const uint32_t*,
                           const uint32_t*, uint32_t*, uint32_t*);
static int seg_index(int G, int K) {
  if (G == 16 && K >= 1 && K <= 4) return K - 1;
  if (G == 32 && K == 1) return 4;
  if (G == 8 && K == 3) return 5;
  if (G == 8 && K == 2) return 6;
  throw HipError{NGIA_ERR_ARG,
                 "no segmented kernel for G=" + std::to_string(G) + ", K=" + std::to_string(K)};
}
static SegKernel seg_kernel(int G, int K) {
  static const SegKernel k[7] = {k_align_seg<16, 1>, k_align_seg<16, 2>, k_align_seg<16, 3>,
                                 k_align_seg<16, 4>, k_align_seg<32, 1>, k_align_seg<8, 3>,
                                 k_align_seg<8, 2>};
  return k[seg_index(G, K)];
}

uint32_t seg_blocks_per_cu(int G, int K) {
  // Resident single-wave blocks per CU as registers and LDS allow (8 per
  // SIMD at 60 VGPRs and 4 KiB): twice the 16 that other bins get, which
  // took 25% off a bin of 2.4M pairs at windows of 17-31 words.
  static int per_cu[7] = {0, 0, 0, 0, 0, 0, 0};
  int& v = per_cu[seg_index(G, K)];
  if (!v) {
    int nb = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, seg_kernel(G, K), 64, 0);
    v = (e == hipSuccess && nb > 0) ? std::min(nb, 32) : 16;
  }
  return (uint32_t)v;
}

void launch_align_seg(int G, int K, const AlignArgs& a, const uint2* pairs, const uint32_t* list,
                      const uint32_t* off_ptr, const uint32_t* count_ptr, uint32_t* work,
                      uint32_t* score, uint32_t blocks, hipStream_t stream) {
  require_lengths(a);
  hipLaunchKernelGGL(seg_kernel(G, K), dim3(blocks), dim3(64), 0, stream, a, pairs, list, off_ptr,
                     count_ptr, work, score);
  NGIA_HIP(hipGetLastError());
}

#define NGIA_WAVE_CASE(Kv)                                                                   \
  case Kv:                                                                                   \
    hipLaunchKernelGGL(k_align_wave<Kv>, dim3(blocks), dim3(64), wave_lds_bytes(Kv), stream, \
                       a, pairs, list, off_ptr, count_ptr, work, score);                     \
    break;

void launch_align_wave(int K, const AlignArgs& a, const uint2* pairs, const uint32_t* list,
                       const uint32_t* off_ptr, const uint32_t* count_ptr, uint32_t* work, uint32_t* score,
                       uint32_t blocks, hipStream_t stream) {
  require_lengths(a);
  switch (K) {
    NGIA_WAVE_CASE(1)
    NGIA_WAVE_CASE(2)
    NGIA_WAVE_CASE(4)
    NGIA_WAVE_CASE(8)
    NGIA_WAVE_CASE(16)
    NGIA_WAVE_CASE(32)
    default:
      throw HipError{NGIA_ERR_ARG, "no wave kernel for K=" + std::to_string(K)};
  }
  NGIA_HIP(hipGetLastError());
}

}  // namespace ngia
