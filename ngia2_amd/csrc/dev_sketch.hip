// dev_sketch.hip -- minhash sketch (makedb.hpp:224-247) and band bucketing
// (makedb.hpp:289-364) on the device.
//
// Sketch mapping: one wavefront per sequence, lane k computes signature k.
// The 8-mer is the same for every lane, so it is built on the scalar unit
// from the sequence's byte codes (scalar loads, four codes per dword); the
// 64 hash functions run one per lane.  The first two steps of the hash fold:
// with h = kmer ^ seed_k,  (h ^ 61) ^ (h >> 16) =
//   (kmer ^ (kmer >> 16)) ^ (seed_k ^ 61 ^ (seed_k >> 16)),
// a wave-uniform term xor a per-lane constant, because a logical right shift
// distributes over xor.  Output is sequence-major: lane k of sequence i
// writes sig[64*i + k], one 256-byte coalesced store per wavefront.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "dev_common.hpp"
#include "dev_graph.hpp"

namespace ngia {

// Hashes the windows ending at positions [j, end) of a record whose byte
// codes start at cw (four per dword), the k-mer of the positions before j in
// kmer; mn keeps the minimum.  Head to a 16-aligned position, then sixteen
// positions per four-dword scalar load: one wait on the scalar cache per
// sixteen windows.  The load reads whole 16-byte groups; a record's codes
// are padded to 32 bytes, so it stays within the record's own padding.
__device__ __forceinline__ void sketch_windows(const uint32_t* cw, uint32_t j, uint32_t end,
                                               uint32_t kmer, uint32_t s1, uint32_t& mn) {
  auto hash = [&](uint32_t k1) {
    uint32_t h = k1 ^ s1;
    h = h + (h << 3);
    h = h ^ (h >> 4);
    h = h * 0x27d4eb2du;
    h = h ^ (h >> 15);
    h = (h ^ (h >> 8)) * 0x9e3779b1u;
    h = h ^ (h >> 14);
    mn = h < mn ? h : mn;
  };
  for (; j < end && (j & 15); j++) {
    const uint32_t w = __builtin_amdgcn_readfirstlane(cw[j >> 2]);
    kmer = (kmer << 4) + ((w >> (8 * (j & 3))) & 0xFu);
    hash(kmer ^ (kmer >> 16));
  }
  for (; j + 16 <= end; j += 16) {
    const uint4 c4 = *(const uint4*)(cw + (j >> 2));
    const uint32_t ws[4] = {(uint32_t)__builtin_amdgcn_readfirstlane(c4.x),
                            (uint32_t)__builtin_amdgcn_readfirstlane(c4.y),
                            (uint32_t)__builtin_amdgcn_readfirstlane(c4.z),
                            (uint32_t)__builtin_amdgcn_readfirstlane(c4.w)};
    #pragma unroll
    for (int q = 0; q < 4; q++) {
      #pragma unroll
      for (int m = 0; m < 4; m++) {
        kmer = (kmer << 4) + ((ws[q] >> (8 * m)) & 0xFu);
        hash(kmer ^ (kmer >> 16));
      }
    }
  }
  for (; j + 4 <= end; j += 4) {
    const uint32_t w = __builtin_amdgcn_readfirstlane(cw[j >> 2]);
    #pragma unroll
    for (int m = 0; m < 4; m++) {
      kmer = (kmer << 4) + ((w >> (8 * m)) & 0xFu);
      hash(kmer ^ (kmer >> 16));
    }
  }
  for (; j < end; j++) {
    const uint32_t w = __builtin_amdgcn_readfirstlane(cw[j >> 2]);
    kmer = (kmer << 4) + ((w >> (8 * (j & 3))) & 0xFu);
    hash(kmer ^ (kmer >> 16));
  }
}

// The codes of record `seq` (the 64-bit offset moves to scalar registers as
// two halves: code areas pass 2^31 bytes at a few million sequences).
__device__ __forceinline__ const uint32_t* record_codes(const uint8_t* codes,
                                                        const uint64_t* code_off, uint32_t seq) {
  const uint64_t co = code_off[seq];
  const uint32_t co_lo = __builtin_amdgcn_readfirstlane((uint32_t)co);
  const uint32_t co_hi = __builtin_amdgcn_readfirstlane((uint32_t)(co >> 32));
  return (const uint32_t*)(codes + (((uint64_t)co_hi << 32) | co_lo));
}

// Hashing starts at position 7, or at the last position of a read shorter
// than 8 (makedb.hpp:234-235).
__device__ __forceinline__ uint32_t first_window(uint32_t L) { return L >= 8 ? 7u : (L ? L - 1 : 0u); }

// Lane k computes signature k: the per-lane half of the folded first two
// hash steps, seed_k ^ 61 ^ (seed_k >> 16).
__device__ __forceinline__ uint32_t lane_seed_term(uint32_t lane) {
  const uint32_t seed = lane * 0x9e3779b1u + 0x85ebca6bu;
  return seed ^ 61u ^ (seed >> 16);
}

__global__ __launch_bounds__(256) void k_sketch(const uint8_t* __restrict__ codes,
                                                const uint64_t* __restrict__ code_off,
                                                const uint32_t* __restrict__ lengths,
                                                uint32_t n, uint32_t* __restrict__ sig) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t seq = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (seq >= n) return;
  const uint32_t L = __builtin_amdgcn_readfirstlane(lengths[seq]);
  const uint32_t* cw = record_codes(codes, code_off, seq);
  const uint32_t first = first_window(L);
  uint32_t kmer = 0;
  for (uint32_t j = 0; j < first; j++) {  // the prefix only feeds the k-mer
    const uint32_t w = __builtin_amdgcn_readfirstlane(cw[j >> 2]);
    kmer = (kmer << 4) + ((w >> (8 * (j & 3))) & 0xFu);
  }
  uint32_t mn = 0xFFFFFFFFu;
  sketch_windows(cw, first, L, kmer, lane_seed_term(lane), mn);
  sig[(uint64_t)seq * 64 + lane] = mn;
}

// Long records in pieces: one wavefront per (record, chunk of positions),
// chunks[c] = (record, first window position); a chunk ends at the next
// multiple of `len` or the record's end; the k-mer is rebuilt from the seven codes
// before the chunk (codes before position 0 count as 0, as in k_sketch).
// The minimum over a record's chunks is the same whatever their order:
// each wavefront folds its chunk's minima into sig with atomicMin, over
// signatures set to all ones first.
__global__ __launch_bounds__(256) void k_sketch_chunks(const uint8_t* __restrict__ codes,
                                                       const uint64_t* __restrict__ code_off,
                                                       const uint32_t* __restrict__ lengths,
                                                       const uint2* __restrict__ chunks,
                                                       uint32_t nchunks, uint32_t len,
                                                       uint32_t* __restrict__ sig) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t c = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  if (c >= nchunks) return;
  const uint2 ch = chunks[c];
  const uint32_t seq = __builtin_amdgcn_readfirstlane(ch.x);
  const uint32_t p0 = __builtin_amdgcn_readfirstlane(ch.y);
  const uint32_t L = __builtin_amdgcn_readfirstlane(lengths[seq]);
  const uint32_t* cw = record_codes(codes, code_off, seq);
  const uint32_t stop = p0 - p0 % len + len;
  const uint32_t end = stop < L ? stop : L;
  uint32_t kmer = 0;
  for (uint32_t j = p0 >= 7 ? p0 - 7 : 0u; j < p0; j++) {
    const uint32_t w = __builtin_amdgcn_readfirstlane(cw[j >> 2]);
    kmer = (kmer << 4) + ((w >> (8 * (j & 3))) & 0xFu);
  }
  uint32_t mn = 0xFFFFFFFFu;
  sketch_windows(cw, p0, end, kmer, lane_seed_term(lane), mn);
  atomicMin(&sig[(uint64_t)seq * 64 + lane], mn);
}

void launch_sketch_chunks(const uint8_t* codes, const uint64_t* code_off, const uint32_t* lengths,
                          const uint2* chunks, uint32_t nchunks, uint32_t len, uint32_t* sig,
                          hipStream_t stream) {
  if (nchunks == 0) return;
  hipLaunchKernelGGL(k_sketch_chunks, dim3((nchunks + 3) / 4), dim3(256), 0, stream, codes,
                     code_off, lengths, chunks, nchunks, len, sig);
  NGIA_HIP(hipGetLastError());
}

void launch_sketch(const uint8_t* codes, const uint64_t* code_off, const uint32_t* lengths,
                   uint32_t n, uint32_t* sig, hipStream_t stream) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_sketch, dim3((n + 3) / 4), dim3(256), 0, stream, codes, code_off, lengths,
                     n, sig);
  NGIA_HIP(hipGetLastError());
}

// ------------------------------------------------------------- bucketing
// Per band the reference stable-sorts (sig, i) by sig (makedb.hpp:320-339)
// and points every element at its predecessor in the equal-signature run
// (makedb.hpp:340-347).  Here all 64 bands are sorted at once by a stable
// radix sort on the 32-bit signature alone, the value carrying (band, i)
// with the band in the top bits: the input is band-major with i ascending,
// so after the stable sort every equal-signature run lists its elements by
// band, then by i -- each band's run is a contiguous, i-ascending slice, and
// an element's predecessor is the previous element when that one has the
// same signature and the same band.  Rows start as the identity; only the
// non-head elements are scattered.  Values need n < 2^26; larger n uses the
// (band << 32 | sig) 38-bit key instead.
constexpr uint32_t kIdxBits = 26;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;

__global__ __launch_bounds__(256) void k_bucket_keys(const uint32_t* __restrict__ sig, uint32_t n,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
  // 64 sequences x 64 bands tile through LDS: coalesced reads of the
  // sequence-major signatures, coalesced writes of band-major keys.
  __shared__ uint32_t tile[64][65];
  const uint32_t i0 = blockIdx.x * 64;
  const uint32_t t = threadIdx.x;
  for (uint32_t e = t; e < 64 * 64; e += 256) {
    const uint32_t si = e >> 6, b = e & 63;
    tile[si][b] = (i0 + si < n) ? sig[(uint64_t)(i0 + si) * 64 + b] : 0u;
  }
  __syncthreads();
  for (uint32_t e = t; e < 64 * 64; e += 256) {
    const uint32_t b = e >> 6, si = e & 63;
    const uint32_t i = i0 + si;
    if (i < n) {
      keys[(uint64_t)b * n + i] = tile[si][b];
      vals[(uint64_t)b * n + i] = (b << kIdxBits) | i;
    }
  }
}

__global__ __launch_bounds__(256) void k_rep_identity(uint32_t n, uint32_t* __restrict__ rep) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;  // blockIdx.y = band
  if (i < n) rep[(uint64_t)blockIdx.y * n + i] = i;
}

__global__ __launch_bounds__(256) void k_bucket_rep(const uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals,
                                                    uint64_t total, uint32_t n,
                                                    uint32_t* __restrict__ rep) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p == 0 || p >= total) return;
  const uint32_t v = vals[p], pv = vals[p - 1];
  if (keys[p] == keys[p - 1] && (v >> kIdxBits) == (pv >> kIdxBits))
    rep[(uint64_t)(v >> kIdxBits) * n + (v & kIdxMask)] = pv & kIdxMask;
}

// Wide form for n >= 2^26: keys (band << 32 | sig), values i.
__global__ __launch_bounds__(256) void k_bucket_keys_wide(const uint32_t* __restrict__ sig,
                                                          uint32_t n, uint64_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals) {
  const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (uint64_t)n * 64) return;
  const uint32_t b = (uint32_t)(e / n), i = (uint32_t)(e % n);
  keys[e] = ((uint64_t)b << 32) | sig[(uint64_t)i * 64 + b];
  vals[e] = i;
}

__global__ __launch_bounds__(256) void k_bucket_rep_wide(const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals,
                                                         uint64_t total, uint32_t n,
                                                         uint32_t* __restrict__ rep) {
  const uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (p == 0 || p >= total) return;
  const uint64_t k = keys[p];
  if (keys[p - 1] == k) rep[(k >> 32) * n + vals[p]] = vals[p - 1];
}

// ------------------------------------------------ partitioned bucketing
// Only equality of signatures within a band matters, so the full sort above
// is replaced, for n <= kPartMaxN, by a partition and a sort in LDS:
//  1. every (band b, record i) gets h = mix(sig), a bijection of the 32-bit
//     signature that spreads its bits (minhash values crowd near 0), the
//     16-bit key (b << pb | top pb bits of h) and the value (h << 32 | i);
//  2. a stable radix sort on the key's low pb bits groups each band's
//     elements into 2^pb partitions, each still in ascending i (the input is
//     band-major, so the stable sort keeps each (bucket, band) group
//     contiguous without sorting the band bits);
//  3. one block per partition sorts it by h in LDS (stable), and every
//     element equal to its predecessor gets rep[b][i] = that predecessor's i
//     -- the element before it in the reference's stable (sig, i) order.
// pb = 8 (one radix pass, partitions average n/256) up to kPart8MaxN records,
// else pb = 10 (two passes, partitions average n/1024).  A partition larger
// than a block's capacity (heavily repeated signatures) is counted; the host
// then redoes the whole bucketing with the full sort.
constexpr int kPartBlock = 512;  // threads per partition sort
constexpr uint32_t kPart8MaxN = 1000000;  // partitions average <= 3907, capacity 8192
constexpr uint32_t kPartMaxN = 2800000;   // partitions average <= 2735, capacity 4096

__device__ inline uint32_t sig_mix(uint32_t x) {
  x *= 0x9E3779B1u;  // odd multiplier: a bijection mod 2^32
  x ^= x >> 16;      // xor with a right shift of itself: a bijection
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  return x;
}

__global__ __launch_bounds__(256) void k_part_keys(const uint32_t* __restrict__ sig, uint32_t n,
                                                   uint32_t pb, uint16_t* __restrict__ keys,
                                                   uint64_t* __restrict__ vals) {
  // 64 records x 64 bands through LDS, as in k_bucket_keys
  __shared__ uint32_t tile[64][65];
  const uint32_t i0 = blockIdx.x * 64;
  const uint32_t t = threadIdx.x;
  for (uint32_t e = t; e < 64 * 64; e += 256) {
    const uint32_t si = e >> 6, b = e & 63;
    tile[si][b] = (i0 + si < n) ? sig[(uint64_t)(i0 + si) * 64 + b] : 0u;
  }
  __syncthreads();
  for (uint32_t e = t; e < 64 * 64; e += 256) {
    const uint32_t b = e >> 6, si = e & 63;
    const uint32_t i = i0 + si;
    if (i < n) {
      const uint32_t h = sig_mix(tile[si][b]);
      keys[(uint64_t)b * n + i] = (uint16_t)((b << pb) | (h >> (32 - pb)));
      vals[(uint64_t)b * n + i] = ((uint64_t)h << 32) | i;
    }
  }
}

// start[k], end[k]: the sorted range of partition k (left 0, 0 when empty).
// Eight keys per thread from one 16-byte load (m is a multiple of 64), the
// neighbours across the group boundary from the adjacent groups.
__global__ __launch_bounds__(256) void k_part_bounds(const uint16_t* __restrict__ keys, uint64_t m,
                                                     uint32_t* __restrict__ start,
                                                     uint32_t* __restrict__ end) {
  const uint64_t groups = m / 8;
  for (uint64_t g = (uint64_t)blockIdx.x * 256 + threadIdx.x; g < groups;
       g += (uint64_t)gridDim.x * 256) {
    const uint64_t p0 = 8 * g;
    const uint4 w = *(const uint4*)(keys + p0);
    uint32_t k[8] = {w.x & 0xFFFFu, w.x >> 16, w.y & 0xFFFFu, w.y >> 16,
                     w.z & 0xFFFFu, w.z >> 16, w.w & 0xFFFFu, w.w >> 16};
    // keys are 16-bit: 0x10000 differs from every key
    uint32_t prev = p0 ? keys[p0 - 1] : 0x10000u;
    const uint32_t after = p0 + 8 < m ? keys[p0 + 8] : 0x10000u;
    #pragma unroll
    for (int j = 0; j < 8; j++) {
      const uint32_t next = j < 7 ? k[j + 1] : after;
      if (prev != k[j]) start[k[j]] = (uint32_t)(p0 + j);
      if (next != k[j]) end[k[j]] = (uint32_t)(p0 + j + 1);
      prev = k[j];
    }
  }
}

// One partition sorted in LDS, stably, by the low kLowBits bits of h alone
// (ITEMS elements per thread in the blocked arrangement): elements of equal h
// then sit in one run of equal low bits, in ascending i, possibly between
// elements of other values with the same low bits (about one per run at the
// partitions' sizes, since h mixes the signature's bits).  An element's
// predecessor -- the element before it in the (h, i) order -- is the nearest
// earlier element of its run with the same h, found by a scan backwards
// through the sorted keys in LDS.  The partition's elements share the top pb
// bits of h; sorting 12 of the other 32 - pb bits takes half the passes.  With
// the scan the bucketing of a million records went from 1.44 to 1.40 ms
// (log-normal lengths) and to 1.43 ms (heavy-tailed); the kernel's loads and
// scattered row writes stay as they were.
constexpr uint32_t kLowBits = 12, kLowMask = (1u << kLowBits) - 1u;
template <int BS, int ITEMS>
__device__ inline void part_sort_tier(const uint64_t* __restrict__ v, uint32_t cnt, uint64_t row,
                                      uint32_t* __restrict__ rep, void* storage_raw) {
  using Sort = rocprim::block_radix_sort<uint32_t, BS, ITEMS, uint32_t>;
  auto& storage = *reinterpret_cast<typename Sort::storage_type*>(storage_raw);
  const uint32_t t = threadIdx.x;
  uint32_t key[ITEMS], val[ITEMS];
  #pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t q = t * ITEMS + j;  // blocked: input order = ascending i
    if (q < cnt) {
      const uint64_t x = v[q];
      key[j] = (uint32_t)(x >> 32);
      val[j] = (uint32_t)x;
    } else {  // padding: the largest low bits, and after every element of them (stable)
      key[j] = 0xFFFFFFFFu;
      val[j] = 0xFFFFFFFFu;
    }
  }
  Sort().sort(key, val, storage, 0, kLowBits);
  __syncthreads();  // the sort's storage now holds the sorted keys and values
  uint32_t* sk = reinterpret_cast<uint32_t*>(storage_raw);
  uint32_t* sv = sk + BS * ITEMS;
  #pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    sk[t * ITEMS + j] = key[j];
    sv[t * ITEMS + j] = val[j];
  }
  __syncthreads();
  // The element just before is in registers (or, for the thread's first
  // item, the previous thread's last); only a run holding other values of
  // the same low bits between equal ones needs the scan through LDS.
  uint32_t pk = t ? sk[t * ITEMS - 1] : 0u, pv = t ? sv[t * ITEMS - 1] : 0u;
  #pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t q = t * ITEMS + j;
    if (q >= 1 && q < cnt && ((pk ^ key[j]) & kLowMask) == 0) {
      if (pk == key[j]) {
        rep[row + val[j]] = pv;
      } else {
        for (uint32_t r = q - 1; r-- > 0;) {
          const uint32_t kr = sk[r];
          if ((kr & kLowMask) != (key[j] & kLowMask)) break;
          if (kr == key[j]) {
            rep[row + val[j]] = sv[r];
            break;
          }
        }
      }
    }
    pk = key[j];
    pv = val[j];
  }
}

// One block of BS threads per partition, sorted at the smallest of three
// capacities (BS x I1, I2 or I3 items) that holds it: the sort's cost follows
// its capacity, and most partitions sit near the average, under the largest.
// The LDS holds the sort's storage, then the sorted keys and values.
template <int BS, int I1, int I2, int I3>
__global__ __launch_bounds__(BS) void k_part_local(const uint64_t* __restrict__ vals,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ end, uint32_t n,
                                                    uint32_t pb, uint32_t* __restrict__ rep,
                                                    uint32_t* __restrict__ overflow) {
  using S1 = rocprim::block_radix_sort<uint32_t, BS, I1, uint32_t>;
  using S2 = rocprim::block_radix_sort<uint32_t, BS, I2, uint32_t>;
  using S3 = rocprim::block_radix_sort<uint32_t, BS, I3, uint32_t>;
  constexpr size_t kStore =
      std::max(std::max(sizeof(typename S1::storage_type),
                        std::max(sizeof(typename S2::storage_type), sizeof(typename S3::storage_type))),
               (size_t)8 * BS * I3);
  __shared__ __align__(16) unsigned char storage[kStore];
  const uint32_t part = blockIdx.x;
  const uint32_t s = start[part], cnt = end[part] - s;
  if (cnt < 2) return;  // a lone element is a bucket head: its row entry stays i
  const uint64_t row = (uint64_t)(part >> pb) * n;
  if (cnt <= (uint32_t)BS * I1)
    part_sort_tier<BS, I1>(vals + s, cnt, row, rep, storage);
  else if (cnt <= (uint32_t)BS * I2)
    part_sort_tier<BS, I2>(vals + s, cnt, row, rep, storage);
  else if (cnt <= (uint32_t)BS * I3)
    part_sort_tier<BS, I3>(vals + s, cnt, row, rep, storage);
  else if (threadIdx.x == 0)
    atomicAdd(overflow, 1u);
}

// Bucketing of at most kBucketSmallMax records (the head of a step, and any
// database that small): one block per band sorts the band's n (mix(sig), i)
// pairs in LDS, stably, and points each element equal to its predecessor at
// that predecessor, as k_part_local does for a partition.  No global sort, no
// look-back between blocks and no host synchronisation, so it runs beside
// other kernels.  The sort's storage holds the threads' last items after the
// sort (8 bytes per thread, well inside it).
template <int BS, int ITEMS>
__global__ __launch_bounds__(BS) void k_bucket_small(const uint32_t* __restrict__ sig, uint32_t n,
                                                     uint32_t* __restrict__ rep) {
  using Sort = rocprim::block_radix_sort<uint32_t, BS, ITEMS, uint32_t>;
  static_assert(sizeof(typename Sort::storage_type) >= 8 * BS, "last items do not fit");
  __shared__ typename Sort::storage_type storage;
  const uint32_t b = blockIdx.x, t = threadIdx.x;
  const uint64_t row = (uint64_t)b * n;
  uint32_t key[ITEMS], val[ITEMS];
  #pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t q = t * ITEMS + j;  // blocked: input order = ascending i
    if (q < n) {
      key[j] = sig_mix(sig[(uint64_t)q * 64 + b]);
      val[j] = q;
      rep[row + q] = q;  // a bucket head keeps its own index
    } else {  // padding sorts after every element (stable)
      key[j] = 0xFFFFFFFFu;
      val[j] = 0xFFFFFFFFu;
    }
  }
  __syncthreads();
  Sort().sort(key, val, storage, 0, 32);
  __syncthreads();  // the storage is free again
  uint32_t* lastk = reinterpret_cast<uint32_t*>(&storage);
  uint32_t* lastv = lastk + BS;
  lastk[t] = key[ITEMS - 1];
  lastv[t] = val[ITEMS - 1];
  __syncthreads();
  uint32_t pk = t ? lastk[t - 1] : 0u, pv = t ? lastv[t - 1] : 0u;
  #pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t q = t * ITEMS + j;
    if (q >= 1 && q < n && key[j] == pk) rep[row + val[j]] = pv;
    pk = key[j];
    pv = val[j];
  }
}

void launch_bucket_small(const uint32_t* sig, uint32_t n, uint32_t* rep, hipStream_t stream) {
  if (n == 0) return;
  if (n > kBucketSmallMax) throw HipError{NGIA_ERR_ARG, "too many records for the small bucketing"};
  if (n <= 2048)
    hipLaunchKernelGGL((k_bucket_small<256, 8>), dim3(64), dim3(256), 0, stream, sig, n, rep);
  else if (n <= 8192)
    hipLaunchKernelGGL((k_bucket_small<1024, 8>), dim3(64), dim3(1024), 0, stream, sig, n, rep);
  else
    hipLaunchKernelGGL((k_bucket_small<1024, 16>), dim3(64), dim3(1024), 0, stream, sig, n, rep);
  NGIA_HIP(hipGetLastError());
}

namespace {
struct BucketLayout {
  bool wide;
  size_t keys_in, keys_out, vals_in, vals_out, temp, temp_bytes, total;
};
// partitioned path: 16-bit keys, 64-bit values, the partition ranges and the
// overflow count
struct PartLayout {
  bool use;
  uint32_t bits;  // partition bits pb
  int tiers;      // capacity tiers: 0 = 4096/5120/8192, 1 = 512/1024/2048, 2 = 1024/2048/4096
  size_t keys_in, keys_out, vals_in, vals_out, start, end, overflow, temp, temp_bytes, total;
};
PartLayout part_layout(uint32_t n, bool force_wide) {
  PartLayout L{};
  L.use = !force_wide && n <= kPartMaxN;
  if (!L.use) return L;
  const uint64_t m = (uint64_t)n * 64;
  // The largest capacity is 1.5 to 2 times the average partition, n >> pb.
  L.bits = n <= kPart8MaxN ? 8 : 10;
  L.tiers = n <= kPart8MaxN ? 0 : n <= 1400000u ? 1 : 2;
  size_t temp_bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, temp_bytes, (uint16_t*)nullptr, (uint16_t*)nullptr,
                                  (uint64_t*)nullptr, (uint64_t*)nullptr, (size_t)m, 0, L.bits);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t parts = (size_t)64 << L.bits;
  L.keys_in = 0;
  L.keys_out = L.keys_in + al(2 * m);
  L.vals_in = L.keys_out + al(2 * m);
  L.vals_out = L.vals_in + al(8 * m);
  L.start = L.vals_out + al(8 * m);
  L.end = L.start + al(4 * parts);
  L.overflow = L.end + al(4 * parts);
  L.temp = L.overflow + 256;
  L.temp_bytes = temp_bytes;
  L.total = L.temp + al(temp_bytes);
  return L;
}
BucketLayout bucket_layout(uint32_t n, bool force_wide) {
  const uint64_t m = (uint64_t)n * 64;
  BucketLayout L;
  L.wide = force_wide || n > kIdxMask;
  const size_t kb = L.wide ? 8 : 4;
  size_t temp_bytes = 0;
  if (L.wide)
    (void)rocprim::radix_sort_pairs(nullptr, temp_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                    (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)m, 0, 38);
  else
    (void)rocprim::radix_sort_pairs(nullptr, temp_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                    (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)m, 0, 32);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  L.keys_in = 0;
  L.keys_out = L.keys_in + al(kb * m);
  L.vals_in = L.keys_out + al(kb * m);
  L.vals_out = L.vals_in + al(4 * m);
  L.temp = L.vals_out + al(4 * m);
  L.temp_bytes = temp_bytes;
  L.total = L.temp + al(temp_bytes);
  return L;
}
}  // namespace

size_t bucket_scratch_bytes(uint32_t n, bool force_wide) {
  return std::max(bucket_layout(n, force_wide).total, part_layout(n, force_wide).total);
}

namespace {
// The partitioned bucketing; returns false when a partition overflowed (rep
// then holds a partial result and the caller redoes it with the full sort).
bool bucket_partitioned(const PartLayout& L, const uint32_t* sig, uint32_t n, uint32_t* rep,
                        char* s, hipStream_t stream, uint32_t* pinned_flag) {
  const uint64_t m = (uint64_t)n * 64;
  const uint32_t pb = L.bits, parts = 64u << pb;
  uint16_t* keys_in = (uint16_t*)(s + L.keys_in);
  uint16_t* keys_out = (uint16_t*)(s + L.keys_out);
  uint64_t* vals_in = (uint64_t*)(s + L.vals_in);
  uint64_t* vals_out = (uint64_t*)(s + L.vals_out);
  uint32_t* start = (uint32_t*)(s + L.start);
  uint32_t* end = (uint32_t*)(s + L.end);
  uint32_t* overflow = (uint32_t*)(s + L.overflow);
  hipLaunchKernelGGL(k_part_keys, dim3((n + 63) / 64), dim3(256), 0, stream, sig, n, pb, keys_in,
                     vals_in);
  NGIA_HIP(hipGetLastError());
  size_t tb = L.temp_bytes;
  NGIA_HIP(rocprim::radix_sort_pairs(s + L.temp, tb, keys_in, keys_out, vals_in, vals_out, (size_t)m,
                                     0, pb, stream));
  fill_u32(start, 0, parts, stream);
  fill_u32(end, 0, parts, stream);
  fill_u32(overflow, 0, 1, stream);
  hipLaunchKernelGGL(k_part_bounds, dim3(std::min<uint64_t>((m / 8 + 255) / 256, 16384)), dim3(256),
                     0, stream, keys_out, m, start, end);
  // capacity tiers around the average partition, n >> pb
  constexpr int B = kPartBlock;
  if (L.tiers == 0)
    hipLaunchKernelGGL((k_part_local<B, 4096 / B, 5120 / B, 8192 / B>), dim3(parts), dim3(B), 0,
                       stream, vals_out, start, end, n, pb, rep, overflow);
  else if (L.tiers == 1)
    hipLaunchKernelGGL((k_part_local<B, 512 / B, 1024 / B, 2048 / B>), dim3(parts), dim3(B), 0,
                       stream, vals_out, start, end, n, pb, rep, overflow);
  else  // 256 threads measured faster here than 512 (2M records: 3.9 against 4.2 ms)
    hipLaunchKernelGGL((k_part_local<256, 4, 8, 16>), dim3(parts), dim3(256), 0, stream, vals_out,
                       start, end, n, pb, rep, overflow);
  NGIA_HIP(hipGetLastError());
  uint32_t local = 0;
  uint32_t* over = pinned_flag ? pinned_flag : &local;
  NGIA_HIP(hipMemcpyAsync(over, overflow, 4, hipMemcpyDeviceToHost, stream));
  NGIA_HIP(hipStreamSynchronize(stream));
  return *over == 0;
}
}  // namespace

void launch_bucket(const uint32_t* sig, uint32_t n, uint32_t* rep, void* scratch,
                   hipStream_t stream, bool force_wide, bool few_values, uint32_t* pinned_flag) {
  if (n == 0) return;
  const BucketLayout L = bucket_layout(n, force_wide);
  char* s = (char*)scratch;
  const uint64_t m = (uint64_t)n * 64;
  uint32_t* vals_in = (uint32_t*)(s + L.vals_in);
  uint32_t* vals_out = (uint32_t*)(s + L.vals_out);
  size_t tb = L.temp_bytes;
  hipLaunchKernelGGL(k_rep_identity, dim3(grid_for(n, 256), 64), dim3(256), 0, stream, n, rep);
  NGIA_HIP(hipGetLastError());
  const PartLayout P = part_layout(n, force_wide);
  if (P.use && !few_values) {
    if (bucket_partitioned(P, sig, n, rep, s, stream, pinned_flag)) return;
    // a partition overflowed: start again from the identity rows, full sort
    hipLaunchKernelGGL(k_rep_identity, dim3(grid_for(n, 256), 64), dim3(256), 0, stream, n, rep);
    NGIA_HIP(hipGetLastError());
  }
  if (!L.wide) {
    uint32_t* keys_in = (uint32_t*)(s + L.keys_in);
    uint32_t* keys_out = (uint32_t*)(s + L.keys_out);
    hipLaunchKernelGGL(k_bucket_keys, dim3((n + 63) / 64), dim3(256), 0, stream, sig, n, keys_in,
                       vals_in);
    NGIA_HIP(hipGetLastError());
    NGIA_HIP(rocprim::radix_sort_pairs(s + L.temp, tb, keys_in, keys_out, vals_in, vals_out,
                                       (size_t)m, 0, 32, stream));
    hipLaunchKernelGGL(k_bucket_rep, dim3(grid_for(m, 256)), dim3(256), 0, stream, keys_out,
                       vals_out, m, n, rep);
  } else {
    uint64_t* keys_in = (uint64_t*)(s + L.keys_in);
    uint64_t* keys_out = (uint64_t*)(s + L.keys_out);
    hipLaunchKernelGGL(k_bucket_keys_wide, dim3(grid_for(m, 256)), dim3(256), 0, stream, sig, n,
                       keys_in, vals_in);
    NGIA_HIP(hipGetLastError());
    NGIA_HIP(rocprim::radix_sort_pairs(s + L.temp, tb, keys_in, keys_out, vals_in, vals_out,
                                       (size_t)m, 0, 38, stream));
    hipLaunchKernelGGL(k_bucket_rep_wide, dim3(grid_for(m, 256)), dim3(256), 0, stream, keys_out,
                       vals_out, m, n, rep);
  }
  NGIA_HIP(hipGetLastError());
}

}  // namespace ngia
