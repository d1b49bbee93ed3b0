// dev_graph.hip -- candidate pairs, pair binning and edge emission for the
// graph build (buildGraph.hpp:256-363) and fast mode (fastClustering.hpp:249-352).
//
// Every kernel here has a grid that depends only on host-known sizes (n,
// bands, a persistent grid) and reads data-dependent counts from device
// memory, so a whole round can be captured in a hipGraph.  Counters are
// cleared with a fill kernel, never with hipMemsetAsync.
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "dev_common.hpp"
#include "dev_graph.hpp"

namespace ngia {

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t count) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t count, hipStream_t s) {
  if (count == 0) return;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, count);
  NGIA_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ scan
// Exclusive scan of n uint32 into n+1 uint64 (out[n] = total), three passes
// of 4096-element tiles.
constexpr uint32_t kScanTile = 4096;

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                     uint64_t* __restrict__ sums) {
  __shared__ uint64_t red[256];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint64_t s = 0;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += 256)
    if (base + k < n) s += in[base + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

// Single block of four wavefronts (a block that needs a whole CU's wave
// slots waits behind any concurrent kernel that fills the CUs): exclusive
// scan of nb block sums in chunks of 256.
__global__ __launch_bounds__(256) void k_scan_sums(uint64_t* __restrict__ sums, uint32_t nb) {
  __shared__ uint64_t sh[256];
  uint64_t carry = 0;
  for (uint32_t c0 = 0; c0 < nb; c0 += 256) {
    const uint32_t i = c0 + threadIdx.x;
    const uint64_t v = i < nb ? sums[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
      const uint64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + sh[threadIdx.x] - v;
    const uint64_t tot = sh[255];
    __syncthreads();
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n,
                                                    const uint64_t* __restrict__ sums,
                                                    uint64_t* __restrict__ out, uint32_t nb) {
  __shared__ uint64_t sh[256];
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  // each thread owns 16 consecutive elements
  const uint64_t mine = base + (uint64_t)threadIdx.x * 16;
  uint32_t v[16];
  uint64_t s = 0;
  #pragma unroll
  for (int k = 0; k < 16; k++) {
    v[k] = (mine + k < n) ? in[mine + k] : 0u;
    s += v[k];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint64_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint64_t run = sums[blockIdx.x] + sh[threadIdx.x] - s;
  #pragma unroll
  for (int k = 0; k < 16; k++) {
    if (mine + k < n) out[mine + k] = run;
    run += v[k];
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = sums[nb];
}

uint64_t scan_scratch_words(uint32_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

void scan_u32(const uint32_t* in, uint32_t n, uint64_t* out, uint64_t* sums, hipStream_t s) {
  const uint32_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) {
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, 0u);
    // out[0] = 0 via a one-element copy of sums[0] (= 0 after k_scan_sums)
    NGIA_HIP(hipMemcpyAsync(out, sums, 8, hipMemcpyDeviceToDevice, s));
    return;
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, in, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, in, n, sums, out, nb);
  NGIA_HIP(hipGetLastError());
}

// ------------------------------------------------------------- predicate
// Candidate predicate of buildGraph.hpp:262-266: rep[i] != i and
// (float)len[i] >= (float)len[rep[i]] * thr.
__device__ inline bool is_job(uint32_t r, uint32_t i, const uint32_t* len, float thr) {
  return r != i && (float)len[i] >= (float)len[r] * thr;
}

// ------------------------------------------------------------- plain form
// Per band: flags, scan, then job triplets in ascending i (buildGraph.hpp:278-321).
__global__ void k_band_flags(const uint32_t* __restrict__ rep, const uint32_t* __restrict__ len,
                             uint32_t n, float thr, const uint32_t* __restrict__ parent,
                             uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t r = rep[i];
  // fast mode (fastClustering.hpp:256-258) has no length filter, only "unclustered"
  flags[i] = parent ? (uint32_t)(r != i && parent[i] == 0xFFFFFFFFu) : (uint32_t)is_job(r, i, len, thr);
}

__global__ void k_band_jobs(const uint32_t* __restrict__ rep, const uint32_t* __restrict__ flags,
                            const uint64_t* __restrict__ pos, uint32_t n,
                            uint32_t* __restrict__ jobs) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const uint64_t j = pos[i];
  jobs[3 * j] = rep[i];
  jobs[3 * j + 1] = i;
  jobs[3 * j + 2] = 0;
}

void band_jobs(const uint32_t* rep, const uint32_t* len, uint32_t n, float thr,
               const uint32_t* parent, uint32_t* flags, uint64_t* pos, uint64_t* sums,
               uint32_t* jobs, hipStream_t s) {
  hipLaunchKernelGGL(k_band_flags, dim3(grid_for(n, 256)), dim3(256), 0, s, rep, len, n, thr,
                     parent, flags);
  NGIA_HIP(hipGetLastError());
  scan_u32(flags, n, pos, sums, s);
  hipLaunchKernelGGL(k_band_jobs, dim3(grid_for(n, 256)), dim3(256), 0, s, rep, flags, pos, n, jobs);
  NGIA_HIP(hipGetLastError());
}

// Slot of this lane in an append-only list: one atomic per wavefront for the
// lanes with `pred` set; returns ~0ull for lanes without it.
__device__ inline unsigned long long wave_append(unsigned long long* counter, bool pred) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(pred);
  if (!m) return ~0ull;
  const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__builtin_popcountll(m));
  base = __shfl(base, (int)leader);
  const uint32_t rank = (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
  return pred ? base + rank : ~0ull;
}

// Append the passing jobs of one band to the raw edge list.
__global__ void k_jobs_to_edges(const uint32_t* __restrict__ jobs, uint32_t count,
                                uint64_t* __restrict__ ekeys, uint32_t* __restrict__ escores,
                                unsigned long long* __restrict__ ecount) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t sc = j < count ? jobs[3 * j + 2] : 0u;
  const unsigned long long at = wave_append(ecount, sc != 0);
  if (sc == 0) return;
  ekeys[at] = ((uint64_t)jobs[3 * j] << 32) | jobs[3 * j + 1];
  escores[at] = sc;
}

void jobs_to_edges(const uint32_t* jobs, uint32_t count, uint64_t* ekeys, uint32_t* escores,
                   unsigned long long* ecount, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_jobs_to_edges, dim3(grid_for(count, 256)), dim3(256), 0, s, jobs, count,
                     ekeys, escores, ecount);
  NGIA_HIP(hipGetLastError());
}

// Fast mode: parent[tar] = rep for each passing job of the band.
__global__ void k_fast_apply(const uint32_t* __restrict__ jobs, uint32_t count,
                             uint32_t* __restrict__ parent) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  if (jobs[3 * j + 2] > 0) parent[jobs[3 * j + 1]] = jobs[3 * j];
}

void fast_apply(const uint32_t* jobs, uint32_t count, uint32_t* parent, hipStream_t s) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_fast_apply, dim3(grid_for(count, 256)), dim3(256), 0, s, jobs, count,
                     parent);
  NGIA_HIP(hipGetLastError());
}

// ------------------------------------------------------------- batched form
// Count pass: a block takes 64 consecutive targets.  The 64 x 64 tile
// rep[b][i0..i0+63] is read row by row (coalesced) into LDS; each wavefront
// handles 16 targets, lane b holding rep[b][i].  The lengths of all 16
// targets and their representatives are gathered before the first is tested,
// so the wave waits on memory once, not 16 times.  The distinct
// representatives passing the candidate predicate are kept in order of their
// first band: keep[i] has bit b set for the band that contributes each kept
// representative, cnt[i] = its popcount, and *total gains the candidate
// count with duplicates.
__device__ inline void load_rep_tile(const uint32_t* __restrict__ rep, uint32_t n, uint32_t bands,
                                     uint32_t i0, uint32_t (&tile)[64][65]) {
  for (uint32_t e = threadIdx.x; e < 64 * 64; e += 256) {
    const uint32_t b = e >> 6, t = e & 63;
    tile[b][t] = (b < bands && i0 + t < n) ? rep[(uint64_t)b * n + i0 + t] : 0u;
  }
}

__global__ __launch_bounds__(256) void k_pair_count(const uint32_t* __restrict__ rep,
                                                    const uint32_t* __restrict__ len, uint32_t n,
                                                    uint32_t t0, uint32_t bands, float thr,
                                                    uint32_t* __restrict__ cnt,
                                                    uint64_t* __restrict__ keep_out,
                                                    unsigned long long* __restrict__ total) {
  __shared__ uint32_t tile[64][65];
  __shared__ unsigned long long btotal;
  const uint32_t i0 = blockIdx.x * 64;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (threadIdx.x == 0) btotal = 0;
  load_rep_tile(rep, n, bands, i0, tile);
  __syncthreads();
  uint32_t lr[16], li[16];
  uint32_t need = 0;  // bit tt: lane's band has a representative other than the target
  #pragma unroll
  for (int tt = 0; tt < 16; tt++) {
    const uint32_t t = wave * 16 + tt, i = i0 + t;
    const uint32_t r = tile[lane][t];
    const bool nd = lane < bands && i < n && i >= t0 && r != i;
    lr[tt] = len[nd ? r : 0u];  // unconditional: rows hold indices < n
    li[tt] = len[i < n ? i : 0u];
    need |= (uint32_t)nd << tt;
  }
  uint32_t wtotal = 0;
  #pragma unroll
  for (int tt = 0; tt < 16; tt++) {
    const uint32_t t = wave * 16 + tt;
    const uint32_t i = i0 + t;
    if (i >= n) break;
    const uint32_t r = tile[lane][t];
    // is_job: r != i and (float)len[i] >= (float)len[r] * thr
    const bool valid = ((need >> tt) & 1u) && (float)li[tt] >= (float)lr[tt] * thr;
    const uint64_t vm = __builtin_amdgcn_ballot_w64(valid);
    uint64_t mask = vm, keep = 0;
    while (mask) {
      const uint32_t b0 = (uint32_t)__builtin_ctzll(mask);
      const uint32_t r0 = __builtin_amdgcn_readlane(r, b0);
      keep |= 1ull << b0;
      mask &= ~__builtin_amdgcn_ballot_w64(valid && r == r0);
    }
    if (lane == 0) {
      cnt[i] = (uint32_t)__builtin_popcountll(keep);
      keep_out[i] = keep;
    }
    wtotal += (uint32_t)__builtin_popcountll(vm);
  }
  if (lane == 0 && wtotal) atomicAdd(&btotal, (unsigned long long)wtotal);
  __syncthreads();
  if (threadIdx.x == 0 && btotal) atomicAdd(total, btotal);
}

// Emit pass: the same tile; target i's kept representatives, in band order,
// go to pairs[off[i] ..].
__global__ __launch_bounds__(256) void k_pair_emit(const uint32_t* __restrict__ rep, uint32_t n,
                                                   uint32_t bands,
                                                   const uint64_t* __restrict__ keep_in,
                                                   const uint64_t* __restrict__ off,
                                                   uint2* __restrict__ pairs) {
  __shared__ uint32_t tile[64][65];
  const uint32_t i0 = blockIdx.x * 64;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  load_rep_tile(rep, n, bands, i0, tile);
  __syncthreads();
  #pragma unroll
  for (int tt = 0; tt < 16; tt++) {
    const uint32_t t = wave * 16 + tt;
    const uint32_t i = i0 + t;
    if (i >= n) break;
    const uint64_t keep = keep_in[i];
    if ((keep >> lane) & 1ull) {
      const uint32_t rank = (uint32_t)__builtin_popcountll(keep & ((1ull << lane) - 1ull));
      pairs[off[i] + rank] = make_uint2(tile[lane][t], i);
    }
  }
}

void pair_count(const uint32_t* rep, const uint32_t* len, uint32_t n, uint32_t t0, uint32_t bands,
                float thr, uint32_t* cnt, uint64_t* keep, unsigned long long* total, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_count, dim3((n + 63) / 64), dim3(256), 0, s, rep, len, n, t0, bands,
                     thr, cnt, keep, total);
  NGIA_HIP(hipGetLastError());
}

void pair_emit(const uint32_t* rep, uint32_t n, uint32_t bands, const uint64_t* keep,
               const uint64_t* off, uint2* pairs, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_emit, dim3((n + 63) / 64), dim3(256), 0, s, rep, n, bands, keep, off,
                     pairs);
  NGIA_HIP(hipGetLastError());
}

// Bin of a pair by its band window: W = a + c words.  Bins 2..kWinMax go to
// k_align_win<W>; windows of up to 60 words to the segmented bins (8 lanes
// per pair with 3 words per lane up to 21 words, 16 lanes with 2 at 22-30,
// 32 lanes with one at 31, 16 lanes with 3 or 4 at 32-60);
// larger windows to k_align_wave<K> with the smallest K in {1,2,4,...,32}
// with 64K >= W + K - 1 (the wave trails the band bottom in steps of K
// words); T == 0 scores 0 at once (its LCS is 0, so the pass test fails);
// anything wider goes to the plain kernel.
__device__ inline uint32_t pair_bin(uint32_t R, uint32_t T, float thr, uint32_t kmax,
                                    uint64_t* tiles) {
  if (length_rejects(R, T, thr) || T == 0) {
    *tiles = 0;
    return kBinZero;
  }
  const Band bd = band_edges(R, T, thr);
  const uint32_t a = bd.lsft / 32, c = (bd.rsft + 31) / 32, W = a + c;
  *tiles = ref_tiles(R, T, thr, (T + 31) / 32);
  if (W <= (uint32_t)kWinMax) return W;  // W >= 2 whenever T > 0
  if (W <= 21) return kBinSeg8x3;
  if (W <= 30) return kBinSeg16x2;
  if (W <= 31) return kBinSeg32;
  if (W <= 45) return kBinSeg16x3;
  if (W <= 60) return kBinSeg16x4;
  for (uint32_t k = 0, K = 1; K <= kmax; k++, K <<= 1)
    if (W + K - 1 <= 64u * (uint32_t)K) return kBinWave0 + k;
  return kBinPlain;
}

__global__ void k_bin_count(const uint2* __restrict__ pairs, const uint64_t* __restrict__ npairs,
                            const uint32_t* __restrict__ len, float thr, uint32_t kmax,
                            uint32_t* __restrict__ bin_of, uint32_t* __restrict__ bin_cnt,
                            unsigned long long* __restrict__ bin_tiles,
                            uint32_t* __restrict__ score) {
  const uint64_t P = *npairs;
  __shared__ uint32_t local[kBins];
  __shared__ unsigned long long ltiles[kBins];
  if (threadIdx.x < kBins) {
    local[threadIdx.x] = 0;
    ltiles[threadIdx.x] = 0;
  }
  __syncthreads();
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 pr = pairs[p];
    uint64_t t;
    const uint32_t b = pair_bin(len[pr.x], len[pr.y], thr, kmax, &t);
    bin_of[p] = b;
    if (b == kBinZero) score[p] = 0;
    atomicAdd(&local[b], 1u);
    if (t) atomicAdd(&ltiles[b], (unsigned long long)t);
  }
  __syncthreads();
  if (threadIdx.x < kBins && local[threadIdx.x]) {
    atomicAdd(&bin_cnt[threadIdx.x], local[threadIdx.x]);
    atomicAdd(&bin_tiles[threadIdx.x], ltiles[threadIdx.x]);
  }
}

// Route window bins by the tuned shapes (remap[b] = the bin that aligns bin
// b's pairs, counts and tiles moved along), then bin_off[b] = exclusive prefix
// of bin_cnt; also resets the scatter cursors.
__global__ void k_bin_offsets(uint32_t* __restrict__ bin_cnt, uint32_t* __restrict__ bin_off,
                              uint32_t* __restrict__ cursor, uint32_t* __restrict__ remap,
                              unsigned long long* __restrict__ bin_tiles, RouteTable route,
                              uint32_t seg2_min) {
  if (threadIdx.x == 0) {
    for (int b = 0; b < kBins; b++) remap[b] = (uint32_t)b;
    for (int W = 2; W <= kWinMax; W++) {
      if (!bin_cnt[W]) continue;
      const uint32_t to = route_window_bin(route, (uint32_t)W, bin_cnt[W]);
      if (to == (uint32_t)W) continue;
      remap[W] = to;
      bin_cnt[to] += bin_cnt[W];
      bin_tiles[to] += bin_tiles[W];
      bin_cnt[W] = 0;
      bin_tiles[W] = 0;
    }
    // Windows of 17-30 words: several words per lane (four or eight pairs
    // per wavefront) when the bin has the pairs to fill the device that way;
    // a smaller bin (long pairs, every segment in flight at once either way)
    // keeps one word per lane, whose per-character chain is shorter.
    // Likewise windows of 32-60 words: three or four words per lane, or a
    // wavefront per pair.
    // The tuned table decides by the bin's size class, or leaves it to the
    // size limit seg2_min (NGIA_LONG_DEFAULT).
    auto fold = [&](int from, int to, int group) {
      const uint32_t n2 = bin_cnt[from];
      const uint32_t choice = route.lng[group][route_class(route, n2)];
      if (n2 && (choice == NGIA_LONG_SINGLE || (choice == NGIA_LONG_DEFAULT && n2 < seg2_min))) {
        remap[from] = (uint32_t)to;
        bin_cnt[to] += n2;
        bin_tiles[to] += bin_tiles[from];
        bin_cnt[from] = 0;
        bin_tiles[from] = 0;
      }
    };
    fold(kBinSeg8x3, kBinSeg32, 0);
    fold(kBinSeg16x2, kBinSeg32, 1);
    fold(kBinSeg16x3, kBinWave0, 2);
    fold(kBinSeg16x4, kBinWave0, 3);
    uint32_t s = 0;
    for (int b = 0; b < kBins; b++) {
      bin_off[b] = s;
      cursor[b] = 0;
      s += bin_cnt[b];
    }
    bin_off[kBins] = s;
  }
}

// Scatter pair indices into their bin's segment.  A block takes tiles of
// 4096 pairs: slots are counted in LDS, reserved with one global atomic per
// (tile, bin), then written.
__global__ __launch_bounds__(256) void k_bin_scatter(const uint64_t* __restrict__ npairs,
                                                     const uint32_t* __restrict__ bin_of,
                                                     const uint32_t* __restrict__ remap,
                                                     const uint32_t* __restrict__ bin_off,
                                                     uint32_t* __restrict__ cursor,
                                                     uint32_t* __restrict__ list) {
  __shared__ uint32_t lcnt[kBins], lbase[kBins];
  const uint64_t P = *npairs;
  for (uint64_t base = (uint64_t)blockIdx.x * 4096; base < P; base += (uint64_t)gridDim.x * 4096) {
    if (threadIdx.x < kBins) lcnt[threadIdx.x] = 0;
    __syncthreads();
    uint32_t b[16], pos[16];
    #pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint64_t p = base + (uint64_t)k * 256 + threadIdx.x;
      b[k] = p < P ? remap[bin_of[p]] : 0u;
      pos[k] = p < P ? atomicAdd(&lcnt[b[k]], 1u) : 0u;
    }
    __syncthreads();
    if (threadIdx.x < kBins)
      lbase[threadIdx.x] = lcnt[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], lcnt[threadIdx.x]) : 0u;
    __syncthreads();
    #pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint64_t p = base + (uint64_t)k * 256 + threadIdx.x;
      if (p < P) list[bin_off[b[k]] + lbase[b[k]] + pos[k]] = (uint32_t)p;
    }
    __syncthreads();
  }
}

void bin_pairs(const uint2* pairs, const uint64_t* npairs, const uint32_t* len, float thr,
               uint32_t* bin_of, uint32_t* bin_cnt, uint32_t* bin_off, uint32_t* cursor,
               uint32_t* list, unsigned long long* bin_tiles, uint32_t* score, uint32_t kmax,
               uint32_t* remap, const RouteTable& route, uint32_t seg2_min, uint32_t grid,
               hipStream_t s) {
  // A sixteenth of the grid: every block ends with one global atomic per
  // nonempty bin, and a few bins take nearly all pairs, so those counters
  // see one atomic per block.
  hipLaunchKernelGGL(k_bin_count, dim3(std::max(grid / 16, 1u)), dim3(256), 0, s, pairs, npairs, len,
                     thr, kmax, bin_of, bin_cnt, bin_tiles, score);
  hipLaunchKernelGGL(k_bin_offsets, dim3(1), dim3(64), 0, s, bin_cnt, bin_off, cursor, remap,
                     bin_tiles, route, seg2_min);
  hipLaunchKernelGGL(k_bin_scatter, dim3(grid), dim3(256), 0, s, npairs, bin_of, remap, bin_off,
                     cursor, list);
  NGIA_HIP(hipGetLastError());
}

// Pairs of the plain-kernel bin as job triplets, scores copied back after.
__global__ void k_bin_to_jobs(const uint2* __restrict__ pairs, const uint32_t* __restrict__ list,
                              const uint32_t* __restrict__ bin_off, const uint32_t* __restrict__ bin_cnt,
                              uint32_t* __restrict__ jobs) {
  const uint32_t cnt = bin_cnt[kBinPlain], off = bin_off[kBinPlain];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x) {
    const uint2 pr = pairs[list[off + j]];
    jobs[3 * j] = pr.x;
    jobs[3 * j + 1] = pr.y;
    jobs[3 * j + 2] = 0;
  }
}
__global__ void k_jobs_to_scores(const uint32_t* __restrict__ list, const uint32_t* __restrict__ bin_off,
                                 const uint32_t* __restrict__ bin_cnt, const uint32_t* __restrict__ jobs,
                                 uint32_t* __restrict__ score) {
  const uint32_t cnt = bin_cnt[kBinPlain], off = bin_off[kBinPlain];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < cnt; j += gridDim.x * blockDim.x)
    score[list[off + j]] = jobs[3 * j + 2];
}
void plain_bin_jobs(const uint2* pairs, const uint32_t* list, const uint32_t* bin_off,
                    const uint32_t* bin_cnt, uint32_t* jobs, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_to_jobs, dim3(64), dim3(256), 0, s, pairs, list, bin_off, bin_cnt, jobs);
  NGIA_HIP(hipGetLastError());
}
void plain_bin_scores(const uint32_t* list, const uint32_t* bin_off, const uint32_t* bin_cnt,
                      const uint32_t* jobs, uint32_t* score, hipStream_t s) {
  hipLaunchKernelGGL(k_jobs_to_scores, dim3(64), dim3(256), 0, s, list, bin_off, bin_cnt, jobs,
                     score);
  NGIA_HIP(hipGetLastError());
}

// Passing pairs -> raw edge list.  A block takes tiles of 4096 pairs (16 per
// thread), ranks its passing pairs with a block prefix sum and reserves their
// slots with one global atomic per tile: passing pairs are a few percent, and
// one atomic per wavefront on the shared counter serialised the kernel.  The
// list's order is free (the sort + unique follows).
__global__ __launch_bounds__(256) void k_pairs_to_edges(const uint2* __restrict__ pairs,
                                                        const uint64_t* __restrict__ npairs,
                                                        const uint32_t* __restrict__ score,
                                                        uint64_t* __restrict__ ekeys,
                                                        uint32_t* __restrict__ escores,
                                                        unsigned long long* __restrict__ ecount) {
  __shared__ uint32_t wsum[4];
  __shared__ unsigned long long bbase;
  const uint64_t P = *npairs;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (uint64_t base = (uint64_t)blockIdx.x * 4096; base < P; base += (uint64_t)gridDim.x * 4096) {
    uint32_t sc[16], mine = 0;
    #pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint64_t p = base + (uint64_t)k * 256 + tid;
      sc[k] = p < P ? score[p] : 0u;
      mine += sc[k] != 0u;
    }
    uint32_t x = mine;  // inclusive prefix over the wavefront
    #pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    if (tid == 0) {
      uint32_t t = 0;
      for (int w = 0; w < 4; w++) {
        const uint32_t v = wsum[w];
        wsum[w] = t;
        t += v;
      }
      bbase = t ? atomicAdd(ecount, (unsigned long long)t) : 0ull;
    }
    __syncthreads();
    unsigned long long at = bbase + wsum[wave] + x - mine;
    #pragma unroll
    for (int k = 0; k < 16; k++) {
      if (sc[k] == 0u) continue;
      const uint2 pr = pairs[base + (uint64_t)k * 256 + tid];
      ekeys[at] = ((uint64_t)pr.x << 32) | pr.y;
      escores[at] = sc[k];
      at++;
    }
    __syncthreads();  // wsum and bbase are rewritten by the next tile
  }
}

void pairs_to_edges(const uint2* pairs, const uint64_t* npairs, const uint32_t* score,
                    uint64_t* ekeys, uint32_t* escores, unsigned long long* ecount, uint32_t grid,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_pairs_to_edges, dim3(grid), dim3(256), 0, s, pairs, npairs, score, ekeys,
                     escores, ecount);
  NGIA_HIP(hipGetLastError());
}

// ------------------------------------------------------------- finalize
// Sort the raw edges by (u,v) and drop duplicates (buildGraph.hpp:355-362):
// the score is a function of (u,v), so unique by key equals the reference's
// unique over (u,v,w) tuples.
size_t finalize_scratch_bytes(uint64_t m) {
  size_t a = 0, b = 0;
  (void)rocprim::radix_sort_pairs(nullptr, a, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                  (uint32_t*)nullptr, (uint32_t*)nullptr, (size_t)m, 0, 64);
  (void)rocprim::unique_by_key(nullptr, b, (uint64_t*)nullptr, (uint32_t*)nullptr,
                               (uint64_t*)nullptr, (uint32_t*)nullptr,
                               (unsigned long long*)nullptr, (size_t)m);
  return (a > b ? a : b) + 256;
}

void finalize_edges(uint64_t* keys, uint32_t* scores, uint64_t m, uint32_t n, uint64_t* keys_tmp,
                    uint32_t* scores_tmp, void* temp, size_t temp_bytes,
                    unsigned long long* d_unique, hipStream_t s) {
  if (m == 0) {
    fill_u32((uint32_t*)d_unique, 0, 2, s);
    return;
  }
  // bits of the largest record index: the keys' bits from 32 + ib on are 0
  const uint32_t ib = n > 1 ? 32u - (uint32_t)__builtin_clz(n - 1) : 1u;
  size_t tb = temp_bytes;
  NGIA_HIP(rocprim::radix_sort_pairs(temp, tb, keys, keys_tmp, scores, scores_tmp, (size_t)m, 0,
                                     32 + ib, s));
  tb = temp_bytes;
  NGIA_HIP(rocprim::unique_by_key(temp, tb, keys_tmp, scores_tmp, keys, scores, d_unique, (size_t)m,
                                  rocprim::equal_to<uint64_t>(), s));
}

}  // namespace ngia
