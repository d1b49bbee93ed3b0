// serve.hip -- the serving loop of include/ngia2.h: a clustered database
// kept resident on the device, streamed query batches assigned to clusters.
//
// Resident snapshot: the database's packed records, byte codes, lengths and
// signatures are copied into arrays with room for `lanes` query regions after
// the residents, so a query is just another index (n + lane*max_queries + j)
// and the length-binned alignment kernels of the batched graph run unchanged
// on resident/query pairs.  Per band, an open-addressing table maps a
// signature to its bucket head: the lowest resident index with that
// signature (built with atomicCAS + atomicMin, so the result does not depend
// on the insertion order).
//
// One batch on one lane stream: copy the packed queries in, byte codes,
// sketch, per-query candidate lookup (one wavefront per query, lane = band,
// distinct heads kept, longer sequence as rep, the isJob length test), the
// length bins, then per query the best (score, lowest index) and its
// cluster, copied back to pinned host memory.  Every kernel's work depends
// only on the query and the snapshot, never on the other queries of the
// batch, so a query's result is the same alone, in any mix and offline.
#include <chrono>
#include <deque>
#include <thread>

#include "capi_internal.hpp"

using namespace ngia;

namespace ngia {
namespace {

constexpr unsigned long long kEmpty = ~0ull;

__device__ __host__ inline uint32_t slot_of_sig(uint32_t sig, uint32_t cap_log2) {
  return (uint32_t)((sig * 0x9E3779B1u) >> (32 - cap_log2));  // Fibonacci hashing
}

// Insert (sig, i) for every resident and band; a slot keeps the lowest i of
// its signature.
__global__ void k_serve_insert(const uint32_t* __restrict__ sig, uint32_t n,
                               unsigned long long* __restrict__ tables, uint32_t cap_log2) {
  const uint64_t total = (uint64_t)n * 64;
  const uint32_t mask = (1u << cap_log2) - 1;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = (uint32_t)(t >> 6), b = (uint32_t)(t & 63);
    const uint32_t key = sig[t];
    unsigned long long* tb = tables + ((uint64_t)b << cap_log2);
    const unsigned long long val = ((unsigned long long)key << 32) | i;
    uint32_t s = slot_of_sig(key, cap_log2);
    // At most n keys per band in 2^cap_log2 >= 2n slots: an empty slot exists.
    for (;;) {
      const unsigned long long old = atomicCAS(&tb[s], kEmpty, val);
      if (old == kEmpty) break;
      if ((uint32_t)(old >> 32) == key) {
        atomicMin(&tb[s], val);
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

// One wavefront per query, lane b = band: the bucket head of the query's
// band-b signature, distinct heads kept once (lowest band), oriented and
// length-tested, appended to the lane's pair list.
__global__ __launch_bounds__(64) void k_serve_lookup(
    const uint32_t* __restrict__ qsig, uint32_t count, uint32_t base,
    const unsigned long long* __restrict__ tables, uint32_t cap_log2,
    const uint32_t* __restrict__ len, float thr, uint2* __restrict__ pairs,
    unsigned long long* __restrict__ npairs, uint32_t* __restrict__ ncand) {
  const uint32_t q = blockIdx.x;
  if (q >= count) return;
  const uint32_t lane = threadIdx.x;
  const uint32_t key = qsig[(uint64_t)q * 64 + lane];
  const unsigned long long* tb = tables + ((uint64_t)lane << cap_log2);
  const uint32_t mask = (1u << cap_log2) - 1;
  uint32_t s = slot_of_sig(key, cap_log2);
  uint32_t head = NGIA_NO_RESIDENT;
  for (uint32_t probe = 0; probe <= mask; probe++) {
    const unsigned long long v = tb[s];
    if (v == kEmpty) break;
    if ((uint32_t)(v >> 32) == key) {
      head = (uint32_t)v;
      break;
    }
    s = (s + 1) & mask;
  }
  bool keep = head != NGIA_NO_RESIDENT;
  for (uint32_t j = 0; j < 63; j++) {
    const uint32_t hj = __shfl(head, (int)j);
    if (j < lane && hj == head) keep = false;
  }
  const uint32_t qi = base + q;
  const uint32_t Lq = len[qi];
  const uint32_t Lr = keep ? len[head] : 0u;
  const uint32_t rep = Lr >= Lq ? head : qi, tar = Lr >= Lq ? qi : head;
  if (keep) keep = !((float)len[tar] < (float)len[rep] * thr);  // buildGraph.hpp:262-266
  const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
  const uint32_t c = (uint32_t)__popcll(m);
  unsigned long long at = 0;
  if (lane == 0 && c) at = atomicAdd(npairs, (unsigned long long)c);
  at = __shfl(at, 0);
  if (keep) {
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    pairs[at + below] = make_uint2(rep, tar);
  }
  if (lane == 0) ncand[q] = c;
}

// best[q] = max over q's passing pairs of (score << 32 | ~resident): the
// highest score, ties to the lowest resident index.
__global__ void k_serve_best(const uint2* __restrict__ pairs, const unsigned long long* __restrict__ npairs,
                             const uint32_t* __restrict__ score, uint32_t n_res, uint32_t base,
                             unsigned long long* __restrict__ best) {
  const uint64_t P = *npairs;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t sc = score[p];
    if (!sc) continue;
    const uint2 pr = pairs[p];
    const uint32_t r = pr.x < n_res ? pr.x : pr.y;
    const uint32_t q = (pr.x < n_res ? pr.y : pr.x) - base;
    atomicMax(&best[q], ((unsigned long long)sc << 32) | (0xFFFFFFFFu - r));
  }
}

__global__ void k_serve_result(const unsigned long long* __restrict__ best,
                               const uint32_t* __restrict__ ncand, uint32_t count,
                               const uint32_t* __restrict__ cluster_of,
                               const uint64_t* __restrict__ qid, uint32_t new_base,
                               ngia_assignment* __restrict__ out) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const unsigned long long b = best[q];
  ngia_assignment a;
  a.candidates = ncand[q];
  if (b == 0) {
    a.cluster = new_base + (uint32_t)qid[q];
    a.resident = NGIA_NO_RESIDENT;
    a.score = 0;
  } else {
    a.resident = 0xFFFFFFFFu - (uint32_t)b;
    a.score = (uint32_t)(b >> 32);
    a.cluster = cluster_of[a.resident];
  }
  out[q] = a;
}

}  // namespace
}  // namespace ngia

namespace {

// Queries [first, first + count) of request `req`, at slot `slot` of a batch.
struct Part {
  uint32_t req, first, count;
};
struct Seg {
  uint32_t req, first, count, slot;
};

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // pinned host staging: packed words, then per-query word_off, code_off, lengths, qid
  uint32_t* h_packed = nullptr;
  uint64_t* h_word_off = nullptr;
  uint64_t* h_code_off = nullptr;
  uint32_t* h_len = nullptr;
  uint64_t* h_qid = nullptr;
  ngia_assignment* h_out = nullptr;
  DevBuf sig, pairs, score, bin_of, list, counters, best, ncand, qid, out, plain_jobs, plain_scratch;
  std::vector<Seg> segs;
  uint32_t count = 0;
  bool busy = false;
  ~Lane() {
    for (void* p : {(void*)h_packed, (void*)h_word_off, (void*)h_code_off, (void*)h_len,
                    (void*)h_qid, (void*)h_out})
      if (p) (void)hipHostFree(p);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace

struct ngia_serve {
  ngia_ctx* ctx = nullptr;
  uint32_t n = 0;            // residents
  uint64_t res_words = 0;    // resident packed words
  uint64_t res_codes = 0;    // resident code bytes
  uint32_t identity = 0, new_base = 0, cap_log2 = 0;
  float thr = 0;
  uint32_t lanes = 2, qmax = 8192;
  uint64_t wmax = 0;         // packed words per lane
  uint64_t cmax = 0;         // code bytes per lane
  uint64_t rmax = 0;         // residues per batch
  uint32_t nw_cap = 0;       // words of the longest sequence allowed (plain-bin scratch)
  DevBuf packed, word_off, lengths, codes, code_off, tables, cluster_of;
  std::vector<std::unique_ptr<Lane>> L;
};

namespace {

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

constexpr uint32_t kPlainHintJobs = 256 * 64;  // grid of the plain bin: 256 blocks

void serve_setup(ngia_serve* sv, ngia_db* db, const uint32_t* cluster_of,
                 const ngia_serve_config* cfg) {
  ngia_ctx* ctx = sv->ctx;
  hipStream_t s = ctx->stream;
  sv->n = db->n;
  sv->lanes = cfg && cfg->lanes ? cfg->lanes : 2;
  sv->qmax = cfg && cfg->max_queries ? cfg->max_queries : 8192;
  if (sv->lanes > 8) throw HipError{NGIA_ERR_ARG, "at most 8 lanes"};
  if (sv->qmax > (1u << 20)) throw HipError{NGIA_ERR_ARG, "max_queries above 2^20"};
  if ((uint64_t)sv->n + (uint64_t)sv->lanes * sv->qmax >= 0xFFFFFFFFull)
    throw HipError{NGIA_ERR_ARG, "residents + query slots exceed 2^32"};
  // Lane capacity in residues from the free HBM: a residue costs 1/8 packed
  // word (4 B per 32 residues x 4 planes / 32) + 1 code byte.
  size_t free_b = 0, total_b = 0;
  NGIA_HIP(hipMemGetInfo(&free_b, &total_b));
  const double frac = cfg && cfg->hbm_fraction > 0 ? cfg->hbm_fraction : 0.25;
  const uint64_t per_query = 256 + 64 * (8 + 4 + 4 + 4 + 12) + 64;  // sig, pairs, bins, jobs
  const uint64_t plain = 4ull * kPlainHintJobs * 2048;               // plain-bin rows
  const double budget = frac * (double)free_b / sv->lanes - (double)per_query * sv->qmax - plain;
  uint64_t rmax = budget > 0 ? (uint64_t)(budget / 2.0) : 0;  // ~2 B per residue with slack
  rmax = std::min<uint64_t>(rmax, 1ull << 26);
  if (cfg && cfg->max_residues) rmax = std::min<uint64_t>(rmax, cfg->max_residues);
  if (rmax < 65535) throw HipError{NGIA_ERR_NOMEM, "free HBM too small for a serving lane"};
  // Room for any qmax queries of rmax residues in all: each record adds at
  // most 4 words and 32 code bytes of padding beyond its residues.
  sv->rmax = rmax;
  sv->wmax = round_up(rmax / 8 + 5ull * sv->qmax + 8, 64);
  // A record's code bytes are 8x its packed words less 8: at most 8 * wmax.
  sv->cmax = round_up(8 * sv->wmax + 64, 256);
  sv->nw_cap = (0xFFFEu + 31) / 32;
  const uint64_t slots = (uint64_t)sv->n + (uint64_t)sv->lanes * sv->qmax;
  sv->res_words = db->n_words;
  uint64_t rc = 0;
  {
    std::vector<uint32_t> len(sv->n);
    if (sv->n) NGIA_HIP(hipMemcpy(len.data(), db->lengths.p, 4ull * sv->n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < sv->n; i++) rc += 32ull * ((len[i] + 31) / 32);
  }
  sv->res_codes = round_up(rc, 256);
  sv->packed.ensure(4 * (sv->res_words + sv->lanes * sv->wmax));
  sv->word_off.ensure(8 * slots);
  sv->lengths.ensure(4 * slots);
  sv->code_off.ensure(8 * slots);
  sv->codes.ensure(sv->res_codes + sv->lanes * sv->cmax + 64);
  if (sv->n) {
    NGIA_HIP(hipMemcpyAsync(sv->packed.p, db->packed.p, 4 * sv->res_words, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->word_off.p, db->word_off.p, 8ull * sv->n, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->lengths.p, db->lengths.p, 4ull * sv->n, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->code_off.p, db->code_off.p, 8ull * sv->n, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->codes.p, db->codes.p, rc, hipMemcpyDeviceToDevice, s));
  }
  // Per-band head tables: 2^cap_log2 >= 2n slots.
  sv->cap_log2 = 10;
  while ((1ull << sv->cap_log2) < 2ull * std::max<uint32_t>(sv->n, 1)) sv->cap_log2++;
  const uint64_t tslots = 64ull << sv->cap_log2;
  sv->tables.ensure(8 * tslots);
  fill_u32(sv->tables.as<uint32_t>(), 0xFFFFFFFFu, 2 * tslots, s);
  if (sv->n)
    hipLaunchKernelGGL(k_serve_insert, dim3(persistent_blocks(ctx, 0) * 4), dim3(256), 0, s,
                       db->sig.as<uint32_t>(), sv->n, sv->tables.as<unsigned long long>(),
                       sv->cap_log2);
  NGIA_HIP(hipGetLastError());
  sv->cluster_of.ensure(4ull * std::max<uint32_t>(sv->n, 1));
  if (sv->n) NGIA_HIP(hipMemcpyAsync(sv->cluster_of.p, cluster_of, 4ull * sv->n, hipMemcpyHostToDevice, s));
  NGIA_HIP(hipStreamSynchronize(s));
  // Lanes.
  const uint64_t pcap = 64ull * sv->qmax;  // at most one pair per band per query
  for (uint32_t k = 0; k < sv->lanes; k++) {
    auto ln = std::make_unique<Lane>();
    NGIA_HIP(hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking));
    NGIA_HIP(hipEventCreate(&ln->t0));
    NGIA_HIP(hipEventCreate(&ln->t1));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_packed, 4 * sv->wmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_word_off, 8ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_code_off, 8ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_len, 4ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_qid, 8ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_out, sizeof(ngia_assignment) * sv->qmax, 0));
    ln->sig.ensure(256ull * sv->qmax);
    ln->pairs.ensure(8 * pcap);
    ln->score.ensure(4 * pcap);
    ln->bin_of.ensure(4 * pcap);
    ln->list.ensure(4 * pcap);
    ln->counters.ensure(sizeof(Counters));
    ln->best.ensure(8ull * sv->qmax);
    ln->ncand.ensure(4ull * sv->qmax);
    ln->qid.ensure(8ull * sv->qmax);
    ln->out.ensure(sizeof(ngia_assignment) * sv->qmax);
    ln->plain_jobs.ensure(12 * pcap);
    ln->plain_scratch.ensure(4 * align_plain_scratch_words(sv->nw_cap) / 4096 * 256);
    sv->L.push_back(std::move(ln));
  }
}

// Pack the parts into lane k's staging, copy them in and launch the batch.
void launch_batch(ngia_serve* sv, uint32_t k, const uint8_t* res, const uint64_t* res_off,
                  const uint32_t* res_len, uint64_t qid0) {
  Lane& ln = *sv->L[k];
  const uint32_t Q = ln.count;
  const uint32_t base = sv->n + k * sv->qmax;
  const uint64_t wbase = sv->res_words + k * sv->wmax;
  const uint64_t cbase = sv->res_codes + k * sv->cmax;
  uint64_t w = 0, c = 0;
  for (const Seg& sg : ln.segs) {
    for (uint32_t j = 0; j < sg.count; j++) {
      const uint32_t q = sg.first + j, slot = sg.slot + j;
      const uint32_t Lq = res_len[q];
      ln.h_word_off[slot] = wbase + w;
      ln.h_code_off[slot] = cbase + c;
      ln.h_len[slot] = Lq;
      ln.h_qid[slot] = qid0 + q;
      w += rec_words(Lq);
      c += 32ull * ((Lq + 31) / 32);
    }
  }
  // Pack the residues (the batch's host work, overlapping the other lanes' kernels).
  for (const Seg& sg : ln.segs)
    pack_records(res, res_off, res_len, sg.first, sg.count, ln.h_word_off + sg.slot, wbase,
                 ln.h_packed);
  hipStream_t s = ln.stream;
  NGIA_HIP(hipEventRecord(ln.t0, s));
  NGIA_HIP(hipMemcpyAsync(sv->packed.as<uint32_t>() + wbase, ln.h_packed, 4 * w, hipMemcpyHostToDevice, s));
  NGIA_HIP(hipMemcpyAsync(sv->word_off.as<uint64_t>() + base, ln.h_word_off, 8ull * Q, hipMemcpyHostToDevice, s));
  NGIA_HIP(hipMemcpyAsync(sv->code_off.as<uint64_t>() + base, ln.h_code_off, 8ull * Q, hipMemcpyHostToDevice, s));
  NGIA_HIP(hipMemcpyAsync(sv->lengths.as<uint32_t>() + base, ln.h_len, 4ull * Q, hipMemcpyHostToDevice, s));
  NGIA_HIP(hipMemcpyAsync(ln.qid.p, ln.h_qid, 8ull * Q, hipMemcpyHostToDevice, s));
  launch_codes(sv->packed.as<uint32_t>(), sv->word_off.as<uint64_t>() + base,
               sv->code_off.as<uint64_t>() + base, Q, sv->codes.as<uint8_t>(), s);
  launch_sketch(sv->codes.as<uint8_t>(), sv->code_off.as<uint64_t>() + base,
                sv->lengths.as<uint32_t>() + base, Q, ln.sig.as<uint32_t>(), s);
  Counters* dc = ln.counters.as<Counters>();
  fill_u32(ln.counters.as<uint32_t>(), 0, sizeof(Counters) / 4, s);
  fill_u32(ln.best.as<uint32_t>(), 0, 2ull * Q, s);
  hipLaunchKernelGGL(k_serve_lookup, dim3(Q), dim3(64), 0, s, ln.sig.as<uint32_t>(), Q, base,
                     sv->tables.as<unsigned long long>(), sv->cap_log2, sv->lengths.as<uint32_t>(),
                     sv->thr, ln.pairs.as<uint2>(), &dc->total_cand, ln.ncand.as<uint32_t>());
  NGIA_HIP(hipGetLastError());
  // The length bins, one after another on the lane stream.
  const uint64_t* np = (const uint64_t*)&dc->total_cand;
  const uint32_t grid = persistent_blocks(sv->ctx, 0);
  bin_pairs(ln.pairs.as<uint2>(), np, sv->lengths.as<uint32_t>(), sv->thr, ln.bin_of.as<uint32_t>(),
            dc->bin_cnt, dc->bin_off, dc->cursor, ln.list.as<uint32_t>(), dc->bin_tiles,
            ln.score.as<uint32_t>(), sv->ctx->wave_kmax, dc->remap, sv->ctx->route, grid, s);
  AlignArgs aa;
  aa.packed = sv->packed.as<uint32_t>();
  aa.word_off = sv->word_off.as<uint64_t>();
  aa.codes = sv->codes.as<uint8_t>();
  aa.code_off = sv->code_off.as<uint64_t>();
  aa.thr = sv->thr;
  for (int kk = 5; kk >= 0; kk--) {
    if ((1u << kk) > sv->ctx->wave_kmax) continue;
    const int bin = kBinWave0 + kk;
    launch_align_wave(1 << kk, aa, ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), &dc->bin_off[bin],
                      &dc->bin_cnt[bin], &dc->work[bin], ln.score.as<uint32_t>(),
                      std::min<uint32_t>(Q * 64, persistent_blocks(sv->ctx, wave_lds_bytes(1 << kk))), s);
  }
  for (int W = kWinMax; W >= 2; W--)
    launch_align_win(W, aa, ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), &dc->bin_off[W],
                     &dc->bin_cnt[W], &dc->work[W], ln.score.as<uint32_t>(),
                     std::min<uint32_t>(Q, persistent_blocks(sv->ctx, win_lds_bytes(W))), s);
  plain_bin_jobs(ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), dc->bin_off, dc->bin_cnt,
                 ln.plain_jobs.as<uint32_t>(), s);
  launch_align_plain(aa, ln.plain_jobs.as<uint32_t>(), kPlainHintJobs, sv->nw_cap,
                     ln.plain_scratch.as<uint32_t>(), ln.plain_scratch.bytes / 4, s,
                     &dc->bin_cnt[kBinPlain]);
  plain_bin_scores(ln.list.as<uint32_t>(), dc->bin_off, dc->bin_cnt, ln.plain_jobs.as<uint32_t>(),
                   ln.score.as<uint32_t>(), s);
  hipLaunchKernelGGL(k_serve_best, dim3(grid), dim3(256), 0, s, ln.pairs.as<uint2>(),
                     (const unsigned long long*)np, ln.score.as<uint32_t>(), sv->n, base,
                     ln.best.as<unsigned long long>());
  hipLaunchKernelGGL(k_serve_result, dim3((Q + 255) / 256), dim3(256), 0, s,
                     ln.best.as<unsigned long long>(), ln.ncand.as<uint32_t>(), Q,
                     sv->cluster_of.as<uint32_t>(), ln.qid.as<uint64_t>(), sv->new_base,
                     ln.out.as<ngia_assignment>());
  NGIA_HIP(hipGetLastError());
  NGIA_HIP(hipMemcpyAsync(ln.h_out, ln.out.p, sizeof(ngia_assignment) * Q, hipMemcpyDeviceToHost, s));
  NGIA_HIP(hipEventRecord(ln.t1, s));
  ln.busy = true;
}

// The engine behind assign and replay: requests enter a FIFO at their
// arrival times; whenever a lane is free the queue head is packed into a
// batch up to the lane's capacity (a request larger than that is split).
int run_stream(ngia_serve* sv, const uint8_t* res, const uint64_t* res_off, const uint32_t* res_len,
               uint32_t nq, const ngia_request* reqs, uint32_t nr, uint64_t qid0,
               ngia_assignment* out, double* latency, ngia_serve_stats* stats) {
  using clk = std::chrono::steady_clock;
  for (uint32_t q = 0; q < nq; q++)
    if (res_len[q] >= 0xFFFF) throw HipError{NGIA_ERR_ARG, "query length >= 65535"};
  for (uint32_t r = 0; r < nr; r++) {
    if ((uint64_t)reqs[r].first + reqs[r].count > nq) throw HipError{NGIA_ERR_ARG, "request outside the query set"};
    if (r && reqs[r].arrival < reqs[r - 1].arrival) throw HipError{NGIA_ERR_ARG, "arrivals must not decrease"};
  }
  const uint64_t rmax_words = sv->wmax;
  std::vector<uint32_t> parts_left(nr, 0);
  std::deque<Part> pending;
  uint64_t batches = 0, pairs = 0, newc = 0;
  double device_s = 0;
  uint32_t next = 0, done = 0;
  const auto t_start = clk::now();
  auto now_s = [&] { return std::chrono::duration<double>(clk::now() - t_start).count(); };
  for (auto& ln : sv->L) ln->busy = false;
  uint32_t empty_done = 0;
  while (done + empty_done < nr) {
    bool progress = false;
    double now = now_s();
    while (next < nr && reqs[next].arrival <= now) {
      if (reqs[next].count == 0) {
        if (latency) latency[next] = 0;
        empty_done++;
      } else {
        pending.push_back(Part{next, reqs[next].first, reqs[next].count});
        parts_left[next] = 1;
      }
      next++;
      progress = true;
    }
    for (uint32_t k = 0; k < sv->lanes; k++) {
      Lane& ln = *sv->L[k];
      if (!ln.busy) continue;
      const hipError_t e = hipEventQuery(ln.t1);
      if (e == hipErrorNotReady) continue;
      NGIA_HIP(e);
      float ms = 0;
      NGIA_HIP(hipEventElapsedTime(&ms, ln.t0, ln.t1));
      device_s += ms * 1e-3;
      const double t = now_s();
      for (const Seg& sg : ln.segs) {
        for (uint32_t j = 0; j < sg.count; j++) {
          out[sg.first + j] = ln.h_out[sg.slot + j];
          newc += ln.h_out[sg.slot + j].resident == NGIA_NO_RESIDENT;
          pairs += ln.h_out[sg.slot + j].candidates;
        }
        if (--parts_left[sg.req] == 0) {
          if (latency) latency[sg.req] = t - reqs[sg.req].arrival;
          done++;
        }
      }
      ln.segs.clear();
      ln.busy = false;
      progress = true;
    }
    for (uint32_t k = 0; k < sv->lanes && !pending.empty(); k++) {
      Lane& ln = *sv->L[k];
      if (ln.busy) continue;
      ln.count = 0;
      uint64_t words = 0;
      while (!pending.empty() && ln.count < sv->qmax) {
        Part& pt = pending.front();
        uint32_t take = 0;
        while (take < pt.count && ln.count + take < sv->qmax) {
          const uint64_t wq = rec_words(res_len[pt.first + take]);
          if (words + wq > rmax_words) break;
          words += wq;
          take++;
        }
        if (take == 0) break;
        ln.segs.push_back(Seg{pt.req, pt.first, take, ln.count});
        ln.count += take;
        if (take == pt.count) {
          pending.pop_front();
        } else {  // the rest waits for the next free lane
          pt.first += take;
          pt.count -= take;
          parts_left[pt.req]++;
          break;
        }
      }
      if (ln.count == 0) throw HipError{NGIA_ERR_ARG, "a query does not fit a serving lane"};
      launch_batch(sv, k, res, res_off, res_len, qid0);
      batches++;
      progress = true;
    }
    if (!progress) {
      bool idle = true;
      for (auto& ln : sv->L) idle = idle && !ln->busy;
      if (idle && pending.empty() && next < nr) {
        const double wait = reqs[next].arrival - now_s();
        if (wait > 2e-4) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
      } else {
        std::this_thread::yield();
      }
    }
  }
  if (stats) {
    stats->queries = nq;
    stats->requests = nr;
    stats->batches = batches;
    stats->pairs = pairs;
    stats->new_clusters = newc;
    stats->lanes = sv->lanes;
    stats->max_queries = sv->qmax;
    stats->max_residues = sv->rmax;
    stats->wall_s = now_s();
    stats->device_s = device_s;
  }
  return NGIA_OK;
}

}  // namespace

extern "C" int ngia_serve_create(ngia_ctx* ctx, ngia_db* db, const uint32_t* cluster_of,
                                 uint32_t identity, uint32_t new_base, const ngia_serve_config* cfg,
                                 ngia_serve** out) {
  if (!ctx || !db || !out || (db->n && !cluster_of)) return fail(ctx, NGIA_ERR_ARG, "null argument");
  *out = nullptr;
  if (identity < 1 || identity > 99) return fail(ctx, NGIA_ERR_ARG, "identity should be 1-99");
  if (!db->have_sig) return fail(ctx, NGIA_ERR_STATE, "database has no signatures (ngia_db_sketch)");
  auto sv = std::make_unique<ngia_serve>();
  sv->ctx = ctx;
  sv->identity = identity;
  sv->thr = (float)identity / 100.0f;
  sv->new_base = new_base;
  const int rc = guarded(ctx, [&] {
    NGIA_HIP(hipSetDevice(ctx->device));
    serve_setup(sv.get(), db, cluster_of, cfg);
    return NGIA_OK;
  });
  if (rc != NGIA_OK) return rc;
  *out = sv.release();
  ctx->serves.push_back(*out);
  return NGIA_OK;
}

void serve_detach(ngia_serve* sv) { sv->ctx = nullptr; }

extern "C" void ngia_serve_destroy(ngia_serve* sv) {
  if (!sv) return;
  if (sv->ctx) {
    (void)hipSetDevice(sv->ctx->device);
    auto& v = sv->ctx->serves;
    v.erase(std::remove(v.begin(), v.end(), sv), v.end());
  }
  for (auto& ln : sv->L)
    if (ln->stream) (void)hipStreamSynchronize(ln->stream);
  delete sv;
}

extern "C" int ngia_serve_assign(ngia_serve* sv, const uint8_t* res, const uint64_t* res_off,
                                 const uint32_t* res_len, uint32_t n_queries, uint64_t first_qid,
                                 ngia_assignment* out, ngia_serve_stats* stats) {
  if (!sv) return fail(nullptr, NGIA_ERR_ARG, "null serving handle");
  if (!sv->ctx) return fail(nullptr, NGIA_ERR_STATE, "the serving handle's context is closed");
  if (n_queries && (!res || !res_off || !res_len || !out)) return fail(sv->ctx, NGIA_ERR_ARG, "null argument");
  return guarded(sv->ctx, [&] {
    NGIA_HIP(hipSetDevice(sv->ctx->device));
    const ngia_request r{0.0, 0, n_queries};
    return run_stream(sv, res, res_off, res_len, n_queries, &r, 1, first_qid, out, nullptr, stats);
  });
}

extern "C" int ngia_serve_replay(ngia_serve* sv, const uint8_t* res, const uint64_t* res_off,
                                 const uint32_t* res_len, uint32_t n_queries,
                                 const ngia_request* reqs, uint32_t n_reqs, ngia_assignment* out,
                                 double* latency_s, ngia_serve_stats* stats) {
  if (!sv) return fail(nullptr, NGIA_ERR_ARG, "null serving handle");
  if (!sv->ctx) return fail(nullptr, NGIA_ERR_STATE, "the serving handle's context is closed");
  if ((n_queries && (!res || !res_off || !res_len || !out)) || (n_reqs && !reqs))
    return fail(sv->ctx, NGIA_ERR_ARG, "null argument");
  return guarded(sv->ctx, [&] {
    NGIA_HIP(hipSetDevice(sv->ctx->device));
    return run_stream(sv, res, res_off, res_len, n_queries, reqs, n_reqs, 0, out, latency_s, stats);
  });
}
