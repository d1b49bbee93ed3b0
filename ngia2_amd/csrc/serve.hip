// serve.hip -- the serving loop of include/ngia2.h: a clustered database
// kept resident on the device, streamed query batches assigned to clusters.
//
// Snapshot: the database's packed records, byte codes, lengths and
// signatures, followed by a region that grows with the stream.  Every served
// query is appended in arrival order: query number p of the session is
// snapshot index n + p, its record and codes stay on the device, and its 64
// band signatures enter the per-band head tables.  A table slot maps a
// signature to its bucket head, the lowest snapshot index with that
// signature (atomicCAS + atomicMin, so the result does not depend on the
// order of insertion within a batch).  The length-binned alignment kernels of
// the batched graph run unchanged on snapshot pairs.
//
// One batch on one lane stream: copy the packed queries into their snapshot
// slots, byte codes, sketch; insert the batch's signatures (after the
// previous batch's insertion, whatever lane it ran on); per query the heads
// of its 64 signatures that lie before it (one wavefront per query, lane =
// band, distinct heads kept, longer sequence as rep, the isJob length test);
// the length bins; per query the best (score, lowest index); then its
// cluster (after the previous batch's clusters): a resident's cluster, an
// earlier query's cluster, or a new one.  A query's result depends only on
// the residents and the queries before it, never on how the stream was cut
// into batches, so it is the same alone, in any mix and offline.
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

// Insert (sig, index0 + i) for i < count and every band (signatures
// sequence-major, 64 per sequence); a slot keeps the lowest index of its
// signature.
__global__ void k_serve_insert(const uint32_t* __restrict__ sig, uint32_t count, uint32_t index0,
                               unsigned long long* __restrict__ tables, uint32_t cap_log2) {
  const uint64_t total = (uint64_t)count * 64;
  const uint32_t mask = (1u << cap_log2) - 1;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = (uint32_t)(t >> 6), b = (uint32_t)(t & 63);
    const uint32_t key = sig[t];
    unsigned long long* tb = tables + ((uint64_t)b << cap_log2);
    const unsigned long long val = ((unsigned long long)key << 32) | (index0 + i);
    uint32_t s = slot_of_sig(key, cap_log2);
    // At most half the 2^cap_log2 slots of a band ever fill: an empty slot exists.
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

// Back to the residents: empty every slot whose head is a grown index (>= n).
// The residents were inserted before any query, so no resident signature's
// probe sequence passes through a grown slot, and a resident slot never
// holds a grown index (atomicMin keeps the lower one).
__global__ void k_serve_unwind(unsigned long long* __restrict__ tables, uint64_t slots, uint32_t n) {
  for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < slots;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long v = tables[s];
    if (v != kEmpty && (uint32_t)v >= n) tables[s] = kEmpty;
  }
}

// One wavefront per query, lane b = band: the bucket head of the query's
// band-b signature when it lies before the query (the batch is inserted
// first, so the head is at most the query's own index), distinct heads kept
// once (lowest band), oriented (longer as rep, ties to the earlier index) and
// length-tested, appended to the lane's pair list.
__global__ __launch_bounds__(64) void k_serve_lookup(
    const uint32_t* __restrict__ qsig, uint32_t count, uint32_t base,
    const unsigned long long* __restrict__ tables, uint32_t cap_log2,
    const uint32_t* __restrict__ len, float thr, uint2* __restrict__ pairs,
    unsigned long long* __restrict__ npairs, uint32_t* __restrict__ ncand) {
  const uint32_t q = blockIdx.x;
  if (q >= count) return;
  const uint32_t lane = threadIdx.x;
  const uint32_t qi = base + q;
  const uint32_t key = qsig[(uint64_t)q * 64 + lane];
  const unsigned long long* tb = tables + ((uint64_t)lane << cap_log2);
  const uint32_t mask = (1u << cap_log2) - 1;
  uint32_t s = slot_of_sig(key, cap_log2);
  uint32_t head = NGIA_NO_RESIDENT;
  for (uint32_t probe = 0; probe <= mask; probe++) {
    const unsigned long long v = __hip_atomic_load(&tb[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == kEmpty) break;
    if ((uint32_t)(v >> 32) == key) {
      head = (uint32_t)v;
      break;
    }
    s = (s + 1) & mask;
  }
  bool keep = head < qi;
  for (uint32_t j = 0; j < 63; j++) {
    const uint32_t hj = __shfl(head, (int)j);
    if (j < lane && hj == head) keep = false;
  }
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

// best[q] = max over q's passing pairs of (score << 32 | ~head): the highest
// score, ties to the lowest index.  In every pair the head is the lower index
// and the query the higher.
__global__ void k_serve_best(const uint2* __restrict__ pairs, const unsigned long long* __restrict__ npairs,
                             const uint32_t* __restrict__ score, uint32_t base,
                             unsigned long long* __restrict__ best) {
  const uint64_t P = *npairs;
  for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P;
       p += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t sc = score[p];
    if (!sc) continue;
    const uint2 pr = pairs[p];
    const uint32_t h = min(pr.x, pr.y);
    const uint32_t q = max(pr.x, pr.y) - base;
    atomicMax(&best[q], ((unsigned long long)sc << 32) | (0xFFFFFFFFu - h));
  }
}

// The batch's clusters, in one workgroup: a query joins the cluster of its
// best (a resident's cluster_of, the cluster of an earlier query of the
// session, kept in qcluster by stream position) or opens new_base + its id.
// A best that is an earlier query of the same batch is resolved in a later
// round than that query; each round resolves at least the first unresolved
// query, so the loop ends.  done[] starts at zero.
__global__ __launch_bounds__(1024) void k_serve_result(
    const unsigned long long* __restrict__ best, const uint32_t* __restrict__ ncand,
    uint32_t count, uint32_t n, uint32_t pos0, const uint32_t* __restrict__ cluster_of,
    uint32_t* __restrict__ qcluster, uint32_t* __restrict__ done, const uint64_t* __restrict__ qid,
    uint32_t new_base, ngia_assignment* __restrict__ out) {
  __shared__ int pending;
  for (;;) {
    if (threadIdx.x == 0) pending = 0;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < count; q += blockDim.x) {
      if (__hip_atomic_load(&done[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) continue;
      const unsigned long long b = best[q];
      ngia_assignment a;
      a.candidates = ncand[q];
      if (b == 0) {
        a.cluster = new_base + (uint32_t)qid[q];
        a.resident = NGIA_NO_RESIDENT;
        a.score = 0;
      } else {
        const uint32_t h = 0xFFFFFFFFu - (uint32_t)b;
        a.resident = h;
        a.score = (uint32_t)(b >> 32);
        if (h < n) {
          a.cluster = cluster_of[h];
        } else if (h - n < pos0) {
          a.cluster = qcluster[h - n];  // an earlier batch, resolved before this kernel
        } else {
          const uint32_t j = h - n - pos0;  // earlier in this batch: j < q
          if (!__hip_atomic_load(&done[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) {
            pending = 1;
            continue;
          }
          a.cluster = __hip_atomic_load(&qcluster[pos0 + j], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
      out[q] = a;
      __hip_atomic_store(&qcluster[pos0 + q], a.cluster, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(&done[q], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __syncthreads();
    const int more = pending;
    __syncthreads();
    if (!more) break;
  }
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
  hipEvent_t inserted = nullptr, resolved = nullptr;  // ordering marks for the next batch
  // pinned host staging: packed words, then per-query word_off, code_off, lengths, qid
  uint32_t* h_packed = nullptr;
  uint64_t* h_word_off = nullptr;
  uint64_t* h_code_off = nullptr;
  uint32_t* h_len = nullptr;
  uint64_t* h_qid = nullptr;
  ngia_assignment* h_out = nullptr;
  // device buffers of qcap queries (grown at admission, by free HBM)
  uint32_t qcap = 0;
  DevBuf sig, pairs, score, bin_of, list, counters, best, ncand, qid, out, done, plain_jobs,
      plain_scratch;
  std::vector<Seg> segs;
  uint32_t count = 0;
  bool busy = false;
  ~Lane() {
    for (void* p : {(void*)h_packed, (void*)h_word_off, (void*)h_code_off, (void*)h_len,
                    (void*)h_qid, (void*)h_out})
      if (p) (void)hipHostFree(p);
    for (hipEvent_t e : {t0, t1, inserted, resolved})
      if (e) (void)hipEventDestroy(e);
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
  double hbm_fraction = 0.25;
  uint64_t wmax = 0;         // packed words of one batch (staging)
  uint64_t rmax = 0;         // residues of one batch
  uint32_t nw_cap = 0;       // words of the longest sequence allowed (plain-bin scratch)
  // the grown part: capacity and use (queries, packed words, code bytes)
  uint32_t grow_q = 0;
  uint64_t grow_words = 0, grow_codes = 0;
  uint32_t pos = 0;
  uint64_t wcur = 0, ccur = 0;
  hipEvent_t last_inserted = nullptr, last_resolved = nullptr;
  DevBuf packed, word_off, lengths, codes, code_off, tables, cluster_of, qcluster;
  std::vector<std::unique_ptr<Lane>> L;
};

namespace {

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// The plain bin (band windows wider than 64 x 32 words: very long pairs at low
// identity) runs on 64 blocks of 64 threads per lane, rows of up to 2,048 words.
constexpr uint32_t kServePlainBlocks = 64;

// Device bytes of a lane's buffers for q queries (at most one pair per band
// per query).
uint64_t lane_bytes(uint32_t q) {
  return (uint64_t)q * (256 + 64 * (8 + 4 + 4 + 4 + 12) + 8 + 4 + 8 + sizeof(ngia_assignment) + 4);
}

void lane_alloc(Lane& ln, uint32_t q) {
  const uint64_t pcap = 64ull * q;
  ln.sig.ensure(256ull * q);
  ln.pairs.ensure(8 * pcap);
  ln.score.ensure(4 * pcap);
  ln.bin_of.ensure(4 * pcap);
  ln.list.ensure(4 * pcap);
  ln.best.ensure(8ull * q);
  ln.ncand.ensure(4ull * q);
  ln.qid.ensure(8ull * q);
  ln.out.ensure(sizeof(ngia_assignment) * q);
  ln.done.ensure(4ull * q);
  ln.plain_jobs.ensure(12 * pcap);
  ln.qcap = q;
}

void serve_setup(ngia_serve* sv, ngia_db* db, const uint32_t* cluster_of,
                 const ngia_serve_config* cfg) {
  ngia_ctx* ctx = sv->ctx;
  hipStream_t s = ctx->stream;
  sv->n = db->n;
  sv->lanes = cfg && cfg->lanes ? cfg->lanes : 2;
  sv->qmax = cfg && cfg->max_queries ? cfg->max_queries : 8192;
  sv->hbm_fraction = cfg && cfg->hbm_fraction > 0 ? cfg->hbm_fraction : 0.25;
  sv->grow_q = cfg && cfg->grow_queries ? cfg->grow_queries : (1u << 20);
  const uint64_t grow_res = cfg && cfg->grow_residues ? cfg->grow_residues : (1ull << 28);
  if (sv->lanes > 8) throw HipError{NGIA_ERR_ARG, "at most 8 lanes"};
  if (sv->qmax > (1u << 20)) throw HipError{NGIA_ERR_ARG, "max_queries above 2^20"};
  if (sv->grow_q > (1u << 26)) throw HipError{NGIA_ERR_ARG, "grow_queries above 2^26"};
  if ((uint64_t)sv->n + sv->grow_q >= 0xFFFFFFFFull)
    throw HipError{NGIA_ERR_ARG, "residents + grown queries reach 2^32"};
  // Residues of one batch: the configured bound, else 2^24, within the grown part.
  uint64_t rmax = cfg && cfg->max_residues ? cfg->max_residues : (1ull << 24);
  rmax = std::min<uint64_t>(rmax, grow_res);
  if (rmax < 65535) throw HipError{NGIA_ERR_ARG, "a batch must hold one sequence of 65,534 residues"};
  sv->rmax = rmax;
  // Staging for any qmax queries of rmax residues: each record adds at most 4
  // words of padding and its length word.
  sv->wmax = round_up(rmax / 8 + 5ull * sv->qmax + 8, 64);
  sv->grow_words = round_up(grow_res / 8 + 5ull * sv->grow_q + 8, 64);
  sv->grow_codes = round_up(grow_res + 32ull * sv->grow_q + 64, 256);
  sv->nw_cap = (0xFFFEu + 31) / 32;
  const uint64_t slots = (uint64_t)sv->n + sv->grow_q;
  sv->res_words = db->n_words;
  uint64_t rc = 0;
  {
    std::vector<uint32_t> len(sv->n);
    if (sv->n) NGIA_HIP(hipMemcpy(len.data(), db->lengths.p, 4ull * sv->n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < sv->n; i++) rc += 32ull * ((len[i] + 31) / 32);
  }
  sv->res_codes = round_up(rc, 256);
  // Per-band head tables: 2^cap_log2 >= 2 (n + grow_q) slots.
  sv->cap_log2 = 10;
  while ((1ull << sv->cap_log2) < 2 * slots) sv->cap_log2++;
  const uint64_t tslots = 64ull << sv->cap_log2;
  // The snapshot and its grown part must fit the free HBM.
  size_t free_b = 0, total_b = 0;
  NGIA_HIP(hipMemGetInfo(&free_b, &total_b));
  const uint64_t need = 4 * (sv->res_words + sv->grow_words) + 24 * slots + sv->res_codes +
                        sv->grow_codes + 8 * tslots + 4ull * sv->grow_q;
  if (need > free_b)
    throw HipError{NGIA_ERR_NOMEM, "snapshot and grown part (" + std::to_string(need >> 20) +
                                       " MiB) exceed the free HBM (" + std::to_string(free_b >> 20) +
                                       " MiB); lower grow_queries / grow_residues"};
  sv->packed.ensure(4 * (sv->res_words + sv->grow_words));
  sv->word_off.ensure(8 * slots);
  sv->lengths.ensure(4 * slots);
  sv->code_off.ensure(8 * slots);
  sv->codes.ensure(sv->res_codes + sv->grow_codes + 64);
  sv->qcluster.ensure(4ull * std::max<uint32_t>(sv->grow_q, 1));
  if (sv->n) {
    NGIA_HIP(hipMemcpyAsync(sv->packed.p, db->packed.p, 4 * sv->res_words, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->word_off.p, db->word_off.p, 8ull * sv->n, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->lengths.p, db->lengths.p, 4ull * sv->n, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->code_off.p, db->code_off.p, 8ull * sv->n, hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(sv->codes.p, db->codes.p, rc, hipMemcpyDeviceToDevice, s));
  }
  sv->tables.ensure(8 * tslots);
  fill_u32(sv->tables.as<uint32_t>(), 0xFFFFFFFFu, 2 * tslots, s);
  if (sv->n)
    hipLaunchKernelGGL(k_serve_insert, dim3(persistent_blocks(ctx, 0) * 4), dim3(256), 0, s,
                       db->sig.as<uint32_t>(), sv->n, 0u, sv->tables.as<unsigned long long>(),
                       sv->cap_log2);
  NGIA_HIP(hipGetLastError());
  sv->cluster_of.ensure(4ull * std::max<uint32_t>(sv->n, 1));
  if (sv->n) NGIA_HIP(hipMemcpyAsync(sv->cluster_of.p, cluster_of, 4ull * sv->n, hipMemcpyHostToDevice, s));
  NGIA_HIP(hipStreamSynchronize(s));
  // Lanes: host staging for qmax queries; device buffers for up to 256 at
  // first, grown at admission while the free HBM allows.
  for (uint32_t k = 0; k < sv->lanes; k++) {
    auto ln = std::make_unique<Lane>();
    NGIA_HIP(hipStreamCreateWithFlags(&ln->stream, hipStreamNonBlocking));
    NGIA_HIP(hipEventCreate(&ln->t0));
    NGIA_HIP(hipEventCreate(&ln->t1));
    NGIA_HIP(hipEventCreateWithFlags(&ln->inserted, hipEventDisableTiming));
    NGIA_HIP(hipEventCreateWithFlags(&ln->resolved, hipEventDisableTiming));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_packed, 4 * sv->wmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_word_off, 8ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_code_off, 8ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_len, 4ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_qid, 8ull * sv->qmax, 0));
    NGIA_HIP(hipHostMalloc((void**)&ln->h_out, sizeof(ngia_assignment) * sv->qmax, 0));
    ln->counters.ensure(sizeof(Counters));
    ln->plain_scratch.ensure(4 * align_plain_scratch_words(sv->nw_cap, kServePlainBlocks));
    lane_alloc(*ln, std::min<uint32_t>(sv->qmax, 256));
    sv->L.push_back(std::move(ln));
  }
}

// Admission by free HBM: a lane whose buffers hold fewer queries than the
// batch wants grows them (doubling, at most qmax) while its share of the free
// HBM allows; otherwise the batch is cut to what the lane holds.  Called with
// the lane idle.
uint32_t lane_admit(ngia_serve* sv, Lane& ln, uint32_t want) {
  if (want <= ln.qcap) return want;
  uint32_t q = ln.qcap;
  while (q < want && q < sv->qmax) q = std::min<uint32_t>(sv->qmax, 2 * q);
  size_t free_b = 0, total_b = 0;
  NGIA_HIP(hipMemGetInfo(&free_b, &total_b));
  const double share = sv->hbm_fraction * (double)free_b / sv->lanes;
  while (q > ln.qcap && (double)(lane_bytes(q) - lane_bytes(ln.qcap)) > share) q = std::max(ln.qcap, q / 2);
  if (q > ln.qcap) lane_alloc(ln, q);
  return std::min(want, ln.qcap);
}

// Pack the parts into lane k's staging, copy them into their snapshot slots
// and launch the batch.
void launch_batch(ngia_serve* sv, uint32_t k, const uint8_t* res, const uint64_t* res_off,
                  const uint32_t* res_len, uint64_t qid0) {
  Lane& ln = *sv->L[k];
  const uint32_t Q = ln.count;
  const uint32_t pos0 = sv->pos;
  const uint32_t base = sv->n + pos0;
  const uint64_t wbase = sv->res_words + sv->wcur;
  const uint64_t cbase = sv->res_codes + sv->ccur;
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
  sv->pos += Q;  // capacity checked at admission
  sv->wcur += w;
  sv->ccur += c;
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
  // The batch's signatures join the tables after the previous batch's (its
  // records and codes are in place by then), so every lookup below sees all
  // earlier queries of the session.
  if (sv->last_inserted) NGIA_HIP(hipStreamWaitEvent(s, sv->last_inserted, 0));
  const uint32_t grid = persistent_blocks(sv->ctx, 0);
  hipLaunchKernelGGL(k_serve_insert, dim3(grid), dim3(256), 0, s, ln.sig.as<uint32_t>(), Q, base,
                     sv->tables.as<unsigned long long>(), sv->cap_log2);
  NGIA_HIP(hipGetLastError());
  NGIA_HIP(hipEventRecord(ln.inserted, s));
  sv->last_inserted = ln.inserted;
  Counters* dc = ln.counters.as<Counters>();
  fill_u32(ln.counters.as<uint32_t>(), 0, sizeof(Counters) / 4, s);
  fill_u32(ln.best.as<uint32_t>(), 0, 2ull * Q, s);
  fill_u32(ln.done.as<uint32_t>(), 0, Q, s);
  hipLaunchKernelGGL(k_serve_lookup, dim3(Q), dim3(64), 0, s, ln.sig.as<uint32_t>(), Q, base,
                     sv->tables.as<unsigned long long>(), sv->cap_log2, sv->lengths.as<uint32_t>(),
                     sv->thr, ln.pairs.as<uint2>(), &dc->total_cand, ln.ncand.as<uint32_t>());
  NGIA_HIP(hipGetLastError());
  // The length bins, one after another on the lane stream (the snapshot's
  // window kernels use the 16-code table, so its routing applies).
  const uint64_t* np = (const uint64_t*)&dc->total_cand;
  bin_pairs(ln.pairs.as<uint2>(), np, sv->lengths.as<uint32_t>(), sv->thr, ln.bin_of.as<uint32_t>(),
            dc->bin_cnt, dc->bin_off, dc->cursor, ln.list.as<uint32_t>(), dc->bin_tiles,
            ln.score.as<uint32_t>(), sv->ctx->wave_kmax, dc->remap, sv->ctx->route,
            sv->ctx->seg2_min, grid, s);
  AlignArgs aa;
  aa.packed = sv->packed.as<uint32_t>();
  aa.word_off = sv->word_off.as<uint64_t>();
  aa.codes = sv->codes.as<uint8_t>();
  aa.code_off = sv->code_off.as<uint64_t>();
  aa.lengths = sv->lengths.as<uint32_t>();
  aa.thr = sv->thr;
  for (int kk = 5; kk >= 0; kk--) {
    if ((1u << kk) > sv->ctx->wave_kmax) continue;
    const int bin = kBinWave0 + kk;
    launch_align_wave(1 << kk, aa, ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), &dc->bin_off[bin],
                      &dc->bin_cnt[bin], dc->queue[bin], ln.score.as<uint32_t>(),
                      std::min<uint32_t>(Q * 64, persistent_blocks(sv->ctx, wave_lds_bytes(1 << kk))), s);
  }
  for (const int bin :
       {kBinSeg16x4, kBinSeg16x3, kBinSeg8x3, kBinSeg16x2, kBinSeg32, kBinSeg16, kBinSeg8x2}) {
    const int G = seg_lanes(bin), K = seg_words(bin);
    launch_align_seg(G, K, aa, ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), &dc->bin_off[bin],
                     &dc->bin_cnt[bin], dc->queue[bin], ln.score.as<uint32_t>(),
                     std::min<uint32_t>(Q * 64,
                                        sv->ctx->prop.multiProcessorCount * seg_blocks_per_cu(G, K)),
                     s);
  }
  for (int W = kWinMax; W >= 2; W--)
    launch_align_win(W, aa, ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), &dc->bin_off[W],
                     &dc->bin_cnt[W], dc->queue[W], ln.score.as<uint32_t>(),
                     std::min<uint32_t>(Q, persistent_blocks(sv->ctx, win_lds_bytes(W))), s);
  plain_bin_jobs(ln.pairs.as<uint2>(), ln.list.as<uint32_t>(), dc->bin_off, dc->bin_cnt,
                 ln.plain_jobs.as<uint32_t>(), s);
  launch_align_plain(aa, ln.plain_jobs.as<uint32_t>(), kServePlainBlocks * 64, sv->nw_cap,
                     ln.plain_scratch.as<uint32_t>(), ln.plain_scratch.bytes / 4, s,
                     &dc->bin_cnt[kBinPlain], kServePlainBlocks);
  plain_bin_scores(ln.list.as<uint32_t>(), dc->bin_off, dc->bin_cnt, ln.plain_jobs.as<uint32_t>(),
                   ln.score.as<uint32_t>(), s);
  hipLaunchKernelGGL(k_serve_best, dim3(grid), dim3(256), 0, s, ln.pairs.as<uint2>(),
                     (const unsigned long long*)np, ln.score.as<uint32_t>(), base,
                     ln.best.as<unsigned long long>());
  // Clusters after the previous batch's (a best can be one of its queries).
  if (sv->last_resolved) NGIA_HIP(hipStreamWaitEvent(s, sv->last_resolved, 0));
  hipLaunchKernelGGL(k_serve_result, dim3(1), dim3(1024), 0, s, ln.best.as<unsigned long long>(),
                     ln.ncand.as<uint32_t>(), Q, sv->n, pos0, sv->cluster_of.as<uint32_t>(),
                     sv->qcluster.as<uint32_t>(), ln.done.as<uint32_t>(), ln.qid.as<uint64_t>(),
                     sv->new_base, ln.out.as<ngia_assignment>());
  NGIA_HIP(hipGetLastError());
  NGIA_HIP(hipEventRecord(ln.resolved, s));
  sv->last_resolved = ln.resolved;
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
      // Room left in the session's grown part, and in this lane (free HBM).
      if (sv->pos >= sv->grow_q)
        throw HipError{NGIA_ERR_NOMEM, "serving snapshot full: grow_queries reached; "
                                       "ngia_serve_reset starts a new session"};
      const uint32_t lane_q = lane_admit(sv, ln, std::min<uint32_t>(sv->qmax, sv->grow_q - sv->pos));
      ln.count = 0;
      uint64_t words = 0, codes = 0;
      while (!pending.empty() && ln.count < lane_q) {
        Part& pt = pending.front();
        uint32_t take = 0;
        while (take < pt.count && ln.count + take < lane_q) {
          const uint32_t Lq = res_len[pt.first + take];
          const uint64_t wq = rec_words(Lq), cq = 32ull * ((Lq + 31) / 32);
          if (words + wq > sv->wmax || sv->wcur + words + wq > sv->grow_words ||
              sv->ccur + codes + cq > sv->grow_codes)
            break;
          words += wq;
          codes += cq;
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
      if (ln.count == 0) {
        if (sv->wcur + rec_words(res_len[pending.front().first]) > sv->grow_words ||
            sv->ccur + 32ull * ((res_len[pending.front().first] + 31) / 32) > sv->grow_codes)
          throw HipError{NGIA_ERR_NOMEM, "serving snapshot full: grow_residues reached; "
                                         "ngia_serve_reset starts a new session"};
        throw HipError{NGIA_ERR_ARG, "a query does not fit a serving lane"};
      }
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
    stats->grown = sv->pos;
  }
  return NGIA_OK;
}

// Wait for every lane (after a failed call some may still run).
void drain(ngia_serve* sv) {
  for (auto& ln : sv->L)
    if (ln->stream) (void)hipStreamSynchronize(ln->stream);
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
  drain(sv);
  delete sv;
}

extern "C" int ngia_serve_reset(ngia_serve* sv) {
  if (!sv) return fail(nullptr, NGIA_ERR_ARG, "null serving handle");
  if (!sv->ctx) return fail(nullptr, NGIA_ERR_STATE, "the serving handle's context is closed");
  return guarded(sv->ctx, [&] {
    NGIA_HIP(hipSetDevice(sv->ctx->device));
    drain(sv);
    hipStream_t s = sv->ctx->stream;
    const uint64_t tslots = 64ull << sv->cap_log2;
    hipLaunchKernelGGL(k_serve_unwind, dim3(persistent_blocks(sv->ctx, 0) * 4), dim3(256), 0, s,
                       sv->tables.as<unsigned long long>(), tslots, sv->n);
    NGIA_HIP(hipGetLastError());
    NGIA_HIP(hipStreamSynchronize(s));
    sv->pos = 0;
    sv->wcur = sv->ccur = 0;
    sv->last_inserted = sv->last_resolved = nullptr;
    return NGIA_OK;
  });
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
    try {
      return run_stream(sv, res, res_off, res_len, n_queries, &r, 1, first_qid, out, nullptr, stats);
    } catch (...) {
      drain(sv);
      throw;
    }
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
    try {
      return run_stream(sv, res, res_off, res_len, n_queries, reqs, n_reqs, 0, out, latency_s,
                        stats);
    } catch (...) {
      drain(sv);
      throw;
    }
  });
}
