// capi.hip -- the C-ABI of include/ngia2.h: context, device-resident
// database, and the plain / batched graph pipelines.
//
// Plain form (buildGraph.hpp:274-350): per band, candidates on the device,
// the job count copied to the host, one alignment launch, a host sync, the
// passing jobs appended to the raw edge list; then one sort+unique.
// Batched form: every band's candidates at once, deduplicated per target,
// binned by band-window width, aligned by the length-specialised persistent
// kernels, passing pairs appended to the raw edge list -- all captured once
// in a hipGraph and replayed; then the same sort+unique.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "capi_internal.hpp"

using namespace ngia;

namespace {

// Capacity of the raw edge list: every candidate of every band, duplicates included.
void ensure_edge_space(ngia_db* db, uint64_t m) {
  m = std::max<uint64_t>(m, 1);
  db->ekeys.ensure(8 * m);
  db->escores.ensure(4 * m);
  db->ekeys_tmp.ensure(8 * m);
  db->escores_tmp.ensure(4 * m);
  db->fin_temp.ensure(finalize_scratch_bytes(m));
}

void finalize(ngia_ctx* ctx, ngia_db* db, uint64_t m) {
  Counters* dc = db->counters.as<Counters>();
  finalize_edges(db->ekeys.as<uint64_t>(), db->escores.as<uint32_t>(), m, db->n,
                 db->ekeys_tmp.as<uint64_t>(), db->escores_tmp.as<uint32_t>(), db->fin_temp.p,
                 db->fin_temp.bytes, &dc->unique, ctx->stream);
  unsigned long long* u = (unsigned long long*)(ctx->pin + kPinUnique);
  d2h(u, &dc->unique, 8, ctx->stream);
  NGIA_HIP(hipStreamSynchronize(ctx->stream));
  db->n_edges = *u;
}

// Workspace of the plain form: per-band flags, scan, job triplets and the
// alignment row scratch.
void ensure_plain_space(ngia_db* db) {
  const uint32_t n = db->n;
  db->flags.ensure(4ull * n);
  db->pos.ensure(8ull * (n + 1));
  db->sums.ensure(8 * scan_scratch_words(n));
  db->jobs.ensure(12ull * n);
  db->plain_scratch.ensure(4 * align_plain_scratch_words((db->max_len + 31) / 32));
}

// One band of the plain form up to the alignment (buildGraph.hpp:274-333):
// the band's job triplets (rep[i], i, 0) for ascending i with isJob, their
// count read back by the host (a stream synchronisation), then one
// kernel_dynamic-contract launch scoring them in place.  Returns the count.
uint64_t plain_band(ngia_ctx* ctx, ngia_db* db, uint32_t band, float thr, const uint32_t* parent,
                    float* ms_cand) {
  hipStream_t s = ctx->stream;
  const uint32_t n = db->n;
  const uint32_t* rep = db->rep.as<uint32_t>() + (uint64_t)band * n;
  NGIA_HIP(hipEventRecord(ctx->ev[0], s));
  band_jobs(rep, db->lengths.as<uint32_t>(), n, thr, parent, db->flags.as<uint32_t>(),
            db->pos.as<uint64_t>(), db->sums.as<uint64_t>(), db->jobs.as<uint32_t>(), s);
  uint64_t count = 0;
  d2h(&count, db->pos.as<uint64_t>() + n, 8, s);
  NGIA_HIP(hipEventRecord(ctx->ev[1], s));
  NGIA_HIP(hipStreamSynchronize(s));  // the host needs the count to size the launch
  if (ms_cand) *ms_cand += event_ms(ctx->ev[0], ctx->ev[1]);
  if (count == 0) return 0;
  const uint32_t nw_max = (db->max_len + 31) / 32;
  launch_align_plain(align_args(db, thr), db->jobs.as<uint32_t>(), (uint32_t)count, nw_max,
                     db->plain_scratch.as<uint32_t>(), align_plain_scratch_words(nw_max), s);
  return count;
}

void graph_plain(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                 ngia_graph_stats* st, const uint32_t* parent /* fast mode */) {
  hipStream_t s = ctx->stream;
  const uint32_t n = db->n;
  const float thr = (float)identity / 100.0f;
  db->have_bins = false;
  ensure_plain_space(db);
  db->counters.ensure(sizeof(Counters));
  if (!parent) ensure_edge_space(db, (uint64_t)n * bands);
  fill_u32(db->counters.as<uint32_t>(), 0, sizeof(Counters) / 4, s);
  Counters* dc = db->counters.as<Counters>();
  uint64_t total = 0;
  float ms_cand = 0, ms_align = 0, ms_edges = 0;
  uint32_t launches = 0, syncs = 0;
  for (uint32_t b = 0; b < bands; b++) {
    const uint64_t count = plain_band(ctx, db, b, thr, parent, &ms_cand);
    syncs++;
    if (count == 0) continue;
    total += count;
    launches++;
    NGIA_HIP(hipEventRecord(ctx->ev[2], s));
    NGIA_HIP(hipStreamSynchronize(s));  // buildGraph.hpp:333
    syncs++;
    if (parent) {
      fast_apply(db->jobs.as<uint32_t>(), (uint32_t)count, const_cast<uint32_t*>(parent), s);
    } else {
      jobs_to_edges(db->jobs.as<uint32_t>(), (uint32_t)count, db->ekeys.as<uint64_t>(),
                    db->escores.as<uint32_t>(), &dc->ecount, s);
    }
    NGIA_HIP(hipEventRecord(ctx->ev[3], s));
    ms_align += event_ms(ctx->ev[1], ctx->ev[2]);
    NGIA_HIP(hipEventSynchronize(ctx->ev[3]));
    ms_edges += event_ms(ctx->ev[2], ctx->ev[3]);
  }
  if (st) {
    st->total_pairs = total;
    st->distinct_pairs = total;
    st->ms_candidates = ms_cand;
    st->ms_align = ms_align;
    st->align_launches = launches;
    st->host_syncs = syncs;
  }
  if (parent) return;
  unsigned long long ecount = 0;
  d2h(&ecount, &dc->ecount, 8, s);
  NGIA_HIP(hipEventRecord(ctx->ev[4], s));
  NGIA_HIP(hipStreamSynchronize(s));
  finalize(ctx, db, ecount);
  NGIA_HIP(hipEventRecord(ctx->ev[5], s));
  NGIA_HIP(hipEventSynchronize(ctx->ev[5]));
  const float ms_fin = event_ms(ctx->ev[4], ctx->ev[5]);
  ctx->timing.graph_ms = ms_cand + ms_align + ms_edges;
  ctx->timing.finalize_ms = ms_fin;
  ms_edges += ms_fin;
  if (st) {
    st->ms_edges = ms_edges;
    st->n_edges = db->n_edges;
  }
}

// A batched round runs in two captured parts.  The first finds the distinct
// candidate pairs and bins them; the host then reads the bin counts, and the
// second launches only the alignment kernels whose bins hold pairs.  An empty
// bin's kernel would still have to place all its persistent blocks, and a
// wide window kernel (55-68 KiB of LDS per block) waits for CUs the busy bins
// hold, stalling every kernel queued behind it on its stream.

// Part 1: candidates and bins (capturable: fixed grids, device-side counts),
// on stream s; targets from db->t0 on.
void enqueue_bins(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands, hipStream_t s) {
  const uint32_t n = db->n;
  const float thr = (float)identity / 100.0f;
  Counters* dc = db->counters.as<Counters>();
  const uint32_t* len = db->lengths.as<uint32_t>();
  fill_u32(db->counters.as<uint32_t>(), 0, sizeof(Counters) / 4, s);
  pair_count(db->rep.as<uint32_t>(), len, n, db->t0, bands, thr, db->cnt.as<uint32_t>(),
             db->keep.as<uint64_t>(), &dc->total_cand, s);
  scan_u32(db->cnt.as<uint32_t>(), n, db->off.as<uint64_t>(), db->sums.as<uint64_t>(), s);
  pair_emit(db->rep.as<uint32_t>(), n, bands, db->keep.as<uint64_t>(), db->off.as<uint64_t>(),
            db->pairs.as<uint2>(), s);
  const uint64_t* npairs = db->off.as<uint64_t>() + n;
  const uint32_t grid = persistent_blocks(ctx, 0) * 4;
  bin_pairs(db->pairs.as<uint2>(), npairs, len, thr, db->bin_of.as<uint32_t>(), dc->bin_cnt,
            dc->bin_off, dc->cursor, db->list.as<uint32_t>(), dc->bin_tiles, db->score.as<uint32_t>(),
            ctx->wave_kmax, dc->remap, db->compact ? ctx->route_compact : ctx->route, ctx->seg2_min,
            grid, s);
}

// Part 2: the alignment kernels of the bins in `mask` (bit b: bin b holds
// pairs), then the passing pairs as raw edges (capturable), on stream s and,
// unless `serial`, the side streams.
void enqueue_align(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t mask,
                   bool time_align, bool tally, hipStream_t s, bool serial = false) {
  const uint32_t n = db->n;
  const float thr = (float)identity / 100.0f;
  Counters* dc = db->counters.as<Counters>();
  const uint64_t* npairs = db->off.as<uint64_t>() + n;
  const uint32_t grid = persistent_blocks(ctx, 0) * 4;
  const AlignArgs a0 = align_args(db, thr);
  auto holds = [&](int bin) { return ((mask >> bin) & 1u) != 0; };
  // The arguments of bin b's kernel: with the tally on, it counts the
  // reference tiles of the blocks it ran into run_tiles[b].
  auto args = [&](int bin) {
    AlignArgs a = a0;
    if (tally) a.tally = &dc->run_tiles[bin];
    return a;
  };
  // The bins run concurrently, forked onto the side streams and joined back.
  // Every wave bin below K = 16 is latency bound (one pair per wavefront, the
  // bin lasts as long as its longest pair), so each gets a side stream of its
  // own and starts at the fork: K = 8 on side stream 0, K = 4 on 1, K = 2 on
  // 2, K = 1 on 3.  The window bins (throughput bound) that hold pairs follow
  // round-robin on side streams 0, 3 and 2, widest first.  K = 32 and 16 (128
  // and 64 KiB of LDS per block), when they hold pairs, run before the fork:
  // launched beside the others, their blocks would wait for a CU with that
  // much free LDS, which the other bins' persistent blocks hold until they
  // finish.
  if (time_align) NGIA_HIP(hipEventRecord(ctx->ev[2], s));
  const bool fork = !ctx->serial_bins && !serial;
  auto mark = [&](int bin, int end, hipStream_t q) {
    if (time_align) NGIA_HIP(hipEventRecord(ctx->bin_ev[2 * bin + end], q));
  };
  auto wave = [&](int k, hipStream_t q) {
    const int bin = kBinWave0 + k;
    if ((1u << k) > ctx->wave_kmax || !holds(bin)) return;
    mark(bin, 0, q);
    launch_align_wave(1 << k, args(bin), db->pairs.as<uint2>(), db->list.as<uint32_t>(), &dc->bin_off[bin],
                      &dc->bin_cnt[bin], dc->queue[bin], db->score.as<uint32_t>(),
                      persistent_blocks(ctx, wave_lds_bytes(1 << k)), q);
    mark(bin, 1, q);
  };
  wave(5, s);
  wave(4, s);
  if (fork) {
    NGIA_HIP(hipEventRecord(ctx->fork, s));
    for (int k = 0; k < kSideStreams; k++) NGIA_HIP(hipStreamWaitEvent(ctx->side[k], ctx->fork, 0));
  }
  for (int k = 3; k >= 0; k--) wave(k, fork ? ctx->side[3 - k] : s);
  // The segmented bins (several pairs per wavefront) follow K = 8 and K = 1.
  auto seg = [&](int bin, hipStream_t q) {
    if (!holds(bin)) return;
    mark(bin, 0, q);
    const int G = seg_lanes(bin), K = seg_words(bin);
    launch_align_seg(G, K, args(bin), db->pairs.as<uint2>(), db->list.as<uint32_t>(),
                     &dc->bin_off[bin], &dc->bin_cnt[bin], dc->queue[bin], db->score.as<uint32_t>(),
                     ctx->prop.multiProcessorCount * seg_blocks_per_cu(G, K), q);
    mark(bin, 1, q);
  };
  seg(kBinSeg16x4, fork ? ctx->side[2] : s);
  seg(kBinSeg16x3, fork ? ctx->side[3] : s);
  seg(kBinSeg8x3, fork ? ctx->side[1] : s);
  seg(kBinSeg8x2, fork ? ctx->side[2] : s);
  seg(kBinSeg16x2, fork ? ctx->side[0] : s);
  seg(kBinSeg32, fork ? ctx->side[0] : s);
  seg(kBinSeg16, fork ? ctx->side[3] : s);
  static constexpr int kWinStreams[3] = {0, 3, 2};
  int slot = 0;
  for (int W = kWinMax; W >= 2; W--) {
    if (!holds(W)) continue;
    hipStream_t q = fork ? ctx->side[kWinStreams[slot++ % 3]] : s;
    mark(W, 0, q);
    launch_align_win(W, args(W), db->pairs.as<uint2>(), db->list.as<uint32_t>(), &dc->bin_off[W],
                     &dc->bin_cnt[W], dc->queue[W], db->score.as<uint32_t>(),
                     persistent_blocks(ctx, win_lds_bytes(W, a0.slots != nullptr)), q);
    mark(W, 1, q);
  }
  for (int k = 0; fork && k < kSideStreams; k++) {
    NGIA_HIP(hipEventRecord(ctx->join[k], ctx->side[k]));
    NGIA_HIP(hipStreamWaitEvent(s, ctx->join[k], 0));
  }
  if (time_align) NGIA_HIP(hipEventRecord(ctx->ev[3], s));
  // windows beyond 64*32 words: the plain kernel over job triplets
  if (holds(kBinPlain)) {
    plain_bin_jobs(db->pairs.as<uint2>(), db->list.as<uint32_t>(), dc->bin_off, dc->bin_cnt,
                   db->plain_jobs.as<uint32_t>(), s);
    const uint32_t nw_max = (db->max_len + 31) / 32;
    mark(kBinPlain, 0, s);
    launch_align_plain(args(kBinPlain), db->plain_jobs.as<uint32_t>(), 0xFFFFFFFFu, nw_max,
                       db->plain_scratch.as<uint32_t>(), db->plain_scratch.bytes / 4, s,
                       &dc->bin_cnt[kBinPlain], kPlainBinBlocks);
    mark(kBinPlain, 1, s);
    plain_bin_scores(db->list.as<uint32_t>(), dc->bin_off, dc->bin_cnt,
                     db->plain_jobs.as<uint32_t>(), db->score.as<uint32_t>(), s);
  }
  pairs_to_edges(db->pairs.as<uint2>(), npairs, db->score.as<uint32_t>(), db->ekeys.as<uint64_t>(),
                 db->escores.as<uint32_t>(), &dc->ecount, grid, s);
}

// Capture what fn enqueues on stream s into an instantiated graph.
template <class F>
hipGraphExec_t capture(hipStream_t s, F fn) {
  hipGraph_t g = nullptr;
  NGIA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    fn();
  } catch (...) {
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  NGIA_HIP(hipStreamEndCapture(s, &g));
  hipGraphExec_t ge = nullptr;
  const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  NGIA_HIP(e);
  return ge;
}

// Workspace of one batched round over db (sized for every target).
void ensure_batched_space(ngia_db* db, uint32_t bands) {
  const uint32_t n = db->n;
  const uint64_t cap = std::max<uint64_t>((uint64_t)n * bands, 1);  // distinct <= candidates <= n*bands
  db->counters.ensure(sizeof(Counters));
  db->off.ensure(8ull * (n + 1));
  db->cnt.ensure(4ull * n);
  db->keep.ensure(8ull * n);
  db->sums.ensure(8 * scan_scratch_words(n));
  db->pairs.ensure(8 * cap);
  db->score.ensure(4 * cap);
  db->bin_of.ensure(4 * cap);
  db->list.ensure(4 * cap);
  ensure_edge_space(db, cap);
}

// The plain bin holds only windows wider than 64 x 32 words (records of
// tens of thousands of residues at low identity), so its job list and row
// scratch are allocated the first time a round's bin holds pairs, not with
// the rest of the workspace: the row scratch alone is kPlainBinBlocks x 64
// threads x ceil(R/32) words.  Returns true when either buffer moved.
bool ensure_plain_bin_space(ngia_db* db, uint32_t bands) {
  const void* j0 = db->plain_jobs.p;
  const void* s0 = db->plain_scratch.p;
  const uint64_t cap = std::max<uint64_t>((uint64_t)db->n * bands, 1);
  db->plain_jobs.ensure(12 * cap);
  db->plain_scratch.ensure(4 * align_plain_scratch_words((db->max_len + 31) / 32, kPlainBinBlocks));
  return db->plain_jobs.p != j0 || db->plain_scratch.p != s0;
}

// The head round of a step (ngia_db_step), whose raw edges join the main
// round's before the sort: its database and the event recorded after them.
struct HeadRound {
  ngia_db* db = nullptr;
  hipEvent_t done = nullptr;
};

// The batched round: its first part captured once per (identity, bands), its
// second once per set of bins that hold pairs, both replayed (or launched
// eagerly, with the alignment phase timed).  With a head round, the main
// round's targets start at db->t0 and the head's raw edges are appended
// before the sort.
void graph_batched(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                   ngia_graph_stats* st, bool eager, bool tally, const HeadRound* head = nullptr) {
  hipStream_t s = ctx->stream;
  const uint32_t n = db->n;
  ensure_batched_space(db, bands);
  // Everything the captured kernels were given: a replay is valid while none
  // of it changed (a grown buffer moves, a new tuning changes the routing).
  std::vector<uintptr_t> key;
  for (const DevBuf* b : {&db->packed, &db->word_off, &db->lengths, &db->codes, &db->code_off,
                          &db->rep, &db->counters, &db->cnt, &db->keep, &db->off, &db->sums, &db->pairs,
                          &db->score, &db->bin_of, &db->list, &db->ekeys, &db->escores})
    key.push_back((uintptr_t)b->p);
  {
    const ngia::RouteTable& rt = db->compact ? ctx->route_compact : ctx->route;
    for (uint32_t w : rt.class_size) key.push_back(w);
    for (int W = 0; W <= kWinMax; W++)
      for (int k = 0; k < kTuneClasses; k++) key.push_back(rt.shape[W][k]);
    for (int g = 0; g < kLongGroups; g++)
      for (int k = 0; k < kTuneClasses; k++) key.push_back(rt.lng[g][k]);
  }
  key.push_back(db->t0);
  key.push_back(ctx->seg2_min);
  const bool reuse = db->gexec && db->g_identity == identity && db->g_bands == bands &&
                     db->g_key == key;
  if (!eager && !reuse) {
    drop_graphs(db);
    db->gexec = capture(s, [&] { enqueue_bins(ctx, db, identity, bands, s); });
    db->g_identity = identity;
    db->g_bands = bands;
    db->g_key = key;
  }
  NGIA_HIP(hipEventRecord(ctx->ev[0], s));
  if (eager)
    enqueue_bins(ctx, db, identity, bands, s);
  else
    NGIA_HIP(hipGraphLaunch(db->gexec, s));
  // the bins that hold pairs pick the second part
  uint32_t cnt[kBins];
  d2h(ctx->pin + kPinCnt, db->counters.as<Counters>()->bin_cnt, sizeof(cnt), s);
  NGIA_HIP(hipStreamSynchronize(s));
  std::memcpy(cnt, ctx->pin + kPinCnt, sizeof(cnt));
  uint32_t mask = 0;
  for (int b = 0; b < kBins; b++)
    if (cnt[b]) mask |= 1u << b;
  if (((mask >> kBinPlain) & 1u) && ensure_plain_bin_space(db, bands)) {
    // captures that launch the plain bin hold the old buffers
    std::vector<std::pair<uint32_t, hipGraphExec_t>> keep;
    for (auto& e : db->gexec2) {
      if ((e.first >> kBinPlain) & 1u)
        (void)hipGraphExecDestroy(e.second);
      else
        keep.push_back(e);
    }
    db->gexec2.swap(keep);
  }
  if (eager) {
    // Same kernels without the graph, with events around the alignment phase.
    enqueue_align(ctx, db, identity, mask, /*time_align=*/true, tally, s);
  } else {
    hipGraphExec_t g2 = nullptr;
    for (const auto& e : db->gexec2)
      if (e.first == mask) g2 = e.second;
    if (!g2) {
      if (db->gexec2.size() >= 16) {  // bounded: start afresh when the bins keep changing
        for (auto& e : db->gexec2) (void)hipGraphExecDestroy(e.second);
        db->gexec2.clear();
      }
      g2 = capture(s, [&] { enqueue_align(ctx, db, identity, mask, false, false, s); });
      db->gexec2.emplace_back(mask, g2);
    }
    NGIA_HIP(hipGraphLaunch(g2, s));
  }
  NGIA_HIP(hipEventRecord(ctx->ev[1], s));
  Counters hc, hh{};
  uint64_t npairs = 0, hpairs = 0;
  d2h(ctx->pin + kPinCounters, db->counters.p, kCounterStats, s);
  d2h(ctx->pin + kPinPairs, db->off.as<uint64_t>() + n, 8, s);
  if (head) {
    NGIA_HIP(hipStreamWaitEvent(s, head->done, 0));
    d2h(ctx->pin + kPinHeadCounters, head->db->counters.p, kCounterStats, s);
    d2h(ctx->pin + kPinHeadPairs, head->db->off.as<uint64_t>() + head->db->n, 8, s);
  }
  NGIA_HIP(hipStreamSynchronize(s));
  std::memcpy(&hc, ctx->pin + kPinCounters, kCounterStats);
  std::memcpy(&npairs, ctx->pin + kPinPairs, 8);
  if (head) {
    std::memcpy(&hh, ctx->pin + kPinHeadCounters, kCounterStats);
    std::memcpy(&hpairs, ctx->pin + kPinHeadPairs, 8);
  }
  db->n_pairs = npairs;
  NGIA_HIP(hipEventRecord(ctx->ev[4], s));
  if (hh.ecount) {
    // the head's pairs (targets below db->t0) are disjoint from the main
    // round's; their raw edges follow the main round's
    NGIA_HIP(hipMemcpyAsync(db->ekeys.as<uint64_t>() + hc.ecount, head->db->ekeys.p, 8 * hh.ecount,
                            hipMemcpyDeviceToDevice, s));
    NGIA_HIP(hipMemcpyAsync(db->escores.as<uint32_t>() + hc.ecount, head->db->escores.p,
                            4 * hh.ecount, hipMemcpyDeviceToDevice, s));
  }
  finalize(ctx, db, hc.ecount + hh.ecount);
  NGIA_HIP(hipEventRecord(ctx->ev[5], s));
  NGIA_HIP(hipEventSynchronize(ctx->ev[5]));
  ctx->timing.graph_ms = event_ms(ctx->ev[0], ctx->ev[1]);
  ctx->timing.finalize_ms = event_ms(ctx->ev[4], ctx->ev[5]);
  ctx->timing.align_ms = eager ? event_ms(ctx->ev[2], ctx->ev[3]) : 0.0;
  db->last = hc;
  db->have_bins = true;
  uint32_t launches = 0;
  for (int b = 0; b < kBins; b++) {
    const bool aligner = b == kBinPlain || is_seg_bin(b) ||
                         (b >= 2 && b <= kWinMax) ||
                         (b >= kBinWave0 && b < kBinPlain && (1u << (b - kBinWave0)) <= ctx->wave_kmax);
    const bool launched = aligner && ((mask >> b) & 1u);
    launches += launched;
    db->bin_ms[b] = eager && launched ? event_ms(ctx->bin_ev[2 * b], ctx->bin_ev[2 * b + 1]) : 0.f;
  }
  uint64_t tiles = 0;
  for (int b = 0; b < kBins; b++) tiles += hc.bin_tiles[b] + hh.bin_tiles[b];
  if (st) {
    st->total_pairs = hc.total_cand + hh.total_cand;
    st->distinct_pairs = npairs + hpairs;
    st->n_edges = db->n_edges;
    st->ref_tiles = tiles;
    st->ms_candidates = 0;
    st->ms_align = event_ms(ctx->ev[0], ctx->ev[1]);
    st->ms_edges = event_ms(ctx->ev[4], ctx->ev[5]);
    st->align_launches = launches;
    st->host_syncs = 3;
  }
}

// The database's representative rows from its signatures: one block per band
// (no scratch, no host synchronisation) up to kBucketSmallMax records, the
// partitioned or full sort beyond (or always, under NGIA_BUCKET_WIDE).
void bucket_rows(ngia_ctx* ctx, ngia_db* db, hipStream_t s) {
  if (!ctx->bucket_wide && db->n <= kBucketSmallMax)
    launch_bucket_small(db->sig.as<uint32_t>(), db->n, db->rep.as<uint32_t>(), s);
  else
    launch_bucket(db->sig.as<uint32_t>(), db->n, db->rep.as<uint32_t>(), db->bucket_scratch.p, s,
                  ctx->bucket_wide, db->compact, (uint32_t*)(ctx->pin + kPinOverflow));
}

// The head of a step: the leading run of records whose own band window (the
// record against itself) is at least min_window words (64 by default, at
// least 32) at thr, in non-increasing length order, at most kHeadMax of
// them.  A pair (r, i) of the head then has R >= T and a window at least
// that of (T, T): every head pair goes to a wavefront bin, to the segmented
// bins of 32-60 words (a min_window below 61) or, past the widest, to the
// plain bin.  With 64 the head holds the pairs of K >= 2, the
// long latency-bound ones; with 32 (the K = 1 pairs too) the heavy-tailed
// step measured the same, and the log-normal set's 8,192-record head at
// -i 50 lost 1%; with 128, 11.6 against 11.3 ms on the heavy-tailed set.
constexpr uint32_t kHeadMax = 8192;
uint32_t head_records(const ngia_db* db, float thr, uint32_t min_window) {
  uint32_t h = 0;
  const uint32_t cap = std::min<uint32_t>(db->n, kHeadMax);
  while (h < cap) {
    const uint32_t L = db->host_len[h];
    if (h && L > db->host_len[h - 1]) break;
    const Band b = band_edges(L, L, thr);
    if (b.lsft / 32 + (b.rsft + 31) / 32 < min_window) break;
    h++;
  }
  return h;
}

// Positions per wavefront of the head's chunked sketch.
constexpr uint32_t kSketchChunk = 1024;
// Fewer head records than this are not split off (NGIA_HEAD_MIN overrides
// it, for tests on small sets): the split's own launches and waits cost about
// 0.2 ms (50 head records of the 10,000-protein set at -i 50: 1.60 against
// 1.35 ms per step).
constexpr uint32_t kHeadMin = 256;

// The head database of a step: records [0, h) of db, sharing its records,
// codes and signatures (views), with rows, round workspace and the chunk
// list of its sketch of its own.
ngia_db* step_head(ngia_ctx* ctx, ngia_db* db, uint32_t h) {
  if (!db->head || db->head->n != h) {
    db->head = std::make_unique<ngia_db>();
    std::vector<uint2> ch;
    for (uint32_t i = 0; i < h; i++) {
      const uint32_t L = db->host_len[i];
      if (L == 0) continue;  // no window: its signature stays all ones
      const uint32_t first = L >= 8 ? 7u : L - 1;
      for (uint32_t p = 0; p < L; p += kSketchChunk)
        ch.push_back(make_uint2(i, std::max(p, first)));
    }
    ngia_db* hd = db->head.get();
    hd->sketch_chunks.ensure(8 * std::max<size_t>(ch.size(), 1));
    if (!ch.empty())
      NGIA_HIP(hipMemcpy(hd->sketch_chunks.p, ch.data(), 8 * ch.size(), hipMemcpyHostToDevice));
    hd->n_sketch_chunks = (uint32_t)ch.size();
  }
  ngia_db* hd = db->head.get();
  hd->ctx = ctx;
  hd->n = h;
  hd->n_words = db->n_words;
  hd->max_len = *std::max_element(db->host_len.begin(), db->host_len.begin() + h);
  hd->packed.view(db->packed.p, db->packed.bytes);
  hd->word_off.view(db->word_off.p, 8ull * h);
  hd->lengths.view(db->lengths.p, 4ull * h);
  hd->codes.view(db->codes.p, db->codes.bytes);
  hd->code_off.view(db->code_off.p, 8ull * h);
  hd->sig.view(db->sig.p, 256ull * h);
  hd->compact = db->compact;
  hd->slot_code = db->slot_code;
  if (db->compact)
    hd->slot_codes.view(db->slot_codes.p, db->slot_codes.bytes);
  else
    hd->slot_codes.release();
  hd->have_sig = true;
  hd->rep_bands = NGIA_BANDS;
  hd->t0 = 0;
  return hd;
}

// Whether a head pair's window can exceed the widest wavefront kernel's:
// with R >= T >= R * thr (shorter targets are length-rejected), the window
// is at most lsft(R)/32 + ceil((R - R thr^2 + 32)/32), one word of slack.
bool head_may_need_plain(const ngia_ctx* ctx, uint32_t max_len, float thr) {
  const double R = max_len;
  const double ws = std::ceil(R * (1.0 - thr) / 32.0) + std::ceil((R * (1.0 - (double)thr * thr) + 33.0) / 32.0) + 1.0;
  const double K = ctx->wave_kmax;
  return ws + K - 1.0 > 64.0 * K;
}

// Sketch, bucketing and the batched graph of the whole database, with the
// head's graph beside the rest's sketch.  The pairs (r, i), r < i, of targets
// i < h depend on the first h records alone (a record's representative in a
// band precedes it in the stable (signature, index) order, so bucketing the
// first h signatures gives the full bucketing's rows for them), and the main
// round takes targets i >= h with the full rows: the two rounds' pairs are
// disjoint and together every pair of the one-round graph.
//
// Side stream 1, with no host synchronisation: the head's signatures (one
// wavefront per chunk of 1,024 positions: a few thousand long records, one
// wavefront each, would leave most SIMDs idle), its rows (one block per
// band), its candidates and its wavefront bins, longest first.  Context
// stream, beside it: the rest's sketch, then after the head's signatures the
// full bucketing, the main round from target h on, the head's raw edges, and
// the sort.  The head's bins are latency bound (a few long pairs, one
// wavefront each) and run on the SIMDs the VALU-bound sketch leaves between
// its instructions.
void graph_step(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands, bool split,
                ngia_graph_stats* st) {
  hipStream_t s = ctx->stream;
  const uint32_t n = db->n;
  const float thr = (float)identity / 100.0f;
  db->sig.ensure(256ull * n);
  db->rep.ensure(256ull * n);
  db->bucket_scratch.ensure(bucket_scratch_bytes(n, ctx->bucket_wide));
  uint32_t h = split ? head_records(db, thr, ctx->head_window) : 0u;
  if (h < ctx->head_min || h >= n) h = 0;  // too few to pay for the split, or nothing beside them
  const uint8_t* codes = db->codes.as<uint8_t>();
  const uint64_t* coff = db->code_off.as<uint64_t>();
  const uint32_t* len = db->lengths.as<uint32_t>();
  uint32_t* sig = db->sig.as<uint32_t>();
  NGIA_HIP(hipEventRecord(ctx->ev[6], s));
  HeadRound hr;
  if (h) {
    ngia_db* hd = step_head(ctx, db, h);
    hd->rep.ensure(256ull * h);
    ensure_batched_space(hd, bands);
    uint32_t mask = 0;
    for (uint32_t k = 0; (1u << k) <= ctx->wave_kmax; k++) mask |= 1u << (kBinWave0 + k);
    // a head window below 61 words lets head pairs into the 3- and 4-word bins
    if (ctx->head_window < 61) mask |= (1u << kBinSeg16x3) | (1u << kBinSeg16x4);
    if (head_may_need_plain(ctx, hd->max_len, thr)) {
      ensure_plain_bin_space(hd, bands);
      mask |= 1u << kBinPlain;
    }
    // the rest's sketch first, so that the context stream is busy while the
    // host enqueues the head's round
    launch_sketch(codes, coff + h, len + h, n - h, sig + 64ull * h, s);
    hipStream_t q = ctx->side[1];
    NGIA_HIP(hipStreamWaitEvent(q, ctx->ev[6], 0));  // after the previous step's use
    fill_u32(sig, 0xFFFFFFFFu, 64ull * h, q);
    launch_sketch_chunks(codes, coff, len, hd->sketch_chunks.as<uint2>(), hd->n_sketch_chunks,
                         kSketchChunk, sig, q);
    NGIA_HIP(hipEventRecord(ctx->head_rows, q));
    launch_bucket_small(sig, h, hd->rep.as<uint32_t>(), q);
    enqueue_bins(ctx, hd, identity, bands, q);
    enqueue_align(ctx, hd, identity, mask, false, false, q, /*serial=*/true);
    NGIA_HIP(hipEventRecord(ctx->head_done, q));
    hr.db = hd;
    hr.done = ctx->head_done;
    NGIA_HIP(hipStreamWaitEvent(s, ctx->head_rows, 0));
  } else {
    launch_sketch(codes, coff, len, n, sig, s);
  }
  bucket_rows(ctx, db, s);
  db->have_sig = true;
  db->rep_bands = NGIA_BANDS;
  NGIA_HIP(hipEventRecord(ctx->ev[7], s));
  db->t0 = h;
  try {
    graph_batched(ctx, db, identity, bands, st, false, false, h ? &hr : nullptr);
  } catch (...) {
    db->t0 = 0;
    throw;
  }
  db->t0 = 0;
  ctx->timing.step_head = h;
}

}  // namespace

// ====================================================================== ABI
extern "C" int ngia_open(int device, ngia_ctx** out) {
  if (!out) return fail(nullptr, NGIA_ERR_ARG, "null ctx pointer");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(nullptr, NGIA_ERR_NODEV, "no HIP device");
  if (device < 0 || device >= count) return fail(nullptr, NGIA_ERR_ARG, "device out of range");
  auto ctx = std::make_unique<ngia_ctx>();
  ctx->device = device;
  int rc = guarded(ctx.get(), [&] {
    NGIA_HIP(hipSetDevice(device));
    NGIA_HIP(hipGetDeviceProperties(&ctx->prop, device));
    ctx->wave_kmax = (uint32_t)align_kernels_setup(ctx->prop.sharedMemPerBlock);
    ctx->noalloc_scratch = noalloc_setup(device);
    const char* bw = std::getenv("NGIA_BUCKET_WIDE");
    ctx->bucket_wide = bw && bw[0] == '1';
    const char* hw = std::getenv("NGIA_HEAD_WINDOW");
    ctx->head_window = hw ? std::max<uint32_t>(32u, (uint32_t)std::strtoul(hw, nullptr, 10)) : 64u;
    const char* hm = std::getenv("NGIA_HEAD_MIN");
    ctx->head_min = hm ? (uint32_t)std::strtoul(hm, nullptr, 10) : kHeadMin;
    if (ctx->head_min < 1) ctx->head_min = 1;
    const char* bs = std::getenv("NGIA_BUCKET_SMALL");
    ctx->bucket_small = bs && bs[0] == '1';
    const char* s2 = std::getenv("NGIA_SEG2_MIN");
    ctx->seg2_min = s2 ? (uint32_t)std::strtoul(s2, nullptr, 10) : kSeg2MinPairs;
    const char* sb = std::getenv("NGIA_SERIAL_BINS");
    ctx->serial_bins = sb && sb[0] == '1';
    const char* rl = std::getenv("NGIA_RAW_LCS");
    ctx->raw_lcs = rl && rl[0] == '1';
    const char* fa = std::getenv("NGIA_FULL_ALPHABET");
    ctx->full_alphabet = fa && fa[0] == '1';
    NGIA_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    NGIA_HIP(hipHostMalloc((void**)&ctx->pin, kPinBytes, hipHostMallocDefault));
    std::memset(ctx->pin, 0, kPinBytes);
    for (auto& e : ctx->ev) NGIA_HIP(hipEventCreate(&e));
    for (auto& q : ctx->side) NGIA_HIP(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    NGIA_HIP(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
    for (auto& e : ctx->join) NGIA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto& e : ctx->bin_ev) NGIA_HIP(hipEventCreate(&e));
    NGIA_HIP(hipEventCreateWithFlags(&ctx->head_rows, hipEventDisableTiming));
    NGIA_HIP(hipEventCreateWithFlags(&ctx->head_done, hipEventDisableTiming));
    return NGIA_OK;
  });
  if (rc != NGIA_OK) return rc;
  *out = ctx.release();
  return NGIA_OK;
}

extern "C" void ngia_close(ngia_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (ngia_db* db : ctx->dbs) db->ctx = nullptr;
  for (ngia_serve* sv : ctx->serves) serve_detach(sv);
  delete ctx;
}

extern "C" const char* ngia_last_error(const ngia_ctx* ctx) {
  return ctx ? ctx->err.c_str() : thread_error();
}

extern "C" int ngia_device_query(ngia_ctx* ctx, ngia_device_info* info) {
  if (!ctx || !info) return fail(ctx, NGIA_ERR_ARG, "null argument");
  return guarded(ctx, [&] {
    std::memset(info, 0, sizeof(*info));
    info->device = ctx->device;
    info->cu_count = ctx->prop.multiProcessorCount;
    info->clock_khz = ctx->prop.clockRate;
    info->warp_size = ctx->prop.warpSize;
    size_t fr = 0, tot = 0;
    NGIA_HIP(hipMemGetInfo(&fr, &tot));
    info->hbm_free = fr;
    info->hbm_total = tot;
    std::snprintf(info->name, sizeof(info->name), "%s", ctx->prop.name);
    std::snprintf(info->arch, sizeof(info->arch), "%s", ctx->prop.gcnArchName);
    return NGIA_OK;
  });
}

extern "C" int ngia_get_timing(ngia_ctx* ctx, ngia_timing* out) {
  if (!ctx || !out) return fail(ctx, NGIA_ERR_ARG, "null argument");
  *out = ctx->timing;
  return NGIA_OK;
}

extern "C" int ngia_sync(ngia_ctx* ctx) {
  if (!ctx) return fail(ctx, NGIA_ERR_ARG, "null ctx");
  return guarded(ctx, [&] {
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    return NGIA_OK;
  });
}

extern "C" int ngia_db_create(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                              const uint64_t* word_off, const uint32_t* lengths, uint32_t n,
                              ngia_db** out) {
  if (!ctx || !out || (n && (!packed || !word_off))) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (n >= 0x7FFFFFFFu) return fail(ctx, NGIA_ERR_ARG, "too many sequences");
  *out = nullptr;
  auto db = std::make_unique<ngia_db>();
  db->ctx = ctx;
  int rc = guarded(ctx, [&] {
    db_upload(ctx, db.get(), packed, n_words, word_off, lengths, n);
    return NGIA_OK;
  });
  if (rc != NGIA_OK) return rc;
  *out = db.release();
  ctx->dbs.push_back(*out);
  return NGIA_OK;
}

extern "C" void ngia_db_destroy(ngia_db* db) {
  if (!db) return;
  if (db->ctx) {
    (void)hipSetDevice(db->ctx->device);
    (void)hipStreamSynchronize(db->ctx->stream);
    auto& v = db->ctx->dbs;
    v.erase(std::remove(v.begin(), v.end(), db), v.end());
  }
  delete db;
}

extern "C" int ngia_db_sketch(ngia_ctx* ctx, ngia_db* db) {
  if (!ctx || !db) return fail(ctx, NGIA_ERR_ARG, "null argument");
  return guarded(ctx, [&] {
    db->sig.ensure(256ull * std::max<uint32_t>(db->n, 1));
    NGIA_HIP(hipEventRecord(ctx->ev[6], ctx->stream));
    launch_sketch(db->codes.as<uint8_t>(), db->code_off.as<uint64_t>(), db->lengths.as<uint32_t>(),
                  db->n, db->sig.as<uint32_t>(), ctx->stream);
    NGIA_HIP(hipEventRecord(ctx->ev[7], ctx->stream));
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timing.sketch_ms = event_ms(ctx->ev[6], ctx->ev[7]);
    db->have_sig = true;
    return NGIA_OK;
  });
}

extern "C" int ngia_db_bucket(ngia_ctx* ctx, ngia_db* db) {
  if (!ctx || !db) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (!db->have_sig) return fail(ctx, NGIA_ERR_STATE, "bucket before sketch");
  return guarded(ctx, [&] {
    db->rep.ensure(256ull * std::max<uint32_t>(db->n, 1));
    db->bucket_scratch.ensure(bucket_scratch_bytes(db->n, ctx->bucket_wide));
    NGIA_HIP(hipEventRecord(ctx->ev[6], ctx->stream));
    bucket_rows(ctx, db, ctx->stream);
    NGIA_HIP(hipEventRecord(ctx->ev[7], ctx->stream));
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timing.bucket_ms = event_ms(ctx->ev[6], ctx->ev[7]);
    db->rep_bands = NGIA_BANDS;
    return NGIA_OK;
  });
}

extern "C" int ngia_db_set_reps(ngia_ctx* ctx, ngia_db* db, const uint32_t* rep_rows,
                                uint32_t bands) {
  if (!ctx || !db || !rep_rows) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (bands < 1 || bands > NGIA_BANDS) return fail(ctx, NGIA_ERR_ARG, "bands should be 1-64");
  const uint64_t m = (uint64_t)bands * db->n;
  for (uint64_t k = 0; k < m; k++)
    if (rep_rows[k] >= db->n) return fail(ctx, NGIA_ERR_ARG, "representative index out of range");
  return guarded(ctx, [&] {
    db->rep.ensure(256ull * std::max<uint32_t>(db->n, 1));
    h2d(db->rep.p, rep_rows, 4 * m, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    db->rep_bands = bands;
    drop_graphs(db);
    return NGIA_OK;
  });
}

extern "C" int ngia_db_get_sigs(ngia_ctx* ctx, ngia_db* db, uint32_t* sig) {
  if (!ctx || !db || !sig) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (!db->have_sig) return fail(ctx, NGIA_ERR_STATE, "no signatures");
  return guarded(ctx, [&] {
    d2h(sig, db->sig.p, 256ull * db->n, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    return NGIA_OK;
  });
}

extern "C" int ngia_db_get_reps(ngia_ctx* ctx, ngia_db* db, uint32_t* rep) {
  if (!ctx || !db || !rep) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (!db->rep_bands) return fail(ctx, NGIA_ERR_STATE, "no representative rows");
  return guarded(ctx, [&] {
    d2h(rep, db->rep.p, 4ull * db->rep_bands * db->n, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    return NGIA_OK;
  });
}

extern "C" int ngia_db_graph(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                             uint32_t flags, ngia_graph_stats* st) {
  if (!ctx || !db) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (identity < 1 || identity > 99) return fail(ctx, NGIA_ERR_ARG, "identity should be 1-99");
  if (bands < 1 || bands > NGIA_BANDS) return fail(ctx, NGIA_ERR_ARG, "bands should be 1-64");
  if (bands > db->rep_bands) return fail(ctx, NGIA_ERR_STATE, "fewer representative rows than bands");
  if (db->n == 0) return fail(ctx, NGIA_ERR_ARG, "empty database");
  if (st) std::memset(st, 0, sizeof(*st));
  return guarded(ctx, [&] {
    auto t0 = std::chrono::steady_clock::now();
    if (flags == NGIA_GRAPH_BATCHED || flags == NGIA_GRAPH_BATCHED_EAGER ||
        flags == NGIA_GRAPH_BATCHED_TALLY)
      graph_batched(ctx, db, identity, bands, st, flags != NGIA_GRAPH_BATCHED,
                    flags == NGIA_GRAPH_BATCHED_TALLY);
    else
      graph_plain(ctx, db, identity, bands, st, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    if (st) st->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return NGIA_OK;
  });
}

extern "C" int ngia_db_step(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                            uint32_t flags, ngia_graph_stats* st) {
  if (!ctx || !db) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (identity < 1 || identity > 99) return fail(ctx, NGIA_ERR_ARG, "identity should be 1-99");
  if (bands < 1 || bands > NGIA_BANDS) return fail(ctx, NGIA_ERR_ARG, "bands should be 1-64");
  if (db->n == 0) return fail(ctx, NGIA_ERR_ARG, "empty database");
  if (flags & ~NGIA_STEP_NO_SPLIT) return fail(ctx, NGIA_ERR_ARG, "unknown step flags");
  if (st) std::memset(st, 0, sizeof(*st));
  return guarded(ctx, [&] {
    auto t0 = std::chrono::steady_clock::now();
    graph_step(ctx, db, identity, bands, !(flags & NGIA_STEP_NO_SPLIT), st);
    ctx->timing.step_ms = event_ms(ctx->ev[6], ctx->ev[5]);
    auto t1 = std::chrono::steady_clock::now();
    if (st) st->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
    return NGIA_OK;
  });
}

extern "C" int ngia_db_band_jobs(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t band,
                                 uint32_t* jobs, uint64_t cap, uint64_t* count) {
  if (!ctx || !db || !count) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (identity < 1 || identity > 99) return fail(ctx, NGIA_ERR_ARG, "identity should be 1-99");
  if (band >= db->rep_bands) return fail(ctx, NGIA_ERR_STATE, "no representative row for band");
  if (db->n == 0) return fail(ctx, NGIA_ERR_ARG, "empty database");
  *count = 0;
  return guarded(ctx, [&] {
    ensure_plain_space(db);
    db->have_bins = false;
    const uint64_t c = plain_band(ctx, db, band, (float)identity / 100.0f, nullptr, nullptr);
    *count = c;
    if (!jobs || c == 0) return NGIA_OK;
    if (cap < c) return fail(ctx, NGIA_ERR_ARG, "job buffer too small");
    d2h(jobs, db->jobs.p, 12 * c, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    return NGIA_OK;
  });
}

extern "C" int ngia_db_get_edges(ngia_ctx* ctx, ngia_db* db, ngia_edge* out, uint64_t cap,
                                 uint64_t* n_edges) {
  if (!ctx || !db || !n_edges) return fail(ctx, NGIA_ERR_ARG, "null argument");
  *n_edges = db->n_edges;
  if (!out) return NGIA_OK;
  if (cap < db->n_edges) return fail(ctx, NGIA_ERR_ARG, "edge buffer too small");
  return guarded(ctx, [&] {
    const uint64_t m = db->n_edges;
    std::vector<uint64_t> k(m);
    std::vector<uint32_t> sc(m);
    d2h(k.data(), db->ekeys.p, 8 * m, ctx->stream);
    d2h(sc.data(), db->escores.p, 4 * m, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t e = 0; e < m; e++) {
      out[e].u = (int32_t)(uint32_t)(k[e] >> 32);
      out[e].v = (int32_t)(uint32_t)k[e];
      out[e].w = (float)sc[e] / 100.0f;  // buildGraph.hpp:346
    }
    return NGIA_OK;
  });
}

extern "C" int ngia_db_get_pair_scores(ngia_ctx* ctx, ngia_db* db, uint64_t* keys,
                                       uint32_t* scores, uint64_t cap, uint64_t* n_pairs) {
  if (!ctx || !db || !n_pairs) return fail(ctx, NGIA_ERR_ARG, "null argument");
  *n_pairs = db->n_pairs;
  if (!keys && !scores) return NGIA_OK;
  if (cap < db->n_pairs) return fail(ctx, NGIA_ERR_ARG, "pair buffer too small");
  return guarded(ctx, [&] {
    const uint64_t m = db->n_pairs;
    std::vector<uint2> pr(m);
    std::vector<uint32_t> sc(m);
    d2h(pr.data(), db->pairs.p, 8 * m, ctx->stream);
    d2h(sc.data(), db->score.p, 4 * m, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<uint64_t> order(m);
    for (uint64_t i = 0; i < m; i++) order[i] = ((uint64_t)pr[i].x << 32) | pr[i].y;
    std::vector<uint64_t> idx(m);
    for (uint64_t i = 0; i < m; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return order[a] < order[b]; });
    for (uint64_t i = 0; i < m; i++) {
      if (keys) keys[i] = order[idx[i]];
      if (scores) scores[i] = sc[idx[i]];
    }
    return NGIA_OK;
  });
}

extern "C" int ngia_db_bin_stats(ngia_ctx* ctx, ngia_db* db, ngia_bin_stat* out, uint32_t cap,
                                 uint32_t* n_bins) {
  if (!ctx || !db || !n_bins) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (!db->have_bins) return fail(ctx, NGIA_ERR_STATE, "no batched graph has run on this database");
  *n_bins = kBins;
  if (!out) return NGIA_OK;
  if (cap < (uint32_t)kBins) return fail(ctx, NGIA_ERR_ARG, "bin buffer too small");
  for (int b = 0; b < kBins; b++) {
    ngia_bin_stat& o = out[b];
    if (b == kBinZero) {
      o.kind = NGIA_BIN_ZERO;
      o.width = 0;
    } else if (b <= kWinMax) {
      o.kind = NGIA_BIN_WINDOW;
      o.width = (uint32_t)b;
    } else if (b < kBinPlain) {
      o.kind = NGIA_BIN_WAVE;
      o.width = 1u << (b - kBinWave0);
    } else if (b == kBinPlain) {
      o.kind = NGIA_BIN_PLAIN;
      o.width = 0;
    } else if (b >= kBinSeg16x2) {
      o.kind = NGIA_BIN_SEG2 + (uint32_t)(seg_words(b) - 2);
      o.width = (uint32_t)seg_lanes(b);
    } else {
      o.kind = NGIA_BIN_SEG;
      o.width = b == kBinSeg16 ? 16u : 32u;
    }
    o.pairs = db->last.bin_cnt[b];
    o.ref_tiles = db->last.bin_tiles[b];
    o.ms = db->bin_ms[b];
    o.run_tiles = db->last.run_tiles[b];
  }
  return NGIA_OK;
}

extern "C" int ngia_db_fast(ngia_ctx* ctx, ngia_db* db, uint32_t identity, uint32_t bands,
                            uint32_t* cluster, uint64_t* total_pairs) {
  if (!ctx || !db || !cluster) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (identity < 1 || identity > 99) return fail(ctx, NGIA_ERR_ARG, "identity should be 1-99");
  if (bands < 1 || bands > db->rep_bands) return fail(ctx, NGIA_ERR_ARG, "bad band count");
  return guarded(ctx, [&] {
    const uint32_t n = db->n;
    DevBuf parent;
    parent.ensure(4ull * std::max<uint32_t>(n, 1));
    fill_u32(parent.as<uint32_t>(), 0xFFFFFFFFu, n, ctx->stream);
    ngia_graph_stats st{};
    graph_plain(ctx, db, identity, bands, &st, parent.as<uint32_t>());
    std::vector<uint32_t> par(n);
    d2h(par.data(), parent.p, 4ull * n, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    // fastClustering.hpp:338-347: parent[i] < i, resolve in ascending order
    for (uint32_t i = 0; i < n; i++) cluster[i] = par[i] == 0xFFFFFFFFu ? i : cluster[par[i]];
    if (total_pairs) *total_pairs = st.total_pairs;
    return NGIA_OK;
  });
}

// ------------------------------------------------------ host-pointer seams
extern "C" int ngia_sketch(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                           const uint64_t* word_off, uint32_t n, uint32_t* sig) {
  if (!ctx || !sig || (n && (!packed || !word_off))) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (n == 0) return NGIA_OK;
  ngia_db* db = nullptr;
  int rc = ngia_db_create(ctx, packed, n_words, word_off, nullptr, n, &db);
  if (rc == NGIA_OK) rc = ngia_db_sketch(ctx, db);
  if (rc == NGIA_OK) rc = ngia_db_get_sigs(ctx, db, sig);
  ngia_db_destroy(db);
  return rc;
}

extern "C" int ngia_bucket(ngia_ctx* ctx, const uint32_t* sig, uint32_t n, uint32_t* rep) {
  if (!ctx || !sig || !rep) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (n == 0) return NGIA_OK;
  return guarded(ctx, [&] {
    DevBuf dsig, drep, scratch;
    dsig.ensure(256ull * n);
    drep.ensure(256ull * n);
    scratch.ensure(bucket_scratch_bytes(n, ctx->bucket_wide));
    h2d(dsig.p, sig, 256ull * n, ctx->stream);
    if (ctx->bucket_small && n <= kBucketSmallMax)
      launch_bucket_small(dsig.as<uint32_t>(), n, drep.as<uint32_t>(), ctx->stream);
    else
      launch_bucket(dsig.as<uint32_t>(), n, drep.as<uint32_t>(), scratch.p, ctx->stream,
                    ctx->bucket_wide);
    d2h(rep, drep.p, 256ull * n, ctx->stream);
    NGIA_HIP(hipStreamSynchronize(ctx->stream));
    return NGIA_OK;
  });
}

extern "C" int ngia_align(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                          const uint64_t* word_off, uint32_t n_seqs, uint32_t* jobs,
                          uint32_t count, float threshold) {
  if (!ctx || !jobs || !packed || !word_off) return fail(ctx, NGIA_ERR_ARG, "null argument");
  for (uint64_t j = 0; j < count; j++)
    if (jobs[3 * j] >= n_seqs || jobs[3 * j + 1] >= n_seqs)
      return fail(ctx, NGIA_ERR_ARG, "job index out of range");
  if (count == 0) return NGIA_OK;
  ngia_db* db = nullptr;
  int rc = ngia_db_create(ctx, packed, n_words, word_off, nullptr, n_seqs, &db);
  if (rc != NGIA_OK) return rc;
  rc = guarded(ctx, [&] {
    hipStream_t s = ctx->stream;
    DevBuf djobs;
    djobs.ensure(12ull * count);
    h2d(djobs.p, jobs, 12ull * count, s);
    const uint32_t nw_max = (db->max_len + 31) / 32;
    const uint64_t sw = align_plain_scratch_words(nw_max);
    db->plain_scratch.ensure(4 * sw);
    launch_align_plain(align_args(db, threshold), djobs.as<uint32_t>(), count, nw_max,
                       db->plain_scratch.as<uint32_t>(), sw, s);
    d2h(jobs, djobs.p, 12ull * count, s);
    NGIA_HIP(hipStreamSynchronize(s));
    return NGIA_OK;
  });
  ngia_db_destroy(db);
  return rc;
}

extern "C" int ngia_build_graph(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                                const uint64_t* word_off, const uint32_t* lengths, uint32_t n,
                                const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                                uint32_t flags, ngia_edge** edges, uint64_t* n_edges,
                                ngia_graph_stats* stats) {
  if (!ctx || !edges || !n_edges || !rep_rows) return fail(ctx, NGIA_ERR_ARG, "null argument");
  *edges = nullptr;
  *n_edges = 0;
  ngia_db* db = nullptr;
  int rc = ngia_db_create(ctx, packed, n_words, word_off, lengths, n, &db);
  if (rc == NGIA_OK) rc = ngia_db_set_reps(ctx, db, rep_rows, bands);
  if (rc == NGIA_OK) rc = ngia_db_graph(ctx, db, identity, bands, flags, stats);
  if (rc == NGIA_OK) {
    uint64_t m = 0;
    ngia_db_get_edges(ctx, db, nullptr, 0, &m);
    ngia_edge* e = (ngia_edge*)std::malloc(sizeof(ngia_edge) * std::max<uint64_t>(m, 1));
    if (!e) rc = fail(ctx, NGIA_ERR_NOMEM, "host allocation failed");
    if (rc == NGIA_OK) rc = ngia_db_get_edges(ctx, db, e, m, &m);
    if (rc == NGIA_OK) {
      *edges = e;
      *n_edges = m;
    } else {
      std::free(e);
    }
  }
  ngia_db_destroy(db);
  return rc;
}
