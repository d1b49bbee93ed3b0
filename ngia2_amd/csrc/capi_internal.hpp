// capi_internal.hpp -- the context and device-resident database behind the
// C-ABI handles, shared by the graph pipelines (capi.hip) and the serving
// loop (serve.hip).
#pragma once

#include <algorithm>
#include <functional>
#include <memory>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "dev_common.hpp"
#include "dev_graph.hpp"
#include "host_common.hpp"

namespace ngia {

constexpr int kSideStreams = 4;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = true;  // false: a view of another buffer's memory
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = true;
  }
  // A view of b bytes at ptr, which the buffer does not free.
  void view(void* ptr, size_t b) {
    release();
    p = ptr;
    bytes = b;
    owned = false;
  }
  // Grow-only; contents are not preserved.
  void ensure(size_t b) {
    if (b <= bytes && p) return;
    if (!owned) throw HipError{NGIA_ERR_ARG, "a view buffer cannot grow"};
    release();
    size_t want = b < 256 ? 256 : (b + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError{NGIA_ERR_NOMEM, "hipMalloc(" + std::to_string(want) + " bytes) failed: " +
                                         hipGetErrorString(e)};
    }
    bytes = want;
  }
  template <class T>
  T* as() const { return (T*)p; }
};

// Device counters of one graph round, cleared by a fill kernel.
struct Counters {
  unsigned long long total_cand;  // candidates with duplicates
  unsigned long long ecount;      // raw edges
  unsigned long long unique;      // edges after unique
  unsigned long long bin_tiles[kBins];  // reference tiles of the pairs in each bin
  unsigned long long run_tiles[kBins];  // tally: reference tiles of the blocks run
  uint32_t bin_cnt[kBins];
  uint32_t bin_off[kBins + 1];
  uint32_t cursor[kBins];
  uint32_t remap[kBins];
  // the work queues of the bins' kernels; last, so that the statistics the
  // host reads (kCounterStats bytes) stay small
  alignas(256) uint32_t queue[kBins][kQueueWords];
};
constexpr size_t kCounterStats = offsetof(Counters, queue);
// Offsets into ngia_ctx::pin.
constexpr size_t kPinOverflow = 0, kPinUnique = 8, kPinPairs = 16, kPinHeadPairs = 24,
                 kPinCnt = 64, kPinCounters = 256, kPinHeadCounters = kPinCounters + kCounterStats,
                 kPinBytes = kPinHeadCounters + kCounterStats;
static_assert(kPinCnt + 4 * kBins <= kPinCounters, "pinned slots overlap");

}  // namespace ngia

struct ngia_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
  std::string err;
  hipEvent_t ev[8] = {};
  // side streams (default priority) for the concurrent alignment bins
  hipStream_t side[ngia::kSideStreams] = {};
  hipEvent_t fork = nullptr, join[ngia::kSideStreams] = {};
  // start/end of each bin's kernel in the eager batched form
  hipEvent_t bin_ev[2 * ngia::kBins] = {};
  // a step's head: its rows ready, and its raw edges written
  hipEvent_t head_rows = nullptr, head_done = nullptr;
  uint32_t wave_kmax = 8;  // largest k_align_wave<K> usable on this device
  // shapes of the window bins by size (autotuner): the 16-code table, and
  // databases of a compact alphabet
  ngia::RouteTable route{}, route_compact{};
  // NGIA_BUCKET_WIDE=1 selects the 38-bit-key bucketing that n >= 2^26 uses
  bool bucket_wide = false;
  // NGIA_BUCKET_SMALL=1: ngia_bucket takes the one-block-per-band bucketing
  // of a step's head whenever n <= kBucketSmallMax (a check of that path)
  bool bucket_small = false;
  // fewest head records that ngia_db_step splits off (env NGIA_HEAD_MIN),
  // and the narrowest own band window of a head record, in words (env
  // NGIA_HEAD_WINDOW, at least 32: every head pair must go to a wavefront bin)
  uint32_t head_min = 256;
  uint32_t head_window = 64;
  // NGIA_SERIAL_BINS=1 runs the alignment bins one after another on the
  // context stream (each bin's kernel time is then its time alone)
  bool serial_bins = false;
  // Bins of 17-30-word windows with fewer pairs than this go to the 32-lane
  // kernel, one word per lane (env NGIA_SEG2_MIN; 0 keeps every such bin on
  // two words per lane)
  uint32_t seg2_min = ngia::kSeg2MinPairs;
  bool raw_lcs = false;  // env NGIA_RAW_LCS=1: alignment kernels store LCS lengths, not scores
  // env NGIA_FULL_ALPHABET=1: the window kernels keep the 16-code table even
  // for a compact alphabet
  bool full_alphabet = false;
  ngia_timing timing{};    // device time of the last sketch / bucket / graph
  // this device's copy of the row scratch of ngia_align_noalloc (noalloc.hip)
  uint32_t* noalloc_scratch = nullptr;
  // Pinned host staging for the small device-to-host reads of a round (bin
  // counts, counters, the edge count, the bucketing's overflow flag): a copy
  // into pageable memory is staged and waited for one at a time, about 30 us
  // each; into pinned memory the copies queue and one synchronisation covers
  // them.  Slots (bytes): kPinOverflow, kPinUnique, kPinCnt, kPinPairs,
  // kPinHeadPairs, kPinCounters, kPinHeadCounters.
  unsigned char* pin = nullptr;
  // live handles made on this context; ngia_close detaches them, so a handle
  // destroyed after its context only frees its own memory
  std::vector<ngia_db*> dbs;
  std::vector<ngia_serve*> serves;
  ~ngia_ctx() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : join)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : bin_ev)
      if (e) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (head_rows) (void)hipEventDestroy(head_rows);
    if (head_done) (void)hipEventDestroy(head_done);
    for (auto& q : side)
      if (q) (void)hipStreamDestroy(q);
    if (stream) (void)hipStreamDestroy(stream);
    if (pin) (void)hipHostFree(pin);
  }
};

struct ngia_db {
  ngia_ctx* ctx = nullptr;
  uint32_t n = 0;
  uint64_t n_words = 0;
  uint32_t max_len = 0;
  uint32_t rep_bands = 0;
  bool have_sig = false;
  ngia::DevBuf packed, word_off, lengths, codes, code_off, sig, rep;
  // compact alphabet (at most kSlotCodes nonzero codes occur): slot of every
  // residue, laid out as codes; slot_code as in AlignArgs
  bool compact = false;
  ngia::DevBuf slot_codes;
  uint32_t slot_code = 0;
  // graph workspace
  ngia::DevBuf flags, pos, sums, jobs, ekeys, escores, ekeys_tmp, escores_tmp, fin_temp;
  ngia::DevBuf counters, cnt, keep, off, pairs, score, bin_of, list, plain_scratch, plain_jobs,
      bucket_scratch;
  uint64_t n_edges = 0, n_pairs = 0;
  // the captured batched round; g_key holds the buffer addresses and the
  // routing it was captured with
  hipGraphExec_t gexec = nullptr;  // candidates and bins
  uint32_t g_identity = 0, g_bands = 0;
  std::vector<uintptr_t> g_key;
  // the alignment part, one capture per set of bins that hold pairs
  std::vector<std::pair<uint32_t, hipGraphExec_t>> gexec2;
  // per-bin figures of the last batched graph
  bool have_bins = false;
  ngia::Counters last{};
  float bin_ms[ngia::kBins] = {};
  // record lengths (host copy), for the split of a step
  std::vector<uint32_t> host_len;
  // first target of the candidate pass: the rounds of a step split the
  // targets at the head's end
  uint32_t t0 = 0;
  // the head of a step (ngia_db_step): records [0, head->n), with views of
  // this database's records, codes and signatures and its own rows and
  // round workspace
  std::unique_ptr<ngia_db> head;
  // (head only) its records' sketch in chunks of positions: (record, first
  // position) per chunk
  ngia::DevBuf sketch_chunks;
  uint32_t n_sketch_chunks = 0;
  ~ngia_db() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (auto& e : gexec2) (void)hipGraphExecDestroy(e.second);
  }
};

// Drop the captured rounds of db (its buffers or routing changed).
inline void drop_graphs(ngia_db* db) {
  if (db->gexec) (void)hipGraphExecDestroy(db->gexec);
  db->gexec = nullptr;
  for (auto& e : db->gexec2) (void)hipGraphExecDestroy(e.second);
  db->gexec2.clear();
}

// Called by ngia_close for each live serving handle (serve.hip).
void serve_detach(ngia_serve* sv);

namespace ngia {
// Device address of the current device's copy of the static row scratch of
// ngia_align_noalloc (noalloc.hip); looked up once per device.
uint32_t* noalloc_setup(int device);
}  // namespace ngia

namespace ngia {

inline int fail(ngia_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  set_thread_error(msg);
  return code;
}

template <class F>
inline int guarded(ngia_ctx* ctx, F&& f) {
  try {
    return f();
  } catch (const HipError& e) {
    return fail(ctx, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(ctx, NGIA_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ctx, NGIA_ERR_ARG, e.what());
  }
}

inline void h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes) NGIA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
}
inline void d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes) NGIA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
}

// Validate the packed records before any kernel touches them: every record
// lies inside the area, its length word matches `lengths` (when given) and is
// below 0xFFFF (the reference's lines[2048] bound).
inline void validate_records(const uint32_t* packed, uint64_t n_words, const uint64_t* word_off,
                      const uint32_t* lengths, uint32_t n, uint32_t* max_len) {
  uint32_t mx = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (word_off[i] >= n_words) throw HipError{NGIA_ERR_ARG, "record offset outside packed area"};
    const uint32_t L = packed[word_off[i]];
    if (L >= 0xFFFF) throw HipError{NGIA_ERR_ARG, "record length >= 65535"};
    if (word_off[i] + rec_words(L) > n_words)
      throw HipError{NGIA_ERR_ARG, "record extends past packed area"};
    if (lengths && lengths[i] != L) throw HipError{NGIA_ERR_ARG, "lengths[] disagrees with record"};
    mx = std::max(mx, L);
  }
  *max_len = mx;
}

inline float event_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  NGIA_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

inline uint32_t persistent_blocks(const ngia_ctx* ctx, size_t lds_bytes) {
  const uint32_t cu = (uint32_t)std::max(1, ctx->prop.multiProcessorCount);
  uint32_t per_cu = lds_bytes ? (uint32_t)std::min<size_t>(16, (160 * 1024) / lds_bytes) : 16;
  if (per_cu < 1) per_cu = 1;
  return cu * per_cu;
}

inline void db_upload(ngia_ctx* ctx, ngia_db* db, const uint32_t* packed, uint64_t n_words,
               const uint64_t* word_off, const uint32_t* lengths, uint32_t n) {
  std::vector<uint32_t> len_v;
  if (!lengths) {
    len_v.resize(n);
    for (uint32_t i = 0; i < n; i++) len_v[i] = packed[word_off[i]];
    lengths = len_v.data();
  }
  validate_records(packed, n_words, word_off, lengths, n, &db->max_len);
  db->host_len.assign(lengths, lengths + n);
  db->head.reset();
  db->n = n;
  db->n_words = n_words;
  hipStream_t s = ctx->stream;
  db->packed.ensure(4 * std::max<uint64_t>(n_words, 1));
  db->word_off.ensure(8ull * std::max<uint32_t>(n, 1));
  db->lengths.ensure(4ull * std::max<uint32_t>(n, 1));
  h2d(db->packed.p, packed, 4 * n_words, s);
  h2d(db->word_off.p, word_off, 8ull * n, s);
  h2d(db->lengths.p, lengths, 4ull * n, s);
  // byte codes: 32 bytes per 32-residue block, sequence i at code_off[i]
  std::vector<uint64_t> coff(std::max<uint32_t>(n, 1));
  uint64_t acc = 0;
  for (uint32_t i = 0; i < n; i++) {
    coff[i] = acc;
    acc += 32ull * ((lengths[i] + 31) / 32);
  }
  db->code_off.ensure(8ull * std::max<uint32_t>(n, 1));
  db->codes.ensure(acc + 64);
  h2d(db->code_off.p, coff.data(), 8ull * n, s);
  launch_codes(db->packed.as<uint32_t>(), db->word_off.as<uint64_t>(), db->code_off.as<uint64_t>(),
               n, db->codes.as<uint8_t>(), s);
  // Which codes occur.  With at most kSlotCodes nonzero ones (DNA: A C G T,
  // and N, which the code table reads as asparagine), the
  // window kernels index their match-mask table by slot (code 0 -> slot 0, the
  // present codes in ascending order -> 1..), a table of kSlotCodes + 1
  // entries per word in place of 16.
  db->compact = false;
  db->slot_codes.release();
  if (acc && !(ctx->full_alphabet)) {
    DevBuf present;
    present.ensure(4);
    fill_u32(present.as<uint32_t>(), 0, 1, s);
    launch_code_set(db->codes.as<uint8_t>(), acc, present.as<uint32_t>(), s);
    uint32_t bits = 0;
    d2h(&bits, present.p, 4, s);
    NGIA_HIP(hipStreamSynchronize(s));
    if (__builtin_popcount(bits & ~1u) <= kSlotCodes) {
      uint64_t lut = 0;
      uint32_t slot = 0, slot_code = 0;
      for (uint32_t c = 1; c < 16; c++) {
        if (!((bits >> c) & 1u)) continue;
        slot++;
        lut |= (uint64_t)slot << (4 * c);
        slot_code |= c << (4 * (slot - 1));
      }
      db->slot_codes.ensure(acc + 64);
      launch_code_slots(db->codes.as<uint8_t>(), acc, lut, db->slot_codes.as<uint8_t>(), s);
      db->compact = true;
      db->slot_code = slot_code;
    }
  }
  NGIA_HIP(hipStreamSynchronize(s));
  db->have_sig = false;
  db->rep_bands = 0;
  drop_graphs(db);
}

inline AlignArgs align_args(const ngia_db* db, float thr) {
  AlignArgs a;
  a.packed = db->packed.as<uint32_t>();
  a.word_off = db->word_off.as<uint64_t>();
  a.codes = db->codes.as<uint8_t>();
  a.code_off = db->code_off.as<uint64_t>();
  a.lengths = db->lengths.as<uint32_t>();
  a.thr = thr;
  a.raw_lcs = db->ctx && db->ctx->raw_lcs ? 1u : 0u;
  if (db->compact) {
    a.slots = db->slot_codes.as<uint8_t>();
    a.slot_code = db->slot_code;
  }
  return a;
}

}  // namespace ngia
