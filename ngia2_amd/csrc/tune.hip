// tune.hip -- the autotuner of the batched form's window bins.
//
// A window bin (band window W = 2..16 words) can be aligned in five shapes:
// k_align_win<W>, one thread per pair (the most work per instruction, but a
// lone pair's T x W word-characters run serially on one lane, a bin needs
// hundreds of thousands of pairs to fill the device, and the table of W x 4
// KiB per wavefront leaves a CU two or three wavefronts at W >= 10);
// k_align_seg<16> and <32>, four and two pairs per wavefront with one word
// per lane; k_align_seg<8, 2>, eight pairs per wavefront with two words per
// lane (W <= 14); and k_align_wave<1>, a wavefront per pair.  For each W the tuner times every
// shape on synthetic pairs whose band window is exactly W (equal-length
// random sequences) at bin sizes of 16 to 262,144 pairs, and installs the
// fastest shape per (W, size class); k_bin_offsets then moves each window
// bin's pairs to the bin of its shape on the device.  Windows of 17..60
// words come in four groups, each with a shape of several words per lane
// (eight or four pairs per wavefront) and one of one word per lane (32 lanes,
// or a wavefront per pair); the tuner times both per group and size class
// the same way and installs the faster.
#include <random>

#include "capi_internal.hpp"

using namespace ngia;

namespace {

constexpr uint32_t kTuneCounts[] = {16, 64, 256, 1024, 4096, 16384, 65536, 262144};
constexpr int kNumCounts = 8;
static_assert(kNumCounts <= kTuneClasses, "size classes");
constexpr uint32_t kTunePairs = 4096;  // distinct synthetic pairs; larger bins repeat them
constexpr uint32_t kTuneList = 262144;

// Band window of a pair of lengths R >= T.
uint32_t window_of(uint32_t R, uint32_t T, float thr) {
  const Band b = band_edges(R, T, thr);
  return b.lsft / 32 + (b.rsft + 31) / 32;
}

float time_kernel(ngia_ctx* ctx, const std::function<void()>& launch, uint32_t* work) {
  hipStream_t s = ctx->stream;
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    fill_u32(work, 0, kQueueWords, s);
    NGIA_HIP(hipEventRecord(ctx->ev[6], s));
    launch();
    NGIA_HIP(hipEventRecord(ctx->ev[7], s));
    NGIA_HIP(hipEventSynchronize(ctx->ev[7]));
    best = std::min(best, event_ms(ctx->ev[6], ctx->ev[7]));
  }
  return best;
}

// Whether shape k can align a window of W words.
bool shape_fits(uint32_t k, uint32_t W) {
  return k == kShapeSeg16 ? W <= 15 : k == kShapeSeg8 ? W <= 14 : true;
}

// The groups of 17..60-word windows: the window the tuner times each at, and
// its two shapes as (lanes, words per lane), lanes 64 meaning k_align_wave<1>.
struct LongGroup {
  uint32_t W;
  int g_multi, k_multi, g_single;
};
constexpr LongGroup kLong[kLongGroups] = {{19, 8, 3, 32}, {26, 16, 2, 32}, {38, 16, 3, 64},
                                          {53, 16, 4, 64}};

}  // namespace

extern "C" int ngia_get_tuning(ngia_ctx* ctx, ngia_tuning* out) {
  if (!ctx || !out) return fail(ctx, NGIA_ERR_ARG, "null argument");
  std::memset(out, 0, sizeof(*out));
  for (int k = 0; k < kTuneClasses; k++) out->class_size[k] = ctx->route.class_size[k];
  for (int W = 0; W <= kWinMax; W++)
    for (int k = 0; k < kTuneClasses; k++) {
      out->shape[W][k] = ctx->route.shape[W][k];
      out->shape_compact[W][k] = ctx->route_compact.shape[W][k];
    }
  for (int g = 0; g < kLongGroups; g++)
    for (int k = 0; k < kTuneClasses; k++) out->long_shape[g][k] = ctx->route.lng[g][k];
  return NGIA_OK;
}

extern "C" int ngia_set_tuning(ngia_ctx* ctx, const ngia_tuning* t) {
  if (!ctx || !t) return fail(ctx, NGIA_ERR_ARG, "null argument");
  ngia::RouteTable rt{}, rc{};
  for (int k = 0; k < kTuneClasses; k++) {
    rt.class_size[k] = rc.class_size[k] = t->class_size[k];
    if (k && t->class_size[k] && t->class_size[k] <= t->class_size[k - 1])
      return fail(ctx, NGIA_ERR_ARG, "class sizes must ascend");
  }
  for (int W = 2; W <= kWinMax; W++)
    for (int k = 0; k < kTuneClasses; k++) {
      for (int tab = 0; tab < 2; tab++) {
        const uint32_t sh = tab ? t->shape_compact[W][k] : t->shape[W][k];
        if (sh >= kShapes) return fail(ctx, NGIA_ERR_ARG, "unknown shape");
        if (!shape_fits(sh, (uint32_t)W))
          return fail(ctx, NGIA_ERR_ARG, "16-lane segments hold W <= 15, 8-lane W <= 14");
        (tab ? rc : rt).shape[W][k] = (uint8_t)sh;
      }
    }
  for (int g = 0; g < kLongGroups; g++)
    for (int k = 0; k < kTuneClasses; k++) {
      const uint32_t sh = t->long_shape[g][k];
      if (sh > NGIA_LONG_SINGLE) return fail(ctx, NGIA_ERR_ARG, "unknown long-window shape");
      rt.lng[g][k] = rc.lng[g][k] = (uint8_t)sh;
    }
  ctx->route = rt;
  ctx->route_compact = rc;
  return NGIA_OK;
}

extern "C" int ngia_autotune(ngia_ctx* ctx, uint32_t identity, ngia_tune_point* points, uint32_t cap,
                             uint32_t* n_points) {
  if (!ctx) return fail(nullptr, NGIA_ERR_ARG, "null context");
  if (identity < 1 || identity > 99) return fail(ctx, NGIA_ERR_ARG, "identity should be 1-99");
  return guarded(ctx, [&] {
    NGIA_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const float thr = (float)identity / 100.0f;
    std::mt19937_64 rng(identity * 7919ull + 1);
    static const char kAmino[] = "ACDEFGHIKLMNPQRSTVWY";
    DevBuf pairs, list, score, ctr;
    pairs.ensure(8ull * kTunePairs);
    list.ensure(4ull * kTuneList);
    score.ensure(4ull * kTunePairs);
    ctr.ensure(4ull * (kQueueStride + kQueueWords));  // [0] offset, [1] count; the work queue
    std::vector<uint2> hp(kTunePairs);
    std::vector<uint32_t> hl(kTuneList);
    for (uint32_t i = 0; i < kTunePairs; i++) hp[i] = make_uint2(2 * i, 2 * i + 1);
    for (uint32_t i = 0; i < kTuneList; i++) hl[i] = i % kTunePairs;
    h2d(pairs.p, hp.data(), 8ull * kTunePairs, s);
    h2d(list.p, hl.data(), 4ull * kTuneList, s);
    uint32_t* off = ctr.as<uint32_t>();
    uint32_t* cnt = off + 1;
    uint32_t* work = off + kQueueStride;
    uint32_t np = 0;
    ngia::RouteTable route{}, route_compact{};
    for (int k = 0; k < kNumCounts; k++) route.class_size[k] = route_compact.class_size[k] = kTuneCounts[k];
    // Synthetic pairs whose band window is exactly W: lengths R >= T passing
    // the length test (the median R among T = R, R - 8, ..., R - 64); pair
    // i = (2i, 2i + 1), a random rep of length R and a target of length T,
    // its first T residues with 5% substitutions.  Returns R, or 0 when no
    // such lengths exist.
    auto synth = [&](uint32_t W, const char* alpha, uint32_t na, ngia_db& db) -> uint32_t {
      std::vector<std::pair<uint32_t, uint32_t>> cand;
      for (uint32_t R = 8; R < 0xFFFF && cand.size() < 4096; R++) {
        if (window_of(R, R, thr) > W + 1) break;
        for (uint32_t d = 0; d <= 64 && d < R; d += 8) {
          const uint32_t T = R - d;
          if (!length_rejects(R, T, thr) && window_of(R, T, thr) == W) cand.push_back({R, T});
        }
      }
      if (cand.empty()) return 0;
      const uint32_t R = cand[cand.size() / 2].first, T = cand[cand.size() / 2].second;
      const uint32_t nseq = 2 * kTunePairs;
      const uint64_t rw = rec_words(R), tw = rec_words(T);
      std::vector<uint32_t> packed((rw + tw) * kTunePairs);
      std::vector<uint64_t> woff(nseq);
      std::vector<uint8_t> res(R);
      for (uint32_t i = 0; i < kTunePairs; i++) {
        for (auto& ch : res) ch = (uint8_t)alpha[rng() % na];
        woff[2 * i] = (rw + tw) * i;
        pack_record(res.data(), R, packed.data() + woff[2 * i]);
        for (uint32_t j = 0; j < T; j++)
          if (rng() % 20 == 0) res[j] = (uint8_t)alpha[rng() % na];
        woff[2 * i + 1] = woff[2 * i] + rw;
        pack_record(res.data(), T, packed.data() + woff[2 * i + 1]);
      }
      db.ctx = ctx;
      db_upload(ctx, &db, packed.data(), packed.size(), woff.data(), nullptr, nseq);
      return R;
    };
    // Two passes: proteins (the 16-code table) and DNA, whose window kernels
    // index a compact table (at most five codes) and fit more waves per CU.
    for (int dna = 0; dna < 2; dna++)
    for (int W = 2; W <= kWinMax; W++) {
      const char* alpha = dna ? "ACGT" : kAmino;
      const uint32_t na = dna ? 4 : 20;
      ngia::RouteTable& rt = dna ? route_compact : route;
      ngia_db db;
      const uint32_t L = synth((uint32_t)W, alpha, na, db);
      if (!L) continue;
      AlignArgs aa = align_args(&db, thr);
      for (int k = 0; k < kNumCounts; k++) {
        const uint32_t n = kTuneCounts[k];
        const uint32_t zero = 0;
        h2d(off, &zero, 4, s);
        h2d(cnt, &n, 4, s);
        double ms[kShapes] = {0, 0, 0, 0, 0};
        for (uint32_t sh = kShapeWindow; sh < kShapes; sh++) {
          if (!shape_fits(sh, (uint32_t)W)) continue;
          ms[sh] = time_kernel(ctx, [&] {
            if (sh == kShapeWindow)
              launch_align_win(W, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                               score.as<uint32_t>(), persistent_blocks(ctx, win_lds_bytes(W, aa.slots != nullptr)), s);
            else if (sh == kShapeWave)
              launch_align_wave(1, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                                score.as<uint32_t>(), persistent_blocks(ctx, wave_lds_bytes(1)), s);
            else {
              const int G = sh == kShapeSeg8 ? 8 : sh == kShapeSeg16 ? 16 : 32;
              const int K = sh == kShapeSeg8 ? 2 : 1;
              launch_align_seg(G, K, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                               score.as<uint32_t>(),
                               ctx->prop.multiProcessorCount * seg_blocks_per_cu(G, K), s);
            }
          }, work);
        }
        uint32_t bestk = kShapeWindow;
        for (uint32_t sh = kShapeSeg16; sh < kShapes; sh++)
          if (ms[sh] > 0 && ms[sh] < ms[bestk]) bestk = sh;
        rt.shape[W][k] = (uint8_t)bestk;
        if (points && np < cap) {
          ngia_tune_point& pt = points[np];
          pt.width = (uint32_t)W;
          pt.dna = (uint32_t)dna;
          pt.length = L;
          pt.pairs = n;
          for (uint32_t q = 0; q < kShapes; q++) pt.ms[q] = ms[q];
        }
        np++;
      }
    }
    // Windows of 17..60 words: several words per lane or one, per group and
    // size class, on proteins (these kernels use the 16-code table for any
    // alphabet); both tables take the result.
    for (int g = 0; g < kLongGroups; g++) {
      const LongGroup& lg = kLong[g];
      ngia_db db;
      const uint32_t L = synth(lg.W, kAmino, 20, db);
      if (!L) continue;
      AlignArgs aa = align_args(&db, thr);
      for (int k = 0; k < kNumCounts; k++) {
        const uint32_t n = kTuneCounts[k];
        const uint32_t zero = 0;
        h2d(off, &zero, 4, s);
        h2d(cnt, &n, 4, s);
        double ms[kShapes] = {0, 0, 0, 0, 0};
        for (int which = 0; which < 2; which++) {
          ms[which] = time_kernel(ctx, [&] {
            if (which == 0)
              launch_align_seg(lg.g_multi, lg.k_multi, aa, pairs.as<uint2>(), list.as<uint32_t>(),
                               off, cnt, work, score.as<uint32_t>(),
                               ctx->prop.multiProcessorCount * seg_blocks_per_cu(lg.g_multi, lg.k_multi),
                               s);
            else if (lg.g_single == 64)
              launch_align_wave(1, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                                score.as<uint32_t>(), persistent_blocks(ctx, wave_lds_bytes(1)), s);
            else
              launch_align_seg(32, 1, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                               score.as<uint32_t>(),
                               ctx->prop.multiProcessorCount * seg_blocks_per_cu(32, 1), s);
          }, work);
        }
        route.lng[g][k] = route_compact.lng[g][k] =
            (uint8_t)(ms[1] < ms[0] ? NGIA_LONG_SINGLE : NGIA_LONG_MULTI);
        if (points && np < cap) {
          ngia_tune_point& pt = points[np];
          pt.width = lg.W;
          pt.dna = 0;
          pt.length = L;
          pt.pairs = n;
          for (uint32_t q = 0; q < kShapes; q++) pt.ms[q] = ms[q];
        }
        np++;
      }
    }
    NGIA_HIP(hipStreamSynchronize(s));
    ctx->route = route;
    ctx->route_compact = route_compact;
    if (n_points) *n_points = np;
    return NGIA_OK;
  });
}
