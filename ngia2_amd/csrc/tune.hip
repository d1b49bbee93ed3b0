// tune.hip -- the autotuner of the batched form's small window bins.
//
// A window bin (band window W = 2..16 words) is aligned by k_align_win<W>,
// one thread per pair: the most work per instruction, but a lone pair takes
// T x W serial word-characters on one lane.  k_align_wave<1> gives each pair
// a wavefront, which finishes a handful of pairs much sooner.  For each W the
// tuner times both kernels on synthetic pairs whose band window is exactly W
// (equal-length random sequences) at several bin sizes and records the
// smallest bin size at which the window kernel is no slower; smaller bins
// are then routed to the wavefront kernel by k_bin_offsets.
#include <random>

#include "capi_internal.hpp"

using namespace ngia;

namespace {

constexpr uint32_t kTuneCounts[] = {16, 64, 256, 1024, 4096};
constexpr int kNumCounts = 5;
constexpr uint32_t kTunePairs = 4096;

// Band window of a pair of lengths R >= T.
uint32_t window_of(uint32_t R, uint32_t T, float thr) {
  const Band b = band_edges(R, T, thr);
  return b.lsft / 32 + (b.rsft + 31) / 32;
}

float time_kernel(ngia_ctx* ctx, const std::function<void()>& launch, uint32_t* work) {
  hipStream_t s = ctx->stream;
  float best = 1e30f;
  for (int rep = 0; rep < 3; rep++) {
    fill_u32(work, 0, 1, s);
    NGIA_HIP(hipEventRecord(ctx->ev[6], s));
    launch();
    NGIA_HIP(hipEventRecord(ctx->ev[7], s));
    NGIA_HIP(hipEventSynchronize(ctx->ev[7]));
    best = std::min(best, event_ms(ctx->ev[6], ctx->ev[7]));
  }
  return best;
}

}  // namespace

extern "C" int ngia_get_tuning(ngia_ctx* ctx, ngia_tuning* out) {
  if (!ctx || !out) return fail(ctx, NGIA_ERR_ARG, "null argument");
  for (int W = 0; W <= kWinMax; W++) out->route_below[W] = ctx->route.route_below[W];
  return NGIA_OK;
}

extern "C" int ngia_set_tuning(ngia_ctx* ctx, const ngia_tuning* t) {
  if (!ctx || !t) return fail(ctx, NGIA_ERR_ARG, "null argument");
  for (int W = 0; W <= kWinMax; W++) ctx->route.route_below[W] = W >= 2 ? t->route_below[W] : 0u;
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
    list.ensure(4ull * kTunePairs);
    score.ensure(4ull * kTunePairs);
    ctr.ensure(64);  // [0] offset, [1] count, [2] work
    std::vector<uint2> hp(kTunePairs);
    std::vector<uint32_t> hl(kTunePairs);
    for (uint32_t i = 0; i < kTunePairs; i++) {
      hp[i] = make_uint2(2 * i, 2 * i + 1);
      hl[i] = i;
    }
    h2d(pairs.p, hp.data(), 8ull * kTunePairs, s);
    h2d(list.p, hl.data(), 4ull * kTunePairs, s);
    uint32_t* off = ctr.as<uint32_t>();
    uint32_t* cnt = off + 1;
    uint32_t* work = off + 2;
    uint32_t np = 0;
    ngia::RouteTable route{};
    for (int W = 2; W <= kWinMax; W++) {
      // lengths R >= T passing the length test with window exactly W: the
      // median R among T = R, R - 8, ..., R - 64
      std::vector<std::pair<uint32_t, uint32_t>> cand;
      for (uint32_t R = 8; R < 0xFFFF && cand.size() < 4096; R++) {
        if (window_of(R, R, thr) > (uint32_t)W + 1) break;
        for (uint32_t d = 0; d <= 64 && d < R; d += 8) {
          const uint32_t T = R - d;
          if (!length_rejects(R, T, thr) && window_of(R, T, thr) == (uint32_t)W) cand.push_back({R, T});
        }
      }
      if (cand.empty()) continue;
      const uint32_t R = cand[cand.size() / 2].first, T = cand[cand.size() / 2].second;
      const uint32_t L = R;
      // pair i = (2i, 2i + 1): a random rep of length R and a target of
      // length T, its first T residues with 5% substitutions
      const uint32_t nseq = 2 * kTunePairs;
      const uint64_t rw = rec_words(R), tw = rec_words(T);
      std::vector<uint32_t> packed((rw + tw) * kTunePairs);
      std::vector<uint64_t> woff(nseq);
      std::vector<uint8_t> res(R);
      for (uint32_t i = 0; i < kTunePairs; i++) {
        for (auto& ch : res) ch = (uint8_t)kAmino[rng() % 20];
        woff[2 * i] = (rw + tw) * i;
        pack_record(res.data(), R, packed.data() + woff[2 * i]);
        for (uint32_t j = 0; j < T; j++)
          if (rng() % 20 == 0) res[j] = (uint8_t)kAmino[rng() % 20];
        woff[2 * i + 1] = woff[2 * i] + rw;
        pack_record(res.data(), T, packed.data() + woff[2 * i + 1]);
      }
      ngia_db db;
      db.ctx = ctx;
      db_upload(ctx, &db, packed.data(), packed.size(), woff.data(), nullptr, nseq);
      AlignArgs aa = align_args(&db, thr);
      uint32_t route_below = 0;
      bool found = false;
      for (int k = 0; k < kNumCounts; k++) {
        const uint32_t n = kTuneCounts[k];
        const uint32_t zero = 0;
        h2d(off, &zero, 4, s);
        h2d(cnt, &n, 4, s);
        const float win_ms = time_kernel(ctx, [&] {
          launch_align_win(W, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                           score.as<uint32_t>(), persistent_blocks(ctx, win_lds_bytes(W)), s);
        }, work);
        const float wave_ms = time_kernel(ctx, [&] {
          launch_align_wave(1, aa, pairs.as<uint2>(), list.as<uint32_t>(), off, cnt, work,
                            score.as<uint32_t>(), persistent_blocks(ctx, wave_lds_bytes(1)), s);
        }, work);
        if (!found && win_ms <= wave_ms) {
          route_below = n;
          found = true;
        }
        if (points && np < cap) points[np] = ngia_tune_point{(uint32_t)W, L, n, win_ms, wave_ms};
        np++;
      }
      // the window kernel lost at every timed size: route up to 4x the largest
      route.route_below[W] = found ? route_below : 4 * kTuneCounts[kNumCounts - 1];
    }
    NGIA_HIP(hipStreamSynchronize(s));
    ctx->route = route;
    if (n_points) *n_points = np;
    return NGIA_OK;
  });
}
