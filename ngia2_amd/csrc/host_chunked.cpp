// host_chunked.cpp -- edge emission for databases larger than one device
// holds.  The reference keeps the whole packed area in managed memory and
// lets the driver page it in and out (README.md:8, buildGraph.hpp:143, 164).
// Here the records are cut into blocks of consecutive indices, and every
// candidate pair (rep r, target t; buildGraph.hpp:262-266) is aligned in the
// chunk made of r's block and t's block -- the only records uploaded for it.
// Each chunk is an ordinary device-resident database (ngia_db_*) whose rows
// keep the candidates of that block pair and no others; the chunks' edges,
// mapped back to global indices, merge into the sorted unique list that
// ngia_build_graph returns for the whole database.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <string>
#include <vector>

#include "host_common.hpp"

using namespace ngia;

namespace {

int fail_msg(int code, const std::string& msg) {
  set_thread_error(msg);
  return code;
}

// Device bytes a chunk needs for a batched graph, per record and per packed
// word: the packed words, one byte code per residue (about eight per word)
// with the byte-code offsets, signatures and rows (256 + 256 bytes), and the
// pair workspaces, bands possible pairs per record at about 48 bytes each.
constexpr uint64_t kBytesPerWord = 4 + 8;
inline uint64_t bytes_per_record(uint32_t bands) { return 8 + 12 + 512 + 48ull * bands; }

}  // namespace

extern "C" int ngia_plan_blocks(const uint32_t* lengths, uint32_t n, uint64_t block_words,
                                uint32_t block_records, uint32_t* bounds, uint32_t* n_blocks) {
  if ((!lengths && n) || !bounds || !n_blocks) return fail_msg(NGIA_ERR_ARG, "null argument");
  if (block_words == 0) return fail_msg(NGIA_ERR_ARG, "block_words must be positive");
  uint32_t k = 0;
  bounds[0] = 0;
  uint64_t w = 0;
  for (uint32_t i = 0; i < n; i++) {
    const uint64_t ri = rec_words(lengths[i]);
    const bool full = w + ri > block_words || (block_records && i - bounds[k] >= block_records);
    if (i > bounds[k] && full) {
      bounds[++k] = i;
      w = 0;
    }
    w += ri;
  }
  if (n) bounds[++k] = n;
  *n_blocks = k;
  return NGIA_OK;
}

extern "C" int ngia_build_graph_chunked(ngia_ctx* ctx, const uint32_t* packed, uint64_t n_words,
                                        const uint64_t* word_off, const uint32_t* lengths,
                                        uint32_t n, const uint32_t* rep_rows, uint32_t bands,
                                        uint32_t identity, uint64_t block_words,
                                        ngia_edge** edges, uint64_t* n_edges,
                                        ngia_graph_stats* stats, uint32_t* n_chunks) {
  if (!ctx || !packed || !word_off || !rep_rows || !edges || !n_edges)
    return fail_msg(NGIA_ERR_ARG, "null argument");
  *edges = nullptr;
  *n_edges = 0;
  if (n_chunks) *n_chunks = 0;
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return fail_msg(NGIA_ERR_ARG, "empty database");
  if (identity < 1 || identity > 99) return fail_msg(NGIA_ERR_ARG, "identity should be 1-99");
  if (bands < 1 || bands > NGIA_BANDS) return fail_msg(NGIA_ERR_ARG, "bands should be 1-64");
  const auto t0 = std::chrono::steady_clock::now();
  // Lengths and record bounds, checked before any index is used.
  std::vector<uint32_t> len(n);
  for (uint32_t i = 0; i < n; i++) {
    if (word_off[i] >= n_words) return fail_msg(NGIA_ERR_ARG, "word offset out of range");
    len[i] = lengths ? lengths[i] : packed[word_off[i]];
    if (word_off[i] + rec_words(len[i]) > n_words)
      return fail_msg(NGIA_ERR_ARG, "record runs past the packed area");
  }
  // Blocks: given, or sized so that two blocks take at most half the free
  // device memory (one block when the whole database fits that).
  uint64_t bw = block_words;
  uint32_t brec = 0;
  if (bw == 0) {
    ngia_device_info info;
    int rc = ngia_device_query(ctx, &info);
    if (rc != NGIA_OK) return rc;
    const uint64_t budget = info.hbm_free / 4;  // per block
    bw = std::max<uint64_t>(budget / kBytesPerWord, 1);
    brec = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(budget / bytes_per_record(bands), 1),
                                        0xFFFFFFFFu);
  }
  std::vector<uint32_t> bounds(n + 1);
  uint32_t nb = 0;
  int rc = ngia_plan_blocks(len.data(), n, bw, brec, bounds.data(), &nb);
  if (rc != NGIA_OK) return rc;
  if (nb == 1) {  // the whole database at once, from the caller's arrays
    rc = ngia_build_graph(ctx, packed, n_words, word_off, lengths, n, rep_rows, bands, identity,
                          NGIA_GRAPH_BATCHED, edges, n_edges, stats);
    if (rc == NGIA_OK && n_chunks) *n_chunks = 1;
    return rc;
  }
  for (uint64_t k = 0; k < (uint64_t)bands * n; k++)
    if (rep_rows[k] >= n) return fail_msg(NGIA_ERR_ARG, "representative index out of range");
  std::vector<uint32_t> blk(n);
  for (uint32_t k = 0; k < nb; k++)
    for (uint32_t i = bounds[k]; i < bounds[k + 1]; i++) blk[i] = k;
  const float thr = (float)identity / 100.0f;  // buildGraph.hpp:259
  std::vector<ngia_edge> all;
  std::vector<uint32_t> lpacked, llen, lrep, glob;
  std::vector<uint64_t> loff;
  std::vector<ngia_edge> part;
  std::vector<char> need(nb);
  uint32_t chunks = 0;
  for (uint32_t j = 0; j < nb; j++) {
    // rep blocks holding a candidate of a target in block j
    std::fill(need.begin(), need.end(), 0);
    for (uint32_t t = bounds[j]; t < bounds[j + 1]; t++)
      for (uint32_t b = 0; b < bands; b++) {
        const uint32_t r = rep_rows[(uint64_t)b * n + t];
        if (r != t && (float)len[t] >= (float)len[r] * thr) need[blk[r]] = 1;
      }
    for (uint32_t i = 0; i < nb; i++) {
      if (!need[i]) continue;
      // local records in global order: the lower block first
      const uint32_t lo = std::min(i, j), hi = std::max(i, j);
      glob.clear();
      for (uint32_t x = bounds[lo]; x < bounds[lo + 1]; x++) glob.push_back(x);
      if (hi != lo)
        for (uint32_t x = bounds[hi]; x < bounds[hi + 1]; x++) glob.push_back(x);
      const uint32_t nl = (uint32_t)glob.size();
      const uint32_t tj0 = hi == lo || j == lo ? 0u : bounds[lo + 1] - bounds[lo];  // block j's first local index
      const uint32_t ri0 = hi == lo || i == lo ? 0u : bounds[lo + 1] - bounds[lo];  // block i's first local index
      lpacked.clear();
      loff.resize(nl);
      llen.resize(nl);
      for (uint32_t x = 0; x < nl; x++) {
        const uint32_t g = glob[x];
        loff[x] = lpacked.size();
        llen[x] = len[g];
        lpacked.insert(lpacked.end(), packed + word_off[g], packed + word_off[g] + rec_words(len[g]));
      }
      // rows: a target of block j keeps the representatives in block i
      lrep.resize((uint64_t)bands * nl);
      for (uint32_t b = 0; b < bands; b++) {
        uint32_t* row = lrep.data() + (uint64_t)b * nl;
        for (uint32_t x = 0; x < nl; x++) row[x] = x;
        for (uint32_t t = bounds[j]; t < bounds[j + 1]; t++) {
          const uint32_t r = rep_rows[(uint64_t)b * n + t];
          if (r != t && blk[r] == i) row[tj0 + (t - bounds[j])] = ri0 + (r - bounds[i]);
        }
      }
      ngia_db* db = nullptr;
      ngia_graph_stats st;
      rc = ngia_db_create(ctx, lpacked.data(), lpacked.size(), loff.data(), llen.data(), nl, &db);
      if (rc == NGIA_OK) rc = ngia_db_set_reps(ctx, db, lrep.data(), bands);
      if (rc == NGIA_OK) rc = ngia_db_graph(ctx, db, identity, bands, NGIA_GRAPH_BATCHED, &st);
      uint64_t m = 0;
      if (rc == NGIA_OK) rc = ngia_db_get_edges(ctx, db, nullptr, 0, &m);
      if (rc == NGIA_OK) {
        part.resize(m);
        rc = ngia_db_get_edges(ctx, db, part.data(), m, &m);
      }
      ngia_db_destroy(db);
      if (rc != NGIA_OK) return rc;
      for (uint64_t e = 0; e < m; e++)
        all.push_back({(int32_t)glob[part[e].u], (int32_t)glob[part[e].v], part[e].w});
      if (stats) {
        stats->total_pairs += st.total_pairs;
        stats->distinct_pairs += st.distinct_pairs;
        stats->ref_tiles += st.ref_tiles;
        stats->ms_candidates += st.ms_candidates;
        stats->ms_align += st.ms_align;
        stats->ms_edges += st.ms_edges;
        stats->align_launches += st.align_launches;
        stats->host_syncs += st.host_syncs;
      }
      chunks++;
    }
  }
  // buildGraph.hpp:355-362: sorted unique (u, v, w); w is a function of (u, v)
  std::sort(all.begin(), all.end(), [](const ngia_edge& a, const ngia_edge& b) {
    return a.u != b.u ? a.u < b.u : a.v < b.v;
  });
  all.erase(std::unique(all.begin(), all.end(),
                        [](const ngia_edge& a, const ngia_edge& b) { return a.u == b.u && a.v == b.v; }),
            all.end());
  ngia_edge* out = (ngia_edge*)std::malloc(sizeof(ngia_edge) * std::max<size_t>(all.size(), 1));
  if (!out) return fail_msg(NGIA_ERR_NOMEM, "host allocation failed");
  std::copy(all.begin(), all.end(), out);
  *edges = out;
  *n_edges = all.size();
  if (n_chunks) *n_chunks = chunks;
  if (stats) {
    stats->n_edges = all.size();
    stats->ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return NGIA_OK;
}
