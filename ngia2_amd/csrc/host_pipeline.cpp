// host_pipeline.cpp -- the two end-to-end commands on top of the C-ABI:
// makeDB (main.cu:54-83 -> makedb.hpp:59-364) and cluster (main.cu:86-127),
// plus the result writer (saveResult.hpp:31-83).  Host work is C++/OpenMP;
// sketching, bucketing, alignment and edge emission run on the device.
#include <fcntl.h>
#include <omp.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "host_common.hpp"

using namespace ngia;

namespace {

struct Mapped {
  const uint8_t* d = nullptr;
  size_t n = 0;
  int fd = -1;
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    n = (size_t)st.st_size;
    if (n == 0) return true;
    void* p = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) return false;
    d = (const uint8_t*)p;
    return true;
  }
  ~Mapped() {
    if (d) munmap((void*)d, n);
    if (fd >= 0) ::close(fd);
  }
};

int err(int code, const std::string& msg) {
  set_thread_error(msg);
  return code;
}

bool write_all(const char* path, const uint8_t* buf, size_t bytes) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return false;
  size_t off = 0;
  while (off < bytes) {
    size_t w = std::fwrite(buf + off, 1, std::min<size_t>(bytes - off, 1 << 26), f);
    if (w == 0) break;
    off += w;
  }
  bool ok = off == bytes && std::fflush(f) == 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok;
}

}  // namespace

extern "C" int ngia_save_result(const uint32_t* cluster, uint32_t n, const char* packed_path,
                                const char* result_path) {
  Mapped db;
  if (!db.open(packed_path) || db.n < 4) return err(NGIA_ERR_IO, std::string("cannot read ") + packed_path);
  uint32_t fn;
  std::memcpy(&fn, db.d, 4);
  if (n > fn || db.n < 4 + 24ull * fn) return err(NGIA_ERR_IO, "packed file header too short");
  const uint8_t* foff = db.d + 4 + 16ull * fn;  // fastaOffsets (saveResult.hpp:40-45)
  uint32_t maxid = 0;
  for (uint32_t i = 0; i < n; i++) maxid = std::max(maxid, cluster[i]);
  // members grouped by cluster id, ascending index inside a cluster
  std::vector<uint32_t> start((size_t)maxid + 2, 0);
  for (uint32_t i = 0; i < n; i++) start[cluster[i] + 1]++;
  for (size_t c = 0; c <= maxid; c++) start[c + 1] += start[c];
  std::vector<uint32_t> fill(start.begin(), start.end() - 1), members(n);
  for (uint32_t i = 0; i < n; i++) members[fill[cluster[i]]++] = i;
  std::string out;
  out.reserve((size_t)n * 24);
  auto line_at = [&](uint64_t off, size_t* len) -> const char* {
    if (off >= db.n) {
      *len = 0;
      return "";
    }
    const void* e = std::memchr(db.d + off, '\n', db.n - off);
    *len = e ? (size_t)((const uint8_t*)e - (db.d + off)) : db.n - off;
    return (const char*)db.d + off;
  };
  for (size_t c = 0; c <= maxid; c++) {
    for (uint32_t k = start[c]; k < start[c + 1]; k++) {
      uint64_t off;
      std::memcpy(&off, foff + 8ull * members[k], 8);
      size_t l1;
      const char* name = line_at(off, &l1);
      if (k == start[c]) {  // representative: name line and residues line
        out.append(name, l1).push_back('\n');
        size_t l2;
        const char* res = line_at(off + l1 + 1, &l2);
        out.append(res, l2).push_back('\n');
      } else {
        out.append("  ").append(name, l1).push_back('\n');
      }
    }
  }
  if (!write_all(result_path, (const uint8_t*)out.data(), out.size()))
    return err(NGIA_ERR_IO, std::string("cannot write ") + result_path);
  return NGIA_OK;
}

extern "C" int ngia_fasta_check(const char* fasta) {
  std::string why;
  if (!fasta) return err(NGIA_ERR_ARG, "null path");
  if (fasta_check(fasta, &why) != 0) return err(NGIA_ERR_ARG, why);
  return NGIA_OK;
}

extern "C" int ngia_makedb(ngia_ctx* ctx, const char* fasta, const char* packed_path) {
  std::string why;
  if (fasta_check(fasta, &why) != 0) return err(NGIA_ERR_ARG, why);
  std::vector<SeqIndex> idx;
  uint64_t total = 0;
  if (fasta_index(fasta, &idx, &total) != 0) return err(NGIA_ERR_IO, std::string("cannot read ") + fasta);
  std::printf("total:\t%llu\nvalid:\t%zu\n", (unsigned long long)total, idx.size());
  const uint64_t n = idx.size();
  if (n == 0) return err(NGIA_ERR_STATE, "no valid sequence (empty file or all reads >= 65535)");
  if (n >= 0x7FFFFFFF) return err(NGIA_ERR_STATE, "too many sequences");
  const uint64_t size = assign_offsets(&idx);
  Mapped fa;
  if (!fa.open(fasta)) return err(NGIA_ERR_IO, std::string("cannot read ") + fasta);
  std::unique_ptr<uint8_t[]> buf(new (std::nothrow) uint8_t[size]);
  if (!buf) return err(NGIA_ERR_NOMEM, "host allocation failed");
  uint8_t* out = buf.get();
  // header (makedb.hpp:179-197)
  const uint32_t n32 = (uint32_t)n;
  std::memcpy(out, &n32, 4);
  std::vector<uint64_t> word_off(n);
  std::vector<uint32_t> lengths(n);
  const uint64_t pack_base = db_pack_base(n);
  for (uint64_t i = 0; i < n; i++) {
    std::memcpy(out + 4 + 4 * i, &idx[i].name_len, 4);
    std::memcpy(out + 4 + 4 * n + 4 * i, &idx[i].read_len, 4);
    std::memcpy(out + 4 + 8 * n + 8 * i, &idx[i].pack_off, 8);
    std::memcpy(out + 4 + 16 * n + 8 * i, &idx[i].text_off, 8);
    word_off[i] = (idx[i].pack_off - pack_base) / 4;
    lengths[i] = idx[i].read_len;
  }
  // packed records and text records (makedb.hpp:212-261), in parallel
  #pragma omp parallel
  {
    std::vector<uint8_t> res;
    #pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      const SeqIndex& r = idx[i];
      res.resize(r.read_len);
      uint32_t L = 0;
      const uint8_t* p = fa.d + r.fasta_off + r.name_len + 1;
      const uint8_t* e = fa.d + std::min<uint64_t>(r.fasta_off + r.seq_len, fa.n);
      for (; p < e; p++)
        if (*p != '\n') res[L++] = *p;
      pack_record(res.data(), L, (uint32_t*)(out + r.pack_off));
      uint8_t* t = out + r.text_off;
      std::memcpy(t, fa.d + r.fasta_off, r.name_len);
      t[r.name_len] = '\n';
      std::memcpy(t + r.name_len + 1, res.data(), L);
      t[r.name_len + 1 + L] = '\n';
    }
  }
  // sketch + bucket on the device, representative rows into the sign area
  const uint64_t n_words = (idx[0].text_off - pack_base) / 4;
  ngia_db* db = nullptr;
  int rc = ngia_db_create(ctx, (const uint32_t*)(out + pack_base), n_words, word_off.data(),
                          lengths.data(), n32, &db);
  if (rc == NGIA_OK) rc = ngia_db_sketch(ctx, db);
  if (rc == NGIA_OK) rc = ngia_db_bucket(ctx, db);
  if (rc == NGIA_OK) rc = ngia_db_get_reps(ctx, db, (uint32_t*)(out + db_sign_base(n)));
  ngia_db_destroy(db);
  if (rc != NGIA_OK) return rc;
  if (!write_all(packed_path, out, size))
    return err(NGIA_ERR_IO, "packData write incomplete (disk full?)");
  struct stat st;
  if (stat(packed_path, &st) != 0 || (uint64_t)st.st_size != size)
    return err(NGIA_ERR_IO, "packData write incomplete");
  std::printf("file size:\t%llu bytes\n", (unsigned long long)size);
  return NGIA_OK;
}

extern "C" int ngia_cluster(ngia_ctx* ctx, const char* packed_path, const char* result_path,
                            uint32_t identity, uint32_t bands, uint32_t flags,
                            ngia_graph_stats* stats) {
  // BuildGraph::check (buildGraph.hpp:74-99)
  {
    std::string why;
    FILE* f = std::fopen(packed_path, "rb");
    if (!f) why += std::string(packed_path) + " does not exists.\n";
    if (identity < 1 || identity > 99) why += "identity should be 1-99\n";
    if (bands < 1 || bands > NGIA_BANDS) why += "bands should be 1-64\n";
    if (f && why.empty()) {
      uint32_t c = 0;
      if (std::fread(&c, 4, 1, f) != 1 || c == 0) why += std::string(packed_path) + " contains 0 sequences.\n";
    }
    if (f) std::fclose(f);
    if (!why.empty()) return err(NGIA_ERR_ARG, why);
  }
  ngia_packed_db pdb;
  int rc = ngia_read_db(packed_path, bands, &pdb);
  if (rc != NGIA_OK) return rc;
  std::printf("reads count:\t%u\nlongest:\t%u\nshortest:\t%u\n", pdb.n, pdb.lengths[0],
              pdb.lengths[pdb.n - 1]);
  std::vector<ngia_edge> edges;
  uint64_t m = 0;
  if (flags == NGIA_GRAPH_BATCHED) {
    // The whole database on the device when it fits; else blocks of it, the
    // part the reference leaves to managed-memory paging (README.md:8).
    // NGIA_BLOCK_WORDS=N forces blocks of N packed words (0 or unset: from
    // free device memory), so the chunked path can be run on a small database.
    const char* bwe = std::getenv("NGIA_BLOCK_WORDS");
    const uint64_t bw = bwe ? std::strtoull(bwe, nullptr, 10) : 0;
    ngia_edge* e = nullptr;
    rc = ngia_build_graph_chunked(ctx, pdb.packed, pdb.n_words, pdb.word_off, pdb.lengths, pdb.n,
                                  pdb.rep_rows, bands, identity, bw, &e, &m, stats, nullptr);
    if (rc == NGIA_OK) edges.assign(e, e + m);
    ngia_free(e);
  } else {
    ngia_db* db = nullptr;
    rc = ngia_db_create(ctx, pdb.packed, pdb.n_words, pdb.word_off, pdb.lengths, pdb.n, &db);
    if (rc == NGIA_OK) rc = ngia_db_set_reps(ctx, db, pdb.rep_rows, bands);
    if (rc == NGIA_OK) rc = ngia_db_graph(ctx, db, identity, bands, flags, stats);
    if (rc == NGIA_OK) rc = ngia_db_get_edges(ctx, db, nullptr, 0, &m);
    if (rc == NGIA_OK) {
      edges.resize(m);
      rc = ngia_db_get_edges(ctx, db, edges.data(), m, &m);
    }
    ngia_db_destroy(db);
  }
  const uint32_t n = pdb.n;
  ngia_release_db(&pdb);
  if (rc != NGIA_OK) return rc;
  std::vector<uint32_t> cluster(n);
  rc = ngia_leiden(edges.data(), m, (int32_t)n, ngia_resolution(identity), 42, 1, cluster.data());
  if (rc != NGIA_OK) return rc;
  return ngia_save_result(cluster.data(), n, packed_path, result_path);
}
