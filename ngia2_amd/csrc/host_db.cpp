// host_db.cpp -- FASTA index, packed-record layout and packed-DB reader.
//
// The packed DB byte layout is the reference's (makedb.hpp:1-25, written at
// makedb.hpp:179-261, read at buildGraph.hpp:134-177): it has to be identical
// for the drop-in, so the layout code below follows it field for field.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <numeric>

#include "host_common.hpp"

namespace ngia {

static thread_local std::string g_thread_error;
void set_thread_error(const std::string& msg) { g_thread_error = msg; }
const char* thread_error() { return g_thread_error.c_str(); }

CodeTable::CodeTable() {
  std::memset(t, 0, sizeof(t));
  // makedb.hpp:161-171: reduced 15-letter alphabet, case-insensitive.
  const char* groups[] = {"A", "C", "D", "EQ", "FWY", "G", "H", "IV", "K", "LM",
                          "N", "P", "R", "S", "TU"};
  for (int code = 1; code <= 15; code++)
    for (const char* p = groups[code - 1]; *p; p++) {
      t[(uint8_t)*p] = (uint8_t)code;
      t[(uint8_t)(*p | 0x20)] = (uint8_t)code;
    }
}
const CodeTable kCodes;

void pack_record(const uint8_t* res, uint32_t L, uint32_t* rec) {
  const uint32_t nb = (L + 31) / 32;
  rec[0] = L;
  for (uint32_t b = 0; b < nb; b++) {
    uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    const uint32_t lim = std::min<uint32_t>(32, L - 32 * b);
    const uint8_t* s = res + 32 * b;
    for (uint32_t k = 0; k < lim; k++) {
      const uint32_t c = kCodes.t[s[k]];
      p0 |= (c & 1u) << k;
      p1 |= ((c >> 1) & 1u) << k;
      p2 |= ((c >> 2) & 1u) << k;
      p3 |= ((c >> 3) & 1u) << k;
    }
    rec[1 + 4 * b + 0] = p0;
    rec[1 + 4 * b + 1] = p1;
    rec[1 + 4 * b + 2] = p2;
    rec[1 + 4 * b + 3] = p3;
  }
}

void pack_records(const uint8_t* res, const uint64_t* res_off, const uint32_t* res_len,
                  uint32_t first, uint32_t count, const uint64_t* word_pos, uint64_t word_base,
                  uint32_t* dst) {
  #pragma omp parallel for schedule(dynamic, 64) if (count > 512)
  for (int64_t j = 0; j < (int64_t)count; j++) {
    const uint32_t q = first + (uint32_t)j;
    pack_record(res + res_off[q], res_len[q], dst + (word_pos[j] - word_base));
  }
}

namespace {
struct MappedFile {
  const uint8_t* data = nullptr;
  size_t size = 0;
  int fd = -1;
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    size = (size_t)st.st_size;
    if (size == 0) return true;
    void* p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) return false;
    data = (const uint8_t*)p;
    madvise(p, size, MADV_SEQUENTIAL);
    return true;
  }
  ~MappedFile() {
    if (data) munmap((void*)data, size);
    if (fd >= 0) ::close(fd);
  }
};
}  // namespace

int fasta_check(const char* path, std::string* why) {
  // makedb.hpp:59-91: the file exists, its first line has no trailing '\r',
  // and a non-empty file ends with '\n'.
  MappedFile f;
  if (!f.open(path)) {
    *why = std::string(path) + " does not exist.";
    return -1;
  }
  size_t e = 0;
  while (e < f.size && f.data[e] != '\n') e++;
  if (e > 0 && f.data[e - 1] == '\r') {
    *why = "Please remove the \\r at the end of lines.";
    return -1;
  }
  if (f.size > 0 && f.data[f.size - 1] != '\n') {
    *why = "File must end with \\n.";
    return -1;
  }
  return 0;
}

int fasta_index(const char* path, std::vector<SeqIndex>* idx, uint64_t* total) {
  MappedFile f;
  if (!f.open(path)) return -1;
  idx->clear();
  const uint8_t* d = f.data;
  const size_t n = f.size;
  size_t pos = 0;
  uint64_t whole = 0;
  auto line_end = [&](size_t p) {
    const void* q = p < n ? std::memchr(d + p, '\n', n - p) : nullptr;
    return q ? (size_t)((const uint8_t*)q - d) : n;
  };
  // makedb.hpp:106-122: a record is one line (the name, whatever it starts
  // with) followed by every line up to the next line that starts with '>'.
  while (pos < n) {
    SeqIndex r{};
    r.fasta_off = pos;
    size_t e = line_end(pos);
    r.name_len = (uint32_t)(e - pos);
    r.seq_len = r.name_len + 1;
    pos = e < n ? e + 1 : n;
    uint64_t rl = 0, sl = r.seq_len;
    while (pos < n && d[pos] != '>') {
      e = line_end(pos);
      rl += e - pos;
      sl += e - pos + 1;
      pos = e < n ? e + 1 : n;
    }
    whole++;
    if (rl < 0xFFFF) {
      r.read_len = (uint32_t)rl;
      r.seq_len = (uint32_t)sl;
      idx->push_back(r);
    }
  }
  // makedb.hpp:130-134: stable sort by read length, descending.
  std::stable_sort(idx->begin(), idx->end(),
                   [](const SeqIndex& a, const SeqIndex& b) { return a.read_len > b.read_len; });
  if (total) *total = whole;
  return 0;
}

uint64_t assign_offsets(std::vector<SeqIndex>* idx) {
  const uint64_t n = idx->size();
  uint64_t off = db_pack_base(n);
  for (auto& r : *idx) {
    r.pack_off = off;
    off += 4ull * rec_words(r.read_len);
  }
  for (auto& r : *idx) {
    r.text_off = off;
    off += (uint64_t)r.name_len + r.read_len + 2;
  }
  return off;
}

}  // namespace ngia

using namespace ngia;

extern "C" int ngia_read_db(const char* path, uint32_t bands, ngia_packed_db* out) {
  std::memset(out, 0, sizeof(*out));
  if (bands < 1 || bands > NGIA_BANDS) {
    set_thread_error("bands should be 1-64");
    return NGIA_ERR_ARG;
  }
  MappedFile f;
  if (!f.open(path) || f.size < 4) {
    set_thread_error(std::string(path) + " cannot be read");
    return NGIA_ERR_IO;
  }
  uint32_t n;
  std::memcpy(&n, f.data, 4);
  if (n == 0 || f.size < db_pack_base(n)) {
    set_thread_error(std::string(path) + " contains 0 sequences or is truncated");
    return NGIA_ERR_IO;
  }
  out->n = n;
  out->name_len = (uint32_t*)std::malloc(4ull * n);
  out->lengths = (uint32_t*)std::malloc(4ull * n);
  out->pack_off = (uint64_t*)std::malloc(8ull * n);
  out->text_off = (uint64_t*)std::malloc(8ull * n);
  out->word_off = (uint64_t*)std::malloc(8ull * n);
  out->rep_rows = (uint32_t*)std::malloc(4ull * n * bands);
  // The header arrays are unaligned 8-byte fields: copy, never cast.
  std::memcpy(out->name_len, f.data + 4, 4ull * n);
  std::memcpy(out->lengths, f.data + 4 + 4ull * n, 4ull * n);
  std::memcpy(out->pack_off, f.data + 4 + 8ull * n, 8ull * n);
  std::memcpy(out->text_off, f.data + 4 + 16ull * n, 8ull * n);
  std::memcpy(out->rep_rows, f.data + db_sign_base(n), 4ull * n * bands);
  const uint64_t pack0 = out->pack_off[0], text0 = out->text_off[0];
  if (text0 < pack0 || text0 > f.size || (text0 - pack0) % 4 != 0) {
    set_thread_error(std::string(path) + ": inconsistent packed area");
    ngia_release_db(out);
    return NGIA_ERR_IO;
  }
  out->n_words = (text0 - pack0) / 4;
  out->packed = (uint32_t*)std::malloc(std::max<uint64_t>(4, text0 - pack0));
  std::memcpy(out->packed, f.data + pack0, text0 - pack0);
  // buildGraph.hpp:173-177: byte offsets -> uint32 word offsets in the area.
  for (uint32_t i = 0; i < n; i++) {
    const uint64_t w = (out->pack_off[i] - pack0) / 4;
    if (out->pack_off[i] < pack0 || w + rec_words(out->lengths[i]) > out->n_words) {
      set_thread_error(std::string(path) + ": record outside the packed area");
      ngia_release_db(out);
      return NGIA_ERR_IO;
    }
    out->word_off[i] = w;
  }
  return NGIA_OK;
}

extern "C" void ngia_release_db(ngia_packed_db* db) {
  if (!db) return;
  std::free(db->name_len);
  std::free(db->lengths);
  std::free(db->pack_off);
  std::free(db->text_off);
  std::free(db->rep_rows);
  std::free(db->packed);
  std::free(db->word_off);
  std::memset(db, 0, sizeof(*db));
}

extern "C" void ngia_free(void* p) { std::free(p); }
