// host_gen.cpp -- seeded synthetic sequence families for the benchmark suite.
// No reference counterpart.  Every family draws from its own RNG stream
// (seeded from (seed, family)), so the output does not depend on the thread
// count; the FASTA order is a seeded shuffle of all members.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "host_common.hpp"

namespace {

inline uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Rng {  // xoshiro256**
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; i++) s[i] = splitmix64(&seed);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  uint32_t below(uint32_t k) { return (uint32_t)(((next() >> 32) * (uint64_t)k) >> 32); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

const char kAmino[] = "ACDEFGHIKLMNPQRSTVWY";
const char kDna[] = "ACGT";

uint32_t draw_length(const ngia_gen_params& p, Rng& r) {
  double L;
  if (p.tail_frac > 0 && r.uniform() < p.tail_frac) {
    // Pareto(alpha) above tail_xmin, truncated at tail_max.
    const double u = 1.0 - r.uniform();
    L = p.tail_xmin * std::pow(u, -1.0 / p.tail_alpha);
    if (L > p.tail_max) L = p.tail_max;
  } else if (p.len_normal) {
    L = p.len_a + p.len_b * r.normal();
  } else {
    L = p.len_a * std::exp(p.len_b * r.normal());
  }
  L = std::floor(L + 0.5);
  if (L < p.len_min) L = p.len_min;
  if (L > p.len_max) L = p.len_max;
  return (uint32_t)L;
}

// Ancestor + members of one family, each member as a residue string.
void make_family(const ngia_gen_params& p, uint32_t fam, std::vector<std::string>* out) {
  Rng r(p.seed * 0x9e3779b97f4a7c15ULL + fam + 1);
  const char* alpha = p.dna ? kDna : kAmino;
  const uint32_t na = p.dna ? 4 : 20;
  const uint32_t L = draw_length(p, r);
  std::string anc(L, 'A');
  for (uint32_t j = 0; j < L; j++) anc[j] = alpha[r.below(na)];
  out->resize(p.members);
  for (uint32_t m = 0; m < p.members; m++) {
    const double sub = p.sub_lo + (p.sub_hi - p.sub_lo) * r.uniform();
    const double ind = p.indel_lo + (p.indel_hi - p.indel_lo) * r.uniform();
    std::string& s = (*out)[m];
    s.clear();
    s.reserve(L + L / 8 + 8);
    for (uint32_t j = 0; j < L; j++) {
      if (r.uniform() < ind) {
        if (r.uniform() < 0.5) continue;       // deletion
        s.push_back(alpha[r.below(na)]);       // insertion before j
      }
      char c = anc[j];
      if (r.uniform() < sub) {                 // substitution by another letter
        uint32_t k = r.below(na - 1);
        const uint32_t cur = (uint32_t)(std::strchr(alpha, c) - alpha);
        if (k >= cur) k++;
        c = alpha[k];
      }
      s.push_back(c);
    }
    if (s.size() >= 0xFFFF) s.resize(0xFFFE);
  }
}

// All members in FASTA order: a seeded permutation of family-major order.
void generate(const ngia_gen_params& p, std::vector<std::string>* seqs,
              std::vector<std::string>* names) {
  const uint64_t N = (uint64_t)p.n_families * p.members;
  std::vector<std::string> all(N);
  #pragma omp parallel
  {
    std::vector<std::string> fam;
    #pragma omp for schedule(dynamic, 16)
    for (int64_t f = 0; f < (int64_t)p.n_families; f++) {
      make_family(p, (uint32_t)f, &fam);
      for (uint32_t m = 0; m < p.members; m++) all[(uint64_t)f * p.members + m].swap(fam[m]);
    }
  }
  std::vector<uint32_t> perm(N);
  for (uint64_t i = 0; i < N; i++) perm[i] = (uint32_t)i;
  Rng r(p.seed ^ 0x5851f42d4c957f2dULL);
  for (uint64_t i = N; i > 1; i--) std::swap(perm[i - 1], perm[r.below((uint32_t)i)]);
  seqs->resize(N);
  names->resize(N);
  for (uint64_t i = 0; i < N; i++) {
    const uint32_t g = perm[i];
    (*seqs)[i].swap(all[g]);
    (*names)[i] = ">f" + std::to_string(g / p.members) + "_m" + std::to_string(g % p.members);
  }
}

}  // namespace

extern "C" int ngia_gen_config(int config, double scale, uint64_t seed, ngia_gen_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->seed = seed;
  switch (config) {
    case 1:  // 10,000 proteins in 200 families, log-normal median 300 in [50,1000]
      p->n_families = 200; p->members = 50;
      p->len_a = 300; p->len_b = 0.5; p->len_min = 50; p->len_max = 1000;
      p->sub_lo = 0.05; p->sub_hi = 0.40; p->indel_lo = p->indel_hi = 0.02;
      break;
    case 2:  // 1,000,000 proteins in 50,000 families, median 350 in [30,5000]
      p->n_families = 50000; p->members = 20;
      p->len_a = 350; p->len_b = 0.5; p->len_min = 30; p->len_max = 5000;
      p->sub_lo = 0.0; p->sub_hi = 0.50; p->indel_lo = 0.01; p->indel_hi = 0.03;
      break;
    case 3:  // 2,000,000 DNA in 20,000 families, normal(1400,100) in [1000,1800]
      p->n_families = 20000; p->members = 100; p->dna = 1; p->len_normal = 1;
      p->len_a = 1400; p->len_b = 100; p->len_min = 1000; p->len_max = 1800;
      p->sub_lo = 0.0; p->sub_hi = 0.05; p->indel_lo = p->indel_hi = 0.005;
      break;
    case 4:  // 1,000,000 heavy-tailed proteins: log-normal body + Pareto(1.5) tail to 35,000
    case 5:  // serving: the config-4 database
      p->n_families = 50000; p->members = 20;
      p->len_a = 300; p->len_b = 0.5; p->len_min = 30; p->len_max = 35000;
      p->tail_frac = 0.02; p->tail_xmin = 1000; p->tail_alpha = 1.5; p->tail_max = 35000;
      p->sub_lo = 0.0; p->sub_hi = 0.40; p->indel_lo = p->indel_hi = 0.01;
      break;
    default:
      ngia::set_thread_error("config must be 1..5");
      return NGIA_ERR_ARG;
  }
  if (scale > 0 && scale != 1.0) {
    double f = std::floor(p->n_families * scale + 0.5);
    p->n_families = f < 1 ? 1 : (uint32_t)f;
  }
  return NGIA_OK;
}

extern "C" int64_t ngia_gen_fasta(const ngia_gen_params* p, const char* path) {
  std::vector<std::string> seqs, names;
  generate(*p, &seqs, &names);
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    ngia::set_thread_error(std::string("cannot write ") + path);
    return NGIA_ERR_IO;
  }
  static char buf[1 << 20];
  std::setvbuf(f, buf, _IOFBF, sizeof(buf));
  for (size_t i = 0; i < seqs.size(); i++) {
    std::fwrite(names[i].data(), 1, names[i].size(), f);
    std::fputc('\n', f);
    std::fwrite(seqs[i].data(), 1, seqs[i].size(), f);
    std::fputc('\n', f);
  }
  if (std::fclose(f) != 0) return NGIA_ERR_IO;
  return (int64_t)seqs.size();
}

extern "C" int ngia_gen_packed(const ngia_gen_params* p, ngia_packed_db* out) {
  using namespace ngia;
  std::memset(out, 0, sizeof(*out));
  std::vector<std::string> seqs, names;
  generate(*p, &seqs, &names);
  // Index order = stable length-descending sort of the FASTA order.
  std::vector<SeqIndex> idx(seqs.size());
  std::vector<uint32_t> order(seqs.size());
  for (size_t i = 0; i < seqs.size(); i++) {
    idx[i].name_len = (uint32_t)names[i].size();
    idx[i].read_len = (uint32_t)seqs[i].size();
    order[i] = (uint32_t)i;
  }
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return idx[a].read_len > idx[b].read_len;
  });
  std::vector<SeqIndex> sorted(seqs.size());
  for (size_t i = 0; i < seqs.size(); i++) sorted[i] = idx[order[i]];
  assign_offsets(&sorted);
  const uint32_t n = (uint32_t)seqs.size();
  out->n = n;
  out->name_len = (uint32_t*)std::malloc(4ull * n);
  out->lengths = (uint32_t*)std::malloc(4ull * n);
  out->pack_off = (uint64_t*)std::malloc(8ull * n);
  out->text_off = (uint64_t*)std::malloc(8ull * n);
  out->word_off = (uint64_t*)std::malloc(8ull * n);
  uint64_t words = 0;
  for (uint32_t i = 0; i < n; i++) {
    out->name_len[i] = sorted[i].name_len;
    out->lengths[i] = sorted[i].read_len;
    out->pack_off[i] = sorted[i].pack_off;
    out->text_off[i] = sorted[i].text_off;
    out->word_off[i] = words;
    words += rec_words(sorted[i].read_len);
  }
  out->n_words = words;
  out->packed = (uint32_t*)std::malloc(std::max<uint64_t>(4, 4 * words));
  #pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const std::string& s = seqs[order[i]];
    pack_record((const uint8_t*)s.data(), (uint32_t)s.size(), out->packed + out->word_off[i]);
  }
  return NGIA_OK;
}

extern "C" int ngia_gen_queries(const uint32_t* packed, const uint64_t* word_off, uint32_t n,
                                const ngia_query_params* p, uint8_t** res, uint64_t** res_off,
                                uint32_t** res_len, uint8_t** kind) {
  if (!p || !res || !res_off || !res_len || !kind || (n && (!packed || !word_off)))
    return NGIA_ERR_ARG;
  if (p->mutated > 0 && n == 0) return NGIA_ERR_ARG;
  const uint32_t nq = p->n_queries;
  // A letter for each code (the first letter of makedb.hpp:164-169 mapping to it).
  static const char kLetter[16] = {'X', 'A', 'C', 'D', 'E', 'F', 'G', 'H',
                                   'I', 'K', 'L', 'N', 'P', 'R', 'S', 'T'};
  std::vector<std::string> q(nq);
  std::vector<uint8_t> kd(nq);
  #pragma omp parallel for schedule(dynamic, 64)
  for (int64_t k = 0; k < (int64_t)nq; k++) {
    Rng r(p->seed * 0xd1342543de82ef95ULL + (uint64_t)k + 7);
    std::string& s = q[k];
    if (r.uniform() < p->mutated) {
      kd[k] = 1;
      const uint32_t* rec = packed + word_off[r.below(n)];
      const uint32_t L = rec[0];
      const double sub = p->sub_lo + (p->sub_hi - p->sub_lo) * r.uniform();
      s.reserve(L + L / 8 + 8);
      for (uint32_t j = 0; j < L; j++) {
        const uint32_t* pl = rec + 1 + 4 * (j / 32);
        const uint32_t b = j % 32;
        const uint32_t code = ((pl[0] >> b) & 1u) | (((pl[1] >> b) & 1u) << 1) |
                              (((pl[2] >> b) & 1u) << 2) | (((pl[3] >> b) & 1u) << 3);
        if (r.uniform() < p->indel) {
          if (r.uniform() < 0.5) continue;   // deletion
          s.push_back(kAmino[r.below(20)]);  // insertion before j
        }
        s.push_back(r.uniform() < sub ? kAmino[r.below(20)] : kLetter[code]);
      }
    } else {
      kd[k] = 0;
      double L = std::floor(p->len_median * std::exp(p->len_sigma * r.normal()) + 0.5);
      L = std::min<double>(std::max<double>(L, p->len_min), p->len_max);
      s.resize((size_t)L);
      for (auto& c : s) c = kAmino[r.below(20)];
    }
    if (s.size() >= 0xFFFF) s.resize(0xFFFE);
  }
  uint64_t total = 0;
  for (const auto& s : q) total += s.size();
  *res = (uint8_t*)std::malloc(std::max<uint64_t>(total, 1));
  *res_off = (uint64_t*)std::malloc(8ull * std::max<uint32_t>(nq, 1));
  *res_len = (uint32_t*)std::malloc(4ull * std::max<uint32_t>(nq, 1));
  *kind = (uint8_t*)std::malloc(std::max<uint32_t>(nq, 1));
  if (!*res || !*res_off || !*res_len || !*kind) return NGIA_ERR_NOMEM;
  uint64_t off = 0;
  for (uint32_t k = 0; k < nq; k++) {
    (*res_off)[k] = off;
    (*res_len)[k] = (uint32_t)q[k].size();
    (*kind)[k] = kd[k];
    std::memcpy(*res + off, q[k].data(), q[k].size());
    off += q[k].size();
  }
  return NGIA_OK;
}
