/*
 * ngia2_oracle.c -- CPU restatement of the nGIA2 clustering hot path.
 * TEST INFRASTRUCTURE ONLY (see ngia2_oracle.h): used by tests/, the smoke
 * check and the CPU-baseline leg of bench.py; never by the product.
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "ngia2_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* B1: reduced 15-letter code, makedb.hpp:161-171.  Case-insensitive; every  */
/* byte not listed maps to 0.                                                */
uint32_t orc_code(uint8_t c) {
  /* c|0x20 lands in 'a'..'z' exactly when c is an ASCII letter */
  switch (c | 0x20) {
    case 'a': return 1;  case 'c': return 2;  case 'd': return 3;
    case 'e': return 4;  case 'q': return 4;  case 'f': return 5;
    case 'w': return 5;  case 'y': return 5;  case 'g': return 6;
    case 'h': return 7;  case 'i': return 8;  case 'v': return 8;
    case 'k': return 9;  case 'l': return 10; case 'm': return 10;
    case 'n': return 11; case 'p': return 12; case 'r': return 13;
    case 's': return 14; case 't': return 15; case 'u': return 15;
    default: return 0;
  }
}

/* B2: 4-bit-plane record, makedb.hpp:222-231.  rec[0]=L; bit j%32 of word  */
/* 1+4*(j/32)+e is bit e of code(residue j).                                 */
void orc_pack(const uint8_t* residues, uint32_t L, uint32_t* rec) {
  uint32_t nw = 1 + (L + 31) / 32 * 4;
  memset(rec, 0, (size_t)nw * 4);
  rec[0] = L;
  for (uint32_t j = 0; j < L; j++) {
    uint32_t code = orc_code(residues[j]);
    for (uint32_t e = 0; e < 4; e++)
      rec[1 + j / 32 * 4 + e] |= ((code >> e) & 1u) << (j % 32);
  }
}

static inline uint32_t rec_code(const uint32_t* rec, uint32_t j) {
  const uint32_t* blk = rec + 1 + (j / 32) * 4;
  uint32_t k = j % 32, c = 0;
  for (uint32_t e = 0; e < 4; e++) c |= ((blk[e] >> k) & 1u) << e;
  return c;
}

/* B3: minhash sketch, makedb.hpp:224-247.  kmer is the last 8 codes (4 bits */
/* each, mod 2^32); the first 7 positions are skipped unless the sequence is */
/* shorter than 8, in which case only its last position is hashed.          */
void orc_sketch_record(const uint32_t* rec, uint32_t sig[ORC_BANDS]) {
  uint32_t L = rec[0];
  for (int k = 0; k < ORC_BANDS; k++) sig[k] = 0xFFFFFFFFu;
  uint32_t kmer = 0;
  for (uint32_t j = 0; j < L; j++) {
    kmer = (kmer << 4) + rec_code(rec, j);
    if (j < 7 && j < L - 1) continue;
    for (uint32_t k = 0; k < ORC_BANDS; k++) {
      uint32_t h = kmer ^ (k * 0x9e3779b1u + 0x85ebca6bu);
      h = (h ^ 61u) ^ (h >> 16);
      h = h + (h << 3);
      h = h ^ (h >> 4);
      h = h * 0x27d4eb2du;
      h = h ^ (h >> 15);
      h = (h ^ (h >> 8)) * 0x9e3779b1u;
      h = h ^ (h >> 14);
      if (h < sig[k]) sig[k] = h;
    }
  }
}

void orc_sketch_all(const uint32_t* packed, const uint64_t* word_off, uint32_t n,
                    uint32_t* sig, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  #pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < (int64_t)n; i++)
    orc_sketch_record(packed + word_off[i], sig + (size_t)i * ORC_BANDS);
}

/* B4: band bucketing, makedb.hpp:289-364.  Per band a stable LSD radix sort */
/* of (sig<<32 | i) by the 32-bit signature, then rep[i] = the previous      */
/* index of the same signature run, or i itself for a run head.             */
void orc_bucket(const uint32_t* sig, uint32_t n, uint32_t* rep, int nthreads) {
  if (n == 0) return;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  if (nthreads > ORC_BANDS) nthreads = ORC_BANDS;
  #pragma omp parallel num_threads(nthreads)
  {
    uint64_t* a = (uint64_t*)malloc((size_t)n * 8);
    uint64_t* b = (uint64_t*)malloc((size_t)n * 8);
    #pragma omp for schedule(dynamic, 1)
    for (int band = 0; band < ORC_BANDS; band++) {
      for (uint32_t i = 0; i < n; i++)
        a[i] = ((uint64_t)sig[(size_t)i * ORC_BANDS + band] << 32) | i;
      uint64_t *src = a, *dst = b;
      for (int pass = 0; pass < 4; pass++) {
        uint32_t cnt[256] = {0};
        int sh = 32 + 8 * pass;
        for (uint32_t i = 0; i < n; i++) cnt[(src[i] >> sh) & 0xFF]++;
        uint32_t s = 0;
        for (int d = 0; d < 256; d++) { uint32_t c = cnt[d]; cnt[d] = s; s += c; }
        for (uint32_t i = 0; i < n; i++) dst[cnt[(src[i] >> sh) & 0xFF]++] = src[i];
        uint64_t* t = src; src = dst; dst = t;
      }
      uint32_t* row = rep + (size_t)band * n;
      uint32_t pk = (uint32_t)(src[0] >> 32), pi = (uint32_t)src[0];
      row[pi] = pi;
      for (uint32_t y = 1; y < n; y++) {
        uint32_t key = (uint32_t)(src[y] >> 32), idx = (uint32_t)src[y];
        row[idx] = (key == pk) ? pi : idx;
        pk = key; pi = idx;
      }
    }
    free(a); free(b);
  }
}

/* B5: candidate predicate, buildGraph.hpp:259-266 and 278-300.             */
static inline int is_job(const uint32_t* rep_row, const uint32_t* len, uint32_t i, float thr) {
  uint32_t r = rep_row[i];
  if (r == i) return 0;
  float lim = (float)len[r] * thr;
  return (float)len[i] >= lim;
}

uint64_t orc_candidates(const uint32_t* rep_row, const uint32_t* len, uint32_t n,
                        uint32_t identity, uint64_t* out) {
  float thr = (float)identity / 100.0f;
  uint64_t c = 0;
  for (uint32_t i = 0; i < n; i++)
    if (is_job(rep_row, len, i, thr)) {
      if (out) out[c] = ((uint64_t)rep_row[i] << 32) | i;
      c++;
    }
  return c;
}

/* float -> uint32 the way both GPU ISAs convert: negative saturates to 0.  */
static inline uint32_t f2u_sat(float x) {
  if (!(x > 0.0f)) return 0;
  if (x >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)x;
}

/* Band edges, buildGraph.hpp:208-210.  fused=1 restates fmaf(-T,thr,T),    */
/* the contraction nvcc performs by default; fused=0 the unfused form.      */
static void band_edges(uint32_t R, uint32_t T, float thr, int fused,
                       uint32_t* lsft, uint32_t* rsft) {
  float fT = (float)T, fR = (float)R;
  float l, r;
  if (fused) {
    l = fmaf(-fT, thr, fT);
    r = fmaf(-fT, thr, fR);
  } else {
    float p = fT * thr;
    l = fT - p;
    r = fR - p;
  }
  *lsft = (f2u_sat(ceilf(l)) + 31u) & ~31u;
  *rsft = f2u_sat(ceilf(r)) + 32u;
}

/* B6: kernel_dynamic, buildGraph.hpp:195-252, as the same 32x32 tile loop.
 * With lcs_out set, also the LCS after each 32-character block (the zero bits
 * of all rows), and with tiles_out the tiles visited through it. */
static uint32_t align_loop(const uint32_t* represent, const uint32_t* target, float thr,
                           int fused, uint32_t* lcs_out, uint64_t* tiles_out) {
  const uint32_t R = represent[0], T = target[0];
  if ((float)T < (float)R * thr) return 0;  /* job keeps its initial 0 */
  const uint32_t nw = (R + 31) / 32;
  uint64_t tiles = 0;
  uint32_t lines[2048];
  for (uint32_t w = 0; w < nw; w++) lines[w] = 0xFFFFFFFFu;
  uint32_t lsft, rsft;
  band_edges(R, T, thr, fused, &lsft, &rsft);
  for (uint32_t i = 0; i < T; i += 32) {
    uint32_t carrys = 0, cols[4];
    for (int e = 0; e < 4; e++) cols[e] = target[1 + i / 32 * 4 + e];
    int32_t js = (int32_t)i - (int32_t)lsft;
    if (js < 0) js = 0;
    int32_t je = (int32_t)i + (int32_t)rsft;
    if (je > (int32_t)R) je = (int32_t)R;
    for (uint32_t j = (uint32_t)js; j < (uint32_t)je; j += 32) {
      uint32_t rows[4], matchs[16];
      for (int e = 0; e < 4; e++) rows[e] = represent[1 + j / 32 * 4 + e];
      matchs[0] = 0;
      for (uint32_t k = 1; k < 16; k++) {
        uint32_t m = 0xFFFFFFFFu;
        for (uint32_t e = 0; e < 4; e++) m &= ((k >> e) & 1u) ? rows[e] : ~rows[e];
        matchs[k] = m;
      }
      uint32_t row = lines[j / 32];
      for (uint32_t k = 0; k < 32; k++) {
        uint32_t order = 0;
        for (uint32_t e = 0; e < 4; e++) order |= ((cols[e] >> k) & 1u) << e;
        uint32_t match = matchs[order];
        uint32_t carry = carrys & 1u;
        uint32_t t0 = row & match, t1 = row & ~match;
        uint32_t s = row + carry;
        carry = s < row;
        s += t0;
        carry |= s < t0;
        row = s | t1;
        carrys = (carrys >> 1) | (carry << 31);
      }
      lines[j / 32] = row;
      tiles++;
    }
    if (lcs_out) {
      uint32_t z = 0;
      for (uint32_t w = 0; w < nw; w++) z += 32u - (uint32_t)__builtin_popcount(lines[w]);
      lcs_out[i / 32] = z;
    }
    if (tiles_out) tiles_out[i / 32] = tiles;
  }
  uint32_t lcs = 0;
  for (uint32_t w = 0; w < nw; w++) lcs += 32u - (uint32_t)__builtin_popcount(lines[w]);
  return lcs;
}

uint32_t orc_align_lcs(const uint32_t* represent, const uint32_t* target, float thr, int fused) {
  return align_loop(represent, target, thr, fused, NULL, NULL);
}

uint32_t orc_align_blocks(const uint32_t* represent, const uint32_t* target, float thr, int fused,
                          uint32_t* lcs_out, uint64_t* tiles_out) {
  const uint32_t R = represent[0], T = target[0];
  if ((float)T < (float)R * thr) return 0;
  align_loop(represent, target, thr, fused, lcs_out, tiles_out);
  return (T + 31) / 32;
}

uint32_t orc_band_window(uint32_t R, uint32_t T, float thr, int fused) {
  uint32_t lsft, rsft;
  band_edges(R, T, thr, fused, &lsft, &rsft);
  return lsft / 32 + (rsft + 31) / 32;
}

uint32_t orc_align_record(const uint32_t* represent, const uint32_t* target, float thr,
                          int fused) {
  const uint32_t R = represent[0], T = target[0];
  if ((float)T < (float)R * thr) return 0;  /* job keeps its initial 0 */
  const uint32_t lcs = orc_align_lcs(represent, target, thr, fused);
  uint32_t denom = R + T - lcs;
  if ((float)lcs > (float)denom * thr) return lcs * 100u / denom;
  return 0;
}

uint64_t orc_align_tiles(uint32_t R, uint32_t T, float thr, int fused) {
  if ((float)T < (float)R * thr) return 0;
  uint32_t lsft, rsft;
  band_edges(R, T, thr, fused, &lsft, &rsft);
  uint64_t tiles = 0;
  for (uint32_t i = 0; i < T; i += 32) {
    int32_t js = (int32_t)i - (int32_t)lsft;
    if (js < 0) js = 0;
    int32_t je = (int32_t)i + (int32_t)rsft;
    if (je > (int32_t)R) je = (int32_t)R;
    if (je > js) tiles += ((uint32_t)je - (uint32_t)js + 31) / 32;
  }
  return tiles;
}

void orc_align_jobs(const uint32_t* packed, const uint64_t* word_off, uint32_t* jobs,
                    uint32_t count, float thr, int fused, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t j = 0; j < (int64_t)count; j++) {
    uint32_t* job = jobs + 3 * (size_t)j;
    job[2] = orc_align_record(packed + word_off[job[0]], packed + word_off[job[1]], thr, fused);
  }
}

void orc_align_jobs_lcs(const uint32_t* packed, const uint64_t* word_off, uint32_t* jobs,
                        uint32_t count, float thr, int fused, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  #pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t j = 0; j < (int64_t)count; j++) {
    uint32_t* job = jobs + 3 * (size_t)j;
    job[2] = orc_align_lcs(packed + word_off[job[0]], packed + word_off[job[1]], thr, fused);
  }
}

/* Textbook LCS DP, two rows. */
uint32_t orc_lcs_full(const uint8_t* a, uint32_t R, const uint8_t* b, uint32_t T) {
  uint32_t* prev = (uint32_t*)calloc((size_t)T + 1, 4);
  uint32_t* cur = (uint32_t*)calloc((size_t)T + 1, 4);
  for (uint32_t i = 1; i <= R; i++) {
    cur[0] = 0;
    for (uint32_t j = 1; j <= T; j++) {
      if (a[i - 1] != 0 && a[i - 1] == b[j - 1]) cur[j] = prev[j - 1] + 1;
      else cur[j] = prev[j] > cur[j - 1] ? prev[j] : cur[j - 1];
    }
    uint32_t* t = prev; prev = cur; cur = t;
  }
  uint32_t r = prev[T];
  free(prev); free(cur);
  return r;
}

/* B7: buildGraph.hpp:256-363.  Band loop, edges where score>0, then sort   */
/* by (u,v,w) and unique.                                                   */
static int edge_cmp(const void* x, const void* y) {
  const orc_edge* a = (const orc_edge*)x;
  const orc_edge* b = (const orc_edge*)y;
  if (a->u != b->u) return a->u < b->u ? -1 : 1;
  if (a->v != b->v) return a->v < b->v ? -1 : 1;
  if (a->w != b->w) return a->w < b->w ? -1 : 1;
  return 0;
}

int orc_build_graph(const uint32_t* packed, const uint64_t* word_off, const uint32_t* len,
                    uint32_t n, const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                    int fused, int nthreads, orc_edge** edges, uint64_t* n_edges,
                    uint64_t* total_pairs) {
  const float thr = (float)identity / 100.0f;
  size_t cap = 1024, m = 0;
  orc_edge* g = (orc_edge*)malloc(cap * sizeof(orc_edge));
  uint64_t* cand = (uint64_t*)malloc(((size_t)n + 1) * 8);
  uint32_t* sc = (uint32_t*)malloc(((size_t)n + 1) * 4);
  uint64_t tot = 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  for (uint32_t band = 0; band < bands; band++) {
    const uint32_t* rep = rep_rows + (size_t)band * n;
    uint64_t cnt = orc_candidates(rep, len, n, identity, cand);
    tot += cnt;
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t j = 0; j < (int64_t)cnt; j++) {
      uint32_t r = (uint32_t)(cand[j] >> 32), t = (uint32_t)cand[j];
      sc[j] = orc_align_record(packed + word_off[r], packed + word_off[t], thr, fused);
    }
    for (uint64_t j = 0; j < cnt; j++) {
      if (sc[j] == 0) continue;
      if (m == cap) { cap *= 2; g = (orc_edge*)realloc(g, cap * sizeof(orc_edge)); }
      g[m].u = (int32_t)(cand[j] >> 32);
      g[m].v = (int32_t)(uint32_t)cand[j];
      g[m].w = (float)sc[j] / 100.0f;
      m++;
    }
  }
  free(cand); free(sc);
  qsort(g, m, sizeof(orc_edge), edge_cmp);
  size_t k = 0;
  for (size_t i = 0; i < m; i++)
    if (k == 0 || edge_cmp(&g[k - 1], &g[i]) != 0) g[k++] = g[i];
  *edges = g;
  *n_edges = k;
  if (total_pairs) *total_pairs = tot;
  return 0;
}

/* main.cu:114-116 */
float orc_resolution(uint32_t identity) {
  float r = (100.0f - fabsf((float)identity - 50.0f) * 2.0f) / 100.0f;
  if (r < 0.0f) r = 0.0f;
  if (r > 1.0f) r = 1.0f;
  return r;
}

/* ------------------------------------------------------------------------ */
/* mt19937_64, parameters from the C++ standard ([rand.predef]).             */
typedef struct { uint64_t mt[312]; int idx; } mt64;
static void mt64_seed(mt64* s, uint64_t seed) {
  s->mt[0] = seed;
  for (int i = 1; i < 312; i++)
    s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
  s->idx = 312;
}
static uint64_t mt64_next(mt64* s) {
  if (s->idx >= 312) {
    for (int i = 0; i < 312; i++) {
      uint64_t x = (s->mt[i] & 0xFFFFFFFF80000000ULL) | (s->mt[(i + 1) % 312] & 0x7FFFFFFFULL);
      uint64_t xa = x >> 1;
      if (x & 1) xa ^= 0xB5026F5AA96619E9ULL;
      s->mt[i] = s->mt[(i + 156) % 312] ^ xa;
    }
    s->idx = 0;
  }
  uint64_t y = s->mt[s->idx++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= y >> 43;
  return y;
}
uint64_t orc_mt19937_64_nth(uint64_t seed, uint64_t nth) {
  mt64 s;
  mt64_seed(&s, seed);
  uint64_t v = 0;
  for (uint64_t i = 0; i < nth; i++) v = mt64_next(&s);
  return v;
}

/* ------------------------------------------------------------------------ */
/* B8: Leiden, leidenParallel.hpp:39-322.                                   */
typedef struct {
  int n;
  int* row;     /* n+1 */
  int* adj;     /* 2m */
  double* awt;  /* 2m */
  double* deg;  /* n */
  double total;
} csr_t;

typedef struct { int u, v; double w; } wedge;

static int wedge_cmp(const void* x, const void* y) {
  const wedge* a = (const wedge*)x;
  const wedge* b = (const wedge*)y;
  if (a->u != b->u) return a->u < b->u ? -1 : 1;
  if (a->v != b->v) return a->v < b->v ? -1 : 1;
  if (a->w != b->w) return a->w < b->w ? -1 : 1;
  return 0;
}

static void csr_free(csr_t* g) {
  free(g->row); free(g->adj); free(g->awt); free(g->deg);
  memset(g, 0, sizeof(*g));
}

/* leidenParallel.hpp:48-72; es is sorted and unique by (u,v), u<=v */
static csr_t csr_from_sorted(int n, const wedge* es, int m) {
  csr_t g;
  g.n = n;
  g.row = (int*)calloc((size_t)n + 1, sizeof(int));
  for (int e = 0; e < m; e++) { g.row[es[e].u + 1]++; g.row[es[e].v + 1]++; }
  for (int v = 0; v < n; v++) g.row[v + 1] += g.row[v];
  g.adj = (int*)malloc(sizeof(int) * (size_t)(2 * m + 1));
  g.awt = (double*)malloc(sizeof(double) * (size_t)(2 * m + 1));
  g.deg = (double*)calloc((size_t)n + 1, sizeof(double));
  int* cur = (int*)malloc(sizeof(int) * ((size_t)n + 1));
  for (int v = 0; v < n; v++) cur[v] = g.row[v];
  for (int e = 0; e < m; e++) {
    int u = es[e].u, v = es[e].v;
    double w = es[e].w;
    g.adj[cur[u]] = v; g.awt[cur[u]] = w; cur[u]++;
    g.adj[cur[v]] = u; g.awt[cur[v]] = w; cur[v]++;
  }
  free(cur);
  for (int v = 0; v < n; v++) {
    double s = 0;
    for (int k = g.row[v]; k < g.row[v + 1]; k++) s += g.awt[k];
    g.deg[v] = s;
  }
  double t = 0.0;
  for (int v = 0; v < n; v++) t += g.deg[v];
  g.total = t;
  return g;
}

/* sort + merge equal (u,v) summing weights in sorted order (:85-94, :250-258) */
static int merge_sorted(wedge* es, int m) {
  qsort(es, (size_t)m, sizeof(wedge), wedge_cmp);
  int k = 0;
  for (int i = 0; i < m;) {
    int j = i;
    double s = 0;
    while (j < m && es[j].u == es[i].u && es[j].v == es[i].v) { s += es[j].w; j++; }
    es[k].u = es[i].u; es[k].v = es[i].v; es[k].w = s;
    k++;
    i = j;
  }
  return k;
}

/* leidenParallel.hpp:99-113 */
static void neigh_weights(const csr_t* g, const int* memb, int v, int* cids, double* wsum,
                          int* nc) {
  int c_n = 0;
  for (int k = g->row[v]; k < g->row[v + 1]; k++) {
    int u = g->adj[k];
    if (u == v) continue;
    double w = g->awt[k];
    int c = memb[u];
    int dup = 0;
    for (int i = 0; i < c_n; i++)
      if (cids[i] == c) { wsum[i] += w; dup = 1; break; }
    if (!dup) { cids[c_n] = c; wsum[c_n] = w; c_n++; }
  }
  *nc = c_n;
}

static int max_degree_count(const csr_t* g) {
  int mx = 1;
  for (int v = 0; v < g->n; v++) {
    int d = g->row[v + 1] - g->row[v];
    if (d > mx) mx = d;
  }
  return mx;
}

/* leidenParallel.hpp:119-157 */
static int local_moving(const csr_t* g, int* memb, double* tot, double gamma) {
  const double M = g->total;
  if (M <= 0.0) return 0;
  int md = max_degree_count(g);
  int* lc = (int*)malloc(sizeof(int) * (size_t)md);
  double* lw = (double*)malloc(sizeof(double) * (size_t)md);
  int total_moves = 0;
  for (int sweep = 0; sweep < 300; sweep++) {
    int moves = 0;
    for (int v = 0; v < g->n; v++) {
      if (g->row[v] == g->row[v + 1]) continue;
      const int old = memb[v];
      const double k = g->deg[v];
      int nc;
      neigh_weights(g, memb, v, lc, lw, &nc);
      double w_old = 0.0;
      for (int i = 0; i < nc; i++) if (lc[i] == old) w_old = lw[i];
      int best = old;
      double best_gain = 0.0;
      for (int i = 0; i < nc; i++) {
        int c = lc[i];
        if (c == old) continue;
        double gain = (lw[i] - gamma * k * (tot[c] + k) / M) - (w_old - gamma * k * tot[old] / M);
        if (gain > best_gain) { best_gain = gain; best = c; }
      }
      if (best != old && best_gain > 0.0) {
        memb[v] = best;
        tot[old] -= k;
        tot[best] += k;
        moves++;
      }
    }
    total_moves += moves;
    if (moves == 0) break;
  }
  free(lc); free(lw);
  return total_moves;
}

/* leidenParallel.hpp:162-232; returns nSub, smemb renumbered */
static int refine(const csr_t* g, const int* coarse, int* smemb, double gamma, uint64_t seed) {
  const int n = g->n;
  const double M = g->total;
  if (M <= 0.0 || n == 0) return 0;
  double* stot = (double*)malloc(sizeof(double) * (size_t)n);
  int* sz = (int*)malloc(sizeof(int) * (size_t)n);
  int* order = (int*)malloc(sizeof(int) * (size_t)n);
  for (int v = 0; v < n; v++) { smemb[v] = v; stot[v] = g->deg[v]; sz[v] = 1; order[v] = v; }
  mt64 rng;
  mt64_seed(&rng, seed);
  for (size_t idx = (size_t)n - 1; idx > 0; idx--) {
    size_t ri = (size_t)(mt64_next(&rng) % (uint64_t)(idx + 1));
    int t = order[idx]; order[idx] = order[ri]; order[ri] = t;
  }
  int md = max_degree_count(g);
  int* cand = (int*)malloc(sizeof(int) * (size_t)md);
  for (int oi = 0; oi < n; oi++) {
    int v = order[oi];
    int sc = smemb[v];
    if (sz[sc] != 1) continue;
    const int cc = coarse[v];
    int kk = 0;
    for (int k = g->row[v]; k < g->row[v + 1]; k++) {
      int u = g->adj[k];
      if (u == v) continue;
      if (coarse[u] != cc) continue;
      int s2 = smemb[u];
      int dup = 0;
      for (int x = 0; x < kk; x++) if (cand[x] == s2) { dup = 1; break; }
      if (!dup) cand[kk++] = s2;
    }
    if (kk == 0) continue;
    int target = cand[(size_t)(mt64_next(&rng) % (uint64_t)kk)];
    if (target == sc) continue;
    double w_old = 0, w_new = 0;
    for (int k = g->row[v]; k < g->row[v + 1]; k++) {
      int u = g->adj[k];
      if (u == v) continue;
      double w = g->awt[k];
      int s2 = smemb[u];
      if (s2 == sc) w_old += w;
      if (s2 == target) w_new += w;
    }
    double kv = g->deg[v];
    double gain = (w_new - gamma * kv * (stot[target] + kv) / M) - (w_old - gamma * kv * stot[sc] / M);
    if (gain >= 0) {
      smemb[v] = target;
      stot[sc] -= g->deg[v];
      stot[target] += g->deg[v];
      sz[sc] = 0;
      sz[target] += 1;
    }
  }
  int* map2 = (int*)malloc(sizeof(int) * (size_t)n);
  for (int v = 0; v < n; v++) map2[v] = -1;
  int k2 = 0;
  for (int v = 0; v < n; v++) if (map2[smemb[v]] < 0) map2[smemb[v]] = k2++;
  for (int v = 0; v < n; v++) smemb[v] = map2[smemb[v]];
  free(map2); free(cand); free(order); free(sz); free(stot);
  return k2;
}

/* leidenParallel.hpp:235-260 */
static csr_t collapse(const csr_t* g, const int* memb, int ncomm) {
  int cap = g->row[g->n] + 1, m = 0;
  wedge* es = (wedge*)malloc(sizeof(wedge) * (size_t)cap);
  for (int u = 0; u < g->n; u++)
    for (int e = g->row[u]; e < g->row[u + 1]; e++) {
      int v = g->adj[e];
      if (u > v) continue;
      int cu = memb[u], cv = memb[v];
      if (cu > cv) { int t = cu; cu = cv; cv = t; }
      es[m].u = cu; es[m].v = cv; es[m].w = g->awt[e];
      m++;
    }
  m = merge_sorted(es, m);
  csr_t r = csr_from_sorted(ncomm, es, m);
  free(es);
  return r;
}

int orc_leiden(const orc_edge* edges, uint64_t m_in, int32_t n, float resolution,
               uint32_t seed, int32_t iters, uint32_t* result) {
  if (n <= 0) return 0;
  const double gamma = (double)resolution;
  wedge* es = (wedge*)malloc(sizeof(wedge) * (size_t)(m_in + 1));
  int m = 0;
  for (uint64_t i = 0; i < m_in; i++) {
    int u = edges[i].u, v = edges[i].v;
    if (u < 0 || v < 0 || u >= n || v >= n) continue;
    if (u == v) continue;
    es[m].u = u < v ? u : v;
    es[m].v = u < v ? v : u;
    es[m].w = 1.0;
    m++;
  }
  m = merge_sorted(es, m);
  csr_t g = csr_from_sorted(n, es, m);
  free(es);
  if (iters < 1) iters = 1;
  int* cur_id = (int*)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) cur_id[i] = i;
  for (int iter = 0; iter < iters; iter++) {
    for (int level = 0; level < 32; level++) {
      int* memb = (int*)malloc(sizeof(int) * ((size_t)g.n + 1));
      double* tot = (double*)malloc(sizeof(double) * ((size_t)g.n + 1));
      for (int v = 0; v < g.n; v++) { memb[v] = v; tot[v] = g.deg[v]; }
      int moved = local_moving(&g, memb, tot, gamma);
      if (moved == 0) { free(memb); free(tot); break; }
      int* smemb = (int*)malloc(sizeof(int) * ((size_t)g.n + 1));
      uint64_t s = (uint64_t)seed + (uint64_t)iter * 0x85ebca6bULL + (uint64_t)level * 0x9e3779b1ULL;
      int nsub = refine(&g, memb, smemb, gamma, s);
      csr_t g2 = collapse(&g, smemb, nsub);
      for (int i = 0; i < n; i++) cur_id[i] = smemb[cur_id[i]];
      free(memb); free(tot); free(smemb);
      csr_free(&g);
      g = g2;
      if (g.n <= 1) break;
    }
    if (g.n <= 1) break;
  }
  int* map2 = (int*)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) map2[i] = -1;
  int k = 0;
  for (int i = 0; i < n; i++) if (map2[cur_id[i]] < 0) map2[cur_id[i]] = k++;
  for (int i = 0; i < n; i++) result[i] = (uint32_t)map2[cur_id[i]];
  free(map2); free(cur_id);
  csr_free(&g);
  return 0;
}

/* B9: fastClustering.hpp:249-352 */
int orc_fast_cluster(const uint32_t* packed, const uint64_t* word_off, const uint32_t* len,
                     uint32_t n, const uint32_t* rep_rows, uint32_t bands, uint32_t identity,
                     int fused, int nthreads, uint32_t* cluster, uint64_t* total_pairs) {
  (void)len;
  const float thr = (float)identity / 100.0f;
  uint32_t* parent = (uint32_t*)malloc(((size_t)n + 1) * 4);
  uint64_t* cand = (uint64_t*)malloc(((size_t)n + 1) * 8);
  uint32_t* sc = (uint32_t*)malloc(((size_t)n + 1) * 4);
  for (uint32_t i = 0; i < n; i++) parent[i] = 0xFFFFFFFFu;
  uint64_t tot = 0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  for (uint32_t round = 0; round < bands; round++) {
    const uint32_t* rep = rep_rows + (size_t)round * n;
    uint64_t cnt = 0;
    for (uint32_t i = 0; i < n; i++)
      if (rep[i] != i && parent[i] == 0xFFFFFFFFu) cand[cnt++] = ((uint64_t)rep[i] << 32) | i;
    tot += cnt;
    #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
    for (int64_t j = 0; j < (int64_t)cnt; j++) {
      uint32_t r = (uint32_t)(cand[j] >> 32), t = (uint32_t)cand[j];
      sc[j] = orc_align_record(packed + word_off[r], packed + word_off[t], thr, fused);
    }
    for (uint64_t j = 0; j < cnt; j++)
      if (sc[j] > 0) parent[(uint32_t)cand[j]] = (uint32_t)(cand[j] >> 32);
  }
  for (uint32_t i = 0; i < n; i++)
    cluster[i] = parent[i] == 0xFFFFFFFFu ? i : cluster[parent[i]];
  free(parent); free(cand); free(sc);
  if (total_pairs) *total_pairs = tot;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* makeDB end to end, makedb.hpp:59-364, in one process with whole-file I/O. */
typedef struct {
  uint64_t fasta_off, pack_off, text_off;
  uint32_t name_len, read_len, seq_len;
  uint32_t file_order;
} orc_index;

static int index_cmp(const void* x, const void* y) {  /* stable: length desc, file order */
  const orc_index* a = (const orc_index*)x;
  const orc_index* b = (const orc_index*)y;
  if (a->read_len != b->read_len) return a->read_len > b->read_len ? -1 : 1;
  return a->file_order < b->file_order ? -1 : (a->file_order > b->file_order);
}

static int read_file(const char* path, uint8_t** data, size_t* size) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  uint8_t* d = (uint8_t*)malloc((size_t)sz + 1);
  if (sz > 0 && fread(d, 1, (size_t)sz, f) != (size_t)sz) { fclose(f); free(d); return -1; }
  fclose(f);
  *data = d;
  *size = (size_t)sz;
  return 0;
}

int orc_makedb(const char* fasta, const char* packed_path, int nthreads) {
  uint8_t* fa;
  size_t fs;
  if (read_file(fasta, &fa, &fs) != 0) return 2;
  /* makedb.hpp:68-85: first line without \r, file ends with \n */
  size_t p = 0;
  while (p < fs && fa[p] != '\n') p++;
  if (p > 0 && fa[p - 1] == '\r') { free(fa); return 2; }
  if (fs > 0 && fa[fs - 1] != '\n') { free(fa); return 2; }
  /* makedb.hpp:94-127 */
  size_t cap = 1024, n = 0;
  orc_index* idx = (orc_index*)malloc(cap * sizeof(orc_index));
  size_t pos = 0, whole = 0;
  while (pos < fs) {
    orc_index r;
    memset(&r, 0, sizeof(r));
    r.fasta_off = pos;
    size_t e = pos;
    while (e < fs && fa[e] != '\n') e++;
    r.name_len = (uint32_t)(e - pos);
    r.seq_len = r.name_len + 1;
    pos = e < fs ? e + 1 : e;
    while (pos < fs && fa[pos] != '>') {
      e = pos;
      while (e < fs && fa[e] != '\n') e++;
      r.read_len += (uint32_t)(e - pos);
      r.seq_len += (uint32_t)(e - pos) + 1;
      pos = e < fs ? e + 1 : e;
    }
    r.file_order = (uint32_t)whole;
    whole++;
    if (r.read_len < 0xFFFF) {
      if (n == cap) { cap *= 2; idx = (orc_index*)realloc(idx, cap * sizeof(orc_index)); }
      idx[n++] = r;
    }
  }
  if (n == 0 || n >= 0x7FFFFFFF) { free(idx); free(fa); return 1; }
  qsort(idx, n, sizeof(orc_index), index_cmp);
  /* makedb.hpp:137-148 */
  uint64_t base = 4 + (uint64_t)n * 24 + (uint64_t)n * 4 * ORC_BANDS;
  for (size_t i = 0; i < n; i++) {
    idx[i].pack_off = base;
    base += (uint64_t)(1 + (idx[i].read_len + 31) / 32 * 4) * 4;
  }
  for (size_t i = 0; i < n; i++) {
    idx[i].text_off = base;
    base += (uint64_t)idx[i].name_len + idx[i].read_len + 2;
  }
  uint8_t* out = (uint8_t*)calloc(base, 1);
  uint32_t n32 = (uint32_t)n;
  memcpy(out, &n32, 4);
  for (size_t i = 0; i < n; i++) {
    memcpy(out + 4 + 4 * i, &idx[i].name_len, 4);
    memcpy(out + 4 + 4 * n + 4 * i, &idx[i].read_len, 4);
    memcpy(out + 4 + 8 * n + 8 * i, &idx[i].pack_off, 8);
    memcpy(out + 4 + 16 * n + 8 * i, &idx[i].text_off, 8);
  }
  const uint64_t sign_base = 4 + (uint64_t)n * 24;
  uint32_t* sig = (uint32_t*)malloc(n * ORC_BANDS * 4);
  uint64_t pack_base = 4 + (uint64_t)n * 24 + (uint64_t)n * 4 * ORC_BANDS;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  #pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    const orc_index* r = &idx[i];
    uint8_t* res = (uint8_t*)malloc((size_t)r->read_len + 1);
    uint32_t L = 0;
    size_t q = r->fasta_off + r->name_len + 1;
    size_t end = r->fasta_off + r->seq_len;
    for (; q < end; q++) if (fa[q] != '\n') res[L++] = fa[q];
    uint32_t* rec = (uint32_t*)(out + r->pack_off);  /* 4-aligned: base and sizes are */
    orc_pack(res, L, rec);
    orc_sketch_record(rec, sig + (size_t)i * ORC_BANDS);
    uint8_t* t = out + r->text_off;
    memcpy(t, fa + r->fasta_off, r->name_len);
    t[r->name_len] = '\n';
    memcpy(t + r->name_len + 1, res, L);
    t[r->name_len + 1 + L] = '\n';
    free(res);
  }
  (void)pack_base;
  orc_bucket(sig, (uint32_t)n, (uint32_t*)(out + sign_base), nthreads);
  free(sig);
  FILE* f = fopen(packed_path, "wb");
  int rc = 0;
  if (!f || fwrite(out, 1, base, f) != base) rc = 1;
  if (f) fclose(f);
  free(out); free(idx); free(fa);
  return rc;
}

/* Appendix C, saveResult.hpp:31-83 */
int orc_save_result(const uint32_t* cluster, uint32_t n, const char* packed_path,
                    const char* result_path) {
  uint8_t* db;
  size_t ds;
  if (read_file(packed_path, &db, &ds) != 0) return 1;
  uint32_t fn;
  memcpy(&fn, db, 4);
  const uint8_t* foff = db + 4 + (size_t)fn * 16;
  uint32_t maxid = 0;
  for (uint32_t i = 0; i < n; i++) if (cluster[i] > maxid) maxid = cluster[i];
  uint32_t* cnt = (uint32_t*)calloc((size_t)maxid + 2, 4);
  for (uint32_t i = 0; i < n; i++) cnt[cluster[i]]++;
  uint32_t* start = (uint32_t*)calloc((size_t)maxid + 2, 4);
  for (uint32_t c = 0; c <= maxid; c++) start[c + 1] = start[c] + cnt[c];
  uint32_t* fill = (uint32_t*)malloc(((size_t)maxid + 1) * 4);
  memcpy(fill, start, ((size_t)maxid + 1) * 4);
  uint32_t* members = (uint32_t*)malloc(((size_t)n + 1) * 4);
  for (uint32_t i = 0; i < n; i++) members[fill[cluster[i]]++] = i;
  FILE* f = fopen(result_path, "wb");
  if (!f) { free(db); return 1; }
  for (uint32_t c = 0; c <= maxid; c++) {
    if (cnt[c] == 0) continue;
    for (uint32_t k = start[c]; k < start[c] + cnt[c]; k++) {
      uint64_t off;
      memcpy(&off, foff + 8 * (size_t)members[k], 8);
      const uint8_t* s = db + off;
      size_t l1 = 0;
      while (off + l1 < ds && s[l1] != '\n') l1++;
      if (k == start[c]) {
        fwrite(s, 1, l1, f);
        fputc('\n', f);
        const uint8_t* s2 = s + l1 + 1;
        size_t l2 = 0;
        while (off + l1 + 1 + l2 < ds && s2[l2] != '\n') l2++;
        fwrite(s2, 1, l2, f);
        fputc('\n', f);
      } else {
        fputs("  ", f);
        fwrite(s, 1, l1, f);
        fputc('\n', f);
      }
    }
  }
  fclose(f);
  free(cnt); free(start); free(fill); free(members); free(db);
  return 0;
}

/* ------------------------------------------------------------- serving */
static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : x > y;
}

/* Lowest entry of a signature in a per-band array of (sig << 32 | index)
 * sorted ascending: its index, or 0xFFFFFFFF when the signature is absent. */
static uint32_t first_of(const uint64_t* row, uint32_t len, uint32_t key) {
  const uint64_t k = (uint64_t)key << 32;
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (row[mid] < k) lo = mid + 1; else hi = mid;
  }
  if (lo == len || (uint32_t)(row[lo] >> 32) != key) return 0xFFFFFFFFu;
  return (uint32_t)row[lo];
}

void orc_serve_assign(const uint32_t* packed, const uint64_t* word_off, uint32_t n,
                      const uint32_t* sig, const uint32_t* cluster_of, const uint32_t* qpacked,
                      const uint64_t* qword_off, uint32_t nq, uint32_t identity,
                      uint32_t new_base, uint64_t first_qid, uint32_t* out, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  const float thr = (float)identity / 100.0f;
  /* Query q is snapshot index n + q: it sees the residents and queries 0..q-1. */
  uint32_t* qsig = (uint32_t*)malloc(sizeof(uint32_t) * ORC_BANDS * (nq ? nq : 1));
  #pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t q = 0; q < (int64_t)nq; q++)
    orc_sketch_record(qpacked + qword_off[q], qsig + (size_t)q * ORC_BANDS);
  /* per band: residents' and queries' (sig << 32 | i) sorted, so the first
   * entry of a signature is its lowest index (the bucket head rule) */
  uint64_t* rsorted = (uint64_t*)malloc(sizeof(uint64_t) * ORC_BANDS * (n ? n : 1));
  uint64_t* qsorted = (uint64_t*)malloc(sizeof(uint64_t) * ORC_BANDS * (nq ? nq : 1));
  #pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int b = 0; b < ORC_BANDS; b++) {
    uint64_t* row = rsorted + (size_t)b * n;
    for (uint32_t i = 0; i < n; i++) row[i] = ((uint64_t)sig[(size_t)i * ORC_BANDS + b] << 32) | i;
    qsort(row, n, sizeof(uint64_t), cmp_u64);
    uint64_t* qrow = qsorted + (size_t)b * nq;
    for (uint32_t q = 0; q < nq; q++)
      qrow[q] = ((uint64_t)qsig[(size_t)q * ORC_BANDS + b] << 32) | q;
    qsort(qrow, nq, sizeof(uint64_t), cmp_u64);
  }
  /* best[q] = (score, snapshot index) of q's best candidate; the scores need
   * no resolution, so every query is scored in parallel */
  uint32_t* best = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (nq ? nq : 1));
  #pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
  for (int64_t q = 0; q < (int64_t)nq; q++) {
    const uint32_t* qrec = qpacked + qword_off[q];
    const uint32_t Lq = qrec[0];
    const uint32_t* qs = qsig + (size_t)q * ORC_BANDS;
    uint32_t heads[ORC_BANDS];
    uint32_t nh = 0;
    for (uint32_t b = 0; b < ORC_BANDS; b++) {
      uint32_t h = first_of(rsorted + (size_t)b * n, n, qs[b]);
      if (h == 0xFFFFFFFFu) {  /* no resident: the first earlier query, if any */
        const uint32_t p = first_of(qsorted + (size_t)b * nq, nq, qs[b]);
        if (p >= (uint32_t)q) continue;
        h = n + p;
      }
      int seen = 0;
      for (uint32_t k = 0; k < nh; k++) seen |= heads[k] == h;
      if (!seen) heads[nh++] = h;
    }
    uint32_t best_score = 0, best_idx = 0xFFFFFFFFu, cand = 0;
    for (uint32_t k = 0; k < nh; k++) {
      const uint32_t h = heads[k];
      const uint32_t* hrec = h < n ? packed + word_off[h] : qpacked + qword_off[h - n];
      const uint32_t Lr = hrec[0];
      const uint32_t* rep = Lr >= Lq ? hrec : qrec;  /* ties: the earlier index */
      const uint32_t* tar = Lr >= Lq ? qrec : hrec;
      if ((float)tar[0] < (float)rep[0] * thr) continue;
      cand++;
      const uint32_t sc = orc_align_record(rep, tar, thr, 1);
      if (sc > best_score || (sc == best_score && sc > 0 && h < best_idx)) {
        best_score = sc;
        best_idx = h;
      }
    }
    best[2 * q] = best_score;
    best[2 * q + 1] = best_idx;
    out[4 * (size_t)q + 2] = best_score;
    out[4 * (size_t)q + 3] = cand;
  }
  /* clusters in arrival order: a query's best is a resident or an earlier query */
  for (uint32_t q = 0; q < nq; q++) {
    uint32_t* o = out + 4 * (size_t)q;
    const uint32_t h = best[2 * q + 1];
    if (best[2 * q] == 0) {
      o[0] = new_base + (uint32_t)(first_qid + (uint64_t)q);
      o[1] = 0xFFFFFFFFu;
    } else {
      o[0] = h < n ? cluster_of[h] : out[4 * (size_t)(h - n)];
      o[1] = h;
    }
  }
  free(best);
  free(qsorted);
  free(rsorted);
  free(qsig);
}
