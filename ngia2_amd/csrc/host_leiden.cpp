// host_leiden.cpp -- Leiden community detection on the edge list, host C++.
//
// Same partition as LeidenOmp::cluster (leidenParallel.hpp:265-322): an
// unweighted undirected graph, RB-configuration quality with resolution
// gamma, sequential local moving accepting strictly positive gains (ties to
// the first neighbouring community met), a refinement that merges singleton
// sub-communities in an mt19937_64-shuffled order accepting gains >= 0, and
// aggregation on the refined partition, for at most 32 levels.
//
// Representation: every graph is kept in CSR form with each node's neighbours
// in ascending id order and a self-loop listed twice -- exactly the adjacency
// order the reference builds from its sorted canonical edge list
// (leidenParallel.hpp:48-72), so neighbour-discovery order, and with it every
// tie-break, is the same.  Edge weights are integer counts held in doubles;
// every sum of them is exact, so only the gain expressions round, and those
// are evaluated in the reference's operation order.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "host_common.hpp"

namespace {

struct Graph {
  int n = 0;
  std::vector<int> row;       // n+1
  std::vector<int> adj;       // neighbour ids, ascending per node, self-loop twice
  std::vector<double> wt;     // weight per adjacency entry
  std::vector<double> deg;    // sum of wt over the row
  double total = 0.0;
};

// Build from canonical (u <= v) weighted pairs that are sorted by (u,v) and
// unique.  Entries of node x: pairs (u,x) with u < x in u order, then the
// self-loop (x,x) twice, then pairs (x,v) in v order.
Graph build_graph(int n, const std::vector<std::pair<uint64_t, double>>& pairs) {
  Graph g;
  g.n = n;
  g.row.assign((size_t)n + 1, 0);
  for (const auto& p : pairs) {
    const int u = (int)(p.first >> 32), v = (int)(uint32_t)p.first;
    g.row[u + 1]++;
    g.row[v + 1]++;
  }
  for (int v = 0; v < n; v++) g.row[v + 1] += g.row[v];
  g.adj.resize(g.row[n]);
  g.wt.resize(g.row[n]);
  std::vector<int> cur(g.row.begin(), g.row.end() - 1);
  for (const auto& p : pairs) {
    const int u = (int)(p.first >> 32), v = (int)(uint32_t)p.first;
    g.adj[cur[u]] = v;
    g.wt[cur[u]++] = p.second;
    g.adj[cur[v]] = u;
    g.wt[cur[v]++] = p.second;
  }
  g.deg.assign(n, 0.0);
  double t = 0.0;
  for (int v = 0; v < n; v++) {
    double s = 0;
    for (int k = g.row[v]; k < g.row[v + 1]; k++) s += g.wt[k];
    g.deg[v] = s;
    t += s;
  }
  g.total = t;
  return g;
}

// Sort (key = u<<32|v, w) and merge equal keys, summing weights.
void merge_pairs(std::vector<std::pair<uint64_t, double>>* es) {
  std::sort(es->begin(), es->end());
  size_t k = 0;
  for (size_t i = 0; i < es->size();) {
    size_t j = i;
    double s = 0;
    while (j < es->size() && (*es)[j].first == (*es)[i].first) s += (*es)[j++].second;
    (*es)[k++] = {(*es)[i].first, s};
    i = j;
  }
  es->resize(k);
}

// Neighbouring communities of v with summed weights, in discovery order.
struct NeighCollector {
  std::vector<int> cid;
  std::vector<double> w;
  std::vector<int> slot;  // community -> index in cid, -1 when absent
  explicit NeighCollector(int n) : slot(n, -1) {}
  void collect(const Graph& g, const std::vector<int>& memb, int v) {
    for (int c : cid) slot[c] = -1;
    cid.clear();
    w.clear();
    for (int k = g.row[v]; k < g.row[v + 1]; k++) {
      const int u = g.adj[k];
      if (u == v) continue;
      const int c = memb[u];
      if (slot[c] < 0) {
        slot[c] = (int)cid.size();
        cid.push_back(c);
        w.push_back(g.wt[k]);
      } else {
        w[slot[c]] += g.wt[k];
      }
    }
  }
};

// leidenParallel.hpp:119-157
int local_moving(const Graph& g, std::vector<int>& memb, std::vector<double>& tot, double gamma) {
  const double M = g.total;
  if (M <= 0.0) return 0;
  NeighCollector nc(g.n);
  int total_moves = 0;
  for (int sweep = 0; sweep < 300; sweep++) {
    int moves = 0;
    for (int v = 0; v < g.n; v++) {
      if (g.row[v] == g.row[v + 1]) continue;
      const int old = memb[v];
      const double k = g.deg[v];
      nc.collect(g, memb, v);
      const double w_old = nc.slot[old] >= 0 ? nc.w[nc.slot[old]] : 0.0;
      int best = old;
      double best_gain = 0.0;
      for (size_t i = 0; i < nc.cid.size(); i++) {
        const int c = nc.cid[i];
        if (c == old) continue;
        const double gain = (nc.w[i] - gamma * k * (tot[c] + k) / M) - (w_old - gamma * k * tot[old] / M);
        if (gain > best_gain) {
          best_gain = gain;
          best = c;
        }
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
  return total_moves;
}

// leidenParallel.hpp:162-232; returns the refined partition, renumbered by
// first appearance, and its community count.
int refine(const Graph& g, const std::vector<int>& coarse, std::vector<int>& sub, double gamma,
           uint64_t seed) {
  const int n = g.n;
  const double M = g.total;
  sub.resize(n);
  for (int v = 0; v < n; v++) sub[v] = v;
  if (M <= 0.0 || n == 0) return n;
  std::vector<double> stot(g.deg);
  std::vector<int> size(n, 1);
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::mt19937_64 rng(seed);
  for (size_t idx = order.size() - 1; idx > 0; idx--) {
    const size_t ri = (size_t)(rng() % (idx + 1));
    std::swap(order[idx], order[ri]);
  }
  std::vector<int> cand;
  for (int v : order) {
    const int sc = sub[v];
    if (size[sc] != 1) continue;
    const int cc = coarse[v];
    cand.clear();
    for (int k = g.row[v]; k < g.row[v + 1]; k++) {
      const int u = g.adj[k];
      if (u == v || coarse[u] != cc) continue;
      const int s2 = sub[u];
      if (std::find(cand.begin(), cand.end(), s2) == cand.end()) cand.push_back(s2);
    }
    if (cand.empty()) continue;
    const int target = cand[(size_t)(rng() % cand.size())];
    if (target == sc) continue;
    double w_old = 0, w_new = 0;
    for (int k = g.row[v]; k < g.row[v + 1]; k++) {
      const int u = g.adj[k];
      if (u == v) continue;
      const int s2 = sub[u];
      if (s2 == sc) w_old += g.wt[k];
      if (s2 == target) w_new += g.wt[k];
    }
    const double kv = g.deg[v];
    const double gain = (w_new - gamma * kv * (stot[target] + kv) / M) - (w_old - gamma * kv * stot[sc] / M);
    if (gain >= 0) {
      sub[v] = target;
      stot[sc] -= g.deg[v];
      stot[target] += g.deg[v];
      size[sc] = 0;
      size[target] += 1;
    }
  }
  std::vector<int> map(n, -1);
  int k2 = 0;
  for (int v = 0; v < n; v++)
    if (map[sub[v]] < 0) map[sub[v]] = k2++;
  for (int v = 0; v < n; v++) sub[v] = map[sub[v]];
  return k2;
}

// leidenParallel.hpp:235-260: one canonical pair per adjacency entry with
// u <= v (a self-loop's two entries both count), mapped through the partition.
Graph aggregate(const Graph& g, const std::vector<int>& part, int ncomm) {
  std::vector<std::pair<uint64_t, double>> es;
  es.reserve(g.adj.size() / 2 + 1);
  for (int u = 0; u < g.n; u++)
    for (int e = g.row[u]; e < g.row[u + 1]; e++) {
      const int v = g.adj[e];
      if (u > v) continue;
      int cu = part[u], cv = part[v];
      if (cu > cv) std::swap(cu, cv);
      es.emplace_back(((uint64_t)(uint32_t)cu << 32) | (uint32_t)cv, g.wt[e]);
    }
  merge_pairs(&es);
  return build_graph(ncomm, es);
}

}  // namespace

extern "C" float ngia_resolution(uint32_t identity) {
  // main.cu:114-116
  float r = (100.0f - std::fabs((float)identity - 50.0f) * 2.0f) / 100.0f;
  if (r < 0.0f) r = 0.0f;
  return std::min(r, 1.0f);
}

extern "C" int ngia_leiden(const ngia_edge* edges, uint64_t m, int32_t n, float resolution,
                           uint32_t seed, int32_t iters, uint32_t* result) {
  if (n <= 0) return NGIA_OK;
  if (!result || (m && !edges)) {
    ngia::set_thread_error("null argument");
    return NGIA_ERR_ARG;
  }
  const double gamma = (double)resolution;
  std::vector<std::pair<uint64_t, double>> es;
  es.reserve(m);
  for (uint64_t i = 0; i < m; i++) {
    int u = edges[i].u, v = edges[i].v;
    if (u < 0 || v < 0 || u >= n || v >= n || u == v) continue;
    if (u > v) std::swap(u, v);
    es.emplace_back(((uint64_t)(uint32_t)u << 32) | (uint32_t)v, 1.0);
  }
  merge_pairs(&es);
  Graph g = build_graph(n, es);
  if (iters < 1) iters = 1;
  std::vector<int> cur(n);
  for (int i = 0; i < n; i++) cur[i] = i;
  for (int iter = 0; iter < iters; iter++) {
    for (int level = 0; level < 32; level++) {
      std::vector<int> memb(g.n);
      for (int v = 0; v < g.n; v++) memb[v] = v;
      std::vector<double> tot(g.deg);
      if (local_moving(g, memb, tot, gamma) == 0) break;
      std::vector<int> sub;
      const uint64_t s = (uint64_t)seed + (uint64_t)iter * 0x85ebca6bULL + (uint64_t)level * 0x9e3779b1ULL;
      const int nsub = refine(g, memb, sub, gamma, s);
      Graph g2 = aggregate(g, sub, nsub);
      for (int i = 0; i < n; i++) cur[i] = sub[cur[i]];
      g = std::move(g2);
      if (g.n <= 1) break;
    }
    if (g.n <= 1) break;
  }
  std::vector<int> map(n, -1);
  int k = 0;
  for (int i = 0; i < n; i++)
    if (map[cur[i]] < 0) map[cur[i]] = k++;
  for (int i = 0; i < n; i++) result[i] = (uint32_t)map[cur[i]];
  return NGIA_OK;
}
