// ref_driver.cpp -- entry points into the reference's host-only headers,
// compiled from /root/reference/src as they lie (no stand-ins: these headers
// need only the C++ standard library and OpenMP).  The CUDA half of the
// reference (main.cu, buildGraph.hpp, fastClustering.hpp) needs cuda.h and
// nvcc and is not built.  TEST INFRASTRUCTURE ONLY: the product never loads
// this library.
#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

#include "makedb.hpp"
#include "leidenParallel.hpp"
#include "saveResult.hpp"

extern "C" {

// main.cu:54-83, the makedb branch.  0 ok, 1 runtime failure, 2 check failure
// (which main.cu reports with the usage text and exit status 0).
int ref_makedb(const char* fasta, const char* packed) {
  if (!MakeDB::check(fasta, packed)) return 2;
  std::vector<MakeDB::seqIndex> idx = MakeDB::makeIndex(fasta);
  size_t n = idx.size();
  if (n == 0 || n >= 0x7FFFFFFF) return 1;
  if (!MakeDB::packData(fasta, packed, idx)) return 1;
  if (!MakeDB::groupSignatures(packed, n)) return 1;
  return 0;
}

// leidenParallel.hpp:265-322 on an edge list given as three arrays.
int ref_leiden(const int32_t* u, const int32_t* v, const float* w, uint64_t m, int32_t n,
               float resolution, uint32_t seed, int32_t iters, uint32_t* out) {
  std::vector<std::tuple<int, int, float>> edges;
  edges.reserve(m);
  for (uint64_t i = 0; i < m; i++) edges.emplace_back(u[i], v[i], w[i]);
  std::vector<uint32_t> r = LeidenOmp::cluster(edges, n, resolution, seed, iters);
  for (size_t i = 0; i < r.size(); i++) out[i] = r[i];
  return (int)r.size();
}

// saveResult.hpp:31-83
void ref_save(const uint32_t* cluster, uint32_t n, const char* packed, const char* result) {
  std::vector<uint32_t> c(cluster, cluster + n);
  SaveResult::save(c, packed, result);
}

}  // extern "C"
