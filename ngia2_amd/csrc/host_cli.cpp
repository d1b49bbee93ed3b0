// host_cli.cpp -- the ngia2 command line (main.cu:27-164), drop-in for both
// spellings the reference documents:
//   ngia2 -m makedb|clustering|fast -f fasta -p packed -r result -i identity -b bands
//   makeDB -f fasta -p packed            (README.md:27; this binary named makeDB)
//   cluster -i identity -p packed -r result   (README.md:36; named cluster)
// Flag parsing keeps the reference parser's behaviour (parser.hpp:35-67):
// argv[1..argc-2] is scanned, the last occurrence of a flag wins, unknown
// flags are ignored, a trailing flag without a value is ignored and values
// convert with istringstream.  Exit codes: usage / check failures 0, runtime
// failures 1.  Extra flags of this build: -g plain|batched, -d device.
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "ngia2.h"

namespace {

struct Param {
  bool required;
  std::string name, describe, value;
};

struct Parser {
  std::vector<Param> ps;
  void add(bool req, const std::string& name, const std::string& d, const std::string& def) {
    ps.push_back({req, name, d, def});
  }
  bool parse(int argc, char** argv) {
    for (Param& p : ps) {
      for (int i = 1; i < argc - 1; i++)
        if (p.name == argv[i]) p.value = argv[i + 1];
      if (p.required && p.value.empty()) {
        std::cout << "usage: " << argv[0] << " ... (* is necessary)\n";
        for (const Param& q : ps)
          if (q.required) std::cout << "* " << q.name << "\t" << q.describe << "\n";
        for (const Param& q : ps)
          if (!q.required) std::cout << "  " << q.name << "\t" << q.describe << "\n";
        return false;
      }
    }
    return true;
  }
  template <class T>
  T get(const std::string& name) const {
    std::string content;
    for (const Param& p : ps)
      if (p.name == name) content = p.value;
    T v{};
    std::istringstream iss(content);
    iss >> v;
    return v;
  }
};

std::string base_name(const char* p) {
  const char* s = std::strrchr(p, '/');
  return s ? s + 1 : p;
}

}  // namespace

int main(int argc, char** argv) {
  // Drop-in aliases: the program name, or a leading subcommand word.
  std::string alias = base_name(argv[0]);
  if (argc > 1 && argv[1][0] != '-') alias = argv[1];
  std::string mode;
  if (alias == "makeDB" || alias == "makedb") mode = "makedb";
  if (alias == "cluster" || alias == "clustering") mode = "clustering";
  if (alias == "fast") mode = "fast";
  const bool mode_from_name = !mode.empty();

  Parser parser;
  parser.add(!mode_from_name, "-m", "mode (string) makedb/clustering/fast", mode);
  parser.add(false, "-f", "fasta file (string)", "data.fasta");
  parser.add(false, "-p", "packed file (string)", "packed.bin");
  parser.add(false, "-r", "result file (string)", "result.txt");
  parser.add(false, "-i", "identity (1-99)", "95");
  parser.add(false, "-b", "bands (1-64)", "64");
  parser.add(false, "-g", "graph form (string) batched/plain", "batched");
  parser.add(false, "-d", "device (int)", "0");
  if (!parser.parse(argc, argv)) return 0;
  mode = parser.get<std::string>("-m");
  const std::string fasta = parser.get<std::string>("-f");
  const std::string packed = parser.get<std::string>("-p");
  const std::string result = parser.get<std::string>("-r");
  const uint32_t identity = parser.get<uint32_t>("-i");
  const uint32_t bands = parser.get<uint32_t>("-b");
  const uint32_t gflags = parser.get<std::string>("-g") == "plain" ? NGIA_GRAPH_PLAIN : NGIA_GRAPH_BATCHED;
  const int device = parser.get<int>("-d");

  if (mode != "makedb" && mode != "clustering" && mode != "fast") {
    std::cout << "Usage:\n" << argv[0] << " -m makedb/clustering/fast\n";
    return 0;
  }
  ngia_ctx* ctx = nullptr;
  auto open_ctx = [&]() -> bool {
    int rc = ngia_open(device, &ctx);
    if (rc != NGIA_OK) {
      std::cout << "device " << device << ": " << ngia_last_error(nullptr) << "\n";
      return false;
    }
    ngia_device_info info;
    if (ngia_device_query(ctx, &info) == NGIA_OK) std::cout << "use gpu:\t" << info.name << " (" << info.arch << ")\n";
    return true;
  };

  if (mode == "makedb") {
    std::cout << "fasta:\t" << fasta << "\npacked:\t" << packed << "\n";
    // MakeDB::check runs before any device work (main.cu:56-60).
    if (ngia_fasta_check(fasta.c_str()) != NGIA_OK) {
      std::cout << ngia_last_error(nullptr) << "\n\nUsage:\n" << argv[0]
                << " -m makedb -f fastaFile -p packedFile\n";
      return 0;
    }
    if (!open_ctx()) return 1;
    int rc = ngia_makedb(ctx, fasta.c_str(), packed.c_str());
    if (rc == NGIA_ERR_ARG) {  // MakeDB::check failed (makedb.hpp:59-91, main.cu:56-60)
      std::cout << ngia_last_error(nullptr) << "\n\nUsage:\n" << argv[0]
                << " -m makedb -f fastaFile -p packedFile\n";
      ngia_close(ctx);
      return 0;
    }
    if (rc != NGIA_OK) {
      std::cout << ngia_last_error(ctx) << ngia_last_error(nullptr) << "\nmakedb failed, abort.\n";
      ngia_close(ctx);
      return 1;
    }
    ngia_close(ctx);
    return 0;
  }

  std::cout << "packed file:\t" << packed << "\nresult file:\t" << result << "\nidentity:\t" << identity
            << "\n";
  // Parameter checks come before any device work (buildGraph.hpp:74-99).
  if (identity < 1 || identity > 99 || bands < 1 || bands > NGIA_BANDS) {
    if (identity < 1 || identity > 99) std::cout << "identity should be 1-99\n";
    if (bands < 1 || bands > NGIA_BANDS) std::cout << "bands should be 1-" << NGIA_BANDS << "\n";
    std::cout << "\nUsage:\n" << argv[0] << " -m " << mode << " -p packedFile -r resultFile -i identity -b bands\n";
    return 0;
  }
  {
    FILE* f = std::fopen(packed.c_str(), "rb");
    uint32_t c = 0;
    const bool ok = f && std::fread(&c, 4, 1, f) == 1 && c > 0;
    if (f) std::fclose(f);
    if (!ok) {
      std::cout << packed << (f ? " contains 0 sequences.\n" : " does not exists.\n");
      std::cout << "\nUsage:\n" << argv[0] << " -m " << mode << " -p packedFile -r resultFile -i identity -b bands\n";
      return 0;
    }
  }
  if (!open_ctx()) return 1;
  int rc = NGIA_OK;
  if (mode == "clustering") {
    std::cout << "Building graph..." << std::endl;
    ngia_graph_stats st{};
    rc = ngia_cluster(ctx, packed.c_str(), result.c_str(), identity, bands, gflags, &st);
    if (rc == NGIA_OK)
      std::cout << "compare pairs:\t" << st.total_pairs << "\ndistinct pairs:\t" << st.distinct_pairs
                << "\nedges:\t" << st.n_edges << "\n[align " << bands << " bands]\t" << st.ms_total
                << " ms\n";
  } else {  // fast (main.cu:130-156)
    ngia_packed_db pdb;
    rc = ngia_read_db(packed.c_str(), bands, &pdb);
    if (rc == NGIA_OK) {
      ngia_db* db = nullptr;
      std::vector<uint32_t> cluster(pdb.n);
      uint64_t pairs = 0;
      rc = ngia_db_create(ctx, pdb.packed, pdb.n_words, pdb.word_off, pdb.lengths, pdb.n, &db);
      if (rc == NGIA_OK) rc = ngia_db_set_reps(ctx, db, pdb.rep_rows, bands);
      if (rc == NGIA_OK) rc = ngia_db_fast(ctx, db, identity, bands, cluster.data(), &pairs);
      ngia_db_destroy(db);
      if (rc == NGIA_OK) {
        uint64_t nc = 0;
        for (uint32_t i = 0; i < pdb.n; i++) nc += cluster[i] == i;
        std::cout << "compare pairs:\t" << pairs << "\nclusters:\t" << nc << "\n";
        rc = ngia_save_result(cluster.data(), pdb.n, packed.c_str(), result.c_str());
      }
      ngia_release_db(&pdb);
    }
  }
  if (rc != NGIA_OK) {
    std::cout << ngia_last_error(ctx) << " " << ngia_last_error(nullptr) << "\n" << mode << " failed, abort.\n";
    ngia_close(ctx);
    return 1;
  }
  ngia_close(ctx);
  return 0;
}
