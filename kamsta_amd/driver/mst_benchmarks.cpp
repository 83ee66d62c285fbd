// mst_benchmarks.cpp -- command-line driver with the reference driver's flag
// names (benchmarks/mst.cpp:40-127) over the C++ drop-in header.
//
//   mst_benchmarks --log_n 16 --log_m 20 -a filter_hybridBoruvka -i 3 -c
//   mst_benchmarks -g INFILE_WEIGHTED -f graph.bin --symmetrize -a hybridBoruvka
//
// Graphs: GNM from the library's deterministic generator, or an
// INFILE_WEIGHTED binary edge list (9-byte records: u32 src, u32 dst,
// u8 weight; evaluation/README.md).  Each iteration times the entry call
// alone with steady_clock, the window of the reference's `mst_complete`
// timer (benchmarks/mst.cpp:212-230), and prints one RESULT line in the
// reference timer's key=value style (include/util/timer.cpp:204-235).
// --checks runs the other algorithm on the same input and requires an
// identical edge set, a forest (union-find over the result) and u64 weight
// agreement.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/kmst_hybrid_mst.hpp"

namespace {

struct WEdge {  // layout and accessors of hybridMST::WEdge (include/definitions.hpp:66-84)
  uint64_t src_, dst_;
  uint32_t weight_;
  WEdge() = default;
  WEdge(uint64_t s, uint64_t d, uint32_t w) : src_(s), dst_(d), weight_(w) {}
  uint64_t get_src() const { return src_; }
  uint64_t get_dst() const { return dst_; }
  uint32_t get_weight() const { return weight_; }
};

struct AlgorithmConfig {  // include/algorithms/algorithm_configurations.hpp:28-34
  int local_preprocessing = 0;
  int compression = 0;
  std::size_t filter_threshold = 4;
};

struct Params {
  std::string algorithm = "hybridBoruvka";
  std::string graphtype = "GNM";
  std::vector<std::string> infiles;
  unsigned log_n = 10, log_m = 12;
  unsigned iterations = 1;
  unsigned max_edge_weight = 254;
  std::size_t filter_threshold = 4;
  unsigned local_kernelization_level = 1;
  unsigned compression_level = 0;
  uint64_t seed = 1;
  bool checks = false;
  bool symmetrize = false;
};

void usage() {
  std::puts(
      "mst_benchmarks [-a hybridBoruvka|filter_hybridBoruvka] [-g GNM|INFILE_WEIGHTED] [-f file]\n"
      "               [--log_n N] [--log_m M] [-i iterations] [--filter_threshold T]\n"
      "               [--max_edge_weight W] [--local_kernelization_level L]\n"
      "               [--compression_level 0|1] [--seed S] [--symmetrize] [-c|--checks]");
}

bool parse(int argc, char** argv, Params& p) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* name) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", name);
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "-h" || a == "--help") return false;
    if (a == "-a" || a == "--algorithm") p.algorithm = next("--algorithm");
    else if (a == "-g" || a == "--graphtype") p.graphtype = next("--graphtype");
    else if (a == "-f" || a == "--infile") p.infiles.push_back(next("--infile"));
    else if (a == "--log_n") p.log_n = std::stoul(next("--log_n"));
    else if (a == "--log_m") p.log_m = std::stoul(next("--log_m"));
    else if (a == "-i" || a == "--iterations") p.iterations = std::stoul(next("--iterations"));
    else if (a == "--filter_threshold") p.filter_threshold = std::stoull(next("--filter_threshold"));
    else if (a == "--max_edge_weight") p.max_edge_weight = std::stoul(next("--max_edge_weight"));
    else if (a == "--local_kernelization_level")
      p.local_kernelization_level = std::stoul(next("--local_kernelization_level"));
    else if (a == "--compression_level") p.compression_level = std::stoul(next("--compression_level"));
    else if (a == "--seed") p.seed = std::stoull(next("--seed"));
    else if (a == "--symmetrize") p.symmetrize = true;
    else if (a == "-c" || a == "--checks") p.checks = true;
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return false;
    }
  }
  return true;
}

std::vector<WEdge> read_weighted_binary(const std::vector<std::string>& files, bool symmetrize) {
  std::vector<WEdge> edges;
  for (const auto& f : files) {
    std::ifstream in(f, std::ios::binary);
    if (!in) {
      std::fprintf(stderr, "cannot open %s\n", f.c_str());
      std::exit(2);
    }
    unsigned char rec[9];
    while (in.read(reinterpret_cast<char*>(rec), 9)) {
      uint32_t s, d;
      std::memcpy(&s, rec, 4);
      std::memcpy(&d, rec + 4, 4);
      if (s == d) continue;  // the MST input contract has no self-loops
      edges.emplace_back(s, d, rec[8]);
      if (symmetrize) edges.emplace_back(d, s, rec[8]);
    }
  }
  std::sort(edges.begin(), edges.end(), [](const WEdge& a, const WEdge& b) {
    return std::tie(a.src_, a.dst_, a.weight_) < std::tie(b.src_, b.dst_, b.weight_);
  });
  if (symmetrize) {  // identical records appear once per direction
    edges.erase(std::unique(edges.begin(), edges.end(),
                            [](const WEdge& a, const WEdge& b) {
                              return a.src_ == b.src_ && a.dst_ == b.dst_ && a.weight_ == b.weight_;
                            }),
                edges.end());
  }
  return edges;
}

std::vector<WEdge> gnm(const Params& p) {
  const uint64_t n = uint64_t{1} << p.log_n, m = uint64_t{1} << p.log_m;
  std::vector<uint32_t> s(2 * m), d(2 * m), w(2 * m);
  if (kmst_gen_gnm_host(n, m, p.max_edge_weight, p.seed, s.data(), d.data(), w.data()) != KMST_OK) {
    std::fprintf(stderr, "generator: %s\n", kmst_last_error());
    std::exit(2);
  }
  std::vector<WEdge> e(2 * m);
  for (uint64_t i = 0; i < 2 * m; ++i) e[i] = WEdge(s[i], d[i], w[i]);
  return e;
}

uint64_t weight_sum(const std::vector<WEdge>& e) {  // sum_edge_weights (include/util/utils.hpp:53-59)
  uint64_t t = 0;
  for (const auto& x : e) t += x.weight_;
  return t;
}

bool is_forest(const std::vector<WEdge>& mst) {
  uint64_t n = 0;
  for (const auto& e : mst) n = std::max<uint64_t>(n, std::max(e.src_, e.dst_) + 1);
  std::vector<uint64_t> p(n);
  std::iota(p.begin(), p.end(), 0);
  auto find = [&](uint64_t x) {
    while (p[x] != x) x = p[x] = p[p[x]];
    return x;
  };
  for (const auto& e : mst) {
    const uint64_t a = find(e.src_), b = find(e.dst_);
    if (a == b) return false;
    p[a] = b;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  Params p;
  if (!parse(argc, argv, p)) {
    usage();
    return argc > 1 && std::string(argv[1]) != "-h" && std::string(argv[1]) != "--help" ? 2 : 0;
  }
  if (p.algorithm != "hybridBoruvka" && p.algorithm != "filter_hybridBoruvka") {
    std::fprintf(stderr, "unsupported algorithm %s\n", p.algorithm.c_str());
    return 2;
  }
  std::vector<WEdge> input;
  if (p.graphtype == "GNM") input = gnm(p);
  else if (p.graphtype == "INFILE_WEIGHTED") input = read_weighted_binary(p.infiles, p.symmetrize);
  else {
    std::fprintf(stderr, "unsupported graph type %s (GNM, INFILE_WEIGHTED)\n", p.graphtype.c_str());
    return 2;
  }
  AlgorithmConfig cfg;
  cfg.local_preprocessing = static_cast<int>(p.local_kernelization_level);
  cfg.compression = static_cast<int>(p.compression_level);
  cfg.filter_threshold = p.filter_threshold;
  const bool filter = p.algorithm == "filter_hybridBoruvka";
  for (unsigned it = 0; it < p.iterations; ++it) {
    std::vector<WEdge> copy = input;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<WEdge> mst = filter ? kmst_hybrid::filter_boruvka(std::move(copy), cfg)
                                    : kmst_hybrid::boruvka(std::move(copy), cfg);
    const auto t1 = std::chrono::steady_clock::now();
    const double secs = std::chrono::duration<double>(t1 - t0).count();
    const uint64_t total = weight_sum(mst);
    std::printf("RESULT algorithm=%s graphtype=%s log_n=%u log_m=%u iteration=%u edges=%zu "
                "mst_edges=%zu mst_weight=%llu mst_complete=%.6f input_edges_per_s=%.1f\n",
                p.algorithm.c_str(), p.graphtype.c_str(), p.log_n, p.log_m, it, input.size(),
                mst.size(), static_cast<unsigned long long>(total), secs,
                double(input.size() / 2) / secs);
    if (p.checks) {
      std::vector<WEdge> other = filter ? kmst_hybrid::boruvka(input, cfg)
                                        : kmst_hybrid::filter_boruvka(input, cfg);
      bool same = other.size() == mst.size();
      for (std::size_t i = 0; same && i < mst.size(); ++i)
        same = mst[i].src_ == other[i].src_ && mst[i].dst_ == other[i].dst_ &&
               mst[i].weight_ == other[i].weight_;
      const bool forest = is_forest(mst);
      std::printf("CHECK identical_across_algorithms=%d forest=%d weight=%llu\n", same ? 1 : 0,
                  forest ? 1 : 0, static_cast<unsigned long long>(total));
      if (!same || !forest) {
        std::fprintf(stderr, "wrong result\n");
        return 1;
      }
    }
  }
  return 0;
}
