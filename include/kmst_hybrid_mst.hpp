// kmst_hybrid_mst.hpp -- C++ drop-in for KaMSTa's algorithm entry points.
//
// Reproduces the two template signatures the reference driver calls
// (benchmarks/mst.cpp:215-228):
//
//   template <typename InputEdges, typename AlgorithmConfigType>
//   WEdgeList hybridMST::boruvka(InputEdges, const AlgorithmConfigType&);
//       include/algorithms/hybrid_boruvka.hpp:263-301
//   template <typename InputEdges, typename AlgorithmConfigType>
//   WEdgeList hybridMST::filter_boruvka(InputEdges, const AlgorithmConfigType&);
//       include/algorithms/hybrid_boruvka_edgefilter.hpp:420-462
//
// InputEdges is a std::vector of an edge type with get_src / get_dst /
// get_weight and an (src, dst, weight) constructor, e.g. hybridMST::WEdge
// (include/definitions.hpp:66-84).  The input is taken by value and
// consumed, as in the reference.  The result is a vector of the same edge
// type holding each MST edge once as (min, max, weight), sorted by
// (min, max).  Like the reference, a failure aborts the process (after
// printing the library's message); kmst_boruvka / kmst_filter_boruvka return
// status codes for callers that prefer them.
//
// One process, one GPU: the MPI rank layout of the reference does not apply;
// call this from a single rank.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "kmst.h"

namespace kmst_hybrid {

namespace detail {

template <typename Config>
kmst_config to_config(const Config& c, int algorithm) {
  kmst_config k;
  kmst_config_default(&k);
  k.algorithm = algorithm;
  k.local_preprocessing = static_cast<int32_t>(c.local_preprocessing);
  k.compression = static_cast<int32_t>(c.compression);
  k.filter_threshold = static_cast<uint64_t>(c.filter_threshold);
  return k;
}

template <typename InputEdges, typename Config>
std::vector<typename InputEdges::value_type> solve(InputEdges input_edges, const Config& cfg,
                                                   int algorithm) {
  using Edge = typename InputEdges::value_type;
  std::vector<kmst_wedge> in(input_edges.size());
  for (std::size_t i = 0; i < input_edges.size(); ++i) {
    const auto& e = input_edges[i];
    in[i] = kmst_wedge{static_cast<uint64_t>(e.get_src()), static_cast<uint64_t>(e.get_dst()),
                       static_cast<uint32_t>(e.get_weight()), 0u};
  }
  InputEdges().swap(input_edges);  // consumed, like the reference's move into the graph
  std::vector<kmst_wedge> out(in.size() / 2 + 1);
  uint64_t n_out = 0, total = 0;
  const kmst_config k = to_config(cfg, algorithm);
  const kmst_status st = algorithm == KMST_ALGO_BORUVKA
                             ? kmst_boruvka(in.data(), in.size(), &k, out.data(), &n_out, &total)
                             : kmst_filter_boruvka(in.data(), in.size(), &k, out.data(), &n_out,
                                                   &total);
  if (st != KMST_OK) {
    std::fprintf(stderr, "kmst: status %d: %s\n", static_cast<int>(st), kmst_last_error());
    std::abort();
  }
  std::vector<Edge> result;
  result.reserve(n_out);
  for (uint64_t i = 0; i < n_out; ++i) result.emplace_back(out[i].src, out[i].dst, out[i].weight);
  return result;
}

}  // namespace detail

template <typename InputEdges, typename AlgorithmConfigType>
std::vector<typename InputEdges::value_type> boruvka(InputEdges input_edges,
                                                     const AlgorithmConfigType& cfg) {
  return detail::solve(std::move(input_edges), cfg, KMST_ALGO_BORUVKA);
}

template <typename InputEdges, typename AlgorithmConfigType>
std::vector<typename InputEdges::value_type> filter_boruvka(InputEdges input_edges,
                                                            const AlgorithmConfigType& cfg) {
  return detail::solve(std::move(input_edges), cfg, KMST_ALGO_FILTER_BORUVKA);
}

}  // namespace kmst_hybrid
