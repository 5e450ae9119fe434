// Initial partitioning of the coarsest hypergraph on the host.
//
// The reference runs a recursive-bisection portfolio (BFS, greedy growing,
// random, LP, each followed by 2-way FM) on the coarsest hypergraph
// (recursive_bisection_initial_partitioner.cpp:475-492,
// flat/pool_initial_partitioner.h:36-75).  The coarsest hypergraph has at most
// 160*k vertices, so this stays a sequential host routine: a portfolio of
// seeded greedy-growing and random starts, each improved by sequential k-way
// km1 hill climbing, best balanced km1 wins.
#pragma once
#include <cstdint>
#include <vector>

#include "host_hypergraph.h"

namespace mt_kahypar_amd {

struct InitialPartitioningParams {
  int k = 2;
  std::vector<int64_t> max_w;   // per block
  uint64_t seed = 0;
  int runs = 20;                // fast_preset.ini:29 (i-runs)
  int refinement_passes = 8;
};

// Returns a block id per vertex; *km1_out receives its km1.
std::vector<int32_t> initial_partition(const HostHypergraph& hg, const InitialPartitioningParams& prm,
                                       int64_t* km1_out);

}  // namespace mt_kahypar_amd
