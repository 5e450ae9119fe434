// Initial partitioning of the coarsest hypergraph on the host.
//
// The reference runs recursive bisection with a portfolio of flat 2-way
// algorithms (BFS, greedy growing, random, LP, each followed by 2-way FM) on
// the coarsest hypergraph (recursive_bisection_initial_partitioner.cpp:475-492,
// flat/pool_initial_partitioner.h:36-75, i-mode=recursive_bisection in
// fast_preset.ini).  The coarsest hypergraph has at most 160*k vertices, so
// this stays a sequential host routine: every bisection takes the best
// balanced cut of seeded greedy-growing, BFS and random starts, each improved
// by 2-way gain moves; cut nets are split between the two sides (the km1
// recursion), and each level gets the adaptive imbalance
// (1 + eps)^(1 / ceil(log2 k)) - 1.
#pragma once
#include <cstdint>
#include <vector>

#include "host_hypergraph.h"

namespace mt_kahypar_amd {

struct InitialPartitioningParams {
  int k = 2;
  std::vector<int64_t> max_w;   // per block
  uint64_t seed = 0;
  double epsilon = -1;          // < 0: max_w[0] / ceil(W / k) - 1
  int runs = 20;                // fast_preset.ini:29 (i-runs), per bisection
  int refinement_passes = 4;    // 2-way gain-move passes per run
  int threads = 4;              // host threads for the runs of a large bisection (result independent of it)
};

// Returns a block id per vertex; *km1_out receives its km1.
std::vector<int32_t> initial_partition(const HostHypergraph& hg, const InitialPartitioningParams& prm,
                                       int64_t* km1_out);

}  // namespace mt_kahypar_amd
