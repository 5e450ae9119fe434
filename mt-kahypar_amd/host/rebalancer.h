// km1 rebalancer on the host (Rebalancer<Km1Policy>, rebalancer.cpp:36-262).
//
// The reference runs it once, at the top level, when the partition is over
// its bounds after refinement (uncoarsening.cpp:146-180).  This restatement is
// sequential with a fixed order, so its result is deterministic:
//   1. empty blocks are given a vertex each (repairEmptyBlocks,
//      rebalancer.cpp:182-262);
//   2. in id order, every border vertex of an overloaded block computes its
//      best move to a block it fits into (Km1Policy::computeMaxGainMove with
//      rebalance = true, gain_policy.h:108-162); a move that does not worsen
//      km1 (gain <= 0) is made at once, the others wait in a priority queue
//      (rebalancer.cpp:73-86);
//   3. while the partition is still over its bounds, the queue is processed
//      in increasing gain order: a vertex whose block is still overloaded
//      recomputes its move and makes it if it is no worse than the queued
//      one, or goes back into the queue with the new gain
//      (rebalancer.cpp:94-163).
// Ties between target blocks go to the smaller block id (the reference
// flips coins); ties in the queue go to the smaller vertex id.
#pragma once
#include <cstdint>
#include <vector>

#include "host_hypergraph.h"

namespace mt_kahypar_amd {

struct RebalanceResult {
  int64_t km1_delta = 0;  // change of km1 (negative: better)
  int64_t moves = 0;
  bool balanced = false;  // every block within its bound afterwards
};

// part (n entries, blocks 0..k-1) is rebalanced in place; hg needs its
// incident-net arrays (build_host_incidence).
RebalanceResult rebalance_km1(const HostHypergraph& hg, int k, const std::vector<int64_t>& max_w,
                              std::vector<int32_t>& part);

}  // namespace mt_kahypar_amd
