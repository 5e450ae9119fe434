// sanitize.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// The steps on either side of the multilevel partitioner in partition()
// (partitioner.cpp:74-86 and 190-195):
//   * degree-zero vertex removal (degree_zero_hn_remover.h:46-61): in id
//     order, every vertex without nets is removed until the remaining count
//     reaches the contraction limit;
//   * large-net removal (large_he_remover.h:47-58): every net with more pins
//     than max(max(N * maxnet-remove-factor, 100), 50000) is removed
//     (threshold from partitioner.cpp:37-38 and large_he_remover.h:90-92,
//     N the vertex count of the input);
//   * restore: large nets come back with their pins (their blocks follow from
//     the vertices'), and the removed vertices are bin-packed into the blocks
//     (degree_zero_hn_remover.h:64-83).
// The removals are restated as one compaction: kept vertices and kept nets
// keep their relative order, pins keep their order within a net.
#include "oracle.h"

#include <algorithm>
#include <numeric>

namespace orc {

uint32_t large_net_threshold(uint32_t n, double factor) {
  const double t = std::max(double(n) * factor, 100.0);
  const uint32_t ti = t >= 4294967295.0 ? 0xffffffffu : uint32_t(t);  // HypernodeID truncation
  return std::max<uint32_t>(ti, 50000u);
}

Hypergraph reduce_hypergraph(const Hypergraph& hg, uint32_t contraction_limit, uint32_t net_threshold,
                             std::vector<uint32_t>& vmap, uint32_t& removed_vertices, uint32_t& removed_nets) {
  // degree_zero_hn_remover.h:50-59: stop once the remaining count is at or
  // below the contraction limit
  uint32_t current = hg.n, removed = 0;
  std::vector<uint8_t> keep_v(hg.n, 1);
  for (uint32_t u = 0; u < hg.n; ++u) {
    if (current - removed <= contraction_limit) break;
    if (hg.degree(u) == 0) {
      keep_v[u] = 0;
      ++removed;
    }
  }
  removed_vertices = removed;
  vmap.assign(hg.n, 0xffffffffu);
  Hypergraph r;
  for (uint32_t u = 0; u < hg.n; ++u) {
    if (!keep_v[u]) continue;
    vmap[u] = r.n++;
    r.vw.push_back(hg.vw[u]);
    r.comm.push_back(hg.comm.empty() ? 0 : hg.comm[u]);
  }
  removed_nets = 0;
  r.noff.assign(1, 0);
  for (uint32_t e = 0; e < hg.m; ++e) {
    if (hg.esize(e) > net_threshold) {  // large_he_remover.h:50
      ++removed_nets;
      continue;
    }
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) r.pins.push_back(vmap[hg.pins[i]]);
    r.noff.push_back(r.pins.size());
    r.nw.push_back(hg.nw[e]);
    r.m += 1;
    r.max_esize = std::max(r.max_esize, hg.esize(e));
  }
  r.total_w = 0;
  for (int32_t w : r.vw) r.total_w += w;
  build_incidence(r);
  return r;
}

// degree_zero_hn_remover.h:64-83.  The removed vertices go in order of
// decreasing weight (ties: smaller id first) to the currently lightest block,
// kept at the front of a list sorted by increasing block weight (ties:
// smaller block id first) by bubbling the block that received a vertex
// forward past every lighter block.
void restore_degree_zero(const std::vector<uint32_t>& removed, const std::vector<int32_t>& weight, int k,
                         std::vector<int64_t>& part_w, std::vector<int32_t>& part) {
  std::vector<uint32_t> order(removed);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return weight[a] > weight[b]; });
  std::vector<int32_t> blocks(k);
  std::iota(blocks.begin(), blocks.end(), 0);
  std::stable_sort(blocks.begin(), blocks.end(), [&](int32_t a, int32_t b) { return part_w[a] < part_w[b]; });
  for (uint32_t u : order) {
    const int32_t to = blocks.front();
    part[u] = to;
    part_w[to] += weight[u];
    int i = 0;
    while (i + 1 < k && part_w[blocks[i]] > part_w[blocks[i + 1]]) {
      std::swap(blocks[i], blocks[i + 1]);
      ++i;
    }
  }
}

}  // namespace orc
