// Incident-net arrays of a host hypergraph (see host_hypergraph.h).
#include "host_hypergraph.h"

#include <cstddef>

namespace mt_kahypar_amd {

void build_host_incidence(HostHypergraph& hg) {
  hg.voff.assign(size_t(hg.n) + 1, 0);
  for (uint64_t i = 0; i < hg.pins.size(); ++i) hg.voff[hg.pins[i] + 1]++;
  for (uint32_t u = 0; u < hg.n; ++u) hg.voff[u + 1] += hg.voff[u];
  hg.inc.assign(hg.pins.size(), 0);
  std::vector<uint64_t> pos(hg.voff.begin(), hg.voff.end() - 1);
  for (uint32_t e = 0; e < hg.m; ++e)
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) hg.inc[pos[hg.pins[i]]++] = e;
}

}  // namespace mt_kahypar_amd
