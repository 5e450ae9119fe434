// Host-side hypergraph in CSR form (pin lists per net plus incident nets per
// vertex), used by the readers, the generators and the initial partitioner.
#pragma once
#include <cstdint>
#include <vector>

namespace mt_kahypar_amd {

struct HostHypergraph {
  uint32_t n = 0, m = 0;
  std::vector<uint64_t> noff;  // m + 1 net offsets into pins
  std::vector<uint32_t> pins;
  std::vector<uint64_t> voff;  // n + 1 vertex offsets into inc (after build_host_incidence)
  std::vector<uint32_t> inc;
  std::vector<int32_t> vw, nw;
};

// Builds the incident-net arrays of a host hypergraph from its pin lists.
void build_host_incidence(HostHypergraph& hg);

}  // namespace mt_kahypar_amd

// Opaque handle of the C-ABI (mkp_host_hypergraph).
struct mkp_host_hypergraph {
  mt_kahypar_amd::HostHypergraph h;
};
