// hMetis hypergraph files and partition files.
//
// Reference: io/hypergraph_io.cpp:95-177 (readHGRHeader, readHyperedges,
// readHypernodeWeights, readHypergraphFile) and :306-333 (readPartitionFile,
// writePartitionFile).
//
// Format: '%' starts a comment line; the header is "E N [type]" with type 1
// (net weights), 10 (vertex weights) or 11 (both); then E net lines
// "[weight] pin pin ..." with 1-based pins, then N vertex weights when the
// type asks for them.  Nets with fewer than two pins are dropped while
// reading and counted, as the reference does.
#pragma once
#include <string>
#include <vector>

#include "host_hypergraph.h"

namespace mt_kahypar_amd {

struct HgrFile {
  HostHypergraph hg;               // vertex and net weights always filled
  uint64_t removed_single_pin = 0; // nets dropped for having fewer than two pins
  int type = 0;                    // header type field (0, 1, 10, 11)
};

// Throws std::runtime_error with the file name and line on malformed input.
HgrFile read_hgr(const std::string& path);
// Writes hg with the weight sections that differ from 1 (type 0/1/10/11).
void write_hgr(const HostHypergraph& hg, const std::string& path);
// One block id per line.
std::vector<int32_t> read_partition_file(const std::string& path);
void write_partition_file(const std::vector<int32_t>& part, const std::string& path);

}  // namespace mt_kahypar_amd
