// Internal object layouts behind the opaque C handles, and the device-wide
// primitives (rocPRIM radix sort and scans) shared by the kernel files.
#pragma once
#include <memory>

#include "common.h"
#include "../../include/mtkahypar_amd.h"

struct mkp_context {
  mkp::Context c;
};

// Static CSR hypergraph in HBM.  32-bit offsets (checked at creation),
// incident nets ascending per vertex.
struct mkp_hypergraph {
  mkp_context* ctx = nullptr;
  uint32_t n = 0, m = 0;
  uint64_t p = 0;
  uint32_t max_esize = 0;
  int64_t total_w = 0;
  mkp::DevBuf<uint32_t> noff, pins, voff, inc;
  mkp::DevBuf<uint32_t> pin_net;  // net of every pin position
  mkp::DevBuf<uint32_t> inc_vtx;  // vertex of every incidence position
  mkp::DevBuf<int32_t> vw, nw, comm;
  bool unit_net_weights = false;
  // locality ranks of the nets (the star expansion numbers net e as N +
  // rank(e)); built on first use: order[r] = e, rank[e] = r, inc_rank[j] =
  // rank(inc[j])
  mkp::DevBuf<uint32_t> net_order, net_rank, inc_rank;
  // pins stored in rank order: net of rank r holds rpins[rpoff[r] ..
  // rpoff[r+1]); rpin_rank[j] = rank of the net of ranked pin slot j
  mkp::DevBuf<uint32_t> rpoff, rpins, rpin_rank;
  bool ranked = false;
};

namespace mkp {
// Packed pin counts: b bits per (net, block), epw entries per 64-bit word,
// vpc words per net (pin_count_in_part.h:198-215).  Fields never carry into
// each other because a count stays within [0, |e|] <= 2^b - 1, so a 64-bit
// atomic add of +-(1 << shift) is an exact per-field update.
// The partition state of a net is one record of `stride` words: its vpc
// pin-count words, then its vcs connectivity words (at coff), then its net
// weight (at woff), so a gain or apply step touches one record per incident
// net instead of three arrays.
struct PcLayout {
  int b = 1, epw = 1, vpc = 1;
  int vcs = 1, coff = 1, woff = 2, stride = 3;
  uint64_t mask = 1;
};
PcLayout pc_layout(int k, uint32_t max_esize);
}  // namespace mkp

namespace mkp { struct LpWork; void lp_work_free(LpWork* w); }

struct mkp_partition {
  ~mkp_partition() { mkp::lp_work_free(lp); }
  mkp_hypergraph* hg = nullptr;
  mkp::LpWork* lp = nullptr;  // label propagation workspace, kept across calls
  int k = 0;
  int vcs = 1;  // connectivity words per net
  mkp::PcLayout L;
  mkp::DevBuf<int32_t> part;
  mkp::DevBuf<int64_t> part_w;
  mkp::DevBuf<uint64_t> pc;  // net records (PcLayout): pin counts, connectivity, weight
  mkp::DevBuf<int64_t> km1_acc;
};

namespace mkp {
// rocPRIM wrappers on the context stream (temporary storage from ctx.tmp).
void sort_keys_u64(Context& c, const uint64_t* in, uint64_t* out, size_t n, int begin_bit, int end_bit);
void sort_pairs_u64_u32(Context& c, const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout,
                        size_t n, int begin_bit, int end_bit);
// out[i] = sum(in[0..i-1]); out has n+1 entries, out[n] = total
void exclusive_scan_u32(Context& c, const uint32_t* in, uint32_t* out, size_t n);
void exclusive_scan_u64(Context& c, const uint64_t* in, uint64_t* out, size_t n);
void inclusive_scan_i64(Context& c, const int64_t* in, int64_t* out, size_t n);
void inclusive_max_scan_u64(Context& c, const uint64_t* in, uint64_t* out, size_t n);

// Builds voff/inc of hg from noff/pins on the device.
void build_incidence(mkp_hypergraph* hg);
// Partition initialisation kernels (block weights, pin counts,
// connectivity, km1) on the device-resident part ids.
void partition_initialize(mkp_partition* p);
int64_t partition_km1(mkp_partition* p);
// batch forms: block weights per (member, block), km1 per member (host out)
void partition_initialize_batch(mkp_partition* p, const uint32_t* vgraph, const uint32_t* gkbase, uint32_t ktot);
void partition_km1_batch(mkp_partition* p, const uint32_t* ngraph, uint32_t members, int64_t* out);

// clustering pass / contraction / LP entry points shared by the drivers
struct ClusterResult {
  uint32_t nodes_after = 0;
  int64_t visits = 0;
};
ClusterResult cluster_pass(mkp_hypergraph* hg, const mkp_coarsen_params& prm, int pass, uint32_t* cid_out);
mkp_hypergraph* contract(mkp_hypergraph* hg, const uint32_t* cid, uint32_t* cmap);
// A batch of independent member hypergraphs stored as one disjoint union
// (device arrays): vertex -> member, member -> first vertex, first block in
// the batch's block-weight array, k, seed.
struct LpBatch {
  uint32_t members = 0;
  uint32_t ktot = 0;  // blocks of all members
  int kmax = 0;
  const uint32_t* vgraph = nullptr;
  const uint32_t* gvbase = nullptr;
  const uint32_t* gkbase = nullptr;
  const int32_t* gk = nullptr;
  const uint64_t* gseed = nullptr;
};
// Builds the net locality ranks of hg once (the hypergraph is immutable).
void ensure_net_ranks(mkp_hypergraph* hg);
// Louvain over a batch stored as one disjoint union hg: every member runs
// local_moving_contract_recurse as it would alone (its own seed and total
// volume, its own round and level stops); hg->comm receives member-local
// community ids, stats one entry per member.
struct LouvainBatch {
  uint32_t members = 0;
  const uint32_t* vbase = nullptr;   // host, members + 1: first vertex
  const uint32_t* nbase = nullptr;   // host, members + 1: first net
  const uint64_t* pbase = nullptr;   // host, members + 1: first pin
  const uint32_t* dvbase = nullptr;  // device copies of vbase / nbase
  const uint32_t* dnbase = nullptr;
  const uint32_t* vgraph = nullptr;  // device, vertex -> member
  const uint32_t* ngraph = nullptr;  // device, net -> member
  const uint64_t* seeds = nullptr;   // host, per member
  bool plain = false;                // every member is a graph (all nets have two pins)
};
void louvain_batch(mkp_hypergraph* hg, const mkp_louvain_params& prm, int wt, const LouvainBatch& d,
                   mkp_louvain_stats* stats);
// nodes (host, distinct, ascending) replaces the initial active set if given.
// With a batch, prm's weight arrays hold ktot entries, stats holds one entry
// per member and every member refines exactly as it would alone.
void fm_refine(mkp_partition* p, const mkp_fm_params& prm, mkp_fm_stats* stats);
void lp_refine(mkp_partition* p, const mkp_lp_params& prm, mkp_lp_stats* stats, const uint32_t* nodes = nullptr,
               uint32_t num_nodes = 0, const LpBatch* batch = nullptr);

// Initial partitioning on the device (recursive bisection, a portfolio of
// `runs` runs per bisection): block ids of hg's vertices into part (device).
void initial_partition(mkp_hypergraph* hg, int k, const std::vector<int64_t>& max_w, uint64_t seed, int runs,
                       int passes, int32_t* part);
// The runs of one bisection of hg (sides: runs x n on the host, results:
// runs x {cut, balanced, side-0 weight, side-1 weight}).
void ip_bisect_runs(mkp_hypergraph* hg, int64_t max0, int64_t max1, int64_t target0, uint64_t stream, int runs,
                    int passes, uint8_t* sides, int64_t* results);
void ip_side_limits(int64_t W, int k, double eps_level, int64_t* max0, int64_t* max1, int64_t* target0);
uint64_t ip_stream(uint64_t seed, int first, int k);

// Rejects partition parameters the pipeline cannot run (k, epsilon, preset
// bounds, sub-round counts above 255) with an invalid-argument error.
void validate_partition_params(const mkp_partition_params& p);

// Degree-zero vertex and large-net removal (degree_zero_hn_remover.h:46-61,
// large_he_remover.h:47-58) as one compaction: hg is null when nothing was
// removed, vmap maps every vertex of the input to its reduced id
// (0xffffffff: removed).
struct Reduction {
  std::unique_ptr<mkp_hypergraph> hg;
  DevBuf<uint32_t> vmap;
  uint32_t removed_vertices = 0, removed_nets = 0;
};
uint32_t large_net_threshold(uint32_t n, double factor);
void reduce(mkp_hypergraph* hg, uint32_t contraction_limit, uint32_t net_threshold, Reduction& r);
// Block ids of the input's vertices (host, n) from the reduced partition,
// with the removed vertices bin-packed (degree_zero_hn_remover.h:64-83).
void restore(mkp_hypergraph* hg, const Reduction& r, const mkp_partition* reduced, int32_t* parts);
}  // namespace mkp
