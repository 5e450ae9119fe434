// oracle.h -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of Mt-KaHyPar's data-parallel core (Louvain local moving on
// the star expansion, community-restricted heavy-edge clustering and
// contraction, km1 label propagation over the partitioned hypergraph).  It is
// the checker for the HIP implementation under mt-kahypar_amd/ and the timed
// CPU baseline; nothing in the product links or calls it.
//
// Two semantics are restated:
//   * reference semantics ("ref"): the asynchronous algorithms as written in
//     the reference (moves visible immediately, atomics, per-net spin locks,
//     insertion-ordered sparse maps, coin flips from per-thread generators).
//     With one thread and randomization disabled this is the deterministic
//     Gauss-Seidel order the reference's own tests pin (karate club, gain
//     vectors, coarsener chain).
//   * synchronous semantics ("sync"): every round/sub-round computes on the
//     state at its start and then applies; ties are broken by a stateless
//     hash of (seed, round, vertex, candidate).  The device kernels implement
//     exactly this, so integer state is compared bit for bit.
//
// Reference files are cited as path:line relative to the reference root.
#pragma once
#include <cstdint>
#include <cstddef>
#include <vector>
#include <atomic>

namespace orc {

// ---------------------------------------------------------------- hashing --
// SplitMix64 finalizer (Steele, Lea, Flood: "Fast splittable pseudorandom
// number generators", OOPSLA 2014).  The same function is the tie-breaking
// specification of the synchronous semantics.
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  uint64_t h = mix64(seed ^ (a * 0xD6E8FEB86659FD93ull));
  h = mix64(h ^ (b * 0xA0761D6478BD642Full));
  return mix64(h ^ (c * 0xE7037ED1A0B428DBull));
}
// Sub-round bucket of a vertex: MurmurHash3's 32-bit finalizer over the id
// and a per-(seed, round) salt.
inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
  return h;
}
inline uint32_t bucket_salt(uint64_t seed, uint64_t tag) { return uint32_t(mix64(seed ^ tag)); }
inline uint32_t bucket_of(uint32_t salt, uint32_t u, uint32_t buckets) {
  return fmix32(u * 0x9E3779B1u + salt) % buckets;
}
// Louvain sub-round membership is per tile of 2^kLvTileShift consecutive
// node ids: node u takes the bucket of tile u >> kLvTileShift.
constexpr int kLvTileShift = 6;

// ------------------------------------------------------------- hypergraph --
// Static CSR hypergraph (static_hypergraph.h:44); incident nets ascending,
// which is the order static_hypergraph_factory.cpp:124-156 produces when
// stable construction is requested.
struct Hypergraph {
  uint32_t n = 0, m = 0;
  std::vector<uint64_t> noff;   // m+1
  std::vector<uint32_t> pins;   // P
  std::vector<uint64_t> voff;   // n+1
  std::vector<uint32_t> inc;    // P
  std::vector<int32_t> vw;      // n
  std::vector<int32_t> nw;      // m
  std::vector<int32_t> comm;    // n (community ids, 0 if none)
  uint32_t max_esize = 0;
  int64_t total_w = 0;
  uint64_t P() const { return pins.size(); }
  uint32_t esize(uint32_t e) const { return uint32_t(noff[e + 1] - noff[e]); }
  uint32_t degree(uint32_t u) const { return uint32_t(voff[u + 1] - voff[u]); }
};
Hypergraph build_hypergraph(uint32_t n, uint32_t m, const uint64_t* noff, const uint32_t* pins,
                            const int32_t* vw, const int32_t* nw);
void build_incidence(Hypergraph& hg);

// ------------------------------------------------------ partition state --
// PartitionedHypergraph (partitioned_hypergraph.h:46).  Pin counts are kept
// unpacked (values are what is compared; the packing of pin_count_in_part.h
// is a storage detail).
struct Partition {
  const Hypergraph* hg = nullptr;
  int k = 0;
  std::vector<int32_t> part;
  std::vector<int64_t> part_w;     // atomically updated in ref mode
  std::vector<uint32_t> pc;        // m*k
  std::vector<uint64_t> conn;      // m*cw
  int cw = 1;
  uint32_t pcount(uint32_t e, int b) const { return pc[size_t(e) * k + b]; }
  bool has(uint32_t e, int b) const { return (conn[size_t(e) * cw + b / 64] >> (b % 64)) & 1ull; }
  int lambda(uint32_t e) const;
};
void phg_init(Partition& p, const Hypergraph& hg, int k, const int32_t* part);
int64_t km1(const Partition& p);
int64_t cut(const Partition& p);
bool is_border(const Partition& p, uint32_t u);

struct LpParams {
  int k = 2;
  std::vector<int64_t> max_w, perfect_w;
  int max_rounds = 5;                 // fast_preset.ini:58
  uint32_t activation_threshold = 100;// fast_preset.ini:60
  bool rebalancing = true;            // fast_preset.ini:59
  uint64_t seed = 0;
  int sub_rounds = 4;                 // sync only
  bool disable_randomization = false; // ref only
  int threads = 1;                    // ref only
  // initial active set (distinct, ascending); empty = all / border vertices
  // (initializeActiveNodes, label_propagation_refiner.cpp:194-233)
  std::vector<uint32_t> refinement_nodes;
};
struct LpStats {
  int64_t delta = 0;       // km1 change
  int64_t moves = 0;       // kept moves
  int64_t reverts = 0;     // moves reverted by the real-delta check
  int64_t visits = 0;      // incidences swept by the gain computation
  int rounds = 0;
};
// Km1Policy::computeMaxGainMoveImpl (gain_policy.h:108-167) with
// randomization disabled: returns {to, gain}.
void km1_max_gain_move(const Partition& p, uint32_t u, const std::vector<int64_t>& max_w,
                       int32_t* to, int32_t* gain);
LpStats lp_refine_ref(Partition& p, const LpParams& prm);
LpStats lp_refine_sync(Partition& p, const LpParams& prm);

// --------------------------------------------------------------------- FM --
struct FmParams {
  std::vector<int64_t> max_w;
  int rounds = 10;               // r-fm-multitry-rounds (fast_preset.ini:63)
  int seeds = 25;                // r-fm-seed-nodes (fast_preset.ini:67)
  int pq = 64;                   // queue entries per search
  int max_moves = 64;            // moves per search
  uint32_t ignore = 1000;        // nets this large are not expanded
  size_t delta_nets = 256;       // nets a search may change
  uint64_t expand_cap = 1024;    // pins one expansion may see
  int sub_rounds = 8;            // groups of searches per round, applied one after the other
  uint64_t seed = 0;
};
struct FmMove {
  uint32_t v;
  int32_t from, to;
  int64_t gain;
};
struct FmRound {
  std::vector<uint32_t> search_len;     // best local prefix of every search
  std::vector<FmMove> search_moves;     // those prefixes, concatenated
  std::vector<FmMove> sequence;         // merged sequence
  std::vector<int64_t> gains;           // recalculated gains of the sequence
  size_t kept = 0;
  int64_t improvement = 0;
};
struct FmStats { int rounds = 0; int64_t improvement = 0; int64_t moves = 0; int64_t searches = 0; };
FmStats fm_refine_sync(Partition& p, const FmParams& prm, std::vector<FmRound>* trace = nullptr);
std::vector<int64_t> fm_recalculate(const Partition& p, const std::vector<FmMove>& seq);
size_t fm_best_prefix(const Partition& p, const std::vector<FmMove>& seq, const std::vector<int64_t>& gain,
                      const std::vector<int64_t>& max_w, int64_t* best_gain);
void apply_move(Partition& p, const FmMove& m);

// ------------------------------------------------------------------ graph --
// Graph (graph.h:41): star expansion of a hypergraph, or the hypergraph
// itself when every net has two pins.
enum class EdgeWeight { uniform = 0, non_uniform = 1, degree = 2 };
struct Graph {
  uint32_t n = 0;
  std::vector<uint64_t> idx;
  std::vector<uint32_t> head;
  std::vector<double> w;
  std::vector<double> vol;
  double tv = 0;
  uint64_t max_degree = 0;
};
// Star expansion (or the graph itself when every net has two pins).  With
// `ranked`, net node ids are N + rank(e), ranks ordering the nets by
// (locality key, id) with the key the lower median of the first nine pins
// (net_locality_ranks); otherwise N + e as in graph.cpp:331-398.
Graph build_graph(const Hypergraph& hg, EdgeWeight wt, bool ranked = false);
std::vector<uint32_t> net_locality_ranks(const Hypergraph& hg);
Graph contract_graph(const Graph& g, std::vector<int32_t>& C);
double modularity(const Graph& g, const std::vector<int32_t>& C);

struct LouvainParams {
  int max_passes = 5;                  // fast_preset.ini:12
  double min_move_fraction = 0.01;     // fast_preset.ini:13
  uint64_t sampling = 200000;          // fast_preset.ini:14
  uint64_t seed = 0;
  bool disable_randomization = false;  // ref only
  int threads = 1;                     // ref only
  int sub_rounds = 8;                  // sync only
};
struct LouvainStats { int64_t arcs_swept = 0; int rounds = 0; int levels = 0; int64_t moves = 0; };
// computeMaxGainCluster (local_moving_modularity.h:87-129), cluster volumes
// given (initializeClusterVolumes, local_moving_modularity.cpp:211-218).
int32_t louvain_best_cluster(const Graph& g, const std::vector<int32_t>& C,
                             const std::vector<double>& cvol, uint32_t u, uint64_t sampling,
                             double* best_gain);
std::vector<int32_t> louvain_ref(Graph& g, const LouvainParams& prm, LouvainStats* st);
std::vector<int32_t> louvain_sync(Graph& g, const LouvainParams& prm, LouvainStats* st);
// Record of a synchronous local-moving level: the nodes of every sub-round
// (members[member_off[i] .. member_off[i+1])) and its moves (mv_u / mv_to
// over move_off[i] .. move_off[i+1]), sub-rounds in execution order.
struct LouvainTrace {
  std::vector<uint32_t> members, mv_u;
  std::vector<int32_t> mv_to;
  std::vector<uint64_t> member_off, move_off;
};
// One synchronous local-moving level (no contraction); returns whether a
// node moved.  tr (optional) receives the level's sub-rounds.
bool louvain_local_moving_sync(const Graph& g, std::vector<int32_t>& C, const LouvainParams& prm,
                               LouvainStats* st, LouvainTrace* tr = nullptr);

// -------------------------------------------------------------- coarsening --
struct CoarsenParams {
  uint32_t contraction_limit = 320;   // 160*k (context.cpp:256-269)
  int32_t max_node_weight = 1;        // context.cpp:271-284
  double min_shrink = 1.01, max_shrink = 2.5;  // fast_preset.ini:18-19
  uint32_t ignore_he_size = 1000;     // fast_preset.ini:3
  uint64_t sampling = 200000;         // fast_preset.ini:22
  uint64_t seed = 0;
  int sub_rounds = 16;                // sync only
  bool disable_randomization = false; // ref only
  bool prefer_unmatched = true;       // best_prefer_unmatched vs best_without_tie_breaking
  int threads = 1;                    // ref only
};
// One clustering pass; returns cluster representative per vertex
// (cluster_ids, multilevel_coarsener.h:161-239) and the node count after it.
std::vector<uint32_t> cluster_pass_ref(const Hypergraph& hg, const CoarsenParams& prm, int pass,
                                       uint32_t* nodes_after);
std::vector<uint32_t> cluster_pass_sync(const Hypergraph& hg, const CoarsenParams& prm, int pass,
                                        uint32_t* nodes_after, int64_t* pin_visits);
// The rating step of one synchronous sub-round (pass, sub_round) for the given
// proposers on the state (cid, cw, matched): target (UINT32_MAX: none) and its
// 36.28 fixed-point rating per proposer.
void rate_sync(const Hypergraph& hg, const CoarsenParams& prm, int pass, int sub_round,
               const std::vector<uint32_t>& cid, const std::vector<int64_t>& cw,
               const std::vector<uint8_t>& matched, const std::vector<uint32_t>& proposers,
               std::vector<uint32_t>& target, std::vector<uint64_t>& rating);
// Reference-semantics rating of the given vertices on the level-0 state with
// OpenMP threads (a throughput sample, no matching); returns pin visits.
int64_t rating_sample_ref(const Hypergraph& hg, const std::vector<uint32_t>& vertices, uint32_t ignore,
                          int threads);
// StaticHypergraph::contract (static_hypergraph.cpp:58-513), canonical form:
// pins ascending per net, incident nets ascending, survivor of parallel nets
// is the smallest original net id.  cmap receives fine->coarse vertex ids.
Hypergraph contract_hypergraph(const Hypergraph& hg, const std::vector<uint32_t>& cluster,
                               std::vector<uint32_t>& cmap);

struct Level { Hypergraph hg; std::vector<uint32_t> cmap; };
// MultilevelCoarsener::coarsenImpl (multilevel_coarsener.h:143-315)
std::vector<Level> coarsen(const Hypergraph& hg, const CoarsenParams& prm, bool sync,
                           int64_t* pin_visits);

// ---------------------------------------------------- pre- and post-steps --
// sanitize.cpp: degree-zero vertex and large-net removal as one compaction
// (vmap: vertex -> reduced id, 0xffffffff if removed), and the bin-packing
// restore of the removed vertices.
uint32_t large_net_threshold(uint32_t n, double factor);
Hypergraph reduce_hypergraph(const Hypergraph& hg, uint32_t contraction_limit, uint32_t net_threshold,
                             std::vector<uint32_t>& vmap, uint32_t& removed_vertices, uint32_t& removed_nets);
void restore_degree_zero(const std::vector<uint32_t>& removed, const std::vector<int32_t>& weight, int k,
                         std::vector<int64_t>& part_w, std::vector<int32_t>& part);

// ---------------------------------------------------- initial partitioning --
// initial.cpp: recursive bisection with a portfolio of runs per bisection
// (greedy growing, level-synchronous BFS, random fill, each refined by 2-way
// gain moves), the form the device computes.
struct IpParams {
  int k = 2;
  std::vector<int64_t> max_w;  // per block
  uint64_t seed = 0;
  double epsilon = -1;         // < 0: max_w[0] / ceil(W / k) - 1
  int runs = 20;               // fast_preset.ini:29 (i-runs), per bisection
  int passes = 4;              // refinement passes per run
  int threads = 4;             // host threads for the runs (result independent of it)
};
struct IpRun {
  std::vector<uint8_t> side;
  int64_t cut = 0;
  bool balanced = false;
  int64_t w0 = 0;
};
constexpr uint64_t kIpStartTag = 0x243F6A8885A308D3ull;
constexpr uint64_t kIpFillTag = 0x13198A2E03707344ull;
constexpr uint64_t kIpPassTag = 0xA4093822299F31D0ull;
constexpr uint32_t kIpRandomMaxN = 1024;  // random-fill runs up to this many vertices
uint32_t ip_perm(uint64_t key, uint32_t n, uint32_t i);
int ip_variant(int r, uint32_t n);        // 0 greedy growing, 1 BFS, 2 random fill
uint32_t ip_grow_batch(uint32_t n);       // vertices a growing step takes
uint64_t ip_stream(uint64_t seed, int first, int k);
uint64_t ip_run_key(uint64_t stream, int r);
uint32_t ip_start(uint64_t rk, uint32_t n);
uint64_t ip_pass_key(uint64_t rk, int p);
double ip_eps_level(int k, double eps);
void ip_side_limits(int64_t W, int k, double eps_level, int64_t* max0, int64_t* max1, int64_t* target0);
int ip_best_run(const std::vector<IpRun>& res);
IpRun ip_bisect_run(const Hypergraph& h, int64_t max0, int64_t max1, int64_t target0, uint64_t stream, int r,
                    int passes);
// applies moves to the bisection `side` with incremental gain upkeep and
// counts, after every move, the pin counts and gains that differ from a
// recount
int64_t ip_move_check(const Hypergraph& h, const std::vector<uint8_t>& side, const std::vector<uint32_t>& moves);
void ip_repair_balance(const Hypergraph& hg, int k, const std::vector<int64_t>& max_w, std::vector<int32_t>& part);
std::vector<int32_t> initial_partition(const Hypergraph& hg, const IpParams& prm, int64_t* km1_out);

}  // namespace orc
