// C++ host interfaces of the reference's plugin boundary, implemented over the
// C-ABI of mtkahypar_amd.h (header-only; link libmtkahypar_amd.so).
//
//   IRefiner                 partition/refinement/i_refiner.h:35-66
//   ICoarsener               partition/coarsening/i_coarsener.h:31-67
//   run_parallel_louvain     partition/preprocessing/community_detection/parallel_louvain.h:26-27
//   LabelPropagationRefiner  partition/refinement/label_propagation/label_propagation_refiner.h:35-180
//   MultilevelCoarsener      partition/coarsening/multilevel_coarsener.h:51-143
//
// Objects hold device handles; the hypergraph and partition state never
// leave HBM between calls.  Errors throw mt_kahypar_amd::Error carrying the
// message of mkp_last_error().
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mtkahypar_amd.h"

namespace mt_kahypar_amd {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void check(int rc) {
  if (rc != 0) throw Error(mkp_last_error());
}

// One device and its stream (replaces the TBB arena).
class Context {
 public:
  explicit Context(int device = 0) { check(mkp_context_create(device, &c_)); }
  ~Context() { if (c_) mkp_context_destroy(c_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  mkp_context* handle() const { return c_; }

 private:
  mkp_context* c_ = nullptr;
};

// The partition context values this path reads (context.h PartitionParameters
// plus the fast preset of config/fast_preset.ini).
struct PartitionContext {
  mkp_partition_params p{};
  PartitionContext(int32_t k, double epsilon, uint64_t seed) {
    check(mkp_partition_params_default(&p, k, epsilon, seed));
  }
};

// StaticHypergraph in HBM; owning, or a borrowed level of a hierarchy.
class Hypergraph {
 public:
  Hypergraph(Context& ctx, uint32_t num_vertices, uint32_t num_nets, const std::vector<uint64_t>& net_offsets,
             const std::vector<uint32_t>& pins, const int32_t* vertex_weights = nullptr,
             const int32_t* net_weights = nullptr) : owned_(true) {
    check(mkp_hypergraph_create(ctx.handle(), num_vertices, num_nets, net_offsets.data(), pins.data(),
                                vertex_weights, net_weights, &h_));
  }
  static Hypergraph borrow(mkp_hypergraph* h) { return Hypergraph(h, false); }
  Hypergraph(Hypergraph&& o) noexcept : h_(o.h_), owned_(o.owned_) { o.h_ = nullptr; }
  Hypergraph& operator=(Hypergraph&& o) noexcept {
    std::swap(h_, o.h_);
    std::swap(owned_, o.owned_);
    return *this;
  }
  ~Hypergraph() { if (h_ && owned_) mkp_hypergraph_destroy(h_); }
  mkp_hypergraph* handle() const { return h_; }
  mkp_hypergraph_info info() const {
    mkp_hypergraph_info i{};
    check(mkp_hypergraph_get_info(h_, &i));
    return i;
  }
  uint32_t initialNumNodes() const { return info().num_vertices; }
  uint32_t initialNumEdges() const { return info().num_nets; }
  int64_t totalWeight() const { return info().total_weight; }
  void setCommunityIDs(const std::vector<int32_t>& c) { check(mkp_hypergraph_set_communities(h_, c.data())); }

 private:
  Hypergraph(mkp_hypergraph* h, bool owned) : h_(h), owned_(owned) {}
  mkp_hypergraph* h_ = nullptr;
  bool owned_ = false;
};

// PartitionedHypergraph: block ids, block weights, packed pin counts and
// connectivity sets of one hypergraph level.
class PartitionedHypergraph {
 public:
  PartitionedHypergraph(int32_t k, Hypergraph& hg, int level = 0) : k_(k), level_(level), hg_(hg.handle()) {
    check(mkp_partition_create(hg.handle(), k, &p_));
  }
  PartitionedHypergraph(PartitionedHypergraph&& o) noexcept : k_(o.k_), level_(o.level_), hg_(o.hg_), p_(o.p_) {
    o.p_ = nullptr;
  }
  PartitionedHypergraph& operator=(PartitionedHypergraph&& o) noexcept {
    std::swap(k_, o.k_); std::swap(level_, o.level_); std::swap(hg_, o.hg_); std::swap(p_, o.p_);
    return *this;
  }
  ~PartitionedHypergraph() { if (p_) mkp_partition_destroy(p_); }
  mkp_partition* handle() const { return p_; }
  int32_t k() const { return k_; }
  // Level in the hierarchy (0 = input); seeds the per-level LP coins.
  int level() const { return level_; }
  // setOnlyNodePart for all vertices at once, then initializePartition
  // (partitioned_hypergraph.h:522-526,1067-1110).
  void setPartition(const std::vector<int32_t>& parts) {
    check(mkp_partition_set_parts(p_, parts.data()));
    check(mkp_partition_initialize(p_));
  }
  std::vector<int32_t> partIDs() const {
    mkp_hypergraph_info i{};
    check(mkp_hypergraph_get_info(hg_, &i));
    std::vector<int32_t> out(i.num_vertices);
    check(mkp_partition_get_parts(p_, out.data()));
    return out;
  }
  std::vector<int64_t> partWeights() const {
    std::vector<int64_t> w(static_cast<size_t>(k_));
    check(mkp_partition_get_part_weights(p_, w.data()));
    return w;
  }
  int64_t km1() const {
    int64_t v = 0;
    check(mkp_partition_km1(p_, &v));
    return v;
  }

 private:
  int32_t k_ = 0;
  int level_ = 0;
  mkp_hypergraph* hg_ = nullptr;
  mkp_partition* p_ = nullptr;
};

// Initial partitioning of the coarsest level on the host (pool of greedy
// growing and random starts with k-way hill climbing); result set on phg.
inline int64_t initial_partition(PartitionedHypergraph& phg, const PartitionContext& ctx) {
  int64_t km1 = 0;
  check(mkp_initial_partition(phg.handle(), ctx.p.epsilon, ctx.p.seed, ctx.p.ip_runs, &km1));
  return km1;
}

// kahypar::Metrics fields used by the refiners.
struct Metrics {
  int64_t cut = 0;
  int64_t km1 = 0;
  double imbalance = 0.0;
};

// metrics::imbalance (metrics.cpp:114-128)
inline double imbalance(const PartitionedHypergraph& phg) {
  const std::vector<int64_t> w = phg.partWeights();
  int64_t total = 0;
  for (int64_t x : w) total += x;
  const double perfect = double((total + phg.k() - 1) / phg.k());
  double worst = 0.0;
  for (int64_t x : w) worst = std::max(worst, double(x) / perfect);
  return worst - 1.0;
}

// i_refiner.h:35-66
class IRefiner {
 public:
  IRefiner(const IRefiner&) = delete;
  IRefiner(IRefiner&&) = delete;
  IRefiner& operator=(const IRefiner&) = delete;
  IRefiner& operator=(IRefiner&&) = delete;
  virtual ~IRefiner() = default;

  void initialize(PartitionedHypergraph& hypergraph) { initializeImpl(hypergraph); }

  bool refine(PartitionedHypergraph& hypergraph, const std::vector<uint32_t>& refinement_nodes,
              Metrics& best_metrics, const double time_limit) {
    return refineImpl(hypergraph, refinement_nodes, best_metrics, time_limit);
  }

 protected:
  IRefiner() = default;

 private:
  virtual void initializeImpl(PartitionedHypergraph& hypergraph) = 0;
  virtual bool refineImpl(PartitionedHypergraph& hypergraph, const std::vector<uint32_t>& refinement_nodes,
                          Metrics& best_metrics, const double time_limit) = 0;
};

// LabelPropagationRefiner<Km1Policy> on the device, synchronous sub-rounds.
class LabelPropagationRefiner final : public IRefiner {
 public:
  explicit LabelPropagationRefiner(const PartitionContext& ctx) : ctx_(ctx) {}
  mkp_lp_stats lastStats() const { return last_; }

 private:
  void initializeImpl(PartitionedHypergraph&) override {}

  bool refineImpl(PartitionedHypergraph& phg, const std::vector<uint32_t>& nodes, Metrics& best,
                  const double) override {
    const mkp_partition_params& p = ctx_.p;
    if (!p.lp_enabled || p.lp_max_rounds <= 0) return false;
    int64_t W = 0;
    for (int64_t w : phg.partWeights()) W += w;
    // setupPartWeights (context.cpp:237-250)
    const int64_t perfect = (W + phg.k() - 1) / phg.k();
    const int64_t maxw = int64_t(int32_t((1.0 + p.epsilon) * double(perfect)));
    std::vector<int64_t> maxv(size_t(phg.k()), maxw), perv(size_t(phg.k()), perfect);
    mkp_lp_params lp{p.lp_max_rounds, p.lp_sub_rounds, p.lp_activation_threshold, p.lp_rebalancing,
                     p.seed + uint64_t(phg.level()) * 1000003ull, maxv.data(), perv.data()};
    mkp_lp_stats st{};
    check(mkp_lp_refine_nodes(phg.handle(), &lp, nodes.data(), uint32_t(nodes.size()), &st));
    last_ = st;
    best.km1 += st.delta;  // refineImpl: km1 += delta (label_propagation_refiner.cpp:48-67)
    best.imbalance = imbalance(phg);
    return st.delta < 0;
  }

  PartitionContext ctx_;
  mkp_lp_stats last_{};
};

// MultiTryKWayFM<GainCacheStrategy> (multitry_kway_fm.h:40-92) on the
// device, synchronous form; refinement nodes are not used (every border
// vertex seeds a search, as in the log(n)-level case of
// multitry_kway_fm.cpp:160-175).
class MultiTryKWayFM final : public IRefiner {
 public:
  explicit MultiTryKWayFM(const PartitionContext& ctx) : ctx_(ctx) {}
  mkp_fm_stats lastStats() const { return last_; }

 private:
  void initializeImpl(PartitionedHypergraph&) override {}

  bool refineImpl(PartitionedHypergraph& phg, const std::vector<uint32_t>&, Metrics& best,
                  const double) override {
    const mkp_partition_params& p = ctx_.p;
    if (p.fm_rounds <= 0) return false;
    int64_t W = 0;
    for (int64_t w : phg.partWeights()) W += w;
    const int64_t perfect = (W + phg.k() - 1) / phg.k();
    const int64_t maxw = int64_t(int32_t((1.0 + p.epsilon) * double(perfect)));
    std::vector<int64_t> maxv(size_t(phg.k()), maxw);
    mkp_fm_params fp{maxv.data(), p.fm_rounds, p.fm_seeds, p.ignore_net_size,
                     p.seed + uint64_t(phg.level()) * 1000003ull};
    mkp_fm_stats st{};
    check(mkp_fm_refine(phg.handle(), &fp, &st));
    last_ = st;
    best.km1 -= st.improvement;  // refineImpl: metrics.km1 -= overall_improvement (multitry_kway_fm.cpp:139)
    best.imbalance = imbalance(phg);
    return st.improvement > 0;
  }

  PartitionContext ctx_;
  mkp_fm_stats last_{};
};

// Graph(hypergraph, edge weight) (graph.h:66-72): the graph Louvain runs on,
// the star expansion (N + E nodes, net e at node N + e) or the hypergraph
// itself when every net has two pins.  The device builds its arcs when
// Louvain runs; this object names the graph and its node count.
class Graph {
 public:
  explicit Graph(Hypergraph& hg) : hg_(hg) {
    int64_t n = 0;
    check(mkp_louvain_graph_nodes(hg.handle(), &n));
    nodes_ = size_t(n);
  }
  size_t numNodes() const { return nodes_; }
  Hypergraph& hypergraph() { return hg_; }
  // graph.h:123-125: the first N entries are the hypernodes' clusters
  void restrictClusteringToHypernodes(const Hypergraph& hg, std::vector<int32_t>& C) const {
    C.resize(hg.initialNumNodes());
  }

 private:
  Hypergraph& hg_;
  size_t nodes_ = 0;
};

// run_parallel_louvain(Graph&, const Context&, bool) (parallel_louvain.h:27):
// the clustering of every node of the graph; the hypernodes' part is also
// stored as the hypergraph's community ids.
inline std::vector<int32_t> run_parallel_louvain(Graph& graph, const PartitionContext& ctx,
                                                 bool disable_randomization = false) {
  (void)disable_randomization;  // the device path is deterministic for a seed
  const mkp_partition_params& p = ctx.p;
  mkp_louvain_params lp{p.louvain_edge_weight, p.louvain_max_passes, p.louvain_min_move_fraction,
                        p.louvain_sampling, p.seed, p.louvain_sub_rounds};
  std::vector<int32_t> C(graph.numNodes());
  mkp_louvain_stats st{};
  check(mkp_louvain_graph(graph.hypergraph().handle(), &lp, C.data(), &st));
  return C;
}

// run_parallel_louvain + restrictClusteringToHypernodes: community id per
// hypernode, also stored as the hypergraph's community ids.
inline std::vector<int32_t> run_parallel_louvain(Hypergraph& hg, const PartitionContext& ctx,
                                                 bool disable_randomization = false) {
  (void)disable_randomization;  // the device path is deterministic for a seed
  const mkp_partition_params& p = ctx.p;
  mkp_louvain_params lp{p.louvain_edge_weight, p.louvain_max_passes, p.louvain_min_move_fraction,
                        p.louvain_sampling, p.seed, p.louvain_sub_rounds};
  std::vector<int32_t> comm(hg.initialNumNodes());
  mkp_louvain_stats st{};
  check(mkp_louvain(hg.handle(), &lp, comm.data(), &st));
  return comm;
}

// i_coarsener.h:31-67
class ICoarsener {
 public:
  ICoarsener(const ICoarsener&) = delete;
  ICoarsener(ICoarsener&&) = delete;
  ICoarsener& operator=(const ICoarsener&) = delete;
  ICoarsener& operator=(ICoarsener&&) = delete;
  virtual ~ICoarsener() = default;

  void coarsen() { coarsenImpl(); }
  PartitionedHypergraph&& uncoarsen(std::unique_ptr<IRefiner>& label_propagation, std::unique_ptr<IRefiner>& fm) {
    return uncoarsenImpl(label_propagation, fm);
  }
  Hypergraph& coarsestHypergraph() { return coarsestHypergraphImpl(); }
  PartitionedHypergraph& coarsestPartitionedHypergraph() { return coarsestPartitionedHypergraphImpl(); }

 protected:
  ICoarsener() = default;

 private:
  virtual void coarsenImpl() = 0;
  virtual PartitionedHypergraph&& uncoarsenImpl(std::unique_ptr<IRefiner>& label_propagation,
                                                std::unique_ptr<IRefiner>& fm) = 0;
  virtual Hypergraph& coarsestHypergraphImpl() = 0;
  virtual PartitionedHypergraph& coarsestPartitionedHypergraphImpl() = 0;
};

// MultilevelCoarsener with the synchronous device clustering; the level
// hierarchy stays on the device.
class MultilevelCoarsener final : public ICoarsener {
 public:
  MultilevelCoarsener(Hypergraph& hg, const PartitionContext& ctx, bool top_level = true)
      : top_(hg), ctx_(ctx), top_level_(top_level) {}
  ~MultilevelCoarsener() override {
    phgs_.clear();
    levels_.clear();
    if (h_) mkp_hierarchy_destroy(h_);
  }
  int numLevels() const { return int(levels_.size()); }

 private:
  void coarsenImpl() override {
    const mkp_partition_params& p = ctx_.p;
    const int64_t W = top_.totalWeight();
    const int64_t perfect = (W + p.k - 1) / p.k;
    const int64_t maxw = int64_t(int32_t((1.0 + p.epsilon) * double(perfect)));
    const uint32_t limit = p.contraction_limit_multiplier * uint32_t(p.k);  // context.cpp:256-269
    int64_t max_node_w = int64_t(std::ceil(p.max_allowed_weight_multiplier / double(limit) * double(W)));
    max_node_w = std::min(max_node_w, maxw);  // context.cpp:271-284
    mkp_coarsen_params cp{limit, int32_t(max_node_w), p.min_shrink_factor, p.max_shrink_factor,
                          p.ignore_net_size, p.rating_sampling, p.seed, p.coarsening_sub_rounds};
    check(mkp_coarsen(top_.handle(), &cp, &h_));
    int32_t L = 0;
    check(mkp_hierarchy_num_levels(h_, &L));
    levels_.clear();
    for (int32_t i = 0; i < L; ++i) {
      mkp_hypergraph* g = nullptr;
      check(mkp_hierarchy_level(h_, i, &g));
      levels_.push_back(Hypergraph::borrow(g));
    }
    phgs_.clear();
    phgs_.resize(size_t(L));
  }

  PartitionedHypergraph&& uncoarsenImpl(std::unique_ptr<IRefiner>& lp, std::unique_ptr<IRefiner>& fm) override {
    const int L = int(levels_.size());
    PartitionedHypergraph& coarsest = coarsestPartitionedHypergraphImpl();
    Metrics m;
    m.km1 = coarsest.km1();
    // doUncoarsen (uncoarsening.cpp:94-188): refine the coarsest level, then
    // project and refine level by level
    for (int i = L - 1; i >= 0; --i) {
      PartitionedHypergraph& cur = *phgs_[size_t(i)];
      if (i < L - 1) {
        check(mkp_partition_project(cur.handle(), phgs_[size_t(i) + 1]->handle(), h_, i));
        phgs_[size_t(i) + 1].reset();
      }
      if (lp) {
        lp->initialize(cur);
        lp->refine(cur, {}, m, 0.0);
      }
      if (fm) {
        fm->initialize(cur);
        fm->refine(cur, {}, m, 0.0);
      }
      if (i > 0) phgs_[size_t(i) - 1].reset(new PartitionedHypergraph(ctx_.p.k, levels_[size_t(i) - 1], i - 1));
    }
    return std::move(*phgs_[0]);
  }

  Hypergraph& coarsestHypergraphImpl() override { return levels_.empty() ? top_ : levels_.back(); }

  PartitionedHypergraph& coarsestPartitionedHypergraphImpl() override {
    if (levels_.empty()) throw Error("coarsen() has not run");
    auto& slot = phgs_.back();
    if (!slot) slot.reset(new PartitionedHypergraph(ctx_.p.k, levels_.back(), int(levels_.size()) - 1));
    return *slot;
  }

  Hypergraph& top_;
  PartitionContext ctx_;
  bool top_level_;
  mkp_hierarchy* h_ = nullptr;
  std::vector<Hypergraph> levels_;  // borrowed views, level 0 = input
  std::vector<std::unique_ptr<PartitionedHypergraph>> phgs_;
};

}  // namespace mt_kahypar_amd
