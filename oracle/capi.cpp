// capi.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// C entry points used by tests/ (ctypes) and by bench.py's CPU baseline leg,
// plus the multilevel pipeline of partitioner.cpp:163-204 built from the
// restated pieces.
#include "oracle.h"
#include "../mt-kahypar_amd/host/rebalancer.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

using namespace orc;

namespace {
double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
// clusterings handed in from Python index per-node arrays: reject ids out of range
bool ids_in_range(const int32_t* C, uint32_t n) {
  for (uint32_t u = 0; u < n; ++u)
    if (C[u] < 0 || uint32_t(C[u]) >= n) return false;
  return true;
}
mt_kahypar_amd::HostHypergraph to_host(const Hypergraph& hg) {
  mt_kahypar_amd::HostHypergraph h;
  h.n = hg.n; h.m = hg.m; h.noff = hg.noff; h.pins = hg.pins; h.voff = hg.voff; h.inc = hg.inc;
  h.vw = hg.vw; h.nw = hg.nw;
  return h;
}
}  // namespace

extern "C" {

void* orc_hg_new(uint32_t n, uint32_t m, const uint64_t* noff, const uint32_t* pins, const int32_t* vw,
                 const int32_t* nw) {
  return new Hypergraph(build_hypergraph(n, m, noff, pins, vw, nw));
}
void orc_hg_free(void* h) { delete static_cast<Hypergraph*>(h); }
void orc_hg_set_comm(void* h, const int32_t* comm) {
  auto* hg = static_cast<Hypergraph*>(h);
  std::copy(comm, comm + hg->n, hg->comm.begin());
}
void orc_hg_sizes(void* h, uint32_t* n, uint32_t* m, uint64_t* p, uint32_t* max_esize, int64_t* total_w) {
  auto* hg = static_cast<Hypergraph*>(h);
  *n = hg->n; *m = hg->m; *p = hg->P(); *max_esize = hg->max_esize; *total_w = hg->total_w;
}
void orc_hg_get(void* h, uint64_t* noff, uint32_t* pins, uint64_t* voff, uint32_t* inc, int32_t* vw,
                int32_t* nw, int32_t* comm) {
  auto* hg = static_cast<Hypergraph*>(h);
  std::copy(hg->noff.begin(), hg->noff.end(), noff);
  std::copy(hg->pins.begin(), hg->pins.end(), pins);
  std::copy(hg->voff.begin(), hg->voff.end(), voff);
  std::copy(hg->inc.begin(), hg->inc.end(), inc);
  std::copy(hg->vw.begin(), hg->vw.end(), vw);
  std::copy(hg->nw.begin(), hg->nw.end(), nw);
  std::copy(hg->comm.begin(), hg->comm.end(), comm);
}

// ---------------------------------------------------------- partition --
int64_t orc_phg_init(void* h, int k, const int32_t* part, int64_t* part_w, uint32_t* pc, uint64_t* conn) {
  auto* hg = static_cast<Hypergraph*>(h);
  Partition p;
  phg_init(p, *hg, k, part);
  if (part_w) std::copy(p.part_w.begin(), p.part_w.end(), part_w);
  if (pc) std::copy(p.pc.begin(), p.pc.end(), pc);
  if (conn) std::copy(p.conn.begin(), p.conn.end(), conn);
  return km1(p);
}
int64_t orc_cut(void* h, int k, const int32_t* part) {
  Partition p;
  phg_init(p, *static_cast<Hypergraph*>(h), k, part);
  return cut(p);
}
void orc_border(void* h, int k, const int32_t* part, uint8_t* out) {
  auto* hg = static_cast<Hypergraph*>(h);
  Partition p;
  phg_init(p, *hg, k, part);
  for (uint32_t u = 0; u < hg->n; ++u) out[u] = is_border(p, u);
}
void orc_max_gain_move(void* h, int k, const int32_t* part, uint32_t u, const int64_t* max_w, int32_t* to,
                       int32_t* gain) {
  Partition p;
  phg_init(p, *static_cast<Hypergraph*>(h), k, part);
  km1_max_gain_move(p, u, std::vector<int64_t>(max_w, max_w + k), to, gain);
}
// stats: delta, moves, reverts, visits, rounds
int64_t orc_lp_refine(void* h, int k, int32_t* part, const int64_t* max_w, const int64_t* perfect_w,
                      uint64_t seed, int max_rounds, int sub_rounds, int sync, int threads, int disable_rand,
                      int rebalancing, const uint32_t* nodes, uint32_t num_nodes, int64_t* stats) {
  auto* hg = static_cast<Hypergraph*>(h);
  Partition p;
  phg_init(p, *hg, k, part);
  LpParams prm;
  prm.k = k;
  prm.max_w.assign(max_w, max_w + k);
  prm.perfect_w.assign(perfect_w, perfect_w + k);
  prm.seed = seed;
  prm.max_rounds = max_rounds;
  prm.sub_rounds = sub_rounds;
  prm.threads = threads;
  prm.disable_randomization = disable_rand != 0;
  prm.rebalancing = rebalancing != 0;
  if (nodes && num_nodes) {
    prm.refinement_nodes.assign(nodes, nodes + num_nodes);
    std::sort(prm.refinement_nodes.begin(), prm.refinement_nodes.end());
    prm.refinement_nodes.erase(std::unique(prm.refinement_nodes.begin(), prm.refinement_nodes.end()),
                               prm.refinement_nodes.end());
  }
  LpStats st = sync ? lp_refine_sync(p, prm) : lp_refine_ref(p, prm);
  std::copy(p.part.begin(), p.part.end(), part);
  if (stats) { stats[0] = st.delta; stats[1] = st.moves; stats[2] = st.reverts; stats[3] = st.visits; stats[4] = st.rounds; }
  return km1(p);
}

// ---------------------------------------------------------------- FM --
// stats: improvement, kept moves, rounds, searches; returns km1 after
int64_t orc_fm_refine(void* h, int k, int32_t* part, const int64_t* max_w, uint64_t seed, int rounds, int seeds,
                      int pq, int max_moves, uint32_t ignore, uint64_t delta_nets, uint64_t expand_cap, int sub_rounds,
                      int64_t* stats) {
  auto* hg = static_cast<Hypergraph*>(h);
  Partition p;
  phg_init(p, *hg, k, part);
  FmParams prm;
  prm.max_w.assign(max_w, max_w + k);
  prm.seed = seed;
  prm.rounds = rounds;
  prm.seeds = seeds;
  prm.pq = pq;
  prm.max_moves = max_moves;
  prm.ignore = ignore;
  prm.delta_nets = size_t(delta_nets);
  prm.expand_cap = expand_cap;
  prm.sub_rounds = sub_rounds;
  FmStats st = fm_refine_sync(p, prm);
  std::copy(p.part.begin(), p.part.end(), part);
  if (stats) { stats[0] = st.improvement; stats[1] = st.moves; stats[2] = st.rounds; stats[3] = st.searches; }
  return km1(p);
}
// recalculated gains of a move sequence applied in order to (hg, part)
void orc_fm_recalculate(void* h, int k, const int32_t* part, uint32_t nm, const uint32_t* v, const int32_t* from,
                        const int32_t* to, int64_t* gain) {
  Partition p;
  phg_init(p, *static_cast<Hypergraph*>(h), k, part);
  std::vector<FmMove> seq(nm);
  for (uint32_t i = 0; i < nm; ++i) seq[i] = FmMove{v[i], from[i], to[i], 0};
  const std::vector<int64_t> g = fm_recalculate(p, seq);
  std::copy(g.begin(), g.end(), gain);
}
// best balanced prefix of a sequence with the given gains
uint64_t orc_fm_best_prefix(void* h, int k, const int32_t* part, uint32_t nm, const uint32_t* v, const int32_t* from,
                            const int32_t* to, const int64_t* gain, const int64_t* max_w, int64_t* best_gain) {
  Partition p;
  phg_init(p, *static_cast<Hypergraph*>(h), k, part);
  std::vector<FmMove> seq(nm);
  for (uint32_t i = 0; i < nm; ++i) seq[i] = FmMove{v[i], from[i], to[i], 0};
  return fm_best_prefix(p, seq, std::vector<int64_t>(gain, gain + nm), std::vector<int64_t>(max_w, max_w + k),
                        best_gain);
}

// ------------------------------------------------------------- graph --
void* orc_graph_new(void* h, int wt, int ranked) {
  return new Graph(build_graph(*static_cast<Hypergraph*>(h), EdgeWeight(wt), ranked != 0));
}
void orc_graph_free(void* g) { delete static_cast<Graph*>(g); }
void orc_graph_sizes(void* gp, uint32_t* n, uint64_t* arcs, double* tv, uint64_t* max_degree) {
  auto* g = static_cast<Graph*>(gp);
  *n = g->n; *arcs = g->idx[g->n]; *tv = g->tv; *max_degree = g->max_degree;
}
void orc_graph_get(void* gp, uint64_t* idx, uint32_t* head, double* w, double* vol) {
  auto* g = static_cast<Graph*>(gp);
  std::copy(g->idx.begin(), g->idx.end(), idx);
  std::copy(g->head.begin(), g->head.end(), head);
  std::copy(g->w.begin(), g->w.end(), w);
  std::copy(g->vol.begin(), g->vol.end(), vol);
}
void* orc_graph_contract(void* gp, int32_t* C) {
  auto* g = static_cast<Graph*>(gp);
  if (!ids_in_range(C, g->n)) return nullptr;
  std::vector<int32_t> c(C, C + g->n);
  Graph* cg = new Graph(contract_graph(*g, c));
  std::copy(c.begin(), c.end(), C);
  return cg;
}
double orc_modularity(void* gp, const int32_t* C) {
  auto* g = static_cast<Graph*>(gp);
  if (!ids_in_range(C, g->n)) return std::nan("");
  return modularity(*g, std::vector<int32_t>(C, C + g->n));
}
// cluster volumes are initialised from C as in initializeClusterVolumes
int32_t orc_louvain_best_cluster(void* gp, const int32_t* C, uint32_t u, uint64_t sampling, double* best_gain) {
  auto* g = static_cast<Graph*>(gp);
  if (!ids_in_range(C, g->n) || u >= g->n) return -1;
  std::vector<int32_t> c(C, C + g->n);
  std::vector<double> cvol(g->n, 0.0);
  for (uint32_t x = 0; x < g->n; ++x) cvol[c[x]] += g->vol[x];
  return louvain_best_cluster(*g, c, cvol, u, sampling, best_gain);
}
// stats: arcs_swept, rounds, levels, moves
void orc_louvain(void* gp, int sync, uint64_t seed, int max_passes, double min_frac, uint64_t sampling,
                 int sub_rounds, int threads, int disable_rand, int32_t* C_out, int64_t* stats) {
  auto* g = static_cast<Graph*>(gp);
  LouvainParams prm;
  prm.seed = seed; prm.max_passes = max_passes; prm.min_move_fraction = min_frac; prm.sampling = sampling;
  prm.sub_rounds = sub_rounds; prm.threads = threads; prm.disable_randomization = disable_rand != 0;
  LouvainStats st;
  std::vector<int32_t> C = sync ? louvain_sync(*g, prm, &st) : louvain_ref(*g, prm, &st);
  std::copy(C.begin(), C.end(), C_out);
  if (stats) { stats[0] = st.arcs_swept; stats[1] = st.rounds; stats[2] = st.levels; stats[3] = st.moves; }
}
int orc_louvain_level_sync(void* gp, uint64_t seed, int max_passes, double min_frac, uint64_t sampling,
                           int sub_rounds, int32_t* C_out, int64_t* stats) {
  auto* g = static_cast<Graph*>(gp);
  LouvainParams prm;
  prm.seed = seed; prm.max_passes = max_passes; prm.min_move_fraction = min_frac; prm.sampling = sampling;
  prm.sub_rounds = sub_rounds;
  LouvainStats st;
  std::vector<int32_t> C;
  const bool changed = louvain_local_moving_sync(*g, C, prm, &st);
  std::copy(C.begin(), C.end(), C_out);
  if (stats) { stats[0] = st.arcs_swept; stats[1] = st.rounds; stats[2] = st.levels; stats[3] = st.moves; }
  return changed ? 1 : 0;
}

// Level-0 synchronous local moving with its sub-rounds recorded.  Returns the
// number of sub-rounds, or -1 if a capacity (members / moves entries,
// sub-rounds + 1 offsets) is too small.
int orc_louvain_level_sync_trace(void* gp, uint64_t seed, int max_passes, double min_frac, uint64_t sampling,
                                 int sub_rounds, int32_t* C_out, uint32_t* members, uint64_t members_cap,
                                 uint64_t* member_off, uint32_t* mv_u, int32_t* mv_to, uint64_t moves_cap,
                                 uint64_t* move_off, uint64_t offs_cap) {
  auto* g = static_cast<Graph*>(gp);
  LouvainParams prm;
  prm.seed = seed; prm.max_passes = max_passes; prm.min_move_fraction = min_frac; prm.sampling = sampling;
  prm.sub_rounds = sub_rounds;
  LouvainTrace tr;
  std::vector<int32_t> C;
  louvain_local_moving_sync(*g, C, prm, nullptr, &tr);
  if (tr.members.size() > members_cap || tr.mv_u.size() > moves_cap || tr.member_off.size() > offs_cap) return -1;
  std::copy(C.begin(), C.end(), C_out);
  std::copy(tr.members.begin(), tr.members.end(), members);
  std::copy(tr.member_off.begin(), tr.member_off.end(), member_off);
  std::copy(tr.mv_u.begin(), tr.mv_u.end(), mv_u);
  std::copy(tr.mv_to.begin(), tr.mv_to.end(), mv_to);
  std::copy(tr.move_off.begin(), tr.move_off.end(), move_off);
  return int(tr.member_off.size()) - 1;
}

// --------------------------------------------------------- coarsening --
static CoarsenParams coarsen_params(uint32_t contraction_limit, int32_t max_node_weight, double min_shrink,
                                    double max_shrink, uint32_t ignore, uint64_t sampling, uint64_t seed,
                                    int sub_rounds, int disable_rand, int prefer_unmatched, int threads) {
  CoarsenParams p;
  p.contraction_limit = contraction_limit; p.max_node_weight = max_node_weight; p.min_shrink = min_shrink;
  p.max_shrink = max_shrink; p.ignore_he_size = ignore; p.sampling = sampling; p.seed = seed;
  p.sub_rounds = sub_rounds; p.disable_randomization = disable_rand != 0;
  p.prefer_unmatched = prefer_unmatched != 0; p.threads = threads;
  return p;
}
uint32_t orc_cluster_pass(void* h, int sync, int pass, uint32_t contraction_limit, int32_t max_node_weight,
                          double max_shrink, uint32_t ignore, uint64_t sampling, uint64_t seed, int sub_rounds,
                          int disable_rand, int prefer_unmatched, uint32_t* cluster_out, int64_t* pin_visits) {
  auto* hg = static_cast<Hypergraph*>(h);
  CoarsenParams p = coarsen_params(contraction_limit, max_node_weight, 1.01, max_shrink, ignore, sampling, seed,
                                   sub_rounds, disable_rand, prefer_unmatched, 1);
  uint32_t after = 0;
  std::vector<uint32_t> c = sync ? cluster_pass_sync(*hg, p, pass, &after, pin_visits)
                                 : cluster_pass_ref(*hg, p, pass, &after);
  std::copy(c.begin(), c.end(), cluster_out);
  return after;
}
// The rating step of one synchronous clustering sub-round on a given state.
void orc_rate_sync(void* h, int pass, int sub_round, int32_t max_node_weight, uint32_t ignore, uint64_t sampling,
                   uint64_t seed, int prefer_unmatched, const uint32_t* cid, const int64_t* cw,
                   const uint8_t* matched, const uint32_t* proposers, uint32_t count, uint32_t* target_out,
                   uint64_t* rating_out) {
  auto* hg = static_cast<Hypergraph*>(h);
  CoarsenParams p = coarsen_params(0, max_node_weight, 1.01, 2.5, ignore, sampling, seed, 1, 0, prefer_unmatched, 1);
  std::vector<uint32_t> c(cid, cid + hg->n), pr(proposers, proposers + count), t;
  std::vector<int64_t> w(cw, cw + hg->n);
  std::vector<uint8_t> m(matched, matched + hg->n);
  std::vector<uint64_t> r;
  rate_sync(*hg, p, pass, sub_round, c, w, m, pr, t, r);
  std::copy(t.begin(), t.end(), target_out);
  std::copy(r.begin(), r.end(), rating_out);
}
int64_t orc_rating_sample_ref(void* h, const uint32_t* vertices, uint32_t count, uint32_t ignore, int threads) {
  auto* hg = static_cast<Hypergraph*>(h);
  return rating_sample_ref(*hg, std::vector<uint32_t>(vertices, vertices + count), ignore, threads);
}
void* orc_hg_contract(void* h, const uint32_t* cluster, uint32_t* cmap) {
  auto* hg = static_cast<Hypergraph*>(h);
  std::vector<uint32_t> cl(cluster, cluster + hg->n), cm;
  Hypergraph* c = new Hypergraph(contract_hypergraph(*hg, cl, cm));
  std::copy(cm.begin(), cm.end(), cmap);
  return c;
}
// Runs the coarsening driver and returns the number of levels; the coarsest
// hypergraph is returned as a new handle in *coarsest (or null).
// km1 rebalancer (host code shared with the device pipeline): part is
// rebalanced in place; returns the km1 change, moves in *moves
int64_t orc_rebalance(void* h, int k, int32_t* part, const int64_t* max_w, int64_t* moves, int32_t* balanced) {
  const auto* hg = static_cast<const Hypergraph*>(h);
  std::vector<int32_t> p(part, part + hg->n);
  mt_kahypar_amd::RebalanceResult r =
      mt_kahypar_amd::rebalance_km1(to_host(*hg), k, std::vector<int64_t>(max_w, max_w + k), p);
  std::copy(p.begin(), p.end(), part);
  if (moves) *moves = r.moves;
  if (balanced) *balanced = r.balanced ? 1 : 0;
  return r.km1_delta;
}

// degree-zero vertex and large-net removal (sanitize.cpp)
void* orc_reduce(void* h, uint32_t contraction_limit, uint32_t net_threshold, uint32_t* vmap, uint32_t* removed) {
  const auto* hg = static_cast<const Hypergraph*>(h);
  std::vector<uint32_t> vm;
  Hypergraph r = reduce_hypergraph(*hg, contraction_limit, net_threshold, vm, removed[0], removed[1]);
  std::copy(vm.begin(), vm.end(), vmap);
  return new Hypergraph(std::move(r));
}
uint32_t orc_large_net_threshold(uint32_t n, double factor) { return large_net_threshold(n, factor); }
// bin packing of removed vertices (ids in `removed`, weights indexed by id)
void orc_restore_degree_zero(uint32_t nr, const uint32_t* removed, uint32_t n, const int32_t* weight, int k,
                             int64_t* part_w, int32_t* part) {
  std::vector<uint32_t> rm(removed, removed + nr);
  std::vector<int32_t> w(weight, weight + n);
  std::vector<int64_t> pw(part_w, part_w + k);
  std::vector<int32_t> pt(part, part + n);
  restore_degree_zero(rm, w, k, pw, pt);
  std::copy(pw.begin(), pw.end(), part_w);
  std::copy(pt.begin(), pt.end(), part);
}

int orc_coarsen(void* h, int sync, uint32_t contraction_limit, int32_t max_node_weight, double min_shrink,
                double max_shrink, uint32_t ignore, uint64_t sampling, uint64_t seed, int sub_rounds,
                int disable_rand, int prefer_unmatched, void** coarsest) {
  auto* hg = static_cast<Hypergraph*>(h);
  CoarsenParams p = coarsen_params(contraction_limit, max_node_weight, min_shrink, max_shrink, ignore, sampling,
                                   seed, sub_rounds, disable_rand, prefer_unmatched, 1);
  std::vector<Level> lv = coarsen(*hg, p, sync != 0, nullptr);
  *coarsest = lv.empty() ? nullptr : new Hypergraph(lv.back().hg);
  return int(lv.size());
}

// ----------------------------------------------------------- pipeline --
// partition (partitioner.cpp:163-204) with the fast preset: community
// detection, removal of degree-zero vertices and large nets, multilevel
// coarsening, initial partitioning on the coarsest hypergraph, uncoarsening
// with LP at every level (FM disabled), restore of the removed vertices.
// times: louvain, coarsening, initial partitioning, refinement, total
// counters: louvain arcs swept, lp incidences visited, levels, rating pin
// visits, removed degree-zero vertices, removed large nets, FM improvement
int64_t orc_partition(void* h, int k, double eps, uint64_t seed, int sync, int threads, int fm_rounds,
                      int32_t* part_out, double* times, int64_t* counters) {
  const double t0 = now();
  Hypergraph hg = *static_cast<Hypergraph*>(h);
  // setupPartWeights / setupContractionLimit / setupMaximumAllowedNodeWeight
  // (context.cpp:213-284)
  const int64_t W = hg.total_w;
  const int64_t perfect = int64_t(std::ceil(double(W) / double(k)));
  const int64_t maxw = int64_t(int32_t((1.0 + eps) * double(perfect)));
  const uint32_t climit = 160u * uint32_t(k);
  int32_t max_node_w = int32_t(std::ceil(1.0 / double(climit) * double(W)));
  max_node_w = int32_t(std::min<int64_t>(max_node_w, maxw));
  // preprocess (partitioner.cpp:62-72,106-126)
  const double density = double(hg.m) / double(hg.n);
  const EdgeWeight wt = density < 0.75 ? EdgeWeight::degree : EdgeWeight::uniform;
  LouvainParams lp;
  lp.seed = seed; lp.threads = threads;
  LouvainStats lst;
  {
    Graph g = build_graph(hg, wt, true);
    std::vector<int32_t> C = sync ? louvain_sync(g, lp, &lst) : louvain_ref(g, lp, &lst);
    for (uint32_t u = 0; u < hg.n; ++u) hg.comm[u] = C[u];
  }
  const double t1 = now();
  // sanitize (partitioner.cpp:74-86): degree-zero vertices and large nets
  std::vector<uint32_t> vmap;
  uint32_t removed_vertices = 0, removed_nets = 0;
  Hypergraph red = reduce_hypergraph(hg, climit, large_net_threshold(hg.n, 0.01), vmap, removed_vertices,
                                     removed_nets);
  const bool reduced = removed_vertices > 0 || removed_nets > 0;
  const Hypergraph& top = reduced ? red : hg;
  CoarsenParams cp;
  cp.contraction_limit = climit; cp.max_node_weight = max_node_w; cp.seed = seed; cp.threads = threads;
  int64_t rating_visits = 0;
  std::vector<Level> levels = coarsen(top, cp, sync != 0, &rating_visits);
  const double t2 = now();
  const Hypergraph& coarsest = levels.empty() ? top : levels.back().hg;
  IpParams ip;
  ip.k = k; ip.max_w.assign(k, maxw); ip.seed = seed; ip.threads = std::max(1, threads);
  std::vector<int32_t> part = initial_partition(coarsest, ip, nullptr);
  const double t3 = now();
  LpParams prm;
  prm.k = k; prm.max_w.assign(k, maxw); prm.perfect_w.assign(k, perfect); prm.seed = seed; prm.threads = threads;
  int64_t lp_visits = 0;
  // LP and then FM at every level (uncoarsening.cpp:94-188); FM has one
  // form, the synchronous one the device computes, in both modes
  FmParams fprm;
  fprm.max_w.assign(k, maxw);
  fprm.rounds = fm_rounds;
  int64_t fm_improvement = 0;
  auto refine = [&](const Hypergraph& g, std::vector<int32_t>& pt, int level) {
    Partition p;
    phg_init(p, g, k, pt.data());
    prm.seed = seed + uint64_t(level) * 1000003ull;
    LpStats st = sync ? lp_refine_sync(p, prm) : lp_refine_ref(p, prm);
    lp_visits += st.visits;
    if (fm_rounds > 0) {
      fprm.seed = prm.seed;
      fm_improvement += fm_refine_sync(p, fprm).improvement;
    }
    pt = p.part;
  };
  refine(coarsest, part, int(levels.size()));
  for (int i = int(levels.size()) - 1; i >= 0; --i) {
    const Hypergraph& fine = (i == 0) ? top : levels[i - 1].hg;
    std::vector<int32_t> fp(fine.n);
    for (uint32_t u = 0; u < fine.n; ++u) fp[u] = part[levels[i].cmap[u]];
    part.swap(fp);
    refine(fine, part, i);
  }
  {
    // uncoarsening.cpp:146-180: an imbalanced top-level partition is
    // rebalanced (host code shared with the device pipeline)
    std::vector<int64_t> pw(k, 0);
    for (uint32_t u = 0; u < top.n; ++u) pw[part[u]] += top.vw[u];
    bool over = false;
    for (int b = 0; b < k; ++b) over |= pw[b] > maxw;
    if (over) mt_kahypar_amd::rebalance_km1(to_host(top), k, std::vector<int64_t>(k, maxw), part);
  }
  if (reduced) {
    // restore (partitioner.cpp:190-195): large nets return with their pins,
    // removed vertices are bin-packed
    Partition p;
    phg_init(p, top, k, part.data());
    std::vector<int64_t> pw(p.part_w.begin(), p.part_w.end());
    std::vector<int32_t> full(hg.n, -1);
    std::vector<uint32_t> removed;
    for (uint32_t u = 0; u < hg.n; ++u) {
      if (vmap[u] != 0xffffffffu) full[u] = part[vmap[u]];
      else removed.push_back(u);
    }
    restore_degree_zero(removed, hg.vw, k, pw, full);
    part.swap(full);
  }
  const double t4 = now();
  Partition fin;
  phg_init(fin, hg, k, part.data());
  std::copy(part.begin(), part.end(), part_out);
  if (times) { times[0] = t1 - t0; times[1] = t2 - t1; times[2] = t3 - t2; times[3] = t4 - t3; times[4] = t4 - t0; }
  if (counters) {
    counters[0] = lst.arcs_swept; counters[1] = lp_visits; counters[2] = int64_t(levels.size());
    counters[3] = rating_visits; counters[4] = removed_vertices; counters[5] = removed_nets;
    counters[6] = fm_improvement;
  }
  return km1(fin);
}

// ---------------------------------------------------- initial partitioning --
int64_t orc_initial_partition(void* h, int k, const int64_t* max_w, uint64_t seed, int runs, int threads,
                              int32_t* part_out) {
  const Hypergraph& hg = *static_cast<Hypergraph*>(h);
  IpParams ip;
  ip.k = k; ip.max_w.assign(max_w, max_w + k); ip.seed = seed; ip.runs = runs; ip.threads = std::max(1, threads);
  int64_t km1v = 0;
  std::vector<int32_t> part = initial_partition(hg, ip, &km1v);
  std::copy(part.begin(), part.end(), part_out);
  return km1v;
}
// one run of a bisection: side per vertex, out = {cut, balanced, side-0 weight}
void orc_ip_bisect_run(void* h, int64_t max0, int64_t max1, int64_t target0, uint64_t stream, int r, int passes,
                       uint8_t* side_out, int64_t* out) {
  const Hypergraph& hg = *static_cast<Hypergraph*>(h);
  IpRun run = ip_bisect_run(hg, max0, max1, target0, stream, r, passes);
  std::copy(run.side.begin(), run.side.end(), side_out);
  out[0] = run.cut; out[1] = run.balanced; out[2] = run.w0;
}
void orc_ip_perm(uint64_t key, uint32_t n, uint32_t* out) {
  for (uint32_t i = 0; i < n; ++i) out[i] = ip_perm(key, n, i);
}
uint64_t orc_ip_stream(uint64_t seed, int first, int k) { return ip_stream(seed, first, k); }
uint64_t orc_ip_run_key(uint64_t stream, int r) { return ip_run_key(stream, r); }
uint32_t orc_ip_start(uint64_t rk, uint32_t n) { return ip_start(rk, n); }
int64_t orc_ip_move_check(void* h, const uint8_t* side, const uint32_t* moves, uint32_t count) {
  const Hypergraph& hg = *static_cast<Hypergraph*>(h);
  return ip_move_check(hg, std::vector<uint8_t>(side, side + hg.n), std::vector<uint32_t>(moves, moves + count));
}

}  // extern "C"
