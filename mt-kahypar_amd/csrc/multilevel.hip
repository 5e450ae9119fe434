// Level hierarchy, projection and the multilevel partitioner.
//
// Reference: MultilevelCoarsener::coarsenImpl (multilevel_coarsener.h:143-315),
// MultilevelCoarsenerBase::doUncoarsen / refine (uncoarsening.cpp:94-188,
// 404-457), partition() (partitioner.cpp:163-204) and the context set-up of
// context.cpp:213-284.
#include <chrono>
#include <cmath>
#include <memory>

#include "../host/rebalancer.h"
#include "internal.h"

struct mkp_hierarchy {
  mkp_hypergraph* top = nullptr;             // borrowed input
  std::vector<mkp_hypergraph*> levels;       // owned coarse levels 1..L
  std::vector<mkp::DevBuf<uint32_t>> cmaps;  // cmaps[i]: level i -> level i+1
  int64_t rating_visits = 0;
  ~mkp_hierarchy() {
    for (auto* h : levels) delete h;
  }
  mkp_hypergraph* level(int i) const { return i == 0 ? top : levels[size_t(i) - 1]; }
  int num_levels() const { return int(levels.size()) + 1; }
};

namespace mkp {

__global__ void project_kernel(const int32_t* __restrict__ coarse_part, const uint32_t* __restrict__ cmap, uint32_t n,
                               int32_t* __restrict__ part) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    part[u] = coarse_part[cmap[u]];
}

mkp_hierarchy* coarsen(mkp_hypergraph* hg, const mkp_coarsen_params& prm) {
  auto* h = new mkp_hierarchy();
  try {
    h->top = hg;
    int pass = 0;
    mkp_hypergraph* cur = hg;
    while (cur->n > prm.contraction_limit) {
      DevBuf<uint32_t> cid(std::max<uint32_t>(cur->n, 1));
      ClusterResult r = cluster_pass(cur, prm, pass, cid.p);
      h->rating_visits += r.visits;
      // stop if the pass shrinks by at most the minimum factor
      // (multilevel_coarsener.h:275-280)
      if (r.nodes_after == 0 || double(cur->n) / double(r.nodes_after) <= prm.min_shrink_factor) break;
      DevBuf<uint32_t> cmap(std::max<uint32_t>(cur->n, 1));
      mkp_hypergraph* coarse = contract(cur, cid.p, cmap.p);
      h->levels.push_back(coarse);
      h->cmaps.push_back(std::move(cmap));
      cur = coarse;
      ++pass;
    }
  } catch (...) {
    delete h;
    throw;
  }
  return h;
}

void project(mkp_partition* fine, const mkp_partition* coarse, const mkp_hierarchy* h, int level) {
  Context& c = fine->hg->ctx->c;
  project_kernel<<<grid_for(fine->hg->n, 256, 65536), 256, 0, c.stream>>>(coarse->part.p, h->cmaps[level].p,
                                                                          fine->hg->n, fine->part.p);
  MKP_LAUNCH_CHECK();
  partition_initialize(fine);
}

mt_kahypar_amd::HostHypergraph download(mkp_hypergraph* hg) {
  Context& c = hg->ctx->c;
  mt_kahypar_amd::HostHypergraph h;
  h.n = hg->n;
  h.m = hg->m;
  std::vector<uint32_t> off(size_t(hg->m) + 1);
  h.pins.resize(hg->p);
  h.vw.resize(hg->n);
  h.nw.resize(hg->m);
  MKP_HIP(hipMemcpyAsync(off.data(), hg->noff.p, off.size() * 4, hipMemcpyDeviceToHost, c.stream));
  if (hg->p) MKP_HIP(hipMemcpyAsync(h.pins.data(), hg->pins.p, hg->p * 4, hipMemcpyDeviceToHost, c.stream));
  if (hg->n) MKP_HIP(hipMemcpyAsync(h.vw.data(), hg->vw.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
  if (hg->m) MKP_HIP(hipMemcpyAsync(h.nw.data(), hg->nw.p, size_t(hg->m) * 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  h.noff.assign(off.begin(), off.end());
  mt_kahypar_amd::build_host_incidence(h);
  return h;
}

// Km1Rebalancer on the host (rebalancer.cpp:36-179) for the partition p:
// the hypergraph and block ids go to the host, the result comes back and the
// partition state is rebuilt.
mt_kahypar_amd::RebalanceResult rebalance_partition(mkp_partition* p, const std::vector<int64_t>& max_w) {
  Context& c = p->hg->ctx->c;
  const uint32_t n = p->hg->n;
  std::vector<int32_t> part(n);
  if (n) MKP_HIP(hipMemcpyAsync(part.data(), p->part.p, size_t(n) * 4, hipMemcpyDeviceToHost, c.stream));
  mt_kahypar_amd::HostHypergraph h = download(p->hg);  // synchronises
  mt_kahypar_amd::RebalanceResult r = mt_kahypar_amd::rebalance_km1(h, p->k, max_w, part);
  if (r.moves) {
    MKP_HIP(hipMemcpyAsync(p->part.p, part.data(), size_t(n) * 4, hipMemcpyHostToDevice, c.stream));
    partition_initialize(p);
    c.sync();
  }
  return r;
}

// Checks every parameter the pipeline forwards before any kernel runs.
// Sub-round counts go into 8-bit fields of the round tags and index 256-entry
// bucket tables, so each must lie in [1, 255] (0 selects the default).
void validate_partition_params(const mkp_partition_params& p) {
  MKP_CHECK(p.k >= 2 && p.epsilon >= 0, "bad k or epsilon");
  MKP_CHECK(p.contraction_limit_multiplier >= 1 && p.max_shrink_factor > 1.0 &&
                p.max_allowed_weight_multiplier > 0 && p.lp_max_rounds >= 0,
            "bad preset parameters");
  MKP_CHECK(p.lp_sub_rounds >= 0 && p.lp_sub_rounds <= 255, "lp_sub_rounds must be in [1, 255] (0: default)");
  MKP_CHECK(p.louvain_sub_rounds >= 0 && p.louvain_sub_rounds <= 255,
            "louvain_sub_rounds must be in [1, 255] (0: default)");
  MKP_CHECK(p.coarsening_sub_rounds >= 0 && p.coarsening_sub_rounds <= 255,
            "coarsening_sub_rounds must be in [1, 255] (0: default)");
  MKP_CHECK(p.louvain_max_passes >= 0 && p.ip_threads >= 0 && p.ip_threads <= 256, "bad preset parameters");
  MKP_CHECK(p.fm_rounds >= 0 && p.fm_rounds <= 1000, "fm_rounds must be in [0, 1000]");
  MKP_CHECK(p.fm_rounds == 0 || (p.fm_seeds >= 1 && p.fm_seeds <= 64), "fm_seeds must be in [1, 64]");
  MKP_CHECK(p.fm_rounds == 0 || p.k <= 128, "FM supports at most 128 blocks (set fm_rounds = 0)");
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace mkp

using namespace mkp;

extern "C" {

int mkp_coarsen(mkp_hypergraph* hg, const mkp_coarsen_params* prm, mkp_hierarchy** out) {
  return guarded([&] {
    MKP_CHECK(hg && prm && out, "null argument");
    MKP_CHECK(prm->sub_rounds >= 1 && prm->sub_rounds <= 255 && prm->max_shrink_factor > 0, "bad parameters");
    Bind bind_(hg->ctx->c);
    *out = coarsen(hg, *prm);
  });
}

int mkp_hierarchy_num_levels(const mkp_hierarchy* h, int32_t* levels) {
  return guarded([&] {
    MKP_CHECK(h && levels, "null argument");
    *levels = h->num_levels();
  });
}

int mkp_hierarchy_level(mkp_hierarchy* h, int32_t level, mkp_hypergraph** hg) {
  return guarded([&] {
    MKP_CHECK(h && hg, "null argument");
    MKP_CHECK(level >= 0 && level < h->num_levels(), "level out of range");
    *hg = h->level(level);
  });
}

int mkp_hierarchy_destroy(mkp_hierarchy* h) {
  return guarded([&] { delete h; });
}

int mkp_partition_project(mkp_partition* fine, const mkp_partition* coarse, const mkp_hierarchy* h, int32_t level) {
  return guarded([&] {
    MKP_CHECK(fine && coarse && h, "null argument");
    MKP_CHECK(level >= 0 && level + 1 < h->num_levels(), "level out of range");
    MKP_CHECK(fine->hg == h->level(level) && coarse->hg == h->level(level + 1), "partitions not on levels i, i+1");
    MKP_CHECK(fine->k == coarse->k, "different k");
    Bind bind_(fine->hg->ctx->c);
    project(fine, coarse, h, level);
    fine->hg->ctx->c.sync();
  });
}

int mkp_initial_partition(mkp_partition* p, double epsilon, uint64_t seed, int32_t runs, int64_t* km1) {
  return guarded([&] {
    MKP_CHECK(p, "null argument");
    MKP_CHECK(epsilon >= 0 && runs >= 1, "bad epsilon or runs");
    Context& c = p->hg->ctx->c;
    Bind bind_(c);
    const int k = p->k;
    const int64_t W = p->hg->total_w;
    const int64_t perfect = int64_t(std::ceil(double(W) / double(k)));
    const std::vector<int64_t> max_w(k, int64_t(int32_t((1.0 + epsilon) * double(perfect))));
    initial_partition(p->hg, k, max_w, seed, runs, 4, p->part.p);
    partition_initialize(p);
    c.sync();
    if (km1) *km1 = partition_km1(p);
  });
}

int mkp_ip_bisect_runs(mkp_hypergraph* hg, int64_t max0, int64_t max1, int64_t target0, uint64_t stream,
                       int32_t runs, int32_t passes, uint8_t* sides, int64_t* results) {
  return guarded([&] {
    MKP_CHECK(hg && sides && results, "null argument");
    MKP_CHECK(runs >= 1 && runs <= 1024 && passes >= 0 && passes <= 64, "runs must be in [1, 1024], passes in [0, 64]");
    Bind bind_(hg->ctx->c);
    ip_bisect_runs(hg, max0, max1, target0, stream, runs, passes, sides, results);
  });
}

int mkp_rebalance(mkp_partition* p, const int64_t* max_part_weights, int64_t* km1_delta, int64_t* moves) {
  return guarded([&] {
    MKP_CHECK(p && max_part_weights, "null argument");
    Context& c = p->hg->ctx->c;
    Bind bind_(c);
    std::vector<int64_t> mw(max_part_weights, max_part_weights + p->k);
    mt_kahypar_amd::RebalanceResult r = rebalance_partition(p, mw);
    c.sync();
    if (km1_delta) *km1_delta = r.km1_delta;
    if (moves) *moves = r.moves;
  });
}

int mkp_partition_params_default(mkp_partition_params* p, int32_t k, double epsilon, uint64_t seed) {
  return guarded([&] {
    MKP_CHECK(p, "null argument");
    *p = mkp_partition_params{};
    p->k = k;
    p->epsilon = epsilon;
    p->seed = seed;
    p->use_community_detection = 1;
    p->lp_sub_rounds = 4;
    p->louvain_sub_rounds = 8;
    p->coarsening_sub_rounds = 16;
    p->ip_runs = 20;
    p->louvain_edge_weight = -1;
    p->louvain_max_passes = 5;
    p->louvain_min_move_fraction = 0.01;
    p->louvain_sampling = 200000;
    p->min_shrink_factor = 1.01;
    p->max_shrink_factor = 2.5;
    p->max_allowed_weight_multiplier = 1.0;
    p->contraction_limit_multiplier = 160;
    p->ignore_net_size = 1000;
    p->rating_sampling = 200000;
    p->lp_enabled = 1;
    p->lp_max_rounds = 5;
    p->lp_activation_threshold = 100;
    p->lp_rebalancing = 1;
    p->ip_threads = 4;
    p->large_net_remove_factor = 0.01;
    p->fm_rounds = 0;  // FM is switched on by r-fm-type=fm_gain_cache (10 rounds) or fm_rounds
    p->fm_seeds = 25;
  });
}

int mkp_partition_hypergraph(mkp_hypergraph* hg, const mkp_partition_params* prm, int32_t* parts,
                             mkp_partition_stats* stats) {
  return guarded([&] {
    MKP_CHECK(hg && prm && parts, "null argument");
    validate_partition_params(*prm);
    Context& c = hg->ctx->c;
    Bind bind_(c);
    const double t0 = now_ms();
    const int k = prm->k;
    // setupPartWeights / setupContractionLimit / setupMaximumAllowedNodeWeight
    const int64_t W = hg->total_w;
    const int64_t perfect = int64_t(std::ceil(double(W) / double(k)));
    const int64_t maxw = int64_t(int32_t((1.0 + prm->epsilon) * double(perfect)));
    const uint32_t climit = prm->contraction_limit_multiplier * uint32_t(k);
    int64_t max_node_w = int64_t(std::ceil(prm->max_allowed_weight_multiplier / double(climit) * double(W)));
    max_node_w = std::min(max_node_w, maxw);
    mkp_partition_stats st{};
    // preprocess: community detection (partitioner.cpp:106-126)
    if (prm->use_community_detection) {
      mkp_louvain_params lp{prm->louvain_edge_weight, prm->louvain_max_passes, prm->louvain_min_move_fraction,
                            prm->louvain_sampling, prm->seed,
                            prm->louvain_sub_rounds > 0 ? prm->louvain_sub_rounds : 8};
      mkp_louvain_stats ls{};
      MKP_CHECK(mkp_louvain(hg, &lp, nullptr, &ls) == 0, mkp_last_error());
      st.louvain_arcs = ls.arcs_swept;
    } else {
      fill_async<int32_t>(hg->comm.p, 0, hg->n, c.stream);
    }
    // sanitize (partitioner.cpp:74-86): degree-zero vertices and large nets
    // are set aside for the multilevel phase
    Reduction red;
    reduce(hg, climit, large_net_threshold(hg->n, prm->large_net_remove_factor), red);
    st.removed_vertices = red.removed_vertices;
    st.removed_nets = red.removed_nets;
    mkp_hypergraph* work = red.hg ? red.hg.get() : hg;
    const double t1 = now_ms();
    mkp_coarsen_params cp{climit, int32_t(max_node_w), prm->min_shrink_factor, prm->max_shrink_factor,
                          prm->ignore_net_size, prm->rating_sampling, prm->seed,
                          prm->coarsening_sub_rounds > 0 ? prm->coarsening_sub_rounds : 16};
    mkp_hierarchy* h = coarsen(work, cp);
    std::unique_ptr<mkp_hierarchy> hold(h);
    st.rating_visits = h->rating_visits;
    const double t2 = now_ms();
    const int L = h->num_levels();
    mkp_hypergraph* coarsest = h->level(L - 1);
    std::vector<int64_t> maxv(k, maxw), perv(k, perfect);
    std::vector<std::unique_ptr<mkp_partition>> phg(L);
    {
      mkp_partition* p = nullptr;
      MKP_CHECK(mkp_partition_create(coarsest, k, &p) == 0, mkp_last_error());
      phg[L - 1].reset(p);
      initial_partition(coarsest, k, maxv, prm->seed, prm->ip_runs > 0 ? prm->ip_runs : 20, 4, p->part.p);
      c.sync();
    }
    const double t3 = now_ms();
    mkp_lp_params lpp{prm->lp_max_rounds, prm->lp_sub_rounds > 0 ? prm->lp_sub_rounds : 4,
                      prm->lp_activation_threshold, prm->lp_rebalancing, 0, maxv.data(), perv.data()};
    for (int i = L - 1; i >= 0; --i) {
      mkp_partition* p = phg[i].get();
      if (i == L - 1) {
        partition_initialize(p);
      } else {
        MKP_CHECK(mkp_partition_create(h->level(i), k, &p) == 0, mkp_last_error());
        phg[i].reset(p);
        project(p, phg[i + 1].get(), h, i);
        phg[i + 1].reset();
      }
      if (prm->lp_enabled && prm->lp_max_rounds > 0) {
        lpp.seed = prm->seed + uint64_t(i) * 1000003ull;
        mkp_lp_stats ls{};
        lp_refine(p, lpp, &ls);
        st.lp_visits += ls.visits;
      }
      if (prm->fm_rounds > 0) {
        const double f0 = now_ms();
        mkp_fm_params fp{maxv.data(), prm->fm_rounds, prm->fm_seeds, prm->ignore_net_size,
                         prm->seed + uint64_t(i) * 1000003ull};
        mkp_fm_stats fs{};
        fm_refine(p, fp, &fs);
        st.fm_improvement += fs.improvement;
        st.fm_ms += now_ms() - f0;
      }
    }
    mkp_partition* top = phg[0].get();
    {
      // uncoarsening.cpp:146-180: an imbalanced top-level partition is
      // rebalanced (the initial partitioning repairs balance and LP keeps
      // it, so this runs only when the bounds cannot all be met)
      std::vector<int64_t> tw(k);
      MKP_HIP(hipMemcpyAsync(tw.data(), top->part_w.p, size_t(k) * 8, hipMemcpyDeviceToHost, c.stream));
      c.sync();
      bool over = false;
      for (int b = 0; b < k; ++b) over |= tw[b] > maxw;
      if (over) rebalance_partition(top, maxv);
    }
    std::unique_ptr<mkp_partition> full;
    if (red.hg) {
      // restore (partitioner.cpp:190-195): large nets return with their
      // pins, the removed vertices are bin-packed into the lightest blocks
      restore(hg, red, top, parts);
      phg[0].reset();
      mkp_partition* p = nullptr;
      MKP_CHECK(mkp_partition_create(hg, k, &p) == 0, mkp_last_error());
      full.reset(p);
      MKP_HIP(hipMemcpyAsync(p->part.p, parts, size_t(hg->n) * 4, hipMemcpyHostToDevice, c.stream));
      partition_initialize(p);
      top = p;
    } else {
      MKP_HIP(hipMemcpyAsync(parts, top->part.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
    }
    std::vector<int64_t> pw(k);
    MKP_HIP(hipMemcpyAsync(pw.data(), top->part_w.p, size_t(k) * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    st.km1 = partition_km1(top);
    double imb = 0;
    for (int b = 0; b < k; ++b) imb = std::max(imb, double(pw[b]) / double(perfect));
    st.imbalance = imb - 1.0;
    st.levels = L;
    const double t4 = now_ms();
    st.louvain_ms = t1 - t0; st.coarsening_ms = t2 - t1; st.initial_ms = t3 - t2; st.refinement_ms = t4 - t3;
    st.total_ms = t4 - t0;
    if (stats) *stats = st;
  });
}

}  // extern "C"
