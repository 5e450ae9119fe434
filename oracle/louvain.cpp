// louvain.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Star expansion, graph contraction, modularity and Louvain local moving.
// Compiled with -ffp-contract=off: the gain expressions keep the reference's
// operation order w - (recip*volU)*vol without fused multiply-adds.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <omp.h>

namespace orc {

// Graph::Graph / construct / constructBipartiteGraph / constructGraph
// (graph.cpp:36-78, 290-446)
std::vector<uint32_t> net_locality_ranks(const Hypergraph& hg) {
  std::vector<std::pair<uint32_t, uint32_t>> key(hg.m);
  for (uint32_t e = 0; e < hg.m; ++e) {
    const uint32_t s = std::min<uint32_t>(hg.esize(e), 9);
    std::vector<uint32_t> p(hg.pins.begin() + int64_t(hg.noff[e]), hg.pins.begin() + int64_t(hg.noff[e] + s));
    std::sort(p.begin(), p.end());
    key[e] = {s ? p[(s - 1) / 2] : 0u, e};
  }
  std::sort(key.begin(), key.end());
  std::vector<uint32_t> rank(hg.m);
  for (uint32_t r = 0; r < hg.m; ++r) rank[key[r].second] = r;
  return rank;
}

Graph build_graph(const Hypergraph& hg, EdgeWeight wt, bool ranked) {
  auto f = [&](int32_t w, uint32_t size, uint32_t deg) -> double {
    switch (wt) {
      case EdgeWeight::uniform: return double(w);
      case EdgeWeight::non_uniform: return double(w) / double(size);
      default: return double(w) * (double(deg) / double(size));
    }
  };
  bool is_graph = true;
  for (uint32_t e = 0; e < hg.m && is_graph; ++e) is_graph = hg.esize(e) == 2;
  Graph g;
  if (is_graph) {
    g.n = hg.n;
    g.idx.assign(size_t(g.n) + 1, 0);
    for (uint32_t u = 0; u < hg.n; ++u) g.idx[u + 1] = g.idx[u] + hg.degree(u);
    g.head.resize(g.idx[g.n]);
    g.w.resize(g.idx[g.n]);
    for (uint32_t u = 0; u < hg.n; ++u) {
      uint64_t pos = g.idx[u];
      g.max_degree = std::max<uint64_t>(g.max_degree, hg.degree(u));
      for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        uint32_t v = UINT32_MAX;
        for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j)
          if (hg.pins[j] != u) { v = hg.pins[j]; break; }
        g.head[pos] = v;
        g.w[pos++] = f(hg.nw[e], 2, hg.degree(u));
      }
    }
  } else {
    const uint32_t N = hg.n;
    std::vector<uint32_t> rank(hg.m);
    if (ranked) rank = net_locality_ranks(hg);
    else for (uint32_t e = 0; e < hg.m; ++e) rank[e] = e;
    g.n = hg.n + hg.m;
    g.idx.assign(size_t(g.n) + 1, 0);
    for (uint32_t u = 0; u < N; ++u) g.idx[u + 1] = hg.degree(u);
    for (uint32_t e = 0; e < hg.m; ++e) g.idx[N + rank[e] + 1] = hg.esize(e);
    for (uint32_t x = 0; x < g.n; ++x) g.idx[x + 1] += g.idx[x];
    g.head.resize(g.idx[g.n]);
    g.w.resize(g.idx[g.n]);
    for (uint32_t u = 0; u < N; ++u) {
      uint64_t pos = g.idx[u];
      g.max_degree = std::max<uint64_t>(g.max_degree, hg.degree(u));
      for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        g.head[pos] = N + rank[e];
        g.w[pos++] = f(hg.nw[e], hg.esize(e), hg.degree(u));
      }
    }
    for (uint32_t e = 0; e < hg.m; ++e) {
      uint64_t pos = g.idx[N + rank[e]];
      g.max_degree = std::max<uint64_t>(g.max_degree, hg.esize(e));
      for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j) {
        const uint32_t v = hg.pins[j];
        g.head[pos] = v;
        g.w[pos++] = f(hg.nw[e], hg.esize(e), hg.degree(v));
      }
    }
  }
  // node volumes and total volume (graph.cpp:319-327)
  g.vol.assign(g.n, 0.0);
  g.tv = 0;
  for (uint32_t u = 0; u < g.n; ++u) {
    double v = 0;
    for (uint64_t a = g.idx[u]; a < g.idx[u + 1]; ++a) v += g.w[a];
    g.vol[u] = v;
    g.tv += v;
  }
  return g;
}

// Graph::contract (graph.cpp:128-268).  Coarse ids are the ranks of the used
// cluster ids; arcs of a coarse node are sorted by head with equal heads
// merged; self loops dropped; volumes summed; total volume kept.
Graph contract_graph(const Graph& g, std::vector<int32_t>& C) {
  Graph cg;
  cg.tv = g.tv;
  std::vector<uint64_t> mapping(g.n, 0);
  for (uint32_t u = 0; u < g.n; ++u) mapping[C[u]] = 1;
  uint64_t run = 0;
  for (uint32_t u = 0; u < g.n; ++u) { const uint64_t x = mapping[u]; mapping[u] = run; run += x; }
  cg.n = uint32_t(run);
  for (uint32_t u = 0; u < g.n; ++u) C[u] = int32_t(mapping[C[u]]);
  std::vector<uint64_t> cnt(size_t(cg.n) + 1, 0);
  cg.vol.assign(cg.n, 0.0);
  for (uint32_t u = 0; u < g.n; ++u) {
    cg.vol[C[u]] += g.vol[u];
    for (uint64_t a = g.idx[u]; a < g.idx[u + 1]; ++a)
      if (C[g.head[a]] != C[u]) cnt[C[u] + 1]++;
  }
  for (uint32_t c = 0; c < cg.n; ++c) cnt[c + 1] += cnt[c];
  std::vector<std::pair<uint32_t, double>> tmp(cnt[cg.n]);
  std::vector<uint64_t> pos(cnt.begin(), cnt.end() - 1);
  for (uint32_t u = 0; u < g.n; ++u)
    for (uint64_t a = g.idx[u]; a < g.idx[u + 1]; ++a) {
      const uint32_t cv = uint32_t(C[g.head[a]]);
      if (cv != uint32_t(C[u])) tmp[pos[C[u]]++] = {cv, g.w[a]};
    }
  cg.idx.assign(size_t(cg.n) + 1, 0);
  for (uint32_t c = 0; c < cg.n; ++c) {
    auto b = tmp.begin() + cnt[c], e = tmp.begin() + cnt[c + 1];
    std::stable_sort(b, e, [](const auto& x, const auto& y) { return x.first < y.first; });
    uint64_t deg = 0;
    for (auto it = b; it != e; ++it) {
      if (it != b && it->first == (it - 1)->first) { cg.w.back() += it->second; continue; }
      cg.head.push_back(it->first);
      cg.w.push_back(it->second);
      ++deg;
    }
    cg.idx[c + 1] = cg.idx[c] + deg;
    cg.max_degree = std::max(cg.max_degree, deg);
  }
  return cg;
}

// metrics::modularity (local_moving_modularity.cpp:28-56)
double modularity(const Graph& g, const std::vector<int32_t>& C) {
  std::vector<double> internal(g.n, 0.0), total(g.n, 0.0);
  for (uint32_t u = 0; u < g.n; ++u) {
    total[C[u]] += g.vol[u];
    internal[C[u]] += g.vol[u];
    for (uint64_t a = g.idx[u]; a < g.idx[u + 1]; ++a)
      if (C[g.head[a]] != C[u]) internal[C[u]] -= g.w[a];
  }
  double q = 0;
  for (uint32_t c = 0; c < g.n; ++c)
    if (total[c] > 0.0) q += internal[c] - (total[c] * total[c]) / g.tv;
  return q / g.tv;
}

// computeMaxGainCluster + modularityGain (local_moving_modularity.h:87-137).
// The insertion-ordered sparse map (sparse_map.h:247-430) is a dense
// (key, weight) list in first-occurrence order.
namespace {
struct InsertionMap {
  std::vector<int32_t> slot;  // cluster -> position, -1 if absent
  std::vector<std::pair<int32_t, double>> items;
  explicit InsertionMap(size_t n) : slot(n, -1) {}
  double& operator[](int32_t c) {
    if (slot[c] < 0) { slot[c] = int32_t(items.size()); items.push_back({c, 0.0}); }
    return items[slot[c]].second;
  }
  void clear() { for (auto& it : items) slot[it.first] = -1; items.clear(); }
};

int32_t best_cluster_with(const Graph& g, const int32_t* C, const double* cvol, uint32_t u,
                          uint64_t sampling, double recip, InsertionMap& map, double* best_gain) {
  const int32_t from = C[u];
  int32_t best = from;
  uint64_t end = g.idx[u + 1];
  if (sampling < end - g.idx[u]) end = g.idx[u] + sampling;
  for (uint64_t a = g.idx[u]; a < end; ++a) map[C[g.head[a]]] += g.w[a];
  const double volume_from = cvol[from];
  const double volU = g.vol[u];
  const double weight_from = map[from];
  const double mult = recip * volU;
  double bg = weight_from - mult * (volume_from - volU);
  for (const auto& it : map.items) {
    const int32_t to = it.first;
    if (to == from) continue;
    const double gain = it.second - mult * cvol[to];
    if (gain > bg) { best = to; bg = gain; }
  }
  map.clear();
  if (best_gain) *best_gain = bg;
  return best;
}
}  // namespace

int32_t louvain_best_cluster(const Graph& g, const std::vector<int32_t>& C,
                             const std::vector<double>& cvol, uint32_t u, uint64_t sampling,
                             double* best_gain) {
  InsertionMap map(g.n);
  return best_cluster_with(g, C.data(), cvol.data(), u, sampling, 1.0 / g.tv, map, best_gain);
}

// ===================================================================== ref ==
namespace {
// ParallelLocalMovingModularity::localMoving (local_moving_modularity.cpp:61-130)
bool local_moving_ref(const Graph& g, std::vector<int32_t>& C, const LouvainParams& prm,
                      std::mt19937& shuffler, LouvainStats* st) {
  const double recip = 1.0 / g.tv;
  std::vector<uint32_t> nodes(g.n);
  std::vector<double> cvol(g.n);
  for (uint32_t u = 0; u < g.n; ++u) { nodes[u] = u; C[u] = int32_t(u); cvol[u] = g.vol[u]; }
  bool changed = false;
  if (g.idx[g.n] == 0) return false;
  const int T = std::max(1, prm.threads);
  size_t moved = g.n;
  for (int r = 0; double(moved) >= prm.min_move_fraction * g.n && r < prm.max_passes; ++r) {
    if (!prm.disable_randomization) std::shuffle(nodes.begin(), nodes.end(), shuffler);
    moved = 0;
    auto move_node = [&](uint32_t u, InsertionMap& map, size_t& mv) {
      const double volU = g.vol[u];
      const int32_t from = __atomic_load_n(&C[u], __ATOMIC_RELAXED);
      const int32_t best = best_cluster_with(g, C.data(), cvol.data(), u, prm.sampling, recip, map, nullptr);
      if (best != from) {
        // double CAS (atomic_wrapper.h:109-120)
        auto add = [&](double* x, double d) {
          double old, nv;
          __atomic_load(x, &old, __ATOMIC_RELAXED);
          do { nv = old + d; } while (!__atomic_compare_exchange(x, &old, &nv, true,
                                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED));
        };
        add(&cvol[best], volU);
        add(&cvol[from], -volU);
        __atomic_store_n(&C[u], best, __ATOMIC_RELAXED);
        ++mv;
      }
    };
    if (T == 1) {
      InsertionMap map(g.n);
      for (uint32_t i = 0; i < g.n; ++i) move_node(nodes[i], map, moved);
    } else {
      size_t total = 0;
#pragma omp parallel num_threads(T) reduction(+ : total)
      {
        InsertionMap map(g.n);
        size_t mv = 0;
#pragma omp for schedule(dynamic, 512)
        for (int64_t i = 0; i < int64_t(g.n); ++i) move_node(nodes[i], map, mv);
        total += mv;
      }
      moved = total;
    }
    if (st) { st->arcs_swept += int64_t(g.idx[g.n]); st->rounds++; st->moves += int64_t(moved); }
    changed |= moved > 0;
  }
  return changed;
}

// local_moving_contract_recurse (parallel_louvain.cpp:27-60)
std::vector<int32_t> recurse_ref(Graph& g, const LouvainParams& prm, std::mt19937& sh,
                                 LouvainStats* st) {
  std::vector<int32_t> C(g.n);
  const bool changed = local_moving_ref(g, C, prm, sh, st);
  if (st) st->levels++;
  if (changed) {
    Graph cg = contract_graph(g, C);
    std::vector<int32_t> cc = recurse_ref(cg, prm, sh, st);
    for (uint32_t u = 0; u < g.n; ++u) C[u] = cc[C[u]];
  }
  return C;
}
}  // namespace

std::vector<int32_t> louvain_ref(Graph& g, const LouvainParams& prm, LouvainStats* st) {
  std::mt19937 sh(uint32_t(prm.seed));
  return recurse_ref(g, prm, sh, st);
}

// ==================================================================== sync ==
// Synchronous local moving: round r is split into sub-rounds by a hash of
// (seed, level, r, tile of u), a tile being 64 consecutive node ids; within a
// sub-round every node of the bucket computes
// its best cluster (same gain, same first-occurrence tie rule) on the state at
// the start of the sub-round, then all moves are applied.  A singleton cluster
// moves into another singleton cluster only towards the smaller id, which
// removes the two-node swap of a fully synchronous update.
namespace {
constexpr uint64_t kLvBucketTag = 0x4C56420000000000ull;

bool local_moving_sync_level(const Graph& g, std::vector<int32_t>& C, const LouvainParams& prm,
                             int level, LouvainStats* st, LouvainTrace* tr = nullptr) {
  const double recip = 1.0 / g.tv;
  std::vector<double> cvol(g.n);
  std::vector<int32_t> csize(g.n, 1);
  for (uint32_t u = 0; u < g.n; ++u) { C[u] = int32_t(u); cvol[u] = g.vol[u]; }
  if (g.idx[g.n] == 0) return false;
  const int B = std::max(1, prm.sub_rounds);
  InsertionMap map(g.n);
  bool changed = false;
  size_t moved = g.n;
  std::vector<std::pair<uint32_t, int32_t>> moves;
  for (int r = 0; double(moved) >= prm.min_move_fraction * g.n && r < prm.max_passes; ++r) {
    moved = 0;
    const uint32_t salt = bucket_salt(prm.seed, kLvBucketTag | (uint64_t(level) << 8) | uint64_t(r));
    for (int s = 0; s < B; ++s) {
      moves.clear();
      for (uint32_t u = 0; u < g.n; ++u) {
        if (B > 1 && bucket_of(salt, u >> kLvTileShift, uint32_t(B)) != uint32_t(s)) continue;
        if (tr) tr->members.push_back(u);
        const int32_t from = C[u];
        const int32_t best = best_cluster_with(g, C.data(), cvol.data(), u, prm.sampling, recip, map, nullptr);
        if (best == from) continue;
        if (csize[from] == 1 && csize[best] == 1 && best > from) continue;
        moves.push_back({u, best});
      }
      if (tr) {
        tr->member_off.push_back(tr->members.size());
        for (const auto& mv : moves) { tr->mv_u.push_back(mv.first); tr->mv_to.push_back(mv.second); }
        tr->move_off.push_back(tr->mv_u.size());
      }
      for (const auto& mv : moves) {
        const uint32_t u = mv.first;
        const int32_t from = C[u], to = mv.second;
        C[u] = to;
        cvol[to] += g.vol[u];
        cvol[from] -= g.vol[u];
        csize[to]++;
        csize[from]--;
      }
      moved += moves.size();
    }
    if (st) { st->arcs_swept += int64_t(g.idx[g.n]); st->rounds++; st->moves += int64_t(moved); }
    changed |= moved > 0;
  }
  return changed;
}

std::vector<int32_t> recurse_sync(Graph& g, const LouvainParams& prm, int level, LouvainStats* st) {
  std::vector<int32_t> C(g.n);
  const bool changed = local_moving_sync_level(g, C, prm, level, st);
  if (st) st->levels++;
  if (changed) {
    Graph cg = contract_graph(g, C);
    std::vector<int32_t> cc = recurse_sync(cg, prm, level + 1, st);
    for (uint32_t u = 0; u < g.n; ++u) C[u] = cc[C[u]];
  }
  return C;
}
}  // namespace

bool louvain_local_moving_sync(const Graph& g, std::vector<int32_t>& C, const LouvainParams& prm,
                               LouvainStats* st, LouvainTrace* tr) {
  C.assign(g.n, 0);
  if (tr) { tr->member_off.assign(1, 0); tr->move_off.assign(1, 0); }
  return local_moving_sync_level(g, C, prm, 0, st, tr);
}

std::vector<int32_t> louvain_sync(Graph& g, const LouvainParams& prm, LouvainStats* st) {
  return recurse_sync(g, prm, 0, st);
}

}  // namespace orc
