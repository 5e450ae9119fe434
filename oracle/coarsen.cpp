// coarsen.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Heavy-edge rating, clustering (reference CAS state machine and the
// synchronous sub-round form) and hypergraph contraction.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <random>
#include <unordered_map>

namespace orc {

namespace {
uint64_t next_pow2(uint64_t x) {  // align_to_next_power_of_two (multilevel_vertex_pair_rater.h:324-326)
  return uint64_t(std::pow(2.0, std::ceil(std::log2(double(x)))));
}
}  // namespace

// ===================================================================== ref ==
// MultilevelCoarsener pass (multilevel_coarsener.h:158-245) with
// MultilevelVertexPairRater::rate (multilevel_vertex_pair_rater.h:109-279),
// HeavyEdgeScore, NoWeightPenalty and BestRatingPreferringUnmatched /
// BestRatingWithoutTieBreaking (policies/rating_*.h), matchVertices
// (multilevel_coarsener.h:333-438).  Single-threaded execution follows the
// vertex order; the CAS protocol is kept so that threads>1 also runs.
std::vector<uint32_t> cluster_pass_ref(const Hypergraph& hg, const CoarsenParams& prm, int pass,
                                       uint32_t* nodes_after) {
  const uint32_t n = hg.n;
  enum : uint8_t { UNMATCHED = 0, IN_PROGRESS = 1, MATCHED = 2 };
  std::vector<uint32_t> order(n), cid(n), partner(n);
  std::vector<std::atomic<uint8_t>> state(n);
  std::vector<std::atomic<int32_t>> cw(n);
  std::vector<uint8_t> already(n, 0);
  for (uint32_t u = 0; u < n; ++u) {
    order[u] = u; cid[u] = u; partner[u] = u;
    state[u].store(UNMATCHED); cw[u].store(hg.vw[u]);
  }
  if (!prm.disable_randomization) {
    std::mt19937 sh(uint32_t(prm.seed + 7919u * uint32_t(pass)));
    std::shuffle(order.begin(), order.end(), sh);
  }
  const uint32_t limit = std::max(uint32_t(double(n) / prm.max_shrink), prm.contraction_limit);
  const uint64_t mask = next_pow2(std::min<uint64_t>(10ull * hg.max_esize, n)) - 1;
  std::vector<uint8_t> bloom(mask + 1, 0);
  std::mt19937 coin_gen(uint32_t(prm.seed));
  std::vector<uint8_t> coins(128);
  { std::uniform_int_distribution<int> bd(0, 1); for (auto& c : coins) c = uint8_t(bd(coin_gen)); }
  size_t coin_pos = 0;
  auto coin = [&]() { return coins[++coin_pos % 128] != 0; };

  uint32_t current = n;
  uint32_t contracted = 0, threshold = 0;
  std::vector<double> rmap_val(n, 0.0);
  std::vector<int32_t> rmap_pos(n, -1);
  std::vector<uint32_t> rmap_keys;

  auto adaptive = [&](uint32_t e) {
    uint32_t s = 0;
    for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j) {
      const uint64_t b = cid[hg.pins[j]] & mask;
      if (!bloom[b]) { ++s; bloom[b] = 1; }
    }
    for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j) bloom[cid[hg.pins[j]] & mask] = 0;
    return s;
  };
  auto add_rating = [&](uint32_t rep, double sc) {
    if (rmap_pos[rep] < 0) { rmap_pos[rep] = int32_t(rmap_keys.size()); rmap_keys.push_back(rep); rmap_val[rep] = 0.0; }
    rmap_val[rep] += sc;
  };

  auto match = [&](uint32_t u, uint32_t v) {
    const int32_t wu = hg.vw[u];
    int32_t wv = cw[v].load();
    if (int64_t(wu) + wv > prm.max_node_weight) return;
    uint8_t un = UNMATCHED;
    if (!state[u].compare_exchange_strong(un, IN_PROGRESS)) return;
    partner[u] = v;
    const uint8_t sv = state[v].load();
    if (sv == MATCHED) {
      if (v == cid[v]) { cid[u] = v; cw[v] += wu; ++contracted; }
      else {
        const uint32_t cv = cid[v];
        if (int64_t(wu) + cw[cv].load() <= prm.max_node_weight) { cid[u] = cv; cw[cv] += wu; ++contracted; }
      }
    } else {
      uint8_t un2 = UNMATCHED;
      if (state[v].compare_exchange_strong(un2, IN_PROGRESS)) {
        cid[u] = v; cw[v] += wu; ++contracted; state[v].store(MATCHED);
      } else {
        while (state[v].load() == IN_PROGRESS) {
          uint32_t cur = u, smallest = u;
          while (partner[cur] != u && partner[cur] != cur) { cur = partner[cur]; smallest = std::min(smallest, cur); }
          if (partner[cur] == u && u == smallest) {
            cid[u] = v; cw[v] += wu; ++contracted; state[v].store(MATCHED); state[u].store(MATCHED);
          }
        }
        if (state[u].load() == IN_PROGRESS) {
          const uint32_t cv = cid[v];
          if (int64_t(wu) + cw[cv].load() <= prm.max_node_weight) { cid[u] = cv; cw[cv] += wu; ++contracted; }
        }
      }
    }
    already[u] = 1; already[v] = 1;
    partner[u] = u;
    state[u].store(MATCHED);
  };

  const uint64_t vdb_size = 3ull * prm.sampling;
  const uint64_t small_map = std::min<uint64_t>(vdb_size, 32768);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t u = order[i];
    if (state[u].load() != UNMATCHED || current <= limit) continue;
    // getRatingMapTypeForRatingOfHypernode (multilevel_vertex_pair_rater.h:281-318)
    bool sampling = false;
    if (n >= small_map) {
      uint64_t ub = 0;
      for (uint64_t j = hg.voff[u]; j < hg.voff[u + 1]; ++j) {
        const uint32_t sz = hg.esize(hg.inc[j]);
        ub += sz < prm.ignore_he_size ? sz : 0;
        if (ub > 32768 / 3) { sampling = vdb_size < n; break; }
      }
    }
    uint64_t accesses = 0;
    for (uint64_t j = hg.voff[u]; j < hg.voff[u + 1]; ++j) {
      const uint32_t e = hg.inc[j];
      const uint32_t esz = hg.esize(e);
      if (esz >= prm.ignore_he_size) continue;
      const uint32_t sz = std::max<uint32_t>(adaptive(e), 2);
      if (sampling && accesses + sz > prm.sampling) break;
      const double score = double(hg.nw[e]) / double(sz - 1);
      for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) {
        const uint32_t rep = cid[hg.pins[q]];
        const uint64_t b = rep & mask;
        if (!bloom[b]) { add_rating(rep, score); bloom[b] = 1; ++accesses; }
      }
      for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) bloom[cid[hg.pins[q]] & mask] = 0;
    }
    const int32_t wu = cw[u].load();
    double max_rating = std::numeric_limits<double>::min();
    uint32_t target = UINT32_MAX;
    for (size_t q = rmap_keys.size(); q-- > 0;) {
      const uint32_t t = rmap_keys[q];
      const int32_t wt = cw[t].load();
      if (t != u && int64_t(wu) + wt <= prm.max_node_weight) {
        const double rating = rmap_val[t] / 1.0;
        bool acc;
        if (prm.prefer_unmatched) {
          acc = max_rating < rating ||
                (max_rating == rating &&
                 ((already[target] && !already[t]) ||
                  (already[target] && already[t] && !prm.disable_randomization && coin()) ||
                  (!already[target] && !already[t] && !prm.disable_randomization && coin())));
        } else {
          acc = max_rating < rating || (max_rating == rating && target < t);
        }
        if (hg.comm[u] == hg.comm[t] && acc) { max_rating = rating; target = t; }
      }
    }
    for (uint32_t t : rmap_keys) rmap_pos[t] = -1;
    rmap_keys.clear();
    if (target != UINT32_MAX) {
      match(u, target);
      if (contracted >= threshold) {
        current = n - contracted;
        threshold += (current > limit ? current - limit : 0) / std::max(1, prm.threads);
      }
    }
  }
  *nodes_after = n - contracted;
  return cid;
}

// ==================================================================== sync ==
// Synchronous clustering pass.  Vertices are split into sub-rounds by a hash
// of (seed, pass, u).  In a sub-round every unmatched vertex of the bucket
// rates the clusters around it on the state at the start of the sub-round:
// exact distinct cluster count per net (adaptive edge size), heavy-edge score
// w(e)/(size'-1) in 36.28 fixed point so sums are associative, acceptance by
// (rating, target not yet matched, hash).  A vertex whose target does not move
// in the sub-round joins the target's cluster (joiners admitted per target in
// (rating desc, hash) order up to the node weight limit); two vertices that
// chose each other pair up under the smaller id; all other proposals wait.
namespace {
constexpr uint64_t kCoBucketTag = 0x434F420000000000ull;
constexpr uint64_t kCoTieTag = 0x434F540000000000ull;
constexpr uint64_t kCoPrioTag = 0x434F500000000000ull;
constexpr int kFixShift = 28;

// Rating of one proposer (multilevel_vertex_pair_rater.h:109-199 with exact
// distinct counts and fixed-point scores) and its acceptance
// (rating_acceptance_policy.h:56-72 with the hash as the coin).
struct SyncRater {
  const Hypergraph& hg;
  const CoarsenParams& prm;
  std::vector<uint64_t> rval;
  std::vector<int32_t> rpos;
  std::vector<uint32_t> rkeys, distinct;
  std::vector<uint8_t> seen;
  int64_t visits = 0;
  SyncRater(const Hypergraph& h, const CoarsenParams& p) : hg(h), prm(p), rval(h.n, 0), rpos(h.n, -1), seen(h.n, 0) {}
  // returns the target (UINT32_MAX: none) and sets its rating
  uint32_t rate(uint32_t u, uint64_t tag, const std::vector<uint32_t>& cid, const std::vector<int64_t>& cw,
                const std::vector<uint8_t>& matched, uint64_t& rating) {
    uint64_t accesses = 0;
    for (uint64_t j = hg.voff[u]; j < hg.voff[u + 1]; ++j) {
      const uint32_t e = hg.inc[j];
      const uint32_t esz = hg.esize(e);
      if (esz >= prm.ignore_he_size) continue;
      distinct.clear();
      for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) {
        const uint32_t rep = cid[hg.pins[q]];
        if (!seen[rep]) { seen[rep] = 1; distinct.push_back(rep); }
      }
      for (uint32_t rep : distinct) seen[rep] = 0;
      visits += 2 * int64_t(esz);
      const uint32_t sz = std::max<uint32_t>(uint32_t(distinct.size()), 2);
      if (accesses + sz > prm.sampling) break;
      const uint64_t score = (uint64_t(uint32_t(hg.nw[e])) << kFixShift) / uint64_t(sz - 1);
      for (uint32_t rep : distinct) {
        if (rpos[rep] < 0) { rpos[rep] = int32_t(rkeys.size()); rkeys.push_back(rep); rval[rep] = 0; }
        rval[rep] += score;
        ++accesses;
      }
    }
    uint64_t best_val = 0, best_h = 0;
    int best_un = -1;
    uint32_t best = UINT32_MAX;
    for (uint32_t t : rkeys) {
      if (t == u || cw[u] + cw[t] > prm.max_node_weight || hg.comm[t] != hg.comm[u]) continue;
      const uint64_t v = rval[t];
      const int un = prm.prefer_unmatched ? (matched[t] ? 0 : 1) : 0;
      const uint64_t h = hash4(prm.seed, kCoTieTag | tag, u, t);
      if (best == UINT32_MAX || v > best_val || (v == best_val && (un > best_un || (un == best_un && h > best_h)))) {
        best = t; best_val = v; best_un = un; best_h = h;
      }
    }
    for (uint32_t t : rkeys) rpos[t] = -1;
    rkeys.clear();
    rating = best_val;
    return best;
  }
};
}  // namespace

void rate_sync(const Hypergraph& hg, const CoarsenParams& prm, int pass, int sub_round,
               const std::vector<uint32_t>& cid, const std::vector<int64_t>& cw,
               const std::vector<uint8_t>& matched, const std::vector<uint32_t>& proposers,
               std::vector<uint32_t>& target, std::vector<uint64_t>& rating) {
  SyncRater r(hg, prm);
  const uint64_t tag = uint64_t(pass) << 8 | uint64_t(sub_round);
  target.assign(proposers.size(), UINT32_MAX);
  rating.assign(proposers.size(), 0);
  for (size_t i = 0; i < proposers.size(); ++i) target[i] = r.rate(proposers[i], tag, cid, cw, matched, rating[i]);
}

std::vector<uint32_t> cluster_pass_sync(const Hypergraph& hg, const CoarsenParams& prm, int pass,
                                        uint32_t* nodes_after, int64_t* pin_visits) {
  const uint32_t n = hg.n;
  const int S = std::max(1, prm.sub_rounds);
  std::vector<uint32_t> cid(n), target(n, UINT32_MAX);
  std::vector<int64_t> cw(n);
  std::vector<uint8_t> matched(n, 0);
  std::vector<uint64_t> rating(n, 0);
  for (uint32_t u = 0; u < n; ++u) { cid[u] = u; cw[u] = hg.vw[u]; }
  const uint32_t limit = std::max(uint32_t(double(n) / prm.max_shrink), prm.contraction_limit);
  uint32_t current = n;
  SyncRater rater(hg, prm);
  for (int s = 0; s < S; ++s) {
    if (current <= limit) break;
    const uint64_t tag = uint64_t(pass) << 8 | uint64_t(s);
    std::vector<uint32_t> proposers;
    const uint32_t salt = bucket_salt(prm.seed, kCoBucketTag | uint64_t(pass));
    for (uint32_t u = 0; u < n; ++u)
      if (!matched[u] && (S == 1 || bucket_of(salt, u, uint32_t(S)) == uint32_t(s))) proposers.push_back(u);
    for (uint32_t u : proposers) target[u] = rater.rate(u, tag, cid, cw, matched, rating[u]);
    // resolution
    std::vector<uint8_t> moving(n, 0);
    for (uint32_t u : proposers) if (target[u] != UINT32_MAX) moving[u] = 1;
    std::map<uint32_t, std::vector<uint32_t>> joiners;
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t u : proposers) {
      const uint32_t t = target[u];
      if (t == UINT32_MAX) continue;
      if (!moving[t]) joiners[t].push_back(u);
      else if (target[t] == u && u < t) pairs.push_back({u, t});
    }
    for (auto& pr : pairs) {
      const uint32_t a = pr.first, b = pr.second;  // a < b, a is the representative
      cid[b] = a; cw[a] += hg.vw[b]; matched[a] = matched[b] = 1; --current;
    }
    for (auto& kv : joiners) {
      const uint32_t t = kv.first;
      std::vector<uint32_t>& J = kv.second;
      int64_t sum = 0;
      for (uint32_t u : J) sum += hg.vw[u];
      if (cw[t] + sum > prm.max_node_weight) {
        std::sort(J.begin(), J.end(), [&](uint32_t x, uint32_t y) {
          if (rating[x] != rating[y]) return rating[x] > rating[y];
          const uint64_t hx = hash4(prm.seed, kCoPrioTag | tag, x, 0), hy = hash4(prm.seed, kCoPrioTag | tag, y, 0);
          if (hx != hy) return hx < hy;
          return x < y;
        });
      }
      int64_t cum = cw[t];
      bool full = false;
      for (uint32_t u : J) {
        if (!full && cum + hg.vw[u] <= prm.max_node_weight) {
          cum += hg.vw[u];
          cid[u] = t; matched[u] = 1; matched[t] = 1; --current;
        } else {
          full = true;
        }
      }
      cw[t] = cum;
    }
  }
  if (pin_visits) *pin_visits += rater.visits;
  *nodes_after = current;
  return cid;
}

// StaticHypergraph::contract (static_hypergraph.cpp:58-513) in canonical form.
Hypergraph contract_hypergraph(const Hypergraph& hg, const std::vector<uint32_t>& cluster,
                               std::vector<uint32_t>& cmap) {
  // stage 1: coarse ids are the ranks of the used cluster ids (:95-121)
  std::vector<uint32_t> mapping(hg.n, 0);
  for (uint32_t u = 0; u < hg.n; ++u) mapping[cluster[u]] = 1;
  uint32_t run = 0;
  for (uint32_t u = 0; u < hg.n; ++u) { const uint32_t x = mapping[u]; mapping[u] = run; run += x; }
  cmap.resize(hg.n);
  for (uint32_t u = 0; u < hg.n; ++u) cmap[u] = mapping[cluster[u]];
  Hypergraph c;
  c.n = run;
  c.vw.assign(c.n, 0);
  c.comm.assign(c.n, 0);
  for (uint32_t u = 0; u < hg.n; ++u) { c.vw[cmap[u]] += hg.vw[u]; c.comm[cmap[u]] = hg.comm[u]; }
  // stage 2: map, sort, unique pins; single-pin nets vanish (:156-207)
  std::vector<std::vector<uint32_t>> npins(hg.m);
  std::vector<uint8_t> valid(hg.m, 0);
  std::vector<int64_t> weight(hg.nw.begin(), hg.nw.end());
  for (uint32_t e = 0; e < hg.m; ++e) {
    auto& v = npins[e];
    for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j) v.push_back(cmap[hg.pins[j]]);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    valid[e] = v.size() > 1;
  }
  // stage 3: parallel nets merge into the smallest original id (:315-373)
  std::map<std::vector<uint32_t>, uint32_t> first;
  for (uint32_t e = 0; e < hg.m; ++e) {
    if (!valid[e]) continue;
    auto it = first.find(npins[e]);
    if (it == first.end()) first.emplace(npins[e], e);
    else { weight[it->second] += weight[e]; valid[e] = 0; }
  }
  // stage 4: compaction (:375-507)
  c.noff.assign(1, 0);
  for (uint32_t e = 0; e < hg.m; ++e) {
    if (!valid[e]) continue;
    c.pins.insert(c.pins.end(), npins[e].begin(), npins[e].end());
    c.noff.push_back(c.pins.size());
    c.nw.push_back(int32_t(weight[e]));
  }
  c.m = uint32_t(c.nw.size());
  build_incidence(c);
  return c;
}

// MultilevelCoarsener::coarsenImpl driver (multilevel_coarsener.h:143-315)
std::vector<Level> coarsen(const Hypergraph& hg, const CoarsenParams& prm, bool sync,
                           int64_t* pin_visits) {
  std::vector<Level> levels;
  int pass = 0;
  const Hypergraph* cur = &hg;
  while (cur->n > prm.contraction_limit) {
    uint32_t after = cur->n;
    std::vector<uint32_t> cl = sync ? cluster_pass_sync(*cur, prm, pass, &after, pin_visits)
                                    : cluster_pass_ref(*cur, prm, pass, &after);
    if (after == 0 || double(cur->n) / double(after) <= prm.min_shrink) break;
    Level lv;
    lv.hg = contract_hypergraph(*cur, cl, lv.cmap);
    levels.push_back(std::move(lv));
    cur = &levels.back().hg;
    ++pass;
  }
  return levels;
}

}  // namespace orc

namespace orc {
// Rating throughput sample of the reference-semantics rater: fillRatingMap
// (bloom-filter adaptive edge size, heavy-edge score, multilevel_vertex_pair_
// rater.h:109-279) and the acceptance scan for the given vertices on the
// level-0 state (every vertex its own cluster), on `threads` OpenMP threads.
// No matching: the vertices are independent.  Returns the pin visits (two
// per pin of a rated net, as the rating sweeps do).
int64_t rating_sample_ref(const Hypergraph& hg, const std::vector<uint32_t>& vertices, uint32_t ignore,
                          int threads) {
  const uint32_t n = hg.n;
  const uint64_t mask = next_pow2(std::min<uint64_t>(10ull * hg.max_esize, n)) - 1;
  int64_t visits = 0;
  int64_t sink = 0;
#pragma omp parallel num_threads(std::max(1, threads)) reduction(+ : visits, sink)
  {
    std::vector<uint8_t> bloom(mask + 1, 0);
    std::vector<double> val(n, 0.0);
    std::vector<int32_t> pos(n, -1);
    std::vector<uint32_t> keys;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(vertices.size()); ++i) {
      const uint32_t u = vertices[size_t(i)];
      for (uint64_t j = hg.voff[u]; j < hg.voff[u + 1]; ++j) {
        const uint32_t e = hg.inc[j];
        const uint32_t esz = hg.esize(e);
        if (esz >= ignore) continue;
        uint32_t s = 0;
        for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) {
          const uint64_t b = hg.pins[q] & mask;
          if (!bloom[b]) { ++s; bloom[b] = 1; }
        }
        for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) bloom[hg.pins[q] & mask] = 0;
        const double score = double(hg.nw[e]) / double(std::max<uint32_t>(s, 2) - 1);
        for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) {
          const uint32_t rep = hg.pins[q];
          const uint64_t b = rep & mask;
          if (!bloom[b]) {
            if (pos[rep] < 0) { pos[rep] = int32_t(keys.size()); keys.push_back(rep); val[rep] = 0.0; }
            val[rep] += score;
            bloom[b] = 1;
          }
        }
        for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) bloom[hg.pins[q] & mask] = 0;
        visits += 2 * int64_t(esz);
      }
      double best = 0.0;
      uint32_t target = UINT32_MAX;
      for (size_t q = keys.size(); q-- > 0;) {
        const uint32_t t = keys[q];
        if (t != u && val[t] > best) { best = val[t]; target = t; }
        pos[t] = -1;
      }
      keys.clear();
      sink += target != UINT32_MAX;
    }
  }
  return visits + 0 * sink;
}
}  // namespace orc
