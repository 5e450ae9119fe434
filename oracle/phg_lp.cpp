// phg_lp.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Hypergraph construction, partition state, km1 and label propagation.
#include "oracle.h"

#include <algorithm>
#include <numeric>
#include <random>
#include <omp.h>

namespace orc {

// static_hypergraph_factory.cpp:32-166 (construction; incident nets in
// ascending net order, i.e. the stable construction of :124-156).
Hypergraph build_hypergraph(uint32_t n, uint32_t m, const uint64_t* noff, const uint32_t* pins,
                            const int32_t* vw, const int32_t* nw) {
  Hypergraph hg;
  hg.n = n;
  hg.m = m;
  hg.noff.assign(noff, noff + m + 1);
  hg.pins.assign(pins, pins + noff[m]);
  hg.vw.assign(n, 1);
  hg.nw.assign(m, 1);
  if (vw) std::copy(vw, vw + n, hg.vw.begin());
  if (nw) std::copy(nw, nw + m, hg.nw.begin());
  hg.comm.assign(n, 0);
  build_incidence(hg);
  return hg;
}

void build_incidence(Hypergraph& hg) {
  hg.voff.assign(size_t(hg.n) + 1, 0);
  hg.max_esize = 0;
  for (uint32_t e = 0; e < hg.m; ++e) {
    hg.max_esize = std::max(hg.max_esize, hg.esize(e));
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) hg.voff[hg.pins[i] + 1]++;
  }
  for (uint32_t u = 0; u < hg.n; ++u) hg.voff[u + 1] += hg.voff[u];
  hg.inc.assign(hg.pins.size(), 0);
  std::vector<uint64_t> pos(hg.voff.begin(), hg.voff.end() - 1);
  for (uint32_t e = 0; e < hg.m; ++e)
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) hg.inc[pos[hg.pins[i]]++] = e;
  hg.total_w = 0;
  for (uint32_t u = 0; u < hg.n; ++u) hg.total_w += hg.vw[u];
}

int Partition::lambda(uint32_t e) const {
  int l = 0;
  for (int i = 0; i < cw; ++i) l += __builtin_popcountll(conn[size_t(e) * cw + i]);
  return l;
}

// initializePartition = initializeBlockWeights + initializePinCountInPart
// (partitioned_hypergraph.h:693-698, 1067-1110).
void phg_init(Partition& p, const Hypergraph& hg, int k, const int32_t* part) {
  p.hg = &hg;
  p.k = k;
  p.cw = (k + 63) / 64;
  p.part.assign(part, part + hg.n);
  p.part_w.assign(k, 0);
  for (uint32_t u = 0; u < hg.n; ++u) p.part_w[p.part[u]] += hg.vw[u];
  p.pc.assign(size_t(hg.m) * k, 0);
  p.conn.assign(size_t(hg.m) * p.cw, 0);
  for (uint32_t e = 0; e < hg.m; ++e) {
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) p.pc[size_t(e) * k + p.part[hg.pins[i]]]++;
    for (int b = 0; b < k; ++b)
      if (p.pc[size_t(e) * k + b] > 0) p.conn[size_t(e) * p.cw + b / 64] |= 1ull << (b % 64);
  }
}

// metrics::km1 (metrics.cpp:49-65)
int64_t km1(const Partition& p) {
  int64_t s = 0;
  for (uint32_t e = 0; e < p.hg->m; ++e) s += int64_t(std::max(p.lambda(e) - 1, 0)) * p.hg->nw[e];
  return s;
}
// metrics::hyperedgeCut (metrics.cpp:29-47)
int64_t cut(const Partition& p) {
  int64_t s = 0;
  for (uint32_t e = 0; e < p.hg->m; ++e) if (p.lambda(e) > 1) s += p.hg->nw[e];
  return s;
}

// isBorderNode (partitioned_hypergraph.h:605-620); HIGH_DEGREE_THRESHOLD :67
bool is_border(const Partition& p, uint32_t u) {
  const Hypergraph& hg = *p.hg;
  if (hg.degree(u) > 100000) return false;
  for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i)
    if (p.lambda(hg.inc[i]) > 1) return true;
  return false;
}

// Km1Policy::computeMaxGainMoveImpl (gain_policy.h:108-167) without coins.
void km1_max_gain_move(const Partition& p, uint32_t u, const std::vector<int64_t>& max_w,
                       int32_t* to_out, int32_t* gain_out) {
  const Hypergraph& hg = *p.hg;
  const int from = p.part[u];
  std::vector<int64_t> score(p.k, 0);
  int64_t internal = 0;
  for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
    const uint32_t e = hg.inc[i];
    if (p.pcount(e, from) > 1) internal += hg.nw[e];
    for (int b = 0; b < p.k; ++b)
      if (b != from && p.has(e, b)) score[b] -= hg.nw[e];
  }
  int best_to = from;
  int64_t best_gain = 0;
  for (int b = 0; b < p.k; ++b) {
    if (b == from) continue;
    const int64_t s = score[b] + internal;
    if (s < best_gain && p.part_w[b] + hg.vw[u] <= max_w[b]) { best_to = b; best_gain = s; }
  }
  *to_out = best_to;
  *gain_out = int32_t(best_gain);
}

// ===================================================================== ref ==
// LabelPropagationRefiner (label_propagation_refiner.h:75-143,
// label_propagation_refiner.cpp:33-164) with changeNodePart
// (partitioned_hypergraph.h:539-563) and km1Delta (hypergraph_common.h:133-140).
namespace {

struct RefLp {
  Partition& p;
  const LpParams& prm;
  const Hypergraph& hg;
  std::vector<std::atomic<uint8_t>> lock;
  std::vector<std::atomic<uint8_t>> visited_he, next_flag;
  std::vector<int64_t> tl_delta;          // per thread (padded)
  std::vector<std::mt19937> gen;
  std::vector<std::vector<uint8_t>> coins;
  std::vector<size_t> coin_pos;
  std::vector<std::vector<uint32_t>> tl_next;
  int64_t moves = 0, reverts = 0, visits = 0;

  RefLp(Partition& p_, const LpParams& prm_)
      : p(p_), prm(prm_), hg(*p_.hg), lock(hg.m), visited_he(hg.m), next_flag(hg.n) {
    const int T = std::max(1, prm.threads);
    tl_delta.assign(size_t(T) * 8, 0);
    tl_next.resize(T);
    // Randomize: per-CPU mt19937(seed+cpu) and 128 precomputed coins
    // (randomize.h:39,62-64,85-90,112-121)
    for (int t = 0; t < T; ++t) {
      gen.emplace_back(uint32_t(prm.seed + t));
      std::uniform_int_distribution<int> bd(0, 1);
      std::vector<uint8_t> c(128);
      for (auto& x : c) x = uint8_t(bd(gen.back()));
      coins.push_back(c);
      coin_pos.push_back(0);
    }
    for (auto& l : lock) l.store(0);
  }
  bool coin(int t) { return coins[t][++coin_pos[t] % 128]; }

  int32_t part_of(uint32_t u) const { return __atomic_load_n(&p.part[u], __ATOMIC_RELAXED); }
  uint32_t pcount(uint32_t e, int b) const {
    return __atomic_load_n(&p.pc[size_t(e) * p.k + b], __ATOMIC_RELAXED);
  }
  bool has(uint32_t e, int b) const {
    return (__atomic_load_n(&p.conn[size_t(e) * p.cw + b / 64], __ATOMIC_RELAXED) >> (b % 64)) & 1ull;
  }
  int lambda(uint32_t e) const {
    int l = 0;
    for (int i = 0; i < p.cw; ++i)
      l += __builtin_popcountll(__atomic_load_n(&p.conn[size_t(e) * p.cw + i], __ATOMIC_RELAXED));
    return l;
  }
  int64_t pw(int b) const { return __atomic_load_n(&p.part_w[b], __ATOMIC_RELAXED); }

  bool border(uint32_t u) const {
    if (hg.degree(u) > 100000) return false;
    for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i)
      if (lambda(hg.inc[i]) > 1) return true;
    return false;
  }

  // gain_policy.h:108-167
  void best_move(uint32_t u, int t, std::vector<int64_t>& score, int& to, int64_t& gain) {
    const int from = part_of(u);
    int64_t internal = 0;
    for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
      const uint32_t e = hg.inc[i];
      if (pcount(e, from) > 1) internal += hg.nw[e];
      for (int b = 0; b < p.k; ++b)
        if (b != from && has(e, b)) score[b] -= hg.nw[e];
    }
    to = from;
    gain = 0;
    for (int b = 0; b < p.k; ++b) {
      if (b != from) {
        const int64_t s = score[b] + internal;
        const bool nb = s < gain || (s == gain && !prm.disable_randomization && coin(t));
        if (nb && pw(b) + hg.vw[u] <= prm.max_w[b]) { to = b; gain = s; }
      }
      score[b] = 0;
    }
  }

  // partitioned_hypergraph.h:539-563 and :1132-1167
  bool change(uint32_t u, int from, int to, int t) {
    const int64_t wu = hg.vw[u];
    const int64_t to_after = __atomic_add_fetch(&p.part_w[to], wu, __ATOMIC_RELAXED);
    const int64_t from_old = __atomic_fetch_sub(&p.part_w[from], wu, __ATOMIC_RELAXED);
    if (to_after <= prm.max_w[to] && from_old > 0) {
      __atomic_store_n(&p.part[u], to, __ATOMIC_RELAXED);
      for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        uint8_t z = 0;
        while (!lock[e].compare_exchange_weak(z, 1, std::memory_order_acquire)) z = 0;
        uint32_t* pf = &p.pc[size_t(e) * p.k + from];
        uint32_t* pt = &p.pc[size_t(e) * p.k + to];
        const uint32_t from_after = __atomic_sub_fetch(pf, 1u, __ATOMIC_RELAXED);
        if (from_after == 0)
          __atomic_fetch_xor(&p.conn[size_t(e) * p.cw + from / 64], 1ull << (from % 64), __ATOMIC_RELAXED);
        const uint32_t to_after_pc = __atomic_add_fetch(pt, 1u, __ATOMIC_RELAXED);
        if (to_after_pc == 1)
          __atomic_fetch_xor(&p.conn[size_t(e) * p.cw + to / 64], 1ull << (to % 64), __ATOMIC_RELAXED);
        lock[e].store(0, std::memory_order_release);
        tl_delta[size_t(t) * 8] += (to_after_pc == 1 ? hg.nw[e] : 0) - (from_after == 0 ? hg.nw[e] : 0);
      }
      return true;
    }
    __atomic_fetch_sub(&p.part_w[to], wu, __ATOMIC_RELAXED);
    __atomic_fetch_add(&p.part_w[from], wu, __ATOMIC_RELAXED);
    return false;
  }

  void activate(uint32_t v, int t) {
    uint8_t z = 0;
    if (next_flag[v].compare_exchange_strong(z, 1)) tl_next[t].push_back(v);
  }

  // label_propagation_refiner.h:75-143
  bool move_vertex(uint32_t u, int t, std::vector<int64_t>& score) {
    if (hg.degree(u) <= 100000) __atomic_fetch_add(&visits, int64_t(hg.degree(u)), __ATOMIC_RELAXED);
    if (!border(u)) return false;
    int to;
    int64_t gain;
    best_move(u, t, score, to, gain);
    const int from = part_of(u);
    const bool positive = gain < 0;
    const bool zero = prm.rebalancing && gain == 0 && pw(from) - 1 > pw(to) + 1 &&
                      pw(to) < prm.perfect_w[to];
    if (to == from || !(positive || zero)) return false;
    const int64_t before = tl_delta[size_t(t) * 8];
    if (!change(u, from, to, t)) return false;
    const int64_t real = tl_delta[size_t(t) * 8] - before;
    if (real == gain || real <= 0) {
      for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        if (hg.esize(e) <= prm.activation_threshold) {
          uint8_t z = 0;
          if (visited_he[e].compare_exchange_strong(z, 1)) {
            for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j) activate(hg.pins[j], t);
          }
        }
      }
      activate(u, t);
      __atomic_fetch_add(&moves, 1, __ATOMIC_RELAXED);
      return true;
    }
    change(u, to, from, t);
    __atomic_fetch_add(&reverts, 1, __ATOMIC_RELAXED);
    return false;
  }
};

}  // namespace

LpStats lp_refine_ref(Partition& p, const LpParams& prm) {
  const Hypergraph& hg = *p.hg;
  RefLp lp(p, prm);
  const int T = std::max(1, prm.threads);
  // initializeActiveNodes (label_propagation_refiner.cpp:193-233)
  std::vector<uint32_t> active;
  for (uint32_t u = 0; u < hg.n; ++u)
    if (prm.rebalancing || is_border(p, u)) active.push_back(u);
  LpStats st;
  std::mt19937 shuffler(uint32_t(prm.seed));
  for (int r = 0; r < prm.max_rounds && !active.empty(); ++r) {
    for (auto& f : lp.visited_he) f.store(0, std::memory_order_relaxed);
    for (auto& f : lp.next_flag) f.store(0, std::memory_order_relaxed);
    if (!prm.disable_randomization) std::shuffle(active.begin(), active.end(), shuffler);
    const int64_t na = int64_t(active.size());
    if (T == 1) {
      std::vector<int64_t> score(p.k, 0);
      for (int64_t j = 0; j < na; ++j) lp.move_vertex(active[j], 0, score);
    } else {
#pragma omp parallel num_threads(T)
      {
        const int t = omp_get_thread_num();
        std::vector<int64_t> score(p.k, 0);
#pragma omp for schedule(dynamic, 256)
        for (int64_t j = 0; j < na; ++j) lp.move_vertex(active[j], t, score);
      }
    }
    active.clear();
    for (int t = 0; t < T; ++t) {
      active.insert(active.end(), lp.tl_next[t].begin(), lp.tl_next[t].end());
      lp.tl_next[t].clear();
    }
    st.rounds++;
  }
  for (int t = 0; t < T; ++t) st.delta += lp.tl_delta[size_t(t) * 8];
  st.moves = lp.moves;
  st.reverts = lp.reverts;
  st.visits = lp.visits;
  return st;
}

// ==================================================================== sync ==
// Synchronous restatement.  Per round r and sub-round s, every active vertex
// of bucket s evaluates Km1Policy on the state at the start of the sub-round,
// candidate moves are admitted per target block in (gain, hash) order up to
// the block's remaining capacity, applied, and each applied move is kept iff
// its real km1 delta on the post-move state is <= 0 (the optimistic check of
// label_propagation_refiner.h:100-137 evaluated after all moves of the
// sub-round).  A sub-round whose total km1 delta is still positive is undone
// and ends the refinement.
namespace {
constexpr uint64_t kLpBucketTag = 0x4C50420000000000ull;
constexpr uint64_t kLpCoinTag = 0x4C50430000000000ull;
constexpr uint64_t kLpPrioTag = 0x4C50500000000000ull;

struct Cand { uint32_t u; int32_t from, to, gain; uint64_t prio; };

// applies u: from -> to, returns km1 delta from the pin-count transitions
int64_t apply_move(Partition& p, uint32_t u, int from, int to) {
  const Hypergraph& hg = *p.hg;
  int64_t d = 0;
  p.part[u] = to;
  p.part_w[from] -= hg.vw[u];
  p.part_w[to] += hg.vw[u];
  for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
    const uint32_t e = hg.inc[i];
    uint32_t& pf = p.pc[size_t(e) * p.k + from];
    uint32_t& pt = p.pc[size_t(e) * p.k + to];
    if (--pf == 0) { p.conn[size_t(e) * p.cw + from / 64] ^= 1ull << (from % 64); d -= hg.nw[e]; }
    if (++pt == 1) { p.conn[size_t(e) * p.cw + to / 64] ^= 1ull << (to % 64); d += hg.nw[e]; }
  }
  return d;
}
}  // namespace

LpStats lp_refine_sync(Partition& p, const LpParams& prm) {
  const Hypergraph& hg = *p.hg;
  const int k = p.k;
  const int B = std::max(1, prm.sub_rounds);
  LpStats st;
  std::vector<uint32_t> active;
  if (!prm.refinement_nodes.empty()) {
    active = prm.refinement_nodes;
  } else {
    for (uint32_t u = 0; u < hg.n; ++u)
      if (prm.rebalancing || is_border(p, u)) active.push_back(u);
  }
  std::vector<uint8_t> next_flag(hg.n, 0);
  std::vector<int64_t> score(k, 0);
  bool stop = false;
  for (int r = 0; r < prm.max_rounds && !active.empty() && !stop; ++r) {
    std::fill(next_flag.begin(), next_flag.end(), 0);
    for (int s = 0; s < B && !stop; ++s) {
      std::vector<Cand> cands;
      const uint64_t rs = (uint64_t(r) << 8) | uint64_t(s);
      for (uint32_t u : active) {
        if (B > 1 && bucket_of(bucket_salt(prm.seed, kLpBucketTag | uint64_t(r)), u, uint32_t(B)) != uint32_t(s))
          continue;
        if (hg.degree(u) > 100000) continue;
        st.visits += hg.degree(u);  // the gain sweep reads every incidence once
        if (!is_border(p, u)) continue;
        const int from = p.part[u];
        int64_t internal = 0;
        for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
          const uint32_t e = hg.inc[i];
          if (p.pcount(e, from) > 1) internal += hg.nw[e];
          for (int b = 0; b < k; ++b)
            if (b != from && p.has(e, b)) score[b] -= hg.nw[e];
        }
        int to = from;
        int64_t gain = 0;
        for (int b = 0; b < k; ++b) {
          if (b != from) {
            const int64_t sc = score[b] + internal;
            const bool nb = sc < gain || (sc == gain && (hash4(prm.seed, kLpCoinTag | rs, u, b) & 1ull));
            if (nb && p.part_w[b] + hg.vw[u] <= prm.max_w[b]) { to = b; gain = sc; }
          }
          score[b] = 0;
        }
        const bool positive = gain < 0;
        const bool zero = prm.rebalancing && gain == 0 && p.part_w[from] - 1 > p.part_w[to] + 1 &&
                          p.part_w[to] < prm.perfect_w[to];
        if (to != from && (positive || zero))
          cands.push_back({u, from, to, int32_t(gain), hash4(prm.seed, kLpPrioTag | rs, u, 0)});
      }
      if (cands.empty()) continue;
      // admission per target block
      std::vector<int64_t> in_w(k, 0);
      for (const Cand& c : cands) in_w[c.to] += hg.vw[c.u];
      std::vector<uint8_t> admit(cands.size(), 1);
      for (int b = 0; b < k; ++b) {
        if (p.part_w[b] + in_w[b] <= prm.max_w[b]) continue;
        std::vector<size_t> idx;
        for (size_t i = 0; i < cands.size(); ++i) if (cands[i].to == b) idx.push_back(i);
        // order: gain clamped to 24 bits, then the top 30 hash bits, then id
        auto clampg = [](int32_t g) { return std::min<int64_t>(std::max<int64_t>(int64_t(g) + (1 << 23), 0), (1 << 24) - 1); };
        std::sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
          const Cand& a = cands[x]; const Cand& c = cands[y];
          if (clampg(a.gain) != clampg(c.gain)) return clampg(a.gain) < clampg(c.gain);
          if ((a.prio >> 34) != (c.prio >> 34)) return (a.prio >> 34) < (c.prio >> 34);
          return a.u < c.u;
        });
        int64_t cum = p.part_w[b];
        bool full = false;
        for (size_t i : idx) {
          if (!full && cum + hg.vw[cands[i].u] <= prm.max_w[b]) cum += hg.vw[cands[i].u];
          else { full = true; admit[i] = 0; }
        }
      }
      std::vector<Cand> moved;
      int64_t D = 0;
      for (size_t i = 0; i < cands.size(); ++i)
        if (admit[i]) { D += apply_move(p, cands[i].u, cands[i].from, cands[i].to); moved.push_back(cands[i]); }
      // real delta of every applied move on the post-move state
      std::vector<uint8_t> keep(moved.size(), 1);
      for (size_t i = 0; i < moved.size(); ++i) {
        const Cand& c = moved[i];
        int64_t real = 0;
        for (uint64_t j = hg.voff[c.u]; j < hg.voff[c.u + 1]; ++j) {
          const uint32_t e = hg.inc[j];
          if (p.pcount(e, c.to) == 1) real += hg.nw[e];
          if (p.pcount(e, c.from) == 0) real -= hg.nw[e];
        }
        if (!(real == c.gain || real <= 0)) keep[i] = 0;
      }
      for (size_t i = 0; i < moved.size(); ++i)
        if (!keep[i]) { D += apply_move(p, moved[i].u, moved[i].to, moved[i].from); st.reverts++; }
      if (D > 0) {
        for (size_t i = 0; i < moved.size(); ++i)
          if (keep[i]) { D += apply_move(p, moved[i].u, moved[i].to, moved[i].from); keep[i] = 0; }
        stop = true;
      }
      st.delta += D;
      for (size_t i = 0; i < moved.size(); ++i) {
        if (!keep[i]) continue;
        st.moves++;
        const uint32_t u = moved[i].u;
        for (uint64_t j = hg.voff[u]; j < hg.voff[u + 1]; ++j) {
          const uint32_t e = hg.inc[j];
          if (hg.esize(e) <= prm.activation_threshold)
            for (uint64_t q = hg.noff[e]; q < hg.noff[e + 1]; ++q) next_flag[hg.pins[q]] = 1;
        }
        next_flag[u] = 1;
      }
    }
    active.clear();
    for (uint32_t u = 0; u < hg.n; ++u) if (next_flag[u]) active.push_back(u);
    st.rounds++;
  }
  return st;
}

}  // namespace orc
