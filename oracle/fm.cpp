// fm.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
// Multitry localized k-way FM with a gain cache and global rollback
// (multitry_kway_fm.cpp:30-150, localized_kway_fm_core.cpp:26-252,
// global_rollback.cpp:146-300, gain_cache_strategy.h:72-160, stop_rule.h:25-60,
// partitioned_hypergraph.h:647-687,707-809) in the synchronous form the
// device computes:
//
//   * A round orders the border vertices of the state at its start by a
//     hash and hands them to searches, `seeds` per search.
//   * Every search runs on the state at the start of its group (below), with
//     the gain cache of that state (benefit, incident weight and connectivity
//     weight per block of every vertex) and a delta of its own
//     (pin counts of the nets it touched, block weights): a priority queue of
//     at most `pq` vertices keyed by their exact gain on the delta state
//     (computeBestTargetBlock), the best one moves, the keys of the queue are
//     recomputed, and the pins of the nets whose gains changed enter the
//     queue (the best `pq` by (gain, hash, id) stay).  The adaptive stopping
//     rule and the best local prefix (km1, or equal km1 with a lighter
//     heaviest block) follow the reference.  A vertex whose key would need
//     corrections in more than 8 blocks is not queued.  A search also stops after
//     `max_moves` moves, or when its delta would exceed `delta_nets` nets
//     or 4 blocks in one net, or an expansion would see more than
//     `expand_cap` pins.
//   * The searches of a round run in `sub_rounds` groups of consecutive
//     searches; a group runs on the state the earlier groups left (with the
//     gain cache of that state), as the reference's later searches see the
//     moves that earlier ones applied to the shared partition.  Within a
//     group, the best prefixes are concatenated in search order; a vertex
//     already moved by an earlier search of the group is dropped.  The gains
//     of the sequence are recalculated exactly (recalculateGains), the best
//     balanced prefix is applied (BalanceAndBestIndexScan) and the rest is
//     not.
#include "oracle.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <map>
#include <unordered_map>
#include <unordered_set>

namespace orc {

namespace {

constexpr uint64_t kFmSeedTag = 0xF5000000ull;
constexpr uint64_t kFmTieTag = 0xF6000000ull;
constexpr int kFmNetBlocks = 4;  // blocks a search may change in one net
constexpr size_t kFmCorr = 8;    // corrected blocks a queue key may take

struct Cache {
  std::vector<int64_t> benefit, incw, connw;  // connw: n*k
};

Cache gain_cache(const Partition& p) {
  const Hypergraph& hg = *p.hg;
  Cache c;
  c.benefit.assign(hg.n, 0);
  c.incw.assign(hg.n, 0);
  c.connw.assign(size_t(hg.n) * p.k, 0);
  for (uint32_t u = 0; u < hg.n; ++u) {
    const int f = p.part[u];
    for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
      const uint32_t e = hg.inc[i];
      const int64_t w = hg.nw[e];
      c.incw[u] += w;
      if (p.pcount(e, f) == 1) c.benefit[u] += w;
      for (int b = 0; b < p.k; ++b)
        if (p.has(e, b)) c.connw[size_t(u) * p.k + b] += w;
    }
  }
  return c;
}

// stop_rule.h:25-60
struct StopRule {
  explicit StopRule(uint32_t n) : beta(std::log(double(n))) {}
  bool should_stop() const {
    return (double(steps) > beta) && (Mk == 0 || double(steps) >= (variance / (Mk * Mk)) * stop_factor);
  }
  void update(int64_t gain) {
    ++steps;
    const double g = double(gain);
    if (steps == 1) {
      Mk = g; Mk_prev = g; Sk_prev = 0.0; variance = 0.0;
    } else {
      Mk = Mk_prev + (g - Mk_prev) / double(steps);
      Sk = Sk_prev + (g - Mk_prev) * (g - Mk);
      variance = Sk / (double(steps) - 1.0);
      Mk_prev = Mk;
      Sk_prev = Sk;
    }
  }
  void reset() { steps = 0; variance = 0.0; }
  uint64_t steps = 0;
  double variance = 0.0, Mk = 0.0, Mk_prev = 0.0, Sk = 0.0, Sk_prev = 0.0;
  const double stop_factor = 1.0 / 2.0 - 0.25;
  double beta;
};

struct Entry {
  uint32_t v;
  int64_t key;
  int32_t to;
  uint64_t h;
};
bool better(const Entry& a, const Entry& b) {
  if (a.key != b.key) return a.key > b.key;
  if (a.h != b.h) return a.h > b.h;
  return a.v < b.v;
}

struct Search {
  const Partition& p;
  const Cache& c;
  const FmParams& prm;
  uint64_t tie;  // tie tag of the round
  std::map<uint32_t, std::vector<std::pair<int, int32_t>>> delta;  // net -> (block, pin delta)
  std::vector<int64_t> dW;
  std::unordered_set<uint32_t> moved;
  std::vector<Entry> pq;
  std::vector<FmMove> moves;

  Search(const Partition& p_, const Cache& c_, const FmParams& prm_, uint64_t tie_)
      : p(p_), c(c_), prm(prm_), tie(tie_), dW(p_.k, 0) {}

  int32_t d(uint32_t e, int b) const {
    auto it = delta.find(e);
    if (it == delta.end()) return 0;
    for (auto& x : it->second)
      if (x.first == b) return x.second;
    return 0;
  }
  int64_t pc(uint32_t e, int b) const { return int64_t(p.pcount(e, b)) + d(e, b); }
  int64_t W(int b) const { return p.part_w[b] + dW[b]; }

  // computeBestTargetBlock (gain_cache_strategy.h:129-155) on the delta
  // state for an unmoved vertex; false when no block can take it
  bool key(uint32_t y, Entry& out) const {
    const Hypergraph& hg = *p.hg;
    const int f = p.part[y];
    int64_t benefit = c.benefit[y];
    std::map<int, int64_t> corr;  // block -> connectivity weight correction
    for (uint64_t i = hg.voff[y]; i < hg.voff[y + 1]; ++i) {
      const uint32_t e = hg.inc[i];
      auto it = delta.find(e);
      if (it == delta.end()) continue;
      const int64_t w = hg.nw[e];
      benefit += w * ((pc(e, f) == 1 ? 1 : 0) - (p.pcount(e, f) == 1 ? 1 : 0));
      for (auto& x : it->second) {
        if (x.first == f) continue;
        const int ch = (int64_t(p.pcount(e, x.first)) + x.second >= 1 ? 1 : 0) - (p.pcount(e, x.first) >= 1 ? 1 : 0);
        if (ch) corr[x.first] += w * ch;
      }
    }
    if (corr.size() > kFmCorr) return false;
    const int64_t wy = hg.vw[y];
    int32_t to = -1;
    int64_t best_pen = INT64_MAX, best_tw = W(f) - wy;
    for (int b = 0; b < p.k; ++b) {
      if (b == f) continue;
      auto ci = corr.find(b);
      const int64_t pen = c.incw[y] - (c.connw[size_t(y) * p.k + b] + (ci == corr.end() ? 0 : ci->second));
      const int64_t tw = W(b);
      if ((pen < best_pen || (pen == best_pen && tw < best_tw)) && tw + wy <= prm.max_w[b]) {
        best_pen = pen;
        to = b;
        best_tw = tw;
      }
    }
    if (to < 0) return false;
    out = Entry{y, benefit - best_pen, to, hash4(prm.seed, tie, y, 0)};
    return true;
  }

  // keeps the best pq entries of pq + cand (no vertex twice)
  void merge(std::vector<Entry>& cand) {
    for (auto& x : cand) pq.push_back(x);
    std::sort(pq.begin(), pq.end(), better);
    pq.erase(std::unique(pq.begin(), pq.end(), [](const Entry& a, const Entry& b) { return a.v == b.v; }),
             pq.end());
    if (pq.size() > size_t(prm.pq)) pq.resize(prm.pq);
  }

  // returns the length of the best local prefix
  size_t run(const std::vector<uint32_t>& seeds) {
    const Hypergraph& hg = *p.hg;
    std::vector<Entry> cand;
    for (uint32_t v : seeds) {
      Entry x;
      if (key(v, x)) cand.push_back(x);
    }
    merge(cand);
    StopRule stop(hg.n);
    int64_t est = 0, best = 0;
    size_t best_idx = 0;
    while (!stop.should_stop() && moves.size() < size_t(prm.max_moves) && !pq.empty()) {
      const Entry top = pq.front();
      const uint32_t v = top.v;
      const int from = p.part[v], to = top.to;
      // capacity of the delta
      size_t new_nets = 0;
      bool fits = true;
      for (uint64_t i = hg.voff[v]; i < hg.voff[v + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        auto it = delta.find(e);
        if (it == delta.end()) { ++new_nets; continue; }
        int need = int(it->second.size());
        bool hf = false, ht = false;
        for (auto& x : it->second) { hf |= x.first == from; ht |= x.first == to; }
        need += (hf ? 0 : 1) + (ht ? 0 : 1);
        if (need > kFmNetBlocks) fits = false;
      }
      if (!fits || delta.size() + new_nets > prm.delta_nets) break;
      int64_t heaviest = W(0);
      for (int b = 1; b < p.k; ++b) heaviest = std::max(heaviest, W(b));
      const int64_t from_w = W(from), to_w = W(to), wv = hg.vw[v];
      for (uint64_t i = hg.voff[v]; i < hg.voff[v + 1]; ++i) {
        auto& rec = delta[hg.inc[i]];
        bool hf = false, ht = false;
        for (auto& x : rec) {
          if (x.first == from) { x.second -= 1; hf = true; }
          if (x.first == to) { x.second += 1; ht = true; }
        }
        if (!hf) rec.push_back({from, -1});
        if (!ht) rec.push_back({to, 1});
      }
      dW[from] -= wv;
      dW[to] += wv;
      moved.insert(v);
      moves.push_back(FmMove{v, from, to, top.key});
      pq.erase(pq.begin());
      est += top.key;
      stop.update(top.key);
      if (est > best || (est >= best && from_w == heaviest && to_w + wv < heaviest)) {
        stop.reset();
        best = est;
        best_idx = moves.size();
      }
      // exact keys of the queue on the new state
      std::vector<Entry> kept;
      for (auto& x : pq) {
        Entry y;
        if (key(x.v, y)) kept.push_back(y);
      }
      pq.clear();
      // pins of the nets whose gains changed (localized_kway_fm_core.cpp:
      // 103-109, 57-82), nets of at least `ignore` pins skipped
      std::vector<uint32_t> nets;
      uint64_t seen = 0;
      for (uint64_t i = hg.voff[v]; i < hg.voff[v + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        if (hg.esize(e) >= prm.ignore) continue;
        const int64_t cf = pc(e, from), ct = pc(e, to);
        if (cf == 0 || cf == 1 || ct == 1 || ct == 2) {
          nets.push_back(e);
          seen += hg.esize(e);
        }
      }
      if (seen > prm.expand_cap) {
        merge(kept);
        break;
      }
      std::unordered_set<uint32_t> inq;
      for (auto& x : kept) inq.insert(x.v);
      std::unordered_set<uint32_t> newc;
      for (uint32_t e : nets)
        for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) {
          const uint32_t y = hg.pins[i];
          if (moved.count(y) || inq.count(y)) continue;
          newc.insert(y);
        }
      for (uint32_t y : newc) {
        Entry x;
        if (key(y, x)) kept.push_back(x);
      }
      merge(kept);
    }
    return best_idx;
  }
};

}  // namespace

// recalculateGains (global_rollback.cpp:207-283): exact gain of every move of
// the sequence given all of them are applied in order.  State p is the one
// before the sequence.
std::vector<int64_t> fm_recalculate(const Partition& p, const std::vector<FmMove>& seq) {
  const Hypergraph& hg = *p.hg;
  std::vector<int64_t> gain(seq.size(), 0);
  // move ids are 1-based, as the reference's (fm_commons.h:40-45), so that
  // last_out = 0 means that no pin left the block
  std::unordered_map<uint32_t, uint32_t> move_of;
  for (size_t i = 0; i < seq.size(); ++i) move_of[seq[i].v] = uint32_t(i) + 1;
  std::unordered_set<uint32_t> nets;
  for (auto& m : seq)
    for (uint64_t i = hg.voff[m.v]; i < hg.voff[m.v + 1]; ++i) nets.insert(hg.inc[i]);
  std::vector<uint32_t> first_in(p.k), last_out(p.k), remaining(p.k);
  for (uint32_t e : nets) {
    std::fill(first_in.begin(), first_in.end(), UINT32_MAX);
    std::fill(last_out.begin(), last_out.end(), 0u);
    std::fill(remaining.begin(), remaining.end(), 0u);
    std::vector<uint32_t> mv;
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) {
      const uint32_t v = hg.pins[i];
      auto it = move_of.find(v);
      if (it != move_of.end()) {
        const FmMove& m = seq[it->second - 1];
        first_in[m.to] = std::min(first_in[m.to], it->second);
        last_out[m.from] = std::max(last_out[m.from], it->second);
        mv.push_back(it->second);
      } else {
        remaining[p.part[v]]++;
      }
    }
    const int64_t w = hg.nw[e];
    for (uint32_t id : mv) {
      const FmMove& m = seq[id - 1];
      const bool benefit = last_out[m.from] == id && first_in[m.from] > id && remaining[m.from] == 0;
      const bool penalty = first_in[m.to] == id && last_out[m.to] < id && remaining[m.to] == 0;
      if (benefit && !penalty) gain[id - 1] += w;
      if (!benefit && penalty) gain[id - 1] -= w;
    }
  }
  return gain;
}

// BalanceAndBestIndexScan (global_rollback.cpp:29-141) in sequential form:
// the prefix with the largest gain sum and no block above its bound, ties
// to the lighter heaviest block, then to the shorter prefix
size_t fm_best_prefix(const Partition& p, const std::vector<FmMove>& seq, const std::vector<int64_t>& gain,
                      const std::vector<int64_t>& max_w, int64_t* best_gain) {
  const Hypergraph& hg = *p.hg;
  std::vector<int64_t> w(p.part_w.begin(), p.part_w.end());
  int over = 0;
  for (int b = 0; b < p.k; ++b) over += w[b] > max_w[b] ? 1 : 0;
  int64_t bg = 0, bh = *std::max_element(w.begin(), w.end());
  size_t bi = 0;
  int64_t sum = 0;
  for (size_t i = 0; i < seq.size(); ++i) {
    const FmMove& m = seq[i];
    sum += gain[i];
    const bool fo = w[m.from] > max_w[m.from];
    w[m.from] -= hg.vw[m.v];
    if (fo && w[m.from] <= max_w[m.from]) --over;
    const bool to_ = w[m.to] > max_w[m.to];
    w[m.to] += hg.vw[m.v];
    if (!to_ && w[m.to] > max_w[m.to]) ++over;
    if (over == 0 && sum >= bg) {
      const int64_t h = *std::max_element(w.begin(), w.end());
      if (sum > bg || h < bh) {
        bg = sum;
        bh = h;
        bi = i + 1;
      }
    }
  }
  *best_gain = bg;
  return bi;
}

void apply_move(Partition& p, const FmMove& m) {
  const Hypergraph& hg = *p.hg;
  for (uint64_t i = hg.voff[m.v]; i < hg.voff[m.v + 1]; ++i) {
    const uint32_t e = hg.inc[i];
    uint32_t& cf = p.pc[size_t(e) * p.k + m.from];
    uint32_t& ct = p.pc[size_t(e) * p.k + m.to];
    if (--cf == 0) p.conn[size_t(e) * p.cw + m.from / 64] ^= 1ull << (m.from % 64);
    if (ct++ == 0) p.conn[size_t(e) * p.cw + m.to / 64] ^= 1ull << (m.to % 64);
  }
  p.part[m.v] = m.to;
  p.part_w[m.from] -= hg.vw[m.v];
  p.part_w[m.to] += hg.vw[m.v];
}

FmStats fm_refine_sync(Partition& p, const FmParams& prm, std::vector<FmRound>* trace) {
  const Hypergraph& hg = *p.hg;
  FmStats st;
  for (int r = 0; r < prm.rounds; ++r) {
    std::vector<std::pair<uint64_t, uint32_t>> order;
    for (uint32_t u = 0; u < hg.n; ++u)
      if (is_border(p, u)) order.push_back({hash4(prm.seed, kFmSeedTag | uint64_t(r), u, 0), u});
    if (order.empty()) break;
    std::sort(order.begin(), order.end());
    const size_t searches = (order.size() + prm.seeds - 1) / prm.seeds;
    const size_t S = size_t(std::max(1, prm.sub_rounds));
    int64_t round_improvement = 0;
    // sub-round j runs searches [j * searches / S, (j + 1) * searches / S)
    // on the state the earlier sub-rounds left
    for (size_t j = 0; j < S; ++j) {
      const size_t s0 = j * searches / S, s1 = (j + 1) * searches / S;
      if (s0 == s1) continue;
      const Cache c = gain_cache(p);
      std::vector<FmMove> seq;
      std::unordered_set<uint32_t> owner;
      FmRound tr;
      for (size_t s = s0; s < s1; ++s) {
        std::vector<uint32_t> seeds;
        for (size_t i = s * prm.seeds; i < std::min(order.size(), (s + 1) * prm.seeds); ++i)
          seeds.push_back(order[i].second);
        Search Sr(p, c, prm, kFmTieTag | uint64_t(r));
        const size_t len = Sr.run(seeds);
        tr.search_len.push_back(uint32_t(len));
        for (size_t i = 0; i < len; ++i) {
          tr.search_moves.push_back(Sr.moves[i]);
          if (owner.insert(Sr.moves[i].v).second) seq.push_back(Sr.moves[i]);
        }
      }
      const std::vector<int64_t> gain = fm_recalculate(p, seq);
      int64_t improvement = 0;
      const size_t keep = fm_best_prefix(p, seq, gain, prm.max_w, &improvement);
      for (size_t i = 0; i < keep; ++i) apply_move(p, seq[i]);
      tr.sequence = seq;
      tr.gains = gain;
      tr.kept = keep;
      tr.improvement = improvement;
      if (trace) trace->push_back(std::move(tr));
      st.moves += int64_t(keep);
      round_improvement += improvement;
    }
    st.rounds += 1;
    st.searches += int64_t(searches);
    st.improvement += round_improvement;
    if (round_improvement <= 0) break;
  }
  return st;
}

}  // namespace orc
