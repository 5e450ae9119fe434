// Initial partitioning of the coarsest hypergraph (see initial_partitioning.h).
#include "initial_partitioning.h"

#include <algorithm>
#include <numeric>
#include <queue>
#include <random>

namespace mt_kahypar_amd {

void build_host_incidence(HostHypergraph& hg) {
  hg.voff.assign(size_t(hg.n) + 1, 0);
  for (uint64_t i = 0; i < hg.pins.size(); ++i) hg.voff[hg.pins[i] + 1]++;
  for (uint32_t u = 0; u < hg.n; ++u) hg.voff[u + 1] += hg.voff[u];
  hg.inc.assign(hg.pins.size(), 0);
  std::vector<uint64_t> pos(hg.voff.begin(), hg.voff.end() - 1);
  for (uint32_t e = 0; e < hg.m; ++e)
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) hg.inc[pos[hg.pins[i]]++] = e;
}

namespace {

struct State {
  const HostHypergraph& hg;
  int k;
  std::vector<int32_t> part;
  std::vector<int64_t> pw;
  std::vector<uint32_t> pc;
  State(const HostHypergraph& h, int kk, const std::vector<int32_t>& p) : hg(h), k(kk), part(p) {
    pw.assign(k, 0);
    pc.assign(size_t(hg.m) * k, 0);
    for (uint32_t u = 0; u < hg.n; ++u) pw[part[u]] += hg.vw[u];
    for (uint32_t e = 0; e < hg.m; ++e)
      for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) pc[size_t(e) * k + part[hg.pins[i]]]++;
  }
  int64_t km1() const {
    int64_t s = 0;
    for (uint32_t e = 0; e < hg.m; ++e) {
      int l = 0;
      for (int b = 0; b < k; ++b) l += pc[size_t(e) * k + b] > 0;
      s += int64_t(std::max(l - 1, 0)) * hg.nw[e];
    }
    return s;
  }
  // km1 delta of moving u to every block (delta[from] = 0)
  void deltas(uint32_t u, std::vector<int64_t>& d) const {
    const int from = part[u];
    std::fill(d.begin(), d.end(), 0);
    int64_t leave = 0, total = 0;
    for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
      const uint32_t e = hg.inc[i];
      const int64_t w = hg.nw[e];
      total += w;
      if (pc[size_t(e) * k + from] == 1) leave += w;
      for (int b = 0; b < k; ++b)
        if (b != from && pc[size_t(e) * k + b] > 0) d[b] -= w;
    }
    for (int b = 0; b < k; ++b) d[b] = (b == from) ? 0 : d[b] + total - leave;
  }
  void move(uint32_t u, int to) {
    const int from = part[u];
    part[u] = to;
    pw[from] -= hg.vw[u];
    pw[to] += hg.vw[u];
    for (uint64_t i = hg.voff[u]; i < hg.voff[u + 1]; ++i) {
      const uint32_t e = hg.inc[i];
      pc[size_t(e) * k + from]--;
      pc[size_t(e) * k + to]++;
    }
  }
};

std::vector<int32_t> greedy_growing(const HostHypergraph& hg, int k, const std::vector<int64_t>& max_w,
                                    std::mt19937_64& rng) {
  std::vector<int32_t> part(hg.n, -1);
  int64_t W = 0;
  for (uint32_t u = 0; u < hg.n; ++u) W += hg.vw[u];
  const int64_t perfect = (W + k - 1) / k;
  std::vector<int64_t> pw(k, 0);
  std::vector<double> score(hg.n, 0.0);
  std::vector<uint32_t> unassigned(hg.n);
  std::iota(unassigned.begin(), unassigned.end(), 0u);
  std::shuffle(unassigned.begin(), unassigned.end(), rng);
  // part: block id, -1 free, -2 free but too heavy for the block being grown
  // (free again for the next block)
  size_t next_seed = 0;
  auto pick_seed = [&]() -> int64_t {
    while (next_seed < unassigned.size() && part[unassigned[next_seed]] >= 0) ++next_seed;
    for (size_t i = next_seed; i < unassigned.size(); ++i)
      if (part[unassigned[i]] == -1) return int64_t(unassigned[i]);
    return -1;
  };
  for (int b = 0; b + 1 < k; ++b) {
    std::priority_queue<std::pair<double, uint32_t>> pq;
    std::vector<uint32_t> touched;
    while (pw[b] < perfect) {
      int64_t v = -1;
      while (!pq.empty()) {
        const auto top = pq.top();
        pq.pop();
        if (part[top.second] == -1 && top.first == score[top.second]) { v = top.second; break; }
      }
      if (v < 0) v = pick_seed();
      if (v < 0) break;
      if (pw[b] + hg.vw[v] > max_w[b]) { part[v] = -2; touched.push_back(uint32_t(v)); continue; }
      part[v] = b;
      pw[b] += hg.vw[v];
      for (uint64_t i = hg.voff[v]; i < hg.voff[v + 1]; ++i) {
        const uint32_t e = hg.inc[i];
        const uint64_t sz = hg.noff[e + 1] - hg.noff[e];
        const double add = double(hg.nw[e]) / double(std::max<uint64_t>(sz - 1, 1));
        for (uint64_t j = hg.noff[e]; j < hg.noff[e + 1]; ++j) {
          const uint32_t x = hg.pins[j];
          if (part[x] < 0 && part[x] != -2) {
            if (score[x] == 0.0) touched.push_back(x);
            score[x] += add;
            pq.push({score[x], x});
          }
        }
      }
    }
    for (uint32_t x : touched) { score[x] = 0.0; if (part[x] == -2) part[x] = -1; }
  }
  // the rest goes to the last block, or the lightest block that fits
  for (uint32_t u = 0; u < hg.n; ++u) {
    if (part[u] >= 0) continue;
    int b = k - 1;
    if (pw[b] + hg.vw[u] > max_w[b]) b = int(std::min_element(pw.begin(), pw.end()) - pw.begin());
    part[u] = b;
    pw[b] += hg.vw[u];
  }
  return part;
}

std::vector<int32_t> random_balanced(const HostHypergraph& hg, int k, std::mt19937_64& rng) {
  std::vector<uint32_t> order(hg.n);
  std::iota(order.begin(), order.end(), 0u);
  std::shuffle(order.begin(), order.end(), rng);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return hg.vw[a] > hg.vw[b]; });
  std::vector<int64_t> pw(k, 0);
  std::vector<int32_t> part(hg.n, 0);
  for (uint32_t u : order) {
    const int b = int(std::min_element(pw.begin(), pw.end()) - pw.begin());
    part[u] = b;
    pw[b] += hg.vw[u];
  }
  return part;
}

void refine(State& s, const std::vector<int64_t>& max_w, int passes, std::mt19937_64& rng) {
  const int k = s.k;
  std::vector<int64_t> d(k);
  std::vector<uint32_t> order(s.hg.n);
  std::iota(order.begin(), order.end(), 0u);
  // restore balance first: move the cheapest vertices out of overloaded blocks
  for (int guard = 0; guard < int(s.hg.n); ++guard) {
    int over = -1;
    for (int b = 0; b < k; ++b) if (s.pw[b] > max_w[b]) { over = b; break; }
    if (over < 0) break;
    int64_t best = INT64_MAX; uint32_t bu = 0; int bt = -1;
    for (uint32_t u = 0; u < s.hg.n; ++u) {
      if (s.part[u] != over) continue;
      s.deltas(u, d);
      for (int b = 0; b < k; ++b)
        if (b != over && s.pw[b] + s.hg.vw[u] <= max_w[b] && d[b] < best) { best = d[b]; bu = u; bt = b; }
    }
    if (bt < 0) break;
    s.move(bu, bt);
  }
  for (int p = 0; p < passes; ++p) {
    std::shuffle(order.begin(), order.end(), rng);
    bool improved = false;
    for (uint32_t u : order) {
      const int from = s.part[u];
      s.deltas(u, d);
      int bt = from;
      int64_t bd = 0;
      for (int b = 0; b < k; ++b) {
        if (b == from || s.pw[b] + s.hg.vw[u] > max_w[b]) continue;
        const bool better = d[b] < bd ||
            (d[b] == bd && bd == 0 && bt == from && s.pw[from] > s.pw[b] + s.hg.vw[u]);
        if (better) { bt = b; bd = d[b]; }
      }
      if (bt != from) { s.move(u, bt); improved |= bd < 0; }
    }
    if (!improved) break;
  }
}

bool balanced(const State& s, const std::vector<int64_t>& max_w) {
  for (int b = 0; b < s.k; ++b) if (s.pw[b] > max_w[b]) return false;
  return true;
}

}  // namespace

std::vector<int32_t> initial_partition(const HostHypergraph& hg, const InitialPartitioningParams& prm,
                                       int64_t* km1_out) {
  const int k = prm.k;
  std::mt19937_64 rng(prm.seed * 0x9E3779B97F4A7C15ull + 12345);
  std::vector<int32_t> best;
  int64_t best_km1 = INT64_MAX;
  bool best_bal = false;
  const int runs = std::max(1, prm.runs);
  for (int r = 0; r < runs; ++r) {
    std::vector<int32_t> p = (r % 4 == 3) ? random_balanced(hg, k, rng) : greedy_growing(hg, k, prm.max_w, rng);
    State s(hg, k, p);
    refine(s, prm.max_w, prm.refinement_passes, rng);
    const int64_t km = s.km1();
    const bool bal = balanced(s, prm.max_w);
    if (best.empty() || (bal && !best_bal) || (bal == best_bal && km < best_km1)) {
      best = s.part; best_km1 = km; best_bal = bal;
    }
  }
  if (km1_out) *km1_out = best_km1;
  return best;
}

}  // namespace mt_kahypar_amd
