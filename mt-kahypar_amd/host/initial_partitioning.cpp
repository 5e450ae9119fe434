// Initial partitioning of the coarsest hypergraph (see initial_partitioning.h).
#include "initial_partitioning.h"

#include <algorithm>
#include <atomic>
#include <numeric>
#include <queue>
#include <cmath>
#include <cstdint>
#include <deque>
#include <random>
#include <thread>

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

// Sub-hypergraph of one recursion step: CSR both ways, weights, and the id of
// every vertex in the coarsest hypergraph.
struct Sub {
  uint32_t n = 0, m = 0;
  std::vector<uint64_t> noff, voff;
  std::vector<uint32_t> pins, inc;
  std::vector<int32_t> vw, nw;
  std::vector<uint32_t> orig;
  int64_t W = 0;
  void finish() {
    voff.assign(size_t(n) + 1, 0);
    for (uint32_t v : pins) voff[v + 1]++;
    for (uint32_t u = 0; u < n; ++u) voff[u + 1] += voff[u];
    inc.assign(pins.size(), 0);
    std::vector<uint64_t> pos(voff.begin(), voff.end() - 1);
    for (uint32_t e = 0; e < m; ++e)
      for (uint64_t i = noff[e]; i < noff[e + 1]; ++i) inc[pos[pins[i]]++] = e;
    W = 0;
    for (int32_t w : vw) W += w;
  }
};

// nets larger than this do not steer greedy growing (their pins would be
// rescored once per added vertex)
constexpr uint64_t kGrowNetLimit = 100;

// Max-heap over vertex scores with one entry per vertex (increase-key in
// place); ties go to the smaller vertex id.
struct ScoreHeap {
  std::vector<uint32_t> heap;
  std::vector<int32_t> pos;  // index in heap, -1 when absent
  std::vector<double> key;
  explicit ScoreHeap(uint32_t n) : pos(n, -1), key(n, 0.0) {}
  bool above(uint32_t a, uint32_t b) const { return key[a] > key[b] || (key[a] == key[b] && a < b); }
  void up(size_t i) {
    const uint32_t v = heap[i];
    while (i > 0) {
      const size_t p = (i - 1) / 2;
      if (!above(v, heap[p])) break;
      heap[i] = heap[p];
      pos[heap[i]] = int32_t(i);
      i = p;
    }
    heap[i] = v;
    pos[v] = int32_t(i);
  }
  void down(size_t i) {
    const uint32_t v = heap[i];
    const size_t n = heap.size();
    while (true) {
      size_t c = 2 * i + 1;
      if (c >= n) break;
      if (c + 1 < n && above(heap[c + 1], heap[c])) ++c;
      if (!above(heap[c], v)) break;
      heap[i] = heap[c];
      pos[heap[i]] = int32_t(i);
      i = c;
    }
    heap[i] = v;
    pos[v] = int32_t(i);
  }
  void add(uint32_t v, double d) {  // key[v] += d, inserting v if absent
    key[v] += d;
    if (pos[v] < 0) {
      heap.push_back(v);
      up(heap.size() - 1);
    } else {
      up(size_t(pos[v]));
    }
  }
  // Deferred form of add: key[v] += d now, the heap position is restored by
  // settle() once per touched vertex (a grown vertex's nets may add to the
  // same vertex thousands of times).  Keys are summed in the same order as
  // by add(), so the heap holds the same keys and pops the same vertices.
  std::vector<uint32_t> touched;
  std::vector<uint8_t> is_touched;
  void add_deferred(uint32_t v, double d) {
    if (is_touched.empty()) is_touched.assign(pos.size(), 0);
    key[v] += d;
    if (!is_touched[v]) {
      is_touched[v] = 1;
      touched.push_back(v);
    }
  }
  void settle() {
    for (uint32_t v : touched) {
      is_touched[v] = 0;
      if (pos[v] < 0) {
        heap.push_back(v);
        up(heap.size() - 1);
      } else {
        up(size_t(pos[v]));
      }
    }
    touched.clear();
  }
  bool empty() const { return heap.empty(); }
  uint32_t pop() {
    const uint32_t top = heap[0];
    pos[top] = -1;
    const uint32_t last = heap.back();
    heap.pop_back();
    if (!heap.empty()) {
      heap[0] = last;
      down(0);
    }
    return top;
  }
};

struct Bisector {
  const Sub& h;
  int64_t max_w[2];
  int64_t target0;
  std::vector<uint8_t> side;
  std::vector<uint32_t> pc[2];  // pins per side and net
  int64_t w[2] = {0, 0};

  Bisector(const Sub& s, int64_t m0, int64_t m1, int64_t t0) : h(s), max_w{m0, m1}, target0(t0) {}

  void counts() {
    pc[0].assign(h.m, 0);
    pc[1].assign(h.m, 0);
    w[0] = w[1] = 0;
    for (uint32_t e = 0; e < h.m; ++e)
      for (uint64_t i = h.noff[e]; i < h.noff[e + 1]; ++i) pc[side[h.pins[i]]][e]++;
    for (uint32_t u = 0; u < h.n; ++u) w[side[u]] += h.vw[u];
  }
  int64_t cut() const {
    int64_t c = 0;
    for (uint32_t e = 0; e < h.m; ++e) c += (pc[0][e] && pc[1][e]) ? h.nw[e] : 0;
    return c;
  }
  // cut reduction of moving u to the other side
  int64_t gain(uint32_t u) const {
    const int f = side[u], t = 1 - f;
    int64_t g = 0;
    for (uint64_t i = h.voff[u]; i < h.voff[u + 1]; ++i) {
      const uint32_t e = h.inc[i];
      if (pc[f][e] == 1) g += h.nw[e];
      if (pc[t][e] == 0) g -= h.nw[e];
    }
    return g;
  }
  void move(uint32_t u) {
    const int f = side[u], t = 1 - f;
    side[u] = uint8_t(t);
    w[f] -= h.vw[u];
    w[t] += h.vw[u];
    for (uint64_t i = h.voff[u]; i < h.voff[u + 1]; ++i) {
      const uint32_t e = h.inc[i];
      pc[f][e]--;
      pc[t][e]++;
    }
  }

  // greedy growing of side 0 from a random seed: the next vertex is the one
  // with the largest connection to side 0 (sum of w(e) / (|e| - 1))
  void grow_greedy(std::mt19937_64& rng) {
    side.assign(h.n, 1);
    std::vector<uint8_t> skip(h.n, 0);
    ScoreHeap pq(h.n);
    int64_t w0 = 0;
    std::uniform_int_distribution<uint32_t> pick(0, h.n - 1);
    uint32_t scan = 0;
    const uint32_t start = pick(rng);
    while (w0 < target0) {
      int64_t v = pq.empty() ? -1 : int64_t(pq.pop());
      while (v < 0 && scan < h.n) {  // a new seed: the next free vertex after `start`
        const uint32_t x = (start + scan++) % h.n;
        if (side[x] == 1 && !skip[x]) v = x;
      }
      if (v < 0) break;
      if (w0 + h.vw[v] > max_w[0]) { skip[v] = 1; continue; }
      side[v] = 0;
      w0 += h.vw[v];
      for (uint64_t i = h.voff[v]; i < h.voff[v + 1]; ++i) {
        const uint32_t e = h.inc[i];
        const uint64_t sz = h.noff[e + 1] - h.noff[e];
        if (sz > kGrowNetLimit) continue;
        const double add = double(h.nw[e]) / double(std::max<uint64_t>(sz - 1, 1));
        for (uint64_t j = h.noff[e]; j < h.noff[e + 1]; ++j) {
          const uint32_t x = h.pins[j];
          if (side[x] == 1 && !skip[x]) pq.add_deferred(x, add);
        }
      }
      pq.settle();
    }
  }

  // breadth-first growing of side 0 from a random seed (a net's pins are
  // all seen after its first scan, so each net is scanned once)
  void grow_bfs(std::mt19937_64& rng) {
    side.assign(h.n, 1);
    std::vector<uint8_t> seen(h.n, 0), scanned(h.m, 0);
    std::deque<uint32_t> q;
    int64_t w0 = 0;
    const uint32_t start = std::uniform_int_distribution<uint32_t>(0, h.n - 1)(rng);
    uint32_t scan = 0;
    while (w0 < target0) {
      if (q.empty()) {
        while (scan < h.n && seen[(start + scan) % h.n]) ++scan;
        if (scan >= h.n) break;
        const uint32_t s = (start + scan) % h.n;
        seen[s] = 1;
        q.push_back(s);
      }
      const uint32_t v = q.front();
      q.pop_front();
      if (w0 + h.vw[v] > max_w[0]) continue;
      side[v] = 0;
      w0 += h.vw[v];
      for (uint64_t i = h.voff[v]; i < h.voff[v + 1]; ++i) {
        const uint32_t e = h.inc[i];
        if (scanned[e] || h.noff[e + 1] - h.noff[e] > kGrowNetLimit) continue;
        scanned[e] = 1;
        for (uint64_t j = h.noff[e]; j < h.noff[e + 1]; ++j) {
          const uint32_t x = h.pins[j];
          if (!seen[x]) { seen[x] = 1; q.push_back(x); }
        }
      }
    }
  }

  void random_fill(std::mt19937_64& rng) {
    side.assign(h.n, 1);
    std::vector<uint32_t> order(h.n);
    std::iota(order.begin(), order.end(), 0u);
    std::shuffle(order.begin(), order.end(), rng);
    int64_t w0 = 0;
    for (uint32_t u : order) {
      if (w0 >= target0) break;
      if (w0 + h.vw[u] > max_w[0]) continue;
      side[u] = 0;
      w0 += h.vw[u];
    }
  }

  // Gains of all vertices kept current across moves (integers, so always
  // equal to gain(u) recomputed): moving v changes the gain of another pin
  // of a net only when the net's count on v's side is 1 or 2 or on the
  // other side 0 or 1 before the move.  v's own gain becomes the negation
  // of its gain before the move (per net, [b == 0] - [a == 1] with a, b the
  // counts on v's old and new side before the move).
  std::vector<int64_t> gc;
  void move_cached(uint32_t v) {
    const int f = side[v], t = 1 - f;
    const int64_t gv = gc[v];
    for (uint64_t i = h.voff[v]; i < h.voff[v + 1]; ++i) {
      const uint32_t e = h.inc[i];
      const uint32_t a = pc[f][e], b = pc[t][e];
      const int64_t w = h.nw[e];
      // gain change of a pin on v's side and of a pin on the other side
      const int64_t df = w * ((a == 2) - (a == 1) + (b == 0));
      const int64_t dt = w * ((b == 0) - (a == 1) - (b == 1));
      if (df || dt)
        for (uint64_t j = h.noff[e]; j < h.noff[e + 1]; ++j) {
          const uint32_t x = h.pins[j];
          if (x != v) gc[x] += side[x] == f ? df : dt;
        }
    }
    move(v);
    gc[v] = -gv;
  }

  // restore balance, then passes of positive-gain moves (zero-gain moves
  // only towards the lighter side) in random order
  void refine(int passes, std::mt19937_64& rng) {
    counts();
    gc.resize(h.n);
    for (uint32_t u = 0; u < h.n; ++u) gc[u] = gain(u);
    for (int f = 0; f < 2; ++f) {
      if (w[f] <= max_w[f]) continue;
      std::vector<std::pair<int64_t, uint32_t>> cand;
      for (uint32_t u = 0; u < h.n; ++u)
        if (side[u] == f) cand.push_back({-gc[u], u});
      std::sort(cand.begin(), cand.end());
      for (const auto& c : cand) {
        if (w[f] <= max_w[f]) break;
        if (w[1 - f] + h.vw[c.second] <= max_w[1 - f]) move_cached(c.second);
      }
    }
    std::vector<uint32_t> order(h.n);
    std::iota(order.begin(), order.end(), 0u);
    for (int p = 0; p < passes; ++p) {
      std::shuffle(order.begin(), order.end(), rng);
      bool improved = false;
      for (uint32_t u : order) {
        const int f = side[u], t = 1 - f;
        if (w[t] + h.vw[u] > max_w[t]) continue;
        const int64_t g = gc[u];
        if (g > 0 || (g == 0 && w[f] - h.vw[u] > w[t] + h.vw[u] && w[f] > w[t] + 2 * int64_t(h.vw[u]))) {
          move_cached(u);
          improved |= g > 0;
        }
      }
      if (!improved) break;
    }
  }
};

uint64_t mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Best balanced cut of the run portfolio (greedy growing, BFS, random; run r
// uses variant r % 3 and a random stream of its own, so runs may execute on
// parallel threads with the same result).
std::vector<uint8_t> bisect(const Sub& h, int64_t max0, int64_t max1, int64_t target0, int runs, int passes,
                            uint64_t stream, int threads) {
  struct Run {
    std::vector<uint8_t> side;
    int64_t cut = 0;
    bool bal = false;
  };
  std::vector<Run> res;
  auto one = [&](int r) {
    Bisector b(h, max0, max1, target0);
    std::mt19937_64 rng(mix(stream ^ mix(uint64_t(r) + 1)));
    if (r % 3 == 0) b.grow_greedy(rng);
    else if (r % 3 == 1) b.grow_bfs(rng);
    else b.random_fill(rng);
    b.refine(passes, rng);
    res[r].cut = b.cut();
    res[r].bal = b.w[0] <= max0 && b.w[1] <= max1;
    res[r].side = std::move(b.side);
  };
  auto better = [&](int a, int b) {  // run a beats run b
    return (res[a].bal && !res[b].bal) || (res[a].bal == res[b].bal && res[a].cut < res[b].cut);
  };
  // adaptive runs: batches of kBatch runs; another batch only while the last
  // one improved the best cut (batch bounds do not depend on the threads)
  constexpr int kBatch = 3;
  int best = -1;
  for (int lo = 0; lo < runs; lo += kBatch) {
    const int hi = std::min(runs, lo + kBatch);
    res.resize(hi);
    const int t = std::max(1, std::min(threads, hi - lo));
    if (t == 1) {
      for (int r = lo; r < hi; ++r) one(r);
    } else {
      std::atomic<int> next{lo};
      std::vector<std::thread> pool;
      for (int i = 0; i < t; ++i)
        pool.emplace_back([&] {
          for (int r = next++; r < hi; r = next++) one(r);
        });
      for (auto& th : pool) th.join();
    }
    bool improved = best < 0;
    for (int r = lo; r < hi; ++r) {
      if (best < 0) { best = r; continue; }
      if (better(r, best)) { best = r; improved = true; }
    }
    if (!improved) break;
  }
  return std::move(res[best].side);
}

// the vertices of side s with the pins of every net on that side (nets left
// with fewer than two pins are dropped)
Sub extract(const Sub& h, const std::vector<uint8_t>& side, int s) {
  Sub o;
  std::vector<uint32_t> id(h.n, UINT32_MAX);
  for (uint32_t u = 0; u < h.n; ++u)
    if (side[u] == s) {
      id[u] = o.n++;
      o.vw.push_back(h.vw[u]);
      o.orig.push_back(h.orig[u]);
    }
  o.noff.assign(1, 0);
  for (uint32_t e = 0; e < h.m; ++e) {
    const size_t before = o.pins.size();
    for (uint64_t i = h.noff[e]; i < h.noff[e + 1]; ++i)
      if (side[h.pins[i]] == s) o.pins.push_back(id[h.pins[i]]);
    if (o.pins.size() - before < 2) {
      o.pins.resize(before);
      continue;
    }
    o.noff.push_back(o.pins.size());
    o.nw.push_back(h.nw[e]);
    o.m += 1;
  }
  o.finish();
  return o;
}

struct Recursion {
  const InitialPartitioningParams& prm;
  double eps_level;      // adaptive imbalance per bisection level
  int64_t final_perfect; // ceil(W / k) of the coarsest hypergraph
  std::vector<int32_t>& part;
  std::atomic<int> helpers{0};  // extra threads running a recursion half

  void run(const Sub& h, int k, int first) {
    if (h.n == 0) return;
    if (k == 1) {
      for (uint32_t u = 0; u < h.n; ++u) part[h.orig[u]] = first;
      return;
    }
    const int k0 = (k + 1) / 2, k1 = k - k0;
    // side limits: (1 + eps') times the side's share of this sub-hypergraph
    const double share = double(h.W) / double(k);
    const int64_t max0 = int64_t((1.0 + eps_level) * std::ceil(share * k0));
    const int64_t max1 = int64_t((1.0 + eps_level) * std::ceil(share * k1));
    const int64_t target0 = int64_t(std::ceil(share * k0));
    int runs = std::max(1, prm.runs);
    const uint64_t P = h.pins.size();
    // the sub-problem (first block, block count) keys the random streams
    const uint64_t stream = mix(prm.seed ^ mix((uint64_t(first) << 32) | uint32_t(k)));
    const int threads = P >= 100000 ? prm.threads : 1;
    std::vector<uint8_t> side = bisect(h, max0, max1, target0, runs, prm.refinement_passes, stream, threads);
    Sub a = extract(h, side, 0), b = extract(h, side, 1);
    std::vector<uint8_t>().swap(side);
    // the halves write disjoint vertices: a large left half may run on a
    // helper thread (bounded by prm.threads - 1 helpers)
    std::thread left;
    if (a.pins.size() >= 50000 && b.pins.size() >= 50000 && helpers.fetch_add(1) < prm.threads - 1) {
      left = std::thread([&] { run(a, k0, first); });
    } else {
      if (a.pins.size() >= 50000 && b.pins.size() >= 50000) helpers.fetch_sub(1);
      run(a, k0, first);
    }
    run(b, k1, first + k0);
    if (left.joinable()) {
      left.join();
      helpers.fetch_sub(1);
    }
  }
};

// moves the lightest vertices of overloaded blocks to the lightest blocks
// they fit into
void repair_balance(const HostHypergraph& hg, int k, const std::vector<int64_t>& max_w, std::vector<int32_t>& part) {
  std::vector<int64_t> pw(k, 0);
  for (uint32_t u = 0; u < hg.n; ++u) pw[part[u]] += hg.vw[u];
  for (int b = 0; b < k; ++b) {
    if (pw[b] <= max_w[b]) continue;
    std::vector<std::pair<int32_t, uint32_t>> vs;
    for (uint32_t u = 0; u < hg.n; ++u)
      if (part[u] == b) vs.push_back({hg.vw[u], u});
    std::sort(vs.begin(), vs.end());
    for (const auto& x : vs) {
      if (pw[b] <= max_w[b]) break;
      int t = -1;
      for (int c = 0; c < k; ++c)
        if (c != b && pw[c] + x.first <= max_w[c] && (t < 0 || pw[c] < pw[t])) t = c;
      if (t < 0) continue;
      part[x.second] = t;
      pw[b] -= x.first;
      pw[t] += x.first;
    }
  }
}

int64_t km1_of(const HostHypergraph& hg, int k, const std::vector<int32_t>& part) {
  std::vector<uint32_t> stamp(k, UINT32_MAX);
  int64_t s = 0;
  for (uint32_t e = 0; e < hg.m; ++e) {
    int l = 0;
    for (uint64_t i = hg.noff[e]; i < hg.noff[e + 1]; ++i) {
      const int b = part[hg.pins[i]];
      if (stamp[b] != e) { stamp[b] = e; ++l; }
    }
    s += int64_t(std::max(l - 1, 0)) * hg.nw[e];
  }
  return s;
}

}  // namespace

std::vector<int32_t> initial_partition(const HostHypergraph& hg, const InitialPartitioningParams& prm,
                                       int64_t* km1_out) {
  const int k = prm.k;
  std::vector<int32_t> part(hg.n, 0);
  Sub root;
  root.n = hg.n;
  root.m = hg.m;
  root.noff = hg.noff;
  root.pins = hg.pins;
  root.vw = hg.vw;
  root.nw = hg.nw;
  root.orig.resize(hg.n);
  std::iota(root.orig.begin(), root.orig.end(), 0u);
  root.finish();
  const int64_t perfect = (root.W + k - 1) / std::max(k, 1);
  double eps = prm.epsilon;
  if (eps < 0) eps = perfect > 0 ? double(prm.max_w[0]) / double(perfect) - 1.0 : 0.0;
  int levels = 0;
  while ((1 << levels) < k) ++levels;
  const double eps_level = levels > 0 ? std::pow(1.0 + std::max(eps, 0.0), 1.0 / levels) - 1.0 : eps;
  Recursion rec{prm, eps_level, perfect, part};
  rec.run(root, k, 0);
  repair_balance(hg, k, prm.max_w, part);
  if (km1_out) *km1_out = km1_of(hg, k, part);
  return part;
}

}  // namespace mt_kahypar_amd
