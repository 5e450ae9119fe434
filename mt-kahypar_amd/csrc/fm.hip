// Multitry localized k-way FM with a gain cache and global rollback.
//
// Reference: MultiTryKWayFM::refineImpl / roundInitialization
// (refinement/fm/multitry_kway_fm.cpp:30-187), LocalizedKWayFM::findMoves /
// internalFindMoves / acquireOrUpdateNeighbors (localized_kway_fm_core.cpp:
// 26-190), GainCacheStrategy (fm/strategies/gain_cache_strategy.h:72-160),
// StopRule (fm/stop_rule.h:25-60), GlobalRollback::recalculateGains and
// BalanceAndBestIndexScan (global_rollback.cpp:29-141, 146-283), the gain
// cache of PartitionedHypergraph (partitioned_hypergraph.h:647-687,707-809).
//
// Synchronous form.  A round orders the border vertices by a hash and gives
// `seeds` of them to each search; the searches run in `sub_rounds` groups of
// consecutive searches, each group on the state the earlier groups left,
// with the gain cache of that state (benefit, incident weight, connectivity
// weight per block).  A search is one wave and runs on that state plus a
// delta of its own
// (pin-count changes of at most 256 nets, at most 4 blocks per net, in LDS;
// block weight changes): its queue (one entry per lane, at most 64) holds
// vertices keyed by their exact gain on the delta state
// (computeBestTargetBlock: least penalty, ties to the lighter block, the
// block bound respected), ordered by (gain, hash, id).  The best entry moves;
// then every entry is re-keyed on the new state and the pins of the moved
// vertex's nets whose gains changed (fewer than `ignore` pins, pin count in
// the source block 0 or 1 or in the target block 1 or 2 after the move) join
// (the best 64 stay).  The adaptive stopping rule and the best local prefix
// follow the reference.  A search stops after 64 moves, when its delta would
// overflow, or when an expansion would see more than `expand_cap` pins.  The
// best prefixes of a group are concatenated in search order, a vertex already
// moved by an earlier search of the group dropped; the sequence's gains are
// recalculated exactly, and the best balanced prefix is applied.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace mkp {

constexpr uint64_t kFmSeedTag = 0xF5000000ull;
constexpr uint64_t kFmTieTag = 0xF6000000ull;
constexpr uint32_t kFmEmpty = 0xffffffffu;
constexpr int kFmNetBlocks = 4;     // blocks a search may change in one net
constexpr int kFmCorr = 8;          // corrected blocks a queue key may take
constexpr int kFmDeltaSlots = 512;  // LDS net records per search (256 nets at most)
constexpr uint32_t kFmDeltaNets = 256;
constexpr int kFmSeenSlots = 2048;  // LDS set of the vertices of one expansion
constexpr uint32_t kFmExpandCap = 1024;
constexpr int kFmMaxMoves = 64;
constexpr int kFmWaves = 4;         // searches per block
constexpr int kFmMaxK = 128;

struct FmArgs {
  const uint32_t* voff;
  const uint32_t* inc;
  const uint32_t* noff;
  const uint32_t* pins;
  const int32_t* vw;
  const int32_t* nw;
  const int32_t* part;
  const int64_t* part_w;
  const uint64_t* pc;
  PcLayout L;
  int k;
  const int64_t* benefit;
  const int64_t* incw;
  const int32_t* connw;  // n * k
  const int64_t* max_w;
  const uint32_t* order;  // border vertices in seed order
  uint32_t nb;
  uint32_t seeds;
  uint32_t s0, s1;  // searches of the group
  uint64_t seed;
  uint64_t tie;
  uint32_t ignore;
  double beta;  // log(n) of the stopping rule
  uint32_t* out_len;
  uint32_t* out_v;
  int32_t* out_from;
  int32_t* out_to;
};

__device__ inline uint32_t fm_pc(const FmArgs& a, uint32_t e, int b) {
  const int wi = b / a.L.epw, sh = (b - wi * a.L.epw) * a.L.b;
  return uint32_t((a.pc[uint64_t(e) * a.L.stride + wi] >> sh) & a.L.mask);
}

// ------------------------------------------------------------ gain cache --
// benefit(u) = sum of w(e) over nets with exactly one pin of u's block,
// incw(u) = sum of w(e), connw(u, b) = sum of w(e) over nets connected to b
// (initializeGainCache, partitioned_hypergraph.h:707-809)
__global__ void __launch_bounds__(256)
fm_cache_kernel(FmArgs a, uint32_t n, int64_t* __restrict__ benefit, int64_t* __restrict__ incw,
                int32_t* __restrict__ connw) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    int32_t* row = connw + uint64_t(u) * a.k;
    for (int b = 0; b < a.k; ++b) row[b] = 0;
    const int f = a.part[u];
    int64_t ben = 0, iw = 0;
    for (uint32_t j = a.voff[u]; j < a.voff[u + 1]; ++j) {
      const uint32_t e = a.inc[j];
      const int32_t w = a.nw[e];
      iw += w;
      if (fm_pc(a, e, f) == 1) ben += w;
      for (int cw = 0; cw < a.L.vcs; ++cw) {
        uint64_t bits = a.pc[uint64_t(e) * a.L.stride + a.L.coff + cw];
        while (bits) {
          const int b = cw * 64 + __ffsll(static_cast<long long>(bits)) - 1;
          bits &= bits - 1;
          row[b] += w;
        }
      }
    }
    benefit[u] = ben;
    incw[u] = iw;
  }
}

// border vertices (isBorderNode, partitioned_hypergraph.h:605-620) with their
// seed keys
__global__ void fm_border_kernel(FmArgs a, uint32_t n, uint32_t round, uint32_t* __restrict__ flag,
                                 uint64_t* __restrict__ key) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    bool border = false;
    const uint32_t jb = a.voff[u], je = a.voff[u + 1];
    if (je - jb <= 100000)
      for (uint32_t j = jb; j < je && !border; ++j) {
        const uint32_t e = a.inc[j];
        int lam = 0;
        for (int cw = 0; cw < a.L.vcs; ++cw) lam += __popcll(a.pc[uint64_t(e) * a.L.stride + a.L.coff + cw]);
        border = lam > 1;
      }
    flag[u] = border ? 1u : 0u;
    key[u] = hash4(a.seed, kFmSeedTag | round, u, 0);
  }
}

__global__ void fm_compact_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                  const uint64_t* __restrict__ key, uint32_t n, uint64_t* __restrict__ okey,
                                  uint32_t* __restrict__ ov) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    if (flag[u]) {
      okey[pos[u]] = key[u];
      ov[pos[u]] = u;
    }
}

// ---------------------------------------------------------------- search --
struct FmEntry {
  uint32_t v;
  int64_t key;
  int32_t to;
  uint64_t h;
};
__device__ inline bool fm_better(const FmEntry& x, const FmEntry& y) {
  if (x.v == kFmEmpty) return false;
  if (y.v == kFmEmpty) return true;
  if (x.key != y.key) return x.key > y.key;
  if (x.h != y.h) return x.h > y.h;
  return x.v < y.v;
}
__device__ inline FmEntry fm_shfl(const FmEntry& x, int src) {
  FmEntry r;
  r.v = __shfl(x.v, src, 64);
  r.key = __shfl(x.key, src, 64);
  r.to = __shfl(x.to, src, 64);
  r.h = __shfl(x.h, src, 64);
  return r;
}

// per-search LDS state
struct FmLds {
  uint32_t dkey[kFmDeltaSlots];   // net id or kFmEmpty
  uint32_t dblk[kFmDeltaSlots];   // 4 block ids, one byte each (0xff: none)
  int16_t dval[kFmDeltaSlots * 4];
  uint32_t seen[kFmSeenSlots];
  int32_t dW[kFmMaxK];
  FmEntry tmp[64];
};

__device__ inline uint32_t fm_dslot0(uint32_t e) { return (e * 0x9E3779B1u) >> (32 - 9); }
static_assert(kFmDeltaSlots == 512, "fm_dslot0 hashes into 2^9 slots");
static_assert(kFmMaxK <= 256, "block pairs are staged as two bytes");

// slot of net e in the delta or -1
__device__ inline int fm_dfind(const FmLds& s, uint32_t e) {
  uint32_t h = fm_dslot0(e);
  while (true) {
    const uint32_t k = s.dkey[h];
    if (k == e) return int(h);
    if (k == kFmEmpty) return -1;
    h = (h + 1) & (kFmDeltaSlots - 1);
  }
}
__device__ inline int fm_dget(const FmLds& s, int slot, int b) {
  if (slot < 0) return 0;
  const uint32_t bl = s.dblk[slot];
#pragma unroll
  for (int i = 0; i < kFmNetBlocks; ++i)
    if (((bl >> (8 * i)) & 0xffu) == uint32_t(b)) return s.dval[slot * 4 + i];
  return 0;
}

// exact key of an unmoved vertex on the delta state (the search's W via dW)
__device__ inline bool fm_key(const FmArgs& a, const FmLds& s, uint32_t y, FmEntry& out) {
  const int f = a.part[y];
  int64_t ben = a.benefit[y];
  int cb[kFmCorr];
  int64_t cwt[kFmCorr];
  int nc = 0;
#pragma unroll
  for (int i = 0; i < kFmCorr; ++i) { cb[i] = -1; cwt[i] = 0; }
  bool over = false;
  // the incident nets four at a time (their loads issued together); only
  // nets in the delta contribute
  const uint32_t jb = a.voff[y], je = a.voff[y + 1];
  for (uint32_t j0 = jb; j0 < je; j0 += 4) {
    uint32_t ee[4];
    int sl[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) ee[t] = j0 + t < je ? a.inc[j0 + t] : kFmEmpty;
#pragma unroll
    for (int t = 0; t < 4; ++t) sl[t] = ee[t] != kFmEmpty ? fm_dfind(s, ee[t]) : -1;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
    const uint32_t e = ee[t];
    const int slot = sl[t];
    if (slot < 0) continue;
    const int64_t w = a.nw[e];
    const int pf = int(fm_pc(a, e, f));
    const int df = fm_dget(s, slot, f);
    ben += w * (int64_t(pf + df == 1 ? 1 : 0) - int64_t(pf == 1 ? 1 : 0));
    const uint32_t bl = s.dblk[slot];
#pragma unroll
    for (int i = 0; i < kFmNetBlocks; ++i) {
      const uint32_t b = (bl >> (8 * i)) & 0xffu;
      if (b == 0xffu || int(b) == f) continue;
      const int pb = int(fm_pc(a, e, int(b)));
      const int d = s.dval[slot * 4 + i];
      const int ch = (pb + d >= 1 ? 1 : 0) - (pb >= 1 ? 1 : 0);
      if (!ch) continue;
      bool found = false;
#pragma unroll
      for (int c = 0; c < kFmCorr; ++c)
        if (cb[c] == int(b)) { cwt[c] += w * ch; found = true; }
      if (!found) {
        if (nc == kFmCorr) { over = true; continue; }
#pragma unroll
        for (int c = 0; c < kFmCorr; ++c)
          if (c == nc) { cb[c] = int(b); cwt[c] = w * ch; }
        ++nc;
      }
    }
    }
  }
  if (over) return false;
  const int64_t wy = a.vw[y];
  int to = -1;
  int64_t best_pen = INT64_MAX, best_tw = a.part_w[f] + s.dW[f] - wy;
  const int32_t* row = a.connw + uint64_t(y) * a.k;
  for (int b = 0; b < a.k; ++b) {
    if (b == f) continue;
    int64_t corr = 0;
#pragma unroll
    for (int c = 0; c < kFmCorr; ++c)
      if (cb[c] == b) corr = cwt[c];
    const int64_t pen = a.incw[y] - (int64_t(row[b]) + corr);
    const int64_t tw = a.part_w[b] + s.dW[b];
    if ((pen < best_pen || (pen == best_pen && tw < best_tw)) && tw + wy <= a.max_w[b]) {
      best_pen = pen;
      to = b;
      best_tw = tw;
    }
  }
  if (to < 0) return false;
  out.v = y;
  out.key = ben - best_pen;
  out.to = to;
  out.h = hash4(a.seed, a.tie, y, 0);
  return true;
}

// queue := the best 64 of queue + candidates (one of each per lane; no
// vertex in both); every lane calls it
__device__ inline FmEntry fm_merge(FmEntry q, const FmEntry& c, FmLds& s) {
  const int lane = threadIdx.x & 63;
  if (!__ballot(c.v != kFmEmpty)) return q;
  int rq = 0, rc = 0;  // entries better than q / c
  for (int j = 0; j < 64; ++j) {
    const FmEntry qj = fm_shfl(q, j), cj = fm_shfl(c, j);
    rq += (fm_better(qj, q) ? 1 : 0) + (fm_better(cj, q) ? 1 : 0);
    rc += (fm_better(qj, c) ? 1 : 0) + (fm_better(cj, c) ? 1 : 0);
  }
  if (q.v != kFmEmpty && rq < 64) s.tmp[rq] = q;
  if (c.v != kFmEmpty && rc < 64) s.tmp[rc] = c;
  // number of valid entries kept
  const int nq = __popcll(__ballot(q.v != kFmEmpty)) + __popcll(__ballot(c.v != kFmEmpty));
  lds_wave_sync();
  FmEntry r;
  if (lane < min(nq, 64)) {
    r = s.tmp[lane];
  } else {
    r.v = kFmEmpty; r.key = 0; r.to = -1; r.h = 0;
  }
  lds_wave_sync();
  return r;
}

// stop_rule.h:25-60 (every lane holds the same values)
struct FmStop {
  uint64_t steps = 0;
  double variance = 0.0, Mk = 0.0, Mk_prev = 0.0, Sk = 0.0, Sk_prev = 0.0;
  __device__ bool should_stop(double beta) const {
    return (double(steps) > beta) && (Mk == 0 || double(steps) >= (variance / (Mk * Mk)) * (1.0 / 2.0 - 0.25));
  }
  __device__ void update(int64_t gain) {
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
  __device__ void reset() { steps = 0; variance = 0.0; }
};

__device__ inline int64_t fm_wsum(int64_t x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__global__ void __launch_bounds__(64 * kFmWaves)
fm_search_kernel(FmArgs a) {
  __shared__ FmLds lds[kFmWaves];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  FmLds& s = lds[wid];
  for (uint32_t sr = a.s0 + blockIdx.x * kFmWaves + wid; sr < a.s1; sr += gridDim.x * kFmWaves) {
    for (int i = lane; i < kFmDeltaSlots; i += 64) {
      s.dkey[i] = kFmEmpty;
      s.dblk[i] = 0xffffffffu;
    }
    for (int b = lane; b < a.k; b += 64) s.dW[b] = 0;
    lds_wave_sync();
    uint32_t dnets = 0;
    // seeds
    FmEntry q;
    q.v = kFmEmpty; q.key = 0; q.to = -1; q.h = 0;
    {
      FmEntry c;
      c.v = kFmEmpty; c.key = 0; c.to = -1; c.h = 0;
      const uint32_t i = sr * a.seeds + uint32_t(lane);
      if (lane < int(a.seeds) && i < a.nb) {
        FmEntry x;
        if (fm_key(a, s, a.order[i], x)) c = x;
      }
      q = fm_merge(q, c, s);
    }
    // moves: lane i holds move i
    uint32_t mv = kFmEmpty;
    int32_t mfrom = 0, mto = 0;
    int nmoves = 0;
    FmStop stop;
    int64_t est = 0, best = 0;
    int best_idx = 0;
    while (!stop.should_stop(a.beta) && nmoves < kFmMaxMoves) {
      if (!__ballot(q.v != kFmEmpty)) break;
      // the best entry
      FmEntry top = q;
      int owner = lane;
      for (int o = 32; o > 0; o >>= 1) {
        const FmEntry other = fm_shfl(top, lane ^ o);
        const int oo = __shfl_xor(owner, o, 64);
        if (fm_better(other, top) || (!fm_better(top, other) && oo < owner)) { top = other; owner = oo; }
      }
      top = fm_shfl(top, 0);
      owner = __shfl(owner, 0, 64);
      const uint32_t v = top.v;
      const int from = a.part[v], to = top.to;
      const uint32_t jb = a.voff[v], je = a.voff[v + 1];
      // capacity of the delta
      uint32_t newn = 0;
      bool fits = true;
      for (uint32_t j = jb + lane; j < je; j += 64) {
        const uint32_t e = a.inc[j];
        const int slot = fm_dfind(s, e);
        if (slot < 0) { ++newn; continue; }
        const uint32_t bl = s.dblk[slot];
        int used = 0;
        bool hf = false, ht = false;
#pragma unroll
        for (int i = 0; i < kFmNetBlocks; ++i) {
          const uint32_t b = (bl >> (8 * i)) & 0xffu;
          if (b != 0xffu) ++used;
          hf |= b == uint32_t(from);
          ht |= b == uint32_t(to);
        }
        if (used + (hf ? 0 : 1) + (ht ? 0 : 1) > kFmNetBlocks) fits = false;
      }
      newn = uint32_t(fm_wsum(int64_t(newn)));
      if (__ballot(!fits) || dnets + newn > kFmDeltaNets) break;
      // weights before the move
      int64_t hv = INT64_MIN;
      for (int b = lane; b < a.k; b += 64) hv = max(hv, a.part_w[b] + s.dW[b]);
      for (int o = 32; o > 0; o >>= 1) hv = max(hv, __shfl_xor(hv, o, 64));
      const int64_t from_w = a.part_w[from] + s.dW[from], to_w = a.part_w[to] + s.dW[to];
      const int64_t wv = a.vw[v];
      lds_wave_sync();
      // apply to the delta (the nets of v are distinct: one lane per record)
      for (uint32_t j = jb + lane; j < je; j += 64) {
        const uint32_t e = a.inc[j];
        uint32_t h = fm_dslot0(e);
        int slot = -1;
        while (true) {
          const uint32_t old = atomicCAS(&s.dkey[h], kFmEmpty, e);
          if (old == kFmEmpty || old == e) { slot = int(h); break; }
          h = (h + 1) & (kFmDeltaSlots - 1);
        }
        uint32_t bl = s.dblk[slot];
        int fi = -1, ti = -1, freei = 0;
#pragma unroll
        for (int i = kFmNetBlocks - 1; i >= 0; --i) {
          const uint32_t b = (bl >> (8 * i)) & 0xffu;
          if (b == uint32_t(from)) fi = i;
          if (b == uint32_t(to)) ti = i;
        }
        // blocks are appended in order: the first free byte follows the used ones
#pragma unroll
        for (int i = 0; i < kFmNetBlocks; ++i)
          if (((bl >> (8 * i)) & 0xffu) != 0xffu) freei = i + 1;
        if (fi < 0) {
          fi = freei++;
          bl = (bl & ~(0xffu << (8 * fi))) | (uint32_t(from) << (8 * fi));
          s.dval[slot * 4 + fi] = 0;
        }
        if (ti < 0) {
          ti = freei++;
          bl = (bl & ~(0xffu << (8 * ti))) | (uint32_t(to) << (8 * ti));
          s.dval[slot * 4 + ti] = 0;
        }
        s.dblk[slot] = bl;
        s.dval[slot * 4 + fi] = int16_t(s.dval[slot * 4 + fi] - 1);
        s.dval[slot * 4 + ti] = int16_t(s.dval[slot * 4 + ti] + 1);
      }
      if (lane == 0) {
        s.dW[from] -= int32_t(wv);
        s.dW[to] += int32_t(wv);
      }
      dnets += newn;
      if (lane == nmoves) {
        mv = v;
        mfrom = from;
        mto = to;
      }
      ++nmoves;
      if (lane == owner) q.v = kFmEmpty;
      lds_wave_sync();
      est += top.key;
      stop.update(top.key);
      if (est > best || (est >= best && from_w == hv && to_w + wv < hv)) {
        stop.reset();
        best = est;
        best_idx = nmoves;
      }
      // exact keys of the queue on the new state
      if (q.v != kFmEmpty) {
        FmEntry x;
        if (fm_key(a, s, q.v, x)) q = x;
        else q.v = kFmEmpty;
      }
      // nets whose gains changed
      uint64_t seen = 0;
      for (uint32_t j = jb + lane; j < je; j += 64) {
        const uint32_t e = a.inc[j];
        const uint32_t sz = a.noff[e + 1] - a.noff[e];
        if (sz >= a.ignore) continue;
        const int slot = fm_dfind(s, e);
        const int cf = int(fm_pc(a, e, from)) + fm_dget(s, slot, from);
        const int ct = int(fm_pc(a, e, to)) + fm_dget(s, slot, to);
        if (cf == 0 || cf == 1 || ct == 1 || ct == 2) seen += sz;
      }
      seen = uint64_t(fm_wsum(int64_t(seen)));
      if (seen > kFmExpandCap) break;
      // the vertices of the expansion: queue and moved vertices first
      for (int i = lane; i < kFmSeenSlots; i += 64) s.seen[i] = kFmEmpty;
      lds_wave_sync();
      auto seen_insert = [&](uint32_t y) -> bool {
        uint32_t h = (y * 0x9E3779B1u) >> (32 - 11);
        while (true) {
          const uint32_t old = atomicCAS(&s.seen[h], kFmEmpty, y);
          if (old == kFmEmpty) return true;
          if (old == y) return false;
          h = (h + 1) & (kFmSeenSlots - 1);
        }
      };
      if (q.v != kFmEmpty) seen_insert(q.v);
      if (lane < nmoves) seen_insert(mv);
      lds_wave_sync();
      for (uint32_t j0 = jb; j0 < je; j0 += 64) {
        const uint32_t j = j0 + uint32_t(lane);
        bool ch = false;
        uint32_t e = 0;
        if (j < je) {
          e = a.inc[j];
          const uint32_t sz = a.noff[e + 1] - a.noff[e];
          if (sz < a.ignore) {
            const int slot = fm_dfind(s, e);
            const int cf = int(fm_pc(a, e, from)) + fm_dget(s, slot, from);
            const int ct = int(fm_pc(a, e, to)) + fm_dget(s, slot, to);
            ch = cf == 0 || cf == 1 || ct == 1 || ct == 2;
          }
        }
        uint64_t nets = __ballot(ch);
        while (nets) {
          const int src = __ffsll(static_cast<long long>(nets)) - 1;
          nets &= nets - 1;
          const uint32_t ne = __shfl(e, src, 64);
          const uint32_t pb = a.noff[ne], pe = a.noff[ne + 1];
          for (uint32_t p0 = pb; p0 < pe; p0 += 64) {
            FmEntry c;
            c.v = kFmEmpty; c.key = 0; c.to = -1; c.h = 0;
            if (p0 + lane < pe) {
              const uint32_t y = a.pins[p0 + lane];
              if (seen_insert(y)) {
                FmEntry x;
                if (fm_key(a, s, y, x)) c = x;
              }
            }
            q = fm_merge(q, c, s);
          }
        }
      }
    }
    // the best local prefix
    if (lane < best_idx) {
      const uint64_t o = uint64_t(sr) * kFmMaxMoves + uint64_t(lane);
      a.out_v[o] = mv;
      a.out_from[o] = mfrom;
      a.out_to[o] = mto;
    }
    if (lane == 0) a.out_len[sr] = uint32_t(best_idx);
    lds_wave_sync();
  }
}

// ------------------------------------------------------------- rollback --
// first occurrence of every vertex over the searches' prefixes (position
// s * 64 + i) wins
// slots x of the group's searches (x = s * 64 + i, relative to the group's
// first slot x0)
__global__ void fm_owner_kernel(const uint32_t* __restrict__ out_len, const uint32_t* __restrict__ out_v,
                                uint64_t x0, uint64_t total, uint32_t* __restrict__ owner) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < total; r += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t x = x0 + r;
    if (uint32_t(x % kFmMaxMoves) < out_len[x / kFmMaxMoves]) atomicMin(&owner[out_v[x]], uint32_t(r));
  }
}
__global__ void fm_keep_kernel(const uint32_t* __restrict__ out_len, const uint32_t* __restrict__ out_v,
                               uint64_t x0, uint64_t total, const uint32_t* __restrict__ owner,
                               uint32_t* __restrict__ keep) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < total; r += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t x = x0 + r;
    keep[r] = (uint32_t(x % kFmMaxMoves) < out_len[x / kFmMaxMoves] && owner[out_v[x]] == uint32_t(r)) ? 1u : 0u;
  }
}
// the sequence in position order; move ids are 1-based (fm_commons.h:40-45)
__global__ void fm_sequence_kernel(const uint32_t* __restrict__ out_v, const int32_t* __restrict__ out_from,
                                   const int32_t* __restrict__ out_to, const uint32_t* __restrict__ keep,
                                   const uint32_t* __restrict__ pos, uint64_t x0, uint64_t total,
                                   uint32_t* __restrict__ sv, int32_t* __restrict__ sf, int32_t* __restrict__ st,
                                   uint32_t* __restrict__ move_of, uint32_t* __restrict__ owner) {
  for (uint64_t r = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; r < total; r += uint64_t(gridDim.x) * blockDim.x)
    if (keep[r]) {
      const uint64_t x = x0 + r;
      const uint32_t i = pos[r];
      sv[i] = out_v[x];
      sf[i] = out_from[x];
      st[i] = out_to[x];
      move_of[out_v[x]] = i + 1;
      owner[out_v[x]] = kFmEmpty;
    }
}

// nets of the moved vertices, each once
__global__ void fm_nets_kernel(const uint32_t* __restrict__ voff, const uint32_t* __restrict__ inc,
                               const uint32_t* __restrict__ sv, uint32_t nm, uint32_t stamp,
                               uint32_t* __restrict__ nmark, uint32_t* __restrict__ nlist,
                               uint32_t* __restrict__ ncount) {
  const int lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < nm; i += waves) {
    const uint32_t v = sv[i];
    for (uint32_t j0 = voff[v]; j0 < voff[v + 1]; j0 += 64) {
      const uint32_t j = j0 + uint32_t(lane);
      bool add = false;
      uint32_t e = 0;
      if (j < voff[v + 1]) {
        e = inc[j];
        add = nmark[e] != stamp && atomicExch(&nmark[e], stamp) != stamp;
      }
      const uint64_t bal = __ballot(add);
      if (!bal) continue;
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(ncount, uint32_t(__popcll(bal)));
      base = __shfl(base, 0, 64);
      if (add) nlist[base + __popcll(bal & ((1ull << lane) - 1))] = e;
    }
  }
}

// recalculateGains (global_rollback.cpp:207-283) per net, one wave per net:
// a pass over the pins collects, per block, the first move into it and the
// last move out of it and the moved pins leaving it (LDS); the unmoved pins of
// block b are its pin count at the group start minus the pins leaving it.  A
// second pass gives every moved pin (id, from f, to t)
//   benefit if it is the last to leave f, none enters f before it and no
//   unmoved pin is in f;
//   penalty if it is the first to enter t, every pin leaving t leaves before
//   it and no unmoved pin is in t.
__global__ void __launch_bounds__(256)
fm_recalc_kernel(FmArgs a, const uint32_t* __restrict__ nlist, const uint32_t* __restrict__ ncount,
                 const uint32_t* __restrict__ move_of, const int32_t* __restrict__ sf, const int32_t* __restrict__ st,
                 int64_t* __restrict__ gain) {
  __shared__ uint32_t s_in[4][kFmMaxK], s_out[4][kFmMaxK], s_left[4][kFmMaxK];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t* first_in = s_in[wid];
  uint32_t* last_out = s_out[wid];
  uint32_t* left = s_left[wid];
  for (int b = lane; b < a.k; b += 64) {
    first_in[b] = 0xffffffffu;
    last_out[b] = 0;
    left[b] = 0;
  }
  lds_wave_sync();
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t cnt = *ncount;
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + wid; i < cnt; i += waves) {
    const uint32_t e = nlist[i];
    const uint32_t pb = a.noff[e], pe = a.noff[e + 1];
    const int64_t w = a.nw[e];
    for (uint32_t q0 = pb; q0 < pe; q0 += 64) {
      if (q0 + lane < pe) {
        const uint32_t id = move_of[a.pins[q0 + lane]];
        if (id) {
          atomicMin(&first_in[st[id - 1]], id);
          atomicMax(&last_out[sf[id - 1]], id);
          atomicAdd(&left[sf[id - 1]], 1u);
        }
      }
    }
    lds_wave_sync();
    for (uint32_t q0 = pb; q0 < pe; q0 += 64) {
      if (q0 + lane < pe) {
        const uint32_t id = move_of[a.pins[q0 + lane]];
        if (id) {
          const int f = sf[id - 1], t = st[id - 1];
          const bool ben = last_out[f] == id && first_in[f] > id && fm_pc(a, e, f) == left[f];
          const bool pen = first_in[t] == id && last_out[t] < id && fm_pc(a, e, t) == left[t];
          if (ben && !pen)
            atomicAdd(reinterpret_cast<unsigned long long*>(&gain[id - 1]), static_cast<unsigned long long>(w));
          if (!ben && pen)
            atomicAdd(reinterpret_cast<unsigned long long*>(&gain[id - 1]), static_cast<unsigned long long>(-w));
        }
      }
    }
    lds_wave_sync();
    // reset the blocks this net touched
    for (uint32_t q0 = pb; q0 < pe; q0 += 64) {
      if (q0 + lane < pe) {
        const uint32_t id = move_of[a.pins[q0 + lane]];
        if (id) {
          first_in[st[id - 1]] = 0xffffffffu;
          last_out[sf[id - 1]] = 0;
          left[sf[id - 1]] = 0;
        }
      }
    }
    lds_wave_sync();
  }
}

// BalanceAndBestIndexScan (global_rollback.cpp:29-141) in sequential form:
// one wave stages 1024 moves at a time in LDS (block pair, vertex weight,
// gain), then one lane walks them, keeping the heaviest block weight up to
// date (rescanned only when the heaviest block loses weight): the prefix with
// the largest gain sum and no block above its bound, ties to the lighter
// heaviest block.  res[0] = length, res[1] = gain.
constexpr uint32_t kFmPrefixChunk = 1024;
__global__ void __launch_bounds__(64)
fm_prefix_kernel(FmArgs a, const uint32_t* __restrict__ sv, const int32_t* __restrict__ sf,
                 const int32_t* __restrict__ st, const int64_t* __restrict__ gain, uint32_t nm,
                 int64_t* __restrict__ res) {
  __shared__ int64_t w[kFmMaxK], mw[kFmMaxK];
  __shared__ int64_t c_gain[kFmPrefixChunk], c_w[kFmPrefixChunk];
  __shared__ uint16_t c_ft[kFmPrefixChunk];
  const int lane = threadIdx.x;
  int over = 0;
  int64_t hv = INT64_MIN, bg = 0, bh = 0, sum = 0;
  uint32_t bi = 0;
  if (lane == 0) {
    for (int b = 0; b < a.k; ++b) {
      w[b] = a.part_w[b];
      mw[b] = a.max_w[b];
      over += w[b] > mw[b] ? 1 : 0;
      hv = max(hv, w[b]);
    }
    bh = hv;
  }
  for (uint32_t c0 = 0; c0 < nm; c0 += kFmPrefixChunk) {
    const uint32_t cn = min(kFmPrefixChunk, nm - c0);
    __syncthreads();
    for (uint32_t i = lane; i < cn; i += 64) {
      c_gain[i] = gain[c0 + i];
      c_w[i] = a.vw[sv[c0 + i]];
      c_ft[i] = uint16_t((uint32_t(sf[c0 + i]) << 8) | uint32_t(st[c0 + i]));
    }
    __syncthreads();
    if (lane != 0) continue;
    for (uint32_t i = 0; i < cn; ++i) {
      const int f = c_ft[i] >> 8, t = c_ft[i] & 0xff;
      const int64_t wv = c_w[i];
      sum += c_gain[i];
      const int64_t wf = w[f], wt = w[t];
      if (wf > mw[f] && wf - wv <= mw[f]) --over;
      if (wt <= mw[t] && wt + wv > mw[t]) ++over;
      w[f] = wf - wv;
      w[t] = wt + wv;
      if (wt + wv > hv) {
        hv = wt + wv;
      } else if (wf == hv && wv > 0) {
        hv = INT64_MIN;
        for (int b = 0; b < a.k; ++b) hv = max(hv, w[b]);
      }
      if (over == 0 && sum >= bg && (sum > bg || hv < bh)) {
        bg = sum;
        bh = hv;
        bi = c0 + i + 1;
      }
    }
  }
  if (lane == 0) {
    res[0] = bi;
    res[1] = bg;
  }
}

// applies moves [0, count) of the sequence: pin counts, connectivity,
// blocks and block weights (changeNodePart, partitioned_hypergraph.h:539-563),
// and keeps the gain cache (gainCacheUpdate, partitioned_hypergraph.h:
// 1016-1047): a block entering or leaving a net changes the connectivity
// weight of every pin of the net; a pin count of the source or target block
// crossing between 1 and 2 marks the net, whose pins then get their benefit
// recomputed (fm_benefit_kernel) on the final state, as does every moved
// vertex.  The crossings of one net are ordered by its atomics, so the
// connectivity weights end exact in any order of the moves.
__global__ void fm_apply_kernel(FmArgs a, int32_t* __restrict__ part, int64_t* __restrict__ part_w,
                                uint64_t* __restrict__ pc, int32_t* __restrict__ connw,
                                const uint32_t* __restrict__ sv, const int32_t* __restrict__ sf,
                                const int32_t* __restrict__ st, uint32_t count, uint32_t stamp,
                                uint32_t* __restrict__ vdirty) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t v = sv[i];
    const int from = sf[i], to = st[i];
    const int fwi = from / a.L.epw, fsh = (from - fwi * a.L.epw) * a.L.b;
    const int twi = to / a.L.epw, tsh = (to - twi * a.L.epw) * a.L.b;
    for (uint32_t j = a.voff[v]; j < a.voff[v + 1]; ++j) {
      const uint32_t e = a.inc[j];
      unsigned long long* rec = reinterpret_cast<unsigned long long*>(pc + uint64_t(e) * a.L.stride);
      const uint64_t of = (atomicAdd(rec + fwi, static_cast<unsigned long long>(-(int64_t(1) << fsh))) >> fsh) & a.L.mask;
      const uint64_t ot = (atomicAdd(rec + twi, 1ull << tsh) >> tsh) & a.L.mask;
      if (of == 1) atomicXor(rec + a.L.coff + from / 64, 1ull << (from % 64));
      if (ot == 0) atomicXor(rec + a.L.coff + to / 64, 1ull << (to % 64));
      const int32_t w = a.nw[e];
      if (of == 1 || ot == 0 || of == 2 || ot == 1) {
        for (uint32_t q = a.noff[e]; q < a.noff[e + 1]; ++q) {
          const uint32_t y = a.pins[q];
          if (of == 1) atomicSub(&connw[uint64_t(y) * a.k + from], w);
          if (ot == 0) atomicAdd(&connw[uint64_t(y) * a.k + to], w);
          if (of == 2 || ot == 1) vdirty[y] = stamp;
        }
      }
    }
    vdirty[v] = stamp;
    part[v] = to;
    atomicAdd(reinterpret_cast<unsigned long long*>(&part_w[from]), static_cast<unsigned long long>(-int64_t(a.vw[v])));
    atomicAdd(reinterpret_cast<unsigned long long*>(&part_w[to]), static_cast<unsigned long long>(int64_t(a.vw[v])));
  }
}

// benefit of the marked vertices on the current state
__global__ void fm_benefit_kernel(FmArgs a, uint32_t n, const uint32_t* __restrict__ vdirty, uint32_t stamp,
                                  int64_t* __restrict__ benefit) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    if (vdirty[u] != stamp) continue;
    const int f = a.part[u];
    int64_t ben = 0;
    for (uint32_t j = a.voff[u]; j < a.voff[u + 1]; ++j) {
      const uint32_t e = a.inc[j];
      if (fm_pc(a, e, f) == 1) ben += a.nw[e];
    }
    benefit[u] = ben;
  }
}

__global__ void fm_clear_kernel(const uint32_t* __restrict__ sv, uint32_t nm, uint32_t* __restrict__ move_of) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nm; i += gridDim.x * blockDim.x) move_of[sv[i]] = 0;
}

// ------------------------------------------------------------------ host --
struct FmWork {
  DevBuf<int64_t> benefit, incw, gain, res;
  DevBuf<int32_t> connw, out_from, out_to, sf, st;
  DevBuf<uint64_t> key, bkey, skey;
  DevBuf<uint32_t> vdirty, flag, pos, bv, order, out_len, out_v, owner, keep, kpos, sv, move_of, nmark, nlist, ncount;
};

void fm_refine(mkp_partition* p, const mkp_fm_params& prm, mkp_fm_stats* stats) {
  mkp_hypergraph* hg = p->hg;
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t n = hg->n, m = hg->m;
  MKP_CHECK(p->k <= kFmMaxK, "FM supports at most 128 blocks");
  MKP_CHECK(prm.seeds >= 1 && prm.seeds <= 64, "seeds per search must be in [1, 64]");
  mkp_fm_stats st{};
  const auto t0 = std::chrono::steady_clock::now();
  if (n == 0 || m == 0 || prm.rounds <= 0) {
    if (stats) *stats = st;
    return;
  }
  FmWork w;
  w.benefit.alloc(n);
  w.incw.alloc(n);
  w.connw.alloc(uint64_t(n) * p->k);
  w.flag.alloc(n);
  w.key.alloc(n);
  w.pos.alloc(uint64_t(n) + 1);
  w.bkey.alloc(n);
  w.bv.alloc(n);
  w.skey.alloc(n);
  w.order.alloc(n);
  w.owner.alloc(n);
  w.move_of.alloc(n);
  w.nmark.alloc(m);
  w.nlist.alloc(m);
  w.ncount.alloc(1);
  w.res.alloc(2);
  fill_async<uint32_t>(w.owner.p, kFmEmpty, n, s);
  fill_async<uint32_t>(w.move_of.p, 0u, n, s);
  fill_async<uint32_t>(w.nmark.p, 0u, m, s);
  w.vdirty.alloc(n);
  fill_async<uint32_t>(w.vdirty.p, 0u, n, s);
  DevBuf<int64_t> maxw(p->k);
  MKP_HIP(hipMemcpyAsync(maxw.p, prm.max_part_weights, size_t(p->k) * 8, hipMemcpyHostToDevice, s));
  FmArgs a{};
  a.voff = hg->voff.p; a.inc = hg->inc.p; a.noff = hg->noff.p; a.pins = hg->pins.p; a.vw = hg->vw.p;
  a.nw = hg->nw.p; a.part = p->part.p; a.part_w = p->part_w.p; a.pc = p->pc.p; a.L = p->L; a.k = p->k;
  a.benefit = w.benefit.p; a.incw = w.incw.p; a.connw = w.connw.p; a.max_w = maxw.p;
  a.seeds = uint32_t(prm.seeds); a.seed = prm.seed; a.ignore = prm.ignore_net_size;
  a.beta = std::log(double(n));
  int64_t* pin = c.pinned;
  const int S = prm.sub_rounds > 0 ? prm.sub_rounds : 8;
  uint32_t stamp = 0;   // net marks of the groups (nmark starts at 0)
  uint32_t vstamp = 0;  // vertex marks of the applies (vdirty starts at 0)
  // the gain cache of the starting state, kept up to date by every apply
  fm_cache_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(a, n, w.benefit.p, w.incw.p, w.connw.p);
  for (int r = 0; r < prm.rounds; ++r) {
    a.tie = kFmTieTag | uint64_t(r);
    fm_border_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(a, n, uint32_t(r), w.flag.p, w.key.p);
    MKP_LAUNCH_CHECK();
    exclusive_scan_u32(c, w.flag.p, w.pos.p, n);
    MKP_HIP(hipMemcpyAsync(pin, w.pos.p + n, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    const uint32_t nb = uint32_t(reinterpret_cast<uint32_t*>(pin)[0]);
    if (nb == 0) break;
    fm_compact_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(w.flag.p, w.pos.p, w.key.p, n, w.bkey.p, w.bv.p);
    MKP_LAUNCH_CHECK();
    sort_pairs_u64_u32(c, w.bkey.p, w.skey.p, w.bv.p, w.order.p, nb, 0, 64);
    const uint32_t searches = (nb + uint32_t(prm.seeds) - 1) / uint32_t(prm.seeds);
    w.out_len.ensure(searches);
    w.out_v.ensure(uint64_t(searches) * kFmMaxMoves);
    w.out_from.ensure(uint64_t(searches) * kFmMaxMoves);
    w.out_to.ensure(uint64_t(searches) * kFmMaxMoves);
    a.order = w.order.p; a.nb = nb;
    a.out_len = w.out_len.p; a.out_v = w.out_v.p; a.out_from = w.out_from.p; a.out_to = w.out_to.p;
    st.searches += searches;
    st.rounds += 1;
    int64_t round_improvement = 0;
    // group j: searches [j * searches / S, (j + 1) * searches / S) on the
    // state the earlier groups left
    for (int j = 0; j < S; ++j) {
      const uint32_t s0 = uint32_t(uint64_t(j) * searches / uint64_t(S));
      const uint32_t s1 = uint32_t(uint64_t(j + 1) * searches / uint64_t(S));
      if (s0 == s1) continue;
      a.s0 = s0; a.s1 = s1;
      {
        TimedRegion tr(c.prof, s, kFamFm, 0.0, "fm_search_kernel");
        fm_search_kernel<<<grid_for(s1 - s0, kFmWaves, unsigned(c.num_cus) * 8), 64 * kFmWaves, 0, s>>>(a);
      }
      MKP_LAUNCH_CHECK();
      const uint64_t x0 = uint64_t(s0) * kFmMaxMoves, slots = uint64_t(s1 - s0) * kFmMaxMoves;
      w.keep.ensure(slots);
      w.kpos.ensure(slots + 1);
      fm_owner_kernel<<<grid_for(slots, 256, 65536), 256, 0, s>>>(w.out_len.p, w.out_v.p, x0, slots, w.owner.p);
      fm_keep_kernel<<<grid_for(slots, 256, 65536), 256, 0, s>>>(w.out_len.p, w.out_v.p, x0, slots, w.owner.p,
                                                                 w.keep.p);
      MKP_LAUNCH_CHECK();
      exclusive_scan_u32(c, w.keep.p, w.kpos.p, slots);
      MKP_HIP(hipMemcpyAsync(pin, w.kpos.p + slots, 4, hipMemcpyDeviceToHost, s));
      c.sync();
      const uint32_t nm = uint32_t(reinterpret_cast<uint32_t*>(pin)[0]);
      if (nm == 0) continue;
      w.sv.ensure(nm);
      w.sf.ensure(nm);
      w.st.ensure(nm);
      w.gain.ensure(nm);
      fm_sequence_kernel<<<grid_for(slots, 256, 65536), 256, 0, s>>>(w.out_v.p, w.out_from.p, w.out_to.p,
                                                                     w.keep.p, w.kpos.p, x0, slots, w.sv.p, w.sf.p,
                                                                     w.st.p, w.move_of.p, w.owner.p);
      fill_async<int64_t>(w.gain.p, 0, nm, s);
      fill_async<uint32_t>(w.ncount.p, 0u, 1, s);
      fm_nets_kernel<<<grid_for(nm, 4, 65536), 256, 0, s>>>(hg->voff.p, hg->inc.p, w.sv.p, nm, ++stamp, w.nmark.p,
                                                            w.nlist.p, w.ncount.p);
      {
        TimedRegion tr(c.prof, s, kFamFm, 0.0, "fm_recalc_kernel");
        fm_recalc_kernel<<<unsigned(c.num_cus) * 8, 256, 0, s>>>(a, w.nlist.p, w.ncount.p, w.move_of.p, w.sf.p,
                                                                 w.st.p, w.gain.p);
      }
      fm_prefix_kernel<<<1, 64, 0, s>>>(a, w.sv.p, w.sf.p, w.st.p, w.gain.p, nm, w.res.p);
      fm_clear_kernel<<<grid_for(nm, 256, 65536), 256, 0, s>>>(w.sv.p, nm, w.move_of.p);
      MKP_LAUNCH_CHECK();
      MKP_HIP(hipMemcpyAsync(pin, w.res.p, 16, hipMemcpyDeviceToHost, s));
      c.sync();
      const uint32_t keep = uint32_t(pin[0]);
      const int64_t improvement = pin[1];
      if (keep) {
        ++vstamp;
        fm_apply_kernel<<<grid_for(keep, 256, 65536), 256, 0, s>>>(a, p->part.p, p->part_w.p, p->pc.p, w.connw.p,
                                                                    w.sv.p, w.sf.p, w.st.p, keep, vstamp,
                                                                    w.vdirty.p);
        fm_benefit_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(a, n, w.vdirty.p, vstamp, w.benefit.p);
        MKP_LAUNCH_CHECK();
      }
      st.moves += keep;
      round_improvement += improvement;
    }
    st.improvement += round_improvement;
    if (round_improvement <= 0) break;
  }
  c.sync();
  st.device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
}

}  // namespace mkp

using namespace mkp;

extern "C" int mkp_fm_refine(mkp_partition* p, const mkp_fm_params* prm, mkp_fm_stats* stats) {
  return guarded([&] {
    MKP_CHECK(p && prm && prm->max_part_weights, "null argument");
    MKP_CHECK(prm->rounds >= 0 && prm->rounds <= 1000, "rounds must be in [0, 1000]");
    MKP_CHECK(prm->ignore_net_size >= 1, "ignore_net_size must be at least 1");
    MKP_CHECK(prm->sub_rounds >= 0 && prm->sub_rounds <= 255, "sub_rounds must be in [0, 255] (0: 8)");
    Bind bind_(p->hg->ctx->c);
    fm_refine(p, *prm, stats);
  });
}
