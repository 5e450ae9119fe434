// Community-restricted heavy-edge clustering and hypergraph contraction.
//
// Reference: MultilevelCoarsener::coarsenImpl / matchVertices
// (multilevel_coarsener.h:143-438), MultilevelVertexPairRater::rate /
// fillRatingMap / adaptiveEdgeSize (multilevel_vertex_pair_rater.h:109-279),
// HeavyEdgeScore, NoWeightPenalty, BestRatingPreferringUnmatched
// (policies/rating_*.h), StaticHypergraph::contract
// (static_hypergraph.cpp:58-513).
//
// Clustering pass, synchronous form: vertices are split into sub-rounds by a
// hash bucket; in a sub-round every unmatched vertex of the bucket rates the
// clusters of its neighbours on the state at the sub-round start:
//   size'(e) = max(#distinct clusters among the pins of e, 2), nets of at
//   least 1000 pins ignored, score(e) = w(e) / (size'(e) - 1) in 36.28 fixed
//   point (exact, associative sums), every distinct cluster of e gains
//   score(e) once;
// the target is the best cluster by (rating, not yet matched, hash) among
// clusters of the same community whose weight fits.  A vertex whose target
// does not move in the sub-round joins it (joiners of a target are admitted
// in (rating desc, hash, id) order up to the node weight limit); two vertices
// that chose each other pair up under the smaller id.
//
// Contraction, canonical form: coarse ids are the ranks of the cluster
// representatives, pins are sorted and unique per net, single-pin nets
// vanish, parallel nets (equal pin sets) merge into the smallest original net
// id with summed weight, coarse nets keep the original order, incident nets
// are ascending.
#include <algorithm>
#include <cstring>
#include <map>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.h"

namespace mkp {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kFixShift = 28;

struct CoArgs {
  uint32_t n;
  const uint32_t* voff;
  const uint32_t* inc;
  const uint32_t* noff;
  const uint32_t* pins;
  const int32_t* nw;
  const int32_t* comm;
  const uint64_t* ub;  // rating upper bound per vertex (sizes the global map)
  const uint32_t* cid;
  const int64_t* cw;
  const uint8_t* matched;
  uint32_t* target;
  uint64_t* rating;
  uint32_t ignore;
  uint64_t sampling;
  int64_t max_w;
  uint64_t seed;
  uint64_t tag;  // (pass << 8) | sub-round
  int prefer_unmatched;
  int64_t* visits;
};

__device__ inline uint32_t co_slot(uint32_t key, int bits) { return (key * 0x9E3779B1u) >> (32 - bits); }

// (value, unmatched, hash) order of the target choice
struct Choice {
  uint64_t val;
  int un;
  uint32_t t;
};
__device__ inline bool choice_better(const CoArgs& a, uint32_t u, const Choice& x, const Choice& y) {
  if (x.t == kNone) return false;
  if (y.t == kNone) return true;
  if (x.val != y.val) return x.val > y.val;
  if (x.un != y.un) return x.un > y.un;
  return hash4(a.seed, kCoTieTag | a.tag, u, x.t) > hash4(a.seed, kCoTieTag | a.tag, u, y.t);
}

// ------------------------------------------------------------- rating --
// A sub-round first marks the nets (fewer than `ignore` pins) of its
// proposers, then lists the distinct clusters of every marked net once
// (dclus[noff[e] ..], dcount[e]); score(e) follows from dcount.  A proposer
// then sums score(e) into a map keyed by cluster for every distinct cluster
// of its nets.  The fixed-point sums are exact and order-free, so any lane
// order gives the same ratings.

// The nets a proposer sweeps, in incidence order, stop before the first net
// whose distinct clusters would push the accessed keys past `sampling`
// (fillRatingMap's vertex degree sampling); visits counts 2|e| for every
// swept net and for the net that stops the sweep.  When even all nets stay
// within the threshold (ub + 2 <= sampling) the sweep takes every net.
__device__ inline uint32_t co_sweep_end(const CoArgs& a, const uint32_t* __restrict__ dcount, uint32_t u,
                                        int64_t& visits) {
  const uint32_t jb = a.voff[u], je = a.voff[u + 1];
  if (a.ub[u] + 2 <= a.sampling) {
    visits += 2 * int64_t(a.ub[u]);
    return je;
  }
  uint64_t acc = 0;
  for (uint32_t j = jb; j < je; ++j) {
    const uint32_t e = a.inc[j];
    const uint32_t s = a.noff[e + 1] - a.noff[e];
    if (s >= a.ignore) continue;
    visits += 2 * int64_t(s);
    const uint32_t d = dcount[e];
    if (acc + max(d, 2u) > a.sampling) return j;
    acc += d;
  }
  return je;
}

// nets of the proposers in `list`; the first marker of a net appends it to
// mlist (each net once per sub-round); G lanes per proposer
template <int G>
__global__ void __launch_bounds__(256)
co_mark_nets_kernel(CoArgs a, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, uint32_t stamp,
                    uint32_t* __restrict__ nmark, uint32_t* __restrict__ mlist, uint32_t* __restrict__ mcount) {
  const uint32_t cnt = *count;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t groups = gridDim.x * (blockDim.x / G);
  // the loops run in whole waves (trip counts are wave-uniform); a wave
  // gathers its appends in LDS and moves them to mlist with one atomic per
  // kMarkStage entries (an atomic on one address per step of every wave
  // would queue in memory)
  constexpr uint32_t kMarkStage = 256;
  __shared__ uint32_t stage[4][kMarkStage];
  const uint32_t wlane = threadIdx.x & 63;
  uint32_t* st = stage[threadIdx.x >> 6];
  uint32_t fill = 0;  // wave-uniform
  auto flush = [&]() {
    lds_wave_sync();
    uint32_t base = 0;
    if (wlane == 0) base = atomicAdd(mcount, fill);
    base = __shfl(base, 0, 64);
    for (uint32_t t = wlane; t < fill; t += 64) mlist[base + t] = st[t];
    lds_wave_sync();
    fill = 0;
  };
  const uint32_t first = blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
  const uint32_t wfirst = first - (wlane / G);  // group of the wave's lane 0
  for (uint32_t i0 = wfirst; i0 < cnt; i0 += groups) {
    const uint32_t i = i0 + wlane / G;
    const uint32_t u = i < cnt ? list[i] : 0u;
    const uint32_t jb = i < cnt ? a.voff[u] : 0u, je = i < cnt ? a.voff[u + 1] : 0u;
    uint32_t len = je - jb;
    for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o, 64));
    for (uint32_t k0 = 0; k0 < len; k0 += G) {
      const uint32_t k = k0 + lane;
      bool add = false;
      uint32_t e = 0;
      if (jb + k < je) {
        e = a.inc[jb + k];
        if (a.noff[e + 1] - a.noff[e] < a.ignore && nmark[e] != stamp) add = atomicExch(&nmark[e], stamp) != stamp;
      }
      const uint64_t bal = __ballot(add);
      if (!bal) continue;
      const uint32_t nadd = uint32_t(__popcll(bal));
      if (fill + nadd > kMarkStage) flush();
      if (add) st[fill + __popcll(bal & ((1ull << wlane) - 1))] = e;
      fill += nadd;
    }
  }
  if (fill) flush();
}

// distinct clusters of the marked nets with at most 16 pins: one net per
// thread, pins in registers
__global__ void __launch_bounds__(256)
co_dedup_small_kernel(CoArgs a, const uint32_t* __restrict__ mlist, const uint32_t* __restrict__ mcount,
                      uint32_t* __restrict__ dclus, uint32_t* __restrict__ dcount) {
  constexpr int D = 16;
  const uint32_t cnt = *mcount;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const uint32_t e = mlist[i];
    const uint32_t pb = a.noff[e], s = a.noff[e + 1] - pb;
    if (s > uint32_t(D)) continue;
    uint32_t r[D];
#pragma unroll
    for (int i = 0; i < D; ++i) r[i] = uint32_t(i) < s ? a.pins[pb + i] : 0u;
#pragma unroll
    for (int i = 0; i < D; ++i) r[i] = uint32_t(i) < s ? a.cid[r[i]] : kNone;
    uint32_t cnt = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      bool first = r[i] != kNone;
#pragma unroll
      for (int j = 0; j < i; ++j) first = first && r[j] != r[i];
      if (first) dclus[pb + cnt++] = r[i];
    }
    dcount[e] = cnt;
  }
}

// the larger marked nets (17 to ignore - 1 pins): one net per wave, an LDS
// set of 2^ceil(log2(2|e|)) slots
__global__ void __launch_bounds__(256)
co_dedup_wave_kernel(CoArgs a, const uint32_t* __restrict__ mlist, const uint32_t* __restrict__ mcount,
                     uint32_t* __restrict__ dclus, uint32_t* __restrict__ dcount) {
  constexpr int NS = 2048;
  __shared__ uint32_t sets[4 * NS];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t* set = sets + wid * NS;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const uint32_t cnt = *mcount;
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + wid; i < cnt; i += waves) {
    const uint32_t e = mlist[i];  // uniform over the wave
    const uint32_t pb = a.noff[e], s = a.noff[e + 1] - pb;
    if (s <= 16) continue;
    int nbits = 1;
    while ((1u << nbits) < 2 * s) ++nbits;
    const uint32_t NT = 1u << nbits;
    for (uint32_t t = lane; t < NT; t += 64) set[t] = kNone;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (uint32_t q = lane; q < s; q += 64) {
      const uint32_t rep = a.cid[a.pins[pb + q]];
      uint32_t h = co_slot(rep, nbits);
      while (true) {
        const uint32_t old = atomicCAS(&set[h], kNone, rep);
        if (old == kNone || old == rep) break;
        h = (h + 1) & (NT - 1);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint32_t base = 0;
    for (uint32_t t0 = 0; t0 < NT; t0 += 64) {
      const uint32_t t = t0 + lane;
      const uint32_t key = t < NT ? set[t] : kNone;
      const uint64_t bal = __ballot(key != kNone);
      if (key != kNone) dclus[pb + base + __popcll(bal & ((1ull << lane) - 1))] = key;
      base += __popcll(bal);
    }
    if (lane == 0) dcount[e] = base;
    __builtin_amdgcn_wave_barrier();
  }
}

// A slot's key never changes once set within a sub-round, so a plain read
// that shows a key is exact; only a slot that reads empty needs the CAS.
__device__ inline void co_map_add(uint32_t* K, uint64_t* V, int bits, uint32_t rep, uint64_t score) {
  const uint32_t T = 1u << bits;
  uint32_t h = co_slot(rep, bits);
  while (true) {
    uint32_t cur = __atomic_load_n(&K[h], __ATOMIC_RELAXED);
    if (cur == kNone) cur = atomicCAS(&K[h], kNone, rep);
    if (cur == kNone || cur == rep) break;
    h = (h + 1) & (T - 1);
  }
  atomicAdd(reinterpret_cast<unsigned long long*>(&V[h]), static_cast<unsigned long long>(score));
}

// Sums score(e) of u's swept nets [jb, jend) into the map.  A group of G
// lanes (G <= 64, inside one wave) takes the nets G at a time starting at
// jb + first with stride `step`: an exclusive scan of the distinct counts
// over the lanes lays the (net, cluster) items of the chunk out in a row,
// and lane l inserts items l, l + G, ...; the owner net of an item is found
// by a binary search over the lanes' running sums.
template <int G>
__device__ inline void co_fill_map(const CoArgs& a, const uint32_t* __restrict__ dclus,
                                   const uint32_t* __restrict__ dcount, uint32_t jb, uint32_t jend, uint32_t first,
                                   uint32_t step, uint32_t lane, uint32_t* K, uint64_t* V, int bits) {
  for (uint32_t j0 = jb + first; j0 < jend; j0 += step) {
    const uint32_t j = j0 + lane;
    uint32_t d = 0, pb = 0;
    uint64_t score = 0;
    if (j < jend) {
      const uint32_t e = a.inc[j];
      pb = a.noff[e];
      const uint32_t s = a.noff[e + 1] - pb;
      if (s < a.ignore) {
        d = dcount[e];
        score = (uint64_t(uint32_t(a.nw[e])) << kFixShift) / uint64_t(max(d, 2u) - 1);
      }
    }
    uint32_t incl = d;
    for (int o = 1; o < G; o <<= 1) {
      const uint32_t t = __shfl_up(incl, o, G);
      if (int(lane) >= o) incl += t;
    }
    const uint32_t total = __shfl(incl, G - 1, G);
    for (uint32_t i0 = 0; i0 < total; i0 += G) {
      const uint32_t i = i0 + lane;
      // owner: first lane whose running sum exceeds i
      int lo = 0, hi = G - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const uint32_t x = __shfl(incl, mid, G);
        if (x > i) hi = mid; else lo = mid + 1;
      }
      const uint32_t o_incl = __shfl(incl, lo, G);
      const uint32_t o_d = __shfl(d, lo, G);
      const uint32_t o_pb = __shfl(pb, lo, G);
      const uint64_t o_score = __shfl(score, lo, G);
      if (i < total) co_map_add(K, V, bits, dclus[o_pb + (i - (o_incl - o_d))], o_score);
    }
  }
}

// best target over map slots [t0, T) strided by `lanes`
__device__ inline Choice co_scan_map(const CoArgs& a, uint32_t u, const uint32_t* K, const uint64_t* V, uint32_t T,
                                     uint32_t lane, uint32_t lanes) {
  const int64_t cwu = a.cw[u];
  const int32_t cu = a.comm[u];
  Choice best{0, 0, kNone};
  for (uint32_t t = lane; t < T; t += lanes) {
    const uint32_t rep = K[t];
    if (rep == kNone || rep == u) continue;
    if (cwu + a.cw[rep] > a.max_w || a.comm[rep] != cu) continue;
    Choice c{V[t], a.prefer_unmatched ? (a.matched[rep] ? 0 : 1) : 0, rep};
    if (choice_better(a, u, c, best)) best = c;
  }
  return best;
}

// G lanes per proposer, a 2^TBITS-slot map per group in LDS (proposers with
// ub <= 2^(TBITS-1) distinct keys at most)
template <int G, int TBITS, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
co_rate_group_kernel(CoArgs a, const uint32_t* __restrict__ list, const uint32_t* __restrict__ list_count,
                     const uint32_t* __restrict__ dclus, const uint32_t* __restrict__ dcount) {
  constexpr int T = 1 << TBITS;
  constexpr int GPB = BLOCK / G;
  __shared__ uint32_t keys[GPB * T];
  __shared__ uint64_t vals[GPB * T];
  __shared__ uint32_t s_end[GPB];
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  uint32_t* K = keys + grp * T;
  uint64_t* V = vals + grp * T;
  const uint32_t count = *list_count;
  int64_t visits = 0;
  for (uint32_t base = blockIdx.x * GPB; base < count; base += gridDim.x * GPB) {
    const uint32_t idx = base + grp;
    const bool in = idx < count;
    const uint32_t u = in ? list[idx] : 0;
    int tb = 1;
    const uint64_t ub = in ? a.ub[u] : 0;
    while (tb < TBITS && (uint64_t(1) << tb) < 2 * ub) ++tb;
    const uint32_t Tn = 1u << tb;
    for (uint32_t t = lane; t < Tn; t += G) { K[t] = kNone; V[t] = 0; }
    if (in && lane == 0) s_end[grp] = co_sweep_end(a, dcount, u, visits);
    __syncthreads();
    if (in) co_fill_map<G>(a, dclus, dcount, a.voff[u], s_end[grp], 0, G, lane, K, V, tb);
    __syncthreads();
    if (in) {
      Choice best = co_scan_map(a, u, K, V, Tn, lane, G);
      for (int o = G / 2; o > 0; o >>= 1) {
        Choice other;
        other.val = __shfl_xor(best.val, o, G);
        other.un = __shfl_xor(best.un, o, G);
        other.t = __shfl_xor(best.t, o, G);
        if (choice_better(a, u, other, best)) best = other;
      }
      if (lane == 0) {
        a.target[u] = best.t;
        a.rating[u] = best.val;
      }
    }
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) visits += __shfl_xor(visits, o, 64);
  if ((threadIdx.x & 63) == 0 && visits)
    atomicAdd(reinterpret_cast<unsigned long long*>(a.visits), static_cast<unsigned long long>(visits));
}

// ------------------------------------------------- heavy proposers --
// Proposers with ub > 1024: their rating items are spread over the whole
// grid.  Each owns a table of T = pow2 >= max(kHrChunk, 2 L) slots (L = its
// distinct (net, cluster) items) in one arena; tables are multiples of
// kHrChunk, so a chunk of the arena belongs to one proposer.
//   hr_prep     per proposer: sweep end, visits, L, T, swept incidences
//   (scans)     item, table and incidence offsets
//   hr_inc      per swept incidence: net and item count
//   (scan)      item offset of every incidence
//   hr_clear    empty tables
//   hr_insert   per item: score(e) into the owner's table
//   hr_scan     best target per chunk
//   hr_decide   per proposer: best over its chunks
constexpr uint32_t kHrChunk = 1024;

struct HrArgs {
  const uint32_t* list;  // heavy proposers of the sub-round
  uint32_t count;
  uint64_t* tsize;       // T per proposer
  uint64_t* isize;       // swept incidences per proposer
  uint64_t* ioff;        // exclusive scan of isize (count + 1)
  uint64_t* toff;        // exclusive scan of tsize (count + 1)
  uint32_t* inc_net;     // per swept incidence: net (kNone: ignored net)
  uint32_t* inc_own;     // per swept incidence: index of its proposer in list
  uint64_t* inc_cnt;     // per swept incidence: items
  uint64_t* inc_off;     // exclusive scan of inc_cnt (incidences + 1)
  uint32_t* keys;
  uint64_t* vals;
};

// co_sweep_end by a whole wave: 64 nets at a time, a prefix scan of their
// distinct counts, the first net that stops the sweep by ballot
__device__ inline uint32_t co_sweep_end_wave(const CoArgs& a, const uint32_t* __restrict__ dcount, uint32_t u,
                                             uint32_t lane, int64_t& visits) {
  const uint32_t jb = a.voff[u], je = a.voff[u + 1];
  if (a.ub[u] + 2 <= a.sampling) {
    if (lane == 0) visits += 2 * int64_t(a.ub[u]);
    return je;
  }
  uint64_t acc = 0;  // distinct keys of the nets before the chunk
  for (uint32_t j0 = jb; j0 < je; j0 += 64) {
    const uint32_t j = j0 + lane;
    uint32_t s = 0, d = 0;
    bool take = false;
    if (j < je) {
      const uint32_t e = a.inc[j];
      s = a.noff[e + 1] - a.noff[e];
      take = s < a.ignore;
      if (take) d = dcount[e];
    }
    uint64_t incl = d;
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t t = __shfl_up(incl, o, 64);
      if (int(lane) >= o) incl += t;
    }
    const bool stop = take && acc + (incl - d) + max(d, 2u) > a.sampling;
    const uint64_t bal = __ballot(stop);
    const int f = bal ? __ffsll(static_cast<long long>(bal)) - 1 : 64;
    int64_t v = (take && int(lane) <= f) ? 2 * int64_t(s) : 0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) visits += v;
    if (bal) return j0 + uint32_t(f);
    acc += __shfl(incl, 63, 64);
  }
  return je;
}

__global__ void __launch_bounds__(256)
hr_prep_kernel(CoArgs a, HrArgs h, const uint32_t* __restrict__ dcount) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  int64_t visits = 0;
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < h.count; i += waves) {
    const uint32_t u = h.list[i];
    const uint32_t jend = co_sweep_end_wave(a, dcount, u, lane, visits);
    const uint32_t jb = a.voff[u];
    uint64_t L = 0;
    for (uint32_t j = jb + lane; j < jend; j += 64) {
      const uint32_t e = a.inc[j];
      if (a.noff[e + 1] - a.noff[e] < a.ignore) L += dcount[e];
    }
    for (int o = 32; o > 0; o >>= 1) L += __shfl_xor(L, o, 64);
    if (lane == 0) {
      uint64_t T = kHrChunk;
      while (T < 2 * L) T <<= 1;
      h.tsize[i] = T;
      h.isize[i] = jend - jb;
    }
  }
  if (lane == 0 && visits)
    atomicAdd(reinterpret_cast<unsigned long long*>(a.visits), static_cast<unsigned long long>(visits));
}

// largest i with off[i] <= x
__device__ inline uint32_t hr_owner(const uint64_t* off, uint32_t count, uint64_t x) {
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void hr_inc_kernel(CoArgs a, HrArgs h, uint64_t total, const uint32_t* __restrict__ dcount) {
  for (uint64_t x = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; x < total;
       x += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t i = hr_owner(h.ioff, h.count, x);
    const uint32_t j = a.voff[h.list[i]] + uint32_t(x - h.ioff[i]);
    const uint32_t e = a.inc[j];
    const bool take = a.noff[e + 1] - a.noff[e] < a.ignore;
    h.inc_net[x] = take ? e : kNone;
    h.inc_own[x] = i;
    h.inc_cnt[x] = take ? dcount[e] : 0;
  }
}

__global__ void hr_clear_kernel(HrArgs h, uint64_t slots) {
  for (uint64_t x = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; x < slots;
       x += uint64_t(gridDim.x) * blockDim.x) {
    h.keys[x] = kNone;
    h.vals[x] = 0;
  }
}

// Item t belongs to the largest incidence x with inc_off[x] <= t, which owns
// items: a later incidence without items would start at the same offset.  A
// block takes kHrItems consecutive items; one thread finds the first and
// last incidence of the chunk in global memory, the offsets in between go to
// LDS (relative to the chunk start), and every item finds its incidence
// there, so an item costs one LDS search instead of a global one.
constexpr uint32_t kHrItems = 2048;  // items per block and chunk
constexpr uint32_t kHrSpan = 2048;   // incidences a chunk may span in LDS

__device__ inline uint64_t hr_last_le(const uint64_t* off, uint64_t lo, uint64_t hi, uint64_t t) {
  // largest x in [lo, hi) with off[x] <= t, given off[lo] <= t
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(256)
hr_insert_kernel(CoArgs a, HrArgs h, uint64_t total_items, uint64_t total_inc, const uint32_t* __restrict__ dclus,
                 const uint32_t* __restrict__ dcount) {
  __shared__ uint32_t s_off[kHrSpan];
  __shared__ uint64_t s_x0, s_x1;
  for (uint64_t t0 = blockIdx.x * uint64_t(kHrItems); t0 < total_items; t0 += uint64_t(gridDim.x) * kHrItems) {
    const uint64_t t1 = min(t0 + kHrItems, total_items);
    if (threadIdx.x == 0) {
      const uint64_t x0 = hr_last_le(h.inc_off, 0, total_inc, t0);
      s_x0 = x0;
      s_x1 = hr_last_le(h.inc_off, x0, total_inc, t1 - 1);
    }
    __syncthreads();
    const uint64_t x0 = s_x0, x1 = s_x1;
    const uint64_t nx = x1 - x0 + 1;
    const bool in_lds = nx <= kHrSpan;
    if (in_lds) {
      // x0 is the last incidence starting at or before t0: every later one
      // starts after t0, so the relative offsets are exact except the first
      for (uint32_t j = threadIdx.x; j < nx; j += blockDim.x) {
        const uint64_t o = h.inc_off[x0 + j];
        s_off[j] = o <= t0 ? 0u : uint32_t(o - t0);
      }
    }
    __syncthreads();
    for (uint64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
      uint64_t x;
      if (in_lds) {
        const uint32_t rel = uint32_t(t - t0);
        uint32_t lo = 0, hi = uint32_t(nx);
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (s_off[mid] <= rel) lo = mid; else hi = mid;
        }
        x = x0 + lo;
      } else {
        x = hr_last_le(h.inc_off, x0, x1 + 1, t);
      }
      const uint32_t e = h.inc_net[x];
      const uint32_t q = uint32_t(t - h.inc_off[x]);
      const uint32_t i = h.inc_own[x];
      const uint32_t d = dcount[e];
      const uint64_t score = (uint64_t(uint32_t(a.nw[e])) << kFixShift) / uint64_t(max(d, 2u) - 1);
      const uint64_t base = h.toff[i];
      const int bits = __builtin_ctzll(h.toff[i + 1] - base);
      co_map_add(h.keys + base, h.vals + base, bits, dclus[a.noff[e] + q], score);
    }
    __syncthreads();
  }
}

// Heavy proposers rated by one block each, grid-strided, with the rating map
// in LDS.  The items of up to BLOCK nets at a time are spread over the whole
// block: a block scan of the distinct counts, and each item finds its net by
// a binary search over the scan in LDS.  A proposer with at most BLOCK swept
// nets loads them once for all its passes.
//
// hr_hash_kernel: a hash map of 2^tb slots (2^tb >= 2 min(L, n) when that
// fits in T / 2, L the proposer's (net, cluster) items and n the clusters of
// the level, else T).  With P passes the keys are split by a hash into P
// classes and pass p inserts only class p, so every key collects its whole
// rating in its own pass and the best target is the best over the passes.
// A pass reserves a count per new key and stops at 3/4 of the slots; the
// proposer then starts over with twice the passes, and with more than
// `max_passes` it goes to `over` (the arena path, which counts its visits
// itself).
//
// hr_dense_kernel (levels with few clusters): the map is an array indexed by
// cluster id, in passes over id ranges of T, with no probing and no
// overflow.  Scores are positive except for nets of weight 0, whose keys are
// kept in a presence bit set.
constexpr uint32_t kHbDense = 8192;  // dense slots per pass

__device__ inline uint32_t hb_class(uint32_t key, uint32_t P) { return fmix32(key ^ 0x5bd1e995u) % P; }

// adds score at key unless the key is new and the map already holds `cap`
// keys (then *over is set)
__device__ inline void hb_map_add(uint32_t* K, uint64_t* V, int bits, uint32_t rep, uint64_t score,
                                  uint32_t* nkeys, uint32_t cap, int* over) {
  const uint32_t T = 1u << bits;
  uint32_t h = co_slot(rep, bits);
  bool reserved = false;
  while (true) {
    uint32_t cur = __atomic_load_n(&K[h], __ATOMIC_RELAXED);
    if (cur == kNone) {
      if (!reserved) {
        if (atomicAdd(nkeys, 1u) >= cap) {
          *over = 1;
          return;
        }
        reserved = true;
      }
      cur = atomicCAS(&K[h], kNone, rep);
    }
    if (cur == kNone || cur == rep) break;
    h = (h + 1) & (T - 1);
  }
  atomicAdd(reinterpret_cast<unsigned long long*>(&V[h]), static_cast<unsigned long long>(score));
}

// LDS state of the per-block item walk
template <int BLOCK>
struct HbWalk {
  uint32_t incl[BLOCK];  // running item count per net of the chunk
  uint32_t pb[BLOCK];    // pin offset of the net
  uint64_t sc[BLOCK];    // score of the net
  uint32_t w[BLOCK / 64];
  uint64_t l[BLOCK / 64];
  uint32_t end;
};

// loads nets [j0, min(j0 + BLOCK, jend)) of the proposer; returns their items
template <int BLOCK>
__device__ inline uint32_t hb_load(const CoArgs& a, const uint32_t* __restrict__ dcount, uint32_t j0, uint32_t jend,
                                   HbWalk<BLOCK>& sh) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const uint32_t j = j0 + tid;
  uint32_t d = 0, pb = 0;
  uint64_t score = 0;
  if (j < jend) {
    const uint32_t e = a.inc[j];
    pb = a.noff[e];
    if (a.noff[e + 1] - pb < a.ignore) {
      d = dcount[e];
      score = (uint64_t(uint32_t(a.nw[e])) << kFixShift) / uint64_t(max(d, 2u) - 1);
    }
  }
  uint32_t incl = d;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(incl, o, 64);
    if (int(lane) >= o) incl += t;
  }
  if (lane == 63) sh.w[wid] = incl;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int w = 0; w < BLOCK / 64; ++w) {
    const uint32_t x = sh.w[w];
    if (w < int(wid)) before += x;
    total += x;
  }
  sh.incl[tid] = before + incl;
  sh.pb[tid] = pb;
  sh.sc[tid] = score;
  __syncthreads();
  return total;
}

// calls f(key, score) for the items [0, total) of the loaded chunk that
// this thread takes (tid, tid + BLOCK, ...), eight at a time so that their
// net searches and key loads are issued together; stops when *stop is set
// (checked per eight); returns the items taken
template <int BLOCK, class F>
__device__ inline uint32_t hb_walk(const uint32_t* __restrict__ dclus, const HbWalk<BLOCK>& sh, uint32_t total,
                                   const int* stop, F&& f) {
  constexpr int Q = 8;
  uint32_t done = 0;
  for (uint32_t base = threadIdx.x; base < total; base += Q * BLOCK) {
    if (stop && __atomic_load_n(stop, __ATOMIC_RELAXED)) break;
    uint32_t key[Q];
    uint64_t sc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint32_t it = min(base + uint32_t(q) * BLOCK, total - 1);
      uint32_t lo = 0, hi = BLOCK - 1;  // first net whose running count exceeds it
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sh.incl[mid] > it) hi = mid; else lo = mid + 1;
      }
      const uint32_t start = lo ? sh.incl[lo - 1] : 0u;
      key[q] = dclus[sh.pb[lo] + (it - start)];
      sc[q] = sh.sc[lo];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
      if (base + uint32_t(q) * BLOCK < total) {
        f(key[q], sc[q]);
        ++done;
      }
  }
  return done;
}

// sweep end and item count L of proposer u; with at most BLOCK swept nets
// they stay loaded in sh (*single set, *total their items)
template <int BLOCK>
__device__ inline uint64_t hb_prologue(const CoArgs& a, const uint32_t* __restrict__ dcount, uint32_t u,
                                       HbWalk<BLOCK>& sh, int64_t& vu, uint32_t& jend, bool& single,
                                       uint32_t& total) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  if (wid == 0) {
    const uint32_t je = co_sweep_end_wave(a, dcount, u, lane, vu);
    if (lane == 0) sh.end = je;
  }
  __syncthreads();
  const uint32_t jb = a.voff[u];
  jend = sh.end;
  single = jend - jb <= uint32_t(BLOCK);
  if (single) {
    total = hb_load<BLOCK>(a, dcount, jb, jend, sh);
    return total;
  }
  uint64_t L = 0;
  for (uint32_t j = jb + tid; j < jend; j += BLOCK) {
    const uint32_t e = a.inc[j];
    if (a.noff[e + 1] - a.noff[e] < a.ignore) L += dcount[e];
  }
  for (int o = 32; o > 0; o >>= 1) L += __shfl_xor(L, o, 64);
  if (lane == 0) sh.l[wid] = L;
  __syncthreads();
  L = 0;
  for (int w = 0; w < BLOCK / 64; ++w) L += sh.l[w];
  return L;
}

// best of the block's per-thread choices; written by thread 0
template <int BLOCK>
__device__ inline void hb_finish(const CoArgs& a, uint32_t u, Choice best, uint64_t* s_val, int* s_un,
                                 uint32_t* s_t) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    Choice other;
    other.val = __shfl_xor(best.val, o, 64);
    other.un = __shfl_xor(best.un, o, 64);
    other.t = __shfl_xor(best.t, o, 64);
    if (choice_better(a, u, other, best)) best = other;
  }
  if (lane == 0) {
    s_val[wid] = best.val;
    s_un[wid] = best.un;
    s_t[wid] = best.t;
  }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < BLOCK / 64; ++w) {
      Choice other{s_val[w], s_un[w], s_t[w]};
      if (choice_better(a, u, other, best)) best = other;
    }
    a.target[u] = best.t;
    a.rating[u] = best.val;
  }
  __syncthreads();
}

template <int TBITS, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
hr_hash_kernel(CoArgs a, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
               const uint32_t* __restrict__ dclus, const uint32_t* __restrict__ dcount, uint32_t max_passes,
               uint32_t* __restrict__ over, uint32_t* __restrict__ over_count) {
  constexpr int T = 1 << TBITS;
  constexpr int NW = BLOCK / 64;
  __shared__ uint32_t K[T];
  __shared__ uint64_t V[T];
  __shared__ HbWalk<BLOCK> sh;
  __shared__ uint64_t s_val[NW];
  __shared__ int s_un[NW];
  __shared__ uint32_t s_t[NW];
  __shared__ uint32_t s_nkeys, s_done;
  __shared__ int s_over;
  const uint32_t tid = threadIdx.x;
  const uint32_t cnt = *count;
  int64_t visits = 0;
  for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
    const uint32_t u = list[i];
    int64_t vu = 0;
    uint32_t jend, total0;
    bool single;
    const uint64_t L = hb_prologue<BLOCK>(a, dcount, u, sh, vu, jend, single, total0);
    const uint32_t jb = a.voff[u];
    const uint64_t D = min<uint64_t>(L, a.n);
    int tb = TBITS;
    if (D <= uint64_t(T / 2)) {
      tb = 6;
      while ((uint64_t(1) << tb) < 2 * D) ++tb;
    }
    uint32_t P = 1;
    bool ok = max_passes > 1 || D <= uint64_t(T / 2);  // one-pass tier: only maps that cannot overflow
    Choice best{0, 0, kNone};
    while (ok) {
      const uint32_t Tn = 1u << tb;
      const uint32_t cap = 3 * Tn / 4;
      best = Choice{0, 0, kNone};
      auto add = [&](uint32_t pass) {
        return [&, pass](uint32_t key, uint64_t score) {
          if (P == 1 || hb_class(key, P) == pass) hb_map_add(K, V, tb, key, score, &s_nkeys, cap, &s_over);
        };
      };
      for (uint32_t pass = 0; pass < P; ++pass) {
        uint32_t done = 0;  // items this thread walked in the pass
        for (uint32_t t = tid; t < Tn; t += BLOCK) {
          K[t] = kNone;
          V[t] = 0;
        }
        if (tid == 0) {
          s_nkeys = 0;
          s_over = 0;
          s_done = 0;
        }
        __syncthreads();
        if (single) {
          done += hb_walk<BLOCK>(dclus, sh, total0, &s_over, add(pass));
        } else {
          for (uint32_t j0 = jb; j0 < jend; j0 += BLOCK) {
            const uint32_t total = hb_load<BLOCK>(a, dcount, j0, jend, sh);
            done += hb_walk<BLOCK>(dclus, sh, total, &s_over, add(pass));
            __syncthreads();
            if (s_over) break;
          }
        }
        __syncthreads();
        const bool failed = s_over != 0;
        if (!failed) {
          const Choice c = co_scan_map(a, u, K, V, Tn, tid, BLOCK);
          if (choice_better(a, u, c, best)) best = c;
        } else if (done) {
          atomicAdd(&s_done, done);
        }
        __syncthreads();
        if (failed) {
          ok = false;
          break;
        }
      }
      if (ok) break;
      // the failed pass saw `cap` keys of its class within s_done of the L
      // items: size the classes for about half a map each
      const uint64_t seen = max<uint64_t>(s_done, 1);
      __syncthreads();  // s_done is cleared by the next pass
      const uint64_t est = uint64_t(cap) * P * L / seen;
      const uint64_t per = 3 * uint64_t(T) / 8;
      const uint64_t next = max<uint64_t>(2 * uint64_t(P), (est + per - 1) / per);
      if (next > max_passes) break;
      P = uint32_t(next);
      tb = TBITS;
      ok = true;
    }
    if (!ok) {
      if (tid == 0) over[atomicAdd(over_count, 1u)] = u;
      continue;
    }
    visits += vu;
    hb_finish<BLOCK>(a, u, best, s_val, s_un, s_t);
  }
  if ((tid & 63) == 0 && visits)
    atomicAdd(reinterpret_cast<unsigned long long*>(a.visits), static_cast<unsigned long long>(visits));
}

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK)
hr_dense_kernel(CoArgs a, const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                const uint32_t* __restrict__ dclus, const uint32_t* __restrict__ dcount) {
  constexpr uint32_t T = kHbDense;
  constexpr int NW = BLOCK / 64;
  __shared__ uint64_t V[T];
  __shared__ uint32_t bits[T / 32];
  __shared__ HbWalk<BLOCK> sh;
  __shared__ uint64_t s_val[NW];
  __shared__ int s_un[NW];
  __shared__ uint32_t s_t[NW];
  const uint32_t tid = threadIdx.x;
  const uint32_t cnt = *count;
  int64_t visits = 0;
  for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
    const uint32_t u = list[i];
    int64_t vu = 0;
    uint32_t jend, total0;
    bool single;
    hb_prologue<BLOCK>(a, dcount, u, sh, vu, jend, single, total0);
    const uint32_t jb = a.voff[u];
    const int64_t cwu = a.cw[u];
    const int32_t cu = a.comm[u];
    Choice best{0, 0, kNone};
    for (uint32_t lo = 0; lo < a.n; lo += T) {
      const uint32_t span = min(T, a.n - lo);
      for (uint32_t t = tid; t < span; t += BLOCK) V[t] = 0;
      for (uint32_t t = tid; t < (span + 31) / 32; t += BLOCK) bits[t] = 0;
      __syncthreads();
      auto add = [&](uint32_t key, uint64_t score) {
        const uint32_t r = key - lo;
        if (r < span) {
          if (score) atomicAdd(reinterpret_cast<unsigned long long*>(&V[r]), static_cast<unsigned long long>(score));
          else atomicOr(&bits[r >> 5], 1u << (r & 31));
        }
      };
      if (single) {
        hb_walk<BLOCK>(dclus, sh, total0, nullptr, add);
      } else {
        for (uint32_t j0 = jb; j0 < jend; j0 += BLOCK) {
          const uint32_t total = hb_load<BLOCK>(a, dcount, j0, jend, sh);
          hb_walk<BLOCK>(dclus, sh, total, nullptr, add);
          __syncthreads();
        }
      }
      __syncthreads();
      for (uint32_t t = tid; t < span; t += BLOCK) {
        const uint64_t v = V[t];
        if (v == 0 && !((bits[t >> 5] >> (t & 31)) & 1u)) continue;
        const uint32_t rep = lo + t;
        if (rep == u) continue;
        if (cwu + a.cw[rep] > a.max_w || a.comm[rep] != cu) continue;
        Choice c{v, a.prefer_unmatched ? (a.matched[rep] ? 0 : 1) : 0, rep};
        if (choice_better(a, u, c, best)) best = c;
      }
      __syncthreads();
    }
    visits += vu;
    hb_finish<BLOCK>(a, u, best, s_val, s_un, s_t);
  }
  if ((tid & 63) == 0 && visits)
    atomicAdd(reinterpret_cast<unsigned long long*>(a.visits), static_cast<unsigned long long>(visits));
}

__global__ void __launch_bounds__(256) hr_scan_kernel(CoArgs a, HrArgs h, uint64_t chunks, uint64_t* __restrict__ cval,
                                                      int* __restrict__ cun, uint32_t* __restrict__ ct) {
  __shared__ uint64_t s_val[4];
  __shared__ int s_un[4];
  __shared__ uint32_t s_t[4];
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const uint64_t s0 = ch * kHrChunk;
    const uint32_t i = hr_owner(h.toff, h.count, s0);
    const uint32_t u = h.list[i];
    Choice best = co_scan_map(a, u, h.keys + s0, h.vals + s0, kHrChunk, threadIdx.x, blockDim.x);
    for (int o = 32; o > 0; o >>= 1) {
      Choice other;
      other.val = __shfl_xor(best.val, o, 64);
      other.un = __shfl_xor(best.un, o, 64);
      other.t = __shfl_xor(best.t, o, 64);
      if (choice_better(a, u, other, best)) best = other;
    }
    if ((threadIdx.x & 63) == 0) {
      s_val[threadIdx.x >> 6] = best.val;
      s_un[threadIdx.x >> 6] = best.un;
      s_t[threadIdx.x >> 6] = best.t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < int(blockDim.x >> 6); ++w) {
        Choice other{s_val[w], s_un[w], s_t[w]};
        if (choice_better(a, u, other, best)) best = other;
      }
      cval[ch] = best.val;
      cun[ch] = best.un;
      ct[ch] = best.t;
    }
    __syncthreads();
  }
}

__global__ void hr_decide_kernel(CoArgs a, HrArgs h, const uint64_t* __restrict__ cval, const int* __restrict__ cun,
                                 const uint32_t* __restrict__ ct) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < h.count; i += gridDim.x * blockDim.x) {
    const uint32_t u = h.list[i];
    Choice best{0, 0, kNone};
    for (uint64_t ch = h.toff[i] / kHrChunk; ch < h.toff[i + 1] / kHrChunk; ++ch) {
      Choice c{cval[ch], cun[ch], ct[ch]};
      if (choice_better(a, u, c, best)) best = c;
    }
    a.target[u] = best.t;
    a.rating[u] = best.val;
  }
}

// rating upper bound: sum of the sizes of the nets the rating sweeps; G
// lanes per vertex (G divides 64)
template <int G>
__global__ void co_ub_kernel(const uint32_t* __restrict__ voff, const uint32_t* __restrict__ inc,
                             const uint32_t* __restrict__ noff, uint32_t n, uint32_t ignore,
                             uint64_t* __restrict__ ub, uint64_t* __restrict__ ub_max) {
  uint64_t mx = 0;
  const uint32_t lane = threadIdx.x % G;
  const uint32_t groups = gridDim.x * (blockDim.x / G);
  // whole waves iterate together: the group reduction shuffles over G lanes
  const uint32_t g0 = blockIdx.x * (blockDim.x / G) + threadIdx.x / G;
  for (uint32_t b = g0 - (threadIdx.x & 63) / G; b < n; b += groups) {
    const uint32_t u = b + (threadIdx.x & 63) / G;
    uint64_t s = 0;
    if (u < n)
      for (uint32_t j = voff[u] + lane; j < voff[u + 1]; j += G) {
        const uint32_t e = inc[j];
        const uint32_t sz = noff[e + 1] - noff[e];
        s += sz < ignore ? sz : 0;
      }
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (u < n && lane == 0) ub[u] = s;
    if (u < n) mx = max(mx, s);
  }
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
  if ((threadIdx.x & 63) == 0 && mx) atomicMax(reinterpret_cast<unsigned long long*>(ub_max), static_cast<unsigned long long>(mx));
}

// proposers of the sub-round by tier (rating map in LDS if ub <= 1024)
__global__ void __launch_bounds__(256)
co_select_kernel(uint32_t n, const uint8_t* __restrict__ matched, const uint64_t* __restrict__ ub, uint32_t salt,
                 uint32_t buckets, uint32_t bucket, uint32_t* __restrict__ target, uint32_t* __restrict__ list_a,
                 uint32_t* __restrict__ list_b, uint32_t* __restrict__ list_c, uint32_t* __restrict__ counts) {
  // tiers: c: ub <= 64 (counts[6]), a: ub <= 1024 (counts[0]), b: heavier
  // (counts[1])
  __shared__ uint32_t ca, cb, cc, ba, bb, bc;
  if (threadIdx.x == 0) { ca = 0; cb = 0; cc = 0; }
  __syncthreads();
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  int tier = -1;
  uint32_t l = 0;
  if (u < n) {
    target[u] = kNone;
    if (!matched[u] && (buckets <= 1 || bucket_of(salt, u, buckets) == bucket)) {
      const uint64_t b = ub[u];
      if (b <= 64) { tier = 2; l = atomicAdd(&cc, 1u); }
      else if (b <= 1024) { tier = 0; l = atomicAdd(&ca, 1u); }
      else { tier = 1; l = atomicAdd(&cb, 1u); }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ba = ca ? atomicAdd(&counts[0], ca) : 0;
    bb = cb ? atomicAdd(&counts[1], cb) : 0;
    bc = cc ? atomicAdd(&counts[6], cc) : 0;
  }
  __syncthreads();
  if (tier == 0) list_a[ba + l] = u;
  if (tier == 1) list_b[bb + l] = u;
  if (tier == 2) list_c[bc + l] = u;
}

// joiners and pairs (counts[4]: overflowing targets)
__global__ void co_resolve_kernel(uint32_t n, const uint32_t* __restrict__ target, const int32_t* __restrict__ vw,
                                  int64_t* __restrict__ jw, uint8_t* __restrict__ role) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint32_t t = target[u];
    uint8_t r = 0;
    if (t != kNone) {
      if (target[t] == kNone) {
        r = 1;  // joiner
        atomicAdd(reinterpret_cast<unsigned long long*>(&jw[t]), static_cast<unsigned long long>(int64_t(vw[u])));
      } else if (target[t] == u && u < t) {
        r = 2;  // pair representative
      }
    }
    role[u] = r;
  }
}

__global__ void co_overflow_kernel(uint32_t n, const uint32_t* __restrict__ target, const uint8_t* __restrict__ role,
                                   const int64_t* __restrict__ cw, const int64_t* __restrict__ jw, int64_t max_w,
                                   uint32_t* __restrict__ olist, uint32_t* __restrict__ ocount) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    if (role[u] != 1) continue;
    const uint32_t t = target[u];
    if (cw[t] + jw[t] > max_w) olist[atomicAdd(ocount, 1u)] = u;
  }
}

// admit: role 1 joiners join unless listed as rejected (role 3)
__global__ void co_apply_kernel(uint32_t n, const uint32_t* __restrict__ target, const uint8_t* __restrict__ role,
                                const int32_t* __restrict__ vw, uint32_t* __restrict__ cid, int64_t* __restrict__ cw,
                                uint8_t* __restrict__ matched, uint32_t* __restrict__ contracted) {
  uint32_t local = 0;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint8_t r = role[u];
    if (r == 1) {
      const uint32_t t = target[u];
      cid[u] = t;
      atomicAdd(reinterpret_cast<unsigned long long*>(&cw[t]), static_cast<unsigned long long>(int64_t(vw[u])));
      matched[u] = 1;
      matched[t] = 1;
      ++local;
    } else if (r == 2) {
      const uint32_t t = target[u];  // u < t, u represents the pair
      cid[t] = u;
      atomicAdd(reinterpret_cast<unsigned long long*>(&cw[u]), static_cast<unsigned long long>(int64_t(vw[t])));
      matched[u] = 1;
      matched[t] = 1;
      ++local;
    }
  }
  for (int o = 32; o > 0; o >>= 1) local += __shfl_xor(local, o, 64);
  if ((threadIdx.x & 63) == 0 && local) atomicAdd(contracted, local);
}

__global__ void co_init_kernel(uint32_t n, const int32_t* __restrict__ vw, uint32_t* __restrict__ cid,
                               int64_t* __restrict__ cw, uint8_t* __restrict__ matched) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    cid[u] = u;
    cw[u] = vw[u];
    matched[u] = 0;
  }
}

// ---------------------------------------------------------- contraction --
__global__ void co_mark_kernel(const uint32_t* __restrict__ cid, uint32_t n, uint32_t* __restrict__ used) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) used[cid[u]] = 1;
}

__global__ void co_cmap_kernel(const uint32_t* __restrict__ cid, const uint32_t* __restrict__ map, uint32_t n,
                               const int32_t* __restrict__ vw, uint32_t* __restrict__ cmap, int32_t* __restrict__ cvw,
                               uint32_t* __restrict__ last) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint32_t c = map[cid[u]];
    cmap[u] = c;
    atomicAdd(&cvw[c], vw[u]);
    atomicMax(&last[c], u);
  }
}

// A coarse vertex takes the community of its highest-numbered member (all
// members share one community when clustering respects communities).
__global__ void co_comm_kernel(uint32_t nc, const uint32_t* __restrict__ last, const int32_t* __restrict__ comm,
                               int32_t* __restrict__ ccomm) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
    ccomm[c] = comm[last[c]];
}

__global__ void co_pin_keys_kernel(const uint32_t* __restrict__ pins, const uint32_t* __restrict__ pin_net,
                                   const uint32_t* __restrict__ cmap, uint64_t p, int bits,
                                   uint64_t* __restrict__ keys) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < p; i += uint64_t(gridDim.x) * blockDim.x)
    keys[i] = (uint64_t(pin_net[i]) << bits) | cmap[pins[i]];
}

// unique flags on the sorted (net, coarse pin) keys; per net distinct count
// and an order-free 64-bit fingerprint of the pin set (sum of mixed pin ids;
// it plays the role of the footprint of static_hypergraph.cpp:152,193-196)
__global__ void co_unique_kernel(const uint64_t* __restrict__ keys, uint64_t p, int bits, uint32_t* __restrict__ flag,
                                 uint32_t* __restrict__ size, uint64_t* __restrict__ foot) {
  const uint64_t mask = (uint64_t(1) << bits) - 1;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < p; i += uint64_t(gridDim.x) * blockDim.x) {
    const bool first = i == 0 || keys[i - 1] != keys[i];
    flag[i] = first ? 1u : 0u;
    if (first) {
      const uint32_t e = uint32_t(keys[i] >> bits);
      const uint32_t v = uint32_t(keys[i] & mask);
      atomicAdd(&size[e], 1u);
      atomicAdd(reinterpret_cast<unsigned long long*>(&foot[e]), static_cast<unsigned long long>(mix64(v)));
    }
  }
}

__global__ void co_net_keys_kernel(uint32_t m, const uint32_t* __restrict__ size, const uint64_t* __restrict__ foot,
                                   uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                   uint32_t* __restrict__ nvalid) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const bool valid = size[e] > 1;
    // invalid nets sort last: valid keys keep the top bit clear
    keys[e] = valid ? ((foot[e] ^ uint64_t(size[e])) >> 1) : ~0ull;
    vals[e] = e;
    if (valid) atomicAdd(nvalid, 1u);
  }
}

// start index of every run of equal fingerprints (max-scanned afterwards)
__global__ void co_run_start_kernel(const uint64_t* __restrict__ skeys, uint32_t nvalid, uint64_t* __restrict__ rs) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nvalid; i += gridDim.x * blockDim.x)
    rs[i] = (i == 0 || skeys[i - 1] != skeys[i]) ? i : 0;
}

// For every valid net in (fingerprint, id) order: the first earlier net of
// the same run with identical pins is its representative (smallest id).
__global__ void co_parallel_kernel(const uint32_t* __restrict__ sids, uint32_t nvalid,
                                   const uint64_t* __restrict__ run_start, const uint32_t* __restrict__ size,
                                   const uint32_t* __restrict__ ustart, const uint64_t* __restrict__ pk, int bits,
                                   uint32_t* __restrict__ rep) {
  const uint64_t mask = (uint64_t(1) << bits) - 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nvalid; i += gridDim.x * blockDim.x) {
    const uint32_t e = sids[i];
    uint32_t r = e;
    const uint32_t sz = size[e];
    for (uint32_t j = uint32_t(run_start[i]); j < i; ++j) {
      const uint32_t f = sids[j];
      if (size[f] != sz) continue;
      bool same = true;
      for (uint32_t q = 0; q < sz && same; ++q) same = (pk[ustart[e] + q] & mask) == (pk[ustart[f] + q] & mask);
      if (same) { r = f; break; }
    }
    rep[e] = r;
  }
}

__global__ void co_net_weight_kernel(uint32_t m, const uint32_t* __restrict__ size, const uint32_t* __restrict__ rep,
                                     const int32_t* __restrict__ nw, int64_t* __restrict__ acc,
                                     uint32_t* __restrict__ survivor) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const bool valid = size[e] > 1;
    const bool surv = valid && rep[e] == e;
    survivor[e] = surv ? 1u : 0u;
    if (valid)
      atomicAdd(reinterpret_cast<unsigned long long*>(&acc[rep[e]]), static_cast<unsigned long long>(int64_t(nw[e])));
  }
}

__global__ void co_net_offsets_kernel(uint32_t m, const uint32_t* __restrict__ survivor,
                                      const uint32_t* __restrict__ newid, const uint32_t* __restrict__ size,
                                      const int64_t* __restrict__ acc, uint32_t* __restrict__ csize,
                                      int32_t* __restrict__ cnw) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    if (!survivor[e]) continue;
    csize[newid[e]] = size[e];
    cnw[newid[e]] = int32_t(acc[e]);
  }
}

__global__ void co_copy_pins_kernel(uint32_t m, const uint32_t* __restrict__ survivor,
                                    const uint32_t* __restrict__ newid, const uint32_t* __restrict__ ustart,
                                    const uint32_t* __restrict__ size, const uint64_t* __restrict__ pk, int bits,
                                    const uint32_t* __restrict__ coff, uint32_t* __restrict__ cpins) {
  const uint64_t mask = (uint64_t(1) << bits) - 1;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    if (!survivor[e]) continue;
    const uint32_t c = newid[e], sz = size[e], src = ustart[e], dst = coff[c];
    for (uint32_t q = 0; q < sz; ++q) cpins[dst + q] = uint32_t(pk[src + q] & mask);
  }
}

__global__ void co_compact_keys_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flag,
                                       const uint32_t* __restrict__ pos, uint64_t p, uint64_t* __restrict__ out,
                                       uint32_t* __restrict__ ustart, int bits) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < p; i += uint64_t(gridDim.x) * blockDim.x) {
    if (!flag[i]) continue;
    const uint32_t j = pos[i];
    out[j] = keys[i];
    const uint32_t e = uint32_t(keys[i] >> bits);
    if (i == 0 || uint32_t(keys[i - 1] >> bits) != e) ustart[e] = j;
  }
}

}  // namespace mkp

// ------------------------------------------------------------ host side --
namespace mkp {

// work arrays of the heavy-proposer rating, kept across sub-rounds
struct HeavyRating {
  DevBuf<uint64_t> tsize, isize, ioff, toff, inc_cnt, inc_off, cval;
  DevBuf<uint32_t> inc_net, inc_own, keys, ct, over1, over2, ocount;
  DevBuf<uint64_t> vals;
  DevBuf<int> cun;
};

// rating of the heavy proposers (list, *count on the device), one block per
// proposer with the map in LDS: on levels of at most 2 * kHbDense clusters
// by cluster id (hr_dense_kernel); otherwise by hash, first in one pass with
// maps of 4096 slots (blocks of 256, three per CU), then in up to
// `rating_block_passes` passes with maps of 8192 slots (blocks of 1024); the
// rest in the global arena (rating_block_passes = 0: every one there)
void rate_heavy(Context& c, const CoArgs& a, const uint32_t* list, const uint32_t* count_dev, const uint32_t* dclus,
                const uint32_t* dcount, HeavyRating& w) {
  const hipStream_t s = c.stream;
  const uint32_t max_passes = uint32_t(c.tune.rating_block_passes);
  if (max_passes > 0 && a.n <= 2 * kHbDense) {
    TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_dense_kernel<1024>");
    hr_dense_kernel<1024><<<unsigned(c.num_cus), 1024, 0, s>>>(a, list, count_dev, dclus, dcount);
    MKP_LAUNCH_CHECK();
    return;
  }
  if (max_passes > 0) {
    w.over1.ensure(a.n);
    w.over2.ensure(a.n);
    w.ocount.ensure(2);
    fill_async<uint32_t>(w.ocount.p, 0u, 2, s);
    {
      TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_hash_kernel<12, 256>");
      hr_hash_kernel<12, 256><<<unsigned(c.num_cus) * 3, 256, 0, s>>>(a, list, count_dev, dclus, dcount, 1u,
                                                                      w.over1.p, w.ocount.p);
    }
    {
      TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_hash_kernel<13, 1024>");
      hr_hash_kernel<13, 1024><<<unsigned(c.num_cus), 1024, 0, s>>>(a, w.over1.p, w.ocount.p, dclus, dcount,
                                                                    max_passes, w.over2.p, w.ocount.p + 1);
    }
    MKP_LAUNCH_CHECK();
    list = w.over2.p;
    count_dev = w.ocount.p + 1;
  }
  uint32_t* pin = reinterpret_cast<uint32_t*>(c.pinned);
  MKP_HIP(hipMemcpyAsync(pin, count_dev, 4, hipMemcpyDeviceToHost, s));
  c.sync();
  const uint32_t hb = *pin;
  if (!hb) return;
  w.tsize.ensure(hb);
  w.isize.ensure(hb);
  w.ioff.ensure(size_t(hb) + 1);
  w.toff.ensure(size_t(hb) + 1);
  HrArgs h{};
  h.list = list; h.count = hb; h.tsize = w.tsize.p; h.isize = w.isize.p; h.ioff = w.ioff.p; h.toff = w.toff.p;
  {
    TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_prep_kernel");
    hr_prep_kernel<<<(hb + 3) / 4, 256, 0, s>>>(a, h, dcount);
    MKP_LAUNCH_CHECK();
  }
  exclusive_scan_u64(c, w.isize.p, w.ioff.p, hb);
  exclusive_scan_u64(c, w.tsize.p, w.toff.p, hb);
  uint64_t* pin64 = reinterpret_cast<uint64_t*>(c.pinned);
  MKP_HIP(hipMemcpyAsync(pin64, w.ioff.p + hb, 8, hipMemcpyDeviceToHost, s));
  MKP_HIP(hipMemcpyAsync(pin64 + 1, w.toff.p + hb, 8, hipMemcpyDeviceToHost, s));
  c.sync();
  const uint64_t total_inc = pin64[0], slots = pin64[1];
  MKP_CHECK(slots < (uint64_t(1) << 32), "heavy rating tables exceed 2^32 slots");
  w.inc_net.ensure(std::max<uint64_t>(total_inc, 1));
  w.inc_own.ensure(std::max<uint64_t>(total_inc, 1));
  w.inc_cnt.ensure(std::max<uint64_t>(total_inc, 1));
  w.inc_off.ensure(total_inc + 1);
  h.inc_net = w.inc_net.p; h.inc_own = w.inc_own.p; h.inc_cnt = w.inc_cnt.p; h.inc_off = w.inc_off.p;
  if (total_inc) {
    TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_inc_kernel");
    hr_inc_kernel<<<grid_for(total_inc, 256, 65536), 256, 0, s>>>(a, h, total_inc, dcount);
    MKP_LAUNCH_CHECK();
  }
  exclusive_scan_u64(c, w.inc_cnt.p, w.inc_off.p, total_inc);
  MKP_HIP(hipMemcpyAsync(pin64, w.inc_off.p + total_inc, 8, hipMemcpyDeviceToHost, s));
  c.sync();
  const uint64_t total_items = pin64[0];
  w.keys.ensure(slots);
  w.vals.ensure(slots);
  h.keys = w.keys.p; h.vals = w.vals.p;
  {
    TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_clear_kernel");
    hr_clear_kernel<<<grid_for(slots, 256, 65536), 256, 0, s>>>(h, slots);
  }
  if (total_items) {
    TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_insert_kernel");
    hr_insert_kernel<<<grid_for(total_items, kHrItems, 65536), 256, 0, s>>>(a, h, total_items, total_inc, dclus,
                                                                             dcount);
  }
  MKP_LAUNCH_CHECK();
  const uint64_t chunks = slots / kHrChunk;
  w.cval.ensure(chunks);
  w.cun.ensure(chunks);
  w.ct.ensure(chunks);
  {
    TimedRegion tr(c.prof, s, kFamRating, 0.0, "hr_scan_kernel");
    hr_scan_kernel<<<unsigned(std::min<uint64_t>(chunks, uint64_t(c.num_cus) * 16)), 256, 0, s>>>(a, h, chunks,
                                                                                                   w.cval.p, w.cun.p,
                                                                                                   w.ct.p);
  }
  hr_decide_kernel<<<grid_for(hb, 256, 4096), 256, 0, s>>>(a, h, w.cval.p, w.cun.p, w.ct.p);
  MKP_LAUNCH_CHECK();
}

// One clustering pass on hg; cid (device, n) receives the representatives.
ClusterResult cluster_pass(mkp_hypergraph* hg, const mkp_coarsen_params& prm, int pass, uint32_t* cid_out) {
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t n = hg->n;
  ClusterResult res;
  res.nodes_after = n;
  // the per-net cluster set of the rating kernel holds 2048 slots
  MKP_CHECK(prm.ignore_net_size >= 1 && prm.ignore_net_size <= 1025, "ignore_net_size must be in [1, 1025]");
  if (n == 0) return res;
  DevBuf<int64_t> cw(n), jw(n), visits(1);
  DevBuf<uint8_t> matched(n), role(n);
  DevBuf<uint32_t> target(n), list_a(n), list_b(n), list_c(n), counts(8), olist(n);
  // per sub-round distinct clusters of the proposers' nets
  const uint32_t m = hg->m;
  DevBuf<uint32_t> nmark(std::max<uint32_t>(m, 1)), mlist(std::max<uint32_t>(m, 1)), dcount(std::max<uint32_t>(m, 1)),
      dclus(std::max<uint64_t>(hg->p, 1));
  fill_async<uint32_t>(nmark.p, 0u, std::max<uint32_t>(m, 1), s);
  DevBuf<uint64_t> rating(n), ub(n);
  co_init_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(n, hg->vw.p, cid_out, cw.p, matched.p);
  DevBuf<uint64_t> ub_max(1);
  fill_async<uint64_t>(ub_max.p, 0ull, 1, s);
  if (hg->p > 64ull * n)  // long incidence lists: a wave per vertex
    co_ub_kernel<64><<<grid_for(n, 4, 65536), 256, 0, s>>>(hg->voff.p, hg->inc.p, hg->noff.p, n, prm.ignore_net_size,
                                                           ub.p, ub_max.p);
  else if (hg->p > 16ull * n)
    co_ub_kernel<8><<<grid_for(n, 32, 65536), 256, 0, s>>>(hg->voff.p, hg->inc.p, hg->noff.p, n, prm.ignore_net_size,
                                                           ub.p, ub_max.p);
  else
    co_ub_kernel<1><<<grid_for(n, 256, 65536), 256, 0, s>>>(hg->voff.p, hg->inc.p, hg->noff.p, n, prm.ignore_net_size,
                                                            ub.p, ub_max.p);
  fill_async<int64_t>(visits.p, 0, 1, s);
  MKP_LAUNCH_CHECK();
  // the heavy tier runs only when some proposer has more than 1024 rating keys
  uint64_t max_ub = 0;
  MKP_HIP(hipMemcpyAsync(c.pinned, ub_max.p, 8, hipMemcpyDeviceToHost, s));
  c.sync();
  max_ub = uint64_t(c.pinned[0]);
  HeavyRating hw;
  const uint32_t limit = std::max<uint32_t>(uint32_t(double(n) / prm.max_shrink_factor), prm.contraction_limit);
  uint32_t current = n;
  const int S = std::max(1, prm.sub_rounds);
  const uint32_t salt = bucket_salt(prm.seed, kCoBucketTag | uint64_t(pass));
  CoArgs a;
  a.n = n; a.voff = hg->voff.p; a.inc = hg->inc.p; a.noff = hg->noff.p; a.pins = hg->pins.p; a.nw = hg->nw.p;
  a.comm = hg->comm.p; a.ub = ub.p; a.cid = cid_out; a.cw = cw.p; a.matched = matched.p; a.target = target.p;
  a.rating = rating.p; a.ignore = prm.ignore_net_size; a.sampling = prm.sampling; a.max_w = prm.max_node_weight;
  a.seed = prm.seed; a.prefer_unmatched = 1; a.visits = visits.p;
  for (int sr = 0; sr < S; ++sr) {
    if (current <= limit) break;
    a.tag = (uint64_t(pass) << 8) | uint64_t(sr);
    fill_async<uint32_t>(counts.p, 0u, 8, s);
    co_select_kernel<<<(n + 255) / 256, 256, 0, s>>>(n, matched.p, ub.p, salt, uint32_t(S), uint32_t(sr), target.p,
                                                     list_a.p, list_b.p, list_c.p, counts.p);
    MKP_LAUNCH_CHECK();
    {
      const uint32_t stamp = uint32_t(sr) + 1;  // nmark is cleared once per pass
      const unsigned mg = unsigned(c.num_cus) * 8;
      {
        TimedRegion tr(c.prof, s, kFamRating, 0.0, "co_mark_nets_kernel");
        co_mark_nets_kernel<8><<<mg, 256, 0, s>>>(a, list_c.p, counts.p + 6, stamp, nmark.p, mlist.p, counts.p + 7);
        co_mark_nets_kernel<16><<<mg, 256, 0, s>>>(a, list_a.p, counts.p + 0, stamp, nmark.p, mlist.p, counts.p + 7);
        co_mark_nets_kernel<64><<<mg, 256, 0, s>>>(a, list_b.p, counts.p + 1, stamp, nmark.p, mlist.p, counts.p + 7);
      }
      {
        TimedRegion tr(c.prof, s, kFamRating, 0.0, "co_dedup_small_kernel");
        co_dedup_small_kernel<<<unsigned(c.num_cus) * 16, 256, 0, s>>>(a, mlist.p, counts.p + 7, dclus.p, dcount.p);
      }
      {
        TimedRegion tr(c.prof, s, kFamRating, 0.0, "co_dedup_wave_kernel");
        co_dedup_wave_kernel<<<unsigned(c.num_cus) * 8, 256, 0, s>>>(a, mlist.p, counts.p + 7, dclus.p, dcount.p);
      }
      MKP_LAUNCH_CHECK();
      {
        TimedRegion tr(c.prof, s, kFamRating, 0.0, "co_rate_group_kernel<16, 7, 256>");
        co_rate_group_kernel<16, 7, 256><<<unsigned(c.num_cus) * 16, 256, 0, s>>>(a, list_c.p, counts.p + 6, dclus.p,
                                                                                 dcount.p);
      }
      {
        TimedRegion tr(c.prof, s, kFamRating, 0.0, "co_rate_group_kernel<64, 11, 64>");
        co_rate_group_kernel<64, 11, 64><<<unsigned(c.num_cus) * 16, 64, 0, s>>>(a, list_a.p, counts.p + 0, dclus.p,
                                                                                dcount.p);
      }
      MKP_LAUNCH_CHECK();
      if (max_ub > 1024) rate_heavy(c, a, list_b.p, counts.p + 1, dclus.p, dcount.p, hw);
    }
    fill_async<int64_t>(jw.p, 0, n, s);
    co_resolve_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(n, target.p, hg->vw.p, jw.p, role.p);
    co_overflow_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(n, target.p, role.p, cw.p, jw.p,
                                                               prm.max_node_weight, olist.p, counts.p + 4);
    MKP_LAUNCH_CHECK();
    uint32_t oc = 0;
    MKP_HIP(hipMemcpyAsync(&oc, counts.p + 4, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    if (oc) {
      // joiners of targets whose capacity would be exceeded: admit in
      // (rating desc, hash, id) order up to the limit (host side; rare)
      std::vector<uint32_t> ol(oc);
      MKP_HIP(hipMemcpyAsync(ol.data(), olist.p, size_t(oc) * 4, hipMemcpyDeviceToHost, s));
      std::vector<uint32_t> tg(n);
      std::vector<uint64_t> rt(n);
      std::vector<int64_t> cwh(n);
      std::vector<int32_t> vwh(n);
      MKP_HIP(hipMemcpyAsync(tg.data(), target.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
      MKP_HIP(hipMemcpyAsync(rt.data(), rating.p, size_t(n) * 8, hipMemcpyDeviceToHost, s));
      MKP_HIP(hipMemcpyAsync(cwh.data(), cw.p, size_t(n) * 8, hipMemcpyDeviceToHost, s));
      MKP_HIP(hipMemcpyAsync(vwh.data(), hg->vw.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
      c.sync();
      std::map<uint32_t, std::vector<uint32_t>> groups;
      for (uint32_t u : ol) groups[tg[u]].push_back(u);
      std::vector<uint32_t> rejected;
      for (auto& kv : groups) {
        std::vector<uint32_t>& J = kv.second;
        std::sort(J.begin(), J.end(), [&](uint32_t x, uint32_t y) {
          if (rt[x] != rt[y]) return rt[x] > rt[y];
          const uint64_t hx = hash4(prm.seed, kCoPrioTag | a.tag, x, 0), hy = hash4(prm.seed, kCoPrioTag | a.tag, y, 0);
          if (hx != hy) return hx < hy;
          return x < y;
        });
        int64_t cum = cwh[kv.first];
        bool full = false;
        for (uint32_t u : J) {
          if (!full && cum + vwh[u] <= prm.max_node_weight) cum += vwh[u];
          else { full = true; rejected.push_back(u); }
        }
      }
      if (!rejected.empty()) {
        DevBuf<uint32_t> rej(rejected.size());
        MKP_HIP(hipMemcpyAsync(rej.p, rejected.data(), rejected.size() * 4, hipMemcpyHostToDevice, s));
        // role 3 = rejected joiner
        std::vector<uint8_t> roles(n);
        MKP_HIP(hipMemcpyAsync(roles.data(), role.p, n, hipMemcpyDeviceToHost, s));
        c.sync();
        for (uint32_t u : rejected) roles[u] = 3;
        MKP_HIP(hipMemcpyAsync(role.p, roles.data(), n, hipMemcpyHostToDevice, s));
      }
    }
    fill_async<uint32_t>(counts.p + 5, 0u, 1, s);
    co_apply_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(n, target.p, role.p, hg->vw.p, cid_out, cw.p, matched.p,
                                                            counts.p + 5);
    MKP_LAUNCH_CHECK();
    uint32_t contracted = 0;
    MKP_HIP(hipMemcpyAsync(&contracted, counts.p + 5, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    current -= contracted;
  }
  MKP_HIP(hipMemcpyAsync(&res.visits, visits.p, 8, hipMemcpyDeviceToHost, s));
  c.sync();
  res.nodes_after = current;
  if (c.prof.enabled) {
    // pin id + cluster gather for the adaptive-size sweep and the scoring
    // sweep, net id/weight/size per incidence, offsets and weight per vertex
    c.prof.bytes[kFamRating] += 8.0 * double(res.visits) + 12.0 * double(hg->p) + 12.0 * n;
  }
  return res;
}

// StaticHypergraph::contract in canonical form; returns the coarse
// hypergraph and writes the fine -> coarse map into cmap (device, n).
mkp_hypergraph* contract(mkp_hypergraph* hg, const uint32_t* cid, uint32_t* cmap) {
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t n = hg->n, m = hg->m;
  const uint64_t p = hg->p;
  TimedRegion tr(c.prof, s, kFamContraction, 12.0 * double(p) + 8.0 * double(p) + 16.0 * m + 16.0 * n);
  auto* cg = new mkp_hypergraph();
  try {
    cg->ctx = hg->ctx;
    DevBuf<uint32_t> used(std::max<uint32_t>(n, 1)), map(size_t(n) + 1), last(std::max<uint32_t>(n, 1));
    fill_async<uint32_t>(used.p, 0u, n, s);
    co_mark_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(cid, n, used.p);
    exclusive_scan_u32(c, used.p, map.p, n);
    uint32_t nc = 0;
    MKP_HIP(hipMemcpyAsync(&nc, map.p + n, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    cg->n = nc;
    cg->vw.alloc(std::max<uint32_t>(nc, 1));
    cg->comm.alloc(std::max<uint32_t>(nc, 1));
    fill_async<int32_t>(cg->vw.p, 0, nc, s);
    fill_async<uint32_t>(last.p, 0u, nc, s);
    co_cmap_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(cid, map.p, n, hg->vw.p, cmap, cg->vw.p, last.p);
    if (nc) co_comm_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(nc, last.p, hg->comm.p, cg->comm.p);
    MKP_LAUNCH_CHECK();
    cg->total_w = hg->total_w;
    int bits = 1;
    while ((uint64_t(1) << bits) < nc) ++bits;
    int nbits = 1;
    while ((uint64_t(1) << nbits) < m) ++nbits;
    MKP_CHECK(bits + nbits <= 64, "hypergraph too large for 64-bit pin keys");
    // stage 2: sorted unique coarse pins per net
    DevBuf<uint64_t> keys(std::max<uint64_t>(p, 1)), skeys(std::max<uint64_t>(p, 1)), ukeys(std::max<uint64_t>(p, 1));
    DevBuf<uint32_t> flag(std::max<uint64_t>(p, 1)), pos(p + 1), size(std::max<uint32_t>(m, 1)),
        ustart(std::max<uint32_t>(m, 1));
    DevBuf<uint64_t> foot(std::max<uint32_t>(m, 1));
    fill_async<uint32_t>(size.p, 0u, m, s);
    fill_async<uint64_t>(foot.p, 0ull, m, s);
    uint32_t U = 0;
    if (p) {
      co_pin_keys_kernel<<<grid_for(p, 256, 65536), 256, 0, s>>>(hg->pins.p, hg->pin_net.p, cmap, p, bits, keys.p);
      MKP_LAUNCH_CHECK();
      sort_keys_u64(c, keys.p, skeys.p, p, 0, bits + nbits);
      co_unique_kernel<<<grid_for(p, 256, 65536), 256, 0, s>>>(skeys.p, p, bits, flag.p, size.p, foot.p);
      MKP_LAUNCH_CHECK();
      exclusive_scan_u32(c, flag.p, pos.p, p);
      co_compact_keys_kernel<<<grid_for(p, 256, 65536), 256, 0, s>>>(skeys.p, flag.p, pos.p, p, ukeys.p, ustart.p,
                                                                     bits);
      MKP_LAUNCH_CHECK();
      MKP_HIP(hipMemcpyAsync(&U, pos.p + p, 4, hipMemcpyDeviceToHost, s));
    }
    // stage 3: parallel nets
    DevBuf<uint64_t> nkeys(std::max<uint32_t>(m, 1)), snkeys(std::max<uint32_t>(m, 1));
    DevBuf<uint32_t> nvals(std::max<uint32_t>(m, 1)), snvals(std::max<uint32_t>(m, 1)), rep(std::max<uint32_t>(m, 1)),
        survivor(std::max<uint32_t>(m, 1)), newid(size_t(m) + 1), nvalid(1);
    DevBuf<int64_t> acc(std::max<uint32_t>(m, 1));
    fill_async<uint32_t>(nvalid.p, 0u, 1, s);
    fill_async<int64_t>(acc.p, 0, m, s);
    uint32_t nv = 0;
    if (m) {
      co_net_keys_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(m, size.p, foot.p, nkeys.p, nvals.p, nvalid.p);
      MKP_LAUNCH_CHECK();
      sort_pairs_u64_u32(c, nkeys.p, snkeys.p, nvals.p, snvals.p, m, 0, 64);
      MKP_HIP(hipMemcpyAsync(&nv, nvalid.p, 4, hipMemcpyDeviceToHost, s));
      c.sync();
      if (nv) {
        DevBuf<uint64_t> rs(nv), rsm(nv);
        co_run_start_kernel<<<grid_for(nv, 256, 65536), 256, 0, s>>>(snkeys.p, nv, rs.p);
        MKP_LAUNCH_CHECK();
        inclusive_max_scan_u64(c, rs.p, rsm.p, nv);
        co_parallel_kernel<<<grid_for(nv, 256, 65536), 256, 0, s>>>(snvals.p, nv, rsm.p, size.p, ustart.p, ukeys.p,
                                                                    bits, rep.p);
      }
      co_net_weight_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(m, size.p, rep.p, hg->nw.p, acc.p, survivor.p);
      MKP_LAUNCH_CHECK();
      exclusive_scan_u32(c, survivor.p, newid.p, m);
    } else {
      fill_async<uint32_t>(newid.p, 0u, 1, s);
    }
    uint32_t mc = 0;
    MKP_HIP(hipMemcpyAsync(&mc, newid.p + m, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    // stage 4: coarse CSR in original net order
    cg->m = mc;
    cg->nw.alloc(std::max<uint32_t>(mc, 1));
    DevBuf<uint32_t> csize(std::max<uint32_t>(mc, 1));
    cg->noff.alloc(size_t(mc) + 1);
    if (m) {
      co_net_offsets_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(m, survivor.p, newid.p, size.p, acc.p, csize.p,
                                                                    cg->nw.p);
      MKP_LAUNCH_CHECK();
    }
    exclusive_scan_u32(c, csize.p, cg->noff.p, mc);
    uint32_t cp = 0;
    MKP_HIP(hipMemcpyAsync(&cp, cg->noff.p + mc, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    cg->p = cp;
    cg->pins.alloc(std::max<uint32_t>(cp, 1));
    if (m && cp) {
      co_copy_pins_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(m, survivor.p, newid.p, ustart.p, size.p, ukeys.p,
                                                                  bits, cg->noff.p, cg->pins.p);
      MKP_LAUNCH_CHECK();
    }
    // max net size on the host from the offsets
    {
      std::vector<uint32_t> off(size_t(mc) + 1);
      MKP_HIP(hipMemcpyAsync(off.data(), cg->noff.p, off.size() * 4, hipMemcpyDeviceToHost, s));
      c.sync();
      uint32_t me = 0;
      for (uint32_t e = 0; e < mc; ++e) me = std::max(me, off[e + 1] - off[e]);
      cg->max_esize = me;
    }
    build_incidence(cg);
  } catch (...) {
    delete cg;
    throw;
  }
  return cg;
}

}  // namespace mkp

using namespace mkp;

extern "C" {

int mkp_cluster_pass(mkp_hypergraph* hg, const mkp_coarsen_params* prm, int32_t pass, uint32_t* clusters,
                     uint32_t* nodes_after, int64_t* pin_visits) {
  return guarded([&] {
    MKP_CHECK(hg && prm && clusters, "null argument");
    MKP_CHECK(prm->sub_rounds >= 1 && prm->sub_rounds <= 255 && prm->max_shrink_factor > 0, "bad parameters");
    Context& c = hg->ctx->c;
    Bind bind_(c);
    DevBuf<uint32_t> cid(std::max<uint32_t>(hg->n, 1));
    ClusterResult r = cluster_pass(hg, *prm, pass, cid.p);
    MKP_HIP(hipMemcpyAsync(clusters, cid.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    if (nodes_after) *nodes_after = r.nodes_after;
    if (pin_visits) *pin_visits = r.visits;
  });
}

int mkp_hypergraph_contract(mkp_hypergraph* hg, const uint32_t* clusters, mkp_hypergraph** coarse, uint32_t* cmap) {
  return guarded([&] {
    MKP_CHECK(hg && clusters && coarse, "null argument");
    for (uint32_t u = 0; u < hg->n; ++u) MKP_CHECK(clusters[u] < hg->n, "cluster id out of range");
    Context& c = hg->ctx->c;
    Bind bind_(c);
    DevBuf<uint32_t> cid(std::max<uint32_t>(hg->n, 1)), cm(std::max<uint32_t>(hg->n, 1));
    MKP_HIP(hipMemcpyAsync(cid.p, clusters, size_t(hg->n) * 4, hipMemcpyHostToDevice, c.stream));
    *coarse = contract(hg, cid.p, cm.p);
    if (cmap) MKP_HIP(hipMemcpyAsync(cmap, cm.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

}  // extern "C"
