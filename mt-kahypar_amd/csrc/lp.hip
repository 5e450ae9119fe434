// km1 label propagation in synchronous sub-rounds.
//
// Reference: LabelPropagationRefiner (label_propagation_refiner.h:75-143,
// label_propagation_refiner.cpp:33-164), Km1Policy::computeMaxGainMoveImpl
// (gain_policy.h:108-167), changeNodePart (partitioned_hypergraph.h:539-563)
// and km1Delta (hypergraph_common.h:133-140).
//
// One sub-round (r, s):
//   gain      every active vertex of bucket s computes its Km1Policy move on
//             the state at the start of the sub-round (coin flips replaced by
//             hash bits of (seed, r, s, u, block));
//   admit     candidates into a block whose capacity would be exceeded are
//             admitted in (gain, hash, id) order up to the capacity;
//   apply     block ids, block weights, packed pin counts (64-bit atomics on
//             the packed words) and connectivity bits (atomic xor on 0<->1
//             transitions, as connectivity_set.h:143-150 does);
//   verify    each applied move is kept iff its km1 delta on the post-move
//             state is <= 0 (label_propagation_refiner.h:100-137), otherwise
//             moved back; a sub-round with positive total delta is undone;
//   activate  kept moves activate their own vertex and the pins of incident
//             nets with at most 100 pins for the next round.
#include <algorithm>

#include "internal.h"

namespace mkp {

constexpr uint32_t kHighDegree = 100000;   // partitioned_hypergraph.h:67
// thread-per-vertex gains (k <= 32) up to this degree: a wave runs as long
// as its heaviest vertex, and at coarse levels degrees spread from tens to
// thousands, so heavier vertices take a wave each
constexpr uint32_t kThreadDegreeLimit = 64;
// group-per-vertex gains (32 < k <= 128) up to this degree
constexpr uint32_t kGroupDegreeLimit = 512;
constexpr int kGainThreads = 128;
constexpr int kLpStripes = 32, kLpStripeWords = 16;  // striped counters (see stripe_add)

struct LpWork {
  uint32_t n = 0;
  int k = 0;
  DevBuf<uint32_t> active, active_next, flag, pos, cand_flag, cand_u, heavy;
  DevBuf<int32_t> cand_to_a, cand_gain_a, cand_from, cand_to, cand_gain;
  DevBuf<uint8_t> admit, keep;
  DevBuf<uint64_t> keys, keys_sorted;
  DevBuf<uint32_t> vals, vals_sorted;
  DevBuf<int64_t> wsort, wscan, seg_base, in_w, max_w, perfect_w, counters;
  DevBuf<uint8_t> overflow;
  DevBuf<uint32_t> net_stamp;  // last round stamp that expanded each net
  DevBuf<uint64_t> bkeys, bkeys_sorted;
  DevBuf<int64_t> bcount;
  DevBuf<int64_t> stripes;  // striped counter copies (single hypergraph)
  DevBuf<int64_t> rb;       // read-back words of a sub-round (2k + 3)
  uint32_t m = 0, stamp = 0;
  void ensure_nets(uint32_t mm, hipStream_t s) {
    if (mm <= m && net_stamp.p) return;
    m = std::max(mm, m);
    net_stamp.alloc(std::max<uint32_t>(m, 1));
    fill_async<uint32_t>(net_stamp.p, 0u, std::max<uint32_t>(m, 1), s);
    stamp = 0;
  }
  uint32_t members = 0;
  // nn vertices, kk blocks (all members together), mm member hypergraphs
  void ensure(uint32_t nn, int kk, uint32_t mm = 1) {
    if (mm > members) {
      members = mm;
      counters.alloc(size_t(kGcStrideWords) * members);
    }
    if (!stripes.p) stripes.alloc(size_t(kLpStripes) * kLpStripeWords);
    if (nn <= n && kk <= k) return;
    n = std::max(nn, n);
    k = std::max(kk, k);
    const size_t N = std::max<uint32_t>(n, 1);
    active.alloc(N); active_next.alloc(N); flag.alloc(N); pos.alloc(N + 1); cand_flag.alloc(N);
    cand_u.alloc(N); heavy.alloc(N);
    cand_to_a.alloc(N); cand_gain_a.alloc(N); cand_from.alloc(N); cand_to.alloc(N); cand_gain.alloc(N);
    admit.alloc(N); keep.alloc(N);
    keys.alloc(N); keys_sorted.alloc(N); vals.alloc(N); vals_sorted.alloc(N);
    wsort.alloc(N); wscan.alloc(N); bkeys.alloc(N); bkeys_sorted.alloc(N); bcount.alloc(4 * 256);
    seg_base.alloc(k); in_w.alloc(k); max_w.alloc(k); perfect_w.alloc(k); overflow.alloc(k);
    rb.alloc(2 * size_t(k) + 3);
  }
  static constexpr int kGcStrideWords = 16;
};

// counters, kGcStride words per member hypergraph (one member unless batched):
//   0 = delta of the current sub-round, 1 = visits, 2 = reverts,
//   3 = heavy-list length (member 0's word, shared), 4 = moves,
//   5 = vertices swept by the gain kernels, 6 = stopped (rolled back),
//   7 = total delta, 8 = rounds, 9 = active in the current round,
//   10 = kept moves of the current sub-round (batched form: counted per
//        member by the verify kernel, added into word 4 by the check kernel),
//   11 = incidences of the current sub-round's candidates (member 0's word,
//        shared; profiling only)
constexpr int kGcStride = LpWork::kGcStrideWords;

struct LpArgs {
  const uint32_t* voff;
  const uint32_t* inc;
  const uint32_t* noff;
  const uint32_t* pins;
  const int32_t* vw;
  const int32_t* nw;
  int32_t* part;
  int64_t* part_w;
  uint64_t* pc;
  const int64_t* max_w;
  const int64_t* perfect_w;
  PcLayout L;
  int vcs;
  int k;
  int rebalancing;
  uint64_t seed;
  uint64_t rs;       // (round << 8) | sub-round
  uint32_t round;
  int buckets;
  int bucket;
  uint32_t bucket_salt;
  int64_t* counters;
  // single hypergraph: words 0, 1, 2, 4 and 5 of the counters are added into
  // striped copies and folded into the counters by lp_check_kernel (see
  // stripe_add)
  int64_t* stripes;
  // batch of independent hypergraphs stored as one disjoint union: block ids
  // stay local to a member, its block weights live at gkbase[g] + b, and every
  // hash uses the member's seed and local vertex id, so each member evolves
  // exactly as it would alone (vgraph == nullptr: a single hypergraph)
  const uint32_t* vgraph;
  const uint32_t* gvbase;
  const uint32_t* gkbase;
  const int32_t* gk;
  const uint64_t* gseed;
};

struct LpMember {
  uint32_t g, lu, kb;
  int k;
  uint64_t seed;
};
__device__ inline LpMember member_of(const LpArgs& a, uint32_t u) {
  if (!a.vgraph) return LpMember{0, u, 0, a.k, a.seed};
  const uint32_t g = a.vgraph[u];
  return LpMember{g, u - a.gvbase[g], a.gkbase[g], a.gk[g], a.gseed[g]};
}
__device__ inline int64_t* member_counters(const LpArgs& a, uint32_t g) { return a.counters + kGcStride * g; }

// Device atomics execute in memory, and atomics on one address one after
// another (about 11 ns each on MI355X), so a counter that every wave of a
// launch updates is kept in kLpStripes copies 128 bytes apart, one per
// residue of the block index, and summed once per sub-round.
constexpr int kLpFoldWords[5] = {0, 1, 2, 4, 5};
__device__ inline void stripe_add(int64_t* stripes, int word, int64_t v) {
  if (v)
    atomicAdd(reinterpret_cast<unsigned long long*>(&stripes[(blockIdx.x % kLpStripes) * kLpStripeWords + word]),
              static_cast<unsigned long long>(v));
}
// adds the stripes into the counters and clears them (one thread)
__device__ inline void stripe_fold(int64_t* counters, int64_t* stripes) {
  for (int w : kLpFoldWords) {
    int64_t sum = 0;
    for (int t = 0; t < kLpStripes; ++t) {
      sum += stripes[t * kLpStripeWords + w];
      stripes[t * kLpStripeWords + w] = 0;
    }
    counters[w] += sum;
  }
}
__device__ inline void counter_add(int64_t* c, int64_t v) {
  if (v) atomicAdd(reinterpret_cast<unsigned long long*>(c), static_cast<unsigned long long>(v));
}

// Adds v to arr[idx] for every lane with on set.  Up to `passes` times, the
// lanes that share the index of the first pending lane add their sum with one
// atomic (consecutive moves and vertices mostly share a member, and a move
// often a block); the rest add alone.  Every lane of the wave calls it.
__device__ inline void wave_add_i64(int64_t* arr, bool on, uint64_t idx, int64_t v, int passes) {
  const int lane = threadIdx.x & 63;
  bool pending = on && v != 0;
  for (int it = 0; it < passes; ++it) {
    const uint64_t act = __ballot(pending);
    if (!act) return;
    const int leader = __ffsll(static_cast<long long>(act)) - 1;
    const uint64_t li = __shfl(idx, leader, 64);
    const bool mine = pending && idx == li;
    int64_t sum = mine ? v : 0;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (lane == leader) counter_add(arr + li, sum);
    if (mine) pending = false;
  }
  if (pending) counter_add(arr + idx, v);
}

// counter `slot` of member g (kGcStride words per member)
__device__ inline void member_add_wave(int64_t* counters, bool on, uint32_t g, int slot, int64_t v) {
  wave_add_i64(counters, on, uint64_t(kGcStride) * g + uint64_t(slot), v, 1);
}

__device__ inline uint32_t pc_field(const uint64_t* pc, const PcLayout& L, uint32_t e, int wi, int sh) {
  return uint32_t((pc[uint64_t(e) * L.stride + wi] >> sh) & L.mask);
}

__device__ inline bool in_bucket(const LpArgs& a, uint32_t u) {
  return a.buckets <= 1 || bucket_of(a.bucket_salt, u, uint32_t(a.buckets)) == uint32_t(a.bucket);
}

__device__ inline bool lp_coin(const LpArgs& a, const LpMember& m, int b) {
  return hash4(m.seed, kLpCoinTag | a.rs, m.lu, uint64_t(b)) & 1ull;
}

// perform condition of moveVertex (label_propagation_refiner.h:84-95)
__device__ inline bool lp_perform(const LpArgs& a, const LpMember& m, int from, int to, int32_t gain) {
  if (to == from) return false;
  if (gain < 0) return true;
  const int64_t* pw = a.part_w + m.kb;
  return a.rebalancing && gain == 0 && pw[from] - 1 > pw[to] + 1 && pw[to] < a.perfect_w[m.kb + to];
}

template <typename T>
__device__ inline T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline int wave_min_i32(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ inline int wave_max_i32(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// ------------------------------------------------------------ gain (A) --
// One thread per vertex, K <= 32 scores in registers (fully unrolled), for
// vertices of degree <= kThreadDegreeLimit; heavier vertices are appended to
// the heavy list for the wave kernel.
template <int K>
__global__ void __launch_bounds__(kGainThreads)
lp_gain_thread_kernel(LpArgs a, const uint32_t* __restrict__ active, uint32_t na, uint32_t* __restrict__ cand_flag,
                      int32_t* __restrict__ cand_to, int32_t* __restrict__ cand_gain, uint32_t* __restrict__ heavy) {
  if (!a.vgraph && a.counters[6]) return;  // a rolled-back sub-round stopped the refinement
  int64_t visits = 0, swept = 0;
  const int lane = threadIdx.x & 63;
  // whole waves step together: the per-member counts combine across lanes
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < na; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + uint32_t(lane);
    const bool valid = i < na;
    const uint32_t u = valid ? active[i] : 0u;
    const LpMember mb = member_of(a, u);
    uint32_t flag = 0;
    int64_t member_visits = 0;
    if (valid && in_bucket(a, u) && !(a.vgraph && member_counters(a, mb.g)[6])) {
      const uint32_t jb = a.voff[u], je = a.voff[u + 1];
      const uint32_t deg = je - jb;
      if (deg > kThreadDegreeLimit) {
        if (deg <= kHighDegree) {
          const uint32_t h = atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[3]), 1ull);
          heavy[h] = i;
        }
      } else {
        const int from = a.part[u];
        const int fwi = from / a.L.epw;
        const int fsh = (from - fwi * a.L.epw) * a.L.b;
        int32_t score[K];
#pragma unroll
        for (int b = 0; b < K; ++b) score[b] = 0;
        int32_t internal = 0;
        bool border = false;
        // batches of 8 incident nets: the net ids, then their weight,
        // connectivity and pin-count words, are each fetched together
        constexpr int NB = 8;
        for (uint32_t j0 = jb; j0 < je; j0 += NB) {
          uint32_t e[NB];
#pragma unroll
          for (int t = 0; t < NB; ++t) e[t] = a.inc[j0 + t < je ? j0 + t : jb];
          int32_t w[NB];
          uint64_t cw[NB];
          uint32_t pf[NB];
#pragma unroll
          for (int t = 0; t < NB; ++t) {
            const uint64_t* r = a.pc + uint64_t(e[t]) * a.L.stride;  // one record per net
            w[t] = int32_t(r[a.L.woff]);
            cw[t] = r[a.L.coff];
            pf[t] = uint32_t((r[fwi] >> fsh) & a.L.mask);
          }
#pragma unroll
          for (int t = 0; t < NB; ++t) {
            if (j0 + t >= je) break;
            const uint64_t c = cw[t];
            if (pf[t] > 1) internal += w[t];
            border |= (c & (c - 1)) != 0;
#pragma unroll
            for (int b = 0; b < K; ++b) score[b] -= ((c >> b) & 1ull) ? w[t] : 0;
          }
        }
        if (a.vgraph) {
          member_visits = deg;
        } else {
          visits += deg;
          swept += 1;
        }
        if (border) {
          const int64_t wu = a.vw[u];
          const int64_t* pw = a.part_w + mb.kb;
          const int64_t* mw = a.max_w + mb.kb;
          int to = from;
          int32_t gain = 0;
#pragma unroll
          for (int b = 0; b < K; ++b) {
            if (b < mb.k && b != from) {
              const int32_t sc = score[b] + internal;
              const bool nb = sc < gain || (sc == gain && lp_coin(a, mb, b));
              if (nb && pw[b] + wu <= mw[b]) { to = b; gain = sc; }
            }
          }
          if (lp_perform(a, mb, from, to, gain)) {
            flag = 1;
            cand_to[i] = to;
            cand_gain[i] = gain;
          }
        }
      }
    }
    if (a.vgraph) {
      member_add_wave(a.counters, member_visits > 0, mb.g, 1, member_visits);
      member_add_wave(a.counters, member_visits > 0, mb.g, 5, 1);
    }
    if (valid) cand_flag[i] = flag;
  }
  if (a.vgraph) return;  // batched: counted per member above
  visits = wave_sum(visits);
  swept = wave_sum(swept);
  if ((threadIdx.x & 63) == 0 && visits) {
    stripe_add(a.stripes, 1, visits);
    stripe_add(a.stripes, 5, swept);
  }
}

// ------------------------------------------------------------ gain (G) --
// G lanes per vertex for 32 < k <= KMAX and degree <= kGroupDegreeLimit:
// the lanes take the incident nets and subtract w(e) from the LDS score of
// every block in the net's connectivity set; the Km1Policy scan with its coin
// flips is evaluated as group reductions exactly as in lp_gain_wave_kernel.
// Heavier vertices go to the heavy list for the wave kernel.  Each group's
// scores are private to its wave, so waves never wait for each other.
template <int G, int KMAX>
__global__ void __launch_bounds__(256)
lp_gain_group_kernel(LpArgs a, const uint32_t* __restrict__ active, uint32_t na, uint32_t* __restrict__ cand_flag,
                     int32_t* __restrict__ cand_to, int32_t* __restrict__ cand_gain, uint32_t* __restrict__ heavy) {
  if (!a.vgraph && a.counters[6]) return;  // a rolled-back sub-round stopped the refinement
  constexpr int GPB = 256 / G, GPW = 64 / G;
  __shared__ int32_t scores[GPB * KMAX];
  // one hypergraph: its block weights and bounds, read by every score scan,
  // are staged in LDS once per block (the weights do not change during the
  // gain sweep)
  __shared__ int64_t s_pw[KMAX], s_mw[KMAX];
  if (!a.vgraph) {
    for (int b = threadIdx.x; b < a.k; b += blockDim.x) { s_pw[b] = a.part_w[b]; s_mw[b] = a.max_w[b]; }
    __syncthreads();
  }
  const int grp = threadIdx.x / G, gl = threadIdx.x % G;
  int32_t* sc = scores + grp * KMAX;
  const uint32_t wid = threadIdx.x >> 6, wgrp = (threadIdx.x & 63) / G;
  int64_t visits = 0, swept = 0;
  for (uint32_t base = (blockIdx.x * 4 + wid) * GPW; base < na; base += gridDim.x * 4 * GPW) {
    const uint32_t i = base + wgrp;
    bool live = i < na;
    const uint32_t u = live ? active[i] : 0u;
    const LpMember mb = member_of(a, u);
    if (live && a.vgraph && member_counters(a, mb.g)[6]) {
      if (gl == 0) cand_flag[i] = 0;
      live = false;
    }
    const uint32_t jb = live ? a.voff[u] : 0u, je = live ? a.voff[u + 1] : 0u;
    if (live && je - jb > kGroupDegreeLimit) {
      if (gl == 0) {
        if (je - jb <= kHighDegree) {
          const uint32_t h = atomicAdd(reinterpret_cast<unsigned long long*>(&a.counters[3]), 1ull);
          heavy[h] = i;
        } else {
          cand_flag[i] = 0;
        }
      }
      live = false;
    }
    const int k = mb.k;
    for (int b = gl; b < k; b += G) sc[b] = 0;
    lds_wave_sync();
    const int from = live ? a.part[u] : 0;
    const int fwi = from / a.L.epw;
    const int fsh = (from - fwi * a.L.epw) * a.L.b;
    int32_t internal = 0;
    int border = 0;
    if (live) {
      for (uint32_t j = jb + gl; j < je; j += G) {
        const uint32_t e = a.inc[j];
        const uint64_t* r = a.pc + uint64_t(e) * a.L.stride;
        const int32_t w = int32_t(r[a.L.woff]);
        if (uint32_t((r[fwi] >> fsh) & a.L.mask) > 1) internal += w;
        int lambda = 0;
        for (int ci = 0; ci < a.vcs; ++ci) {
          uint64_t bits = r[a.L.coff + ci];
          lambda += __popcll(bits);
          while (bits) {
            const int b = ci * 64 + __ffsll(static_cast<long long>(bits)) - 1;
            bits &= bits - 1;
            atomicAdd(&sc[b], -w);
          }
        }
        border |= lambda > 1;
      }
    }
    for (int o = G / 2; o > 0; o >>= 1) {
      internal += __shfl_xor(internal, o, G);
      border |= __shfl_xor(border, o, G);
    }
    lds_wave_sync();
    int flag = 0, to = from;
    int32_t gain = 0;
    if (live) {
      if (a.vgraph) {
        if (gl == 0) {
          counter_add(member_counters(a, mb.g) + 1, int64_t(je - jb));
          counter_add(member_counters(a, mb.g) + 5, 1);
        }
      } else if (gl == 0) {
        visits += int64_t(je - jb);
        swept += 1;
      }
    }
    // the group reductions run on every lane of the wave
    const int64_t wu = live ? a.vw[u] : 0;
    const int64_t* pw = a.vgraph ? a.part_w + mb.kb : s_pw;
    const int64_t* mw = a.vgraph ? a.max_w + mb.kb : s_mw;
    const bool scan = live && border;
    const int INF = 0x7fffffff;
    int m = INF;
    if (scan)
      for (int b = gl; b < k; b += G)
        if (b != from && pw[b] + wu <= mw[b]) m = min(m, sc[b] + internal);
    for (int o = G / 2; o > 0; o >>= 1) m = min(m, __shfl_xor(m, o, G));
    int i0 = INF;
    if (scan && m != INF && m < 0)
      for (int b = gl; b < k; b += G)
        if (b != from && pw[b] + wu <= mw[b] && sc[b] + internal == m) i0 = min(i0, b);
    for (int o = G / 2; o > 0; o >>= 1) i0 = min(i0, __shfl_xor(i0, o, G));
    int i1 = -1;
    if (scan && m != INF && m <= 0) {
      // m < 0: last coin-winning block with score m after the first one;
      // m == 0: last coin-winning block with score 0
      const int lo = m < 0 ? i0 : -1;
      for (int b = gl; b < k; b += G)
        if (b > lo && b != from && pw[b] + wu <= mw[b] && sc[b] + internal == m && lp_coin(a, mb, b)) i1 = max(i1, b);
    }
    for (int o = G / 2; o > 0; o >>= 1) i1 = max(i1, __shfl_xor(i1, o, G));
    if (scan && m != INF) {
      if (m < 0) {
        to = i1 >= 0 ? i1 : i0;
        gain = m;
      } else if (m == 0) {
        if (i1 >= 0) to = i1;
        gain = 0;
      }
      flag = lp_perform(a, mb, from, to, gain) ? 1 : 0;
    }
    if (live && gl == 0) {
      cand_flag[i] = uint32_t(flag);
      if (flag) { cand_to[i] = to; cand_gain[i] = gain; }
    }
    lds_wave_sync();
  }
  if (a.vgraph) return;
  visits = wave_sum(visits);
  swept = wave_sum(swept);
  if ((threadIdx.x & 63) == 0 && visits) {
    stripe_add(a.stripes, 1, visits);
    stripe_add(a.stripes, 5, swept);
  }
}

// ------------------------------------------------------------ gain (B) --
// One wave per vertex, any k: lanes split the incident nets and accumulate
// negative scores into a per-wave LDS array with LDS atomics; the sequential
// coin scan of Km1Policy is evaluated as wave reductions: with m the minimum
// eligible score, the result is the last coin-winning block with score m after
// the first block reaching m (m < 0), or the last coin-winning block with
// score 0 (otherwise).
__global__ void __launch_bounds__(64)
lp_gain_wave_kernel(LpArgs a, const uint32_t* __restrict__ active, const uint32_t* __restrict__ list,
                    const int64_t* __restrict__ list_count, uint32_t count_val, uint32_t* __restrict__ cand_flag,
                    int32_t* __restrict__ cand_to, int32_t* __restrict__ cand_gain) {
  if (!a.vgraph && a.counters[6]) return;  // a rolled-back sub-round stopped the refinement
  extern __shared__ int32_t sc[];
  const int lane = threadIdx.x;
  const uint32_t count = list_count ? uint32_t(*list_count) : count_val;
  int64_t visits = 0, swept = 0;
  for (uint32_t item = blockIdx.x; item < count; item += gridDim.x) {
    const uint32_t i = list ? list[item] : item;
    const uint32_t u = active[i];
    const LpMember mb = member_of(a, u);
    if ((!list && !in_bucket(a, u)) || (a.vgraph && member_counters(a, mb.g)[6])) {
      if (lane == 0) cand_flag[i] = 0;
      continue;
    }
    const uint32_t jb = a.voff[u], je = a.voff[u + 1];
    if (je - jb > kHighDegree) {
      if (lane == 0) cand_flag[i] = 0;
      continue;
    }
    const int64_t* pw = a.part_w + mb.kb;
    const int64_t* mw = a.max_w + mb.kb;
    for (int b = lane; b < mb.k; b += 64) sc[b] = 0;
    __syncthreads();
    const int from = a.part[u];
    const int fwi = from / a.L.epw;
    const int fsh = (from - fwi * a.L.epw) * a.L.b;
    int32_t internal = 0;
    int border = 0;
    for (uint32_t j = jb + lane; j < je; j += 64) {
      const uint32_t e = a.inc[j];
      const int32_t w = int32_t(a.pc[uint64_t(e) * a.L.stride + a.L.woff]);
      if (pc_field(a.pc, a.L, e, fwi, fsh) > 1) internal += w;
      int lambda = 0;
      for (int ci = 0; ci < a.vcs; ++ci) {
        uint64_t bits = a.pc[uint64_t(e) * a.L.stride + a.L.coff + ci];
        lambda += __popcll(bits);
        while (bits) {
          const int b = ci * 64 + __ffsll(static_cast<long long>(bits)) - 1;
          bits &= bits - 1;
          atomicAdd(&sc[b], -w);
        }
      }
      border |= lambda > 1;
    }
    if (a.vgraph) {
      if (lane == 0) {
        counter_add(member_counters(a, mb.g) + 1, int64_t(je - jb));
        counter_add(member_counters(a, mb.g) + 5, 1);
      }
    } else {
      visits += (lane == 0) ? int64_t(je - jb) : 0;
      swept += (lane == 0) ? 1 : 0;
    }
    internal = wave_sum(internal);
    border = wave_max_i32(border);
    __syncthreads();
    int flag = 0, to = from;
    int32_t gain = 0;
    if (border) {
      const int64_t wu = a.vw[u];
      const int INF = 0x7fffffff;
      int m = INF;
      for (int b = lane; b < mb.k; b += 64)
        if (b != from && pw[b] + wu <= mw[b]) m = min(m, sc[b] + internal);
      m = wave_min_i32(m);
      if (m != INF && m < 0) {
        int i0 = INF;
        for (int b = lane; b < mb.k; b += 64)
          if (b != from && pw[b] + wu <= mw[b] && sc[b] + internal == m) i0 = min(i0, b);
        i0 = wave_min_i32(i0);
        int i1 = -1;
        for (int b = lane; b < mb.k; b += 64)
          if (b > i0 && b != from && pw[b] + wu <= mw[b] && sc[b] + internal == m && lp_coin(a, mb, b))
            i1 = max(i1, b);
        i1 = wave_max_i32(i1);
        to = i1 >= 0 ? i1 : i0;
        gain = m;
      } else if (m == 0) {
        int i1 = -1;
        for (int b = lane; b < mb.k; b += 64)
          if (b != from && pw[b] + wu <= mw[b] && sc[b] + internal == 0 && lp_coin(a, mb, b))
            i1 = max(i1, b);
        i1 = wave_max_i32(i1);
        if (i1 >= 0) to = i1;
        gain = 0;
      }
      flag = lp_perform(a, mb, from, to, gain) ? 1 : 0;
    }
    if (lane == 0) {
      cand_flag[i] = uint32_t(flag);
      if (flag) { cand_to[i] = to; cand_gain[i] = gain; }
    }
    __syncthreads();
  }
  if (lane == 0 && visits) {
    stripe_add(a.stripes, 1, visits);
    stripe_add(a.stripes, 5, swept);
  }
}

// ------------------------------------------------------------- compact --
// Block weights and arrivals of a single hypergraph with up to kLpApplyLdsK
// blocks are summed per workgroup in LDS and added once per (workgroup,
// block): every move changes two of only k words, and device atomics on one
// address run one after another.
constexpr int kLpApplyLdsK = 256;

__global__ void lp_scatter_candidates_kernel(const uint32_t* __restrict__ active, uint32_t na,
                                             const uint32_t* __restrict__ cand_flag, const uint32_t* __restrict__ pos,
                                             const int32_t* __restrict__ cto_a, const int32_t* __restrict__ cgain_a,
                                             const int32_t* __restrict__ part, const int32_t* __restrict__ vw,
                                             uint32_t* __restrict__ cu, int32_t* __restrict__ cfrom,
                                             int32_t* __restrict__ cto, int32_t* __restrict__ cgain,
                                             int64_t* __restrict__ in_w, const uint32_t* __restrict__ vgraph,
                                             const uint32_t* __restrict__ gkbase, const uint32_t* __restrict__ voff,
                                             int64_t* __restrict__ inc_sum, int lds_k) {
  // whole waves step together (the incidence count combines across lanes,
  // then across the block: one atomic per block).  With lds_k > 0 (a single
  // hypergraph of lds_k <= kLpApplyLdsK blocks) the arrivals per block are
  // summed in LDS and added once per (workgroup, block).
  __shared__ unsigned long long binc;
  __shared__ unsigned long long s_in[kLpApplyLdsK];
  if (threadIdx.x == 0) binc = 0;
  for (int t = threadIdx.x; t < lds_k; t += blockDim.x) s_in[t] = 0ull;
  __syncthreads();
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < na; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + (threadIdx.x & 63u);
    int64_t d = 0;
    if (i < na && cand_flag[i]) {
      const uint32_t j = pos[i];
      const uint32_t u = active[i];
      cu[j] = u;
      cfrom[j] = part[u];
      cto[j] = cto_a[i];
      cgain[j] = cgain_a[i];
      if (lds_k) {
        atomicAdd(&s_in[cto_a[i]], static_cast<unsigned long long>(vw[u]));
      } else {
        const uint32_t kb = vgraph ? gkbase[vgraph[u]] : 0u;  // arrivals per (member, block)
        atomicAdd(reinterpret_cast<unsigned long long*>(&in_w[kb + cto_a[i]]),
                  static_cast<unsigned long long>(vw[u]));
      }
      if (inc_sum) d = int64_t(voff[u + 1] - voff[u]);
    }
    if (inc_sum) {
      d = wave_sum(d);
      if ((threadIdx.x & 63) == 0 && d) atomicAdd(&binc, static_cast<unsigned long long>(d));
    }
  }
  if (inc_sum || lds_k) __syncthreads();
  if (inc_sum && threadIdx.x == 0 && binc) counter_add(inc_sum, int64_t(binc));
  for (int t = threadIdx.x; t < lds_k; t += blockDim.x)
    if (s_in[t]) counter_add(in_w + t, int64_t(s_in[t]));
}

// admission sort key: target block, gain clamped to 24 bits, 30 hash bits.
// The block is the member's block index in the batch (its own block for one
// hypergraph).  With more than 1024 blocks (`split`), keys hold only gain and
// hash, and the block goes to `blocks` for a second, stable sort.
__global__ void lp_admission_keys_kernel(LpArgs a, const uint32_t* __restrict__ cu, const int32_t* __restrict__ cto,
                                         const int32_t* __restrict__ cgain, uint32_t nc, uint64_t* __restrict__ keys,
                                         uint32_t* __restrict__ vals, int split) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nc; j += gridDim.x * blockDim.x) {
    const LpMember mb = member_of(a, cu[j]);
    const int64_t g = min<int64_t>(max<int64_t>(int64_t(cgain[j]) + (1 << 23), 0), (1 << 24) - 1);
    const uint64_t prio = hash4(mb.seed, kLpPrioTag | a.rs, mb.lu, 0) >> 34;
    const uint64_t low = (uint64_t(g) << 30) | prio;
    keys[j] = split ? low : ((uint64_t(mb.kb + cto[j]) << 54) | low);
    vals[j] = j;
  }
}

// second pass of a split admission sort: block of every candidate in the
// (gain, hash) order
__global__ void lp_sorted_blocks_kernel(LpArgs a, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ cu,
                                        const int32_t* __restrict__ cto, uint32_t nc, uint64_t* __restrict__ blocks) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const uint32_t j = vals[i];
    blocks[i] = member_of(a, cu[j]).kb + uint32_t(cto[j]);
  }
}

__global__ void lp_sorted_weights_kernel(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ cu,
                                         const int32_t* __restrict__ vw, uint32_t nc, int64_t* __restrict__ w) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x)
    w[i] = vw[cu[vals[i]]];
}

__device__ inline uint32_t sorted_block(const uint64_t* keys, const uint64_t* blocks, uint32_t i) {
  return blocks ? uint32_t(blocks[i]) : uint32_t(keys[i] >> 54);
}

__global__ void lp_segment_base_kernel(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ blocks,
                                       const int64_t* __restrict__ w, const int64_t* __restrict__ scan, uint32_t nc,
                                       int64_t* __restrict__ base) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const uint32_t to = sorted_block(keys, blocks, i);
    if (i == 0 || sorted_block(keys, blocks, i - 1) != to) base[to] = scan[i] - w[i];
  }
}

__global__ void lp_admit_kernel(LpArgs a, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ blocks,
                                const uint32_t* __restrict__ vals, const int64_t* __restrict__ scan,
                                const int64_t* __restrict__ base, const uint8_t* __restrict__ overflow, uint32_t nc,
                                uint8_t* __restrict__ admit) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const uint32_t to = sorted_block(keys, blocks, i);
    const bool ok = !overflow[to] || a.part_w[to] + (scan[i] - base[to]) <= a.max_w[to];
    admit[vals[i]] = ok ? 1 : 0;
  }
}

// --------------------------------------------------------------- apply --
// mode 0: apply admitted moves; mode 1: undo admitted && !keep;
// mode 2: undo admitted && keep (sub-round rollback)
//
// Vertices of more than kLpWaveMove incident nets are moved by their whole
// wave (lanes over the nets) after the lanes' own light moves; a single
// thread would otherwise walk thousands of nets alone while the rest of the
// launch waits for it.
constexpr uint32_t kLpWaveMove = 32;

// changeNodePart's pin-count updates for nets j = jb + t0, jb + t0 + step, ...
// of u moving from -> to: 64-bit atomic adds on the packed counters (eight
// nets in flight at a time), connectivity toggled on 0<->1 transitions;
// returns the km1 delta of these nets.
__device__ inline int64_t lp_move_nets(const LpArgs& a, uint32_t u, int from, int to, uint32_t t0, uint32_t step) {
  const int fwi = from / a.L.epw, fsh = (from - fwi * a.L.epw) * a.L.b;
  const int twi = to / a.L.epw, tsh = (to - twi * a.L.epw) * a.L.b;
  const unsigned long long dec = static_cast<unsigned long long>(-(int64_t(1) << fsh));
  const unsigned long long incv = 1ull << tsh;
  constexpr int NB = 8;
  const uint32_t jb = a.voff[u] + t0, je = a.voff[u + 1];
  int64_t du = 0;
  for (uint32_t j0 = jb; j0 < je; j0 += NB * step) {
    uint32_t e[NB];
    uint64_t of[NB], ot[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) e[t] = j0 + t * step < je ? a.inc[j0 + t * step] : 0u;
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      if (j0 + t * step < je) {
        unsigned long long* rec = reinterpret_cast<unsigned long long*>(a.pc + uint64_t(e[t]) * a.L.stride);
        of[t] = atomicAdd(rec + fwi, dec);
        ot[t] = atomicAdd(rec + twi, incv);
      }
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      if (j0 + t * step >= je) break;
      const int32_t w = int32_t(a.pc[uint64_t(e[t]) * a.L.stride + a.L.woff]);
      if (((of[t] >> fsh) & a.L.mask) == 1) {
        atomicXor(reinterpret_cast<unsigned long long*>(&a.pc[uint64_t(e[t]) * a.L.stride + a.L.coff + from / 64]),
                  1ull << (from % 64));
        du -= w;
      }
      if (((ot[t] >> tsh) & a.L.mask) == 0) {
        atomicXor(reinterpret_cast<unsigned long long*>(&a.pc[uint64_t(e[t]) * a.L.stride + a.L.coff + to / 64]),
                  1ull << (to % 64));
        du += w;
      }
    }
  }
  return du;
}

// real km1 delta of the move u: from -> to on the post-move state over nets
// j = jb + t0, jb + t0 + step, ...
__device__ inline int64_t lp_real_delta(const LpArgs& a, uint32_t u, int from, int to, uint32_t t0, uint32_t step) {
  const int fwi = from / a.L.epw, fsh = (from - fwi * a.L.epw) * a.L.b;
  const int twi = to / a.L.epw, tsh = (to - twi * a.L.epw) * a.L.b;
  int64_t real = 0;
  constexpr int NB = 8;  // batches of 8 incident nets, loads issued together
  const uint32_t jb = a.voff[u] + t0, je = a.voff[u + 1];
  for (uint32_t j0 = jb; j0 < je; j0 += NB * step) {
    uint32_t e[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) e[t] = a.inc[j0 + t * step < je ? j0 + t * step : jb];
    int32_t w[NB];
    uint32_t pt[NB], pf[NB];
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      const uint64_t* r = a.pc + uint64_t(e[t]) * a.L.stride;  // one record per net
      w[t] = int32_t(r[a.L.woff]);
      pt[t] = uint32_t((r[twi] >> tsh) & a.L.mask);
      pf[t] = uint32_t((r[fwi] >> fsh) & a.L.mask);
    }
#pragma unroll
    for (int t = 0; t < NB; ++t) {
      if (j0 + t * step >= je) break;
      if (pt[t] == 1) real += w[t];
      if (pf[t] == 0) real -= w[t];
    }
  }
  return real;
}


__global__ void __launch_bounds__(256)
lp_apply_kernel(LpArgs a, const uint32_t* __restrict__ cu, const int32_t* __restrict__ cfrom,
                const int32_t* __restrict__ cto, const uint8_t* __restrict__ admit,
                const uint8_t* __restrict__ keep, uint32_t nc, int mode) {
  // roll back only a sub-round that worsened km1 (mode 2 adds nothing to the
  // delta: the sub-round's delta was already set aside by the check)
  if (mode == 2 && !a.vgraph && !a.counters[6]) return;
  __shared__ unsigned long long s_pw[kLpApplyLdsK];
  const bool lds_pw = !a.vgraph && a.k <= kLpApplyLdsK;
  if (lds_pw) {
    for (int t = threadIdx.x; t < a.k; t += blockDim.x) s_pw[t] = 0ull;
    __syncthreads();
  }
  int64_t d = 0;
  const int lane = threadIdx.x & 63;
  // whole waves step together: block weights and member deltas combine
  // across lanes
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < nc; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + uint32_t(lane);
    bool on = j < nc && !(admit && !admit[j]) && !(mode == 1 && keep[j]) && !(mode == 2 && !keep[j]);
    const uint32_t u = on ? cu[j] : 0u;
    const LpMember mb = member_of(a, u);
    if (mode == 2 && a.vgraph && on && !member_counters(a, mb.g)[6]) on = false;
    const int from = on ? (mode == 0 ? cfrom[j] : cto[j]) : 0;
    const int to = on ? (mode == 0 ? cto[j] : cfrom[j]) : 0;
    const int64_t wu = on ? int64_t(a.vw[u]) : 0;
    if (lds_pw) {
      if (on && wu) {
        atomicAdd(&s_pw[to], static_cast<unsigned long long>(wu));
        atomicAdd(&s_pw[from], static_cast<unsigned long long>(-wu));
      }
    } else {
      wave_add_i64(a.part_w, on, uint64_t(mb.kb) + uint64_t(to), wu, 3);
      wave_add_i64(a.part_w, on, uint64_t(mb.kb) + uint64_t(from), -wu, 3);
    }
    int64_t du = 0;
    const bool wide = on && a.voff[u + 1] - a.voff[u] > kLpWaveMove;
    if (on) {
      a.part[u] = to;
      if (!wide) du = lp_move_nets(a, u, from, to, 0, 1);
    }
    // the wide moves of the wave, one after another, all lanes on its nets
    uint64_t wides = __ballot(wide);
    while (wides) {
      const int owner = __ffsll(static_cast<long long>(wides)) - 1;
      wides &= wides - 1;
      const uint32_t wu_ = __shfl(u, owner, 64);
      const int wf = __shfl(from, owner, 64), wt = __shfl(to, owner, 64);
      const int64_t part = wave_sum(lp_move_nets(a, wu_, wf, wt, uint32_t(lane), 64u));
      if (lane == owner) du = part;
    }
    if (mode != 2) {
      if (a.vgraph) member_add_wave(a.counters, on, mb.g, 0, du);
      else d += du;
    }
  }
  if (lds_pw) {
    __syncthreads();
    for (int t = threadIdx.x; t < a.k; t += blockDim.x)
      if (s_pw[t]) counter_add(a.part_w + t, static_cast<int64_t>(s_pw[t]));
  }
  if (a.vgraph) return;
  d = wave_sum(d);
  if ((threadIdx.x & 63) == 0) stripe_add(a.stripes, 0, d);
}

// The read-back words of a sub-round gathered into one buffer, so one copy
// brings them to the host: the rollback flag of the previous sub-round, the
// candidate count, arrivals and weights per block and, with profiling, the
// candidates' incidences.
__global__ void lp_readback_kernel(const int64_t* __restrict__ counters, const uint32_t* __restrict__ pos,
                                   uint32_t ns, const int64_t* __restrict__ in_w,
                                   const int64_t* __restrict__ part_w, int K, int prof, int64_t* __restrict__ rb) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x), stride = int(gridDim.x * blockDim.x);
  if (t == 0) {
    rb[0] = counters[6];
    rb[1] = int64_t(pos[ns]);
    rb[2 + 2 * K] = prof ? counters[11] : 0;
  }
  for (int b = t; b < K; b += stride) {
    rb[2 + b] = in_w[b];
    rb[2 + K + b] = part_w[b];
  }
}

// real km1 delta of each applied move on the post-move state
__global__ void lp_verify_kernel(LpArgs a, const uint32_t* __restrict__ cu, const int32_t* __restrict__ cfrom,
                                 const int32_t* __restrict__ cto, const int32_t* __restrict__ cgain,
                                 const uint8_t* __restrict__ admit, uint32_t nc, uint8_t* __restrict__ keep) {
  int64_t reverts = 0;
  const int lane = threadIdx.x & 63;
  // whole waves step together: the per-member counts combine across lanes
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < nc; base += gridDim.x * blockDim.x) {
    const uint32_t j = base + uint32_t(lane);
    const bool on = j < nc && !(admit && !admit[j]);
    const uint32_t u = on ? cu[j] : 0u;
    const int from = on ? cfrom[j] : 0, to = on ? cto[j] : 0;
    const bool wide = on && a.voff[u + 1] - a.voff[u] > kLpWaveMove;
    int64_t real = 0;
    if (on && !wide) real = lp_real_delta(a, u, from, to, 0, 1);
    uint64_t wides = __ballot(wide);
    while (wides) {
      const int owner = __ffsll(static_cast<long long>(wides)) - 1;
      wides &= wides - 1;
      const uint32_t wu_ = __shfl(u, owner, 64);
      const int wf = __shfl(from, owner, 64), wt = __shfl(to, owner, 64);
      const int64_t part = wave_sum(lp_real_delta(a, wu_, wf, wt, uint32_t(lane), 64u));
      if (lane == owner) real = part;
    }
    bool k = false;
    uint32_t g = 0;
    if (on) {
      k = real == int64_t(cgain[j]) || real <= 0;
      keep[j] = k ? 1 : 0;
      if (a.vgraph) g = a.vgraph[u];
      else reverts += k ? 0 : 1;
    }
    if (a.vgraph) {
      // reverts, and the kept moves, which count once the member's
      // sub-round is not rolled back (lp_check_kernel)
      member_add_wave(a.counters, on && !k, g, 2, 1);
      member_add_wave(a.counters, on && k, g, 10, 1);
    }
  }
  if (a.vgraph) return;
  reverts = wave_sum(reverts);
  if ((threadIdx.x & 63) == 0) stripe_add(a.stripes, 2, reverts);
}

// One wave per kept move; lanes take the incident nets.  A net of at most
// `threshold` pins is expanded by the first move that stamps it in this round
// (the visited-net flags of label_propagation_refiner.h:113-125).
__global__ void __launch_bounds__(64)
lp_activate_kernel(LpArgs a, const uint32_t* __restrict__ cu, const uint8_t* __restrict__ admit,
                   const uint8_t* __restrict__ keep, uint32_t nc, uint32_t threshold, uint32_t stamp,
                   uint32_t* __restrict__ net_stamp, uint32_t* __restrict__ flag) {
  if (!a.vgraph && a.counters[6]) return;  // the sub-round was rolled back
  int64_t moves = 0;
  const int lane = threadIdx.x;
  for (uint32_t j = blockIdx.x; j < nc; j += gridDim.x) {
    if ((admit && !admit[j]) || !keep[j]) continue;
    const uint32_t u = cu[j];
    if (a.vgraph) {
      const uint32_t g = a.vgraph[u];
      if (member_counters(a, g)[6]) continue;  // the member's sub-round was rolled back
      if (lane == 0) flag[u] = 1;  // counted by the verify and check kernels
    } else if (lane == 0) {
      flag[u] = 1;
      ++moves;
    }
    const uint32_t jb = a.voff[u], je = a.voff[u + 1];
    for (uint32_t jj = jb + lane; jj < je; jj += 64) {
      const uint32_t e = a.inc[jj];
      const uint32_t pb = a.noff[e], pe = a.noff[e + 1];
      if (pe - pb <= threshold && atomicExch(&net_stamp[e], stamp) != stamp)
        for (uint32_t q0 = pb; q0 < pe; q0 += 8) {  // eight pin loads in flight at once
          uint32_t xs[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) xs[t] = q0 + t < pe ? a.pins[q0 + t] : 0xffffffffu;
#pragma unroll
          for (int t = 0; t < 8; ++t)
            if (xs[t] != 0xffffffffu) flag[xs[t]] = 1;
        }
    }
  }
  if (lane == 0) stripe_add(a.stripes, 4, moves);
}

__global__ void lp_compact_active_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                         uint32_t n, uint32_t* __restrict__ out) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    if (flag[u]) out[pos[u]] = u;
}

// isBorderNode as a 0/1 word per vertex (partitioned_hypergraph.h:605-620)
__global__ void lp_border_flag_kernel(LpArgs a, uint32_t n, uint32_t* __restrict__ flag) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint32_t b = a.voff[u], e = a.voff[u + 1];
    uint32_t border = 0;
    if (e - b <= kHighDegree) {
      for (uint32_t j = b; j < e && !border; ++j) {
        int lambda = 0;
        for (int ci = 0; ci < a.vcs; ++ci) lambda += __popcll(a.pc[uint64_t(a.inc[j]) * a.L.stride + a.L.coff + ci]);
        border = lambda > 1;
      }
    }
    flag[u] = border;
  }
}

// After a sub-round's moves and reverts: if its total real delta is positive
// the sub-round is rolled back and the refinement stops (flag counters[6]),
// else the delta joins the total (counters[7]).
__global__ void lp_check_kernel(int64_t* __restrict__ counters, uint32_t members, int64_t* __restrict__ stripes) {
  if (stripes) {  // single hypergraph: one thread folds the stripes first
    if (blockIdx.x == 0 && threadIdx.x == 0) stripe_fold(counters, stripes);
    else return;
  }
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < members; g += gridDim.x * blockDim.x) {
    int64_t* c = counters + kGcStride * g;
    const int64_t d = c[0];
    c[0] = 0;
    if (d > 0) {
      c[6] = 1;
    } else {
      c[7] += d;
      c[4] += c[10];  // the sub-round's kept moves (batched form)
    }
    c[10] = 0;
  }
}

// batched rounds: a member takes part in a round while it has active
// vertices and has not stopped
__global__ void lp_round_mark_kernel(LpArgs a, const uint32_t* __restrict__ active, uint32_t na) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < na; i += gridDim.x * blockDim.x) {
    int64_t* c = member_counters(a, a.vgraph[active[i]]);
    if (!c[6] && !c[9]) c[9] = 1;
  }
}
__global__ void lp_round_count_kernel(int64_t* __restrict__ counters, uint32_t members) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < members; g += gridDim.x * blockDim.x) {
    int64_t* c = counters + kGcStride * g;
    c[8] += c[9];
    c[9] = 0;
  }
}

__global__ void lp_fold_kernel(int64_t* __restrict__ counters, int64_t* __restrict__ stripes) {
  if (blockIdx.x == 0 && threadIdx.x == 0) stripe_fold(counters, stripes);
}

__global__ void iota_kernel(uint32_t* out, uint32_t n) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) out[u] = u;
}

// --------------------------------------------------------------- driver --
// bucket key per active entry; per-block LDS histograms, one global atomic per
// (block, bucket)
// With `stats` set (profiling), count[B + b], count[2B + b] and count[3B + b]
// also receive the incidences of bucket b's vertices that the thread-per-vertex
// gain kernel takes, and the vertices and incidences of those the
// wave-per-vertex kernel takes: the gain launches' algorithmic bytes.
__global__ void __launch_bounds__(256)
lp_bucket_keys_kernel(LpArgs a, const uint32_t* __restrict__ active, uint32_t na, uint64_t* __restrict__ keys,
                      int64_t* __restrict__ count, int stats) {
  __shared__ uint32_t h[256];
  __shared__ unsigned long long hs[3 * 256];
  for (int i = threadIdx.x; i < a.buckets; i += blockDim.x) h[i] = 0;
  if (stats)
    for (int i = threadIdx.x; i < 3 * a.buckets; i += blockDim.x) hs[i] = 0;
  __syncthreads();
  // whole waves step together: the lanes of one bucket count with one LDS
  // atomic per wave
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < na; base += gridDim.x * blockDim.x) {
    const uint32_t i = base + (threadIdx.x & 63u);
    int b = -1;
    if (i < na) {
      if (a.vgraph) {  // the member's own (seed, round) salt and local id
        const LpMember mb = member_of(a, active[i]);
        b = int(bucket_of(bucket_salt(mb.seed, kLpBucketTag | uint64_t(a.round)), mb.lu, uint32_t(a.buckets)));
      } else {
        b = int(bucket_of(a.bucket_salt, active[i], uint32_t(a.buckets)));
      }
      if (keys) keys[i] = uint64_t(b);
    }
    const SlotGroup sg = slot_group(b);
    if (sg.leader) atomicAdd(&h[b], sg.count);
    if (stats) {
      // per bucket of the wave: incidences of thread-kernel vertices,
      // wave-kernel vertices and their incidences
      uint64_t li = 0, wv = 0, wi = 0;
      if (i < na) {
        const uint32_t u = active[i];
        const uint32_t d = a.voff[u + 1] - a.voff[u];
        if (d <= kHighDegree) {
          if (a.k > 128 || d > (a.k > 32 ? kGroupDegreeLimit : kThreadDegreeLimit)) { wv = 1; wi = d; } else { li = d; }
        }
      }
      uint64_t pending = __ballot(b >= 0);
      while (pending) {
        const int first = __ffsll(static_cast<long long>(pending)) - 1;
        const int bb = __shfl(b, first, 64);
        const bool mine = b == bb && b >= 0;
        uint64_t x = mine ? li : 0, y = mine ? wv : 0, z = mine ? wi : 0;
        for (int o = 32; o > 0; o >>= 1) {
          x += __shfl_xor(x, o, 64);
          y += __shfl_xor(y, o, 64);
          z += __shfl_xor(z, o, 64);
        }
        if ((threadIdx.x & 63) == uint32_t(first)) {
          if (x) atomicAdd(&hs[bb], static_cast<unsigned long long>(x));
          if (y) atomicAdd(&hs[a.buckets + bb], static_cast<unsigned long long>(y));
          if (z) atomicAdd(&hs[2 * a.buckets + bb], static_cast<unsigned long long>(z));
        }
        pending &= ~__ballot(mine);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < a.buckets; i += blockDim.x)
    if (h[i]) atomicAdd(reinterpret_cast<unsigned long long*>(&count[i]), static_cast<unsigned long long>(h[i]));
  if (stats)
    for (int i = threadIdx.x; i < 3 * a.buckets; i += blockDim.x)
      if (hs[i]) atomicAdd(reinterpret_cast<unsigned long long*>(&count[a.buckets + i]), hs[i]);
}

// --------------------------------------------------------------- driver --
template <int K>
static void launch_thread_gain(const LpArgs& a, LpWork& w, const uint32_t* act, uint32_t na, hipStream_t s) {
  lp_gain_thread_kernel<K><<<grid_for(na, kGainThreads, 65536), kGainThreads, 0, s>>>(
      a, act, na, w.cand_flag.p, w.cand_to_a.p, w.cand_gain_a.p, w.heavy.p);
}

// Algorithmic bytes of the two gain launches of a sub-round (Km1Policy
// sweep): per vertex its two offsets, block id and weight (16 B); per
// incidence the net id, net weight, the pin-count word of the source block
// and the connectivity words.
struct GainBytes {
  double thread = 0, wave = 0;
};

static void launch_gain(Context& c, const LpArgs& a, LpWork& w, const uint32_t* act, uint32_t na, hipStream_t s,
                        const GainBytes& gb) {
  const int num_cus = c.num_cus;
  if (a.k <= 32) {
    {
      const int K = a.k <= 2 ? 2 : a.k <= 4 ? 4 : a.k <= 8 ? 8 : a.k <= 16 ? 16 : 32;
      TimedRegion tr(c.prof, s, kFamLpGain, gb.thread, "lp_gain_thread_kernel<" + std::to_string(K) + ">");
      if (K == 2) launch_thread_gain<2>(a, w, act, na, s);
      else if (K == 4) launch_thread_gain<4>(a, w, act, na, s);
      else if (K == 8) launch_thread_gain<8>(a, w, act, na, s);
      else if (K == 16) launch_thread_gain<16>(a, w, act, na, s);
      else launch_thread_gain<32>(a, w, act, na, s);
      MKP_LAUNCH_CHECK();
    }
    // heavy vertices: persistent grid over the device-side list
    TimedRegion tr(c.prof, s, kFamLpGain, gb.wave, "lp_gain_wave_kernel");
    lp_gain_wave_kernel<<<unsigned(num_cus) * 8, 64, a.k * sizeof(int32_t), s>>>(
        a, act, w.heavy.p, a.counters + 3, 0, w.cand_flag.p, w.cand_to_a.p, w.cand_gain_a.p);
    MKP_LAUNCH_CHECK();
  } else if (a.k <= 128) {
    {
      // 8 lanes per vertex: mean degrees are small next to the k = 64..128
      // scores each group scans
      TimedRegion tr(c.prof, s, kFamLpGain, gb.thread, "lp_gain_group_kernel<8, 128>");
      lp_gain_group_kernel<8, 128><<<grid_for(na, 32, 65536), 256, 0, s>>>(a, act, na, w.cand_flag.p, w.cand_to_a.p,
                                                                        w.cand_gain_a.p, w.heavy.p);
      MKP_LAUNCH_CHECK();
    }
    TimedRegion tr(c.prof, s, kFamLpGain, gb.wave, "lp_gain_wave_kernel");
    lp_gain_wave_kernel<<<unsigned(num_cus) * 8, 64, a.k * sizeof(int32_t), s>>>(
        a, act, w.heavy.p, a.counters + 3, 0, w.cand_flag.p, w.cand_to_a.p, w.cand_gain_a.p);
    MKP_LAUNCH_CHECK();
  } else {
    TimedRegion tr(c.prof, s, kFamLpGain, gb.wave, "lp_gain_wave_kernel");
    lp_gain_wave_kernel<<<grid_for(na, 1, unsigned(num_cus) * 64), 64, a.k * sizeof(int32_t), s>>>(
        a, act, nullptr, nullptr, na, w.cand_flag.p, w.cand_to_a.p, w.cand_gain_a.p);
    MKP_LAUNCH_CHECK();
  }
}

void lp_work_free(LpWork* w) { delete w; }

static LpWork& work_of(mkp_partition* p) {
  if (!p->lp) p->lp = new LpWork();
  return *p->lp;
}

void lp_refine(mkp_partition* p, const mkp_lp_params& prm, mkp_lp_stats* stats, const uint32_t* nodes,
               uint32_t num_nodes, const LpBatch* batch) {
  mkp_hypergraph* hg = p->hg;
  Context& c = hg->ctx->c;
  hipStream_t s = c.stream;
  const uint32_t n = hg->n;
  const uint32_t G = batch ? batch->members : 1;
  const int K = batch ? int(batch->ktot) : p->k;  // entries of the block-weight arrays
  LpWork& w = work_of(p);
  w.ensure(n, K, G);
  w.ensure_nets(hg->m, s);
  MKP_HIP(hipMemcpyAsync(w.max_w.p, prm.max_part_weights, size_t(K) * 8, hipMemcpyHostToDevice, s));
  MKP_HIP(hipMemcpyAsync(w.perfect_w.p, prm.perfect_part_weights, size_t(K) * 8, hipMemcpyHostToDevice, s));
  fill_async<int64_t>(w.counters.p, 0, size_t(kGcStride) * G, s);
  fill_async<int64_t>(w.stripes.p, 0, size_t(kLpStripes) * kLpStripeWords, s);
  MKP_HIP(hipEventRecord(c.ev_a, s));

  LpArgs a;
  a.voff = hg->voff.p; a.inc = hg->inc.p; a.noff = hg->noff.p; a.pins = hg->pins.p;
  a.vw = hg->vw.p; a.nw = hg->nw.p; a.part = p->part.p; a.part_w = p->part_w.p;
  a.pc = p->pc.p; a.max_w = w.max_w.p; a.perfect_w = w.perfect_w.p;
  a.L = p->L; a.vcs = p->vcs; a.k = batch ? batch->kmax : p->k; a.rebalancing = prm.rebalancing;
  a.seed = prm.seed;
  a.counters = w.counters.p;
  a.stripes = w.stripes.p;
  a.vgraph = batch ? batch->vgraph : nullptr;
  a.gvbase = batch ? batch->gvbase : nullptr;
  a.gkbase = batch ? batch->gkbase : nullptr;
  a.gk = batch ? batch->gk : nullptr;
  a.gseed = batch ? batch->gseed : nullptr;
  const int B = std::max(1, prm.sub_rounds);
  a.buckets = B;

  // initializeActiveNodes (label_propagation_refiner.cpp:194-233): the given
  // refinement nodes (distinct, ascending), else all vertices with
  // rebalancing (the non-border ones drop out in the gain kernel), else the
  // border vertices
  uint32_t na = n;
  if (nodes) {
    na = num_nodes;
    if (na) MKP_HIP(hipMemcpyAsync(w.active.p, nodes, size_t(na) * 4, hipMemcpyHostToDevice, s));
  } else if (prm.rebalancing) {
    iota_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(w.active.p, n);
    MKP_LAUNCH_CHECK();
  } else {
    lp_border_flag_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(a, n, w.flag.p);
    exclusive_scan_u32(c, w.flag.p, w.pos.p, n);
    lp_compact_active_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(w.flag.p, w.pos.p, n, w.active.p);
    MKP_LAUNCH_CHECK();
    uint32_t na32 = 0;
    MKP_HIP(hipMemcpyAsync(&na32, w.pos.p + n, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    na = na32;
  }
  int kbits = 1;  // bits of a block index in the admission keys
  while ((1 << kbits) < K) ++kbits;
  const bool split = kbits > 10;
  int64_t* host = c.pinned;
  const bool prof = c.prof.enabled;
  int rounds = 0;
  bool stop = false;
  for (int r = 0; r < prm.max_rounds && na > 0 && !stop; ++r) {
    a.round = uint32_t(r);
    if (++w.stamp == 0) {  // stamp wrapped: clear the net stamps once
      fill_async<uint32_t>(w.net_stamp.p, 0u, std::max<uint32_t>(w.m, 1), s);
      w.stamp = 1;
    }
    const uint32_t round_stamp = w.stamp;
    fill_async<uint32_t>(w.flag.p, 0u, n, s);
    if (batch) {  // rounds per member
      lp_round_mark_kernel<<<grid_for(na, 256, 65536), 256, 0, s>>>(a, w.active.p, na);
      lp_round_count_kernel<<<grid_for(G, 256, 4096), 256, 0, s>>>(w.counters.p, G);
      MKP_LAUNCH_CHECK();
    }
    // stable partition of the active list by bucket: each sub-round sweeps a
    // contiguous ascending slice
    const uint32_t* base = w.active.p;
    std::vector<uint32_t> boff(size_t(B) + 1, 0);
    boff[B] = na;
    a.buckets = B;
    a.bucket_salt = bucket_salt(a.seed, kLpBucketTag | uint64_t(r));
    // per bucket: vertices; with profiling also the incidences of the thread
    // kernel's vertices, the wave kernel's vertices and their incidences
    std::vector<int64_t> cnt(4 * size_t(B), 0);
    if (B > 1 || prof) {
      fill_async<int64_t>(w.bcount.p, 0, 4 * size_t(B), s);
      lp_bucket_keys_kernel<<<grid_for(na, 256, 65536), 256, 0, s>>>(a, w.active.p, na, B > 1 ? w.bkeys.p : nullptr,
                                                                      w.bcount.p, prof ? 1 : 0);
      MKP_LAUNCH_CHECK();
      if (B > 1) {
        int bb = 1;
        while ((1 << bb) < B) ++bb;
        sort_pairs_u64_u32(c, w.bkeys.p, w.bkeys_sorted.p, w.active.p, w.active_next.p, na, 0, bb);
      }
      int64_t* stg_b = c.staging<int64_t>(4 * size_t(B));
      MKP_HIP(hipMemcpyAsync(stg_b, w.bcount.p, 4 * size_t(B) * 8, hipMemcpyDeviceToHost, s));
      c.sync();
      std::copy(stg_b, stg_b + 4 * B, cnt.begin());
      if (B > 1) {
        for (int b = 0; b < B; ++b) boff[b + 1] = boff[b] + uint32_t(cnt[b]);
        base = w.active_next.p;
      }
    }
    a.buckets = 1;  // slices are pre-partitioned
    for (int sr = 0; sr < B && !stop; ++sr) {
      a.bucket = sr;
      a.rs = (uint64_t(r) << 8) | uint64_t(sr);
      const uint32_t* act = base + boff[sr];
      const uint32_t ns = boff[sr + 1] - boff[sr];
      if (ns == 0) continue;
      fill_async<int64_t>(w.counters.p + 3, 0, 1, s);
      if (prof) fill_async<int64_t>(w.counters.p + 11, 0, 1, s);
      {
        GainBytes gb;
        if (prof) {
          const double per_inc = 16.0 + 8.0 * p->vcs;
          const int sb = B > 1 ? sr : 0;
          gb.thread = 16.0 * double(int64_t(ns) - cnt[2 * B + sb]) + per_inc * double(cnt[B + sb]);
          gb.wave = 16.0 * double(cnt[2 * B + sb]) + per_inc * double(cnt[3 * B + sb]);
        }
        launch_gain(c, a, w, act, ns, s, gb);
      }
      exclusive_scan_u32(c, w.cand_flag.p, w.pos.p, ns);
      fill_async<int64_t>(w.in_w.p, 0, K, s);
      lp_scatter_candidates_kernel<<<grid_for(ns, 256, prof ? 512 : 65536), 256, 0, s>>>(
          act, ns, w.cand_flag.p, w.pos.p, w.cand_to_a.p, w.cand_gain_a.p, p->part.p, hg->vw.p, w.cand_u.p,
          w.cand_from.p, w.cand_to.p, w.cand_gain.p, w.in_w.p, a.vgraph, a.gkbase, hg->voff.p,
          prof ? w.counters.p + 11 : nullptr, (!a.vgraph && a.k <= kLpApplyLdsK) ? a.k : 0);
      MKP_LAUNCH_CHECK();
      // read back (pinned staging): stop flag of the previous sub-round,
      // number of candidates, arrivals per block, block weights (and, with
      // profiling, the candidates' incidences)
      const size_t over_words = (size_t(K) + 7) / 8;
      int64_t* stg = c.staging<int64_t>(2 * size_t(K) + 3 + over_words);
      int64_t* stg_inc = stg + 2 * K + 2;
      uint8_t* stg_over = reinterpret_cast<uint8_t*>(stg + 2 * K + 3);
      lp_readback_kernel<<<unsigned(std::min<int>((K + 255) / 256, 256)), 256, 0, s>>>(
          w.counters.p, w.pos.p, ns, w.in_w.p, p->part_w.p, K, prof ? 1 : 0, w.rb.p);
      MKP_LAUNCH_CHECK();
      MKP_HIP(hipMemcpyAsync(stg, w.rb.p, (2 * size_t(K) + 3) * 8, hipMemcpyDeviceToHost, s));
      c.sync();
      if (!batch && stg[0]) {  // the previous sub-round was rolled back on the device
        stop = true;
        break;
      }
      const uint32_t nc = uint32_t(stg[1]);
      if (nc == 0) continue;
      const int64_t* inw = stg + 2;
      const int64_t* pw = stg + 2 + K;
      bool any_over = false;
      for (int b = 0; b < K; ++b) {
        stg_over[b] = pw[b] + inw[b] > prm.max_part_weights[b];
        any_over |= stg_over[b] != 0;
      }
      const uint8_t* admit = nullptr;
      if (any_over) {
        MKP_HIP(hipMemcpyAsync(w.overflow.p, stg_over, K, hipMemcpyHostToDevice, s));
        lp_admission_keys_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(a, w.cand_u.p, w.cand_to.p, w.cand_gain.p,
                                                                          nc, w.keys.p, w.vals.p, split ? 1 : 0);
        MKP_LAUNCH_CHECK();
        const uint32_t* order = w.vals_sorted.p;
        const uint64_t* blocks = nullptr;
        if (!split) {
          sort_pairs_u64_u32(c, w.keys.p, w.keys_sorted.p, w.vals.p, w.vals_sorted.p, nc, 0, 54 + kbits);
        } else {
          // (gain, hash) order first, then a stable sort by block
          sort_pairs_u64_u32(c, w.keys.p, w.keys_sorted.p, w.vals.p, w.vals_sorted.p, nc, 0, 54);
          lp_sorted_blocks_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(a, w.vals_sorted.p, w.cand_u.p,
                                                                           w.cand_to.p, nc, w.bkeys.p);
          MKP_LAUNCH_CHECK();
          sort_pairs_u64_u32(c, w.bkeys.p, w.bkeys_sorted.p, w.vals_sorted.p, w.vals.p, nc, 0, kbits);
          order = w.vals.p;
          blocks = w.bkeys_sorted.p;
        }
        lp_sorted_weights_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(order, w.cand_u.p, hg->vw.p, nc,
                                                                          w.wsort.p);
        MKP_LAUNCH_CHECK();
        inclusive_scan_i64(c, w.wsort.p, w.wscan.p, nc);
        lp_segment_base_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(w.keys_sorted.p, blocks, w.wsort.p,
                                                                        w.wscan.p, nc, w.seg_base.p);
        MKP_LAUNCH_CHECK();
        lp_admit_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(a, w.keys_sorted.p, blocks, order, w.wscan.p,
                                                                 w.seg_base.p, w.overflow.p, nc, w.admit.p);
        MKP_LAUNCH_CHECK();
        admit = w.admit.p;
      }
      const unsigned gc = grid_for(nc, 256, 65536);
      {
        // changeNodePart per candidate: its id, blocks and admission flag
        // (16 B), per incidence the net id and a read and a write of the
        // packed pin-count word (20 B); the verification reads the net id
        // and the word again (12 B per incidence)
        const double cand_inc = prof ? double(*stg_inc) : 0.0;
        {
          TimedRegion tr(c.prof, s, kFamLpApply, 16.0 * nc + 20.0 * cand_inc, "lp_apply_kernel");
          lp_apply_kernel<<<gc, 256, 0, s>>>(a, w.cand_u.p, w.cand_from.p, w.cand_to.p, admit, nullptr, nc, 0);
          MKP_LAUNCH_CHECK();
        }
        {
          TimedRegion tr(c.prof, s, kFamLpApply, 16.0 * nc + 12.0 * cand_inc, "lp_verify_kernel");
          lp_verify_kernel<<<gc, 256, 0, s>>>(a, w.cand_u.p, w.cand_from.p, w.cand_to.p, w.cand_gain.p, admit, nc,
                                              w.keep.p);
          MKP_LAUNCH_CHECK();
        }
        TimedRegion tr(c.prof, s, kFamLpApply, 0.0, "lp_revert_kernel");
        lp_apply_kernel<<<gc, 256, 0, s>>>(a, w.cand_u.p, w.cand_from.p, w.cand_to.p, admit, w.keep.p, nc, 1);
        MKP_LAUNCH_CHECK();
      }
      // decided on the device, per member: keep the sub-round (delta into the
      // total) or roll it back and stop that member
      lp_check_kernel<<<grid_for(G, 256, 4096), 256, 0, s>>>(w.counters.p, G, batch ? nullptr : w.stripes.p);
      lp_apply_kernel<<<gc, 256, 0, s>>>(a, w.cand_u.p, w.cand_from.p, w.cand_to.p, admit, w.keep.p, nc, 2);
      MKP_LAUNCH_CHECK();
      lp_activate_kernel<<<grid_for(nc, 1, unsigned(c.num_cus) * 16), 64, 0, s>>>(
          a, w.cand_u.p, admit, w.keep.p, nc, prm.activation_threshold, round_stamp, w.net_stamp.p, w.flag.p);
      MKP_LAUNCH_CHECK();
    }
    ++rounds;
    // next active set = flagged vertices in ascending order
    exclusive_scan_u32(c, w.flag.p, w.pos.p, n);
    lp_compact_active_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(w.flag.p, w.pos.p, n, w.active.p);
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipMemcpyAsync(host, w.pos.p + n, 4, hipMemcpyDeviceToHost, s));
    MKP_HIP(hipMemcpyAsync(host + 1, w.counters.p + 6, 8, hipMemcpyDeviceToHost, s));
    c.sync();
    na = *reinterpret_cast<const uint32_t*>(host);
    if (!batch && host[1]) stop = true;  // the round's last sub-round was rolled back
  }
  if (!batch) {  // the moves of the last sub-round are still striped
    lp_fold_kernel<<<1, 64, 0, s>>>(w.counters.p, w.stripes.p);
    MKP_LAUNCH_CHECK();
  }
  MKP_HIP(hipEventRecord(c.ev_b, s));
  std::vector<int64_t> cnt(size_t(kGcStride) * G);
  MKP_HIP(hipMemcpyAsync(cnt.data(), w.counters.p, cnt.size() * 8, hipMemcpyDeviceToHost, s));
  c.sync();
  int64_t all_visits = 0, all_swept = 0;
  for (uint32_t g = 0; g < G; ++g) {
    all_visits += cnt[kGcStride * g + 1];
    all_swept += cnt[kGcStride * g + 5];
  }
  if (stats) {
    float ms = 0;
    MKP_HIP(hipEventElapsedTime(&ms, c.ev_a, c.ev_b));
    for (uint32_t g = 0; g < G; ++g) {
      const int64_t* cg = cnt.data() + kGcStride * g;
      stats[g].delta = cg[7];
      stats[g].moves = cg[4];
      stats[g].reverts = cg[2];
      stats[g].visits = cg[1];
      stats[g].rounds = batch ? cg[8] : rounds;
      stats[g].device_ms = ms;
    }
  }
}

}  // namespace mkp

using namespace mkp;

extern "C" int mkp_lp_refine(mkp_partition* p, const mkp_lp_params* prm, mkp_lp_stats* stats) {
  return guarded([&] {
    MKP_CHECK(p && prm && prm->max_part_weights && prm->perfect_part_weights, "null argument");
    MKP_CHECK(prm->max_rounds >= 0 && prm->sub_rounds >= 1 && prm->sub_rounds <= 255, "bad round parameters");
    Bind bind_(p->hg->ctx->c);
    lp_refine(p, *prm, stats, nullptr, 0);
  });
}

extern "C" int mkp_lp_refine_nodes(mkp_partition* p, const mkp_lp_params* prm, const uint32_t* nodes,
                                   uint32_t num_nodes, mkp_lp_stats* stats) {
  return guarded([&] {
    MKP_CHECK(p && prm && prm->max_part_weights && prm->perfect_part_weights, "null argument");
    MKP_CHECK(nodes || num_nodes == 0, "null refinement nodes");
    MKP_CHECK(prm->max_rounds >= 0 && prm->sub_rounds >= 1 && prm->sub_rounds <= 255, "bad round parameters");
    Bind bind_(p->hg->ctx->c);
    if (num_nodes == 0) {  // empty list: all vertices, as in the reference
      lp_refine(p, *prm, stats, nullptr, 0);
      return;
    }
    std::vector<uint32_t> v(nodes, nodes + num_nodes);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    MKP_CHECK(v.back() < p->hg->n, "refinement node out of range");
    lp_refine(p, *prm, stats, v.data(), uint32_t(v.size()));
  });
}
