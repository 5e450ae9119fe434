// Louvain modularity local moving on the star expansion, graph contraction
// and prolongation.
//
// Reference: ParallelLocalMovingModularity::localMoving and
// computeMaxGainCluster (local_moving_modularity.cpp:61-130,
// local_moving_modularity.h:87-137), Graph construction and contraction
// (graph.cpp:36-78,128-268,290-446), local_moving_contract_recurse
// (parallel_louvain.cpp:27-66).
//
// Level 0 is never materialised: the star expansion's arcs are read straight
// from the hypergraph CSR (hypernode u -> net nodes N+e of its incident nets,
// net node N+e -> its pins).  Coarser levels are explicit CSR graphs built by
// a sort of (cluster(u), cluster(v)) keys.
//
// A round r is split into sub-rounds by hash(seed, level, r, u); within a
// sub-round every node of the bucket picks the cluster of maximum modularity
// gain on the state at the sub-round start (ties: first occurrence in the arc
// order, i.e. the insertion order of the reference's sparse map), then all
// moves are applied.  A singleton moves into another singleton only towards
// the smaller cluster id.  Gains are FP64 with the reference's operation
// order (w - (recip*volU)*vol), built without contraction into FMA.
#include <cfloat>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>

#include "internal.h"

namespace mkp {

// ------------------------------------------------------------ graph views --
// Every view exposes degree(u), arc(u, i) and source(a) (the tail of arc a in
// the global arc order).  `unit` marks hypergraphs whose net weights are all
// one, so the uniform arc weight needs no gather.
// Net nodes are numbered N + rank(e), ranks ordering the nets by a locality
// key (lower median of their first nine pins, then id): net nodes close in id
// then have pins close in id, and the incident nets of a vertex get nearby
// node ids, so the cluster gathers of neighbouring nodes share cache lines.
struct StarView {  // star expansion of a hypergraph (graph.cpp:331-398)
  static constexpr const char* kName = "StarView";
  uint32_t N, M;
  uint64_t P;
  const uint32_t* voff;
  const uint32_t* inc;
  const uint32_t* noff;
  const uint32_t* pins;
  const uint32_t* inc_vtx;
  const uint32_t* pin_net;
  const int32_t* nw;
  int wt;  // 0 uniform, 1 non_uniform, 2 degree
  bool unit;
  const uint32_t* order;      // rank -> net
  const uint32_t* inc_rank;   // per incidence: rank of the net
  const uint32_t* rpoff;      // rank-ordered pin lists
  const uint32_t* rpins;
  const uint32_t* rpin_rank;  // per rank-ordered pin slot: rank of its net
  __device__ uint64_t num_nodes() const { return uint64_t(N) + M; }
  __device__ uint64_t offset(uint64_t u) const { return u < N ? voff[u] : P + rpoff[u - N]; }
  __device__ uint32_t degree(uint32_t u) const {
    return u < N ? voff[u + 1] - voff[u] : rpoff[u - N + 1] - rpoff[u - N];
  }
  __device__ uint32_t source(uint64_t a) const { return a < P ? inc_vtx[a] : N + rpin_rank[a - P]; }
  // algorithmic bytes of one arc of u: head id + cluster gather, plus the
  // weight inputs the weighting reads per arc
  __device__ uint32_t arc_bytes(uint32_t u) const {
    if (u < N) return 8u + (wt == 0 ? (unit ? 0u : 8u) : 16u);
    return 8u + (wt == 0 ? 0u : 8u);
  }
  // unit arc weights (uniform weighting, all net weights one): the move
  // kernels then need the heads only
  static constexpr bool kHasUnit = true;
  __host__ __device__ bool unit_arcs() const { return wt == 0 && unit; }
  __device__ uint32_t head_of(uint32_t u, uint32_t i) const {
    return u < N ? N + inc_rank[voff[u] + i] : rpins[rpoff[u - N] + i];
  }
  __device__ double weight(int32_t w, uint32_t size, uint32_t deg) const {
    if (wt == 0) return double(w);
    if (wt == 1) return double(w) / double(size);
    return double(w) * (double(deg) / double(size));
  }
  __device__ void arc(uint32_t u, uint32_t i, uint32_t& head, double& w) const {
    if (u < N) {
      const uint32_t j = voff[u] + i;
      head = N + inc_rank[j];
      if (wt == 0 && unit) {
        w = 1.0;
      } else {
        const uint32_t e = inc[j];
        w = wt == 0 ? double(nw[e]) : weight(nw[e], noff[e + 1] - noff[e], voff[u + 1] - voff[u]);
      }
    } else {
      const uint32_t r = u - N;
      const uint32_t v = rpins[rpoff[r] + i];
      head = v;
      if (wt == 0 && unit) {
        w = 1.0;
        return;
      }
      const uint32_t e = order[r];
      w = wt == 0 ? (unit ? 1.0 : double(nw[e])) : weight(nw[e], noff[e + 1] - noff[e], voff[v + 1] - voff[v]);
    }
  }
};

struct PlainView {  // all nets have two pins (graph.cpp:401-446)
  static constexpr const char* kName = "PlainView";
  uint32_t N;
  const uint32_t* voff;
  const uint32_t* inc;
  const uint32_t* noff;
  const uint32_t* pins;
  const uint32_t* inc_vtx;
  const int32_t* nw;
  int wt;
  bool unit;
  __device__ uint64_t num_nodes() const { return N; }
  __device__ uint64_t offset(uint64_t u) const { return voff[u]; }
  __device__ uint32_t degree(uint32_t u) const { return voff[u + 1] - voff[u]; }
  __device__ uint32_t source(uint64_t a) const { return inc_vtx[a]; }
  // net id, net offset, both pins, cluster gather (+ net weight)
  __device__ uint32_t arc_bytes(uint32_t) const { return 20u + (unit ? 0u : 4u); }
  static constexpr bool kHasUnit = true;
  __host__ __device__ bool unit_arcs() const { return wt == 0 && unit; }
  __device__ uint32_t head_of(uint32_t u, uint32_t i) const {
    const uint32_t e = inc[voff[u] + i];
    const uint32_t p0 = pins[noff[e]], p1 = pins[noff[e] + 1];
    return p0 != u ? p0 : p1;
  }
  __device__ void arc(uint32_t u, uint32_t i, uint32_t& head, double& w) const {
    const uint32_t e = inc[voff[u] + i];
    const uint32_t p0 = pins[noff[e]], p1 = pins[noff[e] + 1];
    head = p0 != u ? p0 : p1;
    const double deg = double(voff[u + 1] - voff[u]);
    const double we = unit ? 1.0 : double(nw[e]);
    w = wt == 0 ? we : (wt == 1 ? we / 2.0 : we * (deg / 2.0));
  }
};

struct ExplicitView {  // contracted levels
  static constexpr const char* kName = "ExplicitView";
  uint32_t n;
  const uint64_t* idx;
  const uint32_t* head;
  const uint32_t* src;
  const double* w;
  __device__ uint64_t num_nodes() const { return n; }
  __device__ uint64_t offset(uint64_t u) const { return idx[u]; }
  __device__ uint32_t degree(uint32_t u) const { return uint32_t(idx[u + 1] - idx[u]); }
  __device__ uint32_t source(uint64_t a) const { return src[a]; }
  // head id, f64 weight, cluster gather
  __device__ uint32_t arc_bytes(uint32_t) const { return 16u; }
  static constexpr bool kHasUnit = false;
  __host__ __device__ bool unit_arcs() const { return false; }
  __device__ uint32_t head_of(uint32_t u, uint32_t i) const { return head[idx[u] + i]; }
  __device__ void arc(uint32_t u, uint32_t i, uint32_t& h, double& ww) const {
    const uint64_t a = idx[u] + i;
    h = head[a];
    ww = w[a];
  }
};

struct ExplicitGraph {
  uint32_t n = 0;
  uint64_t arcs = 0;
  DevBuf<uint64_t> idx;
  DevBuf<uint32_t> head, src;
  DevBuf<double> w;
  DevBuf<double> vol;
  ExplicitView view() const { return ExplicitView{n, idx.p, head.p, src.p, w.p}; }
};

constexpr int kLvClassesMax = 6;  // degree classes of local moving (see kLvClasses)
constexpr uint32_t kLvBigArcs = 6144;  // at most 3/4 of the 8192-slot map of a block

struct LvArgs {
  uint64_t seed;
  uint32_t level, round;
  int buckets, bucket;
  uint32_t bucket_salt;
  uint32_t sampling;
  double recip;
  const int32_t* C;
  const double* cvol;
  const int32_t* csize;
  const double* nvol;
  // decision per position of the launch's list: the target cluster, or -1
  // (written for every node of the sub-round, so no compaction is needed)
  int32_t* dec;
  // evaluation only (nullptr otherwise): the best gain per list position,
  // max(gain of staying, best candidate's gain) in FP64
  double* gain_out = nullptr;
  // batch of member hypergraphs as one disjoint union (nullptr: one
  // hypergraph): member and member-local id of every node, the member's
  // 1 / total volume, seed and whether it takes part in the round
  const uint32_t* nmember = nullptr;
  const uint32_t* nlocal = nullptr;
  const double* grecip = nullptr;
  const uint8_t* gactive = nullptr;
  const uint64_t* gseed = nullptr;
  // nodes of more than 1024 and at most big_arcs arcs take a block each
  // (class 4), heavier ones the arena path; big_arcs = 1024: every node above
  // 1024 arcs takes the arena path
  uint32_t big_arcs = kLvBigArcs;
};

__device__ inline double lv_recip(const LvArgs& a, uint32_t u) { return a.nmember ? a.grecip[a.nmember[u]] : a.recip; }

__device__ inline uint32_t slot_hash(int32_t key, int bits) {
  return (uint32_t(key) * 0x9E3779B1u) >> (32 - bits);
}

// Inserts (key, w, pos) of every lane of a WIDTH-lane segment into an
// open-addressing map.  Up to `passes` times, the lanes holding the key of the
// lowest pending lane are combined (weight sum, minimum position) and inserted
// once by that lane, so arcs of a heavy node that fall into one cluster cost
// one atomic per pass instead of one each; the remaining lanes insert alone.
template <int WIDTH>
__device__ inline void map_insert_aggregated(int32_t* K, double* Wt, uint32_t* F, int tbits, bool valid,
                                             int32_t key, double w, uint32_t pos, int passes) {
  const int lane = threadIdx.x & 63;
  const int seg_lo = lane & ~(WIDTH - 1);
  const uint64_t seg_mask = (WIDTH == 64) ? ~0ull : (((1ull << (WIDTH & 63)) - 1) << seg_lo);
  bool pending = valid;
  auto insert = [&](int32_t k, double ww, uint32_t pp) {
    const uint32_t T = 1u << tbits;
    uint32_t h = slot_hash(k, tbits);
    while (true) {
      const int32_t old = atomicCAS(&K[h], -1, k);
      if (old == -1 || old == k) break;
      h = (h + 1) & (T - 1);
    }
    atomicAdd(&Wt[h], ww);
    atomicMin(&F[h], pp);
  };
  // every lane of the wave runs every pass (the shuffles need the whole wave)
  for (int it = 0; it < passes; ++it) {
    const uint64_t act = __ballot(pending) & seg_mask;
    const int leader = act ? __ffsll(static_cast<long long>(act)) - 1 : seg_lo;
    const int32_t lk = __shfl(key, leader, 64);
    const bool mine = act != 0 && pending && key == lk;
    double sum = mine ? w : 0.0;
    uint32_t mn = mine ? pos : 0xffffffffu;
    for (int o = WIDTH / 2; o > 0; o >>= 1) {
      sum += __shfl_xor(sum, o, WIDTH);
      mn = min(mn, __shfl_xor(mn, o, WIDTH));
    }
    if (act != 0 && lane == leader) insert(lk, sum, mn);
    if (mine) pending = false;
  }
  if (pending) insert(key, w, pos);
}


// best (gain, first) pair: larger gain, then smaller first occurrence
__device__ inline bool better(double g, uint32_t f, double bg, uint32_t bf) {
  return g > bg || (g == bg && f < bf);
}

__device__ inline void lv_decide(const LvArgs& a, uint32_t pos, uint32_t u, double wfrom, double best_gain,
                                 int32_t best_key) {
  const int32_t from = a.C[u];
  const double volU = a.nvol[u];
  const double mult = lv_recip(a, u) * volU;
  const double bg = wfrom - mult * (a.cvol[from] - volU);
  int32_t to = -1;
  if (best_key >= 0 && best_gain > bg && !(a.csize[from] == 1 && a.csize[best_key] == 1 && best_key > from))
    to = best_key;
  a.dec[pos] = to;
  if (a.gain_out) a.gain_out[pos] = best_key >= 0 && best_gain > bg ? best_gain : bg;
}

// The node's own inputs of the decision, loaded ahead of the arc sweep so
// their latency overlaps it (same values and operation order as lv_decide).
struct LvNode {
  int32_t from;
  double volU, mult, cvol_from;
  int32_t csize_from;
};
__device__ inline void lv_decide(const LvArgs& a, uint32_t pos, const LvNode& nd, double wfrom, double best_gain,
                                 int32_t best_key) {
  const double bg = wfrom - nd.mult * (nd.cvol_from - nd.volU);
  int32_t to = -1;
  if (best_key >= 0 && best_gain > bg && !(nd.csize_from == 1 && a.csize[best_key] == 1 && best_key > nd.from))
    to = best_key;
  a.dec[pos] = to;
  if (a.gain_out) a.gain_out[pos] = best_key >= 0 && best_gain > bg ? best_gain : bg;
}

// ---------------------------------------------------------- move kernels --
// Degree <= 16: one thread per node, no map at all.  The (cluster, weight)
// pairs of the arcs stay in registers; a cluster's weight is summed in arc
// order starting at its first occurrence, and the first occurrence with a
// strictly larger gain wins: the insertion-ordered scan of
// computeMaxGainCluster, summation order included.  The loads are issued in
// three independent batches (arcs, clusters of the heads, cluster volumes) so
// a node waits for three memory round trips, not one per arc.
// Every move body runs as block `bid` of `nb` blocks (a multiple of 8): on its
// own in a kernel of its class, or as one segment of a fused launch.
template <class View, bool UNIT>
__device__ __forceinline__ void lv_small_body(const View& g, const LvArgs& a, const uint32_t* __restrict__ list,
                                              uint32_t count, uint32_t bid, uint32_t nb, uint32_t dbase) {
  constexpr int D = 16;
  // one contiguous chunk of the list per block, chunks dealt per XCD
  const uint32_t per = uint32_t((uint64_t(count) + nb - 1) / nb);
  const uint64_t lo = uint64_t(xcd_chunk(bid, nb)) * per;
  const uint32_t hi = uint32_t(min(uint64_t(count), lo + per));
  for (uint32_t idx = uint32_t(lo) + threadIdx.x; idx < hi; idx += blockDim.x) {
    const uint32_t u = list[idx];
    const uint32_t deg = min(g.degree(u), a.sampling);
    uint32_t head[D];
    // with unit arc weights every weight is 1.0 and none is held (the
    // registers of the weights would cost a wave per SIMD of occupancy)
    double w[UNIT ? 1 : D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      head[i] = u;  // a valid index for the unconditional gathers below
      if constexpr (UNIT) {
        if (uint32_t(i) < deg) head[i] = g.head_of(u, uint32_t(i));
      } else {
        w[i] = 0.0;
        if (uint32_t(i) < deg) g.arc(u, uint32_t(i), head[i], w[i]);
      }
    }
    LvNode nd;
    nd.from = a.C[u];
    nd.volU = a.nvol[u];
    const double rc = lv_recip(a, u);
    int32_t key[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const int32_t c = a.C[head[i]];
      key[i] = uint32_t(i) < deg ? c : -1;
    }
    const int32_t from = nd.from;
    double cv[D];
#pragma unroll
    for (int i = 0; i < D; ++i) cv[i] = a.cvol[key[i] >= 0 ? key[i] : from];
    nd.cvol_from = a.cvol[from];
    nd.csize_from = a.csize[from];
    nd.mult = rc * nd.volU;
    const double mult = nd.mult;
    double wfrom = 0.0, bgain = -DBL_MAX;
    int32_t bkey = -1;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      bool first = key[i] >= 0;
#pragma unroll
      for (int j = 0; j < i; ++j) first = first && key[j] != key[i];
      if (first) {
        double sum = UNIT ? 1.0 : w[UNIT ? 0 : i];
#pragma unroll
        for (int j = i + 1; j < D; ++j)
          if (key[j] == key[i]) sum += UNIT ? 1.0 : w[UNIT ? 0 : j];
        if (key[i] == from) {
          wfrom = sum;
        } else {
          const double gain = sum - mult * cv[i];
          if (gain > bgain) { bgain = gain; bkey = key[i]; }
        }
      }
    }
    lv_decide(a, dbase + idx, nd, wfrom, bgain, bkey);
  }
}

template <class View, bool UNIT>
__global__ void __launch_bounds__(256)
louvain_move_small_kernel(View g, LvArgs a, const uint32_t* __restrict__ list, uint32_t count) {
  lv_small_body<View, UNIT>(g, a, list, count, blockIdx.x, gridDim.x, 0);
}

// G lanes per node, T-slot open-addressing map per group in LDS (16 bytes a
// slot: weight, key, first position).
template <int G, int TBITS, int BLOCK>
constexpr int lv_group_lds() { return (BLOCK / G) * (1 << TBITS) * 16; }

template <class View, int G, int TBITS, int BLOCK>
__device__ __forceinline__ void lv_group_body(const View& g, const LvArgs& a, const uint32_t* __restrict__ list,
                                              uint32_t count, uint32_t bid, uint32_t nb, uint32_t dbase,
                                              unsigned char* lds) {
  constexpr int T = 1 << TBITS;
  constexpr int GPB = BLOCK / G;
  double* wsum = reinterpret_cast<double*>(lds);
  int32_t* keys = reinterpret_cast<int32_t*>(lds + GPB * T * 8);
  uint32_t* first = reinterpret_cast<uint32_t*>(lds + GPB * T * 12);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  int32_t* K = keys + grp * T;
  double* Wt = wsum + grp * T;
  uint32_t* F = first + grp * T;
  // every wave walks its own nodes (64 / G groups per wave): a group's map is
  // touched by its own lanes only, so waves never wait for each other
  constexpr int GPW = 64 / G, WPB = BLOCK / 64;
  const uint32_t wid = threadIdx.x >> 6, wgrp = (threadIdx.x & 63) / G;
  // one contiguous chunk of the list per block, chunks dealt per XCD
  const uint32_t per = uint32_t((uint64_t(count) + nb - 1) / nb);
  const uint64_t lo = uint64_t(xcd_chunk(bid, nb)) * per;
  const uint32_t hi = uint32_t(min(uint64_t(count), lo + per));
  for (uint32_t base = uint32_t(lo) + wid * GPW; base < hi; base += WPB * GPW) {
    const uint32_t idx = base + wgrp;
    const bool in = idx < hi;  // lists hold exactly the nodes of this sub-round
    const uint32_t u = in ? list[idx] : 0;
    const uint32_t deg = in ? min(g.degree(u), a.sampling) : 0;
    // the node's own inputs of the decision, issued ahead of the sweep
    LvNode nd{0, 0.0, 0.0, 0.0, 0};
    double rc = 0.0;
    if (in) {
      nd.from = a.C[u];
      nd.volU = a.nvol[u];
      rc = lv_recip(a, u);
    }
    // the group's map uses the first Tn <= T slots, Tn >= 2 deg
    int tb = 1;
    while (tb < TBITS && (1u << tb) < 2 * deg) ++tb;
    const int Tn = 1 << tb;
    for (int t = lane; t < Tn; t += G) { K[t] = -1; Wt[t] = 0.0; F[t] = 0xffffffffu; }
    lds_wave_sync();
    {
      // trip count uniform over the wave: the aggregation shuffles need it
      uint32_t trips = (deg + G - 1) / G;
      for (int o = 32; o > 0; o >>= 1) trips = max(trips, __shfl_xor(trips, o, 64));
      for (uint32_t t = 0; t < trips; ++t) {
        const uint32_t i = t * G + lane;
        const bool valid = i < deg;
        int32_t c = -1;
        double w = 0.0;
        if (valid) {
          uint32_t v;
          g.arc(u, i, v, w);
          c = a.C[v];
        }
        map_insert_aggregated<G>(K, Wt, F, tb, valid, c, w, i, 2);
      }
    }
    lds_wave_sync();
    if (in) {
      const int32_t from = nd.from;
      if (lane == 0) {  // needed by the decision only; in flight during the scan
        nd.cvol_from = a.cvol[from];
        nd.csize_from = a.csize[from];
      }
      nd.mult = rc * nd.volU;
      const double mult = nd.mult;
      double wfrom = 0.0, bgain = -DBL_MAX;
      uint32_t bfirst = 0xffffffffu;
      int32_t bkey = -1;
      for (int t = lane; t < Tn; t += G) {
        const int32_t k = K[t];
        if (k < 0) continue;
        if (k == from) { wfrom = Wt[t]; continue; }
        const double gain = Wt[t] - mult * a.cvol[k];
        if (better(gain, F[t], bgain, bfirst)) { bgain = gain; bfirst = F[t]; bkey = k; }
      }
      for (int o = G / 2; o > 0; o >>= 1) {
        const double og = __shfl_xor(bgain, o, G);
        const uint32_t of = __shfl_xor(bfirst, o, G);
        const int32_t ok = __shfl_xor(bkey, o, G);
        wfrom += __shfl_xor(wfrom, o, G);
        if (ok >= 0 && (bkey < 0 || better(og, of, bgain, bfirst))) { bgain = og; bfirst = of; bkey = ok; }
      }
      if (lane == 0) lv_decide(a, dbase + idx, nd, wfrom, bgain, bkey);
    }
    lds_wave_sync();
  }
}

template <class View, int G, int TBITS, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
louvain_move_group_kernel(View g, LvArgs a, const uint32_t* __restrict__ list, uint32_t count) {
  __shared__ __align__(16) unsigned char lds[lv_group_lds<G, TBITS, BLOCK>()];
  lv_group_body<View, G, TBITS, BLOCK>(g, a, list, count, blockIdx.x, gridDim.x, 0, lds);
}

// 129..1024 arcs: one block per node; its waves insert into one map of
// T >= 2 deg slots in LDS (up to 2^TBITS), so a node takes deg / BLOCK trips
// instead of deg / 64.  Several waves adding to one slot sum in no fixed
// order: exact under the integer arc weights (the fractional weightings are
// compared by quality only, as in the heavy path).  The best slot is the same
// whatever the order of the scan: larger gain, then smaller first position.
template <int TBITS>
constexpr int lv_block_lds() { return (1 << TBITS) * 16; }

template <class View, int TBITS, int BLOCK>
__device__ __forceinline__ void lv_block_body(const View& g, const LvArgs& a, const uint32_t* __restrict__ list,
                                              uint32_t count, uint32_t bid, uint32_t nb, uint32_t dbase,
                                              unsigned char* lds) {
  constexpr int T = 1 << TBITS;
  constexpr int WPB = BLOCK / 64;
  double* Wt = reinterpret_cast<double*>(lds);
  int32_t* K = reinterpret_cast<int32_t*>(lds + T * 8);
  uint32_t* F = reinterpret_cast<uint32_t*>(lds + T * 12);
  // the per-wave results reuse the start of the map once it has been read
  double* r_gain = reinterpret_cast<double*>(lds);
  double* r_wfrom = r_gain + WPB;
  uint32_t* r_first = reinterpret_cast<uint32_t*>(r_wfrom + WPB);
  int32_t* r_key = reinterpret_cast<int32_t*>(r_first + WPB);
  static_assert(WPB * 24 <= T * 16, "per-wave results fit in the map");
  const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t per = uint32_t((uint64_t(count) + nb - 1) / nb);
  const uint64_t lo = uint64_t(xcd_chunk(bid, nb)) * per;
  const uint32_t hi = uint32_t(min(uint64_t(count), lo + per));
  for (uint32_t idx = uint32_t(lo); idx < hi; ++idx) {
    const uint32_t u = list[idx];
    const uint32_t deg = min(g.degree(u), a.sampling);
    LvNode nd{0, 0.0, 0.0, 0.0, 0};
    nd.from = a.C[u];
    nd.volU = a.nvol[u];
    const double rc = lv_recip(a, u);
    int tb = 1;
    while (tb < TBITS && (1u << tb) < 2 * deg) ++tb;
    const int Tn = 1 << tb;
    for (int t = threadIdx.x; t < Tn; t += BLOCK) { K[t] = -1; Wt[t] = 0.0; F[t] = 0xffffffffu; }
    __syncthreads();
    // the trip count is uniform over the block
    for (uint32_t i0 = 0; i0 < deg; i0 += BLOCK) {
      const uint32_t i = i0 + threadIdx.x;
      const bool valid = i < deg;
      int32_t c = -1;
      double w = 0.0;
      if (valid) {
        uint32_t v;
        g.arc(u, i, v, w);
        c = a.C[v];
      }
      map_insert_aggregated<64>(K, Wt, F, tb, valid, c, w, i, 2);
    }
    if (threadIdx.x == 0) {
      nd.cvol_from = a.cvol[nd.from];
      nd.csize_from = a.csize[nd.from];
    }
    __syncthreads();
    nd.mult = rc * nd.volU;
    double wfrom = 0.0, bgain = -DBL_MAX;
    uint32_t bfirst = 0xffffffffu;
    int32_t bkey = -1;
    for (int t = threadIdx.x; t < Tn; t += BLOCK) {
      const int32_t k = K[t];
      if (k < 0) continue;
      if (k == nd.from) { wfrom = Wt[t]; continue; }
      const double gain = Wt[t] - nd.mult * a.cvol[k];
      if (better(gain, F[t], bgain, bfirst)) { bgain = gain; bfirst = F[t]; bkey = k; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double og = __shfl_xor(bgain, o, 64);
      const uint32_t of = __shfl_xor(bfirst, o, 64);
      const int32_t ok = __shfl_xor(bkey, o, 64);
      wfrom += __shfl_xor(wfrom, o, 64);
      if (ok >= 0 && (bkey < 0 || better(og, of, bgain, bfirst))) { bgain = og; bfirst = of; bkey = ok; }
    }
    __syncthreads();  // every wave has read the map
    if (lane == 0) { r_gain[wid] = bgain; r_first[wid] = bfirst; r_key[wid] = bkey; r_wfrom[wid] = wfrom; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < WPB; ++w) {
        wfrom += r_wfrom[w];
        if (r_key[w] >= 0 && (bkey < 0 || better(r_gain[w], r_first[w], bgain, bfirst))) {
          bgain = r_gain[w]; bfirst = r_first[w]; bkey = r_key[w];
        }
      }
      lv_decide(a, dbase + idx, nd, wfrom, bgain, bkey);
    }
    __syncthreads();  // the map is cleared for the next node
  }
}

template <class View, int TBITS, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
louvain_move_block_kernel(View g, LvArgs a, const uint32_t* __restrict__ list, uint32_t count) {
  __shared__ __align__(16) unsigned char lds[lv_block_lds<TBITS>()];
  lv_block_body<View, TBITS, BLOCK>(g, a, list, count, blockIdx.x, gridDim.x, 0, lds);
}

// One launch for the classes up to 1024 arcs of a sub-round: the grid is cut
// into segments, in this order 129..1024 arcs (a block per node), 65..128
// (32 lanes, 256-slot maps), 17..64 (G1 lanes, 128-slot maps; G1 = 0: that
// class has its own launch) and <= 16 arcs (a thread per node).  Every
// segment has a multiple of 8 blocks, so each keeps its XCD-aware chunking,
// and the long-latency nodes are dispatched first.
struct LvFused {
  uint32_t off[4];     // first list position (and decision slot) per segment
  uint32_t cnt[4];
  uint32_t blocks[4];
};
constexpr int kLvFusedLds = 32768;
constexpr uint64_t kLvFuse1Arcs = 1ull << 21;
static_assert(lv_block_lds<11>() <= kLvFusedLds && lv_group_lds<32, 8, 256>() <= kLvFusedLds &&
                  lv_group_lds<16, 7, 256>() <= kLvFusedLds, "fused move maps fit");

template <class View, int G1, bool UNIT>
__global__ void __launch_bounds__(256)
louvain_move_fused_kernel(View g, LvArgs a, const uint32_t* __restrict__ list, LvFused f) {
  __shared__ __align__(16) unsigned char lds[kLvFusedLds];
  uint32_t b = blockIdx.x;
  if (b < f.blocks[0]) {
    lv_block_body<View, 11, 256>(g, a, list + f.off[0], f.cnt[0], b, f.blocks[0], f.off[0], lds);
    return;
  }
  b -= f.blocks[0];
  if (b < f.blocks[1]) {
    lv_group_body<View, 32, 8, 256>(g, a, list + f.off[1], f.cnt[1], b, f.blocks[1], f.off[1], lds);
    return;
  }
  b -= f.blocks[1];
  if constexpr (G1 != 0) {
    if (b < f.blocks[2]) {
      lv_group_body<View, G1, 7, 256>(g, a, list + f.off[2], f.cnt[2], b, f.blocks[2], f.off[2], lds);
      return;
    }
    b -= f.blocks[2];
  }
  lv_small_body<View, UNIT>(g, a, list + f.off[3], f.cnt[3], b, f.blocks[3], f.off[3]);
}

// ------------------------------------------------------- heavy nodes --
// Nodes with more than 1024 (sampled) arcs.  Their arcs are spread over the
// whole grid instead of one block per node.  Every heavy node of a sub-round
// owns a table of T = pow2 >= 2d slots in one arena.  The arena starts empty
// (key kHvEmpty, weight 0, first position ~0), and the scan that reads a
// sub-round's tables leaves every slot it read empty again, so a pass over
// the arcs can claim a slot and add to it at once.
//   hv_prep     d and T per node            (then exclusive scans: arc and
//                                             table offsets)
//   hv_insert   per arc: find or claim the slot of its cluster, then add the
//               weight and the arc position (wave-aggregated per slot)
//   hv_scan     per chunk of slots: best candidate; the slots are emptied
//   hv_decide   per node: best cluster over its chunks, then the move rule
constexpr uint64_t kHvEmpty = ~0ull;
struct HeavyArgs {
  const uint32_t* list;  // heavy nodes of the sub-round
  uint32_t count;
  const uint64_t* aoff;  // count + 1 arc offsets
  const uint64_t* toff;  // count + 1 table offsets
  uint64_t* keys;        // cluster id, or kHvEmpty
  double* wsum;
  uint32_t* first;
};

template <class View>
__global__ void hv_prep_kernel(View g, LvArgs a, const uint32_t* __restrict__ list, uint32_t count,
                               uint64_t* __restrict__ dsz, uint64_t* __restrict__ tsz) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const uint32_t d = min(g.degree(list[i]), a.sampling);
    uint64_t T = 2;
    while (T < 2ull * d) T <<= 1;
    dsz[i] = d;
    tsz[i] = T;
  }
}

// largest i with aoff[i] <= j
__device__ inline uint32_t hv_node_of(const uint64_t* aoff, uint32_t count, uint64_t j) {
  uint32_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (aoff[mid] <= j) lo = mid; else hi = mid;
  }
  return lo;
}

// aoff / toff hold offsets of the whole round's heavy nodes; a sub-round's
// arcs and table slots are counted from its first node's offsets.  A block
// takes kHvArcs consecutive arcs: every heavy node has more, so the chunk
// spans at most two nodes, found with one search per block.
constexpr uint32_t kHvArcs = 1024;

__device__ inline uint32_t hv_chunk_node(const HeavyArgs& h, uint64_t a0, uint64_t c0, uint32_t* s_node) {
  if (threadIdx.x == 0) *s_node = hv_node_of(h.aoff, h.count, a0 + c0);
  __syncthreads();
  return *s_node;
}

template <class View>
__global__ void __launch_bounds__(256) hv_insert_kernel(View g, LvArgs a, HeavyArgs h) {
  __shared__ uint32_t s_node;
  const uint64_t a0 = h.aoff[0], t0 = h.toff[0];
  const uint64_t total = h.aoff[h.count] - a0;
  const int lane = threadIdx.x & 63;
  for (uint64_t c0 = blockIdx.x * uint64_t(kHvArcs); c0 < total; c0 += uint64_t(gridDim.x) * kHvArcs) {
    const uint32_t i0 = hv_chunk_node(h, a0, c0, &s_node);
    const uint64_t c1 = min(c0 + kHvArcs, total), split = h.aoff[i0 + 1] - a0;
    // whole waves iterate together (the aggregation shuffles need every lane)
    for (uint64_t j0 = c0 + (threadIdx.x & ~63u); j0 < c1; j0 += blockDim.x) {
      const uint64_t j = j0 + lane;
      bool pending = j < c1;
      uint32_t slot = 0xffffffffu, pos = 0;
      double w = 0.0;
      if (pending) {
        const uint32_t i = j < split ? i0 : i0 + 1;
        pos = uint32_t(a0 + j - h.aoff[i]);
        uint32_t v;
        g.arc(h.list[i], pos, v, w);
        const uint64_t c = uint32_t(a.C[v]);
        const uint64_t base = h.toff[i] - t0;
        const uint32_t T = uint32_t(h.toff[i + 1] - h.toff[i]);
        uint32_t t = slot_hash(int32_t(c), __builtin_ctz(T));
        // a slot's key goes from empty to a cluster once per sub-round, so a
        // plain read that shows a cluster is exact; an empty one is settled
        // by the CAS
        uint64_t cur = h.keys[base + t];
        while (cur != c) {
          if (cur != kHvEmpty) {
            t = (t + 1) & (T - 1);
            cur = h.keys[base + t];
            continue;
          }
          cur = atomicCAS(reinterpret_cast<unsigned long long*>(&h.keys[base + t]),
                          static_cast<unsigned long long>(kHvEmpty), static_cast<unsigned long long>(c));
          if (cur == kHvEmpty) break;  // claimed
        }
        slot = uint32_t(base + t);
      }
      for (int it = 0; it < 4; ++it) {  // combine lanes of the lowest pending slot
        const uint64_t act = __ballot(pending);
        if (!act) break;
        const int leader = __ffsll(static_cast<long long>(act)) - 1;
        const uint32_t ls = __shfl(slot, leader, 64);
        const bool mine = pending && slot == ls;
        double sum = mine ? w : 0.0;
        uint32_t mn = mine ? pos : 0xffffffffu;
        for (int o = 32; o > 0; o >>= 1) {
          sum += __shfl_xor(sum, o, 64);
          mn = min(mn, __shfl_xor(mn, o, 64));
        }
        if (lane == leader) {
          atomicAdd(&h.wsum[ls], sum);
          atomicMin(&h.first[ls], mn);
        }
        if (mine) pending = false;
      }
      if (pending) {
        atomicAdd(&h.wsum[slot], w);
        atomicMin(&h.first[slot], pos);
      }
    }
    __syncthreads();  // s_node is rewritten for the next chunk
  }
}

// Best candidate of one chunk of kHvChunk table slots (chunks never straddle
// two tables: T is a power of two >= 2048 for heavy nodes, and tables start
// at multiples of their own size's divisor kHvChunk).
constexpr uint32_t kHvChunk = 1024;
struct HvBest {
  double gain;
  double wfrom;
  uint32_t first;
  int32_t key;
};

template <class View>
__global__ void __launch_bounds__(256) hv_scan_kernel(View g, LvArgs a, HeavyArgs h, HvBest* __restrict__ out) {
  __shared__ HvBest s_b[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t t0 = h.toff[0];
  const uint64_t chunks = (h.toff[h.count] - t0) / kHvChunk;
  for (uint64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const uint64_t s0 = ch * kHvChunk;
    const uint32_t i = hv_node_of(h.toff, h.count, t0 + s0);
    const uint32_t u = h.list[i];
    const int32_t from = a.C[u];
    const double mult = lv_recip(a, u) * a.nvol[u];
    double wfrom = 0.0, bgain = -DBL_MAX;
    uint32_t bfirst = 0xffffffffu;
    int32_t bkey = -1;
    for (uint32_t t = threadIdx.x; t < kHvChunk; t += blockDim.x) {
      const uint64_t kk = h.keys[s0 + t];
      if (kk == kHvEmpty) continue;
      const int32_t k = int32_t(uint32_t(kk));
      const double ws = h.wsum[s0 + t];
      const uint32_t f = h.first[s0 + t];
      h.keys[s0 + t] = kHvEmpty;  // empty again for the next sub-round
      h.wsum[s0 + t] = 0.0;
      h.first[s0 + t] = 0xffffffffu;
      if (k == from) { wfrom = ws; continue; }
      const double gain = ws - mult * a.cvol[k];
      if (better(gain, f, bgain, bfirst)) { bgain = gain; bfirst = f; bkey = k; }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double og = __shfl_xor(bgain, o, 64);
      const uint32_t of = __shfl_xor(bfirst, o, 64);
      const int32_t ok = __shfl_xor(bkey, o, 64);
      wfrom += __shfl_xor(wfrom, o, 64);
      if (ok >= 0 && (bkey < 0 || better(og, of, bgain, bfirst))) { bgain = og; bfirst = of; bkey = ok; }
    }
    if (lane == 0) s_b[wid] = HvBest{bgain, wfrom, bfirst, bkey};
    __syncthreads();
    if (threadIdx.x == 0) {
      HvBest r{-DBL_MAX, 0.0, 0xffffffffu, -1};
      for (int x = 0; x < int(blockDim.x >> 6); ++x) {
        r.wfrom += s_b[x].wfrom;
        if (s_b[x].key >= 0 && (r.key < 0 || better(s_b[x].gain, s_b[x].first, r.gain, r.first))) {
          r.gain = s_b[x].gain; r.first = s_b[x].first; r.key = s_b[x].key;
        }
      }
      out[ch] = r;
    }
    __syncthreads();
  }
}

// per heavy node: combine its chunks, then the move rule
__global__ void hv_decide_kernel(LvArgs a, HeavyArgs h, const HvBest* __restrict__ best) {
  const uint64_t t0 = h.toff[0];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < h.count; i += gridDim.x * blockDim.x) {
    const uint64_t c0 = (h.toff[i] - t0) / kHvChunk, c1 = (h.toff[i + 1] - t0) / kHvChunk;
    HvBest r{-DBL_MAX, 0.0, 0xffffffffu, -1};
    for (uint64_t ch = c0; ch < c1; ++ch) {
      const HvBest b = best[ch];
      r.wfrom += b.wfrom;
      if (b.key >= 0 && (r.key < 0 || better(b.gain, b.first, r.gain, r.first))) {
        r.gain = b.gain; r.first = b.first; r.key = b.key;
      }
    }
    lv_decide(a, i, h.list[i], r.wfrom, r.gain, r.key);
  }
}

// Adds (dv, dn) to cvol[key] / csize[key] for every valid lane of the wave.
// The lanes sharing the key of the first valid lane add their sum once (the
// moves of neighbouring nodes often share a cluster); the others add alone.
// Volumes are integer-valued under the integer arc weights of the star
// expansion, so the order of the additions does not change them.
__device__ inline void lv_volume_add(double* __restrict__ cvol, int32_t* __restrict__ csize, bool valid, int32_t key,
                                     double dv, int dn) {
  const int lane = threadIdx.x & 63;
  bool pending = valid;
  // up to four passes, each combining the lanes of the lowest pending key
  for (int it = 0; it < 4; ++it) {
    const uint64_t act = __ballot(pending);
    if (!act) return;
    const int leader = __ffsll(static_cast<long long>(act)) - 1;
    const int32_t lk = __shfl(key, leader, 64);
    const bool mine = pending && key == lk;
    double sum = mine ? dv : 0.0;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    const int cnt = __popcll(__ballot(mine));
    if (lane == leader) {
      atomicAdd(&cvol[lk], sum);
      atomicAdd(&csize[lk], dn * cnt);
    }
    if (mine) pending = false;
  }
  if (pending) {
    atomicAdd(&cvol[key], dv);
    atomicAdd(&csize[key], dn);
  }
}

// The work-list ranges of one sub-round (one per degree class).
struct LvRanges {
  uint32_t off[kLvClassesMax], cnt[kLvClassesMax];
  uint32_t total;
};

// Applies a sub-round's moves: every position of its list ranges whose
// decision is a cluster moves its node.  Whole waves step together (the
// volume and member updates combine across lanes); each block adds its move
// count once.
__global__ void __launch_bounds__(256)
louvain_apply_kernel(const uint32_t* __restrict__ list, const int32_t* __restrict__ dec, LvRanges rg,
                     const double* __restrict__ nvol, int32_t* __restrict__ C, double* __restrict__ cvol,
                     int32_t* __restrict__ csize, uint32_t* __restrict__ moved_total,
                     const uint32_t* __restrict__ nmember, uint32_t* __restrict__ gmoved) {
  __shared__ uint32_t bmoves;
  if (threadIdx.x == 0) bmoves = 0;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  const int lane = threadIdx.x & 63;
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < rg.total; base += stride) {
    const uint32_t j = base + uint32_t(lane);
    bool valid = false;
    uint32_t u = 0;
    int32_t to = 0, from = 0;
    double vu = 0.0;
    if (j < rg.total) {
      uint32_t r = 0, k = j;
      while (k >= rg.cnt[r]) { k -= rg.cnt[r]; ++r; }
      const uint32_t pos = rg.off[r] + k;
      to = dec[pos];
      if (to >= 0) {
        valid = true;
        u = list[pos];
        from = C[u];
        C[u] = to;
        vu = nvol[u];
      }
    }
    const uint64_t movers = __ballot(valid);
    if (lane == 0 && movers) atomicAdd(&bmoves, uint32_t(__popcll(movers)));
    if (nmember) {  // moves per member (batch): consecutive moves mostly share the member
      const uint32_t g = valid ? nmember[u] : 0u;
      const uint64_t act = movers;
      const int leader = act ? __ffsll(static_cast<long long>(act)) - 1 : 0;
      const uint32_t lg = __shfl(g, leader, 64);
      const bool mine = valid && g == lg;
      const int c = __popcll(__ballot(mine));
      if (act && lane == leader) atomicAdd(&gmoved[lg], uint32_t(c));
      if (valid && !mine) atomicAdd(&gmoved[g], 1u);
    }
    lv_volume_add(cvol, csize, valid, to, vu, 1);
    lv_volume_add(cvol, csize, valid, from, -vu, -1);
  }
  __syncthreads();
  if (threadIdx.x == 0 && bmoves) atomicAdd(moved_total, bmoves);
}

// Maximum of v over the block, added to *out with one atomic per block
// (atomics on one address run one after another in memory, so the grids
// that use it are kept to a few hundred blocks).
__device__ inline void block_max_to(uint32_t* out, uint32_t v) {
  __shared__ uint32_t wmax[16];
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = 0;
    for (uint32_t w = 0; w < (blockDim.x + 63) / 64; ++w) m = max(m, wmax[w]);
    atomicMax(out, m);
  }
}

// node volumes (graph.cpp:319-327), sequential sum in arc order per node
template <class View>
__device__ inline bool constant_arc_weights(const View&, uint32_t) { return false; }
__device__ inline bool constant_arc_weights(const StarView& g, uint32_t u) { return g.wt == 0 && u >= g.N; }
// Under the uniform weighting a hypernode's arcs weigh w(e) of its incident
// nets: the sequential sum is the sum of the net weights (its degree for
// unit weights), read without the arc heads.
template <class View>
__device__ inline bool net_weight_sum(const View&, uint32_t, double&) { return false; }
__device__ inline bool net_weight_sum(const StarView& g, uint32_t u, double& s) {
  if (g.wt != 0 || u >= g.N) return false;
  const uint32_t jb = g.voff[u], je = g.voff[u + 1];
  if (g.unit) {
    s = double(je - jb);
    return true;
  }
  double acc = 0.0;
  for (uint32_t j = jb; j < je; ++j) acc += double(g.nw[g.inc[j]]);
  s = acc;
  return true;
}

template <class View>
__global__ void node_volume_kernel(View g, uint32_t n, double* __restrict__ vol, uint32_t* __restrict__ deg_max) {
  uint32_t dm = 0;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint32_t d = g.degree(u);
    double s = 0.0;
    if (constant_arc_weights(g, u) && d > 0) {
      // a net node under uniform weights: |e| arcs of weight w(e); the
      // sequential sum of equal integer-valued terms is exactly |e| * w(e)
      uint32_t v;
      double w;
      g.arc(u, 0, v, w);
      s = double(d) * w;
    } else if (net_weight_sum(g, u, s)) {
    } else {
      for (uint32_t i = 0; i < d; ++i) {
        uint32_t v;
        double w;
        g.arc(u, i, v, w);
        s += w;
      }
    }
    vol[u] = s;
    dm = max(dm, d);
  }
  block_max_to(deg_max, dm);
}

template <class View>
__global__ void max_degree_kernel(View g, uint32_t n, uint32_t* __restrict__ deg_max) {
  uint32_t dm = 0;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    dm = max(dm, g.degree(u));
  block_max_to(deg_max, dm);
}

// degree classes: 0: <= 16 arcs, 1: <= 64, 2: <= 128, 3: <= 1024, 4: <=
// kLvBigArcs (a 1024-thread block per node, 8192-slot LDS map), 5: heavier
// (the arc-parallel arena path); the work list of a round is grouped by
// (class, bucket) so that every sub-round launch covers exactly its own
// nodes.  Classes 1 and 2 run the same group kernel with maps of 128 and 256
// slots: the smaller maps of the 17..64-arc nodes leave room in LDS for twice
// the waves.
constexpr int kLvClasses = 6;
constexpr int kLvHeavy = 5;  // class of the arena path
static_assert(kLvClasses <= kLvClassesMax, "class ranges of a sub-round");
// In a batch, nodes of members that sit the round out get no slot (-1), and
// the bucket hash takes the member's seed and local id.
template <class View>
__device__ inline int lv_slot(const View& g, const LvArgs& a, uint32_t u) {
  const uint32_t d = min(g.degree(u), a.sampling);
  const int c = d <= 16 ? 0 : (d <= 64 ? 1 : (d <= 128 ? 2 : (d <= 1024 ? 3 : (d <= a.big_arcs ? 4 : kLvHeavy))));
  int b = 0;
  if (a.nmember) {
    const uint32_t gm = a.nmember[u];
    if (!a.gactive[gm]) return -1;
    if (a.buckets > 1) {
      const uint32_t salt = bucket_salt(a.gseed[gm], kLvBucketTag | (uint64_t(a.level) << 8) | uint64_t(a.round));
      b = int(bucket_of(salt, a.nlocal[u] >> kLvTileShift, uint32_t(a.buckets)));
    }
  } else if (a.buckets > 1) {
    b = int(bucket_of(a.bucket_salt, u >> kLvTileShift, uint32_t(a.buckets)));
  }
  return c * a.buckets + b;
}

// Per-block LDS histograms over a contiguous range of nodes; a block flushes
// once, so global atomics stay at one per (block, slot) for a grid of at most
// kLvListBlocks blocks whatever the node count.  The block also stores its
// count of every slot in bmat (slot-major), whose exclusive scan places the
// blocks' ranges of a slot in block order: every work list is then in
// ascending node order up to the order inside one block's range, so list
// neighbours are id neighbours and the XCD-chunked move kernels gather from
// one region of the cluster arrays per XCD.
constexpr unsigned kLvListBlocks = 1024;
inline unsigned lv_list_blocks(uint32_t n) { return grid_for(n, 256, kLvListBlocks); }
// the block's node range: tile-aligned chunks (a wave of 64 consecutive
// nodes then holds one tile, whose nodes share their bucket)
__device__ inline void lv_block_range(uint32_t n, uint32_t& lo, uint32_t& hi) {
  uint64_t per = (uint64_t(n) + gridDim.x - 1) / gridDim.x;
  per = (per + 63) & ~uint64_t(63);
  lo = uint32_t(min(uint64_t(blockIdx.x) * per, uint64_t(n)));
  hi = uint32_t(min(uint64_t(lo) + per, uint64_t(n)));
}

template <class View>
__global__ void __launch_bounds__(256)
lv_count_kernel(View g, LvArgs a, uint32_t n, uint32_t* __restrict__ bmat, uint64_t* __restrict__ amat,
                uint64_t* __restrict__ bymat) {
  __shared__ uint32_t h[kLvClasses * 256];
  __shared__ unsigned long long ha[kLvClasses * 256];
  __shared__ unsigned long long hb[kLvClasses * 256];
  const int slots = kLvClasses * a.buckets;
  for (int i = threadIdx.x; i < slots; i += blockDim.x) { h[i] = 0; ha[i] = 0; hb[i] = 0; }
  __syncthreads();
  uint32_t lo, hi;
  lv_block_range(n, lo, hi);
  // whole waves step together: lanes of one slot add with one LDS atomic
  for (uint32_t base = lo + (threadIdx.x & ~63u); base < hi; base += blockDim.x) {
    const uint32_t u = base + (threadIdx.x & 63u);
    const int slot = u < hi ? lv_slot(g, a, u) : -1;
    const uint32_t d = slot >= 0 ? min(g.degree(u), a.sampling) : 0u;
    const uint64_t by = slot >= 0 && bymat ? 28ull + uint64_t(d) * g.arc_bytes(u) : 0ull;
    const SlotGroup sg = slot_group(slot);
    // per-slot sums of the arcs and bytes: one shuffle reduction per
    // distinct slot of the wave, led by its lowest lane
    uint64_t pending = __ballot(slot >= 0);
    while (pending) {
      const int first = __ffsll(static_cast<long long>(pending)) - 1;
      const int s = __shfl(slot, first, 64);
      const bool mine = slot == s && slot >= 0;
      uint64_t sa = mine ? uint64_t(d) : 0ull, sb = mine ? by : 0ull;
      for (int o = 32; o > 0; o >>= 1) {
        sa += __shfl_xor(sa, o, 64);
        sb += __shfl_xor(sb, o, 64);
      }
      if ((threadIdx.x & 63) == uint32_t(first)) {
        atomicAdd(&ha[s], static_cast<unsigned long long>(sa));
        if (bymat) atomicAdd(&hb[s], static_cast<unsigned long long>(sb));
      }
      pending &= ~__ballot(mine);
    }
    if (sg.leader) atomicAdd(&h[slot], sg.count);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < slots; i += blockDim.x) {
    const size_t at = size_t(i) * gridDim.x + blockIdx.x;
    bmat[at] = h[i];
    amat[at] = ha[i];
    if (bymat) bymat[at] = hb[i];
  }
}

// One block per slot: the slot's node count, arc count and bytes over the
// list blocks, and the exclusive scan of its per-block node counts in place
// (offsets inside the slot's list range).  No global atomics: a grid of
// counting blocks would otherwise queue on each slot's counter.
__global__ void __launch_bounds__(256)
lv_totals_kernel(uint32_t nb, uint32_t* __restrict__ bmat, const uint64_t* __restrict__ amat,
                 const uint64_t* __restrict__ bymat, uint32_t* __restrict__ counts, uint64_t* __restrict__ sarcs,
                 uint64_t* __restrict__ sbytes) {
  __shared__ uint32_t part[256];
  __shared__ unsigned long long ra[4], rb[4];
  const uint32_t slot = blockIdx.x, t = threadIdx.x;
  uint32_t* row = bmat + size_t(slot) * nb;
  const uint32_t per = (nb + blockDim.x - 1) / blockDim.x;
  const uint32_t b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  uint32_t sum = 0;
  uint64_t sa = 0, sb = 0;
  for (uint32_t b = b0; b < b1; ++b) {
    sum += row[b];
    sa += amat[size_t(slot) * nb + b];
    if (bymat) sb += bymat[size_t(slot) * nb + b];
  }
  part[t] = sum;
  for (int o = 32; o > 0; o >>= 1) {
    sa += __shfl_xor(sa, o, 64);
    sb += __shfl_xor(sb, o, 64);
  }
  if ((t & 63) == 0) { ra[t >> 6] = sa; rb[t >> 6] = sb; }
  __syncthreads();
  if (t == 0) {  // exclusive scan of the thread sums (256 values)
    uint32_t run = 0;
    for (uint32_t i = 0; i < blockDim.x; ++i) {
      const uint32_t x = part[i];
      part[i] = run;
      run += x;
    }
    counts[slot] = run;
    sarcs[slot] = ra[0] + ra[1] + ra[2] + ra[3];
    if (sbytes) sbytes[slot] = rb[0] + rb[1] + rb[2] + rb[3];
  }
  __syncthreads();
  uint32_t run = part[t];
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t x = row[b];
    row[b] = run;
    run += x;
  }
}

// Places the nodes of the block's range (the range of the same block of
// lv_count_kernel, same grid) at the block's scanned position of their slot.
template <class View>
__global__ void __launch_bounds__(256)
lv_scatter_kernel(View g, LvArgs a, uint32_t n, const uint32_t* __restrict__ bpos, const uint32_t* __restrict__ counts,
                  uint32_t* __restrict__ list) {
  __shared__ uint32_t h[kLvClasses * 256];
  const int slots = kLvClasses * a.buckets;
  if (threadIdx.x == 0) {  // start of every slot's range: the counts of the slots before it
    uint32_t run = 0;
    for (int i = 0; i < slots; ++i) {
      h[i] = run;
      run += counts[i];
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < slots; i += blockDim.x) h[i] += bpos[size_t(i) * gridDim.x + blockIdx.x];
  __syncthreads();
  uint32_t lo, hi;
  lv_block_range(n, lo, hi);
  // whole waves step together: the lanes of one slot reserve their places
  // with one LDS atomic and keep their lane order inside it
  for (uint32_t base = lo + (threadIdx.x & ~63u); base < hi; base += blockDim.x) {
    const uint32_t u = base + (threadIdx.x & 63u);
    const int slot = u < hi ? lv_slot(g, a, u) : -1;
    const SlotGroup sg = slot_group(slot);
    uint32_t start = 0;
    if (sg.leader) start = atomicAdd(&h[slot], sg.count);
    // the slot's start travels from its leader to its lanes
    uint64_t pending = __ballot(slot >= 0);
    while (pending) {
      const int first = __ffsll(static_cast<long long>(pending)) - 1;
      const int s = __shfl(slot, first, 64);
      const uint32_t st = __shfl(start, first, 64);
      const bool mine = slot == s && slot >= 0;
      if (mine) list[st + sg.rank] = u;
      pending &= ~__ballot(mine);
    }
  }
}

__global__ void lv_init_kernel(uint32_t n, const double* __restrict__ nvol, int32_t* __restrict__ C,
                               double* __restrict__ cvol, int32_t* __restrict__ csize) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    C[u] = int32_t(u);
    cvol[u] = nvol[u];
    csize[u] = 1;
  }
}

// ----------------------------------------------------------- contraction --
__global__ void mark_used_kernel(const int32_t* __restrict__ C, uint32_t n, uint32_t* __restrict__ used) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) used[C[u]] = 1;
}

__global__ void remap_kernel(int32_t* __restrict__ C, uint32_t n, const uint32_t* __restrict__ map) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) C[u] = int32_t(map[C[u]]);
}

__global__ void coarse_volume_kernel(const uint32_t* __restrict__ used, const uint32_t* __restrict__ map,
                                     const double* __restrict__ cvol, uint32_t n, double* __restrict__ out) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    if (used[c]) out[map[c]] = cvol[c];
}

template <class View>
__device__ inline uint32_t arc_source(const View& g, uint64_t a) {
  uint64_t lo = 0, hi = g.num_nodes();  // largest u with offset(u) <= a
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (g.offset(mid) <= a) lo = mid; else hi = mid;
  }
  return uint32_t(lo);
}

template <class View>
__global__ void emit_arcs_kernel(View g, uint64_t num_arcs, const int32_t* __restrict__ C, int bits,
                                 uint64_t* __restrict__ keys, double* __restrict__ vals,
                                 const uint32_t* __restrict__ nmember, const uint8_t* __restrict__ gkeep,
                                 uint32_t* __restrict__ keep) {
  const uint64_t sentinel = uint64_t(1) << (2 * bits);
  for (uint64_t a = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; a < num_arcs;
       a += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t u = g.source(a);
    if (nmember && !gkeep[nmember[u]]) {  // batch: the member ends at this level
      keys[a] = sentinel;
      vals[a] = 0.0;
      if (keep) keep[a] = 0;
      continue;
    }
    uint32_t v;
    double w;
    g.arc(u, uint32_t(a - g.offset(u)), v, w);
    const uint64_t cu = uint32_t(C[u]), cv = uint32_t(C[v]);
    keys[a] = cu == cv ? sentinel : ((cu << bits) | cv);
    vals[a] = w;
    if (keep) keep[a] = cu == cv ? 0u : 1u;
  }
}

// stable compaction of the kept arcs (emission order preserved)
__global__ void compact_arcs_kernel(const uint64_t* __restrict__ keys, const double* __restrict__ vals,
                                    const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos, uint64_t n,
                                    uint64_t* __restrict__ okeys, double* __restrict__ ovals) {
  for (uint64_t a = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; a < n; a += uint64_t(gridDim.x) * blockDim.x) {
    if (keep[a]) {
      okeys[pos[a]] = keys[a];
      ovals[pos[a]] = vals[a];
    }
  }
}

__global__ void coarse_csr_kernel(const uint64_t* __restrict__ ukeys, uint64_t ucount, int bits,
                                  uint32_t* __restrict__ head, uint32_t* __restrict__ src,
                                  uint64_t* __restrict__ ends) {
  const uint64_t mask = (uint64_t(1) << bits) - 1;
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < ucount;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t cu = uint32_t(ukeys[i] >> bits);
    head[i] = uint32_t(ukeys[i] & mask);
    src[i] = cu;
    // keys are sorted by source: the last arc of each source writes its degree
    if (i + 1 == ucount || uint32_t(ukeys[i + 1] >> bits) != cu) ends[cu] = i + 1;
  }
}

__global__ void u32_to_u64_kernel(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = in[i];
}

__global__ void prolong_kernel(int32_t* __restrict__ C, uint32_t n, const int32_t* __restrict__ coarse) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) C[u] = coarse[C[u]];
}

__global__ void all_size_two_kernel(const uint32_t* __restrict__ noff, uint32_t m, uint32_t* __restrict__ flag) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
    if (noff[e + 1] - noff[e] != 2) *flag = 0;
}

__global__ void sum_volume_kernel(const double* __restrict__ vol, uint32_t n, double* __restrict__ out) {
  // fixed-shape reduction; exact for integer-valued volumes
  __shared__ double part[256];
  double s = 0.0;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) s += vol[u];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

// ------------------------------------------------------------ evaluation --
// Cluster volumes and sizes of a given clustering (sums of node volumes;
// exact under integer arc weights).
__global__ void cluster_state_kernel(const int32_t* __restrict__ C, const double* __restrict__ nvol, uint32_t n,
                                     double* __restrict__ cvol, int32_t* __restrict__ csize) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    atomicAdd(&cvol[C[u]], nvol[u]);
    atomicAdd(&csize[C[u]], 1);
  }
}

// metrics::modularity (local_moving_modularity.cpp:28-56), first pass: per
// cluster its total volume and its internal volume (node volumes minus the
// weights of arcs leaving the cluster).
template <class View>
__global__ void modularity_volume_kernel(View g, const int32_t* __restrict__ C, const double* __restrict__ nvol,
                                         uint32_t n, double* __restrict__ internal, double* __restrict__ total) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const int32_t cu = C[u];
    double cut = 0.0;
    const uint32_t d = g.degree(u);
    for (uint32_t i = 0; i < d; ++i) {
      uint32_t v;
      double w;
      g.arc(u, i, v, w);
      if (C[v] != cu) cut += w;
    }
    atomicAdd(&total[cu], nvol[u]);
    atomicAdd(&internal[cu], nvol[u] - cut);
  }
}

// second pass: sum over clusters with volume of internal - total^2 / TV in a
// fixed shape (64 blocks of 256 threads, the partials summed on the host)
__global__ void __launch_bounds__(256)
modularity_sum_kernel(const double* __restrict__ internal, const double* __restrict__ total, uint32_t n, double tv,
                      double* __restrict__ out) {
  __shared__ double part[256];
  double s = 0.0;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
    if (total[c] > 0.0) s += internal[c] - (total[c] * total[c]) / tv;
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (int(threadIdx.x) < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

// --------------------------------------------------------------- driver --
struct LouvainRun {
  Context& c;
  const mkp_louvain_params& prm;
  mkp_louvain_stats st{};
  explicit LouvainRun(Context& ctx, const mkp_louvain_params& p) : c(ctx), prm(p) {}

  double total_volume(const double* vol, uint32_t n) {
    DevBuf<double> partial(64);
    sum_volume_kernel<<<64, 256, 0, c.stream>>>(vol, n, partial.p);
    MKP_LAUNCH_CHECK();
    double h[64];
    MKP_HIP(hipMemcpyAsync(h, partial.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    double s = 0.0;
    for (double x : h) s += x;
    return s;
  }

  // launches the four degree classes of sub-round `sr`; lists are grouped by
  // slot = class * B + bucket with offsets off[] and sizes cnt[]
  // (sb: algorithmic bytes per slot, profiling only); each launch is timed
  // on its own so the per-kernel averages match a kernel trace
  // heavy-node arena and per-sub-round arrays (see hv_* kernels)
  DevBuf<uint64_t> hkeys_, hdsz_, htsz_, haoff_, htoff_;
  DevBuf<double> hw_;
  DevBuf<uint32_t> hf_;
  DevBuf<HvBest> hbest_;

  // reserves the arena for tables of up to 4 * arcs slots (T < 4d); new
  // memory is filled empty, and every scan leaves its slots empty
  void reserve_heavy(uint64_t arcs, uint32_t nodes) {
    const hipStream_t s = c.stream;
    const uint64_t slots = std::max<uint64_t>(4 * arcs, 2);
    MKP_CHECK(slots < (uint64_t(1) << 32), "heavy-node tables exceed 2^32 slots");
    if (slots > hkeys_.n) {
      hkeys_.alloc(slots);
      hw_.alloc(slots);
      hf_.alloc(slots);
      fill_async<uint64_t>(hkeys_.p, kHvEmpty, slots, s);
      fill_async<double>(hw_.p, 0.0, slots, s);
      fill_async<uint32_t>(hf_.p, 0xffffffffu, slots, s);
    }
    hbest_.ensure(slots / kHvChunk + 1);
    hdsz_.ensure(nodes + 1);
    htsz_.ensure(nodes + 1);
    haoff_.ensure(nodes + 1);
    htoff_.ensure(nodes + 1);
  }

  // arc and table offsets of all heavy nodes of a round (their list slots
  // are contiguous: class 3 of every bucket)
  template <class View>
  void prepare_heavy(const View& g, const LvArgs& a, const uint32_t* list, uint32_t count) {
    const hipStream_t s = c.stream;
    hv_prep_kernel<View><<<grid_for(count, 256, 4096), 256, 0, s>>>(g, a, list, count, hdsz_.p, htsz_.p);
    MKP_LAUNCH_CHECK();
    exclusive_scan_u64(c, hdsz_.p, haoff_.p, count);
    exclusive_scan_u64(c, htsz_.p, htoff_.p, count);
  }

  // first: index of the sub-round's first heavy node in the round's list
  template <class View>
  void launch_heavy(const View& g, const LvArgs& a, const uint32_t* list, uint32_t first, uint32_t count,
                    uint64_t arcs, double bytes, hipStream_t s) {
    HeavyArgs h{list, count, haoff_.p + first, htoff_.p + first, hkeys_.p, hw_.p, hf_.p};
    const unsigned grid = grid_for(arcs, kHvArcs, unsigned(c.num_cus) * 16);
    {
      // the compulsory bytes of the heavy moves (node data, head id and
      // cluster gather per arc) are attributed to the gather pass
      TimedRegion t1(c.prof, s, kFamLouvainMove, bytes, std::string("hv_insert_kernel<") + View::kName + ">");
      hv_insert_kernel<View><<<grid, 256, 0, s>>>(g, a, h);
      MKP_LAUNCH_CHECK();
    }
    {
      TimedRegion t3(c.prof, s, kFamLouvainMove, 0.0, std::string("hv_scan_kernel<") + View::kName + ">");
      const unsigned chunks = unsigned(std::min<uint64_t>(4 * arcs / kHvChunk + 1, uint64_t(c.num_cus) * 16));
      hv_scan_kernel<View><<<chunks, 256, 0, s>>>(g, a, h, hbest_.p);
      MKP_LAUNCH_CHECK();
    }
    hv_decide_kernel<<<grid_for(count, 256, 4096), 256, 0, s>>>(a, h, hbest_.p);
    MKP_LAUNCH_CHECK();
  }

  // ----- autotuner of the lanes per node of the 17..128-arc class.  The
  // shapes compute the same result; in the first round of a level with an
  // untuned key, sub-rounds 0..7 run them in the order 8 16 32 64 64 32 16 8
  // with an event pair around the class launch (mirrored, so that a drift of
  // the per-arc cost over the round weighs on every shape alike), and the
  // shape with the least time per arc over its two sub-rounds is cached in
  // the context.  With fewer than 8 sub-rounds every shape runs once.
  static constexpr int kLaneTrials[4] = {8, 16, 32, 64};
  static constexpr int kTrials = 8;
  static int trial_shape(int sr) { return sr < 4 ? sr : 7 - sr; }
  hipEvent_t tev_[2 * kTrials] = {};
  ~LouvainRun() {
    for (auto& e : tev_)
      if (e) (void)hipEventDestroy(e);
  }
  struct LaneChoice {
    int lanes = 32;
    bool trial = false;
    int trials = 0;  // sub-rounds that time a shape
    uint64_t key = 0;
  };
  // Round 0 is not timed: every node is then a singleton and each arc a
  // cluster of its own, while later rounds insert fewer, repeated keys, so
  // the shapes are compared on the first round that follows (round 0 runs 8
  // lanes per node).
  template <class View>
  LaneChoice choose_lanes(const std::vector<uint32_t>& cnt, const std::vector<uint64_t>& sa, int B, int round) {
    LaneChoice ch;
    if (c.tune.lv_group_lanes) {
      ch.lanes = c.tune.lv_group_lanes;
      return ch;
    }
    uint64_t nodes = 0, arcs = 0;
    for (int sr = 0; sr < B; ++sr) { nodes += cnt[B + sr]; arcs += sa[B + sr]; }
    int ln = 0, ld = 0;
    while ((uint64_t(1) << (ln + 1)) <= nodes) ++ln;
    while ((uint64_t(1) << (ld + 1)) <= arcs / std::max<uint64_t>(nodes, 1)) ++ld;
    ch.key = (uint64_t(uint8_t(View::kName[0])) << 32) | (uint64_t(ln) << 8) | uint64_t(ld);
    auto it = c.tune.lv_group_choice.find(ch.key);
    if (it != c.tune.lv_group_choice.end()) {
      ch.lanes = it->second;
      return ch;
    }
    if (round == 0 && prm.max_passes > 1) {
      // 8 lanes would take the class out of the fused launch
      ch.lanes = c.tune.lv_fused_classes ? 32 : 8;
      return ch;
    }
    ch.trials = B >= kTrials ? kTrials : (B >= 4 ? 4 : 0);
    ch.trial = ch.trials > 0;
    for (int i = 0; i < ch.trials && ch.trial; ++i) ch.trial = cnt[B + i] != 0;
    if (ch.trial)
      for (auto& e : tev_)
        if (!e) MKP_HIP(hipEventCreate(&e));
    return ch;
  }
  // after the round's synchronisation: keep the fastest shape per arc
  void finish_lanes(const LaneChoice& ch, const std::vector<uint64_t>& sa, int B) {
    if (!ch.trial) return;
    double ms_sum[4] = {0, 0, 0, 0}, arcs_sum[4] = {0, 0, 0, 0};
    for (int i = 0; i < ch.trials; ++i) {
      float ms = 0;
      MKP_HIP(hipEventElapsedTime(&ms, tev_[2 * i], tev_[2 * i + 1]));
      ms_sum[trial_shape(i)] += double(ms);
      arcs_sum[trial_shape(i)] += double(sa[B + i]);
    }
    // 32 lanes unless another shape is faster by more than 5%: shapes within
    // the timing noise of each other would otherwise be chosen at random
    // from run to run.  With fused launches 8 lanes are not chosen: the
    // class would leave the fused launch, which its own timing does not see.
    int best = 2;
    double best_t = ms_sum[2] / std::max(arcs_sum[2], 1.0);
    for (int j = c.tune.lv_fused_classes ? 1 : 0; j < 4; ++j) {
      const double t = ms_sum[j] / std::max(arcs_sum[j], 1.0);
      if (j != 2 && t < 0.95 * best_t) { best = j; best_t = t; }
    }
    c.tune.lv_group_choice[ch.key] = kLaneTrials[best];
  }
  int lanes_for(const LaneChoice& ch, int sr, hipEvent_t* ta, hipEvent_t* tb) {
    if (ch.trial && sr < ch.trials) {
      *ta = tev_[2 * sr];
      *tb = tev_[2 * sr + 1];
      return kLaneTrials[trial_shape(sr)];
    }
    *ta = *tb = nullptr;
    return ch.lanes;
  }

  // applies the decisions of sub-round sr (its list ranges of every class)
  void launch_apply(const uint32_t* list, const int32_t* dec, const std::vector<uint32_t>& off,
                    const std::vector<uint32_t>& cnt, int B, int sr, const double* nvol, int32_t* C, double* cvol,
                    int32_t* csize, uint32_t* moved, const uint32_t* nmember, uint32_t* gmoved) {
    LvRanges rg{};
    for (int cl = 0; cl < kLvClasses; ++cl) {
      rg.off[cl] = off[cl * B + sr];
      rg.cnt[cl] = cnt[cl * B + sr];
      rg.total += rg.cnt[cl];
    }
    if (!rg.total) return;
    louvain_apply_kernel<<<grid_for(rg.total, 256, 1024), 256, 0, c.stream>>>(list, dec, rg, nvol, C, cvol, csize,
                                                                            moved, nmember, gmoved);
    MKP_LAUNCH_CHECK();
  }

  template <class View, bool UNIT>
  static void launch_fused(const View& g, const LvArgs& a, const uint32_t* list, const LvFused& f, unsigned blocks,
                           int g1, hipStream_t s) {
    if (g1 == 16)
      louvain_move_fused_kernel<View, 16, UNIT><<<blocks, 256, 0, s>>>(g, a, list, f);
    else if (g1 == 32)
      louvain_move_fused_kernel<View, 32, UNIT><<<blocks, 256, 0, s>>>(g, a, list, f);
    else if (g1 == 64)
      louvain_move_fused_kernel<View, 64, UNIT><<<blocks, 256, 0, s>>>(g, a, list, f);
    else
      louvain_move_fused_kernel<View, 0, UNIT><<<blocks, 256, 0, s>>>(g, a, list, f);
  }

  template <class View>
  void launch_move(const View& g, const LvArgs& a, const uint32_t* list, const std::vector<uint32_t>& off,
                   const std::vector<uint32_t>& cnt, const std::vector<uint64_t>& sa,
                   const std::vector<uint64_t>& sb, int sr, int lanes = 32, hipEvent_t ta = nullptr,
                   hipEvent_t tb = nullptr) {
    const int B = a.buckets;
    const int s0 = 0 * B + sr, s1 = 1 * B + sr, s1w = 2 * B + sr, s2 = 3 * B + sr, s2b = 4 * B + sr,
              s3 = kLvHeavy * B + sr;
    // each launch writes the decisions of its own list range
    auto at = [&](int slot) {
      LvArgs x = a;
      x.dec = a.dec + off[slot];
      if (a.gain_out) x.gain_out = a.gain_out + off[slot];
      return x;
    };
    // The classes decide on the same state and write disjoint decision
    // ranges.  With fusion on, the classes up to 1024 arcs share one launch
    // (one launch of a class rarely fills the chip, and their tails
    // overlap).  The 17..64-arc class keeps its own launch while the
    // autotuner times its shapes, with 8 lanes per node (128-thread blocks),
    // and when it has more than kLvFuse1Arcs arcs: the fused kernel runs at
    // the occupancy of its heaviest body (the registers of the <= 16-arc
    // nodes), which a large class of group maps would pay for.
    const hipStream_t s = c.stream;
    const int G = lanes;
    const int mode = c.tune.lv_fused_classes;
    const bool fused = mode != 0;
    const bool join1 = mode == 3 || (mode == 1 && sa[s1] <= kLvFuse1Arcs);
    const bool own1 = cnt[s1] && (!fused || ta || G == 8 || !join1);
    if (fused && (cnt[s0] || cnt[s1w] || cnt[s2] || (cnt[s1] && !own1))) {
      LvFused f{};
      const int slot[4] = {s2, s1w, s1, s0};
      const unsigned per1 = G == 16 ? 16 : (G == 64 ? 4 : 8);
      const unsigned per[4] = {1, 8, per1, 256};
      double bytes = 0.0;
      unsigned blocks = 0;
      for (int k = 0; k < 4; ++k) {
        if (k == 2 && own1) continue;
        f.off[k] = off[slot[k]];
        f.cnt[k] = cnt[slot[k]];
        f.blocks[k] = f.cnt[k] ? xcd_grid(f.cnt[k], per[k], 65536) : 0;
        blocks += f.blocks[k];
        bytes += double(sb[slot[k]]);
      }
      const int g1 = (cnt[s1] && !own1) ? G : 0;
      const bool unit = g.unit_arcs();
      TimedRegion tr(c.prof, s, kFamLouvainMove, bytes,
                     std::string("louvain_move_fused_kernel<") + View::kName + ", " + std::to_string(g1) +
                         (unit ? ", true>" : ", false>"));
      if (unit)
        launch_fused<View, View::kHasUnit>(g, a, list, f, blocks, g1, s);
      else
        launch_fused<View, false>(g, a, list, f, blocks, g1, s);
      MKP_LAUNCH_CHECK();
    }
    if (!fused && cnt[s0]) {
      const bool unit = g.unit_arcs();
      TimedRegion tr(c.prof, s, kFamLouvainMove, double(sb[s0]),
                     std::string("louvain_move_small_kernel<") + View::kName + (unit ? ", true>" : ", false>"));
      if (unit)
        louvain_move_small_kernel<View, View::kHasUnit><<<xcd_grid(cnt[s0], 256, 65536), 256, 0, s>>>(
            g, at(s0), list + off[s0], cnt[s0]);
      else
        louvain_move_small_kernel<View, false><<<xcd_grid(cnt[s0], 256, 65536), 256, 0, s>>>(
            g, at(s0), list + off[s0], cnt[s0]);
    }
    if (own1) {  // 17..64 arcs: 128-slot maps, lanes per node autotuned
      if (ta) MKP_HIP(hipEventRecord(ta, s));
      TimedRegion tr(c.prof, s, kFamLouvainMove, double(sb[s1]),
                     std::string("louvain_move_group_kernel<") + View::kName + ", " + std::to_string(G) + ", 7, " +
                         (G == 8 ? "128>" : "256>"));
      if (G == 8)
        louvain_move_group_kernel<View, 8, 7, 128><<<xcd_grid(cnt[s1], 16, 65536), 128, 0, s>>>(g, at(s1), list + off[s1],
                                                                                                cnt[s1]);
      else if (G == 16)
        louvain_move_group_kernel<View, 16, 7, 256><<<xcd_grid(cnt[s1], 16, 65536), 256, 0, s>>>(g, at(s1), list + off[s1],
                                                                                                 cnt[s1]);
      else if (G == 64)
        louvain_move_group_kernel<View, 64, 7, 256><<<xcd_grid(cnt[s1], 4, 65536), 256, 0, s>>>(g, at(s1), list + off[s1],
                                                                                                cnt[s1]);
      else
        louvain_move_group_kernel<View, 32, 7, 256><<<xcd_grid(cnt[s1], 8, 65536), 256, 0, s>>>(g, at(s1), list + off[s1],
                                                                                                cnt[s1]);
      if (tb) MKP_HIP(hipEventRecord(tb, s));
    }
    if (!fused && cnt[s1w]) {  // 65..128 arcs: 256-slot maps
      TimedRegion tr(c.prof, s, kFamLouvainMove, double(sb[s1w]),
                     std::string("louvain_move_group_kernel<") + View::kName + ", 32, 8, 256>");
      louvain_move_group_kernel<View, 32, 8, 256><<<xcd_grid(cnt[s1w], 8, 65536), 256, 0, s>>>(g, at(s1w), list + off[s1w], cnt[s1w]);
    }
    if (!fused && cnt[s2]) {  // 129..1024 arcs: a block per node, 2048-slot map
      TimedRegion tr(c.prof, s, kFamLouvainMove, double(sb[s2]),
                     std::string("louvain_move_block_kernel<") + View::kName + ", 11, 256>");
      louvain_move_block_kernel<View, 11, 256><<<xcd_grid(cnt[s2], 1, 65536), 256, 0, s>>>(g, at(s2), list + off[s2],
                                                                                           cnt[s2]);
    }
    if (cnt[s2b]) {  // 1025..kLvBigArcs arcs: a 1024-thread block per node, 8192-slot map
      TimedRegion tr(c.prof, s, kFamLouvainMove, double(sb[s2b]),
                     std::string("louvain_move_block_kernel<") + View::kName + ", 13, 1024>");
      louvain_move_block_kernel<View, 13, 1024><<<xcd_grid(cnt[s2b], 1, 65536), 1024, 0, s>>>(
          g, at(s2b), list + off[s2b], cnt[s2b]);
    }
    if (cnt[s3])
      launch_heavy(g, at(s3), list + off[s3], off[s3] - off[kLvHeavy * B], cnt[s3], sa[s3], double(sb[s3]), s);
    MKP_LAUNCH_CHECK();
  }

  // one level of local moving; returns whether any node moved
  template <class View>
  bool local_moving(const View& g, uint32_t n, uint64_t num_arcs, const double* nvol, double tv, uint32_t level,
                    int32_t* C) {
    const hipStream_t s = c.stream;
    const int B = std::max(1, prm.sub_rounds);
    const int slots = kLvClasses * B;
    cvol_.alloc(std::max<uint32_t>(n, 1));
    DevBuf<double>& cvol = cvol_;
    DevBuf<int32_t> csize(std::max<uint32_t>(n, 1));
    DevBuf<uint32_t> list(std::max<uint32_t>(n, 1));
    DevBuf<int32_t> dec(std::max<uint32_t>(n, 1));  // decision per list position
    // per-(slot, block) node counts of the work lists and their scan
    const unsigned lblocks = lv_list_blocks(n);
    DevBuf<uint32_t> bmat(size_t(slots) * lblocks);
    DevBuf<uint64_t> amat(size_t(slots) * lblocks), bymat(c.prof.enabled ? size_t(slots) * lblocks : 1);
    // the round's counters in one buffer, cleared by one launch and read back
    // by one copy: [sarcs u64][sbytes u64][counts u32][unused u32][two unused
    // u32][moved u32 of even rounds, of odd rounds].  The moves of
    // round r are read back with the counts of round r + 1, which are
    // computed before the loop condition is known and dropped if it fails.
    const size_t cw = (size_t(slots) + 1) / 2;  // words of slots u32 entries
    const size_t rwords = 2 * size_t(slots) + 2 * cw + 2;
    DevBuf<uint64_t> rb(rwords);
    uint64_t* sarcs = rb.p;
    uint64_t* sbytes = rb.p + slots;
    uint32_t* counts = reinterpret_cast<uint32_t*>(rb.p + 2 * slots);
    uint32_t* moved_rounds = reinterpret_cast<uint32_t*>(rb.p + 2 * slots + 2 * cw) + 2;
    std::vector<uint64_t> sb(slots, 0), sa(slots, 0);
    lv_init_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(n, nvol, C, cvol.p, csize.p);
    MKP_LAUNCH_CHECK();
    if (num_arcs == 0) {
      c.sync();
      return false;
    }
    LvArgs a;
    a.seed = prm.seed; a.level = level; a.sampling = uint32_t(std::min<uint64_t>(prm.sampling, 0xffffffffu));
    // a level with nodes above kLvBigArcs arcs sends every node above 1024
    // arcs to the arena path: one set of launches per sub-round, not two
    a.big_arcs = std::min<uint32_t>(max_degree_, a.sampling) > kLvBigArcs ? 1024u : kLvBigArcs;
    a.recip = 1.0 / tv; a.C = C; a.cvol = cvol.p; a.csize = csize.p; a.nvol = nvol;
    a.dec = dec.p;
    a.buckets = B;
    bool changed = false;
    uint64_t moved = n;
    std::vector<uint32_t> cnt(slots), off(slots);
    // pinned staging: the device buffer up to and including the moved
    // counters
    uint64_t* stg = c.staging<uint64_t>(rwords);
    uint32_t* stg_cnt = reinterpret_cast<uint32_t*>(stg + 2 * slots);
    const uint32_t* stg_moved = reinterpret_cast<const uint32_t*>(stg + 2 * slots + 2 * cw) + 2;
    // group round r's nodes by (degree class, sub-round bucket), clear its
    // counters and read the buffer back (the moved counter of round r - 1
    // with it); the moved counter of round r itself is cleared here too
    auto count_round = [&](int r) {
      a.round = uint32_t(r);
      a.bucket_salt = bucket_salt(a.seed, kLvBucketTag | (uint64_t(level) << 8) | uint64_t(r));
      fill_async<uint64_t>(rb.p, 0ull, rwords - 1, s);
      fill_async<uint32_t>(moved_rounds + (r & 1), 0u, 1, s);
      uint64_t* by = c.prof.enabled ? bymat.p : nullptr;
      lv_count_kernel<View><<<lblocks, 256, 0, s>>>(g, a, n, bmat.p, amat.p, by);
      lv_totals_kernel<<<slots, 256, 0, s>>>(lblocks, bmat.p, amat.p, by, counts, sarcs, by ? sbytes : nullptr);
      MKP_LAUNCH_CHECK();
      MKP_HIP(hipMemcpyAsync(stg, rb.p, rwords * 8, hipMemcpyDeviceToHost, s));
      c.sync();
    };
    count_round(0);
    for (int r = 0; double(moved) >= prm.min_move_fraction * double(n) && r < prm.max_passes; ++r) {
      const bool prof = c.prof.enabled;
      std::copy(stg_cnt, stg_cnt + slots, cnt.begin());
      std::copy(stg, stg + slots, sa.begin());
      if (prof) std::copy(stg + slots, stg + 2 * slots, sb.begin());
      uint64_t harcs = 0;
      uint32_t hnodes = 0;
      for (int sr = 0; sr < B; ++sr) {
        harcs = std::max(harcs, sa[kLvHeavy * B + sr]);
        hnodes += cnt[kLvHeavy * B + sr];
      }
      if (hnodes) reserve_heavy(harcs, hnodes);
      uint32_t run = 0;
      for (int i = 0; i < slots; ++i) { off[i] = run; run += cnt[i]; }
      lv_scatter_kernel<View><<<lblocks, 256, 0, s>>>(g, a, n, bmat.p, counts, list.p);
      MKP_LAUNCH_CHECK();
      if (hnodes) prepare_heavy(g, a, list.p + off[kLvHeavy * B], hnodes);
      const LaneChoice lanes = choose_lanes<View>(cnt, sa, B, r);
      for (int sr = 0; sr < B; ++sr) {
        a.bucket = sr;
        hipEvent_t ta, tb;
        const int lanes_now = lanes_for(lanes, sr, &ta, &tb);
        launch_move(g, a, list.p, off, cnt, sa, sb, sr, lanes_now, ta, tb);
        launch_apply(list.p, dec.p, off, cnt, B, sr, nvol, C, cvol.p, csize.p, moved_rounds + (r & 1), nullptr,
                     nullptr);
      }
      count_round(r + 1);
      finish_lanes(lanes, sa, B);
      moved = stg_moved[r & 1];
      st.arcs_swept += int64_t(num_arcs);
      st.rounds += 1;
      st.moves += int64_t(moved);
      changed |= moved > 0;
    }
    return changed;
  }

  // One evaluation of computeMaxGainCluster for every node of the level-0
  // graph g on the clustering C (host, n entries): per node the decision of
  // the move kernels (target or -1) and the best gain in FP64, and
  // metrics::modularity of C.  Nothing moves.
  template <class View>
  void evaluate(const View& g, uint32_t n, const int32_t* C_host, int32_t* target, double* gain, double* q) {
    const hipStream_t s = c.stream;
    DevBuf<double> vol(std::max<uint32_t>(n, 1)), cvol(std::max<uint32_t>(n, 1));
    DevBuf<int32_t> C(std::max<uint32_t>(n, 1)), csize(std::max<uint32_t>(n, 1));
    prepare(g, n, vol.p);
    const double tv = total_volume(vol.p, n);
    MKP_HIP(hipMemcpyAsync(C.p, C_host, size_t(n) * 4, hipMemcpyHostToDevice, s));
    fill_async<double>(cvol.p, 0.0, n, s);
    fill_async<int32_t>(csize.p, 0, n, s);
    cluster_state_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(C.p, vol.p, n, cvol.p, csize.p);
    MKP_LAUNCH_CHECK();
    if (q) {
      DevBuf<double> internal(std::max<uint32_t>(n, 1)), total(std::max<uint32_t>(n, 1)), part(64);
      fill_async<double>(internal.p, 0.0, n, s);
      fill_async<double>(total.p, 0.0, n, s);
      modularity_volume_kernel<View><<<grid_for(n, 256, 65536), 256, 0, s>>>(g, C.p, vol.p, n, internal.p, total.p);
      modularity_sum_kernel<<<64, 256, 0, s>>>(internal.p, total.p, n, tv, part.p);
      MKP_LAUNCH_CHECK();
      double* h = c.staging<double>(64);
      MKP_HIP(hipMemcpyAsync(h, part.p, 64 * 8, hipMemcpyDeviceToHost, s));
      c.sync();
      double sum = 0.0;
      for (int i = 0; i < 64; ++i) sum += h[i];
      *q = sum / tv;
    }
    if (!target && !gain) return;
    // every node in one sub-round
    const int B = 1, slots = kLvClasses;
    const unsigned lblocks = lv_list_blocks(n);
    DevBuf<uint32_t> list(std::max<uint32_t>(n, 1)), bmat(size_t(slots) * lblocks), counts(slots);
    DevBuf<uint64_t> amat(size_t(slots) * lblocks), sarcs(slots);
    DevBuf<int32_t> dec(std::max<uint32_t>(n, 1));
    DevBuf<double> dg(std::max<uint32_t>(n, 1));
    LvArgs a;
    a.seed = prm.seed; a.level = 0; a.round = 0; a.buckets = B; a.bucket = 0; a.bucket_salt = 0;
    a.sampling = uint32_t(std::min<uint64_t>(prm.sampling, 0xffffffffu));
    a.recip = 1.0 / tv; a.C = C.p; a.cvol = cvol.p; a.csize = csize.p; a.nvol = vol.p;
    a.dec = dec.p; a.gain_out = dg.p;
    lv_count_kernel<View><<<lblocks, 256, 0, s>>>(g, a, n, bmat.p, amat.p, nullptr);
    lv_totals_kernel<<<slots, 256, 0, s>>>(lblocks, bmat.p, amat.p, nullptr, counts.p, sarcs.p, nullptr);
    lv_scatter_kernel<View><<<lblocks, 256, 0, s>>>(g, a, n, bmat.p, counts.p, list.p);
    MKP_LAUNCH_CHECK();
    std::vector<uint32_t> cnt(slots), off(slots);
    std::vector<uint64_t> sa(slots), sb(slots, 0);
    MKP_HIP(hipMemcpyAsync(cnt.data(), counts.p, slots * 4, hipMemcpyDeviceToHost, s));
    MKP_HIP(hipMemcpyAsync(sa.data(), sarcs.p, slots * 8, hipMemcpyDeviceToHost, s));
    c.sync();
    uint32_t run = 0;
    for (int i = 0; i < slots; ++i) { off[i] = run; run += cnt[i]; }
    if (cnt[kLvHeavy * B]) {
      reserve_heavy(sa[kLvHeavy * B], cnt[kLvHeavy * B]);
      prepare_heavy(g, a, list.p + off[kLvHeavy * B], cnt[kLvHeavy * B]);
    }
    launch_move(g, a, list.p, off, cnt, sa, sb, 0, c.tune.lv_group_lanes ? c.tune.lv_group_lanes : 8);
    std::vector<uint32_t> hl(n);
    std::vector<int32_t> hd(n);
    std::vector<double> hg(n);
    MKP_HIP(hipMemcpyAsync(hl.data(), list.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
    MKP_HIP(hipMemcpyAsync(hd.data(), dec.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
    MKP_HIP(hipMemcpyAsync(hg.data(), dg.p, size_t(n) * 8, hipMemcpyDeviceToHost, s));
    c.sync();
    for (uint32_t i = 0; i < run; ++i) {
      if (target) target[hl[i]] = hd[i];
      if (gain) gain[hl[i]] = hg[i];
    }
  }

  uint32_t max_degree_ = 0;
  DevBuf<double> cvol_;  // cluster volumes left by the last local-moving level

  // level 0: node volumes from the arcs; every level: maximum degree
  template <class View>
  void prepare(const View& g, uint32_t n, double* nvol) {
    DevBuf<uint32_t> dmax(1);
    fill_async<uint32_t>(dmax.p, 0u, 1, c.stream);
    if (nvol)
      node_volume_kernel<View><<<grid_for(n, 256, 256), 256, 0, c.stream>>>(g, n, nvol, dmax.p);
    else
      max_degree_kernel<View><<<grid_for(n, 256, 256), 256, 0, c.stream>>>(g, n, dmax.p);
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipMemcpyAsync(&max_degree_, dmax.p, 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  }

  // Graph::contract; C is remapped to coarse ids in place
  template <class View>
  ExplicitGraph contract(const View& g, uint32_t n, uint64_t num_arcs, const double* nvol, int32_t* C) {
    const hipStream_t s = c.stream;
    TimedRegion tr(c.prof, s, kFamLouvainContract, (12.0 + 4.0 + 12.0) * double(num_arcs) + 20.0 * n);
    ExplicitGraph cg;
    DevBuf<uint32_t> used(std::max<uint32_t>(n, 1)), map(size_t(n) + 1);
    fill_async<uint32_t>(used.p, 0u, n, s);
    mark_used_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(C, n, used.p);
    MKP_LAUNCH_CHECK();
    exclusive_scan_u32(c, used.p, map.p, n);
    uint32_t nc = 0;
    MKP_HIP(hipMemcpyAsync(&nc, map.p + n, 4, hipMemcpyDeviceToHost, s));
    c.sync();
    cg.n = nc;
    cg.vol.alloc(std::max<uint32_t>(nc, 1));
    // coarse volume = tracked volume of the cluster (the sum of its members)
    coarse_volume_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(used.p, map.p, cvol_.p, n, cg.vol.p);
    remap_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(C, n, map.p);
    MKP_LAUNCH_CHECK();
    coarse_arcs(g, num_arcs, C, cg);
    return cg;
  }

  // Coarse CSR of a clustering whose coarse ids (cg.n of them) are already in
  // C: arcs keyed (C[u], C[v]), self loops dropped, parallel arcs summed.  In
  // a batch, the arcs of members that end at this level are dropped as well.
  template <class View>
  void coarse_arcs(const View& g, uint64_t num_arcs, const int32_t* C, ExplicitGraph& cg,
                   const uint32_t* nmember = nullptr, const uint8_t* gkeep = nullptr) {
    const hipStream_t s = c.stream;
    const uint32_t nc = cg.n;
    int bits = 1;
    while ((uint64_t(1) << bits) < nc) ++bits;
    MKP_CHECK(2 * bits + 1 <= 64, "graph too large for 64-bit arc keys");
    cg.idx.alloc(size_t(nc) + 1);
    if (num_arcs == 0 || nc <= 1) {
      fill_async<uint64_t>(cg.idx.p, 0, size_t(nc) + 1, s);
      cg.head.alloc(1);
      cg.src.alloc(1);
      cg.w.alloc(1);
      cg.arcs = 0;
      c.sync();
      return;
    }
    DevBuf<uint64_t> keys(num_arcs), skeys(num_arcs), ukeys(num_arcs);
    DevBuf<double> vals(num_arcs), svals(num_arcs), uvals(num_arcs);
    // arcs inside one cluster (self loops of the coarse graph) are dropped
    // before the sort by a stable compaction into ukeys / uvals, which are
    // free until the reduction
    const bool compact = num_arcs < 0xffffffffull;
    DevBuf<uint32_t> keep(compact ? num_arcs : 1), pos(compact ? num_arcs + 1 : 1);
    emit_arcs_kernel<View><<<grid_for(num_arcs, 256, 65536), 256, 0, s>>>(g, num_arcs, C, bits, keys.p, vals.p,
                                                                           nmember, gkeep,
                                                                           compact ? keep.p : nullptr);
    MKP_LAUNCH_CHECK();
    uint64_t ns = num_arcs;
    const uint64_t* sk = keys.p;
    const double* sv = vals.p;
    if (compact) {
      exclusive_scan_u32(c, keep.p, pos.p, num_arcs);
      compact_arcs_kernel<<<grid_for(num_arcs, 256, 65536), 256, 0, s>>>(keys.p, vals.p, keep.p, pos.p, num_arcs,
                                                                          ukeys.p, uvals.p);
      MKP_LAUNCH_CHECK();
      uint32_t kept = 0;
      MKP_HIP(hipMemcpyAsync(&kept, pos.p + num_arcs, 4, hipMemcpyDeviceToHost, s));
      c.sync();
      ns = kept;
      sk = ukeys.p;
      sv = uvals.p;
    }
    DevBuf<uint64_t> ucount(1);
    fill_async<uint64_t>(ucount.p, 0ull, 1, s);
    if (ns) {
      size_t bytes = 0;
      MKP_HIP(rocprim::radix_sort_pairs(nullptr, bytes, sk, skeys.p, sv, svals.p, ns, 0u, unsigned(2 * bits + 1), s));
      MKP_HIP(rocprim::radix_sort_pairs(c.temp(bytes), bytes, sk, skeys.p, sv, svals.p, ns, 0u,
                                        unsigned(2 * bits + 1), s));
      bytes = 0;
      MKP_HIP(rocprim::reduce_by_key(nullptr, bytes, skeys.p, svals.p, ns, ukeys.p, uvals.p, ucount.p,
                                     rocprim::plus<double>(), rocprim::equal_to<uint64_t>(), s));
      MKP_HIP(rocprim::reduce_by_key(c.temp(bytes), bytes, skeys.p, svals.p, ns, ukeys.p, uvals.p, ucount.p,
                                     rocprim::plus<double>(), rocprim::equal_to<uint64_t>(), s));
    }
    uint64_t U = 0, last = 0;
    MKP_HIP(hipMemcpyAsync(&U, ucount.p, 8, hipMemcpyDeviceToHost, s));
    c.sync();
    if (U && !compact) {  // the compaction already dropped the self loops
      MKP_HIP(hipMemcpyAsync(&last, ukeys.p + (U - 1), 8, hipMemcpyDeviceToHost, s));
      c.sync();
      if (last == (uint64_t(1) << (2 * bits))) --U;  // self-loop sentinel segment
    }
    cg.arcs = U;
    cg.head.alloc(std::max<uint64_t>(U, 1));
    cg.src.alloc(std::max<uint64_t>(U, 1));
    cg.w.alloc(std::max<uint64_t>(U, 1));
    // idx[c + 1] = end of the arcs of source c; sources without arcs carry the
    // previous end forward (inclusive max scan)
    DevBuf<uint64_t> ends(std::max<uint32_t>(nc, 1));
    fill_async<uint64_t>(ends.p, 0, nc, s);
    if (U) {
      coarse_csr_kernel<<<grid_for(U, 256, 65536), 256, 0, s>>>(ukeys.p, U, bits, cg.head.p, cg.src.p, ends.p);
      MKP_LAUNCH_CHECK();
      MKP_HIP(hipMemcpyAsync(cg.w.p, uvals.p, U * 8, hipMemcpyDeviceToDevice, s));
    }
    fill_async<uint64_t>(cg.idx.p, 0, 1, s);
    inclusive_max_scan_u64(c, ends.p, cg.idx.p + 1, nc);
    // the temporaries go back to the context's cache in stream order: no wait
  }

  // local_moving_contract_recurse (parallel_louvain.cpp:27-60), iterative
  template <class View>
  void run_from(const View& g0, uint32_t n0, uint64_t arcs0, int32_t* C0) {
    const hipStream_t s = c.stream;
    DevBuf<double> vol0(std::max<uint32_t>(n0, 1));
    prepare(g0, n0, vol0.p);
    const double tv = total_volume(vol0.p, n0);
    std::vector<ExplicitGraph> levels;
    std::vector<DevBuf<int32_t>> cl;  // clustering of each contracted level
    bool changed = local_moving(g0, n0, arcs0, vol0.p, tv, 0, C0);
    st.levels = 1;
    if (changed) {
      levels.push_back(contract(g0, n0, arcs0, vol0.p, C0));
      while (true) {
        ExplicitGraph& g = levels.back();
        cl.emplace_back(std::max<uint32_t>(g.n, 1));
        int32_t* C = cl.back().p;
        prepare(g.view(), g.n, nullptr);  // volumes are the summed fine volumes of the contraction
        bool ch = local_moving(g.view(), g.n, g.arcs, g.vol.p, tv, uint32_t(levels.size()), C);
        st.levels += 1;
        if (!ch) break;
        levels.push_back(contract(g.view(), g.n, g.arcs, g.vol.p, C));
      }
      // prolong from the deepest clustering back to level 0
      for (int i = int(cl.size()) - 1; i >= 1; --i)
        prolong_kernel<<<grid_for(levels[i - 1].n, 256, 65536), 256, 0, s>>>(cl[i - 1].p, levels[i - 1].n, cl[i].p);
      prolong_kernel<<<grid_for(n0, 256, 65536), 256, 0, s>>>(C0, n0, cl[0].p);
      MKP_LAUNCH_CHECK();
    }
    c.sync();
  }
};

// ------------------------------------------------------------ batch form --
// Many member hypergraphs as one disjoint union.  Level 0 (the star
// expansion of the union) numbers the vertices of all members first, then
// the net nodes of all members in rank order; a member's nets keep their
// relative ranks because its locality keys lie in its own vertex range.
// Every node knows its member and its member-local id (the id it has when
// the member runs alone); contraction numbers the coarse nodes member by
// member in local order, so from level 1 on every member is one contiguous
// id range and its coarse graph is the one it builds alone.
__global__ void batch_layout0_kernel(uint32_t N, uint32_t n, const uint32_t* __restrict__ vgraph,
                                     const uint32_t* __restrict__ ngraph, const uint32_t* __restrict__ vbase,
                                     const uint32_t* __restrict__ nbase, uint32_t* __restrict__ nmember,
                                     uint32_t* __restrict__ nlocal) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    if (u < N) {
      const uint32_t g = vgraph[u];
      nmember[u] = g;
      nlocal[u] = u - vbase[g];
    } else {
      const uint32_t r = u - N;
      const uint32_t g = ngraph[r];
      nmember[u] = g;
      nlocal[u] = (vbase[g + 1] - vbase[g]) + (r - nbase[g]);
    }
  }
}

// contiguous members: member of node u = largest g with cb[g] <= u
__global__ void batch_layout_kernel(const uint32_t* __restrict__ cb, uint32_t G, uint32_t n,
                                    uint32_t* __restrict__ nmember, uint32_t* __restrict__ nlocal) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (cb[mid] <= u) lo = mid; else hi = mid;
    }
    nmember[u] = lo;
    nlocal[u] = u - cb[lo];
  }
}

// member-major position of a node: mbase[member] + local id
template <typename T>
__global__ void batch_to_positions_kernel(const T* __restrict__ in, uint32_t n, const uint32_t* __restrict__ nmember,
                                          const uint32_t* __restrict__ nlocal, const uint32_t* __restrict__ mbase,
                                          T* __restrict__ out) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    out[mbase[nmember[u]] + nlocal[u]] = in[u];
}

__global__ void batch_from_positions_kernel(const uint32_t* __restrict__ in, uint32_t n,
                                            const uint32_t* __restrict__ nmember, const uint32_t* __restrict__ nlocal,
                                            const uint32_t* __restrict__ mbase, uint32_t* __restrict__ out) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    out[u] = in[mbase[nmember[u]] + nlocal[u]];
}

__global__ void batch_mark_used_kernel(const int32_t* __restrict__ C, uint32_t n, const uint32_t* __restrict__ nmember,
                                       const uint8_t* __restrict__ keep, uint32_t* __restrict__ used) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    if (keep[nmember[u]]) used[C[u]] = 1;
}

__global__ void batch_remap_kernel(int32_t* __restrict__ C, uint32_t n, const uint32_t* __restrict__ map,
                                   const uint32_t* __restrict__ nmember, const uint8_t* __restrict__ keep) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    if (keep[nmember[u]]) C[u] = int32_t(map[C[u]]);
}

// out[g] = in[idx[g]], g < count
template <typename T>
__global__ void gather_kernel(const T* __restrict__ in, const uint32_t* __restrict__ idx, uint32_t count,
                              T* __restrict__ out) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) out[g] = in[idx[g]];
}

// Total volume of every member with the shape of sum_volume_kernel over the
// member's own nodes in local order (64 blocks of 256 threads, then the 64
// partials summed in order on the host), so it is the member's value alone.
__global__ void __launch_bounds__(256)
batch_member_volume_kernel(const double* __restrict__ volp, const uint32_t* __restrict__ mbase, uint32_t G,
                           double* __restrict__ out) {
  __shared__ double part[256];
  for (uint32_t g = blockIdx.y; g < G; g += gridDim.y) {
    const uint32_t b = mbase[g], n = mbase[g + 1] - b;
    double s = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += volp[b + i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (int(threadIdx.x) < o) part[threadIdx.x] += part[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[size_t(g) * gridDim.x + blockIdx.x] = part[0];
    __syncthreads();
  }
}

// final community of every node of a level: the member's local id at the
// level where the member ended, else the community of its coarse node
__global__ void batch_final_kernel(int32_t* __restrict__ C, uint32_t n, const uint32_t* __restrict__ nmember,
                                   const uint32_t* __restrict__ nlocal, const int32_t* __restrict__ gend,
                                   int32_t level, const int32_t* __restrict__ coarse) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    C[u] = gend[nmember[u]] == level ? int32_t(nlocal[u]) : coarse[C[u]];
}

struct LouvainBatchRun : LouvainRun {
  const LouvainBatch& d;
  const uint32_t G;
  std::vector<mkp_louvain_stats> gst;
  DevBuf<uint8_t> gflag_;  // per member: active in the round / kept for the next level
  DevBuf<uint32_t> gmoved_;
  DevBuf<double> grecip_;
  DevBuf<uint64_t> gseed_;

  struct Level {
    uint32_t n = 0;
    uint64_t arcs = 0;
    bool member_major = false;  // node ids are member-major positions
    DevBuf<uint32_t> nmember, nlocal, mbase;
    std::vector<uint32_t> nodes;  // per member
    std::vector<uint64_t> marcs;  // per member
  };

  LouvainBatchRun(Context& ctx, const mkp_louvain_params& p, const LouvainBatch& desc)
      : LouvainRun(ctx, p), d(desc), G(desc.members), gst(desc.members) {
    gflag_.alloc(G);
    gmoved_.alloc(G);
    grecip_.alloc(G);
    gseed_.alloc(G);
    uint64_t* stg = c.staging<uint64_t>(G);
    std::copy(d.seeds, d.seeds + G, stg);
    MKP_HIP(hipMemcpyAsync(gseed_.p, stg, size_t(G) * 8, hipMemcpyHostToDevice, c.stream));
    c.sync();
  }

  // uploads one flag per member through the pinned staging
  void upload_flags(const std::vector<uint8_t>& f) {
    uint8_t* stg = c.staging<uint8_t>(G);
    std::copy(f.begin(), f.end(), stg);
    MKP_HIP(hipMemcpyAsync(gflag_.p, stg, G, hipMemcpyHostToDevice, c.stream));
    c.sync();
  }

  template <class View>
  void local_moving_batch(const View& g, const Level& L, const double* nvol, uint32_t level, int32_t* C,
                          const std::vector<uint8_t>& alive, std::vector<uint8_t>& changed) {
    const hipStream_t s = c.stream;
    const uint32_t n = L.n;
    const int B = std::max(1, prm.sub_rounds);
    const int slots = kLvClasses * B;
    changed.assign(G, 0);
    cvol_.alloc(std::max<uint32_t>(n, 1));
    DevBuf<int32_t> csize(std::max<uint32_t>(n, 1));
    DevBuf<uint32_t> list(std::max<uint32_t>(n, 1));
    DevBuf<int32_t> dec(std::max<uint32_t>(n, 1));  // decision per list position
    const unsigned lblocks = lv_list_blocks(n);
    DevBuf<uint32_t> bmat(size_t(slots) * lblocks);
    DevBuf<uint64_t> amat(size_t(slots) * lblocks), bymat(c.prof.enabled ? size_t(slots) * lblocks : 1);
    // the round's counters in one buffer, laid out as in local_moving
    const size_t cw = (size_t(slots) + 1) / 2;
    const size_t rwords = 2 * size_t(slots) + 2 * cw + 2;
    DevBuf<uint64_t> rb(rwords);
    uint64_t* sarcs = rb.p;
    uint64_t* sbytes = rb.p + slots;
    uint32_t* counts = reinterpret_cast<uint32_t*>(rb.p + 2 * slots);
    uint32_t* moved_total = reinterpret_cast<uint32_t*>(rb.p + 2 * slots + 2 * cw);
    std::vector<uint64_t> sb(slots, 0), sa(slots, 0);
    std::vector<uint32_t> cnt(slots), off(slots);
    if (n == 0) return;
    lv_init_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(n, nvol, C, cvol_.p, csize.p);
    MKP_LAUNCH_CHECK();
    LvArgs a;
    a.seed = 0; a.level = level; a.sampling = uint32_t(std::min<uint64_t>(prm.sampling, 0xffffffffu));
    a.recip = 0.0; a.C = C; a.cvol = cvol_.p; a.csize = csize.p; a.nvol = nvol;
    a.dec = dec.p;
    a.buckets = B; a.bucket_salt = 0;
    a.nmember = L.nmember.p; a.nlocal = L.nlocal.p; a.grecip = grecip_.p; a.gactive = gflag_.p; a.gseed = gseed_.p;
    std::vector<uint64_t> moved(L.nodes.begin(), L.nodes.end());
    std::vector<uint8_t> act(G);
    for (int r = 0;; ++r) {
      // the loop condition of local moving, per member
      bool any = false;
      for (uint32_t gm = 0; gm < G; ++gm) {
        act[gm] = alive[gm] && L.marcs[gm] > 0 && r < prm.max_passes &&
                  double(moved[gm]) >= prm.min_move_fraction * double(L.nodes[gm]);
        any |= act[gm] != 0;
      }
      if (!any) break;
      upload_flags(act);
      a.round = uint32_t(r);
      fill_async<uint64_t>(rb.p, 0ull, rwords, s);
      fill_async<uint32_t>(gmoved_.p, 0u, G, s);
      const bool prof = c.prof.enabled;
      uint64_t* by = prof ? bymat.p : nullptr;
      lv_count_kernel<View><<<lblocks, 256, 0, s>>>(g, a, n, bmat.p, amat.p, by);
      lv_totals_kernel<<<slots, 256, 0, s>>>(lblocks, bmat.p, amat.p, by, counts, sarcs, by ? sbytes : nullptr);
      MKP_LAUNCH_CHECK();
      uint64_t* stg = c.staging<uint64_t>(2 * size_t(slots) + 2 * cw);
      uint32_t* stg_cnt = reinterpret_cast<uint32_t*>(stg + 2 * slots);
      MKP_HIP(hipMemcpyAsync(stg, rb.p, (2 * size_t(slots) + cw) * 8, hipMemcpyDeviceToHost, s));
      c.sync();
      std::copy(stg_cnt, stg_cnt + slots, cnt.begin());
      std::copy(stg, stg + slots, sa.begin());
      if (prof) std::copy(stg + slots, stg + 2 * slots, sb.begin());
      uint64_t harcs = 0;
      uint32_t hnodes = 0;
      for (int sr = 0; sr < B; ++sr) {
        harcs = std::max(harcs, sa[kLvHeavy * B + sr]);
        hnodes += cnt[kLvHeavy * B + sr];
      }
      if (hnodes) reserve_heavy(harcs, hnodes);
      uint32_t run = 0;
      for (int i = 0; i < slots; ++i) { off[i] = run; run += cnt[i]; }
      lv_scatter_kernel<View><<<lblocks, 256, 0, s>>>(g, a, n, bmat.p, counts, list.p);
      MKP_LAUNCH_CHECK();
      if (hnodes) prepare_heavy(g, a, list.p + off[kLvHeavy * B], hnodes);
      const LaneChoice lanes = choose_lanes<View>(cnt, sa, B, r);
      for (int sr = 0; sr < B; ++sr) {
        a.bucket = sr;
        hipEvent_t ta, tb;
        const int lanes_now = lanes_for(lanes, sr, &ta, &tb);
        launch_move(g, a, list.p, off, cnt, sa, sb, sr, lanes_now, ta, tb);
        launch_apply(list.p, dec.p, off, cnt, B, sr, nvol, C, cvol_.p, csize.p, moved_total, L.nmember.p,
                     gmoved_.p);
      }
      uint32_t* stg_moved = c.staging<uint32_t>(G);
      MKP_HIP(hipMemcpyAsync(stg_moved, gmoved_.p, size_t(G) * 4, hipMemcpyDeviceToHost, s));
      c.sync();
      finish_lanes(lanes, sa, B);
      for (uint32_t gm = 0; gm < G; ++gm) {
        if (!act[gm]) continue;
        moved[gm] = stg_moved[gm];
        gst[gm].arcs_swept += int64_t(L.marcs[gm]);
        gst[gm].rounds += 1;
        gst[gm].moves += int64_t(moved[gm]);
        changed[gm] |= moved[gm] > 0;
      }
    }
  }

  // Graph::contract of the members in `keep` (the others end at this level);
  // coarse nodes are numbered member by member in local order.
  template <class View>
  void contract_batch(const View& g, const Level& L, int32_t* C, const std::vector<uint8_t>& keep, Level& out,
                      ExplicitGraph& cg) {
    const hipStream_t s = c.stream;
    const uint32_t n = L.n;
    TimedRegion tr(c.prof, s, kFamLouvainContract, (12.0 + 4.0 + 12.0) * double(L.arcs) + 20.0 * n);
    upload_flags(keep);
    DevBuf<uint32_t> used(std::max<uint32_t>(n, 1)), scan(size_t(n) + 1), map;
    fill_async<uint32_t>(used.p, 0u, n, s);
    batch_mark_used_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(C, n, L.nmember.p, gflag_.p, used.p);
    MKP_LAUNCH_CHECK();
    const uint32_t* mp = scan.p;
    if (L.member_major) {
      exclusive_scan_u32(c, used.p, scan.p, n);
    } else {
      DevBuf<uint32_t> upos(std::max<uint32_t>(n, 1));
      batch_to_positions_kernel<uint32_t><<<grid_for(n, 256, 65536), 256, 0, s>>>(used.p, n, L.nmember.p,
                                                                                 L.nlocal.p, L.mbase.p, upos.p);
      MKP_LAUNCH_CHECK();
      exclusive_scan_u32(c, upos.p, scan.p, n);
      map.alloc(std::max<uint32_t>(n, 1));
      batch_from_positions_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(scan.p, n, L.nmember.p, L.nlocal.p,
                                                                          L.mbase.p, map.p);
      MKP_LAUNCH_CHECK();
      mp = map.p;
    }
    // coarse member bases: the scan at every member's first position
    out.mbase.alloc(size_t(G) + 1);
    gather_kernel<uint32_t><<<grid_for(G + 1, 256, 4096), 256, 0, s>>>(scan.p, L.mbase.p, G + 1, out.mbase.p);
    MKP_LAUNCH_CHECK();
    uint32_t* stg = c.staging<uint32_t>(size_t(G) + 1);
    MKP_HIP(hipMemcpyAsync(stg, out.mbase.p, (size_t(G) + 1) * 4, hipMemcpyDeviceToHost, s));
    c.sync();
    std::vector<uint32_t> cb(stg, stg + G + 1);
    const uint32_t nc = cb[G];
    out.n = nc;
    out.member_major = true;
    out.nodes.resize(G);
    for (uint32_t gm = 0; gm < G; ++gm) out.nodes[gm] = cb[gm + 1] - cb[gm];
    cg.n = nc;
    cg.vol.alloc(std::max<uint32_t>(nc, 1));
    coarse_volume_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(used.p, mp, cvol_.p, n, cg.vol.p);
    batch_remap_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(C, n, mp, L.nmember.p, gflag_.p);
    MKP_LAUNCH_CHECK();
    out.nmember.alloc(std::max<uint32_t>(nc, 1));
    out.nlocal.alloc(std::max<uint32_t>(nc, 1));
    if (nc) {
      batch_layout_kernel<<<grid_for(nc, 256, 65536), 256, 0, s>>>(out.mbase.p, G, nc, out.nmember.p, out.nlocal.p);
      MKP_LAUNCH_CHECK();
    }
    coarse_arcs(g, L.arcs, C, cg, L.nmember.p, gflag_.p);
    out.arcs = cg.arcs;
    // arcs per member: idx at the member bases
    DevBuf<uint64_t> ab(size_t(G) + 1);
    gather_kernel<uint64_t><<<grid_for(G + 1, 256, 4096), 256, 0, s>>>(cg.idx.p, out.mbase.p, G + 1, ab.p);
    MKP_LAUNCH_CHECK();
    uint64_t* stg64 = c.staging<uint64_t>(size_t(G) + 1);
    MKP_HIP(hipMemcpyAsync(stg64, ab.p, (size_t(G) + 1) * 8, hipMemcpyDeviceToHost, s));
    c.sync();
    out.marcs.resize(G);
    for (uint32_t gm = 0; gm < G; ++gm) out.marcs[gm] = stg64[gm + 1] - stg64[gm];
  }

  // local_moving_contract_recurse of every member, all members level by
  // level in the same launches; C0 receives member-local communities
  template <class View>
  void run_batch(const View& g0, Level& L0, int32_t* C0) {
    const hipStream_t s = c.stream;
    const uint32_t n0 = L0.n;
    DevBuf<double> vol0(std::max<uint32_t>(n0, 1));
    prepare(g0, n0, vol0.p);
    {
      // total volume per member in its local node order
      const double* volp = vol0.p;
      DevBuf<double> vpos;
      if (!L0.member_major && n0) {
        vpos.alloc(n0);
        batch_to_positions_kernel<double><<<grid_for(n0, 256, 65536), 256, 0, s>>>(vol0.p, n0, L0.nmember.p,
                                                                                  L0.nlocal.p, L0.mbase.p, vpos.p);
        MKP_LAUNCH_CHECK();
        volp = vpos.p;
      }
      DevBuf<double> part(size_t(G) * 64);
      batch_member_volume_kernel<<<dim3(64, std::min<uint32_t>(G, 65535)), 256, 0, s>>>(volp, L0.mbase.p, G,
                                                                                        part.p);
      MKP_LAUNCH_CHECK();
      double* stg = c.staging<double>(size_t(G) * 65);  // partials, then 1 / volume
      double* rc = stg + size_t(G) * 64;
      MKP_HIP(hipMemcpyAsync(stg, part.p, size_t(G) * 64 * 8, hipMemcpyDeviceToHost, s));
      c.sync();
      for (uint32_t gm = 0; gm < G; ++gm) {
        double tv = 0.0;
        for (int i = 0; i < 64; ++i) tv += stg[size_t(gm) * 64 + i];
        rc[gm] = 1.0 / tv;
      }
      MKP_HIP(hipMemcpyAsync(grecip_.p, rc, size_t(G) * 8, hipMemcpyHostToDevice, s));
      c.sync();
    }
    std::vector<uint8_t> alive(G, 1), changed;
    std::vector<int32_t> gend(G, -1);
    // levels[i] / graphs[i - 1] / cl[i]: layout, graph and clustering of level i
    std::vector<std::unique_ptr<Level>> levels;
    std::vector<std::unique_ptr<ExplicitGraph>> graphs;
    std::vector<DevBuf<int32_t>> cl;
    uint32_t level = 0;
    local_moving_batch(g0, L0, vol0.p, 0, C0, alive, changed);
    auto settle = [&](uint32_t lv) {  // members without a move end at this level
      bool more = false;
      for (uint32_t gm = 0; gm < G; ++gm) {
        if (!alive[gm]) continue;
        gst[gm].levels += 1;
        if (!changed[gm]) {
          alive[gm] = 0;
          gend[gm] = int32_t(lv);
        } else {
          more = true;
        }
      }
      return more;
    };
    bool more = settle(0);
    while (more) {
      levels.emplace_back(new Level());
      graphs.emplace_back(new ExplicitGraph());
      if (level == 0) {
        contract_batch(g0, L0, C0, alive, *levels.back(), *graphs.back());
      } else {
        contract_batch(graphs[level - 1]->view(), *levels[level - 1], cl[level - 1].p, alive, *levels.back(),
                       *graphs.back());
      }
      ++level;
      ExplicitGraph& g = *graphs.back();
      cl.emplace_back(std::max<uint32_t>(g.n, 1));
      local_moving_batch(g.view(), *levels.back(), g.vol.p, level, cl.back().p, alive, changed);
      more = settle(level);
    }
    // communities from the deepest level back to level 0
    DevBuf<int32_t> dend(G);
    {
      int32_t* stg = c.staging<int32_t>(G);
      std::copy(gend.begin(), gend.end(), stg);
      MKP_HIP(hipMemcpyAsync(dend.p, stg, size_t(G) * 4, hipMemcpyHostToDevice, s));
    }
    for (int lv = int(level); lv >= 1; --lv) {
      const Level& L = *levels[lv - 1];
      if (L.n)
        batch_final_kernel<<<grid_for(L.n, 256, 65536), 256, 0, s>>>(
            cl[lv - 1].p, L.n, L.nmember.p, L.nlocal.p, dend.p, lv, lv == int(level) ? nullptr : cl[lv].p);
    }
    if (n0)
      batch_final_kernel<<<grid_for(n0, 256, 65536), 256, 0, s>>>(C0, n0, L0.nmember.p, L0.nlocal.p, dend.p, 0,
                                                                  level == 0 ? nullptr : cl[0].p);
    MKP_LAUNCH_CHECK();
    c.sync();
  }
};

void louvain_batch(mkp_hypergraph* hg, const mkp_louvain_params& prm, int wt, const LouvainBatch& d,
                   mkp_louvain_stats* stats) {
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t G = d.members;
  MKP_HIP(hipEventRecord(c.ev_a, s));
  LouvainBatchRun run(c, prm, d);
  LouvainBatchRun::Level L0;
  L0.nodes.resize(G);
  L0.marcs.resize(G);
  std::vector<uint32_t> mb(size_t(G) + 1);
  for (uint32_t g = 0; g <= G; ++g) mb[g] = d.plain ? d.vbase[g] : d.vbase[g] + d.nbase[g];
  for (uint32_t g = 0; g < G; ++g) {
    const uint64_t pins = d.pbase[g + 1] - d.pbase[g];
    L0.nodes[g] = mb[g + 1] - mb[g];
    L0.marcs[g] = d.plain ? pins : 2 * pins;
  }
  L0.n = mb[G];
  L0.arcs = d.plain ? hg->p : 2 * hg->p;
  L0.member_major = d.plain;
  L0.nmember.alloc(std::max<uint32_t>(L0.n, 1));
  L0.nlocal.alloc(std::max<uint32_t>(L0.n, 1));
  L0.mbase.alloc(size_t(G) + 1);
  {
    uint32_t* stg = c.staging<uint32_t>(size_t(G) + 1);
    std::copy(mb.begin(), mb.end(), stg);
    MKP_HIP(hipMemcpyAsync(L0.mbase.p, stg, (size_t(G) + 1) * 4, hipMemcpyHostToDevice, s));
    c.sync();
  }
  if (L0.n) {
    batch_layout0_kernel<<<grid_for(L0.n, 256, 65536), 256, 0, s>>>(d.plain ? L0.n : hg->n, L0.n, d.vgraph, d.ngraph,
                                                                    d.dvbase, d.dnbase, L0.nmember.p, L0.nlocal.p);
    MKP_LAUNCH_CHECK();
  }
  DevBuf<int32_t> C(std::max<uint32_t>(L0.n, 1));
  if (d.plain) {
    PlainView g{hg->n, hg->voff.p, hg->inc.p, hg->noff.p, hg->pins.p, hg->inc_vtx.p, hg->nw.p, wt,
                hg->unit_net_weights};
    run.run_batch(g, L0, C.p);
  } else {
    ensure_net_ranks(hg);
    StarView g{hg->n,         hg->m,         hg->p,    hg->voff.p, hg->inc.p,          hg->noff.p,
               hg->pins.p,    hg->inc_vtx.p, hg->pin_net.p, hg->nw.p, wt, hg->unit_net_weights,
               hg->net_order.p, hg->inc_rank.p, hg->rpoff.p, hg->rpins.p, hg->rpin_rank.p};
    run.run_batch(g, L0, C.p);
  }
  MKP_HIP(hipMemcpyAsync(hg->comm.p, C.p, size_t(hg->n) * 4, hipMemcpyDeviceToDevice, s));
  MKP_HIP(hipEventRecord(c.ev_b, s));
  c.sync();
  if (stats) {
    float ms = 0;
    MKP_HIP(hipEventElapsedTime(&ms, c.ev_a, c.ev_b));
    for (uint32_t g = 0; g < G; ++g) {
      stats[g] = run.gst[g];
      stats[g].device_ms = ms;
    }
  }
}

// Locality key of every net: the lower median of its first nine pins, packed
// above the net id so a radix sort yields the (key, id) order.
__global__ void net_key_kernel(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ pins, uint32_t m,
                               uint64_t* __restrict__ keys) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const uint32_t b = noff[e];
    const uint32_t s = min(noff[e + 1] - b, 9u);
    uint32_t p[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) p[i] = i < int(s) ? pins[b + i] : 0xffffffffu;
#pragma unroll
    for (int i = 1; i < 9; ++i)  // insertion sort of nine registers
#pragma unroll
      for (int j = i; j > 0; --j)
        if (p[j] < p[j - 1]) { const uint32_t t = p[j]; p[j] = p[j - 1]; p[j - 1] = t; }
    uint32_t key = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i)
      if (s > 0 && uint32_t(i) == (s - 1) / 2) key = p[i];
    keys[e] = (uint64_t(key) << 32) | e;
  }
}

__global__ void net_rank_kernel(const uint64_t* __restrict__ sorted, uint32_t m, uint32_t* __restrict__ order,
                                uint32_t* __restrict__ rank) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
    const uint32_t e = uint32_t(sorted[r]);
    order[r] = e;
    rank[e] = r;
  }
}

__global__ void inc_rank_kernel(const uint32_t* __restrict__ inc, const uint32_t* __restrict__ rank, uint64_t p,
                                uint32_t* __restrict__ out) {
  for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < p; j += uint64_t(gridDim.x) * blockDim.x)
    out[j] = rank[inc[j]];
}

__global__ void ranked_size_kernel(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ order, uint32_t m,
                                   uint32_t* __restrict__ size) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x)
    size[r] = noff[order[r] + 1] - noff[order[r]];
}

// pin slot j of net e moves to slot rpoff[rank(e)] + (j - noff[e])
__global__ void ranked_pins_kernel(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ pins,
                                   const uint32_t* __restrict__ pin_net, const uint32_t* __restrict__ rank,
                                   const uint32_t* __restrict__ rpoff, uint64_t p, uint32_t* __restrict__ rpins,
                                   uint32_t* __restrict__ rpin_rank) {
  for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < p; j += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t e = pin_net[j];
    const uint32_t r = rank[e];
    const uint32_t dst = rpoff[r] + uint32_t(j - noff[e]);
    rpins[dst] = pins[j];
    rpin_rank[dst] = r;
  }
}

// Builds the net locality ranks of hg once (the hypergraph is immutable).
void ensure_net_ranks(mkp_hypergraph* hg) {
  if (hg->ranked) return;
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t m = hg->m;
  hg->net_order.alloc(std::max<uint32_t>(m, 1));
  hg->net_rank.alloc(std::max<uint32_t>(m, 1));
  hg->inc_rank.alloc(std::max<uint64_t>(hg->p, 1));
  if (m) {
    DevBuf<uint64_t> keys(m), sorted(m);
    net_key_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(hg->noff.p, hg->pins.p, m, keys.p);
    MKP_LAUNCH_CHECK();
    int bits = 1;  // key bits above the 32-bit net id
    while ((uint64_t(1) << bits) < hg->n) ++bits;
    sort_keys_u64(c, keys.p, sorted.p, m, 0, 32 + bits);
    net_rank_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(sorted.p, m, hg->net_order.p, hg->net_rank.p);
    MKP_LAUNCH_CHECK();
  }
  hg->rpoff.alloc(size_t(m) + 1);
  hg->rpins.alloc(std::max<uint64_t>(hg->p, 1));
  hg->rpin_rank.alloc(std::max<uint64_t>(hg->p, 1));
  {
    DevBuf<uint32_t> size(std::max<uint32_t>(m, 1));
    if (m) {
      ranked_size_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(hg->noff.p, hg->net_order.p, m, size.p);
      MKP_LAUNCH_CHECK();
    }
    exclusive_scan_u32(c, size.p, hg->rpoff.p, m);
  }
  if (hg->p) {
    inc_rank_kernel<<<grid_for(hg->p, 256, 65536), 256, 0, s>>>(hg->inc.p, hg->net_rank.p, hg->p, hg->inc_rank.p);
    ranked_pins_kernel<<<grid_for(hg->p, 256, 65536), 256, 0, s>>>(hg->noff.p, hg->pins.p, hg->pin_net.p,
                                                                   hg->net_rank.p, hg->rpoff.p, hg->p, hg->rpins.p,
                                                                   hg->rpin_rank.p);
    MKP_LAUNCH_CHECK();
  }
  hg->ranked = true;
}

}  // namespace mkp

using namespace mkp;

extern "C" int mkp_louvain_evaluate(mkp_hypergraph* hg, const mkp_louvain_params* prm,
                                    const int32_t* node_communities, int32_t* targets, double* gains,
                                    double* modularity) {
  return guarded([&] {
    MKP_CHECK(hg && prm && node_communities, "null argument");
    Context& c = hg->ctx->c;
    Bind bind_(c);
    int wt = prm->edge_weight;
    if (wt < 0) wt = (double(hg->m) / double(std::max<uint32_t>(hg->n, 1)) < 0.75) ? 2 : 0;
    LouvainRun run(c, *prm);
    bool is_graph = hg->m > 0;
    {
      std::vector<uint32_t> noff(size_t(hg->m) + 1);
      MKP_HIP(hipMemcpyAsync(noff.data(), hg->noff.p, noff.size() * 4, hipMemcpyDeviceToHost, c.stream));
      c.sync();
      for (uint32_t e = 0; e < hg->m && is_graph; ++e) is_graph = noff[e + 1] - noff[e] == 2;
    }
    const uint32_t nodes = is_graph ? hg->n : hg->n + hg->m;
    for (uint32_t u = 0; u < nodes; ++u)
      MKP_CHECK(node_communities[u] >= 0 && uint32_t(node_communities[u]) < nodes, "community id out of range");
    if (is_graph) {
      PlainView g{hg->n, hg->voff.p, hg->inc.p, hg->noff.p, hg->pins.p, hg->inc_vtx.p, hg->nw.p, wt,
                  hg->unit_net_weights};
      run.evaluate(g, nodes, node_communities, targets, gains, modularity);
    } else {
      ensure_net_ranks(hg);
      StarView g{hg->n,         hg->m,         hg->p,    hg->voff.p, hg->inc.p,          hg->noff.p,
                 hg->pins.p,    hg->inc_vtx.p, hg->pin_net.p, hg->nw.p, wt, hg->unit_net_weights,
                 hg->net_order.p, hg->inc_rank.p, hg->rpoff.p, hg->rpins.p, hg->rpin_rank.p};
      run.evaluate(g, nodes, node_communities, targets, gains, modularity);
    }
  });
}

// Clustering of net node N + rank(e) in the numbering of the reference's
// Graph, net e at node N + e (graph.cpp: the star expansion's net nodes
// follow the hypernodes in net order).
__global__ void star_net_order_kernel(const int32_t* __restrict__ C, uint32_t N, uint32_t M,
                                      const uint32_t* __restrict__ order, int32_t* __restrict__ out) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < M; r += gridDim.x * blockDim.x)
    out[N + order[r]] = C[N + r];
}

// all_size_two over the nets: the Louvain graph is the hypergraph itself
static bool louvain_is_graph(mkp_hypergraph* hg) {
  Context& c = hg->ctx->c;
  if (!hg->m) return false;
  DevBuf<uint32_t> flag(1);
  fill_async<uint32_t>(flag.p, 1u, 1, c.stream);
  all_size_two_kernel<<<grid_for(hg->m, 256, 65536), 256, 0, c.stream>>>(hg->noff.p, hg->m, flag.p);
  MKP_LAUNCH_CHECK();
  uint32_t is_graph = 0;
  MKP_HIP(hipMemcpyAsync(&is_graph, flag.p, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  return is_graph != 0;
}

// graph_clustering (host, one id per node of the Louvain graph) may be NULL
static void louvain_run(mkp_hypergraph* hg, const mkp_louvain_params* prm, int32_t* communities,
                        int32_t* graph_clustering, mkp_louvain_stats* stats);

extern "C" int mkp_louvain_graph_nodes(const mkp_hypergraph* hg, int64_t* num_nodes) {
  return guarded([&] {
    MKP_CHECK(hg && num_nodes, "null argument");
    mkp_hypergraph* h = const_cast<mkp_hypergraph*>(hg);
    Bind bind_(h->ctx->c);
    *num_nodes = louvain_is_graph(h) ? int64_t(hg->n) : int64_t(hg->n) + int64_t(hg->m);
  });
}

extern "C" int mkp_louvain_graph(mkp_hypergraph* hg, const mkp_louvain_params* prm, int32_t* clustering,
                                 mkp_louvain_stats* stats) {
  return guarded([&] {
    MKP_CHECK(hg && prm && clustering, "null argument");
    louvain_run(hg, prm, nullptr, clustering, stats);
  });
}

extern "C" int mkp_louvain(mkp_hypergraph* hg, const mkp_louvain_params* prm, int32_t* communities,
                           mkp_louvain_stats* stats) {
  return guarded([&] {
    MKP_CHECK(hg && prm, "null argument");
    louvain_run(hg, prm, communities, nullptr, stats);
  });
}

static void louvain_run(mkp_hypergraph* hg, const mkp_louvain_params* prm, int32_t* communities,
                        int32_t* graph_clustering, mkp_louvain_stats* stats) {
  {
    MKP_CHECK(hg && prm, "null argument");
    MKP_CHECK(prm->max_passes >= 0 && prm->sub_rounds >= 1 && prm->sub_rounds <= 255, "bad round parameters");
    Context& c = hg->ctx->c;
    Bind bind_(c);
    MKP_HIP(hipEventRecord(c.ev_a, c.stream));
    // configurePreprocessing (partitioner.cpp:62-72): hybrid -> degree if
    // |E|/|V| < 0.75, else uniform
    int wt = prm->edge_weight;
    if (wt < 0) wt = (double(hg->m) / double(std::max<uint32_t>(hg->n, 1)) < 0.75) ? 2 : 0;
    LouvainRun run(c, *prm);
    const bool is_graph = louvain_is_graph(hg);
    const uint32_t nodes = is_graph ? hg->n : hg->n + hg->m;
    DevBuf<int32_t> C(std::max<uint32_t>(nodes, 1));
    if (is_graph) {
      PlainView g{hg->n, hg->voff.p, hg->inc.p, hg->noff.p, hg->pins.p, hg->inc_vtx.p, hg->nw.p, wt,
                  hg->unit_net_weights};
      run.run_from(g, hg->n, hg->p, C.p);
    } else {
      ensure_net_ranks(hg);
      StarView g{hg->n,         hg->m,         hg->p,    hg->voff.p, hg->inc.p,          hg->noff.p,
                 hg->pins.p,    hg->inc_vtx.p, hg->pin_net.p, hg->nw.p, wt, hg->unit_net_weights,
                 hg->net_order.p, hg->inc_rank.p, hg->rpoff.p, hg->rpins.p, hg->rpin_rank.p};
      run.run_from(g, nodes, 2 * hg->p, C.p);
    }
    // restrictClusteringToHypernodes + setCommunityIDs (graph.h:123-125,
    // partitioner.cpp:116-117)
    MKP_HIP(hipMemcpyAsync(hg->comm.p, C.p, size_t(hg->n) * 4, hipMemcpyDeviceToDevice, c.stream));
    if (communities)
      MKP_HIP(hipMemcpyAsync(communities, C.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
    DevBuf<int32_t> G;
    if (graph_clustering && !is_graph && hg->m) {  // net nodes back to net order
      G.alloc(nodes);
      MKP_HIP(hipMemcpyAsync(G.p, C.p, size_t(hg->n) * 4, hipMemcpyDeviceToDevice, c.stream));
      star_net_order_kernel<<<grid_for(hg->m, 256, 65536), 256, 0, c.stream>>>(C.p, hg->n, hg->m,
                                                                               hg->net_order.p, G.p);
      MKP_LAUNCH_CHECK();
    }
    if (graph_clustering)
      MKP_HIP(hipMemcpyAsync(graph_clustering, G.p ? G.p : C.p, size_t(nodes) * 4, hipMemcpyDeviceToHost,
                             c.stream));
    MKP_HIP(hipEventRecord(c.ev_b, c.stream));
    c.sync();
    if (stats) {
      *stats = run.st;
      float ms = 0;
      MKP_HIP(hipEventElapsedTime(&ms, c.ev_a, c.ev_b));
      stats->device_ms = ms;
    }
  }
}
