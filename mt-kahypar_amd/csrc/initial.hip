// Initial partitioning of the coarsest hypergraph on the device.
//
// Recursive bisection (the recursion of
// recursive_bisection_initial_partitioner.cpp:475-492; the flat pool of
// flat/pool_initial_partitioner.h:36-75 is replaced by a portfolio of runs).
// Every bisection makes all of its runs at once, one workgroup of 1024
// threads per run, and keeps the balanced run with the smallest cut (ties to
// the smaller run index).  Run r grows side 0 by variant r % 3 (greedy
// growing, level-synchronous BFS, random fill; r % 2 above 1024 vertices)
// and then improves the bisection with passes of 2-way gain moves.
//
// The steps of a run are sequential -- growing one vertex at a time and
// moving one vertex at a time is what makes them effective -- and the work
// inside a step (a vertex's nets and their pins, the arg-max over the
// vertices, the next positive-gain vertex of a pass) is spread over the
// workgroup.  Growing scores and gains are integers updated by atomics, so a
// run does not depend on the order of the updates.  For up to kIpLdsN
// vertices the per-vertex state (side, flags, score or gain) lives in LDS;
// for up to kIpAdjN vertices the scores a joining vertex adds are one
// precomputed row of clique scores.
//
// The rules, each stated so that the result is one function of the input:
//   * greedy growing: score of a side-1 vertex = sum over its nets of at most
//     100 pins shared with side 0 of floor(w(e) * 2^16 / (|e| - 1)) (64-bit
//     wrap-around); the min(256, max(1, n / 2048)) highest scores join in
//     order, ties to the smaller id; with no scored vertex the next free
//     vertex after a hashed start vertex seeds;
//   * BFS: a level's vertices join in ascending id order; the next level is
//     every unseen pin of an unscanned net of at most 100 pins incident to a
//     vertex that joined;
//   * random fill and refinement passes visit vertices in the order of a
//     keyed four-round Feistel permutation of [0, n) (cycle walking);
//   * refinement: an overloaded side gives up vertices in descending gain
//     order at the time of the repair (ties to the smaller id) while they fit
//     the other side; then every pass moves, in permutation order, each
//     vertex with positive gain, or zero gain towards the lighter side, that
//     fits its target side, until a pass makes no positive-gain move.
// oracle/initial.cpp restates the same rules on the host.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "internal.h"

namespace mkp {
namespace {

constexpr int kIpThreads = 1024;
constexpr int kIpWaves = kIpThreads / 64;
constexpr uint32_t kIpGrowNetLimit = 100;  // larger nets do not steer growing
constexpr int kIpScoreFrac = 16;           // fractional bits of a growing score
constexpr uint32_t kIpLdsN = 8192;         // all per-vertex state in LDS up to this many vertices
constexpr uint32_t kIpSideN = 32768;       // sides and flags in LDS up to this many vertices
constexpr uint32_t kIpAdjN = 4096;         // clique scores precomputed up to this many vertices (128 MB)
constexpr uint64_t kIpAdjBytes = uint64_t(256) << 20;  // clique-score bytes of one recursion level
constexpr uint32_t kIpBigNet = 64;         // nets of a move with more pins are updated by a whole wave
constexpr uint32_t kIpQueue = 512;
constexpr int kIpMoveBatch = 4;
constexpr int kIpScanBatch = 8;            // vertices whose scores one thread loads at once in an arg-max            // incident nets of a move whose records one thread loads at once
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kIpStartTag = 0x243F6A8885A308D3ull;
constexpr uint64_t kIpFillTag = 0x13198A2E03707344ull;
constexpr uint64_t kIpPassTag = 0xA4093822299F31D0ull;

// per-vertex flag bits
constexpr uint8_t kSkip = 1;  // growing: too heavy for side 0
constexpr uint8_t kSeen = 4;  // BFS: reached
constexpr uint8_t kNext = 8;  // BFS: in the next level
constexpr uint8_t kCand = 16; // repair: candidate not yet visited
constexpr uint8_t kPick = 32; // growing: taken in the current step
constexpr uint32_t kIpRandomMaxN = 1024;  // random-fill runs up to this many vertices
constexpr uint32_t kIpMaxBatch = 256;     // growing takes min(256, max(1, n / 2048)) vertices per step

// Pin counts of a net on side 0 / 1 in one 32-bit word (16 bits each) when
// every net has fewer than 2^16 pins ("narrow"), else in two words: a run's
// counts of all nets then stay small enough for the device's last-level
// cache while the nets' pins and weights are shared by all runs.
template <bool kWide>
__device__ inline void pc_get(const uint32_t* pc, uint32_t e, uint32_t& c0, uint32_t& c1) {
  if (kWide) {
    const uint2 v = reinterpret_cast<const uint2*>(pc)[e];
    c0 = v.x;
    c1 = v.y;
  } else {
    const uint32_t v = pc[e];
    c0 = v & 0xFFFFu;
    c1 = v >> 16;
  }
}
template <bool kWide>
__device__ inline void pc_set(uint32_t* pc, uint32_t e, uint32_t c0, uint32_t c1) {
  if (kWide) reinterpret_cast<uint2*>(pc)[e] = make_uint2(c0, c1);
  else pc[e] = c0 | (c1 << 16);
}

struct IpArgs {
  const uint32_t* noff;
  const uint32_t* pins;
  const uint32_t* voff;
  const uint32_t* inc;
  const int32_t* vw;
  const int32_t* nw;
  uint32_t n, m;
  int64_t W;
  int64_t max0, max1, target0;
  uint64_t stream;
  int passes;
  // per-run buffers, run r at r * n (r * m or r * 2m words for pc, r * 4
  // for out)
  uint8_t* side;   // the run's bisection
  uint8_t* flag;   // beyond kIpSideN vertices only
  int64_t* val;    // gains from the gain kernel; growing scores beyond kIpLdsN vertices
  int64_t* snap;
  uint32_t* list;
  uint32_t* pc;    // per net: pins on side 0 / 1
  int64_t* out;    // cut, balanced, side-0 weight, side-1 weight
  // clique scores of growing, n x n (row v: what v's joining adds to every
  // vertex), for up to kIpAdjN vertices; null: growing walks the nets
  const uint64_t* adj;
};

struct IpPerm {
  uint64_t key;
  uint32_t n, hb, mask;
};
__device__ inline IpPerm ip_perm_of(uint64_t key, uint32_t n) {
  uint32_t b = 1;
  while ((uint64_t(1) << b) < n) ++b;
  IpPerm p;
  p.key = key;
  p.n = n;
  p.hb = (b + 1) / 2;
  p.mask = (1u << p.hb) - 1u;
  return p;
}
__device__ inline uint32_t ip_perm(const IpPerm& p, uint32_t i) {
  uint32_t x = i;
  do {
    uint32_t L = x >> p.hb, R = x & p.mask;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r) {
      const uint32_t t = L ^ (uint32_t(mix64(p.key ^ (uint64_t(r) << 32) ^ R)) & p.mask);
      L = R;
      R = t;
    }
    x = (L << p.hb) | R;
  } while (x >= p.n);
  return x;
}

__host__ __device__ inline uint64_t ip_run_key(uint64_t stream, int r) { return mix64(stream ^ mix64(uint64_t(r) + 1)); }

struct IpShared {
  uint64_t rk[kIpWaves];
  uint32_t ri[kIpWaves];
  int64_t rs[kIpWaves];
  uint32_t wc[kIpWaves];
  uint64_t bk;
  uint32_t bi;
  int64_t bs;
  int64_t w[2];
  uint32_t scan, count, qn;
  uint32_t mv_u;
  int64_t mv_g;
  uint32_t qe[kIpQueue];
  int64_t qdf[kIpQueue], qdt[kIpQueue];
  uint32_t buf[kIpThreads];
  uint32_t pick[kIpMaxBatch];
  uint32_t pre[kIpMaxBatch + 1];  // growing: incident-net offsets of the step's joined vertices
};

__device__ inline unsigned long long shfl_xor_u64(unsigned long long x, int off) { return __shfl_xor(x, off, 64); }

// block-wide maximum of (key, -id): the larger key, ties to the smaller id;
// every thread calls it and receives the result
__device__ inline void block_argmax(IpShared& s, uint64_t& key, uint32_t& id) {
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t ok = shfl_xor_u64(key, off);
    const uint32_t oi = __shfl_xor(id, off, 64);
    if (ok > key || (ok == key && oi < id)) { key = ok; id = oi; }
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) { s.rk[w] = key; s.ri[w] = id; }
  __syncthreads();
  if (w == 0) {
    key = lane < kIpWaves ? s.rk[lane] : 0;
    id = lane < kIpWaves ? s.ri[lane] : kNone;
    for (int off = 8; off > 0; off >>= 1) {
      const uint64_t ok = shfl_xor_u64(key, off);
      const uint32_t oi = __shfl_xor(id, off, 64);
      if (ok > key || (ok == key && oi < id)) { key = ok; id = oi; }
    }
    if (lane == 0) { s.bk = key; s.bi = id; }
  }
  __syncthreads();
  key = s.bk;
  id = s.bi;
}
__device__ inline uint32_t block_min(IpShared& s, uint32_t v) {
  uint64_t key = 0;
  block_argmax(s, key, v);
  return v;
}
__device__ inline int64_t block_sum(IpShared& s, int64_t v) {
  for (int off = 32; off > 0; off >>= 1) v += int64_t(shfl_xor_u64((unsigned long long)v, off));
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s.rs[w] = v;
  __syncthreads();
  if (w == 0) {
    v = lane < kIpWaves ? s.rs[lane] : 0;
    for (int off = 8; off > 0; off >>= 1) v += int64_t(shfl_xor_u64((unsigned long long)v, off));
    if (lane == 0) s.bs = v;
  }
  __syncthreads();
  return s.bs;
}

// list[0 .. count) = the ids x < n with kNext set, ascending; clears the bit
__device__ inline uint32_t compact_next(IpShared& s, uint8_t* flag, uint32_t n, uint32_t* list) {
  uint32_t base = 0;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t c0 = 0; c0 < n; c0 += kIpThreads) {
    const uint32_t x = c0 + threadIdx.x;
    bool f = false;
    if (x < n) {
      const uint8_t fl = flag[x];
      f = (fl & kNext) != 0;
      if (f) flag[x] = uint8_t(fl & ~kNext);
    }
    const uint64_t b = __ballot(f);
    if (lane == 0) s.wc[w] = uint32_t(__popcll(b));
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int i = 0; i < kIpWaves; ++i) {
      off += i < w ? s.wc[i] : 0u;
      tot += s.wc[i];
    }
    if (f) list[base + off + uint32_t(__popcll(b & ((1ull << lane) - 1ull)))] = x;
    __syncthreads();
    base += tot;
  }
  return base;
}

// Moves v to the other side and keeps every gain current (all threads call
// it with the same v; the state is settled when it returns).  With a, b the
// pin counts of a net on v's old and new side before the move, a pin on the
// old side gains w * ([a == 2] - [a == 1] + [b == 0]), a pin on the new side
// w * ([b == 0] - [a == 1] - [b == 1]); v's own gain is negated.
// val[x] += d for the pins x of pins[jb, je) other than `skip` for which
// take(x) holds; up to 8 pins are loaded before any is used
template <typename F>
__device__ inline void ip_add_pins(const uint32_t* pins, uint32_t jb, uint32_t je, uint32_t skip, int64_t* val,
                                   F&& delta) {
  for (uint32_t j0 = jb; j0 < je; j0 += 8) {
    uint32_t xs[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) xs[q] = j0 + q < je ? pins[j0 + q] : kNone;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (xs[q] != kNone && xs[q] != skip) {
        const int64_t d = delta(xs[q]);
        if (d) atomicAdd(reinterpret_cast<unsigned long long*>(&val[xs[q]]), (unsigned long long)d);
      }
  }
}

template <bool kWide>
__device__ inline void ip_move(const IpArgs& a, IpShared& s, uint8_t* side, int64_t* val, uint32_t* pc, uint32_t v) {
  const int f = side[v], t = 1 - f;
  const int64_t gv = val[v];
  const uint32_t vb = a.voff[v], ve = a.voff[v + 1];
  for (uint32_t i0 = vb + threadIdx.x; i0 < ve; i0 += kIpMoveBatch * kIpThreads) {
    // the records of up to kIpMoveBatch nets are loaded before any is used
    uint32_t eb[kIpMoveBatch], c0b[kIpMoveBatch], c1b[kIpMoveBatch];
    int32_t wb[kIpMoveBatch];
#pragma unroll
    for (int q = 0; q < kIpMoveBatch; ++q) {
      const uint32_t i = i0 + q * kIpThreads;
      eb[q] = i < ve ? a.inc[i] : kNone;
    }
#pragma unroll
    for (int q = 0; q < kIpMoveBatch; ++q) {
      c0b[q] = c1b[q] = 0;
      wb[q] = 0;
      if (eb[q] != kNone) {
        pc_get<kWide>(pc, eb[q], c0b[q], c1b[q]);
        wb[q] = a.nw[eb[q]];
      }
    }
#pragma unroll
    for (int q = 0; q < kIpMoveBatch; ++q) {
    if (eb[q] == kNone) continue;
    const uint32_t e = eb[q];
    const uint32_t ca = f ? c1b[q] : c0b[q], cb = f ? c0b[q] : c1b[q];
    const int64_t we = wb[q];
    const int64_t df = we * (int64_t(ca == 2) - int64_t(ca == 1) + int64_t(cb == 0));
    const int64_t dt = we * (int64_t(cb == 0) - int64_t(ca == 1) - int64_t(cb == 1));
    if (df || dt) {
      const uint32_t jb = a.noff[e], je = a.noff[e + 1];
      uint32_t slot = kNone;
      if (je - jb > kIpBigNet) {
        slot = atomicAdd(&s.qn, 1u);
        if (slot < kIpQueue) {
          s.qe[slot] = e;
          s.qdf[slot] = df;
          s.qdt[slot] = dt;
        }
      }
      if (slot == kNone || slot >= kIpQueue)
        ip_add_pins(a.pins, jb, je, v, val, [&](uint32_t x) { return side[x] == f ? df : dt; });
    }
    if (f == 0) pc_set<kWide>(pc, e, ca - 1, cb + 1);
    else pc_set<kWide>(pc, e, cb + 1, ca - 1);
    }
  }
  __syncthreads();
  const uint32_t qn = min(s.qn, kIpQueue);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (uint32_t q = w; q < qn; q += kIpWaves) {
    const uint32_t e = s.qe[q];
    const int64_t df = s.qdf[q], dt = s.qdt[q];
    for (uint32_t j = a.noff[e] + lane; j < a.noff[e + 1]; j += 64) {
      const uint32_t x = a.pins[j];
      if (x != v) atomicAdd(reinterpret_cast<unsigned long long*>(&val[x]), (unsigned long long)(side[x] == f ? df : dt));
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    side[v] = uint8_t(t);
    val[v] = -gv;
    s.w[f] -= a.vw[v];
    s.w[t] += a.vw[v];
    s.qn = 0;
  }
  __syncthreads();
}

// Where a run keeps its per-vertex state: kMode 0 all in LDS (up to kIpLdsN
// vertices), 1 sides and flags in LDS and scores / gains in global memory (up
// to kIpSideN), 2 all in global memory.
template <int kMode>
constexpr uint32_t ip_side_slots() { return kMode == 0 ? kIpLdsN : (kMode == 1 ? kIpSideN : 1); }
template <int kMode>
constexpr uint32_t ip_val_slots() { return kMode == 0 ? kIpLdsN : 1; }

// The run kernels take the bisections of one recursion level at once:
// workgroup b works on bisection b / runs (its arguments args[b / runs]).
template <int kMode, bool kWide>
__global__ __launch_bounds__(kIpThreads) void ip_grow_kernel(const IpArgs* args, uint32_t runs) {
  const IpArgs a = args[blockIdx.x / runs];
  __shared__ IpShared s;
  __shared__ uint8_t l_side[ip_side_slots<kMode>()];
  __shared__ uint8_t l_flag[ip_side_slots<kMode>()];
  __shared__ int64_t l_val[ip_val_slots<kMode>()];
  const int r = int(blockIdx.x % runs);
  const uint32_t n = a.n, m = a.m;
  const uint32_t tid = threadIdx.x;
  uint8_t* g_side = a.side + size_t(r) * n;
  uint8_t* side = kMode < 2 ? l_side : g_side;
  uint8_t* flag = kMode < 2 ? l_flag : a.flag + size_t(r) * n;
  int64_t* val = kMode == 0 ? l_val : a.val + size_t(r) * n;
  uint32_t* list = a.list + size_t(r) * n;
  uint32_t* pc = a.pc + size_t(r) * m * (kWide ? 2 : 1);
  const uint64_t rk = ip_run_key(a.stream, r);
  // no random fill beyond kIpRandomMaxN vertices: a random start needs
  // about n / 4 refinement moves there and never wins
  const int variant = n > kIpRandomMaxN ? r % 2 : r % 3;
  const uint32_t batch = min(kIpMaxBatch, max(1u, n / 2048u));
  const uint32_t start = uint32_t(mix64(rk ^ kIpStartTag) % n);
  for (uint32_t x = tid; x < n; x += kIpThreads) {
    side[x] = 1;
    flag[x] = 0;
    val[x] = 0;
  }
  if (variant == 1)
    for (uint32_t e = tid; e < m; e += kIpThreads) pc[kWide ? 2 * e : e] = 0;  // scanned marks
  if (tid == 0) {
    s.w[0] = 0;
    s.scan = 0;
    s.count = 0;
    s.qn = 0;
  }
  __syncthreads();

  // ---------------------------------------------------------- growing --
  if (variant == 0) {
    for (;;) {
      if (s.w[0] >= a.target0) break;
      // the `batch` highest non-zero scores, ties to the smaller id
      uint32_t np = 0;
      while (np < batch) {
        uint64_t key = 0;
        uint32_t v = kNone;
        for (uint32_t x0 = tid; x0 < n; x0 += kIpScanBatch * kIpThreads) {
          // scores of kIpScanBatch vertices are loaded before any is used
          uint64_t kb[kIpScanBatch];
#pragma unroll
          for (int q = 0; q < kIpScanBatch; ++q) {
            const uint32_t x = x0 + q * kIpThreads;
            kb[q] = x < n ? uint64_t(val[x]) : 0;
          }
#pragma unroll
          for (int q = 0; q < kIpScanBatch; ++q) {
            const uint32_t x = x0 + q * kIpThreads;
            if (kb[q] != 0 && side[x] == 1 && !(flag[x] & (kSkip | kPick)) && (v == kNone || kb[q] > key)) {
              key = kb[q];
              v = x;
            }
          }
        }
        block_argmax(s, key, v);
        if (v == kNone) break;
        if (tid == 0) {
          flag[v] = uint8_t(flag[v] | kPick);
          s.pick[np] = v;
        }
        ++np;
        __syncthreads();
      }
      if (np == 0) {  // a seed: the next free vertex after `start`
        const uint32_t sc = s.scan;
        uint32_t first = kNone;
        for (uint32_t i = sc + tid; i < n; i += kIpThreads) {
          const uint32_t x = (start + i) % n;
          if (side[x] == 1 && !(flag[x] & kSkip)) { first = i; break; }
        }
        first = block_min(s, first);
        if (first == kNone) break;
        if (tid == 0) {
          s.pick[0] = (start + first) % n;
          s.scan = first + 1;
        }
        np = 1;
        __syncthreads();
      }
      // the picks join in order while side 0 is below its target
      if (tid == 0) {
        uint32_t na = 0, deg = 0;
        int64_t w0 = s.w[0];
        for (uint32_t i = 0; i < np; ++i) {
          const uint32_t v = s.pick[i];
          flag[v] = uint8_t(flag[v] & ~kPick);
          if (w0 >= a.target0) continue;
          if (w0 + a.vw[v] > a.max0) {
            flag[v] = uint8_t(flag[v] | kSkip);
            continue;
          }
          side[v] = 0;
          w0 += a.vw[v];
          s.pre[na] = deg;
          s.pick[na++] = v;
          deg += a.voff[v + 1] - a.voff[v];
        }
        s.pre[na] = deg;
        s.w[0] = w0;
        s.count = na;
      }
      __syncthreads();
      const uint32_t na = s.count;
      if (a.adj) {
        for (uint32_t i = 0; i < na; ++i) {
          const uint64_t* row = a.adj + size_t(s.pick[i]) * n;
          for (uint32_t x = tid; x < n; x += kIpThreads)
            if (side[x] == 1 && !(flag[x] & kSkip)) val[x] = int64_t(uint64_t(val[x]) + row[x]);
        }
      } else {
        // the incident nets of all joined vertices in one flat range, so
        // that every thread's loads are independent of each other
        const uint32_t total = s.pre[na];
        for (uint32_t k = tid; k < total; k += kIpThreads) {
          uint32_t lo = 0, hi = na - 1;  // the joined vertex whose range holds k
          while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s.pre[mid] <= k) lo = mid;
            else hi = mid - 1;
          }
          const uint32_t e = a.inc[a.voff[s.pick[lo]] + (k - s.pre[lo])];
          const uint32_t jb = a.noff[e], je = a.noff[e + 1];
          if (je - jb > kIpGrowNetLimit) continue;
          const uint64_t add = (uint64_t(uint32_t(a.nw[e])) << kIpScoreFrac) / uint64_t(max(je - jb - 1, 1u));
          ip_add_pins(a.pins, jb, je, kNone, val, [&](uint32_t x) {
            return (side[x] == 1 && !(flag[x] & kSkip)) ? int64_t(add) : int64_t(0);
          });
        }
      }
      __syncthreads();
    }
  } else if (variant == 1) {
    for (;;) {
      if (s.w[0] >= a.target0) break;
      if (s.count == 0) {  // a new seed: the next unseen vertex after `start`
        const uint32_t sc = s.scan;
        uint32_t first = kNone;
        for (uint32_t i = sc + tid; i < n; i += kIpThreads)
          if (!(flag[(start + i) % n] & kSeen)) { first = i; break; }
        first = block_min(s, first);
        if (first == kNone) break;
        if (tid == 0) {
          const uint32_t x = (start + first) % n;
          flag[x] = uint8_t(flag[x] | kSeen);
          list[0] = x;
          s.count = 1;
          s.scan = first + 1;
        }
        __syncthreads();
      }
      // the level joins in ascending id order (one thread; every vertex is
      // in at most one level)
      if (tid == 0) {
        uint32_t na = 0;
        int64_t w0 = s.w[0];
        for (uint32_t i = 0; i < s.count; ++i) {
          if (w0 >= a.target0) break;
          const uint32_t x = list[i];
          if (w0 + a.vw[x] > a.max0) continue;
          side[x] = 0;
          w0 += a.vw[x];
          list[na++] = x;
        }
        s.w[0] = w0;
        s.count = na;
      }
      __syncthreads();
      if (s.w[0] >= a.target0) break;
      const uint32_t na = s.count;
      for (uint32_t i = 0; i < na; ++i) {
        const uint32_t v = list[i];
        for (uint32_t q = a.voff[v] + tid; q < a.voff[v + 1]; q += kIpThreads) {
          const uint32_t e = a.inc[q];
          const uint32_t jb = a.noff[e], je = a.noff[e + 1];
          if (pc[kWide ? 2 * e : e] || je - jb > kIpGrowNetLimit) continue;
          pc[kWide ? 2 * e : e] = 1;
          for (uint32_t j = jb; j < je; ++j) {
            const uint32_t x = a.pins[j];
            const uint8_t fl = flag[x];
            if (!(fl & kSeen)) flag[x] = uint8_t(fl | kSeen | kNext);
          }
        }
      }
      __syncthreads();
      const uint32_t cnt = compact_next(s, flag, n, list);
      if (tid == 0) s.count = cnt;
      __syncthreads();
    }
  } else {
    const IpPerm P = ip_perm_of(mix64(rk ^ kIpFillTag), n);
    for (uint32_t pos = 0; pos < n; pos += kIpThreads) {
      if (s.w[0] >= a.target0) break;
      if (pos + tid < n) s.buf[tid] = ip_perm(P, pos + tid);
      __syncthreads();
      if (tid == 0) {
        int64_t w0 = s.w[0];
        const uint32_t cnt = min(uint32_t(kIpThreads), n - pos);
        for (uint32_t i = 0; i < cnt && w0 < a.target0; ++i) {
          const uint32_t u = s.buf[i];
          if (w0 + a.vw[u] > a.max0) continue;
          side[u] = 0;
          w0 += a.vw[u];
        }
        s.w[0] = w0;
      }
      __syncthreads();
    }
  }

  // the bisection and side-0 weight go to the refinement; its gains are
  // summed into the zeroed global array by the gain kernel
  __syncthreads();
  int64_t* g_val = a.val + size_t(r) * n;
  for (uint32_t x = tid; x < n; x += kIpThreads) {
    if (kMode < 2) g_side[x] = side[x];
    g_val[x] = 0;
  }
  if (tid == 0) a.out[4 * size_t(r) + 2] = s.w[0];
}

// Pin counts, net records and initial gains of every run's bisection, with
// `parts` workgroups per run, each over one range of nets: a pin on side s
// of net e gains w(e) if it is e's only pin on s and loses w(e) if e has no
// pin on the other side.  Nets of more than kIpBigNet pins are done by whole
// waves.
template <int kMode, bool kWide>
__global__ __launch_bounds__(kIpThreads) void ip_gain_kernel(const IpArgs* args, uint32_t runs, uint32_t parts) {
  const IpArgs a = args[blockIdx.x / (runs * parts)];
  __shared__ uint8_t l_side[ip_side_slots<kMode>()];
  __shared__ int64_t l_val[ip_val_slots<kMode>()];
  __shared__ uint32_t qe[kIpQueue];
  __shared__ uint32_t qn;
  const uint32_t r = (blockIdx.x / parts) % runs, g = blockIdx.x % parts;
  const uint32_t n = a.n, m = a.m, tid = threadIdx.x;
  const uint8_t* g_side = a.side + size_t(r) * n;
  int64_t* g_val = a.val + size_t(r) * n;
  uint32_t* pc = a.pc + size_t(r) * m * (kWide ? 2 : 1);
  const uint8_t* side = kMode < 2 ? l_side : g_side;
  int64_t* val = kMode == 0 ? l_val : g_val;
  for (uint32_t x = tid; x < n; x += kIpThreads) {
    if (kMode < 2) l_side[x] = g_side[x];
    if (kMode == 0) l_val[x] = 0;
  }
  if (tid == 0) qn = 0;
  __syncthreads();
  const uint32_t chunk = (m + parts - 1) / parts;
  const uint32_t e0 = min(m, g * chunk), e1 = min(m, e0 + chunk);
  for (uint32_t e = e0 + tid; e < e1; e += kIpThreads) {
    const uint32_t jb = a.noff[e], je = a.noff[e + 1];
    if (je - jb > kIpBigNet) {
      const uint32_t q = atomicAdd(&qn, 1u);
      if (q < kIpQueue) {
        qe[q] = e;
        continue;
      }
    }
    uint32_t c1 = 0;
    for (uint32_t j = jb; j < je; ++j) c1 += side[a.pins[j]];
    const uint32_t c0 = (je - jb) - c1;
    const int32_t w = a.nw[e];
    pc_set<kWide>(pc, e, c0, c1);
    if (w != 0 && (c0 <= 1 || c1 <= 1)) {
      const int64_t d0 = (c0 == 1 ? int64_t(w) : 0) - (c1 == 0 ? int64_t(w) : 0);
      const int64_t d1 = (c1 == 1 ? int64_t(w) : 0) - (c0 == 0 ? int64_t(w) : 0);
      for (uint32_t j = jb; j < je; ++j) {
        const uint32_t x = a.pins[j];
        const int64_t d = side[x] ? d1 : d0;
        if (d) atomicAdd(reinterpret_cast<unsigned long long*>(&val[x]), (unsigned long long)d);
      }
    }
  }
  __syncthreads();
  const uint32_t nq = min(qn, kIpQueue);
  const int wv = tid >> 6, lane = tid & 63;
  for (uint32_t q = wv; q < nq; q += kIpWaves) {
    const uint32_t e = qe[q];
    const uint32_t jb = a.noff[e], je = a.noff[e + 1];
    uint32_t c1 = 0;
    for (uint32_t j = jb + lane; j < je; j += 64) c1 += side[a.pins[j]];
    for (int off = 32; off > 0; off >>= 1) c1 += __shfl_xor(c1, off, 64);
    const uint32_t c0 = (je - jb) - c1;
    const int32_t w = a.nw[e];
    if (lane == 0) pc_set<kWide>(pc, e, c0, c1);
    if (w != 0 && (c0 <= 1 || c1 <= 1)) {
      const int64_t d0 = (c0 == 1 ? int64_t(w) : 0) - (c1 == 0 ? int64_t(w) : 0);
      const int64_t d1 = (c1 == 1 ? int64_t(w) : 0) - (c0 == 0 ? int64_t(w) : 0);
      for (uint32_t j = jb + lane; j < je; j += 64) {
        const uint32_t x = a.pins[j];
        const int64_t d = side[x] ? d1 : d0;
        if (d) atomicAdd(reinterpret_cast<unsigned long long*>(&val[x]), (unsigned long long)d);
      }
    }
  }
  if (kMode == 0) {
    __syncthreads();
    for (uint32_t x = tid; x < n; x += kIpThreads)
      if (l_val[x]) atomicAdd(reinterpret_cast<unsigned long long*>(&g_val[x]), (unsigned long long)l_val[x]);
  }
}

// Balance repair, gain-move passes and the cut of every run (one workgroup
// per run, gains from the gain kernel).
template <int kMode, bool kWide>
__global__ __launch_bounds__(kIpThreads) void ip_refine_kernel(const IpArgs* args, uint32_t runs) {
  const IpArgs a = args[blockIdx.x / runs];
  __shared__ IpShared s;
  __shared__ uint8_t l_side[ip_side_slots<kMode>()];
  __shared__ uint8_t l_flag[ip_side_slots<kMode>()];
  __shared__ int64_t l_val[ip_val_slots<kMode>()];
  const int r = int(blockIdx.x % runs);
  const uint32_t n = a.n, m = a.m;
  const uint32_t tid = threadIdx.x;
  uint8_t* g_side = a.side + size_t(r) * n;
  int64_t* g_val = a.val + size_t(r) * n;
  uint8_t* side = kMode < 2 ? l_side : g_side;
  uint8_t* flag = kMode < 2 ? l_flag : a.flag + size_t(r) * n;
  int64_t* val = kMode == 0 ? l_val : g_val;
  int64_t* snap = a.snap + size_t(r) * n;
  uint32_t* pc = a.pc + size_t(r) * m * (kWide ? 2 : 1);
  const uint64_t rk = ip_run_key(a.stream, r);
  for (uint32_t x = tid; x < n; x += kIpThreads) {
    if (kMode < 2) l_side[x] = g_side[x];
    if (kMode == 0) l_val[x] = g_val[x];
    flag[x] = 0;
  }
  if (tid == 0) {
    s.w[0] = a.out[4 * size_t(r) + 2];
    s.w[1] = a.W - s.w[0];
    s.qn = 0;
  }
  __syncthreads();
  for (int f = 0; f < 2; ++f) {
    const int64_t mf = f == 0 ? a.max0 : a.max1, mt = f == 0 ? a.max1 : a.max0;
    if (s.w[f] <= mf) continue;
    for (uint32_t x = tid; x < n; x += kIpThreads)
      if (side[x] == f) {
        snap[x] = val[x];
        flag[x] = uint8_t(flag[x] | kCand);
      }
    __syncthreads();
    for (;;) {
      if (s.w[f] <= mf) break;
      // descending gain, ties to the smaller id: order-preserving key of
      // the signed gain
      uint64_t key = 0;
      uint32_t c = kNone;
      for (uint32_t x0 = tid; x0 < n; x0 += kIpScanBatch * kIpThreads) {
        uint64_t kb[kIpScanBatch];
#pragma unroll
        for (int q = 0; q < kIpScanBatch; ++q) {
          const uint32_t x = x0 + q * kIpThreads;
          kb[q] = x < n ? uint64_t(snap[x]) ^ (1ull << 63) : 0;
        }
#pragma unroll
        for (int q = 0; q < kIpScanBatch; ++q) {
          const uint32_t x = x0 + q * kIpThreads;
          if (x < n && (flag[x] & kCand) && (c == kNone || kb[q] > key)) {
            key = kb[q];
            c = x;
          }
        }
      }
      block_argmax(s, key, c);
      if (c == kNone) break;
      const bool fits = s.w[1 - f] + a.vw[c] <= mt;
      __syncthreads();
      if (tid == 0) flag[c] = uint8_t(flag[c] & ~kCand);
      __syncthreads();
      if (fits) ip_move<kWide>(a, s, side, val, pc, c);
    }
    for (uint32_t x = tid; x < n; x += kIpThreads) flag[x] = uint8_t(flag[x] & ~kCand);
    __syncthreads();
  }
  for (int p = 0; p < a.passes; ++p) {
    const IpPerm P = ip_perm_of(mix64(rk ^ mix64(kIpPassTag + uint64_t(p))), n);
    bool improved = false;
    uint32_t pos = 0;
    while (pos < n) {
      // the first position from pos on whose vertex moves under the current
      // state (nothing changes before it)
      const uint32_t i = pos + tid;
      uint32_t hit = kNone;
      if (i < n) {
        const uint32_t u = ip_perm(P, i);
        const int f = side[u], t = 1 - f;
        const int64_t vu = a.vw[u];
        if (s.w[t] + vu <= (t == 0 ? a.max0 : a.max1)) {
          const int64_t g = val[u];
          if (g > 0 || (g == 0 && s.w[f] - vu > s.w[t] + vu)) hit = i;
        }
        s.buf[tid] = u;
      }
      hit = block_min(s, hit);
      if (hit == kNone) {
        pos += kIpThreads;
        __syncthreads();
        continue;
      }
      const uint32_t u = s.buf[hit - pos];
      improved |= val[u] > 0;
      __syncthreads();
      ip_move<kWide>(a, s, side, val, pc, u);
      pos = hit + 1;
    }
    if (!improved) break;
  }

  int64_t cut = 0;
  for (uint32_t e = tid; e < m; e += kIpThreads) {
    uint32_t c0, c1;
    pc_get<kWide>(pc, e, c0, c1);
    if (c0 && c1) cut += a.nw[e];
  }
  cut = block_sum(s, cut);
  if (kMode < 2)
    for (uint32_t x = tid; x < n; x += kIpThreads) g_side[x] = side[x];
  if (tid == 0) {
    int64_t* o = a.out + 4 * size_t(r);
    o[0] = cut;
    o[1] = s.w[0] <= a.max0 && s.w[1] <= a.max1;
    o[2] = s.w[0];
    o[3] = s.w[1];
  }
}

// Row v of the clique scores: for every other vertex x, the sum over the
// nets of at most 100 pins holding both of floor(w(e) * 2^16 / (|e| - 1)),
// i.e. what v's joining side 0 adds to x's growing score (built in LDS with
// atomics, written out in one coalesced pass).
__global__ __launch_bounds__(kIpThreads) void ip_adjacency_kernel(const uint32_t* noff, const uint32_t* pins,
                                                                  const uint32_t* voff, const uint32_t* inc,
                                                                  const int32_t* nw, uint32_t n, uint64_t* adj) {
  __shared__ uint64_t row[kIpAdjN];
  for (uint32_t v = blockIdx.x; v < n; v += gridDim.x) {
    for (uint32_t x = threadIdx.x; x < n; x += kIpThreads) row[x] = 0;
    __syncthreads();
    for (uint32_t i = voff[v] + threadIdx.x; i < voff[v + 1]; i += kIpThreads) {
      const uint32_t e = inc[i];
      const uint32_t jb = noff[e], je = noff[e + 1];
      if (je - jb > kIpGrowNetLimit) continue;
      const uint64_t add = (uint64_t(uint32_t(nw[e])) << kIpScoreFrac) / uint64_t(max(je - jb - 1, 1u));
      for (uint32_t j = jb; j < je; ++j) {
        const uint32_t x = pins[j];
        if (x != v) atomicAdd(reinterpret_cast<unsigned long long*>(&row[x]), (unsigned long long)add);
      }
    }
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < n; x += kIpThreads) adj[size_t(v) * n + x] = row[x];
    __syncthreads();
  }
}

// index of the balanced run with the smallest cut (ties: the smaller index),
// one workgroup per bisection
__global__ void ip_best_kernel(const IpArgs* args, int runs, int32_t* bests) {
  if (threadIdx.x != 0) return;
  const int64_t* out = args[blockIdx.x].out;
  int32_t* best = bests + blockIdx.x;
  int b = 0;
  for (int r = 1; r < runs; ++r) {
    const bool ba = out[4 * r + 1] != 0, bb = out[4 * b + 1] != 0;
    if ((ba && !bb) || (ba == bb && out[4 * r] < out[4 * b])) b = r;
  }
  *best = b;
}

// New ids of both halves (order kept), their vertex weights and original
// ids; sizes = {vertices of side 0, weight of side 0, weight of side 1}.
__global__ __launch_bounds__(kIpThreads) void ip_split_vertices_kernel(const uint8_t* sides, const int32_t* best,
                                                                       uint32_t n, const int32_t* vw,
                                                                       const uint32_t* orig, uint32_t* newid,
                                                                       int32_t* vw0, int32_t* vw1, uint32_t* orig0,
                                                                       uint32_t* orig1, int64_t* sizes) {
  __shared__ uint32_t wc[kIpWaves];
  __shared__ int64_t ws[2][kIpWaves];
  const uint8_t* side = sides + size_t(*best) * n;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t base0 = 0;
  int64_t sw[2] = {0, 0};
  for (uint32_t c0 = 0; c0 < n; c0 += kIpThreads) {
    const uint32_t x = c0 + threadIdx.x;
    const bool in = x < n;
    const int sd = in ? side[x] : 1;
    const uint64_t b = __ballot(in && sd == 0);
    if (lane == 0) wc[w] = uint32_t(__popcll(b));
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int i = 0; i < kIpWaves; ++i) {
      off += i < w ? wc[i] : 0u;
      tot += wc[i];
    }
    if (in) {
      const uint32_t r0 = base0 + off + uint32_t(__popcll(b & ((1ull << lane) - 1ull)));
      const uint32_t id = sd == 0 ? r0 : x - r0;
      const uint32_t o = orig ? orig[x] : x;
      newid[x] = id;
      if (sd == 0) { vw0[id] = vw[x]; orig0[id] = o; }
      else { vw1[id] = vw[x]; orig1[id] = o; }
      sw[sd] += vw[x];
    }
    __syncthreads();
    base0 += tot;
  }
  for (int q = 0; q < 2; ++q) {
    int64_t v = sw[q];
    for (int off = 32; off > 0; off >>= 1) v += int64_t(shfl_xor_u64((unsigned long long)v, off));
    if (lane == 0) ws[q][w] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t a0 = 0, a1 = 0;
    for (int i = 0; i < kIpWaves; ++i) { a0 += ws[0][i]; a1 += ws[1][i]; }
    sizes[0] = base0;
    sizes[1] = a0;
    sizes[2] = a1;
  }
}

// pins of every net per side; a side keeps the net when it holds two or more
__global__ void ip_net_count_kernel(const uint8_t* sides, const int32_t* best, uint32_t n, const uint32_t* noff,
                                    const uint32_t* pins, uint32_t m, uint32_t* cnt0, uint32_t* cnt1,
                                    uint32_t* keep0, uint32_t* keep1) {
  const uint8_t* side = sides + size_t(*best) * n;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    uint32_t c1 = 0;
    for (uint32_t j = noff[e]; j < noff[e + 1]; ++j) c1 += side[pins[j]];
    const uint32_t c0 = noff[e + 1] - noff[e] - c1;
    cnt0[e] = c0 >= 2 ? c0 : 0;
    cnt1[e] = c1 >= 2 ? c1 : 0;
    keep0[e] = c0 >= 2;
    keep1[e] = c1 >= 2;
  }
}

__global__ void ip_net_scatter_kernel(const uint8_t* sides, const int32_t* best, uint32_t n, const uint32_t* noff,
                                      const uint32_t* pins, const int32_t* nw, uint32_t m, const uint32_t* newid,
                                      const uint32_t* off0, const uint32_t* off1, const uint32_t* eid0,
                                      const uint32_t* eid1, uint32_t* noff0, uint32_t* noff1, uint32_t* pins0,
                                      uint32_t* pins1, int32_t* nw0, int32_t* nw1) {
  const uint8_t* side = sides + size_t(*best) * n;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const bool k0 = off0[e + 1] > off0[e], k1 = off1[e + 1] > off1[e];
    uint32_t d0 = off0[e], d1 = off1[e];
    if (k0) { noff0[eid0[e]] = d0; nw0[eid0[e]] = nw[e]; }
    if (k1) { noff1[eid1[e]] = d1; nw1[eid1[e]] = nw[e]; }
    for (uint32_t j = noff[e]; j < noff[e + 1]; ++j) {
      const uint32_t x = pins[j];
      if (side[x] == 0) { if (k0) pins0[d0++] = newid[x]; }
      else if (k1) pins1[d1++] = newid[x];
    }
  }
}

__global__ void ip_degree_kernel(const uint32_t* pins, uint64_t P, uint32_t* deg) {
  for (uint64_t j = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; j < P; j += uint64_t(gridDim.x) * blockDim.x)
    atomicAdd(&deg[pins[j]], 1u);
}
// incident nets in any order (every rule of a run reads them as a set)
__global__ void ip_incidence_kernel(const uint32_t* noff, const uint32_t* pins, uint32_t m, const uint32_t* voff,
                                    uint32_t* cursor, uint32_t* inc) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
    for (uint32_t j = noff[e]; j < noff[e + 1]; ++j) {
      const uint32_t x = pins[j];
      inc[voff[x] + atomicAdd(&cursor[x], 1u)] = e;
    }
}
// The same two steps for up to kIpLdsN vertices, counting in LDS: every
// workgroup takes one range of nets, counts its pins per vertex, reserves one
// range per vertex with a single atomic and places its nets from LDS
// cursors (a coarse level has few vertices, so per-pin atomics on the global
// counters would all contend).
__global__ __launch_bounds__(kIpThreads) void ip_degree_lds_kernel(const uint32_t* pins, uint64_t P, uint32_t n,
                                                                   uint32_t* deg) {
  __shared__ uint32_t h[kIpLdsN];
  for (uint32_t x = threadIdx.x; x < n; x += kIpThreads) h[x] = 0;
  __syncthreads();
  for (uint64_t j = blockIdx.x * uint64_t(kIpThreads) + threadIdx.x; j < P; j += uint64_t(gridDim.x) * kIpThreads)
    atomicAdd(&h[pins[j]], 1u);
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < n; x += kIpThreads)
    if (h[x]) atomicAdd(&deg[x], h[x]);
}
__global__ __launch_bounds__(kIpThreads) void ip_incidence_lds_kernel(const uint32_t* noff, const uint32_t* pins,
                                                                      uint32_t m, uint32_t n, const uint32_t* voff,
                                                                      uint32_t* cursor, uint32_t* inc) {
  __shared__ uint32_t h[kIpLdsN];
  const uint32_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const uint32_t e0 = min(m, blockIdx.x * chunk), e1 = min(m, e0 + chunk);
  for (uint32_t x = threadIdx.x; x < n; x += kIpThreads) h[x] = 0;
  __syncthreads();
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += kIpThreads)
    for (uint32_t j = noff[e]; j < noff[e + 1]; ++j) atomicAdd(&h[pins[j]], 1u);
  __syncthreads();
  for (uint32_t x = threadIdx.x; x < n; x += kIpThreads)
    if (h[x]) h[x] = voff[x] + atomicAdd(&cursor[x], h[x]);
  __syncthreads();
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += kIpThreads)
    for (uint32_t j = noff[e]; j < noff[e + 1]; ++j) inc[atomicAdd(&h[pins[j]], 1u)] = e;
}

__global__ void ip_assign_kernel(const uint32_t* orig, uint32_t n, int32_t block, int32_t* part) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x)
    part[orig ? orig[x] : x] = block;
}

// A sub-hypergraph of the recursion: the root borrows the coarsest
// hypergraph's arrays, the others own theirs.
struct IpSub {
  uint32_t n = 0, m = 0;
  uint32_t max_esize = 0;  // bound on the net sizes
  uint64_t P = 0;
  int64_t W = 0;
  const uint32_t *noff = nullptr, *pins = nullptr, *voff = nullptr, *inc = nullptr, *orig = nullptr;
  const int32_t *vw = nullptr, *nw = nullptr;
  DevBuf<uint32_t> b_noff, b_pins, b_voff, b_inc, b_orig;
  DevBuf<int32_t> b_vw, b_nw;
};

// Scratch of the runs of one bisection.
struct IpRuns {
  DevBuf<uint8_t> side, flag;
  DevBuf<int64_t> val, snap, out;
  DevBuf<uint32_t> list, pc;
  DevBuf<uint64_t> adj;
};

// Temporaries of splitting one sub-hypergraph into its halves.
struct IpSplit {
  DevBuf<uint32_t> newid, cnt0, cnt1, keep0, keep1, off0, off1, eid0, eid1;
  DevBuf<int64_t> sizes;
};

// One bisection of a recursion level.
struct IpTask {
  std::unique_ptr<IpSub> h;
  int k = 0, first = 0;
  int mode = 0;
  bool wide = false;
  IpArgs args{};
  IpRuns rb;
  IpSplit sp;
  std::unique_ptr<IpSub> half[2];
};

// Buffers and arguments of the runs of task t (and its clique scores).
void prepare(Context& c, IpTask& t, int64_t max0, int64_t max1, int64_t target0, uint64_t stream, int runs,
             int passes, bool use_adj) {
  const IpSub& h = *t.h;
  IpRuns& rb = t.rb;
  const size_t n = std::max<uint32_t>(h.n, 1), m = h.m;
  t.mode = h.n <= kIpLdsN ? 0 : (h.n <= kIpSideN ? 1 : 2);
  t.wide = h.max_esize > 0xFFFFu;
  rb.side.alloc(size_t(runs) * n);
  rb.snap.alloc(size_t(runs) * n);
  rb.list.alloc(size_t(runs) * n);
  rb.pc.alloc(std::max<size_t>(size_t(runs) * m * (t.wide ? 2 : 1), 1));
  rb.out.alloc(size_t(runs) * 4);
  rb.val.alloc(size_t(runs) * n);
  if (t.mode == 2) rb.flag.alloc(size_t(runs) * n);
  if (use_adj) {
    // clique scores, shared by the growing runs of this bisection
    rb.adj.alloc(n * n);
    TimedRegion tr(c.prof, c.stream, kFamInitial, 0.0, "ip_adjacency_kernel");
    ip_adjacency_kernel<<<unsigned(std::min<size_t>(n, 2048)), kIpThreads, 0, c.stream>>>(h.noff, h.pins, h.voff,
                                                                                       h.inc, h.nw, h.n, rb.adj.p);
    MKP_LAUNCH_CHECK();
  }
  t.args = IpArgs{h.noff, h.pins, h.voff, h.inc, h.vw, h.nw, h.n, h.m, h.W, max0, max1, target0, stream, passes,
                  rb.side.p, t.mode == 2 ? rb.flag.p : nullptr, rb.val.p, rb.snap.p, rb.list.p, rb.pc.p, rb.out.p,
                  use_adj ? rb.adj.p : nullptr};
}

template <int kMode, bool kWide>
void run_group(Context& c, const IpArgs* args, uint32_t count, uint32_t runs, uint32_t parts) {
  const std::string tag = "<" + std::to_string(kMode) + (kWide ? ", true>" : ", false>");
  {
    TimedRegion tr(c.prof, c.stream, kFamInitial, 0.0, std::string("ip_grow_kernel") + tag);
    ip_grow_kernel<kMode, kWide><<<count * runs, kIpThreads, 0, c.stream>>>(args, runs);
    MKP_LAUNCH_CHECK();
  }
  {
    TimedRegion tr(c.prof, c.stream, kFamInitial, 0.0, std::string("ip_gain_kernel") + tag);
    ip_gain_kernel<kMode, kWide><<<count * runs * parts, kIpThreads, 0, c.stream>>>(args, runs, parts);
    MKP_LAUNCH_CHECK();
  }
  {
    TimedRegion tr(c.prof, c.stream, kFamInitial, 0.0, std::string("ip_refine_kernel") + tag);
    ip_refine_kernel<kMode, kWide><<<count * runs, kIpThreads, 0, c.stream>>>(args, runs);
    MKP_LAUNCH_CHECK();
  }
}

// Runs every prepared bisection of `tasks` (reordered by state placement,
// which each launch shares) and leaves the index of each one's best run in
// best[i] for tasks[i].
void run_level(Context& c, std::vector<IpTask*>& tasks, int runs, DevBuf<IpArgs>& dargs, DevBuf<int32_t>& best) {
  std::stable_sort(tasks.begin(), tasks.end(), [](const IpTask* a, const IpTask* b) {
    return a->mode * 2 + int(a->wide) < b->mode * 2 + int(b->wide);
  });
  const size_t T = tasks.size();
  dargs.ensure(T);
  best.ensure(T);
  IpArgs* ha = c.staging<IpArgs>(T);
  uint64_t max_m = 0;
  for (size_t i = 0; i < T; ++i) {
    ha[i] = tasks[i]->args;
    max_m = std::max<uint64_t>(max_m, tasks[i]->h->m);
  }
  MKP_HIP(hipMemcpyAsync(dargs.p, ha, T * sizeof(IpArgs), hipMemcpyHostToDevice, c.stream));
  // enough workgroups per run for the pass over the nets to fill the chip
  const uint32_t R = uint32_t(runs);
  uint32_t parts = std::max<uint32_t>(1, uint32_t(2 * c.num_cus) / uint32_t(T * R));
  parts = std::min<uint32_t>(parts, uint32_t(std::max<uint64_t>(1, (max_m + 4095) / 4096)));
  for (size_t lo = 0; lo < T;) {
    size_t hi = lo;
    while (hi < T && tasks[hi]->mode == tasks[lo]->mode && tasks[hi]->wide == tasks[lo]->wide) ++hi;
    const IpArgs* ga = dargs.p + lo;
    const uint32_t cnt = uint32_t(hi - lo);
    const int mode = tasks[lo]->mode;
    const bool wide = tasks[lo]->wide;
    if (mode == 0) wide ? run_group<0, true>(c, ga, cnt, R, parts) : run_group<0, false>(c, ga, cnt, R, parts);
    else if (mode == 1) wide ? run_group<1, true>(c, ga, cnt, R, parts) : run_group<1, false>(c, ga, cnt, R, parts);
    else wide ? run_group<2, true>(c, ga, cnt, R, parts) : run_group<2, false>(c, ga, cnt, R, parts);
    lo = hi;
  }
  ip_best_kernel<<<unsigned(T), 64, 0, c.stream>>>(dargs.p, runs, best.p);
  MKP_LAUNCH_CHECK();
}

// The halves of every task's sub-hypergraph under its best run (one wait
// for all sizes).
void split_level(Context& c, std::vector<IpTask*>& tasks, const int32_t* best) {
  const hipStream_t st = c.stream;
  const size_t T = tasks.size();
  // sizes of task i: n0, W0, W1 then off0[m], off1[m], eid0[m], eid1[m]
  int64_t* hs = c.staging<int64_t>(8 * T);
  for (size_t i = 0; i < T; ++i) {
    IpTask& t = *tasks[i];
    const IpSub& h = *t.h;
    IpSplit& sp = t.sp;
    const uint32_t n = h.n, m = h.m;
    const size_t ms = std::max<uint32_t>(m, 1);
    t.half[0] = std::make_unique<IpSub>();
    t.half[1] = std::make_unique<IpSub>();
    IpSub &a = *t.half[0], &b = *t.half[1];
    sp.newid.alloc(std::max<uint32_t>(n, 1));
    a.b_vw.alloc(std::max<uint32_t>(n, 1));
    b.b_vw.alloc(std::max<uint32_t>(n, 1));
    a.b_orig.alloc(std::max<uint32_t>(n, 1));
    b.b_orig.alloc(std::max<uint32_t>(n, 1));
    sp.sizes.alloc(4);
    ip_split_vertices_kernel<<<1, kIpThreads, 0, st>>>(t.rb.side.p, best + i, n, h.vw, h.orig, sp.newid.p, a.b_vw.p,
                                                       b.b_vw.p, a.b_orig.p, b.b_orig.p, sp.sizes.p);
    MKP_LAUNCH_CHECK();
    sp.cnt0.alloc(ms); sp.cnt1.alloc(ms); sp.keep0.alloc(ms); sp.keep1.alloc(ms);
    sp.off0.alloc(size_t(m) + 1); sp.off1.alloc(size_t(m) + 1);
    sp.eid0.alloc(size_t(m) + 1); sp.eid1.alloc(size_t(m) + 1);
    if (m) {
      ip_net_count_kernel<<<grid_for(m, 256, 65536), 256, 0, st>>>(t.rb.side.p, best + i, n, h.noff, h.pins, m,
                                                                  sp.cnt0.p, sp.cnt1.p, sp.keep0.p, sp.keep1.p);
      MKP_LAUNCH_CHECK();
    }
    exclusive_scan_u32(c, sp.cnt0.p, sp.off0.p, m);
    exclusive_scan_u32(c, sp.cnt1.p, sp.off1.p, m);
    exclusive_scan_u32(c, sp.keep0.p, sp.eid0.p, m);
    exclusive_scan_u32(c, sp.keep1.p, sp.eid1.p, m);
    int64_t* hi = hs + 8 * i;
    uint32_t* hi32 = reinterpret_cast<uint32_t*>(hi + 4);
    MKP_HIP(hipMemcpyAsync(hi, sp.sizes.p, 3 * 8, hipMemcpyDeviceToHost, st));
    MKP_HIP(hipMemcpyAsync(hi32 + 0, sp.off0.p + m, 4, hipMemcpyDeviceToHost, st));
    MKP_HIP(hipMemcpyAsync(hi32 + 1, sp.off1.p + m, 4, hipMemcpyDeviceToHost, st));
    MKP_HIP(hipMemcpyAsync(hi32 + 2, sp.eid0.p + m, 4, hipMemcpyDeviceToHost, st));
    MKP_HIP(hipMemcpyAsync(hi32 + 3, sp.eid1.p + m, 4, hipMemcpyDeviceToHost, st));
  }
  c.sync();
  for (size_t i = 0; i < T; ++i) {
    IpTask& t = *tasks[i];
    const IpSub& h = *t.h;
    IpSplit& sp = t.sp;
    const uint32_t n = h.n, m = h.m;
    IpSub &a = *t.half[0], &b = *t.half[1];
    const int64_t* hi = hs + 8 * i;
    const uint32_t* hi32 = reinterpret_cast<const uint32_t*>(hi + 4);
    a.max_esize = b.max_esize = h.max_esize;
    a.n = uint32_t(hi[0]);
    b.n = n - a.n;
    a.W = hi[1];
    b.W = hi[2];
    a.P = hi32[0];
    b.P = hi32[1];
    a.m = hi32[2];
    b.m = hi32[3];
    for (IpSub* s : {&a, &b}) {
      s->b_noff.alloc(size_t(s->m) + 1);
      s->b_pins.alloc(std::max<uint64_t>(s->P, 1));
      s->b_nw.alloc(std::max<uint32_t>(s->m, 1));
      fill_async<uint32_t>(s->b_noff.p + s->m, uint32_t(s->P), 1, st);
    }
    if (m) {
      ip_net_scatter_kernel<<<grid_for(m, 256, 65536), 256, 0, st>>>(
          t.rb.side.p, best + i, n, h.noff, h.pins, h.nw, m, sp.newid.p, sp.off0.p, sp.off1.p, sp.eid0.p,
          sp.eid1.p, a.b_noff.p, b.b_noff.p, a.b_pins.p, b.b_pins.p, a.b_nw.p, b.b_nw.p);
      MKP_LAUNCH_CHECK();
    }
    for (IpSub* s : {&a, &b}) {
      s->b_voff.alloc(size_t(s->n) + 1);
      s->b_inc.alloc(std::max<uint64_t>(s->P, 1));
      DevBuf<uint32_t> deg(size_t(s->n) + 1);
      fill_async<uint32_t>(deg.p, 0u, size_t(s->n) + 1, st);
      const bool lds = s->n <= kIpLdsN;
      const unsigned lds_blocks = unsigned(std::min<uint64_t>(2 * c.num_cus, (s->P + 8191) / 8192 + 1));
      if (s->P) {
        if (lds) ip_degree_lds_kernel<<<lds_blocks, kIpThreads, 0, st>>>(s->b_pins.p, s->P, s->n, deg.p);
        else ip_degree_kernel<<<grid_for(s->P, 256, 65536), 256, 0, st>>>(s->b_pins.p, s->P, deg.p);
        MKP_LAUNCH_CHECK();
      }
      exclusive_scan_u32(c, deg.p, s->b_voff.p, s->n);
      fill_async<uint32_t>(deg.p, 0u, size_t(s->n) + 1, st);
      if (s->m) {
        if (lds)
          ip_incidence_lds_kernel<<<lds_blocks, kIpThreads, 0, st>>>(s->b_noff.p, s->b_pins.p, s->m, s->n,
                                                                     s->b_voff.p, deg.p, s->b_inc.p);
        else
          ip_incidence_kernel<<<grid_for(s->m, 256, 65536), 256, 0, st>>>(s->b_noff.p, s->b_pins.p, s->m,
                                                                           s->b_voff.p, deg.p, s->b_inc.p);
        MKP_LAUNCH_CHECK();
      }
      s->noff = s->b_noff.p;
      s->pins = s->b_pins.p;
      s->voff = s->b_voff.p;
      s->inc = s->b_inc.p;
      s->vw = s->b_vw.p;
      s->nw = s->b_nw.p;
      s->orig = s->b_orig.p;
    }
  }
}

std::unique_ptr<IpSub> root_of(mkp_hypergraph* hg) {
  auto r = std::make_unique<IpSub>();
  r->n = hg->n; r->m = hg->m; r->P = hg->p; r->W = hg->total_w; r->max_esize = hg->max_esize;
  r->noff = hg->noff.p; r->pins = hg->pins.p; r->voff = hg->voff.p; r->inc = hg->inc.p;
  r->vw = hg->vw.p; r->nw = hg->nw.p;
  return r;
}

}  // namespace

// Side limits of a bisection of weight W into ceil(k/2) and floor(k/2)
// blocks: (1 + eps') times each side's share, side 0 targets its share.
void ip_side_limits(int64_t W, int k, double eps_level, int64_t* max0, int64_t* max1, int64_t* target0) {
  const int k0 = (k + 1) / 2, k1 = k - k0;
  const double share = double(W) / double(k);
  *max0 = int64_t((1.0 + eps_level) * std::ceil(share * k0));
  *max1 = int64_t((1.0 + eps_level) * std::ceil(share * k1));
  *target0 = int64_t(std::ceil(share * k0));
}
uint64_t ip_stream(uint64_t seed, int first, int k) {
  return mix64(seed ^ mix64((uint64_t(uint32_t(first)) << 32) | uint32_t(k)));
}

void ip_bisect_runs(mkp_hypergraph* hg, int64_t max0, int64_t max1, int64_t target0, uint64_t stream, int runs,
                    int passes, uint8_t* sides, int64_t* results) {
  Context& c = hg->ctx->c;
  if (hg->n == 0) return;
  IpTask t;
  t.h = root_of(hg);
  prepare(c, t, max0, max1, target0, stream, runs, passes, hg->n <= kIpAdjN);
  std::vector<IpTask*> tasks{&t};
  DevBuf<IpArgs> dargs;
  DevBuf<int32_t> best;
  run_level(c, tasks, runs, dargs, best);
  MKP_HIP(hipMemcpyAsync(sides, t.rb.side.p, size_t(runs) * hg->n, hipMemcpyDeviceToHost, c.stream));
  MKP_HIP(hipMemcpyAsync(results, t.rb.out.p, size_t(runs) * 4 * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
}

void initial_partition(mkp_hypergraph* hg, int k, const std::vector<int64_t>& max_w, uint64_t seed, int runs,
                       int passes, int32_t* part) {
  Context& c = hg->ctx->c;
  const int64_t W = hg->total_w;
  const int64_t perfect = (W + k - 1) / std::max(k, 1);
  const double eps = perfect > 0 ? double(max_w[0]) / double(perfect) - 1.0 : 0.0;
  int levels = 0;
  while ((1 << levels) < k) ++levels;
  const double eps_level = levels > 0 ? std::pow(1.0 + std::max(eps, 0.0), 1.0 / levels) - 1.0 : eps;
  runs = std::max(1, runs);
  // the recursion one level at a time: every bisection of a level runs in
  // the same launches
  std::vector<std::unique_ptr<IpTask>> level;
  {
    auto t = std::make_unique<IpTask>();
    t->h = root_of(hg);
    t->k = k;
    t->first = 0;
    level.push_back(std::move(t));
  }
  DevBuf<IpArgs> dargs;
  DevBuf<int32_t> best;
  while (!level.empty()) {
    std::vector<IpTask*> active;
    // clique scores for the bisections of up to kIpAdjN vertices while the
    // level's matrices fit kIpAdjBytes (the others grow by walking the nets,
    // with the same result)
    uint64_t adj_bytes = 0;
    for (auto& t : level) {
      const IpSub& h = *t->h;
      if (h.n == 0) continue;
      if (t->k == 1) {
        ip_assign_kernel<<<grid_for(h.n, 256, 65536), 256, 0, c.stream>>>(h.orig, h.n, t->first, part);
        MKP_LAUNCH_CHECK();
        continue;
      }
      int64_t max0, max1, target0;
      ip_side_limits(h.W, t->k, eps_level, &max0, &max1, &target0);
      const uint64_t ab = uint64_t(h.n) * h.n * 8;
      const bool use_adj = h.n <= kIpAdjN && adj_bytes + ab <= kIpAdjBytes;
      if (use_adj) adj_bytes += ab;
      prepare(c, *t, max0, max1, target0, ip_stream(seed, t->first, t->k), runs, passes, use_adj);
      active.push_back(t.get());
    }
    if (active.empty()) break;
    run_level(c, active, runs, dargs, best);
    split_level(c, active, best.p);
    std::vector<std::unique_ptr<IpTask>> next;
    for (IpTask* t : active) {
      const int k0 = (t->k + 1) / 2, k1 = t->k - k0;
      for (int s2 = 0; s2 < 2; ++s2) {
        auto ch = std::make_unique<IpTask>();
        ch->h = std::move(t->half[s2]);
        ch->k = s2 == 0 ? k0 : k1;
        ch->first = s2 == 0 ? t->first : t->first + k0;
        next.push_back(std::move(ch));
      }
    }
    level.swap(next);  // the finished level's buffers go back to the allocation cache
  }
  // a block still over its bound gives its lightest vertices to the
  // lightest blocks they fit (host: the coarsest hypergraph is small)
  std::vector<int32_t> pt(hg->n), vw(hg->n);
  MKP_HIP(hipMemcpyAsync(pt.data(), part, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
  MKP_HIP(hipMemcpyAsync(vw.data(), hg->vw.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  std::vector<int64_t> pw(k, 0);
  for (uint32_t u = 0; u < hg->n; ++u) pw[pt[u]] += vw[u];
  bool changed = false;
  for (int bl = 0; bl < k; ++bl) {
    if (pw[bl] <= max_w[bl]) continue;
    std::vector<std::pair<int32_t, uint32_t>> vs;
    for (uint32_t u = 0; u < hg->n; ++u)
      if (pt[u] == bl) vs.push_back({vw[u], u});
    std::sort(vs.begin(), vs.end());
    for (const auto& x : vs) {
      if (pw[bl] <= max_w[bl]) break;
      int t = -1;
      for (int q = 0; q < k; ++q)
        if (q != bl && pw[q] + x.first <= max_w[q] && (t < 0 || pw[q] < pw[t])) t = q;
      if (t < 0) continue;
      pt[x.second] = t;
      pw[bl] -= x.first;
      pw[t] += x.first;
      changed = true;
    }
  }
  if (changed) {
    MKP_HIP(hipMemcpyAsync(part, pt.data(), size_t(hg->n) * 4, hipMemcpyHostToDevice, c.stream));
    c.sync();
  }
}

}  // namespace mkp
