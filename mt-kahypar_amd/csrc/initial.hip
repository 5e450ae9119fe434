// Initial partitioning of the coarsest hypergraph on the device.
//
// Recursive bisection (the recursion of
// recursive_bisection_initial_partitioner.cpp:475-492; the flat pool of
// flat/pool_initial_partitioner.h:36-75 is replaced by a portfolio of runs).
// Every bisection makes all of its runs at once, one workgroup of 1024
// threads per run, and keeps the balanced run with the smallest cut (ties to
// the smaller run index).  Run r grows side 0 by variant r % 3 (greedy
// growing, level-synchronous BFS, random fill) and then improves the
// bisection with passes of 2-way gain moves.
//
// The steps of a run are sequential -- growing one vertex at a time and
// moving one vertex at a time is what makes them effective -- and the work
// inside a step (a vertex's nets and their pins, the arg-max over the
// vertices, the next positive-gain vertex of a pass) is spread over the
// workgroup.  Growing scores and gains are integers updated by atomics, so a
// run does not depend on the order of the updates.  For up to kIpLdsN
// vertices the per-vertex state (side, flags, score or gain) lives in LDS.
//
// The rules, each stated so that the result is one function of the input:
//   * greedy growing: score of a side-1 vertex = sum over its nets of at most
//     100 pins shared with side 0 of floor(w(e) * 2^16 / (|e| - 1)) (64-bit
//     wrap-around); the highest score joins, ties to the smaller id; with no
//     scored vertex the next free vertex after a hashed start vertex seeds;
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
constexpr uint32_t kIpLdsN = 8192;         // per-vertex state in LDS up to this many vertices
constexpr uint32_t kIpBigNet = 64;         // nets of a move with more pins are updated by a whole wave
constexpr uint32_t kIpQueue = 512;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kIpStartTag = 0x243F6A8885A308D3ull;
constexpr uint64_t kIpFillTag = 0x13198A2E03707344ull;
constexpr uint64_t kIpPassTag = 0xA4093822299F31D0ull;

// per-vertex flag bits
constexpr uint8_t kSkip = 1;  // growing: too heavy for side 0
constexpr uint8_t kHeap = 2;  // growing: scored and not taken
constexpr uint8_t kSeen = 4;  // BFS: reached
constexpr uint8_t kNext = 8;  // BFS: in the next level
constexpr uint8_t kCand = 16; // repair: candidate not yet visited

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
  // per-run buffers, run r at r * n (r * 2m for pc, r * 4 for out)
  uint8_t* side;   // the run's bisection
  uint8_t* flag;   // beyond kIpLdsN vertices only
  int64_t* val;    // beyond kIpLdsN vertices only
  int64_t* snap;
  uint32_t* list;
  uint32_t* pc;    // pins of net e on side 0 / 1 at 2e / 2e+1
  int64_t* out;    // cut, balanced, side-0 weight, side-1 weight
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
__device__ inline void ip_move(const IpArgs& a, IpShared& s, uint8_t* side, int64_t* val, uint32_t* pc, uint32_t v) {
  const int f = side[v], t = 1 - f;
  const int64_t gv = val[v];
  const uint32_t vb = a.voff[v], ve = a.voff[v + 1];
  for (uint32_t i = vb + threadIdx.x; i < ve; i += kIpThreads) {
    const uint32_t e = a.inc[i];
    const uint32_t ca = pc[2 * e + f], cb = pc[2 * e + t];
    const int64_t we = a.nw[e];
    const int64_t df = we * (int64_t(ca == 2) - int64_t(ca == 1) + int64_t(cb == 0));
    const int64_t dt = we * (int64_t(cb == 0) - int64_t(ca == 1) - int64_t(cb == 1));
    if (df || dt) {
      const uint32_t jb = a.noff[e], je = a.noff[e + 1];
      uint32_t q = kNone;
      if (je - jb > kIpBigNet) {
        q = atomicAdd(&s.qn, 1u);
        if (q < kIpQueue) {
          s.qe[q] = e;
          s.qdf[q] = df;
          s.qdt[q] = dt;
        }
      }
      if (q == kNone || q >= kIpQueue)
        for (uint32_t j = jb; j < je; ++j) {
          const uint32_t x = a.pins[j];
          if (x != v) atomicAdd(reinterpret_cast<unsigned long long*>(&val[x]), (unsigned long long)(side[x] == f ? df : dt));
        }
    }
    pc[2 * e + f] = ca - 1;
    pc[2 * e + t] = cb + 1;
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

template <bool kLds>
__global__ __launch_bounds__(kIpThreads) void ip_run_kernel(IpArgs a) {
  __shared__ IpShared s;
  __shared__ uint8_t l_side[kLds ? kIpLdsN : 1];
  __shared__ uint8_t l_flag[kLds ? kIpLdsN : 1];
  __shared__ int64_t l_val[kLds ? kIpLdsN : 1];
  const int r = blockIdx.x;
  const uint32_t n = a.n, m = a.m;
  const uint32_t tid = threadIdx.x;
  uint8_t* g_side = a.side + size_t(r) * n;
  uint8_t* side = kLds ? l_side : g_side;
  uint8_t* flag = kLds ? l_flag : a.flag + size_t(r) * n;
  int64_t* val = kLds ? l_val : a.val + size_t(r) * n;
  int64_t* snap = a.snap + size_t(r) * n;
  uint32_t* list = a.list + size_t(r) * n;
  uint32_t* pc = a.pc + size_t(r) * 2 * m;
  const uint64_t rk = ip_run_key(a.stream, r);
  const int variant = r % 3;
  const uint32_t start = uint32_t(mix64(rk ^ kIpStartTag) % n);
  for (uint32_t x = tid; x < n; x += kIpThreads) {
    side[x] = 1;
    flag[x] = 0;
    val[x] = 0;
  }
  if (variant == 1)
    for (uint32_t e = tid; e < m; e += kIpThreads) pc[2 * e] = 0;  // scanned marks
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
      uint64_t key = 0;
      uint32_t v = kNone;
      for (uint32_t x = tid; x < n; x += kIpThreads)
        if (flag[x] & kHeap) {
          const uint64_t k = uint64_t(val[x]);
          if (v == kNone || k > key) { key = k; v = x; }
        }
      block_argmax(s, key, v);
      if (v == kNone) {  // a new seed: the next free vertex after `start`
        const uint32_t sc = s.scan;
        uint32_t first = kNone;
        for (uint32_t i = sc + tid; i < n; i += kIpThreads) {
          const uint32_t x = (start + i) % n;
          if (side[x] == 1 && !(flag[x] & kSkip)) { first = i; break; }
        }
        first = block_min(s, first);
        if (first == kNone) break;
        v = (start + first) % n;
        if (tid == 0) s.scan = first + 1;
      }
      const bool fits = s.w[0] + a.vw[v] <= a.max0;
      __syncthreads();
      if (tid == 0) {
        flag[v] = uint8_t((flag[v] & ~kHeap) | (fits ? 0 : kSkip));
        if (fits) {
          side[v] = 0;
          s.w[0] += a.vw[v];
        }
      }
      __syncthreads();
      if (!fits) continue;
      for (uint32_t i = a.voff[v] + tid; i < a.voff[v + 1]; i += kIpThreads) {
        const uint32_t e = a.inc[i];
        const uint32_t jb = a.noff[e], je = a.noff[e + 1];
        if (je - jb > kIpGrowNetLimit) continue;
        const uint64_t add = (uint64_t(uint32_t(a.nw[e])) << kIpScoreFrac) / uint64_t(max(je - jb - 1, 1u));
        for (uint32_t j = jb; j < je; ++j) {
          const uint32_t x = a.pins[j];
          const uint8_t fl = flag[x];
          if (side[x] == 1 && !(fl & kSkip)) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&val[x]), (unsigned long long)add);
            if (!(fl & kHeap)) flag[x] = uint8_t(fl | kHeap);
          }
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
          if (pc[2 * e] || je - jb > kIpGrowNetLimit) continue;
          pc[2 * e] = 1;
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

  // ------------------------------------------------------- refinement --
  for (uint32_t e = tid; e < m; e += kIpThreads) {
    const uint32_t jb = a.noff[e], je = a.noff[e + 1];
    uint32_t c1 = 0;
    for (uint32_t j = jb; j < je; ++j) c1 += side[a.pins[j]];
    pc[2 * e] = (je - jb) - c1;
    pc[2 * e + 1] = c1;
  }
  if (tid == 0) s.w[1] = a.W - s.w[0];
  __syncthreads();
  {
    const int w = tid >> 6, lane = tid & 63;
    for (uint32_t u = w; u < n; u += kIpWaves) {
      const int f = side[u], t = 1 - f;
      int64_t g = 0;
      for (uint32_t i = a.voff[u] + lane; i < a.voff[u + 1]; i += 64) {
        const uint32_t e = a.inc[i];
        const int64_t we = a.nw[e];
        g += (pc[2 * e + f] == 1 ? we : 0) - (pc[2 * e + t] == 0 ? we : 0);
      }
      for (int off = 32; off > 0; off >>= 1) g += int64_t(shfl_xor_u64((unsigned long long)g, off));
      if (lane == 0) val[u] = g;
    }
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
      for (uint32_t x = tid; x < n; x += kIpThreads)
        if (flag[x] & kCand) {
          const uint64_t k = uint64_t(snap[x]) ^ (1ull << 63);
          if (c == kNone || k > key) { key = k; c = x; }
        }
      block_argmax(s, key, c);
      if (c == kNone) break;
      const bool fits = s.w[1 - f] + a.vw[c] <= mt;
      __syncthreads();
      if (tid == 0) flag[c] = uint8_t(flag[c] & ~kCand);
      __syncthreads();
      if (fits) ip_move(a, s, side, val, pc, c);
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
      ip_move(a, s, side, val, pc, u);
      pos = hit + 1;
    }
    if (!improved) break;
  }

  int64_t cut = 0;
  for (uint32_t e = tid; e < m; e += kIpThreads)
    if (pc[2 * e] && pc[2 * e + 1]) cut += a.nw[e];
  cut = block_sum(s, cut);
  if (kLds)
    for (uint32_t x = tid; x < n; x += kIpThreads) g_side[x] = side[x];
  if (tid == 0) {
    int64_t* o = a.out + 4 * size_t(r);
    o[0] = cut;
    o[1] = s.w[0] <= a.max0 && s.w[1] <= a.max1;
    o[2] = s.w[0];
    o[3] = s.w[1];
  }
}

// index of the balanced run with the smallest cut (ties: the smaller index)
__global__ void ip_best_kernel(const int64_t* out, int runs, int32_t* best) {
  if (threadIdx.x != 0) return;
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
__global__ void ip_assign_kernel(const uint32_t* orig, uint32_t n, int32_t block, int32_t* part) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < n; x += gridDim.x * blockDim.x)
    part[orig ? orig[x] : x] = block;
}

// A sub-hypergraph of the recursion: the root borrows the coarsest
// hypergraph's arrays, the others own theirs.
struct IpSub {
  uint32_t n = 0, m = 0;
  uint64_t P = 0;
  int64_t W = 0;
  const uint32_t *noff = nullptr, *pins = nullptr, *voff = nullptr, *inc = nullptr, *orig = nullptr;
  const int32_t *vw = nullptr, *nw = nullptr;
  DevBuf<uint32_t> b_noff, b_pins, b_voff, b_inc, b_orig;
  DevBuf<int32_t> b_vw, b_nw;
};

// Scratch of one bisection's runs.
struct IpRuns {
  DevBuf<uint8_t> side, flag;
  DevBuf<int64_t> val, snap, out;
  DevBuf<uint32_t> list, pc;
  DevBuf<int32_t> best;
};

void launch_runs(Context& c, const IpSub& h, int64_t max0, int64_t max1, int64_t target0, uint64_t stream, int runs,
                 int passes, IpRuns& rb) {
  const size_t n = std::max<uint32_t>(h.n, 1), m = h.m;
  const bool lds = h.n <= kIpLdsN;
  rb.side.ensure(size_t(runs) * n);
  rb.snap.ensure(size_t(runs) * n);
  rb.list.ensure(size_t(runs) * n);
  rb.pc.ensure(std::max<size_t>(size_t(runs) * 2 * m, 1));
  rb.out.ensure(size_t(runs) * 4);
  rb.best.ensure(1);
  if (!lds) {
    rb.flag.ensure(size_t(runs) * n);
    rb.val.ensure(size_t(runs) * n);
  }
  IpArgs a{h.noff, h.pins, h.voff, h.inc, h.vw, h.nw, h.n, h.m, h.W, max0, max1, target0, stream, passes,
           rb.side.p, lds ? nullptr : rb.flag.p, lds ? nullptr : rb.val.p, rb.snap.p, rb.list.p, rb.pc.p, rb.out.p};
  {
    TimedRegion tr(c.prof, c.stream, kFamInitial, 0.0, lds ? "ip_run_kernel<true>" : "ip_run_kernel<false>");
    if (lds) ip_run_kernel<true><<<runs, kIpThreads, 0, c.stream>>>(a);
    else ip_run_kernel<false><<<runs, kIpThreads, 0, c.stream>>>(a);
    MKP_LAUNCH_CHECK();
  }
  ip_best_kernel<<<1, 64, 0, c.stream>>>(rb.out.p, runs, rb.best.p);
  MKP_LAUNCH_CHECK();
}

// the two halves of h under the best run's bisection
void split(Context& c, const IpSub& h, const IpRuns& rb, IpSub& a, IpSub& b) {
  const hipStream_t st = c.stream;
  const uint32_t n = h.n, m = h.m;
  DevBuf<uint32_t> newid(std::max<uint32_t>(n, 1));
  a.b_vw.alloc(std::max<uint32_t>(n, 1));
  b.b_vw.alloc(std::max<uint32_t>(n, 1));
  a.b_orig.alloc(std::max<uint32_t>(n, 1));
  b.b_orig.alloc(std::max<uint32_t>(n, 1));
  DevBuf<int64_t> sizes(4);
  ip_split_vertices_kernel<<<1, kIpThreads, 0, st>>>(rb.side.p, rb.best.p, n, h.vw, h.orig, newid.p, a.b_vw.p,
                                                     b.b_vw.p, a.b_orig.p, b.b_orig.p, sizes.p);
  MKP_LAUNCH_CHECK();
  const size_t ms = std::max<uint32_t>(m, 1);
  DevBuf<uint32_t> cnt0(ms), cnt1(ms), keep0(ms), keep1(ms), off0(size_t(m) + 1), off1(size_t(m) + 1),
      eid0(size_t(m) + 1), eid1(size_t(m) + 1);
  if (m) {
    ip_net_count_kernel<<<grid_for(m, 256, 65536), 256, 0, st>>>(rb.side.p, rb.best.p, n, h.noff, h.pins, m, cnt0.p,
                                                                cnt1.p, keep0.p, keep1.p);
    MKP_LAUNCH_CHECK();
  }
  exclusive_scan_u32(c, cnt0.p, off0.p, m);
  exclusive_scan_u32(c, cnt1.p, off1.p, m);
  exclusive_scan_u32(c, keep0.p, eid0.p, m);
  exclusive_scan_u32(c, keep1.p, eid1.p, m);
  int64_t* hs = c.staging<int64_t>(8);
  uint32_t* hs32 = reinterpret_cast<uint32_t*>(hs + 4);
  MKP_HIP(hipMemcpyAsync(hs, sizes.p, 3 * 8, hipMemcpyDeviceToHost, st));
  MKP_HIP(hipMemcpyAsync(hs32 + 0, off0.p + m, 4, hipMemcpyDeviceToHost, st));
  MKP_HIP(hipMemcpyAsync(hs32 + 1, off1.p + m, 4, hipMemcpyDeviceToHost, st));
  MKP_HIP(hipMemcpyAsync(hs32 + 2, eid0.p + m, 4, hipMemcpyDeviceToHost, st));
  MKP_HIP(hipMemcpyAsync(hs32 + 3, eid1.p + m, 4, hipMemcpyDeviceToHost, st));
  c.sync();
  a.n = uint32_t(hs[0]);
  b.n = n - a.n;
  a.W = hs[1];
  b.W = hs[2];
  a.P = hs32[0];
  b.P = hs32[1];
  a.m = hs32[2];
  b.m = hs32[3];
  for (IpSub* s : {&a, &b}) {
    s->b_noff.alloc(size_t(s->m) + 1);
    s->b_pins.alloc(std::max<uint64_t>(s->P, 1));
    s->b_nw.alloc(std::max<uint32_t>(s->m, 1));
    fill_async<uint32_t>(s->b_noff.p + s->m, uint32_t(s->P), 1, st);
  }
  if (m) {
    ip_net_scatter_kernel<<<grid_for(m, 256, 65536), 256, 0, st>>>(
        rb.side.p, rb.best.p, n, h.noff, h.pins, h.nw, m, newid.p, off0.p, off1.p, eid0.p, eid1.p, a.b_noff.p,
        b.b_noff.p, a.b_pins.p, b.b_pins.p, a.b_nw.p, b.b_nw.p);
    MKP_LAUNCH_CHECK();
  }
  for (IpSub* s : {&a, &b}) {
    s->b_voff.alloc(size_t(s->n) + 1);
    s->b_inc.alloc(std::max<uint64_t>(s->P, 1));
    DevBuf<uint32_t> deg(size_t(s->n) + 1);
    fill_async<uint32_t>(deg.p, 0u, size_t(s->n) + 1, st);
    if (s->P) {
      ip_degree_kernel<<<grid_for(s->P, 256, 65536), 256, 0, st>>>(s->b_pins.p, s->P, deg.p);
      MKP_LAUNCH_CHECK();
    }
    exclusive_scan_u32(c, deg.p, s->b_voff.p, s->n);
    fill_async<uint32_t>(deg.p, 0u, size_t(s->n) + 1, st);
    if (s->m) {
      ip_incidence_kernel<<<grid_for(s->m, 256, 65536), 256, 0, st>>>(s->b_noff.p, s->b_pins.p, s->m, s->b_voff.p,
                                                                       deg.p, s->b_inc.p);
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
  IpSub h;
  h.n = hg->n; h.m = hg->m; h.P = hg->p; h.W = hg->total_w;
  h.noff = hg->noff.p; h.pins = hg->pins.p; h.voff = hg->voff.p; h.inc = hg->inc.p;
  h.vw = hg->vw.p; h.nw = hg->nw.p;
  if (h.n == 0) return;
  IpRuns rb;
  launch_runs(c, h, max0, max1, target0, stream, runs, passes, rb);
  MKP_HIP(hipMemcpyAsync(sides, rb.side.p, size_t(runs) * h.n, hipMemcpyDeviceToHost, c.stream));
  MKP_HIP(hipMemcpyAsync(results, rb.out.p, size_t(runs) * 4 * 8, hipMemcpyDeviceToHost, c.stream));
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
  struct Node {
    std::unique_ptr<IpSub> h;
    int k, first;
  };
  std::vector<Node> stack;
  {
    auto root = std::make_unique<IpSub>();
    root->n = hg->n; root->m = hg->m; root->P = hg->p; root->W = W;
    root->noff = hg->noff.p; root->pins = hg->pins.p; root->voff = hg->voff.p; root->inc = hg->inc.p;
    root->vw = hg->vw.p; root->nw = hg->nw.p;
    stack.push_back(Node{std::move(root), k, 0});
  }
  IpRuns rb;
  while (!stack.empty()) {
    Node nd = std::move(stack.back());
    stack.pop_back();
    const IpSub& h = *nd.h;
    if (h.n == 0) continue;
    if (nd.k == 1) {
      ip_assign_kernel<<<grid_for(h.n, 256, 65536), 256, 0, c.stream>>>(h.orig, h.n, nd.first, part);
      MKP_LAUNCH_CHECK();
      continue;
    }
    const int k0 = (nd.k + 1) / 2, k1 = nd.k - k0;
    int64_t max0, max1, target0;
    ip_side_limits(h.W, nd.k, eps_level, &max0, &max1, &target0);
    launch_runs(c, h, max0, max1, target0, ip_stream(seed, nd.first, nd.k), std::max(1, runs), passes, rb);
    auto a = std::make_unique<IpSub>(), b = std::make_unique<IpSub>();
    split(c, h, rb, *a, *b);
    stack.push_back(Node{std::move(b), k1, nd.first + k0});
    stack.push_back(Node{std::move(a), k0, nd.first});
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
