// Partition state: block weights, packed pin counts, connectivity sets, km1.
//
// initializePinCountInPart (partitioned_hypergraph.h:1086-1110) is the pure
// bandwidth kernel of the path: one coalesced sweep over the pins of a tile of
// nets, a gather of the pin's block id, and LDS 64-bit atomics into the tile's
// packed counters; the tile is then written out once, together with the
// connectivity words and the tile's km1 contribution.
#include "internal.h"

namespace mkp {

constexpr int kInitThreads = 256;

__device__ inline uint32_t find_net(const uint32_t* off, int count, uint32_t pos) {
  // largest i in [0, count) with off[i] <= pos
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= pos) lo = mid; else hi = mid;
  }
  return uint32_t(lo);
}

__global__ void __launch_bounds__(kInitThreads)
pin_count_init_kernel(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ pins,
                      const int32_t* __restrict__ part, const int32_t* __restrict__ nw, uint32_t m, int k,
                      PcLayout L, int vcs, int nets_per_block, uint64_t* __restrict__ rec,
                      int64_t* __restrict__ km1_acc) {
  extern __shared__ uint64_t smem[];
  uint64_t* words = smem;                                                   // nets_per_block * vpc
  uint32_t* off = reinterpret_cast<uint32_t*>(words + size_t(nets_per_block) * L.vpc);  // nets_per_block + 1
  __shared__ int64_t red[kInitThreads / 64];
  const uint32_t e0 = blockIdx.x * uint32_t(nets_per_block);
  const int ne = int(min<uint32_t>(uint32_t(nets_per_block), m - e0));
  const int nwords = ne * L.vpc;
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) words[i] = 0;
  for (int i = threadIdx.x; i <= ne; i += blockDim.x) off[i] = noff[e0 + i];
  __syncthreads();
  const uint32_t pbeg = off[0], pend = off[ne];
  for (uint32_t q = pbeg + threadIdx.x; q < pend; q += blockDim.x) {
    const uint32_t le = find_net(off, ne, q);
    const int blk = part[pins[q]];
    const int wi = blk / L.epw;
    const int sh = (blk - wi * L.epw) * L.b;
    atomicAdd(reinterpret_cast<unsigned long long*>(&words[le * L.vpc + wi]), 1ull << sh);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nwords; i += blockDim.x) {
    const int le = i / L.vpc;
    rec[uint64_t(e0 + le) * L.stride + (i - le * L.vpc)] = words[i];
  }
  int64_t local = 0;
  for (int le = threadIdx.x; le < ne; le += blockDim.x) {
    int lambda = 0;
    for (int ci = 0; ci < vcs; ++ci) {
      uint64_t bits = 0;
      const int bbeg = ci * 64, bend = min(k, bbeg + 64);
      for (int blk = bbeg; blk < bend; ++blk) {
        const int wi = blk / L.epw;
        const int sh = (blk - wi * L.epw) * L.b;
        if ((words[le * L.vpc + wi] >> sh) & L.mask) bits |= 1ull << (blk - bbeg);
      }
      rec[uint64_t(e0 + le) * L.stride + L.coff + ci] = bits;
      lambda += __popcll(bits);
    }
    rec[uint64_t(e0 + le) * L.stride + L.woff] = uint64_t(int64_t(nw[e0 + le]));
    local += int64_t(max(lambda - 1, 0)) * nw[e0 + le];
  }
  for (int o = 32; o > 0; o >>= 1) local += __shfl_down(local, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0 && km1_acc) {
    int64_t s = 0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
    if (s) atomicAdd(reinterpret_cast<unsigned long long*>(km1_acc), static_cast<unsigned long long>(s));
  }
}

__global__ void block_weight_kernel(const int32_t* __restrict__ part, const int32_t* __restrict__ vw, uint32_t n,
                                    int k, int64_t* __restrict__ part_w) {
  extern __shared__ int64_t bins[];
  for (int i = threadIdx.x; i < k; i += blockDim.x) bins[i] = 0;
  __syncthreads();
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    atomicAdd(reinterpret_cast<unsigned long long*>(&bins[part[u]]), static_cast<unsigned long long>(vw[u]));
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x)
    if (bins[i]) atomicAdd(reinterpret_cast<unsigned long long*>(&part_w[i]), static_cast<unsigned long long>(bins[i]));
}

// km1 per block into acc[blockIdx.x] (no atomics: the host sums the
// partials of the at most kKm1Blocks blocks)
constexpr unsigned kKm1Blocks = 512;
__global__ void __launch_bounds__(256)
km1_kernel(const uint64_t* __restrict__ rec, const int32_t* __restrict__ nw, uint32_t m, PcLayout L,
           int64_t* __restrict__ acc) {
  __shared__ int64_t red[4];
  int64_t local = 0;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    int lambda = 0;
    for (int ci = 0; ci < L.vcs; ++ci) lambda += __popcll(rec[uint64_t(e) * L.stride + L.coff + ci]);
    local += int64_t(max(lambda - 1, 0)) * nw[e];
  }
  for (int o = 32; o > 0; o >>= 1) local += __shfl_down(local, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0) acc[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void unpack_pc_kernel(const uint64_t* __restrict__ pc, uint32_t m, int k, PcLayout L,
                                 uint32_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < uint64_t(m) * k;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t e = i / k;
    const int blk = int(i - e * k);
    const int wi = blk / L.epw;
    const int sh = (blk - wi * L.epw) * L.b;
    out[i] = uint32_t((pc[e * L.stride + wi] >> sh) & L.mask);
  }
}

// connectivity words of every net, unpacked from the records
__global__ void conn_words_kernel(const uint64_t* __restrict__ rec, uint32_t m, PcLayout L,
                                  uint64_t* __restrict__ out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < uint64_t(m) * L.vcs;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t e = i / L.vcs;
    out[i] = rec[e * L.stride + L.coff + (i - e * L.vcs)];
  }
}

__global__ void border_kernel(const uint32_t* __restrict__ voff, const uint32_t* __restrict__ inc,
                              const uint64_t* __restrict__ rec, uint32_t n, PcLayout L, uint8_t* __restrict__ out) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    const uint32_t b = voff[u], e = voff[u + 1];
    uint8_t border = 0;
    if (e - b <= 100000u) {
      for (uint32_t j = b; j < e && !border; ++j) {
        int lambda = 0;
        for (int ci = 0; ci < L.vcs; ++ci) lambda += __popcll(rec[uint64_t(inc[j]) * L.stride + L.coff + ci]);
        border = lambda > 1;
      }
    }
    out[u] = border;
  }
}

static int nets_per_block(const PcLayout& L) {
  int npb = 256;
  while (npb > 16 && size_t(npb) * L.vpc * 8 + (npb + 1) * 4 > 64 * 1024) npb >>= 1;
  return npb;
}

void partition_initialize(mkp_partition* p) {
  mkp_hypergraph* hg = p->hg;
  Context& c = hg->ctx->c;
  fill_async<int64_t>(p->part_w.p, 0, p->k, c.stream);
  // algorithmic bytes: block id + weight per vertex and k weights (block
  // weights); net offsets, pin id + block id per pin, packed counters and
  // connectivity written once (pin counts)
  if (hg->n) {
    TimedRegion tr(c.prof, c.stream, kFamPinCountInit, 8.0 * hg->n + 4.0 * p->k, "block_weight_kernel");
    block_weight_kernel<<<grid_for(hg->n, 256, 256), 256, p->k * sizeof(int64_t), c.stream>>>(
        p->part.p, hg->vw.p, hg->n, p->k, p->part_w.p);
    MKP_LAUNCH_CHECK();
  }
  if (hg->m) {
    TimedRegion tr(c.prof, c.stream, kFamPinCountInit,
                   4.0 * (hg->m + 1) + 8.0 * double(hg->p) + 8.0 * hg->m * (p->L.vpc + p->vcs + 1),
                   "pin_count_init_kernel");
    const int npb = nets_per_block(p->L);
    const size_t lds = size_t(npb) * p->L.vpc * 8 + size_t(npb + 1) * 4;
    pin_count_init_kernel<<<(hg->m + npb - 1) / npb, kInitThreads, lds, c.stream>>>(
        hg->noff.p, hg->pins.p, p->part.p, hg->nw.p, hg->m, p->k, p->L, p->vcs, npb, p->pc.p, nullptr);
    MKP_LAUNCH_CHECK();
  }
}

// ------------------------------------------------------------- batches --
// Members of a batch are disjoint parts of one hypergraph with local block
// ids; block weights go to the member's own range of the weight array.
// A block takes kBwChunk consecutive vertices; they belong to few members
// (members are contiguous vertex ranges), so the blocks they touch form a
// short range of the weight array, summed in LDS and flushed once.
constexpr uint32_t kBwChunk = 2048, kBwBins = 2048;

__global__ void __launch_bounds__(256)
block_weight_batch_kernel(const int32_t* __restrict__ part, const int32_t* __restrict__ vw, uint32_t n,
                          const uint32_t* __restrict__ vgraph, const uint32_t* __restrict__ gkbase,
                          int64_t* __restrict__ part_w) {
  __shared__ int64_t bins[kBwBins];
  for (uint32_t c0 = blockIdx.x * kBwChunk; c0 < n; c0 += gridDim.x * kBwChunk) {
    const uint32_t c1 = min(c0 + kBwChunk, n);
    const uint32_t lo = gkbase[vgraph[c0]], hi = gkbase[vgraph[c1 - 1] + 1];
    const bool in_lds = hi - lo <= kBwBins;
    if (in_lds)
      for (uint32_t i = threadIdx.x; i < hi - lo; i += blockDim.x) bins[i] = 0;
    __syncthreads();
    for (uint32_t u = c0 + threadIdx.x; u < c1; u += blockDim.x) {
      const uint32_t key = gkbase[vgraph[u]] + uint32_t(part[u]);
      if (in_lds)
        atomicAdd(reinterpret_cast<unsigned long long*>(&bins[key - lo]), static_cast<unsigned long long>(int64_t(vw[u])));
      else
        atomicAdd(reinterpret_cast<unsigned long long*>(&part_w[key]), static_cast<unsigned long long>(int64_t(vw[u])));
    }
    __syncthreads();
    if (in_lds)
      for (uint32_t i = threadIdx.x; i < hi - lo; i += blockDim.x)
        if (bins[i]) atomicAdd(reinterpret_cast<unsigned long long*>(&part_w[lo + i]), static_cast<unsigned long long>(bins[i]));
    __syncthreads();
  }
}

// km1 per member: nets of a member are contiguous, so the lanes of a wave
// mostly share the member and add their sum with one atomic
__global__ void km1_batch_kernel(const uint64_t* __restrict__ rec, const int32_t* __restrict__ nw, uint32_t m,
                                 PcLayout L, const uint32_t* __restrict__ ngraph, int64_t* __restrict__ acc) {
  const int lane = threadIdx.x & 63;
  for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); base < m; base += gridDim.x * blockDim.x) {
    const uint32_t e = base + uint32_t(lane);
    const bool valid = e < m;
    int64_t v = 0;
    uint32_t g = 0;
    if (valid) {
      int lambda = 0;
      for (int ci = 0; ci < L.vcs; ++ci) lambda += __popcll(rec[uint64_t(e) * L.stride + L.coff + ci]);
      v = lambda > 1 ? int64_t(lambda - 1) * nw[e] : 0;
      g = ngraph[e];
    }
    const uint64_t act = __ballot(valid);
    const int leader = act ? __ffsll(static_cast<long long>(act)) - 1 : 0;
    const uint32_t lg = __shfl(g, leader, 64);
    const bool mine = valid && g == lg;
    int64_t sum = mine ? v : 0;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if (act && lane == leader && sum) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[lg]), static_cast<unsigned long long>(sum));
    if (valid && !mine && v) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[g]), static_cast<unsigned long long>(v));
  }
}

void partition_initialize_batch(mkp_partition* p, const uint32_t* vgraph, const uint32_t* gkbase, uint32_t ktot) {
  mkp_hypergraph* hg = p->hg;
  Context& c = hg->ctx->c;
  fill_async<int64_t>(p->part_w.p, 0, ktot, c.stream);
  if (hg->n) {
    block_weight_batch_kernel<<<grid_for(hg->n, kBwChunk, 65536), 256, 0, c.stream>>>(p->part.p, hg->vw.p, hg->n,
                                                                                     vgraph, gkbase, p->part_w.p);
    MKP_LAUNCH_CHECK();
  }
  if (hg->m) {  // pin counts need no member information: block ids are local
    const int npb = nets_per_block(p->L);
    const size_t lds = size_t(npb) * p->L.vpc * 8 + size_t(npb + 1) * 4;
    pin_count_init_kernel<<<(hg->m + npb - 1) / npb, kInitThreads, lds, c.stream>>>(
        hg->noff.p, hg->pins.p, p->part.p, hg->nw.p, hg->m, p->k, p->L, p->vcs, npb, p->pc.p, nullptr);
    MKP_LAUNCH_CHECK();
  }
}

void partition_km1_batch(mkp_partition* p, const uint32_t* ngraph, uint32_t members, int64_t* out) {
  Context& c = p->hg->ctx->c;
  DevBuf<int64_t> acc(std::max<uint32_t>(members, 1));
  fill_async<int64_t>(acc.p, 0, members, c.stream);
  if (p->hg->m) {
    km1_batch_kernel<<<grid_for(p->hg->m, 256, 65536), 256, 0, c.stream>>>(p->pc.p, p->hg->nw.p, p->hg->m, p->L,
                                                                         ngraph, acc.p);
    MKP_LAUNCH_CHECK();
  }
  MKP_HIP(hipMemcpyAsync(out, acc.p, size_t(members) * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
}

int64_t partition_km1(mkp_partition* p) {
  Context& c = p->hg->ctx->c;
  if (!p->hg->m) return 0;
  const unsigned blocks = grid_for(p->hg->m, 256, kKm1Blocks);
  km1_kernel<<<blocks, 256, 0, c.stream>>>(p->pc.p, p->hg->nw.p, p->hg->m, p->L, p->km1_acc.p);
  MKP_LAUNCH_CHECK();
  int64_t* stg = c.staging<int64_t>(blocks);
  MKP_HIP(hipMemcpyAsync(stg, p->km1_acc.p, size_t(blocks) * 8, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  int64_t v = 0;
  for (unsigned b = 0; b < blocks; ++b) v += stg[b];
  return v;
}

}  // namespace mkp

using namespace mkp;

extern "C" {

int mkp_partition_create(mkp_hypergraph* hg, int32_t k, mkp_partition** out) {
  return guarded([&] {
    MKP_CHECK(hg && out, "null argument");
    MKP_CHECK(k >= 2 && k <= 1024, "k must be in [2, 1024]");
    Bind bind_(hg->ctx->c);
    auto* p = new mkp_partition();
    try {
      p->hg = hg;
      p->k = k;
      p->vcs = (k + 63) / 64;
      p->L = pc_layout(k, std::max<uint32_t>(hg->max_esize, 1));
      p->part.alloc(std::max<uint32_t>(hg->n, 1));
      p->part_w.alloc(k);
      p->pc.alloc(std::max<uint64_t>(uint64_t(hg->m) * p->L.stride, 1));
      p->km1_acc.alloc(kKm1Blocks);
      fill_async<int32_t>(p->part.p, 0, hg->n, hg->ctx->c.stream);
      partition_initialize(p);
      hg->ctx->c.sync();
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
  });
}

int mkp_partition_destroy(mkp_partition* p) {
  return guarded([&] {
    if (!p) return;
    Bind bind_(p->hg->ctx->c);
    (void)hipStreamSynchronize(p->hg->ctx->c.stream);
    delete p;
  });
}

int mkp_partition_set_parts(mkp_partition* p, const int32_t* parts) {
  return guarded([&] {
    MKP_CHECK(p && parts, "null argument");
    for (uint32_t u = 0; u < p->hg->n; ++u) MKP_CHECK(parts[u] >= 0 && parts[u] < p->k, "block id out of range");
    Context& c = p->hg->ctx->c;
    Bind bind_(c);
    MKP_HIP(hipMemcpyAsync(p->part.p, parts, size_t(p->hg->n) * 4, hipMemcpyHostToDevice, c.stream));
    partition_initialize(p);
    c.sync();
  });
}

int mkp_partition_initialize(mkp_partition* p) {
  return guarded([&] {
    MKP_CHECK(p, "null argument");
    Bind bind_(p->hg->ctx->c);
    partition_initialize(p);
    p->hg->ctx->c.sync();
  });
}

int mkp_partition_copy_parts(mkp_partition* dst, const mkp_partition* src) {
  return guarded([&] {
    MKP_CHECK(dst && src, "null argument");
    MKP_CHECK(dst->hg == src->hg && dst->k == src->k, "partitions of different hypergraphs or k");
    Context& c = dst->hg->ctx->c;
    Bind bind_(c);
    MKP_HIP(hipMemcpyAsync(dst->part.p, src->part.p, size_t(dst->hg->n) * 4, hipMemcpyDeviceToDevice, c.stream));
    partition_initialize(dst);
    c.sync();
  });
}

int mkp_partition_get_parts(const mkp_partition* p, int32_t* parts) {
  return guarded([&] {
    MKP_CHECK(p && parts, "null argument");
    Bind bind_(p->hg->ctx->c);
    Context& c = p->hg->ctx->c;
    MKP_HIP(hipMemcpyAsync(parts, p->part.p, size_t(p->hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int mkp_partition_get_part_weights(const mkp_partition* p, int64_t* w) {
  return guarded([&] {
    MKP_CHECK(p && w, "null argument");
    Bind bind_(p->hg->ctx->c);
    Context& c = p->hg->ctx->c;
    MKP_HIP(hipMemcpyAsync(w, p->part_w.p, size_t(p->k) * 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int mkp_partition_get_pin_counts(const mkp_partition* p, uint32_t* counts) {
  return guarded([&] {
    MKP_CHECK(p && counts, "null argument");
    Bind bind_(p->hg->ctx->c);
    Context& c = p->hg->ctx->c;
    const uint64_t total = uint64_t(p->hg->m) * p->k;
    if (!total) return;
    DevBuf<uint32_t> tmp(total);
    unpack_pc_kernel<<<grid_for(total, 256, 65536), 256, 0, c.stream>>>(p->pc.p, p->hg->m, p->k, p->L, tmp.p);
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipMemcpyAsync(counts, tmp.p, total * 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int mkp_partition_get_connectivity(const mkp_partition* p, uint64_t* words) {
  return guarded([&] {
    MKP_CHECK(p && words, "null argument");
    Bind bind_(p->hg->ctx->c);
    Context& c = p->hg->ctx->c;
    const uint64_t total = uint64_t(p->hg->m) * p->vcs;
    if (total) {
      DevBuf<uint64_t> tmp(total);
      conn_words_kernel<<<grid_for(total, 256, 65536), 256, 0, c.stream>>>(p->pc.p, p->hg->m, p->L, tmp.p);
      MKP_LAUNCH_CHECK();
      MKP_HIP(hipMemcpyAsync(words, tmp.p, total * 8, hipMemcpyDeviceToHost, c.stream));
    }
    c.sync();
  });
}

int mkp_partition_get_border(const mkp_partition* p, uint8_t* border) {
  return guarded([&] {
    MKP_CHECK(p && border, "null argument");
    Bind bind_(p->hg->ctx->c);
    Context& c = p->hg->ctx->c;
    if (!p->hg->n) return;
    DevBuf<uint8_t> tmp(p->hg->n);
    border_kernel<<<grid_for(p->hg->n, 256, 65536), 256, 0, c.stream>>>(p->hg->voff.p, p->hg->inc.p, p->pc.p,
                                                                        p->hg->n, p->L, tmp.p);
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipMemcpyAsync(border, tmp.p, p->hg->n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

// ------------------------------------------------ allocation-free border --
// The border flags of the allocation-free entry point are staged in this
// array (one byte per vertex, at most kBorderScratchBytes vertices).  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code never passes the array's name to a kernel or copy: it uses the
// device address that hipGetSymbolAddress returns on that device.
}  // extern "C"
namespace mkp {
constexpr size_t kBorderScratchBytes = size_t(16) << 20;
static __device__ __align__(256) uint8_t scratch[kBorderScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64] = {};  // per device: address of `scratch`

// Address of this device's `scratch`, looked up once per device with the
// device current.
static uint8_t* border_scratch(int device) {
  MKP_CHECK(device >= 0 && device < 64, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    MKP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MKP_CHECK(ptr != nullptr, "hipGetSymbolAddress returned no address for the border scratch");
    g_scratch_ptr[device] = ptr;
  }
  return static_cast<uint8_t*>(g_scratch_ptr[device]);
}
}  // namespace mkp
extern "C" {

int mkp_partition_get_border_noalloc(const mkp_partition* p, uint8_t* border) {
  return guarded([&] {
    MKP_CHECK(p && border, "null argument");
    const uint32_t n = p->hg->n;
    MKP_CHECK(n <= kBorderScratchBytes, "mkp_partition_get_border_noalloc: " + std::to_string(n) +
                                            " vertices exceed the scratch limit of " +
                                            std::to_string(kBorderScratchBytes) + " vertices");
    Context& c = p->hg->ctx->c;
    MKP_HIP(hipSetDevice(c.device));
    if (!n) return;
    uint8_t* ptr = border_scratch(c.device);
    border_kernel<<<grid_for(n, 256, 65536), 256, 0, c.stream>>>(p->hg->voff.p, p->hg->inc.p, p->pc.p, n, p->L,
                                                                  ptr);
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipMemcpyAsync(border, ptr, n, hipMemcpyDeviceToHost, c.stream));
    MKP_HIP(hipStreamSynchronize(c.stream));
  });
}

int mkp_partition_km1(mkp_partition* p, int64_t* km1) {
  return guarded([&] {
    MKP_CHECK(p && km1, "null argument");
    Bind bind_(p->hg->ctx->c);
    *km1 = partition_km1(p);
  });
}

}  // extern "C"
