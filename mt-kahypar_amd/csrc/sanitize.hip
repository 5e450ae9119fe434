// Steps on either side of the multilevel partitioner in partition()
// (partitioner.cpp:74-86, 190-195):
//
//   * degree-zero vertices (degree_zero_hn_remover.h:46-61): in id order,
//     every vertex without nets is removed until the remaining count reaches
//     the contraction limit;
//   * large nets (large_he_remover.h:47-58): nets with more pins than
//     max(max(N * maxnet-remove-factor, 100), 50000) are removed (threshold
//     of partitioner.cpp:37-38 and large_he_remover.h:90-92, N the vertex
//     count of the input);
//   * restore: large nets come back with their pins, so their blocks follow
//     from the vertices', and the removed vertices are bin-packed into the
//     lightest blocks (degree_zero_hn_remover.h:64-83).
//
// Both removals are one compaction on the device: kept vertices and nets keep
// their relative order and a net keeps its pin order, so the reduced
// hypergraph is the input with the removed elements cut out.
#include <algorithm>
#include <numeric>

#include "internal.h"

namespace mkp {

constexpr uint32_t kRemoved = 0xffffffffu;  // vertex map entry of a removed vertex

__global__ void sz_degree_zero_kernel(const uint32_t* __restrict__ voff, uint32_t n, uint32_t* __restrict__ deg0) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    deg0[u] = voff[u + 1] == voff[u] ? 1u : 0u;
}

// the first `removed` degree-zero vertices (rank = exclusive count of
// degree-zero vertices before u) go, every other vertex stays
__global__ void sz_keep_vertices_kernel(const uint32_t* __restrict__ deg0, const uint32_t* __restrict__ rank,
                                        uint32_t n, uint32_t removed, uint32_t* __restrict__ keep) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
    keep[u] = (deg0[u] && rank[u] < removed) ? 0u : 1u;
}

__global__ void sz_keep_nets_kernel(const uint32_t* __restrict__ noff, uint32_t m, uint32_t threshold,
                                    uint32_t* __restrict__ keep, uint32_t* __restrict__ size) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const uint32_t s = noff[e + 1] - noff[e];
    const bool k = s <= threshold;
    keep[e] = k ? 1u : 0u;
    size[e] = k ? s : 0u;
  }
}

__global__ void sz_vertices_kernel(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos, uint32_t n,
                                   const int32_t* __restrict__ vw, const int32_t* __restrict__ comm,
                                   uint32_t* __restrict__ vmap, int32_t* __restrict__ rvw,
                                   int32_t* __restrict__ rcomm) {
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
    if (keep[u]) {
      const uint32_t r = pos[u];
      vmap[u] = r;
      rvw[r] = vw[u];
      rcomm[r] = comm[u];
    } else {
      vmap[u] = kRemoved;
    }
  }
}

// kept net e becomes net epos[e]; its pins start at off[e] (exclusive scan of
// the kept sizes over all nets)
__global__ void sz_nets_kernel(const uint32_t* __restrict__ keep, const uint32_t* __restrict__ epos,
                               const uint32_t* __restrict__ off, uint32_t m, const int32_t* __restrict__ nw,
                               uint32_t* __restrict__ rnoff, int32_t* __restrict__ rnw) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    if (!keep[e]) continue;
    rnoff[epos[e]] = off[e];
    rnw[epos[e]] = nw[e];
  }
}

__global__ void sz_pins_kernel(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ pins,
                               const uint32_t* __restrict__ pin_net, uint64_t p, const uint32_t* __restrict__ keep,
                               const uint32_t* __restrict__ off, const uint32_t* __restrict__ vmap,
                               uint32_t* __restrict__ rpins) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < p; i += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t e = pin_net[i];
    if (keep[e]) rpins[off[e] + uint32_t(i - noff[e])] = vmap[pins[i]];
  }
}

__global__ void sz_max_size_kernel(const uint32_t* __restrict__ noff, uint32_t m, uint32_t* __restrict__ mx) {
  uint32_t best = 0;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
    best = max(best, noff[e + 1] - noff[e]);
  for (int o = 32; o > 0; o >>= 1) best = max(best, static_cast<uint32_t>(__shfl_xor(best, o, 64)));
  if ((threadIdx.x & 63) == 0 && best) atomicMax(mx, best);
}

uint32_t large_net_threshold(uint32_t n, double factor) {
  const double t = std::max(double(n) * factor, 100.0);
  const uint32_t ti = t >= 4294967295.0 ? 0xffffffffu : uint32_t(t);  // HypernodeID truncation
  return std::max<uint32_t>(ti, 50000u);
}

void reduce(mkp_hypergraph* hg, uint32_t contraction_limit, uint32_t net_threshold, Reduction& r) {
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t n = hg->n, m = hg->m;
  r.removed_vertices = 0;
  r.removed_nets = 0;
  // degree-zero vertices (removable only while more than the contraction
  // limit remain), large nets
  DevBuf<uint32_t> deg0(std::max<uint32_t>(n, 1)), rank(size_t(n) + 1), keepe(std::max<uint32_t>(m, 1)),
      esize(std::max<uint32_t>(m, 1)), eoff(size_t(m) + 1), epos(size_t(m) + 1);
  if (n) {
    sz_degree_zero_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(hg->voff.p, n, deg0.p);
    MKP_LAUNCH_CHECK();
  }
  exclusive_scan_u32(c, deg0.p, rank.p, n);
  if (m) {
    sz_keep_nets_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(hg->noff.p, m, net_threshold, keepe.p, esize.p);
    MKP_LAUNCH_CHECK();
  }
  exclusive_scan_u32(c, keepe.p, epos.p, m);
  exclusive_scan_u32(c, esize.p, eoff.p, m);
  uint32_t* pin = reinterpret_cast<uint32_t*>(c.pinned);
  MKP_HIP(hipMemcpyAsync(pin, rank.p + n, 4, hipMemcpyDeviceToHost, s));
  MKP_HIP(hipMemcpyAsync(pin + 1, epos.p + m, 4, hipMemcpyDeviceToHost, s));
  MKP_HIP(hipMemcpyAsync(pin + 2, eoff.p + m, 4, hipMemcpyDeviceToHost, s));
  c.sync();
  const uint32_t zero = pin[0], mr = pin[1], pr = pin[2];
  const uint32_t removable = n > contraction_limit ? n - contraction_limit : 0;
  r.removed_vertices = std::min(zero, removable);
  r.removed_nets = m - mr;
  if (r.removed_vertices == 0 && r.removed_nets == 0) return;
  const uint32_t nr = n - r.removed_vertices;

  auto* g = new mkp_hypergraph();
  r.hg.reset(g);
  g->ctx = hg->ctx;
  g->n = nr;
  g->m = mr;
  g->p = pr;
  DevBuf<uint32_t> keepv(std::max<uint32_t>(n, 1)), vpos(size_t(n) + 1);
  r.vmap.alloc(std::max<uint32_t>(n, 1));
  g->vw.alloc(std::max<uint32_t>(nr, 1));
  g->comm.alloc(std::max<uint32_t>(nr, 1));
  g->noff.alloc(size_t(mr) + 1);
  g->nw.alloc(std::max<uint32_t>(mr, 1));
  g->pins.alloc(std::max<uint32_t>(pr, 1));
  if (n) {
    sz_keep_vertices_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(deg0.p, rank.p, n, r.removed_vertices, keepv.p);
    MKP_LAUNCH_CHECK();
  }
  exclusive_scan_u32(c, keepv.p, vpos.p, n);
  if (n) {
    sz_vertices_kernel<<<grid_for(n, 256, 65536), 256, 0, s>>>(keepv.p, vpos.p, n, hg->vw.p, hg->comm.p, r.vmap.p,
                                                                 g->vw.p, g->comm.p);
    MKP_LAUNCH_CHECK();
  }
  if (m) {
    sz_nets_kernel<<<grid_for(m, 256, 65536), 256, 0, s>>>(keepe.p, epos.p, eoff.p, m, hg->nw.p, g->noff.p, g->nw.p);
    MKP_LAUNCH_CHECK();
  }
  fill_async<uint32_t>(g->noff.p + mr, pr, 1, s);
  if (hg->p) {
    sz_pins_kernel<<<grid_for(hg->p, 256, 65536), 256, 0, s>>>(hg->noff.p, hg->pins.p, hg->pin_net.p, hg->p, keepe.p,
                                                                 eoff.p, r.vmap.p, g->pins.p);
    MKP_LAUNCH_CHECK();
  }
  DevBuf<uint32_t> mx(1);
  fill_async<uint32_t>(mx.p, 0u, 1, s);
  if (mr) {
    sz_max_size_kernel<<<grid_for(mr, 256, 4096), 256, 0, s>>>(g->noff.p, mr, mx.p);
    MKP_LAUNCH_CHECK();
  }
  // weights of the removed vertices leave the total
  std::vector<int32_t> rvw(nr);
  if (nr) MKP_HIP(hipMemcpyAsync(rvw.data(), g->vw.p, size_t(nr) * 4, hipMemcpyDeviceToHost, s));
  MKP_HIP(hipMemcpyAsync(pin + 3, mx.p, 4, hipMemcpyDeviceToHost, s));
  c.sync();
  g->max_esize = pin[3];
  int64_t total = 0;
  for (int32_t w : rvw) total += w;
  g->total_w = total;
  build_incidence(g);
}

// degree_zero_hn_remover.h:64-83.  The removed vertices take, in order of
// decreasing weight (ties: smaller id first), the block at the front of a
// list sorted by increasing block weight (ties: smaller block id first); the
// block that received a vertex moves back past every lighter block.
void bin_pack(const std::vector<uint32_t>& removed, const std::vector<int32_t>& weight, int k,
              std::vector<int64_t>& part_w, std::vector<int32_t>& part) {
  std::vector<uint32_t> order(removed);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return weight[a] > weight[b]; });
  std::vector<int32_t> blocks(static_cast<size_t>(k), 0);
  std::iota(blocks.begin(), blocks.end(), 0);
  std::stable_sort(blocks.begin(), blocks.end(), [&](int32_t a, int32_t b) { return part_w[a] < part_w[b]; });
  for (uint32_t u : order) {
    const int32_t to = blocks[0];
    part[u] = to;
    part_w[size_t(to)] += weight[u];
    for (int i = 0; i + 1 < k && part_w[size_t(blocks[i])] > part_w[size_t(blocks[i + 1])]; ++i)
      std::swap(blocks[i], blocks[i + 1]);
  }
}

void restore(mkp_hypergraph* hg, const Reduction& r, const mkp_partition* reduced, int32_t* parts) {
  Context& c = hg->ctx->c;
  const hipStream_t s = c.stream;
  const uint32_t n = hg->n, nr = r.hg->n;
  const int k = reduced->k;
  std::vector<uint32_t> vmap(n);
  std::vector<int32_t> rp(nr), vw(n);
  std::vector<int64_t> pw(static_cast<size_t>(k), 0);
  if (n) {
    MKP_HIP(hipMemcpyAsync(vmap.data(), r.vmap.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
    MKP_HIP(hipMemcpyAsync(vw.data(), hg->vw.p, size_t(n) * 4, hipMemcpyDeviceToHost, s));
  }
  if (nr) MKP_HIP(hipMemcpyAsync(rp.data(), reduced->part.p, size_t(nr) * 4, hipMemcpyDeviceToHost, s));
  MKP_HIP(hipMemcpyAsync(pw.data(), reduced->part_w.p, size_t(k) * 8, hipMemcpyDeviceToHost, s));
  c.sync();
  std::vector<uint32_t> removed;
  removed.reserve(r.removed_vertices);
  std::vector<int32_t> full(n, -1);
  for (uint32_t u = 0; u < n; ++u) {
    if (vmap[u] != kRemoved) full[u] = rp[vmap[u]];
    else removed.push_back(u);
  }
  bin_pack(removed, vw, k, pw, full);
  std::copy(full.begin(), full.end(), parts);
}

}  // namespace mkp

using namespace mkp;

extern "C" {

int mkp_hypergraph_reduce(mkp_hypergraph* hg, uint32_t contraction_limit, uint32_t net_threshold,
                          mkp_hypergraph** reduced, uint32_t* vmap, uint32_t* removed_vertices,
                          uint32_t* removed_nets) {
  return guarded([&] {
    MKP_CHECK(hg && reduced, "null argument");
    Context& c = hg->ctx->c;
    Bind bind_(c);
    Reduction r;
    reduce(hg, contraction_limit, net_threshold, r);
    if (vmap && hg->n) {
      if (r.hg) {
        MKP_HIP(hipMemcpyAsync(vmap, r.vmap.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
      } else {
        for (uint32_t u = 0; u < hg->n; ++u) vmap[u] = u;
      }
    }
    if (removed_vertices) *removed_vertices = r.removed_vertices;
    if (removed_nets) *removed_nets = r.removed_nets;
    *reduced = r.hg.release();
  });
}

uint32_t mkp_large_net_threshold(uint32_t num_vertices, double factor) {
  return large_net_threshold(num_vertices, factor);
}

}  // extern "C"
