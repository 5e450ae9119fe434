// Context, error state and the device-resident static hypergraph.
#include <cmath>
#include <cstring>
#include <mutex>

#include "internal.h"

namespace mkp {

static thread_local std::string g_last_error;
void set_last_error(const std::string& s) { g_last_error = s; }

// One thread per pin: key = (pin << 32) | net, found by binary search of the
// pin position in the net offsets.
__global__ void incidence_keys_kernel(const uint32_t* __restrict__ noff, const uint32_t* __restrict__ pins,
                                      uint32_t m, uint64_t p, uint64_t* __restrict__ keys,
                                      uint32_t* __restrict__ degree, uint32_t* __restrict__ pin_net) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < p; i += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t lo = 0, hi = m;  // largest e with noff[e] <= i
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (noff[mid] <= uint32_t(i)) lo = mid; else hi = mid;
    }
    const uint32_t v = pins[i];
    keys[i] = (uint64_t(v) << 32) | lo;
    pin_net[i] = lo;
    atomicAdd(&degree[v], 1u);
  }
}

__global__ void split_words_kernel(const uint64_t* __restrict__ keys, uint32_t* __restrict__ lo_out,
                                   uint32_t* __restrict__ hi_out, uint64_t p) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < p; i += uint64_t(gridDim.x) * blockDim.x) {
    lo_out[i] = uint32_t(keys[i]);
    hi_out[i] = uint32_t(keys[i] >> 32);
  }
}

__global__ void unit_weight_kernel(const int32_t* __restrict__ nw, uint32_t m, uint32_t* __restrict__ flag) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
    if (nw[e] != 1) *flag = 0;
}

void build_incidence(mkp_hypergraph* hg) {
  Context& c = hg->ctx->c;
  hg->voff.alloc(size_t(hg->n) + 1);
  hg->inc.alloc(std::max<uint64_t>(hg->p, 1));
  hg->inc_vtx.alloc(std::max<uint64_t>(hg->p, 1));
  hg->pin_net.alloc(std::max<uint64_t>(hg->p, 1));
  DevBuf<uint32_t> degree(size_t(hg->n) + 1), flag(1);
  fill_async<uint32_t>(degree.p, 0u, hg->n, c.stream);
  fill_async<uint32_t>(flag.p, 1u, 1, c.stream);
  if (hg->m) {
    unit_weight_kernel<<<grid_for(hg->m, 256, 65536), 256, 0, c.stream>>>(hg->nw.p, hg->m, flag.p);
    MKP_LAUNCH_CHECK();
  }
  if (hg->p > 0) {
    DevBuf<uint64_t> keys(hg->p), sorted(hg->p);
    incidence_keys_kernel<<<grid_for(hg->p, 256, 65536), 256, 0, c.stream>>>(hg->noff.p, hg->pins.p, hg->m, hg->p,
                                                                            keys.p, degree.p, hg->pin_net.p);
    MKP_LAUNCH_CHECK();
    int pin_bits = 1;
    while ((uint64_t(1) << pin_bits) < hg->n) ++pin_bits;
    sort_keys_u64(c, keys.p, sorted.p, hg->p, 0, 32 + pin_bits);
    split_words_kernel<<<grid_for(hg->p, 256, 65536), 256, 0, c.stream>>>(sorted.p, hg->inc.p, hg->inc_vtx.p, hg->p);
    MKP_LAUNCH_CHECK();
    exclusive_scan_u32(c, degree.p, hg->voff.p, hg->n);
  } else {
    fill_async<uint32_t>(hg->voff.p, 0u, size_t(hg->n) + 1, c.stream);
  }
  uint32_t unit = 0;
  MKP_HIP(hipMemcpyAsync(&unit, flag.p, 4, hipMemcpyDeviceToHost, c.stream));
  c.sync();
  hg->unit_net_weights = unit != 0;
}

PcLayout pc_layout(int k, uint32_t max_esize) {
  // num_bits_per_element / num_entries_per_value / num_values_per_hyperedge
  // (pin_count_in_part.h:198-215)
  PcLayout L;
  L.b = int(std::ceil(std::log2(double(max_esize) + 1.0)));
  if (L.b < 1) L.b = 1;
  L.epw = std::min(64 / L.b, k);
  L.vpc = k / L.epw + (k % L.epw != 0);
  L.mask = (L.b >= 64) ? ~0ull : ((1ull << L.b) - 1);
  L.vcs = (k + 63) / 64;
  L.coff = L.vpc;
  L.woff = L.vpc + L.vcs;
  L.stride = L.vpc + L.vcs + 1;
  return L;
}

// Plain device copy, 16 bytes per lane per iteration.
__global__ void __launch_bounds__(256) stream_copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                          uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = in[i];
}

}  // namespace mkp

using namespace mkp;

extern "C" {

const char* mkp_last_error(void) { return g_last_error.c_str(); }

int mkp_context_create(int device, mkp_context** out) {
  return guarded([&] {
    MKP_CHECK(out, "out is null");
    int count = 0;
    MKP_HIP(hipGetDeviceCount(&count));
    MKP_CHECK(device >= 0 && device < count, "no such device");
    MKP_HIP(hipSetDevice(device));
    auto* ctx = new mkp_context();
    ctx->c.device = device;
    MKP_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
    MKP_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->c.pinned), 64 * sizeof(int64_t), hipHostMallocDefault));
    MKP_HIP(hipEventCreate(&ctx->c.ev_a));
    MKP_HIP(hipEventCreate(&ctx->c.ev_b));
    MKP_HIP(hipEventCreateWithFlags(&ctx->c.ev_wait, hipEventDisableTiming));
    hipDeviceProp_t prop;
    MKP_HIP(hipGetDeviceProperties(&prop, device));
    ctx->c.num_cus = prop.multiProcessorCount;
    // the allocation cache may hold up to 3/5 of the device memory
    ctx->c.pool->max_cached = size_t(double(prop.totalGlobalMem) * 0.6);
    *out = ctx;
  });
}

int mkp_context_destroy(mkp_context* ctx) {
  return guarded([&] {
    if (!ctx) return;
    Bind bind_(ctx->c);
    (void)hipStreamSynchronize(ctx->c.stream);
    ctx->c.tmp.release();
    for (auto& s : ctx->c.scratch) s.release();
    // cached blocks go back to the device; buffers still alive (objects
    // destroyed after their context) are freed directly when released
    ctx->c.pool->close();
    if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
    if (ctx->c.stage) (void)hipHostFree(ctx->c.stage);
    if (ctx->c.ev_a) (void)hipEventDestroy(ctx->c.ev_a);
    if (ctx->c.ev_b) (void)hipEventDestroy(ctx->c.ev_b);
    if (ctx->c.ev_wait) (void)hipEventDestroy(ctx->c.ev_wait);
    (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
  });
}

int mkp_context_synchronize(mkp_context* ctx) {
  return guarded([&] { MKP_CHECK(ctx, "ctx is null"); ctx->c.sync(); });
}

int mkp_context_set_profiling(mkp_context* ctx, int enable) {
  return guarded([&] {
    MKP_CHECK(ctx, "ctx is null");
    ctx->c.sync();
    ctx->c.prof.enabled = enable != 0;
  });
}

int mkp_context_reset_profile(mkp_context* ctx) {
  return guarded([&] {
    MKP_CHECK(ctx, "ctx is null");
    ctx->c.sync();
    ctx->c.prof.reset();
  });
}

int mkp_context_set_tuning(mkp_context* ctx, const char* key, int64_t value) {
  return guarded([&] {
    MKP_CHECK(ctx && key, "null argument");
    const std::string k(key);
    if (k == "louvain_group_lanes") {
      MKP_CHECK(value == 0 || value == 8 || value == 16 || value == 32 || value == 64,
                "louvain_group_lanes must be 0 (autotune), 8, 16, 32 or 64");
      ctx->c.tune.lv_group_lanes = int(value);
    } else if (k == "rating_block_passes") {
      MKP_CHECK(value >= 0 && value <= 1024, "rating_block_passes must be in [0, 1024]");
      ctx->c.tune.rating_block_passes = int(value);
    } else if (k == "louvain_fused_classes") {
      MKP_CHECK(value >= 0 && value <= 3, "louvain_fused_classes must be 0, 1, 2 or 3");
      ctx->c.tune.lv_fused_classes = int(value);
    } else if (k == "sleep_wait") {
      MKP_CHECK(value == 0 || value == 1, "sleep_wait must be 0 (spin) or 1 (poll, then sleep)");
      ctx->c.sleep_wait = value != 0;
    } else {
      MKP_CHECK(false, "unknown tuning key");
    }
  });
}

int mkp_context_set_profile_filter(mkp_context* ctx, const char* kernel_name) {
  return guarded([&] {
    MKP_CHECK(ctx, "ctx is null");
    ctx->c.sync();
    ctx->c.prof.only = kernel_name ? std::string(kernel_name) : std::string();
  });
}

int mkp_context_get_profile(mkp_context* ctx, int family, double* ms, double* bytes, int64_t* launches) {
  return guarded([&] {
    MKP_CHECK(ctx, "ctx is null");
    MKP_CHECK(family >= 0 && family < kNumFamilies, "unknown kernel family");
    ctx->c.sync();
    if (ms) *ms = ctx->c.prof.ms[family];
    if (bytes) *bytes = ctx->c.prof.bytes[family];
    if (launches) *launches = ctx->c.prof.launches[family];
  });
}

int mkp_context_profile_kernel_count(mkp_context* ctx, int32_t* count) {
  return guarded([&] {
    MKP_CHECK(ctx && count, "null argument");
    ctx->c.sync();
    *count = int32_t(ctx->c.prof.kernels.size());
  });
}

int mkp_context_profile_kernel(mkp_context* ctx, int32_t index, const char** name, double* ms, double* bytes,
                               int64_t* launches) {
  return guarded([&] {
    MKP_CHECK(ctx, "null argument");
    ctx->c.sync();
    MKP_CHECK(index >= 0 && size_t(index) < ctx->c.prof.kernels.size(), "kernel record index out of range");
    const auto& k = ctx->c.prof.kernels[size_t(index)];
    if (name) *name = k.name.c_str();
    if (ms) *ms = k.ms;
    if (bytes) *bytes = k.bytes;
    if (launches) *launches = k.launches;
  });
}

int mkp_stream_copy(mkp_context* ctx, uint64_t bytes, int32_t reps, double* ms_per_copy) {
  return guarded([&] {
    MKP_CHECK(ctx && ms_per_copy, "null argument");
    MKP_CHECK(bytes >= 16 && bytes % 16 == 0 && reps >= 1, "bytes must be a positive multiple of 16");
    Context& c = ctx->c;
    Bind bind_(c);
    const uint64_t n = bytes / 16;
    DevBuf<uint4> a(n), b(n);
    fill_async<uint32_t>(reinterpret_cast<uint32_t*>(a.p), 1u, n * 4, c.stream);
    const unsigned grid = unsigned(std::min<uint64_t>((n + 255) / 256, uint64_t(c.num_cus) * 32));
    stream_copy_kernel<<<grid, 256, 0, c.stream>>>(a.p, b.p, n);  // warm-up
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipEventRecord(c.ev_a, c.stream));
    for (int32_t r = 0; r < reps; ++r) stream_copy_kernel<<<grid, 256, 0, c.stream>>>(a.p, b.p, n);
    MKP_LAUNCH_CHECK();
    MKP_HIP(hipEventRecord(c.ev_b, c.stream));
    c.sync();
    float ms = 0;
    MKP_HIP(hipEventElapsedTime(&ms, c.ev_a, c.ev_b));
    *ms_per_copy = double(ms) / reps;
  });
}

const char* mkp_profile_family_name(int family) {
  static const char* names[kNumFamilies] = {"pin_count_init", "lp_gain", "lp_apply", "louvain_move",
                                            "louvain_contract", "rating", "contraction", "fm", "initial"};
  return (family >= 0 && family < kNumFamilies) ? names[family] : "";
}

int mkp_hypergraph_create(mkp_context* ctx, uint32_t n, uint32_t m, const uint64_t* noff, const uint32_t* pins,
                          const int32_t* vw, const int32_t* nw, mkp_hypergraph** out) {
  return guarded([&] {
    MKP_CHECK(ctx && out && noff, "null argument");
    const uint64_t p = noff[m];
    MKP_CHECK(p < (uint64_t(1) << 32), "more than 2^32-1 pins are not supported");
    MKP_CHECK(p == 0 || pins, "pins is null");
    Bind bind_(ctx->c);
    std::vector<uint32_t> off32(size_t(m) + 1);
    uint32_t max_esize = 0;
    for (uint32_t e = 0; e <= m; ++e) {
      off32[e] = uint32_t(noff[e]);
      if (e < m) {
        MKP_CHECK(noff[e + 1] >= noff[e], "net offsets are not monotone");
        max_esize = std::max<uint32_t>(max_esize, uint32_t(noff[e + 1] - noff[e]));
      }
    }
    for (uint64_t i = 0; i < p; ++i) MKP_CHECK(pins[i] < n, "pin id out of range");
    int64_t total = 0;
    if (vw) {
      for (uint32_t u = 0; u < n; ++u) { MKP_CHECK(vw[u] >= 0, "negative vertex weight"); total += vw[u]; }
    } else {
      total = n;
    }
    MKP_CHECK(total < (int64_t(1) << 31), "total vertex weight exceeds int32");
    auto* hg = new mkp_hypergraph();
    try {
      hg->ctx = ctx;
      hg->n = n; hg->m = m; hg->p = p; hg->max_esize = max_esize; hg->total_w = total;
      hipStream_t s = ctx->c.stream;
      hg->noff.alloc(size_t(m) + 1);
      hg->pins.alloc(std::max<uint64_t>(p, 1));
      hg->vw.alloc(std::max<uint32_t>(n, 1));
      hg->nw.alloc(std::max<uint32_t>(m, 1));
      hg->comm.alloc(std::max<uint32_t>(n, 1));
      MKP_HIP(hipMemcpyAsync(hg->noff.p, off32.data(), off32.size() * 4, hipMemcpyHostToDevice, s));
      if (p) MKP_HIP(hipMemcpyAsync(hg->pins.p, pins, p * 4, hipMemcpyHostToDevice, s));
      if (vw) MKP_HIP(hipMemcpyAsync(hg->vw.p, vw, size_t(n) * 4, hipMemcpyHostToDevice, s));
      else fill_async<int32_t>(hg->vw.p, 1, n, s);
      if (nw) MKP_HIP(hipMemcpyAsync(hg->nw.p, nw, size_t(m) * 4, hipMemcpyHostToDevice, s));
      else fill_async<int32_t>(hg->nw.p, 1, m, s);
      fill_async<int32_t>(hg->comm.p, 0, n, s);
      build_incidence(hg);
    } catch (...) {
      delete hg;
      throw;
    }
    *out = hg;
  });
}

int mkp_hypergraph_destroy(mkp_hypergraph* hg) {
  return guarded([&] {
    if (!hg) return;
    Bind bind_(hg->ctx->c);
    (void)hipStreamSynchronize(hg->ctx->c.stream);
    delete hg;
  });
}

int mkp_hypergraph_get_info(const mkp_hypergraph* hg, mkp_hypergraph_info* info) {
  return guarded([&] {
    MKP_CHECK(hg && info, "null argument");
    info->num_vertices = hg->n; info->num_nets = hg->m; info->num_pins = hg->p;
    info->max_net_size = hg->max_esize; info->total_weight = hg->total_w;
  });
}

int mkp_hypergraph_set_communities(mkp_hypergraph* hg, const int32_t* comm) {
  return guarded([&] {
    MKP_CHECK(hg && comm, "null argument");
    Bind bind_(hg->ctx->c);
    MKP_HIP(hipMemcpyAsync(hg->comm.p, comm, size_t(hg->n) * 4, hipMemcpyHostToDevice, hg->ctx->c.stream));
    hg->ctx->c.sync();
  });
}

int mkp_hypergraph_copy_to_host(const mkp_hypergraph* hg, uint64_t* noff, uint32_t* pins, uint64_t* voff,
                                uint32_t* inc, int32_t* vw, int32_t* nw, int32_t* comm) {
  return guarded([&] {
    MKP_CHECK(hg, "null hypergraph");
    Bind bind_(hg->ctx->c);
    hipStream_t s = hg->ctx->c.stream;
    std::vector<uint32_t> t;
    if (noff) {
      t.resize(size_t(hg->m) + 1);
      MKP_HIP(hipMemcpyAsync(t.data(), hg->noff.p, t.size() * 4, hipMemcpyDeviceToHost, s));
      MKP_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < t.size(); ++i) noff[i] = t[i];
    }
    if (voff) {
      t.resize(size_t(hg->n) + 1);
      MKP_HIP(hipMemcpyAsync(t.data(), hg->voff.p, t.size() * 4, hipMemcpyDeviceToHost, s));
      MKP_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < t.size(); ++i) voff[i] = t[i];
    }
    if (pins && hg->p) MKP_HIP(hipMemcpyAsync(pins, hg->pins.p, hg->p * 4, hipMemcpyDeviceToHost, s));
    if (inc && hg->p) MKP_HIP(hipMemcpyAsync(inc, hg->inc.p, hg->p * 4, hipMemcpyDeviceToHost, s));
    if (vw && hg->n) MKP_HIP(hipMemcpyAsync(vw, hg->vw.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, s));
    if (nw && hg->m) MKP_HIP(hipMemcpyAsync(nw, hg->nw.p, size_t(hg->m) * 4, hipMemcpyDeviceToHost, s));
    if (comm && hg->n) MKP_HIP(hipMemcpyAsync(comm, hg->comm.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, s));
    MKP_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
