// Shared definitions of the gfx950 implementation: error handling, device
// buffers, the context (device, stream, scratch), and the stateless hash that
// defines tie-breaking of every synchronous round.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace mkp {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MKP_HIP(call)                                                                       \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::mkp::Error(2, std::string(#call) + " failed: " + hipGetErrorString(_e) +      \
                                " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");    \
  } while (0)
#define MKP_CHECK(cond, msg)                           \
  do {                                                 \
    if (!(cond)) throw ::mkp::Error(1, (msg));         \
  } while (0)
#define MKP_LAUNCH_CHECK() MKP_HIP(hipGetLastError())

void set_last_error(const std::string& s);

// Runs f; maps exceptions to a status code and the last-error string.
template <typename F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return 3;
  }
}

// ------------------------------------------------------------------ hash --
// SplitMix64 finalizer (Steele, Lea, Flood, OOPSLA 2014); hash4 mixes a seed
// and three words.  Every tie-break of the synchronous rounds is a function of
// (seed, round tag, vertex, candidate) through this hash.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t hash4(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
  uint64_t h = mix64(seed ^ (a * 0xD6E8FEB86659FD93ull));
  h = mix64(h ^ (b * 0xA0761D6478BD642Full));
  return mix64(h ^ (c * 0xE7037ED1A0B428DBull));
}

// Sub-round bucket of a vertex: MurmurHash3's 32-bit finalizer over the id
// and a per-(seed, round) salt computed once on the host.
__host__ __device__ inline uint32_t fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85EBCA6Bu; h ^= h >> 13; h *= 0xC2B2AE35u; h ^= h >> 16;
  return h;
}
__host__ __device__ inline uint32_t bucket_salt(uint64_t seed, uint64_t tag) { return uint32_t(mix64(seed ^ tag)); }
__host__ __device__ inline uint32_t bucket_of(uint32_t salt, uint32_t u, uint32_t buckets) {
  return fmix32(u * 0x9E3779B1u + salt) % buckets;
}

// Louvain sub-rounds take whole tiles of 2^kLvTileShift consecutive node ids
// (the bucket of node u is the bucket of tile u >> kLvTileShift), so a
// sub-round reads whole cache lines of the per-node and per-arc arrays
// instead of one id in eight.
constexpr int kLvTileShift = 6;

// Block b of a grid of nb blocks (nb a multiple of 8) works on chunk
// (b % 8) * (nb / 8) + b / 8 of a launch's work.  Workgroups are dealt to the
// eight XCDs round-robin, so every XCD then takes one contiguous eighth of
// the work and its L2 holds the data that neighbouring items gather.
__host__ __device__ inline uint32_t xcd_chunk(uint32_t b, uint32_t nb) { return (b & 7u) * (nb >> 3) + (b >> 3); }
// grid_for rounded up to a multiple of 8 blocks (see xcd_chunk)
inline unsigned xcd_grid(uint64_t work, unsigned per_block, unsigned max_blocks = 1u << 20) {
  uint64_t g = (work + per_block - 1) / per_block;
  if (g == 0) g = 1;
  if (g > max_blocks) g = max_blocks;
  return unsigned((g + 7) & ~uint64_t(7));
}

// Groups the lanes of a wave by a small key ("slot", -1 = none).  Lanes with
// the same slot are served together: up to 64 rounds (one per
// distinct slot of the wave), each with the lanes of the lowest pending slot.
// Returns, for the calling lane, the rank among its slot's lanes and the
// slot's lane count and leader; every lane of the wave calls it.
struct SlotGroup {
  uint32_t rank, count;
  bool leader;
};
__device__ inline SlotGroup slot_group(int slot) {
  const int lane = threadIdx.x & 63;
  uint64_t pending = __ballot(slot >= 0);
  SlotGroup r{0, 0, false};
  while (pending) {
    const int first = __ffsll(static_cast<long long>(pending)) - 1;
    const int s = __shfl(slot, first, 64);
    const uint64_t same = __ballot(slot == s) & pending;
    if (slot == s && ((same >> lane) & 1ull)) {
      r.count = uint32_t(__popcll(same));
      r.rank = uint32_t(__popcll(same & ((1ull << lane) - 1ull)));
      r.leader = lane == first;
    }
    pending &= ~same;
  }
  return r;
}

// orders the LDS accesses of the lanes of one wave (in place of a block
// barrier when every LDS region is private to one wave)
__device__ inline void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// round tags (one per decision family)
constexpr uint64_t kLpBucketTag = 0x4C50420000000000ull;
constexpr uint64_t kLpCoinTag = 0x4C50430000000000ull;
constexpr uint64_t kLpPrioTag = 0x4C50500000000000ull;
constexpr uint64_t kLvBucketTag = 0x4C56420000000000ull;
constexpr uint64_t kCoBucketTag = 0x434F420000000000ull;
constexpr uint64_t kCoTieTag = 0x434F540000000000ull;
constexpr uint64_t kCoPrioTag = 0x434F500000000000ull;

// ------------------------------------------------------------ buffers --
// Cache of device allocations of one context.  Blocks come in power-of-two
// size classes and go back to the cache when their buffer is released, so the
// steady state of repeated calls allocates nothing and, unlike hipFree, a
// release never waits for the device.  A block is reused only by work of the
// same context, i.e. on the same stream, where stream order makes reuse safe.
struct DevicePool {
  std::mutex mu;
  std::multimap<size_t, void*> blocks;  // size class -> cached block
  size_t cached = 0;
  size_t in_use = 0;  // bytes of blocks handed out and not yet returned
  size_t peak = 0;    // most bytes in use at once since reset_peak
  bool closed = false;
  size_t held() {
    std::lock_guard<std::mutex> g(mu);
    return in_use + cached;
  }
  size_t cached_bytes() {
    std::lock_guard<std::mutex> g(mu);
    return cached;
  }
  size_t reset_peak() {
    std::lock_guard<std::mutex> g(mu);
    peak = in_use;
    return peak;
  }
  size_t peak_bytes() {
    std::lock_guard<std::mutex> g(mu);
    return peak;
  }
  // bytes the cache may hold; the context sets it from the device's memory
  // (a step whose working set does not fit would free and re-allocate, and a
  // free waits for the device)
  size_t max_cached = size_t(64) << 30;
  static size_t size_class(size_t bytes) {
    size_t c = 256;
    while (c < bytes) c <<= 1;
    return c;
  }
  void* get(size_t cls) {
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = blocks.find(cls);
      if (it != blocks.end()) {
        void* p = it->second;
        blocks.erase(it);
        cached -= cls;
        in_use += cls;
        peak = std::max(peak, in_use);
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {  // return the cache to the device and retry once
      trim();
      MKP_HIP(hipMalloc(&p, cls));
    }
    std::lock_guard<std::mutex> g(mu);
    in_use += cls;
    peak = std::max(peak, in_use);
    return p;
  }
  void put(void* p, size_t cls) {
    {
      std::lock_guard<std::mutex> g(mu);
      in_use -= cls;
      if (!closed && cached + cls <= max_cached) {
        blocks.emplace(cls, p);
        cached += cls;
        return;
      }
    }
    (void)hipFree(p);
  }
  void trim() {
    std::multimap<size_t, void*> b;
    {
      std::lock_guard<std::mutex> g(mu);
      b.swap(blocks);
      cached = 0;
    }
    for (auto& kv : b) (void)hipFree(kv.second);
  }
  void close() {
    {
      std::lock_guard<std::mutex> g(mu);
      closed = true;
    }
    trim();
  }
};

// Pool of the context whose entry point runs on this thread (set by Bind).
inline thread_local std::shared_ptr<DevicePool> t_pool;

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t cls = 0;                     // pooled size class, 0 = plain allocation
  std::shared_ptr<DevicePool> pool;   // owner of a pooled block
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), cls(o.cls), pool(std::move(o.pool)) {
    o.p = nullptr; o.n = 0; o.cls = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; cls = o.cls; pool = std::move(o.pool);
      o.p = nullptr; o.n = 0; o.cls = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (!count) return;
    if (t_pool) {
      pool = t_pool;
      cls = DevicePool::size_class(count * sizeof(T));
      p = static_cast<T*>(pool->get(cls));
    } else {
      MKP_HIP(hipMalloc(&p, count * sizeof(T)));
    }
  }
  void ensure(size_t count) { if (count > n) alloc(count); }
  void release() {
    if (p) {
      if (pool) pool->put(p, cls);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    cls = 0;
    pool.reset();
  }
  T* get() const { return p; }
};

// ------------------------------------------------------------ profiling --
// Kernel families timed with event pairs on the launch stream when
// profiling is enabled; algorithmic bytes are attributed per launch.
enum Family : int {
  kFamPinCountInit = 0,
  kFamLpGain,
  kFamLpApply,
  kFamLouvainMove,
  kFamLouvainContract,
  kFamRating,
  kFamContraction,
  kFamFm,
  kFamInitial,
  kNumFamilies
};

struct Profiler {
  bool enabled = false;
  std::string only;  // non-empty: time only the launches of this kernel
  // kname: single-kernel record (nullptr for a family-only region)
  struct Rec { hipEvent_t a, b; int fam; double bytes; std::string kname; };
  struct KernelStat { std::string name; double ms = 0, bytes = 0; int64_t launches = 0; };
  std::vector<Rec> open;
  std::vector<hipEvent_t> pool;
  double ms[kNumFamilies] = {};
  double bytes[kNumFamilies] = {};
  int64_t launches[kNumFamilies] = {};
  std::vector<KernelStat> kernels;  // per kernel (instantiation) name
  KernelStat& kernel(const std::string& name) {
    for (auto& k : kernels)
      if (k.name == name) return k;
    kernels.push_back(KernelStat{name});
    return kernels.back();
  }
  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  // Resolves recorded pairs (the stream must be idle).
  void collect() {
    for (auto& r : open) {
      float t = 0;
      if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
        ms[r.fam] += t;
        bytes[r.fam] += r.bytes;
        launches[r.fam] += 1;
        if (!r.kname.empty()) {
          KernelStat& k = kernel(r.kname);
          k.ms += t;
          k.bytes += r.bytes;
          k.launches += 1;
        }
      }
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    open.clear();
  }
  void reset() {
    for (int i = 0; i < kNumFamilies; ++i) { ms[i] = 0; bytes[i] = 0; launches[i] = 0; }
    kernels.clear();
  }
};

// Scoped event pair around the launches of one kernel family; with a kernel
// name it times exactly one launch and is also recorded per kernel.
// Kernel names follow the demangled form without namespaces and arguments,
// e.g. "louvain_move_group_kernel<StarView, 32, 8, 256>".
struct TimedRegion {
  Profiler& p;
  hipStream_t s;
  int fam;
  double bytes;
  std::string kname;
  hipEvent_t a = nullptr;
  TimedRegion(Profiler& prof, hipStream_t st, int f, double b, std::string kernel_name = std::string())
      : p(prof), s(st), fam(f), bytes(b), kname(std::move(kernel_name)) {
    if (p.enabled && (p.only.empty() || kname == p.only)) { a = p.take(); (void)hipEventRecord(a, s); }
  }
  ~TimedRegion() {
    if (a) {
      hipEvent_t b = p.take();
      (void)hipEventRecord(b, s);
      p.open.push_back({a, b, fam, bytes, std::move(kname)});
    }
  }
};

// ------------------------------------------------------------ context --
// Kernel shape choices that do not change results (set by the autotuner or
// mkp_context_set_tuning).
struct Tuning {
  // lanes per node of the Louvain 17..128-arc class (8, 16, 32 or 64); 0 lets
  // the autotuner choose per (graph kind, class size, mean degree)
  int lv_group_lanes = 0;
  std::map<uint64_t, int> lv_group_choice;  // autotuner cache
  // heavy rating proposers are rated by one block each with LDS maps when
  // their keys fit in at most this many passes (0: every one in the arena)
  int rating_block_passes = 4;
  // the Louvain degree classes up to 1024 arcs of a sub-round share one
  // launch: 1 with the 17..64-arc class when it is small, 2 never with it,
  // 3 always with it, 0 one launch per class
  int lv_fused_classes = 1;
};

struct Context {
  Profiler prof;
  Tuning tune;
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf<uint8_t> tmp;          // rocPRIM temporary storage
  DevBuf<uint8_t> scratch[4];   // general scratch arenas
  int64_t* pinned = nullptr;    // small host readback area (64 words)
  // pinned staging for per-round transfers (copies from and to it are truly
  // asynchronous); grows by reallocation, which waits for the stream
  uint8_t* stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  int num_cus = 256;
  template <typename T>
  T* staging(size_t count) {
    const size_t bytes = count * sizeof(T);
    if (bytes > stage_bytes) {
      MKP_HIP(hipStreamSynchronize(stream));
      if (stage) (void)hipHostFree(stage);
      stage = nullptr;
      stage_bytes = 0;
      size_t b = 4096;
      while (b < bytes) b <<= 1;
      MKP_HIP(hipHostMalloc(reinterpret_cast<void**>(&stage), b, hipHostMallocDefault));
      stage_bytes = b;
    }
    return reinterpret_cast<T*>(stage);
  }
  // Host waits: spinning (hipStreamSynchronize) gives the lowest latency to
  // one caller; with `sleep_wait` set, a wait polls the stream for 20 us and
  // then sleeps between polls, so that threads waiting on several contexts
  // leave the host cores to other work.
  bool sleep_wait = false;
  hipEvent_t ev_wait = nullptr;
  void wait_stream() {
    if (!sleep_wait || !ev_wait) {
      MKP_HIP(hipStreamSynchronize(stream));
      return;
    }
    MKP_HIP(hipEventRecord(ev_wait, stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
      const hipError_t q = hipEventQuery(ev_wait);
      if (q == hipSuccess) return;
      if (q != hipErrorNotReady) MKP_HIP(q);
      if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(20))
        std::this_thread::sleep_for(std::chrono::microseconds(40));
    }
  }
  void sync() {
    wait_stream();
    if (!prof.open.empty()) prof.collect();
  }
  void* temp(size_t bytes) { tmp.ensure(bytes ? bytes : 1); return tmp.p; }
  std::shared_ptr<DevicePool> pool = std::make_shared<DevicePool>();
  template <typename T>
  T* scratch_as(int i, size_t count) { scratch[i].ensure(count * sizeof(T)); return reinterpret_cast<T*>(scratch[i].p); }
};

// Entry-point scope on a context: its device is current and device buffers
// allocated on this thread come from its pool.
struct Bind {
  std::shared_ptr<DevicePool> prev;
  explicit Bind(Context& c) : prev(t_pool) {
    MKP_HIP(hipSetDevice(c.device));
    t_pool = c.pool;
  }
  ~Bind() { t_pool = prev; }
};

template <typename T>
__global__ void fill_kernel(T* p, T v, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}
// Fill by kernel rather than hipMemsetAsync, so the same code is safe inside
// captured graphs.
template <typename T>
void fill_async(T* p, T v, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  uint64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  fill_kernel<T><<<unsigned(g), 256, 0, s>>>(p, v, n);
}

inline unsigned grid_for(uint64_t work, unsigned block, unsigned max_blocks = 1u << 20) {
  uint64_t g = (work + block - 1) / block;
  if (g == 0) g = 1;
  return unsigned(g < max_blocks ? g : max_blocks);
}

}  // namespace mkp
