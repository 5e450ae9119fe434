// Shared definitions of the gfx950 implementation: error handling, device
// buffers, the context (device, stream, scratch), and the stateless hash that
// defines tie-breaking of every synchronous round.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
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

// round tags (one per decision family)
constexpr uint64_t kLpBucketTag = 0x4C50420000000000ull;
constexpr uint64_t kLpCoinTag = 0x4C50430000000000ull;
constexpr uint64_t kLpPrioTag = 0x4C50500000000000ull;
constexpr uint64_t kLvBucketTag = 0x4C56420000000000ull;
constexpr uint64_t kCoBucketTag = 0x434F420000000000ull;
constexpr uint64_t kCoTieTag = 0x434F540000000000ull;
constexpr uint64_t kCoPrioTag = 0x434F500000000000ull;

// ------------------------------------------------------------ buffers --
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) MKP_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  void ensure(size_t count) { if (count > n) alloc(count); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
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
  kNumFamilies
};

struct Profiler {
  bool enabled = false;
  struct Rec { hipEvent_t a, b; int fam; double bytes; };
  std::vector<Rec> open;
  std::vector<hipEvent_t> pool;
  double ms[kNumFamilies] = {};
  double bytes[kNumFamilies] = {};
  int64_t launches[kNumFamilies] = {};
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
      }
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    open.clear();
  }
  void reset() {
    for (int i = 0; i < kNumFamilies; ++i) { ms[i] = 0; bytes[i] = 0; launches[i] = 0; }
  }
};

// Scoped event pair around the launches of one kernel family.
struct TimedRegion {
  Profiler& p;
  hipStream_t s;
  int fam;
  double bytes;
  hipEvent_t a = nullptr;
  TimedRegion(Profiler& prof, hipStream_t st, int f, double b) : p(prof), s(st), fam(f), bytes(b) {
    if (p.enabled) { a = p.take(); (void)hipEventRecord(a, s); }
  }
  ~TimedRegion() {
    if (a) {
      hipEvent_t b = p.take();
      (void)hipEventRecord(b, s);
      p.open.push_back({a, b, fam, bytes});
    }
  }
};

// ------------------------------------------------------------ context --
struct Context {
  Profiler prof;
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf<uint8_t> tmp;          // rocPRIM temporary storage
  DevBuf<uint8_t> scratch[4];   // general scratch arenas
  int64_t* pinned = nullptr;    // small host readback area (64 words)
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  int num_cus = 256;
  void sync() {
    MKP_HIP(hipStreamSynchronize(stream));
    if (!prof.open.empty()) prof.collect();
  }
  void* temp(size_t bytes) { tmp.ensure(bytes ? bytes : 1); return tmp.p; }
  template <typename T>
  T* scratch_as(int i, size_t count) { scratch[i].ensure(count * sizeof(T)); return reinterpret_cast<T*>(scratch[i].p); }
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
