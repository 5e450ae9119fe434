// Device-wide primitives on rocPRIM: radix sort and prefix sums.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "internal.h"

namespace mkp {

void sort_keys_u64(Context& c, const uint64_t* in, uint64_t* out, size_t n, int bb, int eb) {
  if (n == 0) return;
  size_t bytes = 0;
  MKP_HIP(rocprim::radix_sort_keys(nullptr, bytes, in, out, n, unsigned(bb), unsigned(eb), c.stream));
  void* t = c.temp(bytes);
  MKP_HIP(rocprim::radix_sort_keys(t, bytes, in, out, n, unsigned(bb), unsigned(eb), c.stream));
}

void sort_pairs_u64_u32(Context& c, const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout,
                        size_t n, int bb, int eb) {
  if (n == 0) return;
  size_t bytes = 0;
  MKP_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, unsigned(bb), unsigned(eb), c.stream));
  void* t = c.temp(bytes);
  MKP_HIP(rocprim::radix_sort_pairs(t, bytes, kin, kout, vin, vout, n, unsigned(bb), unsigned(eb), c.stream));
}

template <typename T>
static void excl_scan(Context& c, const T* in, T* out, size_t n) {
  // out[0..n-1] exclusive, out[n] = total: scan n+1 values where in[n] := 0
  // is emulated by an inclusive scan shifted by one.
  if (n == 0) {
    fill_async<T>(out, T(0), 1, c.stream);
    return;
  }
  size_t bytes = 0;
  MKP_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out + 1, n, rocprim::plus<T>(), c.stream));
  void* t = c.temp(bytes);
  MKP_HIP(rocprim::inclusive_scan(t, bytes, in, out + 1, n, rocprim::plus<T>(), c.stream));
  fill_async<T>(out, T(0), 1, c.stream);
}

void exclusive_scan_u32(Context& c, const uint32_t* in, uint32_t* out, size_t n) { excl_scan(c, in, out, n); }
void exclusive_scan_u64(Context& c, const uint64_t* in, uint64_t* out, size_t n) { excl_scan(c, in, out, n); }

void inclusive_max_scan_u64(Context& c, const uint64_t* in, uint64_t* out, size_t n) {
  if (n == 0) return;
  size_t bytes = 0;
  MKP_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, n, rocprim::maximum<uint64_t>(), c.stream));
  void* t = c.temp(bytes);
  MKP_HIP(rocprim::inclusive_scan(t, bytes, in, out, n, rocprim::maximum<uint64_t>(), c.stream));
}

void inclusive_scan_i64(Context& c, const int64_t* in, int64_t* out, size_t n) {
  if (n == 0) return;
  size_t bytes = 0;
  MKP_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, n, rocprim::plus<int64_t>(), c.stream));
  void* t = c.temp(bytes);
  MKP_HIP(rocprim::inclusive_scan(t, bytes, in, out, n, rocprim::plus<int64_t>(), c.stream));
}

}  // namespace mkp
