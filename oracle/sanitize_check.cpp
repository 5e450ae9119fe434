// Sanitizer driver (test infrastructure): runs the host code that the product
// shares with the oracle -- the km1 rebalancer -- the oracle's initial
// partitioning with its worker threads, and the oracle's pipelines in both modes (OpenMP
// threads in reference mode) on seeded random hypergraphs, so that a build
// with -fsanitize=address,undefined or -fsanitize=thread reports memory
// errors, undefined behaviour and data races of that code.  Prints "ok" and
// returns 0 when every run finished and stayed within its balance bound.
// With the argument "threads" only the std::thread code runs (initial
// partitioning and rebalancer): the thread sanitizer does not see OpenMP's
// own synchronisation, so the OpenMP pipelines are left to the address
// sanitizer build.
#include <cstdint>
#include <cstdio>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "oracle.h"
#include "../mt-kahypar_amd/host/rebalancer.h"

extern "C" {
void* orc_hg_new(uint32_t n, uint32_t m, const uint64_t* noff, const uint32_t* pins, const int32_t* vw,
                 const int32_t* nw);
void orc_hg_free(void* h);
int64_t orc_partition(void* h, int k, double eps, uint64_t seed, int sync, int threads, int fm_rounds,
                      int32_t* part_out, double* times, int64_t* counters);
}

namespace {
// nets with a size drawn from a truncated power law, pins in a window around
// a random centre (80%) or anywhere
mt_kahypar_amd::HostHypergraph random_hypergraph(uint32_t n, uint32_t m, uint32_t max_size, bool weighted,
                                                 uint64_t seed) {
  std::mt19937_64 rng(seed);
  mt_kahypar_amd::HostHypergraph h;
  h.n = n;
  h.m = m;
  h.noff.assign(1, 0);
  std::vector<uint32_t> stamp(n, 0);
  for (uint32_t e = 0; e < m; ++e) {
    const double x = std::uniform_real_distribution<double>(0.0, 1.0)(rng);
    uint32_t s = uint32_t(2.0 / std::pow(1.0 - x * 0.999, 1.0 / 1.3));
    s = std::max<uint32_t>(2, std::min(s, max_size));
    const uint32_t c = uint32_t(rng() % n);
    for (uint32_t i = 0; i < s; ++i) {
      uint32_t v = (rng() % 10 < 8) ? uint32_t((c + rng() % 256) % n) : uint32_t(rng() % n);
      while (stamp[v] == e + 1) v = (v + 1) % n;
      stamp[v] = e + 1;
      h.pins.push_back(v);
    }
    h.noff.push_back(h.pins.size());
  }
  h.vw.assign(n, 1);
  if (weighted)
    for (auto& w : h.vw) w = 1 + int32_t(rng() % 8);
  h.nw.assign(m, 1);
  mt_kahypar_amd::build_host_incidence(h);
  return h;
}

bool balanced(const mt_kahypar_amd::HostHypergraph& h, int k, const std::vector<int32_t>& part, int64_t maxw) {
  std::vector<int64_t> w(k, 0);
  for (uint32_t u = 0; u < h.n; ++u) w[part[u]] += h.vw[u];
  for (int64_t x : w)
    if (x > maxw) return false;
  return true;
}
}  // namespace

int main(int argc, char** argv) {
  const bool threads_only = argc > 1 && std::string(argv[1]) == "threads";
  int failures = 0;
  // initial partitioning with worker threads: bisections of at least
  // 20,000 pins run their runs on threads
  for (int t = 0; t < 3; ++t) {
    const int k = t == 0 ? 2 : (t == 1 ? 8 : 32);
    auto h = random_hypergraph(6000, 30000, 60, t == 2, 11 + t);
    int64_t W = 0;
    for (int32_t w : h.vw) W += w;
    const int64_t perfect = (W + k - 1) / k, maxw = int64_t(1.03 * double(perfect));
    orc::IpParams ip;
    ip.k = k;
    ip.max_w.assign(k, maxw);
    ip.seed = 5 + t;
    ip.threads = 8;
    int64_t km1 = 0;
    const orc::Hypergraph oh = orc::build_hypergraph(h.n, h.m, h.noff.data(), h.pins.data(), h.vw.data(),
                                                     h.nw.data());
    std::vector<int32_t> part = orc::initial_partition(oh, ip, &km1);
    if (!balanced(h, k, part, maxw)) { std::printf("initial partition over its bound (k=%d)\n", k); ++failures; }
    // rebalancer on a partition pushed out of balance
    for (uint32_t u = 0; u < h.n / 4; ++u) part[u] = 0;
    mt_kahypar_amd::RebalanceResult rr = mt_kahypar_amd::rebalance_km1(h, k, ip.max_w, part);
    (void)rr;
  }
  // the oracle's whole pipeline, synchronous and reference mode (OpenMP)
  for (int sync = 0; sync < 2 && !threads_only; ++sync) {
    auto h = random_hypergraph(20000, 24000, 300, sync == 1, 21 + sync);
    void* oh = orc_hg_new(h.n, h.m, h.noff.data(), h.pins.data(), h.vw.data(), h.nw.data());
    std::vector<int32_t> part(h.n);
    std::vector<double> times(5);
    std::vector<int64_t> cnt(7);
    const int k = sync ? 4 : 8;
    orc_partition(oh, k, 0.03, 42, sync, 8, sync ? 10 : 0, part.data(), times.data(), cnt.data());
    int64_t W = 0;
    for (int32_t w : h.vw) W += w;
    const int64_t perfect = (W + k - 1) / k;
    if (!balanced(h, k, part, int64_t(1.03 * double(perfect)))) {
      std::printf("pipeline over its bound (sync=%d)\n", sync);
      ++failures;
    }
    orc_hg_free(oh);
  }
  std::printf(failures ? "failed\n" : "ok\n");
  return failures ? 1 : 0;
}
