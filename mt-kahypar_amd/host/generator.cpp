// Seeded synthetic hypergraphs (host side).  A single SplitMix64 stream
// makes every input a pure function of its parameters.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mtkahypar_amd.h"
#include "../csrc/common.h"
#include "host_hypergraph.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

// discrete power law pmf ~ s^-alpha on [lo, hi]
struct PowerLaw {
  uint32_t lo;
  std::vector<double> cdf;
  PowerLaw(double alpha, uint32_t lo_, uint32_t hi) : lo(lo_) {
    double acc = 0;
    for (uint32_t s = lo; s <= hi; ++s) { acc += std::pow(double(s), -alpha); cdf.push_back(acc); }
    for (auto& x : cdf) x /= acc;
  }
  uint32_t sample(SplitMix64& r) const {
    const double u = r.uniform();
    return lo + uint32_t(std::lower_bound(cdf.begin(), cdf.end() - 1, u) - cdf.begin());
  }
  double mean() const {
    double m = 0, prev = 0;
    for (size_t i = 0; i < cdf.size(); ++i) { m += (cdf[i] - prev) * double(lo + i); prev = cdf[i]; }
    return m;
  }
};

}  // namespace

using namespace mkp;

extern "C" {

int mkp_generate_powerlaw(const mkp_powerlaw_params* prm, mkp_host_hypergraph** out) {
  return guarded([&] {
    MKP_CHECK(prm && out, "null argument");
    MKP_CHECK(prm->num_vertices >= 2 && prm->min_size >= 2 && prm->max_size >= prm->min_size, "bad sizes");
    const uint32_t N = prm->num_vertices;
    SplitMix64 rng(prm->seed);
    PowerLaw pl(prm->exponent, prm->min_size, std::min(prm->max_size, N));
    auto* hh = new mkp_host_hypergraph();
    auto& h = hh->h;
    h.n = N;
    h.m = prm->num_nets;
    h.noff.assign(1, 0);
    h.pins.reserve(size_t(prm->num_nets) * 8);
    std::vector<uint32_t> stamp(N, 0);
    for (uint32_t e = 0; e < prm->num_nets; ++e) {
      const uint32_t s = pl.sample(rng);
      const uint32_t c = uint32_t(rng.below(N));
      for (uint32_t i = 0; i < s; ++i) {
        uint32_t v = 0;
        bool placed = false;
        for (int attempt = 0; attempt < 16 && !placed; ++attempt) {
          if (prm->window && rng.uniform() < prm->local_fraction) {
            const int64_t off = int64_t(rng.below(prm->window)) - int64_t(prm->window / 2);
            v = uint32_t(((int64_t(c) + off) % int64_t(N) + int64_t(N)) % int64_t(N));
          } else {
            v = uint32_t(rng.below(N));
          }
          placed = stamp[v] != e + 1;
        }
        while (!placed) {  // dense fallback: next unused id
          v = (v + 1) % N;
          placed = stamp[v] != e + 1;
        }
        stamp[v] = e + 1;
        h.pins.push_back(v);
      }
      h.noff.push_back(h.pins.size());
    }
    h.vw.assign(N, 1);
    if (prm->vertex_weight_max > prm->vertex_weight_min && prm->vertex_weight_min >= 1) {
      const uint64_t span = uint64_t(prm->vertex_weight_max - prm->vertex_weight_min + 1);
      for (auto& w : h.vw) w = prm->vertex_weight_min + int32_t(rng.below(span));
    } else if (prm->vertex_weight_min > 1) {
      std::fill(h.vw.begin(), h.vw.end(), prm->vertex_weight_min);
    }
    h.nw.assign(h.m, 1);
    *out = hh;
  });
}

int mkp_generate_rownet(const mkp_rownet_params* prm, mkp_host_hypergraph** out) {
  return guarded([&] {
    MKP_CHECK(prm && out, "null argument");
    MKP_CHECK(prm->num_rows >= 2 && prm->num_cols >= 2 && prm->max_nnz >= 1, "bad sizes");
    // exponent with the requested mean on [1, max_nnz]
    double lo = 1.01, hi = 6.0;
    for (int it = 0; it < 60; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (PowerLaw(mid, 1, prm->max_nnz).mean() > prm->mean_nnz) lo = mid; else hi = mid;
    }
    PowerLaw pl(0.5 * (lo + hi), 1, std::min(prm->max_nnz, prm->num_cols));
    SplitMix64 rng(prm->seed);
    const uint32_t R = prm->num_rows, Cn = prm->num_cols;
    std::vector<uint64_t> roff(1, 0);
    std::vector<uint32_t> cols;
    cols.reserve(size_t(R) * 9);
    std::vector<uint32_t> stamp(Cn, 0);
    for (uint32_t r = 0; r < R; ++r) {
      const uint32_t nnz = pl.sample(rng);
      const uint64_t diag = uint64_t(r) * Cn / R;
      for (uint32_t i = 0; i < nnz; ++i) {
        uint32_t c = 0;
        bool placed = false;
        for (int attempt = 0; attempt < 16 && !placed; ++attempt) {
          if (prm->band && rng.uniform() < prm->band_fraction) {
            const int64_t off = int64_t(rng.below(prm->band)) - int64_t(prm->band / 2);
            c = uint32_t(((int64_t(diag) + off) % int64_t(Cn) + int64_t(Cn)) % int64_t(Cn));
          } else {
            c = uint32_t(rng.below(Cn));
          }
          placed = stamp[c] != r + 1;
        }
        while (!placed) { c = (c + 1) % Cn; placed = stamp[c] != r + 1; }
        stamp[c] = r + 1;
        cols.push_back(c);
      }
      roff.push_back(cols.size());
    }
    // columns become nets; empty and single-pin columns are dropped
    // (io/hypergraph_io.cpp:121-137)
    std::vector<uint64_t> ccount(size_t(Cn) + 1, 0);
    for (uint32_t c : cols) ccount[c + 1]++;
    for (uint32_t c = 0; c < Cn; ++c) ccount[c + 1] += ccount[c];
    std::vector<uint32_t> rows(cols.size());
    {
      std::vector<uint64_t> pos(ccount.begin(), ccount.end() - 1);
      for (uint32_t r = 0; r < R; ++r)
        for (uint64_t i = roff[r]; i < roff[r + 1]; ++i) rows[pos[cols[i]]++] = r;
    }
    auto* hh = new mkp_host_hypergraph();
    auto& h = hh->h;
    h.n = R;
    h.noff.assign(1, 0);
    for (uint32_t c = 0; c < Cn; ++c) {
      if (ccount[c + 1] - ccount[c] < 2) continue;
      h.pins.insert(h.pins.end(), rows.begin() + ccount[c], rows.begin() + ccount[c + 1]);
      h.noff.push_back(h.pins.size());
    }
    h.m = uint32_t(h.noff.size() - 1);
    h.vw.resize(R);
    for (uint32_t r = 0; r < R; ++r) h.vw[r] = int32_t(roff[r + 1] - roff[r]);
    h.nw.assign(h.m, 1);
    *out = hh;
  });
}

int mkp_host_hypergraph_info(const mkp_host_hypergraph* hh, mkp_hypergraph_info* info) {
  return guarded([&] {
    MKP_CHECK(hh && info, "null argument");
    const auto& h = hh->h;
    info->num_vertices = h.n;
    info->num_nets = h.m;
    info->num_pins = h.pins.size();
    uint32_t me = 0;
    for (uint32_t e = 0; e < h.m; ++e) me = std::max<uint32_t>(me, uint32_t(h.noff[e + 1] - h.noff[e]));
    info->max_net_size = me;
    int64_t tw = 0;
    for (int32_t w : h.vw) tw += w;
    info->total_weight = tw;
  });
}

int mkp_host_hypergraph_copy(const mkp_host_hypergraph* hh, uint64_t* noff, uint32_t* pins, int32_t* vw,
                             int32_t* nw) {
  return guarded([&] {
    MKP_CHECK(hh, "null argument");
    const auto& h = hh->h;
    if (noff) std::memcpy(noff, h.noff.data(), h.noff.size() * 8);
    if (pins && !h.pins.empty()) std::memcpy(pins, h.pins.data(), h.pins.size() * 4);
    if (vw && h.n) std::memcpy(vw, h.vw.data(), size_t(h.n) * 4);
    if (nw && h.m) std::memcpy(nw, h.nw.data(), size_t(h.m) * 4);
  });
}

int mkp_host_hypergraph_create(uint32_t n, uint32_t m, const uint64_t* noff, const uint32_t* pins, const int32_t* vw,
                               const int32_t* nw, mkp_host_hypergraph** out) {
  return guarded([&] {
    MKP_CHECK(out && noff, "null argument");
    MKP_CHECK(noff[0] == 0, "net offsets must start at 0");
    const uint64_t p = noff[m];
    MKP_CHECK(p == 0 || pins, "pins is null");
    for (uint32_t e = 0; e < m; ++e) MKP_CHECK(noff[e + 1] >= noff[e], "net offsets are not monotone");
    for (uint64_t i = 0; i < p; ++i) MKP_CHECK(pins[i] < n, "pin id out of range");
    if (vw)
      for (uint32_t u = 0; u < n; ++u) MKP_CHECK(vw[u] >= 0, "negative vertex weight");
    auto* hh = new mkp_host_hypergraph();
    auto& h = hh->h;
    h.n = n;
    h.m = m;
    h.noff.assign(noff, noff + size_t(m) + 1);
    h.pins.assign(pins, pins + p);
    h.vw = vw ? std::vector<int32_t>(vw, vw + n) : std::vector<int32_t>(n, 1);
    h.nw = nw ? std::vector<int32_t>(nw, nw + m) : std::vector<int32_t>(m, 1);
    build_host_incidence(h);
    *out = hh;
  });
}

int mkp_host_hypergraph_destroy(mkp_host_hypergraph* hh) {
  return guarded([&] { delete hh; });
}

int mkp_hypergraph_create_from_host(mkp_context* ctx, const mkp_host_hypergraph* hh, mkp_hypergraph** out) {
  if (!hh) {
    set_last_error("null argument");
    return 1;
  }
  const auto& h = hh->h;
  return mkp_hypergraph_create(ctx, h.n, h.m, h.noff.data(), h.pins.data(), h.vw.data(), h.nw.data(), out);
}

}  // extern "C"
