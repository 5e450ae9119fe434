// hMetis reader and writer (io/hypergraph_io.cpp:95-177,306-333).
#include "hgr_io.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <fstream>
#include <stdexcept>

#include "../../include/mtkahypar_amd.h"
#include "../csrc/common.h"

namespace mt_kahypar_amd {
namespace {

// Read-only view of a whole file (mapped, so gigabyte inputs are not copied).
struct MappedFile {
  const char* data = nullptr;
  size_t size = 0;
  int fd = -1;
  explicit MappedFile(const std::string& path) {
    fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("could not open " + path);
    struct stat st;
    if (fstat(fd, &st) != 0) { ::close(fd); throw std::runtime_error("could not stat " + path); }
    size = size_t(st.st_size);
    if (size) {
      void* p = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (p == MAP_FAILED) { ::close(fd); throw std::runtime_error("could not map " + path); }
      data = static_cast<const char*>(p);
    }
  }
  ~MappedFile() {
    if (data) munmap(const_cast<char*>(data), size);
    if (fd >= 0) ::close(fd);
  }
};

// Line cursor that skips '%' comment lines.
class Lines {
 public:
  Lines(const MappedFile& f, const std::string& path) : p_(f.data), end_(f.data + f.size), path_(path) {}
  // Next non-comment line as [b, e); false at end of file.
  bool next(const char*& b, const char*& e) {
    while (p_ < end_) {
      const char* s = p_;
      const char* nl = static_cast<const char*>(memchr(s, '\n', size_t(end_ - s)));
      const char* t = nl ? nl : end_;
      p_ = nl ? nl + 1 : end_;
      ++line_;
      if (s < t && *s == '%') continue;
      b = s;
      e = t;
      return true;
    }
    return false;
  }
  [[noreturn]] void fail(const std::string& what) const {
    throw std::runtime_error(path_ + ":" + std::to_string(line_) + ": " + what);
  }

 private:
  const char* p_;
  const char* end_;
  std::string path_;
  uint64_t line_ = 0;
};

// Parses unsigned decimal numbers separated by blanks; false at line end.
bool next_number(const char*& p, const char* e, uint64_t& v, const Lines& L) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  if (p == e) return false;
  if (*p < '0' || *p > '9') L.fail("expected a non-negative integer");
  uint64_t x = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    x = x * 10 + uint64_t(*p - '0');
    if (x > (uint64_t(1) << 40)) L.fail("number out of range");
    ++p;
  }
  if (p < e && !(*p == ' ' || *p == '\t' || *p == '\r')) L.fail("unexpected character");
  v = x;
  return true;
}

}  // namespace

HgrFile read_hgr(const std::string& path) {
  MappedFile f(path);
  Lines L(f, path);
  HgrFile out;
  const char *b, *e;
  if (!L.next(b, e)) L.fail("missing header");
  uint64_t E = 0, N = 0, type = 0;
  if (!next_number(b, e, E, L) || !next_number(b, e, N, L)) L.fail("header needs 'num_nets num_vertices'");
  if (next_number(b, e, type, L) && type != 0 && type != 1 && type != 10 && type != 11)
    L.fail("unknown hypergraph type " + std::to_string(type));
  if (N >= (uint64_t(1) << 32) || E >= (uint64_t(1) << 32)) L.fail("more than 2^32 - 1 vertices or nets");
  const bool net_w = type == 1 || type == 11;
  const bool vtx_w = type == 10 || type == 11;
  HostHypergraph& h = out.hg;
  h.n = uint32_t(N);
  h.noff.assign(1, 0);
  h.nw.reserve(E);
  std::vector<uint32_t> line_pins;
  for (uint64_t i = 0; i < E; ++i) {
    if (!L.next(b, e)) L.fail("expected " + std::to_string(E) + " net lines");
    uint64_t w = 1, v;
    if (net_w && !next_number(b, e, w, L)) w = 0;  // empty line: dropped below
    line_pins.clear();
    while (next_number(b, e, v, L)) {
      if (v < 1 || v > N) L.fail("pin " + std::to_string(v) + " outside 1.." + std::to_string(N));
      line_pins.push_back(uint32_t(v - 1));
    }
    if (line_pins.size() < 2) {  // isSinglePinHyperedge (hypergraph_io.cpp:120-137)
      ++out.removed_single_pin;
      continue;
    }
    if (w > uint64_t(INT32_MAX)) L.fail("net weight out of range");
    h.pins.insert(h.pins.end(), line_pins.begin(), line_pins.end());
    h.noff.push_back(h.pins.size());
    h.nw.push_back(int32_t(w));
  }
  h.m = uint32_t(h.nw.size());
  h.vw.assign(h.n, 1);
  if (vtx_w) {
    for (uint32_t u = 0; u < h.n; ++u) {
      uint64_t w;
      if (!L.next(b, e) || !next_number(b, e, w, L)) L.fail("expected " + std::to_string(N) + " vertex weights");
      if (w > uint64_t(INT32_MAX)) L.fail("vertex weight out of range");
      h.vw[u] = int32_t(w);
    }
  }
  out.type = int(type);
  build_host_incidence(h);
  return out;
}

void write_hgr(const HostHypergraph& h, const std::string& path) {
  bool net_w = false, vtx_w = false;
  for (int32_t w : h.nw) net_w |= w != 1;
  for (int32_t w : h.vw) vtx_w |= w != 1;
  std::ofstream o(path);
  if (!o) throw std::runtime_error("could not write " + path);
  o << h.m << ' ' << h.n;
  if (net_w || vtx_w) o << ' ' << (vtx_w ? 10 : 0) + (net_w ? 1 : 0);
  o << '\n';
  for (uint32_t e = 0; e < h.m; ++e) {
    if (net_w) o << h.nw[e] << ' ';
    for (uint64_t j = h.noff[e]; j < h.noff[e + 1]; ++j) o << (j > h.noff[e] ? " " : "") << h.pins[j] + 1;
    o << '\n';
  }
  if (vtx_w)
    for (uint32_t u = 0; u < h.n; ++u) o << h.vw[u] << '\n';
}

std::vector<int32_t> read_partition_file(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("could not open " + path);
  std::vector<int32_t> part;
  long long b;
  while (in >> b) part.push_back(int32_t(b));
  return part;
}

void write_partition_file(const std::vector<int32_t>& part, const std::string& path) {
  std::ofstream o(path);
  if (!o) throw std::runtime_error("could not write " + path);
  for (int32_t b : part) o << b << '\n';
}

}  // namespace mt_kahypar_amd

using mt_kahypar_amd::HgrFile;

extern "C" {

int mkp_read_hgr(const char* path, mkp_host_hypergraph** out, uint64_t* removed_single_pin_nets) {
  return mkp::guarded([&] {
    MKP_CHECK(path && out, "null argument");
    HgrFile f = mt_kahypar_amd::read_hgr(path);
    auto* hh = new mkp_host_hypergraph();
    hh->h = std::move(f.hg);
    if (removed_single_pin_nets) *removed_single_pin_nets = f.removed_single_pin;
    *out = hh;
  });
}

int mkp_write_hgr(const mkp_host_hypergraph* h, const char* path) {
  return mkp::guarded([&] {
    MKP_CHECK(h && path, "null argument");
    mt_kahypar_amd::write_hgr(h->h, path);
  });
}

int mkp_write_partition_file(const int32_t* parts, uint32_t n, const char* path) {
  return mkp::guarded([&] {
    MKP_CHECK((parts || n == 0) && path, "null argument");
    mt_kahypar_amd::write_partition_file(std::vector<int32_t>(parts, parts + n), path);
  });
}

}  // extern "C"
