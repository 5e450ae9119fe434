// The reference's C library interface (lib/libkahypar.cc:44-200) over the
// device pipeline.
#include "../../include/libkahypar.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mtkahypar_amd.h"
#include "../csrc/common.h"
#include "hgr_io.h"

struct mt_kahypar_context_s {
  mkp_partition_params prm{};  // fast preset; k, epsilon and seed are set per call
  int device = 0;
};

namespace {

std::mutex g_lock;                 // partition() is serialised
std::map<int, mkp_context*> g_ctx;  // one device context per GPU, kept until exit
int g_return_on_error = 0;

// ERROR(msg) of macros.h:108: print and exit(-1); or return (extension).
bool fail(const std::string& msg) {
  mkp::set_last_error(msg);
  if (!g_return_on_error) {
    std::fprintf(stderr, "[ERROR] %s\n", msg.c_str());
    std::exit(-1);
  }
  return false;
}

mkp_context* device_context(int device) {
  auto it = g_ctx.find(device);
  if (it != g_ctx.end()) return it->second;
  mkp_context* c = nullptr;
  if (mkp_context_create(device, &c) != 0) throw std::runtime_error(mkp_last_error());
  g_ctx[device] = c;
  return c;
}

std::string trim(const std::string& s) {
  const size_t b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return "";
  const size_t e = s.find_last_not_of(" \t\r");
  return s.substr(b, e - b + 1);
}

bool parse_bool(const std::string& v) {
  if (v == "true" || v == "1") return true;
  if (v == "false" || v == "0") return false;
  throw std::runtime_error("expected true/false, got '" + v + "'");
}

// Applies one key of an ini preset (context option names of
// command_line_options.cpp); keys of components outside this path are ignored.
void apply_option(mkp_partition_params& p, const std::string& key, const std::string& v) {
  if (key == "maxnet-ignore") p.ignore_net_size = uint32_t(std::stoul(v));
  else if (key == "p-enable-community-detection") p.use_community_detection = parse_bool(v);
  else if (key == "p-louvain-edge-weight-function") {
    if (v == "hybrid") p.louvain_edge_weight = -1;
    else if (v == "uniform") p.louvain_edge_weight = 0;
    else if (v == "non_uniform") p.louvain_edge_weight = 1;
    else if (v == "degree") p.louvain_edge_weight = 2;
    else throw std::runtime_error("unknown edge weight function '" + v + "'");
  } else if (key == "p-max-louvain-pass-iterations") p.louvain_max_passes = std::stoi(v);
  else if (key == "p-louvain-min-vertex-move-fraction") p.louvain_min_move_fraction = std::stod(v);
  else if (key == "p-vertex-degree-sampling-threshold") p.louvain_sampling = std::stoull(v);
  else if (key == "c-type") {
    if (v != "multilevel_coarsener") throw std::runtime_error("only c-type=multilevel_coarsener is available");
  } else if (key == "c-min-shrink-factor") p.min_shrink_factor = std::stod(v);
  else if (key == "c-max-shrink-factor") p.max_shrink_factor = std::stod(v);
  else if (key == "c-s") p.max_allowed_weight_multiplier = std::stod(v);
  else if (key == "c-t") p.contraction_limit_multiplier = uint32_t(std::stoul(v));
  else if (key == "c-vertex-degree-sampling-threshold") p.rating_sampling = std::stoull(v);
  else if (key == "i-runs") p.ip_runs = std::stoi(v);
  else if (key == "maxnet-remove-factor") p.large_net_remove_factor = std::stod(v);
  else if (key == "r-lp-type") {
    if (v == "label_propagation_km1") p.lp_enabled = 1;
    else if (v == "do_nothing") p.lp_enabled = 0;
    else throw std::runtime_error("unknown r-lp-type '" + v + "'");
  } else if (key == "r-lp-maximum-iterations") p.lp_max_rounds = std::stoi(v);
  else if (key == "r-lp-rebalancing") p.lp_rebalancing = parse_bool(v);
  else if (key == "r-lp-he-size-activation-threshold") p.lp_activation_threshold = uint32_t(std::stoul(v));
  else if (key == "r-fm-type") {
    if (v == "do_nothing") p.fm_rounds = 0;
    else if (v == "fm_gain_cache") p.fm_rounds = p.fm_rounds > 0 ? p.fm_rounds : 10;
    else throw std::runtime_error("unknown r-fm-type '" + v + "'");
  } else if (key == "r-fm-multitry-rounds") p.fm_rounds = std::stoi(v);
  else if (key == "r-fm-seed-nodes") p.fm_seeds = std::stoi(v);
}

}  // namespace

extern "C" {

mt_kahypar_context_t* mt_kahypar_context_new(void) {
  auto* c = new mt_kahypar_context_s();
  mkp_partition_params_default(&c->prm, 2, 0.03, 0);
  return c;
}

void mt_kahypar_context_free(mt_kahypar_context_t* c) { delete c; }

void mt_kahypar_configure_context_from_file(mt_kahypar_context_t* c, const char* ini_file_name) {
  if (!c || !ini_file_name) { fail("null argument"); return; }
  std::ifstream in(ini_file_name);
  if (!in) { fail(std::string("could not open ") + ini_file_name); return; }
  mkp_partition_params p = c->prm;
  std::string line;
  int lineno = 0;
  try {
    while (std::getline(in, line)) {
      ++lineno;
      const size_t hash = line.find('#');
      if (hash != std::string::npos) line.resize(hash);
      line = trim(line);
      if (line.empty()) continue;
      const size_t eq = line.find('=');
      if (eq == std::string::npos) throw std::runtime_error("expected key=value");
      apply_option(p, trim(line.substr(0, eq)), trim(line.substr(eq + 1)));
    }
  } catch (const std::exception& e) {
    fail(std::string(ini_file_name) + ":" + std::to_string(lineno) + ": " + e.what());
    return;
  }
  c->prm = p;
}

void mt_kahypar_initialize_thread_pool(const size_t, const bool) {
  // Every phase runs on the device: there is no host thread pool to size.
}

void mt_kahypar_read_hypergraph_from_file(const char* file_name, mt_kahypar_hypernode_id_t* num_vertices,
                                          mt_kahypar_hyperedge_id_t* num_hyperedges, size_t** hyperedge_indices,
                                          mt_kahypar_hyperedge_id_t** hyperedges,
                                          mt_kahypar_hyperedge_weight_t** hyperedge_weights,
                                          mt_kahypar_hypernode_weight_t** vertex_weights) {
  if (!file_name || !num_vertices || !num_hyperedges || !hyperedge_indices || !hyperedges || !hyperedge_weights ||
      !vertex_weights) {
    fail("null argument");
    return;
  }
  mt_kahypar_amd::HgrFile f;
  try {
    f = mt_kahypar_amd::read_hgr(file_name);
  } catch (const std::exception& e) {
    fail(e.what());
    return;
  }
  const auto& h = f.hg;
  *num_vertices = h.n;
  *num_hyperedges = h.m;
  *hyperedge_indices = new size_t[size_t(h.m) + 1];
  for (uint32_t e = 0; e <= h.m; ++e) (*hyperedge_indices)[e] = size_t(h.noff[e]);
  *hyperedges = new mt_kahypar_hyperedge_id_t[std::max<size_t>(h.pins.size(), 1)];
  for (size_t i = 0; i < h.pins.size(); ++i) (*hyperedges)[i] = h.pins[i];
  *hyperedge_weights = new mt_kahypar_hyperedge_weight_t[std::max<size_t>(h.m, 1)];
  if (h.m) std::memcpy(*hyperedge_weights, h.nw.data(), size_t(h.m) * sizeof(int));
  *vertex_weights = new mt_kahypar_hypernode_weight_t[std::max<size_t>(h.n, 1)];
  if (h.n) std::memcpy(*vertex_weights, h.vw.data(), size_t(h.n) * sizeof(int));
}

void mt_kahypar_partition(const mt_kahypar_hypernode_id_t num_vertices, const mt_kahypar_hyperedge_id_t num_hyperedges,
                          const double epsilon, const mt_kahypar_partition_id_t num_blocks, const int seed,
                          const mt_kahypar_hypernode_weight_t* vertex_weights,
                          const mt_kahypar_hyperedge_weight_t* hyperedge_weights, const size_t* hyperedge_indices,
                          const mt_kahypar_hyperedge_id_t* hyperedges, mt_kahypar_hyperedge_weight_t* objective,
                          mt_kahypar_context_t* kahypar_context, mt_kahypar_partition_id_t* partition,
                          const bool verbose) {
  if (objective) *objective = -1;
  if (!kahypar_context || !partition || !objective || (num_hyperedges && (!hyperedge_indices || !hyperedges))) {
    fail("null argument");
    return;
  }
  // internal ids are 32-bit (hypergraph_common.h:39)
  if (num_vertices >= (1ul << 32) || num_hyperedges >= (1ul << 32)) { fail("more than 2^32 - 1 vertices or nets"); return; }
  if (num_blocks < 2 || num_blocks > (1u << 20)) { fail("number of blocks must be at least 2"); return; }
  if (!(epsilon >= 0)) { fail("epsilon must be non-negative"); return; }
  const uint32_t n = uint32_t(num_vertices), m = uint32_t(num_hyperedges);
  std::vector<uint64_t> noff(size_t(m) + 1, 0);
  const size_t P = m ? hyperedge_indices[m] - hyperedge_indices[0] : 0;
  std::vector<uint32_t> pins(P);
  for (uint32_t e = 0; e <= m; ++e) noff[e] = uint64_t(hyperedge_indices[e] - hyperedge_indices[0]);
  for (uint32_t e = 0; e < m; ++e) {
    if (hyperedge_indices[e + 1] < hyperedge_indices[e]) { fail("hyperedge indices must be non-decreasing"); return; }
  }
  for (size_t i = 0; i < P; ++i) {
    const mt_kahypar_hyperedge_id_t v = hyperedges[hyperedge_indices[0] + i];
    if (v >= num_vertices) { fail("pin id out of range"); return; }
    pins[i] = uint32_t(v);
  }
  std::lock_guard<std::mutex> guard(g_lock);
  mkp_partition_params prm = kahypar_context->prm;
  prm.k = int32_t(num_blocks);
  prm.epsilon = epsilon;
  prm.seed = uint64_t(int64_t(seed));
  std::vector<int32_t> parts(std::max<uint32_t>(n, 1));
  mkp_partition_stats st{};
  mkp_hypergraph* hg = nullptr;
  try {
    mkp_context* ctx = device_context(kahypar_context->device);
    if (mkp_hypergraph_create(ctx, n, m, noff.data(), pins.data(), vertex_weights, hyperedge_weights, &hg) != 0)
      throw std::runtime_error(mkp_last_error());
    if (mkp_partition_hypergraph(hg, &prm, parts.data(), &st) != 0) throw std::runtime_error(mkp_last_error());
  } catch (const std::exception& e) {
    if (hg) mkp_hypergraph_destroy(hg);
    fail(e.what());
    return;
  }
  mkp_hypergraph_destroy(hg);
  for (uint32_t u = 0; u < n; ++u) partition[u] = mt_kahypar_partition_id_t(parts[u]);
  *objective = mt_kahypar_hyperedge_weight_t(st.km1);
  if (verbose) {
    std::fprintf(stdout,
                 "km1 = %lld, imbalance = %.4f, levels = %d, time = %.1f ms (louvain %.1f, coarsening %.1f, "
                 "initial %.1f, refinement %.1f)\n",
                 static_cast<long long>(st.km1), st.imbalance, st.levels, st.total_ms, st.louvain_ms,
                 st.coarsening_ms, st.initial_ms, st.refinement_ms);
  }
}

void mt_kahypar_amd_set_error_mode(int return_on_error) { g_return_on_error = return_on_error; }

void mt_kahypar_amd_set_device(mt_kahypar_context_t* c, int device) {
  if (c) c->device = device;
}

int mt_kahypar_amd_get_params(const mt_kahypar_context_t* c, void* params) {
  if (!c || !params) return -1;
  std::memcpy(params, &c->prm, sizeof(mkp_partition_params));
  return 0;
}

}  // extern "C"
