// The batched form: many independent hypergraphs stored as one disjoint union
// in HBM, so every round of every member runs in the same kernel launches.
// Each member keeps local vertex and block ids for every hash and tie-break,
// its own seed, block weights and stop state: a member's result is the one it
// gets alone.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../host/host_hypergraph.h"
#include "internal.h"

struct mkp_batch {
  mkp_hypergraph* hg = nullptr;  // the disjoint union (owned)
  uint32_t members = 0;
  std::vector<uint32_t> vbase, nbase;  // members + 1 entries
  std::vector<uint64_t> pbase;         // members + 1 entries
  std::vector<int64_t> weight;         // total vertex weight per member
  std::vector<uint8_t> plain;          // per member: has nets and all of them have two pins
  mkp::DevBuf<uint32_t> vgraph, ngraph, gvbase, gnbase;
  ~mkp_batch() { delete hg; }
};

struct mkp_batch_partition {
  mkp_batch* b = nullptr;
  mkp_partition* p = nullptr;  // union partition: k = kmax for the counter layout
  std::vector<int32_t> k;
  std::vector<uint32_t> kbase;  // members + 1 entries
  int kmax = 0;
  mkp::DevBuf<uint32_t> gkbase;
  mkp::DevBuf<int32_t> gk;
  mkp::DevBuf<uint64_t> gseed;
  ~mkp_batch_partition() { delete p; }
  mkp::LpBatch desc() const {
    mkp::LpBatch d;
    d.members = b->members;
    d.ktot = kbase.back();
    d.kmax = kmax;
    d.vgraph = b->vgraph.p;
    d.gvbase = b->gvbase.p;
    d.gkbase = gkbase.p;
    d.gk = gk.p;
    d.gseed = gseed.p;
    return d;
  }
};

using namespace mkp;

extern "C" {

int mkp_batch_create(mkp_context* ctx, uint32_t count, const mkp_host_hypergraph* const* members, mkp_batch** out) {
  return guarded([&] {
    MKP_CHECK(ctx && out && (members || count == 0), "null argument");
    MKP_CHECK(count >= 1, "a batch needs at least one member");
    Bind bind_(ctx->c);
    auto* b = new mkp_batch();
    try {
      b->members = count;
      b->vbase.assign(1, 0);
      b->nbase.assign(1, 0);
      b->pbase.assign(1, 0);
      std::vector<uint64_t> noff(1, 0);
      std::vector<uint32_t> pins, vgraph, ngraph;
      std::vector<int32_t> vw, nw;
      for (uint32_t g = 0; g < count; ++g) {
        MKP_CHECK(members[g], "null member");
        const auto& h = members[g]->h;
        const uint64_t vb = b->vbase.back();
        MKP_CHECK(vb + h.n < (uint64_t(1) << 32) && b->nbase.back() + uint64_t(h.m) < (uint64_t(1) << 32),
                  "batch exceeds 2^32 vertices or nets");
        bool two = h.m > 0;
        for (uint32_t e = 0; e < h.m; ++e) {
          for (uint64_t j = h.noff[e]; j < h.noff[e + 1]; ++j) pins.push_back(uint32_t(vb) + h.pins[j]);
          noff.push_back(pins.size());
          two = two && h.noff[e + 1] - h.noff[e] == 2;
        }
        b->plain.push_back(two ? 1 : 0);
        b->pbase.push_back(pins.size());
        int64_t w = 0;
        for (uint32_t u = 0; u < h.n; ++u) w += h.vw[u];
        vw.insert(vw.end(), h.vw.begin(), h.vw.end());
        nw.insert(nw.end(), h.nw.begin(), h.nw.end());
        vgraph.insert(vgraph.end(), h.n, g);
        ngraph.insert(ngraph.end(), h.m, g);
        b->vbase.push_back(uint32_t(vb + h.n));
        b->nbase.push_back(b->nbase.back() + h.m);
        b->weight.push_back(w);
      }
      const uint32_t N = b->vbase.back(), M = b->nbase.back();
      MKP_CHECK(mkp_hypergraph_create(ctx, N, M, noff.data(), pins.data(), vw.data(), nw.data(), &b->hg) == 0,
                mkp_last_error());
      Context& c = ctx->c;
      b->vgraph.alloc(std::max<uint32_t>(N, 1));
      b->ngraph.alloc(std::max<uint32_t>(M, 1));
      b->gvbase.alloc(count + 1);
      b->gnbase.alloc(count + 1);
      if (N) MKP_HIP(hipMemcpyAsync(b->vgraph.p, vgraph.data(), size_t(N) * 4, hipMemcpyHostToDevice, c.stream));
      if (M) MKP_HIP(hipMemcpyAsync(b->ngraph.p, ngraph.data(), size_t(M) * 4, hipMemcpyHostToDevice, c.stream));
      MKP_HIP(hipMemcpyAsync(b->gvbase.p, b->vbase.data(), size_t(count + 1) * 4, hipMemcpyHostToDevice, c.stream));
      MKP_HIP(hipMemcpyAsync(b->gnbase.p, b->nbase.data(), size_t(count + 1) * 4, hipMemcpyHostToDevice, c.stream));
      c.sync();
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int mkp_batch_destroy(mkp_batch* b) {
  return guarded([&] {
    if (!b) return;
    Bind bind_(b->hg->ctx->c);
    delete b;
  });
}

int mkp_batch_get_info(const mkp_batch* b, uint32_t* members, uint32_t* vertices, uint32_t* nets, uint64_t* pins) {
  return guarded([&] {
    MKP_CHECK(b, "null argument");
    if (members) *members = b->members;
    if (vertices) *vertices = b->hg->n;
    if (nets) *nets = b->hg->m;
    if (pins) *pins = b->hg->p;
  });
}

int mkp_batch_hypergraph(mkp_batch* b, mkp_hypergraph** hg) {
  return guarded([&] {
    MKP_CHECK(b && hg, "null argument");
    *hg = b->hg;
  });
}

int mkp_batch_louvain(mkp_batch* b, const mkp_louvain_params* prm, const uint64_t* seeds, int32_t* communities,
                      mkp_louvain_stats* stats) {
  return guarded([&] {
    MKP_CHECK(b && prm, "null argument");
    MKP_CHECK(prm->max_passes >= 0 && prm->sub_rounds >= 1 && prm->sub_rounds <= 255, "bad round parameters");
    mkp_hypergraph* hg = b->hg;
    Context& c = hg->ctx->c;
    Bind bind_(c);
    const uint32_t G = b->members;
    // every member resolves the edge weighting and the graph case as it
    // would alone (partitioner.cpp:62-72, graph.cpp:36-78); one union holds
    // one view, so these must agree across the members
    int wt = prm->edge_weight;
    bool plain = b->plain[0] != 0;
    for (uint32_t g = 0; g < G; ++g) {
      const uint32_t n = b->vbase[g + 1] - b->vbase[g], m = b->nbase[g + 1] - b->nbase[g];
      const int w = prm->edge_weight >= 0 ? prm->edge_weight
                                          : ((double(m) / double(std::max<uint32_t>(n, 1)) < 0.75) ? 2 : 0);
      if (g == 0) wt = w;
      MKP_CHECK(w == wt, "batch members resolve to different Louvain edge weightings; set edge_weight explicitly");
      MKP_CHECK((b->plain[g] != 0) == plain,
                "a batch cannot mix members whose nets all have two pins with members that have larger nets");
    }
    std::vector<uint64_t> sd(G, prm->seed);
    if (seeds) std::copy(seeds, seeds + G, sd.begin());
    LouvainBatch d;
    d.members = G;
    d.vbase = b->vbase.data();
    d.nbase = b->nbase.data();
    d.pbase = b->pbase.data();
    d.dvbase = b->gvbase.p;
    d.dnbase = b->gnbase.p;
    d.vgraph = b->vgraph.p;
    d.ngraph = b->ngraph.p;
    d.seeds = sd.data();
    d.plain = plain;
    louvain_batch(hg, *prm, wt, d, stats);
    if (communities && hg->n) {
      MKP_HIP(hipMemcpyAsync(communities, hg->comm.p, size_t(hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
      c.sync();
    }
  });
}

int mkp_batch_partition_create(mkp_batch* b, const int32_t* k, mkp_batch_partition** out) {
  return guarded([&] {
    MKP_CHECK(b && k && out, "null argument");
    Context& c = b->hg->ctx->c;
    Bind bind_(c);
    auto* bp = new mkp_batch_partition();
    try {
      bp->b = b;
      bp->k.assign(k, k + b->members);
      bp->kbase.assign(1, 0);
      for (uint32_t g = 0; g < b->members; ++g) {
        MKP_CHECK(k[g] >= 2 && k[g] <= 1024, "k must be in [2, 1024]");
        bp->kmax = std::max(bp->kmax, int(k[g]));
        bp->kbase.push_back(bp->kbase.back() + uint32_t(k[g]));
      }
      MKP_CHECK(mkp_partition_create(b->hg, bp->kmax, &bp->p) == 0, mkp_last_error());
      const uint32_t ktot = bp->kbase.back();
      bp->p->part_w.alloc(ktot);
      bp->gkbase.alloc(b->members + 1);
      bp->gk.alloc(b->members);
      bp->gseed.alloc(b->members);
      MKP_HIP(hipMemcpyAsync(bp->gkbase.p, bp->kbase.data(), bp->kbase.size() * 4, hipMemcpyHostToDevice, c.stream));
      MKP_HIP(hipMemcpyAsync(bp->gk.p, bp->k.data(), bp->k.size() * 4, hipMemcpyHostToDevice, c.stream));
      fill_async<uint64_t>(bp->gseed.p, 0ull, b->members, c.stream);
      partition_initialize_batch(bp->p, b->vgraph.p, bp->gkbase.p, ktot);
      c.sync();
    } catch (...) {
      delete bp;
      throw;
    }
    *out = bp;
  });
}

int mkp_batch_partition_destroy(mkp_batch_partition* bp) {
  return guarded([&] {
    if (!bp) return;
    Bind bind_(bp->b->hg->ctx->c);
    delete bp;
  });
}

int mkp_batch_partition_set_parts(mkp_batch_partition* bp, const int32_t* parts) {
  return guarded([&] {
    MKP_CHECK(bp && parts, "null argument");
    const mkp_batch* b = bp->b;
    for (uint32_t g = 0; g < b->members; ++g)
      for (uint32_t u = b->vbase[g]; u < b->vbase[g + 1]; ++u)
        MKP_CHECK(parts[u] >= 0 && parts[u] < bp->k[g], "block id out of range for its member");
    Context& c = b->hg->ctx->c;
    Bind bind_(c);
    if (b->hg->n)
      MKP_HIP(hipMemcpyAsync(bp->p->part.p, parts, size_t(b->hg->n) * 4, hipMemcpyHostToDevice, c.stream));
    partition_initialize_batch(bp->p, b->vgraph.p, bp->gkbase.p, bp->kbase.back());
    c.sync();
  });
}

int mkp_batch_partition_copy_parts(mkp_batch_partition* dst, const mkp_batch_partition* src) {
  return guarded([&] {
    MKP_CHECK(dst && src, "null argument");
    MKP_CHECK(dst->b == src->b && dst->k == src->k, "partitions of different batches or block counts");
    const mkp_batch* b = dst->b;
    Context& c = b->hg->ctx->c;
    Bind bind_(c);
    if (b->hg->n)
      MKP_HIP(hipMemcpyAsync(dst->p->part.p, src->p->part.p, size_t(b->hg->n) * 4, hipMemcpyDeviceToDevice,
                             c.stream));
    partition_initialize_batch(dst->p, b->vgraph.p, dst->gkbase.p, dst->kbase.back());
    c.sync();
  });
}

int mkp_batch_partition_get_parts(const mkp_batch_partition* bp, int32_t* parts) {
  return guarded([&] {
    MKP_CHECK(bp && parts, "null argument");
    Context& c = bp->b->hg->ctx->c;
    Bind bind_(c);
    if (bp->b->hg->n)
      MKP_HIP(hipMemcpyAsync(parts, bp->p->part.p, size_t(bp->b->hg->n) * 4, hipMemcpyDeviceToHost, c.stream));
    c.sync();
  });
}

int mkp_batch_partition_km1(mkp_batch_partition* bp, int64_t* km1) {
  return guarded([&] {
    MKP_CHECK(bp && km1, "null argument");
    Bind bind_(bp->b->hg->ctx->c);
    partition_km1_batch(bp->p, bp->b->ngraph.p, bp->b->members, km1);
  });
}

int mkp_batch_lp_refine(mkp_batch_partition* bp, const mkp_batch_lp_params* prm, const uint64_t* seeds,
                        mkp_lp_stats* stats) {
  return guarded([&] {
    MKP_CHECK(bp && prm && seeds, "null argument");
    MKP_CHECK(prm->max_rounds >= 0 && prm->sub_rounds >= 1 && prm->sub_rounds <= 255 && prm->epsilon >= 0,
              "bad parameters");
    const mkp_batch* b = bp->b;
    Context& c = b->hg->ctx->c;
    Bind bind_(c);
    // setupPartWeights per member (context.cpp:237-250)
    const uint32_t ktot = bp->kbase.back();
    std::vector<int64_t> maxw(ktot), perfect(ktot);
    for (uint32_t g = 0; g < b->members; ++g) {
      const int64_t pf = int64_t(std::ceil(double(b->weight[g]) / double(bp->k[g])));
      const int64_t mw = int64_t(int32_t((1.0 + prm->epsilon) * double(pf)));
      for (int x = 0; x < bp->k[g]; ++x) {
        maxw[bp->kbase[g] + x] = mw;
        perfect[bp->kbase[g] + x] = pf;
      }
    }
    MKP_HIP(hipMemcpyAsync(bp->gseed.p, seeds, size_t(b->members) * 8, hipMemcpyHostToDevice, c.stream));
    mkp_lp_params lp{prm->max_rounds, prm->sub_rounds, prm->activation_threshold, prm->rebalancing, 0,
                     maxw.data(), perfect.data()};
    std::vector<mkp_lp_stats> st(b->members);
    const LpBatch d = bp->desc();
    lp_refine(bp->p, lp, st.data(), nullptr, 0, &d);
    if (stats) std::copy(st.begin(), st.end(), stats);
  });
}

}  // extern "C"
