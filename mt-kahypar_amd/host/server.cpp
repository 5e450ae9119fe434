// The served form: a request loop over partition requests of uneven size.
//
// Every slot is a host thread with a context of its own (stream, device
// allocation cache, pinned staging), so the kernels of the requests in flight
// overlap on the device while each request keeps its own launch order.
// Admission is first come, first served: the oldest waiting request starts
// when a slot is free and its byte estimate fits the budget beside the
// estimates of the running requests and the bytes the allocation caches hold;
// caches of idle slots are released when that makes the difference.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "../csrc/internal.h"
#include "host_hypergraph.h"

using namespace mkp;

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t ticket = 0;
  const mkp_host_hypergraph* h = nullptr;
  mkp_partition_params prm{};
  uint64_t estimate = 0;
  Clock::time_point t_submit, t_start, t_end;
  std::vector<int32_t> parts;
  mkp_partition_stats st{};
  int status = 0;
  std::string error;
  bool done = false;
  int slot = -1;
  uint64_t peak = 0;
};

}  // namespace

struct mkp_server {
  uint64_t budget = 0;
  int ip_threads = 0;
  int host_threads = 0;
  std::vector<mkp_context*> ctx;
  std::vector<uint8_t> busy;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  uint64_t next_ticket = 1;
  uint64_t reserved = 0;  // estimates of the running requests
  int running = 0;
  bool stop = false;
  mkp_server_stats stats{};

  // bytes the allocation caches hold beyond what running requests use
  uint64_t cached_total() {
    uint64_t c = 0;
    for (auto* x : ctx) c += x->c.pool->cached_bytes();
    return c;
  }

  // called with mu held: whether the oldest waiting request may start now
  bool admissible() {
    const auto& r = queue.front();
    if (reserved + r->estimate + cached_total() <= budget) return true;
    // release the caches of idle slots (their streams have no work queued)
    bool trimmed = false;
    for (size_t s = 0; s < ctx.size(); ++s) {
      if (busy[s] || !ctx[s]->c.pool->cached_bytes()) continue;
      Bind bind_(ctx[s]->c);
      ctx[s]->c.pool->trim();
      trimmed = true;
    }
    if (trimmed) stats.trims += 1;
    return reserved + r->estimate + cached_total() <= budget;
  }

  void work(int s) {
    mkp_context* cx = ctx[s];
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      cv_work.wait(lk, [&] { return (stop && queue.empty()) || (!queue.empty() && admissible()); });
      if (queue.empty()) return;  // stopping
      std::shared_ptr<Request> r = queue.front();
      queue.pop_front();
      reserved += r->estimate;
      busy[s] = 1;
      running += 1;
      stats.max_in_flight = std::max(stats.max_in_flight, running);
      stats.max_reserved_bytes = std::max(stats.max_reserved_bytes, reserved);
      r->slot = s;
      r->t_start = Clock::now();
      const bool more = !queue.empty();
      lk.unlock();
      if (more) cv_work.notify_one();  // the next request may fit as well
      run(cx, *r);
      lk.lock();
      reserved -= r->estimate;
      busy[s] = 0;
      running -= 1;
      r->done = true;
      if (r->status) stats.failed += 1;
      else stats.completed += 1;
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }

  static void run(mkp_context* cx, Request& r) {
    const auto& h = r.h->h;
    r.parts.assign(h.n, 0);
    const size_t base = cx->c.pool->reset_peak();
    mkp_hypergraph* hg = nullptr;
    int st = mkp_hypergraph_create(cx, h.n, h.m, h.noff.data(), h.pins.data(), h.vw.data(), h.nw.data(), &hg);
    if (st == 0) st = mkp_partition_hypergraph(hg, &r.prm, r.parts.data(), &r.st);
    if (st) r.error = mkp_last_error();
    if (hg) mkp_hypergraph_destroy(hg);
    r.status = st;
    r.peak = cx->c.pool->peak_bytes() - base;
    // the context's work areas go back to its cache, which admission counts
    // and may release; nothing stays in use between requests
    cx->c.tmp.release();
    for (auto& b : cx->c.scratch) b.release();
    r.t_end = Clock::now();
  }
};

extern "C" {

namespace {
// device bytes of one initial-partitioning run on a coarsest level of at
// most n vertices and m nets
double ip_run_bytes(uint32_t n, uint32_t m, uint32_t max_net_size) {
  return (max_net_size > 0xFFFFu ? 8.0 : 4.0) * double(m) + 22.0 * double(n);
}
}  // namespace

int mkp_request_estimate(uint32_t n, uint32_t m, uint64_t p, uint32_t max_net_size, int32_t k, uint64_t* bytes) {
  return guarded([&] {
    MKP_CHECK(bytes, "null argument");
    MKP_CHECK(k >= 2, "k must be at least 2");
    const double P = double(p), N = double(n), E = double(m), nodes = N + E, arcs = 2.0 * P;
    const PcLayout L = pc_layout(k, std::max<uint32_t>(max_net_size, 1));
    // resident hypergraph: CSR both ways, pin/incidence owners, net ranks and
    // rank-ordered pins, weights and communities
    const double hyper = 28.0 * P + 12.0 * N + 20.0 * E;
    // community detection: node volumes, clustering and its volumes and
    // sizes, work lists, heavy-node tables, then the contraction (three key
    // and three weight arrays per arc, the kept flags and their scan, sort
    // storage, the coarse graph)
    const double louvain = 40.0 * nodes + 20.0 * arcs + 48.0 * arcs + 8.0 * arcs + 16.0 * arcs + 16.0 * arcs +
                           24.0 * nodes;
    // coarsening: the levels of the hierarchy (at most the input again in
    // total under the shrink bounds) and the rating and contraction work
    const double coarsen = hyper + 96.0 * P + 64.0 * N;
    // partition state and label propagation work at the finest level
    const double partition = 8.0 * E * double(L.stride) + 4.0 * N + 80.0 * N + 24.0 * P;
    // initial partitioning on the coarsest level (at most the input), one
    // recursion level at a time: per run pin counts of every net and
    // per-vertex state, the level's clique scores (at most 256 MB), the
    // halves of the level's sub-hypergraphs and their split temporaries
    const double ip = ip_run_bytes(n, m, max_net_size) * 20.0 + std::min(8.0 * N * N, 268435456.0) + 16.0 * P +
                      48.0 * E;
    // with degree-zero vertices or large nets set aside, the multilevel phase
    // runs on a reduced copy of the hypergraph
    const double need = hyper + std::max(louvain, hyper + coarsen + std::max(partition, ip));
    // every buffer rounds up to a power of two in the allocation cache
    *bytes = uint64_t(2.0 * need) + (uint64_t(64) << 20);
  });
}

int mkp_server_create(const mkp_server_params* prm, mkp_server** out) {
  return guarded([&] {
    MKP_CHECK(prm && out, "null argument");
    MKP_CHECK(prm->slots >= 1 && prm->slots <= 64, "slots must be in [1, 64]");
    auto* s = new mkp_server();
    try {
      for (int i = 0; i < prm->slots; ++i) {
        mkp_context* c = nullptr;
        MKP_CHECK(mkp_context_create(prm->device, &c) == 0, mkp_last_error());
        s->ctx.push_back(c);
        // the slot threads share the host cores with each other and with
        // the submitting threads: their waits sleep unless asked to spin
        c->c.sleep_wait = prm->spin_wait == 0;
      }
      s->budget = prm->hbm_budget_bytes;
      MKP_CHECK(prm->ip_threads >= 0 && prm->ip_threads <= 256, "ip_threads must be in [0, 256]");
      s->ip_threads = prm->ip_threads;
      MKP_CHECK(prm->host_threads >= 0 && prm->host_threads <= 1024, "host_threads must be in [0, 1024]");
      s->host_threads = prm->host_threads;
      if (!s->budget) {
        size_t fr = 0, total = 0;
        MKP_HIP(hipSetDevice(prm->device));
        MKP_HIP(hipMemGetInfo(&fr, &total));
        s->budget = total > (size_t(4) << 30) ? total - (size_t(4) << 30) : total / 2;
      }
      s->stats.budget_bytes = s->budget;
      s->busy.assign(prm->slots, 0);
      for (int i = 0; i < prm->slots; ++i) s->workers.emplace_back([s, i] { s->work(i); });
    } catch (...) {
      for (auto* c : s->ctx) mkp_context_destroy(c);
      delete s;
      throw;
    }
    *out = s;
  });
}

int mkp_server_destroy(mkp_server* s) {
  return guarded([&] {
    if (!s) return;
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto* c : s->ctx) mkp_context_destroy(c);
    delete s;
  });
}

int mkp_server_submit(mkp_server* s, const mkp_host_hypergraph* h, const mkp_partition_params* prm,
                      uint64_t* ticket) {
  return guarded([&] {
    MKP_CHECK(s && h && prm && ticket, "null argument");
    validate_partition_params(*prm);
    auto r = std::make_shared<Request>();
    r->t_submit = Clock::now();
    r->h = h;
    r->prm = *prm;
    uint32_t me = 0;
    for (uint32_t e = 0; e < h->h.m; ++e) me = std::max<uint32_t>(me, uint32_t(h->h.noff[e + 1] - h->h.noff[e]));
    MKP_CHECK(mkp_request_estimate(h->h.n, h->h.m, h->h.pins.size(), me, prm->k, &r->estimate) == 0,
              mkp_last_error());
    if (prm->ip_runs > 20)  // the estimate assumes the default 20 runs per bisection
      r->estimate += uint64_t(2.0 * ip_run_bytes(h->h.n, h->h.m, me) * double(prm->ip_runs - 20));
    std::lock_guard<std::mutex> g(s->mu);
    MKP_CHECK(r->estimate <= s->budget, "request needs more device memory than the whole budget");
    MKP_CHECK(!s->stop, "server is stopping");
    r->ticket = s->next_ticket++;
    s->live[r->ticket] = r;
    s->queue.push_back(r);
    s->stats.submitted += 1;
    *ticket = r->ticket;
    s->cv_work.notify_one();
  });
}

int mkp_server_wait(mkp_server* s, uint64_t ticket, int32_t* parts, mkp_request_result* res) {
  return guarded([&] {
    MKP_CHECK(s, "null argument");
    std::shared_ptr<Request> r;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      auto it = s->live.find(ticket);
      MKP_CHECK(it != s->live.end(), "unknown ticket");
      r = it->second;
      s->cv_done.wait(lk, [&] { return r->done; });
      s->live.erase(ticket);
    }
    if (res) {
      res->km1 = r->st.km1;
      res->imbalance = r->st.imbalance;
      res->levels = r->st.levels;
      res->slot = r->slot;
      res->queue_ms = ms_between(r->t_submit, r->t_start);
      res->service_ms = ms_between(r->t_start, r->t_end);
      res->latency_ms = ms_between(r->t_submit, r->t_end);
      res->estimate_bytes = r->estimate;
      res->peak_bytes = r->peak;
    }
    MKP_CHECK(r->status == 0, "request failed: " + r->error);
    if (parts && !r->parts.empty()) std::copy(r->parts.begin(), r->parts.end(), parts);
  });
}

int mkp_server_get_stats(mkp_server* s, mkp_server_stats* out) {
  return guarded([&] {
    MKP_CHECK(s && out, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
  });
}

}  // extern "C"
