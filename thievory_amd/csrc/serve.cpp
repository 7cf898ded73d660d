// serve.cpp -- continuous-batching query server (no reference counterpart).
//
// One server thread owns multi-source engines (multi.hpp) for BFS and SSSP,
// each on its own HIP stream so their levels overlap on the device.  It
// launches the next level of every idle busy engine, polls the streams, and
// for each engine whose level has finished harvests the lanes that ended
// and admits waiting queries into free lanes (a lane is the unit of
// frontier/scratch memory, so admission is bounded by the memory the
// engines were sized for).  Engines advance independently: a BFS cohort's
// short levels do not wait for a long SSSP level.
//
// BFS on a symmetric graph (direction-optimising) is served in cohorts: a
// BFS engine admits only when none of its lanes is running, so its lanes
// stay at one depth and its dense levels can run bottom-up with early exit
// (lanes at other depths would keep every vertex "missing").  Two BFS
// engines alternate, so a new cohort starts while the previous one is
// still running.  Elsewhere admission is continuous.
//
// Submit and poll are thread-safe; the graph stays resident for the
// server's lifetime (store reconfiguration is refused meanwhile).
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "multi.hpp"

namespace thv {
void set_last_error(const std::string &m);

struct Server {
  Graph *g = nullptr;
  thv_serve_cfg cfg{};
  std::vector<std::unique_ptr<MultiEngine>> eng;  // BFS engine(s), then SSSP
  bool cohort = false;                  // BFS admits only into idle engines
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  struct Req {
    uint64_t ticket;
    uint32_t src;
  };
  std::deque<Req> waiting[2];
  std::unordered_map<uint64_t, thv_result> res;
  std::unordered_map<uint64_t, std::vector<uint32_t>> full;
  uint64_t next_ticket = 1;
  uint64_t outstanding = 0;
  bool stop = false;
  int err = 0;
  std::string errmsg;
  uint64_t t0 = 0;
  static constexpr size_t MAX_WAITING = 1u << 20;
  // keep_full: finished value arrays wait on the host until their poll; at
  // most max(KEEP_MIN, KEEP_BYTES / (4 n)) of them (counting the queries
  // still running) exist at once, beyond that submit returns THV_EBUSY
  static constexpr uint64_t KEEP_MIN = 64, KEEP_BYTES = 4ull << 30;
  uint64_t keep_cap = 0;
  // BFS cohorts: two engines alternate.  While SSSP lanes are running (a
  // mixed load, where SSSP levels slow every cohort down) the other engines
  // also start cohorts, for queries that have waited SPILL_MS already.  When
  // BFS has the GPU to itself they stay idle: more cohorts at once means
  // smaller ones, which cost throughput (4,983 against 4,244 queries/s
  // saturated at R-MAT 25).
  static constexpr int COHORT_ENGINES = 4;
  static constexpr double SPILL_MS = 10.0;

  double now_ms() const { return (now_ns() - t0) * 1e-6; }
  void loop();
  void fail_all(int code, const std::string &msg);
};

void Server::fail_all(int code, const std::string &msg) {
  std::lock_guard<std::mutex> lk(mu);
  err = code;
  errmsg = msg;
  for (auto &kv : res)
    if (kv.second.status == 1) kv.second.status = code;
  for (auto &w : waiting) w.clear();
  outstanding = 0;
  cv_done.notify_all();
}

void Server::loop() {
  std::vector<LaneResult> done;
  // engines whose level is on the device; each runs at its own pace (a
  // BFS level is not held back by a long SSSP level on the other stream)
  std::vector<char> inflight(eng.size(), 0);
  try {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        auto idle = [&] {
          for (size_t i = 0; i < eng.size(); i++)
            if (inflight[i] || eng[i]->busy()) return false;
          return waiting[0].empty() && waiting[1].empty();
        };
        cv_work.wait(lk, [&] { return stop || !idle(); });
        if (stop && idle()) break;
        // admission between an engine's levels only
        bool sssp_busy = false;
        for (size_t i = 0; i < eng.size(); i++)
          if (eng[i]->algo == THV_ALGO_SSSP && (inflight[i] || eng[i]->busy())) sssp_busy = true;
        for (size_t i = 0; i < eng.size(); i++) {
          MultiEngine &e = *eng[i];
          if (inflight[i]) continue;
          const int a = e.algo == THV_ALGO_BFS ? 0 : 1;
          if (a == 0 && cohort && e.active) continue;
          if (a == 0 && cohort && i >= 2 &&
              (!sssp_busy || waiting[0].empty() ||
               now_ms() - res[waiting[0].front().ticket].submit_ms < SPILL_MS))
            continue;
          while (!waiting[a].empty() && e.free_lanes()) {
            const Req r = waiting[a].front();
            waiting[a].pop_front();
            e.admit(r.src, r.ticket);
            res[r.ticket].start_ms = now_ms();
          }
        }
      }
      for (size_t i = 0; i < eng.size(); i++)
        if (!inflight[i] && eng[i]->busy()) {
          eng[i]->launch_level();
          inflight[i] = 1;
        }
      done.clear();
      bool progressed = false;
      for (size_t i = 0; i < eng.size(); i++) {
        if (!inflight[i]) continue;
        MultiEngine &e = *eng[i];
        const hipError_t q = hipStreamQuery(e.stream);
        if (q == hipErrorNotReady) continue;
        HIPC(q);
        inflight[i] = 0;
        progressed = true;
        e.finish_level(done);
        if (cfg.keep_full && e.just_finished) {
          for (uint32_t l = 0; l < e.L; l++) {
            if (!((e.just_finished >> l) & 1ull)) continue;
            std::vector<uint32_t> v(g->n);
            e.copy_full(l, v.data());
            std::lock_guard<std::mutex> lk(mu);
            full[e.tag_of[l]] = std::move(v);
          }
        }
      }
      if (!progressed) std::this_thread::yield();
      if (!done.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now_ms();
        for (const LaneResult &r : done) {
          thv_result &x = res[r.tag];
          x.status = 0;
          x.depth = r.depth;
          x.reached = r.reached;
          x.hash = r.hash;
          x.done_ms = t;
          if (outstanding) outstanding--;
        }
        cv_done.notify_all();
      }
    }
  } catch (const Error &e) {
    fail_all(e.code, e.what());
  } catch (const std::exception &e) {
    fail_all(THV_EHIP, e.what());
  }
}

}  // namespace thv

using namespace thv;

template <class F>
static int guard(F f) {
  try {
    return f();
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return THV_ENOMEM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return THV_EHIP;
  }
}

extern "C" {

int thv_serve_start(thv_graph *gh, const thv_serve_cfg *cfg, thv_server **out) {
  return guard([&] {
    if (!gh || !cfg || !out) fail(THV_EINVAL, "null argument");
    Graph *g = reinterpret_cast<Graph *>(gh);
    std::unique_ptr<Server> s(new Server());
    s->g = g;
    s->cfg = *cfg;
    // the store must not change while the engines are built on it: hold the
    // graph's lock until the server is counted (store_config, set_option,
    // autotune and close refuse while it runs)
    std::lock_guard<std::mutex> glk(g->mu);
    s->keep_cap = std::max<uint64_t>(Server::KEEP_MIN,
                                     Server::KEEP_BYTES / std::max<uint64_t>(1, g->n * 4));
    uint32_t lanes = cfg->max_lanes ? std::min<uint32_t>(cfg->max_lanes, 64) : 64;
    uint64_t budget = cfg->scratch_bytes;
    if (budget == 0) {
      size_t fr = 0, tot = 0;
      HIPC(hipMemGetInfo(&fr, &tot));
      budget = fr / 2;
    }
    const bool kf = cfg->keep_full != 0;
    // BFS lanes first (cheap: 3 masks per vertex), SSSP lanes get the rest,
    // halving until the lane-packed distances fit the scratch budget.
    uint64_t need_bfs = MultiEngine::bytes_needed(*g, THV_ALGO_BFS, lanes, kf);
    if (need_bfs > budget)
      fail(THV_ENOMEM, "scratch budget too small for the BFS engine");
    s->cohort = g->sym && g->dir_opt;
    s->eng.emplace_back(new MultiEngine(*g, THV_ALGO_BFS, lanes, kf));
    const uint64_t one = need_bfs;
    for (int k = 1; s->cohort && k < Server::COHORT_ENGINES && need_bfs + one <= budget; k++) {
      s->eng.emplace_back(new MultiEngine(*g, THV_ALGO_BFS, lanes, kf));
      need_bfs += one;
    }
    if (g->rec == 8) {
      uint32_t sl = lanes;
      while (sl && MultiEngine::bytes_needed(*g, THV_ALGO_SSSP, sl, kf) > budget - need_bfs) sl /= 2;
      uint64_t used = need_bfs;
      const uint32_t want = std::max<uint32_t>(1, cfg->sssp_engines);
      for (uint32_t k = 0; sl && k < want; k++) {
        const uint64_t one_s = MultiEngine::bytes_needed(*g, THV_ALGO_SSSP, sl, kf);
        if (k && used + one_s > budget) break;
        s->eng.emplace_back(new MultiEngine(*g, THV_ALGO_SSSP, sl, kf));
        used += one_s;
      }
    }
    g->servers++;
    s->t0 = now_ns();
    Server *sp = s.get();
    s->th = std::thread([sp] { sp->loop(); });
    *out = reinterpret_cast<thv_server *>(s.release());
    return THV_OK;
  });
}

int thv_serve_submit(thv_server *sh, int algo, uint32_t src, uint64_t *ticket) {
  return guard([&] {
    Server *s = reinterpret_cast<Server *>(sh);
    if (!s || !ticket) fail(THV_EINVAL, "null argument");
    const int a = algo == THV_ALGO_BFS ? 0 : algo == THV_ALGO_SSSP ? 1 : -1;
    if (a < 0) fail(THV_EINVAL, "serving takes BFS or SSSP queries");
    bool have = false;
    for (auto &e : s->eng) have |= e->algo == algo;
    if (!have) fail(THV_ENOTSUP, "no engine for this algorithm (unweighted graph or memory)");
    if (src >= s->g->n) fail(THV_EINVAL, "source vertex out of range");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->err) fail(THV_ESTATE, "server failed: " + s->errmsg);
    if (s->stop) fail(THV_ESTATE, "server is stopping");
    if (s->waiting[a].size() >= Server::MAX_WAITING) fail(THV_EBUSY, "admission queue full");
    if (s->cfg.keep_full && s->outstanding + s->full.size() >= s->keep_cap)
      fail(THV_EBUSY, "kept value arrays at their bound: poll finished queries first");
    const uint64_t t = s->next_ticket++;
    thv_result r;
    memset(&r, 0, sizeof r);
    r.ticket = t;
    r.algo = algo;
    r.status = 1;
    r.src = src;
    r.submit_ms = s->now_ms();
    s->res[t] = r;
    s->waiting[a].push_back({t, src});
    s->outstanding++;
    *ticket = t;
    s->cv_work.notify_one();
    return THV_OK;
  });
}

int thv_serve_poll(thv_server *sh, uint64_t ticket, thv_result *out, uint32_t *fullv) {
  return guard([&] {
    Server *s = reinterpret_cast<Server *>(sh);
    if (!s || !out) fail(THV_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->res.find(ticket);
    if (it == s->res.end()) fail(THV_EINVAL, "unknown ticket (never issued, or already returned)");
    *out = it->second;
    if (fullv && it->second.status == 0) {
      auto f = s->full.find(ticket);
      if (f == s->full.end()) fail(THV_ESTATE, "server does not keep value arrays");
      memcpy(fullv, f->second.data(), f->second.size() * 4);
    }
    // a finished query is handed over once: its entry and value array go
    if (it->second.status != 1) {
      s->full.erase(ticket);
      s->res.erase(it);
    }
    return THV_OK;
  });
}

int thv_serve_drain(thv_server *sh) {
  return guard([&] {
    Server *s = reinterpret_cast<Server *>(sh);
    if (!s) fail(THV_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    if (s->err) fail(s->err, "server failed: " + s->errmsg);
    return THV_OK;
  });
}

void thv_serve_stop(thv_server *sh) {
  Server *s = reinterpret_cast<Server *>(sh);
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  s->g->servers--;
  delete s;
}

}  // extern "C"
