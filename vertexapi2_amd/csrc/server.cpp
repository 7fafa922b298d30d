// Continuous-batching query server over graphs resident in HBM.
//
// Waiting queries that are cheap on the plain engine for their kind and
// graph, or that few others of their kind and graph wait beside, are queued
// to solo host threads that answer them there one at a time each (a batch
// of one); the plain engine's latency-bound supersteps run beside the batch
// rounds while the loop keeps driving the batch engines.
//
// One batch engine per (query kind, graph) -- BFS and SSSP lanes
// (BatchEngine), personalised PageRank lanes (PprBatch) -- each on its own
// stream, so rounds of different engines overlap on the device.  The host
// loop (gas_server_replay) admits arrived queries into free lanes of their
// engine while the per-query state memory fits the budget, adds them between
// rounds of the engines already running, steps every engine that has work,
// and retires lanes whose query finished: the lanes' results are extracted
// into per-lane buffers in HBM in one pass (and optionally copied to the
// caller), the lanes and their memory are released, and waiting queries can
// take them.
// Under a backlog, BFS / SSSP queries join a running batch only during its
// first rounds and otherwise wait for it to drain, so the lanes of a batch
// traverse the graph level by level together and share its edge reads
// (lanes started rounds apart each pay for their own dense levels).
// A query's latency is the time from its arrival to its result being in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gas.h"
#include "gas/batch.hpp"
#include "gas_internal.hpp"

using namespace gas;

namespace {

double now_s() {
  using C = std::chrono::steady_clock;
  static const C::time_point t0 = C::now();
  return std::chrono::duration<double>(C::now() - t0).count();
}

// Rounds an engine runs between two looks at its state when no query is
// waiting for it: a finished lane waits at most this many rounds longer.
constexpr int kRoundsPerPoll = 4;

// A query is answered by the plain engine alone (a batch of one) on the solo
// thread when that takes at most kSoloMaxMs for its kind and graph, or at
// most kSoloIdleMaxMs while few queries of its kind and graph wait, and the
// solo queue would finish it no later than a batch of one would (both costs
// measured at the first replay): a batched round costs about as much for
// one lane as for 64 (R-MAT-22 PPR: 44 ms as a batch of one, 8 ms plain),
// while the plain engine's latency-bound supersteps leave the GPU room for
// batch rounds beside them.
constexpr double kSoloMaxMs = 2.0;
constexpr double kSoloIdleMaxMs = 50.0;
constexpr int kSoloBacklog = 4;
constexpr double kSoloQueueMs = 8.0;
// Solo threads: the plain engine's supersteps are latency-bound, so two of
// them (each with its own plain engines and stream) overlap on the GPU.
constexpr int kSoloThreads = 2;
static_assert(kSoloThreads <= gas_graph::kEngineSets,
              "each solo thread needs an engine set of its own");

size_t query_bytes(int kind, int64_t V) {
  // lane state + the lane's result buffer
  switch (kind) {
    case GAS_BATCH_BFS: return (size_t)(V * 2 + V * 3 / 8 + V * 4);
    case GAS_BATCH_SSSP: return (size_t)(V * 4 + V / 4 + V * 4);
    default: return (size_t)(V * 16 + V * 3 / 64 + V * 4);  // PPR
  }
}

// A solo host thread: queued queries one at a time on the plain engine
// (gas_*_device into a per-graph buffer), each result optionally copied to
// the caller on the thread's own stream.
struct SoloWorker {
  std::mutex mu;
  std::condition_variable cv;
  std::thread th;
  bool stop = false;
  std::deque<int> jobs;                              // queued queries
  std::vector<std::pair<int, double>> finished;      // query, finish time
  int rc = GAS_OK;                                   // first failure
  std::string err;
  double queued_ms = 0;  // loop side: expected plain-engine time handed over
  int pending = 0;       // loop side: queries handed over, not yet taken

  // run(q) answers query q; the thread's whole-run calls use engine set
  // `set` of each graph, so solo threads never share a plain engine.
  template <class F>
  void start(int set, F run) {
    th = std::thread([this, run, set] {
      engine_set() = set;
      for (;;) {
        int q;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return stop || !jobs.empty(); });
          if (stop) return;
          q = jobs.front();
          jobs.pop_front();
        }
        const int r = run(q);
        const double t = now_s();
        std::lock_guard<std::mutex> lk(mu);
        if (r != GAS_OK && rc == GAS_OK) {
          rc = r;
          err = gas_last_error();
        }
        finished.emplace_back(q, t);
      }
    });
  }
  void submit(int q) {
    {
      std::lock_guard<std::mutex> lk(mu);
      jobs.push_back(q);
    }
    cv.notify_one();
  }
  // finished queries since the last call; throws the first failure
  std::vector<std::pair<int, double>> take() {
    std::vector<std::pair<int, double>> out;
    std::lock_guard<std::mutex> lk(mu);
    if (rc != GAS_OK) throw Error(rc, err);
    out.swap(finished);
    return out;
  }
  ~SoloWorker() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_one();
    if (th.joinable()) th.join();
  }
};

struct Engine {
  int kind = 0, graph = 0;
  int gang_start = 0;               // round at which the current batch began
  int64_t V = 0;
  std::unique_ptr<BatchEngine> be;
  std::unique_ptr<PprBatch> pe;
  uint64_t running = 0;             // lanes with an unfinished query
  bool inflight = false;            // a round is queued on the stream
  int lane_query[BW];
  std::vector<int32_t> add_lanes, add_src;
  DevBuf<int32_t> res_i[BW];
  DevBuf<float> res_f[BW];
  gasrt_stream_t stream() const { return be ? be->stream() : pe->stream(); }
};

}  // namespace

struct gas_server {
  std::vector<gas_graph_t> graphs;
  std::vector<int64_t> nV;
  size_t budget = 0, reserved = 0;
  int ppr_iterations = 10;
  int gang_backlog = 16, gang_rounds = 1;        // gas_server_set_gang
  double solo_max_ms = kSoloMaxMs, solo_idle_max_ms = kSoloIdleMaxMs;
  std::vector<std::unique_ptr<Engine>> engines;  // [kind * n_graphs + graph]
  std::vector<double> solo_ms;                   // plain-engine cost per key
  std::vector<double> batch1_ms;                 // batch-of-one cost per key
  bool calibrated = false;
  // [solo thread][graph] plain results
  std::vector<DevBuf<int32_t>> solo_out[kSoloThreads];

  Engine &engine(int kind, int graph) {
    auto &slot = engines[(size_t)kind * graphs.size() + graph];
    if (!slot) {
      std::unique_ptr<Engine> e(new Engine);
      e->kind = kind;
      e->graph = graph;
      e->V = nV[graph];
      if (kind == GAS_BATCH_PPR)
        e->pe.reset(new PprBatch(graphs[graph], ppr_iterations));
      else
        e->be.reset(new BatchEngine(graphs[graph],
                                    kind == GAS_BATCH_BFS
                                        ? BatchEngine::BFS_KIND
                                        : BatchEngine::SSSP_KIND));
      for (auto &q : e->lane_query) q = -1;
      slot = std::move(e);
    }
    return *slot;
  }
};

namespace {

template <class F>
int guarded_s(F f) {
  try {
    clear_error();
    f();
    return GAS_OK;
  } catch (const Error &x) {
    return set_error(x.code ? x.code : GAS_ERR_HIP, "%s", x.what());
  } catch (const std::bad_alloc &) {
    return set_error(GAS_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception &x) {
    return set_error(GAS_ERR_STATE, "%s", x.what());
  }
}

}  // namespace

extern "C" {

size_t gas_server_query_bytes(int kind, int64_t n_vertices) {
  return query_bytes(kind, n_vertices);
}

int gas_server_create(const gas_graph_t *graphs, int32_t n_graphs,
                      size_t memory_budget_bytes, int32_t ppr_iterations,
                      gas_server_t *out) {
  if (!graphs || n_graphs < 1 || !out || ppr_iterations < 1)
    return set_error(GAS_ERR_ARG, "gas_server_create: bad arguments");
  return guarded_s([&] {
    std::unique_ptr<gas_server> s(new gas_server);
    for (int i = 0; i < n_graphs; ++i) {
      if (!graphs[i]) throw Error(GAS_ERR_ARG, "null graph");
      gas_graph_view gv;
      check(gas_graph_view_get(graphs[i], &gv), "graph view");
      s->graphs.push_back(graphs[i]);
      s->nV.push_back(gv.n_vertices);
    }
    s->budget = memory_budget_bytes;
    s->ppr_iterations = ppr_iterations;
    s->engines.resize(3 * (size_t)n_graphs);
    s->solo_ms.assign(3 * (size_t)n_graphs, 1e30);
    s->batch1_ms.assign(3 * (size_t)n_graphs, 0.0);
    for (auto &v : s->solo_out) v.resize((size_t)n_graphs);
    *out = s.release();
  });
}

int gas_server_set_solo(gas_server_t s, double max_ms, double idle_max_ms) {
  if (!s || !(max_ms >= 0) || !(idle_max_ms >= 0))
    return set_error(GAS_ERR_ARG, "gas_server_set_solo: bad arguments");
  s->solo_max_ms = max_ms;
  s->solo_idle_max_ms = idle_max_ms;
  return GAS_OK;
}

int gas_server_set_gang(gas_server_t s, int32_t backlog, int32_t rounds) {
  if (!s || backlog < 0 || rounds < 0)
    return set_error(GAS_ERR_ARG, "gas_server_set_gang: bad arguments");
  s->gang_backlog = backlog;
  s->gang_rounds = rounds;
  return GAS_OK;
}

int gas_server_free(gas_server_t s) {
  delete s;
  return GAS_OK;
}

int gas_server_replay(gas_server_t s, int32_t n, const int32_t *kinds,
                      const int32_t *graph_ids, const int32_t *sources,
                      const double *arrival_s, void *const *host_results,
                      double *latency_ms, gas_serve_stats *stats) {
  if (!s || n < 0 || (n > 0 && (!kinds || !graph_ids || !sources ||
                                !arrival_s || !latency_ms)))
    return set_error(GAS_ERR_ARG, "gas_server_replay: bad arguments");
  for (int i = 0; i < n; ++i) {
    if (kinds[i] < 0 || kinds[i] > 2 || graph_ids[i] < 0 ||
        graph_ids[i] >= (int)s->graphs.size() || sources[i] < 0 ||
        sources[i] >= s->nV[graph_ids[i]])
      return set_error(GAS_ERR_ARG, "gas_server_replay: query %d invalid", i);
    if (query_bytes(kinds[i], s->nV[graph_ids[i]]) > s->budget)
      return set_error(GAS_ERR_FULL,
                       "gas_server_replay: query %d needs %zu bytes, budget "
                       "is %zu",
                       i, query_bytes(kinds[i], s->nV[graph_ids[i]]),
                       s->budget);
    if (i && arrival_s[i] < arrival_s[i - 1])
      return set_error(GAS_ERR_ARG, "gas_server_replay: arrivals not sorted");
  }
  return guarded_s([&] {
    // plain-engine cost of BFS and SSSP per graph, measured once on an idle
    // server (the first run of each builds that plain engine)
    if (!s->calibrated) {
      // every solo thread's plain engines (engine set = thread index) are
      // built here, so no query pays for building one; set 0's second run
      // is the timed one
      struct SetGuard {
        ~SetGuard() { engine_set() = 0; }
      } set_guard;
      for (int set = kSoloThreads - 1; set >= 0; --set) {
        engine_set() = set;
        for (size_t gi = 0; gi < s->graphs.size(); ++gi) {
          auto &buf = s->solo_out[set][gi];
          if (buf.n < (size_t)s->nV[gi]) buf.alloc((size_t)s->nV[gi]);
          for (int kind : {GAS_BATCH_BFS, GAS_BATCH_SSSP, GAS_BATCH_PPR}) {
            auto run = [&] {
              return kind == GAS_BATCH_BFS
                         ? gas_bfs_device(s->graphs[gi], 0, buf.p, nullptr)
                     : kind == GAS_BATCH_SSSP
                         ? gas_sssp_device(s->graphs[gi], 0, buf.p, nullptr)
                         : gas_ppr_device(s->graphs[gi], 0, s->ppr_iterations,
                                          reinterpret_cast<float *>(buf.p), nullptr);
            };
            if (run() != GAS_OK) continue;  // e.g. SSSP on an unweighted graph
            if (set != 0) continue;
            const double ts = now_s();
            if (run() != GAS_OK) continue;
            s->solo_ms[(size_t)kind * s->graphs.size() + gi] = (now_s() - ts) * 1e3;
          }
        }
      }
      // and a batch of one lane on each batch engine (from vertex 0)
      for (size_t gi = 0; gi < s->graphs.size(); ++gi) {
        for (int kind : {GAS_BATCH_BFS, GAS_BATCH_SSSP, GAS_BATCH_PPR}) {
          const size_t key = (size_t)kind * s->graphs.size() + gi;
          if (s->solo_ms[key] >= 1e30) continue;  // kind unavailable
          Engine &e = s->engine(kind, (int)gi);
          const int32_t lane = 0, src = 0;
          const double ts = now_s();
          if (e.be) {
            e.be->addLanes(&lane, &src, 1);
            e.be->runToCompletion();
            e.be->sync();
            e.be->releaseLanes(1);
          } else {
            e.pe->addLanes(&lane, &src, 1);
            e.pe->runToCompletion();
            e.pe->sync();
            e.pe->releaseLanes(1);
          }
          s->batch1_ms[key] = (now_s() - ts) * 1e3;
        }
      }
      clear_error();
      s->calibrated = true;
    }
    std::deque<int> waiting;
    int next = 0, done = 0;
    int64_t rounds = 0, admissions_deferred = 0;
    int max_running = 0, solo = 0;
    const double t0 = now_s();
    struct StreamGuard {
      gasrt_stream_t s = nullptr;
      ~StreamGuard() {
        if (s) gasrt_stream_destroy(s);
      }
    } copy_guard[kSoloThreads];
    for (auto &c : copy_guard) check(gasrt_stream_create(&c.s), "stream");
    SoloWorker workers[kSoloThreads];  // joined before the streams are destroyed
    for (int w = 0; w < kSoloThreads; ++w) {
      workers[w].start(w, [&, w](int q) {
        const int gi = graph_ids[q];
        auto &buf = s->solo_out[w][gi];
        if (buf.n < (size_t)s->nV[gi]) buf.alloc((size_t)s->nV[gi]);
        int rc = kinds[q] == GAS_BATCH_BFS
                     ? gas_bfs_device(s->graphs[gi], sources[q], buf.p, nullptr)
                 : kinds[q] == GAS_BATCH_SSSP
                     ? gas_sssp_device(s->graphs[gi], sources[q], buf.p, nullptr)
                     : gas_ppr_device(s->graphs[gi], sources[q], s->ppr_iterations,
                                      reinterpret_cast<float *>(buf.p), nullptr);
        if (rc == GAS_OK && host_results && host_results[q])
          rc = gasrt_memcpy_d2h(host_results[q], buf.p, 4 * (size_t)s->nV[gi],
                                copy_guard[w].s);
        return rc;
      });
    }
    auto running_total = [&] {
      int r = 0;
      for (auto &ep : s->engines)
        if (ep) r += __builtin_popcountll(ep->running);
      return r;
    };
    while (done < n) {
      bool progressed = false;
      // 1. engines whose round finished: retire the queries that ended
      for (auto &ep : s->engines) {
        if (!ep || !ep->inflight) continue;
        Engine &e = *ep;
        if (!(e.be ? e.be->poll() : e.pe->poll())) continue;
        e.inflight = false;
        progressed = true;
        const uint64_t alive =
            e.be ? e.be->aliveSnapshot() : e.pe->lanesAlive();
        const uint64_t fin_all = e.running & ~alive;
        if (!fin_all) continue;
        int32_t *outs[BW] = {};
        for (uint64_t fin = fin_all; fin; fin &= fin - 1) {
          const int lane = __builtin_ctzll(fin);
          if (e.be) {
            if (e.res_i[lane].n < (size_t)e.V) e.res_i[lane].alloc((size_t)e.V);
            outs[lane] = e.res_i[lane].p;
          } else {
            if (e.res_f[lane].n < (size_t)e.V) e.res_f[lane].alloc((size_t)e.V);
            e.pe->laneResultDevice(lane, e.res_f[lane].p);
          }
        }
        if (e.be) e.be->laneResultsDevice(fin_all, outs);
        check(gasrt_stream_synchronize(e.stream()), "sync");
        const double tf = now_s() - t0;
        for (uint64_t fin = fin_all; fin; fin &= fin - 1) {
          const int lane = __builtin_ctzll(fin);
          const int q = e.lane_query[lane];
          latency_ms[q] = (tf - arrival_s[q]) * 1e3;
          if (host_results && host_results[q]) {
            void *src = e.be ? (void *)e.res_i[lane].p : (void *)e.res_f[lane].p;
            check(gasrt_memcpy_d2h(host_results[q], src, 4 * (size_t)e.V,
                                   e.stream()),
                  "result copy");
          }
          s->reserved -= query_bytes(e.kind, e.V);
          e.lane_query[lane] = -1;
          e.running &= ~(1ull << lane);
          if (e.be)
            e.be->releaseLanes(1ull << lane);
          else
            e.pe->releaseLanes(1ull << lane);
          ++done;
        }
      }
      // 2. arrivals and admission (first come first served per engine; a
      // query that does not fit does not block other engines), only into
      // engines between rounds
      const double t = now_s() - t0;
      while (next < n && arrival_s[next] <= t) waiting.push_back(next++);
      // queries the solo threads finished
      for (auto &worker : workers)
        for (const auto &f : worker.take()) {
          const int q = f.first;
          latency_ms[q] = (f.second - t0 - arrival_s[q]) * 1e3;
          worker.queued_ms -= s->solo_ms[(size_t)kinds[q] * s->graphs.size() +
                                         graph_ids[q]];
          if (--worker.pending == 0) worker.queued_ms = 0;
          ++done;
          ++solo;
          progressed = true;
        }
      // waiting queries go to the solo thread with the shortest queue, in
      // arrival order, when their plain-engine time is at most solo_max_ms,
      // or at most solo_idle_max_ms while fewer than kSoloBacklog queries of
      // their kind and graph wait (more would share a batch), and that queue
      // would finish them no later than a batch of one (an idle solo thread
      // takes one regardless).  Solo threads after the first take queries
      // only while at most kSoloBacklog queries wait in all: under load the
      // batches answer them for less GPU time each.  BFS / SSSP results are exact on either engine; a PPR
      // result from the plain engine is within the same 1e-5 of FP64 as a
      // batch lane but not bit-identical to one.
      {
        std::vector<int> backlog(s->engines.size(), 0);
        for (int q : waiting)
          ++backlog[(size_t)kinds[q] * s->graphs.size() + graph_ids[q]];
        const int n_solo = waiting.size() <= (size_t)kSoloBacklog ? kSoloThreads : 1;
        for (auto it = waiting.begin(); it != waiting.end();) {
          const int q = *it;
          const size_t key = (size_t)kinds[q] * s->graphs.size() + graph_ids[q];
          const double ms = s->solo_ms[key];
          const double cap =
              backlog[key] < kSoloBacklog ? s->solo_idle_max_ms : s->solo_max_ms;
          const double budget = std::max(kSoloQueueMs, s->batch1_ms[key]);
          SoloWorker *wk = &workers[0];
          for (int w = 1; w < n_solo; ++w)
            if (workers[w].queued_ms < wk->queued_ms ||
                (workers[w].pending == 0 && wk->pending > 0))
              wk = &workers[w];
          if (ms > cap || (wk->pending > 0 && wk->queued_ms + ms > budget)) {
            ++it;
            continue;
          }
          wk->queued_ms += ms;
          ++wk->pending;
          wk->submit(q);
          it = waiting.erase(it);
          progressed = true;
        }
      }
      std::vector<char> blocked(s->engines.size(), 0);
      std::vector<int> backlog(s->engines.size(), 0);
      for (int q : waiting)
        ++backlog[(size_t)kinds[q] * s->graphs.size() + graph_ids[q]];
      for (auto it = waiting.begin(); it != waiting.end();) {
        const int q = *it;
        const size_t key = (size_t)kinds[q] * s->graphs.size() + graph_ids[q];
        if (blocked[key]) {
          ++it;
          continue;
        }
        Engine &e = s->engine(kinds[q], graph_ids[q]);
        const size_t need = query_bytes(kinds[q], s->nV[graph_ids[q]]);
        uint64_t busy = e.be ? e.be->lanesUsed() : e.pe->lanesUsed();
        for (int32_t l : e.add_lanes) busy |= 1ull << l;
        // a backlog waits for a BFS / SSSP batch past its first rounds to
        // drain (PPR lanes cost the same dense sweeps at any offset)
        const bool hold = e.be && e.running && s->gang_backlog > 0 &&
                          backlog[key] >= s->gang_backlog &&
                          e.be->round() - e.gang_start > s->gang_rounds;
        if (e.inflight || busy == ~0ull || hold ||
            s->reserved + need > s->budget) {
          blocked[key] = 1;
          if (!e.inflight) ++admissions_deferred;
          ++it;
          continue;
        }
        if (e.be && !e.running && e.add_lanes.empty())
          e.gang_start = e.be->round();
        const int lane = __builtin_ctzll(~busy);
        e.add_lanes.push_back(lane);
        e.add_src.push_back(sources[q]);
        e.lane_query[lane] = q;
        s->reserved += need;
        it = waiting.erase(it);
      }
      // 3. start the next rounds of every idle engine with work: one round
      // when a query waits for the engine (it joins after that round),
      // several otherwise (fewer host round trips per query)
      std::vector<char> wants(s->engines.size(), 0);
      for (int q : waiting)
        wants[(size_t)kinds[q] * s->graphs.size() + graph_ids[q]] = 1;
      for (size_t ei = 0; ei < s->engines.size(); ++ei) {
        auto &ep = s->engines[ei];
        if (!ep || ep->inflight) continue;
        Engine &e = *ep;
        if (!e.add_lanes.empty()) {
          if (e.be)
            e.be->addLanes(e.add_lanes.data(), e.add_src.data(),
                           (int)e.add_lanes.size());
          else
            e.pe->addLanes(e.add_lanes.data(), e.add_src.data(),
                           (int)e.add_lanes.size());
          for (int32_t l : e.add_lanes) e.running |= 1ull << l;
          e.add_lanes.clear();
          e.add_src.clear();
        }
        if (e.running) {
          const int k = wants[ei] ? 1 : kRoundsPerPoll;
          if (e.be)
            e.be->stepAsync(k);
          else
            e.pe->stepAsync(k);
          e.inflight = true;
          ++rounds;
          progressed = true;
        }
      }
      max_running = std::max(max_running, running_total());
      if (!progressed) {
        bool any_inflight = false;
        for (auto &w : workers) any_inflight |= w.pending > 0;
        for (auto &ep : s->engines) any_inflight |= ep && ep->inflight;
        if (!any_inflight && waiting.empty() && next < n) {
          const double wait = arrival_s[next] - (now_s() - t0);
          if (wait > 0.0005)
            std::this_thread::sleep_for(
                std::chrono::duration<double>(wait * 0.8));
        } else {
          std::this_thread::yield();
        }
      }
    }
    if (stats) {
      memset(stats, 0, sizeof(*stats));
      stats->elapsed_s = now_s() - t0;
      stats->rounds = rounds;
      stats->admissions_deferred = admissions_deferred;
      stats->max_concurrent = max_running;
      stats->solo = solo;
    }
  });
}

}  // extern "C"
