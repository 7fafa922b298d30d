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
// stream, so rounds of different engines overlap on the device.  The
// serving loop (Session) admits arrived queries into free lanes of their
// engine, adds them between rounds of the engines already running, steps
// every engine that has work, and retires lanes whose query finished: the
// lanes' results are extracted into per-lane buffers in HBM in one pass
// (and optionally copied to the caller), the lanes are released, and
// waiting queries can take them.  Queries reach the loop from a pre-timed
// trace (gas_server_replay, on the caller's thread) or one at a time from
// callers while the loop runs on the server's own thread
// (gas_server_start / submit / poll / wait / stop).
// Memory: every device allocation the server makes -- a batch engine's
// state for its 64 lanes, a lane's result buffer, the solo threads' result
// buffers -- is charged against the budget when it is made, and admission
// makes one only while it fits, evicting batch engines that have no
// running query (least recently used first) to make room.
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
#include <functional>
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
// measured at calibration): a batched round costs about as much for one
// lane as for 64 (R-MAT-22 PPR: 44 ms as a batch of one, 8 ms plain), while
// the plain engine's latency-bound supersteps leave the GPU room for batch
// rounds beside them.
constexpr double kSoloMaxMs = 2.0;
constexpr double kSoloIdleMaxMs = 50.0;
constexpr int kSoloBacklog = 4;
// While this many queries wait, run in batch lanes or sit in the solo
// queues, none goes solo: under
// load the batches answer each for a fraction of the plain engine's GPU
// time (road1024 + R-MAT-22 mix at 5,000 q/s offered: 2,884 q/s with solo
// queries at any load, 3,097 q/s with none).
constexpr int kSoloMaxInflight = 48;
constexpr double kSoloQueueMs = 8.0;
// Solo threads: the plain engine's supersteps are latency-bound, so two of
// them (each with its own plain engines and stream) overlap on the GPU.
constexpr int kSoloThreads = 2;
static_assert(kSoloThreads <= gas_graph::kEngineSets,
              "each solo thread needs an engine set of its own");

// Share of the free device memory the server may use when no budget is
// given.
constexpr double kDefaultBudgetShare = 0.8;

size_t query_bytes(int kind, int64_t V) {
  // lane state + the lane's result buffer
  switch (kind) {
    case GAS_BATCH_BFS: return (size_t)(V * 2 + V * 3 / 8 + V * 4);
    case GAS_BATCH_SSSP: return (size_t)(V * 4 + V / 4 + V * 4);
    default: return (size_t)(V * 16 + V * 3 / 64 + V * 4);  // PPR
  }
}

// One query as the loop sees it.
struct QRec {
  int32_t kind = 0, graph = 0, source = 0;
  double arrival = 0;        // seconds on the session clock
  void *host_out = nullptr;  // V int32 / float, or null
};

// A solo host thread: queued queries one at a time on the plain engine
// (gas_*_device into a per-graph buffer), each result optionally copied to
// the caller on the thread's own stream.
struct SoloWorker {
  std::mutex mu;
  std::condition_variable cv;
  std::thread th;
  bool stop = false;
  std::deque<std::pair<int, QRec>> jobs;             // queued queries
  std::vector<std::pair<int, double>> finished;      // query, finish time
  int rc = GAS_OK;                                   // first failure
  std::string err;
  double queued_ms = 0;  // loop side: expected plain-engine time handed over
  int pending = 0;       // loop side: queries handed over, not yet taken

  // run(i, q) answers query i; the thread's whole-run calls use engine set
  // `set` of each graph, so solo threads never share a plain engine.
  template <class F>
  void start(int set, F run) {
    th = std::thread([this, run, set] {
      engine_set() = set;
      for (;;) {
        std::pair<int, QRec> job;
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return stop || !jobs.empty(); });
          if (stop) return;
          job = jobs.front();
          jobs.pop_front();
        }
        const int r = run(job.first, job.second);
        const double t = now_s();
        std::lock_guard<std::mutex> lk(mu);
        if (r != GAS_OK && rc == GAS_OK) {
          rc = r;
          err = gas_last_error();
        }
        finished.emplace_back(job.first, t);
      }
    });
  }
  void submit(int i, const QRec &q) {
    {
      std::lock_guard<std::mutex> lk(mu);
      jobs.emplace_back(i, q);
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
  size_t charged = 0;               // bytes of the budget this engine holds
  double last_used = 0;             // clock of its last admission
  // extractions queued on the stream: an event after the extraction and
  // the queries whose results it writes (done when the event is)
  struct Extraction {
    gasrt_event_t ev = nullptr;
    std::vector<int> queries;
  };
  std::deque<Extraction> ext;
  std::vector<gasrt_event_t> spare_ev;
  ~Engine() {
    if (!ext.empty()) gasrt_stream_synchronize(stream());
    for (auto &x : ext) gasrt_event_destroy(x.ev);
    for (auto ev : spare_ev) gasrt_event_destroy(ev);
  }
  gasrt_stream_t stream() const { return be ? be->stream() : pe->stream(); }
  size_t stateBytes() const { return be ? be->deviceBytes() : pe->deviceBytes(); }
  bool idle() const { return !running && !inflight && add_lanes.empty() && ext.empty(); }
  bool hasResult(int lane) const {
    return be ? res_i[lane].n >= (size_t)V : res_f[lane].n >= (size_t)V;
  }
};

}  // namespace

struct gas_server {
  std::vector<gas_graph_t> graphs;
  std::vector<int64_t> nV;
  size_t budget = 0;                 // device bytes the server may hold
  size_t charged = 0;                // device bytes it holds now
  int ppr_iterations = 10;
  int gang_backlog = 16, gang_rounds = 1;        // gas_server_set_gang
  double solo_max_ms = kSoloMaxMs, solo_idle_max_ms = kSoloIdleMaxMs;
  std::vector<std::unique_ptr<Engine>> engines;  // [kind * n_graphs + graph]
  std::vector<size_t> footprint;                 // engine state bytes per key
  std::vector<double> solo_ms;                   // plain-engine cost per key
  std::vector<double> batch1_ms;                 // batch-of-one cost per key
  bool calibrated = false;
  int64_t evictions = 0;
  // [solo thread][graph] plain results
  std::vector<DevBuf<int32_t>> solo_out[kSoloThreads];

  // online serving (gas_server_start ... stop)
  std::mutex omu;
  std::condition_variable ocv;
  std::thread othread;
  bool online = false, ostop = false;
  std::vector<std::pair<int64_t, QRec>> inbox;   // (ticket, query) not yet taken
  struct Ticket {
    int state = 0;        // 0 waiting or running, 1 done, -1 failed
    double latency_ms = -1;
  };
  std::vector<Ticket> tickets;
  int orc = GAS_OK;
  std::string oerr;
  gas_serve_stats ostats{};

  size_t key(int kind, int graph) const { return (size_t)kind * graphs.size() + graph; }

  // Result buffer sizes for graph g.
  size_t resultBytes(int graph) const { return 4 * (size_t)nV[graph]; }

  Engine *find(int kind, int graph) { return engines[key(kind, graph)].get(); }

  // Builds the batch engine of (kind, graph) and charges its state.
  Engine &create(int kind, int graph) {
    auto &slot = engines[key(kind, graph)];
    std::unique_ptr<Engine> e(new Engine);
    e->kind = kind;
    e->graph = graph;
    e->V = nV[graph];
    if (kind == GAS_BATCH_PPR)
      e->pe.reset(new PprBatch(graphs[graph], ppr_iterations));
    else
      e->be.reset(new BatchEngine(graphs[graph],
                                  kind == GAS_BATCH_BFS ? BatchEngine::BFS_KIND
                                                        : BatchEngine::SSSP_KIND));
    for (auto &q : e->lane_query) q = -1;
    e->charged = e->stateBytes();
    footprint[key(kind, graph)] = e->charged;
    charged += e->charged;
    slot = std::move(e);
    return *slot;
  }

  // Frees an idle batch engine and what it was charged.
  void evict(size_t k) {
    auto &slot = engines[k];
    if (!slot) return;
    charged -= slot->charged;
    slot.reset();
    ++evictions;
  }

  // Evicts idle batch engines other than `keep`, least recently used first,
  // until `need` more bytes fit the budget; false if they cannot.
  bool makeRoom(size_t need, size_t keep) {
    while (charged + need > budget) {
      size_t victim = engines.size();
      for (size_t k = 0; k < engines.size(); ++k) {
        const auto &e = engines[k];
        if (!e || k == keep || !e->idle()) continue;
        if (victim == engines.size() || e->last_used < engines[victim]->last_used)
          victim = k;
      }
      if (victim == engines.size()) return false;
      evict(victim);
    }
    return true;
  }

  // Bytes a query of (kind, graph) needs beyond the solo buffers at the
  // least: its engine's state and one result buffer (0 if never measured).
  size_t minimumNeed(int kind, int graph) const {
    return footprint[key(kind, graph)] + resultBytes(graph);
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

size_t solo_bytes(const gas_server *s) {
  size_t b = 0;
  for (const auto &set : s->solo_out)
    for (const auto &buf : set) b += buf.bytes();
  return b;
}

// The solo threads' result buffers, the plain-engine cost of every kind per
// graph, and a batch of one lane on each batch engine, measured once on an
// idle server.  Each batch engine is built for the measurement (its state
// size is recorded) and kept only while all of them fit the budget.
void calibrate(gas_server *s) {
  if (s->calibrated) return;
  struct SetGuard {
    ~SetGuard() { engine_set() = 0; }
  } set_guard;
  for (int set = kSoloThreads - 1; set >= 0; --set) {
    engine_set() = set;
    for (size_t gi = 0; gi < s->graphs.size(); ++gi) {
      auto &buf = s->solo_out[set][gi];
      if (buf.n < (size_t)s->nV[gi]) {
        buf.alloc((size_t)s->nV[gi]);
        s->charged += buf.bytes();
      }
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
        s->solo_ms[s->key(kind, (int)gi)] = (now_s() - ts) * 1e3;
      }
    }
  }
  for (size_t gi = 0; gi < s->graphs.size(); ++gi) {
    for (int kind : {GAS_BATCH_BFS, GAS_BATCH_SSSP, GAS_BATCH_PPR}) {
      const size_t key = s->key(kind, (int)gi);
      if (s->solo_ms[key] >= 1e30) continue;  // kind unavailable
      Engine *e = s->find(kind, (int)gi);
      if (!e) e = &s->create(kind, (int)gi);
      const int32_t lane = 0, src = 0;
      const double ts = now_s();
      if (e->be) {
        e->be->addLanes(&lane, &src, 1);
        e->be->runToCompletion();
        e->be->sync();
        e->be->releaseLanes(1);
      } else {
        e->pe->addLanes(&lane, &src, 1);
        e->pe->runToCompletion();
        e->pe->sync();
        e->pe->releaseLanes(1);
      }
      s->batch1_ms[key] = (now_s() - ts) * 1e3;
      if (s->charged > s->budget) s->evict(key);
    }
  }
  s->evictions = 0;
  clear_error();
  s->calibrated = true;
}

// GAS_ERR_FULL if a query of (kind, graph) can never fit the budget.
int check_fits(const gas_server *s, int kind, int graph, int index) {
  const size_t need = s->minimumNeed(kind, graph) + solo_bytes(s);
  if (need > s->budget)
    return set_error(GAS_ERR_FULL,
                     "query %d (kind %d, graph %d) needs %zu bytes of device memory "
                     "(batch engine state, a result buffer and the solo buffers); the "
                     "budget is %zu",
                     index, kind, graph, need, s->budget);
  return GAS_OK;
}

// One serving session: the loop over the queries handed to it.
struct Session {
  gas_server *s;
  std::vector<QRec> q;
  std::deque<int> waiting;
  int done = 0;
  int64_t rounds = 0, admissions_deferred = 0;
  int max_running = 0, solo = 0;
  double t0 = now_s();
  // called on the loop thread when query i finished: latency in ms
  std::function<void(int, double)> finished;
  struct StreamGuard {
    gasrt_stream_t st = nullptr;
    ~StreamGuard() {
      if (st) gasrt_stream_destroy(st);
    }
  } copy_guard[kSoloThreads];
  std::unique_ptr<SoloWorker[]> workers;  // joined before the streams go

  explicit Session(gas_server *srv) : s(srv) {
    for (auto &c : copy_guard) check(gasrt_stream_create(&c.st), "stream");
    workers.reset(new SoloWorker[kSoloThreads]);
    for (int w = 0; w < kSoloThreads; ++w) {
      workers[w].start(w, [this, w](int, const QRec &r) {
        auto &buf = s->solo_out[w][r.graph];
        int rc = r.kind == GAS_BATCH_BFS
                     ? gas_bfs_device(s->graphs[r.graph], r.source, buf.p, nullptr)
                 : r.kind == GAS_BATCH_SSSP
                     ? gas_sssp_device(s->graphs[r.graph], r.source, buf.p, nullptr)
                     : gas_ppr_device(s->graphs[r.graph], r.source, s->ppr_iterations,
                                      reinterpret_cast<float *>(buf.p), nullptr);
        if (rc == GAS_OK && r.host_out)
          rc = gasrt_memcpy_d2h(r.host_out, buf.p, s->resultBytes(r.graph),
                                copy_guard[w].st);
        return rc;
      });
    }
  }
  ~Session() { workers.reset(); }

  double clock() const { return now_s() - t0; }

  // Hands query r (arrived now or at r.arrival) to the loop; returns its index.
  int add(const QRec &r) {
    q.push_back(r);
    waiting.push_back((int)q.size() - 1);
    return (int)q.size() - 1;
  }

  int runningTotal() const {
    int r = 0;
    for (auto &ep : s->engines)
      if (ep) r += __builtin_popcountll(ep->running);
    return r;
  }

  bool busy() const {
    if (!waiting.empty()) return true;
    for (int w = 0; w < kSoloThreads; ++w)
      if (workers[w].pending > 0) return true;
    for (auto &ep : s->engines)
      if (ep && !ep->idle()) return true;
    return false;
  }

  // One pass of the loop; true if anything happened.
  bool iterate() {
    bool progressed = false;
    // 0. extractions whose results are in HBM: their queries are done
    for (auto &ep : s->engines) {
      if (!ep) continue;
      Engine &e = *ep;
      while (!e.ext.empty()) {
        int ok = 0;
        check(gasrt_event_query(e.ext.front().ev, &ok), "event query");
        if (!ok) break;
        const double tf = clock();
        for (int i : e.ext.front().queries) {
          ++done;
          finished(i, (tf - q[i].arrival) * 1e3);
        }
        e.spare_ev.push_back(e.ext.front().ev);
        e.ext.pop_front();
        progressed = true;
      }
    }
    // 1. engines whose round finished: extract the results of the queries
    // that ended and free their lanes (stream order keeps the extraction
    // ahead of any reuse of the lanes); without a host copy to make, the
    // loop does not wait for the extraction
    for (auto &ep : s->engines) {
      if (!ep || !ep->inflight) continue;
      Engine &e = *ep;
      if (!(e.be ? e.be->poll() : e.pe->poll())) continue;
      e.inflight = false;
      progressed = true;
      const uint64_t alive = e.be ? e.be->aliveSnapshot() : e.pe->lanesAlive();
      const uint64_t fin_all = e.running & ~alive;
      if (!fin_all) continue;
      int32_t *outs[BW] = {};
      float *outs_f[BW] = {};
      for (uint64_t fin = fin_all; fin; fin &= fin - 1) {
        const int lane = __builtin_ctzll(fin);
        if (e.be)
          outs[lane] = e.res_i[lane].p;
        else
          outs_f[lane] = e.res_f[lane].p;
      }
      if (e.be)
        e.be->laneResultsDevice(fin_all, outs);
      else
        e.pe->laneResultsDevice(fin_all, outs_f);
      bool to_host = false;
      for (uint64_t fin = fin_all; fin; fin &= fin - 1)
        to_host |= q[e.lane_query[__builtin_ctzll(fin)]].host_out != nullptr;
      Engine::Extraction x;
      if (to_host) {
        check(gasrt_stream_synchronize(e.stream()), "sync");
      } else {
        if (e.spare_ev.empty()) {
          gasrt_event_t ev = nullptr;
          check(gasrt_event_create(&ev), "event");
          e.spare_ev.push_back(ev);
        }
        x.ev = e.spare_ev.back();
        e.spare_ev.pop_back();
        check(gasrt_event_record(x.ev, e.stream()), "event");
      }
      const double tf = clock();
      for (uint64_t fin = fin_all; fin; fin &= fin - 1) {
        const int lane = __builtin_ctzll(fin);
        const int i = e.lane_query[lane];
        if (q[i].host_out) {
          void *src = e.be ? (void *)e.res_i[lane].p : (void *)e.res_f[lane].p;
          check(gasrt_memcpy_d2h(q[i].host_out, src, 4 * (size_t)e.V, e.stream()),
                "result copy");
        }
        e.lane_query[lane] = -1;
        e.running &= ~(1ull << lane);
        if (to_host) {
          ++done;
          finished(i, (tf - q[i].arrival) * 1e3);
        } else {
          x.queries.push_back(i);
        }
      }
      if (!to_host) e.ext.push_back(std::move(x));
      if (e.be)
        e.be->releaseLanes(fin_all);
      else
        e.pe->releaseLanes(fin_all);
      // state the engine grew while extracting (single-lane scratch)
      const size_t now_b = e.stateBytes() + [&] {
        size_t b = 0;
        for (int l = 0; l < BW; ++l) b += e.res_i[l].bytes() + e.res_f[l].bytes();
        return b;
      }();
      s->charged += now_b - e.charged;
      e.charged = now_b;
    }
    // 2. queries the solo threads finished
    for (int w = 0; w < kSoloThreads; ++w)
      for (const auto &f : workers[w].take()) {
        const int i = f.first;
        workers[w].queued_ms -= s->solo_ms[s->key(q[i].kind, q[i].graph)];
        if (--workers[w].pending == 0) workers[w].queued_ms = 0;
        ++done;
        ++solo;
        progressed = true;
        finished(i, (f.second - t0 - q[i].arrival) * 1e3);
      }
    // 3. waiting queries go to the solo thread with the shortest queue, in
    // arrival order, when their plain-engine time is at most solo_max_ms, or
    // at most solo_idle_max_ms while fewer than kSoloBacklog queries of their
    // kind and graph wait (more would share a batch), fewer than
    // kSoloMaxInflight queries wait or run in batches, and that queue would
    // finish them no later than a batch of one (an idle solo thread takes
    // one regardless; queries above solo_max_ms queue on the first thread up
    // to solo_idle_max_ms of work).  Solo threads after the first take
    // queries only while at most kSoloBacklog queries wait in all: under
    // load the batches answer them for less GPU time each.  BFS / SSSP results are exact on either
    // engine; a PPR result from the plain engine is within the same 1e-5 of
    // FP64 as a batch lane but not bit-identical to one.
    {
      std::vector<int> backlog(s->engines.size(), 0);
      for (int i : waiting) ++backlog[s->key(q[i].kind, q[i].graph)];
      const int n_solo = waiting.size() <= (size_t)kSoloBacklog ? kSoloThreads : 1;
      int solo_pending = 0;
      for (int w = 0; w < kSoloThreads; ++w) solo_pending += workers[w].pending;
      const bool loaded =
          (int)waiting.size() + runningTotal() + solo_pending >= kSoloMaxInflight;
      for (auto it = waiting.begin(); it != waiting.end();) {
        const int i = *it;
        const size_t key = s->key(q[i].kind, q[i].graph);
        const double ms = s->solo_ms[key];
        const double cap =
            backlog[key] < kSoloBacklog ? s->solo_idle_max_ms : s->solo_max_ms;
        const double budget = std::max(kSoloQueueMs, s->batch1_ms[key]);
        // queries above solo_max_ms queue on the first solo thread, one
        // after another, up to solo_idle_max_ms of queued work: two such
        // runs at once share the GPU's bandwidth and both finish late
        const bool heavy = ms > s->solo_max_ms;
        SoloWorker *wk = &workers[0];
        for (int w = 1; w < n_solo && !heavy; ++w)
          if (workers[w].queued_ms < wk->queued_ms ||
              (workers[w].pending == 0 && wk->pending > 0))
            wk = &workers[w];
        const double limit = heavy ? std::max(budget, s->solo_idle_max_ms) : budget;
        if (ms > cap || loaded || (wk->pending > 0 && wk->queued_ms + ms > limit)) {
          ++it;
          continue;
        }
        wk->queued_ms += ms;
        ++wk->pending;
        wk->submit(i, q[i]);
        it = waiting.erase(it);
        progressed = true;
      }
    }
    // 4. admission into batch lanes (first come first served per engine; a
    // query that does not fit does not block other engines), only into
    // engines between rounds; engine state and lane result buffers are
    // charged when made, idle engines evicted to make room
    std::vector<char> blocked(s->engines.size(), 0);
    std::vector<int> backlog(s->engines.size(), 0);
    for (int i : waiting) ++backlog[s->key(q[i].kind, q[i].graph)];
    for (auto it = waiting.begin(); it != waiting.end();) {
      const int i = *it;
      const size_t key = s->key(q[i].kind, q[i].graph);
      if (blocked[key]) {
        ++it;
        continue;
      }
      Engine *ep = s->find(q[i].kind, q[i].graph);
      if (!ep) {
        if (!s->makeRoom(s->minimumNeed(q[i].kind, q[i].graph), key)) {
          blocked[key] = 1;
          ++admissions_deferred;
          ++it;
          continue;
        }
        ep = &s->create(q[i].kind, q[i].graph);
      }
      Engine &e = *ep;
      uint64_t busy_l = e.be ? e.be->lanesUsed() : e.pe->lanesUsed();
      for (int32_t l : e.add_lanes) busy_l |= 1ull << l;
      // a backlog waits for a BFS / SSSP batch past its first rounds to
      // drain (PPR lanes cost the same dense sweeps at any offset)
      const bool hold = e.be && e.running && s->gang_backlog > 0 &&
                        backlog[key] >= s->gang_backlog &&
                        e.be->round() - e.gang_start > s->gang_rounds;
      if (e.inflight || busy_l == ~0ull || hold) {
        blocked[key] = 1;
        if (!e.inflight) ++admissions_deferred;
        ++it;
        continue;
      }
      // the lowest free lane, preferring one whose result buffer exists
      int lane = -1;
      for (uint64_t f = ~busy_l; f; f &= f - 1) {
        const int l = __builtin_ctzll(f);
        if (lane < 0) lane = l;
        if (e.hasResult(l)) {
          lane = l;
          break;
        }
      }
      if (!e.hasResult(lane)) {
        const size_t rb = s->resultBytes(q[i].graph);
        if (!s->makeRoom(rb, key)) {
          blocked[key] = 1;
          ++admissions_deferred;
          ++it;
          continue;
        }
        if (e.be)
          e.res_i[lane].alloc((size_t)e.V);
        else
          e.res_f[lane].alloc((size_t)e.V);
        s->charged += rb;
        e.charged += rb;
      }
      if (e.be && !e.running && e.add_lanes.empty()) e.gang_start = e.be->round();
      e.add_lanes.push_back(lane);
      e.add_src.push_back(q[i].source);
      e.lane_query[lane] = i;
      e.last_used = now_s();
      it = waiting.erase(it);
    }
    // 5. start the next rounds of every idle engine with work: one round
    // when a query waits for the engine (it joins after that round), several
    // otherwise (fewer host round trips per query)
    std::vector<char> wants(s->engines.size(), 0);
    for (int i : waiting) wants[s->key(q[i].kind, q[i].graph)] = 1;
    for (size_t ei = 0; ei < s->engines.size(); ++ei) {
      auto &ep = s->engines[ei];
      if (!ep || ep->inflight) continue;
      Engine &e = *ep;
      if (!e.add_lanes.empty()) {
        if (e.be)
          e.be->addLanes(e.add_lanes.data(), e.add_src.data(), (int)e.add_lanes.size());
        else
          e.pe->addLanes(e.add_lanes.data(), e.add_src.data(), (int)e.add_lanes.size());
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
    max_running = std::max(max_running, runningTotal());
    return progressed;
  }

  bool anyInflight() const {
    for (int w = 0; w < kSoloThreads; ++w)
      if (workers[w].pending > 0) return true;
    for (auto &ep : s->engines)
      if (ep && (ep->inflight || !ep->ext.empty())) return true;
    return false;
  }

  void fillStats(gas_serve_stats *st) const {
    memset(st, 0, sizeof(*st));
    st->elapsed_s = clock();
    st->rounds = rounds;
    st->admissions_deferred = admissions_deferred;
    st->max_concurrent = max_running;
    st->solo = solo;
    st->evictions = s->evictions;
    st->device_bytes = (int64_t)s->charged;
  }
};

bool valid_query(const gas_server *s, int32_t kind, int32_t graph, int32_t source) {
  return kind >= 0 && kind <= 2 && graph >= 0 && graph < (int)s->graphs.size() &&
         source >= 0 && source < s->nV[graph];
}

// The online loop: takes submitted queries as they come, until stopped with
// nothing left to serve.
void online_loop(gas_server *s) {
  std::unique_ptr<Session> ses;
  std::vector<int64_t> ticket_of;  // session index -> ticket
  int rc = guarded_s([&] {
    ses.reset(new Session(s));
    ses->finished = [&](int i, double ms) {
      std::lock_guard<std::mutex> lk(s->omu);
      auto &t = s->tickets[(size_t)ticket_of[(size_t)i]];
      t.state = 1;
      t.latency_ms = ms;
      s->ocv.notify_all();
    };
    for (;;) {
      bool stop = false;
      {
        std::lock_guard<std::mutex> lk(s->omu);
        const double t = ses->clock();
        for (auto &tr : s->inbox) {
          tr.second.arrival = t;
          ses->add(tr.second);
          ticket_of.push_back(tr.first);
        }
        s->inbox.clear();
        stop = s->ostop;
      }
      const bool progressed = ses->iterate();
      if (stop && !ses->busy()) break;
      if (!progressed) {
        if (ses->anyInflight() || !ses->waiting.empty()) {
          std::this_thread::yield();
        } else {
          std::unique_lock<std::mutex> lk(s->omu);
          s->ocv.wait_for(lk, std::chrono::milliseconds(2),
                          [&] { return s->ostop || !s->inbox.empty(); });
        }
      }
    }
    std::lock_guard<std::mutex> lk(s->omu);
    ses->fillStats(&s->ostats);
  });
  std::lock_guard<std::mutex> lk(s->omu);
  if (rc != GAS_OK) {
    s->orc = rc;
    s->oerr = gas_last_error();
    for (auto &t : s->tickets)
      if (t.state == 0) t.state = -1;
  }
  s->ocv.notify_all();
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
    if (memory_budget_bytes == 0) {
      size_t free_b = 0, total_b = 0;
      check(gasrt_mem_info(&free_b, &total_b), "memory info");
      memory_budget_bytes = (size_t)(kDefaultBudgetShare * (double)free_b);
    }
    s->budget = memory_budget_bytes;
    s->ppr_iterations = ppr_iterations;
    s->engines.resize(3 * (size_t)n_graphs);
    s->footprint.assign(3 * (size_t)n_graphs, 0);
    s->solo_ms.assign(3 * (size_t)n_graphs, 1e30);
    s->batch1_ms.assign(3 * (size_t)n_graphs, 0.0);
    for (auto &v : s->solo_out) v.resize((size_t)n_graphs);
    *out = s.release();
  });
}

int gas_server_budget(gas_server_t s, size_t *budget_bytes, size_t *held_bytes) {
  if (!s) return set_error(GAS_ERR_ARG, "gas_server_budget: null");
  std::lock_guard<std::mutex> lk(s->omu);
  if (budget_bytes) *budget_bytes = s->budget;
  if (held_bytes) *held_bytes = s->charged;
  return GAS_OK;
}

int gas_server_set_budget(gas_server_t s, size_t budget_bytes) {
  if (!s) return set_error(GAS_ERR_ARG, "gas_server_set_budget: null");
  std::lock_guard<std::mutex> lk(s->omu);
  if (s->online) return set_error(GAS_ERR_STATE, "gas_server_set_budget: serving online");
  return guarded_s([&] {
    for (size_t k = 0; k < s->engines.size(); ++k) s->evict(k);
    s->evictions = 0;
    s->budget = budget_bytes;
  });
}

int gas_server_engine_bytes(gas_server_t s, int32_t kind, int32_t graph_id, size_t *bytes) {
  if (!s || !bytes || kind < 0 || kind > 2 || graph_id < 0 ||
      graph_id >= (int32_t)s->graphs.size())
    return set_error(GAS_ERR_ARG, "gas_server_engine_bytes: bad arguments");
  std::lock_guard<std::mutex> lk(s->omu);
  *bytes = s->footprint[s->key(kind, graph_id)];
  return GAS_OK;
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

int gas_server_stop(gas_server_t s, gas_serve_stats *stats) {
  if (!s) return set_error(GAS_ERR_ARG, "gas_server_stop: null");
  {
    std::lock_guard<std::mutex> lk(s->omu);
    if (!s->online) return set_error(GAS_ERR_STATE, "gas_server_stop: not started");
    s->ostop = true;
  }
  s->ocv.notify_all();
  if (s->othread.joinable()) s->othread.join();
  std::lock_guard<std::mutex> lk(s->omu);
  s->online = false;
  s->ostop = false;
  if (stats) *stats = s->ostats;
  if (s->orc != GAS_OK) {
    const int rc = s->orc;
    s->orc = GAS_OK;
    return set_error(rc, "%s", s->oerr.c_str());
  }
  return GAS_OK;
}

int gas_server_free(gas_server_t s) {
  if (!s) return GAS_OK;
  bool running = false;
  {
    std::lock_guard<std::mutex> lk(s->omu);
    running = s->online;
  }
  if (running) gas_server_stop(s, nullptr);
  delete s;
  clear_error();
  return GAS_OK;
}

int gas_server_start(gas_server_t s) {
  if (!s) return set_error(GAS_ERR_ARG, "gas_server_start: null");
  return guarded_s([&] {
    {
      std::lock_guard<std::mutex> lk(s->omu);
      if (s->online) throw Error(GAS_ERR_STATE, "gas_server_start: already started");
    }
    calibrate(s);
    std::lock_guard<std::mutex> lk(s->omu);
    s->online = true;
    s->ostop = false;
    s->orc = GAS_OK;
    s->othread = std::thread(online_loop, s);
  });
}

int gas_server_submit(gas_server_t s, int32_t kind, int32_t graph_id, int32_t source,
                      void *host_result, int64_t *ticket) {
  if (!s || !ticket) return set_error(GAS_ERR_ARG, "gas_server_submit: null");
  if (!valid_query(s, kind, graph_id, source))
    return set_error(GAS_ERR_ARG, "gas_server_submit: invalid query");
  std::lock_guard<std::mutex> lk(s->omu);
  if (!s->online || s->ostop)
    return set_error(GAS_ERR_STATE, "gas_server_submit: server not started");
  if (s->orc != GAS_OK) return set_error(s->orc, "%s", s->oerr.c_str());
  GAS_TRY(check_fits(s, kind, graph_id, (int)s->tickets.size()));
  QRec r;
  r.kind = kind;
  r.graph = graph_id;
  r.source = source;
  r.host_out = host_result;
  *ticket = (int64_t)s->tickets.size();
  s->tickets.emplace_back();
  s->inbox.emplace_back(*ticket, r);
  s->ocv.notify_all();
  return GAS_OK;
}

int gas_server_poll(gas_server_t s, int64_t ticket, int32_t *state, double *latency_ms) {
  if (!s || !state) return set_error(GAS_ERR_ARG, "gas_server_poll: null");
  std::lock_guard<std::mutex> lk(s->omu);
  if (ticket < 0 || ticket >= (int64_t)s->tickets.size())
    return set_error(GAS_ERR_ARG, "gas_server_poll: unknown ticket %lld", (long long)ticket);
  const auto &t = s->tickets[(size_t)ticket];
  *state = t.state;
  if (latency_ms) *latency_ms = t.latency_ms;
  if (t.state < 0) return set_error(s->orc ? s->orc : GAS_ERR_STATE, "%s", s->oerr.c_str());
  return GAS_OK;
}

int gas_server_wait(gas_server_t s, int64_t ticket, double timeout_s, double *latency_ms) {
  if (!s) return set_error(GAS_ERR_ARG, "gas_server_wait: null");
  std::unique_lock<std::mutex> lk(s->omu);
  if (ticket < 0 || ticket >= (int64_t)s->tickets.size())
    return set_error(GAS_ERR_ARG, "gas_server_wait: unknown ticket %lld", (long long)ticket);
  const auto deadline = std::chrono::steady_clock::now() +
                        std::chrono::duration<double>(std::max(0.0, timeout_s));
  const bool ok = s->ocv.wait_until(lk, deadline, [&] {
    return s->tickets[(size_t)ticket].state != 0;
  });
  const auto &t = s->tickets[(size_t)ticket];
  if (!ok) return set_error(GAS_ERR_STATE, "gas_server_wait: ticket %lld not done after %g s",
                            (long long)ticket, timeout_s);
  if (t.state < 0) return set_error(s->orc ? s->orc : GAS_ERR_STATE, "%s", s->oerr.c_str());
  if (latency_ms) *latency_ms = t.latency_ms;
  return GAS_OK;
}

int gas_server_replay(gas_server_t s, int32_t n, const int32_t *kinds,
                      const int32_t *graph_ids, const int32_t *sources,
                      const double *arrival_s, void *const *host_results,
                      double *latency_ms, gas_serve_stats *stats) {
  if (!s || n < 0 || (n > 0 && (!kinds || !graph_ids || !sources ||
                                !arrival_s || !latency_ms)))
    return set_error(GAS_ERR_ARG, "gas_server_replay: bad arguments");
  {
    std::lock_guard<std::mutex> lk(s->omu);
    if (s->online)
      return set_error(GAS_ERR_STATE, "gas_server_replay: the server is serving online");
  }
  for (int i = 0; i < n; ++i) {
    if (!valid_query(s, kinds[i], graph_ids[i], sources[i]))
      return set_error(GAS_ERR_ARG, "gas_server_replay: query %d invalid", i);
    if (i && arrival_s[i] < arrival_s[i - 1])
      return set_error(GAS_ERR_ARG, "gas_server_replay: arrivals not sorted");
  }
  return guarded_s([&] {
    calibrate(s);
    for (int i = 0; i < n; ++i) check(check_fits(s, kinds[i], graph_ids[i], i), "admission");
    Session ses(s);
    ses.finished = [&](int i, double ms) { latency_ms[i] = ms; };
    int next = 0;
    while (ses.done < n) {
      const double t = ses.clock();
      while (next < n && arrival_s[next] <= t) {
        QRec r;
        r.kind = kinds[next];
        r.graph = graph_ids[next];
        r.source = sources[next];
        r.arrival = arrival_s[next];
        r.host_out = host_results ? host_results[next] : nullptr;
        ses.add(r);
        ++next;
      }
      if (!ses.iterate()) {
        if (!ses.anyInflight() && ses.waiting.empty() && next < n) {
          const double wait = arrival_s[next] - ses.clock();
          if (wait > 0.0005)
            std::this_thread::sleep_for(std::chrono::duration<double>(wait * 0.8));
        } else {
          std::this_thread::yield();
        }
      }
    }
    if (stats) ses.fillStats(stats);
  });
}

}  // extern "C"
