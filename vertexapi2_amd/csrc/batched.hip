// C ABI of the batched multi-source engines (gas_batch_*): BFS / SSSP lanes
// (BatchEngine) and personalised PageRank lanes (PprBatch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "gas.h"
#include "gas/batch.hpp"
#include "gas_internal.hpp"

using namespace gas;

struct gas_batch {
  std::unique_ptr<BatchEngine> e;   // BFS / SSSP
  std::unique_ptr<PprBatch> pp;     // PPR
};

namespace {

template <class F>
int guarded_b(F f) {
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

#define BATCH_OR_FAIL(b)                                           \
  if (!(b) || (!(b)->e && !(b)->pp))                               \
  return set_error(GAS_ERR_ARG, "%s: null batch", __func__)

extern "C" {

int gas_batch_create(gas_graph_t g, int kind, gas_batch_t *out) {
  if (kind == GAS_BATCH_PPR) return gas_batch_create_ppr(g, 10, out);
  if (!g || !out || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP))
    return set_error(GAS_ERR_ARG, "gas_batch_create: bad arguments");
  return guarded_b([&] {
    std::unique_ptr<gas_batch> b(new gas_batch);
    b->e.reset(new BatchEngine(g, kind == GAS_BATCH_BFS
                                      ? BatchEngine::BFS_KIND
                                      : BatchEngine::SSSP_KIND));
    *out = b.release();
  });
}

int gas_batch_create_ppr(gas_graph_t g, int32_t iterations, gas_batch_t *out) {
  if (!g || !out || iterations < 1)
    return set_error(GAS_ERR_ARG, "gas_batch_create_ppr: bad arguments");
  return guarded_b([&] {
    std::unique_ptr<gas_batch> b(new gas_batch);
    b->pp.reset(new PprBatch(g, iterations));
    *out = b.release();
  });
}

int gas_batch_add(gas_batch_t b, const int32_t *lanes, const int32_t *sources,
                  int32_t n) {
  BATCH_OR_FAIL(b);
  if (n < 0 || (n > 0 && (!lanes || !sources)))
    return set_error(GAS_ERR_ARG, "gas_batch_add: bad arguments");
  return guarded_b([&] {
    if (b->e)
      b->e->addLanes(lanes, sources, n);
    else
      b->pp->addLanes(lanes, sources, n);
  });
}

int gas_batch_step(gas_batch_t b, int32_t rounds) {
  BATCH_OR_FAIL(b);
  return guarded_b([&] {
    if (b->e)
      b->e->advance(rounds);
    else
      for (int i = 0; i < rounds; ++i) b->pp->step();
  });
}

int gas_batch_alive(gas_batch_t b, uint64_t *alive) {
  BATCH_OR_FAIL(b);
  if (!alive) return set_error(GAS_ERR_ARG, "gas_batch_alive: null");
  return guarded_b([&] {
    *alive = b->e ? b->e->lanesAlive() : b->pp->lanesAlive();
  });
}

int gas_batch_used(gas_batch_t b, uint64_t *used) {
  BATCH_OR_FAIL(b);
  if (!used) return set_error(GAS_ERR_ARG, "gas_batch_used: null");
  *used = b->e ? b->e->lanesUsed() : b->pp->lanesUsed();
  return GAS_OK;
}

int gas_batch_release(gas_batch_t b, uint64_t lanes) {
  BATCH_OR_FAIL(b);
  return guarded_b([&] {
    if (b->e)
      b->e->releaseLanes(lanes);
    else
      b->pp->releaseLanes(lanes);
  });
}

int gas_batch_run(gas_batch_t b, gas_run_stats *stats) {
  BATCH_OR_FAIL(b);
  return guarded_b([&] {
    gasrt_stream_t s = b->e ? b->e->stream() : b->pp->stream();
    gasrt_event_t a = nullptr, z = nullptr;
    check(gasrt_event_create(&a), "event");
    check(gasrt_event_create(&z), "event");
    check(gasrt_event_record(a, s), "event");
    const int rounds =
        b->e ? b->e->runToCompletion() : b->pp->runToCompletion();
    check(gasrt_event_record(z, s), "event");
    check(gasrt_event_synchronize(z), "event");
    float ms = 0;
    check(gasrt_event_elapsed_ms(a, z, &ms), "elapsed");
    gasrt_event_destroy(a);
    gasrt_event_destroy(z);
    if (stats) {
      memset(stats, 0, sizeof(*stats));
      stats->supersteps = rounds;
      stats->ms = ms;
    }
  });
}

int gas_batch_result(gas_batch_t b, int32_t lane, int32_t *host_out) {
  BATCH_OR_FAIL(b);
  if (!host_out || !b->e)
    return set_error(GAS_ERR_ARG, "gas_batch_result: null or PPR batch");
  return guarded_b([&] { b->e->laneResult(lane, host_out); });
}

int gas_batch_result_device(gas_batch_t b, int32_t lane, int32_t *dev_out) {
  BATCH_OR_FAIL(b);
  if (!dev_out || !b->e)
    return set_error(GAS_ERR_ARG, "gas_batch_result_device: null or PPR batch");
  return guarded_b([&] {
    b->e->laneResultDevice(lane, dev_out);
    b->e->sync();
  });
}

int gas_batch_results_device(gas_batch_t b, uint64_t lanes,
                             int32_t *const *dev_outs) {
  BATCH_OR_FAIL(b);
  if (!dev_outs || !b->e)
    return set_error(GAS_ERR_ARG, "gas_batch_results_device: null or PPR batch");
  return guarded_b([&] {
    b->e->laneResultsDevice(lanes, dev_outs);
    b->e->sync();
  });
}

int gas_batch_result_f32(gas_batch_t b, int32_t lane, float *host_out) {
  BATCH_OR_FAIL(b);
  if (!host_out || !b->pp)
    return set_error(GAS_ERR_ARG, "gas_batch_result_f32: null or not PPR");
  return guarded_b([&] { b->pp->laneResult(lane, host_out); });
}

int gas_batch_result_f32_device(gas_batch_t b, int32_t lane, float *dev_out) {
  BATCH_OR_FAIL(b);
  if (!dev_out || !b->pp)
    return set_error(GAS_ERR_ARG, "gas_batch_result_f32_device: bad batch");
  return guarded_b([&] {
    b->pp->laneResultDevice(lane, dev_out);
    b->pp->sync();
  });
}

int gas_batch_results_f32_device(gas_batch_t b, uint64_t lanes,
                                 float *const *dev_outs) {
  BATCH_OR_FAIL(b);
  if (!dev_outs || !b->pp)
    return set_error(GAS_ERR_ARG, "gas_batch_results_f32_device: null or not PPR");
  return guarded_b([&] {
    b->pp->laneResultsDevice(lanes, dev_outs);
    b->pp->sync();
  });
}

int gas_batch_set_pull_divisor(gas_batch_t b, int32_t divisor) {
  BATCH_OR_FAIL(b);
  if (divisor < 0)
    return set_error(GAS_ERR_ARG, "gas_batch_set_pull_divisor: need >= 0");
  if (b->e) b->e->setPullDivisor(divisor);
  return GAS_OK;
}

int gas_batch_set_sssp_delta(gas_batch_t b, int32_t delta) {
  BATCH_OR_FAIL(b);
  if (delta < 0)
    return set_error(GAS_ERR_ARG, "gas_batch_set_sssp_delta: need >= 0");
  if (b->e) b->e->setSsspDelta(delta);
  return GAS_OK;
}

int gas_graph_batch_param(gas_graph_t g, int kind, int32_t *param) {
  if (!g || !param || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP))
    return set_error(GAS_ERR_ARG, "gas_graph_batch_param: bad arguments");
  *param = g->batch_param[kind];
  return GAS_OK;
}

int gas_graph_batch_hub_items(gas_graph_t g, int kind, int32_t *items) {
  if (!g || !items || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP))
    return set_error(GAS_ERR_ARG, "gas_graph_batch_hub_items: bad arguments");
  *items = g->batch_hub[kind] ? g->batch_hub[kind] : HUB_ITEMS;
  return GAS_OK;
}

int gas_graph_batch_graph_rounds(gas_graph_t g, int kind, int32_t *rounds) {
  if (!g || !rounds || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP))
    return set_error(GAS_ERR_ARG, "gas_graph_batch_graph_rounds: bad arguments");
  *rounds = g->batch_graph[kind];
  return GAS_OK;
}

int gas_graph_set_batch_graph_rounds(gas_graph_t g, int kind, int32_t rounds) {
  if (!g || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP) || rounds < 0 ||
      rounds > BatchEngine::kMaxGraphRounds)
    return set_error(GAS_ERR_ARG, "gas_graph_set_batch_graph_rounds: bad arguments");
  g->batch_graph[kind] = rounds;
  return GAS_OK;
}

int gas_batch_set_graph_rounds(gas_batch_t b, int32_t rounds) {
  BATCH_OR_FAIL(b);
  if (!b->e)
    return set_error(GAS_ERR_ARG, "gas_batch_set_graph_rounds: not a BFS / SSSP batch");
  if (rounds < 0 || rounds > BatchEngine::kMaxGraphRounds)
    return set_error(GAS_ERR_ARG, "gas_batch_set_graph_rounds: need 0..%d",
                     BatchEngine::kMaxGraphRounds);
  return guarded_b([&] { b->e->setGraphRounds(rounds); });
}

int gas_graph_batch_item_edges(gas_graph_t g, int kind, int32_t *edges) {
  if (!g || !edges || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP))
    return set_error(GAS_ERR_ARG, "gas_graph_batch_item_edges: bad arguments");
  *edges = g->batch_items[kind] ? g->batch_items[kind]
                                : (kind == GAS_BATCH_BFS ? BFS_ITEM_EDGES : SSSP_ITEM_EDGES);
  return GAS_OK;
}

int gas_batch_set_item_edges(gas_batch_t b, int32_t edges) {
  BATCH_OR_FAIL(b);
  if (!b->e)
    return set_error(GAS_ERR_ARG, "gas_batch_set_item_edges: not a BFS / SSSP batch");
  return guarded_b([&] { b->e->setItemEdges(edges); });
}

int gas_batch_set_hub_items(gas_batch_t b, int32_t items) {
  BATCH_OR_FAIL(b);
  if (items < HUB_ITEMS_MIN)
    return set_error(GAS_ERR_ARG, "gas_batch_set_hub_items: need >= %d", HUB_ITEMS_MIN);
  return guarded_b([&] {
    if (b->e) b->e->setHubItems(items);
  });
}

int gas_graph_autotune_batch(gas_graph_t g, int kind, int32_t *param_out) {
  if (!g || (kind != GAS_BATCH_BFS && kind != GAS_BATCH_SSSP))
    return set_error(GAS_ERR_ARG, "gas_graph_autotune_batch: bad arguments");
  if (kind == GAS_BATCH_SSSP && !g->csr_w && g->nE > 0)
    return set_error(GAS_ERR_ARG, "gas_graph_autotune_batch: no weights");
  return guarded_b([&] {
    // kSamples samples of up to 64 sources drawn uniformly (SplitMix64)
    // from the vertices with out-edges -- the distribution batches are
    // served from.  A setting replaces the default (the first candidate)
    // only if it is faster in total and no slower than 1.1x the default on
    // every sample: settings such as pull on skewed graphs win on batches
    // whose lanes all start near the core and lose several-fold on those
    // with a lane far from it.
    constexpr int kSamples = 6;
    const int64_t V = g->nV;
    int64_t with_out = 0;
    for (int64_t v = 0; v < V; ++v) with_out += g->h_csr_off[v + 1] > g->h_csr_off[v];
    std::vector<int32_t> samples[kSamples];
    for (int k = 0; k < kSamples; ++k) {
      // distinct streams: start points far apart on the SplitMix64 sequence
      uint64_t x = 0xD1342543DE82EF95ull * (uint64_t)(k + 1) + 0x2545F4914F6CDD1Dull;
      auto &src = samples[k];
      const int want = (int)std::min<int64_t>(BW, with_out);
      for (int64_t tries = 0; (int)src.size() < want && tries < 64 * V + 1024; ++tries) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const int32_t v = (int32_t)(z % (uint64_t)std::max<int64_t>(V, 1));
        if (g->h_csr_off[v + 1] > g->h_csr_off[v] &&
            std::find(src.begin(), src.end(), v) == src.end())
          src.push_back(v);
      }
      if (src.empty()) src.push_back(0);
    }
    std::vector<int32_t> lanes;
    for (int i = 0; i < BW; ++i) lanes.push_back(i);
    std::vector<int32_t> cand;
    if (kind == GAS_BATCH_BFS) {
      cand = {0, 2, 8, 32};
    } else {
      const int32_t w = g->max_w;
      for (int32_t c : {0, w / 4, w / 2, w})
        if (std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
    }
    BatchEngine e(g, kind == GAS_BATCH_BFS ? BatchEngine::BFS_KIND
                                           : BatchEngine::SSSP_KIND);
    // per-sample times of one setting
    const int ie_default = kind == GAS_BATCH_BFS ? BFS_ITEM_EDGES : SSSP_ITEM_EDGES;
    auto time_one = [&](int32_t c, int32_t hub, int ie) {
      if (kind == GAS_BATCH_BFS)
        e.setPullDivisor(c);
      else
        e.setSsspDelta(c);
      e.setHubItems(hub);
      e.setItemEdges(ie);
      std::vector<double> t;
      for (const auto &src : samples) {
        const uint64_t all = src.size() == 64 ? ~0ull : ((1ull << src.size()) - 1);
        e.addLanes(lanes.data(), src.data(), (int)src.size());
        e.sync();
        const auto t0 = std::chrono::steady_clock::now();
        e.runToCompletion();
        e.sync();
        t.push_back(
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        e.releaseLanes(all);
      }
      return t;
    };
    auto total = [](const std::vector<double> &t) {
      double s = 0;
      for (double x : t) s += x;
      return s;
    };
    time_one(cand[0], HUB_ITEMS, ie_default);  // warm-up
    const std::vector<double> base = time_one(cand[0], HUB_ITEMS, ie_default);
    int32_t best = cand[0];
    double best_s = total(base);
    std::vector<double> best_t = base;
    for (size_t i = 1; i < cand.size(); ++i) {
      const std::vector<double> t = time_one(cand[i], HUB_ITEMS, ie_default);
      bool safe = true;
      for (int k = 0; k < kSamples; ++k) safe = safe && t[k] <= 1.1 * base[k];
      if (safe && total(t) < best_s) {
        best_s = total(t);
        best = cand[i];
        best_t = t;
      }
    }
    g->batch_param[kind] = best;
    // second stage: the hub threshold with that setting, same rule against
    // the default threshold's times
    int32_t best_hub = HUB_ITEMS;
    const std::vector<double> hub_base = best_t;
    for (int32_t h : {16, 64, 128}) {
      const std::vector<double> t = time_one(best, h, ie_default);
      bool safe = true;
      for (int k = 0; k < kSamples; ++k) safe = safe && t[k] <= 1.1 * hub_base[k];
      if (safe && total(t) < best_s) {
        best_s = total(t);
        best_hub = h;
        best_t = t;
      }
    }
    g->batch_hub[kind] = best_hub;
    // third stage: out-edges per push item (the tile a list vertex is cut
    // into: BFS 16 or 8, SSSP 64 or 32), same rule
    int32_t best_ie = ie_default;
    const std::vector<double> ie_base = best_t;
    const int ie_alt = kind == GAS_BATCH_BFS ? BFS_ITEM_EDGES_MIN : SSSP_ITEM_EDGES_MIN;
    {
      const std::vector<double> t = time_one(best, best_hub, ie_alt);
      bool safe = true;
      for (int k = 0; k < kSamples; ++k) safe = safe && t[k] <= 1.1 * ie_base[k];
      if (safe && total(t) < best_s) {
        best_s = total(t);
        best_ie = ie_alt;
      }
    }
    g->batch_items[kind] = best_ie;
    if (param_out) *param_out = best;
  });
}

int gas_batch_free(gas_batch_t b) {
  delete b;
  return GAS_OK;
}

}  // extern "C"
