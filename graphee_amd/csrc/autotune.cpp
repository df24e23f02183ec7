// Autotuner for the PageRank iteration: picks the source-strip width, the destination-bin size
// and the bin-task size of the strip layout (or the row gather, strip width 0), and for the row
// gather the degree-bin threshold (vec_min), the LDS tile (stream_cap rows' edges per wave), the
// hub chunk size and the workgroups per CU, on the loaded graph, by measuring iterations with
// HIP events; then the push/sweep threshold of CC and Bellman-Ford by timing Bellman-Ford, and the
// pull threshold of delta-stepping by timing delta = 64; the batched kernels' knobs have their own
// search (batch_autotune), on the caller's instances.  GraphMP fixes the matching
// knobs at compile time (CMPNUM threads, GraphMP/system/Global.h:58); nothing here changes
// results, only the task plan and the grid.
//
// Search: coordinate descent from the current options over discrete candidate sets, two
// sweeps, each trial = a few PageRank iterations timed on the engine stream; the best point
// is applied and returned.  The result can be saved as JSON and loaded at start-up.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gmp_internal.hpp"

namespace gmp {

namespace {

double time_gather(Engine &e, int iters) {
  gmp_pr_params p{0.85f, iters, 0.0, -1, 0, 0};
  gmp_stats st{};
  run_pagerank(e, p, nullptr, &st);  // warm (plan build, caches)
  run_pagerank(e, p, nullptr, &st);
  return st.launches ? st.ms_kernel / st.launches : 0.0;
}

// Bellman-Ford solve time from src (stored weights, or unit weights on an unweighted graph).
double time_bf(Engine &e, uint32_t src) {
  gmp_sssp_params p{GMP_SSSP_BF, 0, 0, 0};
  gmp_stats st{};
  run_sssp(e, src, p, nullptr, &st);  // warm (layouts)
  run_sssp(e, src, p, nullptr, &st);
  return st.ms_total;
}

// Fixed delta = 64 solve time from src.
double time_delta(Engine &e, uint32_t src) {
  gmp_sssp_params p{GMP_SSSP_DELTA, 64, 0, 0};
  gmp_stats st{};
  run_sssp(e, src, p, nullptr, &st);
  run_sssp(e, src, p, nullptr, &st);
  return st.ms_total;
}

uint32_t *field(Options &o, int k) {
  switch (k) {
    case 0: return &o.strip_log;
    case 1: return &o.bin_log;
    case 2: return &o.bin_task_log;
    case 3: return &o.strip_task_log;
    case 4: return &o.vec_min;
    case 5: return &o.stream_cap;
    case 6: return &o.chunk;
    case 7: return &o.wg_per_cu;
    default: return &o.row_cap;
  }
}
constexpr int kStripKnobs = 4;  // fields 0..3 shape the strip layout, the rest the row gather

// Makes o the engine's options, dropping the strip layout when its shape changes.
void apply(Engine &e, const Options &o) {
  if (o.strip_log != e.opt.strip_log || o.bin_log != e.opt.bin_log || o.bin_task_log != e.opt.bin_task_log ||
      o.strip_task_log != e.opt.strip_task_log)
    e.g.strips.valid = false;
  e.opt = o;
}

}  // namespace

void autotune(Engine &e, int32_t iters, gmp_tune_result *out) {
  if (iters <= 0) iters = 3;
  ensure_work(e);
  const std::vector<std::vector<uint32_t>> cand = {
      {0, 13, 14, 15},                  // strip_log (0: row gather)
      {12, 13, 14},                     // bin_log
      {18, 19, 20, 21},                 // bin_task_log
      {4, 5, 6, 7, 8},                  // strip_task_log
      {16, 32, 64, 128, 256, 512},      // vec_min
      {512, 1024, 2048, 4096},          // stream_cap
      {1024, 2048, 4096, 8192, 16384},  // chunk
      {2, 4, 8, 16, 32},                // wg_per_cu
      {64, 128, 256, 512, 1024},        // row_cap
  };
  Options best = e.opt;
  double best_ms = time_gather(e, iters);
  const double start_ms = best_ms;
  int trials = 1;
  for (int sweep = 0; sweep < 2; sweep++) {
    bool improved = false;
    for (size_t k = 0; k < cand.size(); k++) {
      for (uint32_t v : cand[k]) {
        Options o = best;
        if (*field(o, (int)k) == v) continue;
        // the strip knobs matter only with strips on, the gather knobs only without
        if ((int)k >= kStripKnobs ? best.strip_log != 0 : ((int)k > 0 && best.strip_log == 0)) continue;
        *field(o, (int)k) = v;
        if (o.vec_min > o.stream_cap) continue;
        apply(e, o);
        const double ms = time_gather(e, iters);
        trials++;
        if (ms > 0 && ms < best_ms * 0.995) {
          best_ms = ms;
          best = o;
          improved = true;
        }
      }
    }
    if (!improved) break;
  }
  apply(e, best);
  // the push/sweep switch of CC and Bellman-Ford (strip layout only): Bellman-Ford from the
  // vertex with the most out-edges for each threshold
  if (best.strip_log != 0 && e.g.V > 1) {
    std::vector<uint32_t> od(e.g.V);
    GMP_HIP(hipMemcpy(od.data(), e.g.out_deg.p, (size_t)e.g.V * 4, hipMemcpyDeviceToHost));
    const uint32_t src = (uint32_t)(std::max_element(od.begin(), od.end()) - od.begin());
    double best_bf = time_bf(e, src);
    trials++;
    for (uint32_t v : {0u, 6u, 8u, 10u, 12u, 14u}) {
      if (v == best.push_log) continue;
      e.opt.push_log = v;
      const double ms = time_bf(e, src);
      trials++;
      if (ms > 0 && ms < best_bf * 0.98) {
        best_bf = ms;
        best.push_log = v;
      }
    }
    e.opt.push_log = best.push_log;
    // the pull threshold of delta-stepping, timed on delta = 64 from the same vertex (weighted
    // graphs, or unit weights)
    double best_ds = time_delta(e, src);
    trials++;
    for (uint32_t v : {0u, 2u, 3u, 4u, 5u, 6u}) {
      if (v == best.delta_pull_log) continue;
      e.opt.delta_pull_log = v;
      const double ms = time_delta(e, src);
      trials++;
      if (ms > 0 && ms < best_ds * 0.98) {
        best_ds = ms;
        best.delta_pull_log = v;
      }
    }
    e.opt.delta_pull_log = best.delta_pull_log;
  }
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->strip_log = best.strip_log;
    out->bin_log = best.bin_log;
    out->bin_task_log = best.bin_task_log;
    out->stream_cap = best.stream_cap;
    out->row_cap = best.row_cap;
    out->chunk = best.chunk;
    out->vec_min = best.vec_min;
    out->wg_per_cu = best.wg_per_cu;
    out->trials = trials;
    out->gather_ms_before = start_ms;
    out->gather_ms_after = best_ms;
    out->push_log = best.push_log;
    out->delta_pull_log = best.delta_pull_log;
    out->strip_task_log = best.strip_task_log;
  }
}

// Batched kernels: coordinate descent over each kind's knobs, two sweeps, a trial = one batched
// SSSP run or a few batched PPR iterations on the caller's instances, timed with HIP events.
void batch_autotune(Engine &e, uint32_t ns, const uint32_t *sgraph, const uint32_t *source, uint32_t np,
                    const uint32_t *pgraph, const uint32_t *seed, float damping, int32_t iters,
                    gmp_batch_tune_result *out) {
  if (!e.batch.ready) fail(GMP_ESTATE, "no batch loaded");
  if (ns == 0 && np == 0) fail(GMP_EINVAL, "gmp_batch_autotune needs instances");
  const int ti = iters > 0 ? std::min(iters, 6) : 6;
  int trials = 0;
  auto run = [&](bool sssp) {
    gmp_stats st{};
    if (sssp) batch_sssp(e, ns, sgraph, source, nullptr, &st);
    else batch_ppr(e, np, pgraph, seed, damping, ti, nullptr, &st);
    trials++;
    return sssp ? st.ms_total : st.ms_total / ti;
  };
  // descend over the knobs of one kind; returns {first, best} time
  auto descend = [&](bool sssp, const std::vector<std::pair<uint32_t *, std::vector<uint32_t>>> &knobs) {
    run(sssp);  // warm (kernel code objects, caches)
    double best = run(sssp);
    const double first = best;
    for (int sweep = 0; sweep < 2; sweep++) {
      bool improved = false;
      for (const auto &k : knobs) {
        const uint32_t keep = *k.first;
        uint32_t chosen = keep;
        for (uint32_t v : k.second) {
          if (v == keep) continue;
          *k.first = v;
          const double ms = run(sssp);
          if (ms > 0 && ms < best * 0.98) {
            best = ms;
            chosen = v;
            improved = true;
          }
        }
        *k.first = chosen;
      }
      if (!improved) break;
    }
    return std::make_pair(first, best);
  };
  Options &o = e.opt;
  std::pair<double, double> ts{0.0, 0.0}, tp{0.0, 0.0};
  if (ns)
    ts = descend(true, {{&o.bsssp_scan_log, {0, 4, 6, 8, 10, 12}},
                        {&o.bsssp_hub_log, {0, 8, 10, 11, 12, 14}},
                        {&o.bsssp_unroll, {2, 4, 8}},
                        {&o.bsssp_grid, {1, 2, 4, 8, 16}}});
  if (np)
    tp = descend(false, {{&o.ppr_piece, {256, 512, 1024, 2048, 4096}},
                         {&o.ppr_run_rows, {8, 16, 32, 64, 128}},
                         {&o.ppr_run_edges, {512, 1024, 2048, 4096, 8192}},
                         {&o.ppr_loads, {4, 8, 16}},
                         {&o.ppr_grid, {2, 4, 8, 16, 32}}});
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->sssp_unroll = o.bsssp_unroll;
    out->sssp_grid_per_cu = o.bsssp_grid;
    out->sssp_scan_log = o.bsssp_scan_log;
    out->sssp_hub_log = o.bsssp_hub_log;
    out->ppr_piece = o.ppr_piece;
    out->ppr_run_rows = o.ppr_run_rows;
    out->ppr_run_edges = o.ppr_run_edges;
    out->ppr_loads = o.ppr_loads;
    out->ppr_grid_per_cu = o.ppr_grid;
    out->trials = (uint32_t)trials;
    out->sssp_ms_before = ts.first;
    out->sssp_ms_after = ts.second;
    out->ppr_ms_before = tp.first;
    out->ppr_ms_after = tp.second;
  }
}

void tune_save(Engine &e, const std::string &path) {
  std::ofstream f(path);
  if (!f) fail(GMP_EIO, path + ": cannot open for writing");
  f << "{\"reorder\": " << e.opt.reorder << ", \"stream_cap\": " << e.opt.stream_cap
    << ", \"row_cap\": " << e.opt.row_cap << ", \"chunk\": " << e.opt.chunk << ", \"vec_min\": " << e.opt.vec_min
    << ", \"wg_per_cu\": " << e.opt.wg_per_cu << ", \"strip_log\": " << e.opt.strip_log
    << ", \"bin_log\": " << e.opt.bin_log << ", \"bin_task_log\": " << e.opt.bin_task_log
    << ", \"strip_task_log\": " << e.opt.strip_task_log
    << ", \"push_log\": " << e.opt.push_log << ", \"delta_pull_log\": " << e.opt.delta_pull_log
    << ", \"batch_sssp_unroll\": " << e.opt.bsssp_unroll << ", \"batch_sssp_grid\": " << e.opt.bsssp_grid
    << ", \"batch_sssp_scan_log\": " << e.opt.bsssp_scan_log << ", \"batch_sssp_hub_log\": " << e.opt.bsssp_hub_log
    << ", \"batch_ppr_piece\": " << e.opt.ppr_piece << ", \"batch_ppr_run_rows\": " << e.opt.ppr_run_rows
    << ", \"batch_ppr_run_edges\": " << e.opt.ppr_run_edges << ", \"batch_ppr_loads\": " << e.opt.ppr_loads
    << ", \"batch_ppr_grid\": " << e.opt.ppr_grid << "}\n";
  if (!f) fail(GMP_EIO, path + ": short write");
}

// Minimal reader for the flat integer object tune_save writes.
void tune_load(Engine &e, const std::string &path) {
  std::ifstream f(path);
  if (!f) fail(GMP_EIO, path + ": cannot open");
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string s = ss.str();
  Options o = e.opt;
  auto get = [&](const char *key, uint32_t lo, uint32_t hi, uint32_t &dst) {
    const std::string k = std::string("\"") + key + "\"";
    size_t p = s.find(k);
    if (p == std::string::npos) return;
    p = s.find(':', p);
    if (p == std::string::npos) fail(GMP_EFORMAT, path + ": bad value for " + key);
    long v = std::strtol(s.c_str() + p + 1, nullptr, 10);
    if (v < (long)lo || v > (long)hi) fail(GMP_EFORMAT, path + ": " + key + " out of range");
    dst = (uint32_t)v;
  };
  uint32_t reorder = (uint32_t)o.reorder;
  get("reorder", 0, 1, reorder);
  get("stream_cap", 64, 4096, o.stream_cap);
  get("row_cap", 1, 4096, o.row_cap);
  get("chunk", 64, 1u << 20, o.chunk);
  get("vec_min", 1, 4096, o.vec_min);
  get("wg_per_cu", 1, 64, o.wg_per_cu);
  get("strip_log", 0, 15, o.strip_log);
  get("bin_log", 8, 14, o.bin_log);
  get("bin_task_log", 8, 22, o.bin_task_log);
  get("strip_task_log", 0, 12, o.strip_task_log);
  get("push_log", 0, 31, o.push_log);
  get("delta_pull_log", 0, 31, o.delta_pull_log);
  get("batch_sssp_unroll", 2, 8, o.bsssp_unroll);
  get("batch_sssp_grid", 1, 64, o.bsssp_grid);
  get("batch_sssp_scan_log", 0, 31, o.bsssp_scan_log);
  get("batch_sssp_hub_log", 0, 24, o.bsssp_hub_log);
  get("batch_ppr_piece", 64, 65536, o.ppr_piece);
  get("batch_ppr_run_rows", 1, 1024, o.ppr_run_rows);
  get("batch_ppr_run_edges", 64, 65536, o.ppr_run_edges);
  get("batch_ppr_loads", 4, 16, o.ppr_loads);
  get("batch_ppr_grid", 1, 64, o.ppr_grid);
  if (o.bsssp_unroll != 2 && o.bsssp_unroll != 4 && o.bsssp_unroll != 8)
    fail(GMP_EFORMAT, path + ": batch_sssp_unroll must be 2, 4 or 8");
  if (o.bsssp_hub_log != 0 && o.bsssp_hub_log < 6) fail(GMP_EFORMAT, path + ": batch_sssp_hub_log out of range");
  if (o.ppr_loads != 4 && o.ppr_loads != 8 && o.ppr_loads != 16)
    fail(GMP_EFORMAT, path + ": batch_ppr_loads must be 4, 8 or 16");
  if (o.strip_log != 0 && o.strip_log < 8) fail(GMP_EFORMAT, path + ": strip_log out of range");
  if ((int)reorder != e.opt.reorder) {
    e.g.work_ready = false;
    e.g.wout_ready = false;
    e.g.plan.built = false;
  }
  o.reorder = (int)reorder;
  apply(e, o);
}

}  // namespace gmp

extern "C" {

int gmp_autotune(gmp_engine *h, int32_t iters, gmp_tune_result *out) {
  GMP_API_BEGIN
  if (!h) gmp::fail(GMP_EINVAL, "null engine handle");
  auto *e = reinterpret_cast<gmp::Engine *>(h);
  if (!e->loaded) gmp::fail(GMP_ESTATE, "no graph loaded");
  if (e->g.semi) gmp::fail(GMP_EUNSUPPORTED, "gmp_autotune: the graph is semi-external");
  GMP_HIP(hipSetDevice(e->device));
  gmp::autotune(*e, iters, out);
  GMP_API_END
}

int gmp_tune_save(gmp_engine *h, const char *path) {
  GMP_API_BEGIN
  if (!h || !path) gmp::fail(GMP_EINVAL, "gmp_tune_save: null argument");
  gmp::tune_save(*reinterpret_cast<gmp::Engine *>(h), path);
  GMP_API_END
}

int gmp_tune_load(gmp_engine *h, const char *path) {
  GMP_API_BEGIN
  if (!h || !path) gmp::fail(GMP_EINVAL, "gmp_tune_load: null argument");
  gmp::tune_load(*reinterpret_cast<gmp::Engine *>(h), path);
  GMP_API_END
}

}  // extern "C"
