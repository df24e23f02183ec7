// C ABI of the engine: handle lifetime, graph loading, app entry points.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gmp_internal.hpp"

using gmp::Engine;
using gmp::fail;

namespace {

Engine *check(gmp_engine *h) {
  if (!h) fail(GMP_EINVAL, "null engine handle");
  return reinterpret_cast<Engine *>(h);
}

Engine *check_loaded(gmp_engine *h) {
  Engine *e = check(h);
  if (!e->loaded) fail(GMP_ESTATE, "no graph loaded");
  GMP_HIP(hipSetDevice(e->device));
  return e;
}

std::string join(const char *dir, const std::string &name) {
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d.push_back('/');
  return d + name;
}

}  // namespace

namespace gmp {

hipEvent_t *engine_events(Engine &e, size_t n) {
  while (e.ev_pool.size() < n) {
    hipEvent_t x = nullptr;
    GMP_HIP(hipEventCreate(&x));
    e.ev_pool.push_back(x);
  }
  return e.ev_pool.data();
}

double *engine_host_doubles(Engine &e, size_t n, double **dev) {
  if (e.h_hist_n < n) {
    if (e.h_hist) (void)hipHostFree(e.h_hist);
    e.h_hist = nullptr;
    e.d_hist = nullptr;
    e.h_hist_n = 0;
    void *p = nullptr;
    GMP_HIP(hipHostMalloc(&p, n * sizeof(double), hipHostMallocMapped));
    e.h_hist = static_cast<double *>(p);
    void *d = nullptr;
    GMP_HIP(hipHostGetDevicePointer(&d, p, 0));
    e.d_hist = static_cast<double *>(d);
    e.h_hist_n = n;
  }
  *dev = e.d_hist;
  return e.h_hist;
}

}  // namespace gmp

extern "C" {

int gmp_create(int32_t device, gmp_engine **out) {
  GMP_API_BEGIN
  if (!out) fail(GMP_EINVAL, "gmp_create: out is null");
  *out = nullptr;
  int n = 0;
  GMP_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) fail(GMP_EINVAL, "gmp_create: device ordinal out of range");
  GMP_HIP(hipSetDevice(device));
  auto *e = new Engine();
  try {
    e->device = device;
    int cus = 0;
    GMP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    e->num_cus = cus > 0 ? cus : 256;
    GMP_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    GMP_HIP(hipHostMalloc(&e->h_pinned, 256, hipHostMallocDefault));
    GMP_HIP(hipEventCreate(&e->na_ev[0]));
    GMP_HIP(hipEventCreate(&e->na_ev[1]));
    gmp::noalloc_attach(*e);  // the scratch array of gmp_pagerank_noalloc on this device
  } catch (...) {
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->h_pinned) (void)hipHostFree(e->h_pinned);
    for (hipEvent_t &x : e->na_ev)
      if (x) (void)hipEventDestroy(x);
    delete e;
    throw;
  }
  *out = reinterpret_cast<gmp_engine *>(e);
  GMP_API_END
}

int gmp_destroy(gmp_engine *h) {
  GMP_API_BEGIN
  if (!h) return GMP_OK;
  Engine *e = reinterpret_cast<Engine *>(h);
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  if (e->h_pinned) (void)hipHostFree(e->h_pinned);
  if (e->h_hist) (void)hipHostFree(e->h_hist);
  for (hipEvent_t x : e->ev_pool) (void)hipEventDestroy(x);
  for (hipEvent_t x : e->na_ev)
    if (x) (void)hipEventDestroy(x);
  hipStream_t s = e->stream;
  delete e;
  if (s) (void)hipStreamDestroy(s);
  GMP_API_END
}

int gmp_set_option(gmp_engine *h, int32_t key, int64_t value) {
  GMP_API_BEGIN
  Engine *e = check(h);
  auto in = [&](int64_t lo, int64_t hi, const char *name) {
    if (value < lo || value > hi)
      fail(GMP_EINVAL, std::string("gmp_set_option: ") + name + " must be in [" + std::to_string(lo) + ", " +
                           std::to_string(hi) + "]");
  };
  switch (key) {
    case GMP_OPT_REORDER:
      in(0, 1, "reorder");
      if ((int)value != e->opt.reorder) {
        e->opt.reorder = (int)value;
        e->g.work_ready = false;
        e->g.wout_ready = false;
        e->g.plan.built = false;
      }
      break;
    case GMP_OPT_STREAM_CAP:
      in(64, 4096, "stream_cap");
      e->opt.stream_cap = (uint32_t)value;
      break;
    case GMP_OPT_ROW_CAP:
      in(1, 4096, "row_cap");
      e->opt.row_cap = (uint32_t)value;
      break;
    case GMP_OPT_CHUNK:
      in(64, 1 << 20, "chunk");
      e->opt.chunk = (uint32_t)value;
      break;
    case GMP_OPT_WG_PER_CU:
      in(1, 64, "wg_per_cu");
      e->opt.wg_per_cu = (uint32_t)value;
      break;
    case GMP_OPT_VEC_MIN:
      in(1, 4096, "vec_min");
      e->opt.vec_min = (uint32_t)value;
      break;
    case GMP_OPT_STRIP_LOG:
      if (value != 0) in(8, 15, "strip_log");
      if ((uint32_t)value != e->opt.strip_log) e->g.strips.valid = false;
      e->opt.strip_log = (uint32_t)value;
      break;
    case GMP_OPT_BIN_LOG:
      in(8, 14, "bin_log");
      if ((uint32_t)value != e->opt.bin_log) e->g.strips.valid = false;
      e->opt.bin_log = (uint32_t)value;
      break;
    case GMP_OPT_BIN_TASK_LOG:
      in(8, 22, "bin_task_log");
      if ((uint32_t)value != e->opt.bin_task_log) e->g.strips.valid = false;
      e->opt.bin_task_log = (uint32_t)value;
      break;
    case GMP_OPT_SELECTIVE:
      in(0, 1, "selective");
      e->opt.selective = value != 0;
      break;
    case GMP_OPT_STRIP_TASK_LOG:
      in(0, 12, "strip_task_log");
      if ((uint32_t)value != e->opt.strip_task_log) e->g.strips.valid = false;
      e->opt.strip_task_log = (uint32_t)value;
      break;
    case GMP_OPT_EDGE_BUDGET_MB:
      in(0, 1ll << 30, "edge_budget_mb");
      e->opt.edge_budget_mb = (uint64_t)value;
      break;
    case GMP_OPT_WINDOW_MB:
      in(1, 4096, "window_mb");
      e->opt.window_mb = (uint32_t)value;
      break;
    case GMP_OPT_WINDOW_PACKED:
      in(0, 1, "window_packed");
      e->opt.window_packed = (int)value;
      break;
    case GMP_OPT_PUSH_LOG:
      in(0, 31, "push_log");
      e->opt.push_log = (uint32_t)value;
      break;
    case GMP_OPT_CC_UNION_FIND:
      in(0, 1, "cc_union_find");
      e->opt.cc_union_find = (int)value;
      break;
    case GMP_OPT_DELTA_PULL_LOG:
      in(0, 31, "delta_pull_log");
      e->opt.delta_pull_log = (uint32_t)value;
      break;
    case GMP_OPT_LEARNED_FINISH:
      in(0, 1, "learned_finish");
      e->opt.learned_finish = (int)value;
      break;
    case GMP_OPT_BATCH_SSSP_UNROLL:
      if (value != 2 && value != 4 && value != 8) fail(GMP_EINVAL, "gmp_set_option: batch_sssp_unroll must be 2, 4 or 8");
      e->opt.bsssp_unroll = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_SSSP_GRID:
      in(1, 64, "batch_sssp_grid");
      e->opt.bsssp_grid = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_SSSP_SCAN_LOG:
      in(0, 31, "batch_sssp_scan_log");
      e->opt.bsssp_scan_log = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_SSSP_HUB_LOG:
      if (value != 0) in(6, 24, "batch_sssp_hub_log");
      e->opt.bsssp_hub_log = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_PPR_PIECE:
      in(64, 65536, "batch_ppr_piece");
      e->opt.ppr_piece = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_PPR_RUN_ROWS:
      in(1, 1024, "batch_ppr_run_rows");
      e->opt.ppr_run_rows = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_PPR_RUN_EDGES:
      in(64, 65536, "batch_ppr_run_edges");
      e->opt.ppr_run_edges = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_PPR_LOADS:
      if (value != 4 && value != 8 && value != 16) fail(GMP_EINVAL, "gmp_set_option: batch_ppr_loads must be 4, 8 or 16");
      e->opt.ppr_loads = (uint32_t)value;
      break;
    case GMP_OPT_BATCH_PPR_GRID:
      in(1, 64, "batch_ppr_grid");
      e->opt.ppr_grid = (uint32_t)value;
      break;
    case GMP_OPT_MLP_NET:
      in(0, 1, "mlp_net");
      e->net_sel = (int)value;
      break;
    case GMP_OPT_WINDOW_PREFETCH:
      in(0, 1, "window_prefetch");
      e->opt.window_prefetch = (int)value;
      break;
    default:
      fail(GMP_EINVAL, "gmp_set_option: unknown key " + std::to_string(key));
  }
  GMP_API_END
}

int gmp_get_option(gmp_engine *h, int32_t key, int64_t *value) {
  GMP_API_BEGIN
  Engine *e = check(h);
  if (!value) fail(GMP_EINVAL, "gmp_get_option: value is null");
  switch (key) {
    case GMP_OPT_REORDER: *value = e->opt.reorder; break;
    case GMP_OPT_STREAM_CAP: *value = e->opt.stream_cap; break;
    case GMP_OPT_ROW_CAP: *value = e->opt.row_cap; break;
    case GMP_OPT_CHUNK: *value = e->opt.chunk; break;
    case GMP_OPT_WG_PER_CU: *value = e->opt.wg_per_cu; break;
    case GMP_OPT_VEC_MIN: *value = e->opt.vec_min; break;
    case GMP_OPT_STRIP_LOG: *value = e->opt.strip_log; break;
    case GMP_OPT_BIN_LOG: *value = e->opt.bin_log; break;
    case GMP_OPT_BIN_TASK_LOG: *value = e->opt.bin_task_log; break;
    case GMP_OPT_STRIP_TASK_LOG: *value = e->opt.strip_task_log; break;
    case GMP_OPT_SELECTIVE: *value = e->opt.selective ? 1 : 0; break;
    case GMP_OPT_EDGE_BUDGET_MB: *value = (int64_t)e->opt.edge_budget_mb; break;
    case GMP_OPT_WINDOW_MB: *value = e->opt.window_mb; break;
    case GMP_OPT_WINDOW_PACKED: *value = e->opt.window_packed; break;
    case GMP_OPT_PUSH_LOG: *value = e->opt.push_log; break;
    case GMP_OPT_CC_UNION_FIND: *value = e->opt.cc_union_find; break;
    case GMP_OPT_DELTA_PULL_LOG: *value = e->opt.delta_pull_log; break;
    case GMP_OPT_LEARNED_FINISH: *value = e->opt.learned_finish; break;
    case GMP_OPT_BATCH_SSSP_UNROLL: *value = e->opt.bsssp_unroll; break;
    case GMP_OPT_BATCH_SSSP_GRID: *value = e->opt.bsssp_grid; break;
    case GMP_OPT_BATCH_SSSP_SCAN_LOG: *value = e->opt.bsssp_scan_log; break;
    case GMP_OPT_BATCH_SSSP_HUB_LOG: *value = e->opt.bsssp_hub_log; break;
    case GMP_OPT_BATCH_PPR_PIECE: *value = e->opt.ppr_piece; break;
    case GMP_OPT_BATCH_PPR_RUN_ROWS: *value = e->opt.ppr_run_rows; break;
    case GMP_OPT_BATCH_PPR_RUN_EDGES: *value = e->opt.ppr_run_edges; break;
    case GMP_OPT_BATCH_PPR_LOADS: *value = e->opt.ppr_loads; break;
    case GMP_OPT_BATCH_PPR_GRID: *value = e->opt.ppr_grid; break;
    case GMP_OPT_MLP_NET: *value = e->net_sel; break;
    case GMP_OPT_WINDOW_PREFETCH: *value = e->opt.window_prefetch; break;
    default: fail(GMP_EINVAL, "gmp_get_option: unknown key " + std::to_string(key));
  }
  GMP_API_END
}

int gmp_load_csr(gmp_engine *h, uint32_t V, uint64_t E, const uint64_t *row_off,
                 const uint32_t *col_idx, const uint32_t *weights, uint32_t P,
                 const uint32_t *shard_bounds) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  if (V == 0 || V == 0xFFFFFFFFu) fail(GMP_EINVAL, "gmp_load_csr: V must be in [1, 2^32-2]");
  if (!row_off || (E && !col_idx)) fail(GMP_EINVAL, "gmp_load_csr: null pointer");
  if (row_off[0] != 0 || row_off[V] != E) fail(GMP_EINVAL, "gmp_load_csr: row_off must run from 0 to E");
  for (uint32_t r = 0; r < V; r++)
    if (row_off[r + 1] < row_off[r]) fail(GMP_EINVAL, "gmp_load_csr: row_off must be nondecreasing");
  for (uint64_t k = 0; k < E; k++)
    if (col_idx[k] >= V) fail(GMP_EINVAL, "gmp_load_csr: column id " + std::to_string(col_idx[k]) + " >= V");
  e->loaded = false;
  gmp::graph_upload(*e, V, E, row_off, col_idx, weights, false);
  gmp::graph_finish(*e, P, shard_bounds);
  e->loaded = true;
  GMP_API_END
}

int gmp_load_edge_list(gmp_engine *h, const char *path, uint32_t V, uint32_t P, uint32_t flags) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  if (!path) fail(GMP_EINVAL, "gmp_load_edge_list: null path");
  if (V == 0xFFFFFFFFu) fail(GMP_EINVAL, "gmp_load_edge_list: V must be below 2^32 - 1");
  gmp::EdgeCsr g;
  gmp::read_edge_list(path, flags, V, g);
  const uint32_t np = P ? P : 1;
  std::vector<uint32_t> bounds(np + 1);
  gmp::partition_bounds(g.raw_indeg.data(), V, np, bounds.data());
  e->loaded = false;
  gmp::graph_upload(*e, V, g.off[V], g.off.data(), g.idx.data(), g.w.empty() ? nullptr : g.w.data(), false);
  gmp::graph_finish(*e, np, bounds.data());
  e->g.symmetric = (flags & GMP_EDGES_SYMMETRIZE) != 0;  // every edge was added both ways
  e->loaded = true;
  GMP_API_END
}

int gmp_load_shards(gmp_engine *h, const char *dir, uint32_t V, uint32_t P) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  if (!dir) fail(GMP_EINVAL, "gmp_load_shards: dir is null");
  if (V == 0 || V == 0xFFFFFFFFu || P == 0) fail(GMP_EINVAL, "gmp_load_shards: V and P must be positive");
  std::vector<uint64_t> off;
  off.reserve((size_t)V + 1);
  off.push_back(0);
  std::vector<uint32_t> idx, wts, bounds(P + 1, 0), buf, wbuf;
  bool weighted = false;
  for (uint32_t p = 0; p < P; p++) {
    const std::string path = join(dir, std::to_string(p) + ".edge.npy");
    gmp::npy_read_u32(path, buf);
    // header [edge_num_raw, len(indices), len(indptr), start_id, end_id] (Comp.h:37-42)
    if (buf.size() < 5) fail(GMP_EFORMAT, path + ": shorter than the 5-word shard header");
    const uint64_t nnz = buf[1], nptr = buf[2], start = buf[3], end = buf[4];
    if ((int32_t)buf[1] < 0 || (int32_t)buf[2] < 1 || (int32_t)buf[3] < 0 || (int32_t)buf[4] < 0)
      fail(GMP_EFORMAT, path + ": negative field in shard header");
    if (buf.size() != 5 + nnz + nptr) fail(GMP_EFORMAT, path + ": length != 5 + len(indices) + len(indptr)");
    if (end < start || nptr != end - start + 1) fail(GMP_EFORMAT, path + ": len(indptr) != end - start + 1");
    if (start != off.size() - 1) fail(GMP_EFORMAT, path + ": shard does not start where the previous one ended");
    if (end > V) fail(GMP_EFORMAT, path + ": end_id beyond the vertex count");
    const uint32_t *ind = buf.data() + 5, *ptr = buf.data() + 5 + nnz;
    if (ptr[0] != 0 || ptr[nptr - 1] != nnz) fail(GMP_EFORMAT, path + ": indptr must run from 0 to len(indices)");
    for (uint64_t i = 0; i + 1 < nptr; i++)
      if (ptr[i + 1] < ptr[i]) fail(GMP_EFORMAT, path + ": indptr must be nondecreasing");
    for (uint64_t k = 0; k < nnz; k++)
      if (ind[k] >= V) fail(GMP_EFORMAT, path + ": source id " + std::to_string(ind[k]) + " >= V");
    const uint64_t base = idx.size();
    idx.insert(idx.end(), ind, ind + nnz);
    for (uint64_t i = 1; i < nptr; i++) off.push_back(base + ptr[i]);
    bounds[p] = (uint32_t)start;
    bounds[p + 1] = (uint32_t)end;
    const std::string wpath = join(dir, std::to_string(p) + ".weight.npy");
    if (FILE *f = std::fopen(wpath.c_str(), "rb")) {
      std::fclose(f);
      if (p > 0 && !weighted) fail(GMP_EFORMAT, wpath + ": weights present for some shards only");
      weighted = true;
      gmp::npy_read_u32(wpath, wbuf);
      if (wbuf.size() != nnz) fail(GMP_EFORMAT, wpath + ": weight count != len(indices)");
      wts.insert(wts.end(), wbuf.begin(), wbuf.end());
    } else if (weighted) {
      fail(GMP_EFORMAT, wpath + ": weights present for some shards only");
    }
  }
  if (off.size() != (size_t)V + 1) fail(GMP_EFORMAT, "shards do not cover all V vertices");
  e->loaded = false;
  gmp::graph_upload(*e, V, idx.size(), off.data(), idx.data(), weighted ? wts.data() : nullptr, false);
  gmp::graph_finish(*e, P, bounds.data());
  e->loaded = true;
  GMP_API_END
}

int gmp_generate_rmat(gmp_engine *h, const gmp_rmat_params *p) {
  GMP_API_BEGIN
  Engine *e = check(h);
  if (!p) fail(GMP_EINVAL, "gmp_generate_rmat: params are null");
  GMP_HIP(hipSetDevice(e->device));
  e->loaded = false;
  gmp::graph_generate_rmat(*e, *p);
  gmp::graph_finish(*e, p->num_shards, nullptr);
  gmp::Graph &g = e->g;
  if (gmp::semi_wanted(*e, g.E, g.weighted)) {
    // generated on the device, then moved out to host memory like a loaded semi-external graph
    std::vector<uint32_t> idx(g.E), w(g.weighted ? g.E : 0);
    if (g.E) GMP_HIP(hipMemcpy(idx.data(), g.in.idx.p, g.E * 4, hipMemcpyDeviceToHost));
    if (g.weighted && g.E) GMP_HIP(hipMemcpy(w.data(), g.in.w.p, g.E * 4, hipMemcpyDeviceToHost));
    g.in.idx.release();
    g.in.w.release();
    gmp::semi_setup(*e, idx.data(), g.weighted ? w.data() : nullptr);
  }
  e->loaded = true;
  GMP_API_END
}

int gmp_graph_info(gmp_engine *h, uint32_t *V, uint64_t *E, uint32_t *P, uint32_t *weighted) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  if (V) *V = e->g.V;
  if (E) *E = e->g.E;
  if (P) *P = e->g.P;
  if (weighted) *weighted = e->g.weighted ? 1u : 0u;
  GMP_API_END
}

int gmp_windows(gmp_engine *h, gmp_window_stats *out) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  if (!out) fail(GMP_EINVAL, "gmp_windows: out is null");
  gmp::semi_info(*e, out);
  GMP_API_END
}

int gmp_graph_download(gmp_engine *h, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights,
                       uint32_t *shard_bounds, uint32_t *out_degree) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  gmp::graph_download(*e, row_off, col_idx, weights, out_degree);
  if (shard_bounds) std::memcpy(shard_bounds, e->g.bounds.data(), e->g.bounds.size() * 4);
  GMP_API_END
}

int gmp_pick_source(gmp_engine *h, uint64_t seed, uint32_t salt, uint32_t *vertex) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  if (!vertex) fail(GMP_EINVAL, "gmp_pick_source: vertex is null");
  const uint32_t V = e->g.V;
  std::vector<uint32_t> od(V);
  GMP_HIP(hipMemcpy(od.data(), e->g.out_deg.p, (size_t)V * 4, hipMemcpyDeviceToHost));
  for (uint64_t t = 0; t < (1u << 20); t++) {
    const uint32_t v = (uint32_t)(gmp::mix64(seed + salt + (t + 1) * gmp::kGamma) % V);
    if (od[v] > 0) {
      *vertex = v;
      return GMP_OK;
    }
  }
  fail(GMP_EUNSUPPORTED, "gmp_pick_source: no vertex with out-degree > 0");
  GMP_API_END
}

int gmp_run_app(gmp_engine *h, int32_t app, int32_t max_iter, const void *msg_init,
                const int32_t *vertex_out, void *msg_out, void *value_out, gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  gmp::run_compat(*e, app, max_iter, msg_init, vertex_out, msg_out, value_out, st);
  GMP_API_END
}

int gmp_pagerank(gmp_engine *h, const gmp_pr_params *p, float *rank_out, gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  if (!p) fail(GMP_EINVAL, "gmp_pagerank: params are null");
  gmp::run_pagerank(*e, *p, rank_out, st);
  GMP_API_END
}

int gmp_sssp(gmp_engine *h, uint32_t source, const gmp_sssp_params *p, uint32_t *dist_out,
             gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  gmp_sssp_params d{GMP_SSSP_BF, 0, 0, 0};
  gmp::run_sssp(*e, source, p ? *p : d, dist_out, st);
  GMP_API_END
}

int gmp_sssp_trace(gmp_engine *h, uint32_t source, const gmp_sssp_params *p, gmp_trace_row *rows, uint32_t cap,
                   uint32_t *nrows, uint32_t *dist_out, gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  if (!nrows || (cap && !rows)) fail(GMP_EINVAL, "gmp_sssp_trace: null argument");
  if (e->g.semi) fail(GMP_EUNSUPPORTED, "gmp_sssp_trace: the graph is semi-external");
  gmp_sssp_params d{GMP_SSSP_LEARNED, 0, 0, 0};
  gmp::run_sssp_trace(*e, source, p ? *p : d, rows, cap, nrows, dist_out, st);
  GMP_API_END
}

int gmp_batch_generate(gmp_engine *h, uint32_t ngraphs, const gmp_rmat_params *params) {
  GMP_API_BEGIN
  Engine *e = check(h);
  if (!params) fail(GMP_EINVAL, "gmp_batch_generate: params are null");
  GMP_HIP(hipSetDevice(e->device));
  gmp::batch_generate(*e, ngraphs, params);
  GMP_API_END
}

int gmp_batch_info(gmp_engine *h, uint32_t graph, uint32_t *V, uint64_t *E) {
  GMP_API_BEGIN
  Engine *e = check(h);
  const gmp::BatchGraphs &b = e->batch;
  if (!b.ready) fail(GMP_ESTATE, "no batch loaded");
  if (graph + 1 >= b.vbase.size()) fail(GMP_EINVAL, "gmp_batch_info: graph out of range");
  if (V) *V = b.vbase[graph + 1] - b.vbase[graph];
  if (E) *E = b.ebase[graph + 1] - b.ebase[graph];
  GMP_API_END
}

int gmp_batch_download(gmp_engine *h, uint32_t graph, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  const gmp::BatchGraphs &b = e->batch;
  if (!b.ready) fail(GMP_ESTATE, "no batch loaded");
  if (graph + 1 >= b.vbase.size()) fail(GMP_EINVAL, "gmp_batch_download: graph out of range");
  if (weights && !b.weighted) fail(GMP_EUNSUPPORTED, "batch graphs have no weights");
  const uint32_t v0 = b.vbase[graph], V = b.vbase[graph + 1] - v0;
  const uint64_t e0 = b.ebase[graph], E = b.ebase[graph + 1] - e0;
  // the graph is stored relabelled (stored id = o2n[original id]); rows come back in the original
  // numbering with every row's sources ascending, the order the graph was built in
  const uint32_t *o2n = b.h_o2n.data() + v0;
  std::vector<uint32_t> n2o(V), sidx(E), sw(b.weighted ? E : 0);
  for (uint32_t v = 0; v < V; v++) n2o[o2n[v] - v0] = v;
  if (E) {
    GMP_HIP(hipMemcpy(sidx.data(), b.in.idx.as<uint32_t>() + e0, E * 4, hipMemcpyDeviceToHost));
    if (b.weighted) GMP_HIP(hipMemcpy(sw.data(), b.in.w.as<uint32_t>() + e0, E * 4, hipMemcpyDeviceToHost));
  }
  std::vector<std::pair<uint32_t, uint32_t>> row;
  uint64_t pos = 0;
  for (uint32_t r = 0; r < V; r++) {
    if (row_off) row_off[r] = pos;
    const uint32_t s = o2n[r] - v0;
    const uint64_t a = b.h_off[v0 + s] - e0, z = b.h_off[v0 + s + 1] - e0;
    row.clear();
    for (uint64_t k = a; k < z; k++) row.emplace_back(n2o[sidx[k] - v0], b.weighted ? sw[k] : 0u);
    std::stable_sort(row.begin(), row.end(),
                     [](const std::pair<uint32_t, uint32_t> &x, const std::pair<uint32_t, uint32_t> &y) {
                       return x.first < y.first;
                     });
    for (const auto &q : row) {
      if (col_idx) col_idx[pos] = q.first;
      if (weights) weights[pos] = q.second;
      pos++;
    }
  }
  if (row_off) row_off[V] = pos;
  GMP_API_END
}

int gmp_batch_sssp(gmp_engine *h, uint32_t n, const uint32_t *graph, const uint32_t *source, uint32_t *dist_out,
                   gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  gmp::batch_sssp(*e, n, graph, source, dist_out, st);
  GMP_API_END
}

int gmp_batch_ppr(gmp_engine *h, uint32_t n, const uint32_t *graph, const uint32_t *seed, float damping,
                  int32_t iters, float *rank_out, gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  gmp::batch_ppr(*e, n, graph, seed, damping, iters, rank_out, st);
  GMP_API_END
}

int gmp_window_trace(gmp_engine *h, int32_t kind, uint32_t source, gmp_trace_row *rows, uint32_t cap,
                     uint32_t *nrows, gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  if (!rows || !nrows) fail(GMP_EINVAL, "gmp_window_trace: null output");
  if (!e->g.semi) fail(GMP_EUNSUPPORTED, "gmp_window_trace: the graph is resident, not semi-external");
  GMP_HIP(hipSetDevice(e->device));
  gmp::semi_trace(*e, kind, source, rows, cap, nrows, st);
  GMP_API_END
}

int gmp_batch_run(gmp_engine *h, uint32_t ns, const uint32_t *sssp_graph, const uint32_t *source, uint32_t np,
                  const uint32_t *ppr_graph, const uint32_t *seed, float damping, int32_t iters,
                  uint32_t *dist_out, float *rank_out, gmp_stats *st_sssp, gmp_stats *st_ppr) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  gmp::batch_run(*e, ns, sssp_graph, source, np, ppr_graph, seed, damping, iters, dist_out, rank_out, st_sssp,
                 st_ppr, nullptr);
  GMP_API_END
}

int gmp_batch_autotune(gmp_engine *h, uint32_t ns, const uint32_t *sssp_graph, const uint32_t *source,
                       uint32_t np, const uint32_t *ppr_graph, const uint32_t *seed, float damping, int32_t iters,
                       gmp_batch_tune_result *out) {
  GMP_API_BEGIN
  Engine *e = check(h);
  GMP_HIP(hipSetDevice(e->device));
  gmp::batch_autotune(*e, ns, sssp_graph, source, np, ppr_graph, seed, damping, iters, out);
  GMP_API_END
}

int gmp_cc(gmp_engine *h, uint32_t *label_out, gmp_stats *st) {
  GMP_API_BEGIN
  Engine *e = check_loaded(h);
  gmp::run_cc(*e, label_out, st);
  GMP_API_END
}

}  // extern "C"
