// PageRank on gfx950 and the GraphMP-compatible synchronous apps.
//
// Standard PageRank (gmp_pagerank): per iteration
//   r'[v] = base + d * sum_{u->v} c[u]   (+ extra at the seed for personalised PageRank)
//   c'[v] = r'[v] / outdeg[v]  (0 for dangling v)
// with base = ((1-d) + d*D)/N, D = dangling mass of r.  Ranks and contributions are FP32;
// the per-vertex gather is Kahan-compensated; the dangling mass and the L1 change are
// FP64 reductions with a fixed order (per-workgroup partials, one finalising workgroup).
// One iteration is three launches: the fused gather/apply over the wave task plan, the
// combine of multi-chunk hub rows, and the FP64 finalise that prepares the next base.
//
// Compat apps (gmp_run_app) reproduce GraphMP's per-vertex rules bit for bit:
//   pagerank.cpp:42-48  rel = sum in source order of Msg[u]/Out[u] in FP32, then
//                       (float)(rel*0.85 + 1.0/N) in FP64 (no FMA contraction);
//   sssp.cpp:40-46      min(Msg[v], Msg[u]+1) in int32;
//   cc.cpp:40-45        min(Msg[v], Msg[u]) in int32;
// under the synchronous loop of GraphPS.h:256-386 (changed count against VertexValue,
// stop when nothing changed and step > 1).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "gmp_internal.hpp"
#include "pr_common.hpp"
#include "varint.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kFinalBlock = 256;

template <class OffT>
struct PrArgs {
  const OffT *off;
  const uint32_t *idx;
  const uint32_t *outdeg;
  const float *contrib_in;
  const float *rank_in;
  float *rank_out;
  float *contrib_out;
  const uint32_t *tasks;
  uint32_t ntasks;
  const HubRow *hubs;
  const uint32_t *multi;
  uint32_t nmulti;
  float2 *hub_part;
  const PrScal *sc;
  float d;
  uint32_t seed;
  uint32_t stream_cap;
  uint32_t chunk;
  double *part_resid;
  double *part_dang;
  Packed pk;  // compressed column ids (COMP kernels)
  uint64_t E;
};

template <bool KAHAN>
__device__ inline void acc_add(float &s, float &c, float x) {
  if (KAHAN) kahan_add(s, c, x);
  else s += x;
}

// Gathers contributions of edges [e0, e1) in 64-edge groups aligned to the global edge index
// (lane k takes the edges with e % 64 == k), so the plain and the compressed column streams
// assign every edge to the same lane: the two paths give bit-identical sums.  STREAM tasks
// store each value at buf[e - e0]; CHUNK tasks accumulate per lane.
template <class OffT, bool KAHAN, bool COMP, bool TO_LDS>
__device__ inline void gather_range(const PrArgs<OffT> &a, uint64_t e0, uint64_t e1, float *buf, uint32_t *scratch,
                                    float &s, float &c) {
  const uint32_t lane = lane_id();
  if (e1 <= e0) return;
  if (COMP) {
    // two-stage pipeline: group g's gathers and group g+2's bytes are in flight while group g+1
    // decodes; lanes outside [e0, e1) load a harmless contrib_in[0]
    const uint64_t gend = (e1 + kVarintGroup - 1) / kVarintGroup;
    uint64_t g = e0 / kVarintGroup;
    const VarintGroupWords w0 = varint_fetch(a.pk, g);
    VarintGroupWords w1 = g + 1 < gend ? varint_fetch(a.pk, g + 1) : w0;
    uint64_t e = g * kVarintGroup + lane;
    bool ok = e >= e0 && e < e1;
    uint32_t src = varint_decode_words(w0, a.pk, g, a.E, scratch);
    float x = a.contrib_in[ok ? src : 0u];
    for (; g < gend; g++) {
      const VarintGroupWords w2 = g + 2 < gend ? varint_fetch(a.pk, g + 2) : w1;
      const uint64_t en = e + kVarintGroup;
      const bool okn = en >= e0 && en < e1;
      float xn = 0.0f;
      if (g + 1 < gend) {
        const uint32_t sn = varint_decode_words(w1, a.pk, g + 1, a.E, scratch);
        xn = a.contrib_in[okn ? sn : 0u];
      }
      if (ok) {
        if (TO_LDS) buf[e - e0] = x;
        else acc_add<KAHAN>(s, c, x);
      }
      x = xn;
      ok = okn;
      e = en;
      w1 = w2;
    }
    return;
  }
  uint64_t j = (e0 & ~(uint64_t)(kWave - 1)) + lane;
  // four independent id loads, then four gathers per lane and round (memory-level parallelism)
  for (; j + 3 * kWave < e1; j += 4 * kWave) {
    const bool ok0 = j >= e0, ok1 = j + kWave >= e0;
    const uint32_t i0 = ok0 ? __builtin_nontemporal_load(&a.idx[j]) : 0u;
    const uint32_t i1 = ok1 ? __builtin_nontemporal_load(&a.idx[j + kWave]) : 0u;
    const uint32_t i2 = __builtin_nontemporal_load(&a.idx[j + 2 * kWave]);
    const uint32_t i3 = __builtin_nontemporal_load(&a.idx[j + 3 * kWave]);
    const float c0 = ok0 ? a.contrib_in[i0] : 0.0f, c1 = ok1 ? a.contrib_in[i1] : 0.0f;
    const float c2 = a.contrib_in[i2], c3 = a.contrib_in[i3];
    if (TO_LDS) {
      if (ok0) buf[j - e0] = c0;
      if (ok1) buf[j + kWave - e0] = c1;
      buf[j + 2 * kWave - e0] = c2;
      buf[j + 3 * kWave - e0] = c3;
    } else {
      if (ok0) acc_add<KAHAN>(s, c, c0);
      if (ok1) acc_add<KAHAN>(s, c, c1);
      acc_add<KAHAN>(s, c, c2);
      acc_add<KAHAN>(s, c, c3);
    }
  }
  for (; j < e1; j += kWave) {
    if (j < e0) continue;
    const float x = a.contrib_in[__builtin_nontemporal_load(&a.idx[j])];
    if (TO_LDS) buf[j - e0] = x;
    else acc_add<KAHAN>(s, c, x);
  }
}

// Rows are destinations; every row is applied except multi-chunk hubs (pr_hub_combine_kernel).
template <class OffT, bool KAHAN, bool COMP>
__global__ __launch_bounds__(kBlock) void pr_gather_kernel(PrArgs<OffT> a) {
  extern __shared__ float lds[];
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t per_wave = a.stream_cap + (COMP ? kVarintScratchWords : 0u);
  float *buf = lds + wv * per_wave;
  uint32_t *scratch = reinterpret_cast<uint32_t *>(buf + a.stream_cap);
  const float base = a.sc->base, extra = a.sc->extra;
  PrAcc acc;
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t t = blockIdx.x * kWavesPerBlock + wv; t < a.ntasks; t += nw) {
    const uint32_t x = a.tasks[2 * t], y = a.tasks[2 * t + 1];
    if (!(y & kTaskChunkBit)) {
      // STREAM: stage the gathered contributions of rows [x, x+y) in LDS, one row per lane.
      const uint64_t e0 = a.off[x], e1 = a.off[x + y];
      float s0 = 0, c0 = 0;
      gather_range<OffT, KAHAN, COMP, true>(a, e0, e1, buf, scratch, s0, c0);
      wave_sync();
      for (uint32_t i = lane; i < y; i += kWave) {
        const uint32_t r = x + i;
        const uint32_t b = (uint32_t)((uint64_t)a.off[r] - e0), en = (uint32_t)((uint64_t)a.off[r + 1] - e0);
        float s = 0, c = 0;
        for (uint32_t k = b; k < en; k++) acc_add<KAHAN>(s, c, buf[k]);
        pr_apply(a, r, s - c, base, extra, acc);
      }
      wave_sync();
    } else {
      // CHUNK of a hub row: per-lane compensated sums, then a wave reduction.
      const HubRow h = a.hubs[x];
      const uint32_t k = y & ~kTaskChunkBit;
      const uint64_t rb = a.off[h.row], re = a.off[h.row + 1];
      const uint64_t e0 = rb + (uint64_t)k * a.chunk;
      const uint64_t e1 = min(e0 + a.chunk, re);
      float s = 0, c = 0;
      gather_range<OffT, KAHAN, COMP, false>(a, e0, e1, buf, scratch, s, c);
      float hi = s, lo = -c;
      wave_reduce_dd(hi, lo);
      if (lane == 0) {
        if (h.nchunks == 1) pr_apply(a, h.row, hi, base, extra, acc);
        else a.hub_part[h.part_base + k] = make_float2(hi, lo);
      }
    }
  }
  block_store_acc<kWavesPerBlock>(acc, a.part_resid, a.part_dang);
}

// One wave per multi-chunk hub: combine its chunk partials in chunk order, then apply.
template <class OffT>
__global__ __launch_bounds__(kBlock) void pr_hub_combine_kernel(PrArgs<OffT> a) {
  const uint32_t lane = lane_id(), wv = wave_id();
  const float base = a.sc->base, extra = a.sc->extra;
  PrAcc acc;
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t t = blockIdx.x * kWavesPerBlock + wv; t < a.nmulti; t += nw) {
    const HubRow h = a.hubs[a.multi[t]];
    float hi = 0, lo = 0;
    for (uint32_t k = lane; k < h.nchunks; k += kWave) {
      const float2 p = a.hub_part[h.part_base + k];
      dd_add(hi, lo, p.x, p.y);
    }
    wave_reduce_dd(hi, lo);
    if (lane == 0) pr_apply(a, h.row, hi, base, extra, acc);
  }
  block_store_acc<kWavesPerBlock>(acc, a.part_resid, a.part_dang);
}

// r = teleport vector, c = r / outdeg, dangling partials.
__global__ __launch_bounds__(kBlock) void pr_init_kernel(uint32_t V, uint32_t seed, const uint32_t *outdeg,
                                                         float *rank, float *contrib, double *part_resid,
                                                         double *part_dang) {
  PrAcc acc;
  const float r0 = seed == 0xFFFFFFFFu ? (float)(1.0 / (double)V) : 0.0f;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const float r = (seed != 0xFFFFFFFFu && v == seed) ? 1.0f : r0;
    rank[v] = r;
    const uint32_t od = outdeg[v];
    contrib[v] = od ? r / (float)od : 0.0f;
    if (!od) acc.dang += (double)r;
  }
  block_store_acc<kWavesPerBlock>(acc, part_resid, part_dang);
}

// Sums partials in a fixed order (n1 from the gather, n2 from the hub combine), records
// the iteration's residual and dangling mass, and prepares the next base.
// With tol > 0 it also raises sc->done once the residual is below tol (it < 0, the set-up call,
// clears it), after which the iteration launches queued behind it do nothing.
__global__ __launch_bounds__(kFinalBlock) void pr_finalize_kernel(const double *pr1, const double *pd1, uint32_t n1,
                                                                  const double *pr2, const double *pd2, uint32_t n2,
                                                                  uint32_t V, float d, int personalised,
                                                                  PrScal *sc, double *hist, int32_t it, double tol) {
  __shared__ double sr[kFinalBlock], sd[kFinalBlock];
  if (it >= 0 && sc->done) return;
  double r = 0, dd = 0;
  for (uint32_t i = threadIdx.x; i < n1; i += kFinalBlock) {
    r += pr1[i];
    dd += pd1[i];
  }
  for (uint32_t i = threadIdx.x; i < n2; i += kFinalBlock) {
    r += pr2[i];
    dd += pd2[i];
  }
  sr[threadIdx.x] = r;
  sd[threadIdx.x] = dd;
  __syncthreads();
  for (int s = kFinalBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      sd[threadIdx.x] += sd[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double D = sd[0];
    const double tele = (1.0 - (double)d) + (double)d * D;
    sc->resid = sr[0];
    sc->dang = D;
    if (personalised) {
      sc->base = 0.0f;
      sc->extra = (float)tele;
    } else {
      sc->base = (float)(tele / (double)V);
      sc->extra = 0.0f;
    }
    if (hist && it >= 0) {
      hist[2 * it] = sr[0];
      hist[2 * it + 1] = D;
    }
    sc->done = (it >= 0 && tol > 0.0 && sr[0] < tol) ? 1 : 0;
  }
}

// ---- GraphMP-compatible synchronous step (one thread per destination) ----
// Rows [v0, v1); with stamp set, the block of every row whose message changes is stamped with
// `sweep` (window activation of semi-external graphs).
template <class OffT, int APP>
__global__ __launch_bounds__(kBlock) void compat_step_kernel(const OffT *off, const uint32_t *idx,
                                                             const uint32_t *w, const uint8_t *w8,
                                                             const int32_t *vout, uint32_t v0,
                                                             uint32_t v1, double invN, const void *msg_v,
                                                             void *msgnew_v, void *value_v,
                                                             unsigned long long *changed, uint32_t *stamp,
                                                             uint32_t blk_log, uint32_t sweep) {
  for (uint32_t v = v0 + blockIdx.x * blockDim.x + threadIdx.x; v < v1; v += gridDim.x * blockDim.x) {
    const uint64_t k0 = off[v], k1 = off[v + 1];
    bool ch;
    if (APP == GMP_APP_PAGERANK) {
      const float *msg = (const float *)msg_v;
      float rel = 0.0f;
      for (uint64_t k = k0; k < k1; k++) {
        const uint32_t u = idx[k];
        const int32_t o = vout[u];
        rel += msg[u] / (float)(o == 0 ? 1 : o);
      }
      rel = (float)(__dadd_rn(__dmul_rn((double)rel, 0.85), invN));
      ((float *)msgnew_v)[v] = rel;
      float *val = (float *)value_v;
      ch = rel != val[v];
      if (ch) val[v] = rel;
    } else if (APP == GMP_APP_SSSP || APP == GMP_APP_CC) {
      const int32_t *msg = (const int32_t *)msg_v;
      int32_t best = msg[v];
      for (uint64_t k = k0; k < k1; k++) {
        int32_t t = msg[idx[k]];
        if (APP == GMP_APP_SSSP) t = (int32_t)((uint32_t)t + 1u);
        if (best > t) best = t;
      }
      ((int32_t *)msgnew_v)[v] = best;
      if (stamp && best != msg[v]) stamp[v >> blk_log] = sweep;
      int32_t *val = (int32_t *)value_v;
      ch = best != val[v];
      if (ch) val[v] = best;
    } else {
      const uint32_t *msg = (const uint32_t *)msg_v;
      uint32_t best = msg[v];
      for (uint64_t k = k0; k < k1; k++) {
        const uint64_t t = (uint64_t)msg[idx[k]] + (uint64_t)(w8 ? (uint32_t)w8[k] : w[k]);
        if (t < (uint64_t)best) best = (uint32_t)t;
      }
      ((uint32_t *)msgnew_v)[v] = best;
      if (stamp && best != msg[v]) stamp[v >> blk_log] = sweep;
      uint32_t *val = (uint32_t *)value_v;
      ch = best != val[v];
      if (ch) val[v] = best;
    }
    const uint64_t m = __ballot(ch);
    if (lane_id() == 0 && m) atomicAdd(changed, (unsigned long long)__popcll(m));
  }
}

struct CompatRows {
  const uint32_t *idx, *w;  // ids and weights by global edge
  const uint8_t *w8;        // the weights as bytes (semi-external windows whose weights all fit one)
  uint32_t v0, v1;
  uint32_t *stamp;
  uint32_t blk_log, sweep;
};

template <class OffT>
void launch_compat(Engine &e, int app, uint32_t grid, double invN, const void *msg, void *msgnew,
                   void *value, const int32_t *vout, unsigned long long *changed, const CompatRows &r) {
  const OffT *off = e.g.in.off.as<OffT>();
#define GMP_COMPAT(APP)                                                                                      \
  compat_step_kernel<OffT, APP><<<grid, kBlock, 0, e.stream>>>(off, r.idx, r.w, r.w8, vout, r.v0, r.v1, invN, msg, \
                                                               msgnew, value, changed, r.stamp, r.blk_log,  \
                                                               r.sweep)
  switch (app) {
    case GMP_APP_PAGERANK: GMP_COMPAT(GMP_APP_PAGERANK); break;
    case GMP_APP_SSSP: GMP_COMPAT(GMP_APP_SSSP); break;
    case GMP_APP_CC: GMP_COMPAT(GMP_APP_CC); break;
    default: GMP_COMPAT(GMP_APP_SSSP_W); break;
  }
#undef GMP_COMPAT
  GMP_CHECK_LAUNCH();
}

// One synchronous step of a compat app over the windows of a semi-external graph.  The min
// apps skip a window none of whose source blocks changed in the step before (the stamps of
// step - 1): its rows would recompute their current messages, which are copied instead.
template <class OffT>
void compat_windows_step(Engine &e, int app, int step, uint32_t grid_cap, double invN, const void *msg, void *msgnew,
                         void *value, const int32_t *vout, unsigned long long *changed) {
  SemiExternal &s = *e.g.semi;
  const bool minapp = app != GMP_APP_PAGERANK;
  const uint8_t *act = nullptr;
  if (minapp && step > 0 && e.opt.selective) act = window_activity(e, (uint32_t)step);  // the stop rule is the changed count
  // messages are double-buffered, so the order of the windows changes nothing; window_order
  // spreads the streamed windows between the cached ones, keeping the copy engines busy
  for (const uint32_t w : window_order(s)) {
    const EdgeWindow &x = s.win[w];
    if (act && !act[w]) {
      GMP_HIP(hipMemcpyAsync(static_cast<uint32_t *>(msgnew) + x.r0, static_cast<const uint32_t *>(msg) + x.r0,
                             (size_t)(x.r1 - x.r0) * 4, hipMemcpyDeviceToDevice, e.stream));
      s.windows_skipped++;
      continue;
    }
    const WindowDev d = window_begin(e, w, app == GMP_APP_SSSP_W, true);
    CompatRows r{d.idx, d.w, d.w8, x.r0, x.r1, minapp ? s.stamp.as<uint32_t>() : nullptr, s.blk_log,
                 (uint32_t)step + 1};
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((x.r1 - x.r0 + kBlock - 1) / kBlock, grid_cap));
    launch_compat<OffT>(e, app, grid, invN, msg, msgnew, value, vout, changed, r);
    window_end(e, d);
  }
}

// PageRank over the windows of a semi-external graph: per iteration every window's gather
// plan runs on its rows (row ids relative to the window, the row arrays shifted to its first
// row, ids addressed by global edge), fed from the device cache or a ring slot; packed windows
// are decoded inside the gather.  The gathers give the same per-row sums as the resident row
// gather (same tasks per row, same lane per edge); the FP64 partials are summed over all windows
// in a fixed order.
template <class OffT>
void pagerank_windows(Engine &e, const gmp_pr_params &p, float *rank_out, gmp_stats *st) {
  Graph &g = e.g;
  SemiExternal &s = *g.semi;
  window_plans(e);
  const uint32_t V = g.V, nwin = (uint32_t)s.win.size();
  const bool personalised = p.seed_vertex >= 0;
  const uint32_t seed = personalised ? (uint32_t)p.seed_vertex : 0xFFFFFFFFu;
  const uint32_t G = std::max<uint32_t>(1, (uint32_t)e.num_cus * e.opt.wg_per_cu);
  const int max_iter = std::max(0, p.max_iter);
  const bool comp = s.packed, kahan = !(p.flags & GMP_PR_NO_KAHAN);
  std::vector<uint32_t> g1(nwin), g2(nwin);
  uint32_t n1 = 0, n2 = 0;
  for (uint32_t w = 0; w < nwin; w++) {
    const EdgeWindow &x = s.win[w];
    g1[w] = x.ntasks ? std::min<uint32_t>(G, (x.ntasks + kWavesPerBlock - 1) / kWavesPerBlock) : 0u;
    g2[w] = x.nmulti ? std::min<uint32_t>(G, (x.nmulti + kWavesPerBlock - 1) / kWavesPerBlock) : 0u;
    n1 += g1[w];
    n2 += g2[w];
  }
  // windows in an order that spreads the streamed ones evenly among the cached ones (a cached
  // one first), so the copy engines keep filling the ring while cached windows are gathered;
  // every row's sum is the same in any order
  const std::vector<uint32_t> order = window_order(s);
  const uint32_t np1 = std::max(n1, G), np2 = std::max<uint32_t>(n2, 1);
  e.s_a.alloc((size_t)V * 4);
  e.s_b.alloc((size_t)V * 4);
  e.s_c.alloc((size_t)V * 4);
  e.s_d.alloc((size_t)V * 4);
  e.s_e.alloc(std::max<size_t>(s.max_parts, 1) * sizeof(float2));
  e.s_part.alloc(((size_t)np1 + np2) * 2 * sizeof(double));
  e.s_scal.alloc(sizeof(PrScal) + (size_t)(max_iter + 1) * 2 * sizeof(double));
  float *rank[2] = {e.s_a.as<float>(), e.s_b.as<float>()};
  float *contrib[2] = {e.s_c.as<float>(), e.s_d.as<float>()};
  double *pr1 = e.s_part.as<double>(), *pd1 = pr1 + np1;
  double *pr2 = pd1 + np1, *pd2 = pr2 + np2;
  PrScal *sc = e.s_scal.as<PrScal>();
  double *hist = reinterpret_cast<double *>(e.s_scal.as<char>() + sizeof(PrScal));
  const uint32_t *outdeg = g.out_deg.as<uint32_t>();
  const OffT *off = g.in.off.as<OffT>();
  const uint32_t *tasks = s.tasks.as<uint32_t>(), *multi = s.multi.as<uint32_t>();
  const HubRow *hubs = s.hubs.as<HubRow>();

  std::vector<hipEvent_t> ev(2 * (size_t)max_iter + 2);
  for (auto &x : ev) GMP_HIP(hipEventCreate(&x));
  struct EvGuard {
    std::vector<hipEvent_t> &v;
    ~EvGuard() {
      for (auto x : v) (void)hipEventDestroy(x);
    }
  } evg{ev};

  GMP_HIP(hipEventRecord(ev[0], e.stream));
  pr_init_kernel<<<G, kBlock, 0, e.stream>>>(V, seed, outdeg, rank[0], contrib[0], pr1, pd1);
  GMP_CHECK_LAUNCH();
  pr_finalize_kernel<<<1, kFinalBlock, 0, e.stream>>>(pr1, pd1, G, pr2, pd2, 0, V, p.damping, personalised, sc,
                                                      nullptr, -1, 0.0);
  GMP_CHECK_LAUNCH();
  const size_t lds = (size_t)kWavesPerBlock * (g.plan.stream_cap + (comp ? kVarintScratchWords : 0u)) * sizeof(float);
  int cur = 0, it = 0;
  double resid = 0, dang = 0;
  for (it = 0; it < max_iter; it++) {
    GMP_HIP(hipEventRecord(ev[2 + 2 * it], e.stream));
    uint32_t o1 = 0, o2 = 0;
    for (uint32_t wi = 0; wi < nwin; wi++) {
      const uint32_t w = order[wi];
      const EdgeWindow &x = s.win[w];
      const WindowDev d = window_begin(e, w, false, false);
      PrArgs<OffT> a{};
      a.off = off + x.r0;
      a.idx = d.idx;
      a.pk = d.pk;
      a.E = g.E;
      a.outdeg = outdeg + x.r0;
      a.contrib_in = contrib[cur];
      a.rank_in = rank[cur] + x.r0;
      a.rank_out = rank[cur ^ 1] + x.r0;
      a.contrib_out = contrib[cur ^ 1] + x.r0;
      a.tasks = tasks + 2ull * x.task0;
      a.ntasks = x.ntasks;
      a.hubs = hubs + x.hub0;
      a.multi = multi + x.multi0;
      a.nmulti = x.nmulti;
      a.hub_part = e.s_e.as<float2>();
      a.sc = sc;
      a.d = p.damping;
      a.seed = seed - x.r0;  // matches no row of the window when the seed lies outside it
      a.stream_cap = g.plan.stream_cap;
      a.chunk = g.plan.chunk;
      if (g1[w]) {
        a.part_resid = pr1 + o1;
        a.part_dang = pd1 + o1;
        if (comp) {
          if (kahan) pr_gather_kernel<OffT, true, true><<<g1[w], kBlock, lds, e.stream>>>(a);
          else pr_gather_kernel<OffT, false, true><<<g1[w], kBlock, lds, e.stream>>>(a);
        } else {
          if (kahan) pr_gather_kernel<OffT, true, false><<<g1[w], kBlock, lds, e.stream>>>(a);
          else pr_gather_kernel<OffT, false, false><<<g1[w], kBlock, lds, e.stream>>>(a);
        }
        GMP_CHECK_LAUNCH();
        o1 += g1[w];
      }
      if (g2[w]) {
        a.part_resid = pr2 + o2;
        a.part_dang = pd2 + o2;
        pr_hub_combine_kernel<OffT><<<g2[w], kBlock, 0, e.stream>>>(a);
        GMP_CHECK_LAUNCH();
        o2 += g2[w];
      }
      window_end(e, d);
    }
    GMP_HIP(hipEventRecord(ev[3 + 2 * it], e.stream));
    pr_finalize_kernel<<<1, kFinalBlock, 0, e.stream>>>(pr1, pd1, n1, pr2, pd2, n2, V, p.damping, personalised, sc,
                                                        hist, it, 0.0);
    GMP_CHECK_LAUNCH();
    cur ^= 1;
    if (p.tol > 0) {
      double h[2];
      GMP_HIP(hipMemcpyAsync(h, hist + 2 * it, sizeof h, hipMemcpyDeviceToHost, e.stream));
      GMP_HIP(hipStreamSynchronize(e.stream));
      resid = h[0];
      if (resid < p.tol) {
        it++;
        break;
      }
    }
  }
  GMP_HIP(hipEventRecord(ev[1], e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  double kms = 0;
  for (int i = 0; i < it; i++) {
    float t = 0;
    GMP_HIP(hipEventElapsedTime(&t, ev[2 + 2 * i], ev[3 + 2 * i]));
    kms += t;
  }
  float tot = 0;
  GMP_HIP(hipEventElapsedTime(&tot, ev[0], ev[1]));
  if (it > 0) {
    double h[2];
    GMP_HIP(hipMemcpy(h, hist + 2 * (it - 1), sizeof h, hipMemcpyDeviceToHost));
    resid = h[0];
    dang = h[1];
  }
  if (rank_out) GMP_HIP(hipMemcpy(rank_out, rank[cur], (size_t)V * 4, hipMemcpyDeviceToHost));
  if (st) {
    st->iterations = it;
    st->launches = it;
    st->ms_total = tot;
    st->ms_kernel = kms;
    st->edges = g.E * (uint64_t)it;
    st->vertices = (uint64_t)V * it;
    st->residual = resid;
    st->dangling = dang;
    st->changed = 0;
    const uint64_t o = g.off64 ? 8 : 4;
    st->bytes = (uint64_t)it * (8 * g.E + (o + 16) * (uint64_t)V);
    uint64_t ib = 0;
    for (const auto &x : s.win) ib += x.b1 - x.b0;
    st->index_bytes = ib;
  }
}

template <class OffT>
void pagerank_impl(Engine &e, const gmp_pr_params &p, float *rank_out, gmp_stats *st) {
  Graph &g = e.g;
  if (g.semi) {
    pagerank_windows<OffT>(e, p, rank_out, st);
    return;
  }
  ensure_plan(e);
  GatherPlan &pl = g.plan;
  const uint32_t V = g.V;
  const bool personalised = p.seed_vertex >= 0;
  const uint32_t seed = personalised ? work_id(e, (uint32_t)p.seed_vertex) : 0xFFFFFFFFu;
  const uint32_t G = std::max<uint32_t>(1, (uint32_t)e.num_cus * e.opt.wg_per_cu);
  const uint32_t Gh = pl.nmulti ? std::min<uint32_t>((pl.nmulti + kWavesPerBlock - 1) / kWavesPerBlock, G) : 0;
  const int max_iter = std::max(0, p.max_iter);

  const bool kahan = !(p.flags & GMP_PR_NO_KAHAN);
  // source-strip layout: the default for compensated PageRank on graphs whose edges fit 32-bit
  // indices (compressed: its codes bit-packed); the row gather serves the other cases
  // (compressed: varint column ids)
  const bool packed = (p.flags & GMP_PR_COMPRESSED) && kahan && ensure_strip_pack(e);
  const bool strips = packed || (!(p.flags & GMP_PR_COMPRESSED) && kahan && ensure_strips(e));
  const bool comp = (p.flags & GMP_PR_COMPRESSED) && !strips;
  if (comp) ensure_packed(e);
  // the strip layout's edges as 4-byte source ids in place of the 2-byte codes (same sums)
  const bool ids = strips && !packed && (p.flags & GMP_PR_STRIP_IDS) && ensure_strip_ids(e);
  const uint32_t np = std::max(G, strips ? strip_partials(e) : 0u);
  // with a tolerance, the strip kernels run one iteration ahead of the host's residual check
  const bool ahead = strips && p.tol > 0;

  // buffers: 2 x rank, 2 x contrib, hub partials, FP64 partials, scalars, history
  e.s_a.alloc((size_t)V * 4);
  e.s_b.alloc((size_t)V * 4);
  e.s_c.alloc((size_t)V * 4);
  e.s_d.alloc((size_t)V * 4);
  e.s_e.alloc(std::max<size_t>(pl.nparts, 1) * sizeof(float2));
  e.s_part.alloc((size_t)np * 4 * sizeof(double));
  e.s_scal.alloc(sizeof(PrScal) + (size_t)(max_iter + 1) * 2 * sizeof(double));
  float *rank[2] = {e.s_a.as<float>(), e.s_b.as<float>()};
  float *contrib[2] = {e.s_c.as<float>(), e.s_d.as<float>()};
  double *pr1 = e.s_part.as<double>(), *pd1 = pr1 + np;
  double *pr2 = pd1 + np, *pd2 = pr2 + np;
  PrScal *sc = e.s_scal.as<PrScal>();
  // the residual history lives in pinned host memory, written by the finalize kernel, so the host
  // reads an iteration's residual as soon as that kernel is done, without a copy behind later work
  double *hist = nullptr;
  double *hh = engine_host_doubles(e, 2 * (size_t)max_iter + 2, &hist);
  hipEvent_t *ev = engine_events(e, 2 * (size_t)max_iter + 2);

  PrArgs<OffT> a{};
  const CsrView wv = g.work_in();
  a.off = wv.offs<OffT>();
  a.idx = wv.idx;
  a.outdeg = g.work_deg();
  a.tasks = pl.tasks.as<uint32_t>();
  a.ntasks = pl.ntasks;
  a.hubs = pl.hubs.as<HubRow>();
  a.multi = pl.multi.as<uint32_t>();
  a.nmulti = pl.nmulti;
  a.hub_part = e.s_e.as<float2>();
  a.sc = sc;
  a.d = p.damping;
  a.seed = seed;
  a.stream_cap = pl.stream_cap;
  a.chunk = pl.chunk;

  GMP_HIP(hipEventRecord(ev[0], e.stream));
  pr_init_kernel<<<G, kBlock, 0, e.stream>>>(V, seed, a.outdeg, rank[0], contrib[0], pr1, pd1);
  GMP_CHECK_LAUNCH();
  pr_finalize_kernel<<<1, kFinalBlock, 0, e.stream>>>(pr1, pd1, G, pr2, pd2, 0, V, p.damping, personalised, sc,
                                                      nullptr, -1, 0.0);
  GMP_CHECK_LAUNCH();
  if (comp) {
    a.pk = Packed{g.packed.bytes.as<uint8_t>(), g.packed.gbyte.as<uint64_t>(), g.packed.gmask.as<uint64_t>()};
    a.E = g.E;
  }
  const size_t lds = (size_t)kWavesPerBlock * (pl.stream_cap + (comp ? kVarintScratchWords : 0u)) * sizeof(float);
  int cur = 0, it = 0;
  double resid = 0, dang = 0;
  for (it = 0; it < max_iter; it++) {
    a.contrib_in = contrib[cur];
    a.rank_in = rank[cur];
    a.rank_out = rank[cur ^ 1];
    a.contrib_out = contrib[cur ^ 1];
    a.part_resid = pr1;
    a.part_dang = pd1;
    GMP_HIP(hipEventRecord(ev[2 + 2 * it], e.stream));
    uint32_t nfin1 = G, nfin2 = Gh;
    if (strips) {
      const PrStripIO io{a.outdeg, contrib[cur], rank[cur], rank[cur ^ 1], contrib[cur ^ 1], sc, p.damping, seed,
                         pr1, pd1, pr2, pd2, packed, ids};
      strip_iteration(e, io, nfin1, nfin2);
    } else {
      if (comp) {
        if (kahan) pr_gather_kernel<OffT, true, true><<<G, kBlock, lds, e.stream>>>(a);
        else pr_gather_kernel<OffT, false, true><<<G, kBlock, lds, e.stream>>>(a);
      } else {
        if (kahan) pr_gather_kernel<OffT, true, false><<<G, kBlock, lds, e.stream>>>(a);
        else pr_gather_kernel<OffT, false, false><<<G, kBlock, lds, e.stream>>>(a);
      }
      GMP_CHECK_LAUNCH();
      if (Gh) {
        a.part_resid = pr2;
        a.part_dang = pd2;
        pr_hub_combine_kernel<OffT><<<Gh, kBlock, 0, e.stream>>>(a);
        GMP_CHECK_LAUNCH();
      }
    }
    GMP_HIP(hipEventRecord(ev[3 + 2 * it], e.stream));
    pr_finalize_kernel<<<1, kFinalBlock, 0, e.stream>>>(pr1, pd1, nfin1, pr2, pd2, nfin2, V, p.damping, personalised,
                                                        sc, hist, it, ahead ? (double)p.tol : 0.0);
    GMP_CHECK_LAUNCH();
    cur ^= 1;
    if (p.tol > 0 && !ahead) {
      GMP_HIP(hipStreamSynchronize(e.stream));
      if (hh[2 * it] < p.tol) {
        it++;
        break;
      }
    } else if (p.tol > 0 && it >= 1) {
      // iteration it is queued: read iteration it - 1's residual once its finalize is done (the
      // event recorded at the start of iteration it); when it converged, iteration it finds
      // sc->done set and does nothing, and its buffers keep iteration it - 1's ranks
      GMP_HIP(hipEventSynchronize(ev[2 + 2 * it]));
      if (hh[2 * (it - 1)] < p.tol) {
        cur ^= 1;
        break;
      }
    }
  }
  GMP_HIP(hipEventRecord(ev[1], e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  double kms = 0;
  for (int i = 0; i < it; i++) {
    float t = 0;
    GMP_HIP(hipEventElapsedTime(&t, ev[2 + 2 * i], ev[3 + 2 * i]));
    kms += t;
  }
  float tot = 0;
  GMP_HIP(hipEventElapsedTime(&tot, ev[0], ev[1]));
  if (it > 0) {
    resid = hh[2 * (it - 1)];
    dang = hh[2 * (it - 1) + 1];
  }
  if (rank_out) {
    e.s_f.alloc((size_t)V * 4);
    map_to_original(e, reinterpret_cast<const uint32_t *>(rank[cur]), e.s_f.as<uint32_t>());
    GMP_HIP(hipMemcpyAsync(rank_out, e.s_f.p, (size_t)V * 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
  }
  if (st) {
    st->iterations = it;
    st->launches = it;
    st->ms_total = tot;
    st->ms_kernel = kms;
    st->edges = g.E * (uint64_t)it;
    st->vertices = (uint64_t)V * it;
    st->residual = resid;
    st->dangling = dang;
    st->changed = 0;
    const uint64_t o = g.off64 ? 8 : 4;
    st->bytes = (uint64_t)it * (8 * g.E + (o + 16) * (uint64_t)V);
    const StripCsr &S = g.strips;
    st->index_bytes = packed ? 4 * (S.pk_words + 9 * (uint64_t)S.pk_nblk)
                      : comp ? g.packed.nbytes + 16 * g.packed.ngroups
                      : ids  ? 4 * g.E
                      : strips ? 2 * g.E
                               : 4 * g.E;
  }
}

template <class OffT>
void compat_impl(Engine &e, int app, int max_iter, const void *msg_init, const int32_t *vout,
                 void *msg_out, void *value_out, gmp_stats *st) {
  Graph &g = e.g;
  const uint32_t V = g.V;
  if (app == GMP_APP_SSSP_W && !g.weighted) fail(GMP_EUNSUPPORTED, "weighted SSSP needs a weighted graph");
  e.s_a.alloc((size_t)V * 4);
  e.s_b.alloc((size_t)V * 4);
  e.s_c.alloc((size_t)V * 4);
  e.s_d.alloc(std::max<size_t>((size_t)V * 4, 8));
  e.s_scal.alloc(8);
  void *msg[2] = {e.s_a.p, e.s_b.p};
  void *value = e.s_c.p;
  int32_t *dvout = nullptr;
  if (app == GMP_APP_PAGERANK) {
    if (!vout) fail(GMP_EINVAL, "pagerank needs vertex_out");
    dvout = e.s_d.as<int32_t>();
    GMP_HIP(hipMemcpyAsync(dvout, vout, (size_t)V * 4, hipMemcpyHostToDevice, e.stream));
  }
  GMP_HIP(hipMemcpyAsync(msg[0], msg_init, (size_t)V * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemsetAsync(value, 0, (size_t)V * 4, e.stream));
  if (g.semi) window_stamp_init(e, 0u, -1);
  unsigned long long *changed = e.s_scal.as<unsigned long long>();
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((V + kBlock - 1) / kBlock, e.num_cus * 16));
  const double invN = 1.0 / (double)V;
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  int cur = 0, step = 0;
  unsigned long long nch = 0;
  for (step = 0; step < max_iter; step++) {
    GMP_HIP(hipMemsetAsync(changed, 0, 8, e.stream));
    if (g.semi) {
      compat_windows_step<OffT>(e, app, step, grid, invN, msg[cur], msg[cur ^ 1], value, dvout, changed);
    } else {
      const CompatRows all{g.in.idx.as<uint32_t>(), g.in.w.as<uint32_t>(), nullptr, 0u, V, nullptr, 0u, 0u};
      launch_compat<OffT>(e, app, grid, invN, msg[cur], msg[cur ^ 1], value, dvout, changed, all);
    }
    GMP_HIP(hipMemcpyAsync(&nch, changed, 8, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    cur ^= 1;
    if (nch == 0 && step > 1) {
      step++;
      break;
    }
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (msg_out) GMP_HIP(hipMemcpy(msg_out, msg[cur], (size_t)V * 4, hipMemcpyDeviceToHost));
  if (value_out) GMP_HIP(hipMemcpy(value_out, value, (size_t)V * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = step;
    st->launches = step;
    st->ms_total = tot.ms();
    st->edges = g.E * (uint64_t)step;
    st->vertices = (uint64_t)V * step;
    st->changed = nch;
  }
}

}  // namespace

void run_pagerank(Engine &e, const gmp_pr_params &p, float *rank_out, gmp_stats *st) {
  if (!(p.damping >= 0.0f && p.damping <= 1.0f)) fail(GMP_EINVAL, "damping must be in [0, 1]");
  if (p.seed_vertex >= (int64_t)e.g.V) fail(GMP_EINVAL, "seed vertex out of range");
  if ((p.flags & GMP_PR_STRIP_IDS) && (p.flags & GMP_PR_COMPRESSED))
    fail(GMP_EINVAL, "GMP_PR_STRIP_IDS reads uncompressed ids: not with GMP_PR_COMPRESSED");
  if (e.g.off64) pagerank_impl<uint64_t>(e, p, rank_out, st);
  else pagerank_impl<uint32_t>(e, p, rank_out, st);
}

void run_compat(Engine &e, int app, int max_iter, const void *msg_init, const int32_t *vout,
                void *msg_out, void *value_out, gmp_stats *st) {
  if (app < GMP_APP_PAGERANK || app > GMP_APP_SSSP_W) fail(GMP_EINVAL, "unknown app id");
  if (!msg_init) fail(GMP_EINVAL, "msg_init is required");
  if (e.g.off64) compat_impl<uint64_t>(e, app, max_iter, msg_init, vout, msg_out, value_out, st);
  else compat_impl<uint32_t>(e, app, max_iter, msg_init, vout, msg_out, value_out, st);
}

}  // namespace gmp
