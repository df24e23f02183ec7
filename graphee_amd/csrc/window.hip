// Semi-external graphs: the edges stay in pinned host memory and windows of destination rows
// stream through device memory.
//
// GraphMP keeps the vertex arrays in memory and streams the edge shards from disk through a
// compressed edge cache, skipping shards none of whose sources changed (GraphMP/system/
// GraphPS.h:256-386, Global.h:129-255).  Here the vertex arrays and the in-CSR row offsets live
// in device memory, the column ids (delta + varint packed in the format of varint.hpp, or plain)
// and the weights in pinned host memory.  The rows are cut into windows of about
// GMP_OPT_WINDOW_MB id bytes; the windows that fit into the edge budget stay cached on the
// device and the others are copied into a ring of device slots on a copy stream: each copy waits
// for the kernels of the slot's previous window, each window's kernels wait for its copy, so the
// PCIe transfers overlap the gathers.
//
// Activation: every window keeps a bitmap of the source blocks (2^blk_log consecutive ids) that
// its edges read.  The min-type sweeps stamp the block of every vertex they change with the sweep
// number, and a window runs in sweep s only when one of its blocks carries stamp s - 1 --
// GraphMP's bloom-filter test (GraphPS.h:318-351) made exact at block granularity: no window
// whose sources changed is ever skipped.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "gmp_internal.hpp"
#include "unionfind.hpp"
#include "window_order.hpp"
#include "varint.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr uint32_t kRingSlots = 4;
constexpr uint32_t kCopyStreams = 2;              // slot k copies on stream k % kCopyStreams
constexpr uint64_t kPackPad = 64;                 // zero bytes after the packed stream
constexpr uint64_t kBitmapBytes = 64ull << 20;    // all windows' source-block bitmaps together
constexpr uint64_t kStageEdges = 32ull << 20;     // ids per chunk when counting out-degrees

enum MinMode { kMinLabel = 0, kMinUnit = 1, kMinWeighted = 2 };

unsigned host_threads() {
  const unsigned n = std::thread::hardware_concurrency();
  return std::max(1u, std::min(n ? n : 1u, 16u));
}

// f(lo, hi) over consecutive parts of [0, n), on up to 16 threads; f must not throw.
template <class F>
void parallel_ranges(uint64_t n, uint64_t min_part, F f) {
  const uint64_t T = std::min<uint64_t>(host_threads(), std::max<uint64_t>(1, n / std::max<uint64_t>(min_part, 1)));
  if (T <= 1) {
    f((uint64_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (uint64_t i = 0; i < T; i++) th.emplace_back([&, i] { f(n * i / T, n * (i + 1) / T); });
  for (auto &t : th) t.join();
}

template <class T>
const T *shift_back(const void *p, uint64_t bytes) {
  return reinterpret_cast<const T *>(reinterpret_cast<uintptr_t>(p) - bytes);
}

uint8_t *pinned_alloc(uint64_t n) {
  void *p = nullptr;
  GMP_HIP(hipHostMalloc(&p, std::max<uint64_t>(n, 64), hipHostMallocDefault));
  return static_cast<uint8_t *>(p);
}

// Visits the edges of groups [g0, g1) in order: body(g, lane, value, restart) with the varint
// value and restart flag of varint.hpp (restart at a group's first edge, a row's first edge
// and wherever the ids do not ascend).
template <class F>
void for_group_values(const uint64_t *off, uint32_t V, const uint32_t *idx, uint64_t E, uint64_t g0, uint64_t g1,
                      F body) {
  if (g0 >= g1) return;
  const uint64_t first = g0 * kVarintGroup;
  uint64_t r = (uint64_t)(std::upper_bound(off, off + (uint64_t)V + 1, first) - off) - 1;
  for (uint64_t g = g0; g < g1; g++) {
    for (uint32_t lane = 0; lane < kVarintGroup; lane++) {
      const uint64_t e = g * kVarintGroup + lane;
      if (e >= E) break;
      while (off[r + 1] <= e) r++;
      const uint32_t src = idx[e];
      const uint32_t prev = lane ? idx[e - 1] : 0u;
      const bool restart = lane == 0 || off[r] == e || src <= prev;
      body(g, lane, restart ? src : src - prev - 1u, restart);
    }
  }
}

// ---- kernels ----

__global__ void outdeg_add_kernel(const uint32_t *idx, uint64_t n, uint32_t *deg) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&deg[idx[j]], 1u);
}

__global__ void fill_kernel(uint32_t *x, uint64_t n, uint32_t value) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    x[j] = value;
}

__global__ void iota_kernel(uint32_t *x, uint32_t n) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) x[j] = j;
}

__global__ void set_one_kernel(uint32_t *x, uint64_t i, uint32_t value) {
  if (blockIdx.x == 0 && threadIdx.x == 0) x[i] = value;
}

// out[e - e0] = id of edge e, e in [e0, e1); one wave per 64-edge group.
__global__ __launch_bounds__(kBlock) void window_decode_kernel(Packed pk, uint64_t E, uint64_t e0, uint64_t e1,
                                                               uint32_t *out) {
  __shared__ uint32_t scratch[kWavesPerBlock][kVarintScratchWords];
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint64_t g0 = e0 / kVarintGroup, g1 = (e1 + kVarintGroup - 1) / kVarintGroup;
  for (uint64_t g = g0 + blockIdx.x * (uint64_t)kWavesPerBlock + wv; g < g1;
       g += (uint64_t)gridDim.x * kWavesPerBlock) {
    const uint32_t id = varint_decode_group(pk, g, E, scratch[wv]);
    const uint64_t e = g * kVarintGroup + lane;
    if (e >= e0 && e < e1) out[e - e0] = id;
  }
}

// active[w] = 1 when a source block of window w carries stamp `sweep`.
__global__ __launch_bounds__(kBlock) void window_activity_kernel(const uint64_t *srcblk, uint32_t words,
                                                                 const uint32_t *stamp, uint32_t sweep,
                                                                 uint8_t *active) {
  const uint64_t *bm = srcblk + (uint64_t)blockIdx.x * words;
  int hit = 0;
  for (uint32_t i = threadIdx.x; i < words && !hit; i += blockDim.x) {
    uint64_t b = bm[i];
    while (b) {
      const uint32_t k = (uint32_t)__ffsll((long long)b) - 1u;
      b &= b - 1;
      if (stamp[(uint64_t)i * 64 + k] == sweep) {
        hit = 1;
        break;
      }
    }
  }
  hit = __syncthreads_or(hit);
  if (threadIdx.x == 0) active[blockIdx.x] = hit ? 1 : 0;
}

// Counts behind the window predictor's features, added into cnt[0..14] (cleared before): [0..7]
// active windows in each eighth of the window list, [8] active windows, [9] windows active in the
// sweep before, [10] source blocks changed in sweep `last`, [11..14] those per quarter of the blocks.
__global__ __launch_bounds__(kBlock) void window_count_kernel(const uint8_t *active, const uint8_t *prev,
                                                              uint32_t nwin, const uint32_t *stamp, uint32_t nblk,
                                                              uint32_t last, uint32_t *cnt) {
  __shared__ uint32_t sh[16];
  if (threadIdx.x < 16) sh[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x, i0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t w = i0; w < nwin; w += stride) {
    if (active[w]) {
      atomicAdd(&sh[(uint32_t)((uint64_t)w * 8 / nwin)], 1u);
      atomicAdd(&sh[8], 1u);
    }
    if (prev[w]) atomicAdd(&sh[9], 1u);
  }
  for (uint32_t b = i0; b < nblk; b += stride)
    if (stamp[b] == last) {
      atomicAdd(&sh[10], 1u);
      atomicAdd(&sh[11 + (uint32_t)((uint64_t)b * 4 / nblk)], 1u);
    }
  __syncthreads();
  if (threadIdx.x < 15 && sh[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], sh[threadIdx.x]);
}

// The window predictor (one block of 64 threads): the 16 features of the sweep -- active fraction
// of each eighth of the windows, overall and in the sweep before, changed fraction of the source
// blocks, overall and per quarter, log of the sweep number -- through the network's hidden layer to
// its 64 shard logits (shard j = rows [j V / 64, (j + 1) V / 64)); a window is predicted active in
// the sweep after next when a shard it overlaps has a positive logit.  Writes the features too.
__global__ __launch_bounds__(64) void window_predict_kernel(const float *P, const uint32_t *cnt, uint32_t nwin,
                                                            uint32_t nblk, uint32_t sweep, const uint32_t *wrow,
                                                            uint32_t V, uint8_t *pred, float *feat) {
  __shared__ float f[kMlpIn], hsh[kMlpHid], lg[kMlpShards];
  const uint32_t j = threadIdx.x;
  if (j == 0) {
    for (uint32_t i = 0; i < 8; i++) {
      const uint32_t lo = (i * nwin + 7) / 8, hi = ((i + 1) * nwin + 7) / 8;
      f[i] = hi > lo ? (float)cnt[i] / (float)(hi - lo) : 0.0f;
    }
    f[8] = (float)cnt[8] / (float)max(nwin, 1u);
    f[9] = (float)cnt[9] / (float)max(nwin, 1u);
    f[10] = (float)cnt[10] / (float)max(nblk, 1u);
    for (uint32_t q = 0; q < 4; q++) {
      const uint32_t lo = (uint32_t)(((uint64_t)q * nblk + 3) / 4), hi = (uint32_t)(((uint64_t)(q + 1) * nblk + 3) / 4);
      f[11 + q] = hi > lo ? (float)cnt[11 + q] / (float)(hi - lo) : 0.0f;
    }
    f[15] = log2f(1.0f + (float)sweep) / 6.0f;
  }
  __syncthreads();
  if (j < (uint32_t)kMlpIn) feat[j] = f[j];
  if (!P) {  // no network yet (trace recording): features only
    for (uint32_t w = j; w < nwin; w += 64) pred[w] = 0;
    return;
  }
  const float *W1 = P, *b1 = P + kMlpIn * kMlpHid, *W2 = b1 + kMlpHid, *b2 = W2 + kMlpHid * kMlpOut;
  float a = b1[j];
  for (int i = 0; i < kMlpIn; i++) a = fmaf(f[i], W1[i * kMlpHid + j], a);
  hsh[j] = a > 0.0f ? a : 0.0f;
  __syncthreads();
  float y = b2[kMlpClasses + j];
  for (int k = 0; k < kMlpHid; k++) y = fmaf(hsh[k], W2[k * kMlpOut + kMlpClasses + j], y);
  lg[j] = y;
  __syncthreads();
  for (uint32_t w = j; w < nwin; w += 64) {
    const uint32_t r0 = wrow[w], r1 = wrow[w + 1];
    const uint32_t s0 = (uint32_t)((uint64_t)r0 * kMlpShards / V);
    const uint32_t s1 = r1 > r0 ? (uint32_t)((uint64_t)(r1 - 1) * kMlpShards / V) : s0;
    bool on = false;
    for (uint32_t s = s0; s <= s1 && s < (uint32_t)kMlpShards; s++) on = on || lg[s] > 0.0f;
    pred[w] = on ? 1 : 0;
  }
}

__global__ void copy_u8_kernel(const uint8_t *src, uint8_t *dst, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

template <class OffT>
struct MinWin {
  const OffT *off;      // row offsets, shifted to the window's first row
  const uint32_t *idx;  // ids by global edge
  const uint32_t *w;    // weights by global edge (kMinWeighted), or
  const uint8_t *w8;    // the same as bytes
  uint32_t *val;        // labels or distances, global ids
  uint32_t r0;          // window's first row
  const uint32_t *tasks;
  uint32_t ntasks;
  const HubRow *hubs;
  uint32_t stream_cap, chunk;
  uint32_t *stamp;
  uint32_t blk_log, sweep;
};

template <int MODE, class OffT>
__device__ inline uint32_t min_candidate(const MinWin<OffT> &a, uint64_t j) {
  const uint32_t d = a.val[a.idx[j]];
  if (MODE == kMinLabel) return d;
  if (d == kInf) return kInf;
  const uint64_t t = (uint64_t)d + (MODE == kMinWeighted ? (a.w8 ? (uint32_t)a.w8[j] : a.w[j]) : 1u);
  return t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
}

// One sweep of a window, in place: val[v] = min(val[v], min over in-edges u->v of the candidate
// from u).  Low-degree rows (STREAM tasks) stage their candidates in LDS and take one row per
// lane; hub chunks reduce wave-wide and lower the row with atomicMin.  Any interleaving of the
// in-place updates reaches the same fixed point.
template <int MODE, class OffT>
__global__ __launch_bounds__(kBlock) void window_min_kernel(MinWin<OffT> a) {
  extern __shared__ uint32_t lds_u32[];
  const uint32_t lane = lane_id(), wv = wave_id();
  uint32_t *buf = lds_u32 + wv * a.stream_cap;
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t t = blockIdx.x * kWavesPerBlock + wv; t < a.ntasks; t += nw) {
    const uint32_t x = a.tasks[2 * t], y = a.tasks[2 * t + 1];
    if (!(y & kTaskChunkBit)) {
      const uint64_t e0 = a.off[x], e1 = a.off[x + y];
      for (uint64_t j = e0 + lane; j < e1; j += kWave) buf[j - e0] = min_candidate<MODE>(a, j);
      wave_sync();
      for (uint32_t i = lane; i < y; i += kWave) {
        const uint32_t r = x + i;
        const uint32_t b = (uint32_t)((uint64_t)a.off[r] - e0), en = (uint32_t)((uint64_t)a.off[r + 1] - e0);
        uint32_t m = kInf;
        for (uint32_t k = b; k < en; k++) m = min(m, buf[k]);
        const uint32_t v = a.r0 + r;
        if (m < a.val[v]) {
          a.val[v] = m;
          a.stamp[v >> a.blk_log] = a.sweep;
        }
      }
      wave_sync();
    } else {
      const HubRow h = a.hubs[x];
      const uint32_t k = y & ~kTaskChunkBit;
      const uint64_t rb = a.off[h.row], re = a.off[h.row + 1];
      const uint64_t e0 = rb + (uint64_t)k * a.chunk;
      const uint64_t e1 = min(e0 + a.chunk, re);
      uint32_t m = kInf;
      for (uint64_t j = e0 + lane; j < e1; j += kWave) m = min(m, min_candidate<MODE>(a, j));
      m = wave_reduce_min(m);
      if (lane == 0) {
        const uint32_t v = a.r0 + h.row;
        if (m < a.val[v] && m < atomicMin(&a.val[v], m)) a.stamp[v >> a.blk_log] = a.sweep;
      }
    }
  }
}

inline uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 1u << 16));
}

SemiExternal &semi_of(Engine &e) {
  if (!e.g.semi) fail(GMP_ESTATE, "the graph is resident, not semi-external");
  return *e.g.semi;
}

template <int MODE, class OffT>
void launch_min(Engine &e, const MinWin<OffT> &a, uint32_t grid) {
  const size_t lds = (size_t)kWavesPerBlock * a.stream_cap * 4;
  window_min_kernel<MODE, OffT><<<grid, kBlock, lds, e.stream>>>(a);
  GMP_CHECK_LAUNCH();
}

// Min-type sweeps over the windows until no window is active.  With learned prefetch (a trained
// window predictor and GMP_OPT_WINDOW_PREFETCH), the windows the predictor expects in the next
// sweep have their copies issued as soon as this sweep is enqueued -- before the host waits for
// the next sweep's exact activation -- and run first; a prefetched window the activation does
// not want is dropped (its bytes counted as wasted).  With `rows` set, a trace row is recorded
// per sweep: the predictor's features and the shard mask of the windows active one sweep later.
template <class OffT>
void semi_min(Engine &e, int mode, uint32_t src, uint32_t *out, gmp_stats *st, gmp_trace_row *rows = nullptr,
              uint32_t cap = 0, uint32_t *nrows = nullptr) {
  Graph &g = e.g;
  SemiExternal &s = semi_of(e);
  window_plans(e);
  const uint32_t V = g.V, nwin = (uint32_t)s.win.size();
  e.s_a.alloc((size_t)V * 4);
  uint32_t *val = e.s_a.as<uint32_t>();
  const bool learn = e.opt.window_prefetch && e.mlp_win.trained && s.nslots > 0;
  const bool predict = learn || rows;
  if (nrows) *nrows = 0;
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  if (mode == kMinLabel) {
    iota_kernel<<<grid_for(V), kBlock, 0, e.stream>>>(val, V);
    GMP_CHECK_LAUNCH();
    window_stamp_init(e, 0u, -1);
  } else {
    fill_kernel<<<grid_for(V), kBlock, 0, e.stream>>>(val, V, kInf);
    GMP_CHECK_LAUNCH();
    set_one_kernel<<<1, 64, 0, e.stream>>>(val, src, 0u);
    GMP_CHECK_LAUNCH();
    window_stamp_init(e, kInf, src);
  }
  const uint32_t G = std::max<uint32_t>(1, (uint32_t)e.num_cus * e.opt.wg_per_cu);
  MinWin<OffT> a{};
  a.val = val;
  a.stream_cap = g.plan.stream_cap;
  a.chunk = g.plan.chunk;
  a.stamp = s.stamp.as<uint32_t>();
  a.blk_log = s.blk_log;
  const OffT *off = g.in.off.as<OffT>();
  const uint32_t *tasks = s.tasks.as<uint32_t>();
  const HubRow *hubs = s.hubs.as<HubRow>();
  const bool weights = mode == kMinWeighted;
  // values are updated in place, so a window may see this sweep's earlier windows; the minimum
  // the sweeps converge to does not depend on the order
  const std::vector<uint32_t> order = window_order(s);
  std::vector<uint8_t> pred(nwin, 0);
  std::vector<int> pf_slot(nwin, -1);  // ring slot of a window copied ahead, -1 otherwise
  std::vector<uint32_t> pf;             // those windows, in the order their copies were issued
  float feat[kMlpIn] = {}, prev_feat[kMlpIn] = {};
  bool have_prev = false;
  int sweeps = 0;
  uint64_t edges = 0;
  auto run_window = [&](uint32_t w, const WindowDev &d) {
    const EdgeWindow &x = s.win[w];
    a.off = off + x.r0;
    a.idx = d.idx;
    a.w = d.w;
    a.w8 = d.w8;
    a.r0 = x.r0;
    a.tasks = tasks + 2ull * x.task0;
    a.ntasks = x.ntasks;
    a.hubs = hubs + x.hub0;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(G, (x.ntasks + kWavesPerBlock - 1) / kWavesPerBlock));
    if (x.ntasks) {
      if (mode == kMinLabel) launch_min<kMinLabel>(e, a, grid);
      else if (mode == kMinUnit) launch_min<kMinUnit>(e, a, grid);
      else launch_min<kMinWeighted>(e, a, grid);
    }
    window_end(e, d);
    edges += x.e1 - x.e0;
  };
  for (uint32_t sweep = 1;; sweep++) {
    const uint8_t *act = predict ? window_activity_predicted(e, sweep - 1, pred.data(), feat)
                                 : window_activity(e, sweep - 1);
    uint32_t nact = 0;
    for (uint32_t w = 0; w < nwin; w++) nact += act[w];
    if (rows && have_prev && *nrows < cap) {  // the features of the sweep before and what followed
      gmp_trace_row r{};
      std::memcpy(r.feat, prev_feat, sizeof r.feat);
      r.best = -1;
      r.epoch = sweep;
      r.shard_mask = window_shard_mask(s, act, V);
      rows[(*nrows)++] = r;
    }
    std::memcpy(prev_feat, feat, sizeof feat);
    have_prev = true;
    a.sweep = sweep;
    // windows copied ahead first: run the active ones, drop the others.  The sweeps end when no
    // window is active (no value changed in the sweep before); without selective scheduling
    // every window runs until then
    for (uint32_t w : pf) {
      if ((act[w] || !e.opt.selective) && nact) run_window(w, window_begin_slot(e, w, pf_slot[w], weights, true));
      else window_release(e, w, pf_slot[w], weights);
    }
    if (nact == 0) break;
    for (uint32_t wi = 0; wi < nwin; wi++) {
      const uint32_t w = order[wi];
      if (pf_slot[w] >= 0) continue;
      if (!act[w] && e.opt.selective) {
        s.windows_skipped++;
        continue;
      }
      run_window(w, window_begin(e, w, weights, true));
    }
    for (uint32_t w : pf) pf_slot[w] = -1;
    pf.clear();
    sweeps++;
    if (learn) {
      // the windows the predictor expects after this sweep: their copies start now, into the ring
      // slots this sweep's last windows release
      for (uint32_t wi = 0; wi < nwin && pf.size() < s.nslots; wi++) {
        const uint32_t w = order[wi];
        if (!pred[w] || s.win[w].cache >= 0) continue;
        pf_slot[w] = window_prefetch(e, w, weights);
        pf.push_back(w);
      }
    }
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (out) GMP_HIP(hipMemcpy(out, val, (size_t)V * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = sweeps;
    st->launches = sweeps;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = edges;
    st->bytes = edges * 8;
    st->policy = learn ? 1 : 0;
  }
}

// The first kAfSample column ids of rows [r0, r1) (ids by global edge index), kAfNone where a row
// is shorter.
template <class OffT>
__global__ void first_ids_kernel(const OffT *off, const uint32_t *idx, uint32_t r0, uint32_t r1, uint32_t *first) {
  for (uint32_t v = r0 + blockIdx.x * blockDim.x + threadIdx.x; v < r1; v += gridDim.x * blockDim.x) {
    const uint64_t b = off[v], d = (uint64_t)off[v + 1] - b;
#pragma unroll
    for (uint32_t r = 0; r < kAfSample; r++) first[(uint64_t)v * kAfSample + r] = r < d ? idx[b + r] : kAfNone;
  }
}

// flags[w] = 1 when a row of window w lies outside the giant component and has edges past the
// sampled ones (one block per window).
template <class OffT>
__global__ __launch_bounds__(kBlock) void finish_flags_kernel(const OffT *off, const uint32_t *p, uint32_t giant,
                                                              const uint32_t *wrow, uint8_t *flags) {
  const uint32_t r0 = wrow[blockIdx.x], r1 = wrow[blockIdx.x + 1];
  int hit = 0;
  for (uint32_t v = r0 + threadIdx.x; v < r1 && !hit; v += blockDim.x)
    if (p[v] != giant && (uint64_t)off[v + 1] - (uint64_t)off[v] > kAfSample) hit = 1;
  hit = __syncthreads_or(hit);
  if (threadIdx.x == 0) flags[blockIdx.x] = hit ? 1 : 0;
}

// CC of a symmetric semi-external graph by sampled union-find: the sampled neighbours come from a
// table of every row's first ids (extracted by one pass over the windows and kept: vertex-sized
// data), and the finish phase streams only the windows flagged above.
template <class OffT>
void semi_cc_union_find(Engine &e, uint32_t *label_out, gmp_stats *st) {
  Graph &g = e.g;
  SemiExternal &s = semi_of(e);
  const uint32_t V = g.V, nwin = (uint32_t)s.win.size();
  const OffT *off = g.in.off.as<OffT>();
  if (!s.first.p) {
    s.first.alloc((size_t)V * kAfSample * 4);
    std::vector<uint32_t> wrow(nwin + 1);
    for (uint32_t w = 0; w < nwin; w++) wrow[w] = s.win[w].r0;
    wrow[nwin] = V;
    s.wrow.alloc((size_t)(nwin + 1) * 4);
    GMP_HIP(hipMemcpy(s.wrow.p, wrow.data(), (size_t)(nwin + 1) * 4, hipMemcpyHostToDevice));
    for (const uint32_t w : window_order(s)) {  // rows are independent: any window order
      const EdgeWindow &x = s.win[w];
      const WindowDev d = window_begin(e, w, false, true);
      first_ids_kernel<OffT><<<grid_for(x.r1 - x.r0), kBlock, 0, e.stream>>>(off, d.idx, x.r0, x.r1,
                                                                              s.first.as<uint32_t>());
      GMP_CHECK_LAUNCH();
      window_end(e, d);
    }
    GMP_HIP(hipStreamSynchronize(e.stream));
  }
  e.s_a.alloc((size_t)V * 4);
  e.s_b.alloc((size_t)V * 4);
  e.s_c.alloc((size_t)kAfProbe * 4 + 8);
  uint32_t *label = e.s_a.as<uint32_t>(), *p = e.s_b.as<uint32_t>(), *probe = e.s_c.as<uint32_t>();
  unsigned long long *dedges = reinterpret_cast<unsigned long long *>(probe + kAfProbe);
  const uint32_t gv = grid_for(V);
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  af_init_kernel<<<gv, kBlock, 0, e.stream>>>(p, V);
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemsetAsync(dedges, 0, 8, e.stream));
  for (uint32_t r = 0; r < kAfSample; r++) {
    af_sample_table_kernel<<<gv, kBlock, 0, e.stream>>>(s.first.as<uint32_t>(), V, r, p);
    af_compress_kernel<<<gv, kBlock, 0, e.stream>>>(p, V);
  }
  af_probe_kernel<<<(kAfProbe + kBlock - 1) / kBlock, kBlock, 0, e.stream>>>(p, V, probe);
  GMP_CHECK_LAUNCH();
  std::vector<uint32_t> roots(kAfProbe);
  GMP_HIP(hipMemcpyAsync(roots.data(), probe, kAfProbe * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  const uint32_t giant = af_giant(roots);
  finish_flags_kernel<OffT><<<nwin, kBlock, 0, e.stream>>>(off, p, giant, s.wrow.as<uint32_t>(), s.active.as<uint8_t>());
  GMP_CHECK_LAUNCH();
  std::vector<uint8_t> flags(nwin);
  GMP_HIP(hipMemcpyAsync(flags.data(), s.active.p, nwin, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  // the links' order changes the forest but not the components, whose least ids become labels
  for (const uint32_t w : window_order(s)) {
    const EdgeWindow &x = s.win[w];
    if (!flags[w]) {
      s.windows_skipped++;
      continue;
    }
    const WindowDev d = window_begin(e, w, false, true);
    af_finish_kernel<OffT><<<grid_for(x.r1 - x.r0), kBlock, 0, e.stream>>>(off, d.idx, x.r0, x.r1, giant, p, dedges);
    GMP_CHECK_LAUNCH();
    window_end(e, d);
  }
  af_compress_kernel<<<gv, kBlock, 0, e.stream>>>(p, V);
  fill_kernel<<<gv, kBlock, 0, e.stream>>>(label, V, kInf);
  af_min_kernel<<<gv, kBlock, 0, e.stream>>>(p, nullptr, V, giant, label);
  af_label_kernel<<<gv, kBlock, 0, e.stream>>>(p, label, V, label);
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  uint64_t h = 0;
  GMP_HIP(hipMemcpyAsync(&h, dedges, 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (label_out) GMP_HIP(hipMemcpy(label_out, label, (size_t)V * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = (int)kAfSample + 1;
    st->launches = st->iterations;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = (uint64_t)V * kAfSample + h;
    st->bytes = st->edges * 8;
  }
}

}  // namespace

SemiExternal::~SemiExternal() {
  for (auto c : copy) (void)hipStreamSynchronize(c);
  for (auto x : ready) (void)hipEventDestroy(x);
  for (auto x : freed) (void)hipEventDestroy(x);
  for (auto c : copy) (void)hipStreamDestroy(c);
  if (h_ids) (void)hipHostFree(h_ids);
  if (h_w) (void)hipHostFree(h_w);
  if (h_w8) (void)hipHostFree(h_w8);
}

bool semi_wanted(const Engine &e, uint64_t E, bool weighted) {
  if (e.opt.edge_budget_mb == 0) return false;
  return E * (weighted ? 8ull : 4ull) > (e.opt.edge_budget_mb << 20);
}

void pack_host(const uint64_t *off, uint32_t V, const uint32_t *idx, uint64_t E, std::vector<uint64_t> &gbyte,
               std::vector<uint64_t> &gmask, uint8_t *bytes, uint64_t &nbytes, bool measure_only) {
  const uint64_t G = (E + kVarintGroup - 1) / kVarintGroup;
  if (measure_only) {
    gbyte.assign(G + 1, 0);
    gmask.assign(std::max<uint64_t>(G, 1), 0);
    parallel_ranges(G, 1 << 14, [&](uint64_t g0, uint64_t g1) {
      for_group_values(off, V, idx, E, g0, g1, [&](uint64_t g, uint32_t lane, uint32_t v, bool restart) {
        gbyte[g] += varint_len(v);
        if (restart) gmask[g] |= 1ull << lane;
      });
    });
    uint64_t acc = 0;
    for (uint64_t g = 0; g <= G; g++) {
      const uint64_t n = gbyte[g];
      gbyte[g] = acc;
      acc += n;
    }
    nbytes = acc;
    return;
  }
  parallel_ranges(G, 1 << 14, [&](uint64_t g0, uint64_t g1) {
    uint64_t pos = 0, cur = ~0ull;
    for_group_values(off, V, idx, E, g0, g1, [&](uint64_t g, uint32_t, uint32_t v, bool) {
      if (g != cur) {
        cur = g;
        pos = gbyte[g];
      }
      const uint32_t len = varint_len(v);
      for (uint32_t i = 0; i < len; i++)
        bytes[pos++] = (uint8_t)(((v >> (7 * i)) & 0x7fu) | (i + 1 < len ? 0x80u : 0u));
    });
  });
}

void unpack_host(const uint8_t *bytes, const uint64_t *gbyte, const uint64_t *gmask, uint64_t E, uint64_t e0,
                 uint64_t e1, uint32_t *out) {
  if (e1 <= e0) return;
  for (uint64_t g = e0 / kVarintGroup; g * kVarintGroup < e1; g++) {
    uint64_t p = gbyte[g];
    uint32_t prev = 0;
    for (uint32_t lane = 0; lane < kVarintGroup; lane++) {
      const uint64_t e = g * kVarintGroup + lane;
      if (e >= E) break;
      uint32_t v = 0, sh = 0;
      for (;;) {
        const uint8_t b = bytes[p++];
        v |= (uint32_t)(b & 0x7f) << sh;
        sh += 7;
        if (!(b & 0x80)) break;
      }
      const uint32_t id = ((gmask[g] >> lane) & 1ull) ? v : prev + v + 1u;
      prev = id;
      if (e >= e0 && e < e1) out[e - e0] = id;
    }
  }
}

void semi_setup(Engine &e, const uint32_t *idx, const uint32_t *w) {
  Graph &g = e.g;
  auto sp = std::make_unique<SemiExternal>();
  SemiExternal &s = *sp;
  const uint32_t V = g.V;
  const uint64_t E = g.E;
  const uint64_t *off = g.h_in_off.data();
  s.packed = e.opt.window_packed != 0;
  s.budget = e.opt.edge_budget_mb << 20;
  // host id stream
  std::vector<uint64_t> gbyte, gmask;
  uint64_t stream_bytes = 0;
  if (s.packed) {
    pack_host(off, V, idx, E, gbyte, gmask, nullptr, stream_bytes, true);
    s.h_ids = pinned_alloc(stream_bytes + kPackPad);
    std::memset(s.h_ids + stream_bytes, 0, kPackPad);
    pack_host(off, V, idx, E, gbyte, gmask, s.h_ids, stream_bytes, false);
    const uint64_t G = gbyte.size() - 1;
    s.gbyte.alloc((G + 1) * 8);
    s.gmask.alloc(std::max<uint64_t>(G, 1) * 8);
    GMP_HIP(hipMemcpy(s.gbyte.p, gbyte.data(), (G + 1) * 8, hipMemcpyHostToDevice));
    GMP_HIP(hipMemcpy(s.gmask.p, gmask.data(), std::max<uint64_t>(G, 1) * 8, hipMemcpyHostToDevice));
  } else {
    stream_bytes = E * 4;
    s.h_ids = pinned_alloc(stream_bytes);
    parallel_ranges(E, 1 << 22, [&](uint64_t a, uint64_t b) { std::memcpy(s.h_ids + 4 * a, idx + a, (b - a) * 4); });
  }
  s.host_bytes = stream_bytes;
  if (w) {
    // one byte per weight when every weight fits (R-MAT weights are 1..255): a quarter of the
    // weight bytes to stream and to cache
    std::vector<uint8_t> over(host_threads() + 1, 0);
    std::atomic<uint32_t> part{0};
    parallel_ranges(E, 1 << 22, [&](uint64_t a, uint64_t b) {
      uint32_t m = 0;
      for (uint64_t i = a; i < b; i++) m |= w[i];
      over[part++] = m > 0xFFu;
    });
    const bool fits = std::find(over.begin(), over.end(), (uint8_t)1) == over.end();
    if (fits) {
      s.wsize = 1;
      s.h_w8 = pinned_alloc(std::max<uint64_t>(E, 1));
      parallel_ranges(E, 1 << 22, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; i++) s.h_w8[i] = (uint8_t)w[i];
      });
    } else {
      s.wsize = 4;
      s.h_w = reinterpret_cast<uint32_t *>(pinned_alloc(E * 4));
      parallel_ranges(E, 1 << 22, [&](uint64_t a, uint64_t b) { std::memcpy(s.h_w + a, w + a, (b - a) * 4); });
    }
    s.host_bytes += E * s.wsize;
  }
  // windows: consecutive rows up to the target id bytes (a longer row makes a window alone)
  const uint64_t target = (uint64_t)e.opt.window_mb << 20;
  auto extent = [&](uint64_t e0, uint64_t e1, uint64_t &b0, uint64_t &b1) {
    if (!s.packed) {
      b0 = 4 * e0;
      b1 = 4 * e1;
      return;
    }
    // the decoder reads whole words from the word holding a group's first byte up to 3 bytes
    // past the group's last one
    const uint64_t g0 = e0 / kVarintGroup, g1 = (e1 + kVarintGroup - 1) / kVarintGroup;
    b0 = gbyte[g0] & ~3ull;
    b1 = std::min<uint64_t>(gbyte[std::max(g1, g0)] + 4, stream_bytes + kPackPad);
  };
  uint32_t r0 = 0;
  while (r0 < V) {
    uint32_t r1 = r0 + 1;
    uint64_t b0, b1;
    while (r1 < V) {
      extent(off[r0], off[r1 + 1], b0, b1);
      if (b1 - b0 > target) break;
      r1++;
    }
    EdgeWindow x;
    x.r0 = r0;
    x.r1 = r1;
    x.e0 = off[r0];
    x.e1 = off[r1];
    extent(x.e0, x.e1, x.b0, x.b1);
    s.win.push_back(x);
    r0 = r1;
  }
  const uint32_t nwin = (uint32_t)s.win.size();
  for (const auto &x : s.win) {
    s.slot_bytes = std::max(s.slot_bytes, x.b1 - x.b0);
    s.max_edges = std::max(s.max_edges, x.e1 - x.e0);
  }
  s.slot_bytes = (s.slot_bytes + 255) & ~255ull;
  // device: ring slots, decode buffer, cache of the leading windows that fit the budget
  s.nslots = kRingSlots;
  const uint64_t wslot = w ? (s.max_edges * s.wsize + 255) & ~255ull : 0;
  s.wslot_bytes = wslot;
  const uint64_t ring_bytes = s.nslots * (s.slot_bytes + wslot) + (s.packed ? s.max_edges * 4 : 0);
  uint64_t cache_ids = 0, cache_w = 0;
  uint32_t ncached = 0;
  while (ncached < nwin) {
    const EdgeWindow &x = s.win[ncached];
    const uint64_t ids = ((x.b1 - x.b0) + 255) & ~255ull, wb = w ? (x.e1 - x.e0) * s.wsize : 0;
    if (ring_bytes + cache_ids + cache_w + ids + wb > s.budget) break;
    s.win[ncached].cache = (int64_t)cache_ids;
    cache_ids += ids;
    cache_w += wb;
    ncached++;
  }
  if (ncached == nwin) s.nslots = 0;  // everything cached: no ring
  s.ring.alloc(std::max<uint64_t>((uint64_t)s.nslots * s.slot_bytes, 256));
  if (w && s.nslots) s.wring.alloc((uint64_t)s.nslots * wslot);
  if (s.packed) s.plain.alloc(std::max<uint64_t>(s.max_edges, 1) * 4);
  s.cache_bytes = cache_ids + cache_w;
  if (ncached) {
    s.cache.alloc(cache_ids);
    for (uint32_t k = 0; k < ncached; k++) {
      const EdgeWindow &x = s.win[k];
      GMP_HIP(hipMemcpy(s.cache.as<uint8_t>() + x.cache, s.h_ids + x.b0, x.b1 - x.b0, hipMemcpyHostToDevice));
    }
    if (w && cache_w) {
      s.wcache.alloc(cache_w);
      s.wcache_e0 = 0;
      GMP_HIP(hipMemcpy(s.wcache.p, s.h_w8 ? (const void *)s.h_w8 : (const void *)s.h_w, cache_w,
                        hipMemcpyHostToDevice));
    }
  }
  // out-degrees: ids counted on the device in chunks
  {
    g.out_deg.alloc((size_t)std::max<uint32_t>(V, 1) * 4);
    GMP_HIP(hipMemsetAsync(g.out_deg.p, 0, (size_t)V * 4, e.stream));
    DevBuf stage;
    const uint64_t chunk = std::min<uint64_t>(std::max<uint64_t>(E, 1), kStageEdges);
    stage.alloc(chunk * 4);
    for (uint64_t a = 0; a < E; a += chunk) {
      const uint64_t n = std::min(chunk, E - a);
      GMP_HIP(hipMemcpyAsync(stage.p, idx + a, n * 4, hipMemcpyHostToDevice, e.stream));
      outdeg_add_kernel<<<grid_for(n), kBlock, 0, e.stream>>>(stage.as<uint32_t>(), n, g.out_deg.as<uint32_t>());
      GMP_CHECK_LAUNCH();
      GMP_HIP(hipStreamSynchronize(e.stream));
    }
  }
  // activation bitmaps: blocks small enough for a useful test, all bitmaps within kBitmapBytes
  {
    uint32_t lg = 6;
    auto words_for = [&](uint32_t l) { return (uint64_t)(((uint64_t)V >> l) + 1 + 63) / 64; };
    while (lg < 32 && (uint64_t)nwin * words_for(lg) * 8 > kBitmapBytes) lg++;
    s.blk_log = lg;
    s.nblk = (uint32_t)(((uint64_t)V + (1ull << lg) - 1) >> lg);
    s.blk_words = (s.nblk + 63) / 64;
    std::vector<uint64_t> bm((uint64_t)nwin * s.blk_words, 0);
    parallel_ranges(nwin, 1, [&](uint64_t w0, uint64_t w1) {
      for (uint64_t k = w0; k < w1; k++) {
        uint64_t *b = bm.data() + k * s.blk_words;
        for (uint64_t j = s.win[k].e0; j < s.win[k].e1; j++) {
          const uint32_t blk = idx[j] >> lg;
          b[blk >> 6] |= 1ull << (blk & 63);
        }
      }
    });
    s.srcblk.alloc(std::max<uint64_t>(bm.size(), 1) * 8);
    if (!bm.empty()) GMP_HIP(hipMemcpy(s.srcblk.p, bm.data(), bm.size() * 8, hipMemcpyHostToDevice));
    s.stamp.alloc((uint64_t)std::max<uint32_t>(s.nblk, 1) * 4);
    s.active.alloc(std::max<uint32_t>(nwin, 1));
  }
  s.copy.resize(kCopyStreams);
  for (auto &c : s.copy) GMP_HIP(hipStreamCreateWithFlags(&c, hipStreamNonBlocking));
  s.ready.resize(std::max<uint32_t>(s.nslots, 1));
  s.freed.resize(std::max<uint32_t>(s.nslots, 1));
  for (auto &x : s.ready) GMP_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
  for (auto &x : s.freed) {
    GMP_HIP(hipEventCreateWithFlags(&x, hipEventDisableTiming));
    GMP_HIP(hipEventRecord(x, e.stream));
  }
  GMP_HIP(hipStreamSynchronize(e.stream));
  g.semi = std::move(sp);
  g.work_identity = true;
  g.work_ready = true;
}

void window_plans(Engine &e) {
  Graph &g = e.g;
  SemiExternal &s = semi_of(e);
  const uint32_t key[4] = {e.opt.stream_cap, e.opt.row_cap, e.opt.chunk, e.opt.vec_min};
  if (s.tasks.p && std::equal(key, key + 4, s.plan_key)) return;
  GatherPlan &pl = g.plan;  // carries the options; the tasks live in the windows' arrays
  pl.stream_cap = key[0];
  pl.row_cap = key[1];
  pl.chunk = key[2];
  pl.vec_min = key[3];
  std::vector<uint32_t> tasks, multi, t1, m1;
  std::vector<HubRow> hubs, h1;
  s.max_parts = 0;
  for (auto &x : s.win) {
    GatherPlan p1 = GatherPlan();
    p1.stream_cap = key[0];
    p1.row_cap = key[1];
    p1.chunk = key[2];
    p1.vec_min = key[3];
    build_gather_plan_host(g.h_in_off.data() + x.r0, x.r1 - x.r0, p1, t1, h1, m1);
    x.task0 = (uint32_t)(tasks.size() / 2);
    x.ntasks = p1.ntasks;
    x.hub0 = (uint32_t)hubs.size();
    x.nhubs = p1.nhubs;
    x.multi0 = (uint32_t)multi.size();
    x.nmulti = p1.nmulti;
    x.nparts = p1.nparts;
    s.max_parts = std::max(s.max_parts, p1.nparts);
    tasks.insert(tasks.end(), t1.begin(), t1.end());
    hubs.insert(hubs.end(), h1.begin(), h1.end());
    multi.insert(multi.end(), m1.begin(), m1.end());
  }
  s.tasks.release();
  s.tasks.alloc(std::max<size_t>(tasks.size(), 2) * 4);
  s.hubs.alloc(std::max<size_t>(hubs.size(), 1) * sizeof(HubRow));
  s.multi.alloc(std::max<size_t>(multi.size(), 1) * 4);
  if (!tasks.empty()) GMP_HIP(hipMemcpy(s.tasks.p, tasks.data(), tasks.size() * 4, hipMemcpyHostToDevice));
  if (!hubs.empty()) GMP_HIP(hipMemcpy(s.hubs.p, hubs.data(), hubs.size() * sizeof(HubRow), hipMemcpyHostToDevice));
  if (!multi.empty()) GMP_HIP(hipMemcpy(s.multi.p, multi.data(), multi.size() * 4, hipMemcpyHostToDevice));
  std::copy(key, key + 4, s.plan_key);
}

std::vector<uint32_t> window_order(const SemiExternal &s) {
  std::vector<bool> cached(s.win.size());
  for (size_t w = 0; w < s.win.size(); w++) cached[w] = s.win[w].cache >= 0;
  return interleave_windows(cached);
}

namespace {

// Copies window x's ids (and weights) into the next ring slot on that slot's copy stream, after
// the slot's previous window is done with it; returns the slot.
uint32_t ring_copy(SemiExternal &s, const EdgeWindow &x, bool need_w) {
  const uint32_t k = (uint32_t)(s.slot_seq++ % s.nslots);
  uint8_t *dst = s.ring.as<uint8_t>() + (uint64_t)k * s.slot_bytes;
  hipStream_t cs = s.copy[k % s.copy.size()];
  GMP_HIP(hipStreamWaitEvent(cs, s.freed[k], 0));
  if (x.b1 > x.b0) GMP_HIP(hipMemcpyAsync(dst, s.h_ids + x.b0, x.b1 - x.b0, hipMemcpyHostToDevice, cs));
  s.streamed_bytes += x.b1 - x.b0;
  if (need_w) {
    uint8_t *wd = s.wring.as<uint8_t>() + (uint64_t)k * s.wslot_bytes;
    const uint8_t *hw = s.h_w8 ? s.h_w8 : reinterpret_cast<const uint8_t *>(s.h_w);
    if (x.e1 > x.e0)
      GMP_HIP(hipMemcpyAsync(wd, hw + x.e0 * s.wsize, (x.e1 - x.e0) * s.wsize, hipMemcpyHostToDevice, cs));
    s.streamed_bytes += (x.e1 - x.e0) * s.wsize;
  }
  GMP_HIP(hipEventRecord(s.ready[k], cs));
  return k;
}

// Device pointers of window wi for work enqueued next on e.stream: from the cache (slot < 0) or
// from ring slot `slot`, whose copy the stream waits for.
WindowDev window_use(Engine &e, uint32_t wi, int slot, bool need_w, bool plain) {
  SemiExternal &s = semi_of(e);
  const EdgeWindow &x = s.win[wi];
  WindowDev d;
  const uint8_t *ids;
  const uint8_t *wts = nullptr;  // the window's first weight
  if (slot < 0) {
    ids = s.cache.as<uint8_t>() + x.cache;
    if (need_w) wts = s.wcache.as<uint8_t>() + (x.e0 - s.wcache_e0) * s.wsize;
  } else {
    GMP_HIP(hipStreamWaitEvent(e.stream, s.ready[slot], 0));
    ids = s.ring.as<uint8_t>() + (uint64_t)slot * s.slot_bytes;
    if (need_w) wts = s.wring.as<uint8_t>() + (uint64_t)slot * s.wslot_bytes;
    d.slot = slot;
  }
  if (s.packed) {
    d.pk = Packed{shift_back<uint8_t>(ids, x.b0), s.gbyte.as<uint64_t>(), s.gmask.as<uint64_t>()};
    if (plain) {
      if (x.e1 > x.e0) {
        const uint64_t groups = (x.e1 + kVarintGroup - 1) / kVarintGroup - x.e0 / kVarintGroup;
        const uint32_t grid = (uint32_t)std::min<uint64_t>((groups + kWavesPerBlock - 1) / kWavesPerBlock, 1u << 16);
        window_decode_kernel<<<grid, kBlock, 0, e.stream>>>(d.pk, e.g.E, x.e0, x.e1, s.plain.as<uint32_t>());
        GMP_CHECK_LAUNCH();
      }
      d.idx = shift_back<uint32_t>(s.plain.p, 4 * x.e0);
      if (d.slot >= 0 && !need_w) {  // the ids are decoded: the slot is free once the decode ran
        GMP_HIP(hipEventRecord(s.freed[d.slot], e.stream));
        d.slot = -1;
      }
    }
  } else {
    d.idx = shift_back<uint32_t>(ids, x.b0);
  }
  if (wts) {
    if (s.wsize == 1) d.w8 = shift_back<uint8_t>(wts, x.e0);
    else d.w = shift_back<uint32_t>(wts, 4 * x.e0);
  }
  s.windows_run++;
  return d;
}

}  // namespace

WindowDev window_begin(Engine &e, uint32_t wi, bool weights, bool plain) {
  SemiExternal &s = semi_of(e);
  const EdgeWindow &x = s.win[wi];
  const bool need_w = weights && s.has_w();
  const int slot = x.cache >= 0 ? -1 : (int)ring_copy(s, x, need_w);
  return window_use(e, wi, slot, need_w, plain);
}

int window_prefetch(Engine &e, uint32_t wi, bool weights) {
  SemiExternal &s = semi_of(e);
  const EdgeWindow &x = s.win[wi];
  if (x.cache >= 0) return -1;
  s.prefetched++;
  return (int)ring_copy(s, x, weights && s.has_w());
}

WindowDev window_begin_slot(Engine &e, uint32_t wi, int slot, bool weights, bool plain) {
  SemiExternal &s = semi_of(e);
  return window_use(e, wi, slot, weights && s.has_w(), plain);
}

void window_release(Engine &e, uint32_t wi, int slot, bool weights) {
  SemiExternal &s = semi_of(e);
  if (slot < 0) return;
  const EdgeWindow &x = s.win[wi];
  s.prefetch_wasted_bytes += (x.b1 - x.b0) + (weights && s.has_w() ? (x.e1 - x.e0) * s.wsize : 0);
  s.windows_skipped++;
  GMP_HIP(hipEventRecord(s.freed[slot], e.stream));
}

void window_end(Engine &e, const WindowDev &d) {
  SemiExternal &s = semi_of(e);
  if (d.slot >= 0) GMP_HIP(hipEventRecord(s.freed[d.slot], e.stream));
}

void window_stamp_init(Engine &e, uint32_t value, int64_t only_vertex) {
  SemiExternal &s = semi_of(e);
  fill_kernel<<<grid_for(s.nblk), kBlock, 0, e.stream>>>(s.stamp.as<uint32_t>(), s.nblk, value);
  GMP_CHECK_LAUNCH();
  if (only_vertex >= 0) {
    set_one_kernel<<<1, 64, 0, e.stream>>>(s.stamp.as<uint32_t>(), (uint64_t)only_vertex >> s.blk_log, 0u);
    GMP_CHECK_LAUNCH();
  }
}

const uint8_t *window_activity(Engine &e, uint32_t sweep) {
  SemiExternal &s = semi_of(e);
  const uint32_t nwin = (uint32_t)s.win.size();
  static thread_local std::vector<uint8_t> flags;
  flags.assign(nwin, 0);
  if (!nwin) return flags.data();
  window_activity_kernel<<<nwin, kBlock, 0, e.stream>>>(s.srcblk.as<uint64_t>(), s.blk_words,
                                                        s.stamp.as<uint32_t>(), sweep, s.active.as<uint8_t>());
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemcpyAsync(flags.data(), s.active.p, nwin, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  return flags.data();
}

void window_rows(Engine &e) {
  SemiExternal &s = semi_of(e);
  if (s.wrow.p) return;
  const uint32_t nwin = (uint32_t)s.win.size();
  std::vector<uint32_t> wrow(nwin + 1);
  for (uint32_t w = 0; w < nwin; w++) wrow[w] = s.win[w].r0;
  wrow[nwin] = e.g.V;
  s.wrow.alloc((size_t)(nwin + 1) * 4);
  GMP_HIP(hipMemcpy(s.wrow.p, wrow.data(), (size_t)(nwin + 1) * 4, hipMemcpyHostToDevice));
}

const uint8_t *window_activity_predicted(Engine &e, uint32_t sweep, uint8_t *pred, float *feat) {
  SemiExternal &s = semi_of(e);
  const uint32_t nwin = (uint32_t)s.win.size();
  static thread_local std::vector<uint8_t> flags;
  flags.assign(nwin, 0);
  if (!nwin) return flags.data();
  window_rows(e);
  if (!s.prev.p) {
    s.prev.alloc(nwin);
    s.pred.alloc(nwin);
    s.wcnt.alloc(16 * 4);
    s.wfeat.alloc(16 * 4);
    GMP_HIP(hipMemsetAsync(s.prev.p, 0, nwin, e.stream));
  }
  if (sweep == 0) GMP_HIP(hipMemsetAsync(s.prev.p, 0, nwin, e.stream));  // no sweep before the first
  window_activity_kernel<<<nwin, kBlock, 0, e.stream>>>(s.srcblk.as<uint64_t>(), s.blk_words,
                                                        s.stamp.as<uint32_t>(), sweep, s.active.as<uint8_t>());
  fill_kernel<<<1, 64, 0, e.stream>>>(s.wcnt.as<uint32_t>(), 16, 0u);
  const uint32_t gc = grid_for(std::max<uint32_t>(s.nblk, nwin));
  window_count_kernel<<<std::min<uint32_t>(gc, 1024), kBlock, 0, e.stream>>>(
      s.active.as<uint8_t>(), s.prev.as<uint8_t>(), nwin, s.stamp.as<uint32_t>(), s.nblk, sweep, s.wcnt.as<uint32_t>());
  window_predict_kernel<<<1, 64, 0, e.stream>>>(e.mlp_win.params.as<float>(), s.wcnt.as<uint32_t>(), nwin, s.nblk,
                                                sweep + 1, s.wrow.as<uint32_t>(), e.g.V, s.pred.as<uint8_t>(),
                                                s.wfeat.as<float>());
  copy_u8_kernel<<<grid_for(nwin), kBlock, 0, e.stream>>>(s.active.as<uint8_t>(), s.prev.as<uint8_t>(), nwin);
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemcpyAsync(flags.data(), s.active.p, nwin, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipMemcpyAsync(pred, s.pred.p, nwin, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipMemcpyAsync(feat, s.wfeat.p, 16 * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  return flags.data();
}

uint64_t window_shard_mask(const SemiExternal &s, const uint8_t *act, uint32_t V) {
  uint64_t m = 0;
  for (size_t w = 0; w < s.win.size(); w++) {
    if (!act[w] || s.win[w].r1 <= s.win[w].r0) continue;
    const uint32_t s0 = (uint32_t)((uint64_t)s.win[w].r0 * kMlpShards / V);
    const uint32_t s1 = (uint32_t)((uint64_t)(s.win[w].r1 - 1) * kMlpShards / V);
    for (uint32_t j = s0; j <= s1 && j < (uint32_t)kMlpShards; j++) m |= 1ull << j;
  }
  return m;
}

void semi_cc(Engine &e, uint32_t *label_out, gmp_stats *st) {
  if (e.g.symmetric && e.opt.cc_union_find) {
    if (e.g.off64) semi_cc_union_find<uint64_t>(e, label_out, st);
    else semi_cc_union_find<uint32_t>(e, label_out, st);
    return;
  }
  if (e.g.off64) semi_min<uint64_t>(e, kMinLabel, 0, label_out, st);
  else semi_min<uint32_t>(e, kMinLabel, 0, label_out, st);
}

void semi_trace(Engine &e, int kind, uint32_t src, gmp_trace_row *rows, uint32_t cap, uint32_t *nrows,
                gmp_stats *st) {
  if (kind < 0 || kind > 2) fail(GMP_EINVAL, "gmp_window_trace: kind must be 0 (labels), 1 or 2 (distances)");
  if (kind && src >= e.g.V) fail(GMP_EINVAL, "source vertex out of range");
  if (kind == 2 && !e.g.weighted) fail(GMP_EINVAL, "gmp_window_trace: the graph has no weights");
  const int mode = kind == 0 ? kMinLabel : kind == 1 ? kMinUnit : kMinWeighted;
  if (e.g.off64) semi_min<uint64_t>(e, mode, src, nullptr, st, rows, cap, nrows);
  else semi_min<uint32_t>(e, mode, src, nullptr, st, rows, cap, nrows);
}

void semi_sssp(Engine &e, uint32_t src, bool unit, uint32_t *dist_out, gmp_stats *st) {
  const int mode = (unit || !e.g.weighted) ? kMinUnit : kMinWeighted;
  if (e.g.off64) semi_min<uint64_t>(e, mode, src, dist_out, st);
  else semi_min<uint32_t>(e, mode, src, dist_out, st);
}

void semi_download(Engine &e, uint32_t *col_idx, uint32_t *weights) {
  SemiExternal &s = semi_of(e);
  const uint64_t E = e.g.E;
  if (col_idx) {
    if (s.packed) {
      std::vector<uint64_t> gbyte((E + kVarintGroup - 1) / kVarintGroup + 1), gmask(std::max<uint64_t>(gbyte.size() - 1, 1));
      GMP_HIP(hipMemcpy(gbyte.data(), s.gbyte.p, gbyte.size() * 8, hipMemcpyDeviceToHost));
      GMP_HIP(hipMemcpy(gmask.data(), s.gmask.p, gmask.size() * 8, hipMemcpyDeviceToHost));
      const uint64_t G = gbyte.size() - 1;
      parallel_ranges(G, 1 << 14, [&](uint64_t g0, uint64_t g1) {
        unpack_host(s.h_ids, gbyte.data(), gmask.data(), E, g0 * kVarintGroup, std::min(E, g1 * kVarintGroup),
                    col_idx + g0 * kVarintGroup);
      });
    } else {
      std::memcpy(col_idx, s.h_ids, E * 4);
    }
  }
  if (weights) {
    if (!s.has_w()) fail(GMP_EUNSUPPORTED, "the graph has no weights");
    if (s.h_w8) std::copy(s.h_w8, s.h_w8 + E, weights);
    else std::memcpy(weights, s.h_w, E * 4);
  }
}

void semi_info(Engine &e, gmp_window_stats *out) {
  std::memset(out, 0, sizeof *out);
  if (!e.g.semi) return;
  const SemiExternal &s = *e.g.semi;
  out->windows = (uint32_t)s.win.size();
  for (const auto &x : s.win) out->cached += x.cache >= 0 ? 1u : 0u;
  out->ring_slots = s.nslots;
  out->packed = s.packed ? 1u : 0u;
  out->host_bytes = s.host_bytes;
  out->device_bytes = s.cache_bytes + s.ring.bytes + s.wring.bytes + s.plain.bytes;
  out->streamed_bytes = s.streamed_bytes;
  out->windows_run = s.windows_run;
  out->windows_skipped = s.windows_skipped;
  out->prefetched = s.prefetched;
  out->prefetch_wasted_bytes = s.prefetch_wasted_bytes;
}

}  // namespace gmp
