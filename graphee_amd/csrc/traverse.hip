// Frontier-driven SSSP and connected components on gfx950.
//
// GraphMP sweeps every shard each step and skips shards none of whose sources changed
// (selective scheduling, GraphMP/system/GraphPS.h:318-368, with per-shard bloom filters).
// Here the changed vertices themselves form the device frontier: a relax kernel pushes each
// frontier vertex's value along its out-edges with atomicMin, and every vertex whose value
// dropped is appended to the next frontier exactly once (iteration-stamped dedup) through a
// wave-aggregated append -- ballot, popcount prefix, one atomic per wave.  Within a wave the
// 64 frontier vertices' edge lists are processed cooperatively: an exclusive prefix sum of
// their degrees turns the union of their edges into one coalesced index space, and each lane
// finds its edge's owner by binary search over the wave's 64 prefix entries in LDS.
//
// Results equal the reference's fixed points regardless of processing order: SSSP distances
// are shortest-path lengths (min over relaxations), CC labels are the minimum vertex id of
// each component (min-label rule of GraphMP/app/cc.cpp:238-244 on a symmetric graph).
//
// Delta-stepping (fixed or learned bucket width) is the near/far form: vertices whose new
// distance lies below the current threshold join the near frontier, others the far pile;
// when the near frontier empties, the threshold jumps to the bucket of the smallest live far
// distance and the far pile is split.
#include <algorithm>
#include <cstring>

#include "gmp_internal.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kInf = 0xFFFFFFFFu;

enum RelaxMode { kModeWeighted = 0, kModeUnit = 1, kModeMinLabel = 2 };

struct Counters {
  uint32_t near_cnt;
  uint32_t far_cnt;
  uint32_t far_min;
  uint32_t pad;
  unsigned long long edges;
};

template <class OffT>
struct RelaxArgs {
  const OffT *off;
  const uint32_t *idx;
  const uint32_t *w;
  const uint32_t *front;  // nullptr: frontier is 0..nfront-1
  uint32_t nfront;
  uint32_t *dist;
  uint32_t *near_stamp;
  uint32_t stamp;
  uint32_t *near_out;
  uint32_t threshold;     // values >= threshold go to the far pile
  uint32_t *far_flag;
  uint32_t *far_out;
  Counters *cnt;
};

template <class OffT, int MODE>
__global__ __launch_bounds__(kBlock) void relax_kernel(RelaxArgs<OffT> a) {
  __shared__ uint32_t s_start[kWavesPerBlock][kWave];
  __shared__ uint64_t s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_val[kWavesPerBlock][kWave];
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nchunks = (a.nfront + kWave - 1) / kWave;
  unsigned long long edges = 0;
  for (uint32_t c = blockIdx.x * kWavesPerBlock + wv; c < nchunks; c += gridDim.x * kWavesPerBlock) {
    const uint32_t i = c * kWave + lane;
    const bool ok = i < a.nfront;
    const uint32_t u = ok ? (a.front ? a.front[i] : i) : 0u;
    uint64_t b = 0;
    uint32_t deg = 0, du = kInf;
    if (ok) {
      b = a.off[u];
      deg = (uint32_t)((uint64_t)a.off[u + 1] - b);
      du = a.dist[u];
      if (MODE != kModeMinLabel && du == kInf) deg = 0;
    }
    const uint32_t incl = wave_inclusive_scan(deg);
    const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
    s_start[wv][lane] = incl - deg;
    s_base[wv][lane] = b;
    s_val[wv][lane] = du;
    wave_sync();
    edges += total;
    for (uint32_t jb = 0; jb < total; jb += kWave) {
      const uint32_t j = jb + lane;
      bool to_near = false, to_far = false;
      uint32_t v = 0;
      if (j < total) {
        // owner = last lane whose start <= j (zero-degree lanes never win)
        uint32_t lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (s_start[wv][mid] <= j) lo = mid;
          else hi = mid - 1;
        }
        const uint64_t edge = s_base[wv][lo] + (j - s_start[wv][lo]);
        v = a.idx[edge];
        uint32_t nd;
        if (MODE == kModeMinLabel) {
          nd = s_val[wv][lo];
        } else {
          const uint64_t t = (uint64_t)s_val[wv][lo] + (MODE == kModeUnit ? 1u : a.w[edge]);
          nd = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
        }
        if (nd < a.dist[v]) {
          const uint32_t old = atomicMin(&a.dist[v], nd);
          if (nd < old) {
            if (nd < a.threshold) to_near = atomicExch(&a.near_stamp[v], a.stamp) != a.stamp;
            else to_far = atomicExch(&a.far_flag[v], 1u) == 0u;
          }
        }
      }
      const uint32_t sn = wave_append(to_near, &a.cnt->near_cnt);
      if (to_near) a.near_out[sn] = v;
      if (a.far_out) {
        const uint32_t sf = wave_append(to_far, &a.cnt->far_cnt);
        if (to_far) a.far_out[sf] = v;
      }
    }
    wave_sync();
  }
  if (lane == 0 && edges) atomicAdd(&a.cnt->edges, edges);
}

// Smallest distance among far entries still at or above the old threshold.
__global__ __launch_bounds__(kBlock) void far_min_kernel(const uint32_t *far, uint32_t n, const uint32_t *dist,
                                                         uint32_t thr_old, Counters *cnt) {
  uint32_t m = kInf;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t d = dist[far[i]];
    if (d >= thr_old) m = min(m, d);
  }
  m = wave_reduce_min(m);
  if (lane_id() == 0 && m != kInf) atomicMin(&cnt->far_min, m);
}

// Splits the far pile: entries settled below thr_old are dropped, entries below thr_new move
// to the near frontier, the rest stay (compacted into far_next).
__global__ __launch_bounds__(kBlock) void far_split_kernel(const uint32_t *far, uint32_t n, const uint32_t *dist,
                                                           uint32_t thr_old, uint32_t thr_new, uint32_t *far_flag,
                                                           uint32_t *near_stamp, uint32_t stamp, uint32_t *near_out,
                                                           uint32_t *far_next, Counters *cnt) {
  const uint32_t nfull = (n + kBlock - 1) / kBlock * kBlock;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nfull; i += gridDim.x * blockDim.x) {
    bool to_near = false, keep = false;
    uint32_t v = 0;
    if (i < n) {
      v = far[i];
      const uint32_t d = dist[v];
      if (d < thr_old) {
        far_flag[v] = 0;
      } else if (d < thr_new) {
        far_flag[v] = 0;
        to_near = atomicExch(&near_stamp[v], stamp) != stamp;
      } else {
        keep = true;
      }
    }
    const uint32_t sn = wave_append(to_near, &cnt->near_cnt);
    if (to_near) near_out[sn] = v;
    const uint32_t sf = wave_append(keep, &cnt->far_cnt);
    if (keep) far_next[sf] = v;
  }
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// CC labels start as the original vertex ids (so the minimum is taken over original ids
// whatever the work numbering).
__global__ void init_labels_kernel(uint32_t *p, const uint32_t *n2o, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = n2o ? n2o[i] : i;
}

// Maps a work-id result back to original ids and copies it to the host.
void finish_output(Engine &e, const uint32_t *x_work, uint32_t *host_out) {
  if (!host_out) return;
  e.s_f.alloc((size_t)e.g.V * 4);
  map_to_original(e, x_work, e.s_f.as<uint32_t>());
  GMP_HIP(hipMemcpyAsync(host_out, e.s_f.p, (size_t)e.g.V * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
}

struct TraverseBufs {
  uint32_t *dist, *stamp, *farflag, *q0, *q1, *f0, *f1;
  Counters *cnt;
};

TraverseBufs alloc_traverse(Engine &e, bool need_far) {
  const uint32_t V = e.g.V;
  const size_t n = std::max<size_t>(V, 1) * 4;
  e.s_a.alloc(n);
  e.s_b.alloc(n);
  e.s_c.alloc(n);
  e.s_d.alloc(n);
  if (need_far) {
    e.s_e.alloc(n);
    e.s_f.alloc(2 * n);
  }
  e.s_scal.alloc(sizeof(Counters));
  TraverseBufs t{};
  t.dist = e.s_a.as<uint32_t>();
  t.stamp = e.s_b.as<uint32_t>();
  t.q0 = e.s_c.as<uint32_t>();
  t.q1 = e.s_d.as<uint32_t>();
  if (need_far) {
    t.farflag = e.s_e.as<uint32_t>();
    t.f0 = e.s_f.as<uint32_t>();
    t.f1 = t.f0 + std::max<uint32_t>(V, 1);
  }
  t.cnt = e.s_scal.as<Counters>();
  return t;
}

inline uint32_t grid_fill(Engine &e, uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, (uint64_t)e.num_cus * 16));
}

template <class OffT>
void launch_relax(Engine &e, int mode, const RelaxArgs<OffT> &a) {
  const uint32_t nchunks = (a.nfront + kWave - 1) / kWave;
  const uint32_t grid = std::max<uint32_t>(
      1, std::min<uint32_t>((nchunks + kWavesPerBlock - 1) / kWavesPerBlock, (uint32_t)e.num_cus * 16));
  if (mode == kModeWeighted) relax_kernel<OffT, kModeWeighted><<<grid, kBlock, 0, e.stream>>>(a);
  else if (mode == kModeUnit) relax_kernel<OffT, kModeUnit><<<grid, kBlock, 0, e.stream>>>(a);
  else relax_kernel<OffT, kModeMinLabel><<<grid, kBlock, 0, e.stream>>>(a);
  GMP_CHECK_LAUNCH();
}

void read_counters(Engine &e, Counters *dev, Counters &h) {
  GMP_HIP(hipMemcpyAsync(e.h_pinned, dev, sizeof(Counters), hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  std::memcpy(&h, e.h_pinned, sizeof(Counters));
}

template <class OffT>
void sssp_impl(Engine &e, uint32_t src, const gmp_sssp_params &p, uint32_t *dist_out, gmp_stats *st) {
  Graph &g = e.g;
  ensure_work_out(e);
  const bool unit = p.unit_weights || !g.weighted;
  const int mode = unit ? kModeUnit : kModeWeighted;
  const bool delta_mode = p.method == GMP_SSSP_DELTA || p.method == GMP_SSSP_LEARNED;
  const uint32_t delta = p.delta ? p.delta : 64u;
  const uint32_t V = g.V;
  src = work_id(e, src);
  TraverseBufs b = alloc_traverse(e, delta_mode);
  const CsrView ov = g.work_out();
  const OffT *ooff = ov.offs<OffT>();
  const uint32_t *oidx = ov.idx;
  const uint32_t *ow = ov.w;

  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.dist, V, kInf);
  fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.stamp, V, 0u);
  if (delta_mode) fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.farflag, V, 0u);
  GMP_CHECK_LAUNCH();
  const uint32_t zero = 0;
  GMP_HIP(hipMemcpyAsync(b.dist + src, &zero, 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(b.q0, &src, 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemsetAsync(b.cnt, 0, sizeof(Counters), e.stream));

  RelaxArgs<OffT> a{};
  a.off = ooff;
  a.idx = oidx;
  a.w = ow;
  a.dist = b.dist;
  a.near_stamp = b.stamp;
  a.far_flag = b.farflag;
  a.cnt = b.cnt;
  uint32_t nnear = 1, nfar = 0, stamp = 1;
  uint32_t thr = delta_mode ? delta : kInf;
  uint32_t *near_in = b.q0, *near_out = b.q1, *far_cur = b.f0, *far_nxt = b.f1;
  int iters = 0;
  Counters h{};
  for (;;) {
    while (nnear) {
      a.front = near_in;
      a.nfront = nnear;
      a.stamp = stamp++;
      a.near_out = near_out;
      a.threshold = thr;
      a.far_out = delta_mode ? far_cur + nfar : nullptr;
      GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 8, e.stream));  // near_cnt, far_cnt
      launch_relax<OffT>(e, mode, a);
      read_counters(e, b.cnt, h);
      nnear = h.near_cnt;
      nfar += h.far_cnt;
      std::swap(near_in, near_out);
      iters++;
    }
    if (!delta_mode || nfar == 0) break;
    // next bucket: the smallest live far distance decides the new threshold
    const uint32_t inf = kInf;
    GMP_HIP(hipMemcpyAsync(&b.cnt->far_min, &inf, 4, hipMemcpyHostToDevice, e.stream));
    far_min_kernel<<<grid_fill(e, nfar), kBlock, 0, e.stream>>>(far_cur, nfar, b.dist, thr, b.cnt);
    GMP_CHECK_LAUNCH();
    read_counters(e, b.cnt, h);
    const uint32_t thr_old = thr;
    if (h.far_min == kInf) {
      thr = kInf;
    } else {
      const uint64_t t = ((uint64_t)h.far_min / delta + 1) * delta;
      thr = t >= kInf ? kInf : (uint32_t)t;
    }
    GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 8, e.stream));
    far_split_kernel<<<grid_fill(e, nfar), kBlock, 0, e.stream>>>(far_cur, nfar, b.dist, thr_old, thr, b.farflag,
                                                                 b.stamp, stamp++, near_in, far_nxt, b.cnt);
    GMP_CHECK_LAUNCH();
    read_counters(e, b.cnt, h);
    nnear = h.near_cnt;
    nfar = h.far_cnt;
    std::swap(far_cur, far_nxt);
    if (h.far_min == kInf) break;
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  read_counters(e, b.cnt, h);
  finish_output(e, b.dist, dist_out);
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = iters;
    st->launches = iters;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = h.edges;
    const uint64_t o = g.off64 ? 8 : 4;
    st->bytes = h.edges * (unit ? 8 : 12);
    (void)o;
  }
}

template <class OffT>
void cc_impl(Engine &e, uint32_t *label_out, gmp_stats *st) {
  Graph &g = e.g;
  ensure_work_out(e);
  const uint32_t V = g.V;
  TraverseBufs b = alloc_traverse(e, false);
  const CsrView ov = g.work_out();
  const OffT *ooff = ov.offs<OffT>();
  const uint32_t *oidx = ov.idx;
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  init_labels_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.dist, g.work_identity ? nullptr : g.n2o.as<uint32_t>(),
                                                              V);
  fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.stamp, V, 0u);
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemsetAsync(b.cnt, 0, sizeof(Counters), e.stream));
  RelaxArgs<OffT> a{};
  a.off = ooff;
  a.idx = oidx;
  a.dist = b.dist;
  a.near_stamp = b.stamp;
  a.threshold = kInf;
  a.cnt = b.cnt;
  uint32_t n = V, stamp = 1;
  const uint32_t *front = nullptr;  // every vertex starts active
  uint32_t *qa = b.q0, *qb = b.q1;
  int iters = 0;
  Counters h{};
  while (n) {
    a.front = front;
    a.nfront = n;
    a.stamp = stamp++;
    a.near_out = qa;
    GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 8, e.stream));
    launch_relax<OffT>(e, kModeMinLabel, a);
    read_counters(e, b.cnt, h);
    n = h.near_cnt;
    front = qa;
    std::swap(qa, qb);
    iters++;
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  finish_output(e, b.dist, label_out);
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = iters;
    st->launches = iters;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = h.edges;
    st->bytes = h.edges * 8;
  }
}

}  // namespace

void run_sssp(Engine &e, uint32_t src, const gmp_sssp_params &p, uint32_t *dist_out, gmp_stats *st) {
  if (src >= e.g.V) fail(GMP_EINVAL, "source vertex out of range");
  if (p.method < GMP_SSSP_BF || p.method > GMP_SSSP_LEARNED) fail(GMP_EINVAL, "unknown SSSP method");
  if (e.g.off64) sssp_impl<uint64_t>(e, src, p, dist_out, st);
  else sssp_impl<uint32_t>(e, src, p, dist_out, st);
}

void run_cc(Engine &e, uint32_t *label_out, gmp_stats *st) {
  if (e.g.off64) cc_impl<uint64_t>(e, label_out, st);
  else cc_impl<uint32_t>(e, label_out, st);
}

}  // namespace gmp
