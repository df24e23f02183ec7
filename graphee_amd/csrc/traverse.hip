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
// each component (min-label rule of GraphMP/app/cc.cpp:40-45 on a symmetric graph).
//
// Delta-stepping (fixed or learned bucket width) is the near/far form: vertices whose new
// distance lies below the current threshold join the near frontier, others the far pile;
// when the near frontier empties, the threshold jumps to the bucket of the smallest live far
// distance and the far pile is split.
#include <algorithm>
#include <cstring>

#include "gmp_internal.hpp"
#include "unionfind.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kInf = 0xFFFFFFFFu;
constexpr uint32_t kHubDegree = 1024;  // vertices above this out-degree relax in slices
constexpr uint32_t kHubSlice = 2048;   // edges per hub slice (one wave)

enum RelaxMode { kModeWeighted = 0, kModeUnit = 1, kModeMinLabel = 2 };

struct Counters {
  uint32_t near_cnt;
  uint32_t far_cnt;
  unsigned long long near_deg;      // out-edges of the vertices appended to the near frontier
  uint32_t far_min;
  uint32_t choice;                  // policy output: bucket-width class
  uint32_t hub_cnt;                 // flagged hub slices of the current relax launch
  uint32_t pad;
  unsigned long long edges;
  unsigned long long shard_mask;    // OR of 1 << (64 v / V) over vertices appended to the near frontier
  uint32_t hist[16];                // far-pile occupancy / degree histograms, [13] = live entries
};
static_assert(sizeof(Counters) <= 256, "counters must fit the pinned readback buffer");

// Feature vector of a bucket epoch from the far-pile histograms (shared by the recorder on the
// host and the policy kernel on the device).
__host__ __device__ inline void epoch_features(const uint32_t *hist, unsigned long long edges, uint32_t far_min,
                                               float *f) {
  const float live = (float)(hist[13] ? hist[13] : 1u);
  for (int i = 0; i < 13; i++) f[i] = (float)hist[i] / live;
  f[13] = log2f(1.0f + (float)hist[13]) / 32.0f;
  f[14] = log2f(1.0f + (float)edges) / 40.0f;
  f[15] = log2f(1.0f + (float)far_min) / 24.0f;
}

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
  uint32_t V;
  int track_shards;       // record which 64ths of the id range join the near frontier
  // hubs (out-degree > kHubDegree) leave the wave chunks; hub_relax_kernel splits their edges
  const uint32_t *hub_ids;  // ascending work ids
  uint32_t nhubs;
  uint32_t *hub_flag;       // = stamp when the hub is in this launch's frontier
  const uint2 *slices;      // {hub index, slice number}
  uint32_t nslices;
  uint32_t *hub_list;       // flagged slices of this launch (hub_compact_kernel)
  const uint32_t *deg;      // set: out-degrees, summed over the near appends into cnt->near_deg
};

constexpr int kRelaxUnroll = 4;  // edges per lane in flight in the relax loops

// Appends of improved targets are staged per wave in LDS (WaveAppend, wave.hpp).
constexpr uint32_t kAppendCap = 512;  // entries per wave and list

// Relaxes up to kRelaxUnroll edges per lane (edge[q], source value du[q]; entries with bit q
// of `live` clear are skipped) and appends the improved targets to the near frontier (first
// time this launch) or the far pile.  Each phase is issued for all edges before its results
// are read: the edge and weight loads, the distance loads, the atomicMin of the improving
// edges, then the stamp or flag exchange of those whose atomicMin lowered the distance.
template <class OffT, int MODE>
__device__ inline void relax_batch(const RelaxArgs<OffT> &a, const uint64_t (&edge)[kRelaxUnroll],
                                   const uint32_t (&du)[kRelaxUnroll], uint32_t live, unsigned long long &smask,
                                   unsigned long long &ndeg, WaveAppend &wn, WaveAppend &wf) {
  uint32_t v[kRelaxUnroll], wt[kRelaxUnroll];
#pragma unroll
  for (int q = 0; q < kRelaxUnroll; q++) {
    v[q] = a.idx[edge[q]];
    wt[q] = MODE == kModeWeighted ? a.w[edge[q]] : 1u;
  }
  uint32_t nd[kRelaxUnroll], cd[kRelaxUnroll];
#pragma unroll
  for (int q = 0; q < kRelaxUnroll; q++) {
    if (MODE == kModeMinLabel) {
      nd[q] = du[q];
    } else {
      const uint64_t t = (uint64_t)du[q] + wt[q];
      nd[q] = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
    }
    cd[q] = a.dist[v[q]];
  }
  bool tried[kRelaxUnroll];
  uint32_t old[kRelaxUnroll];
#pragma unroll
  for (int q = 0; q < kRelaxUnroll; q++) {
    tried[q] = ((live >> q) & 1u) && nd[q] < cd[q];
    old[q] = 0u;
    if (tried[q]) old[q] = atomicMin(&a.dist[v[q]], nd[q]);
  }
  // the results are first read here, after every atomicMin has been issued (an empty asm
  // taking them keeps the compares below from being moved up next to each atomic)
#pragma unroll
  for (int q = 0; q < kRelaxUnroll; q++) asm volatile("" : "+v"(old[q]));
  bool tn[kRelaxUnroll], tf[kRelaxUnroll];
  uint32_t ex[kRelaxUnroll];
#pragma unroll
  for (int q = 0; q < kRelaxUnroll; q++) {
    const bool imp = tried[q] && nd[q] < old[q];
    tn[q] = imp && nd[q] < a.threshold;
    tf[q] = imp && !(nd[q] < a.threshold);
    if (tn[q]) ex[q] = atomicExch(&a.near_stamp[v[q]], a.stamp);
    else if (tf[q]) ex[q] = atomicExch(&a.far_flag[v[q]], 1u);
  }
  uint32_t nmask = 0, fmask = 0;
#pragma unroll
  for (int q = 0; q < kRelaxUnroll; q++) {
    if (tn[q] && ex[q] != a.stamp) nmask |= 1u << q;
    if (tf[q] && ex[q] == 0u) fmask |= 1u << q;
  }
  // staged appends (LDS), flushed with one atomic when the next batch could overflow them
  const uint32_t nn = (uint32_t)__popc(nmask), nf = (uint32_t)__popc(fmask);
  const uint32_t in = wave_scan_dpp(nn);
  const uint32_t tnn = read_lane_u32(in, kWave - 1);
  if (tnn) {
    uint32_t pos = wn.n + in - nn;
#pragma unroll
    for (int q = 0; q < kRelaxUnroll; q++)
      if ((nmask >> q) & 1u) {
        wn.buf[pos++] = v[q];
        if (a.track_shards) smask |= 1ull << (uint32_t)(((uint64_t)v[q] * kMlpShards) / a.V);
        if (a.deg) ndeg += a.deg[v[q]];
      }
    wn.n += tnn;
    if (wn.n > kAppendCap - kWave * kRelaxUnroll) append_flush(wn, &a.cnt->near_cnt, a.near_out);
  }
  if (a.far_out) {
    const uint32_t inf = wave_scan_dpp(nf);
    const uint32_t tff = read_lane_u32(inf, kWave - 1);
    if (tff) {
      uint32_t pos = wf.n + inf - nf;
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++)
        if ((fmask >> q) & 1u) wf.buf[pos++] = v[q];
      wf.n += tff;
      if (wf.n > kAppendCap - kWave * kRelaxUnroll) append_flush(wf, &a.cnt->far_cnt, a.far_out);
    }
  }
}

template <class OffT>
__device__ inline void relax_flush_counts(const RelaxArgs<OffT> &a, unsigned long long edges, unsigned long long smask,
                                          unsigned long long ndeg, WaveAppend &wn, WaveAppend &wf) {
  append_flush(wn, &a.cnt->near_cnt, a.near_out);
  if (a.far_out) append_flush(wf, &a.cnt->far_cnt, a.far_out);
  if (lane_id() == 0 && edges) atomicAdd(&a.cnt->edges, edges);
  if (a.deg) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ndeg += (unsigned long long)__shfl_xor((long long)ndeg, o, kWave);
    if (lane_id() == 0 && ndeg) atomicAdd(&a.cnt->near_deg, ndeg);
  }
  if (a.track_shards) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) smask |= (unsigned long long)__shfl_xor((long long)smask, o, kWave);
    if (lane_id() == 0 && smask) atomicOr(&a.cnt->shard_mask, smask);
  }
}

template <class OffT, int MODE>
__global__ __launch_bounds__(kBlock) void relax_kernel(RelaxArgs<OffT> a) {
  __shared__ uint32_t s_start[kWavesPerBlock][kWave];
  __shared__ uint64_t s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_val[kWavesPerBlock][kWave];
  __shared__ uint32_t s_app[kWavesPerBlock][2][kAppendCap];
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nchunks = (a.nfront + kWave - 1) / kWave;
  unsigned long long edges = 0, smask = 0, ndeg = 0;
  WaveAppend wn{s_app[wv][0], 0u}, wf{s_app[wv][1], 0u};
  if (blockIdx.x == 0 && threadIdx.x == 0) a.cnt->hub_cnt = 0;  // hub_compact_kernel counts after us
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
      // a vertex whose value already dropped again in this launch is in the next frontier,
      // where it relaxes its edges from the lower value: skip it here
      if (a.near_stamp[u] == a.stamp) deg = 0;
      if (deg > kHubDegree && a.nhubs) {
        // a hub: its edges go to hub_relax_kernel's slices, spread over many waves
        uint32_t lo = 0, hi = a.nhubs - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (a.hub_ids[mid] < u) lo = mid + 1;
          else hi = mid;
        }
        a.hub_flag[lo] = a.stamp;
        deg = 0;
      }
    }
    const uint32_t incl = wave_scan_dpp(deg);
    const uint32_t total = read_lane_u32(incl, kWave - 1);
    s_start[wv][lane] = incl - deg;
    s_base[wv][lane] = b;
    s_val[wv][lane] = du;
    wave_sync();
    edges += total;
    for (uint32_t jb = 0; jb < total; jb += kWave * kRelaxUnroll) {
      uint64_t e[kRelaxUnroll];
      uint32_t du[kRelaxUnroll], live = 0;
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        const uint32_t j = jb + lane + (uint32_t)q * kWave;
        if (j < total) live |= 1u << q;
        const uint32_t jj = j < total ? j : 0u;  // past the end: the wave's first edge, skipped
        // owner = last lane whose start <= jj (zero-degree lanes never win)
        uint32_t lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (s_start[wv][mid] <= jj) lo = mid;
          else hi = mid - 1;
        }
        e[q] = s_base[wv][lo] + (jj - s_start[wv][lo]);
        du[q] = s_val[wv][lo];
      }
      relax_batch<OffT, MODE>(a, e, du, live, smask, ndeg, wn, wf);
    }
    wave_sync();
  }
  relax_flush_counts(a, edges, smask, ndeg, wn, wf);
}

// Lists the slices of the hubs flagged by this launch's relax_kernel (one thread per slice).
template <class OffT>
__global__ __launch_bounds__(kBlock) void hub_compact_kernel(RelaxArgs<OffT> a) {
  const uint32_t n = (a.nslices + kBlock - 1) / kBlock * kBlock;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const bool f = t < a.nslices && a.hub_flag[a.slices[t].x] == a.stamp;
    const uint32_t p = wave_append(f, &a.cnt->hub_cnt);
    if (f) a.hub_list[p] = t;
  }
}

// The listed hub slices: one wave per slice of kHubSlice edges, so a hub's edge list is
// spread over the whole chip.
template <class OffT, int MODE>
__global__ __launch_bounds__(kBlock) void hub_relax_kernel(RelaxArgs<OffT> a) {
  __shared__ uint32_t s_app[kWavesPerBlock][2][kAppendCap];
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nlist = a.cnt->hub_cnt;
  unsigned long long edges = 0, smask = 0, ndeg = 0;
  WaveAppend wn{s_app[wv][0], 0u}, wf{s_app[wv][1], 0u};
  for (uint32_t i = blockIdx.x * kWavesPerBlock + wv; i < nlist; i += gridDim.x * kWavesPerBlock) {
    const uint2 sl = a.slices[a.hub_list[i]];
    const uint32_t u = a.hub_ids[sl.x];
    const uint64_t rb = a.off[u], re = a.off[u + 1];
    const uint64_t e0 = rb + (uint64_t)sl.y * kHubSlice, e1 = min(e0 + kHubSlice, re);
    const uint32_t du = a.dist[u];
    edges += e1 - e0;
    for (uint64_t jb = e0; jb < e1; jb += (uint64_t)kWave * kRelaxUnroll) {
      uint64_t e[kRelaxUnroll];
      uint32_t dv[kRelaxUnroll], live = 0;
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        const uint64_t j = jb + lane + (uint64_t)q * kWave;
        if (j < e1) live |= 1u << q;
        e[q] = j < e1 ? j : e0;
        dv[q] = du;
      }
      relax_batch<OffT, MODE>(a, e, dv, live, smask, ndeg, wn, wf);
    }
  }
  relax_flush_counts(a, edges, smask, ndeg, wn, wf);
}

// Histograms of the live far pile for the epoch features: distance above the pile minimum in
// bins [0,8),[8,16),...,[512,inf) -> hist[0..7]; out-degree <=1, 2-7, 8-63, 64-511, >=512 ->
// hist[8..12]; live count -> hist[13].
template <class OffT>
__global__ __launch_bounds__(kBlock) void far_features_kernel(const uint32_t *far, uint32_t n, const uint32_t *dist,
                                                              const OffT *off, uint32_t thr_old, Counters *cnt) {
  __shared__ uint32_t h[16];
  if (threadIdx.x < 16) h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t m = cnt->far_min;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = far[i];
    const uint32_t d = dist[v];
    if (d < thr_old) continue;
    const uint32_t x = d - m;
    const uint32_t ob = x < 8 ? 0u : min(7u, 31u - (uint32_t)__clz((int)x) - 2u);
    const uint32_t deg = (uint32_t)((uint64_t)off[v + 1] - (uint64_t)off[v]);
    const uint32_t db = deg <= 1 ? 0u : deg < 8 ? 1u : deg < 64 ? 2u : deg < 512 ? 3u : 4u;
    atomicAdd(&h[ob], 1u);
    atomicAdd(&h[8 + db], 1u);
    atomicAdd(&h[13], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 14 && h[threadIdx.x]) atomicAdd(&cnt->hist[threadIdx.x], h[threadIdx.x]);
}

// Inference of the scheduling MLP for one epoch: features -> 64 hidden -> 8 width logits;
// the argmax class goes to cnt->choice (ties to the smaller width).
__global__ __launch_bounds__(64) void policy_kernel(const float *P, Counters *cnt) {
  __shared__ float f[kMlpIn], hsh[kMlpHid], logits[kMlpClasses];
  if (threadIdx.x == 0) epoch_features(cnt->hist, cnt->edges, cnt->far_min, f);
  __syncthreads();
  const float *W1 = P, *b1 = P + kMlpIn * kMlpHid, *W2 = b1 + kMlpHid, *b2 = W2 + kMlpHid * kMlpOut;
  const uint32_t j = threadIdx.x;
  float a = b1[j];
  for (int i = 0; i < kMlpIn; i++) a = fmaf(f[i], W1[i * kMlpHid + j], a);
  hsh[j] = a > 0.0f ? a : 0.0f;
  __syncthreads();
  if (j < (uint32_t)kMlpClasses) {
    float y = b2[j];
    for (int k = 0; k < kMlpHid; k++) y = fmaf(hsh[k], W2[k * kMlpOut + j], y);
    logits[j] = y;
  }
  __syncthreads();
  if (j == 0) {
    uint32_t best = 0;
    for (uint32_t c = 1; c < (uint32_t)kMlpClasses; c++)
      if (logits[c] > logits[best]) best = c;
    cnt->choice = best;
  }
}

// Smallest distance among far entries still at or above the old threshold.
__global__ __launch_bounds__(kBlock) void far_min_kernel(const uint32_t *far, uint32_t n, const uint32_t *dist,
                                                         uint32_t thr_old, Counters *cnt) {
  __shared__ uint32_t sm[kWavesPerBlock];
  uint32_t m = kInf;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t d = dist[far[i]];
    if (d >= thr_old) m = min(m, d);
  }
  m = wave_reduce_min(m);
  if (lane_id() == 0) sm[wave_id()] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kWavesPerBlock; w++) m = min(m, sm[w]);
    if (m != kInf) atomicMin(&cnt->far_min, m);
  }
}

// Splits the far pile at a new threshold: entries below the old threshold are settled and
// dropped, entries below the new one join the near frontier (once per launch), the rest stay.
// kSplitPer entries per thread, their loads in flight together; one atomic per block and list.
constexpr uint32_t kSplitPer = 4;
__global__ __launch_bounds__(kBlock) void far_split_kernel(const uint32_t *far, uint32_t n, const uint32_t *dist,
                                                           uint32_t thr_old, uint32_t thr_new, uint32_t *far_flag,
                                                           uint32_t *near_stamp, uint32_t stamp, uint32_t *near_out,
                                                           uint32_t *far_next, Counters *cnt, const uint32_t *deg) {
  __shared__ uint32_t shn[kWavesPerBlock + 1], shk[kWavesPerBlock + 1];
  const uint64_t per_block = (uint64_t)kBlock * kSplitPer;
  unsigned long long ndeg = 0;
  for (uint64_t b0 = blockIdx.x * per_block; b0 < n; b0 += (uint64_t)gridDim.x * per_block) {
    uint32_t v[kSplitPer], d[kSplitPer];
#pragma unroll
    for (uint32_t k = 0; k < kSplitPer; k++) {
      const uint64_t i = b0 + k * kBlock + threadIdx.x;
      v[k] = i < n ? far[i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kSplitPer; k++) d[k] = dist[v[k]];
    uint32_t to_near = 0, keep = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSplitPer; k++) {
      if (b0 + k * kBlock + threadIdx.x >= n) continue;
      if (d[k] < thr_old) {
        far_flag[v[k]] = 0;
      } else if (d[k] < thr_new) {
        far_flag[v[k]] = 0;
        if (atomicExch(&near_stamp[v[k]], stamp) != stamp) to_near |= 1u << k;
      } else {
        keep |= 1u << k;
      }
    }
    uint32_t pn = block_append<kWavesPerBlock>((uint32_t)__popc(to_near), &cnt->near_cnt, shn);
    uint32_t pk = block_append<kWavesPerBlock>((uint32_t)__popc(keep), &cnt->far_cnt, shk);
#pragma unroll
    for (uint32_t k = 0; k < kSplitPer; k++) {
      if ((to_near >> k) & 1u) {
        near_out[pn++] = v[k];
        if (deg) ndeg += deg[v[k]];
      }
      if ((keep >> k) & 1u) far_next[pk++] = v[k];
    }
  }
  if (deg) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ndeg += (unsigned long long)__shfl_xor((long long)ndeg, o, kWave);
    if (lane_id() == 0 && ndeg) atomicAdd(&cnt->near_deg, ndeg);
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

// hub candidates: {work id, out-degree} of every vertex above kHubDegree (unordered)
template <class OffT>
__global__ void hub_collect_kernel(const OffT *off, uint32_t V, uint32_t cap, uint32_t *cnt, uint2 *out) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const uint64_t d = (uint64_t)off[v + 1] - (uint64_t)off[v];
    if (d > kHubDegree) {
      const uint32_t i = atomicAdd(cnt, 1u);
      if (i < cap) out[i] = make_uint2(v, d > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)d);
    }
  }
}

// Builds e.g.hubs for the work out-CSR (once per layout).
template <class OffT>
void ensure_hubs(Engine &e) {
  Graph &g = e.g;
  HubTable &T = g.hubs;
  if (T.valid) return;
  T = HubTable();
  const CsrView ov = g.work_out();
  const uint32_t V = g.V;
  const uint32_t cap = (uint32_t)std::min<uint64_t>(g.E / (kHubDegree + 1) + 1, V);
  DevBuf cand;
  cand.alloc((size_t)cap * sizeof(uint2) + 16);
  uint32_t *cnt = reinterpret_cast<uint32_t *>(cand.as<char>() + (size_t)cap * sizeof(uint2));
  GMP_HIP(hipMemsetAsync(cnt, 0, 4, e.stream));
  hub_collect_kernel<OffT><<<grid_fill(e, V), kBlock, 0, e.stream>>>(ov.offs<OffT>(), V, cap, cnt, cand.as<uint2>());
  GMP_CHECK_LAUNCH();
  uint32_t H = 0;
  GMP_HIP(hipMemcpyAsync(&H, cnt, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (H > cap) fail(GMP_ESTATE, "hub table overflow");
  std::vector<uint2> hub(H);
  if (H) GMP_HIP(hipMemcpy(hub.data(), cand.p, (size_t)H * sizeof(uint2), hipMemcpyDeviceToHost));
  std::sort(hub.begin(), hub.end(), [](const uint2 &p, const uint2 &q) { return p.x < q.x; });
  std::vector<uint32_t> ids(H);
  std::vector<uint2> sl;
  for (uint32_t h = 0; h < H; h++) {
    ids[h] = hub[h].x;
    for (uint32_t k = 0; (uint64_t)k * kHubSlice < hub[h].y; k++) sl.push_back(make_uint2(h, k));
  }
  T.H = H;
  T.nslices = (uint32_t)sl.size();
  if (H) {
    T.ids.alloc((size_t)H * 4);
    T.flag.alloc((size_t)H * 4);
    T.slices.alloc(sl.size() * sizeof(uint2));
    T.list.alloc(sl.size() * 4);
    GMP_HIP(hipMemcpy(T.ids.p, ids.data(), (size_t)H * 4, hipMemcpyHostToDevice));
    GMP_HIP(hipMemcpy(T.slices.p, sl.data(), sl.size() * sizeof(uint2), hipMemcpyHostToDevice));
  }
  T.valid = true;
}

// Points the relax arguments at the hub table and clears its flags (stamps restart per run).
template <class OffT>
void attach_hubs(Engine &e, RelaxArgs<OffT> &a) {
  ensure_hubs<OffT>(e);
  HubTable &T = e.g.hubs;
  a.nhubs = T.H;
  a.nslices = T.nslices;
  a.hub_ids = T.ids.as<uint32_t>();
  a.hub_flag = T.flag.as<uint32_t>();
  a.slices = T.slices.as<uint2>();
  a.hub_list = T.list.as<uint32_t>();
  if (T.H) {
    fill_u32_kernel<<<grid_fill(e, T.H), kBlock, 0, e.stream>>>(a.hub_flag, T.H, 0u);
    GMP_CHECK_LAUNCH();
  }
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
  if (a.nslices) {
    hub_compact_kernel<OffT><<<grid_fill(e, a.nslices), kBlock, 0, e.stream>>>(a);
    GMP_CHECK_LAUNCH();
    const uint32_t gh = std::max<uint32_t>(
        1, std::min<uint32_t>((a.nslices + kWavesPerBlock - 1) / kWavesPerBlock, (uint32_t)e.num_cus * 8));
    if (mode == kModeWeighted) hub_relax_kernel<OffT, kModeWeighted><<<gh, kBlock, 0, e.stream>>>(a);
    else if (mode == kModeUnit) hub_relax_kernel<OffT, kModeUnit><<<gh, kBlock, 0, e.stream>>>(a);
    else hub_relax_kernel<OffT, kModeMinLabel><<<gh, kBlock, 0, e.stream>>>(a);
    GMP_CHECK_LAUNCH();
  }
}

void read_counters(Engine &e, Counters *dev, Counters &h) {
  GMP_HIP(hipMemcpyAsync(e.h_pinned, dev, sizeof(Counters), hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  std::memcpy(&h, e.h_pinned, sizeof(Counters));
}

// Shards (64ths of the id range) holding a vertex whose value differs between before and after,
// OR-ed into cnt->shard_mask: the trace's shard mask of an epoch relaxed without shard tracking.
__global__ __launch_bounds__(kBlock) void changed_shards_kernel(const uint32_t *before, const uint32_t *after,
                                                                uint32_t V, Counters *cnt) {
  unsigned long long m = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
    if (before[v] != after[v]) m |= 1ull << (uint32_t)(((uint64_t)v * kMlpShards) / V);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m |= (unsigned long long)__shfl_xor((long long)m, o, kWave);
  if (lane_id() == 0 && m) atomicOr(&cnt->shard_mask, m);
}

// Vertices whose stamp equals `sweep`, appended to list (count = cnt->near_cnt): kCollectPer
// vertices per thread and one atomic per block and pass.
constexpr uint32_t kCollectPer = 16;
__global__ __launch_bounds__(kBlock) void collect_stamped_kernel(const uint32_t *vstamp, uint32_t V, uint32_t sweep,
                                                                 uint32_t *list, Counters *cnt) {
  __shared__ uint32_t sh[kWavesPerBlock + 1];
  const uint64_t per_block = (uint64_t)kBlock * kCollectPer;
  for (uint64_t b0 = blockIdx.x * per_block; b0 < V; b0 += (uint64_t)gridDim.x * per_block) {
    uint32_t hit = 0;
#pragma unroll
    for (uint32_t k = 0; k < kCollectPer; k++) {
      const uint64_t v = b0 + k * kBlock + threadIdx.x;
      if (v < V && vstamp[v] == sweep) hit |= 1u << k;
    }
    uint32_t pos = block_append<kWavesPerBlock>((uint32_t)__popc(hit), &cnt->near_cnt, sh);
#pragma unroll
    for (uint32_t k = 0; k < kCollectPer; k++)
      if ((hit >> k) & 1u) list[pos++] = (uint32_t)(b0 + k * kBlock + threadIdx.x);
  }
}

// Direction-optimizing min solve (CC labels, Bellman-Ford distances) on the work layout: frontier
// pushes over the out-CSR while at most V >> push_log vertices are active, dense pull sweeps over
// the source strips otherwise (GraphPS's sparse/dense switch, GraphPS/include/defines.h:9 and
// GraphPS/graphps.h:336-390, on one GPU).  SSSP starts pushing from the source, CC sweeping over
// every vertex; either may switch any number of times.  A sweep stamps the vertices it changes,
// so its frontier is exact; entering sweeps marks only the frontier's strips.  Every phase keeps
// relaxing until nothing drops, so the fixed point is the pure sweeps' or pure pushes'.  With
// front0 set, the values already in e.s_a (plain) are continued from the n0 vertices of front0
// (one of e.s_c, e.s_d: changed, not yet propagated) instead of the starting values.  Returns
// false, having launched nothing, when the strip layout does not apply; the values (plain, work
// order) are left in e.s_a.
template <class OffT>
bool hybrid_min(Engine &e, int kind, uint32_t src, uint32_t *front0, uint32_t n0, int &rounds, int &sweeps, double &ms,
                uint64_t &edges) {
  Graph &g = e.g;
  const uint32_t V = g.V;
  TraverseBufs b = alloc_traverse(e, false);
  e.s_e.alloc(std::max<size_t>(V, 1) * 4);
  uint32_t *vstamp = e.s_e.as<uint32_t>();
  const uint32_t T = e.opt.push_log ? (V >> e.opt.push_log) : 0u;  // largest frontier that pushes
  SweepRunPtr sw = sweep_open(e, kind, b.dist, T ? vstamp : nullptr);
  if (!sw) return false;
  RelaxArgs<OffT> a{};
  if (T) {
    ensure_work_out(e);
    ensure_hubs<OffT>(e);
    const CsrView ov = g.work_out();
    a.off = ov.offs<OffT>();
    a.idx = ov.idx;
    a.w = ov.w;
  }
  const int mode = kind == 0 ? kModeMinLabel : kind == 1 ? kModeUnit : kModeWeighted;
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  GMP_HIP(hipMemsetAsync(b.cnt, 0, sizeof(Counters), e.stream));
  if (T) {
    fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(vstamp, V, 0u);
    fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.stamp, V, 0u);
    GMP_CHECK_LAUNCH();
    a.dist = b.dist;
    a.near_stamp = b.stamp;
    a.threshold = kInf;
    a.cnt = b.cnt;
    a.V = V;
    attach_hubs(e, a);
  }
  uint32_t *qa = b.q0, *qb = b.q1;
  uint32_t n = 0, stamp = 1;
  bool pushing = false;
  if (front0) {  // continue: a frontier larger than T switches to sweeps at once
    qa = front0;
    qb = front0 == b.q0 ? b.q1 : b.q0;
    n = n0;
    pushing = true;
  } else if (kind == 0 || !T) {
    sweep_start(*sw, src);  // complemented starting values, every strip (CC) or the source's active
  } else {
    fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.dist, V, kInf);
    GMP_CHECK_LAUNCH();
    const uint32_t zero = 0;
    GMP_HIP(hipMemcpyAsync(b.dist + src, &zero, 4, hipMemcpyHostToDevice, e.stream));
    GMP_HIP(hipMemcpyAsync(qa, &src, 4, hipMemcpyHostToDevice, e.stream));
    n = 1;
    pushing = true;
  }
  rounds = 0;
  sweeps = 0;
  Counters h{};
  for (;;) {
    if (pushing) {
      if (n == 0) break;
      if (n > T) {  // dense: continue with sweeps from the frontier
        complement_values(e, b.dist, V);
        sweep_enter(*sw, qa, n);
        pushing = false;
        continue;
      }
      a.front = qa;
      a.nfront = n;
      a.stamp = stamp++;
      a.near_out = qb;
      GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 8, e.stream));  // near_cnt, far_cnt
      launch_relax<OffT>(e, mode, a);
      read_counters(e, b.cnt, h);
      n = h.near_cnt;
      std::swap(qa, qb);
      rounds++;
    } else {
      const uint32_t changed = sweep_step(*sw);
      sweeps++;
      if (changed == 0) break;
      if (changed <= T) {  // sparse: push from the vertices this sweep changed
        GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 4, e.stream));
        collect_stamped_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(vstamp, V, sweep_last(*sw), qa, b.cnt);
        GMP_CHECK_LAUNCH();
        complement_values(e, b.dist, V);
        read_counters(e, b.cnt, h);
        n = h.near_cnt;  // the frontier the list holds (every changed vertex is stamped once)
        pushing = true;
      }
    }
  }
  if (!pushing) complement_values(e, b.dist, V);
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  read_counters(e, b.cnt, h);
  edges = sweep_edges(*sw) + h.edges;
  ms = tot.ms();
  return true;
}

// Bucket epochs of delta-stepping.  FIXED: aligned buckets [k*delta, (k+1)*delta).  LEARNED:
// the MLP picks each epoch's bucket from far-pile features (m = smallest live far distance):
// class c the bucket [m, m + 8 * 2^c), widths 8 .. 1024; with the learned-finish option, class 7
// instead finishes the solve as Bellman-Ford pushes and strip sweeps (hybrid_min) from the live
// far pile.  EXPLORE records traces: at checkpoint epochs, from a snapshot of the state, every
// class runs the epoch and then fixed width 64 runs to the end of the solve; the class with the
// shortest such rollout is the label and the run continues with it.  Every policy relaxes until
// no distance drops, so distances are the same.
enum Policy { kPolicyBF = 0, kPolicyFixed = 1, kPolicyLearned = 2, kPolicyExplore = 3 };
constexpr uint32_t kFinishClass = 7;  // the finishing class, when the option enables it

template <class OffT>
void sssp_impl(Engine &e, uint32_t src, const gmp_sssp_params &p, int policy, gmp_trace_row *rows, uint32_t cap,
               uint32_t *nrows, uint32_t *dist_out, gmp_stats *st) {
  Graph &g = e.g;
  const bool unit = p.unit_weights || !g.weighted;
  if (policy == kPolicyBF) {
    // Bellman-Ford: frontier pushes while the frontier is small, sweeps over the source strips
    // while it is large (when the strip layout applies); the frontier form below otherwise
    ensure_work(e);
    int rounds = 0, sweeps = 0;
    double ms = 0.0;
    uint64_t edges = 0;
    if (hybrid_min<OffT>(e, unit ? 1 : 2, work_id(e, src), nullptr, 0u, rounds, sweeps, ms, edges)) {
      finish_output(e, e.s_a.as<uint32_t>(), dist_out);
      if (st) {
        std::memset(st, 0, sizeof *st);
        st->iterations = rounds + sweeps;
        st->launches = rounds + sweeps;
        st->ms_total = ms;
        st->ms_kernel = ms;
        st->edges = edges;
        st->bytes = edges * (unit ? 8 : 12);
        st->epochs = sweeps;
      }
      return;
    }
  }
  ensure_work_out(e);
  const int mode = unit ? kModeUnit : kModeWeighted;
  if (policy == kPolicyLearned && !e.mlp.trained) policy = kPolicyFixed;  // untrained network: plain schedule
  if (policy == kPolicyLearned) std::fill(std::begin(e.last_actions), std::end(e.last_actions), 0u);
  const bool bucketed = policy != kPolicyBF;
  const uint32_t delta = p.delta ? p.delta : 64u;
  const uint32_t V = g.V;
  src = work_id(e, src);
  ensure_hubs<OffT>(e);
  // pull rounds need the source-strip layout (and its weights): opened before the clock starts
  SweepRunPtr pr;
  if (e.opt.delta_pull_log && g.E) pr = pull_open(e, unit);
  TraverseBufs b = alloc_traverse(e, true);
  const CsrView ov = g.work_out();
  const OffT *ooff = ov.offs<OffT>();
  const uint32_t *oidx = ov.idx;
  const uint32_t *ow = ov.w;
  if (nrows) *nrows = 0;

  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.dist, V, kInf);
  fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.stamp, V, 0u);
  fill_u32_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(b.farflag, V, 0u);
  GMP_CHECK_LAUNCH();
  const uint32_t zero = 0, one = 1;
  GMP_HIP(hipMemcpyAsync(b.dist + src, &zero, 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemsetAsync(b.cnt, 0, sizeof(Counters), e.stream));
  uint32_t nnear = 0, nfar = 0, stamp = 1, thr = 0;
  uint32_t *near_in = b.q0, *near_out = b.q1, *far_cur = b.f0, *far_nxt = b.f1;
  if (bucketed) {  // the source starts in the far pile; the first split opens its bucket
    GMP_HIP(hipMemcpyAsync(far_cur, &src, 4, hipMemcpyHostToDevice, e.stream));
    GMP_HIP(hipMemcpyAsync(b.farflag + src, &one, 4, hipMemcpyHostToDevice, e.stream));
    nfar = 1;
  } else {
    GMP_HIP(hipMemcpyAsync(near_in, &src, 4, hipMemcpyHostToDevice, e.stream));
    nnear = 1;
    thr = kInf;
  }

  RelaxArgs<OffT> a{};
  a.off = ooff;
  a.idx = oidx;
  a.w = ow;
  a.dist = b.dist;
  a.near_stamp = b.stamp;
  a.far_flag = b.farflag;
  a.cnt = b.cnt;
  a.V = V;
  attach_hubs(e, a);
  int iters = 0, epochs = 0, pulls = 0;
  Counters h{};
  // rounds whose frontier has more than E >> delta_pull_log out-edges pull over the source strips
  const uint64_t pull_edges = e.opt.delta_pull_log ? (g.E >> e.opt.delta_pull_log) : ~0ull;
  if (pr) a.deg = g.work_deg();
  uint64_t front_deg = 0;  // out-edges of the current near frontier (when pull rounds are possible)

  auto relax_until_empty = [&](bool track) {
    while (nnear) {
      a.front = near_in;
      a.nfront = nnear;
      a.stamp = stamp++;
      a.near_out = near_out;
      a.threshold = thr;
      a.far_out = bucketed ? far_cur + nfar : nullptr;
      a.track_shards = track ? 1 : 0;
      GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 16, e.stream));  // near_cnt, far_cnt, near_deg
      if (pr && front_deg > pull_edges && !track) {
        // pull: the frontier's vertices carry the stamp of the launch that appended them
        PullRound q{};
        q.dist = b.dist;
        q.near_stamp = b.stamp;
        q.far_flag = b.farflag;
        q.near_out = near_out;
        q.far_out = a.far_out;
        q.near_cnt = &b.cnt->near_cnt;
        q.far_cnt = &b.cnt->far_cnt;
        q.near_deg = &b.cnt->near_deg;
        q.front = near_in;
        q.deg = a.deg;
        q.nfront = nnear;
        q.fstamp = a.stamp - 1u;
        q.next_stamp = a.stamp;
        q.thr = thr;
        pull_round(*pr, q);
        pulls++;
      } else {
        launch_relax<OffT>(e, mode, a);
      }
      read_counters(e, b.cnt, h);
      nnear = h.near_cnt;
      nfar += h.far_cnt;
      front_deg = h.near_deg;
      std::swap(near_in, near_out);
      iters++;
    }
  };
  auto split = [&](uint32_t thr_old, uint32_t thr_new) {
    GMP_HIP(hipMemsetAsync(&b.cnt->near_cnt, 0, 16, e.stream));
    far_split_kernel<<<grid_fill(e, nfar), kBlock, 0, e.stream>>>(far_cur, nfar, b.dist, thr_old, thr_new, b.farflag,
                                                                 b.stamp, stamp++, near_in, far_nxt, b.cnt, a.deg);
    GMP_CHECK_LAUNCH();
    read_counters(e, b.cnt, h);
    nnear = h.near_cnt;
    nfar = h.far_cnt;
    front_deg = h.near_deg;
    std::swap(far_cur, far_nxt);
    thr = thr_new;
  };
  auto width_end = [](uint32_t m, uint32_t c) {
    const uint64_t t = (uint64_t)m + (8ull << c);
    return t >= kInf ? kInf : (uint32_t)t;
  };
  // the finish action: every live far entry joins the frontier and hybrid_min relaxes to the
  // fixed point (plain pushes without the strip layout); the edge counter stays cumulative
  auto finish = [&]() {
    read_counters(e, b.cnt, h);
    const unsigned long long before = h.edges;
    split(thr, kInf);
    int r = 0, sw = 0;
    double fms = 0.0;
    uint64_t fed = 0;
    if (hybrid_min<OffT>(e, unit ? 1 : 2, 0u, near_in, nnear, r, sw, fms, fed)) {
      const unsigned long long tot_edges = before + fed;
      GMP_HIP(hipMemcpy(&b.cnt->edges, &tot_edges, 8, hipMemcpyHostToDevice));
      iters += r + sw;
      nnear = 0;
      nfar = 0;
    } else {
      relax_until_empty(false);
    }
  };
  // snapshot of the traversal state for trace exploration
  DevBuf sd, ss, sf, sfar, scnt;
  if (policy == kPolicyExplore) {
    sd.alloc((size_t)V * 4);
    ss.alloc((size_t)V * 4);
    sf.alloc((size_t)V * 4);
    sfar.alloc((size_t)V * 4);
    scnt.alloc(sizeof(Counters));
  }

  // smallest live far distance, or kInf when only settled entries are left
  auto far_min = [&]() {
    const uint32_t inf = kInf;
    GMP_HIP(hipMemcpyAsync(&b.cnt->far_min, &inf, 4, hipMemcpyHostToDevice, e.stream));
    far_min_kernel<<<grid_fill(e, nfar), kBlock, 0, e.stream>>>(far_cur, nfar, b.dist, thr, b.cnt);
    GMP_CHECK_LAUNCH();
    read_counters(e, b.cnt, h);
    return h.far_min;
  };
  // trace rollout: one epoch of width class c, then width 64 (class 3, the fixed schedule's) to the
  // end of the run
  auto run_rollout = [&](uint32_t m, uint32_t c) {
    if (c == kFinishClass && e.opt.learned_finish) {
      finish();
      return;
    }
    for (;;) {
      split(thr, width_end(m, c));
      relax_until_empty(false);
      if (nfar == 0) return;
      m = far_min();
      if (m == kInf) return;
      c = 3u;
    }
  };
  // explored epochs: the first four, then powers of two
  auto checkpoint = [](int ep) { return ep <= 4 || (ep & (ep - 1)) == 0; };
  uint32_t cur_class = 3;  // width 64 until the first exploration

  relax_until_empty(false);
  while (bucketed && nfar) {
    const uint32_t thr_old = thr;
    const uint32_t m = far_min();
    if (m == kInf) break;  // only settled entries were left
    epochs++;
    if (policy == kPolicyFixed) {
      const uint64_t t = ((uint64_t)m / delta + 1) * delta;
      split(thr_old, t >= kInf ? kInf : (uint32_t)t);
      relax_until_empty(false);
      continue;
    }
    const bool explore = policy == kPolicyExplore && checkpoint(epochs);
    if (policy == kPolicyLearned || explore) {
      GMP_HIP(hipMemsetAsync(b.cnt->hist, 0, sizeof(h.hist), e.stream));
      far_features_kernel<OffT><<<grid_fill(e, nfar), kBlock, 0, e.stream>>>(far_cur, nfar, b.dist, ooff, thr_old,
                                                                            b.cnt);
      GMP_CHECK_LAUNCH();
    }
    if (policy == kPolicyLearned) {
      policy_kernel<<<1, 64, 0, e.stream>>>(e.mlp.params.as<float>(), b.cnt);
      GMP_CHECK_LAUNCH();
      read_counters(e, b.cnt, h);
      cur_class = h.choice;
    }
    gmp_trace_row row{};
    if (explore) {
      // time every width class from this state to the end of the run
      read_counters(e, b.cnt, h);
      epoch_features(h.hist, h.edges, m, row.feat);
      row.epoch = (uint32_t)epochs;
      const uint32_t snap_nfar = nfar;
      GMP_HIP(hipMemcpyAsync(sd.p, b.dist, (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
      GMP_HIP(hipMemcpyAsync(ss.p, b.stamp, (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
      GMP_HIP(hipMemcpyAsync(sf.p, b.farflag, (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
      GMP_HIP(hipMemcpyAsync(sfar.p, far_cur, (size_t)snap_nfar * 4, hipMemcpyDeviceToDevice, e.stream));
      GMP_HIP(hipMemcpyAsync(scnt.p, b.cnt, sizeof(Counters), hipMemcpyDeviceToDevice, e.stream));
      auto restore = [&]() {
        GMP_HIP(hipMemcpyAsync(b.dist, sd.p, (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
        GMP_HIP(hipMemcpyAsync(b.stamp, ss.p, (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
        GMP_HIP(hipMemcpyAsync(b.farflag, sf.p, (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
        GMP_HIP(hipMemcpyAsync(far_cur, sfar.p, (size_t)snap_nfar * 4, hipMemcpyDeviceToDevice, e.stream));
        GMP_HIP(hipMemcpyAsync(b.cnt, scnt.p, sizeof(Counters), hipMemcpyDeviceToDevice, e.stream));
        nfar = snap_nfar;
        nnear = 0;
        thr = thr_old;
      };
      double best_ms = 0;
      const unsigned long long snap_edges = h.edges;
      for (uint32_t c = 0; c < (uint32_t)kMlpClasses; c++) {
        restore();
        EventPair tr;
        GMP_HIP(hipEventRecord(tr.a, e.stream));
        run_rollout(m, c);
        GMP_HIP(hipEventRecord(tr.b, e.stream));
        GMP_HIP(hipStreamSynchronize(e.stream));
        read_counters(e, b.cnt, h);
        row.ms[c] = std::max(1e-3f, tr.ms());
        row.relaxed[c] = (float)(h.edges - snap_edges);  // relaxations the rollout needed
        if (c == 0 || row.ms[c] < best_ms) {
          best_ms = row.ms[c];
          cur_class = c;
        }
      }
      restore();
      row.best = (int32_t)cur_class;
      GMP_HIP(hipMemsetAsync(&b.cnt->shard_mask, 0, 8, e.stream));
    }
    if (policy == kPolicyLearned) e.last_actions[cur_class]++;
    if (cur_class == kFinishClass && e.opt.learned_finish) {
      finish();
      if (explore) {  // the finish does not track shards: take them from the changed distances
        changed_shards_kernel<<<grid_fill(e, V), kBlock, 0, e.stream>>>(sd.as<uint32_t>(), b.dist, V, b.cnt);
        GMP_CHECK_LAUNCH();
      }
    } else {
      split(thr_old, width_end(m, cur_class));
      relax_until_empty(explore);
    }
    if (explore) {
      read_counters(e, b.cnt, h);
      row.shard_mask = h.shard_mask;
      if (rows && nrows && *nrows < cap) rows[(*nrows)++] = row;
    }
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  read_counters(e, b.cnt, h);
  const uint64_t pulled = pr ? sweep_edges(*pr) : 0;  // edges of the strips the pull rounds gathered
  finish_output(e, b.dist, dist_out);
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = iters;
    st->launches = iters;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = h.edges + pulled;
    st->bytes = st->edges * (unit ? 8 : 12);
    st->epochs = epochs;
    st->policy = policy;
    st->changed = (uint64_t)pulls;
  }
}

// Labels (work order, original-id values) into e.s_a.
template <class OffT>
void afforest(Engine &e, int &phases, double &ms, uint64_t &edges) {
  Graph &g = e.g;
  const uint32_t V = g.V;
  const CsrView w = g.work_in();  // symmetric: the in-lists are the out-lists
  const OffT *off = w.offs<OffT>();
  e.s_a.alloc((size_t)V * 4);
  e.s_b.alloc((size_t)V * 4);
  e.s_c.alloc((size_t)kAfProbe * 4 + 8);
  uint32_t *label = e.s_a.as<uint32_t>(), *p = e.s_b.as<uint32_t>();
  uint32_t *probe = e.s_c.as<uint32_t>();
  unsigned long long *dedges = reinterpret_cast<unsigned long long *>(probe + kAfProbe);
  const uint32_t gv = grid_fill(e, V);
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  af_init_kernel<<<gv, kBlock, 0, e.stream>>>(p, V);
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemsetAsync(dedges, 0, 8, e.stream));
  for (uint32_t r = 0; r < kAfSample; r++) {
    af_sample_kernel<OffT><<<gv, kBlock, 0, e.stream>>>(off, w.idx, V, r, p);
    af_compress_kernel<<<gv, kBlock, 0, e.stream>>>(p, V);
  }
  GMP_CHECK_LAUNCH();
  af_probe_kernel<<<(kAfProbe + kBlock - 1) / kBlock, kBlock, 0, e.stream>>>(p, V, probe);
  GMP_CHECK_LAUNCH();
  std::vector<uint32_t> roots(kAfProbe);
  GMP_HIP(hipMemcpyAsync(roots.data(), probe, kAfProbe * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  const uint32_t giant = af_giant(roots);
  af_finish_kernel<OffT><<<gv, kBlock, 0, e.stream>>>(off, w.idx, 0u, V, giant, p, dedges);
  af_compress_kernel<<<gv, kBlock, 0, e.stream>>>(p, V);
  // label first holds each root's least original id (slot = root), then every vertex takes its
  // root's: a root's slot is rewritten with its own value, other slots are never read as roots
  fill_u32_kernel<<<gv, kBlock, 0, e.stream>>>(label, V, kInf);
  af_min_kernel<<<gv, kBlock, 0, e.stream>>>(p, g.work_identity ? nullptr : g.n2o.as<uint32_t>(), V, giant, label);
  af_label_kernel<<<gv, kBlock, 0, e.stream>>>(p, label, V, label);
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  uint64_t h = 0;
  GMP_HIP(hipMemcpyAsync(&h, dedges, 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  ms = tot.ms();
  phases = (int)kAfSample + 1;
  edges = (uint64_t)V * kAfSample + h;
}

template <class OffT>
void cc_impl(Engine &e, uint32_t *label_out, gmp_stats *st) {
  Graph &g = e.g;
  if (g.symmetric && e.opt.cc_union_find) {
    ensure_work(e);
    int phases = 0;
    double ms = 0.0;
    uint64_t edges = 0;
    afforest<OffT>(e, phases, ms, edges);
    finish_output(e, e.s_a.as<uint32_t>(), label_out);
    if (st) {
      std::memset(st, 0, sizeof *st);
      st->iterations = phases;
      st->launches = phases;
      st->ms_total = ms;
      st->ms_kernel = ms;
      st->edges = edges;
      st->bytes = edges * 8;
    }
    return;
  }
  {
    // dense sweeps over the source-strip layout while many labels change (every sweep streams the
    // edges once: 2-byte codes, labels staged in LDS), frontier pushes for the sparse tail
    ensure_work(e);
    int rounds = 0, sweeps = 0;
    double ms = 0.0;
    uint64_t edges = 0;
    if (hybrid_min<OffT>(e, 0, 0u, nullptr, 0u, rounds, sweeps, ms, edges)) {
      finish_output(e, e.s_a.as<uint32_t>(), label_out);
      if (st) {
        std::memset(st, 0, sizeof *st);
        st->iterations = rounds + sweeps;
        st->launches = rounds + sweeps;
        st->ms_total = ms;
        st->ms_kernel = ms;
        st->edges = edges;
        st->bytes = edges * 8;
        st->epochs = sweeps;
      }
      return;
    }
  }
  ensure_work_out(e);
  const uint32_t V = g.V;
  ensure_hubs<OffT>(e);
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
  attach_hubs(e, a);
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
  if (e.g.semi) {  // every method gives the same distances; windows run Bellman-Ford sweeps
    semi_sssp(e, src, p.unit_weights != 0, dist_out, st);
    return;
  }
  int policy;
  switch (p.method) {
    case GMP_SSSP_BF: policy = kPolicyBF; break;
    case GMP_SSSP_DELTA: policy = kPolicyFixed; break;
    case GMP_SSSP_LEARNED: policy = kPolicyLearned; break;
    default: fail(GMP_EINVAL, "unknown SSSP method");
  }
  if (e.g.off64) sssp_impl<uint64_t>(e, src, p, policy, nullptr, 0, nullptr, dist_out, st);
  else sssp_impl<uint32_t>(e, src, p, policy, nullptr, 0, nullptr, dist_out, st);
}

void run_sssp_trace(Engine &e, uint32_t src, const gmp_sssp_params &p, gmp_trace_row *rows, uint32_t cap,
                    uint32_t *nrows, uint32_t *dist_out, gmp_stats *st) {
  if (src >= e.g.V) fail(GMP_EINVAL, "source vertex out of range");
  if (e.g.off64) sssp_impl<uint64_t>(e, src, p, kPolicyExplore, rows, cap, nrows, dist_out, st);
  else sssp_impl<uint32_t>(e, src, p, kPolicyExplore, rows, cap, nrows, dist_out, st);
}

void run_cc(Engine &e, uint32_t *label_out, gmp_stats *st) {
  if (e.g.semi) {
    semi_cc(e, label_out, st);
    return;
  }
  if (e.g.off64) cc_impl<uint64_t>(e, label_out, st);
  else cc_impl<uint32_t>(e, label_out, st);
}

}  // namespace gmp
