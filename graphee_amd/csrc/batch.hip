// Batched multi-instance runs: many graphs resident at once, many SSSP sources and
// personalised-PageRank seeds spread unevenly across them, each kind executed as one
// persistent launch per round whose rounds are captured in a HIP graph.
//
// The graphs are stored as one disjoint union (graph g owns union ids [vbase_g, vbase_g+V_g)),
// in-CSR for PageRank and out-CSR for SSSP.  An SSSP instance i on graph g owns the distance
// slots [dbase_i, dbase_i + V_g); a frontier entry is a slot index, so every instance's frontier
// lives in one queue and one relax launch advances all instances together.  Personalised
// PageRank instances on the same graph share the edge stream: their ranks are stored
// instance-minor ([v][k]) so one gathered source yields the contributions of all k instances.
//
// HIP has no conditional graph nodes, so SSSP captures two rounds (ping-pong queues, counters
// reset by a kernel -- no memset nodes) and the host replays the graph until the device
// frontier counter reads zero; PageRank captures all of its fixed iterations once.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "gmp_internal.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kInf = 0xFFFFFFFFu;

#define GRID_STRIDE32(i, n) \
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

// ---- SSSP ------------------------------------------------------------------------------

struct BatchSsspArgs {
  const uint64_t *off;        // union out-CSR offsets
  const uint32_t *idx;        // union ids
  const uint32_t *w;
  const uint32_t *inst_slot;  // [n+1] slot bases
  const uint32_t *inst_vbase; // [n] union vertex base of the instance's graph
  uint32_t ninst;
  uint32_t *dist;             // [total slots]
  uint32_t *stamp;            // [total slots]
  uint32_t *queue[2];
  uint32_t *count;            // [0]=queue0 size, [1]=queue1 size, [2]=round, [3]=pad
  unsigned long long *edges;
};

__device__ inline uint32_t instance_of(const uint32_t *slot_base, uint32_t n, uint32_t s) {
  uint32_t lo = 0, hi = n - 1;  // last i with slot_base[i] <= s
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (slot_base[mid] <= s) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One round for every instance: reads queue[round&1] (size count[round&1]), appends improved
// slots to the other queue, then the last workgroup to finish (ticket counter) flips the
// round and clears the consumed counter, so the next captured round needs no host work.
__global__ __launch_bounds__(kBlock) void batch_relax_kernel(BatchSsspArgs a, uint32_t *ticket) {
  __shared__ uint32_t s_start[kWavesPerBlock][kWave];
  __shared__ uint64_t s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_val[kWavesPerBlock][kWave];
  __shared__ uint32_t s_sbase[kWavesPerBlock][kWave];  // slot base minus union vertex base
  const uint32_t round = a.count[2];
  const uint32_t cur = round & 1u, nxt = cur ^ 1u;
  const uint32_t nfront = a.count[cur];
  const uint32_t *front = a.queue[cur];
  uint32_t *out = a.queue[nxt];
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nchunks = (nfront + kWave - 1) / kWave;
  unsigned long long edges = 0;
  for (uint32_t c = blockIdx.x * kWavesPerBlock + wv; c < nchunks; c += gridDim.x * kWavesPerBlock) {
    const uint32_t i = c * kWave + lane;
    const bool ok = i < nfront;
    uint64_t b = 0;
    uint32_t deg = 0, du = kInf, sb = 0;
    if (ok) {
      const uint32_t s = front[i];
      const uint32_t inst = instance_of(a.inst_slot, a.ninst, s);
      const uint32_t u = a.inst_vbase[inst] + (s - a.inst_slot[inst]);
      sb = a.inst_slot[inst] - a.inst_vbase[inst];
      b = a.off[u];
      deg = (uint32_t)(a.off[u + 1] - b);
      du = a.dist[s];
      if (du == kInf) deg = 0;
    }
    const uint32_t incl = wave_inclusive_scan(deg);
    const uint32_t total = (uint32_t)__shfl((int)incl, kWave - 1, kWave);
    s_start[wv][lane] = incl - deg;
    s_base[wv][lane] = b;
    s_val[wv][lane] = du;
    s_sbase[wv][lane] = sb;
    wave_sync();
    edges += total;
    for (uint32_t jb = 0; jb < total; jb += kWave) {
      const uint32_t j = jb + lane;
      bool app = false;
      uint32_t slot = 0;
      if (j < total) {
        uint32_t lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (s_start[wv][mid] <= j) lo = mid;
          else hi = mid - 1;
        }
        const uint64_t e = s_base[wv][lo] + (j - s_start[wv][lo]);
        slot = s_sbase[wv][lo] + a.idx[e];
        const uint64_t t = (uint64_t)s_val[wv][lo] + a.w[e];
        const uint32_t nd = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
        if (nd < a.dist[slot]) {
          const uint32_t old = atomicMin(&a.dist[slot], nd);
          if (nd < old) app = atomicExch(&a.stamp[slot], round + 1) != round + 1;
        }
      }
      const uint32_t pos = wave_append(app, &a.count[nxt]);
      if (app) out[pos] = slot;
    }
    wave_sync();
  }
  if (lane == 0 && edges) atomicAdd(a.edges, edges);
  // round epilogue by the last workgroup
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t t = atomicAdd(ticket, 1u);
    if (t == gridDim.x - 1) {
      a.count[cur] = 0;
      a.count[2] = round + 1;
      *ticket = 0;
      __threadfence();
    }
  }
}

__global__ void batch_sssp_init_kernel(uint32_t *dist, uint32_t *stamp, uint32_t total) {
  GRID_STRIDE32(i, total) {
    dist[i] = kInf;
    stamp[i] = 0;
  }
}

__global__ void batch_sssp_seed_kernel(const uint32_t *slots, uint32_t n, uint32_t *dist, uint32_t *queue,
                                       uint32_t *count) {
  GRID_STRIDE32(i, n) {
    dist[slots[i]] = 0;
    queue[i] = slots[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    count[0] = n;
    count[1] = 0;
    count[2] = 0;
  }
}

__global__ void batch_gather_out_kernel(const uint32_t *dist, const uint32_t *inst_slot, const uint32_t *perm_base,
                                        uint32_t ninst, uint32_t total, uint32_t *out) {
  GRID_STRIDE32(i, total) out[i] = dist[i];
  (void)inst_slot;
  (void)perm_base;
  (void)ninst;
}

// ---- personalised PageRank -------------------------------------------------------------

struct PprGroup {       // the instances of one graph
  uint32_t vbase;       // union vertex base
  uint32_t V;
  uint32_t k;           // instances on this graph
  uint32_t kk;          // k rounded up to a power of two, <= 64
  uint64_t rbase;       // base of this group's [V][k] rank/contrib block
  uint32_t first_inst;  // global instance id of column 0
  uint32_t pad;
};

struct PprTask {        // a run of rows of one group
  uint32_t group;
  uint32_t row0;
  uint32_t nrows;
  uint32_t pad;
};

struct PprArgs {
  const uint64_t *off;  // union in-CSR
  const uint32_t *idx;
  const uint32_t *outdeg;   // union out-degree
  const PprGroup *groups;
  const PprTask *tasks;
  uint32_t ntasks;
  const float *contrib_in;
  float *rank_out;
  float *contrib_out;
  const uint32_t *seed_local;  // per instance: seed (local id in its graph)
  const double *dang;          // per instance: dangling mass of the current ranks
  float d;
};

// Lanes are (edge slot, instance) pairs: lane l serves instance l % kk and edges
// row_begin + l / kk, stepping by 64 / kk; lanes of one instance are combined by a
// butterfly over the edge-slot bits.  Kahan per lane, hi/lo combine.
__global__ __launch_bounds__(kBlock) void ppr_gather_kernel(PprArgs a) {
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t t = blockIdx.x * kWavesPerBlock + wv; t < a.ntasks; t += nw) {
    const PprTask tk = a.tasks[t];
    const PprGroup g = a.groups[tk.group];
    const uint32_t inst = lane % g.kk, slot = lane / g.kk, step = kWave / g.kk;
    for (uint32_t r = tk.row0; r < tk.row0 + tk.nrows; r++) {
      const uint32_t u = g.vbase + r;
      const uint64_t e0 = a.off[u], e1 = a.off[u + 1];
      float s = 0, c = 0;
      if (inst < g.k) {
        for (uint64_t e = e0 + slot; e < e1; e += step) {
          const uint32_t src = a.idx[e] - g.vbase;
          kahan_add(s, c, a.contrib_in[g.rbase + (uint64_t)src * g.k + inst]);
        }
      }
      float hi = s, lo = -c;
      for (uint32_t off = kWave >> 1; off >= g.kk; off >>= 1) {
        const float oh = __shfl_xor(hi, (int)off, kWave), ol = __shfl_xor(lo, (int)off, kWave);
        dd_add(hi, lo, oh, ol);
      }
      if (slot == 0 && inst < g.k) {
        const uint32_t gi = g.first_inst + inst;
        float v = a.d * hi;
        if (r == a.seed_local[gi]) v += (float)((1.0 - (double)a.d) + (double)a.d * a.dang[gi]);
        const uint64_t pos = g.rbase + (uint64_t)r * g.k + inst;
        a.rank_out[pos] = v;
        const uint32_t od = a.outdeg[u];
        a.contrib_out[pos] = od ? v / (float)od : 0.0f;
      }
    }
  }
}

// Per instance, the FP64 mass on dangling vertices (fixed-order block reduction).
__global__ __launch_bounds__(kBlock) void ppr_dangling_kernel(const PprGroup *groups, const uint32_t *group_of,
                                                              const uint32_t *dlist, const uint32_t *dbase,
                                                              const float *rank, double *dang) {
  __shared__ double sh[kBlock];
  const uint32_t i = blockIdx.x;
  const PprGroup g = groups[group_of[i]];
  const uint32_t col = i - g.first_inst;
  const uint32_t b = dbase[group_of[i]], e = dbase[group_of[i] + 1];
  double acc = 0;
  for (uint32_t j = b + threadIdx.x; j < e; j += kBlock) acc += (double)rank[g.rbase + (uint64_t)dlist[j] * g.k + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) dang[i] = sh[0];
}

__global__ void ppr_init_kernel(const PprGroup *groups, const uint32_t *group_of, const uint32_t *seed_local,
                                const uint32_t *outdeg, uint32_t ninst, float *rank, float *contrib) {
  // one block per instance: e_seed and its contribution
  const uint32_t i = blockIdx.x;
  if (i >= ninst) return;
  const PprGroup g = groups[group_of[i]];
  const uint32_t col = i - g.first_inst;
  for (uint32_t v = threadIdx.x; v < g.V; v += blockDim.x) {
    const float r = v == seed_local[i] ? 1.0f : 0.0f;
    const uint64_t pos = g.rbase + (uint64_t)v * g.k + col;
    rank[pos] = r;
    const uint32_t od = outdeg[g.vbase + v];
    contrib[pos] = od ? r / (float)od : 0.0f;
  }
}

}  // namespace

// ---- host side -------------------------------------------------------------------------

struct BatchGraphs {
  std::vector<uint32_t> vbase;  // [G+1]
  std::vector<uint64_t> ebase;  // [G+1]
  DevBuf in_off, in_idx, in_w, out_off, out_idx, out_w, outdeg;
  uint64_t E = 0;
  uint32_t V = 0;
  bool weighted = false;
};

namespace {
BatchGraphs *g_batch_of(Engine &e);
}

}  // namespace gmp
