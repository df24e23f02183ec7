// Batched multi-instance runs: many graphs resident at once, many SSSP sources and
// personalised-PageRank seeds spread unevenly across them, each kind advanced by one
// persistent launch per round, with the rounds captured in a HIP graph.
//
// The graphs are stored as one disjoint union (graph k owns union ids [vbase_k, vbase_{k+1})),
// in-CSR for PageRank and out-CSR for SSSP.  An SSSP instance i on graph k owns the distance
// slots [slot_i, slot_i + V_k); a frontier entry is a slot, so the frontiers of all instances
// share one queue and one relax launch advances every instance.  Personalised-PageRank
// instances on the same graph share the edge stream: their ranks are stored instance-minor
// ([v][k]) so one gathered source yields the contributions of all k instances.
//
// HIP graphs have no conditional nodes, so SSSP captures two rounds (ping-pong queues; the last
// workgroup of each round flips the round counter and clears the consumed queue's count --
// no memset nodes) and the host replays the graph until the frontier is empty; PageRank runs a
// fixed number of iterations and captures all of them once.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "gmp_internal.hpp"
#include "pr_common.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kInf = 0xFFFFFFFFu;

#define GRID_STRIDE32(i, n) \
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

// ---- union construction ----------------------------------------------------------------

template <class OffT>
__global__ void shift_offsets_kernel(const OffT *off, uint32_t V, uint64_t ebase, uint64_t *out) {
  GRID_STRIDE32(r, V + 1) out[r] = (uint64_t)off[r] + ebase;
}

// out[vbase + i] = vbase + map[i] (map null: the identity).
__global__ void shift_map_kernel(const uint32_t *map, uint32_t n, uint32_t vbase, uint32_t *out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    out[vbase + i] = vbase + (map ? map[i] : i);
}

__global__ void shift_ids_kernel(const uint32_t *idx, uint64_t E, uint32_t vbase, uint32_t *out) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E; j += (uint64_t)gridDim.x * blockDim.x)
    out[j] = idx[j] + vbase;
}

inline uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 1u << 20));
}

// ---- SSSP ------------------------------------------------------------------------------

struct BatchSsspArgs {
  const uint64_t *off;         // union out-CSR offsets
  const uint32_t *idx;         // union ids
  const uint32_t *w;
  const uint32_t *inst_slot;   // [n+1] slot bases
  const uint32_t *inst_vbase;  // [n] union vertex base of the instance's graph
  uint32_t ninst;
  uint32_t *dist;              // [total slots]
  uint32_t *stamp;             // [total slots]
  uint32_t *queue0, *queue1;
  uint32_t *count;             // [0]=|queue0|, [1]=|queue1|, [2]=round
  uint32_t *ticket;
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

constexpr int kRelaxUnroll = 4;    // edges per lane in flight in the relax loop
constexpr uint32_t kLdsInst = 1024;  // instance slot bases kept in LDS up to this many instances
constexpr uint32_t kAppendCap = 512;  // staged appends per wave

// One round for every instance: reads queue[round&1] and appends improved slots to the other
// queue; the last workgroup to finish flips the round and clears the consumed count.  A wave
// takes 64 frontier slots and expands their edges kRelaxUnroll per lane at a time: the edge,
// weight and distance loads of those edges are issued together (edges past the end reload the
// wave's first edge and are dropped), then the improving ones are applied with atomics.
__global__ __launch_bounds__(kBlock) void batch_relax_kernel(BatchSsspArgs a) {
  __shared__ uint32_t s_start[kWavesPerBlock][kWave];
  __shared__ uint64_t s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_val[kWavesPerBlock][kWave];
  __shared__ uint32_t s_sb[kWavesPerBlock][kWave];  // slot base minus union vertex base
  __shared__ uint32_t s_islot[kLdsInst];
  __shared__ uint32_t s_app[kWavesPerBlock][kAppendCap];
  const bool lds_inst = a.ninst <= kLdsInst;
  if (lds_inst)
    for (uint32_t i = threadIdx.x; i < a.ninst; i += blockDim.x) s_islot[i] = a.inst_slot[i];
  __syncthreads();
  const uint32_t round = a.count[2];
  const uint32_t cur = round & 1u, nxt = cur ^ 1u;
  const uint32_t nfront = a.count[cur];
  const uint32_t *front = cur ? a.queue1 : a.queue0;
  uint32_t *out = cur ? a.queue0 : a.queue1;
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nchunks = (nfront + kWave - 1) / kWave;
  unsigned long long edges = 0;
  WaveAppend wa{s_app[wv], 0u};
  for (uint32_t c = blockIdx.x * kWavesPerBlock + wv; c < nchunks; c += gridDim.x * kWavesPerBlock) {
    const uint32_t i = c * kWave + lane;
    uint64_t b = 0;
    uint32_t deg = 0, du = kInf, sb = 0;
    if (i < nfront) {
      const uint32_t s = front[i];
      const uint32_t inst = lds_inst ? instance_of(s_islot, a.ninst, s) : instance_of(a.inst_slot, a.ninst, s);
      const uint32_t u = a.inst_vbase[inst] + (s - a.inst_slot[inst]);
      sb = a.inst_slot[inst] - a.inst_vbase[inst];
      b = a.off[u];
      deg = (uint32_t)(a.off[u + 1] - b);
      du = a.dist[s];
      // a slot whose distance already dropped again in this round is queued for the next one,
      // where it relaxes its edges from the lower distance: skip it here
      if (du == kInf || a.stamp[s] == round + 1) deg = 0;
    }
    const uint32_t incl = wave_scan_dpp(deg);
    const uint32_t total = read_lane_u32(incl, kWave - 1);
    s_start[wv][lane] = incl - deg;
    s_base[wv][lane] = b;
    s_val[wv][lane] = du;
    s_sb[wv][lane] = sb;
    wave_sync();
    edges += total;
    for (uint32_t jb = 0; jb < total; jb += kWave * kRelaxUnroll) {
      uint64_t e[kRelaxUnroll];
      uint32_t src[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        const uint32_t j = jb + lane + (uint32_t)q * kWave;
        const uint32_t jj = j < total ? j : 0u;
        // the frontier lane holding edge jj: the last lane whose first edge is <= jj
        uint32_t lo = 0, hi = kWave - 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (s_start[wv][mid] <= jj) lo = mid;
          else hi = mid - 1;
        }
        src[q] = lo;
        e[q] = s_base[wv][lo] + (jj - s_start[wv][lo]);
      }
      uint32_t to[kRelaxUnroll], wt[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        to[q] = a.idx[e[q]];
        wt[q] = a.w[e[q]];
      }
      uint32_t slot[kRelaxUnroll], nd[kRelaxUnroll], cd[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        slot[q] = s_sb[wv][src[q]] + to[q];
        const uint64_t t = (uint64_t)s_val[wv][src[q]] + wt[q];
        nd[q] = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
        cd[q] = a.dist[slot[q]];
      }
      // nd < cd (read this round) means the slot's distance drops this round, by this lane or
      // by another one, so the slot belongs in the next frontier whichever atomicMin wins; the
      // stamp exchange lets exactly one lane append it.  All atomics are issued before any
      // result is read.
      bool imp[kRelaxUnroll];
      uint32_t old[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        imp[q] = jb + lane + (uint32_t)q * kWave < total && nd[q] < cd[q];
        if (imp[q]) {
          atomicMin(&a.dist[slot[q]], nd[q]);
          old[q] = atomicExch(&a.stamp[slot[q]], round + 1);
        }
      }
      uint32_t app = 0;
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++)
        if (imp[q] && old[q] != round + 1) app |= 1u << q;
      // the lane's improved slots, staged in the wave's LDS buffer
      const uint32_t cnt = (uint32_t)__popc(app);
      const uint32_t inc = wave_scan_dpp(cnt);
      const uint32_t tot = read_lane_u32(inc, kWave - 1);
      if (tot) {
        uint32_t pos = wa.n + inc - cnt;
#pragma unroll
        for (int q = 0; q < kRelaxUnroll; q++)
          if ((app >> q) & 1u) wa.buf[pos++] = slot[q];
        wa.n += tot;
        if (wa.n > kAppendCap - kWave * kRelaxUnroll) append_flush(wa, &a.count[nxt], out);
      }
    }
    wave_sync();
  }
  append_flush(wa, &a.count[nxt], out);
  if (lane == 0 && edges) atomicAdd(a.edges, edges);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t t = atomicAdd(a.ticket, 1u);
    if (t == gridDim.x - 1) {  // every workgroup has read count[cur]
      a.count[cur] = 0;
      a.count[2] = round + 1;
      *a.ticket = 0;
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

__global__ void batch_sssp_seed_kernel(const uint32_t *slots, uint32_t n, uint32_t *dist, uint32_t *queue0,
                                       uint32_t *count, uint32_t *ticket, unsigned long long *edges) {
  GRID_STRIDE32(i, n) {
    dist[slots[i]] = 0;
    queue0[i] = slots[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    count[0] = n;
    count[1] = 0;
    count[2] = 0;
    *ticket = 0;
    *edges = 0;
  }
}

// ---- personalised PageRank -------------------------------------------------------------
//
// A group holds up to 64 instances of one graph; lane l of a wave serves instance l % kk of
// the group (kk = instances rounded up to a power of two) in slot l / kk, so a wave runs
// 64 / kk slots.  Rows of in-degree up to kHubDegree are handed out in runs, one row per slot
// at a time; a slot walks its row's in-edges with kLoads source loads in flight per lane and a
// Kahan sum per lane, and writes the row's k ranks as one contiguous vector.  Rows above
// kHubDegree are cut into pieces of kHubDegree edges spread over the slots of a wave; each
// piece leaves an FP64 partial per instance, and ppr_hub_kernel adds a row's pieces in order.
// Waves take tasks (runs and pieces of every group) from one queue, so groups of any size and
// instance count share the chip without a cost model.  The dangling mass of the new ranks
// (needed by the next iteration's teleport) is summed per instance in an exact 61-bit fixed
// point: the sum does not depend on which wave took which task, so results are reproducible.

constexpr uint32_t kHubDegree = 1024;  // rows above this are split into pieces of this size
constexpr int kLoads = 8;              // source loads in flight per lane
constexpr uint32_t kRunRows = 32;      // rows per slot in a row run
constexpr uint32_t kRunEdges = 2048;   // edges per slot in a row run
constexpr uint32_t kCols = 64;         // partials are stored 64 instances wide

struct PprGroup {      // up to 64 instances on one graph
  uint32_t vbase;      // union vertex base
  uint32_t V;
  uint32_t k;          // instances in the group
  uint32_t kk;         // k rounded up to a power of two
  uint64_t rbase;      // base of the group's [V][k] block of ranks
  uint32_t first;      // group-order id of instance column 0
  uint32_t pad;
};

struct PprTask {       // a run of rows (part == kNoPiece) or one piece of a hub row
  uint32_t group;
  uint32_t row;        // first row of the run, or the hub row (group-local)
  uint32_t n;          // rows of the run, or the piece number
  uint32_t part;       // hub pieces: partial slot
};
constexpr uint32_t kNoPiece = 0xFFFFFFFFu;

struct PprHub {        // a row above kHubDegree
  uint32_t group;
  uint32_t row;        // group-local
  uint32_t part0;      // partial slots [part0, part0 + nparts)
  uint32_t nparts;
};

struct PprArgs {
  const uint64_t *off;  // union in-CSR
  const uint32_t *idx;
  const uint32_t *outdeg;
  const PprGroup *groups;
  const PprTask *tasks;
  uint32_t ntasks;
  uint32_t *next;       // task queue head (cleared before each gather)
  const PprHub *hubs;
  uint32_t nhubs;
  const float *contrib_in;
  float *rank_out;
  float *contrib_out;
  const uint32_t *seed_local;       // per group-order instance: seed (local id)
  const unsigned long long *dang;   // per group-order instance: dangling mass of the current ranks
  unsigned long long *dang_next;    // ... and of the new ranks (fixed point, cleared beforehand)
  double *part;                     // [parts][kCols]: FP64 sums of hub pieces
  float d;
};

// Kahan-sums contrib_in over the in-edges [e, e1) of a row, taking every step-th edge.  The
// loads are unconditional -- an edge past e1 reloads edge e, a lane without an instance reads
// column 0 -- so all kLoads of a batch are in flight together; the extra values are dropped.
__device__ inline void ppr_sum_edges(const PprArgs &a, const PprGroup &g, uint32_t inst, bool on, uint64_t e,
                                     uint64_t e1, uint32_t step, float &s, float &c) {
  const float *col = a.contrib_in + g.rbase + (on ? inst : 0u);
  for (; e < e1; e += (uint64_t)kLoads * step) {
    uint32_t src[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; u++) {
      const uint64_t j = e + (uint64_t)u * step;
      src[u] = a.idx[j < e1 ? j : e];
    }
    float x[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; u++) x[u] = col[(uint64_t)(src[u] - g.vbase) * g.k];
#pragma unroll
    for (int u = 0; u < kLoads; u++)
      if (e + (uint64_t)u * step < e1) kahan_add(s, c, x[u]);
  }
}

// New rank of row r (group-local, out-degree od) from its gathered sum; returns the mass it
// adds to the dangling sum, in fixed point.  tele = (1 - d) + d * (dangling mass), added at
// the instance's seed.
__device__ inline unsigned long long ppr_finish_row(const PprArgs &a, const PprGroup &g, uint32_t inst, uint32_t r,
                                                    float sum, uint32_t seed, float tele, uint32_t od) {
  float v = a.d * sum;
  if (r == seed) v += tele;
  const uint64_t pos = g.rbase + (uint64_t)r * g.k + inst;
  a.rank_out[pos] = v;
  a.contrib_out[pos] = od ? v / (float)od : 0.0f;
  return od ? 0ull : fix_from_float(v);
}

// Adds a wave's fixed-point dangling sums (one per lane) of group g into dang_next: the slots
// of an instance first, then one atomic per instance.
__device__ inline void ppr_flush_dangling(const PprArgs &a, const PprGroup &g, unsigned long long &acc) {
  const uint32_t lane = lane_id();
  for (uint32_t o = kWave >> 1; o >= g.kk; o >>= 1) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)acc, (int)o, kWave);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(acc >> 32), (int)o, kWave);
    acc += ((unsigned long long)hi << 32) | lo;
  }
  if (lane < g.k && acc) atomicAdd(&a.dang_next[g.first + lane], acc);
  acc = 0ull;
}

__global__ __launch_bounds__(kBlock) void ppr_gather_kernel(PprArgs a) {
  const uint32_t lane = lane_id();
  uint32_t gid = 0xFFFFFFFFu;
  PprGroup g{};
  uint32_t inst = 0, slot = 0, S = 1, seed = kInf;
  bool on = false;
  float tele = 0.0f;
  unsigned long long dacc = 0ull;  // fixed-point dangling mass of this lane's instance
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.next, 1u);
    t = (uint32_t)__shfl((int)t, 0, kWave);
    if (t >= a.ntasks) break;
    const PprTask tk = a.tasks[t];
    if (tk.group != gid) {
      if (gid != 0xFFFFFFFFu) ppr_flush_dangling(a, g, dacc);
      gid = tk.group;
      g = a.groups[gid];
      inst = lane & (g.kk - 1);
      slot = lane / g.kk;
      S = kWave / g.kk;
      on = inst < g.k;
      seed = on ? a.seed_local[g.first + inst] : kInf;
      tele = on ? (float)((1.0 - (double)a.d) + (double)a.d * fix_to_double(a.dang[g.first + inst])) : 0.0f;
    }
    if (tk.part == kNoPiece) {
      // a slot's next row (clamped into the run, so the loads need no branch) is looked up
      // while the current row's edges are summed
      const uint32_t rend = tk.row + tk.n, rlast = g.vbase + rend - 1;
      uint32_t q = min(g.vbase + tk.row + slot, rlast);
      uint64_t e0 = a.off[q], e1 = a.off[q + 1];
      uint32_t od = a.outdeg[q];
      for (uint32_t r = tk.row + slot; r < rend; r += S) {
        q = min(g.vbase + r + S, rlast);
        const uint64_t n0 = a.off[q], n1 = a.off[q + 1];
        const uint32_t nod = a.outdeg[q];
        float s = 0.0f, c = 0.0f;
        ppr_sum_edges(a, g, inst, on, e0, e1, 1, s, c);
        if (on) dacc += ppr_finish_row(a, g, inst, r, s - c, seed, tele, od);
        e0 = n0;
        e1 = n1;
        od = nod;
      }
    } else {
      const uint32_t u = g.vbase + tk.row;
      const uint64_t e0 = a.off[u] + (uint64_t)tk.n * kHubDegree;
      const uint64_t e1 = min(a.off[u + 1], e0 + kHubDegree);
      float s = 0.0f, c = 0.0f;
      ppr_sum_edges(a, g, inst, on, e0 + slot, e1, S, s, c);
      // the slots' sums of one instance, combined in a fixed butterfly order
      float hi = s, lo = -c;
      for (uint32_t o = kWave >> 1; o >= g.kk; o >>= 1) {
        const float oh = __shfl_xor(hi, (int)o, kWave), ol = __shfl_xor(lo, (int)o, kWave);
        dd_add(hi, lo, oh, ol);
      }
      if (slot == 0 && on) a.part[(size_t)tk.part * kCols + inst] = (double)hi + (double)lo;
    }
  }
  if (gid != 0xFFFFFFFFu) ppr_flush_dangling(a, g, dacc);
}

// One wave per hub row: its pieces' partials in piece order, then the rank update.
__global__ __launch_bounds__(kBlock) void ppr_hub_kernel(PprArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t nw = gridDim.x * kWavesPerBlock;
  for (uint32_t h = blockIdx.x * kWavesPerBlock + wave_id(); h < a.nhubs; h += nw) {
    const PprHub hb = a.hubs[h];
    const PprGroup g = a.groups[hb.group];
    if (lane < g.k) {
      double s = 0.0;
      for (uint32_t p = 0; p < hb.nparts; p++) s += a.part[(size_t)(hb.part0 + p) * kCols + lane];
      const uint32_t gi = g.first + lane;
      const float tele = (float)((1.0 - (double)a.d) + (double)a.d * fix_to_double(a.dang[gi]));
      const unsigned long long dm =
          ppr_finish_row(a, g, lane, hb.row, (float)s, a.seed_local[gi], tele, a.outdeg[g.vbase + hb.row]);
      if (dm) atomicAdd(&a.dang_next[gi], dm);
    }
  }
}

// Clears the task queue and the dangling sums the next gather accumulates into.
__global__ void ppr_reset_kernel(uint32_t *next, unsigned long long *dang_next, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dang_next[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) *next = 0u;
}

// r = e_seed, c = r / outdeg over the group's [V][k] block (blockIdx.y = group), and the
// dangling mass of r per instance.
__global__ void ppr_init_kernel(const PprGroup *groups, const uint32_t *seed_local, const uint32_t *outdeg,
                                float *rank, float *contrib, unsigned long long *dang) {
  const PprGroup g = groups[blockIdx.y];
  const uint64_t total = (uint64_t)g.V * g.k;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(i / g.k), col = (uint32_t)(i % g.k);
    const float r = v == seed_local[g.first + col] ? 1.0f : 0.0f;
    rank[g.rbase + i] = r;
    const uint32_t od = outdeg[g.vbase + v];
    contrib[g.rbase + i] = od ? r / (float)od : 0.0f;
  }
  if (blockIdx.x == 0 && threadIdx.x < g.k) {
    const uint32_t gi = g.first + threadIdx.x;
    dang[gi] = outdeg[g.vbase + seed_local[gi]] ? 0ull : fix_from_float(1.0f);
  }
}

struct GraphExecGuard {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GraphExecGuard() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

}  // namespace

void batch_generate(Engine &e, uint32_t ngraphs, const gmp_rmat_params *params) {
  if (ngraphs == 0) fail(GMP_EINVAL, "batch needs at least one graph");
  BatchGraphs &b = e.batch;
  b = BatchGraphs();
  Graph saved = std::move(e.g);
  struct Restore {
    Engine &e;
    Graph &saved;
    ~Restore() { e.g = std::move(saved); }
  } restore{e, saved};
  b.vbase.assign(ngraphs + 1, 0);
  b.ebase.assign(ngraphs + 1, 0);
  // each graph in its degree-ordered layout (hubs' values share cache lines in the PageRank
  // gathers) for the in-CSR, and in its own numbering for the out-CSR of the SSSP pushes (spread
  // ids keep the distance atomics of hub neighbours off shared lines)
  std::vector<Csr> parts(ngraphs), orig(ngraphs);
  std::vector<DevBuf> maps(ngraphs);
  std::vector<bool> off64(ngraphs);
  bool weighted = true, relabelled = false;
  for (uint32_t k = 0; k < ngraphs; k++) {
    graph_generate_rmat(e, params[k]);
    if ((uint64_t)b.vbase[k] + e.g.V >= 0xFFFFFFFFull) fail(GMP_EUNSUPPORTED, "batch exceeds 2^32 vertices");
    b.vbase[k + 1] = b.vbase[k] + e.g.V;
    b.ebase[k + 1] = b.ebase[k] + e.g.E;
    weighted = weighted && e.g.weighted;
    off64[k] = e.g.off64;
    graph_finish(e, 1, nullptr);
    ensure_work(e);
    if (e.g.work_identity) {
      parts[k] = std::move(e.g.in);
    } else {
      relabelled = true;
      parts[k] = std::move(e.g.win);
      orig[k] = std::move(e.g.in);
      maps[k] = std::move(e.g.o2n);
    }
  }
  b.V = b.vbase[ngraphs];
  b.E = b.ebase[ngraphs];
  b.weighted = weighted;
  b.o2n.alloc(std::max<size_t>(b.V, 1) * 4);
  for (uint32_t k = 0; k < ngraphs; k++) {
    const uint32_t Vk = b.vbase[k + 1] - b.vbase[k];
    if (Vk) {
      shift_map_kernel<<<grid_for(Vk), kBlock, 0, e.stream>>>(maps[k].as<uint32_t>(), Vk, b.vbase[k],
                                                              b.o2n.as<uint32_t>());
      GMP_CHECK_LAUNCH();
    }
  }
  b.h_o2n.resize(b.V);
  GMP_HIP(hipMemcpyAsync(b.h_o2n.data(), b.o2n.p, (size_t)b.V * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  maps.clear();
  // the union of per-graph CSRs, ids shifted by the graphs' bases (the parts are released)
  auto concat = [&](std::vector<Csr> &from, Csr &to) {
    to.off.alloc(((size_t)b.V + 1) * 8);
    to.idx.alloc(std::max<uint64_t>(b.E, 1) * 4);
    if (weighted) to.w.alloc(std::max<uint64_t>(b.E, 1) * 4);
    for (uint32_t k = 0; k < ngraphs; k++) {
      const uint32_t Vk = b.vbase[k + 1] - b.vbase[k];
      const uint64_t Ek = b.ebase[k + 1] - b.ebase[k];
      uint64_t *o = to.off.as<uint64_t>() + b.vbase[k];
      if (off64[k]) shift_offsets_kernel<uint64_t><<<grid_for(Vk + 1), kBlock, 0, e.stream>>>(from[k].off.as<uint64_t>(), Vk, b.ebase[k], o);
      else shift_offsets_kernel<uint32_t><<<grid_for(Vk + 1), kBlock, 0, e.stream>>>(from[k].off.as<uint32_t>(), Vk, b.ebase[k], o);
      GMP_CHECK_LAUNCH();
      if (Ek) {
        shift_ids_kernel<<<grid_for(Ek), kBlock, 0, e.stream>>>(from[k].idx.as<uint32_t>(), Ek, b.vbase[k],
                                                               to.idx.as<uint32_t>() + b.ebase[k]);
        GMP_CHECK_LAUNCH();
        if (weighted)
          GMP_HIP(hipMemcpyAsync(to.w.as<uint32_t>() + b.ebase[k], from[k].w.p, Ek * 4, hipMemcpyDeviceToDevice,
                                 e.stream));
      }
      GMP_HIP(hipStreamSynchronize(e.stream));
      from[k] = Csr();
    }
    to.valid = true;
  };
  concat(parts, b.in);
  if (relabelled) {
    for (uint32_t k = 0; k < ngraphs; k++)
      if (!orig[k].idx.p && b.ebase[k + 1] > b.ebase[k]) fail(GMP_ESTATE, "batch: mixed graph layouts");
    Csr whole;
    concat(orig, whole);
    transpose_csr(e, CsrView{whole.off.p, whole.idx.as<uint32_t>(), whole.w.as<uint32_t>()}, b.V, b.E, true, weighted,
                  b.out);
  } else {
    transpose_csr(e, CsrView{b.in.off.p, b.in.idx.as<uint32_t>(), b.in.w.as<uint32_t>()}, b.V, b.E, true, weighted,
                  b.out);
  }
  b.outdeg.alloc(std::max<size_t>(b.V, 1) * 4);
  count_degrees(e, b.in.idx.as<uint32_t>(), b.E, b.V, b.outdeg.as<uint32_t>());
  b.h_off.resize((size_t)b.V + 1);
  GMP_HIP(hipMemcpy(b.h_off.data(), b.in.off.p, ((size_t)b.V + 1) * 8, hipMemcpyDeviceToHost));
  b.ready = true;
}

void batch_sssp(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *source, uint32_t *dist_out,
                gmp_stats *st) {
  BatchGraphs &b = e.batch;
  if (!b.ready) fail(GMP_ESTATE, "no batch loaded");
  if (!b.weighted) fail(GMP_EUNSUPPORTED, "batched SSSP needs weighted graphs");
  if (n == 0 || !graph || !source) fail(GMP_EINVAL, "batched SSSP needs instances");
  const uint32_t G = (uint32_t)b.vbase.size() - 1;
  std::vector<uint32_t> slot(n + 1, 0), vb(n), seeds(n);
  for (uint32_t i = 0; i < n; i++) {
    if (graph[i] >= G) fail(GMP_EINVAL, "instance graph out of range");
    const uint32_t Vk = b.vbase[graph[i] + 1] - b.vbase[graph[i]];
    if (source[i] >= Vk) fail(GMP_EINVAL, "instance source out of range");
    if ((uint64_t)slot[i] + Vk >= 0xFFFFFFFFull) fail(GMP_EUNSUPPORTED, "batch exceeds 2^32 distance slots");
    slot[i + 1] = slot[i] + Vk;
    vb[i] = b.vbase[graph[i]];
    seeds[i] = slot[i] + source[i];
  }
  const uint32_t T = slot[n];
  DevBuf d_slot, d_vb, d_seed, dist, stamp, q0, q1, cnt;
  d_slot.alloc((n + 1) * 4);
  d_vb.alloc(n * 4);
  d_seed.alloc(n * 4);
  dist.alloc((size_t)T * 4);
  stamp.alloc((size_t)T * 4);
  q0.alloc((size_t)T * 4);
  q1.alloc((size_t)T * 4);
  cnt.alloc(64);
  GMP_HIP(hipMemcpyAsync(d_slot.p, slot.data(), (n + 1) * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_vb.p, vb.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_seed.p, seeds.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  BatchSsspArgs a{};
  a.off = b.out.off.as<uint64_t>();
  a.idx = b.out.idx.as<uint32_t>();
  a.w = b.out.w.as<uint32_t>();
  a.inst_slot = d_slot.as<uint32_t>();
  a.inst_vbase = d_vb.as<uint32_t>();
  a.ninst = n;
  a.dist = dist.as<uint32_t>();
  a.stamp = stamp.as<uint32_t>();
  a.queue0 = q0.as<uint32_t>();
  a.queue1 = q1.as<uint32_t>();
  a.count = cnt.as<uint32_t>();
  a.ticket = cnt.as<uint32_t>() + 4;
  a.edges = reinterpret_cast<unsigned long long *>(cnt.as<uint32_t>() + 8);
  const uint32_t grid = (uint32_t)e.num_cus * 4;

  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  batch_sssp_init_kernel<<<grid_for(T), kBlock, 0, e.stream>>>(a.dist, a.stamp, T);
  batch_sssp_seed_kernel<<<grid_for(n), kBlock, 0, e.stream>>>(d_seed.as<uint32_t>(), n, a.dist, a.queue0, a.count,
                                                              a.ticket, a.edges);
  GMP_CHECK_LAUNCH();
  // capture two rounds; the counters live on the device, so replays need no host input
  GraphExecGuard ge;
  GMP_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
  batch_relax_kernel<<<grid, kBlock, 0, e.stream>>>(a);
  batch_relax_kernel<<<grid, kBlock, 0, e.stream>>>(a);
  GMP_HIP(hipStreamEndCapture(e.stream, &ge.graph));
  GMP_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
  uint32_t h[4] = {0, 0, 0, 0};
  int launches = 0;
  for (;;) {
    GMP_HIP(hipGraphLaunch(ge.exec, e.stream));
    launches++;
    GMP_HIP(hipMemcpyAsync(e.h_pinned, a.count, 16, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    std::memcpy(h, e.h_pinned, 16);
    if (h[h[2] & 1u] == 0) break;
    if (launches > (1 << 22)) fail(GMP_ESTATE, "batched SSSP did not converge");
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  unsigned long long edges = 0;
  GMP_HIP(hipMemcpy(&edges, a.edges, 8, hipMemcpyDeviceToHost));
  if (dist_out) GMP_HIP(hipMemcpy(dist_out, a.dist, (size_t)T * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = (int32_t)h[2];
    st->launches = launches;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = edges;
    st->bytes = edges * 12;
  }
}

void batch_ppr(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *seed, float damping, int32_t iters,
               float *rank_out, gmp_stats *st) {
  BatchGraphs &b = e.batch;
  if (!b.ready) fail(GMP_ESTATE, "no batch loaded");
  if (n == 0 || !graph || !seed) fail(GMP_EINVAL, "batched PageRank needs instances");
  if (!(damping >= 0.0f && damping <= 1.0f) || iters < 0) fail(GMP_EINVAL, "bad damping or iteration count");
  const uint32_t G = (uint32_t)b.vbase.size() - 1;
  // group instances by graph, at most 64 per group; group order keeps caller order per graph
  std::vector<std::vector<uint32_t>> per(G);
  for (uint32_t i = 0; i < n; i++) {
    if (graph[i] >= G) fail(GMP_EINVAL, "instance graph out of range");
    if (seed[i] >= b.vbase[graph[i] + 1] - b.vbase[graph[i]]) fail(GMP_EINVAL, "instance seed out of range");
    per[graph[i]].push_back(i);
  }
  std::vector<PprGroup> groups;
  std::vector<uint32_t> group_of, seed_local, order;  // per group-order instance
  std::vector<PprTask> pieces, runs;
  std::vector<PprHub> hubs;
  uint64_t rtot = 0;
  uint32_t nparts = 0;
  for (uint32_t k = 0; k < G; k++) {
    const uint32_t Vk = b.vbase[k + 1] - b.vbase[k];
    const uint64_t *off = b.h_off.data() + b.vbase[k];
    for (size_t s0 = 0; s0 < per[k].size(); s0 += kCols) {
      PprGroup g{};
      g.vbase = b.vbase[k];
      g.V = Vk;
      g.k = (uint32_t)std::min<size_t>(kCols, per[k].size() - s0);
      g.kk = 1;
      while (g.kk < g.k) g.kk <<= 1;
      g.rbase = rtot;
      g.first = (uint32_t)order.size();
      const uint32_t gid = (uint32_t)groups.size();
      for (uint32_t c = 0; c < g.k; c++) {
        const uint32_t i = per[k][s0 + c];
        order.push_back(i);
        group_of.push_back(gid);
        seed_local.push_back(b.h_o2n[b.vbase[k] + seed[i]] - b.vbase[k]);
      }
      rtot += (uint64_t)Vk * g.k;
      // tasks: the pieces of the hub rows, and runs of the other rows sized for the slots
      const uint32_t S = kWave / g.kk;
      for (uint32_t r = 0; r < Vk; r++) {
        const uint64_t d = off[r + 1] - off[r];
        if (d <= kHubDegree) continue;
        const uint32_t np = (uint32_t)((d + kHubDegree - 1) / kHubDegree);
        hubs.push_back(PprHub{gid, r, nparts, np});
        for (uint32_t q = 0; q < np; q++) pieces.push_back(PprTask{gid, r, q, nparts + q});
        nparts += np;
      }
      for (uint32_t r0 = 0; r0 < Vk;) {
        uint32_t r1 = r0;
        uint64_t edges = 0;
        while (r1 < Vk && r1 - r0 < kRunRows * S) {
          const uint64_t d = off[r1 + 1] - off[r1];
          if (d > kHubDegree || (r1 > r0 && edges + d > (uint64_t)kRunEdges * S)) break;
          edges += d;
          r1++;
        }
        if (r1 > r0) runs.push_back(PprTask{gid, r0, r1 - r0, kNoPiece});
        r0 = r1 + (r1 < Vk && off[r1 + 1] - off[r1] > kHubDegree ? 1 : 0);
      }
      groups.push_back(g);
    }
  }
  // the queue: hub pieces first (the longest tasks start early), then the runs
  std::vector<PprTask> tasks(pieces);
  tasks.insert(tasks.end(), runs.begin(), runs.end());
  const uint32_t ng = (uint32_t)groups.size();
  DevBuf d_groups, d_gof, d_seed, d_tasks, d_hubs, d_next, d_dang, d_part, rank0, rank1, con0, con1;
  d_groups.alloc(ng * sizeof(PprGroup));
  d_gof.alloc(n * 4);
  d_seed.alloc(n * 4);
  d_tasks.alloc(std::max<size_t>(tasks.size(), 1) * sizeof(PprTask));
  d_hubs.alloc(std::max<size_t>(hubs.size(), 1) * sizeof(PprHub));
  d_next.alloc(4);
  d_dang.alloc((size_t)2 * n * 8);  // two fixed-point dangling sums per instance, ping-pong
  d_part.alloc(std::max<size_t>(nparts, 1) * kCols * 8);
  rank0.alloc(std::max<uint64_t>(rtot, 1) * 4);
  rank1.alloc(std::max<uint64_t>(rtot, 1) * 4);
  con0.alloc(std::max<uint64_t>(rtot, 1) * 4);
  con1.alloc(std::max<uint64_t>(rtot, 1) * 4);
  GMP_HIP(hipMemcpyAsync(d_groups.p, groups.data(), ng * sizeof(PprGroup), hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_gof.p, group_of.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_seed.p, seed_local.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  if (!tasks.empty())
    GMP_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(PprTask), hipMemcpyHostToDevice, e.stream));
  if (!hubs.empty())
    GMP_HIP(hipMemcpyAsync(d_hubs.p, hubs.data(), hubs.size() * sizeof(PprHub), hipMemcpyHostToDevice, e.stream));
  float *rank[2] = {rank0.as<float>(), rank1.as<float>()};
  float *con[2] = {con0.as<float>(), con1.as<float>()};
  unsigned long long *dang[2] = {d_dang.as<unsigned long long>(), d_dang.as<unsigned long long>() + n};
  PprArgs a{};
  a.off = b.in.off.as<uint64_t>();
  a.idx = b.in.idx.as<uint32_t>();
  a.outdeg = b.outdeg.as<uint32_t>();
  a.groups = d_groups.as<PprGroup>();
  a.tasks = d_tasks.as<PprTask>();
  a.ntasks = (uint32_t)tasks.size();
  a.next = d_next.as<uint32_t>();
  a.hubs = d_hubs.as<PprHub>();
  a.nhubs = (uint32_t)hubs.size();
  a.seed_local = d_seed.as<uint32_t>();
  a.part = d_part.as<double>();
  a.d = damping;
  const uint32_t gather_grid = (uint32_t)std::max(1, e.num_cus) * 8;
  const uint32_t hub_grid = std::max<uint32_t>(1, std::min<uint32_t>((a.nhubs + kWavesPerBlock - 1) / kWavesPerBlock,
                                                                      (uint32_t)e.num_cus * 8));
  uint32_t init_x = 1;
  for (const PprGroup &g : groups)
    init_x = std::max<uint32_t>(init_x, (uint32_t)std::min<uint64_t>(((uint64_t)g.V * g.k + kBlock - 1) / kBlock, 1024));

  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  ppr_init_kernel<<<dim3(init_x, ng), kBlock, 0, e.stream>>>(a.groups, a.seed_local, a.outdeg, rank[0], con[0],
                                                              dang[0]);
  GMP_CHECK_LAUNCH();
  GraphExecGuard ge;
  GMP_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < iters; it++) {
    const int cur = it & 1;
    PprArgs ai = a;
    ai.contrib_in = con[cur];
    ai.rank_out = rank[cur ^ 1];
    ai.contrib_out = con[cur ^ 1];
    ai.dang = dang[cur];
    ai.dang_next = dang[cur ^ 1];
    ppr_reset_kernel<<<1, kBlock, 0, e.stream>>>(ai.next, ai.dang_next, n);
    ppr_gather_kernel<<<gather_grid, kBlock, 0, e.stream>>>(ai);
    if (a.nhubs) ppr_hub_kernel<<<hub_grid, kBlock, 0, e.stream>>>(ai);
  }
  GMP_HIP(hipStreamEndCapture(e.stream, &ge.graph));
  if (iters > 0) {
    GMP_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
    GMP_HIP(hipGraphLaunch(ge.exec, e.stream));
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  const float *fin = rank[iters & 1];
  uint64_t total_edges = 0;
  for (uint32_t i = 0; i < n; i++) total_edges += b.ebase[graph[i] + 1] - b.ebase[graph[i]];
  if (rank_out) {
    std::vector<float> h((size_t)rtot);
    GMP_HIP(hipMemcpy(h.data(), fin, (size_t)rtot * 4, hipMemcpyDeviceToHost));
    std::vector<uint64_t> obase(n + 1, 0);
    for (uint32_t i = 0; i < n; i++) obase[i + 1] = obase[i] + (b.vbase[graph[i] + 1] - b.vbase[graph[i]]);
    for (uint32_t gi = 0; gi < n; gi++) {
      const PprGroup &g = groups[group_of[gi]];
      const uint32_t col = gi - g.first, i = order[gi];
      float *dst = rank_out + obase[i];
      const uint32_t *map = b.h_o2n.data() + g.vbase;  // caller's local id -> stored union id
      for (uint32_t v = 0; v < g.V; v++) dst[v] = h[g.rbase + (uint64_t)(map[v] - g.vbase) * g.k + col];
    }
  }
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = iters;
    st->launches = 1;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = total_edges * (uint64_t)iters;
  }
}

}  // namespace gmp
