// Batched multi-instance runs: many graphs resident at once, many SSSP sources and
// personalised-PageRank seeds spread unevenly across them, each kind advanced by one
// persistent launch per round, with the rounds captured in a HIP graph.
//
// The graphs are stored as one disjoint union (graph k owns union ids [vbase_k, vbase_{k+1})),
// in-CSR for PageRank and out-CSR for SSSP.  Instances on the same graph share the edge stream
// and keep their values instance-minor ([v][k]): SSSP groups of up to 32 instances relax an edge
// once for every instance active at its source (a frontier entry is a group vertex with a mask
// of instances), so the frontiers of all instances share one queue and one relax launch advances
// every instance; one gathered PageRank source yields the contributions of all k instances.
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

// flag = 1 when some weight needs more than a byte (a race only writes the same value)
__global__ void weight_over_byte_kernel(const uint32_t *w, uint64_t n, uint32_t *flag) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (w[i] > 0xFFu) *flag = 1u;
}

__global__ void batch_weights_to_bytes_kernel(const uint32_t *w, uint64_t n, uint8_t *b) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = (uint8_t)w[i];
}

// ---- SSSP ------------------------------------------------------------------------------
//
// Instances are grouped by graph, at most kGroupInst to a group.  A group keeps the distances of
// its k instances vertex-major: instance j's distance at vertex v is dist[dslot + v * kk + j],
// kk = k rounded up to a power of two, so the instances of one graph relax an edge with one read
// of the edge and one cache line of destination distances.  A frontier entry is a group vertex
// (gv = gvbase + v, each at most once per round); mask[gv] holds the instances whose distance at
// v dropped in the round before.  Two mask arrays alternate: a round reads and clears the one
// its entries were queued with and ORs the instances it improves into the other; the OR that
// finds a vertex's mask empty queues the vertex.
//
// A round is two launches.  batch_relax_kernel takes the frontier from the queue while it is
// small, and from the mask array itself, in group-vertex order, once it holds more than
// ngv >> scan_log entries: then the waves work through one group at a time, so the random
// destination lines they touch are those of one group's distances (tens to hundreds of MB, held
// by the memory-side cache) instead of the whole batch's.  Entries with more than 2^hub_log
// out-edges are not expanded by the wave that finds them: they go to a hub list, and
// batch_hub_kernel relaxes their edges in pieces of 2^hub_log spread over every wave, so a hub
// (a source of round 1, a popular vertex of a dense round) does not hold one wave while the
// others wait at the end of the round.

constexpr uint32_t kGroupInst = 32;   // instances per group (bits of a mask word)
constexpr uint32_t kLdsGroups = 256;  // group table kept in LDS up to this many groups
constexpr uint32_t kAppendCap = 512;  // staged appends per wave

struct BatchGroup {
  uint32_t gvbase;  // first group vertex
  uint32_t vbase;   // union vertex base of the group's graph
  uint32_t dslot;   // first distance slot
  uint32_t kklog;   // log2 of the distance stride kk
};

struct BatchSsspArgs {
  const uint64_t *off;        // union out-CSR offsets
  const uint32_t *idx;        // union ids
  const uint32_t *w;          // weights, or null when
  const uint8_t *w8;          // every weight fits a byte (R-MAT weights are 1..255)
  const BatchGroup *groups;
  uint32_t ngroups;
  uint32_t *dist;             // group-major, vertex-major inside a group
  uint32_t *mask0, *mask1;    // per group vertex, alternating by round
  uint32_t *queue0, *queue1;  // group vertices
  uint32_t *count;            // [0]=|queue0|, [1]=|queue1|, [2]=round
  uint32_t *ticket;
  unsigned long long *edges;  // instance edges relaxed
  uint4 *hubs;                // the round's hub entries: (group vertex, mask, first piece, group)
  unsigned long long *hubcnt; // hub entries << 32 | pieces, of the round
  uint32_t ngv;               // group vertices
  uint32_t scan_min;          // frontiers of more entries walk the mask array (ngv >> scan_log)
  uint32_t hub_log;           // entries with more than 2^hub_log out-edges go to the hub list (0: none)
};

// the last group whose first group vertex is <= gv
__device__ inline uint32_t group_of(const BatchGroup *g, uint32_t n, uint32_t gv) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (g[mid].gvbase <= gv) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Relaxes edge (to, weight wt) of a frontier entry for the instances in mo: sb and tb are the
// distance slots of instance 0 at the source and at the destination.  Returns the instances
// whose distance at the destination drops: nd < the distance read this round means it drops this
// round, by this lane or by another one, so the instance belongs in the next frontier whichever
// atomicMin wins (waiting for atomicMin's old value instead costs more than the marks it saves).
__device__ inline uint32_t relax_instances(uint32_t *dist, uint32_t sb, uint32_t tb, uint32_t mo, uint32_t wt) {
  uint32_t imp = 0;
  while (mo) {
    const uint32_t k = (uint32_t)__builtin_ctz(mo);
    mo &= mo - 1u;
    const uint64_t t = (uint64_t)dist[sb + k] + wt;
    const uint32_t nd = t > 0xFFFFFFFEull ? 0xFFFFFFFEu : (uint32_t)t;
    if (nd < dist[tb + k]) {
      atomicMin(&dist[tb + k], nd);
      imp |= 1u << k;
    }
  }
  return imp;
}

// Stages the group vertices this lane queued (bit q of app: gvt[q]) in the wave's LDS buffer,
// flushing it to the next queue when the following batch could overflow it.
template <int kN>
__device__ inline void stage_appends(WaveAppend &wa, uint32_t app, const uint32_t (&gvt)[kN], uint32_t *count,
                                     uint32_t *out) {
  const uint32_t cnt = (uint32_t)__popc(app);
  const uint32_t inc = wave_scan_dpp(cnt);
  const uint32_t tot = read_lane_u32(inc, kWave - 1);
  if (!tot) return;
  uint32_t pos = wa.n + inc - cnt;
#pragma unroll
  for (int q = 0; q < kN; q++)
    if ((app >> q) & 1u) wa.buf[pos++] = gvt[q];
  wa.n += tot;
  if (wa.n > kAppendCap - kWave * kN) append_flush(wa, count, out);
}

// First half of a round for every instance: takes the frontier (queue[round&1], or the mask
// array in group-vertex order when the frontier is large), clears the entries' masks, defers
// hub entries to the hub list, and appends improved group vertices to the other queue.  A wave
// takes 64 frontier entries and expands their edges kRelaxUnroll per lane at a time (a tunable:
// 2, 4 or 8); a lane walks the instances of its edge's entry: for each, the improving distance
// goes in by atomicMin, and one atomicOr per edge marks the instances improved at the
// destination.
template <int kRelaxUnroll>
__global__ __launch_bounds__(kBlock) void batch_relax_kernel(BatchSsspArgs a) {
  __shared__ uint32_t s_start[kWavesPerBlock][kWave];
  __shared__ uint64_t s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_mask[kWavesPerBlock][kWave];
  __shared__ uint32_t s_src[kWavesPerBlock][kWave];  // distance slot of the entry's instance 0
  __shared__ uint32_t s_grp[kWavesPerBlock][kWave];
  __shared__ BatchGroup s_g[kLdsGroups];
  __shared__ uint32_t s_app[kWavesPerBlock][kAppendCap];
  const bool lds_g = a.ngroups <= kLdsGroups;
  if (lds_g)
    for (uint32_t i = threadIdx.x; i < a.ngroups; i += blockDim.x) s_g[i] = a.groups[i];
  __syncthreads();
  const BatchGroup *G = lds_g ? s_g : a.groups;
  const uint32_t round = a.count[2];
  const uint32_t cur = round & 1u, nxt = cur ^ 1u;
  const uint32_t nfront = a.count[cur];
  const uint32_t *front = cur ? a.queue1 : a.queue0;
  uint32_t *out = cur ? a.queue0 : a.queue1;
  uint32_t *mcur = cur ? a.mask1 : a.mask0;
  uint32_t *mnxt = cur ? a.mask0 : a.mask1;
  const uint32_t lane = lane_id(), wv = wave_id();
  const bool scan = nfront > a.scan_min;
  const uint32_t nchunks = ((scan ? a.ngv : nfront) + kWave - 1) / kWave;
  unsigned long long edges = 0;
  WaveAppend wa{s_app[wv], 0u};
  for (uint32_t c = blockIdx.x * kWavesPerBlock + wv; c < nchunks; c += gridDim.x * kWavesPerBlock) {
    const uint32_t i = c * kWave + lane;
    uint64_t b = 0;
    uint32_t deg = 0, m = 0, src = 0, gi = 0, gv = 0;
    if (scan) {
      if (i < a.ngv) {
        gv = i;
        m = mcur[gv];
      }
    } else if (i < nfront) {
      gv = front[i];
      m = mcur[gv];
    }
    if (m) {
      gi = group_of(G, a.ngroups, gv);
      const BatchGroup g = G[gi];
      const uint32_t v = gv - g.gvbase, u = g.vbase + v;
      mcur[gv] = 0u;
      // instances whose distance at v dropped again in this round are queued for the next one,
      // where they relax v's edges from the lower distance: skip them here
      m &= ~mnxt[gv];
      b = a.off[u];
      deg = m ? (uint32_t)(a.off[u + 1] - b) : 0u;
      src = g.dslot + (v << g.kklog);
      edges += (unsigned long long)deg * __popc(m);
      if (a.hub_log && deg > (1u << a.hub_log)) {  // relaxed by batch_hub_kernel, in pieces
        const uint32_t np = (deg + (1u << a.hub_log) - 1u) >> a.hub_log;
        const unsigned long long old = atomicAdd(a.hubcnt, (1ull << 32) | np);
        a.hubs[old >> 32] = uint4{gv, m, (uint32_t)old, gi};
        deg = 0;
      }
    }
    const uint32_t incl = wave_scan_dpp(deg);
    const uint32_t total = read_lane_u32(incl, kWave - 1);
    s_start[wv][lane] = incl - deg;
    s_base[wv][lane] = b;
    s_mask[wv][lane] = m;
    s_src[wv][lane] = src;
    s_grp[wv][lane] = gi;
    wave_sync();
    for (uint32_t jb = 0; jb < total; jb += kWave * kRelaxUnroll) {
      uint64_t e[kRelaxUnroll];
      uint32_t own[kRelaxUnroll];
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
        own[q] = lo;
        e[q] = s_base[wv][lo] + (jj - s_start[wv][lo]);
      }
      uint32_t to[kRelaxUnroll], wt[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        to[q] = a.idx[e[q]];
        wt[q] = a.w8 ? (uint32_t)a.w8[e[q]] : a.w[e[q]];
      }
      uint32_t app = 0, gvt[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        gvt[q] = 0u;
        if (jb + lane + (uint32_t)q * kWave >= total) continue;
        const uint32_t o = own[q];
        const BatchGroup g = G[s_grp[wv][o]];
        const uint32_t tl = to[q] - g.vbase;
        const uint32_t imp = relax_instances(a.dist, s_src[wv][o], g.dslot + (tl << g.kklog), s_mask[wv][o], wt[q]);
        if (imp) {
          gvt[q] = g.gvbase + tl;
          if (atomicOr(&mnxt[gvt[q]], imp) == 0u) app |= 1u << q;  // the vertex's first mark
        }
      }
      stage_appends(wa, app, gvt, &a.count[nxt], out);
    }
    wave_sync();
  }
  append_flush(wa, &a.count[nxt], out);
  edges = wave_reduce_sum_u64(edges);
  if (lane == 0 && edges) atomicAdd(a.edges, edges);
}

// The last hub entry whose first piece is <= p, for a wave (p uniform): a 64-way search, each
// step comparing the first pieces at 64 evenly spaced entries of the interval left (hubs[0]'s
// first piece is 0, and first pieces grow with the entry).
__device__ inline uint32_t hub_of_piece(const uint4 *hubs, uint32_t nhub, uint32_t p) {
  const uint32_t lane = lane_id();
  uint32_t lo = 0, n = nhub;  // the entry is in [lo, lo + n); hubs[lo] qualifies
  while (n > 1) {
    const uint32_t step = (n + kWave - 1) / kWave;
    const uint32_t i = lo + lane * step;
    const uint64_t le = __ballot(i < lo + n && hubs[i].z <= p);
    const uint32_t k = 63u - (uint32_t)__clzll(le);
    const uint32_t nlo = lo + k * step;
    n = min(step, lo + n - nlo);
    lo = nlo;
  }
  return lo;
}

// Second half of a round: the out-edges of the hub entries batch_relax_kernel deferred, in
// pieces of 2^hub_log edges.  Pieces are numbered hub by hub (a hub entry's first piece comes
// from the same atomic as its slot, so first pieces grow with the slot), and the waves take
// them interleaved: at any time the grid works on neighbouring pieces, of hubs the relax
// kernel met at about the same time (in group-vertex order in a dense round), so the
// destination lines it touches belong to few groups.  The last workgroup to finish flips the
// round and clears the consumed queue's count and the hub list.
template <int kRelaxUnroll>
__global__ __launch_bounds__(kBlock) void batch_hub_kernel(BatchSsspArgs a) {
  __shared__ BatchGroup s_g[kLdsGroups];
  __shared__ uint32_t s_app[kWavesPerBlock][kAppendCap];
  const bool lds_g = a.ngroups <= kLdsGroups;
  if (lds_g)
    for (uint32_t i = threadIdx.x; i < a.ngroups; i += blockDim.x) s_g[i] = a.groups[i];
  __syncthreads();
  const BatchGroup *G = lds_g ? s_g : a.groups;
  const uint32_t round = a.count[2];
  const uint32_t cur = round & 1u, nxt = cur ^ 1u;
  uint32_t *out = cur ? a.queue0 : a.queue1;
  uint32_t *mnxt = cur ? a.mask0 : a.mask1;
  const uint32_t lane = lane_id(), wv = wave_id();
  const unsigned long long hc = *a.hubcnt;
  const uint32_t nhub = (uint32_t)(hc >> 32), npiece = (uint32_t)hc;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  WaveAppend wa{s_app[wv], 0u};
  for (uint32_t p = blockIdx.x * kWavesPerBlock + wv; p < npiece; p += nwaves) {
    const uint4 hub = a.hubs[hub_of_piece(a.hubs, nhub, p)];  // (group vertex, mask, first piece, group)
    const BatchGroup g = G[hub.w];
    const uint32_t v = hub.x - g.gvbase, u = g.vbase + v, sb = g.dslot + (v << g.kklog);
    const uint64_t eb = a.off[u], ee = a.off[u + 1];
    const uint64_t e0 = eb + ((uint64_t)(p - hub.z) << a.hub_log);
    const uint64_t e1 = min(e0 + (1ull << a.hub_log), ee);
    for (uint64_t jb = e0; jb < e1; jb += (uint64_t)kWave * kRelaxUnroll) {
      uint32_t to[kRelaxUnroll], wt[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        const uint64_t e = jb + lane + (uint64_t)q * kWave;
        const uint64_t ee2 = e < e1 ? e : e0;
        to[q] = a.idx[ee2];
        wt[q] = a.w8 ? (uint32_t)a.w8[ee2] : a.w[ee2];
      }
      uint32_t app = 0, gvt[kRelaxUnroll];
#pragma unroll
      for (int q = 0; q < kRelaxUnroll; q++) {
        gvt[q] = 0u;
        if (jb + lane + (uint64_t)q * kWave >= e1) continue;
        const uint32_t tl = to[q] - g.vbase;
        const uint32_t imp = relax_instances(a.dist, sb, g.dslot + (tl << g.kklog), hub.y, wt[q]);
        if (imp) {
          gvt[q] = g.gvbase + tl;
          if (atomicOr(&mnxt[gvt[q]], imp) == 0u) app |= 1u << q;  // the vertex's first mark
        }
      }
      stage_appends(wa, app, gvt, &a.count[nxt], out);
    }
  }
  append_flush(wa, &a.count[nxt], out);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const uint32_t t = atomicAdd(a.ticket, 1u);
    if (t == gridDim.x - 1) {  // every workgroup has read count[cur] and the hub list
      a.count[cur] = 0;
      a.count[2] = round + 1;
      *a.hubcnt = 0ull;
      *a.ticket = 0;
      __threadfence();
    }
  }
}

// Every distance unreached, every mask empty, the counters cleared.
__global__ void batch_sssp_init_kernel(uint32_t *dist, uint32_t total, uint32_t *mask0, uint32_t *mask1,
                                       uint32_t ngv, uint32_t *count, uint32_t *ticket, unsigned long long *edges,
                                       unsigned long long *hubcnt) {
  GRID_STRIDE32(i, total) dist[i] = kInf;
  GRID_STRIDE32(i, ngv) {
    mask0[i] = 0u;
    mask1[i] = 0u;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    count[0] = count[1] = count[2] = 0u;
    *ticket = 0u;
    *edges = 0ull;
    *hubcnt = 0ull;
  }
}

// The sources: seed i = (distance slot, group vertex, instance bit); instances of one group
// that share a source vertex queue it once.
__global__ void batch_sssp_seed_kernel(const uint4 *seeds, uint32_t n, uint32_t *dist, uint32_t *mask0,
                                       uint32_t *queue0, uint32_t *count) {
  GRID_STRIDE32(i, n) {
    const uint4 sd = seeds[i];
    dist[sd.x] = 0u;
    if (atomicOr(&mask0[sd.y], sd.z) == 0u) queue0[atomicAdd(&count[0], 1u)] = sd.y;
  }
}

// Distances back in instance order: out[oslot_i + v] = dist[dslot_i + v * kk_i + j_i].
__global__ void batch_sssp_out_kernel(const uint32_t *dist, const uint4 *inst, const uint32_t *oslot, uint32_t n,
                                      uint32_t total, uint32_t *out) {
  GRID_STRIDE32(s, total) {
    uint32_t lo = 0, hi = n - 1;  // last instance with oslot <= s
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (oslot[mid] <= s) lo = mid;
      else hi = mid - 1;
    }
    const uint4 in = inst[lo];  // (dslot, kklog, j, -)
    out[s] = dist[in.x + ((s - oslot[lo]) << in.y) + in.z];
  }
}

// ---- personalised PageRank -------------------------------------------------------------
//
// A group holds up to 64 instances of one graph; lane l of a wave serves instance l % kk of
// the group (kk = instances rounded up to a power of two) in slot l / kk, so a wave runs
// 64 / kk slots.  Rows of in-degree up to the piece size are handed out in runs (of at most
// run_rows rows and run_edges edges per slot), one row per slot at a time; a slot walks its
// row's in-edges with kLoads source loads in flight per lane (a tunable: 4, 8 or 16) and a Kahan
// sum per lane, and writes the row's k ranks as one contiguous vector.  Longer rows are cut into
// pieces of `piece` edges spread over the slots of a wave; each piece leaves an FP64 partial per
// instance, and ppr_hub_kernel adds a row's pieces in order.  The piece size, run sizes, loads in
// flight and grid are the autotuner's (gmp_batch_autotune).
// Waves take tasks (runs and pieces of every group) from one queue, so groups of any size and
// instance count share the chip without a cost model.  The dangling mass of the new ranks
// (needed by the next iteration's teleport) is summed per instance in an exact 61-bit fixed
// point: the sum does not depend on which wave took which task, so results are reproducible.

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

struct PprHub {        // a row above the piece size
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
  uint32_t piece;                   // edges per hub piece
  // set when the iterations share a replayed graph with other work: an iteration whose number
  // (*iter) has reached iter_limit does nothing, and ppr_advance_kernel counts iterations
  uint32_t *iter;
  uint32_t iter_limit;
};

__device__ inline bool ppr_skip(const PprArgs &a) { return a.iter && *a.iter >= a.iter_limit; }

// Kahan-sums contrib_in over the in-edges [e, e1) of a row, taking every step-th edge.  The
// loads are unconditional -- an edge past e1 reloads edge e, a lane without an instance reads
// column 0 -- so all kLoads of a batch are in flight together; the extra values are dropped.
template <int kLoads>
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

template <int kLoads>
__global__ __launch_bounds__(kBlock) void ppr_gather_kernel(PprArgs a) {
  if (ppr_skip(a)) return;
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
        ppr_sum_edges<kLoads>(a, g, inst, on, e0, e1, 1, s, c);
        if (on) dacc += ppr_finish_row(a, g, inst, r, s - c, seed, tele, od);
        e0 = n0;
        e1 = n1;
        od = nod;
      }
    } else {
      const uint32_t u = g.vbase + tk.row;
      const uint64_t e0 = a.off[u] + (uint64_t)tk.n * a.piece;
      const uint64_t e1 = min(a.off[u + 1], e0 + a.piece);
      float s = 0.0f, c = 0.0f;
      ppr_sum_edges<kLoads>(a, g, inst, on, e0 + slot, e1, S, s, c);
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
  if (ppr_skip(a)) return;
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
__global__ void ppr_reset_kernel(uint32_t *next, unsigned long long *dang_next, uint32_t n, const uint32_t *iter,
                                 uint32_t iter_limit) {
  if (iter && *iter >= iter_limit) return;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dang_next[i] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) *next = 0u;
}

// Counts a finished iteration (one thread; after the iteration's kernels on the same stream).
__global__ void ppr_advance_kernel(uint32_t *iter, uint32_t iter_limit) {
  if (threadIdx.x == 0 && *iter < iter_limit) *iter += 1u;
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
  // the SSSP pushes read one byte per weight when every weight fits one
  b.w8 = DevBuf();
  if (weighted && b.E) {
    DevBuf flag;
    flag.alloc(4);
    GMP_HIP(hipMemsetAsync(flag.p, 0, 4, e.stream));
    weight_over_byte_kernel<<<grid_for(b.E), kBlock, 0, e.stream>>>(b.out.w.as<uint32_t>(), b.E, flag.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    uint32_t over = 1;
    GMP_HIP(hipMemcpyAsync(&over, flag.p, 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    if (!over) {
      b.w8.alloc(b.E);
      batch_weights_to_bytes_kernel<<<grid_for(b.E), kBlock, 0, e.stream>>>(b.out.w.as<uint32_t>(), b.E,
                                                                            b.w8.as<uint8_t>());
      GMP_CHECK_LAUNCH();
      GMP_HIP(hipStreamSynchronize(e.stream));
      b.out.w.release();
    }
  }
  b.outdeg.alloc(std::max<size_t>(b.V, 1) * 4);
  count_degrees(e, b.in.idx.as<uint32_t>(), b.E, b.V, b.outdeg.as<uint32_t>());
  b.h_off.resize((size_t)b.V + 1);
  GMP_HIP(hipMemcpy(b.h_off.data(), b.in.off.p, ((size_t)b.V + 1) * 8, hipMemcpyDeviceToHost));
  b.ready = true;
}

namespace {

// ---- prepared runs ---------------------------------------------------------------------
// A batched SSSP or PPR run split into its set-up (instances checked, device buffers filled), its
// start (initial values) and its rounds or iterations, so that both kinds can be captured
// into one HIP graph by gmp_batch_run as well as run alone.

struct SsspRun {
  uint32_t n = 0, T = 0, ngv = 0, D = 0;  // instances, output slots, group vertices, distance slots
  std::vector<uint32_t> oslot;            // [n+1] output slot bases (instance order)
  DevBuf d_groups, d_seed, d_inst, d_oslot, dist, mask0, mask1, q0, q1, cnt, hubs;
  BatchSsspArgs a{};
  uint32_t grid = 0, nseed = 0;
};

void sssp_prepare(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *source, SsspRun &r) {
  BatchGraphs &b = e.batch;
  if (!b.ready) fail(GMP_ESTATE, "no batch loaded");
  if (!b.weighted) fail(GMP_EUNSUPPORTED, "batched SSSP needs weighted graphs");
  if (n == 0 || !graph || !source) fail(GMP_EINVAL, "batched SSSP needs instances");
  const uint32_t G = (uint32_t)b.vbase.size() - 1;
  r.n = n;
  r.oslot.assign(n + 1, 0);
  std::vector<std::vector<uint32_t>> per(G);
  for (uint32_t i = 0; i < n; i++) {
    if (graph[i] >= G) fail(GMP_EINVAL, "instance graph out of range");
    const uint32_t Vk = b.vbase[graph[i] + 1] - b.vbase[graph[i]];
    if (source[i] >= Vk) fail(GMP_EINVAL, "instance source out of range");
    if ((uint64_t)r.oslot[i] + Vk >= 0xFFFFFFFFull) fail(GMP_EUNSUPPORTED, "batch exceeds 2^32 distance slots");
    r.oslot[i + 1] = r.oslot[i] + Vk;
    per[graph[i]].push_back(i);
  }
  r.T = r.oslot[n];
  // groups of at most kGroupInst instances of one graph, in graph order
  std::vector<BatchGroup> groups;
  std::vector<uint4> seeds(n), inst(n);
  uint64_t gv = 0, ds = 0, hub_cap = 1, pieces = 0;
  const uint32_t hub_log = e.opt.bsssp_hub_log;
  for (uint32_t k = 0; k < G; k++) {
    const uint32_t Vk = b.vbase[k + 1] - b.vbase[k];
    for (size_t s0 = 0; s0 < per[k].size(); s0 += kGroupInst) {
      const uint32_t cnt = (uint32_t)std::min<size_t>(kGroupInst, per[k].size() - s0);
      uint32_t kklog = 0;
      while ((1u << kklog) < cnt) kklog++;
      if (gv + Vk >= 0xFFFFFFFFull || ds + ((uint64_t)Vk << kklog) >= 0xFFFFFFFFull)
        fail(GMP_EUNSUPPORTED, "batch exceeds 2^32 distance slots");
      groups.push_back(BatchGroup{(uint32_t)gv, b.vbase[k], (uint32_t)ds, kklog});
      // a round's hub entries: group vertices of more than 2^hub_log out-edges, at most E >> hub_log
      const uint64_t Ek = b.ebase[k + 1] - b.ebase[k];
      if (hub_log) {
        hub_cap += Ek >> hub_log;
        pieces += (Ek >> hub_log) * 2;
      }
      for (uint32_t j = 0; j < cnt; j++) {
        const uint32_t i = per[k][s0 + j];
        seeds[i] = uint4{(uint32_t)ds + (source[i] << kklog) + j, (uint32_t)gv + source[i], 1u << j, 0u};
        inst[i] = uint4{(uint32_t)ds, kklog, j, 0u};
      }
      gv += Vk;
      ds += (uint64_t)Vk << kklog;
    }
  }
  if (pieces >= 0xFFFFFFFFull) fail(GMP_EUNSUPPORTED, "batch: too many hub pieces per round");
  r.ngv = (uint32_t)gv;
  r.D = (uint32_t)ds;
  r.nseed = n;
  r.d_groups.alloc(groups.size() * sizeof(BatchGroup));
  r.d_seed.alloc(n * sizeof(uint4));
  r.d_inst.alloc(n * sizeof(uint4));
  r.d_oslot.alloc((n + 1) * 4);
  r.dist.alloc((size_t)r.D * 4);
  r.mask0.alloc((size_t)r.ngv * 4);
  r.mask1.alloc((size_t)r.ngv * 4);
  r.q0.alloc((size_t)r.ngv * 4);
  r.q1.alloc((size_t)r.ngv * 4);
  r.cnt.alloc(64);
  r.hubs.alloc(hub_cap * sizeof(uint4));
  GMP_HIP(hipMemcpyAsync(r.d_groups.p, groups.data(), groups.size() * sizeof(BatchGroup), hipMemcpyHostToDevice,
                         e.stream));
  GMP_HIP(hipMemcpyAsync(r.d_seed.p, seeds.data(), n * sizeof(uint4), hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(r.d_inst.p, inst.data(), n * sizeof(uint4), hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(r.d_oslot.p, r.oslot.data(), (n + 1) * 4, hipMemcpyHostToDevice, e.stream));
  BatchSsspArgs &a = r.a;
  a.off = b.out.off.as<uint64_t>();
  a.idx = b.out.idx.as<uint32_t>();
  a.w = b.out.w.as<uint32_t>();
  a.w8 = b.w8.as<uint8_t>();
  a.groups = r.d_groups.as<BatchGroup>();
  a.ngroups = (uint32_t)groups.size();
  a.dist = r.dist.as<uint32_t>();
  a.mask0 = r.mask0.as<uint32_t>();
  a.mask1 = r.mask1.as<uint32_t>();
  a.queue0 = r.q0.as<uint32_t>();
  a.queue1 = r.q1.as<uint32_t>();
  a.count = r.cnt.as<uint32_t>();
  a.ticket = r.cnt.as<uint32_t>() + 4;
  a.edges = reinterpret_cast<unsigned long long *>(r.cnt.as<uint32_t>() + 8);
  a.hubcnt = reinterpret_cast<unsigned long long *>(r.cnt.as<uint32_t>() + 10);
  a.hubs = r.hubs.as<uint4>();
  a.ngv = r.ngv;
  a.scan_min = e.opt.bsssp_scan_log ? r.ngv >> e.opt.bsssp_scan_log : 0xFFFFFFFFu;
  a.hub_log = hub_log;
  r.grid = std::max<uint32_t>(1, (uint32_t)e.num_cus * e.opt.bsssp_grid);
}

void sssp_start(SsspRun &r, hipStream_t s) {
  batch_sssp_init_kernel<<<grid_for(std::max(r.D, r.ngv)), kBlock, 0, s>>>(r.a.dist, r.D, r.a.mask0, r.a.mask1, r.ngv,
                                                                            r.a.count, r.a.ticket, r.a.edges,
                                                                            r.a.hubcnt);
  batch_sssp_seed_kernel<<<grid_for(r.nseed), kBlock, 0, s>>>(r.d_seed.as<uint4>(), r.nseed, r.a.dist, r.a.mask0,
                                                              r.a.queue0, r.a.count);
  GMP_CHECK_LAUNCH();
}

void sssp_rounds(Engine &e, SsspRun &r, hipStream_t s, int rounds) {
  for (int k = 0; k < rounds; k++) {
    switch (e.opt.bsssp_unroll) {
      case 2:
        batch_relax_kernel<2><<<r.grid, kBlock, 0, s>>>(r.a);
        batch_hub_kernel<2><<<r.grid, kBlock, 0, s>>>(r.a);
        break;
      case 8:
        batch_relax_kernel<8><<<r.grid, kBlock, 0, s>>>(r.a);
        batch_hub_kernel<8><<<r.grid, kBlock, 0, s>>>(r.a);
        break;
      default:
        batch_relax_kernel<4><<<r.grid, kBlock, 0, s>>>(r.a);
        batch_hub_kernel<4><<<r.grid, kBlock, 0, s>>>(r.a);
        break;
    }
    GMP_CHECK_LAUNCH();
  }
}

void sssp_results(Engine &e, SsspRun &r, uint32_t *dist_out, uint64_t &edges) {
  unsigned long long h = 0;
  GMP_HIP(hipMemcpy(&h, r.a.edges, 8, hipMemcpyDeviceToHost));
  edges = h;
  if (dist_out) {
    DevBuf out;
    out.alloc((size_t)r.T * 4);
    batch_sssp_out_kernel<<<grid_for(r.T), kBlock, 0, e.stream>>>(r.a.dist, r.d_inst.as<uint4>(),
                                                                  r.d_oslot.as<uint32_t>(), r.n, r.T,
                                                                  out.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    GMP_HIP(hipMemcpyAsync(dist_out, out.p, (size_t)r.T * 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
  }
}

struct PprRun {
  uint32_t n = 0, ng = 0;
  uint64_t rtot = 0, instance_edges = 0;
  std::vector<PprGroup> groups;
  std::vector<uint32_t> group_of, order;
  DevBuf d_groups, d_gof, d_seed, d_tasks, d_hubs, d_next, d_dang, d_part, rank0, rank1, con0, con1, d_iter;
  PprArgs a{};
  float *rank[2] = {nullptr, nullptr};
  float *con[2] = {nullptr, nullptr};
  unsigned long long *dang[2] = {nullptr, nullptr};
  uint32_t gather_grid = 0, hub_grid = 0, init_x = 1;
};

void ppr_prepare(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *seed, float damping, PprRun &r) {
  BatchGraphs &b = e.batch;
  if (!b.ready) fail(GMP_ESTATE, "no batch loaded");
  if (n == 0 || !graph || !seed) fail(GMP_EINVAL, "batched PageRank needs instances");
  if (!(damping >= 0.0f && damping <= 1.0f)) fail(GMP_EINVAL, "bad damping");
  const uint32_t G = (uint32_t)b.vbase.size() - 1;
  const uint32_t piece = e.opt.ppr_piece, run_rows = e.opt.ppr_run_rows, run_edges = e.opt.ppr_run_edges;
  // group instances by graph, at most 64 per group; group order keeps caller order per graph
  std::vector<std::vector<uint32_t>> per(G);
  for (uint32_t i = 0; i < n; i++) {
    if (graph[i] >= G) fail(GMP_EINVAL, "instance graph out of range");
    if (seed[i] >= b.vbase[graph[i] + 1] - b.vbase[graph[i]]) fail(GMP_EINVAL, "instance seed out of range");
    per[graph[i]].push_back(i);
    r.instance_edges += b.ebase[graph[i] + 1] - b.ebase[graph[i]];
  }
  r.n = n;
  std::vector<uint32_t> seed_local;
  std::vector<PprTask> pieces, runs;
  std::vector<PprHub> hubs;
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
      g.rbase = r.rtot;
      g.first = (uint32_t)r.order.size();
      const uint32_t gid = (uint32_t)r.groups.size();
      for (uint32_t c = 0; c < g.k; c++) {
        const uint32_t i = per[k][s0 + c];
        r.order.push_back(i);
        r.group_of.push_back(gid);
        seed_local.push_back(b.h_o2n[b.vbase[k] + seed[i]] - b.vbase[k]);
      }
      r.rtot += (uint64_t)Vk * g.k;
      // tasks: the pieces of the hub rows, and runs of the other rows sized for the slots
      const uint32_t S = kWave / g.kk;
      for (uint32_t row = 0; row < Vk; row++) {
        const uint64_t d = off[row + 1] - off[row];
        if (d <= piece) continue;
        const uint32_t np = (uint32_t)((d + piece - 1) / piece);
        hubs.push_back(PprHub{gid, row, nparts, np});
        for (uint32_t q = 0; q < np; q++) pieces.push_back(PprTask{gid, row, q, nparts + q});
        nparts += np;
      }
      for (uint32_t r0 = 0; r0 < Vk;) {
        uint32_t r1 = r0;
        uint64_t edges = 0;
        while (r1 < Vk && r1 - r0 < run_rows * S) {
          const uint64_t d = off[r1 + 1] - off[r1];
          if (d > piece || (r1 > r0 && edges + d > (uint64_t)run_edges * S)) break;
          edges += d;
          r1++;
        }
        if (r1 > r0) runs.push_back(PprTask{gid, r0, r1 - r0, kNoPiece});
        r0 = r1 + (r1 < Vk && off[r1 + 1] - off[r1] > piece ? 1 : 0);
      }
      r.groups.push_back(g);
    }
  }
  // the queue: hub pieces first (the longest tasks start early), then the runs
  std::vector<PprTask> tasks(pieces);
  tasks.insert(tasks.end(), runs.begin(), runs.end());
  r.ng = (uint32_t)r.groups.size();
  r.d_groups.alloc(r.ng * sizeof(PprGroup));
  r.d_gof.alloc(n * 4);
  r.d_seed.alloc(n * 4);
  r.d_tasks.alloc(std::max<size_t>(tasks.size(), 1) * sizeof(PprTask));
  r.d_hubs.alloc(std::max<size_t>(hubs.size(), 1) * sizeof(PprHub));
  r.d_next.alloc(4);
  r.d_iter.alloc(4);
  r.d_dang.alloc((size_t)2 * n * 8);  // two fixed-point dangling sums per instance, ping-pong
  r.d_part.alloc(std::max<size_t>(nparts, 1) * kCols * 8);
  r.rank0.alloc(std::max<uint64_t>(r.rtot, 1) * 4);
  r.rank1.alloc(std::max<uint64_t>(r.rtot, 1) * 4);
  r.con0.alloc(std::max<uint64_t>(r.rtot, 1) * 4);
  r.con1.alloc(std::max<uint64_t>(r.rtot, 1) * 4);
  GMP_HIP(hipMemcpyAsync(r.d_groups.p, r.groups.data(), r.ng * sizeof(PprGroup), hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(r.d_gof.p, r.group_of.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(r.d_seed.p, seed_local.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  if (!tasks.empty())
    GMP_HIP(hipMemcpyAsync(r.d_tasks.p, tasks.data(), tasks.size() * sizeof(PprTask), hipMemcpyHostToDevice, e.stream));
  if (!hubs.empty())
    GMP_HIP(hipMemcpyAsync(r.d_hubs.p, hubs.data(), hubs.size() * sizeof(PprHub), hipMemcpyHostToDevice, e.stream));
  r.rank[0] = r.rank0.as<float>();
  r.rank[1] = r.rank1.as<float>();
  r.con[0] = r.con0.as<float>();
  r.con[1] = r.con1.as<float>();
  r.dang[0] = r.d_dang.as<unsigned long long>();
  r.dang[1] = r.dang[0] + n;
  PprArgs &a = r.a;
  a.off = b.in.off.as<uint64_t>();
  a.idx = b.in.idx.as<uint32_t>();
  a.outdeg = b.outdeg.as<uint32_t>();
  a.groups = r.d_groups.as<PprGroup>();
  a.tasks = r.d_tasks.as<PprTask>();
  a.ntasks = (uint32_t)tasks.size();
  a.next = r.d_next.as<uint32_t>();
  a.hubs = r.d_hubs.as<PprHub>();
  a.nhubs = (uint32_t)hubs.size();
  a.seed_local = r.d_seed.as<uint32_t>();
  a.part = r.d_part.as<double>();
  a.d = damping;
  a.piece = piece;
  r.gather_grid = (uint32_t)std::max(1, e.num_cus) * e.opt.ppr_grid;
  r.hub_grid = std::max<uint32_t>(1, std::min<uint32_t>((a.nhubs + kWavesPerBlock - 1) / kWavesPerBlock,
                                                         (uint32_t)e.num_cus * 8));
  for (const PprGroup &g : r.groups)
    r.init_x = std::max<uint32_t>(r.init_x, (uint32_t)std::min<uint64_t>(((uint64_t)g.V * g.k + kBlock - 1) / kBlock, 1024));
}

__global__ void set_u32_kernel(uint32_t *p, uint32_t v) {
  if (threadIdx.x == 0) *p = v;
}

void ppr_start(PprRun &r, hipStream_t s) {
  ppr_init_kernel<<<dim3(r.init_x, r.ng), kBlock, 0, s>>>(r.a.groups, r.a.seed_local, r.a.outdeg, r.rank[0], r.con[0],
                                                          r.dang[0]);
  set_u32_kernel<<<1, 64, 0, s>>>(r.d_iter.as<uint32_t>(), 0u);
  GMP_CHECK_LAUNCH();
}

// Iterations it0 .. it0 + count - 1 (ping-pong buffers by parity).  guard: each iteration checks
// the iteration counter against `limit` and counts itself (for graphs replayed past the end).
void ppr_iterations(Engine &e, PprRun &r, hipStream_t s, int it0, int count, bool guard, uint32_t limit) {
  for (int it = it0; it < it0 + count; it++) {
    const int cur = it & 1;
    PprArgs ai = r.a;
    ai.contrib_in = r.con[cur];
    ai.rank_out = r.rank[cur ^ 1];
    ai.contrib_out = r.con[cur ^ 1];
    ai.dang = r.dang[cur];
    ai.dang_next = r.dang[cur ^ 1];
    ai.iter = guard ? r.d_iter.as<uint32_t>() : nullptr;
    ai.iter_limit = limit;
    ppr_reset_kernel<<<1, kBlock, 0, s>>>(ai.next, ai.dang_next, r.n, ai.iter, limit);
    switch (e.opt.ppr_loads) {
      case 4: ppr_gather_kernel<4><<<r.gather_grid, kBlock, 0, s>>>(ai); break;
      case 16: ppr_gather_kernel<16><<<r.gather_grid, kBlock, 0, s>>>(ai); break;
      default: ppr_gather_kernel<8><<<r.gather_grid, kBlock, 0, s>>>(ai); break;
    }
    if (r.a.nhubs) ppr_hub_kernel<<<r.hub_grid, kBlock, 0, s>>>(ai);
    if (guard) ppr_advance_kernel<<<1, 64, 0, s>>>(ai.iter, limit);
    GMP_CHECK_LAUNCH();
  }
}

void ppr_results(Engine &e, PprRun &r, const uint32_t *graph, int iters, float *rank_out) {
  if (!rank_out) return;
  BatchGraphs &b = e.batch;
  const float *fin = r.rank[iters & 1];
  std::vector<float> h((size_t)r.rtot);
  GMP_HIP(hipMemcpy(h.data(), fin, (size_t)r.rtot * 4, hipMemcpyDeviceToHost));
  std::vector<uint64_t> obase(r.n + 1, 0);
  for (uint32_t i = 0; i < r.n; i++) obase[i + 1] = obase[i] + (b.vbase[graph[i] + 1] - b.vbase[graph[i]]);
  for (uint32_t gi = 0; gi < r.n; gi++) {
    const PprGroup &g = r.groups[r.group_of[gi]];
    const uint32_t col = gi - g.first, i = r.order[gi];
    float *dst = rank_out + obase[i];
    const uint32_t *map = b.h_o2n.data() + g.vbase;  // caller's local id -> stored union id
    for (uint32_t v = 0; v < g.V; v++) dst[v] = h[g.rbase + (uint64_t)(map[v] - g.vbase) * g.k + col];
  }
}

// Reads the SSSP counters ({|queue0|, |queue1|, round}) and the PPR iteration count.
void read_batch_state(Engine &e, const SsspRun *sr, const PprRun *pr, uint32_t h[4], uint32_t &it) {
  uint32_t *p = static_cast<uint32_t *>(e.h_pinned);
  if (sr) GMP_HIP(hipMemcpyAsync(p, sr->a.count, 16, hipMemcpyDeviceToHost, e.stream));
  if (pr) GMP_HIP(hipMemcpyAsync(p + 4, pr->d_iter.p, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (sr) std::memcpy(h, p, 16);
  if (pr) it = p[4];
}

}  // namespace

void batch_sssp(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *source, uint32_t *dist_out,
                gmp_stats *st) {
  SsspRun r;
  sssp_prepare(e, n, graph, source, r);
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  sssp_start(r, e.stream);
  // capture two rounds; the counters live on the device, so replays need no host input
  GraphExecGuard ge;
  GMP_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
  sssp_rounds(e, r, e.stream, 2);
  GMP_HIP(hipStreamEndCapture(e.stream, &ge.graph));
  GMP_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
  uint32_t h[4] = {0, 0, 0, 0}, unused = 0;
  int launches = 0;
  for (;;) {
    GMP_HIP(hipGraphLaunch(ge.exec, e.stream));
    launches++;
    read_batch_state(e, &r, nullptr, h, unused);
    if (h[h[2] & 1u] == 0) break;
    if (launches > (1 << 22)) fail(GMP_ESTATE, "batched SSSP did not converge");
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  uint64_t edges = 0;
  sssp_results(e, r, dist_out, edges);
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
  if (iters < 0) fail(GMP_EINVAL, "bad iteration count");
  PprRun r;
  ppr_prepare(e, n, graph, seed, damping, r);
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  ppr_start(r, e.stream);
  GraphExecGuard ge;
  GMP_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
  ppr_iterations(e, r, e.stream, 0, iters, false, 0u);
  GMP_HIP(hipStreamEndCapture(e.stream, &ge.graph));
  if (iters > 0) {
    GMP_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
    GMP_HIP(hipGraphLaunch(ge.exec, e.stream));
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  ppr_results(e, r, graph, iters, rank_out);
  if (st) {
    std::memset(st, 0, sizeof *st);
    st->iterations = iters;
    st->launches = 1;
    st->ms_total = tot.ms();
    st->ms_kernel = st->ms_total;
    st->edges = r.instance_edges * (uint64_t)iters;
  }
}

// SSSP and PPR instances in one replayed HIP graph: a fork into two branches -- two SSSP rounds on
// one stream, two PPR iterations (guarded by the iteration counter) on the other -- joined again,
// replayed until the SSSP frontier is empty and every PPR iteration ran.  The branches run at the
// same time, so the PPR gathers fill the chip while the SSSP rounds wait on their atomics.
void batch_run(Engine &e, uint32_t ns, const uint32_t *sgraph, const uint32_t *source, uint32_t np,
               const uint32_t *pgraph, const uint32_t *seed, float damping, int32_t iters, uint32_t *dist_out,
               float *rank_out, gmp_stats *st_sssp, gmp_stats *st_ppr, double *ms_total) {
  if (iters < 0) fail(GMP_EINVAL, "bad iteration count");
  if (ns == 0 && np == 0) fail(GMP_EINVAL, "batched run needs instances");
  SsspRun sr;
  PprRun pr;
  if (ns) sssp_prepare(e, ns, sgraph, source, sr);
  if (np) ppr_prepare(e, np, pgraph, seed, damping, pr);
  const bool has_ppr = np && iters > 0;
  struct StreamGuard {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    ~StreamGuard() {
      if (fork) (void)hipEventDestroy(fork);
      if (join) (void)hipEventDestroy(join);
      if (s) (void)hipStreamDestroy(s);
    }
  } sg;
  GMP_HIP(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  GMP_HIP(hipEventCreateWithFlags(&sg.fork, hipEventDisableTiming));
  GMP_HIP(hipEventCreateWithFlags(&sg.join, hipEventDisableTiming));
  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  if (ns) sssp_start(sr, e.stream);
  if (np) ppr_start(pr, e.stream);
  GraphExecGuard ge;
  GMP_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
  GMP_HIP(hipEventRecord(sg.fork, e.stream));
  GMP_HIP(hipStreamWaitEvent(sg.s, sg.fork, 0));
  if (ns) sssp_rounds(e, sr, e.stream, 2);
  if (has_ppr) ppr_iterations(e, pr, sg.s, 0, 2, true, (uint32_t)iters);
  GMP_HIP(hipEventRecord(sg.join, sg.s));
  GMP_HIP(hipStreamWaitEvent(e.stream, sg.join, 0));
  GMP_HIP(hipStreamEndCapture(e.stream, &ge.graph));
  GMP_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
  uint32_t h[4] = {0, 0, 0, 0}, it = 0;
  int launches = 0;
  for (;;) {
    GMP_HIP(hipGraphLaunch(ge.exec, e.stream));
    launches++;
    read_batch_state(e, ns ? &sr : nullptr, has_ppr ? &pr : nullptr, h, it);
    const bool sssp_done = !ns || h[h[2] & 1u] == 0;
    const bool ppr_done = !has_ppr || it >= (uint32_t)iters;
    if (sssp_done && ppr_done) break;
    if (launches > (1 << 22)) fail(GMP_ESTATE, "batched run did not converge");
  }
  GMP_HIP(hipEventRecord(tot.b, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  const double ms = tot.ms();
  if (ms_total) *ms_total = ms;
  if (ns) {
    uint64_t edges = 0;
    sssp_results(e, sr, dist_out, edges);
    if (st_sssp) {
      std::memset(st_sssp, 0, sizeof *st_sssp);
      st_sssp->iterations = (int32_t)h[2];
      st_sssp->launches = launches;
      st_sssp->ms_total = ms;
      st_sssp->ms_kernel = ms;
      st_sssp->edges = edges;
      st_sssp->bytes = edges * 12;
    }
  }
  if (np) {
    ppr_results(e, pr, pgraph, iters, rank_out);
    if (st_ppr) {
      std::memset(st_ppr, 0, sizeof *st_ppr);
      st_ppr->iterations = iters;
      st_ppr->launches = launches;
      st_ppr->ms_total = ms;
      st_ppr->ms_kernel = ms;
      st_ppr->edges = pr.instance_edges * (uint64_t)iters;
    }
  }
}

}  // namespace gmp
