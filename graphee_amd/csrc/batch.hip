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

// One round for every instance: reads queue[round&1] and appends improved slots to the other
// queue; the last workgroup to finish flips the round and clears the consumed count.
__global__ __launch_bounds__(kBlock) void batch_relax_kernel(BatchSsspArgs a) {
  __shared__ uint32_t s_start[kWavesPerBlock][kWave];
  __shared__ uint64_t s_base[kWavesPerBlock][kWave];
  __shared__ uint32_t s_val[kWavesPerBlock][kWave];
  __shared__ uint32_t s_sb[kWavesPerBlock][kWave];  // slot base minus union vertex base
  const uint32_t round = a.count[2];
  const uint32_t cur = round & 1u, nxt = cur ^ 1u;
  const uint32_t nfront = a.count[cur];
  const uint32_t *front = cur ? a.queue1 : a.queue0;
  uint32_t *out = cur ? a.queue0 : a.queue1;
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t nchunks = (nfront + kWave - 1) / kWave;
  unsigned long long edges = 0;
  for (uint32_t c = blockIdx.x * kWavesPerBlock + wv; c < nchunks; c += gridDim.x * kWavesPerBlock) {
    const uint32_t i = c * kWave + lane;
    uint64_t b = 0;
    uint32_t deg = 0, du = kInf, sb = 0;
    if (i < nfront) {
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
    s_sb[wv][lane] = sb;
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
        slot = s_sb[wv][lo] + a.idx[e];
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

struct PprGroup {       // up to 64 instances on one graph
  uint32_t vbase;       // union vertex base
  uint32_t V;
  uint32_t k;           // instances in the group
  uint32_t kk;          // k rounded up to a power of two
  uint64_t rbase;       // base of the group's [V][k] block
  uint32_t first;       // group-order instance id of column 0
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
  const uint32_t *outdeg;
  const PprGroup *groups;
  const PprTask *tasks;
  uint32_t ntasks;
  const float *contrib_in;
  float *rank_out;
  float *contrib_out;
  const uint32_t *seed_local;  // per group-order instance: seed (local id)
  const double *dang;          // per group-order instance: dangling mass of the current ranks
  float d;
};

// Lane l serves instance l % kk over edge slots l / kk, l / kk + 64/kk, ...; the lanes of one
// instance are combined by a butterfly over the edge-slot bits (Kahan per lane, hi/lo combine).
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
        const uint32_t gi = g.first + inst;
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

// One block per instance: FP64 mass on the dangling vertices of its graph (fixed order).
__global__ __launch_bounds__(kBlock) void ppr_dangling_kernel(const PprGroup *groups, const uint32_t *group_of,
                                                              const uint32_t *dlist, const uint64_t *dbase,
                                                              const float *rank, double *dang) {
  __shared__ double sh[kBlock];
  const uint32_t i = blockIdx.x;
  const uint32_t gid = group_of[i];
  const PprGroup g = groups[gid];
  const uint32_t col = i - g.first;
  double acc = 0;
  for (uint64_t j = dbase[gid] + threadIdx.x; j < dbase[gid + 1]; j += kBlock)
    acc += (double)rank[g.rbase + (uint64_t)dlist[j] * g.k + col];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) dang[i] = sh[0];
}

// One block per instance: r = e_seed, c = r / outdeg.
__global__ void ppr_init_kernel(const PprGroup *groups, const uint32_t *group_of, const uint32_t *seed_local,
                                const uint32_t *outdeg, float *rank, float *contrib) {
  const uint32_t i = blockIdx.x;
  const PprGroup g = groups[group_of[i]];
  const uint32_t col = i - g.first;
  for (uint32_t v = threadIdx.x; v < g.V; v += blockDim.x) {
    const float r = v == seed_local[i] ? 1.0f : 0.0f;
    const uint64_t pos = g.rbase + (uint64_t)v * g.k + col;
    rank[pos] = r;
    const uint32_t od = outdeg[g.vbase + v];
    contrib[pos] = od ? r / (float)od : 0.0f;
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
  std::vector<Csr> parts(ngraphs);
  std::vector<bool> off64(ngraphs);
  bool weighted = true;
  for (uint32_t k = 0; k < ngraphs; k++) {
    graph_generate_rmat(e, params[k]);
    if ((uint64_t)b.vbase[k] + e.g.V >= 0xFFFFFFFFull) fail(GMP_EUNSUPPORTED, "batch exceeds 2^32 vertices");
    b.vbase[k + 1] = b.vbase[k] + e.g.V;
    b.ebase[k + 1] = b.ebase[k] + e.g.E;
    weighted = weighted && e.g.weighted;
    off64[k] = e.g.off64;
    parts[k] = std::move(e.g.in);
  }
  b.V = b.vbase[ngraphs];
  b.E = b.ebase[ngraphs];
  b.weighted = weighted;
  b.in.off.alloc(((size_t)b.V + 1) * 8);
  b.in.idx.alloc(std::max<uint64_t>(b.E, 1) * 4);
  if (weighted) b.in.w.alloc(std::max<uint64_t>(b.E, 1) * 4);
  for (uint32_t k = 0; k < ngraphs; k++) {
    const uint32_t Vk = b.vbase[k + 1] - b.vbase[k];
    const uint64_t Ek = b.ebase[k + 1] - b.ebase[k];
    uint64_t *o = b.in.off.as<uint64_t>() + b.vbase[k];
    if (off64[k]) shift_offsets_kernel<uint64_t><<<grid_for(Vk + 1), kBlock, 0, e.stream>>>(parts[k].off.as<uint64_t>(), Vk, b.ebase[k], o);
    else shift_offsets_kernel<uint32_t><<<grid_for(Vk + 1), kBlock, 0, e.stream>>>(parts[k].off.as<uint32_t>(), Vk, b.ebase[k], o);
    GMP_CHECK_LAUNCH();
    if (Ek) {
      shift_ids_kernel<<<grid_for(Ek), kBlock, 0, e.stream>>>(parts[k].idx.as<uint32_t>(), Ek, b.vbase[k],
                                                             b.in.idx.as<uint32_t>() + b.ebase[k]);
      GMP_CHECK_LAUNCH();
      if (weighted)
        GMP_HIP(hipMemcpyAsync(b.in.w.as<uint32_t>() + b.ebase[k], parts[k].w.p, Ek * 4, hipMemcpyDeviceToDevice,
                               e.stream));
    }
    GMP_HIP(hipStreamSynchronize(e.stream));
    parts[k] = Csr();
  }
  b.in.valid = true;
  b.outdeg.alloc(std::max<size_t>(b.V, 1) * 4);
  count_degrees(e, b.in.idx.as<uint32_t>(), b.E, b.V, b.outdeg.as<uint32_t>());
  transpose_csr(e, CsrView{b.in.off.p, b.in.idx.as<uint32_t>(), b.in.w.as<uint32_t>()}, b.V, b.E, true, weighted,
                b.out);
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
  std::vector<PprTask> tasks;
  std::vector<uint32_t> dlist;
  std::vector<uint64_t> dbase(1, 0);
  std::vector<uint32_t> od((size_t)b.V);
  GMP_HIP(hipMemcpy(od.data(), b.outdeg.p, (size_t)b.V * 4, hipMemcpyDeviceToHost));
  uint64_t rtot = 0;
  for (uint32_t k = 0; k < G; k++) {
    const uint32_t Vk = b.vbase[k + 1] - b.vbase[k];
    for (size_t s0 = 0; s0 < per[k].size(); s0 += 64) {
      PprGroup g{};
      g.vbase = b.vbase[k];
      g.V = Vk;
      g.k = (uint32_t)std::min<size_t>(64, per[k].size() - s0);
      g.kk = 1;
      while (g.kk < g.k) g.kk <<= 1;
      g.rbase = rtot;
      g.first = (uint32_t)order.size();
      const uint32_t gid = (uint32_t)groups.size();
      for (uint32_t c = 0; c < g.k; c++) {
        const uint32_t i = per[k][s0 + c];
        order.push_back(i);
        group_of.push_back(gid);
        seed_local.push_back(seed[i]);
      }
      rtot += (uint64_t)Vk * g.k;
      groups.push_back(g);
      for (uint32_t v = 0; v < Vk; v++)
        if (od[g.vbase + v] == 0) dlist.push_back(v);
      dbase.push_back(dlist.size());
      // row runs of at most 2048 edges or 256 rows
      uint32_t r0 = 0;
      while (r0 < Vk) {
        uint32_t r1 = r0;
        uint64_t edges = 0;
        while (r1 < Vk && r1 - r0 < 256) {
          const uint64_t d = b.h_off[g.vbase + r1 + 1] - b.h_off[g.vbase + r1];
          if (r1 > r0 && edges + d > 2048) break;
          edges += d;
          r1++;
        }
        tasks.push_back(PprTask{gid, r0, r1 - r0, 0});
        r0 = r1;
      }
    }
  }
  const uint32_t ng = (uint32_t)groups.size();
  DevBuf d_groups, d_gof, d_seed, d_tasks, d_dlist, d_dbase, d_dang, rank0, rank1, con0, con1;
  d_groups.alloc(ng * sizeof(PprGroup));
  d_gof.alloc(n * 4);
  d_seed.alloc(n * 4);
  d_tasks.alloc(std::max<size_t>(tasks.size(), 1) * sizeof(PprTask));
  d_dlist.alloc(std::max<size_t>(dlist.size(), 1) * 4);
  d_dbase.alloc(dbase.size() * 8);
  d_dang.alloc(n * 8);
  rank0.alloc(std::max<uint64_t>(rtot, 1) * 4);
  rank1.alloc(std::max<uint64_t>(rtot, 1) * 4);
  con0.alloc(std::max<uint64_t>(rtot, 1) * 4);
  con1.alloc(std::max<uint64_t>(rtot, 1) * 4);
  GMP_HIP(hipMemcpyAsync(d_groups.p, groups.data(), ng * sizeof(PprGroup), hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_gof.p, group_of.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_seed.p, seed_local.data(), n * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(PprTask), hipMemcpyHostToDevice, e.stream));
  if (!dlist.empty())
    GMP_HIP(hipMemcpyAsync(d_dlist.p, dlist.data(), dlist.size() * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(d_dbase.p, dbase.data(), dbase.size() * 8, hipMemcpyHostToDevice, e.stream));
  float *rank[2] = {rank0.as<float>(), rank1.as<float>()};
  float *con[2] = {con0.as<float>(), con1.as<float>()};
  PprArgs a{};
  a.off = b.in.off.as<uint64_t>();
  a.idx = b.in.idx.as<uint32_t>();
  a.outdeg = b.outdeg.as<uint32_t>();
  a.groups = d_groups.as<PprGroup>();
  a.tasks = d_tasks.as<PprTask>();
  a.ntasks = (uint32_t)tasks.size();
  a.seed_local = d_seed.as<uint32_t>();
  a.dang = d_dang.as<double>();
  a.d = damping;
  const uint32_t grid = (uint32_t)e.num_cus * 8;

  EventPair tot;
  GMP_HIP(hipEventRecord(tot.a, e.stream));
  ppr_init_kernel<<<n, kBlock, 0, e.stream>>>(a.groups, d_gof.as<uint32_t>(), a.seed_local, a.outdeg, rank[0],
                                               con[0]);
  GMP_CHECK_LAUNCH();
  GraphExecGuard ge;
  GMP_HIP(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
  for (int it = 0; it < iters; it++) {
    const int cur = it & 1;
    ppr_dangling_kernel<<<n, kBlock, 0, e.stream>>>(a.groups, d_gof.as<uint32_t>(), d_dlist.as<uint32_t>(),
                                                     d_dbase.as<uint64_t>(), rank[cur], d_dang.as<double>());
    PprArgs ai = a;
    ai.contrib_in = con[cur];
    ai.rank_out = rank[cur ^ 1];
    ai.contrib_out = con[cur ^ 1];
    ppr_gather_kernel<<<grid, kBlock, 0, e.stream>>>(ai);
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
      for (uint32_t v = 0; v < g.V; v++) dst[v] = h[g.rbase + (uint64_t)v * g.k + col];
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
