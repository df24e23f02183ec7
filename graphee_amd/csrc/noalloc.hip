// PageRank that allocates nothing, for callers that cannot allocate device memory (work enqueued
// under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// It computes what gmp_pagerank computes -- the standard formula with FP32 ranks, a compensated
// (Kahan) per-vertex gather, FP64 dangling-mass and L1-residual reductions in a fixed order --
// over a destination-row CSR the caller already holds in device memory, and writes the ranks to a
// caller-owned device buffer.  Everything else it needs (the second rank buffer, both contribution
// buffers, the FP64 partials and the per-iteration scalars) lives in `scratch`, a file-scope device
// array of kScratchBytes per GPU, held from the library's load until the process ends.  Host code
// never passes `scratch` itself to a kernel (in host code the name is a stand-in, not a device
// address): gmp_create resolves the array's address on its device once, with hipGetSymbolAddress,
// and every launch takes that pointer.
//
// Calls on one GPU share its array and leave their contents behind: each call writes every scratch
// word it later reads before reading it, and calls that overlap on one GPU (two engines, or two
// host threads, on the same device) race on the array.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "gmp_internal.hpp"
#include "pr_common.hpp"
#include "wave.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;
constexpr uint32_t kMaxBlocks = 4096;                    // gather grid cap: one FP64 partial pair each
constexpr uint64_t kPartBytes = 2ull * kMaxBlocks * sizeof(double);
constexpr uint64_t kScalBytes = 256;                     // PrScal, padded
constexpr uint64_t kScratchBytes = 3ull * GMP_NOALLOC_MAX_VERTICES * sizeof(float) + kPartBytes + kScalBytes;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, resolved once per device by noalloc_attach.
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

struct NaArgs {
  const uint64_t *off;
  const uint32_t *idx;
  const uint32_t *outdeg;
  const float *contrib_in;
  const float *rank_in;
  float *rank_out;
  float *contrib_out;
  const PrScal *sc;
  float d;
  uint32_t seed;  // 0xFFFFFFFF: uniform teleport
  uint32_t V;
  int kahan;
  double *pr, *pd;
};

// Ranks start at the teleport vector (1/V, or 1 at the seed), contributions rank / out-degree;
// the FP64 dangling mass of the start goes to per-block partials (pr = 0).
__global__ __launch_bounds__(kBlock) void na_init_kernel(uint32_t V, uint32_t seed, const uint32_t *outdeg,
                                                         float *rank, float *contrib, double *pr, double *pd) {
  PrAcc acc;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const float r = seed == 0xFFFFFFFFu ? 1.0f / (float)V : (v == seed ? 1.0f : 0.0f);
    const uint32_t od = outdeg[v];
    rank[v] = r;
    contrib[v] = od ? r / (float)od : 0.0f;
    if (!od) acc.dang += (double)r;
  }
  acc.resid = 0.0;
  block_store_acc<kWaves>(acc, pr, pd);
}

// The FP64 partials of n blocks summed in a fixed order: the residual and dangling mass of the
// ranks just written, and the next iteration's teleport terms.
__global__ __launch_bounds__(kBlock) void na_finalize_kernel(const double *pr, const double *pd, uint32_t n,
                                                             uint32_t V, float d, int personalised, PrScal *sc) {
  __shared__ double sr[kBlock], sd[kBlock];
  double r = 0, dd = 0;
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
    r += pr[i];
    dd += pd[i];
  }
  sr[threadIdx.x] = r;
  sd[threadIdx.x] = dd;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
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
    sc->base = personalised ? 0.0f : (float)(tele / (double)V);
    sc->extra = personalised ? (float)tele : 0.0f;
  }
}

// One iteration, a wave per destination row: lanes take every 64th in-edge with a Kahan sum each,
// the hi/lo pairs are combined by a butterfly with a fixed order, and lane 0 applies the update.
__global__ __launch_bounds__(kBlock) void na_gather_kernel(NaArgs a) {
  const uint32_t lane = lane_id();
  const float base = a.sc->base, extra = a.sc->extra;
  PrAcc acc;
  for (uint32_t r = blockIdx.x * kWaves + wave_id(); r < a.V; r += gridDim.x * kWaves) {
    const uint64_t e0 = a.off[r], e1 = a.off[r + 1];
    float s = 0.0f, c = 0.0f;
    if (a.kahan) {
      for (uint64_t k = e0 + lane; k < e1; k += kWave) kahan_add(s, c, a.contrib_in[a.idx[k]]);
    } else {
      for (uint64_t k = e0 + lane; k < e1; k += kWave) s += a.contrib_in[a.idx[k]];
    }
    float hi = s, lo = -c;  // true sum ~= s - c
    wave_reduce_dd(hi, lo);
    if (lane == 0) pr_apply(a, r, hi, base, extra, acc);
  }
  block_store_acc<kWaves>(acc, a.pr, a.pd);
}

__global__ void na_copy_kernel(const float *src, float *dst, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

}  // namespace

void noalloc_attach(Engine &e) {
  if (e.device < 0 || e.device >= kMaxDevices) fail(GMP_EINVAL, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[e.device]) {
    void *p = nullptr;
    GMP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the caller made e.device current
    if (!p) fail(GMP_EHIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch[e.device] = p;
  }
  e.scratch = g_scratch[e.device];
}

}  // namespace gmp

extern "C" int gmp_pagerank_noalloc(gmp_engine *h, uint32_t V, uint64_t E, const uint64_t *d_row_off,
                                    const uint32_t *d_col_idx, const uint32_t *d_out_degree,
                                    const gmp_pr_params *p, float *d_rank_out, gmp_stats *st) {
  using namespace gmp;
  GMP_API_BEGIN
  if (!h || !p) fail(GMP_EINVAL, "gmp_pagerank_noalloc: null engine or parameters");
  Engine &e = *reinterpret_cast<Engine *>(h);
  if (V == 0 || V > GMP_NOALLOC_MAX_VERTICES)
    fail(GMP_EINVAL, "gmp_pagerank_noalloc: V = " + std::to_string(V) + " is outside 1.." +
                         std::to_string(GMP_NOALLOC_MAX_VERTICES) + " (the size of the scratch array)");
  if (!d_row_off || !d_out_degree || !d_rank_out || (E && !d_col_idx))
    fail(GMP_EINVAL, "gmp_pagerank_noalloc: null device pointer");
  if (!(p->damping >= 0.0f && p->damping <= 1.0f)) fail(GMP_EINVAL, "damping must be in [0, 1]");
  if (p->seed_vertex >= (int64_t)V) fail(GMP_EINVAL, "seed vertex out of range");
  if (p->flags & ~GMP_PR_NO_KAHAN) fail(GMP_EUNSUPPORTED, "gmp_pagerank_noalloc: the CSR holds plain column ids");
  if (!e.scratch) fail(GMP_ESTATE, "gmp_pagerank_noalloc: the engine has no scratch array");
  GMP_HIP(hipSetDevice(e.device));

  // scratch: rank buffer, two contribution buffers, FP64 partials, scalars
  char *base = static_cast<char *>(e.scratch);
  float *rank_b = reinterpret_cast<float *>(base);
  float *contrib[2] = {rank_b + GMP_NOALLOC_MAX_VERTICES, rank_b + 2ull * GMP_NOALLOC_MAX_VERTICES};
  double *pr = reinterpret_cast<double *>(base + 3ull * GMP_NOALLOC_MAX_VERTICES * sizeof(float));
  double *pd = pr + kMaxBlocks;
  PrScal *sc = reinterpret_cast<PrScal *>(base + 3ull * GMP_NOALLOC_MAX_VERTICES * sizeof(float) + kPartBytes);
  float *rank[2] = {d_rank_out, rank_b};

  const bool personalised = p->seed_vertex >= 0;
  const uint32_t seed = personalised ? (uint32_t)p->seed_vertex : 0xFFFFFFFFu;
  const uint32_t gi = (uint32_t)std::min<uint64_t>(((uint64_t)V + kBlock - 1) / kBlock, kMaxBlocks);
  const uint32_t gg = (uint32_t)std::min<uint64_t>(((uint64_t)V + kWaves - 1) / kWaves, kMaxBlocks);
  const int max_iter = std::max(0, p->max_iter);
  hipStream_t s = e.stream;

  GMP_HIP(hipEventRecord(e.na_ev[0], s));
  na_init_kernel<<<gi, kBlock, 0, s>>>(V, seed, d_out_degree, rank[0], contrib[0], pr, pd);
  na_finalize_kernel<<<1, kBlock, 0, s>>>(pr, pd, gi, V, p->damping, personalised, sc);
  GMP_CHECK_LAUNCH();
  NaArgs a{};
  a.off = d_row_off;
  a.idx = d_col_idx;
  a.outdeg = d_out_degree;
  a.sc = sc;
  a.d = p->damping;
  a.seed = seed;
  a.V = V;
  a.kahan = (p->flags & GMP_PR_NO_KAHAN) ? 0 : 1;
  a.pr = pr;
  a.pd = pd;
  int cur = 0, it = 0;
  PrScal h{};
  for (it = 0; it < max_iter;) {
    a.contrib_in = contrib[cur];
    a.rank_in = rank[cur];
    a.rank_out = rank[cur ^ 1];
    a.contrib_out = contrib[cur ^ 1];
    na_gather_kernel<<<gg, kBlock, 0, s>>>(a);
    na_finalize_kernel<<<1, kBlock, 0, s>>>(pr, pd, gg, V, p->damping, personalised, sc);
    GMP_CHECK_LAUNCH();
    cur ^= 1;
    it++;
    if (p->tol > 0) {  // the L1 change of this iteration decides whether another one runs
      GMP_HIP(hipMemcpyAsync(e.h_pinned, sc, sizeof(PrScal), hipMemcpyDeviceToHost, s));
      GMP_HIP(hipStreamSynchronize(s));
      std::memcpy(&h, e.h_pinned, sizeof(PrScal));
      if (h.resid < p->tol) break;
    }
  }
  if (cur == 1) {  // the last ranks are in the scratch buffer
    na_copy_kernel<<<gi, kBlock, 0, s>>>(rank[1], d_rank_out, V);
    GMP_CHECK_LAUNCH();
  }
  GMP_HIP(hipEventRecord(e.na_ev[1], s));
  GMP_HIP(hipMemcpyAsync(e.h_pinned, sc, sizeof(PrScal), hipMemcpyDeviceToHost, s));
  GMP_HIP(hipStreamSynchronize(s));
  std::memcpy(&h, e.h_pinned, sizeof(PrScal));
  if (st) {
    std::memset(st, 0, sizeof *st);
    float ms = 0.0f;
    GMP_HIP(hipEventElapsedTime(&ms, e.na_ev[0], e.na_ev[1]));
    st->iterations = it;
    st->launches = it;
    st->ms_total = ms;
    st->ms_kernel = ms;
    st->edges = E * (uint64_t)it;
    st->vertices = (uint64_t)V * (uint64_t)it;
    st->residual = it ? h.resid : 0.0;
    st->dangling = h.dang;
    st->bytes = (8 * E + 24ull * V) * (uint64_t)it;
  }
  GMP_API_END
}
