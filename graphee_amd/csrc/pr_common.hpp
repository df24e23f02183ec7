// Pieces shared by the PageRank kernels of pagerank.hip (row gather) and strip.hip (source
// strips): the per-iteration scalars, the update of one vertex, the deterministic FP64 block
// reductions, and the exact fixed-point form of a contribution.
#pragma once

#include "gmp_internal.hpp"
#include "wave.hpp"

namespace gmp {

struct PrScal {
  float base;    // added to every vertex
  float extra;   // added to the seed vertex only
  double resid;  // L1 change of the last iteration
  double dang;   // dangling mass of the current ranks
  int32_t done;  // the solve has converged: iteration launches queued after that do nothing
};

struct PrAcc {
  double resid = 0, dang = 0;
};

// r'[v] = d*s + base (+ extra at the seed); c'[v] = r'[v] / outdeg[v], given outdeg[v] = od
// and the old rank.  `A` carries outdeg, rank_in, rank_out, contrib_out, d and seed.
template <class A>
__device__ inline void pr_update(const A &a, uint32_t r, float s, float base, float extra, uint32_t od, float old,
                                 PrAcc &acc) {
  float v = fmaf(a.d, s, base);
  if (r == a.seed) v += extra;
  __builtin_nontemporal_store(v, &a.rank_out[r]);
  a.contrib_out[r] = od ? v / (float)od : 0.0f;
  if (!od) acc.dang += (double)v;
  acc.resid += fabs((double)v - (double)old);
}

template <class A>
__device__ inline void pr_apply(const A &a, uint32_t r, float s, float base, float extra, PrAcc &acc) {
  const uint32_t od = __builtin_nontemporal_load(&a.outdeg[r]);
  const float old = __builtin_nontemporal_load(&a.rank_in[r]);
  pr_update(a, r, s, base, extra, od, old, acc);
}

// pr_apply over rows v0 + i, i = first, first + stride, ... < n, U rows' loads in flight at a
// time; sum(i) gives the row's gathered sum.
template <int U, class A, class F>
__device__ inline void pr_apply_rows(const A &a, uint32_t v0, uint32_t n, uint32_t first, uint32_t stride, F sum,
                                     float base, float extra, PrAcc &acc) {
  for (uint32_t i0 = first; i0 < n; i0 += stride * U) {
    uint32_t od[U];
    float old[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = i0 + u * stride;
      if (i < n) {
        od[u] = __builtin_nontemporal_load(&a.outdeg[v0 + i]);
        old[u] = __builtin_nontemporal_load(&a.rank_in[v0 + i]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = i0 + u * stride;
      if (i < n) pr_update(a, v0 + i, sum(i), base, extra, od[u], old[u], acc);
    }
  }
}

// Deterministic reduction of the two FP64 accumulators of a block of WAVES wavefronts into
// pr[blockIdx.x], pd[blockIdx.x].
template <int WAVES>
__device__ inline void block_store_acc(PrAcc acc, double *pr, double *pd) {
  __shared__ double sr[WAVES], sd[WAVES];
  double r = wave_reduce_sum(acc.resid), d = wave_reduce_sum(acc.dang);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sr[w] = r;
    sd[w] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0, td = 0;
    for (int i = 0; i < WAVES; i++) {
      tr += sr[i];
      td += sd[i];
    }
    pr[blockIdx.x] = tr;
    pd[blockIdx.x] = td;
  }
}

// Fixed point with 61 fraction bits: a non-negative float x < 4 maps to floor(x * 2^61), which
// is exact for x >= 2^-38 (all 24 mantissa bits land above bit 0).  Sums of such integers are
// exact and independent of order, so unordered accumulation stays deterministic.
constexpr int kFixBits = 61;

// Without branches (a message's conversion sits inside the bin kernel's unrolled loop): x >= 2,
// and any bit pattern with the sign set, saturates to 2^62 through the clamp to 2.0f's bits;
// zero and subnormals (< 2^-126, below the grid) shift right by 63 to 0.
__device__ inline unsigned long long fix_from_float(float x) {
  const uint32_t bits = min(__float_as_uint(x), 0x40000000u);
  const int32_t ex = (int32_t)(bits >> 23);
  const unsigned long long m = (bits & 0x7FFFFFu) | 0x800000u;
  // x = m * 2^(ex-150); x * 2^61 = m * 2^(ex-89)
  const unsigned long long l = m << (uint32_t)(max(ex, 89) - 89);
  const unsigned long long r = m >> (uint32_t)min(max(89 - ex, 0), 63);
  return ex >= 89 ? l : r;
}

__device__ inline double fix_to_double(unsigned long long s) { return (double)s * 0x1p-61; }

// One PageRank iteration over the source-strip layout (strip.hip): what it reads and writes.
struct PrStripIO {
  const uint32_t *outdeg;
  const float *contrib_in;
  const float *rank_in;
  float *rank_out;
  float *contrib_out;
  const PrScal *sc;
  float d;
  uint32_t seed;
  double *pr1, *pd1;  // FP64 residual/dangling partials of the bin kernel, one per block
  double *pr2, *pd2;  // and of the split-bin kernel
  bool packed;        // the gather reads the packed codes (ensure_strip_pack)
  bool ids;           // the gather reads 4-byte source ids (ensure_strip_ids)
};

// FP64 partials that strip_iteration may write into each of pr1, pd1, pr2, pd2.
uint32_t strip_partials(Engine &e);
// Launches one iteration on e.stream; n1 and n2 receive the partial counts written.
void strip_iteration(Engine &e, const PrStripIO &io, uint32_t &n1, uint32_t &n2);

}  // namespace gmp
