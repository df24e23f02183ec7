// Connected components of a symmetric graph by sampled union-find, shared by the resident
// engine (traverse.hip) and semi-external graphs (window.hip).
//
// Afforest (Sutton, Ben-Nun, Barak: "Optimizing Parallel Graph Connectivity Computation via
// Subgraph Sampling", IPDPS 2018): every vertex is linked to its first kAfSample neighbours, the
// trees are compressed, the component holding most of a vertex sample is taken as the giant one,
// and only vertices outside it link their remaining neighbours -- on a symmetric graph an edge
// into the giant component is also an edge out of the other endpoint, so nothing is lost.  Trees
// hang from their least (work) id; the label of a component is its least original id, the fixed
// point of GraphMP's min-label rule (cc.cpp:40-45).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "gmp_internal.hpp"
#include "wave.hpp"

namespace gmp {
namespace {

constexpr uint32_t kAfSample = 2;
constexpr uint32_t kAfProbe = 1024;
constexpr uint32_t kAfNone = 0xFFFFFFFFu;  // no such neighbour
constexpr int kAfBlock = 256;
constexpr int kAfWaves = kAfBlock / kWave;

// Hooks the trees of u and v: the larger root is swung to the smaller with a compare-and-swap;
// otherwise both climb.  Every value ever stored in p[x] is <= x, so whatever (possibly stale)
// values the plain loads return, max(a, b) drops on every pass and the loop ends; after a failed
// swing the climb starts from the value the compare-and-swap found, not from a reload.
__device__ inline void af_link(uint32_t *p, uint32_t u, uint32_t v) {
  uint32_t a = p[u], b = p[v];
  while (a != b) {
    const uint32_t hi = max(a, b), lo = min(a, b);
    uint32_t ph = p[hi];
    if (ph == lo) return;  // hi hangs from lo: the trees are one
    if (ph == hi) {
      ph = atomicCAS(&p[hi], hi, lo);
      if (ph == hi) return;
    }
    a = p[ph];  // ph < hi
    b = p[lo];
  }
}

__global__ void af_init_kernel(uint32_t *p, uint32_t V) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) p[v] = v;
}

// Neighbour r of every vertex, from the CSR.
template <class OffT>
__global__ void af_sample_kernel(const OffT *off, const uint32_t *idx, uint32_t V, uint32_t r, uint32_t *p) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const uint64_t b = off[v];
    if ((uint64_t)off[v + 1] - b > r) af_link(p, v, idx[b + r]);
  }
}

// Neighbour r of every vertex, from a table of the first kAfSample neighbours (kAfNone: absent).
__global__ void af_sample_table_kernel(const uint32_t *first, uint32_t V, uint32_t r, uint32_t *p) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const uint32_t u = first[(uint64_t)v * kAfSample + r];
    if (u != kAfNone) af_link(p, v, u);
  }
}

// Pointer jumping until every vertex points at its root (the chain strictly decreases).
__global__ void af_compress_kernel(uint32_t *p, uint32_t V) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    uint32_t x = p[v];
    while (p[x] != x) x = p[x];
    p[v] = x;
  }
}

// Roots of kAfProbe vertices spread over the id range (after compression p is the root).
__global__ void af_probe_kernel(const uint32_t *p, uint32_t V, uint32_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kAfProbe) out[i] = p[(uint32_t)(((uint64_t)mix64(i + 1) >> 11) % V)];
}

// Rows [v0, v1) outside the giant component link their neighbours past the sampled ones (ids by
// global edge index); one lane per short list, the whole wave per list longer than a wave.
template <class OffT>
__global__ __launch_bounds__(kAfBlock) void af_finish_kernel(const OffT *off, const uint32_t *idx, uint32_t v0,
                                                             uint32_t v1, uint32_t giant, uint32_t *p,
                                                             unsigned long long *edges) {
  const uint32_t lane = lane_id();
  unsigned long long n = 0;
  for (uint32_t b0 = v0 + blockIdx.x * blockDim.x; b0 < v1; b0 += gridDim.x * blockDim.x) {
    const uint32_t v = b0 + threadIdx.x;
    uint64_t b = 0, en = 0;
    bool mine = false;
    if (v < v1 && p[v] != giant) {
      b = (uint64_t)off[v] + kAfSample;
      en = off[v + 1];
      mine = en > b;
    }
    const bool wide = mine && en - b > kWave;
    if (mine && !wide) {
      for (uint64_t k = b; k < en; k++) af_link(p, v, idx[k]);
      n += en - b;
    }
    uint64_t wmask = __ballot(wide);
    while (wmask) {
      const uint32_t l = (uint32_t)__ffsll((long long)wmask) - 1u;
      wmask &= wmask - 1;
      const uint32_t u = read_lane_u32(v, (int)l);
      const uint64_t lb = ((uint64_t)read_lane_u32((uint32_t)(b >> 32), (int)l) << 32) | read_lane_u32((uint32_t)b, (int)l);
      const uint64_t le = ((uint64_t)read_lane_u32((uint32_t)(en >> 32), (int)l) << 32) | read_lane_u32((uint32_t)en, (int)l);
      for (uint64_t k = lb + lane; k < le; k += kWave) af_link(p, u, idx[k]);
      if (lane == 0) n += le - lb;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) n += (unsigned long long)__shfl_xor((long long)n, o, kWave);
  if (lane == 0 && n) atomicAdd(edges, n);
}

// minorig[root] = least original id of the component (the giant component's minimum reduced
// per block first: one atomic per block on its root).
__global__ __launch_bounds__(kAfBlock) void af_min_kernel(const uint32_t *p, const uint32_t *n2o, uint32_t V,
                                                          uint32_t giant, uint32_t *minorig) {
  __shared__ uint32_t sm[kAfWaves];
  uint32_t g = 0xFFFFFFFFu;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const uint32_t r = p[v], o = n2o ? n2o[v] : v;
    if (r == giant) g = min(g, o);
    else atomicMin(&minorig[r], o);
  }
  g = wave_reduce_min(g);
  if (lane_id() == 0) sm[wave_id()] = g;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kAfWaves; w++) g = min(g, sm[w]);
    if (g != 0xFFFFFFFFu) atomicMin(&minorig[giant], g);
  }
}

// label[v] = minorig[p[v]]; label may be minorig itself (a root's slot keeps its own value and no
// other slot is read as a root).
__global__ void af_label_kernel(const uint32_t *p, const uint32_t *minorig, uint32_t V, uint32_t *label) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) label[v] = minorig[p[v]];
}

// The root most of the probes hang from.
inline uint32_t af_giant(std::vector<uint32_t> roots) {
  std::sort(roots.begin(), roots.end());
  uint32_t giant = roots.empty() ? 0u : roots[0], best = 0;
  for (size_t i = 0; i < roots.size();) {
    size_t j = i;
    while (j < roots.size() && roots[j] == roots[i]) j++;
    if (j - i > best) {
      best = (uint32_t)(j - i);
      giant = roots[i];
    }
    i = j;
  }
  return giant;
}

}  // namespace
}  // namespace gmp
