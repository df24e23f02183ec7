// Device-resident graph construction: upload, R-MAT generation, CSR assembly, transpose.
//
// The whole graph lives in HBM as a destination-row CSR (the layout of GraphMP's shards,
// GraphMP/tools/format_spark.py:137-157) with row offsets in uint32 when E < 2^32 and uint64
// otherwise, uint32 column ids and optional uint32 weights.  Generation and transposition
// are sort-based on the device: one 64-bit key per edge (row << shift | column), an LSD radix
// sort over exactly 2*shift bits, a unique pass, and a single kernel that turns the sorted
// keys into row offsets.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "gmp_internal.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;

inline uint32_t grid_for(uint64_t n, uint32_t block = kBlock, uint32_t cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

inline uint32_t bits_for(uint32_t V) {
  uint32_t b = 0;
  while (b < 32 && (1ull << b) < V) b++;
  return std::max<uint32_t>(b, 1);
}

struct RmatConsts {
  uint64_t base;
  uint64_t pk[4];  // permutation keys
  uint32_t t1, t2, t3;
  uint32_t scale;
  uint32_t half;   // (scale + 1) / 2
  uint64_t mask;
};

__device__ inline uint32_t perm_dev(uint64_t x, const RmatConsts &c) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    x ^= c.pk[j] & c.mask;
    x = (x * ((c.pk[j] >> 32) | 1ull)) & c.mask;
    x ^= x >> c.half;
  }
  return (uint32_t)x;
}

// One sample per thread (grid-stride).  Self-loops become the key of (V-1, V-1), itself a
// self-loop, so they sort to the end and the unique pass leaves at most one to drop.
__global__ void rmat_keys_kernel(RmatConsts c, uint64_t m, uint64_t *keys, int undirected) {
  const uint64_t sentinel = (c.mask << c.scale) | c.mask;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t st = c.base ^ i;
    uint64_t src = 0, dst = 0;
    for (uint32_t k = 0; k < c.half; k++) {
      uint64_t z = mix64(st + (uint64_t)(k + 1) * kGamma);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        uint32_t lvl = 2 * k + h;
        if (lvl < c.scale) {
          uint32_t r = (uint32_t)(z >> (32 * h));
          uint64_t sb = r >= c.t2;
          uint64_t db = (r >= c.t1 && r < c.t2) || r >= c.t3;
          src = (src << 1) | sb;
          dst = (dst << 1) | db;
        }
      }
    }
    uint64_t s = perm_dev(src, c), d = perm_dev(dst, c);
    bool loop = (s == d);
    keys[i] = loop ? sentinel : ((d << c.scale) | s);
    if (undirected) keys[m + i] = loop ? sentinel : ((s << c.scale) | d);
  }
}

// off[r] = first j with row(key[j]) >= r, for r in [0, V]; key rows are nondecreasing.
template <class OffT>
__global__ void offsets_from_keys_kernel(const uint64_t *keys, uint64_t E, uint32_t shift,
                                         uint32_t V, OffT *off) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j <= E;
       j += (uint64_t)gridDim.x * blockDim.x) {
    int64_t r = j < E ? (int64_t)(keys[j] >> shift) : (int64_t)V;
    int64_t rp = j > 0 ? (int64_t)(keys[j - 1] >> shift) : -1;
    for (int64_t rr = rp + 1; rr <= r; rr++) off[rr] = (OffT)j;
  }
}

__global__ void cols_from_keys_kernel(const uint64_t *keys, uint64_t E, uint64_t mask,
                                      uint32_t *idx) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
       j += (uint64_t)gridDim.x * blockDim.x)
    idx[j] = (uint32_t)(keys[j] & mask);
}

__global__ void rmat_weights_kernel(const uint64_t *keys, uint64_t E, uint32_t shift,
                                    uint64_t mask, uint64_t seed2, uint32_t *w) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
       j += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t dst = keys[j] >> shift, src = keys[j] & mask;
    w[j] = 1u + (uint32_t)(mix64(seed2 ^ (src << 32) ^ dst) % 255ull);
  }
}

__global__ void out_degree_kernel(const uint32_t *idx, uint64_t E, uint32_t *deg) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
       j += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&deg[idx[j]], 1u);
}

// rowid[off[r]] = r for every non-empty row (unique writer per position); a max-scan
// then labels every edge with its row.
template <class OffT>
__global__ void mark_rows_kernel(const OffT *off, uint32_t V, uint32_t *rowid) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x)
    if (off[r + 1] > off[r]) rowid[off[r]] = (uint32_t)r;
}

__global__ void transpose_keys_kernel(const uint32_t *idx, const uint32_t *rowid, uint64_t E,
                                      uint32_t shift, uint64_t *keys) {
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < E;
       j += (uint64_t)gridDim.x * blockDim.x)
    keys[j] = ((uint64_t)idx[j] << shift) | rowid[j];
}

template <class OffT>
__global__ void widen_offsets_kernel(const OffT *off, uint32_t n, uint64_t *out) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n;
       r += (uint64_t)gridDim.x * blockDim.x)
    out[r] = (uint64_t)off[r];
}

uint32_t thresh(double x) {
  double t = std::floor(x * 4294967296.0);
  if (t < 0) t = 0;
  if (t > 4294967295.0) t = 4294967295.0;
  return (uint32_t)t;
}

// Sorted 64-bit keys (row << shift | col), already unique -> CSR arrays in g (in or out side).
void csr_from_sorted_keys(Engine &e, const uint64_t *keys, uint64_t E, uint32_t shift, uint32_t V,
                          bool off64, DevBuf &off, DevBuf &idx) {
  off.alloc((size_t)(V + 1) * (off64 ? 8 : 4));
  idx.alloc(std::max<uint64_t>(E, 1) * 4);
  uint32_t g = grid_for(E + 1);
  if (off64)
    offsets_from_keys_kernel<uint64_t><<<g, kBlock, 0, e.stream>>>(keys, E, shift, V, off.as<uint64_t>());
  else
    offsets_from_keys_kernel<uint32_t><<<g, kBlock, 0, e.stream>>>(keys, E, shift, V, off.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  if (E) {
    cols_from_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(keys, E, (1ull << shift) - 1, idx.as<uint32_t>());
    GMP_CHECK_LAUNCH();
  }
}

}  // namespace

void graph_upload(Engine &e, uint32_t V, uint64_t E, const uint64_t *row_off,
                  const uint32_t *col_idx, const uint32_t *weights, bool force_off64) {
  Graph &g = e.g;
  g = Graph();
  g.V = V;
  g.E = E;
  g.off64 = force_off64 || E > 0xFFFFFFFFull;
  g.weighted = weights != nullptr;
  g.in_off.alloc((size_t)(V + 1) * (g.off64 ? 8 : 4));
  if (g.off64) {
    GMP_HIP(hipMemcpyAsync(g.in_off.p, row_off, (size_t)(V + 1) * 8, hipMemcpyHostToDevice, e.stream));
  } else {
    std::vector<uint32_t> o32(V + 1);
    for (uint64_t r = 0; r <= V; r++) o32[r] = (uint32_t)row_off[r];
    GMP_HIP(hipMemcpyAsync(g.in_off.p, o32.data(), (size_t)(V + 1) * 4, hipMemcpyHostToDevice, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
  }
  g.in_idx.alloc(std::max<uint64_t>(E, 1) * 4);
  if (E) GMP_HIP(hipMemcpyAsync(g.in_idx.p, col_idx, E * 4, hipMemcpyHostToDevice, e.stream));
  if (weights) {
    g.in_w.alloc(std::max<uint64_t>(E, 1) * 4);
    if (E) GMP_HIP(hipMemcpyAsync(g.in_w.p, weights, E * 4, hipMemcpyHostToDevice, e.stream));
  }
  g.h_in_off.assign(row_off, row_off + V + 1);
  GMP_HIP(hipStreamSynchronize(e.stream));
}

void graph_generate_rmat(Engine &e, const gmp_rmat_params &p) {
  if (p.scale < 1 || p.scale > 31) fail(GMP_EINVAL, "rmat: scale must be in [1, 31]");
  if (p.edge_factor < 1) fail(GMP_EINVAL, "rmat: edge_factor must be >= 1");
  if (!(p.a >= 0 && p.b >= 0 && p.c >= 0 && p.a + p.b + p.c <= 1.0))
    fail(GMP_EINVAL, "rmat: probabilities must be >= 0 with a + b + c <= 1");
  Graph &g = e.g;
  g = Graph();
  const uint32_t V = 1u << p.scale;
  const uint64_t m = (uint64_t)p.edge_factor << p.scale;
  const uint64_t nkeys = p.undirected ? 2 * m : m;
  RmatConsts c;
  c.base = mix64(p.seed + kGamma);
  for (int j = 0; j < 4; j++) c.pk[j] = mix64(p.seed + 1 + (uint64_t)(j + 1) * kGamma);
  c.t1 = thresh(p.a);
  c.t2 = thresh(p.a + p.b);
  c.t3 = thresh(p.a + p.b + p.c);
  c.scale = p.scale;
  c.half = (p.scale + 1) / 2;
  c.mask = (1ull << p.scale) - 1;

  DevBuf kA, kB, tmp, cnt;
  kA.alloc(nkeys * 8);
  kB.alloc(nkeys * 8);
  rmat_keys_kernel<<<grid_for(m), kBlock, 0, e.stream>>>(c, m, kA.as<uint64_t>(), p.undirected ? 1 : 0);
  GMP_CHECK_LAUNCH();
  rocprim::double_buffer<uint64_t> db(kA.as<uint64_t>(), kB.as<uint64_t>());
  size_t tb = 0;
  GMP_HIP(rocprim::radix_sort_keys(nullptr, tb, db, nkeys, 0, 2 * p.scale, e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::radix_sort_keys(tmp.p, tb, db, nkeys, 0, 2 * p.scale, e.stream));
  uint64_t *sorted = db.current();
  uint64_t *uniq = (sorted == kA.as<uint64_t>()) ? kB.as<uint64_t>() : kA.as<uint64_t>();
  cnt.alloc(8);
  size_t tb2 = 0;
  GMP_HIP(rocprim::unique(nullptr, tb2, sorted, uniq, cnt.as<uint64_t>(), nkeys,
                          rocprim::equal_to<uint64_t>(), e.stream));
  tmp.alloc(tb2);
  GMP_HIP(rocprim::unique(tmp.p, tb2, sorted, uniq, cnt.as<uint64_t>(), nkeys,
                          rocprim::equal_to<uint64_t>(), e.stream));
  uint64_t nuniq = 0, last = 0;
  GMP_HIP(hipMemcpyAsync(&nuniq, cnt.p, 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (nuniq) {
    GMP_HIP(hipMemcpy(&last, uniq + nuniq - 1, 8, hipMemcpyDeviceToHost));
    const uint64_t sentinel = (c.mask << p.scale) | c.mask;
    if (last == sentinel) nuniq--;
  }
  g.V = V;
  g.E = nuniq;
  g.off64 = (p.flags & GMP_GRAPH_OFF64) || nuniq > 0xFFFFFFFFull;
  g.symmetric = p.undirected != 0;
  csr_from_sorted_keys(e, uniq, nuniq, p.scale, V, g.off64, g.in_off, g.in_idx);
  if (p.weighted) {
    g.weighted = true;
    g.in_w.alloc(std::max<uint64_t>(nuniq, 1) * 4);
    if (nuniq) {
      rmat_weights_kernel<<<grid_for(nuniq), kBlock, 0, e.stream>>>(uniq, nuniq, p.scale, c.mask, p.seed + 2,
                                                                   g.in_w.as<uint32_t>());
      GMP_CHECK_LAUNCH();
    }
  }
  GMP_HIP(hipStreamSynchronize(e.stream));
  // host copy of the offsets for partitioning and gather planning
  g.h_in_off.resize((size_t)V + 1);
  if (g.off64) {
    GMP_HIP(hipMemcpy(g.h_in_off.data(), g.in_off.p, (size_t)(V + 1) * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint32_t> o32((size_t)V + 1);
    GMP_HIP(hipMemcpy(o32.data(), g.in_off.p, (size_t)(V + 1) * 4, hipMemcpyDeviceToHost));
    for (size_t r = 0; r <= V; r++) g.h_in_off[r] = o32[r];
  }
}

void graph_finish(Engine &e, uint32_t P, const uint32_t *bounds) {
  Graph &g = e.g;
  g.out_deg.alloc((size_t)std::max<uint32_t>(g.V, 1) * 4);
  GMP_HIP(hipMemsetAsync(g.out_deg.p, 0, (size_t)g.V * 4, e.stream));
  if (g.E) {
    out_degree_kernel<<<grid_for(g.E), kBlock, 0, e.stream>>>(g.in_idx.as<uint32_t>(), g.E, g.out_deg.as<uint32_t>());
    GMP_CHECK_LAUNCH();
  }
  if (P == 0) P = 1;
  g.P = P;
  g.bounds.assign(P + 1, 0);
  if (bounds) {
    if (bounds[0] != 0 || bounds[P] != g.V) fail(GMP_EINVAL, "shard bounds must run from 0 to V");
    for (uint32_t i = 0; i < P; i++)
      if (bounds[i + 1] < bounds[i]) fail(GMP_EINVAL, "shard bounds must be nondecreasing");
    std::copy(bounds, bounds + P + 1, g.bounds.begin());
  } else {
    std::vector<uint64_t> indeg(g.V);
    for (uint32_t v = 0; v < g.V; v++) indeg[v] = g.h_in_off[v + 1] - g.h_in_off[v];
    partition_bounds(indeg.data(), g.V, P, g.bounds.data());
  }
  GMP_HIP(hipStreamSynchronize(e.stream));
}

void graph_build_out(Engine &e) {
  Graph &g = e.g;
  if (g.has_out) return;
  if (g.symmetric && !g.weighted) {
    // a symmetric graph is its own transpose: share the in-CSR
    g.has_out = true;
    return;
  }
  const uint32_t shift = bits_for(g.V);
  const uint64_t E = g.E;
  DevBuf rowid, kA, kB, wB, tmp;
  rowid.alloc(std::max<uint64_t>(E, 1) * 4);
  kA.alloc(std::max<uint64_t>(E, 1) * 8);
  kB.alloc(std::max<uint64_t>(E, 1) * 8);
  if (E) {
    GMP_HIP(hipMemsetAsync(rowid.p, 0, E * 4, e.stream));
    if (g.off64)
      mark_rows_kernel<uint64_t><<<grid_for(g.V), kBlock, 0, e.stream>>>(g.in_off.as<uint64_t>(), g.V, rowid.as<uint32_t>());
    else
      mark_rows_kernel<uint32_t><<<grid_for(g.V), kBlock, 0, e.stream>>>(g.in_off.as<uint32_t>(), g.V, rowid.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    size_t tb = 0;
    GMP_HIP(rocprim::inclusive_scan(nullptr, tb, rowid.as<uint32_t>(), rowid.as<uint32_t>(), E,
                                    rocprim::maximum<uint32_t>(), e.stream));
    tmp.alloc(tb);
    GMP_HIP(rocprim::inclusive_scan(tmp.p, tb, rowid.as<uint32_t>(), rowid.as<uint32_t>(), E,
                                    rocprim::maximum<uint32_t>(), e.stream));
    transpose_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(g.in_idx.as<uint32_t>(), rowid.as<uint32_t>(), E,
                                                               shift, kA.as<uint64_t>());
    GMP_CHECK_LAUNCH();
  }
  rocprim::double_buffer<uint64_t> dk(kA.as<uint64_t>(), kB.as<uint64_t>());
  if (E) {
    size_t tb = 0;
    if (g.weighted) {
      wB.alloc(E * 4);
      // the in-CSR weights are consumed as the value buffer's first half
      DevBuf wA;
      wA.alloc(E * 4);
      GMP_HIP(hipMemcpyAsync(wA.p, g.in_w.p, E * 4, hipMemcpyDeviceToDevice, e.stream));
      rocprim::double_buffer<uint32_t> dv(wA.as<uint32_t>(), wB.as<uint32_t>());
      GMP_HIP(rocprim::radix_sort_pairs(nullptr, tb, dk, dv, E, 0, 2 * shift, e.stream));
      tmp.alloc(tb);
      GMP_HIP(rocprim::radix_sort_pairs(tmp.p, tb, dk, dv, E, 0, 2 * shift, e.stream));
      if (dv.current() == wA.as<uint32_t>()) std::swap(wA, wB);
      g.out_w = std::move(wB);
      GMP_HIP(hipStreamSynchronize(e.stream));
    } else {
      GMP_HIP(rocprim::radix_sort_keys(nullptr, tb, dk, E, 0, 2 * shift, e.stream));
      tmp.alloc(tb);
      GMP_HIP(rocprim::radix_sort_keys(tmp.p, tb, dk, E, 0, 2 * shift, e.stream));
    }
  }
  csr_from_sorted_keys(e, dk.current(), E, shift, g.V, g.off64, g.out_off, g.out_idx);
  GMP_HIP(hipStreamSynchronize(e.stream));
  g.has_out = true;
}

void graph_download(Engine &e, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights,
                    uint32_t *out_degree) {
  Graph &g = e.g;
  if (row_off) std::memcpy(row_off, g.h_in_off.data(), (size_t)(g.V + 1) * 8);
  if (col_idx && g.E) GMP_HIP(hipMemcpy(col_idx, g.in_idx.p, g.E * 4, hipMemcpyDeviceToHost));
  if (weights) {
    if (!g.weighted) fail(GMP_EUNSUPPORTED, "graph has no weights");
    if (g.E) GMP_HIP(hipMemcpy(weights, g.in_w.p, g.E * 4, hipMemcpyDeviceToHost));
  }
  if (out_degree && g.V) GMP_HIP(hipMemcpy(out_degree, g.out_deg.p, (size_t)g.V * 4, hipMemcpyDeviceToHost));
}

}  // namespace gmp
