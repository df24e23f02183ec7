// Device-resident graph construction: upload, R-MAT generation, CSR assembly, transpose,
// and the degree-ordered work layout.
//
// The whole graph lives in HBM as a destination-row CSR (the layout of GraphMP's shards,
// GraphMP/tools/format_spark.py:137-157) with row offsets in uint32 when E < 2^32 and uint64
// otherwise, uint32 column ids and optional uint32 weights.  Generation, transposition and
// renumbering are sort-based on the device: one 64-bit key per edge (row << shift | column),
// an LSD radix sort over exactly 2*shift bits, and one kernel that turns sorted keys into
// row offsets.
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

inline uint32_t bits_for(uint64_t n) {
  uint32_t b = 0;
  while (b < 63 && (1ull << b) < n) b++;
  return std::max<uint32_t>(b, 1);
}

#define GRID_STRIDE(i, n) \
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

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
  GRID_STRIDE(i, m) {
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

// A run of rows [r0, r1] whose offsets are all `val`.
struct GapRun {
  uint32_t r0, r1;
  uint64_t val;
};
constexpr uint32_t kGapSerial = 256;     // rows an edge's thread fills itself
constexpr uint32_t kGapPiece = 1u << 16;  // rows of one gap_fill_kernel block

// off[r] = first j with row(key[j]) >= r, for r in [0, V]; key rows are nondecreasing.  The
// thread of edge j (j = E: the end) owns rows (row(key[j-1]), row(key[j])]: up to kGapSerial of
// them it writes itself, longer runs it leaves as pieces of kGapPiece rows in `gaps` (at most
// cap of them; *ngaps counts them) for gap_fill_kernel.
template <class OffT>
__global__ void offsets_from_keys_kernel(const uint64_t *keys, uint64_t E, uint32_t shift, uint32_t V, OffT *off,
                                         GapRun *gaps, uint32_t *ngaps) {
  GRID_STRIDE(j, E + 1) {
    const int64_t r = j < E ? (int64_t)(keys[j] >> shift) : (int64_t)V;
    const int64_t rp = j > 0 ? (int64_t)(keys[j - 1] >> shift) : -1;
    if (r - rp <= (int64_t)kGapSerial) {
      for (int64_t rr = rp + 1; rr <= r; rr++) off[rr] = (OffT)j;
    } else {
      const uint32_t np = (uint32_t)((r - rp + kGapPiece - 1) / kGapPiece);
      const uint32_t first = atomicAdd(ngaps, np);
      for (uint32_t k = 0; k < np; k++) {
        const int64_t a = rp + 1 + (int64_t)k * kGapPiece;
        gaps[first + k] = GapRun{(uint32_t)a, (uint32_t)std::min<int64_t>(a + kGapPiece - 1, r), (uint64_t)j};
      }
    }
  }
}

// One block per run left by offsets_from_keys_kernel.
template <class OffT>
__global__ void gap_fill_kernel(const GapRun *gaps, const uint32_t *ngaps, OffT *off) {
  const uint32_t n = *ngaps;
  for (uint32_t d = blockIdx.x; d < n; d += gridDim.x) {
    const GapRun g = gaps[d];
    for (uint64_t rr = (uint64_t)g.r0 + threadIdx.x; rr <= g.r1; rr += blockDim.x) off[rr] = (OffT)g.val;
  }
}

__global__ void cols_from_keys_kernel(const uint64_t *keys, uint64_t E, uint64_t mask, uint32_t *idx) {
  GRID_STRIDE(j, E) idx[j] = (uint32_t)(keys[j] & mask);
}

__global__ void rmat_weights_kernel(const uint64_t *keys, uint64_t E, uint32_t shift, uint64_t mask, uint64_t seed2,
                                    uint32_t *w) {
  GRID_STRIDE(j, E) {
    uint64_t dst = keys[j] >> shift, src = keys[j] & mask;
    w[j] = 1u + (uint32_t)(mix64(seed2 ^ (src << 32) ^ dst) % 255ull);
  }
}

__global__ void out_degree_kernel(const uint32_t *idx, uint64_t E, uint32_t *deg) {
  GRID_STRIDE(j, E) atomicAdd(&deg[idx[j]], 1u);
}

// rowid[off[r]] = r for every non-empty row (unique writer per position); a max-scan then
// labels every edge with its row.
template <class OffT>
__global__ void mark_rows_kernel(const OffT *off, uint32_t V, uint32_t *rowid) {
  GRID_STRIDE(r, V) if (off[r + 1] > off[r]) rowid[off[r]] = (uint32_t)r;
}

// key = map(col) << shift | map(row); map == nullptr is the identity.
__global__ void edge_keys_kernel(const uint32_t *idx, const uint32_t *rowid, uint64_t E, uint32_t shift,
                                 const uint32_t *map, int col_major, uint64_t *keys) {
  GRID_STRIDE(j, E) {
    uint64_t r = rowid[j], c = idx[j];
    if (map) {
      r = map[r];
      c = map[c];
    }
    keys[j] = col_major ? ((c << shift) | r) : ((r << shift) | c);
  }
}

// Work-layout order key: shard, then out-degree descending (stable sort keeps id order).
__global__ void order_keys_kernel(const uint32_t *outdeg, const uint32_t *bounds, uint32_t P, uint32_t V,
                                  uint64_t *keys, uint32_t *vals) {
  GRID_STRIDE(v, V) {
    uint32_t lo = 0, hi = P - 1;  // shard p with bounds[p] <= v < bounds[p+1]
    while (lo < hi) {
      uint32_t mid = (lo + hi + 1) >> 1;
      if (bounds[mid] <= v) lo = mid;
      else hi = mid - 1;
    }
    keys[v] = ((uint64_t)lo << 32) | (uint64_t)(0xFFFFFFFFu - outdeg[v]);
    vals[v] = (uint32_t)v;
  }
}

__global__ void scatter_inverse_kernel(const uint32_t *n2o, uint32_t V, uint32_t *o2n) {
  GRID_STRIDE(i, V) o2n[n2o[i]] = (uint32_t)i;
}

__global__ void gather_u32_kernel(const uint32_t *x, const uint32_t *map, uint32_t n, uint32_t *out) {
  GRID_STRIDE(i, n) out[i] = x[map[i]];
}

uint32_t thresh(double x) {
  double t = std::floor(x * 4294967296.0);
  if (t < 0) t = 0;
  if (t > 4294967295.0) t = 4294967295.0;
  return (uint32_t)t;
}

// Sorted unique keys (row << shift | col) -> CSR offsets and column ids.
void csr_from_sorted_keys(Engine &e, const uint64_t *keys, uint64_t E, uint32_t shift, uint32_t V, bool off64,
                          Csr &c) {
  c.off.alloc((size_t)(V + 1) * (off64 ? 8 : 4));
  c.idx.alloc(std::max<uint64_t>(E, 1) * 4);
  uint32_t g = grid_for(E + 1);
  // runs longer than kGapSerial rows are disjoint, each over kGapSerial rows long: at most
  // (V + 1) / kGapSerial of them, plus one piece boundary per kGapPiece rows
  const size_t cap = ((size_t)V + 1) / kGapSerial + ((size_t)V + 1) / kGapPiece + 2;
  e.s_gaps.alloc(256 + cap * sizeof(GapRun));
  uint32_t *ngaps = e.s_gaps.as<uint32_t>();
  GapRun *gaps = reinterpret_cast<GapRun *>(static_cast<char *>(e.s_gaps.p) + 256);
  GMP_HIP(hipMemsetAsync(ngaps, 0, 4, e.stream));
  if (off64) {
    offsets_from_keys_kernel<uint64_t><<<g, kBlock, 0, e.stream>>>(keys, E, shift, V, c.off.as<uint64_t>(), gaps, ngaps);
    gap_fill_kernel<uint64_t><<<1024, kBlock, 0, e.stream>>>(gaps, ngaps, c.off.as<uint64_t>());
  } else {
    offsets_from_keys_kernel<uint32_t><<<g, kBlock, 0, e.stream>>>(keys, E, shift, V, c.off.as<uint32_t>(), gaps, ngaps);
    gap_fill_kernel<uint32_t><<<1024, kBlock, 0, e.stream>>>(gaps, ngaps, c.off.as<uint32_t>());
  }
  GMP_CHECK_LAUNCH();
  if (E) {
    cols_from_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(keys, E, (1ull << shift) - 1, c.idx.as<uint32_t>());
    GMP_CHECK_LAUNCH();
  }
  c.valid = true;
}

// Row id of every edge of a CSR (mark + inclusive max-scan).
void edge_rows(Engine &e, const CsrView &src, uint32_t V, uint64_t E, bool off64, DevBuf &rowid, DevBuf &tmp) {
  rowid.alloc(std::max<uint64_t>(E, 1) * 4);
  if (!E) return;
  GMP_HIP(hipMemsetAsync(rowid.p, 0, E * 4, e.stream));
  if (off64) mark_rows_kernel<uint64_t><<<grid_for(V), kBlock, 0, e.stream>>>(src.offs<uint64_t>(), V, rowid.as<uint32_t>());
  else mark_rows_kernel<uint32_t><<<grid_for(V), kBlock, 0, e.stream>>>(src.offs<uint32_t>(), V, rowid.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  size_t tb = 0;
  GMP_HIP(rocprim::inclusive_scan(nullptr, tb, rowid.as<uint32_t>(), rowid.as<uint32_t>(), E,
                                  rocprim::maximum<uint32_t>(), e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::inclusive_scan(tmp.p, tb, rowid.as<uint32_t>(), rowid.as<uint32_t>(), E,
                                  rocprim::maximum<uint32_t>(), e.stream));
}

// Re-keys every edge of `src` (optionally renumbering vertices through `map` and/or
// swapping rows and columns), sorts, and builds `dst` -- weights travel with their edge.
void rebuild_csr(Engine &e, const CsrView &src, uint32_t V, uint64_t E, bool off64, bool weighted,
                 const uint32_t *map, bool transpose, Csr &dst) {
  const uint32_t shift = bits_for(V);
  DevBuf rowid, kA, kB, wA, wB, tmp;
  edge_rows(e, src, V, E, off64, rowid, tmp);
  kA.alloc(std::max<uint64_t>(E, 1) * 8);
  kB.alloc(std::max<uint64_t>(E, 1) * 8);
  if (E) {
    edge_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(src.idx, rowid.as<uint32_t>(), E, shift, map,
                                                          transpose ? 1 : 0, kA.as<uint64_t>());
    GMP_CHECK_LAUNCH();
  }
  rowid.release();
  rocprim::double_buffer<uint64_t> dk(kA.as<uint64_t>(), kB.as<uint64_t>());
  if (E) {
    size_t tb = 0;
    if (weighted) {
      wA.alloc(E * 4);
      wB.alloc(E * 4);
      GMP_HIP(hipMemcpyAsync(wA.p, src.w, E * 4, hipMemcpyDeviceToDevice, e.stream));
      rocprim::double_buffer<uint32_t> dv(wA.as<uint32_t>(), wB.as<uint32_t>());
      GMP_HIP(rocprim::radix_sort_pairs(nullptr, tb, dk, dv, E, 0, 2 * shift, e.stream));
      tmp.alloc(tb);
      GMP_HIP(rocprim::radix_sort_pairs(tmp.p, tb, dk, dv, E, 0, 2 * shift, e.stream));
      if (dv.current() == wA.as<uint32_t>()) std::swap(wA, wB);
      dst.w = std::move(wB);
    } else {
      GMP_HIP(rocprim::radix_sort_keys(nullptr, tb, dk, E, 0, 2 * shift, e.stream));
      tmp.alloc(tb);
      GMP_HIP(rocprim::radix_sort_keys(tmp.p, tb, dk, E, 0, 2 * shift, e.stream));
    }
  }
  csr_from_sorted_keys(e, dk.current(), E, shift, V, off64, dst);
  GMP_HIP(hipStreamSynchronize(e.stream));
}

void download_offsets(Engine &e, const Csr &c, uint32_t V, bool off64, std::vector<uint64_t> &h) {
  h.resize((size_t)V + 1);
  if (off64) {
    GMP_HIP(hipMemcpyAsync(h.data(), c.off.p, (size_t)(V + 1) * 8, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
  } else {
    std::vector<uint32_t> o32((size_t)V + 1);
    GMP_HIP(hipMemcpyAsync(o32.data(), c.off.p, (size_t)(V + 1) * 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    for (size_t r = 0; r <= V; r++) h[r] = o32[r];
  }
}

}  // namespace

void graph_upload(Engine &e, uint32_t V, uint64_t E, const uint64_t *row_off, const uint32_t *col_idx,
                  const uint32_t *weights, bool force_off64) {
  Graph &g = e.g;
  g = Graph();
  g.V = V;
  g.E = E;
  g.off64 = force_off64 || E > 0xFFFFFFFFull;
  g.weighted = weights != nullptr;
  g.in.off.alloc((size_t)(V + 1) * (g.off64 ? 8 : 4));
  if (g.off64) {
    GMP_HIP(hipMemcpyAsync(g.in.off.p, row_off, (size_t)(V + 1) * 8, hipMemcpyHostToDevice, e.stream));
  } else {
    std::vector<uint32_t> o32((size_t)V + 1);
    for (uint64_t r = 0; r <= V; r++) o32[r] = (uint32_t)row_off[r];
    GMP_HIP(hipMemcpyAsync(g.in.off.p, o32.data(), (size_t)(V + 1) * 4, hipMemcpyHostToDevice, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
  }
  g.h_in_off.assign(row_off, row_off + V + 1);
  if (semi_wanted(e, E, weights != nullptr)) {
    // the edges stay in pinned host memory; windows stream them (window.hip)
    GMP_HIP(hipStreamSynchronize(e.stream));
    g.in.valid = true;
    semi_setup(e, col_idx, weights);
    return;
  }
  g.in.idx.alloc(std::max<uint64_t>(E, 1) * 4);
  if (E) GMP_HIP(hipMemcpyAsync(g.in.idx.p, col_idx, E * 4, hipMemcpyHostToDevice, e.stream));
  if (weights) {
    g.in.w.alloc(std::max<uint64_t>(E, 1) * 4);
    if (E) GMP_HIP(hipMemcpyAsync(g.in.w.p, weights, E * 4, hipMemcpyHostToDevice, e.stream));
  }
  g.in.valid = true;
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
  GMP_HIP(rocprim::unique(nullptr, tb2, sorted, uniq, cnt.as<uint64_t>(), nkeys, rocprim::equal_to<uint64_t>(),
                          e.stream));
  tmp.alloc(tb2);
  GMP_HIP(rocprim::unique(tmp.p, tb2, sorted, uniq, cnt.as<uint64_t>(), nkeys, rocprim::equal_to<uint64_t>(),
                          e.stream));
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
  csr_from_sorted_keys(e, uniq, nuniq, p.scale, V, g.off64, g.in);
  if (p.weighted) {
    g.weighted = true;
    g.in.w.alloc(std::max<uint64_t>(nuniq, 1) * 4);
    if (nuniq) {
      rmat_weights_kernel<<<grid_for(nuniq), kBlock, 0, e.stream>>>(uniq, nuniq, p.scale, c.mask, p.seed + 2,
                                                                   g.in.w.as<uint32_t>());
      GMP_CHECK_LAUNCH();
    }
  }
  GMP_HIP(hipStreamSynchronize(e.stream));
  download_offsets(e, g.in, V, g.off64, g.h_in_off);
}

void graph_finish(Engine &e, uint32_t P, const uint32_t *bounds) {
  Graph &g = e.g;
  if (!g.semi) {  // semi-external graphs counted their out-degrees while setting up
    g.out_deg.alloc((size_t)std::max<uint32_t>(g.V, 1) * 4);
    GMP_HIP(hipMemsetAsync(g.out_deg.p, 0, (size_t)g.V * 4, e.stream));
  }
  if (g.E && !g.semi) {
    out_degree_kernel<<<grid_for(g.E), kBlock, 0, e.stream>>>(g.in.idx.as<uint32_t>(), g.E, g.out_deg.as<uint32_t>());
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
  if (g.out.valid) return;
  rebuild_csr(e, g.view(g.in), g.V, g.E, g.off64, g.weighted, nullptr, true, g.out);
}

void ensure_work(Engine &e) {
  Graph &g = e.g;
  if (g.work_ready) return;
  g.work_identity = !e.opt.reorder || g.semi != nullptr;
  g.wout_ready = false;
  g.packed.valid = false;
  g.strips.valid = false;
  g.hubs.valid = false;
  g.plan.built = false;
  if (g.work_identity) {
    g.work_ready = true;
    return;
  }
  const uint32_t V = g.V;
  // order: out-degree descending over the whole graph, ties by id (one interval: the work
  // layout's hottest sources then form one dense prefix for the source-strip gather)
  DevBuf keys, vals, kB, vB, dbounds, tmp;
  keys.alloc((size_t)V * 8);
  kB.alloc((size_t)V * 8);
  vals.alloc((size_t)V * 4);
  vB.alloc((size_t)V * 4);
  const uint32_t whole[2] = {0, V};
  dbounds.alloc(8);
  GMP_HIP(hipMemcpyAsync(dbounds.p, whole, 8, hipMemcpyHostToDevice, e.stream));
  order_keys_kernel<<<grid_for(V), kBlock, 0, e.stream>>>(g.out_deg.as<uint32_t>(), dbounds.as<uint32_t>(), 1u, V,
                                                         keys.as<uint64_t>(), vals.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  rocprim::double_buffer<uint64_t> dk(keys.as<uint64_t>(), kB.as<uint64_t>());
  rocprim::double_buffer<uint32_t> dv(vals.as<uint32_t>(), vB.as<uint32_t>());
  const uint32_t endbit = 32;
  size_t tb = 0;
  GMP_HIP(rocprim::radix_sort_pairs(nullptr, tb, dk, dv, (size_t)V, 0, endbit, e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::radix_sort_pairs(tmp.p, tb, dk, dv, (size_t)V, 0, endbit, e.stream));
  g.n2o.alloc((size_t)V * 4);
  g.o2n.alloc((size_t)V * 4);
  GMP_HIP(hipMemcpyAsync(g.n2o.p, dv.current(), (size_t)V * 4, hipMemcpyDeviceToDevice, e.stream));
  scatter_inverse_kernel<<<grid_for(V), kBlock, 0, e.stream>>>(g.n2o.as<uint32_t>(), V, g.o2n.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  g.wdeg.alloc((size_t)V * 4);
  gather_u32_kernel<<<grid_for(V), kBlock, 0, e.stream>>>(g.out_deg.as<uint32_t>(), g.n2o.as<uint32_t>(), V,
                                                         g.wdeg.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  keys.release();
  kB.release();
  vals.release();
  vB.release();
  rebuild_csr(e, g.view(g.in), V, g.E, g.off64, g.weighted, g.o2n.as<uint32_t>(), false, g.win);
  download_offsets(e, g.win, V, g.off64, g.h_win_off);
  g.plan.built = false;
  g.work_ready = true;
}

void ensure_work_out(Engine &e) {
  Graph &g = e.g;
  ensure_work(e);
  if (g.wout_ready) return;
  if (g.work_identity) {
    if (!(g.symmetric && !g.weighted)) graph_build_out(e);
  } else if (!(g.symmetric && !g.weighted)) {
    rebuild_csr(e, g.view(g.win), g.V, g.E, g.off64, g.weighted, nullptr, true, g.wout);
  }
  g.wout_ready = true;
}

void edge_row_ids(Engine &e, const CsrView &src, uint32_t V, uint64_t E, bool off64, DevBuf &rowid) {
  DevBuf tmp;
  edge_rows(e, src, V, E, off64, rowid, tmp);
  GMP_HIP(hipStreamSynchronize(e.stream));
}

void transpose_csr(Engine &e, const CsrView &src, uint32_t V, uint64_t E, bool off64, bool weighted, Csr &dst) {
  rebuild_csr(e, src, V, E, off64, weighted, nullptr, true, dst);
}

void count_degrees(Engine &e, const uint32_t *idx, uint64_t E, uint32_t V, uint32_t *deg) {
  GMP_HIP(hipMemsetAsync(deg, 0, (size_t)V * 4, e.stream));
  if (E) {
    out_degree_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(idx, E, deg);
    GMP_CHECK_LAUNCH();
  }
}

void map_to_original(Engine &e, const uint32_t *x_work, uint32_t *out) {
  Graph &g = e.g;
  if (g.work_identity) {
    GMP_HIP(hipMemcpyAsync(out, x_work, (size_t)g.V * 4, hipMemcpyDeviceToDevice, e.stream));
    return;
  }
  gather_u32_kernel<<<grid_for(g.V), kBlock, 0, e.stream>>>(x_work, g.o2n.as<uint32_t>(), g.V, out);
  GMP_CHECK_LAUNCH();
}

uint32_t work_id(Engine &e, uint32_t original) {
  Graph &g = e.g;
  if (g.work_identity) return original;
  uint32_t w = 0;
  GMP_HIP(hipMemcpy(&w, g.o2n.as<uint32_t>() + original, 4, hipMemcpyDeviceToHost));
  return w;
}

void graph_download(Engine &e, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights, uint32_t *out_degree) {
  Graph &g = e.g;
  if (row_off) std::memcpy(row_off, g.h_in_off.data(), (size_t)(g.V + 1) * 8);
  if (g.semi) {
    if (weights && !g.weighted) fail(GMP_EUNSUPPORTED, "graph has no weights");
    semi_download(e, col_idx, weights);
    if (out_degree && g.V) GMP_HIP(hipMemcpy(out_degree, g.out_deg.p, (size_t)g.V * 4, hipMemcpyDeviceToHost));
    return;
  }
  if (col_idx && g.E) GMP_HIP(hipMemcpy(col_idx, g.in.idx.p, g.E * 4, hipMemcpyDeviceToHost));
  if (weights) {
    if (!g.weighted) fail(GMP_EUNSUPPORTED, "graph has no weights");
    if (g.E) GMP_HIP(hipMemcpy(weights, g.in.w.p, g.E * 4, hipMemcpyDeviceToHost));
  }
  if (out_degree && g.V) GMP_HIP(hipMemcpy(out_degree, g.out_deg.p, (size_t)g.V * 4, hipMemcpyDeviceToHost));
}

}  // namespace gmp
