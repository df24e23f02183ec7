// Source-strip layout of the PageRank in-CSR (see StripCsr).
//
// Why: with the whole graph resident, a PageRank iteration is bound by its random gather of
// contributions.  Gathering row by row touches every source at once, so every cold source
// costs an L2 miss and a full cache line.  Sorting the edges by (source strip, destination,
// source) lets the gather walk one strip of 2^wlog consecutive (degree-ordered) sources at a
// time -- a strip's contributions fit one XCD's L2 -- while every (destination, strip) pair
// leaves one FP32 hi/lo partial that a row-block combine adds in strip order.
//
// Build: one 64-bit key per edge, strip << 2b | dst << b | src (b = id bits), an LSD radix
// sort, head flags where (strip, dst) changes, an exclusive scan into pair ids, and binary
// searches for the first pair of every (row block, strip).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "gmp_internal.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;

#define GS64(i, n) \
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * blockDim.x)

inline uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 1u << 20));
}

inline uint32_t bits_for(uint64_t n) {
  uint32_t b = 0;
  while (b < 63 && (1ull << b) < n) b++;
  return std::max<uint32_t>(b, 1);
}

__global__ void strip_keys_kernel(const uint32_t *idx, const uint32_t *rowid, uint64_t E, uint32_t b, uint32_t wlog,
                                  uint64_t *keys) {
  GS64(j, E) {
    const uint64_t src = idx[j], dst = rowid[j];
    keys[j] = ((src >> wlog) << (2 * b)) | (dst << b) | src;
  }
}

__global__ void head_flags_kernel(const uint64_t *keys, uint64_t E, uint32_t b, uint32_t *flag) {
  GS64(j, E) flag[j] = (j == 0 || (keys[j] >> b) != (keys[j - 1] >> b)) ? 1u : 0u;
}

__global__ void pairs_from_keys_kernel(const uint64_t *keys, const uint32_t *pid, uint64_t E, uint32_t b,
                                       uint32_t *pair_off, uint32_t *pair_row, uint32_t *pair_strip, uint32_t *idx) {
  const uint64_t mask = (1ull << b) - 1;
  GS64(j, E) {
    const uint64_t k = keys[j];
    idx[j] = (uint32_t)(k & mask);
    if (j == 0 || (k >> b) != (keys[j - 1] >> b)) {
      const uint32_t p = pid[j];
      pair_off[p] = (uint32_t)j;
      pair_row[p] = (uint32_t)((k >> b) & mask);
      pair_strip[p] = (uint32_t)(k >> (2 * b));
    }
  }
}

// strip_base[s] = first pair whose strip >= s, s in [0, nstrips]
__global__ void strip_base_kernel(const uint32_t *pair_strip, uint64_t P, uint32_t nstrips, uint32_t *base) {
  GS64(i, P + 1) {
    const int64_t s = i < P ? (int64_t)pair_strip[i] : (int64_t)nstrips;
    const int64_t sp = i > 0 ? (int64_t)pair_strip[i - 1] : -1;
    for (int64_t q = sp + 1; q <= s; q++) base[q] = (uint32_t)i;
  }
}

// blk[k * nstrips + s] = first pair of strip s whose row >= k * rows_per_block, k in [0, nblk]
__global__ void block_table_kernel(const uint32_t *pair_row, const uint32_t *strip_base, uint32_t nstrips,
                                   uint32_t nblk, uint32_t rows_per_block, uint32_t *blk) {
  GS64(t, (uint64_t)(nblk + 1) * nstrips) {
    const uint32_t s = (uint32_t)(t % nstrips), k = (uint32_t)(t / nstrips);
    const uint64_t target = (uint64_t)k * rows_per_block;
    uint32_t lo = strip_base[s], hi = strip_base[s + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)pair_row[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    blk[t] = lo;
  }
}

}  // namespace

bool ensure_strips(Engine &e) {
  ensure_work(e);
  Graph &g = e.g;
  StripCsr &S = g.strips;
  const uint32_t wlog = e.opt.strip_log;
  if (wlog == 0 || g.off64 || g.E == 0) return false;
  const uint32_t V = g.V;
  const uint64_t E = g.E;
  const uint32_t b = bits_for(V);
  const uint32_t nstrips = (uint32_t)(((uint64_t)V + (1ull << std::min<uint32_t>(wlog, 40)) - 1) >> std::min<uint32_t>(wlog, 40));
  const uint32_t sbits = bits_for((uint64_t)nstrips + 1);
  if (2 * b + sbits > 64) return false;
  if (!(S.valid && S.wlog == wlog)) {
    S = StripCsr();
    const CsrView w = g.work_in();
    DevBuf rowid, kA, kB, tmp, flag;
    edge_row_ids(e, w, V, E, g.off64, rowid);
    kA.alloc(E * 8);
    kB.alloc(E * 8);
    strip_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(w.idx, rowid.as<uint32_t>(), E, b, wlog, kA.as<uint64_t>());
    GMP_CHECK_LAUNCH();
    rowid.release();
    rocprim::double_buffer<uint64_t> dk(kA.as<uint64_t>(), kB.as<uint64_t>());
    size_t tb = 0;
    GMP_HIP(rocprim::radix_sort_keys(nullptr, tb, dk, E, 0, 2 * b + sbits, e.stream));
    tmp.alloc(tb);
    GMP_HIP(rocprim::radix_sort_keys(tmp.p, tb, dk, E, 0, 2 * b + sbits, e.stream));
    const uint64_t *keys = dk.current();
    flag.alloc(E * 4);
    head_flags_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(keys, E, b, flag.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    // exclusive scan of the head flags into pair ids, in place
    DevBuf pid;
    pid.alloc(E * 4);
    size_t tb2 = 0;
    GMP_HIP(rocprim::exclusive_scan(nullptr, tb2, flag.as<uint32_t>(), pid.as<uint32_t>(), 0u, E,
                                    rocprim::plus<uint32_t>(), e.stream));
    tmp.alloc(tb2);
    GMP_HIP(rocprim::exclusive_scan(tmp.p, tb2, flag.as<uint32_t>(), pid.as<uint32_t>(), 0u, E,
                                    rocprim::plus<uint32_t>(), e.stream));
    uint32_t last_pid = 0, last_flag = 0;
    GMP_HIP(hipMemcpyAsync(&last_pid, pid.as<uint32_t>() + E - 1, 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipMemcpyAsync(&last_flag, flag.as<uint32_t>() + E - 1, 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    flag.release();
    const uint64_t P = (uint64_t)last_pid + last_flag;
    S.npairs = P;
    S.pair_off.alloc((P + 1) * 4);
    S.pair_row.alloc(std::max<uint64_t>(P, 1) * 4);
    S.idx.alloc(E * 4);
    DevBuf pstrip, sbase;
    pstrip.alloc(std::max<uint64_t>(P, 1) * 4);
    pairs_from_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(keys, pid.as<uint32_t>(), E, b,
                                                                S.pair_off.as<uint32_t>(), S.pair_row.as<uint32_t>(),
                                                                pstrip.as<uint32_t>(), S.idx.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    const uint32_t e32 = (uint32_t)E;
    GMP_HIP(hipMemcpyAsync(S.pair_off.as<uint32_t>() + P, &e32, 4, hipMemcpyHostToDevice, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    kA.release();
    kB.release();
    pid.release();
    sbase.alloc(((size_t)nstrips + 1) * 4);
    strip_base_kernel<<<grid_for(P + 1), kBlock, 0, e.stream>>>(pstrip.as<uint32_t>(), P, nstrips, sbase.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    S.nstrips = nstrips;
    S.nblk = (V + kCombineRows - 1) / kCombineRows;
    S.blk.alloc(((size_t)S.nblk + 1) * nstrips * 4);
    block_table_kernel<<<grid_for((uint64_t)(S.nblk + 1) * nstrips), kBlock, 0, e.stream>>>(
        S.pair_row.as<uint32_t>(), sbase.as<uint32_t>(), nstrips, S.nblk, kCombineRows, S.blk.as<uint32_t>());
    GMP_CHECK_LAUNCH();
    std::vector<uint32_t> h32(P + 1);
    GMP_HIP(hipMemcpyAsync(h32.data(), S.pair_off.p, (P + 1) * 4, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    S.h_pair_off.assign(h32.begin(), h32.end());
    S.wlog = wlog;
    S.valid = true;
  }
  plan_for(e, S.plan, S.h_pair_off.data(), (uint32_t)S.npairs);
  return true;
}

}  // namespace gmp
