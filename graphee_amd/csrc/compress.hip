// Builds the delta + varint neighbour lists of the work in-CSR (format in varint.hpp).
// One wave per 64-edge group: lane k takes edge 64g + k, decides restart or delta, and the
// group's varint lengths are summed (pass 1), scanned over groups into byte offsets, then
// written at wave-prefix-sum positions (pass 2).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "gmp_internal.hpp"
#include "varint.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;

template <class OffT>
__global__ void mark_row_starts_kernel(const OffT *off, uint32_t V, uint8_t *rs) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < V; r += (uint64_t)gridDim.x * blockDim.x)
    if (off[r + 1] > off[r]) rs[off[r]] = 1;
}

__device__ inline void edge_value(const uint32_t *idx, const uint8_t *rs, uint64_t E, uint64_t g, uint32_t lane,
                                  bool &valid, bool &restart, uint32_t &v) {
  const uint64_t e = g * kVarintGroup + lane;
  valid = e < E;
  restart = true;
  v = 0;
  if (!valid) return;
  const uint32_t src = idx[e];
  const uint32_t prev = lane ? idx[e - 1] : 0u;
  restart = lane == 0 || rs[e] || src <= prev;
  v = restart ? src : src - prev - 1u;
}

__global__ __launch_bounds__(kBlock) void varint_measure_kernel(const uint32_t *idx, const uint8_t *rs, uint64_t E,
                                                                uint64_t G, uint64_t *glen, uint64_t *gmask) {
  const uint32_t lane = lane_id();
  for (uint64_t g = blockIdx.x * (uint64_t)kWavesPerBlock + wave_id(); g < G; g += (uint64_t)gridDim.x * kWavesPerBlock) {
    bool valid, restart;
    uint32_t v;
    edge_value(idx, rs, E, g, lane, valid, restart, v);
    const uint32_t len = valid ? varint_len(v) : 0u;
    uint32_t tot = len;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) tot += (uint32_t)__shfl_xor((int)tot, o, kWave);
    const uint64_t m = __ballot(valid && restart);
    if (lane == 0) {
      glen[g] = tot;
      gmask[g] = m;
    }
  }
}

__global__ __launch_bounds__(kBlock) void varint_write_kernel(const uint32_t *idx, const uint8_t *rs, uint64_t E,
                                                              uint64_t G, const uint64_t *gbyte, uint8_t *bytes) {
  const uint32_t lane = lane_id();
  for (uint64_t g = blockIdx.x * (uint64_t)kWavesPerBlock + wave_id(); g < G; g += (uint64_t)gridDim.x * kWavesPerBlock) {
    bool valid, restart;
    uint32_t v;
    edge_value(idx, rs, E, g, lane, valid, restart, v);
    const uint32_t len = valid ? varint_len(v) : 0u;
    const uint32_t local = wave_inclusive_scan(len) - len;
    uint8_t *p = bytes + gbyte[g] + local;
    for (uint32_t i = 0; i < len; i++) p[i] = (uint8_t)(((v >> (7 * i)) & 0x7fu) | (i + 1 < len ? 0x80u : 0u));
  }
}

inline uint32_t grid_for(uint64_t n) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 1u << 20));
}

}  // namespace

void ensure_packed(Engine &e) {
  ensure_work(e);
  Graph &g = e.g;
  if (g.packed.valid) return;
  const CsrView w = g.work_in();
  const uint64_t E = g.E, G = (E + kVarintGroup - 1) / kVarintGroup;
  DevBuf rs, glen, tmp;
  rs.alloc(std::max<uint64_t>(E, 1));
  glen.alloc((G + 1) * 8);
  g.packed.gbyte.alloc((G + 1) * 8);
  g.packed.gmask.alloc(std::max<uint64_t>(G, 1) * 8);
  GMP_HIP(hipMemsetAsync(rs.p, 0, std::max<uint64_t>(E, 1), e.stream));
  GMP_HIP(hipMemsetAsync(glen.p, 0, (G + 1) * 8, e.stream));
  if (g.off64) mark_row_starts_kernel<uint64_t><<<grid_for(g.V), kBlock, 0, e.stream>>>(w.offs<uint64_t>(), g.V, rs.as<uint8_t>());
  else mark_row_starts_kernel<uint32_t><<<grid_for(g.V), kBlock, 0, e.stream>>>(w.offs<uint32_t>(), g.V, rs.as<uint8_t>());
  GMP_CHECK_LAUNCH();
  const uint32_t wgrid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((G + kWavesPerBlock - 1) / kWavesPerBlock, 1u << 20));
  if (G) {
    varint_measure_kernel<<<wgrid, kBlock, 0, e.stream>>>(w.idx, rs.as<uint8_t>(), E, G, glen.as<uint64_t>(),
                                                          g.packed.gmask.as<uint64_t>());
    GMP_CHECK_LAUNCH();
  }
  size_t tb = 0;
  GMP_HIP(rocprim::exclusive_scan(nullptr, tb, glen.as<uint64_t>(), g.packed.gbyte.as<uint64_t>(), (uint64_t)0, G + 1,
                                  rocprim::plus<uint64_t>(), e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::exclusive_scan(tmp.p, tb, glen.as<uint64_t>(), g.packed.gbyte.as<uint64_t>(), (uint64_t)0, G + 1,
                                  rocprim::plus<uint64_t>(), e.stream));
  uint64_t total = 0;
  GMP_HIP(hipMemcpyAsync(&total, g.packed.gbyte.as<uint64_t>() + G, 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  g.packed.bytes.alloc(total + 64);  // word-granular staging may read a few bytes past the end
  GMP_HIP(hipMemsetAsync(g.packed.bytes.p, 0, total + 64, e.stream));
  if (G) {
    varint_write_kernel<<<wgrid, kBlock, 0, e.stream>>>(w.idx, rs.as<uint8_t>(), E, G, g.packed.gbyte.as<uint64_t>(),
                                                        g.packed.bytes.as<uint8_t>());
    GMP_CHECK_LAUNCH();
  }
  GMP_HIP(hipStreamSynchronize(e.stream));
  g.packed.nbytes = total;
  g.packed.ngroups = G;
  g.packed.valid = true;
}

}  // namespace gmp
