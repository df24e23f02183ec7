// Delta + varint neighbour lists, decoded by a whole wavefront.
//
// GraphMP keeps shards in a snappy/zlib compressed cache and inflates a shard before every
// use (GraphMP/system/Global.h:129-255).  Here the in-CSR's column ids are kept compressed in
// HBM and decoded inside the gather, 64 edges (one per lane) at a time:
//   * edges are cut into global groups of 64 (edge e is lane e % 64 of group e / 64);
//   * inside a row, an edge stores delta = src - prev_src - 1 (rows are ascending and
//     deduplicated); the first edge of a group, the first edge of a row and any edge that
//     does not ascend store the absolute id instead ("restart");
//   * every value is a LEB128 varint (7 bits per byte, high bit = more bytes follow);
//   * per group: the byte offset of its first varint (gbyte, G+1 entries) and the 64-bit mask
//     of restart lanes (gmask).
// Decoding a group: the wave stages its <= 320 bytes in LDS, finds each varint's terminating
// byte with a terminator count + wave prefix sum, assembles the values, and turns deltas back
// into ids with a segmented wave prefix sum that restarts at the mask's set bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gmp_internal.hpp"
#include "wave.hpp"

namespace gmp {

constexpr uint32_t kVarintGroup = 64;
constexpr uint32_t kVarintMaxBytes = 5 * kVarintGroup;     // 320
constexpr uint32_t kVarintStageWords = 96;                  // >= (320 + 6) / 4 + 1
constexpr uint32_t kVarintScratchWords = kVarintStageWords + kVarintGroup;  // staging + varint ends

__host__ __device__ inline uint32_t varint_len(uint32_t v) {
  return v < (1u << 7) ? 1u : v < (1u << 14) ? 2u : v < (1u << 21) ? 3u : v < (1u << 28) ? 4u : 5u;
}

// Returns the source id of edge 64*g + lane (meaningless for lanes beyond the last edge).
// `scratch` is this wave's kVarintScratchWords words of LDS.  All 64 lanes must call it.
__device__ inline uint32_t varint_decode_group(const Packed &pk, uint64_t g, uint64_t E, uint32_t *scratch) {
  const uint32_t lane = lane_id();
  uint32_t *stage = scratch;
  uint32_t *ends = scratch + kVarintStageWords;
  const uint64_t b0 = pk.gbyte[g], b1 = pk.gbyte[g + 1];
  const uint64_t a0 = b0 & ~3ull;
  const uint32_t nw = (uint32_t)((b1 + 3 - a0) >> 2);
  const uint32_t *wsrc = reinterpret_cast<const uint32_t *>(pk.bytes + a0);
  for (uint32_t w = lane; w < nw; w += kWave) stage[w] = __builtin_nontemporal_load(&wsrc[w]);
  wave_sync();
  const uint8_t *by = reinterpret_cast<const uint8_t *>(stage) + (b0 - a0);
  const uint32_t nb = (uint32_t)(b1 - b0);
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t i = 0; i < 5; i++) {
    const uint32_t p = 5 * lane + i;
    cnt += (p < nb && !(by[p] & 0x80)) ? 1u : 0u;
  }
  uint32_t k = wave_inclusive_scan(cnt) - cnt;
#pragma unroll
  for (uint32_t i = 0; i < 5; i++) {
    const uint32_t p = 5 * lane + i;
    if (p < nb && !(by[p] & 0x80)) ends[k++] = p;
  }
  wave_sync();
  const uint64_t first = g * kVarintGroup;
  const uint32_t ne = (uint32_t)min<uint64_t>(kVarintGroup, E - first);
  uint32_t v = 0;
  if (lane < ne) {
    const uint32_t start = lane ? ends[lane - 1] + 1 : 0, end = ends[lane];
    uint32_t sh = 0;
    for (uint32_t p = start; p <= end; p++, sh += 7) v |= (uint32_t)(by[p] & 0x7f) << sh;
  }
  const uint64_t mask = pk.gmask[g];
  const bool restart = (mask >> lane) & 1ull;
  const uint32_t term = restart ? v : v + 1u;
  const uint32_t P = wave_inclusive_scan(term);
  const uint64_t upto = mask & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));  // bit 0 always set
  const uint32_t s = 63u - (uint32_t)__clzll((long long)upto);
  const uint32_t before = (uint32_t)__shfl((int)P, (int)(s ? s - 1 : 0), kWave);
  wave_sync();
  return P - (s ? before : 0u);
}

}  // namespace gmp
