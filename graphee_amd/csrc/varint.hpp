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

// A group's bytes in registers: the words from the one holding its first byte, two per lane
// (a group spans at most 81 words), loaded ahead of its decode.
struct VarintGroupWords {
  uint64_t b0, b1;  // byte range of the group
  uint32_t w0, w1;  // words lane and lane + 64 of the range
};

// Issues the loads of group g's bytes; every lane loads an in-range word (indices are clamped
// to the group's last word), so the loads never wait on a branch.
__device__ inline VarintGroupWords varint_fetch(const Packed &pk, uint64_t g) {
  const uint32_t lane = lane_id();
  VarintGroupWords r;
  r.b0 = pk.gbyte[g];
  r.b1 = pk.gbyte[g + 1];
  const uint64_t a0 = r.b0 & ~3ull;
  const uint32_t nw = (uint32_t)((r.b1 + 3 - a0) >> 2);
  const uint32_t *wsrc = reinterpret_cast<const uint32_t *>(pk.bytes + a0);
  r.w0 = __builtin_nontemporal_load(&wsrc[min(lane, nw - 1)]);
  r.w1 = __builtin_nontemporal_load(&wsrc[min(lane + kWave, nw - 1)]);
  return r;
}

// Returns the source id of edge 64*g + lane (meaningless for lanes beyond the last edge) from
// the group's fetched words.  `scratch` is this wave's kVarintScratchWords words of LDS.  All 64
// lanes must call it.
__device__ inline uint32_t varint_decode_words(const VarintGroupWords &gw, const Packed &pk, uint64_t g, uint64_t E,
                                               uint32_t *scratch) {
  const uint32_t lane = lane_id();
  uint32_t *stage = scratch;
  uint32_t *ends = scratch + kVarintStageWords;
  const uint64_t b0 = gw.b0, b1 = gw.b1;
  const uint64_t a0 = b0 & ~3ull;
  const uint32_t nw = (uint32_t)((b1 + 3 - a0) >> 2);
  stage[lane] = gw.w0;
  if (lane + kWave < nw) stage[lane + kWave] = gw.w1;
  wave_sync();
  // bytes are read as aligned word pairs and shifted (stage words past the group are stale and
  // masked off; indices stay below kVarintStageWords)
  const uint32_t off0 = (uint32_t)(b0 - a0), nb = (uint32_t)(b1 - b0);
  auto bytes5 = [&](uint32_t p) -> uint64_t {  // stage bytes p, p+1, ... of the group, low 40 bits
    const uint32_t q = off0 + p;
    const uint64_t two = ((uint64_t)stage[(q >> 2) + 1] << 32) | stage[q >> 2];
    return two >> ((q & 3u) * 8u);
  };
  // this lane's 5 bytes: which of them end a varint
  const uint64_t mine = bytes5(5 * lane);
  uint32_t ends_here = 0;
#pragma unroll
  for (uint32_t i = 0; i < 5; i++)
    if (5 * lane + i < nb && !((mine >> (8 * i + 7)) & 1ull)) ends_here |= 1u << i;
  const uint32_t cnt = (uint32_t)__popc(ends_here);
  uint32_t k = wave_scan_dpp(cnt) - cnt;
#pragma unroll
  for (uint32_t i = 0; i < 5; i++)
    if ((ends_here >> i) & 1u) ends[k++] = 5 * lane + i;
  wave_sync();
  const uint64_t first = g * kVarintGroup;
  const uint32_t ne = (uint32_t)min<uint64_t>(kVarintGroup, E - first);
  uint32_t v = 0;
  if (lane < ne) {
    const uint32_t start = lane ? ends[lane - 1] + 1 : 0, end = ends[lane];
    uint64_t t = bytes5(start);
    const uint32_t len = end - start + 1;  // 1..5
    if (len < 5) t &= (1ull << (8 * len)) - 1ull;
    // 7 payload bits of byte i land at bit 7i
    v = (uint32_t)((t & 0x7full) | ((t >> 1) & 0x3f80ull) | ((t >> 2) & 0x1fc000ull) | ((t >> 3) & 0xfe00000ull) |
                   ((t >> 4) & 0xf0000000ull));
  }
  const uint64_t mask = pk.gmask[g];
  const bool restart = (mask >> lane) & 1ull;
  const uint32_t term = restart ? v : v + 1u;
  const uint32_t P = wave_scan_dpp(term);
  const uint64_t upto = mask & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));  // bit 0 always set
  const uint32_t s = 63u - (uint32_t)__clzll((long long)upto);
  const uint32_t before = (uint32_t)__shfl((int)P, (int)(s ? s - 1 : 0), kWave);
  wave_sync();
  return P - (s ? before : 0u);
}

// Returns the source id of edge 64*g + lane (meaningless for lanes beyond the last edge).
// `scratch` is this wave's kVarintScratchWords words of LDS.  All 64 lanes must call it.
__device__ inline uint32_t varint_decode_group(const Packed &pk, uint64_t g, uint64_t E, uint32_t *scratch) {
  return varint_decode_words(varint_fetch(pk, g), pk, g, E, scratch);
}

}  // namespace gmp
