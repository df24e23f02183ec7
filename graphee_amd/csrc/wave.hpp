// Wavefront-level helpers for gfx950 (64 lanes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gmp {

constexpr int kWave = 64;

// Orders LDS traffic between the lanes of one wavefront (no workgroup barrier: the
// waves of a workgroup run independent tasks).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t lane_id() { return __lane_id(); }

// gfx950 deals the workgroups of a launch round-robin over its 8 XCDs, each with its own L2.
// xcd_block_id() renumbers them so that each XCD holds a contiguous range of ids (when the grid
// is a multiple of 8): tasks taken in id order then put neighbouring tasks -- which share the
// cache lines at their borders -- on one L2.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_block_id() {
  const uint32_t b = blockIdx.x, g = gridDim.x;
  return g % kXcds ? b : (b % kXcds) * (g / kXcds) + b / kXcds;
}

// Wave index inside the workgroup, made provably uniform so loads indexed by it are
// scalar loads.
__device__ inline uint32_t wave_id() {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Compensated (Kahan) accumulation: true sum ~= s - c.
__device__ inline void kahan_add(float &s, float &c, float x) {
  float y = x - c;
  float t = s + y;
  c = (t - s) - y;
  s = t;
}

// Error-free sum of two floats.
__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// (h1 + l1) + (h2 + l2) as a normalised hi/lo pair.
__device__ inline void dd_add(float &h1, float &l1, float h2, float l2) {
  float s, e;
  two_sum(h1, h2, s, e);
  e += l1 + l2;
  float hi = s + e;
  l1 = e - (hi - s);
  h1 = hi;
}

// Butterfly reduction of hi/lo pairs; lane 0's result has a fixed combination order.
__device__ inline void wave_reduce_dd(float &h, float &l) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    float oh = __shfl_xor(h, off, kWave);
    float ol = __shfl_xor(l, off, kWave);
    dd_add(h, l, oh, ol);
  }
}

__device__ inline double wave_reduce_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

__device__ inline unsigned long long wave_reduce_sum_u64(unsigned long long x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

__device__ inline uint32_t wave_reduce_min(uint32_t x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, off, kWave));
  return x;
}

// Inclusive prefix sum over the 64 lanes (uint32).
__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = (uint32_t)__shfl_up((int)x, off, kWave);
    if (lane >= (uint32_t)off) x += y;
  }
  return x;
}

// ---- DPP forms (GFX9 row shifts / row broadcasts / wave shift; no LDS traffic) ----
// dpp<CTRL, ROWS>(x): x of the lane the DPP control names; 0 where that lane is out of range
// or the lane's row is outside ROWS.
template <int CTRL, int ROWS = 0xF>
__device__ inline uint32_t dpp_u32(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xF, true);
}

constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

// Inclusive prefix sum over the 64 lanes: row-local shifts by 1, 2, 4, 8, then the last lane
// of row 0 (2) is added to row 1 (3), then lane 31 to rows 2 and 3.
__device__ inline uint32_t wave_scan_dpp(uint32_t x) {
  x += dpp_u32<kDppRowShr1>(x);
  x += dpp_u32<kDppRowShr2>(x);
  x += dpp_u32<kDppRowShr4>(x);
  x += dpp_u32<kDppRowShr8>(x);
  x += dpp_u32<kDppRowBcast15, 0xA>(x);
  x += dpp_u32<kDppRowBcast31, 0xC>(x);
  return x;
}

// Segment flags of the scans below (bit l: a segment starts inside lane l's window): after a
// row shift by S the flags move up by S inside each row of 16; after the row broadcast of lane
// 15 into rows 1 and 3 a row takes the flag of its source lane.
template <int S>
__device__ inline uint64_t seg_flags_shift(uint64_t f) {
  constexpr uint64_t kRow = 0x0001000100010001ull;  // lane 0 of every row
  return f | ((f << S) & ~(kRow * ((1ull << S) - 1)));
}
__device__ inline uint64_t seg_flags_bcast15(uint64_t f) {
  return f | ((f >> 15) & 1ull ? 0x00000000FFFF0000ull : 0ull) | ((f >> 47) & 1ull ? 0xFFFF000000000000ull : 0ull);
}

// One step of wave_segmented_scan_f32: lanes whose flag (bit l of f) is set keep x, the others
// add the value the DPP control moves in (one DPP add and one select on a scalar mask).
template <int CTRL, int ROWS = 0xF>
__device__ inline float seg_scan_step_f32(float x, uint64_t f) {
  const float y = __builtin_amdgcn_update_dpp(0.0f, x, CTRL, ROWS, 0xF, true);
  return __builtin_amdgcn_inverse_ballot_w64(f) ? x : x + y;
}

// Inclusive segmented FP32 sum over the 64 lanes with the step pattern of wave_scan_dpp; bit l
// of `heads` starts a segment at lane l.  The flags follow on the scalar unit.  The combination
// tree is fixed, so results are deterministic.
__device__ inline float wave_segmented_scan_f32(float x, uint64_t heads) {
  uint64_t f = heads;
  x = seg_scan_step_f32<kDppRowShr1>(x, f);
  f = seg_flags_shift<1>(f);
  x = seg_scan_step_f32<kDppRowShr2>(x, f);
  f = seg_flags_shift<2>(f);
  x = seg_scan_step_f32<kDppRowShr4>(x, f);
  f = seg_flags_shift<4>(f);
  x = seg_scan_step_f32<kDppRowShr8>(x, f);
  f = seg_flags_shift<8>(f);
  x = seg_scan_step_f32<kDppRowBcast15, 0xA>(x, f);  // lane 15 into row 1, lane 47 into row 3
  return seg_scan_step_f32<kDppRowBcast31, 0xC>(x, seg_flags_bcast15(f));  // lane 31 into rows 2, 3
}

template <int CTRL, int ROWS = 0xF>
__device__ inline uint32_t seg_max_step_u32(uint32_t x, uint64_t f) {
  const uint32_t y = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xF, true);
  return __builtin_amdgcn_inverse_ballot_w64(f) ? x : max(x, y);
}

// Inclusive segmented maximum over the 64 lanes with the steps and flags of
// wave_segmented_scan_f32; 0 (what the DPP moves in past a row's start) is neutral.
__device__ inline uint32_t wave_segmented_max_u32(uint32_t x, uint64_t heads) {
  uint64_t f = heads;
  x = seg_max_step_u32<kDppRowShr1>(x, f);
  f = seg_flags_shift<1>(f);
  x = seg_max_step_u32<kDppRowShr2>(x, f);
  f = seg_flags_shift<2>(f);
  x = seg_max_step_u32<kDppRowShr4>(x, f);
  f = seg_flags_shift<4>(f);
  x = seg_max_step_u32<kDppRowShr8>(x, f);
  f = seg_flags_shift<8>(f);
  x = seg_max_step_u32<kDppRowBcast15, 0xA>(x, f);
  return seg_max_step_u32<kDppRowBcast31, 0xC>(x, seg_flags_bcast15(f));
}

__device__ inline uint32_t read_lane_u32(uint32_t x, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}


// Wave-aggregated append: lanes with `pred` get consecutive slots of `list`, with one
// atomic per wave on `count` (ballot + popcount prefix).
__device__ inline uint32_t wave_append(bool pred, uint32_t *count) {
  const uint64_t mask = __ballot(pred);
  const uint32_t lane = lane_id();
  const uint32_t total = (uint32_t)__popcll(mask);
  uint32_t base = 0;
  if (total) {
    const int leader = __ffsll((long long)mask) - 1;
    if ((int)lane == leader) base = atomicAdd(count, total);
    base = (uint32_t)__shfl((int)base, leader, kWave);
  }
  const uint64_t below = lane ? (mask & ((1ull << lane) - 1)) : 0ull;
  return base + (uint32_t)__popcll(below);
}

// Appends staged in a per-wave LDS buffer and written out with one atomic per flush (when the
// buffer could overflow on the next batch, and at the end of the kernel), coalesced, instead of
// one atomic per wave and batch: a single append counter serialises its atomics.
struct WaveAppend {
  uint32_t *buf;  // this wave's LDS slots
  uint32_t n;     // entries held (the same in every lane)
};

__device__ inline void append_flush(WaveAppend &w, uint32_t *count, uint32_t *out) {
  if (!w.n) return;
  wave_sync();
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(count, w.n);
  base = read_lane_u32(base, 0);
  for (uint32_t i = lane_id(); i < w.n; i += kWave) out[base + i] = w.buf[i];
  wave_sync();
  w.n = 0;
}

// Block-aggregated append: every thread of the block calls it (block-uniform control flow) with
// its number of entries; returns the first of its consecutive slots, with one atomic on `count`
// per block.  `sh` is WAVES + 1 words of LDS.
template <int WAVES>
__device__ inline uint32_t block_append(uint32_t mine, uint32_t *count, uint32_t *sh) {
  const uint32_t lane = lane_id(), wv = wave_id();
  const uint32_t incl = wave_scan_dpp(mine);
  if (lane == kWave - 1) sh[wv] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < WAVES; w++) {
      const uint32_t x = sh[w];
      sh[w] = t;
      t += x;
    }
    sh[WAVES] = t ? atomicAdd(count, t) : 0u;
  }
  __syncthreads();
  const uint32_t base = sh[WAVES] + sh[wv] + incl - mine;
  __syncthreads();  // sh may be reused by the next call
  return base;
}

}  // namespace gmp
