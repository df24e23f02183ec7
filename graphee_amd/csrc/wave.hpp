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

}  // namespace gmp
