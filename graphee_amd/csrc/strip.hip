// Source-strip layout of the PageRank in-CSR and the iteration that runs over it.
//
// With the whole graph resident, a row-by-row PageRank gather is bound by its random reads of
// source contributions: every cold source costs a cache-line miss, and even L2 hits cost one
// L1 fill per edge.  Here the edges are sorted by (source strip, destination, source), strips
// being 2^wlog consecutive work ids (the degree order puts the hubs in the first strips):
//
//   pr_strip_gather_kernel  stages one strip's contributions in LDS and streams the strip's
//                           16-bit source offsets (8 per lane per round).  Every run of edges
//                           with one destination -- a segment -- leaves one FP32 message, its
//                           sum taken in FP64 (lane-sequential, then a segmented wave scan).
//                           Messages go through L2 with plain stores, which merges the lanes'
//                           scattered writes into whole lines.
//   pr_bin_kernel           owns a bin of 2^rlog destinations: its messages form one run per
//                           strip; they are added into LDS accumulators in an exact 61-bit
//                           fixed point with integer atomics -- exact sums do not depend on
//                           the order of the adds, so the result is deterministic -- and the
//                           rank update is applied.
//   pr_split_apply_kernel   finishes bins whose messages were split over several workgroups
//                           (their slices leave fixed-point partials).
//
// Segments end where the strip or destination changes and at every wave-task boundary, so
// each wave task starts with a fresh segment.  While the layout is built, bit 15 of an edge's
// code marks a segment's last edge (hence strips of at most 2^15 sources); the finished layout
// keeps those bits in a byte per 8 edges and the codes as bare offsets.
//
// Build (once per graph and layout options): a 64-bit key per edge, strip << (b + wlog) |
// dst << wlog | source offset (b = id bits), an LSD radix sort, segment ends, an exclusive
// scan into segment ids, and binary searches for the first segment of every (bin, strip).
#include <algorithm>
#include <cstring>
#include <type_traits>

#include <rocprim/rocprim.hpp>

#include "pr_common.hpp"

namespace gmp {

namespace {

constexpr int kBlock = 256;                    // build kernels, split apply
constexpr int kIterBlock = 1024;               // gather and bin kernels
constexpr int kIterWaves = kIterBlock / kWave;
constexpr uint32_t kRound = 8 * kWave;         // edges per wave per round (8 per lane)
constexpr uint32_t kWaveTaskEdges = 16 * kRound;  // edges per wave task
constexpr int kAhead = 4;                      // rounds of codes in flight per wave
constexpr int kUnroll = 16;                    // messages in flight per thread (bin kernel)
constexpr uint32_t kChunk = 512;               // bin kernel: messages per chunk-index entry
constexpr uint32_t kSplitRows = 1024;          // split apply: rows per block task
constexpr uint32_t kCodePad = kRound;          // codes read past the last edge stay in bounds
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kNever = 0xFFFFFFFFu;       // strip flag: no source changed yet
constexpr int kPullUnroll = 4;                 // pull rounds: rows per thread in flight (bin apply)

// What a sweep over the strips computes.  PageRank sums contributions; the others keep
// complemented values (c = ~x), so the least label or distance is the greatest c and 0 -- what
// DPP moves in past a row's start, and the value of an unreached vertex -- is neutral.
enum SweepMode {
  kSweepPR = 0,
  kSweepCC = 1,
  kSweepUnit = 2,
  kSweepWeighted = 3,
  kSweepWeighted8 = 4,
  kSweepPRPacked = 5,  // PageRank over the packed codes
  kSweepPRIds = 6,     // PageRank over 4-byte source ids
};
constexpr bool is_pr(int mode) { return mode == kSweepPR || mode == kSweepPRPacked || mode == kSweepPRIds; }

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

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

struct StripArgs {
  // vertex update (pr_apply)
  const uint32_t *outdeg;
  const float *contrib_in;
  const float *rank_in;
  float *rank_out;
  float *contrib_out;
  const PrScal *sc;
  float d;
  uint32_t seed;
  // layout
  const uint16_t *code;   // source offsets in the strip
  const uint64_t *sbase;  // per strip: the edge its wave tasks' edge offsets count from
  const uint8_t *endb;    // bit k of byte i: edge 8 i + k ends its segment
  const uint32_t *pk;     // packed codes (kSweepPRPacked): the stream,
  const uint32_t *pkhdr;  // each block's 8 width words,
  const uint32_t *pkoff;  // each block's first word,
  const uint32_t *pkblk;  // each strip's first block
  const uint32_t *ids;    // source work ids in layout order (kSweepPRIds)
  const uint4 *wtask;
  const uint4 *gtask;
  uint32_t ngt;
  uint32_t *gq;           // workgroup-task queue: [0] next task, [1] ticket, [2 + b] workgroup b's task
  float *msg;
  const uint16_t *seg_row;
  const uint32_t *runs;
  uint32_t ns;
  const uint4 *btask;
  uint32_t nbt;
  unsigned long long *bpart;
  const uint4 *split;
  uint32_t nsplit;
  uint32_t V, wlog, rlog;
  double *pr1, *pd1, *pr2, *pd2;
  // CC and SSSP sweeps
  uint32_t *nlab;          // complemented labels or distances, updated in place
  uint32_t *changed;       // vertices whose value dropped in the sweep
  const uint32_t *wt;      // edge weights in layout order (weighted SSSP), or
  const uint8_t *wt8;      // the same as bytes when every weight fits one
  uint32_t *sflag;         // per strip: the last sweep that changed one of its vertices
  uint32_t *vstamp;        // per vertex: the last sweep that changed it (null: not kept)
  uint32_t *gflag;         // per strip: the last sweep that gathered it
  uint32_t sweep;          // this sweep's number, from 1
  unsigned long long *edges;  // edges gathered (tasks not skipped)
  // delta-stepping pull rounds (fmask set): the values are plain distances in `dist`; a source
  // offers only when fmask[u] == fstamp (it is in the round's near frontier), so the round
  // relaxes exactly the frontier's out-edges; a destination whose distance drops below thr joins
  // the next near frontier (near stamp next_stamp), one that stays at or above it the far pile
  const uint32_t *fmask;
  uint32_t fstamp, next_stamp, thr;
  uint32_t *dist, *near_stamp, *near_out, *near_cnt, *far_flag, *far_out, *far_cnt;
  const uint32_t *deg;          // out-degree by work id: the next frontier's out-edges
  unsigned long long *near_deg; // are summed here
};

// ---------------------------------------------------------------- build kernels

__global__ void strip_keys_kernel(const uint32_t *idx, const uint32_t *rowid, uint64_t E, uint32_t b, uint32_t wlog,
                                  uint64_t *keys) {
  const uint64_t wmask = (1ull << wlog) - 1;
  GS64(j, E) {
    const uint64_t src = idx[j], dst = rowid[j];
    keys[j] = ((src >> wlog) << (b + wlog)) | (dst << wlog) | (src & wmask);
  }
}

// off[s] = first edge whose strip >= s, s in [0, ns]
__global__ void strip_offsets_kernel(const uint64_t *keys, uint64_t E, uint32_t sh, uint32_t ns, uint64_t *off) {
  GS64(i, E + 1) {
    const int64_t s = i < E ? (int64_t)(keys[i] >> sh) : (int64_t)ns;
    const int64_t sp = i > 0 ? (int64_t)(keys[i - 1] >> sh) : -1;
    for (int64_t q = sp + 1; q <= s; q++) off[q] = i;
  }
}

// code = source offset in the strip | 0x8000 when the next edge has another (strip, dst)
__global__ void codes_kernel(const uint64_t *keys, uint64_t E, uint32_t wlog, uint16_t *code) {
  const uint64_t wmask = (1ull << wlog) - 1;
  GS64(j, E) {
    const uint64_t k = keys[j];
    const bool end = j + 1 == E || (keys[j + 1] >> wlog) != (k >> wlog);
    code[j] = (uint16_t)((k & wmask) | (end ? 0x8000u : 0u));
  }
}

// every wave task starts a segment: end one at the edge before it
__global__ void task_ends_kernel(const uint4 *wt, const uint64_t *sbase, uint32_t n, uint16_t *code) {
  GS64(t, n) {
    const uint64_t e0 = sbase[wt[t].w] + wt[t].x;
    if (e0 > 0) code[e0 - 1] |= (uint16_t)0x8000u;
  }
}

// ids[i]: the source work id of edge i in layout order, the strip's first id plus the edge's
// code; a block per wave task
__global__ void strip_ids_kernel(const uint4 *wt, const uint64_t *sbase, uint32_t n, uint32_t wlog,
                                 const uint16_t *code, uint32_t *ids) {
  for (uint32_t t = blockIdx.x; t < n; t += gridDim.x) {
    const uint4 w = wt[t];
    const uint64_t b = sbase[w.w];
    const uint32_t s0 = w.w << wlog, m = (1u << wlog) - 1u;
    for (uint32_t i = w.x + threadIdx.x; i < w.y; i += blockDim.x) ids[b + i] = s0 | (code[b + i] & m);
  }
}

__global__ void end_flags_kernel(const uint16_t *code, uint64_t E, uint32_t *flag) {
  GS64(j, E) flag[j] = code[j] >> 15;
}

__global__ void task_seg_kernel(uint4 *wt, const uint64_t *sbase, uint32_t n, const uint32_t *segid) {
  GS64(t, n) wt[t].z = segid[sbase[wt[t].w] + wt[t].x];
}

// Moves the end bits out of the codes: endb[i] bit k = bit 15 of code[8 i + k], which is then
// cleared, so a code is the bare source offset (n = groups of 8 codes, padding included).
__global__ void split_ends_kernel(uint16_t *code, uint64_t n, uint8_t *endb) {
  GS64(i, n) {
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint16_t c = code[8 * i + k];
      m |= (uint32_t)(c >> 15) << k;
      code[8 * i + k] = (uint16_t)(c & 0x7FFFu);
    }
    endb[i] = (uint8_t)m;
  }
}

// at each segment's last edge: the destination inside its bin, and the full destination
__global__ void seg_rows_kernel(const uint64_t *keys, const uint16_t *code, const uint32_t *segid, uint64_t E,
                                uint32_t wlog, uint32_t b, uint32_t rmask, uint16_t *seg_row, uint32_t *seg_full) {
  const uint64_t bmask = (1ull << b) - 1;
  GS64(j, E) {
    if (code[j] & 0x8000u) {
      const uint32_t row = (uint32_t)((keys[j] >> wlog) & bmask);
      const uint32_t p = segid[j];
      seg_row[p] = (uint16_t)(row & rmask);
      seg_full[p] = row;
    }
  }
}

// sseg[s] = first segment of strip s (an empty strip gets its successor's), sseg[ns] = nseg
__global__ void strip_seg_kernel(const uint64_t *soff, const uint32_t *segid, uint32_t ns, uint64_t E,
                                 uint32_t nseg, uint32_t *sseg) {
  GS64(s, (uint64_t)ns + 1) sseg[s] = (s < ns && soff[s] < E) ? segid[soff[s]] : nseg;
}

// runs[bb*ns + s] = first segment of strip s whose destination >= bb << rlog, bb in [0, nbins]
__global__ void runs_kernel(const uint32_t *seg_full, const uint32_t *sseg, uint32_t ns, uint32_t nbins,
                            uint32_t rlog, uint32_t *runs) {
  GS64(t, (uint64_t)(nbins + 1) * ns) {
    const uint32_t s = (uint32_t)(t % ns), bb = (uint32_t)(t / ns);
    const uint64_t target = (uint64_t)bb << rlog;
    uint32_t lo = sseg[s], hi = sseg[s + 1];
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)seg_full[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    runs[t] = lo;
  }
}

// cnt[bb] = messages of bin bb (one wave per bin)
__global__ void bin_count_kernel(const uint32_t *runs, uint32_t ns, uint32_t nbins, uint32_t *cnt) {
  const uint32_t lane = lane_id();
  const uint32_t wpb = blockDim.x / kWave;
  for (uint32_t bb = blockIdx.x * wpb + threadIdx.x / kWave; bb < nbins; bb += gridDim.x * wpb) {
    uint32_t c = 0;
    for (uint32_t s = lane; s < ns; s += kWave) c += runs[(size_t)(bb + 1) * ns + s] - runs[(size_t)bb * ns + s];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += (uint32_t)__shfl_xor((int)c, o, kWave);
    if (lane == 0) cnt[bb] = c;
  }
}

// ---------------------------------------------------------------- iteration kernels

// The gather kernels hold no static LDS (the host checks), so their dynamic LDS -- the strip's
// values, then the message buffers -- starts at LDS address 0 and a byte offset is an address.
typedef __attribute__((address_space(3))) float lds_f32;
__device__ inline float lds_ld(uint32_t byte) { return *reinterpret_cast<const lds_f32 *>((uintptr_t)byte); }
__device__ inline void lds_st(uint32_t byte, float v) { *reinterpret_cast<lds_f32 *>((uintptr_t)byte) = v; }

// Byte offsets of the strip values the low and high 16-bit codes of a word name (one SDWA shift
// each: the compiler folds the mask of the low half into the shift otherwise).
__device__ inline uint32_t code_lo4(uint32_t w) {
  uint32_t r;
  asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_0"
      : "=v"(r)
      : "v"(w));
  return r;
}
__device__ inline uint32_t code_hi4(uint32_t w) {
  uint32_t r;
  asm("v_lshlrev_b32_sdwa %0, 2, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:WORD_1"
      : "=v"(r)
      : "v"(w));
  return r;
}

// All ones when bit K of nends is set (edge K is not an end), else 0.  Written as the instruction
// itself: the compiler rewrites the 1.0-or-0 factor built from it into a compare and a select,
// and it converts a signed 1-bit field extract to float as if it were unsigned.
template <int K>
__device__ inline uint32_t nonend_mask(uint32_t nends) {
  uint32_t m;
  asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(m) : "v"(nends), "n"(K));
  return m;
}

// a + 4 m (one instruction; m is 0 or all ones, so this is a or a - 4)
__device__ inline uint32_t lshl2_add(uint32_t m, uint32_t a) {
  uint32_t r;
  asm("v_lshl_add_u32 %0, %1, 2, %2" : "=v"(r) : "v"(m), "v"(a));
  return r;
}

// Edge K of a lane's PageRank round (strip_round): the run restarts after an end, and the store
// goes to the slot of the lane's next message.
template <int K>
__device__ inline void pr_edge(float (&x)[8], float &run, uint32_t nends, uint32_t &A) {
  const uint32_t m = nonend_mask<K - 1>(nends);
  run = fmaf(run, __uint_as_float(m & 0x3F800000u), x[K]);
  A = lshl2_add(m, A);
  lds_st(A + 4u * K, run);
  x[K] = run;
}

// Message buffer of one wave in LDS: messages [base, base + fill) of the wave task wait there
// until a round would overflow it, then leave with coalesced stores.
struct MsgBuf {
  float *msg;  // the message array
  float *mb;
  uint32_t base, fill;
};

__device__ inline void flush_messages(MsgBuf &B) {
  wave_sync();
  // four LDS reads in flight before their stores
  uint32_t i = lane_id();
  for (; i + 3 * kWave < B.fill; i += 4 * kWave) {
    const float m0 = B.mb[i], m1 = B.mb[i + kWave], m2 = B.mb[i + 2 * kWave], m3 = B.mb[i + 3 * kWave];
    B.msg[B.base + i] = m0;
    B.msg[B.base + i + kWave] = m1;
    B.msg[B.base + i + 2 * kWave] = m2;
    B.msg[B.base + i + 3 * kWave] = m3;
  }
  for (; i < B.fill; i += kWave) B.msg[B.base + i] = B.mb[i];
  wave_sync();
  B.base += B.fill;
  B.fill = 0;
}

// One PageRank round: 8 consecutive edges per lane from q (codes cv, end bits eb), the wave's
// message buffer at LDS byte address mbb.  FULL: every lane's 8 edges lie inside [e0, e1) or
// entirely past e1 (only the first and last rounds of a task can cut a lane).  The messages of
// the segments that end in the round go to the buffer in segment order; the sum of the segment
// left open stays in `carry`.
// LDS byte offset of code k of a lane's 8: two to a word as the plain layout holds them, or one
// to a word (the packed codes' decoder).
struct WordCodes {
  uint32_t c[8];
};
__device__ inline uint32_t code_addr(u32x4 cv, int k) { return (k & 1) ? code_hi4(cv[k >> 1]) : code_lo4(cv[k >> 1]); }
__device__ inline uint32_t code_addr(const WordCodes &cv, int k) { return cv.c[k] << 2; }
// 4-byte source ids: the low wlog bits (the offset in the strip) times 4, m4 = (2^wlog - 1) * 4.
// The mask also keeps ids read past a task's end (padding, the next strip's edges) inside the
// strip's values.
struct IdWords {
  u32x4 lo, hi;
  uint32_t m4;
};
__device__ inline uint32_t code_addr(const IdWords &cv, int k) { return ((k < 4 ? cv.lo[k] : cv.hi[k - 4]) << 2) & cv.m4; }

template <bool FULL, class Codes>
__device__ inline void strip_round(MsgBuf &B, uint32_t mbb, Codes cv, uint32_t eb, uint32_t q,
                                   uint32_t e0, uint32_t e1, double &carry) {
  const uint32_t lane = lane_id();
  float x[8];
  uint32_t ends = eb;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const float v = lds_ld(code_addr(cv, k));
    if (FULL) {
      x[k] = v;
    } else {
      const bool ok = q + k >= e0 && q + k < e1;
      x[k] = ok ? v : 0.0f;
      if (!ok) ends &= ~(1u << k);
    }
  }
  // a lane past e1 (last round) reads codes of the padding or the next task: it ends nothing,
  // and its sums reach only the carry out of the task's last round, which is dropped
  if (FULL && q >= e1) ends = 0u;
  const uint32_t cnt = (uint32_t)__popc(ends);
  const uint32_t incl = wave_scan_dpp(cnt);
  const uint32_t total = read_lane_u32(incl, kWave - 1);
  // room for the round plus a free slot (needed only when some edge is not an end)
  if (B.fill + total >= kRound) flush_messages(B);
  const uint32_t o0 = B.fill + incl - cnt;  // this lane's first slot
  // Every edge writes slot slot0 + (ends before it).  In a lane with ends, an end writes its
  // message and a non-end writes the slot of the lane's next message (which overwrites it
  // later in program order) or, after the lane's last end, the first slot of the next lane
  // with an end -- its head slot, rewritten below after the scan -- or the free slot past the
  // round's messages.  A lane without ends writes a free slot of its own, one per lane where the
  // buffer has room (lanes writing one LDS address would serialise).
  const uint32_t spare = B.fill + total + min(lane, kRound - 1 - min(B.fill + total, kRound - 1));
  const uint32_t slot0 = ends ? o0 : spare;
  // Edge k's slot address is mbb + 4 (slot0 + k - non-ends before k) = A + 4 k: A drops by 4 after
  // every non-end and 4 k is the store's immediate offset.  The lane's sums (at most 8 edges) run
  // in FP32 and restart after an end: run = fma(run, keep, x) with keep = 1.0 (after a non-end)
  // or 0 -- exactly the rounded run + x, or exactly x.  Sums across lanes and rounds: below.
  const uint32_t nends = ~ends;
  uint32_t A = mbb + 4u * slot0;
  float run = x[0];
  lds_st(A, run);
  pr_edge<1>(x, run, nends, A);
  pr_edge<2>(x, run, nends, A);
  pr_edge<3>(x, run, nends, A);
  pr_edge<4>(x, run, nends, A);
  pr_edge<5>(x, run, nends, A);
  pr_edge<6>(x, run, nends, A);
  pr_edge<7>(x, run, nends, A);
  run *= __uint_as_float(nonend_mask<7>(nends) & 0x3F800000u);  // the open tail (none after an end at 7)
  // the lane's first message before the carry in: its run at the first end (3-level select)
  const uint32_t kf = (uint32_t)__builtin_ctz(ends | 0x80000000u);
  const float t0 = (kf & 1u) ? x[1] : x[0], t1 = (kf & 1u) ? x[3] : x[2];
  const float t2 = (kf & 1u) ? x[5] : x[4], t3 = (kf & 1u) ? x[7] : x[6];
  const float u0 = (kf & 2u) ? t1 : t0, u1 = (kf & 2u) ? t3 : t2;
  const float first = (kf & 4u) ? u1 : u0;
  // The open sums through each lane (FP32, segments start at lanes with ends); a head lane's
  // message adds the open sum through the lane before it, and the carry of the rounds before
  // joins the round's first head lane.  Head lanes (re)write their first slot; the others
  // write their free slot.  The carry stays in FP64.
  const uint64_t hm = __ballot(ends != 0u);
  const float open = wave_segmented_scan_f32(run, hm);
  const float cin = __builtin_amdgcn_update_dpp(0.0f, open, kDppWaveShr1, 0xF, 0xF, true);
  const bool firsthead = __builtin_amdgcn_inverse_ballot_w64(hm & (0ull - hm));
  const double m0 = (double)cin + (double)first;
  lds_st(mbb + 4u * slot0, (float)(firsthead ? m0 + carry : m0));
  const double tail = (double)__builtin_bit_cast(float, read_lane_u32(__builtin_bit_cast(uint32_t, open), kWave - 1));
  carry = hm ? tail : carry + tail;
  B.fill += total;
}

// Edge K of a lane's sweep round (min_round): the running maximum restarts after an end (m = 0
// clears it; 0 is neutral), and the store goes to the slot of the lane's next message.
template <int K>
__device__ inline void min_edge(uint32_t (&x)[8], uint32_t &run, uint32_t nends, uint32_t &A) {
  const uint32_t m = nonend_mask<K - 1>(nends);
  run = max(run & m, x[K]);
  A = lshl2_add(m, A);
  lds_st(A + 4u * K, __uint_as_float(run));
  x[K] = run;
}

// The round of the CC and SSSP sweeps: the value an edge offers its destination is the source's
// complemented label (CC) or the complement of min(distance + weight, 2^32 - 2) (SSSP; 0 from
// an unreached source), a segment's message is the greatest offer, and the open segment's
// value carries over lanes and rounds by a segmented maximum.  Slots, FULL and the layout of
// the round as in strip_round; w0, w1 hold the lane's 8 weights (MODE kSweepWeighted) or its
// 8 byte weights in w0.x, w0.y (kSweepWeighted8).
template <bool FULL, int MODE>
__device__ inline void min_round(MsgBuf &B, uint32_t mbb, u32x4 cv, uint32_t eb, u32x4 w0, u32x4 w1, uint32_t q,
                                 uint32_t e0, uint32_t e1, uint32_t &carry) {
  const uint32_t lane = lane_id();
  uint32_t x[8];
  uint32_t ends = eb;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t cw = cv[k >> 1];
    uint32_t v = __float_as_uint(lds_ld((k & 1) ? code_hi4(cw) : code_lo4(cw)));
    if (MODE != kSweepCC) {
      const uint32_t w = MODE == kSweepWeighted    ? (k < 4 ? w0[k] : w1[k - 4])
                         : MODE == kSweepWeighted8 ? (w0[k >> 2] >> (8 * (k & 3))) & 0xFFu
                                                   : 1u;
      // in complements c = ~d: ~min(d + w, 2^32 - 2) is c - w, or 1 where d + w saturates
      // (c <= w); an unreached source (c = 0) offers nothing
      v = min(v, max(__builtin_elementwise_sub_sat(v, w), 1u));
    }
    if (FULL) {
      x[k] = v;
    } else {
      const bool ok = q + k >= e0 && q + k < e1;
      x[k] = ok ? v : 0u;
      if (!ok) ends &= ~(1u << k);
    }
  }
  if (FULL && q >= e1) ends = 0u;  // past the task's last edge: only the dropped carry sees it
  const uint32_t cnt = (uint32_t)__popc(ends);
  const uint32_t incl = wave_scan_dpp(cnt);
  const uint32_t total = read_lane_u32(incl, kWave - 1);
  if (B.fill + total >= kRound) flush_messages(B);
  const uint32_t o0 = B.fill + incl - cnt;
  const uint32_t spare = B.fill + total + min(lane, kRound - 1 - min(B.fill + total, kRound - 1));
  const uint32_t slot0 = ends ? o0 : spare;
  const uint32_t nends = ~ends;
  uint32_t A = mbb + 4u * slot0;
  uint32_t run = x[0];
  lds_st(A, __uint_as_float(run));
  min_edge<1>(x, run, nends, A);
  min_edge<2>(x, run, nends, A);
  min_edge<3>(x, run, nends, A);
  min_edge<4>(x, run, nends, A);
  min_edge<5>(x, run, nends, A);
  min_edge<6>(x, run, nends, A);
  min_edge<7>(x, run, nends, A);
  run &= nonend_mask<7>(nends);
  const uint32_t kf = (uint32_t)__builtin_ctz(ends | 0x80000000u);
  const uint32_t t0 = (kf & 1u) ? x[1] : x[0], t1 = (kf & 1u) ? x[3] : x[2];
  const uint32_t t2 = (kf & 1u) ? x[5] : x[4], t3 = (kf & 1u) ? x[7] : x[6];
  const uint32_t u0 = (kf & 2u) ? t1 : t0, u1 = (kf & 2u) ? t3 : t2;
  const uint32_t first = (kf & 4u) ? u1 : u0;
  const uint64_t hm = __ballot(ends != 0u);
  const uint32_t open = wave_segmented_max_u32(run, hm);
  const uint32_t cin = dpp_u32<kDppWaveShr1>(open);
  const bool firsthead = __builtin_amdgcn_inverse_ballot_w64(hm & (0ull - hm));
  const uint32_t m0 = max(cin, first);
  lds_st(mbb + 4u * slot0, __uint_as_float(firsthead ? max(m0, carry) : m0));
  const uint32_t tail = read_lane_u32(open, kWave - 1);
  carry = hm ? tail : max(carry, tail);
  B.fill += total;
}

// Sweep messages of edges [e0, e1) whose first segment has index seg (every segment ends
// inside); loads as in pr_wave_task, weights with them.
template <int MODE>
__device__ inline void min_wave_task(const StripArgs &a, float *mb, uint32_t mbb, uint64_t sb, uint32_t e0,
                                     uint32_t e1, uint32_t seg) {
  const uint32_t lane = lane_id();
  const uint32_t b0 = e0 & ~7u;
  const uint32_t nr = (e1 - b0 + kRound - 1) / kRound;
  const u32x4 *cb = reinterpret_cast<const u32x4 *>(a.code + sb + b0);
  const uint8_t *ebb = a.endb + (sb + b0) / 8;
  const u32x4 *wb = reinterpret_cast<const u32x4 *>(a.wt + sb + b0);   // two words of 4 weights per lane
  const u32x2 *w8b = reinterpret_cast<const u32x2 *>(a.wt8 + sb + b0);  // 8 byte weights per lane
  uint32_t carry = 0;  // the open segment's value
  MsgBuf B{a.msg, mb, seg, 0u};
  u32x4 buf[kAhead], wa[kAhead], wc[kAhead];
  uint32_t eba[kAhead];
  auto load = [&](int i, uint32_t r) {
    buf[i] = cb[r * kWave + lane];
    eba[i] = ebb[r * kWave + lane];
    if (MODE == kSweepWeighted) {
      wa[i] = wb[2 * (r * kWave + lane)];
      wc[i] = wb[2 * (r * kWave + lane) + 1];
    }
    if (MODE == kSweepWeighted8) {
      const u32x2 v = w8b[r * kWave + lane];
      wa[i] = u32x4{v.x, v.y, 0u, 0u};
    }
  };
#pragma unroll
  for (int i = 0; i < kAhead; i++) {
    buf[i] = wa[i] = wc[i] = u32x4{0u, 0u, 0u, 0u};
    eba[i] = 0u;
    if ((uint32_t)i < nr) load(i, (uint32_t)i);
  }
  for (uint32_t r = 0; r < nr; r += kAhead) {
#pragma unroll
    for (int i = 0; i < kAhead; i++) {
      if (r + i < nr) {
        const u32x4 cv = buf[i], w0 = wa[i], w1 = wc[i];
        const uint32_t eb = eba[i];
        if (r + i + kAhead < nr) load(i, r + i + kAhead);
        const uint32_t base = b0 + (r + i) * kRound;
        const uint32_t q = base + 8 * lane;
        const bool cut = ((e0 & 7u) && base == b0) || ((e1 & 7u) && base + kRound > e1);
        if (cut) min_round<false, MODE>(B, mbb, cv, eb, w0, w1, q, e0, e1, carry);
        else min_round<true, MODE>(B, mbb, cv, eb, w0, w1, q, e0, e1, carry);
      }
    }
  }
  flush_messages(B);
}

// PageRank messages of edges [e0, e1) whose first segment has index seg; every segment ends
// inside.  The loads of codes and end bits run kAhead rounds ahead of their use; a round's loads
// are whole (the layout's padding keeps those of the last round in bounds).  They are plain
// loads: a round's codes start anywhere in a 128-byte line (a strip starts at any multiple of 8
// edges), so its first and last lines are shared with the rounds before and after, and end bytes
// are half a line a round; streaming loads fetched those lines twice (6 % more bytes).
__device__ inline void pr_wave_task(const StripArgs &a, float *mb, uint32_t mbb, uint64_t sb, uint32_t e0,
                                    uint32_t e1, uint32_t seg) {
  const uint32_t lane = lane_id();
  const uint32_t b0 = e0 & ~7u;
  const uint32_t nr = (e1 - b0 + kRound - 1) / kRound;
  const u32x4 *cb = reinterpret_cast<const u32x4 *>(a.code + sb + b0);  // a round is kWave words of 8 codes
  const uint8_t *ebb = a.endb + (sb + b0) / 8;
  double carry = 0.0;  // the open segment's sum
  MsgBuf B{a.msg, mb, seg, 0u};
  u32x4 buf[kAhead];
  uint32_t eba[kAhead];
#pragma unroll
  for (int i = 0; i < kAhead; i++) {
    buf[i] = u32x4{0u, 0u, 0u, 0u};
    eba[i] = 0u;
    if ((uint32_t)i < nr) {
      buf[i] = cb[(uint32_t)i * kWave + lane];
      eba[i] = ebb[(uint32_t)i * kWave + lane];
    }
  }
  for (uint32_t r = 0; r < nr; r += kAhead) {
#pragma unroll
    for (int i = 0; i < kAhead; i++) {
      if (r + i < nr) {
        const u32x4 cv = buf[i];
        const uint32_t eb = eba[i];
        const uint32_t rn = r + i + kAhead;
        if (rn < nr) {
          buf[i] = cb[rn * kWave + lane];
          eba[i] = ebb[rn * kWave + lane];
        }
        const uint32_t base = b0 + (r + i) * kRound;
        const uint32_t q = base + 8 * lane;
        // lanes are cut only where e0 or e1 is not a multiple of 8 inside this round
        const bool cut = ((e0 & 7u) && base == b0) || ((e1 & 7u) && base + kRound > e1);
        if (cut) strip_round<false>(B, mbb, cv, eb, q, e0, e1, carry);
        else strip_round<true>(B, mbb, cv, eb, q, e0, e1, carry);
      }
    }
  }
  flush_messages(B);
}

// ---- packed codes (kSweepPRPacked) ----------------------------------------------------------
// A block serves one 512-edge round of a strip: 8 header words (a separate array) hold the 64
// lanes' widths (4 bits each, lane l in bits 4 (l % 8) of word l / 8); the block's words in the
// stream hold the lanes' groups bit after bit, least significant bit first.  A group is the
// lane's first code (15 bits) and then, for edges 1..7, a field: the code itself (15 bits) where
// the edge starts a segment (the end byte says the edge before it ends one), else the code minus
// the code before it (codes ascend inside a segment), at the group's width.  A lane finds its
// group from the sizes of the groups before it (a wave prefix sum) and gets its 8 codes back
// exactly, so the round proceeds as with the plain codes and the sums are the same bits.

// Bits of a group of width b whose end byte is eb.
__device__ inline uint32_t pk_group_bits(uint32_t b, uint32_t eb) {
  const uint32_t nst = (uint32_t)__popc(eb & 0x7Fu);
  return 15u + 7u * b + nst * (15u - b);
}

// The 8 codes of the group whose first bit is bit sh of w0; w0..w4 are the five words from there
// (a group has at most 120 bits).  Fields go in pairs (two fit one word: the first code with
// field 1, then fields 2-3, 4-5, 6-7); after a pair the window shifts by their widths and keeps
// only the words the fields left can fill (15 bits each at most).
__device__ inline WordCodes pk_unpack(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t w4, uint32_t sh,
                                      uint32_t b, uint32_t eb) {
  uint32_t d0 = __builtin_amdgcn_alignbit(w1, w0, sh), d1 = __builtin_amdgcn_alignbit(w2, w1, sh);
  uint32_t d2 = __builtin_amdgcn_alignbit(w3, w2, sh), d3 = __builtin_amdgcn_alignbit(w4, w3, sh);
  const uint32_t neb = ~eb;
  const int bm = (int)b - 15;
  // field j: k = 1 for a difference (edge j - 1 does not end a segment) of width b, else a code
  // of 15 bits; code j = k * code (j - 1) + field
  auto keep = [&](int j) { return __builtin_amdgcn_ubfe(neb, (uint32_t)(j - 1), 1u); };
  auto next = [](uint32_t c, uint32_t k, uint32_t v) {  // c * k + v (k is 0 or 1)
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(c), "v"(k), "v"(v));
    return r;
  };
  auto width = [&](uint32_t k) { return (uint32_t)(__mul24((int)k, bm) + 15); };
  const uint32_t k1 = keep(1), k2 = keep(2), k3 = keep(3), k4 = keep(4), k5 = keep(5), k6 = keep(6), k7 = keep(7);
  const uint32_t x1 = width(k1), x2 = width(k2), x3 = width(k3), x4 = width(k4), x5 = width(k5), x6 = width(k6),
                 x7 = width(k7);
  WordCodes o;
  o.c[0] = d0 & 0x7FFFu;
  o.c[1] = next(o.c[0], k1, __builtin_amdgcn_ubfe(d0, 15u, x1));
  uint32_t s = 15u + x1;  // 90 bits left at most
  d0 = __builtin_amdgcn_alignbit(d1, d0, s);
  d1 = __builtin_amdgcn_alignbit(d2, d1, s);
  d2 = __builtin_amdgcn_alignbit(d3, d2, s);
  o.c[2] = next(o.c[1], k2, __builtin_amdgcn_ubfe(d0, 0u, x2));
  o.c[3] = next(o.c[2], k3, __builtin_amdgcn_ubfe(d0, x2, x3));
  s = x2 + x3;  // 60 bits left at most
  d0 = __builtin_amdgcn_alignbit(d1, d0, s);
  d1 = __builtin_amdgcn_alignbit(d2, d1, s);
  o.c[4] = next(o.c[3], k4, __builtin_amdgcn_ubfe(d0, 0u, x4));
  o.c[5] = next(o.c[4], k5, __builtin_amdgcn_ubfe(d0, x4, x5));
  s = x4 + x5;  // 30 bits left at most
  d0 = __builtin_amdgcn_alignbit(d1, d0, s);
  o.c[6] = next(o.c[5], k6, __builtin_amdgcn_ubfe(d0, 0u, x6));
  o.c[7] = next(o.c[6], k7, __builtin_amdgcn_ubfe(d0, x6, x7));
  return o;
}

// A round's group words and what decodes them.
struct PkRound {
  uint32_t w0, w1, w2, w3, w4, sh, bw, eb;
};

// pr_wave_task over packed codes; round r of the task is block blk + r.  The task's width words
// (8 a round, 16 rounds at most) are loaded at its start, two a lane, and handed to the lanes
// that need them by ds_bpermute; a round's end bytes are loaded 6 rounds ahead of its use, its
// group words (whose address needs the widths) 2 rounds ahead.
__device__ inline void pr_wave_task_packed(const StripArgs &a, float *mb, uint32_t mbb, uint64_t sb, uint32_t e0,
                                           uint32_t e1, uint32_t seg, uint32_t blk) {
  const uint32_t lane = lane_id();
  const uint32_t b0 = e0 & ~7u;
  const uint32_t nr = (e1 - b0 + kRound - 1) / kRound;  // at most kWaveTaskEdges / kRound (the packer checks)
  const uint8_t *ebb = a.endb + (sb + b0) / 8;
  static_assert(kWaveTaskEdges / kRound <= 16, "a task's width words must fit two words a lane");
  // the task's block offsets, one per lane, and its width words, two per lane
  const uint32_t pov = lane < nr ? a.pkoff[blk + lane] : 0u;
  const uint32_t *hp = a.pkhdr + (size_t)blk * 8u;
  const uint32_t hv0 = lane < 8u * nr ? hp[lane] : 0u;
  const uint32_t hv1 = lane + kWave < 8u * nr ? hp[lane + kWave] : 0u;
  const int hsrc = (int)(4u * (lane >> 3));  // ds_bpermute byte address of the lane's word in round 0
  const uint32_t hsh = 4u * (lane & 7u);      // the lane's width in that word
  double carry = 0.0;
  MsgBuf B{a.msg, mb, seg, 0u};
  uint32_t eba[kAhead], hda[kAhead], boa[kAhead];  // round q's head in slot q % kAhead
  auto load_head = [&](int i, uint32_t r) {
    eba[i] = ebb[r * kWave + lane];
    boa[i] = read_lane_u32(pov, (int)r);
    hda[i] = (uint32_t)__builtin_amdgcn_ds_bpermute(hsrc + (int)(32u * (r & 7u)), (int)(r < 8u ? hv0 : hv1));
  };
  PkRound G[2];  // round q's groups in G[q % 2]
  auto load_groups = [&](PkRound &g, int i) {
    g.eb = eba[i];
    g.bw = __builtin_amdgcn_ubfe(hda[i], hsh, 4u);
    const uint32_t bits = pk_group_bits(g.bw, g.eb);
    const uint32_t o = wave_scan_dpp(bits) - bits;
    const uint32_t *p = a.pk + boa[i] + (o >> 5);
    g.sh = o & 31u;
    // plain loads: a block's first and last lines are shared with the rounds before and after,
    // which find them in L2 (streaming loads fetched them again, 6 % more bytes)
    g.w0 = p[0];
    g.w1 = p[1];
    g.w2 = p[2];
    g.w3 = p[3];
    g.w4 = p[4];
  };
#pragma unroll
  for (int i = 0; i < kAhead; i++) {
    eba[i] = hda[i] = boa[i] = 0u;
    if ((uint32_t)i < nr) load_head(i, (uint32_t)i);
  }
  G[0] = G[1] = PkRound{0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
  for (int i = 0; i < 2; i++)
    if ((uint32_t)i < nr) {
      load_groups(G[i], i);
      if ((uint32_t)(i + kAhead) < nr) load_head(i, (uint32_t)(i + kAhead));
    }
  static_assert(kAhead % 2 == 0, "round parity must match the slot's");
  for (uint32_t r = 0; r < nr; r += kAhead) {
#pragma unroll
    for (int i = 0; i < kAhead; i++) {
      const uint32_t rr = r + (uint32_t)i;
      if (rr < nr) {
        PkRound &g = G[i & 1];
        const WordCodes cv = pk_unpack(g.w0, g.w1, g.w2, g.w3, g.w4, g.sh, g.bw, g.eb);
        const uint32_t ebr = g.eb;
        if (rr + 2 < nr) {  // round rr + 2's groups, then its head slot refilled kAhead ahead
          const int in = (i + 2) % kAhead;
          load_groups(g, in);
          if (rr + 2 + kAhead < nr) load_head(in, rr + 2 + kAhead);
        }
        const uint32_t base = b0 + rr * kRound;
        const uint32_t q = base + 8 * lane;
        const bool cut = ((e0 & 7u) && base == b0) || ((e1 & 7u) && base + kRound > e1);
        if (cut) strip_round<false>(B, mbb, cv, ebr, q, e0, e1, carry);
        else strip_round<true>(B, mbb, cv, ebr, q, e0, e1, carry);
      }
    }
  }
  flush_messages(B);
}

// pr_wave_task over 4-byte source ids: a lane's 8 ids are two 16-byte loads (a round reads 2 KiB
// of ids against 1 KiB of codes); the sums are the same bits as over the codes.
__device__ inline void pr_wave_task_ids(const StripArgs &a, float *mb, uint32_t mbb, uint64_t sb, uint32_t e0,
                                        uint32_t e1, uint32_t seg) {
  const uint32_t lane = lane_id();
  const uint32_t b0 = e0 & ~7u;
  const uint32_t nr = (e1 - b0 + kRound - 1) / kRound;
  const u32x4 *ib = reinterpret_cast<const u32x4 *>(a.ids + sb + b0);  // a round is 2 kWave words of 4 ids
  const uint8_t *ebb = a.endb + (sb + b0) / 8;
  const uint32_t m4 = ((1u << a.wlog) - 1u) << 2;
  double carry = 0.0;
  MsgBuf B{a.msg, mb, seg, 0u};
  u32x4 blo[kAhead], bhi[kAhead];
  uint32_t eba[kAhead];
#pragma unroll
  for (int i = 0; i < kAhead; i++) {
    blo[i] = bhi[i] = u32x4{0u, 0u, 0u, 0u};
    eba[i] = 0u;
    if ((uint32_t)i < nr) {
      blo[i] = ib[2 * ((uint32_t)i * kWave + lane)];
      bhi[i] = ib[2 * ((uint32_t)i * kWave + lane) + 1];
      eba[i] = ebb[(uint32_t)i * kWave + lane];
    }
  }
  for (uint32_t r = 0; r < nr; r += kAhead) {
#pragma unroll
    for (int i = 0; i < kAhead; i++) {
      if (r + i < nr) {
        const IdWords cv{blo[i], bhi[i], m4};
        const uint32_t eb = eba[i];
        const uint32_t rn = r + i + kAhead;
        if (rn < nr) {
          blo[i] = ib[2 * (rn * kWave + lane)];
          bhi[i] = ib[2 * (rn * kWave + lane) + 1];
          eba[i] = ebb[rn * kWave + lane];
        }
        const uint32_t base = b0 + (r + i) * kRound;
        const uint32_t q = base + 8 * lane;
        const bool cut = ((e0 & 7u) && base == b0) || ((e1 & 7u) && base + kRound > e1);
        if (cut) strip_round<false>(B, mbb, cv, eb, q, e0, e1, carry);
        else strip_round<true>(B, mbb, cv, eb, q, e0, e1, carry);
      }
    }
  }
  flush_messages(B);
}

// Messages of edges [e0, e1) whose first segment has index seg (blk: the task's first block of
// packed codes).
template <int MODE>
__device__ inline void strip_wave_task(const StripArgs &a, float *mb, uint32_t mbb, uint64_t sb, uint32_t e0,
                                       uint32_t e1, uint32_t seg, uint32_t blk) {
  if constexpr (MODE == kSweepPR) pr_wave_task(a, mb, mbb, sb, e0, e1, seg);
  else if constexpr (MODE == kSweepPRPacked) pr_wave_task_packed(a, mb, mbb, sb, e0, e1, seg, blk);
  else if constexpr (MODE == kSweepPRIds) pr_wave_task_ids(a, mb, mbb, sb, e0, e1, seg);
  else min_wave_task<MODE>(a, mb, mbb, sb, e0, e1, seg);
}

// One strip's values (contributions, or complemented labels for CC) staged in LDS, then the
// strip's wave tasks.
template <int MODE>
__device__ inline void strip_gather(const StripArgs &a, float *sv) {
  const uint32_t W = 1u << a.wlog;
  const uint32_t *vals = !is_pr(MODE) ? a.nlab : reinterpret_cast<const uint32_t *>(a.contrib_in);
  const uint32_t wv = wave_id();
  float *mb = sv + W + wv * kRound;
  const uint32_t mbb = 4u * (W + wv * kRound);  // mb as an LDS address
  // Workgroups take tasks from a queue (heaviest first), so they finish together; the task index
  // reaches the whole workgroup through its slot in gq (the LDS is full).
  volatile uint32_t *slot = a.gq + 2 + blockIdx.x;
  for (;;) {
    if (threadIdx.x == 0) *slot = atomicAdd(a.gq, 1u);
    __syncthreads();
    const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)*slot);
    __syncthreads();  // every thread has read the slot before the next task is taken
    if (t >= a.ngt) break;
    const uint4 gt = a.gtask[t];
    // a sweep skips strips none of whose vertices changed in the previous sweep; the bins apply
    // the messages of gathered strips only
    if (!is_pr(MODE)) {
      if (a.sflag[gt.x] + 1u != a.sweep) continue;
      if (threadIdx.x == 0) {
        atomicAdd(a.edges, (unsigned long long)(a.wtask[gt.z - 1].y - a.wtask[gt.y].x));
        a.gflag[gt.x] = a.sweep;
      }
    }
    const uint32_t v0 = gt.x << a.wlog;
    const uint32_t n = min(W, a.V - v0);
    __syncthreads();  // the previous task's readers are done
    if (!is_pr(MODE) && a.fmask) {
      // pull round of delta-stepping: complemented distances of the frontier's sources, 0 (no
      // offer) for every other source
      uint32_t *su = reinterpret_cast<uint32_t *>(sv);
      if (n == W && (v0 & 3u) == 0) {
        const u32x4 *dd = reinterpret_cast<const u32x4 *>(a.dist + v0);
        const u32x4 *ff = reinterpret_cast<const u32x4 *>(a.fmask + v0);
        u32x4 *dst = reinterpret_cast<u32x4 *>(sv);
        const uint32_t fs = a.fstamp;
        for (uint32_t i = threadIdx.x; i < W / 4; i += kIterBlock) {
          const u32x4 d = dd[i], f = ff[i];
          u32x4 o;
          o.x = f.x == fs ? ~d.x : 0u;
          o.y = f.y == fs ? ~d.y : 0u;
          o.z = f.z == fs ? ~d.z : 0u;
          o.w = f.w == fs ? ~d.w : 0u;
          dst[i] = o;
        }
      } else {
        for (uint32_t i = threadIdx.x; i < n; i += kIterBlock) {
          const uint32_t d = a.dist[v0 + i];
          su[i] = a.fmask[v0 + i] == a.fstamp ? ~d : 0u;
        }
      }
    } else if (n == W && (v0 & 3u) == 0) {
      // whole strip: 16-byte loads, four in flight per thread before their stores
      const u32x4 *src = reinterpret_cast<const u32x4 *>(vals + v0);
      u32x4 *dst = reinterpret_cast<u32x4 *>(sv);
      const uint32_t n4 = W / 4;
      uint32_t i = threadIdx.x;
      for (; i + 3 * kIterBlock < n4; i += 4 * kIterBlock) {
        const u32x4 t0 = src[i], t1 = src[i + kIterBlock], t2 = src[i + 2 * kIterBlock], t3 = src[i + 3 * kIterBlock];
        dst[i] = t0;
        dst[i + kIterBlock] = t1;
        dst[i + 2 * kIterBlock] = t2;
        dst[i + 3 * kIterBlock] = t3;
      }
      for (; i < n4; i += kIterBlock) dst[i] = src[i];
    } else {
      for (uint32_t i = threadIdx.x; i < n; i += kIterBlock) reinterpret_cast<uint32_t *>(sv)[i] = vals[v0 + i];
    }
    __syncthreads();
    const uint64_t sb = a.sbase[gt.x];
    const uint32_t blk = MODE == kSweepPRPacked ? a.pkblk[gt.x] : 0u;
    for (uint32_t w = gt.y + wv; w < gt.z; w += kIterWaves) {
      const uint4 wt = a.wtask[w];
      // a task's rounds start at a multiple of kRound from the strip's base (the first task at
      // 0, the others at multiples of kWaveTaskEdges), one packed block each
      strip_wave_task<MODE>(a, mb, mbb, sb, wt.x, wt.y, wt.z, blk + (wt.x & ~7u) / kRound);
    }
  }
  // the last workgroup to leave resets the queue for the next launch
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(a.gq + 1, 1u) == gridDim.x - 1) {
      atomicExch(a.gq, 0u);
      atomicExch(a.gq + 1, 0u);
    }
  }
}

// An iteration queued after the solve converged (sc->done) does nothing, in every kernel of it.
__global__ __launch_bounds__(kIterBlock) void pr_strip_gather_kernel(StripArgs a) {
  extern __shared__ float sv[];  // contributions of the task's strip, then a message buffer per wave
  if (a.sc->done) return;
  strip_gather<kSweepPR>(a, sv);
}

// The same gather over the packed codes.
__global__ __launch_bounds__(kIterBlock) void pr_strip_gather_ids_kernel(StripArgs a) {
  extern __shared__ float sv[];
  if (a.sc->done) return;
  strip_gather<kSweepPRIds>(a, sv);
}

__global__ __launch_bounds__(kIterBlock) void pr_strip_gather_packed_kernel(StripArgs a) {
  extern __shared__ float sv[];
  if (a.sc->done) return;
  strip_gather<kSweepPRPacked>(a, sv);
}

// ---- packing the codes (one wave per block; block k covers edges blkpos[k] .. + kRound - 1)

// Lane group at edge q (a multiple of 8): its codes, end byte and width.
struct PkGroup {
  uint32_t c[8];
  uint32_t eb, b;
  bool desc;  // a code below the one before it inside a segment (the layout never has one)
};

__device__ inline PkGroup pk_group(const uint16_t *code, const uint8_t *endb, uint64_t q) {
  PkGroup g;
  const u32x4 v = *reinterpret_cast<const u32x4 *>(code + q);
#pragma unroll
  for (int k = 0; k < 8; k++) g.c[k] = (k & 1) ? v[k >> 1] >> 16 : v[k >> 1] & 0xFFFFu;
  g.eb = endb[q / 8];
  uint32_t m = 0;  // the width of the widest difference is the width of their OR
  g.desc = false;
#pragma unroll
  for (int j = 1; j < 8; j++)
    if (!((g.eb >> (j - 1)) & 1u)) {
      m |= g.c[j] - g.c[j - 1];
      g.desc |= g.c[j] < g.c[j - 1];
    }
  g.b = g.desc ? 15u : m ? 32u - (uint32_t)__clz(m) : 0u;
  return g;
}

// words[k] = the words of block k; *bad set where codes descend inside a segment
__global__ void pk_size_kernel(const uint16_t *code, const uint8_t *endb, const uint64_t *blkpos, uint32_t nblk,
                               uint32_t *words, uint32_t *bad) {
  const uint32_t lane = lane_id(), wpb = blockDim.x / kWave;
  for (uint32_t k = blockIdx.x * wpb + threadIdx.x / kWave; k < nblk; k += gridDim.x * wpb) {
    const PkGroup g = pk_group(code, endb, blkpos[k] + 8u * lane);
    if (g.desc) atomicOr(bad, 1u);
    const uint32_t tot = read_lane_u32(wave_scan_dpp(pk_group_bits(g.b, g.eb)), kWave - 1);
    if (lane == 0) words[k] = (tot + 31u) / 32u;
  }
}

// v (below 2^w) at bit pos of the wave's block image w
__device__ inline void pk_put(uint32_t *w, uint32_t &pos, uint32_t v, uint32_t wd) {
  if (!wd) return;
  const uint32_t s = pos & 31u;
  atomicOr(&w[pos >> 5], v << s);
  if (s + wd > 32u) atomicOr(&w[(pos >> 5) + 1], v >> (32u - s));
  pos += wd;
}

constexpr uint32_t kPkBlockWords = 8 + (kWave * 120 + 31) / 32;  // a block's most words (248)

// Block k to pk + off[k], its widths to hdr + 8 k: assembled in LDS (groups share words; the
// image's first 8 words are the widths), then stored whole.
__global__ __launch_bounds__(kBlock) void pk_write_kernel(const uint16_t *code, const uint8_t *endb,
                                                          const uint64_t *blkpos, const uint32_t *off, uint32_t nblk,
                                                          uint32_t *pk, uint32_t *hdr) {
  __shared__ uint32_t img[kBlock / kWave][kPkBlockWords];
  const uint32_t lane = lane_id(), wv = threadIdx.x / kWave, wpb = blockDim.x / kWave;
  uint32_t *w = img[wv];
  for (uint32_t k = blockIdx.x * wpb + wv; k < nblk; k += gridDim.x * wpb) {
    for (uint32_t i = lane; i < kPkBlockWords; i += kWave) w[i] = 0u;
    wave_sync();
    const PkGroup g = pk_group(code, endb, blkpos[k] + 8u * lane);
    const uint32_t bits = pk_group_bits(g.b, g.eb);
    const uint32_t incl = wave_scan_dpp(bits);
    const uint32_t tot = read_lane_u32(incl, kWave - 1);
    atomicOr(&w[lane >> 3], g.b << (4u * (lane & 7u)));
    uint32_t pos = 256u + incl - bits;
    pk_put(w, pos, g.c[0], 15u);
#pragma unroll
    for (int j = 1; j < 8; j++) {
      const bool st = (g.eb >> (j - 1)) & 1u;
      pk_put(w, pos, st ? g.c[j] : g.c[j] - g.c[j - 1], st ? 15u : g.b);
    }
    wave_sync();
    const uint32_t nw = (tot + 31u) / 32u;
    if (lane < 8u) hdr[(size_t)k * 8u + lane] = w[lane];
    for (uint32_t i = lane; i < nw; i += kWave) pk[off[k] + i] = w[8u + i];
    wave_sync();
  }
}

__global__ void narrow_offsets_kernel(const uint64_t *in, uint32_t n, uint32_t *out) {
  GS64(i, n) out[i] = (uint32_t)in[i];
}

// complemented labels or distances of the task's strip, then the message buffers
template <int MODE>
__global__ __launch_bounds__(kIterBlock) void min_strip_gather_kernel(StripArgs a) {
  extern __shared__ float sv[];
  strip_gather<MODE>(a, sv);
}

// LDS of a bin task: 2^rlog accumulators of type ACC, then the bin's run table -- the first
// message of each run, the exclusive prefix of run lengths (ns + 1) and the chunk index (the
// run holding message bt.y + c * kChunk).
template <class ACC>
struct BinLds {
  ACC *acc;
  uint32_t *rbase, *pre, *cidx;
};

template <class ACC>
__device__ inline BinLds<ACC> bin_lds(void *lds, uint32_t R, uint32_t ns) {
  BinLds<ACC> L;
  L.acc = reinterpret_cast<ACC *>(lds);
  L.rbase = reinterpret_cast<uint32_t *>(L.acc + R);
  L.pre = L.rbase + ns;
  L.cidx = L.pre + ns + 1;
  return L;
}

// Sets every accumulator to init and builds the run table of bin task bt (all threads of the
// block; thread tid owns strips [s0, s1)).  The task is the slice [bt.y, bt.z) of the bin's
// messages in strip order; the table lists, per strip, where its part of the slice starts in
// msg (rbase) and where it starts among the task's live messages (pre): with IDLE, strips the
// sweep did not gather contribute nothing (their messages were applied when they were made,
// and applying a maximum again changes nothing), so later loops visit live messages only.
// Returns the number of live messages (the same in every thread).
template <bool IDLE, class ACC>
__device__ inline uint32_t bin_prepare(const StripArgs &a, const uint4 &bt, const BinLds<ACC> &L, ACC init,
                                       uint32_t *wtot, uint32_t *live, uint32_t s0, uint32_t s1) {
  const uint32_t R = 1u << a.rlog, ns = a.ns, b = bt.x;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wv = wave_id();
  __syncthreads();  // the previous task is done with the LDS
  for (uint32_t i = tid; i < R; i += kIterBlock) L.acc[i] = init;
  // run lengths and their prefix over the bin: the coordinates of bt.y and bt.z
  uint32_t sum = 0;
  for (uint32_t s = s0; s < s1; s++) {
    const uint32_t r0 = a.runs[(size_t)b * ns + s], r1 = a.runs[(size_t)(b + 1) * ns + s];
    L.rbase[s] = r0;
    L.pre[s] = r1 - r0;
    sum += r1 - r0;
  }
  uint32_t incl = wave_inclusive_scan(sum);
  if (lane == kWave - 1) wtot[wv] = incl;
  __syncthreads();
  uint32_t run = incl - sum;
  for (uint32_t w = 0; w < wv; w++) run += wtot[w];
  // each strip's live part of the slice
  uint32_t sum2 = 0;
  for (uint32_t s = s0; s < s1; s++) {
    const uint32_t p0 = run, p1 = run + L.pre[s];
    run = p1;
    const uint32_t q0 = max(p0, bt.y), q1 = min(p1, bt.z);
    const bool idle = IDLE && a.gflag[s] != a.sweep;  // not gathered in this sweep
    const uint32_t len = (q1 > q0 && !idle) ? q1 - q0 : 0u;
    L.rbase[s] += len ? q0 - p0 : 0u;
    L.pre[s] = len;
    sum2 += len;
  }
  __syncthreads();  // every thread has read wtot
  incl = wave_inclusive_scan(sum2);
  if (lane == kWave - 1) wtot[wv] = incl;
  __syncthreads();
  run = incl - sum2;
  for (uint32_t w = 0; w < wv; w++) run += wtot[w];
  for (uint32_t s = s0; s < s1; s++) {
    const uint32_t l = L.pre[s];
    L.pre[s] = run;
    run += l;
  }
  if (tid == kIterBlock - 1) {
    L.pre[ns] = run;
    *live = run;
  }
  __syncthreads();
  const uint32_t total = *live;
  // chunk index: the strip holding every kChunk-th live message bounds the searches below
  const uint32_t nch = (total + kChunk - 1) / kChunk;
  for (uint32_t s = s0; s < s1; s++) {
    const uint32_t p0 = L.pre[s], p1 = L.pre[s + 1];
    for (uint32_t c = (p0 + kChunk - 1) / kChunk; c * kChunk < p1; c++) L.cidx[c] = s;
  }
  if (tid == 0) L.cidx[nch] = ns - 1;
  __syncthreads();
  return total;
}

// Calls add(row, message word) for the n live messages of the task (bin_prepare's count),
// kUnroll messages in flight per thread; the caller synchronises afterwards.
template <class ACC, class F>
__device__ inline void bin_messages(const StripArgs &a, uint32_t n, const BinLds<ACC> &L, F add) {
  const uint32_t *msgw = reinterpret_cast<const uint32_t *>(a.msg);
  for (uint32_t jb = threadIdx.x; jb < n; jb += kIterBlock * kUnroll) {
    uint32_t m[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint32_t j = jb + u * kIterBlock;
      m[u] = 0xFFFFFFFFu;
      if (j < n) {
        // the strip holding live message j: the last s with pre[s] <= j (strips with no live
        // message share the next one's pre and are never the last)
        const uint32_t c = j / kChunk;
        uint32_t lo = L.cidx[c], hi = L.cidx[c + 1];
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (L.pre[mid] <= j) lo = mid;
          else hi = mid - 1;
        }
        m[u] = L.rbase[lo] + (j - L.pre[lo]);
      }
    }
    uint32_t x[kUnroll], r[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      if (m[u] != 0xFFFFFFFFu) {
        // plain loads: the lines at a run's borders are shared with the neighbouring bins
        x[u] = msgw[m[u]];
        r[u] = a.seg_row[m[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++)
      if (m[u] != 0xFFFFFFFFu) add(r[u], x[u]);
  }
}

__global__ __launch_bounds__(kIterBlock) void pr_bin_kernel(StripArgs a) {
  extern __shared__ unsigned long long lds2[];
  __shared__ uint32_t wtot[kIterWaves], live;
  if (a.sc->done) return;
  const uint32_t R = 1u << a.rlog, ns = a.ns;
  const BinLds<unsigned long long> L = bin_lds<unsigned long long>(lds2, R, ns);
  const uint32_t tid = threadIdx.x;
  const float base = a.sc->base, extra = a.sc->extra;
  const uint32_t per = (ns + kIterBlock - 1) / kIterBlock;
  const uint32_t s0 = min(tid * per, ns), s1 = min(s0 + per, ns);
  PrAcc pacc;
  for (uint32_t t = xcd_block_id(); t < a.nbt; t += gridDim.x) {
    const uint4 bt = a.btask[t];
    const uint32_t n = bin_prepare<false>(a, bt, L, 0ull, wtot, &live, s0, s1);
    bin_messages(a, n, L, [&](uint32_t r, uint32_t x) { atomicAdd(&L.acc[r], fix_from_float(__uint_as_float(x))); });
    __syncthreads();
    if (bt.w == kNoSlot) {
      const uint32_t v0 = bt.x << a.rlog;
      pr_apply_rows<kUnroll>(
          a, v0, min(R, a.V - v0), tid, kIterBlock, [&](uint32_t i) { return (float)fix_to_double(L.acc[i]); }, base,
          extra, pacc);
    } else {
      unsigned long long *dst = a.bpart + ((size_t)bt.w << a.rlog);
      for (uint32_t i = tid; i < R; i += kIterBlock) dst[i] = L.acc[i];
    }
  }
  block_store_acc<kIterWaves>(pacc, a.pr1, a.pd1);
}

__global__ __launch_bounds__(kBlock) void pr_split_apply_kernel(StripArgs a) {
  if (a.sc->done) return;
  const float base = a.sc->base, extra = a.sc->extra;
  const uint32_t R = 1u << a.rlog;
  PrAcc pacc;
  const uint32_t nck = (R + kSplitRows - 1) / kSplitRows;  // row chunks per split bin
  for (uint32_t t = blockIdx.x; t < a.nsplit * nck; t += gridDim.x) {
    const uint4 sp = a.split[t / nck];
    const uint32_t r0 = (t % nck) * kSplitRows;
    const uint32_t v0 = sp.x << a.rlog;
    const uint32_t n = min(min(R, a.V - v0), r0 + kSplitRows);
    for (uint32_t i0 = r0 + threadIdx.x; i0 < n; i0 += kBlock * 4) {
      unsigned long long s[4] = {0ull, 0ull, 0ull, 0ull};
      for (uint32_t k = 0; k < sp.z; k++) {
        const unsigned long long *part = a.bpart + ((size_t)(sp.y + k) << a.rlog);
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (i0 + u * kBlock < n) s[u] += part[i0 + u * kBlock];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (i0 + u * kBlock < n) pr_apply(a, v0 + i0 + u * kBlock, (float)fix_to_double(s[u]), base, extra, pacc);
    }
  }
  block_store_acc<kBlock / kWave>(pacc, a.pr2, a.pd2);
}

// CC and SSSP bins: LDS atomicMax of the messages (exact, order-free), then c = max(c, bin
// maximum) in place, marking the strip of every vertex that changed; split bins leave their
// maxima in bpart (as 32-bit words).
__device__ inline void min_count_changed(uint32_t n, uint32_t *changed) {
  n = read_lane_u32(wave_scan_dpp(n), kWave - 1);
  if (lane_id() == 0 && n) atomicAdd(changed, n);
}

// Pull round of delta-stepping: distance v takes the round's best offer m (complemented) when it
// is shorter; the caller appends v to the next near frontier (*to_near) or the far pile (*to_far)
// exactly as a push round would.  Each vertex is applied by one thread, so plain stores suffice.
__device__ inline bool pull_apply(const StripArgs &a, uint32_t v, uint32_t m, uint32_t old, bool &to_near,
                                  bool &to_far) {
  to_near = to_far = false;
  const uint32_t nd = ~m;
  if (!m || nd >= old) return false;
  a.dist[v] = nd;
  if (nd < a.thr) {
    a.near_stamp[v] = a.next_stamp;
    to_near = true;
  } else if (a.far_out && a.far_flag[v] == 0u) {
    a.far_flag[v] = 1u;
    to_far = true;
  }
  return true;
}

// End of a pull round's bin or split kernel: the block's changed count and the out-degree sum of
// its near appends, one atomic each (every thread of the block calls it).
template <int WAVES>
__device__ inline void pull_count_block(const StripArgs &a, uint32_t changed, unsigned long long ndeg, uint32_t *sh) {
  __shared__ unsigned long long shd[WAVES];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ndeg += (unsigned long long)__shfl_xor((long long)ndeg, o, kWave);
  changed = read_lane_u32(wave_scan_dpp(changed), kWave - 1);
  __syncthreads();
  if (lane_id() == 0) {
    shd[wave_id()] = ndeg;
    sh[wave_id()] = changed;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long d = 0;
    uint32_t c = 0;
    for (int w = 0; w < WAVES; w++) {
      d += shd[w];
      c += sh[w];
    }
    if (d) atomicAdd(a.near_deg, d);
    if (c) atomicAdd(a.changed, c);
  }
}

__device__ inline void min_apply(const StripArgs &a, uint32_t v, uint32_t m, uint32_t &changed) {
  const uint32_t old = a.nlab[v];
  if (m > old) {
    a.nlab[v] = m;
    a.sflag[v >> a.wlog] = a.sweep;
    if (a.vstamp) a.vstamp[v] = a.sweep;
    changed++;
  }
}

// Messages of strips gathered this sweep in the whole bin b (the same for every thread): reads
// the strip flags and the run bounds of active strips only, so a bin that receives nothing skips
// the run table and the LDS accumulators of bin_prepare.
__device__ inline uint32_t bin_live_whole(const StripArgs &a, uint32_t b, uint32_t *qtot, uint32_t s0, uint32_t s1) {
  __syncthreads();  // every thread is done reading qtot of the previous task
  uint32_t n = 0;
  for (uint32_t s = s0; s < s1; s++)
    if (a.gflag[s] == a.sweep) n += a.runs[(size_t)(b + 1) * a.ns + s] - a.runs[(size_t)b * a.ns + s];
  n = read_lane_u32(wave_inclusive_scan(n), kWave - 1);
  if (lane_id() == 0) qtot[wave_id()] = n;
  __syncthreads();
  uint32_t tot = 0;
  for (uint32_t w = 0; w < kIterWaves; w++) tot += qtot[w];
  return tot;
}

__global__ __launch_bounds__(kIterBlock) void min_bin_kernel(StripArgs a) {
  extern __shared__ uint32_t lds2u[];
  __shared__ uint32_t wtot[kIterWaves], qtot[kIterWaves], live, shp[kIterWaves + 1];
  const uint32_t R = 1u << a.rlog, ns = a.ns;
  const BinLds<uint32_t> L = bin_lds<uint32_t>(lds2u, R, ns);
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (ns + kIterBlock - 1) / kIterBlock;
  const uint32_t s0 = min(tid * per, ns), s1 = min(s0 + per, ns);
  uint32_t changed = 0;
  unsigned long long ndeg = 0;
  for (uint32_t t = xcd_block_id(); t < a.nbt; t += gridDim.x) {
    const uint4 bt = a.btask[t];
    // a whole bin with no message of a gathered strip cannot change; a split bin's slice still
    // writes its partial (zeros), which the split apply reads
    if (bt.w == kNoSlot && bin_live_whole(a, bt.x, qtot, s0, s1) == 0) continue;
    const uint32_t n = bin_prepare<true>(a, bt, L, 0u, wtot, &live, s0, s1);
    if (n == 0 && bt.w == kNoSlot) continue;
    bin_messages(a, n, L, [&](uint32_t r, uint32_t x) { atomicMax(&L.acc[r], x); });
    __syncthreads();
    if (bt.w == kNoSlot && a.fmask) {
      // pull round: thread tid owns rows tid + k * kIterBlock (k < 16, bins of at most 2^14
      // rows); the rows that join the near frontier or the far pile leave with one atomic per
      // list and bin task
      const uint32_t v0 = bt.x << a.rlog, nr = min(R, a.V - v0);
      uint32_t nmask = 0, fmask = 0;
      for (uint32_t k0 = 0; k0 * kIterBlock < nr; k0 += kPullUnroll) {
        uint32_t m[kPullUnroll], old[kPullUnroll];
#pragma unroll
        for (int u = 0; u < kPullUnroll; u++) {
          const uint32_t i = (k0 + u) * kIterBlock + tid;
          m[u] = i < nr ? L.acc[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kPullUnroll; u++) old[u] = m[u] ? a.dist[v0 + (k0 + u) * kIterBlock + tid] : 0u;
#pragma unroll
        for (int u = 0; u < kPullUnroll; u++) {
          const uint32_t v = v0 + (k0 + u) * kIterBlock + tid;
          bool tn, tf;
          if (pull_apply(a, v, m[u], old[u], tn, tf)) changed++;
          nmask |= (tn ? 1u : 0u) << (k0 + u);
          fmask |= (tf ? 1u : 0u) << (k0 + u);
        }
      }
      uint32_t pn = block_append<kIterWaves>((uint32_t)__popc(nmask), a.near_cnt, shp);
      uint32_t pf = a.far_out ? block_append<kIterWaves>((uint32_t)__popc(fmask), a.far_cnt, shp) : 0u;
      for (uint32_t k = 0; nmask | fmask; k++, nmask >>= 1, fmask >>= 1) {
        const uint32_t v = v0 + k * kIterBlock + tid;
        if (nmask & 1u) {
          a.near_out[pn++] = v;
          ndeg += a.deg[v];
        }
        if (fmask & 1u) a.far_out[pf++] = v;
      }
    } else if (bt.w == kNoSlot) {
      // rows without a message hold 0 and cannot change: only the others read their value, kUnroll
      // loads in flight per thread
      const uint32_t v0 = bt.x << a.rlog, n = min(R, a.V - v0);
      for (uint32_t i0 = tid; i0 < n; i0 += kIterBlock * kUnroll) {
        uint32_t m[kUnroll], old[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          const uint32_t i = i0 + u * kIterBlock;
          m[u] = i < n ? L.acc[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) old[u] = m[u] ? a.nlab[v0 + i0 + u * kIterBlock] : ~0u;
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
          if (m[u] > old[u]) {
            const uint32_t v = v0 + i0 + u * kIterBlock;
            a.nlab[v] = m[u];
            a.sflag[v >> a.wlog] = a.sweep;
            if (a.vstamp) a.vstamp[v] = a.sweep;
            changed++;
          }
        }
      }
    } else {
      uint32_t *dst = reinterpret_cast<uint32_t *>(a.bpart) + ((size_t)bt.w << a.rlog);
      for (uint32_t i = tid; i < R; i += kIterBlock) dst[i] = L.acc[i];
    }
  }
  if (a.fmask) pull_count_block<kIterWaves>(a, changed, ndeg, shp);
  else min_count_changed(changed, a.changed);
}

__global__ __launch_bounds__(kBlock) void min_split_apply_kernel(StripArgs a) {
  const uint32_t R = 1u << a.rlog;
  const uint32_t *part = reinterpret_cast<const uint32_t *>(a.bpart);
  const uint32_t nck = (R + kSplitRows - 1) / kSplitRows;  // row chunks per split bin
  __shared__ uint32_t shp[kBlock / kWave + 1];
  uint32_t changed = 0;
  unsigned long long ndeg = 0;
  for (uint32_t t = blockIdx.x; t < a.nsplit * nck; t += gridDim.x) {
    const uint4 sp = a.split[t / nck];
    const uint32_t r0 = (t % nck) * kSplitRows;
    const uint32_t v0 = sp.x << a.rlog;
    const uint32_t n = min(min(R, a.V - v0), r0 + kSplitRows);
    if (a.fmask) {  // pull round: block-uniform steps, one atomic per list and step
      for (uint32_t i0 = r0; i0 < n; i0 += kBlock) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t m = 0;
        if (i < n)
          for (uint32_t k = 0; k < sp.z; k++) m = max(m, part[((size_t)(sp.y + k) << a.rlog) + i]);
        bool tn, tf;
        if (pull_apply(a, v0 + i, m, m ? a.dist[v0 + i] : 0u, tn, tf)) changed++;
        const uint32_t pn = block_append<kBlock / kWave>(tn ? 1u : 0u, a.near_cnt, shp);
        if (tn) {
          a.near_out[pn] = v0 + i;
          ndeg += a.deg[v0 + i];
        }
        if (a.far_out) {
          const uint32_t pf = block_append<kBlock / kWave>(tf ? 1u : 0u, a.far_cnt, shp);
          if (tf) a.far_out[pf] = v0 + i;
        }
      }
      continue;
    }
    for (uint32_t i = r0 + threadIdx.x; i < n; i += kBlock) {
      uint32_t m = 0;
      for (uint32_t k = 0; k < sp.z; k++) m = max(m, part[((size_t)(sp.y + k) << a.rlog) + i]);
      min_apply(a, v0 + i, m, changed);
    }
  }
  if (a.fmask) pull_count_block<kBlock / kWave>(a, changed, ndeg, shp);
  else min_count_changed(changed, a.changed);
}

// Starting values: CC c = ~(original id) and every strip active; SSSP c = 0 (unreached) but
// ~0 at the source, and only the source's strip active.  No strip has been gathered yet.
__global__ void min_init_kernel(uint32_t *c, const uint32_t *n2o, uint32_t n, uint32_t src, uint32_t *sflag,
                                uint32_t *gflag, uint32_t ns, uint32_t wlog) {
  const uint32_t stride = gridDim.x * blockDim.x, i0 = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t i = i0; i < n; i += stride) c[i] = src == kNever ? ~(n2o ? n2o[i] : i) : (i == src ? ~0u : 0u);
  for (uint32_t s = i0; s < ns; s += stride) {
    sflag[s] = (src == kNever || s == (src >> wlog)) ? 0u : kNever;
    gflag[s] = 0u;
  }
}

// Entering sweeps from a frontier: every strip idle, then the strips of the frontier's vertices
// (changed, not yet propagated) marked as changed in sweep `last`.
__global__ void sflag_fill_kernel(uint32_t *sflag, uint32_t n, uint32_t value) {
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) sflag[s] = value;
}

__global__ void sflag_mark_kernel(const uint32_t *front, uint32_t n, uint32_t wlog, uint32_t last, uint32_t *sflag) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    sflag[front[i] >> wlog] = last;
}

__global__ void weights_to_bytes_kernel(const uint32_t *w, uint64_t n, uint8_t *b) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = (uint8_t)w[i];
}

__global__ void complement_kernel(uint32_t *x, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = ~x[i];
}

int lds_limit(Engine &e) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, e.device) != hipSuccess || v <= 0)
    v = 65536;
  return v;
}

// The gather kernels address their dynamic LDS from 0, which holds while they declare no static
// LDS (the dynamic part follows the static part).
void check_gather_lds(const void *kernel) {
  hipFuncAttributes fa{};
  GMP_HIP(hipFuncGetAttributes(&fa, kernel));
  if (fa.sharedSizeBytes != 0) fail(GMP_EHIP, "strip gather kernel holds static LDS");
}

}  // namespace

uint32_t strip_partials(Engine &e) { return 4u * (uint32_t)std::max(1, e.num_cus); }

bool ensure_strips(Engine &e) {
  ensure_work(e);
  Graph &g = e.g;
  StripCsr &S = g.strips;
  const uint32_t wlog = e.opt.strip_log;
  if (wlog == 0 || g.E == 0) return false;
  if (S.valid) return true;
  const uint32_t V = g.V;
  const uint64_t E = g.E;
  const uint32_t b = bits_for(V);
  const uint32_t nsmax = (uint32_t)(((uint64_t)V + (1ull << wlog) - 1) >> wlog);
  const uint32_t sbits = bits_for(nsmax);
  const size_t ldsmax = (size_t)lds_limit(e);
  const size_t ldsmax2 = ldsmax - 512;  // the bin kernel's static LDS (scan and FP64 partials)
  const size_t mbuf = (size_t)kIterWaves * kRound * 4;  // per-wave message buffers
  if (sbits + b + wlog > 64 || (4ull << wlog) + mbuf > ldsmax) return false;

  S = StripCsr();
  const CsrView w = g.work_in();
  DevBuf rowid, kA, kB, tmp;
  edge_row_ids(e, w, V, E, g.off64, rowid);
  kA.alloc(E * 8);
  kB.alloc(E * 8);
  strip_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(w.idx, rowid.as<uint32_t>(), E, b, wlog, kA.as<uint64_t>());
  GMP_CHECK_LAUNCH();
  rowid.release();
  rocprim::double_buffer<uint64_t> dk(kA.as<uint64_t>(), kB.as<uint64_t>());
  size_t tb = 0;
  GMP_HIP(rocprim::radix_sort_keys(nullptr, tb, dk, E, 0, sbits + b + wlog, e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::radix_sort_keys(tmp.p, tb, dk, E, 0, sbits + b + wlog, e.stream));
  const uint64_t *keys = dk.current();
  uint64_t last = 0;
  GMP_HIP(hipMemcpyAsync(&last, keys + E - 1, 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  const uint32_t sh = b + wlog;
  const uint32_t ns = (uint32_t)(last >> sh) + 1;

  // bins: as many rows as the LDS holds next to the run table
  uint32_t rlog = std::min<uint32_t>(e.opt.bin_log, b);
  auto lds2 = [&](uint32_t r) {
    return ((size_t)8 << r) + (size_t)(2 * (size_t)ns + 1) * 4 + ((size_t)(1u << e.opt.bin_task_log) / kChunk + 2) * 4;
  };
  while (rlog > 4 && lds2(rlog) > ldsmax2) rlog--;
  if (lds2(rlog) > ldsmax2) return false;
  const uint32_t R = 1u << rlog;
  const uint32_t nbins = (uint32_t)(((uint64_t)V + R - 1) >> rlog);

  // strip edge ranges -> wave tasks and workgroup tasks
  DevBuf soff;
  soff.alloc(((size_t)ns + 1) * 8);
  strip_offsets_kernel<<<grid_for(E + 1), kBlock, 0, e.stream>>>(keys, E, sh, ns, soff.as<uint64_t>());
  GMP_CHECK_LAUNCH();
  std::vector<uint64_t> hso((size_t)ns + 1), hsb(std::max<uint32_t>(ns, 1), 0);
  GMP_HIP(hipMemcpyAsync(hso.data(), soff.p, hso.size() * 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  std::vector<uint4> wt, gt;
  const uint32_t gwt = 1u << e.opt.strip_task_log;  // wave tasks per workgroup task (one staging)
  for (uint32_t s = 0; s < ns; s++) {
    const uint64_t e0 = hso[s], e1 = hso[s + 1];
    // wave tasks count edges from the strip's base (a multiple of 8, for whole 16-byte code
    // loads), so a graph may have any number of edges while a strip stays under 2^32
    hsb[s] = e0 & ~7ull;
    if (e0 == e1) continue;
    if (e1 - hsb[s] >= 0xFFFFFFFFull - kCodePad) return false;
    const uint32_t w0 = (uint32_t)wt.size();
    // task boundaries inside a strip fall on multiples of 8, so only the strip's own ends can
    // cut a lane's 8 edges
    for (uint64_t x = e0; x < e1;) {
      const uint64_t nx = std::min<uint64_t>(((x & ~7ull) + kWaveTaskEdges), e1);
      wt.push_back(make_uint4((uint32_t)(x - hsb[s]), (uint32_t)(nx - hsb[s]), 0u, s));
      x = nx;
    }
    for (uint32_t k = w0; k < (uint32_t)wt.size(); k += gwt)
      gt.push_back(make_uint4(s, k, std::min<uint32_t>(k + gwt, (uint32_t)wt.size()), 0u));
  }
  // heaviest workgroup tasks first: the gather kernels' queue hands them out longest first
  std::stable_sort(gt.begin(), gt.end(), [&](const uint4 &p, const uint4 &q) {
    return wt[p.z - 1].y - wt[p.y].x > wt[q.z - 1].y - wt[q.y].x;
  });
  S.nwt = (uint32_t)wt.size();
  S.ngt = (uint32_t)gt.size();
  S.wtask.alloc((size_t)S.nwt * sizeof(uint4));
  S.gtask.alloc((size_t)S.ngt * sizeof(uint4));
  GMP_HIP(hipMemcpyAsync(S.wtask.p, wt.data(), (size_t)S.nwt * sizeof(uint4), hipMemcpyHostToDevice, e.stream));
  S.sbase.alloc(hsb.size() * 8);
  GMP_HIP(hipMemcpyAsync(S.sbase.p, hsb.data(), hsb.size() * 8, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipMemcpyAsync(S.gtask.p, gt.data(), (size_t)S.ngt * sizeof(uint4), hipMemcpyHostToDevice, e.stream));

  // codes and segment ids
  S.code.alloc((E + kCodePad) * 2);
  GMP_HIP(hipMemsetAsync(S.code.p, 0, (E + kCodePad) * 2, e.stream));
  codes_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(keys, E, wlog, S.code.as<uint16_t>());
  GMP_CHECK_LAUNCH();
  task_ends_kernel<<<grid_for(S.nwt), kBlock, 0, e.stream>>>(S.wtask.as<uint4>(), S.sbase.as<uint64_t>(), S.nwt,
                                                             S.code.as<uint16_t>());
  GMP_CHECK_LAUNCH();
  DevBuf flag, segid;
  flag.alloc(E * 4);
  segid.alloc(E * 4);
  end_flags_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(S.code.as<uint16_t>(), E, flag.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  {  // segment ids and message slots are 32-bit: count the segments in 64 bits first
    DevBuf total, rtmp;
    total.alloc(8);
    size_t rb = 0;
    GMP_HIP(rocprim::reduce(nullptr, rb, flag.as<uint32_t>(), total.as<uint64_t>(), (uint64_t)0, E,
                            rocprim::plus<uint64_t>(), e.stream));
    rtmp.alloc(rb);
    GMP_HIP(rocprim::reduce(rtmp.p, rb, flag.as<uint32_t>(), total.as<uint64_t>(), (uint64_t)0, E,
                            rocprim::plus<uint64_t>(), e.stream));
    uint64_t hseg = 0;
    GMP_HIP(hipMemcpyAsync(&hseg, total.p, 8, hipMemcpyDeviceToHost, e.stream));
    GMP_HIP(hipStreamSynchronize(e.stream));
    if (hseg >= 0xFFFFFFFFull) return false;
  }
  size_t tb2 = 0;
  GMP_HIP(rocprim::exclusive_scan(nullptr, tb2, flag.as<uint32_t>(), segid.as<uint32_t>(), 0u, E,
                                  rocprim::plus<uint32_t>(), e.stream));
  DevBuf tmp2;
  tmp2.alloc(tb2);
  GMP_HIP(rocprim::exclusive_scan(tmp2.p, tb2, flag.as<uint32_t>(), segid.as<uint32_t>(), 0u, E,
                                  rocprim::plus<uint32_t>(), e.stream));
  uint32_t lastseg = 0;
  GMP_HIP(hipMemcpyAsync(&lastseg, segid.as<uint32_t>() + E - 1, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  flag.release();
  tmp2.release();
  const uint32_t nseg = lastseg + 1;  // the last edge always ends a segment
  S.nseg = nseg;
  task_seg_kernel<<<grid_for(S.nwt), kBlock, 0, e.stream>>>(S.wtask.as<uint4>(), S.sbase.as<uint64_t>(), S.nwt,
                                                            segid.as<uint32_t>());
  GMP_CHECK_LAUNCH();

  // segment destinations and the (bin, strip) run table
  DevBuf segfull, sseg;
  S.seg_row.alloc((size_t)nseg * 2);
  segfull.alloc((size_t)nseg * 4);
  seg_rows_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(keys, S.code.as<uint16_t>(), segid.as<uint32_t>(), E, wlog,
                                                        b, R - 1, S.seg_row.as<uint16_t>(), segfull.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  sseg.alloc(((size_t)ns + 1) * 4);
  strip_seg_kernel<<<grid_for((uint64_t)ns + 1), kBlock, 0, e.stream>>>(soff.as<uint64_t>(), segid.as<uint32_t>(), ns,
                                                                         E, nseg, sseg.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  // the end bits leave the codes for a byte per 8 edges (the padding's bytes stay zero)
  const uint64_t ngrp = (E + kCodePad) / 8;
  S.endb.alloc(ngrp);
  split_ends_kernel<<<grid_for(ngrp), kBlock, 0, e.stream>>>(S.code.as<uint16_t>(), ngrp, S.endb.as<uint8_t>());
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipStreamSynchronize(e.stream));
  kA.release();
  kB.release();
  tmp.release();
  segid.release();
  S.runs.alloc(((size_t)nbins + 1) * ns * 4);
  runs_kernel<<<grid_for((uint64_t)(nbins + 1) * ns), kBlock, 0, e.stream>>>(segfull.as<uint32_t>(), sseg.as<uint32_t>(),
                                                                            ns, nbins, rlog, S.runs.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  DevBuf bcnt;
  bcnt.alloc((size_t)nbins * 4);
  bin_count_kernel<<<std::max<uint32_t>(1, std::min<uint32_t>((nbins + 3) / 4, 65536)), kBlock, 0, e.stream>>>(
      S.runs.as<uint32_t>(), ns, nbins, bcnt.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  std::vector<uint32_t> hc(nbins);
  GMP_HIP(hipMemcpyAsync(hc.data(), bcnt.p, (size_t)nbins * 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));

  // bin tasks: heavier bins are split into slices that leave fixed-point partials
  const uint32_t kBinTaskMsgs = 1u << e.opt.bin_task_log;
  std::vector<uint4> bt, sp;
  uint32_t nslots = 0;
  for (uint32_t bb = 0; bb < nbins; bb++) {
    const uint32_t m = hc[bb];
    if (m <= kBinTaskMsgs) {
      bt.push_back(make_uint4(bb, 0u, m, kNoSlot));
      continue;
    }
    const uint32_t k = (m + kBinTaskMsgs - 1) / kBinTaskMsgs;
    for (uint32_t i = 0; i < k; i++)
      bt.push_back(make_uint4(bb, i * kBinTaskMsgs, std::min<uint32_t>(m, (i + 1) * kBinTaskMsgs), nslots + i));
    sp.push_back(make_uint4(bb, nslots, k, 0u));
    nslots += k;
  }
  std::stable_sort(bt.begin(), bt.end(), [](const uint4 &p, const uint4 &q) { return p.z - p.y > q.z - q.y; });
  S.nbt = (uint32_t)bt.size();
  S.nsplit = (uint32_t)sp.size();
  S.nslots = nslots;
  S.btask.alloc((size_t)S.nbt * sizeof(uint4));
  GMP_HIP(hipMemcpyAsync(S.btask.p, bt.data(), (size_t)S.nbt * sizeof(uint4), hipMemcpyHostToDevice, e.stream));
  if (S.nsplit) {
    S.split.alloc((size_t)S.nsplit * sizeof(uint4));
    GMP_HIP(hipMemcpyAsync(S.split.p, sp.data(), (size_t)S.nsplit * sizeof(uint4), hipMemcpyHostToDevice, e.stream));
    S.bpart.alloc(((size_t)nslots << rlog) * 8);
  }
  S.msg.alloc((size_t)nseg * 4);
  GMP_HIP(hipStreamSynchronize(e.stream));

  S.wlog = wlog;
  S.rlog = rlog;
  S.ns = ns;
  S.nbins = nbins;
  S.lds1 = (uint32_t)((4ull << wlog) + mbuf);
  S.lds2 = (uint32_t)lds2(rlog);
  const uint32_t cus = (uint32_t)std::max(1, e.num_cus);
  const uint32_t per1 = (uint32_t)std::min<size_t>(2, std::max<size_t>(1, ldsmax / S.lds1));
  const uint32_t per2 = (uint32_t)std::min<size_t>(2, std::max<size_t>(1, ldsmax / (S.lds2 + 512)));
  S.g1 = std::max<uint32_t>(1, std::min<uint32_t>(S.ngt, cus * per1));
  S.gq.alloc((size_t)(2 + S.g1) * 4);
  GMP_HIP(hipMemsetAsync(S.gq.p, 0, (size_t)(2 + S.g1) * 4, e.stream));
  S.g2 = std::max<uint32_t>(1, std::min<uint32_t>(S.nbt, cus * per2));
  S.g3 = std::max<uint32_t>(1, std::min<uint32_t>(S.nsplit * ((R + kSplitRows - 1) / kSplitRows), cus * 4));
  // raise the dynamic-LDS ceiling where the runtime keeps one; a refusal leaves the launch
  // itself to report a size the device cannot hold
  check_gather_lds(reinterpret_cast<const void *>(&pr_strip_gather_kernel));
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&pr_strip_gather_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)S.lds1) != hipSuccess)
    (void)hipGetLastError();
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&pr_bin_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)S.lds2) != hipSuccess)
    (void)hipGetLastError();
  S.valid = true;
  return true;
}

void strip_iteration(Engine &e, const PrStripIO &io, uint32_t &n1, uint32_t &n2) {
  const StripCsr &S = e.g.strips;
  if (!S.valid) fail(GMP_ESTATE, "strip layout not built");
  StripArgs a{};
  a.outdeg = io.outdeg;
  a.contrib_in = io.contrib_in;
  a.rank_in = io.rank_in;
  a.rank_out = io.rank_out;
  a.contrib_out = io.contrib_out;
  a.sc = io.sc;
  a.d = io.d;
  a.seed = io.seed;
  a.code = S.code.as<uint16_t>();
  a.endb = S.endb.as<uint8_t>();
  a.wtask = S.wtask.as<uint4>();
  a.sbase = S.sbase.as<uint64_t>();
  a.gtask = S.gtask.as<uint4>();
  a.gq = S.gq.as<uint32_t>();
  a.ngt = S.ngt;
  a.msg = S.msg.as<float>();
  a.seg_row = S.seg_row.as<uint16_t>();
  a.runs = S.runs.as<uint32_t>();
  a.ns = S.ns;
  a.btask = S.btask.as<uint4>();
  a.nbt = S.nbt;
  a.bpart = S.bpart.as<unsigned long long>();
  a.split = S.split.as<uint4>();
  a.nsplit = S.nsplit;
  a.V = e.g.V;
  a.wlog = S.wlog;
  a.rlog = S.rlog;
  a.pr1 = io.pr1;
  a.pd1 = io.pd1;
  a.pr2 = io.pr2;
  a.pd2 = io.pd2;
  if (io.packed) {
    if (!S.pk_valid) fail(GMP_ESTATE, "packed strip codes not built");
    a.pk = S.pk.as<uint32_t>();
    a.pkhdr = S.pkhdr.as<uint32_t>();
    a.pkoff = S.pkoff.as<uint32_t>();
    a.pkblk = S.pkblk.as<uint32_t>();
    pr_strip_gather_packed_kernel<<<S.g1, kIterBlock, S.lds1, e.stream>>>(a);
  } else if (io.ids) {
    if (!S.ids_valid) fail(GMP_ESTATE, "strip source ids not built");
    a.ids = S.ids.as<uint32_t>();
    pr_strip_gather_ids_kernel<<<S.g1, kIterBlock, S.lds1, e.stream>>>(a);
  } else {
    pr_strip_gather_kernel<<<S.g1, kIterBlock, S.lds1, e.stream>>>(a);
  }
  GMP_CHECK_LAUNCH();
  pr_bin_kernel<<<S.g2, kIterBlock, S.lds2, e.stream>>>(a);
  GMP_CHECK_LAUNCH();
  n1 = S.g2;
  n2 = 0;
  if (S.nsplit) {
    pr_split_apply_kernel<<<S.g3, kBlock, 0, e.stream>>>(a);
    GMP_CHECK_LAUNCH();
    n2 = S.g3;
  }
}

bool ensure_strip_pack(Engine &e) {
  if (!ensure_strips(e)) return false;
  StripCsr &S = e.g.strips;
  if (S.pk_valid) return true;
  const uint64_t E = e.g.E;
  // blocks: each strip's rounds from its base, up to its last wave task's end
  std::vector<uint4> wt(S.nwt);
  std::vector<uint64_t> sb(S.ns);
  GMP_HIP(hipMemcpyAsync(wt.data(), S.wtask.p, (size_t)S.nwt * sizeof(uint4), hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipMemcpyAsync(sb.data(), S.sbase.p, (size_t)S.ns * 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  std::vector<uint32_t> send(S.ns, 0u);
  for (const uint4 &t : wt) {
    // the gather finds a task's first block at its first edge / kRound (rounded down) and holds
    // the widths of at most kWaveTaskEdges / kRound rounds
    if ((t.x & ~7u) % kRound || t.y - (t.x & ~7u) > kWaveTaskEdges) return false;
    send[t.w] = std::max(send[t.w], t.y);
  }
  std::vector<uint32_t> pkblk(S.ns);
  std::vector<uint64_t> pos;
  for (uint32_t s = 0; s < S.ns; s++) {
    pkblk[s] = (uint32_t)pos.size();
    const uint32_t nb = (send[s] + kRound - 1) / kRound;
    for (uint32_t r = 0; r < nb; r++) pos.push_back(sb[s] + (uint64_t)r * kRound);
  }
  if (pos.empty() || pos.size() >= 0xFFFFFFFFull || pos.back() + kRound > E + kCodePad) return false;
  const uint32_t nblk = (uint32_t)pos.size();
  DevBuf dpos, words, off64, tmp, bad;
  dpos.alloc((size_t)nblk * 8);
  words.alloc((size_t)nblk * 4);
  bad.alloc(4);
  GMP_HIP(hipMemsetAsync(bad.p, 0, 4, e.stream));
  off64.alloc((size_t)nblk * 8);
  GMP_HIP(hipMemcpyAsync(dpos.p, pos.data(), (size_t)nblk * 8, hipMemcpyHostToDevice, e.stream));
  const uint32_t wg = std::max<uint32_t>(1, std::min<uint32_t>((nblk + 3) / 4, 1u << 16));
  pk_size_kernel<<<wg, kBlock, 0, e.stream>>>(S.code.as<uint16_t>(), S.endb.as<uint8_t>(), dpos.as<uint64_t>(), nblk,
                                             words.as<uint32_t>(), bad.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  size_t tb = 0;
  GMP_HIP(rocprim::exclusive_scan(nullptr, tb, words.as<uint32_t>(), off64.as<uint64_t>(), (uint64_t)0, nblk,
                                  rocprim::plus<uint64_t>(), e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::exclusive_scan(tmp.p, tb, words.as<uint32_t>(), off64.as<uint64_t>(), (uint64_t)0, nblk,
                                  rocprim::plus<uint64_t>(), e.stream));
  uint64_t lo = 0;
  uint32_t lw = 0, hbad = 0;
  GMP_HIP(hipMemcpyAsync(&lo, off64.as<uint64_t>() + nblk - 1, 8, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipMemcpyAsync(&lw, words.as<uint32_t>() + nblk - 1, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (hbad) return false;  // the differences would not be codes
  const uint64_t total = lo + lw;
  constexpr uint32_t kPkPad = 8;  // a group's five-word read from the last block stays in the stream
  if (total + kPkPad >= 0xFFFFFFFFull) return false;
  S.pk.alloc((total + kPkPad) * 4);
  S.pkoff.alloc((size_t)nblk * 4);
  S.pkhdr.alloc((size_t)nblk * 32);
  S.pkblk.alloc((size_t)S.ns * 4);
  GMP_HIP(hipMemsetAsync(S.pk.as<uint32_t>() + total, 0, kPkPad * 4, e.stream));
  narrow_offsets_kernel<<<grid_for(nblk), kBlock, 0, e.stream>>>(off64.as<uint64_t>(), nblk, S.pkoff.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipMemcpyAsync(S.pkblk.p, pkblk.data(), (size_t)S.ns * 4, hipMemcpyHostToDevice, e.stream));
  pk_write_kernel<<<wg, kBlock, 0, e.stream>>>(S.code.as<uint16_t>(), S.endb.as<uint8_t>(), dpos.as<uint64_t>(),
                                              S.pkoff.as<uint32_t>(), nblk, S.pk.as<uint32_t>(),
                                              S.pkhdr.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipStreamSynchronize(e.stream));
  check_gather_lds(reinterpret_cast<const void *>(&pr_strip_gather_packed_kernel));
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&pr_strip_gather_packed_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)S.lds1) != hipSuccess)
    (void)hipGetLastError();
  S.pk_words = total;
  S.pk_nblk = nblk;
  S.pk_valid = true;
  return true;
}

bool ensure_strip_ids(Engine &e) {
  if (!ensure_strips(e)) return false;
  StripCsr &S = e.g.strips;
  if (S.ids_valid) return true;
  const uint64_t E = e.g.E;
  // kCodePad entries past the end keep the last round's 16-byte loads in bounds
  S.ids.alloc((E + kCodePad) * 4);
  GMP_HIP(hipMemsetAsync(S.ids.p, 0, (E + kCodePad) * 4, e.stream));
  strip_ids_kernel<<<std::min<uint32_t>(S.nwt, 1u << 20), kBlock, 0, e.stream>>>(
      S.wtask.as<uint4>(), S.sbase.as<uint64_t>(), S.nwt, S.wlog, S.code.as<uint16_t>(), S.ids.as<uint32_t>());
  GMP_CHECK_LAUNCH();
  GMP_HIP(hipStreamSynchronize(e.stream));
  check_gather_lds(reinterpret_cast<const void *>(&pr_strip_gather_ids_kernel));
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&pr_strip_gather_ids_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)S.lds1) != hipSuccess)
    (void)hipGetLastError();
  S.ids_valid = true;
  return true;
}

namespace {

// Weights of the work in-CSR in layout order (the same key sort as the codes, weights as
// payload); kCodePad zero entries past the end keep the 16-byte round loads in bounds.
void ensure_strip_weights(Engine &e) {
  Graph &g = e.g;
  StripCsr &S = g.strips;
  if (S.wt_valid) return;
  const CsrView w = g.work_in();
  if (!w.w) fail(GMP_ESTATE, "graph has no weights");
  const uint32_t V = g.V;
  const uint64_t E = g.E;
  const uint32_t b = bits_for(V), wlog = S.wlog;
  const uint32_t sbits = bits_for(((uint64_t)V + (1ull << wlog) - 1) >> wlog);
  DevBuf rowid, kA, kB, vB, tmp;
  edge_row_ids(e, w, V, E, g.off64, rowid);
  kA.alloc(E * 8);
  kB.alloc(E * 8);
  vB.alloc(E * 4);
  strip_keys_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(w.idx, rowid.as<uint32_t>(), E, b, wlog, kA.as<uint64_t>());
  GMP_CHECK_LAUNCH();
  rowid.release();
  S.wt.alloc((E + kCodePad) * 4);
  GMP_HIP(hipMemcpyAsync(S.wt.p, w.w, E * 4, hipMemcpyDeviceToDevice, e.stream));
  GMP_HIP(hipMemsetAsync(S.wt.as<uint32_t>() + E, 0, kCodePad * 4, e.stream));
  rocprim::double_buffer<uint64_t> dk(kA.as<uint64_t>(), kB.as<uint64_t>());
  rocprim::double_buffer<uint32_t> dv(S.wt.as<uint32_t>(), vB.as<uint32_t>());
  size_t tb = 0;
  GMP_HIP(rocprim::radix_sort_pairs(nullptr, tb, dk, dv, E, 0, sbits + b + wlog, e.stream));
  tmp.alloc(tb);
  GMP_HIP(rocprim::radix_sort_pairs(tmp.p, tb, dk, dv, E, 0, sbits + b + wlog, e.stream));
  if (dv.current() != S.wt.as<uint32_t>())
    GMP_HIP(hipMemcpyAsync(S.wt.p, dv.current(), E * 4, hipMemcpyDeviceToDevice, e.stream));
  // one byte per weight when every weight fits (R-MAT weights are 1..255)
  DevBuf mx;
  mx.alloc(4);
  size_t tr = 0;
  GMP_HIP(rocprim::reduce(nullptr, tr, S.wt.as<uint32_t>(), mx.as<uint32_t>(), 0u, E, rocprim::maximum<uint32_t>(),
                          e.stream));
  tmp.alloc(tr);
  GMP_HIP(rocprim::reduce(tmp.p, tr, S.wt.as<uint32_t>(), mx.as<uint32_t>(), 0u, E, rocprim::maximum<uint32_t>(),
                          e.stream));
  uint32_t wmax = 0;
  GMP_HIP(hipMemcpyAsync(&wmax, mx.p, 4, hipMemcpyDeviceToHost, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  if (wmax <= 0xFFu) {
    S.wt8.alloc(E + kCodePad);
    GMP_HIP(hipMemsetAsync(S.wt8.as<uint8_t>() + E, 0, kCodePad, e.stream));
    weights_to_bytes_kernel<<<grid_for(E), kBlock, 0, e.stream>>>(S.wt.as<uint32_t>(), E, S.wt8.as<uint8_t>());
    GMP_CHECK_LAUNCH();
    GMP_HIP(hipStreamSynchronize(e.stream));
    S.wt.release();
  }
  S.wt_valid = true;
}

// Synchronous sweeps of one min-type rule over the strip layout until nothing changes; c is
// the complemented result in work order.
}  // namespace

struct SweepRun {
  Engine &e;
  int mode = kSweepCC;
  StripArgs a{};
  DevBuf changed, sflag, gflag, nedges;
  uint32_t nsall = 0, gi = 1;
  size_t lds2 = 0;
  uint32_t last = 0;        // number of the last sweep run
  explicit SweepRun(Engine &en) : e(en) {}
};

void SweepRunDeleter::operator()(SweepRun *s) const { delete s; }

SweepRunPtr sweep_open(Engine &e, int kind, uint32_t *c, uint32_t *vstamp) {
  if (!ensure_strips(e)) return nullptr;
  int mode = kind == 0 ? kSweepCC : kind == 1 ? kSweepUnit : kSweepWeighted;
  if (mode == kSweepWeighted) {
    if (!e.g.work_in().w) return nullptr;
    ensure_strip_weights(e);
    if (e.g.strips.wt8.p) mode = kSweepWeighted8;
  }
  SweepRunPtr sp(new SweepRun(e));
  SweepRun &r = *sp;
  r.mode = mode;
  const StripCsr &S = e.g.strips;
  const uint32_t V = e.g.V, R = 1u << S.rlog;
  r.lds2 = (size_t)4 * R + (2 * (size_t)S.ns + 1) * 4 + ((size_t)(1u << e.opt.bin_task_log) / kChunk + 2) * 4;
  const void *gk = mode == kSweepCC          ? reinterpret_cast<const void *>(&min_strip_gather_kernel<kSweepCC>)
                   : mode == kSweepUnit      ? reinterpret_cast<const void *>(&min_strip_gather_kernel<kSweepUnit>)
                   : mode == kSweepWeighted8 ? reinterpret_cast<const void *>(&min_strip_gather_kernel<kSweepWeighted8>)
                                             : reinterpret_cast<const void *>(&min_strip_gather_kernel<kSweepWeighted>);
  check_gather_lds(gk);
  if (hipFuncSetAttribute(gk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)S.lds1) != hipSuccess)
    (void)hipGetLastError();
  if (hipFuncSetAttribute(reinterpret_cast<const void *>(&min_bin_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)r.lds2) != hipSuccess)
    (void)hipGetLastError();
  r.changed.alloc(4);
  r.nedges.alloc(8);
  GMP_HIP(hipMemsetAsync(r.nedges.p, 0, 8, e.stream));
  // a flag for every strip of the id range: vertices past the last strip holding an edge (no
  // out-edges in the degree order) still change and mark their strip
  r.nsall = (uint32_t)(((uint64_t)V + (1ull << S.wlog) - 1) >> S.wlog);
  r.sflag.alloc((size_t)std::max<uint32_t>(r.nsall, 1) * 4);
  r.gflag.alloc((size_t)std::max<uint32_t>(r.nsall, 1) * 4);
  StripArgs &a = r.a;
  a.code = S.code.as<uint16_t>();
  a.endb = S.endb.as<uint8_t>();
  a.wtask = S.wtask.as<uint4>();
  a.sbase = S.sbase.as<uint64_t>();
  a.gtask = S.gtask.as<uint4>();
  a.gq = S.gq.as<uint32_t>();
  a.ngt = S.ngt;
  a.msg = S.msg.as<float>();
  a.seg_row = S.seg_row.as<uint16_t>();
  a.runs = S.runs.as<uint32_t>();
  a.ns = S.ns;
  a.btask = S.btask.as<uint4>();
  a.nbt = S.nbt;
  a.bpart = S.bpart.as<unsigned long long>();
  a.split = S.split.as<uint4>();
  a.nsplit = S.nsplit;
  a.V = V;
  a.wlog = S.wlog;
  a.rlog = S.rlog;
  a.nlab = c;
  a.changed = r.changed.as<uint32_t>();
  a.wt = S.wt.as<uint32_t>();
  a.wt8 = S.wt8.as<uint8_t>();
  a.sflag = r.sflag.as<uint32_t>();
  a.vstamp = vstamp;
  a.gflag = r.gflag.as<uint32_t>();
  a.edges = r.nedges.as<unsigned long long>();
  r.gi = (uint32_t)std::min<uint64_t>(((uint64_t)V + kBlock - 1) / kBlock, (uint64_t)e.num_cus * 8);
  return sp;
}

void sweep_start(SweepRun &r, uint32_t src) {
  const StripCsr &S = r.e.g.strips;
  const uint32_t *n2o = r.e.g.work_identity ? nullptr : r.e.g.n2o.as<uint32_t>();  // built with the layout
  min_init_kernel<<<r.gi, kBlock, 0, r.e.stream>>>(r.a.nlab, n2o, r.a.V, r.mode == kSweepCC ? kNever : src, r.a.sflag,
                                                   r.a.gflag, r.nsall, S.wlog);
  GMP_CHECK_LAUNCH();
  r.last = 0;
}

void sweep_enter(SweepRun &r, const uint32_t *front, uint32_t n) {
  const StripCsr &S = r.e.g.strips;
  sflag_fill_kernel<<<grid_for(r.nsall), kBlock, 0, r.e.stream>>>(r.a.sflag, r.nsall, kNever);
  GMP_CHECK_LAUNCH();
  if (r.last == 0) {  // no sweep ran yet in this solve
    sflag_fill_kernel<<<grid_for(r.nsall), kBlock, 0, r.e.stream>>>(r.a.gflag, r.nsall, 0u);
    GMP_CHECK_LAUNCH();
  }
  if (n) {
    sflag_mark_kernel<<<grid_for(n), kBlock, 0, r.e.stream>>>(front, n, S.wlog, r.last, r.a.sflag);
    GMP_CHECK_LAUNCH();
  }
}

uint32_t sweep_step(SweepRun &r) {
  const StripCsr &S = r.e.g.strips;
  StripArgs &a = r.a;
  hipStream_t st = r.e.stream;
  a.sweep = ++r.last;
  GMP_HIP(hipMemsetAsync(r.changed.p, 0, 4, st));
  // without selective scheduling every strip counts as changed in the sweep before
  if (!r.e.opt.selective) sflag_fill_kernel<<<grid_for(r.nsall), kBlock, 0, st>>>(a.sflag, r.nsall, a.sweep - 1u);
  if (r.mode == kSweepCC) min_strip_gather_kernel<kSweepCC><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  else if (r.mode == kSweepUnit) min_strip_gather_kernel<kSweepUnit><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  else if (r.mode == kSweepWeighted8) min_strip_gather_kernel<kSweepWeighted8><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  else min_strip_gather_kernel<kSweepWeighted><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  GMP_CHECK_LAUNCH();
  min_bin_kernel<<<S.g2, kIterBlock, r.lds2, st>>>(a);
  GMP_CHECK_LAUNCH();
  if (S.nsplit) {
    min_split_apply_kernel<<<S.g3, kBlock, 0, st>>>(a);
    GMP_CHECK_LAUNCH();
  }
  uint32_t *h = static_cast<uint32_t *>(r.e.h_pinned);
  GMP_HIP(hipMemcpyAsync(h, r.changed.p, 4, hipMemcpyDeviceToHost, st));
  GMP_HIP(hipStreamSynchronize(st));
  return *h;
}

uint32_t sweep_last(const SweepRun &r) { return r.last; }

SweepRunPtr pull_open(Engine &e, bool unit) {
  return sweep_open(e, unit ? 1 : 2, nullptr, nullptr);
}

void pull_round(SweepRun &r, const PullRound &p) {
  const StripCsr &S = r.e.g.strips;
  StripArgs &a = r.a;
  hipStream_t st = r.e.stream;
  a.sweep = ++r.last;
  if (r.last == 1) {
    sflag_fill_kernel<<<grid_for(r.nsall), kBlock, 0, st>>>(a.gflag, r.nsall, 0u);
    GMP_CHECK_LAUNCH();
  }
  // only the strips holding a frontier vertex are gathered (every strip without selective
  // scheduling: the others offer nothing, so the round relaxes the same edges)
  sflag_fill_kernel<<<grid_for(r.nsall), kBlock, 0, st>>>(a.sflag, r.nsall, r.e.opt.selective ? kNever : r.last - 1);
  GMP_CHECK_LAUNCH();
  if (p.nfront && r.e.opt.selective) {
    sflag_mark_kernel<<<grid_for(p.nfront), kBlock, 0, st>>>(p.front, p.nfront, S.wlog, r.last - 1, a.sflag);
    GMP_CHECK_LAUNCH();
  }
  a.fmask = p.near_stamp;
  a.fstamp = p.fstamp;
  a.next_stamp = p.next_stamp;
  a.thr = p.thr;
  a.dist = p.dist;
  a.near_stamp = p.near_stamp;
  a.near_out = p.near_out;
  a.near_cnt = p.near_cnt;
  a.far_flag = p.far_flag;
  a.far_out = p.far_out;
  a.far_cnt = p.far_cnt;
  a.deg = p.deg;
  a.near_deg = p.near_deg;
  a.changed = p.changed ? p.changed : r.changed.as<uint32_t>();
  a.nlab = nullptr;
  a.vstamp = nullptr;
  if (r.mode == kSweepUnit) min_strip_gather_kernel<kSweepUnit><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  else if (r.mode == kSweepWeighted8) min_strip_gather_kernel<kSweepWeighted8><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  else min_strip_gather_kernel<kSweepWeighted><<<S.g1, kIterBlock, S.lds1, st>>>(a);
  GMP_CHECK_LAUNCH();
  min_bin_kernel<<<S.g2, kIterBlock, r.lds2, st>>>(a);
  GMP_CHECK_LAUNCH();
  if (S.nsplit) {
    min_split_apply_kernel<<<S.g3, kBlock, 0, st>>>(a);
    GMP_CHECK_LAUNCH();
  }
}

uint64_t sweep_edges(SweepRun &r) {
  uint64_t *h = static_cast<uint64_t *>(r.e.h_pinned);
  GMP_HIP(hipMemcpyAsync(h, r.nedges.p, 8, hipMemcpyDeviceToHost, r.e.stream));
  GMP_HIP(hipStreamSynchronize(r.e.stream));
  return *h;
}

void complement_values(Engine &e, uint32_t *x, uint32_t n) {
  const uint32_t gi = (uint32_t)std::min<uint64_t>(((uint64_t)n + kBlock - 1) / kBlock, (uint64_t)e.num_cus * 8);
  complement_kernel<<<std::max<uint32_t>(gi, 1), kBlock, 0, e.stream>>>(x, n);
  GMP_CHECK_LAUNCH();
}

}  // namespace gmp
