// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gmp.h"

namespace gmp {

// Error carried from anywhere inside the engine up to the C ABI, where it becomes a
// status code plus the thread's last-error text.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define GMP_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      ::gmp::fail(e_ == hipErrorOutOfMemory ? GMP_ENOMEM : GMP_EHIP,                    \
                  std::string(#call) + ": " + hipGetErrorString(e_));                   \
  } while (0)

#define GMP_CHECK_LAUNCH() GMP_HIP(hipGetLastError())

// Owning device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    if (n <= bytes && p) return;
    release();
    if (n == 0) return;
    GMP_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// PageRank gather plan: wave-sized tasks over the in-CSR.
//   STREAM task  {row0, nrows}: rows [row0,row0+nrows) whose edges (<= stream_cap in total)
//                are staged through LDS and reduced one row per lane.
//   CHUNK task   {hub, k}:      edges [k*chunk, (k+1)*chunk) of hub row hubs[hub], reduced by
//                the whole wave; rows with one chunk are applied in place, longer rows
//                leave per-chunk partials for the hub-combine kernel.
constexpr uint32_t kTaskChunkBit = 0x80000000u;
struct HubRow {
  uint32_t row;
  uint32_t nchunks;
  uint32_t part_base;
  uint32_t pad;
};
struct GatherPlan {
  uint32_t stream_cap = 1024;  // edges per STREAM task (LDS floats per wave)
  uint32_t row_cap = 512;      // rows per STREAM task
  uint32_t chunk = 2048;       // edges per CHUNK task
  uint32_t vec_min = 64;       // rows with more edges are reduced by a whole wave
  uint32_t ntasks = 0, nhubs = 0, nmulti = 0, nparts = 0;
  DevBuf tasks;   // uint2[ntasks]
  DevBuf hubs;    // HubRow[nhubs]
  DevBuf multi;   // uint32[nmulti]: hub ids with nchunks > 1
  bool built = false;
};

// One CSR in device memory; off is uint32[V+1] or uint64[V+1] (Graph::off64).
struct Csr {
  DevBuf off, idx, w;
  bool valid = false;
};

// Non-owning view handed to the kernel launchers.
struct CsrView {
  const void *off = nullptr;
  const uint32_t *idx = nullptr;
  const uint32_t *w = nullptr;
  template <class OffT> const OffT *offs() const { return static_cast<const OffT *>(off); }
};

// Source-strip layout of the work in-CSR (strip.hip): edges sorted by (strip of the source,
// destination, source) where a strip is 2^wlog consecutive work ids.  A segment is the run of
// edges of one destination inside one strip (also cut at wave-task boundaries); the gather
// stages one strip's values in LDS and leaves one message per segment, and bins of 2^rlog
// destinations combine their messages.  PageRank sums; CC and SSSP sweeps take minima.
struct StripCsr {
  uint32_t wlog = 0, rlog = 0;  // strips of 2^wlog sources, bins of 2^rlog destinations
  uint32_t ns = 0, nbins = 0;   // strips up to the last one holding an edge; destination bins
  uint64_t nseg = 0;            // segments = messages per iteration
  DevBuf code;                  // uint16[E+pad]: source offset in its strip
  DevBuf endb;                  // uint8[(E+pad)/8]: bit k of byte i ends the segment of edge 8 i + k
  DevBuf sbase;                 // uint64[ns]: the strip's first edge rounded down to a multiple of 8
  DevBuf wtask;                 // uint32x4[nwt]: {first edge, end edge (both from the strip's sbase),
                                //   first segment, strip}
  DevBuf gtask;                 // uint32x4[ngt]: {strip, first wave task, end wave task, 0}
  DevBuf seg_row;               // uint16[nseg]: destination offset inside its bin
  DevBuf runs;                  // uint32[(nbins+1)*ns]: [b*ns+s] = first segment of strip s in bin >= b
  DevBuf btask;                 // uint32x4[nbt]: {bin, first message, end message, partial slot or ~0}
  DevBuf split;                 // uint32x4[nsplit]: {bin, first slot, slots, 0}
  DevBuf msg;                   // float[nseg]: the iteration's segment sums
  DevBuf bpart;                 // uint64[nslots << rlog]: fixed-point sums of split bins
  DevBuf wt;                    // uint32[E+pad]: edge weights in layout order (weighted SSSP sweeps)
  DevBuf wt8;                   // uint8[E+pad]: the same as bytes, in place of wt when all fit
  DevBuf gq;                    // uint32[2 + g1]: next workgroup task, ticket, each workgroup's task
  // packed codes (GMP_PR_COMPRESSED on this layout): one block per 512-edge gather round of each
  // strip -- 8 words of per-lane widths, and the lanes' groups of 8 codes bit-packed
  DevBuf pk;                    // uint32 stream of the groups, padded
  DevBuf pkhdr;                 // uint32[8 blocks]: the widths
  DevBuf pkoff;                 // uint32[blocks]: first word of each block
  DevBuf pkblk;                 // uint32[ns]: first block of each strip
  DevBuf ids;                   // uint32[E+pad]: each edge's source work id in layout order (the
                                //   index as the in-CSR stores it: GMP_PR_STRIP_IDS)
  bool ids_valid = false;
  uint64_t pk_words = 0;        // words of the stream (padding not counted)
  uint32_t pk_nblk = 0;
  bool pk_valid = false;
  bool wt_valid = false;
  uint32_t nwt = 0, ngt = 0, nbt = 0, nsplit = 0, nslots = 0;
  uint32_t lds1 = 0, lds2 = 0;  // dynamic LDS bytes of the gather and bin kernels
  uint32_t g1 = 0, g2 = 0, g3 = 0;  // grids of the gather, bin and split-apply kernels
  bool valid = false;
};

// Vertices of the work out-CSR whose out-degree exceeds a threshold, with their edge lists cut
// into fixed-size slices, so frontier relaxation spreads a hub over many waves (traverse.hip).
struct HubTable {
  DevBuf ids;     // uint32[H]: hub work ids, ascending
  DevBuf slices;  // uint2[nslices]: {hub index, slice number}
  DevBuf flag;    // uint32[H]: stamp of the relax launch whose frontier holds the hub
  DevBuf list;    // uint32[nslices]: slices flagged in the current launch
  uint32_t H = 0, nslices = 0;
  bool valid = false;
};

// Device view of a varint-packed column stream (format in varint.hpp).
struct Packed {
  const uint8_t *bytes;   // padded by >= 16 bytes
  const uint64_t *gbyte;  // G+1 byte offsets
  const uint64_t *gmask;  // G restart masks
};

// Delta + varint column ids of the work in-CSR (decoded inside the gather).
struct PackedCols {
  DevBuf bytes, gbyte, gmask;
  uint64_t nbytes = 0, ngroups = 0;
  bool valid = false;
};

// Semi-external graphs (window.hip): the column ids (varint-packed or plain) and weights stay in
// pinned host memory; windows of consecutive destination rows are copied into a ring of device
// slots on copy streams while the compute stream works on the windows before, and the windows
// that fit into the device edge budget stay cached in device memory.  This is GraphMP's sliding
// window over shards (GraphPS.h:256-386) with its compressed edge cache (Global.h:129-255).
struct EdgeWindow {
  uint32_t r0 = 0, r1 = 0;  // destination rows
  uint64_t e0 = 0, e1 = 0;  // edges
  uint64_t b0 = 0, b1 = 0;  // bytes of the host id stream the window reads
  int64_t cache = -1;       // byte offset in the device cache, -1: streamed through the ring
  uint32_t task0 = 0, ntasks = 0, hub0 = 0, nhubs = 0, multi0 = 0, nmulti = 0, nparts = 0;
};
struct SemiExternal {
  bool packed = false;
  uint32_t blk_log = 0, nblk = 0, blk_words = 0;  // activation: source blocks of 2^blk_log ids
  uint64_t slot_bytes = 0, max_edges = 0;         // id bytes of one ring slot, edges of the largest window
  uint32_t nslots = 0;
  uint64_t host_bytes = 0, cache_bytes = 0, budget = 0;
  std::vector<EdgeWindow> win;
  uint8_t *h_ids = nullptr;  // pinned: packed bytes (+ padding) or uint32 ids
  uint32_t *h_w = nullptr;   // pinned: weights as uint32, or
  uint8_t *h_w8 = nullptr;   // pinned: weights as bytes, when every weight fits one
  uint32_t wsize = 4;        // bytes per weight in the host array, the ring and the cache
  uint64_t wslot_bytes = 0;  // weight bytes of one ring slot
  bool has_w() const { return h_w || h_w8; }
  DevBuf gbyte, gmask;       // group index of the packed stream
  DevBuf ring, wring;        // nslots slots of ids and of weights
  DevBuf cache, wcache;      // cached windows (ids at EdgeWindow::cache, weights at wsize*(e0 - first cached edge))
  uint64_t wcache_e0 = 0;
  DevBuf plain;              // one window's decoded ids
  DevBuf tasks, hubs, multi; // gather plans of all windows, rows relative to the window
  uint32_t plan_key[4] = {0, 0, 0, 0};
  uint32_t max_parts = 0;
  DevBuf srcblk;             // uint64[nwin * blk_words]: source blocks of every window
  DevBuf stamp;              // uint32[nblk]: last sweep that changed a vertex of the block
  DevBuf active;             // uint8[nwin]
  DevBuf first;              // uint32[2V]: every row's first two ids (union-find CC), built on demand
  DevBuf wrow;               // uint32[nwin + 1]: first row of every window
  std::vector<hipStream_t> copy;  // copy streams of the ring slots
  std::vector<hipEvent_t> ready, freed;
  uint64_t slot_seq = 0;
  uint64_t streamed_bytes = 0, windows_run = 0, windows_skipped = 0;  // since the load
  // learned prefetch (GMP_OPT_WINDOW_PREFETCH): the window predictor's inputs and outputs
  DevBuf prev;               // uint8[nwin]: the activity flags of the sweep before
  DevBuf pred;               // uint8[nwin]: predicted activity of the sweep after next
  DevBuf wcnt;               // uint32[16]: counts behind the predictor's features
  DevBuf wfeat;              // float[16]: the last features
  uint64_t prefetched = 0, prefetch_wasted_bytes = 0;  // windows copied ahead, bytes of those not run
  SemiExternal() = default;
  SemiExternal(const SemiExternal &) = delete;
  SemiExternal &operator=(const SemiExternal &) = delete;
  ~SemiExternal();
};

struct Graph {
  uint32_t V = 0;
  uint64_t E = 0;
  bool off64 = false;
  bool weighted = false;
  bool symmetric = false;
  // layout as loaded: destination rows, sources in shard-file order
  Csr in;
  Csr out;                          // its transpose (built on demand)
  DevBuf out_deg;                   // uint32[V]
  uint32_t P = 1;
  std::vector<uint32_t> bounds;     // P+1
  std::vector<uint64_t> h_in_off;   // host copy of the in-CSR offsets
  // work layout: the vertices are renumbered by descending out-degree (ties by id), so the
  // most-gathered sources share cache lines and PageRank's source strips start with the hubs.
  // With reordering off the work layout is the loaded one.
  bool work_ready = false;
  bool work_identity = true;
  Csr win, wout;                    // work in/out CSR (unused when identity)
  bool wout_ready = false;
  DevBuf o2n, n2o;                  // uint32[V]: work id of an original id and back
  DevBuf wdeg;                      // uint32[V]: out-degree by work id
  std::vector<uint64_t> h_win_off;  // host copy of the work in-CSR offsets (planning)
  GatherPlan plan;                  // over the work in-CSR
  PackedCols packed;                // compressed work in-CSR columns (built on demand)
  StripCsr strips;                  // source-strip layout (built on demand)
  HubTable hubs;                    // hub slices of the work out-CSR (built on demand)
  std::unique_ptr<SemiExternal> semi;  // set: the edges live in host memory (window.hip)

  CsrView view(const Csr &c) const { return CsrView{c.off.p, c.idx.as<uint32_t>(), c.w.as<uint32_t>()}; }
  CsrView work_in() const { return work_identity ? view(in) : view(win); }
  // a symmetric unweighted graph is its own transpose
  CsrView work_out() const {
    if (symmetric && !weighted) return work_in();
    return work_identity ? view(out) : view(wout);
  }
  const uint32_t *work_deg() const { return work_identity ? out_deg.as<uint32_t>() : wdeg.as<uint32_t>(); }
  const std::vector<uint64_t> &work_h_off() const { return work_identity ? h_in_off : h_win_off; }
};

// Tunables (gmp_set_option); the autotuner writes these.
struct Options {
  int reorder = 1;             // build the degree-ordered work layout
  uint32_t stream_cap = 2048;  // gather plan: edges per STREAM task
  uint32_t row_cap = 512;      // gather plan: rows per STREAM task
  uint32_t chunk = 4096;       // gather plan: edges per hub CHUNK task
  uint32_t vec_min = 128;      // gather plan: rows above this many edges go wave-wide
  uint32_t wg_per_cu = 16;     // gather grid = wg_per_cu * CUs
  uint32_t strip_log = 15;     // PageRank source strips of 2^strip_log ids, 8..15 (0: row gather)
  uint32_t bin_log = 14;       // PageRank destination bins of up to 2^bin_log rows, 8..14
  uint32_t bin_task_log = 20;  // PageRank bins with more than 2^k messages are split, 8..22
  uint32_t strip_task_log = 6; // strip gathers: 2^k wave tasks per workgroup task (one staging), 0..12
  bool selective = true;       // sweeps skip strips and windows whose sources did not change
  uint64_t edge_budget_mb = 0; // > 0: graphs whose edge data exceed it load semi-externally
  uint32_t window_mb = 64;     // semi-external windows: id bytes per window
  int window_packed = 1;       // semi-external windows: varint-packed ids
  uint32_t push_log = 10;      // CC / Bellman-Ford: frontiers of at most V >> push_log push (0: sweeps only)
  int cc_union_find = 1;       // CC of a symmetric graph by sampled union-find (0: sweeps and pushes)
  uint32_t delta_pull_log = 3; // delta-stepping rounds with more than E >> k frontier out-edges pull (0: never)
  int learned_finish = 0;      // learned SSSP: class 7 finishes by Bellman-Ford instead of width 1024
  // batched runs (batch.hip; gmp_batch_autotune writes these)
  uint32_t bsssp_unroll = 4;     // SSSP relax: edges per lane in flight (2, 4, 8)
  uint32_t bsssp_grid = 4;       // SSSP relax: workgroups per CU
  uint32_t bsssp_scan_log = 8;   // SSSP: frontiers above ngv >> k walk the mask array (0: always the queue)
  uint32_t bsssp_hub_log = 11;   // SSSP: entries above 2^k out-edges are relaxed in pieces (0: never)
  uint32_t ppr_piece = 1024;     // PPR: rows with more in-edges are cut into pieces of this many
  uint32_t ppr_run_rows = 32;    // PPR: rows per slot in a row run
  uint32_t ppr_run_edges = 2048; // PPR: edges per slot in a row run
  uint32_t ppr_loads = 8;        // PPR: source loads in flight per lane (4, 8, 16)
  uint32_t ppr_grid = 8;         // PPR gather: workgroups per CU
  int window_prefetch = 0;       // semi-external sweeps: learned window prefetch and order
};

// Many graphs resident at once as one disjoint union (graph k owns union ids
// [vbase[k], vbase[k+1])), for batched multi-instance runs.
struct BatchGraphs {
  std::vector<uint32_t> vbase;        // G+1
  std::vector<uint64_t> ebase;        // G+1
  std::vector<uint64_t> h_off;        // union in-CSR offsets (host copy)
  Csr in, out;                        // union in-CSR (offsets always uint64) and its transpose
  DevBuf w8;                          // uint8[E]: out's weights as bytes, in place of out.w, when all fit
  DevBuf outdeg;                      // uint32[Vtot]
  // each graph is stored in its degree-ordered layout (as the engine's work layout): o2n maps a
  // union id in the caller's numbering to the stored one (identity when reordering is off)
  DevBuf o2n;                         // uint32[Vtot]
  std::vector<uint32_t> h_o2n;        // host copy
  uint32_t V = 0;
  uint64_t E = 0;
  bool weighted = false;
  bool ready = false;
};

// Scheduling MLP (mlp.hip): 16 -> 64 (ReLU) -> 8 width logits + 64 shard logits.
constexpr int kMlpIn = 16, kMlpHid = 64, kMlpClasses = 8, kMlpShards = 64;
constexpr int kMlpOut = kMlpClasses + kMlpShards;
constexpr int kMlpParams = kMlpIn * kMlpHid + kMlpHid + kMlpHid * kMlpOut + kMlpOut;  // 5768
struct Mlp {
  DevBuf params, grad, adam_m, adam_v;
  int step = 0;
  bool trained = false;
};

struct Engine {
  int device = 0;
  int num_cus = 256;
  hipStream_t stream = nullptr;
  Graph g;
  BatchGraphs batch;
  Mlp mlp;      // delta-stepping bucket schedule
  Mlp mlp_win;  // semi-external window predictor
  int net_sel = 0;  // network the gmp_mlp_* calls act on (GMP_OPT_MLP_NET)
  Mlp &net() { return net_sel ? mlp_win : mlp; }
  Options opt;
  bool loaded = false;
  // scratch reused across runs
  DevBuf s_a, s_b, s_c, s_d, s_e, s_f, s_part, s_scal;
  DevBuf s_gaps;  // graph build: runs of empty rows left to a kernel of their own (count, then runs)
  void *h_pinned = nullptr;  // small pinned host buffer for per-iteration readbacks
  std::vector<hipEvent_t> ev_pool;  // timing events, created on first use and kept
  double *h_hist = nullptr;         // pinned, device-mapped: per-iteration PageRank residuals
  double *d_hist = nullptr;         // its device address
  size_t h_hist_n = 0;
  uint32_t last_actions[kMlpClasses] = {};  // classes the last learned SSSP run chose, per class
  void *scratch = nullptr;       // device address of noalloc.hip's scratch array on this device
  hipEvent_t na_ev[2] = {};      // timing events of gmp_pagerank_noalloc (created with the engine)
};

// The first n events of the engine's pool (created as needed, destroyed with the engine).
hipEvent_t *engine_events(Engine &e, size_t n);
// n doubles of pinned host memory that kernels write through *dev (kept with the engine).
double *engine_host_doubles(Engine &e, size_t n, double **dev);

void set_last_error(const std::string &m);
const char *last_error();

// Wraps a C ABI body: internal errors become status codes plus the last-error text.
#define GMP_API_BEGIN try {
#define GMP_API_END                                                                    \
  return GMP_OK;                                                                       \
  }                                                                                    \
  catch (const ::gmp::Error &e) {                                                      \
    ::gmp::set_last_error(e.what());                                                   \
    return e.code;                                                                     \
  }                                                                                    \
  catch (const std::bad_alloc &) {                                                     \
    ::gmp::set_last_error("host allocation failed");                                   \
    return GMP_ENOMEM;                                                                 \
  }                                                                                    \
  catch (const std::exception &e) {                                                    \
    ::gmp::set_last_error(e.what());                                                   \
    return GMP_EINVAL;                                                                 \
  }

// ---- format.cpp (host) ----
void npy_read_u32(const std::string &path, std::vector<uint32_t> &out);
void npy_write_i32(const std::string &path, const int32_t *data, uint64_t n);
void npy_write_u32(const std::string &path, const uint32_t *data, uint64_t n);
void partition_bounds(const uint64_t *indeg, uint32_t V, uint32_t P, uint32_t *bounds);

// ---- edgelist.cpp ----
// Destination-row CSR built on the host from a text edge list, with the raw in-edge counts
// (duplicates counted) that GraphMP's partitioner and shard headers use.
struct EdgeCsr {
  std::vector<uint64_t> off;        // V + 1
  std::vector<uint32_t> idx, w;     // E (w empty without weights)
  std::vector<uint64_t> raw_indeg;  // V
};
// V == 0: set to the largest id + 1.  flags: GMP_EDGES_WEIGHTED, GMP_EDGES_SYMMETRIZE.
void read_edge_list(const std::string &path, uint32_t flags, uint32_t &V, EdgeCsr &out);
void write_shard_files(const std::string &dir, uint32_t V, const EdgeCsr &g, uint32_t P, const uint32_t *bounds);

// ---- graph_build.hip ----
void graph_upload(Engine &e, uint32_t V, uint64_t E, const uint64_t *row_off,
                  const uint32_t *col_idx, const uint32_t *weights, bool force_off64);
void graph_generate_rmat(Engine &e, const gmp_rmat_params &p);
void graph_finish(Engine &e, uint32_t P, const uint32_t *bounds);  // out-degree, shards
void graph_build_out(Engine &e);                                   // transpose of `in`
void ensure_work(Engine &e);                                       // degree-ordered layout
void ensure_work_out(Engine &e);                                   // its transpose
// out[v] = x[o2n[v]] for 4-byte values (work ids -> original ids); device pointers.
void map_to_original(Engine &e, const uint32_t *x_work, uint32_t *out);
uint32_t work_id(Engine &e, uint32_t original);
// Transpose of a CSR (weights travel with their edge); rows of dst = columns of src.
void transpose_csr(Engine &e, const CsrView &src, uint32_t V, uint64_t E, bool off64, bool weighted, Csr &dst);
// deg[v] = number of entries equal to v in idx[0..E)
void count_degrees(Engine &e, const uint32_t *idx, uint64_t E, uint32_t V, uint32_t *deg);
void graph_download(Engine &e, uint64_t *row_off, uint32_t *col_idx, uint32_t *weights,
                    uint32_t *out_degree);

// ---- window.hip ----
// True when a graph of E edges (weighted or not) must load semi-externally under the options.
bool semi_wanted(const Engine &e, uint64_t E, bool weighted);
// Takes the host column ids (and weights) of the loaded graph, whose offsets and out-degrees are
// already on the device, into pinned host memory and builds the windows.
void semi_setup(Engine &e, const uint32_t *idx, const uint32_t *w);
// Device pointers of one window: ids (plain, indexed by global edge), pk (packed, indexed by the
// global group index), w (weights, indexed by global edge; null when not requested).
struct WindowDev {
  const uint32_t *idx = nullptr;
  Packed pk{};
  const uint32_t *w = nullptr;   // weights by global edge, or
  const uint8_t *w8 = nullptr;   // the same as bytes
  int slot = -1;
};
// Makes window `w` available to work enqueued next on e.stream: a cached window at once, a
// streamed one by a copy on the copy stream that e.stream waits for.  plain: ids decoded to
// uint32 (always available for plain streams; packed ones are decoded on e.stream).
// Windows with the streamed ones spread evenly among the cached ones, a cached one first: the
// copy engines keep filling the ring while cached windows are gathered.
std::vector<uint32_t> window_order(const SemiExternal &s);
WindowDev window_begin(Engine &e, uint32_t w, bool weights, bool plain);
// Releases the ring slot of window_begin after the work enqueued so far on e.stream.
void window_end(Engine &e, const WindowDev &d);
// Learned prefetch: issues the copy of streamed window w into the next ring slot now and returns
// the slot (-1 for a cached window); window_begin_slot then uses that slot in place of a copy, or
// window_release gives it back unused (its bytes counted as wasted).
int window_prefetch(Engine &e, uint32_t w, bool weights);
WindowDev window_begin_slot(Engine &e, uint32_t w, int slot, bool weights, bool plain);
void window_release(Engine &e, uint32_t w, int slot, bool weights);
// window_activity plus the window predictor: pred (nwin) receives the windows it expects to be
// active in the sweep after, feat (16) the features it read.
const uint8_t *window_activity_predicted(Engine &e, uint32_t sweep, uint8_t *pred, float *feat);
// Bit j: an active window overlaps shard j (rows [j V / 64, (j + 1) V / 64)).
uint64_t window_shard_mask(const SemiExternal &s, const uint8_t *act, uint32_t V);
void window_rows(Engine &e);  // s.wrow
// Min sweeps (kind 0 labels, 1 unit, 2 weighted distances) recording the window predictor's trace.
void semi_trace(Engine &e, int kind, uint32_t src, gmp_trace_row *rows, uint32_t cap, uint32_t *nrows,
                gmp_stats *st);
// Gather plans of the windows for the current plan options.
void window_plans(Engine &e);
// Activation: flags[w] = 1 when a source block of window w holds a vertex changed in `sweep`;
// returns the flags on the host (e.stream synchronised).
const uint8_t *window_activity(Engine &e, uint32_t sweep);
void window_stamp_init(Engine &e, uint32_t value, int64_t only_vertex);
void semi_cc(Engine &e, uint32_t *label_out, gmp_stats *st);
void semi_sssp(Engine &e, uint32_t src, bool unit, uint32_t *dist_out, gmp_stats *st);
void semi_download(Engine &e, uint32_t *col_idx, uint32_t *weights);
void semi_info(Engine &e, gmp_window_stats *out);
// Host varint packing in the format of varint.hpp (compress.hip builds the same on the device).
void pack_host(const uint64_t *off, uint32_t V, const uint32_t *idx, uint64_t E, std::vector<uint64_t> &gbyte,
               std::vector<uint64_t> &gmask, uint8_t *bytes, uint64_t &nbytes, bool measure_only);
void unpack_host(const uint8_t *bytes, const uint64_t *gbyte, const uint64_t *gmask, uint64_t E, uint64_t e0,
                 uint64_t e1, uint32_t *out);

// ---- plan.cpp (host) ----
void build_gather_plan_host(const uint64_t *off, uint32_t V, GatherPlan &plan,
                            std::vector<uint32_t> &tasks, std::vector<HubRow> &hubs,
                            std::vector<uint32_t> &multi);
void ensure_plan(Engine &e);
void plan_for(Engine &e, GatherPlan &pl, const uint64_t *off, uint32_t nrows);

// ---- compress.hip ----
void ensure_packed(Engine &e);

// ---- strip.hip ----
// Builds e.g.strips for e.opt.strip_log / bin_log; returns false when the layout does not apply
// (strips off, or more than 2^32 - 1 edges).
bool ensure_strips(Engine &e);
// The packed form of the layout's codes (built from them on the host); false when there is no
// strip layout or the stream would pass 2^32 words.
bool ensure_strip_pack(Engine &e);
// The layout's edges as 4-byte source work ids (built from the codes on the device); false when
// there is no strip layout.
bool ensure_strip_ids(Engine &e);
// Min-type sweeps one at a time, for switching between sweeps and frontier pushes.  kind: 0 CC,
// 1 unit-weight SSSP, 2 weighted SSSP.  c holds complemented values (c = ~x, 0 = unreached) in
// work order; vstamp (V entries, or null) receives, per vertex, the number of the last sweep that
// changed it.  sweep_open returns null, having launched nothing, when the layout does not apply.
struct SweepRun;
struct SweepRunDeleter {
  void operator()(SweepRun *s) const;
};
using SweepRunPtr = std::unique_ptr<SweepRun, SweepRunDeleter>;
SweepRunPtr sweep_open(Engine &e, int kind, uint32_t *c, uint32_t *vstamp);
// Starting values: CC c = ~(original id), every strip active; SSSP only the source reached.
void sweep_start(SweepRun &r, uint32_t src);
// Continues from a frontier: the n vertices of front (device, work ids) changed and are not yet
// propagated; every other strip is idle.
void sweep_enter(SweepRun &r, const uint32_t *front, uint32_t n);
// Runs one sweep; returns the number of vertices it changed.
uint32_t sweep_step(SweepRun &r);
uint32_t sweep_last(const SweepRun &r);
// Delta-stepping pull rounds over the strip layout (plain distances, unit or stored weights):
// the round relaxes exactly the out-edges of the near frontier -- the vertices whose near stamp
// is fstamp (front lists them, for the strips to gather) -- like a push round would, with the
// same appends: improved vertices below thr to near_out (stamp next_stamp, out-degrees summed
// into near_deg), the others to far_out once (far_flag).  Nothing is synchronised.
struct PullRound {
  uint32_t *dist, *near_stamp, *far_flag, *near_out, *far_out, *near_cnt, *far_cnt, *changed;
  const uint32_t *front, *deg;
  unsigned long long *near_deg;
  uint32_t nfront, fstamp, next_stamp, thr;
};
SweepRunPtr pull_open(Engine &e, bool unit);
void pull_round(SweepRun &r, const PullRound &p);
uint64_t sweep_edges(SweepRun &r);
void complement_values(Engine &e, uint32_t *x, uint32_t n);
// Row id of every edge of a CSR (device array of E entries).
void edge_row_ids(Engine &e, const CsrView &src, uint32_t V, uint64_t E, bool off64, DevBuf &rowid);

// ---- autotune.cpp ----
void autotune(Engine &e, int32_t iters, gmp_tune_result *out);

// ---- mlp.hip ----
void mlp_ensure(Engine &e);
void mlp_init(Engine &e, uint64_t seed);
void mlp_set(Engine &e, const float *params);
void mlp_get(Engine &e, float *params);
void mlp_forward(Engine &e, const float *X, uint32_t n, float *Y);
float mlp_loss_grad(Engine &e, const float *X, const int32_t *y, const uint64_t *mask, uint32_t n, float *grad);
float mlp_train(Engine &e, const float *X, const int32_t *y, const uint64_t *mask, uint32_t n, int32_t epochs,
                float lr);

// ---- noalloc.hip ----
// Resolves the device address of the scratch array on e.device (current) into e.scratch, once per
// device.
void noalloc_attach(Engine &e);

// ---- pagerank.hip ----
void run_pagerank(Engine &e, const gmp_pr_params &p, float *rank_out, gmp_stats *st);
void run_compat(Engine &e, int app, int max_iter, const void *msg_init, const int32_t *vout,
                void *msg_out, void *value_out, gmp_stats *st);

// ---- traverse.hip ----
void run_sssp(Engine &e, uint32_t src, const gmp_sssp_params &p, uint32_t *dist_out,
              gmp_stats *st);
void run_cc(Engine &e, uint32_t *label_out, gmp_stats *st);
void run_sssp_trace(Engine &e, uint32_t src, const gmp_sssp_params &p, gmp_trace_row *rows, uint32_t cap,
                    uint32_t *nrows, uint32_t *dist_out, gmp_stats *st);

// ---- batch.hip ----
void batch_generate(Engine &e, uint32_t ngraphs, const gmp_rmat_params *params);
void batch_sssp(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *source, uint32_t *dist_out,
                gmp_stats *st);
void batch_ppr(Engine &e, uint32_t n, const uint32_t *graph, const uint32_t *seed, float damping, int32_t iters,
               float *rank_out, gmp_stats *st);
void batch_run(Engine &e, uint32_t ns, const uint32_t *sgraph, const uint32_t *source, uint32_t np,
               const uint32_t *pgraph, const uint32_t *seed, float damping, int32_t iters, uint32_t *dist_out,
               float *rank_out, gmp_stats *st_sssp, gmp_stats *st_ppr, double *ms_total);
void batch_autotune(Engine &e, uint32_t ns, const uint32_t *sgraph, const uint32_t *source, uint32_t np,
                    const uint32_t *pgraph, const uint32_t *seed, float damping, int32_t iters,
                    gmp_batch_tune_result *out);

// ---- shared constants of the synthetic-input definition ----
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Elapsed-time helper on the engine stream.
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    GMP_HIP(hipEventCreate(&a));
    GMP_HIP(hipEventCreate(&b));
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  float ms() const {
    float t = 0;
    GMP_HIP(hipEventElapsedTime(&t, a, b));
    return t;
  }
};

}  // namespace gmp
