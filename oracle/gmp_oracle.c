/*
 * gmp_oracle.c -- CPU oracle for the gfx950 graph engine.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the checker (or the timed CPU baseline).  The product path
 * (graphee_amd/) never links, loads or calls it.
 *
 * Every function restates one piece of GraphMP's algorithm (cap-ntu/Graphee,
 * directory GraphMP/) in plain C, or is a textbook method that pins it:
 *
 *   orc_partition_bounds   GraphMP/tools/format_spark.py:85-108 equal-in-edge partitioner
 *   orc_app_sync           GraphMP/system/GraphPS.h:211-386 run() loop (synchronous
 *                          Msg/MsgNew swap, changed count against VertexValue, stop rule
 *                          changed==0 && step>1) with the three per-vertex update rules
 *                          GraphMP/app/pagerank.cpp:42-59, sssp.cpp:40-57, cc.cpp:40-56
 *   orc_pr_f64             textbook FP64 power iteration (uniform or single-seed teleport,
 *                          dangling mass redistributed to the teleport vector)
 *   orc_sssp_dijkstra      textbook binary-heap Dijkstra (pins Bellman-Ford / delta-stepping)
 *   orc_cc_unionfind       textbook union-find, label = minimum id in the component
 *   orc_graphmp_baseline   GraphPS.h:256-386 engine shape for timing: OpenMP guided over
 *                          destination-interval shards, per-shard pull gather, sync swap,
 *                          selective scheduling of shards whose sources changed, by exact
 *                          source sets or by GraphMP's per-shard bloom filters
 *   orc_bf_*               the vendored bloom filter GraphMP uses for that activation
 *                          (GraphMP/system/bloom_filter.hpp) with GraphPS.h:99-108's
 *                          parameters: sizing, salts, the hash of a 4-byte key
 *   orc_pr_f64_step_residual  one FP64 step from FP32 ranks: ||T(x)-x||_1 and _inf
 *   orc_pr_f64_refine      parallel FP64 power iteration from FP32 ranks to an a-posteriori
 *                          bound, the reference for measured L1 / L-inf errors
 *   orc_mlp_loss_grad      FP64 loss and gradient of the scheduling MLP
 *
 * Pinning (which test checks which function) is listed in tests/test_oracle.py.
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 5

/* app ids shared with the engine's compat apps */
#define ORC_APP_PAGERANK 1 /* float, GraphMP/app/pagerank.cpp */
#define ORC_APP_SSSP 2     /* int32, unit weights, GraphMP/app/sssp.cpp */
#define ORC_APP_CC 3       /* int32, min label over in-edges, GraphMP/app/cc.cpp */
#define ORC_APP_SSSP_W 4   /* uint32 weighted generalisation, saturating add */

/* ------------------------------------------------------------------------------------ */
/* format_spark.py:91-108.  vertex_vector = cumsum(indeg) with a leading 0 (:95-96);
 * edge_per_partition = int(math.ceil(edge_num/partition_num)) (:93) -- the script is
 * Python 2 and both operands are integers, so '/' floors before ceil sees it;
 * boundary i (1..P-1) = argmax(vertex_vector >= edge_per_partition*i) (:99-100), the first
 * prefix index reaching the target; the last boundary is vertex_num (:105) and a 0 is
 * prepended (:112).  Shard p covers destinations [bounds[p], bounds[p+1]). */
int orc_partition_bounds(const int64_t *indeg, int64_t V, int32_t P, int64_t *bounds) {
  if (P <= 0 || V < 0) return ORC_EINVAL;
  int64_t E = 0;
  for (int64_t v = 0; v < V; v++) E += indeg[v];
  int64_t epp = E / P;
  int64_t j = 0, cum = 0; /* cum == vertex_vector[j] */
  bounds[0] = 0;
  for (int32_t i = 1; i < P; i++) {
    int64_t target = epp * (int64_t)i;
    while (cum < target) {
      cum += indeg[j];
      j++;
    }
    bounds[i] = j;
  }
  bounds[P] = V;
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* GraphPS<T>::run synchronous loop with the per-app update of the reference.
 * off/idx: in-CSR over destinations (rows) with ascending sources, int64 offsets.
 * msg: in = init_vertex() result (_VertexMsg), out = final _VertexMsg.
 * value: out = final _VertexValue (initialised to 0 as every init_vertex does).
 * vout: _VertexOut as loaded (pagerank only; zeros are replaced by 1, pagerank.cpp:77-80).
 * w: per-edge weights (ORC_APP_SSSP_W only).
 * Returns the number of executed iterations in *iters and the per-iteration changed
 * counts (the reference's _Changed_Vertex, GraphPS.h:292) in changed[0..iters). */
int orc_app_sync(int32_t app, int64_t V, const int64_t *off, const uint32_t *idx,
                 const uint32_t *w, const int32_t *vout, int32_t max_iter, void *msg,
                 void *value, int32_t *iters, int64_t *changed) {
  if (V <= 0) return ORC_EINVAL;
  size_t es = 4;
  void *msgnew = malloc((size_t)V * es);
  int32_t *out = NULL;
  if (!msgnew) return ORC_ENOMEM;
  memset(value, 0, (size_t)V * es);
  memcpy(msgnew, msg, (size_t)V * es); /* GraphPS.h:248 */
  if (app == ORC_APP_PAGERANK) {
    out = (int32_t *)malloc((size_t)V * sizeof(int32_t));
    if (!out) { free(msgnew); return ORC_ENOMEM; }
    for (int64_t v = 0; v < V; v++) out[v] = vout[v] == 0 ? 1 : vout[v];
  }
  int32_t step;
  for (step = 0; step < max_iter; step++) {
    int64_t nchanged = 0;
    if (app == ORC_APP_PAGERANK) {
      const float *m = (const float *)msg;
      float *mn = (float *)msgnew, *val = (float *)value;
      for (int64_t v = 0; v < V; v++) {
        float rel = 0;
        for (int64_t k = off[v]; k < off[v + 1]; k++) {
          uint32_t u = idx[k];
          rel += m[u] / (float)out[u]; /* pagerank.cpp:46: float / int32 -> float */
        }
        rel = (float)((double)rel * 0.85 + 1.0 / (double)V); /* pagerank.cpp:48 */
        mn[v] = rel;
        if (rel != val[v]) { nchanged++; val[v] = rel; }
      }
    } else if (app == ORC_APP_SSSP || app == ORC_APP_CC) {
      const int32_t *m = (const int32_t *)msg;
      int32_t *mn = (int32_t *)msgnew, *val = (int32_t *)value;
      for (int64_t v = 0; v < V; v++) {
        int32_t best = m[v];
        for (int64_t k = off[v]; k < off[v + 1]; k++) {
          int32_t t = m[idx[k]];
          if (app == ORC_APP_SSSP) t = (int32_t)((uint32_t)t + 1u); /* sssp.cpp:43 */
          if (best > t) best = t;
        }
        mn[v] = best;
        if (best != val[v]) { nchanged++; val[v] = best; }
      }
    } else if (app == ORC_APP_SSSP_W) {
      const uint32_t *m = (const uint32_t *)msg;
      uint32_t *mn = (uint32_t *)msgnew, *val = (uint32_t *)value;
      for (int64_t v = 0; v < V; v++) {
        uint32_t best = m[v];
        for (int64_t k = off[v]; k < off[v + 1]; k++) {
          uint64_t t = (uint64_t)m[idx[k]] + (uint64_t)w[k];
          if (t > 0xFFFFFFFFull) t = 0xFFFFFFFFull;
          if ((uint64_t)best > t) best = (uint32_t)t;
        }
        mn[v] = best;
        if (best != val[v]) { nchanged++; val[v] = best; }
      }
    } else {
      free(msgnew);
      free(out);
      return ORC_EINVAL;
    }
    memcpy(msg, msgnew, (size_t)V * es); /* GraphPS.h:334/343 */
    if (changed) changed[step] = nchanged;
    if (nchanged == 0 && step > 1) { step++; break; } /* GraphPS.h:383-385 */
  }
  *iters = step;
  free(msgnew);
  free(out);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* FP64 power iteration.  x_{k+1}[v] = (1-d) t[v] + d (sum_{u->v} x_k[u]/outdeg[u] + D_k t[v]),
 * D_k = sum over dangling u of x_k[u]; t = uniform 1/V, or e_seed when seed >= 0.
 * Stops after max_iter iterations, or as soon as the L1 change ||x_{k+1}-x_k||_1 < tol
 * (tol > 0).  *iters = iterations executed, *resid = the last L1 change.  threads > 1 runs the
 * two loops with OpenMP (the per-vertex sums keep their order; only D and the residual are summed
 * in another order). */
int orc_pr_f64(int64_t V, const int64_t *off, const uint32_t *idx, const uint32_t *outdeg,
               double d, int32_t max_iter, double tol, int64_t seed, const double *x0,
               double *x, int32_t *iters, double *resid, int32_t threads) {
  if (V <= 0) return ORC_EINVAL;
  if (threads <= 0) threads = 1;
  double *c = (double *)malloc((size_t)V * sizeof(double));
  double *xn = (double *)malloc((size_t)V * sizeof(double));
  if (!c || !xn) { free(c); free(xn); return ORC_ENOMEM; }
  for (int64_t v = 0; v < V; v++) x[v] = x0 ? x0[v] : (seed >= 0 ? (v == seed ? 1.0 : 0.0) : 1.0 / (double)V);
  int32_t it = 0;
  double r = 0;
  for (it = 0; it < max_iter;) {
    double D = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : D) if (threads > 1)
    for (int64_t u = 0; u < V; u++) {
      if (outdeg[u] == 0) { D += x[u]; c[u] = 0; }
      else c[u] = x[u] / (double)outdeg[u];
    }
    r = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096) reduction(+ : r) if (threads > 1)
    for (int64_t v = 0; v < V; v++) {
      double s = 0;
      for (int64_t k = off[v]; k < off[v + 1]; k++) s += c[idx[k]];
      double t = seed >= 0 ? (v == seed ? 1.0 : 0.0) : 1.0 / (double)V;
      xn[v] = (1.0 - d) * t + d * (s + D * t);
      r += fabs(xn[v] - x[v]);
    }
    memcpy(x, xn, (size_t)V * sizeof(double));
    it++;
    if (tol > 0 && r < tol) break;
  }
  if (iters) *iters = it;
  if (resid) *resid = r;
  free(c);
  free(xn);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* Dijkstra over an out-CSR (rows = sources).  w == NULL means unit weights.  Unreachable
 * vertices get 0xFFFFFFFF; a path whose length would exceed 0xFFFFFFFE saturates there
 * (the engine's saturating uint32 add). */
typedef struct { uint64_t key; uint32_t v; } orc_heap_item;

static void heap_push(orc_heap_item *h, int64_t *n, orc_heap_item it) {
  int64_t i = (*n)++;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (h[p].key <= it.key) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = it;
}
static orc_heap_item heap_pop(orc_heap_item *h, int64_t *n) {
  orc_heap_item top = h[0], last = h[--(*n)];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    uint64_t mk = last.key;
    if (l < *n && h[l].key < mk) { m = l; mk = h[l].key; }
    if (r < *n && h[r].key < mk) { m = r; }
    if (m == i) break;
    h[i] = h[m];
    i = m;
  }
  h[i] = last;
  return top;
}

int orc_sssp_dijkstra(int64_t V, const int64_t *ooff, const uint32_t *oidx, const uint32_t *ow,
                      uint32_t src, uint32_t *dist) {
  if (V <= 0 || (int64_t)src >= V) return ORC_EINVAL;
  int64_t E = ooff[V];
  orc_heap_item *h = (orc_heap_item *)malloc((size_t)(E + 1) * sizeof(orc_heap_item));
  if (!h) return ORC_ENOMEM;
  for (int64_t v = 0; v < V; v++) dist[v] = 0xFFFFFFFFu;
  int64_t n = 0;
  dist[src] = 0;
  orc_heap_item s0 = {0, src};
  heap_push(h, &n, s0);
  while (n > 0) {
    orc_heap_item it = heap_pop(h, &n);
    if (it.key != dist[it.v]) continue;
    for (int64_t k = ooff[it.v]; k < ooff[it.v + 1]; k++) {
      uint64_t nd = it.key + (ow ? ow[k] : 1u);
      if (nd > 0xFFFFFFFEull) nd = 0xFFFFFFFEull;
      uint32_t v = oidx[k];
      if (nd < dist[v]) {
        dist[v] = (uint32_t)nd;
        orc_heap_item ni = {nd, v};
        heap_push(h, &n, ni);
      }
    }
  }
  free(h);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* Union-find over any edge set given as a CSR; label[v] = minimum vertex id in v's
 * (weakly) connected component -- the unique fixed point of min-label propagation on
 * a symmetric graph (GraphMP/app/cc.cpp:40-45). */
static int64_t uf_find(int64_t *p, int64_t x) {
  while (p[x] != x) {
    p[x] = p[p[x]];
    x = p[x];
  }
  return x;
}

int orc_cc_unionfind(int64_t V, const int64_t *off, const uint32_t *idx, uint32_t *label) {
  if (V <= 0) return ORC_EINVAL;
  int64_t *p = (int64_t *)malloc((size_t)V * sizeof(int64_t));
  if (!p) return ORC_ENOMEM;
  for (int64_t v = 0; v < V; v++) p[v] = v;
  for (int64_t v = 0; v < V; v++)
    for (int64_t k = off[v]; k < off[v + 1]; k++) {
      int64_t a = uf_find(p, v), b = uf_find(p, idx[k]);
      if (a < b) p[b] = a;
      else if (b < a) p[a] = b;
    }
  for (int64_t v = 0; v < V; v++) label[v] = (uint32_t)uf_find(p, v);
  free(p);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* CPU baseline: GraphMP's engine shape with the edges resident in RAM (its cache-level-0
 * best case).  GraphPS.h:256-386: per step, shards in shuffled order under
 * `omp parallel for schedule(guided)` (:263-278), each running the app's per-shard pull
 * gather (pagerank.cpp:42-59 etc.) into MsgNew; then the O(V) swap (:338-345); shards
 * none of whose sources changed are skipped when fewer than BF_THRE=0.001 of the vertices
 * changed (:318-368) -- exact per-shard source membership stands in for the bloom filter,
 * whose false positives only add work.
 *
 * app: ORC_APP_PAGERANK runs the `standard` formula in FP32 (plain sum like the reference,
 * teleport (1-d)/N, dangling mass redistributed) for a fixed max_iter;
 * ORC_APP_SSSP_W and ORC_APP_CC run the min rules until no vertex changes.
 * threads <= 0 uses omp_get_max_threads().  Returns iterations, wall seconds, edges
 * gathered. */
typedef struct {
  int32_t iterations;
  int32_t threads;
  double seconds;
  double edges;
  double shards; /* shard computations, summed over the steps */
} orc_baseline_stats;

/* ---- GraphMP's per-shard bloom filters -------------------------------------------------
 * GraphPS.h:99-108 gives every shard a filter of projected_element_count 1000000,
 * false_positive_probability 0.01 and random_seed 0xA5A5A5A5; GraphPS.h:225-245 inserts the
 * source of every edge of the shard (int32 keys); GraphPS.h:349-368 activates a shard when its
 * filter contains any changed vertex.  The filter (bloom_filter.hpp) restated:
 *   sizing (:107-157): the k in 1, 2, ..., 999 whose m(k) = -k n / ln(1 - p^(1/k)) is
 *     smallest (first minimum), m truncated to an integer and rounded up to whole bytes;
 *   salts (:466-510, k <= the predefined count): the first k predefined salts, then in
 *     place, i = 0..k-1, salt[i] = salt[i] * salt[(i+3) % k] + (uint32)seed', with
 *     seed' = seed * 0xA5A5A5A5 + 1 (the constructor, :176-185), all in 32-bit words;
 *   hash of a 4-byte key (hash_ap, :550-608: no 8-byte rounds, one 4-byte step with an even
 *     step counter): h = salt; h ^= ~((h << 11) + (key ^ (h >> 5)));
 *   bit (compute_indices :460-464, insert :259-270, contains :300-318): index h mod m, byte
 *     index / 8, bit index mod 8.
 * The salt table has to match the reference's word for word; only the first 16 entries are
 * carried, enough for every false-positive probability down to 1e-4 at any element count
 * (k = 14).  Exact bits of this restatement are "parity unpinned" (no reference build);
 * the tests pin the filter's behaviour (no false negatives, the designed false-positive
 * rate, the textbook optimum of k and m). */
#define ORC_BF_SALTS 16
static const uint32_t orc_bf_predef[ORC_BF_SALTS] = {
    0xAAAAAAAAu, 0x55555555u, 0x33333333u, 0xCCCCCCCCu, 0x66666666u, 0x99999999u, 0xB5B5B5B5u, 0x4B4B4B4Bu,
    0xAA55AA55u, 0x55335533u, 0x33CC33CCu, 0xCC66CC66u, 0x66996699u, 0x99B599B5u, 0xB54BB54Bu, 0x4BAA4BAAu};

typedef struct {
  uint32_t k;
  uint64_t m;                     /* bits */
  uint32_t salt[ORC_BF_SALTS];
} orc_bf_shape;

int orc_bf_params(uint64_t n, double p, uint64_t seed, orc_bf_shape *out) {
  if (!out || n == 0 || !(p > 0.0 && p < 1.0)) return ORC_EINVAL;
  double min_m = INFINITY, min_k = 0.0;
  for (double k = 1.0; k < 1000.0; k += 1.0) {
    const double cur = (-k * (double)n) / log(1.0 - pow(p, 1.0 / k));
    if (cur < min_m) {
      min_m = cur;
      min_k = k;
    }
  }
  uint64_t m = (uint64_t)min_m;
  if (m % 8) m += 8 - m % 8;
  const uint32_t k = (uint32_t)min_k;
  if (k < 1 || k > ORC_BF_SALTS || m == 0) return ORC_EINVAL;
  out->k = k;
  out->m = m;
  const uint32_t s2 = (uint32_t)(seed * 0xA5A5A5A5ull + 1ull);
  for (uint32_t i = 0; i < k; i++) out->salt[i] = orc_bf_predef[i];
  for (uint32_t i = 0; i < k; i++) out->salt[i] = out->salt[i] * out->salt[(i + 3) % k] + s2;
  return ORC_OK;
}

static inline uint64_t orc_bf_bit(const orc_bf_shape *f, uint32_t salt, uint32_t key) {
  uint32_t h = salt;
  h ^= ~((h << 11) + (key ^ (h >> 5)));
  return (uint64_t)h % f->m;
}

static inline void orc_bf_insert(const orc_bf_shape *f, uint8_t *bits, uint32_t key) {
  for (uint32_t i = 0; i < f->k; i++) {
    const uint64_t b = orc_bf_bit(f, f->salt[i], key);
    bits[b >> 3] |= (uint8_t)(1u << (b & 7u));
  }
}

static inline int orc_bf_contains(const orc_bf_shape *f, const uint8_t *bits, uint32_t key) {
  for (uint32_t i = 0; i < f->k; i++) {
    const uint64_t b = orc_bf_bit(f, f->salt[i], key);
    if (!(bits[b >> 3] & (1u << (b & 7u)))) return 0;
  }
  return 1;
}

/* One filter of shape f over keys[0..n) (bits: m / 8 bytes, cleared here), then `q` probes:
 * hit[j] = contains(probe[j]).  Exposes the filter to the tests. */
int orc_bf_probe(const orc_bf_shape *f, const uint32_t *keys, int64_t n, const uint32_t *probe, int64_t q,
                 uint8_t *hit) {
  if (!f || f->k == 0 || f->k > ORC_BF_SALTS || f->m == 0) return ORC_EINVAL;
  uint8_t *bits = (uint8_t *)calloc((size_t)(f->m / 8), 1);
  if (!bits) return ORC_ENOMEM;
  for (int64_t i = 0; i < n; i++) orc_bf_insert(f, bits, keys[i]);
  for (int64_t j = 0; j < q; j++) hit[j] = (uint8_t)orc_bf_contains(f, bits, probe[j]);
  free(bits);
  return ORC_OK;
}

static uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static int has_u32(const uint32_t *a, int64_t n, uint32_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && a[lo] == x;
}

int orc_graphmp_baseline(int32_t app, int64_t V, int32_t P, const int64_t *bounds,
                         const int64_t *off, const uint32_t *idx, const uint32_t *w,
                         const uint32_t *outdeg, float damping, int32_t max_iter,
                         int32_t threads, int32_t bloom, void *msg, orc_baseline_stats *st) {
  if (V <= 0 || P <= 0) return ORC_EINVAL;
  /* bloom != 0: GraphMP's filters decide which shards a sparse step activates (a superset of
   * the exact choice: false positives add shards, never drop one) */
  orc_bf_shape bf;
  uint8_t *bfbits = NULL;
  if (bloom) {
    if (orc_bf_params(1000000ull, 0.01, 0xA5A5A5A5ull, &bf) != ORC_OK) return ORC_EINVAL;
    bfbits = (uint8_t *)calloc((size_t)P, (size_t)(bf.m / 8));
    if (!bfbits) return ORC_ENOMEM;
  }
  if (threads <= 0) threads = omp_get_max_threads();
  int64_t E = off[V];
  void *msgnew = malloc((size_t)V * 4);
  float *contrib = NULL;
  uint8_t *pactive = (uint8_t *)malloc((size_t)P);
  int32_t *order = (int32_t *)malloc((size_t)P * sizeof(int32_t));
  /* per-shard sorted unique source sets: the exact stand-in for the per-shard bloom
   * filters built in GraphPS.h:225-245 (built before timing, as the reference builds its
   * filters during init) */
  /* shard p's set: srcset[ss_off[2p], ss_off[2p+1]), inside the shard's own edge range */
  uint32_t *srcset = (uint32_t *)malloc((size_t)(E > 0 ? E : 1) * sizeof(uint32_t));
  int64_t *ss_off = (int64_t *)malloc((size_t)(2 * P) * sizeof(int64_t));
  uint32_t *active = (uint32_t *)malloc((size_t)V * sizeof(uint32_t));
  const int64_t nwords = (V + 63) / 64;
  uint64_t *marks = (uint64_t *)calloc((size_t)threads * (size_t)nwords, sizeof(uint64_t));
  if (!msgnew || !pactive || !order || !srcset || !ss_off || !active || !marks) {
    free(msgnew); free(pactive); free(order); free(srcset); free(ss_off); free(active); free(marks);
    return ORC_ENOMEM;
  }
  if (app == ORC_APP_PAGERANK) contrib = (float *)malloc((size_t)V * sizeof(float));
  /* sorted unique sources per shard: mark them in a bitmap over V, read the marks in order
   * (shards in parallel, one bitmap per thread, cleared as it is read) */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int32_t p = 0; p < P; p++) {
    uint64_t *mk = marks + (size_t)omp_get_thread_num() * (size_t)nwords;
    int64_t a = off[bounds[p]], b = off[bounds[p + 1]], n = 0;
    for (int64_t q = a; q < b; q++) mk[idx[q] >> 6] |= 1ull << (idx[q] & 63u);
    for (int64_t wi = 0; wi < nwords; wi++) {
      uint64_t m = mk[wi];
      mk[wi] = 0;
      while (m) {
        srcset[a + n++] = (uint32_t)(wi * 64 + __builtin_ctzll(m));
        m &= m - 1;
      }
    }
    ss_off[2 * p] = a;
    ss_off[2 * p + 1] = a + n;
  }
  free(marks);
  if (bloom) {
    /* GraphPS.h:225-245: every edge's source into its shard's filter (during init) */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int32_t p = 0; p < P; p++) {
      uint8_t *bits = bfbits + (size_t)p * (size_t)(bf.m / 8);
      for (int64_t q = off[bounds[p]]; q < off[bounds[p + 1]]; q++) orc_bf_insert(&bf, bits, idx[q]);
    }
  }
  memcpy(msgnew, msg, (size_t)V * 4);
  memset(pactive, 1, (size_t)P);
  double t0 = omp_get_wtime();
  double edges = 0, shards = 0;
  int32_t step;
  uint64_t shuf = 0x5eed;
  for (step = 0; step < max_iter; step++) {
    for (int32_t k = 0; k < P; k++) order[k] = k;
    for (int32_t k = P - 1; k > 0; k--) { /* std::random_shuffle, GraphPS.h:263 */
      shuf = orc_mix64(shuf + 0x9E3779B97F4A7C15ull);
      int32_t j = (int32_t)(shuf % (uint64_t)(k + 1));
      int32_t t = order[k]; order[k] = order[j]; order[j] = t;
    }
    float base = 0, d = damping;
    if (app == ORC_APP_PAGERANK) {
      const float *m = (const float *)msg;
      double D = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : D)
      for (int64_t u = 0; u < V; u++) {
        if (outdeg[u]) contrib[u] = m[u] / (float)outdeg[u];
        else { contrib[u] = 0; D += m[u]; }
      }
      base = (float)((1.0 - damping) / (double)V + damping * D / (double)V);
    }
    int64_t nchanged = 0;
    double e_it = 0, s_it = 0;
#pragma omp parallel for num_threads(threads) schedule(guided) reduction(+ : nchanged, e_it, s_it)
    for (int32_t k = 0; k < P; k++) {
      int32_t p = order[k];
      if (!pactive[p]) continue;
      int64_t s = bounds[p], e = bounds[p + 1];
      e_it += (double)(off[e] - off[s]);
      s_it += 1.0;
      if (app == ORC_APP_PAGERANK) {
        const float *m = (const float *)msg;
        float *mn = (float *)msgnew;
        for (int64_t v = s; v < e; v++) {
          float rel = 0;
          for (int64_t q = off[v]; q < off[v + 1]; q++) rel += contrib[idx[q]];
          rel = base + d * rel;
          mn[v] = rel;
          if (rel != m[v]) nchanged++;
        }
      } else {
        const uint32_t *m = (const uint32_t *)msg;
        uint32_t *mn = (uint32_t *)msgnew;
        for (int64_t v = s; v < e; v++) {
          uint32_t best = m[v];
          for (int64_t q = off[v]; q < off[v + 1]; q++) {
            uint64_t t = (uint64_t)m[idx[q]] + (app == ORC_APP_SSSP_W ? (uint64_t)w[q] : 0u);
            if (t < best) best = (uint32_t)t;
          }
          mn[v] = best;
          if (best != m[v]) nchanged++;
        }
      }
    }
    edges += e_it;
    shards += s_it;
    /* GraphPS.h:316-368: collect changed vertices while swapping when the changed ratio
     * is below BF_THRE, then keep only shards holding a changed source */
    double ratio = (double)nchanged / (double)V;
    uint32_t *mo = (uint32_t *)msg;
    const uint32_t *mnew = (const uint32_t *)msgnew;
    if (ratio < 0.001) {
      int64_t na = 0;
      for (int64_t v = 0; v < V; v++) {
        if (mo[v] != mnew[v]) active[na++] = (uint32_t)v;
        mo[v] = mnew[v];
      }
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int32_t p = 0; p < P; p++) {
        uint8_t act = 0;
        if (bloom) {
          const uint8_t *bits = bfbits + (size_t)p * (size_t)(bf.m / 8);
          for (int64_t a = 0; a < na && !act; a++) act = (uint8_t)orc_bf_contains(&bf, bits, active[a]);
        } else {
          for (int64_t a = 0; a < na && !act; a++)
            act = (uint8_t)has_u32(srcset + ss_off[2 * p], ss_off[2 * p + 1] - ss_off[2 * p], active[a]);
        }
        pactive[p] = act;
      }
    } else {
#pragma omp parallel for num_threads(threads) schedule(static)
      for (int64_t v = 0; v < V; v++) mo[v] = mnew[v];
      memset(pactive, 1, (size_t)P);
    }
    if (app != ORC_APP_PAGERANK && nchanged == 0) { step++; break; }
  }
  double t1 = omp_get_wtime();
  if (st) {
    st->iterations = step;
    st->threads = threads;
    st->seconds = t1 - t0;
    st->edges = edges;
    st->shards = shards;
  }
  free(msgnew); free(contrib); free(pactive); free(order); free(srcset); free(ss_off); free(active); free(bfbits);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* A-posteriori FP64 check of an FP32 PageRank vector x (uniform teleport): one FP64 step
 * y = T(x) of the standard formula, returning ||y - x||_1 and ||y - x||_inf.  Since T is
 * a contraction with factor d in L1, ||x - x*||_1 <= ||T(x) - x||_1 / (1 - d) bounds the
 * distance to the exact fixed point. */
int orc_pr_f64_step_residual(int64_t V, const int64_t *off, const uint32_t *idx,
                             const uint32_t *outdeg, double d, const float *x, int32_t threads,
                             double *l1, double *linf) {
  if (V <= 0) return ORC_EINVAL;
  if (threads <= 0) threads = omp_get_max_threads();
  double *c = (double *)malloc((size_t)V * sizeof(double));
  if (!c) return ORC_ENOMEM;
  double D = 0, s1 = 0, mx = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : D)
  for (int64_t u = 0; u < V; u++) {
    if (outdeg[u] == 0) { D += x[u]; c[u] = 0; }
    else c[u] = (double)x[u] / (double)outdeg[u];
  }
  const double t = 1.0 / (double)V;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096) reduction(+ : s1) reduction(max : mx)
  for (int64_t v = 0; v < V; v++) {
    double s = 0;
    for (int64_t k = off[v]; k < off[v + 1]; k++) s += c[idx[k]];
    double y = (1.0 - d) * t + d * (s + D * t);
    double e = fabs(y - (double)x[v]);
    s1 += e;
    if (e > mx) mx = e;
  }
  *l1 = s1;
  *linf = mx;
  free(c);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* FP64 power iteration of the standard formula (uniform teleport, as orc_pr_f64) started from
 * the FP32 ranks x0 and run in parallel, for measuring the error of x0 against the FP64 fixed
 * point x*.  Iterates until the a-posteriori bound ||x_k - x*||_1 <= ||x_k - x_{k-1}||_1 * d/(1-d)
 * drops below `bound` or max_iter steps ran; x receives x_k, *iters the steps, *last_change the
 * final ||x_k - x_{k-1}||_1 (so the caller can state how far x_k itself may be from x*). */
int orc_pr_f64_refine(int64_t V, const int64_t *off, const uint32_t *idx, const uint32_t *outdeg,
                      double d, const float *x0, int32_t threads, int32_t max_iter, double bound,
                      double *x, int32_t *iters, double *last_change) {
  if (V <= 0) return ORC_EINVAL;
  if (threads <= 0) threads = omp_get_max_threads();
  double *c = (double *)malloc((size_t)V * sizeof(double));
  double *xn = (double *)malloc((size_t)V * sizeof(double));
  if (!c || !xn) { free(c); free(xn); return ORC_ENOMEM; }
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t v = 0; v < V; v++) x[v] = (double)x0[v];
  const double t = 1.0 / (double)V;
  int32_t it = 0;
  double r = 0;
  while (it < max_iter) {
    double D = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : D)
    for (int64_t u = 0; u < V; u++) {
      if (outdeg[u] == 0) { D += x[u]; c[u] = 0; }
      else c[u] = x[u] / (double)outdeg[u];
    }
    r = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096) reduction(+ : r)
    for (int64_t v = 0; v < V; v++) {
      double s = 0;
      for (int64_t k = off[v]; k < off[v + 1]; k++) s += c[idx[k]];
      xn[v] = (1.0 - d) * t + d * (s + D * t);
      r += fabs(xn[v] - x[v]);
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t v = 0; v < V; v++) x[v] = xn[v];
    it++;
    if (r * d / (1.0 - d) < bound) break;
  }
  *iters = it;
  *last_change = r;
  free(c);
  free(xn);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* FP64 reference of the scheduling MLP (the engine trains it in FP32 on the GPU):
 * h = relu(x W1 + b1), y = h W2 + b2, 16 -> 64 -> 72.  Loss per sample = cross-entropy of
 * softmax(y[0:8]) against label (no class term when label < 0) + mean over 64 shards of the
 * logistic loss of y[8+j] against mask bit j; averaged over n.  Parameters: W1[16][64], b1[64], W2[64][72], b2[72]. */
#define ORC_IN 16
#define ORC_HID 64
#define ORC_CLS 8
#define ORC_SH 64
#define ORC_OUT 72
int orc_mlp_loss_grad(const double *P, const float *X, const int32_t *label, const uint64_t *mask, int64_t n,
                      double *loss, double *grad) {
  const double *W1 = P, *b1 = P + ORC_IN * ORC_HID, *W2 = b1 + ORC_HID, *b2 = W2 + ORC_HID * ORC_OUT;
  const int np = ORC_IN * ORC_HID + ORC_HID + ORC_HID * ORC_OUT + ORC_OUT;
  if (n <= 0) return ORC_EINVAL;
  if (grad) memset(grad, 0, sizeof(double) * np);
  double *gW1 = grad, *gb1 = grad ? grad + ORC_IN * ORC_HID : NULL;
  double *gW2 = grad ? gb1 + ORC_HID : NULL, *gb2 = grad ? gW2 + ORC_HID * ORC_OUT : NULL;
  double L = 0;
  for (int64_t s = 0; s < n; s++) {
    const float *x = X + s * ORC_IN;
    double h[ORC_HID], y[ORC_OUT], dy[ORC_OUT];
    for (int j = 0; j < ORC_HID; j++) {
      double a = b1[j];
      for (int i = 0; i < ORC_IN; i++) a += (double)x[i] * W1[i * ORC_HID + j];
      h[j] = a > 0 ? a : 0;
    }
    for (int k = 0; k < ORC_OUT; k++) {
      double a = b2[k];
      for (int j = 0; j < ORC_HID; j++) a += h[j] * W2[j * ORC_OUT + k];
      y[k] = a;
    }
    double m = y[0], z = 0;
    for (int c = 1; c < ORC_CLS; c++) m = y[c] > m ? y[c] : m;
    for (int c = 0; c < ORC_CLS; c++) z += exp(y[c] - m);
    /* label < 0: a row without a class trains the shard outputs only */
    if (label[s] >= 0) L += log(z) + m - y[label[s]];
    for (int c = 0; c < ORC_CLS; c++)
      dy[c] = label[s] >= 0 ? (exp(y[c] - m) / z - (c == label[s])) / (double)n : 0.0;
    double b = 0;
    for (int j = 0; j < ORC_SH; j++) {
      const double t = (double)((mask[s] >> j) & 1u), v = y[ORC_CLS + j];
      b += (v > 0 ? v : 0) - t * v + log1p(exp(-fabs(v)));
      dy[ORC_CLS + j] = (1.0 / (1.0 + exp(-v)) - t) / (double)n / ORC_SH;
    }
    L += b / ORC_SH;
    if (!grad) continue;
    for (int k = 0; k < ORC_OUT; k++) {
      gb2[k] += dy[k];
      for (int j = 0; j < ORC_HID; j++) gW2[j * ORC_OUT + k] += h[j] * dy[k];
    }
    for (int j = 0; j < ORC_HID; j++) {
      if (h[j] <= 0) continue;
      double dh = 0;
      for (int k = 0; k < ORC_OUT; k++) dh += dy[k] * W2[j * ORC_OUT + k];
      gb1[j] += dh;
      for (int i = 0; i < ORC_IN; i++) gW1[i * ORC_HID + j] += (double)x[i] * dh;
    }
  }
  *loss = L / (double)n;
  return ORC_OK;
}
