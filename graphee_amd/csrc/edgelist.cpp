// Text edge lists -> the destination-row CSR of GraphMP's shards, on the host.
//
// GraphMP converts a TSV edge list with a Spark job (GraphMP/tools/format_spark.py): it counts
// the raw in-edges of every destination (:54-89), cuts P partitions of equal raw in-edge count
// (:91-108), and writes each partition as a scipy CSR over (destination - start, source) with
// ascending, deduplicated sources and self-loops kept, behind the header [raw edges, nnz,
// rows + 1, start, end] (:137-157).  The same rules here, natively: the file is memory-mapped
// and parsed by one thread per chunk, the CSR is built by a counting sort on the destination
// and a per-row sort + deduplication, and the result is either loaded into the engine
// (gmp_load_edge_list, gmp_api.cpp) or written as shards (gmp_convert_edge_list).
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gmp_internal.hpp"

namespace gmp {

namespace {

unsigned worker_count() {
  const unsigned n = std::thread::hardware_concurrency();
  return std::max(1u, std::min(n ? n : 1u, 16u));
}

struct Chunk {
  std::vector<uint32_t> src, dst, w;
  uint32_t max_id = 0;
  bool any = false;
  std::string err;
};

bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// Parses the complete lines of [p, end) (byte offset base in the file): "src dst [weight]"
// with non-negative integers.
void parse_range(const char *p, const char *end, bool weighted, uint64_t base, Chunk &out) {
  const char *const p0 = p;
  while (p < end) {
    const char *line = p;
    while (p < end && is_space(*p)) p++;
    if (p == end) break;
    if (*p == '\n') {
      p++;
      continue;
    }
    if (*p == '#' || *p == '%') {  // comment line
      while (p < end && *p != '\n') p++;
      continue;
    }
    uint64_t v[3] = {0, 0, 0};
    int n = 0;
    while (p < end && *p != '\n') {
      if (is_space(*p)) {
        p++;
        continue;
      }
      if (*p < '0' || *p > '9' || n == 3) {
        out.err = "edge list: bad token at byte " + std::to_string(base + (uint64_t)(p - p0));
        return;
      }
      uint64_t x = 0;
      while (p < end && *p >= '0' && *p <= '9') {
        x = x * 10 + (uint64_t)(*p - '0');
        if (x > 0xFFFFFFFEull) {
          out.err = "edge list: value above 2^32 - 2 at byte " + std::to_string(base + (uint64_t)(p - p0));
          return;
        }
        p++;
      }
      v[n++] = x;
    }
    if (n < 2 || (weighted && n < 3)) {
      out.err = "edge list: line at byte " + std::to_string(base + (uint64_t)(line - p0)) + " has " +
                std::to_string(n) + " fields";
      return;
    }
    out.src.push_back((uint32_t)v[0]);
    out.dst.push_back((uint32_t)v[1]);
    if (weighted) out.w.push_back((uint32_t)v[2]);
    out.max_id = std::max<uint32_t>(out.max_id, (uint32_t)std::max(v[0], v[1]));
    out.any = true;
  }
}

struct Mapped {
  int fd = -1;
  void *p = MAP_FAILED;
  size_t n = 0;
  ~Mapped() {
    if (p != MAP_FAILED && n) munmap(p, n);
    if (fd >= 0) close(fd);
  }
};

}  // namespace

void read_edge_list(const std::string &path, uint32_t flags, uint32_t &V, EdgeCsr &out) {
  const bool weighted = flags & GMP_EDGES_WEIGHTED, sym = flags & GMP_EDGES_SYMMETRIZE;
  Mapped m;
  m.fd = open(path.c_str(), O_RDONLY);
  if (m.fd < 0) fail(GMP_EIO, "cannot open " + path);
  struct stat st {};
  if (fstat(m.fd, &st) != 0) fail(GMP_EIO, "cannot stat " + path);
  m.n = (size_t)st.st_size;
  const char *text = "";
  if (m.n) {
    m.p = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (m.p == MAP_FAILED) fail(GMP_EIO, "cannot map " + path);
    text = static_cast<const char *>(m.p);
  }
  // chunks of whole lines, one per thread
  const unsigned T = worker_count();
  std::vector<size_t> cut(T + 1, m.n);
  cut[0] = 0;
  for (unsigned i = 1; i < T; i++) {
    size_t c = std::max(cut[i - 1], m.n / T * i);
    while (c > 0 && c < m.n && text[c - 1] != '\n') c++;
    cut[i] = c;
  }
  std::vector<Chunk> ch(T);
  std::vector<std::thread> th;
  for (unsigned i = 0; i < T; i++)
    th.emplace_back([&, i] { parse_range(text + cut[i], text + cut[i + 1], weighted, cut[i], ch[i]); });
  for (auto &t : th) t.join();
  uint64_t E = 0;
  uint32_t max_id = 0;
  bool any = false;
  for (auto &c : ch) {
    if (!c.err.empty()) fail(GMP_EFORMAT, c.err + " of " + path);
    E += c.src.size();
    if (c.any) max_id = std::max(max_id, c.max_id), any = true;
  }
  if (V == 0) V = any ? max_id + 1 : 1;
  else if (any && max_id >= V) fail(GMP_EINVAL, "edge list: vertex id " + std::to_string(max_id) + " >= V");
  const uint64_t Et = sym ? 2 * E : E;
  // counting sort on the destination: raw in-edge counts, offsets, then the sources
  out.raw_indeg.assign(V, 0);
  for (auto &c : ch)
    for (size_t k = 0; k < c.src.size(); k++) {
      out.raw_indeg[c.dst[k]]++;
      if (sym) out.raw_indeg[c.src[k]]++;
    }
  std::vector<uint64_t> pos(V + 1, 0);
  for (uint32_t v = 0; v < V; v++) pos[v + 1] = pos[v] + out.raw_indeg[v];
  std::vector<uint32_t> col(Et), wv(weighted ? Et : 0);
  {
    std::vector<uint64_t> at(pos.begin(), pos.end() - 1);
    for (auto &c : ch) {
      for (size_t k = 0; k < c.src.size(); k++) {
        const uint64_t a = at[c.dst[k]]++;
        col[a] = c.src[k];
        if (weighted) wv[a] = c.w[k];
        if (sym) {
          const uint64_t b = at[c.src[k]]++;
          col[b] = c.dst[k];
          if (weighted) wv[b] = c.w[k];
        }
      }
      std::vector<uint32_t>().swap(c.src);
      std::vector<uint32_t>().swap(c.dst);
      std::vector<uint32_t>().swap(c.w);
    }
  }
  // every row: ascending sources, duplicates merged (the least weight stays)
  std::vector<uint64_t> kept(V, 0);
  std::atomic<uint64_t> next{0};
  const uint64_t rows_per = 4096;
  auto work = [&] {
    std::vector<std::pair<uint32_t, uint32_t>> tmp;
    for (;;) {
      const uint64_t r0 = next.fetch_add(rows_per);
      if (r0 >= V) break;
      const uint32_t r1 = (uint32_t)std::min<uint64_t>(V, r0 + rows_per);
      for (uint32_t r = (uint32_t)r0; r < r1; r++) {
        const uint64_t a = pos[r], b = pos[r + 1];
        uint64_t n = 0;
        if (weighted) {
          tmp.clear();
          for (uint64_t k = a; k < b; k++) tmp.emplace_back(col[k], wv[k]);
          std::sort(tmp.begin(), tmp.end());
          for (size_t k = 0; k < tmp.size(); k++)
            if (k == 0 || tmp[k].first != tmp[k - 1].first) {
              col[a + n] = tmp[k].first;
              wv[a + n] = tmp[k].second;
              n++;
            }
        } else {
          std::sort(col.begin() + a, col.begin() + b);
          n = (uint64_t)(std::unique(col.begin() + a, col.begin() + b) - (col.begin() + a));
        }
        kept[r] = n;
      }
    }
  };
  th.clear();
  for (unsigned i = 0; i < T; i++) th.emplace_back(work);
  for (auto &t : th) t.join();
  out.off.assign(V + 1, 0);
  for (uint32_t v = 0; v < V; v++) out.off[v + 1] = out.off[v] + kept[v];
  out.idx.resize(out.off[V]);
  out.w.resize(weighted ? out.off[V] : 0);
  for (uint32_t v = 0; v < V; v++) {
    std::memcpy(out.idx.data() + out.off[v], col.data() + pos[v], kept[v] * 4);
    if (weighted) std::memcpy(out.w.data() + out.off[v], wv.data() + pos[v], kept[v] * 4);
  }
}

void write_shard_files(const std::string &dir, uint32_t V, const EdgeCsr &g, uint32_t P, const uint32_t *bounds) {
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d.push_back('/');
  std::vector<int32_t> buf;
  for (uint32_t p = 0; p < P; p++) {
    const uint32_t s = bounds[p], e = bounds[p + 1];
    const uint64_t e0 = g.off[s], e1 = g.off[e], nnz = e1 - e0;
    uint64_t raw = 0;
    for (uint32_t r = s; r < e; r++) raw += g.raw_indeg[r];
    if (raw > 0x7fffffffull) fail(GMP_EUNSUPPORTED, "shard exceeds int32 edges");
    buf.resize(5 + nnz + (e - s) + 1);
    buf[0] = (int32_t)raw;
    buf[1] = (int32_t)nnz;
    buf[2] = (int32_t)(e - s + 1);
    buf[3] = (int32_t)s;
    buf[4] = (int32_t)e;
    std::memcpy(buf.data() + 5, g.idx.data() + e0, nnz * 4);
    for (uint32_t r = s; r <= e; r++) buf[5 + nnz + (r - s)] = (int32_t)(g.off[r] - e0);
    npy_write_i32(d + std::to_string(p) + ".edge.npy", buf.data(), buf.size());
    if (!g.w.empty()) npy_write_u32(d + std::to_string(p) + ".weight.npy", g.w.data() + e0, nnz);
  }
  std::vector<int32_t> od(V, 0);
  for (uint32_t c : g.idx) od[c]++;
  npy_write_i32(d + "vertexout.npy", od.data(), V);
}

}  // namespace gmp

extern "C" {

int gmp_convert_edge_list(const char *path, const char *dir, uint32_t V, uint32_t P, uint32_t flags,
                          uint32_t *V_out) {
  GMP_API_BEGIN
  if (!path || !dir) gmp::fail(GMP_EINVAL, "gmp_convert_edge_list: null path");
  if (P == 0) gmp::fail(GMP_EINVAL, "gmp_convert_edge_list: P must be > 0");
  if (V == 0xFFFFFFFFu) gmp::fail(GMP_EINVAL, "gmp_convert_edge_list: V must be below 2^32 - 1");
  gmp::EdgeCsr g;
  gmp::read_edge_list(path, flags, V, g);
  std::vector<uint32_t> bounds(P + 1);
  gmp::partition_bounds(g.raw_indeg.data(), V, P, bounds.data());
  gmp::write_shard_files(dir, V, g, P, bounds.data());
  if (V_out) *V_out = V;
  GMP_API_END
}

}  // extern "C"
