// Host-side shard formats: .npy container, destination-interval shards, partitioner.
//
// The on-disk layout is GraphMP's: one file {p}.edge.npy per destination interval holding
// int32[5 + nnz + (n+1)] = [edge_num_raw, len(indices), len(indptr), start_id, end_id,
// indices..., indptr...] (GraphMP/tools/format_spark.py:137-157, read back by
// GraphMP/system/Comp.h:35-42), inside a 1-D little-endian .npy container whose header is
// parsed under cnpy's rules (GraphMP/system/cnpy.h:73-107).  The engine reads every shard
// once and keeps the whole graph in HBM.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gmp_internal.hpp"

namespace gmp {

namespace {

struct File {
  FILE *f = nullptr;
  File(const std::string &path, const char *mode) : f(std::fopen(path.c_str(), mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// Parses the header dictionary the way cnpy.h:73-107 does: 'fortran_order' value,
// the tuple between '(' and ')' (1-D only), and 'descr' = byte order + kind + word size.
void parse_header(const std::string &path, const std::string &h, uint64_t &count,
                  uint32_t &word, char &kind) {
  size_t fo = h.find("fortran_order");
  size_t sh0 = h.find('('), sh1 = h.find(')');
  size_t de = h.find("descr");
  if (fo == std::string::npos || sh0 == std::string::npos || sh1 == std::string::npos ||
      sh1 < sh0 || de == std::string::npos)
    fail(GMP_EFORMAT, path + ": malformed .npy header");
  if (h.compare(fo + 16, 4, "True") == 0) fail(GMP_EFORMAT, path + ": fortran_order arrays are not supported");
  std::string shape = h.substr(sh0 + 1, sh1 - sh0 - 1);
  while (!shape.empty() && shape.back() == ' ') shape.pop_back();
  size_t ncomma = 0;
  for (char c : shape) ncomma += (c == ',');
  size_t ndims = (!shape.empty() && shape.back() == ',') ? 1 : ncomma + 1;  // cnpy.h:88-89
  if (ndims != 1 || shape.empty()) fail(GMP_EFORMAT, path + ": only 1-D arrays are supported (cnpy.h:90)");
  char *endp = nullptr;
  errno = 0;
  unsigned long long n = std::strtoull(shape.c_str(), &endp, 10);
  if (errno || endp == shape.c_str()) fail(GMP_EFORMAT, path + ": bad shape '" + shape + "'");
  count = n;
  // descr value starts 9 characters after the key: descr': '<i4 (cnpy.h:99)
  size_t q = h.find('\'', de + 6);
  if (q == std::string::npos || q + 3 >= h.size()) fail(GMP_EFORMAT, path + ": bad descr");
  char order = h[q + 1];
  if (order != '<' && order != '|') fail(GMP_EFORMAT, path + ": big-endian .npy is not supported (cnpy.h:100-101)");
  kind = h[q + 2];
  word = (uint32_t)std::strtoul(h.c_str() + q + 3, nullptr, 10);
}

void write_npy(const std::string &path, const void *data, uint64_t n, const char *descr) {
  char dict[160];
  std::snprintf(dict, sizeof dict, "{'descr': '%s', 'fortran_order': False, 'shape': (%llu,), }",
                descr, (unsigned long long)n);
  std::string hdr(dict);
  size_t total = 10 + hdr.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  hdr.append(pad, ' ');
  hdr.push_back('\n');
  if (hdr.size() > 65535) fail(GMP_EFORMAT, path + ": header too long");
  File f(path, "wb");
  if (!f.f) fail(GMP_EIO, path + ": cannot open for writing: " + std::strerror(errno));
  unsigned char pre[10] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, 0, 0};
  pre[8] = (unsigned char)(hdr.size() & 0xff);
  pre[9] = (unsigned char)(hdr.size() >> 8);
  if (std::fwrite(pre, 1, 10, f.f) != 10 || std::fwrite(hdr.data(), 1, hdr.size(), f.f) != hdr.size() ||
      (n && std::fwrite(data, 4, n, f.f) != n))
    fail(GMP_EIO, path + ": short write");
}

}  // namespace

void npy_read_u32(const std::string &path, std::vector<uint32_t> &out) {
  File f(path, "rb");
  if (!f.f) fail(GMP_EIO, path + ": cannot open: " + std::strerror(errno));
  unsigned char pre[10];
  if (std::fread(pre, 1, 10, f.f) != 10) fail(GMP_EFORMAT, path + ": truncated .npy preamble");
  if (std::memcmp(pre, "\x93NUMPY", 6) != 0) fail(GMP_EFORMAT, path + ": not a .npy file");
  uint64_t hlen;
  if (pre[6] == 1) {
    hlen = pre[8] | (pre[9] << 8);
  } else if (pre[6] == 2 || pre[6] == 3) {
    unsigned char ext[2];
    if (std::fread(ext, 1, 2, f.f) != 2) fail(GMP_EFORMAT, path + ": truncated .npy preamble");
    hlen = pre[8] | (pre[9] << 8) | ((uint64_t)ext[0] << 16) | ((uint64_t)ext[1] << 24);
  } else {
    fail(GMP_EFORMAT, path + ": unsupported .npy version");
  }
  std::string h(hlen, '\0');
  if (hlen && std::fread(&h[0], 1, hlen, f.f) != hlen) fail(GMP_EFORMAT, path + ": truncated .npy header");
  if (h.empty() || h.back() != '\n') fail(GMP_EFORMAT, path + ": header does not end in newline (cnpy.h:79)");
  uint64_t n = 0;
  uint32_t word = 0;
  char kind = 0;
  parse_header(path, h, n, word, kind);
  if (word != 4 || (kind != 'i' && kind != 'u'))
    fail(GMP_EFORMAT, path + ": expected 4-byte integers, got kind '" + std::string(1, kind) +
                          "' size " + std::to_string(word));
  out.resize(n);
  if (n && std::fread(out.data(), 4, n, f.f) != n) fail(GMP_EIO, path + ": short read (cnpy.h:122-124)");
}

void npy_write_i32(const std::string &path, const int32_t *data, uint64_t n) {
  write_npy(path, data, n, "<i4");
}
void npy_write_u32(const std::string &path, const uint32_t *data, uint64_t n) {
  write_npy(path, data, n, "<u4");
}

// format_spark.py:91-108: boundary i = first prefix index whose in-edge count reaches
// i * floor(E / P) (the script is Python 2, where integer '/' floors).
void partition_bounds(const uint64_t *indeg, uint32_t V, uint32_t P, uint32_t *bounds) {
  if (P == 0) fail(GMP_EINVAL, "partition_bounds: P must be > 0");
  uint64_t E = 0;
  for (uint32_t v = 0; v < V; v++) E += indeg[v];
  uint64_t epp = E / P, cum = 0;
  uint32_t j = 0;
  bounds[0] = 0;
  for (uint32_t i = 1; i < P; i++) {
    uint64_t target = epp * i;
    while (cum < target) cum += indeg[j++];
    bounds[i] = j;
  }
  bounds[P] = V;
}

}  // namespace gmp

// ---- C ABI: host-side format utilities ------------------------------------------------

namespace {
thread_local std::string g_err;
}

namespace gmp {
void set_last_error(const std::string &m) { g_err = m; }
const char *last_error() { return g_err.c_str(); }
}  // namespace gmp

extern "C" {

const char *gmp_last_error(void) { return gmp::last_error(); }
int gmp_abi_version(void) { return GMP_ABI_VERSION; }

int gmp_partition_bounds(const uint64_t *indeg, uint32_t V, uint32_t P, uint32_t *bounds) {
  GMP_API_BEGIN
  if ((!indeg && V) || !bounds) gmp::fail(GMP_EINVAL, "gmp_partition_bounds: null pointer");
  gmp::partition_bounds(indeg, V, P, bounds);
  GMP_API_END
}

int gmp_npy_read_u32(const char *path, uint32_t *buf, uint64_t cap, uint64_t *n_out) {
  GMP_API_BEGIN
  if (!path || !n_out) gmp::fail(GMP_EINVAL, "gmp_npy_read_u32: null pointer");
  std::vector<uint32_t> v;
  gmp::npy_read_u32(path, v);
  *n_out = v.size();
  if (buf) {
    if (cap < v.size()) gmp::fail(GMP_EINVAL, "gmp_npy_read_u32: buffer too small");
    std::memcpy(buf, v.data(), v.size() * 4);
  }
  GMP_API_END
}

int gmp_npy_write_i32(const char *path, const int32_t *data, uint64_t n) {
  GMP_API_BEGIN
  if (!path || (!data && n)) gmp::fail(GMP_EINVAL, "gmp_npy_write_i32: null pointer");
  gmp::npy_write_i32(path, data, n);
  GMP_API_END
}

int gmp_write_shards(const char *dir, uint32_t V, const uint64_t *row_off,
                     const uint32_t *col_idx, const uint32_t *weights, uint32_t P,
                     const uint32_t *bounds, const int32_t *vertex_out) {
  GMP_API_BEGIN
  if (!dir || !row_off || !bounds || (!col_idx && row_off[V]))
    gmp::fail(GMP_EINVAL, "gmp_write_shards: null pointer");
  if (P == 0 || bounds[0] != 0 || bounds[P] != V)
    gmp::fail(GMP_EINVAL, "gmp_write_shards: bounds must run from 0 to V");
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d.push_back('/');
  std::vector<int32_t> buf;
  for (uint32_t p = 0; p < P; p++) {
    uint32_t s = bounds[p], e = bounds[p + 1];
    if (e < s) gmp::fail(GMP_EINVAL, "gmp_write_shards: bounds must be nondecreasing");
    uint64_t e0 = row_off[s], e1 = row_off[e], nnz = e1 - e0;
    if (nnz > 0x7fffffffull) gmp::fail(GMP_EUNSUPPORTED, "gmp_write_shards: shard exceeds int32 edges");
    buf.resize(5 + nnz + (e - s) + 1);
    buf[0] = (int32_t)nnz;
    buf[1] = (int32_t)nnz;
    buf[2] = (int32_t)(e - s + 1);
    buf[3] = (int32_t)s;
    buf[4] = (int32_t)e;
    std::memcpy(buf.data() + 5, col_idx + e0, nnz * 4);
    for (uint32_t r = s; r <= e; r++) buf[5 + nnz + (r - s)] = (int32_t)(row_off[r] - e0);
    gmp::npy_write_i32(d + std::to_string(p) + ".edge.npy", buf.data(), buf.size());
    if (weights) gmp::npy_write_u32(d + std::to_string(p) + ".weight.npy", weights + e0, nnz);
  }
  if (vertex_out) gmp::npy_write_i32(d + "vertexout.npy", vertex_out, V);
  GMP_API_END
}

}  // extern "C"
