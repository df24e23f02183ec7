// Edge list -> GraphMP shard directory, the native counterpart of GraphMP/tools/format_spark.py
// (no Spark, HDFS or GPU): reads "src dst [weight]" lines and writes DATA_DIR/{p}.edge.npy,
// DATA_DIR/vertexout.npy and, with --weighted, DATA_DIR/{p}.weight.npy.
//   usage: gmp_convert EDGES DATA_DIR SHARDS [VERTICES] [--weighted] [--symmetrize]
// VERTICES 0 or omitted: the largest id + 1.  The shards then load with gmp_load_shards or run
// under the GraphMP-style apps (gmp_pagerank DATA_DIR VERTICES SHARDS ...).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gmp.h"

int main(int argc, char *argv[]) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s EDGES DATA_DIR SHARDS [VERTICES] [--weighted] [--symmetrize]\n", argv[0]);
    return 2;
  }
  uint32_t flags = 0, V = 0;
  int pos = 0;
  for (int i = 4; i < argc; i++) {
    if (!std::strcmp(argv[i], "--weighted")) flags |= GMP_EDGES_WEIGHTED;
    else if (!std::strcmp(argv[i], "--symmetrize")) flags |= GMP_EDGES_SYMMETRIZE;
    else if (pos++ == 0) V = (uint32_t)std::strtoul(argv[i], nullptr, 10);
    else {
      std::fprintf(stderr, "unexpected argument %s\n", argv[i]);
      return 2;
    }
  }
  const uint32_t P = (uint32_t)std::strtoul(argv[3], nullptr, 10);
  uint32_t Vout = 0;
  if (gmp_convert_edge_list(argv[1], argv[2], V, P, flags, &Vout) != GMP_OK) {
    std::fprintf(stderr, "error: %s\n", gmp_last_error());
    return 1;
  }
  std::printf("%u vertices in %u shards written to %s\n", Vout, P, argv[2]);
  return 0;
}
