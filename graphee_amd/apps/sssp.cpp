// SSSP as a GraphMP app on the gfx950 engine: GraphMP/app/sssp.cpp:63-94 on gmp::GraphPS
// (unit weights, GPS_INF = 10000 as in GraphMP/system/Global.h:50).  The source is a command
// line argument instead of the hard-coded vertex 1 of sssp.cpp:72 (default 1).
//   usage: gmp_sssp DATA_DIR VERTICES SHARDS MAX_ITER [OUT.npy] [SOURCE]
#include <cstdio>
#include <cstdlib>

#include "gmp_graphps.hpp"

static const int32_t kGpsInf = 10000;

template <class T>
class SsspPS : public gmp::GraphPS<T> {
 public:
  int32_t source = 1;
  SsspPS() : gmp::GraphPS<T>() { this->_app = GMP_APP_SSSP; }
  void init_vertex() override {
    this->_VertexValue.assign(this->_VertexNum, 0);
    this->_VertexMsg.assign(this->_VertexNum, kGpsInf);
    this->_VertexMsg[source] = 0;
  }
};

int main(int argc, char *argv[]) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s DATA_DIR VERTICES SHARDS MAX_ITER [OUT.npy] [SOURCE]\n", argv[0]);
    return 2;
  }
  try {
    gmp::init_workers();
    SsspPS<int32_t> pg;
    if (argc > 6) pg.source = std::atoi(argv[6]);
    pg.init(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    pg.run();
    std::printf("iterations %d, device time %.3f ms\n", pg._Stats.iterations, pg._Stats.ms_total);
    if (argc > 5) gmp::check(gmp_npy_write_i32(argv[5], pg._VertexMsg.data(), pg._VertexMsg.size()), "write");
    gmp::finalize_workers();
  } catch (const gmp::GraphError &e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
