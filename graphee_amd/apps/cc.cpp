// Connected components (min label) as a GraphMP app on the gfx950 engine:
// GraphMP/app/cc.cpp:62-94 on gmp::GraphPS.
//   usage: gmp_cc DATA_DIR VERTICES SHARDS MAX_ITER [OUT.npy]
#include <cstdio>
#include <cstdlib>

#include "gmp_graphps.hpp"

template <class T>
class CcPS : public gmp::GraphPS<T> {
 public:
  CcPS() : gmp::GraphPS<T>() { this->_app = GMP_APP_CC; }
  void init_vertex() override {
    this->_VertexValue.assign(this->_VertexNum, 0);
    this->_VertexMsg.assign(this->_VertexNum, 0);
    for (int32_t i = 0; i < this->_VertexNum; i++) this->_VertexMsg[i] = i;
  }
};

int main(int argc, char *argv[]) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s DATA_DIR VERTICES SHARDS MAX_ITER [OUT.npy]\n", argv[0]);
    return 2;
  }
  try {
    gmp::init_workers();
    CcPS<int32_t> pg;
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
