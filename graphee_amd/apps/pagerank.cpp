// PageRank as a GraphMP app on the gfx950 engine.  The class is GraphMP/app/pagerank.cpp:68-84
// moved onto gmp::GraphPS: same init_vertex(), the update rule selected by id instead of the
// _comp pointer.  main() takes the dataset from the command line instead of the literals of
// pagerank.cpp:94-98.
//   usage: gmp_pagerank DATA_DIR VERTICES SHARDS MAX_ITER [OUT.npy]
// OUT.npy receives _VertexMsg (float32 bits stored as int32).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gmp_graphps.hpp"

template <class T>
class PagerankPS : public gmp::GraphPS<T> {
 public:
  PagerankPS() : gmp::GraphPS<T>() { this->_app = GMP_APP_PAGERANK; }
  void init_vertex() override {
    this->load_vertex_out();
    for (int32_t i = 0; i < this->_VertexNum; i++)
      if (this->_VertexOut[i] == 0) this->_VertexOut[i] = 1;
    this->_VertexMsg.assign(this->_VertexNum, 1.0 / this->_VertexNum);
    this->_VertexValue.assign(this->_VertexNum, 0);
  }
};

int main(int argc, char *argv[]) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s DATA_DIR VERTICES SHARDS MAX_ITER [OUT.npy]\n", argv[0]);
    return 2;
  }
  try {
    gmp::init_workers();
    PagerankPS<float> pg;
    pg.init(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
    pg.run();
    std::printf("iterations %d, device time %.3f ms\n", pg._Stats.iterations, pg._Stats.ms_total);
    if (argc > 5) {
      std::vector<int32_t> bits(pg._VertexMsg.size());
      std::memcpy(bits.data(), pg._VertexMsg.data(), bits.size() * 4);
      gmp::check(gmp_npy_write_i32(argv[5], bits.data(), bits.size()), "gmp_npy_write_i32");
    }
    gmp::finalize_workers();
  } catch (const gmp::GraphError &e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
