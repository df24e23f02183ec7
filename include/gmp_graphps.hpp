// gmp_graphps.hpp -- header-only C++ facade over libgmp.so with the shape of GraphMP's engine
// class GraphPS<T> (cap-ntu/Graphee GraphMP/system/GraphPS.h:22-65), so GraphMP apps move over
// by changing their include, their base class and the line that installs the update rule.
//
//   reference (GraphMP/app/pagerank.cpp:68-84)        this facade
//   #include "../system/GraphPS.h"                     #include <gmp_graphps.hpp>
//   class PagerankPS : public GraphPS<T>               class PagerankPS : public gmp::GraphPS<T>
//   this->_comp = comp_pagerank<T>;                    this->_app = GMP_APP_PAGERANK;
//   init_vertex(), init(path, V, P, iters), run()      unchanged
//
// The per-shard host function pointer _comp (GraphPS.h:25-34) cannot run on the GPU, so the
// update rule is chosen by app id; each id reproduces the reference's rule bit for bit
// (GMP_APP_* in gmp.h).  init() loads every shard into HBM once (GraphMP re-reads shards each
// iteration, GraphMP/system/Global.h:129-255); run() calls init_vertex() on the host like the
// reference (GraphPS.h:211), runs the synchronous loop on the GPU and fills _VertexMsg,
// _VertexMsgNew and _VertexValue back.  Errors throw gmp::GraphError carrying the C ABI
// status code (the reference asserts or aborts instead, cnpy.h:130-133).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gmp.h"

namespace gmp {

struct GraphError : std::runtime_error {
  int code;
  GraphError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char *what) {
  if (rc != GMP_OK) throw GraphError(rc, std::string(what) + ": " + gmp_last_error());
}

// RAII handle on one engine (one GPU).
class Engine {
 public:
  explicit Engine(int32_t device = 0) { check(gmp_create(device, &h_), "gmp_create"); }
  ~Engine() { gmp_destroy(h_); }
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;
  gmp_engine *handle() const { return h_; }

  void load_shards(const std::string &dir, uint32_t V, uint32_t P) {
    check(gmp_load_shards(h_, dir.c_str(), V, P), "gmp_load_shards");
  }
  void load_csr(const std::vector<uint64_t> &off, const std::vector<uint32_t> &idx,
                const std::vector<uint32_t> *w = nullptr, uint32_t P = 1) {
    check(gmp_load_csr(h_, (uint32_t)(off.size() - 1), off.back(), off.data(), idx.data(), w ? w->data() : nullptr,
                       P, nullptr),
          "gmp_load_csr");
  }
  void generate_rmat(const gmp_rmat_params &p) { check(gmp_generate_rmat(h_, &p), "gmp_generate_rmat"); }
  uint32_t vertices() const {
    uint32_t V = 0;
    check(gmp_graph_info(h_, &V, nullptr, nullptr, nullptr), "gmp_graph_info");
    return V;
  }
  std::vector<float> pagerank(const gmp_pr_params &p, gmp_stats *st = nullptr) {
    std::vector<float> r(vertices());
    check(gmp_pagerank(h_, &p, r.data(), st), "gmp_pagerank");
    return r;
  }
  std::vector<uint32_t> sssp(uint32_t source, const gmp_sssp_params &p, gmp_stats *st = nullptr) {
    std::vector<uint32_t> d(vertices());
    check(gmp_sssp(h_, source, &p, d.data(), st), "gmp_sssp");
    return d;
  }
  std::vector<uint32_t> cc(gmp_stats *st = nullptr) {
    std::vector<uint32_t> l(vertices());
    check(gmp_cc(h_, l.data(), st), "gmp_cc");
    return l;
  }

 private:
  gmp_engine *h_ = nullptr;
};

// GraphMP's init_workers()/finalize_workers() (GraphMP/system/Global.h:336-362, :274-294) set up
// MPI and ZeroMQ; one GPU needs neither, so they only exist to keep app mains unchanged.
inline void init_workers() {}
inline void finalize_workers() {}

template <class T>
class GraphPS {
  static_assert(sizeof(T) == 4, "GraphMP apps use 4-byte vertex values (float or int32_t)");

 public:
  int32_t _app = 0;  // GMP_APP_*: the update rule (replaces _comp, GraphPS.h:25-34)
  std::string _DataPath;
  int32_t _ThreadNum = 1;
  int32_t _VertexNum = 0;
  int32_t _PartitionNum = 0;
  int32_t _MaxIteration = 0;
  std::vector<T> _VertexValue;
  std::vector<int32_t> _VertexOut;
  std::vector<int32_t> _VertexIn;
  std::vector<T> _VertexMsg;
  std::vector<T> _VertexMsgNew;
  gmp_stats _Stats{};  // iterations, device time, changed count of the last step

  GraphPS() = default;
  virtual ~GraphPS() = default;

  // GraphPS.h:77-151: remembers the run parameters; here it also moves every shard into HBM.
  void init(std::string DataPath, const int32_t VertexNum, const int32_t PartitionNum,
            const int32_t MaxIteration = 10) {
    _DataPath = DataPath;
    if (!_DataPath.empty() && _DataPath.back() != '/') _DataPath.push_back('/');
    _VertexNum = VertexNum;
    _PartitionNum = PartitionNum;
    _MaxIteration = MaxIteration;
    _VertexMsgNew.assign(_VertexNum, 0);
    if (!engine_) engine_.reset(new Engine(0));
    engine_->load_shards(_DataPath, (uint32_t)_VertexNum, (uint32_t)_PartitionNum);
  }

  virtual void init_vertex() = 0;
  void set_threadnum(const int32_t ThreadNum) { _ThreadNum = ThreadNum; }

  // GraphPS.h:171-387 with the app's rule on the GPU.
  void run() {
    if (!engine_) throw GraphError(GMP_ESTATE, "run() before init()");
    init_vertex();
    if ((int32_t)_VertexMsg.size() != _VertexNum) throw GraphError(GMP_EINVAL, "init_vertex() must size _VertexMsg to V");
    std::vector<T> msg(_VertexNum), value(_VertexNum);
    const int32_t *vout = _app == GMP_APP_PAGERANK ? _VertexOut.data() : nullptr;
    if (_app == GMP_APP_PAGERANK && (int32_t)_VertexOut.size() != _VertexNum)
      throw GraphError(GMP_EINVAL, "pagerank needs _VertexOut (load_vertex_out() in init_vertex())");
    check(gmp_run_app(engine_->handle(), _app, _MaxIteration, _VertexMsg.data(), vout, msg.data(), value.data(),
                      &_Stats),
          "gmp_run_app");
    _VertexMsg.swap(msg);
    _VertexMsgNew = _VertexMsg;
    _VertexValue.swap(value);
  }

  // GraphPS.h:153-169: int32 per-vertex degree files written next to the shards.
  void load_vertex_in() { load_i32(_DataPath + "vertexin.npy", _VertexIn); }
  void load_vertex_out() { load_i32(_DataPath + "vertexout.npy", _VertexOut); }

 protected:
  std::unique_ptr<Engine> engine_;

 private:
  void load_i32(const std::string &path, std::vector<int32_t> &out) {
    uint64_t n = 0;
    check(gmp_npy_read_u32(path.c_str(), nullptr, 0, &n), "gmp_npy_read_u32");
    if (n < (uint64_t)_VertexNum) throw GraphError(GMP_EFORMAT, path + ": fewer entries than vertices");
    std::vector<uint32_t> tmp(n);
    check(gmp_npy_read_u32(path.c_str(), tmp.data(), n, &n), "gmp_npy_read_u32");
    out.assign(tmp.begin(), tmp.begin() + _VertexNum);
  }
};

}  // namespace gmp
