// Gather planning (host): cuts the destination rows into wave-sized tasks.
//
// GraphMP parallelises one shard per OpenMP thread (GraphMP/system/GraphPS.h:266-278); on
// gfx950 the unit is a 64-lane wavefront, so rows are grouped by in-degree instead:
//   * consecutive low-degree rows are packed into STREAM tasks holding at most stream_cap
//     edges and row_cap rows: the wave streams their column ids coalesced, stages the
//     gathered values in LDS and reduces one row per lane;
//   * a row with more than vec_min edges becomes a hub split into CHUNK tasks of
//     `chunk` edges, each reduced by the whole wave with cross-lane reductions; rows of
//     one chunk are finished by that wave, longer rows leave per-chunk partials that a
//     second kernel combines in chunk order.
// Every task therefore touches a bounded number of edges, so a static grid-stride
// assignment of tasks to waves stays balanced and every reduction has a fixed order.
#include <algorithm>

#include "gmp_internal.hpp"

namespace gmp {

void build_gather_plan_host(const uint64_t *off, uint32_t V, GatherPlan &plan,
                            std::vector<uint32_t> &tasks, std::vector<HubRow> &hubs,
                            std::vector<uint32_t> &multi) {
  tasks.clear();
  hubs.clear();
  multi.clear();
  const uint64_t cap = plan.stream_cap, rcap = plan.row_cap, chunk = plan.chunk;
  // rows above vec_min edges are reduced by a whole wave; below, by one lane from LDS
  const uint64_t vmin = std::min<uint64_t>(plan.vec_min, cap);
  uint32_t parts = 0;
  uint32_t r0 = 0, nr = 0;
  uint64_t acc = 0;
  auto close_stream = [&]() {
    if (nr) {
      tasks.push_back(r0);
      tasks.push_back(nr);
    }
    nr = 0;
    acc = 0;
  };
  for (uint32_t r = 0; r < V; r++) {
    uint64_t deg = off[r + 1] - off[r];
    if (deg > vmin) {
      close_stream();
      uint64_t nch = (deg + chunk - 1) / chunk;
      if (nch > 0x7fffffffull) fail(GMP_EUNSUPPORTED, "row too long for the gather plan");
      uint32_t h = (uint32_t)hubs.size();
      HubRow hr{r, (uint32_t)nch, nch > 1 ? parts : 0u, 0u};
      if (nch > 1) {
        parts += (uint32_t)nch;
        multi.push_back(h);
      }
      hubs.push_back(hr);
      for (uint64_t k = 0; k < nch; k++) {
        tasks.push_back(h);
        tasks.push_back(kTaskChunkBit | (uint32_t)k);
      }
      r0 = r + 1;
      continue;
    }
    if (nr == 0) r0 = r;
    if (nr > 0 && (acc + deg > cap || nr >= rcap)) {
      close_stream();
      r0 = r;
    }
    nr++;
    acc += deg;
  }
  close_stream();
  plan.ntasks = (uint32_t)(tasks.size() / 2);
  plan.nhubs = (uint32_t)hubs.size();
  plan.nmulti = (uint32_t)multi.size();
  plan.nparts = parts;
}

// (Re)builds `pl` over `nrows` rows with offsets `off` when the plan options changed.
void plan_for(Engine &e, GatherPlan &pl, const uint64_t *off, uint32_t nrows) {
  if (pl.built && pl.stream_cap == e.opt.stream_cap && pl.row_cap == e.opt.row_cap && pl.chunk == e.opt.chunk &&
      pl.vec_min == e.opt.vec_min)
    return;
  pl.stream_cap = e.opt.stream_cap;
  pl.row_cap = e.opt.row_cap;
  pl.chunk = e.opt.chunk;
  pl.vec_min = e.opt.vec_min;
  std::vector<uint32_t> tasks, multi;
  std::vector<HubRow> hubs;
  build_gather_plan_host(off, nrows, pl, tasks, hubs, multi);
  pl.tasks.alloc(std::max<size_t>(tasks.size(), 2) * 4);
  pl.hubs.alloc(std::max<size_t>(hubs.size(), 1) * sizeof(HubRow));
  pl.multi.alloc(std::max<size_t>(multi.size(), 1) * 4);
  if (!tasks.empty())
    GMP_HIP(hipMemcpyAsync(pl.tasks.p, tasks.data(), tasks.size() * 4, hipMemcpyHostToDevice, e.stream));
  if (!hubs.empty())
    GMP_HIP(hipMemcpyAsync(pl.hubs.p, hubs.data(), hubs.size() * sizeof(HubRow), hipMemcpyHostToDevice, e.stream));
  if (!multi.empty())
    GMP_HIP(hipMemcpyAsync(pl.multi.p, multi.data(), multi.size() * 4, hipMemcpyHostToDevice, e.stream));
  GMP_HIP(hipStreamSynchronize(e.stream));
  pl.built = true;
}

void ensure_plan(Engine &e) {
  ensure_work(e);
  plan_for(e, e.g.plan, e.g.work_h_off().data(), e.g.V);
}

}  // namespace gmp
