// Window visiting order of a semi-external pass (plain C++, shared with the host tests).
#pragma once

#include <cstdint>
#include <vector>

namespace gmp {

// The windows 0..n-1 (cached[w]: window w stays in device memory) in the order a pass visits
// them: the cached and the streamed windows each keep their index order, and the streamed ones
// are spread evenly over the pass -- among the first k windows visited, floor(k * streamed / n)
// are streamed -- so the copy engines fetch the next streamed window while cached windows run.
// The first window visited is a cached one whenever any is.
inline std::vector<uint32_t> interleave_windows(const std::vector<bool> &cached) {
  const uint64_t n = cached.size();
  std::vector<uint32_t> cw, sw, order;
  order.reserve(n);
  for (uint32_t w = 0; w < n; w++) (cached[w] ? cw : sw).push_back(w);
  size_t ic = 0, is = 0;
  while (ic < cw.size() || is < sw.size()) {
    const bool take_s = is < sw.size() && (ic == cw.size() || (is + 1) * n <= (ic + is + 1) * sw.size());
    order.push_back(take_s ? sw[is++] : cw[ic++]);
  }
  return order;
}

}  // namespace gmp
