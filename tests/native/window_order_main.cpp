// Prints the window order of interleave_windows for a pattern given as a string of 'c'
// (cached) and 's' (streamed) windows.
#include <cstdio>
#include <cstring>

#include "window_order.hpp"

int main(int argc, char **argv) {
  for (int i = 1; i < argc; i++) {
    std::vector<bool> cached;
    for (const char *p = argv[i]; *p; p++) cached.push_back(*p == 'c');
    const std::vector<uint32_t> o = gmp::interleave_windows(cached);
    for (size_t k = 0; k < o.size(); k++) std::printf(k ? " %u" : "%u", o[k]);
    std::printf("\n");
  }
  return 0;
}
