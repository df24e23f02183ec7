// Test harness (CPU only): runs the engine's host varint packer and unpacker on a CSR read from
// a file, for tests/test_window_host.py.
//   input  (argv[1]): uint32 V, uint64 E, uint64 off[V + 1], uint32 idx[E], uint64 e0, uint64 e1
//   output (argv[2]): uint64 nbytes, uint64 gbyte[G + 1], uint64 gmask[G], uint8 bytes[nbytes],
//                     uint32 ids[e1 - e0] (unpacked edges [e0, e1))
#include <cstdint>
#include <cstdio>
#include <vector>

namespace gmp {
void pack_host(const uint64_t *off, uint32_t V, const uint32_t *idx, uint64_t E, std::vector<uint64_t> &gbyte,
               std::vector<uint64_t> &gmask, uint8_t *bytes, uint64_t &nbytes, bool measure_only);
void unpack_host(const uint8_t *bytes, const uint64_t *gbyte, const uint64_t *gmask, uint64_t E, uint64_t e0,
                 uint64_t e1, uint32_t *out);
}  // namespace gmp

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f) return 3;
  uint32_t V = 0;
  uint64_t E = 0, e0 = 0, e1 = 0;
  if (std::fread(&V, 4, 1, f) != 1 || std::fread(&E, 8, 1, f) != 1) return 4;
  std::vector<uint64_t> off(V + 1);
  std::vector<uint32_t> idx(E);
  if (std::fread(off.data(), 8, V + 1, f) != V + 1) return 4;
  if (E && std::fread(idx.data(), 4, E, f) != E) return 4;
  if (std::fread(&e0, 8, 1, f) != 1 || std::fread(&e1, 8, 1, f) != 1) return 4;
  std::fclose(f);
  std::vector<uint64_t> gbyte, gmask;
  uint64_t nbytes = 0;
  gmp::pack_host(off.data(), V, idx.data(), E, gbyte, gmask, nullptr, nbytes, true);
  std::vector<uint8_t> bytes(nbytes + 64, 0);
  gmp::pack_host(off.data(), V, idx.data(), E, gbyte, gmask, bytes.data(), nbytes, false);
  std::vector<uint32_t> ids(e1 - e0);
  gmp::unpack_host(bytes.data(), gbyte.data(), gmask.data(), E, e0, e1, ids.data());
  FILE *o = std::fopen(argv[2], "wb");
  if (!o) return 5;
  const uint64_t G = gbyte.size() - 1;
  std::fwrite(&nbytes, 8, 1, o);
  std::fwrite(gbyte.data(), 8, G + 1, o);
  std::fwrite(gmask.data(), 8, G, o);
  std::fwrite(bytes.data(), 1, nbytes, o);
  std::fwrite(ids.data(), 4, ids.size(), o);
  std::fclose(o);
  return 0;
}
