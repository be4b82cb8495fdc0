// Writes philox_vectors.bin: 256 (counter[4], key[2], output[4]) uint32 records
// computed by rocRAND's Philox4x32-10 (rocrand_philox4x32_10.h, MIT licence,
// after D. E. Shaw Research's Random123), used only to pin the project's own
// Philox implementation.  Host code; build: see make_philox_vectors.sh.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdint>
#include <cstdio>

struct Probe : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main(int argc, char** argv) {
  FILE* f = std::fopen(argc > 1 ? argv[1] : "philox_vectors.bin", "wb");
  if (!f) return 1;
  Probe p;
  uint64_t x = 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(x >> 32);
  };
  for (int i = 0; i < 256; ++i) {
    uint4 c;
    uint2 k;
    if (i == 0) { c = {0, 0, 0, 0}; k = {0, 0}; }
    else if (i == 1) { c = {~0u, ~0u, ~0u, ~0u}; k = {~0u, ~0u}; }
    else { c = {next(), next(), next(), next()}; k = {next(), next()}; }
    const uint4 o = p.run(c, k);
    const uint32_t rec[10] = {c.x, c.y, c.z, c.w, k.x, k.y, o.x, o.y, o.z, o.w};
    std::fwrite(rec, sizeof rec, 1, f);
  }
  std::fclose(f);
  return 0;
}
