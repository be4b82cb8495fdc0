// Host-only check of the strategy-file functions (include/mccfr/strategy.hpp)
// against the reference's behaviour (poker_cuda/src/strategy.cpp:35-181):
// save -> load round trip, placeholder rows skipped, checkpoint / v1 / foreign
// files rejected, table-size mismatch rejected, top-action printing.  Makes no
// device call.  Exit code 0 on success; the failing check's number otherwise.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "mccfr/strategy.hpp"

static bool write_raw(const std::string& path, uint64_t magic, uint32_t ts, uint32_t na, uint32_t n,
                      const uint32_t* keys, const float* probs) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::fwrite(&magic, 8, 1, f);
  std::fwrite(&ts, 4, 1, f);
  std::fwrite(&na, 4, 1, f);
  std::fwrite(&n, 4, 1, f);
  for (uint32_t i = 0; i < n; ++i) {
    std::fwrite(&keys[i], 4, 1, f);
    std::fwrite(probs + (size_t)i * na, 4, na, f);
  }
  std::fclose(f);
  return true;
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  HostStrategyTable a;
  for (uint32_t k : {7u, 3u, 123456u, (uint32_t)GPU_TABLE_SIZE - 1}) {
    StrategyEntry e;
    for (int i = 0; i < GPU_NUM_ACTIONS; ++i) e.probs[i] = (float)((k + 3 * i) % 11) / 40.f;
    a.emplace(k, e);
  }
  const std::string p = dir + "/s.bin";
  if (!strategy_save(a, p)) return 1;
  HostStrategyTable b;
  if (!strategy_load(b, p) || b.size() != a.size()) return 2;
  for (const auto& kv : a) {
    auto it = b.find(kv.first);
    if (it == b.end() || std::memcmp(it->second.probs, kv.second.probs, sizeof kv.second.probs) != 0) return 3;
  }
  // strategy_load_auto on a strategy file reads it directly (no trainer made)
  HostStrategyTable c;
  if (!strategy_load_auto(c, p, 2, 1000, 10, 20) || c.size() != a.size()) return 4;

  // a uniform placeholder row (within 1e-6 of 1/8) is skipped, a near-uniform one kept
  uint32_t keys[3] = {1, 2, 3};
  float probs[24];
  for (int i = 0; i < 24; ++i) probs[i] = 0.125f;
  probs[1] = 0.125f + 5e-7f;       // still a placeholder
  probs[8 + 0] = 0.125f + 1e-4f;   // a real row
  probs[16 + 2] = 0.0f;            // a real row
  const std::string q = dir + "/placeholders.bin";
  if (!write_raw(q, 0x53545241540100ull, GPU_TABLE_SIZE, GPU_NUM_ACTIONS, 3, keys, probs)) return 5;
  HostStrategyTable d;
  if (!strategy_load(d, q) || d.size() != 2 || d.count(1) || !d.count(2) || !d.count(3)) return 6;

  // rejected: checkpoint magic, v1 magic, foreign magic, wrong table size, truncated
  const std::string r = dir + "/bad.bin";
  HostStrategyTable e;
  write_raw(r, 0x43465247505500ull, GPU_TABLE_SIZE, GPU_NUM_ACTIONS, 0, keys, probs);
  if (strategy_load(e, r)) return 7;
  write_raw(r, 0x53545241540000ull, GPU_TABLE_SIZE, GPU_NUM_ACTIONS, 0, keys, probs);
  if (strategy_load(e, r)) return 8;
  write_raw(r, 0x1234ull, GPU_TABLE_SIZE, GPU_NUM_ACTIONS, 0, keys, probs);
  if (strategy_load(e, r) || strategy_load_auto(e, r, 2, 1000, 10, 20)) return 9;
  write_raw(r, 0x53545241540100ull, 1u << 20, GPU_NUM_ACTIONS, 1, keys + 1, probs + 8);
  if (strategy_load(e, r)) return 10;
  // header says 3 entries, the file holds 1
  write_raw(r, 0x53545241540100ull, GPU_TABLE_SIZE, GPU_NUM_ACTIONS, 1, keys + 1, probs + 8);
  std::FILE* f = std::fopen(r.c_str(), "r+b");
  if (!f) return 11;
  const uint32_t three = 3;
  if (std::fseek(f, 16, SEEK_SET) != 0 || std::fwrite(&three, 4, 1, f) != 1) return 12;
  std::fclose(f);
  if (strategy_load(e, r)) return 13;
  if (strategy_load(e, dir + "/missing.bin")) return 14;

  // top actions: probabilities 0.5 / 0.3 / 0.2 on raise_pot / call / fold
  HostStrategyTable t;
  StrategyEntry te;
  te.probs[6] = 0.5f;
  te.probs[2] = 0.3f;
  te.probs[0] = 0.2f;
  t.emplace(42, te);
  strategy_print_top(t, 42, 3);
  strategy_print_top(t, 43, 3);
  return 0;
}
