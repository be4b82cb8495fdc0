// strategy.hpp — the reference's strategy-file functions (poker_cuda/src/strategy.h:5-9)
// over the C ABI, for callers such as the reference CLI (main.cpp:366-377) and
// its evaluation and live-play tools.
//
// File layout (.bin v2, strategy.cpp:6-11): u64 magic 0x53545241540100,
// u32 table_size, u32 num_actions, u32 entries, then per entry u32 key and
// f32 probs[num_actions].  Behaviour follows the reference: strategy_load
// rejects checkpoints, v1 exports and files built for another table size or
// action count (false + a message on stderr) and skips uniform 1/8
// placeholder rows; strategy_load_auto converts a .ckpt through the trainer
// and writes the converted strategy next to it (`x.ckpt` -> `x`, else `x.bin`).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "gpu_cfr_trainer.hpp"
#include "mccfr.h"

// strategy.cpp:35-52.  Rows are written in ascending key order.
inline bool strategy_save(const HostStrategyTable& strat, const std::string& path) {
  std::vector<uint32_t> keys;
  keys.reserve(strat.size());
  for (const auto& kv : strat) keys.push_back(kv.first);
  std::sort(keys.begin(), keys.end());
  std::vector<float> probs(keys.size() * GPU_NUM_ACTIONS);
  for (size_t i = 0; i < keys.size(); ++i) {
    const StrategyEntry& e = strat.at(keys[i]);
    std::copy(e.probs, e.probs + GPU_NUM_ACTIONS, &probs[i * GPU_NUM_ACTIONS]);
  }
  return mccfr_write_strategy(path.c_str(), (uint32_t)GPU_TABLE_SIZE, (uint32_t)GPU_NUM_ACTIONS,
                              (uint32_t)keys.size(), keys.data(), probs.data()) == MCCFR_OK;
}

// strategy.cpp:54-110
inline bool strategy_load(HostStrategyTable& strat, const std::string& path) {
  uint32_t ts = 0, na = 0, n = 0;
  if (mccfr_load_strategy(path.c_str(), &ts, &na, &n, nullptr, nullptr, 0) != MCCFR_OK) {
    std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(nullptr));
    return false;
  }
  if (ts != (uint32_t)GPU_TABLE_SIZE || na != (uint32_t)GPU_NUM_ACTIONS) {
    std::fprintf(stderr,
                 "ERROR: %s was exported for table_size=%u num_actions=%u, but this build expects %u and %u.\n",
                 path.c_str(), ts, na, (unsigned)GPU_TABLE_SIZE, (unsigned)GPU_NUM_ACTIONS);
    return false;
  }
  std::vector<uint32_t> keys(n);
  std::vector<float> probs((size_t)n * na);
  if (mccfr_load_strategy(path.c_str(), &ts, &na, &n, keys.data(), probs.data(), keys.size()) != MCCFR_OK) {
    std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(nullptr));
    return false;
  }
  strat.reserve(strat.size() + n);
  for (uint32_t i = 0; i < n; ++i) {
    StrategyEntry e;
    std::copy(&probs[(size_t)i * na], &probs[(size_t)i * na] + na, e.probs);
    strat.emplace(keys[i], e);
  }
  return true;
}

// strategy.cpp:112-156
inline bool strategy_load_auto(HostStrategyTable& strat, const std::string& path, int n_players, int stack, int sb,
                               int bb) {
  uint64_t magic = 0;
  {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const bool ok = std::fread(&magic, 8, 1, f) == 1;
    std::fclose(f);
    if (!ok) return false;
  }
  constexpr uint64_t kV1 = 0x53545241540000ull, kV2 = 0x53545241540100ull, kCkpt = 0x43465247505500ull;
  if (magic == kV1 || magic == kV2) return strategy_load(strat, path);
  if (magic != kCkpt) {
    std::fprintf(stderr, "ERROR: %s is neither a strategy .bin nor a checkpoint .ckpt file.\n", path.c_str());
    return false;
  }
  std::fprintf(stderr, "NOTE: %s is a checkpoint; converting it to a playable strategy in memory...\n",
               path.c_str());
  GPUCFRTrainer trainer(n_players, stack, sb, bb);
  if (!trainer.load_checkpoint(path)) {
    std::fprintf(stderr, "ERROR: Could not load checkpoint from %s\n", path.c_str());
    return false;
  }
  strat = trainer.get_strategy();
  std::string out = path;
  if (out.size() >= 5 && out.compare(out.size() - 5, 5, ".ckpt") == 0)
    out.erase(out.size() - 5);
  else
    out += ".bin";
  if (strategy_save(strat, out))
    std::fprintf(stderr, "NOTE: Exported playable strategy to %s\n", out.c_str());
  else
    std::fprintf(stderr, "WARNING: Converted checkpoint in memory, but could not save %s\n", out.c_str());
  return true;
}

// strategy.cpp:158-181: the n most likely actions above 0.1%.
inline void strategy_print_top(const HostStrategyTable& strat, uint32_t key, int n = 3) {
  static const char* kNames[GPU_NUM_ACTIONS] = {"fold",       "check",      "call",      "raise_min",
                                                "raise_half", "raise_two_thirds", "raise_pot", "all_in"};
  auto it = strat.find(key);
  if (it == strat.end()) {
    std::printf("(info set not found)\n");
    return;
  }
  const StrategyEntry& e = it->second;
  int order[GPU_NUM_ACTIONS];
  for (int i = 0; i < GPU_NUM_ACTIONS; ++i) order[i] = i;
  // exchange sort on strictly greater, so ties come out in the reference's order
  for (int i = 0; i + 1 < GPU_NUM_ACTIONS; ++i)
    for (int j = i + 1; j < GPU_NUM_ACTIONS; ++j)
      if (e.probs[order[j]] > e.probs[order[i]]) std::swap(order[i], order[j]);
  for (int i = 0; i < n && i < GPU_NUM_ACTIONS; ++i) {
    const int a = order[i];
    if (e.probs[a] > 0.001f) std::printf("  %-16s %.1f%%\n", kNames[a], e.probs[a] * 100.f);
  }
}
