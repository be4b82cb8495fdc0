// gpu_cfr_trainer.hpp — drop-in C++ trainer class over the C ABI (mccfr.h).
//
// Same class name, constructor arguments, defaults and member functions as the
// reference's GPUCFRTrainer (poker_cuda/src/cfr_gpu.cuh:48-132), so callers such
// as the reference CLI (main.cpp:309-377) compile against it unchanged:
//   GPUCFRTrainer(num_players=2, starting_stack=1000, small_blind=10,
//                 big_blind=20, use_cfr_plus=true, use_linear_cfr=true)
//   load_hand_table, train, get_strategy, save_checkpoint, load_checkpoint,
//   num_info_sets_active, export_tables, import_tables.
// The reference constructor gives the reference geometry (2^26 slots x 8
// actions); a second constructor takes any mccfr_game_cfg.  export/import use
// action-major vectors (cfr_gpu.cu:1437-1459).  Errors follow the
// reference: file and format problems return false; a device error prints the
// library message and exits (the reference's CUDA_CHECK, cfr_gpu.cu:61-69) --
// the library itself never exits.
#pragma once
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unordered_map>
#include <vector>

#include "mccfr.h"

static constexpr int GPU_NUM_ACTIONS = 8;          // cfr_gpu.cuh:17
static constexpr int GPU_TABLE_SIZE = 1 << 26;     // cfr_gpu.cuh:18
static constexpr int GPU_MAX_PLAYERS = 6;          // cfr_gpu.cuh:20

struct StrategyEntry {  // cfr_gpu.cuh:30-32
  float probs[GPU_NUM_ACTIONS] = {};
};
using HostStrategyTable = std::unordered_map<uint32_t, StrategyEntry>;

class GPUCFRTrainer {
 public:
  GPUCFRTrainer(int num_players = 2, int starting_stack = 1000, int small_blind = 10, int big_blind = 20,
                bool use_cfr_plus = true, bool use_linear_cfr = true)
      : N_(num_players) {
    mccfr_game_cfg_reference(&cfg_, num_players, starting_stack, small_blind, big_blind, use_cfr_plus ? 1 : 0,
                             use_linear_cfr ? 1 : 0);
    check(mccfr_create(&cfg_, batch_cap_, &ctx_), "create");
    shape();
  }
  // Any game of the C ABI: menu, raise cap, table size and seed set by the caller.
  explicit GPUCFRTrainer(const mccfr_game_cfg& cfg, uint32_t batch_cap = 65536)
      : N_(cfg.num_players), cfg_(cfg), batch_cap_(batch_cap) {
    check(mccfr_create(&cfg_, batch_cap_, &ctx_), "create");
    shape();
  }
  ~GPUCFRTrainer() { mccfr_destroy(ctx_); }
  GPUCFRTrainer(const GPUCFRTrainer&) = delete;
  GPUCFRTrainer& operator=(const GPUCFRTrainer&) = delete;

  // cfr_gpu.cu:1086-1104: false when the file is missing or fails validation.
  bool load_hand_table(const char* path = "data/handranks.dat") {
    const int rc = mccfr_load_hand_table(ctx_, path);
    if (rc == MCCFR_OK) return true;
    if (rc == MCCFR_E_HIP || rc == MCCFR_E_NOMEM) check(rc, "load_hand_table");
    std::fprintf(stderr, "[train] %s\n", mccfr_last_error(ctx_));
    return false;
  }
  // The project generates the table itself (no download needed).
  bool use_generated_hand_table() { return mccfr_use_generated_hand_table(ctx_) == MCCFR_OK; }

  // cfr_gpu.cu:1244-1378: iterations of N launches, progress every 100
  // iterations, checkpoint every ckpt_interval iterations.
  void train(long long total_iterations, int batch_size = 65536, bool verbose = true,
             const std::string& ckpt_path = "", long long ckpt_interval = 10000) {
    if ((uint32_t)batch_size > batch_cap_) {
      check(mccfr_set_max_batch(ctx_, (uint32_t)batch_size), "set_max_batch");
      batch_cap_ = (uint32_t)batch_size;
    }
    if (verbose)
      std::printf("Training: %lld iters, batch=%d, players=%d, CFR+=%s, LCFR=%s\n", total_iterations, batch_size,
                  N_, cfg_.use_cfr_plus ? "on" : "off", cfg_.use_linear_cfr ? "on" : "off");
    const auto t0 = std::chrono::steady_clock::now();
    long long done = 0;
    while (done < total_iterations) {
      long long chunk = total_iterations - done;
      if (verbose && chunk > 100) chunk = 100 - done % 100;
      if (!ckpt_path.empty() && ckpt_interval > 0) {
        const long long to_ckpt = ckpt_interval - done % ckpt_interval;
        if (chunk > to_ckpt) chunk = to_ckpt;
      }
      check(mccfr_train(ctx_, chunk, (uint32_t)batch_size), "train");
      done += chunk;
      if (verbose && done % 100 == 0) {
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("\r  %lld/%lld iters  %.1fM hands/s  ", done, total_iterations,
                    (double)done * batch_size * N_ / sec / 1e6);
        std::fflush(stdout);
      }
      if (!ckpt_path.empty() && ckpt_interval > 0 && done % ckpt_interval == 0 && save_checkpoint(ckpt_path) &&
          verbose)
        std::printf("\n  [checkpoint -> %s]\n", ckpt_path.c_str());
    }
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double hands = (double)total_iterations * batch_size * N_;
    if (verbose) std::printf("\n");
    std::printf("Done: %.0fM hands in %.2fs (%.1fM hands/sec)\n", hands / 1e6, sec, hands / sec / 1e6);
  }

  // cfr_gpu.cu:1383-1412: normalised average strategy of rows with sum >= 1e-6
  // (8-action games; save_strategy writes any game).
  HostStrategyTable get_strategy() const {
    if (row_ != GPU_NUM_ACTIONS) {
      std::fprintf(stderr, "get_strategy: the game has %d actions, StrategyEntry holds %d\n", row_, GPU_NUM_ACTIONS);
      std::exit(1);
    }
    std::vector<float> avg(elems());
    check(mccfr_normalized_strategy(ctx_, avg.data(), avg.size(), 0), "normalized_strategy");
    HostStrategyTable table;
    for (uint64_t s = 0; s < slots_; ++s) {
      const float* row = &avg[(size_t)s * GPU_NUM_ACTIONS];
      float total = 0.f;
      for (int a = 0; a < GPU_NUM_ACTIONS; ++a) total += row[a];
      if (total < 1e-6f) continue;
      StrategyEntry e;
      for (int a = 0; a < GPU_NUM_ACTIONS; ++a) e.probs[a] = row[a];
      table.emplace((uint32_t)s, e);
    }
    return table;
  }
  // get_strategy() written as a .bin v2 file (main.cpp:371-375), for any game.
  bool save_strategy(const std::string& path) const {
    const int rc = mccfr_save_strategy(ctx_, 0, path.c_str());
    if (rc != MCCFR_OK) std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(ctx_));
    return rc == MCCFR_OK;
  }

  bool save_checkpoint(const std::string& path) const { return mccfr_save_checkpoint(ctx_, path.c_str()) == MCCFR_OK; }
  bool load_checkpoint(const std::string& path) {
    const int rc = mccfr_load_checkpoint(ctx_, path.c_str());
    if (rc != MCCFR_OK) std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(ctx_));
    return rc == MCCFR_OK;
  }

  // cfr_gpu.cu:1414-1426: rows with any positive strategy sum.
  int num_info_sets_active() const {
    std::vector<float> S(elems());
    check(mccfr_download_tables(ctx_, nullptr, S.data(), nullptr, S.size(), 0), "download");
    int count = 0;
    for (size_t s = 0; s < slots_; ++s)
      for (int a = 0; a < row_; ++a)
        if (S[s * row_ + a] > 0.f) {
          ++count;
          break;
        }
    return count;
  }

  // cfr_gpu.cu:1437-1459 (action-major vectors of actions x slots floats)
  void export_tables(std::vector<float>& regrets, std::vector<float>& strategy_sum) const {
    const size_t n = elems();
    regrets.resize(n);
    strategy_sum.resize(n);
    check(mccfr_download_tables(ctx_, regrets.data(), strategy_sum.data(), nullptr, n, 1), "export_tables");
  }
  void import_tables(const std::vector<float>& regrets, const std::vector<float>& strategy_sum) {
    const size_t n = elems();
    if (regrets.size() != n || strategy_sum.size() != n) {
      std::fprintf(stderr, "import_tables: expected %zu floats per table\n", n);
      std::exit(1);
    }
    check(mccfr_upload_tables(ctx_, regrets.data(), strategy_sum.data(), n, 1), "import_tables");
  }

  mccfr_ctx* context() const { return ctx_; }
  const mccfr_game_cfg& config() const { return cfg_; }
  uint64_t table_slots() const { return slots_; }
  int row_actions() const { return row_; }

 private:
  void shape() {
    uint32_t row = 0;
    check(mccfr_tenant_table_shape(ctx_, 0, &slots_, &row), "table_shape");
    row_ = (int)row;
  }
  size_t elems() const { return (size_t)slots_ * row_; }
  void check(int rc, const char* what) const {
    if (rc == MCCFR_OK) return;
    std::fprintf(stderr, "mccfr %s failed (%d): %s\n", what, rc, mccfr_last_error(ctx_));
    std::exit(1);
  }

  int N_;
  mccfr_game_cfg cfg_{};
  mccfr_ctx* ctx_ = nullptr;
  uint32_t batch_cap_ = 65536;
  uint64_t slots_ = 0;
  int row_ = GPU_NUM_ACTIONS;
};
