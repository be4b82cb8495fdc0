// poker_mccfr — command-line trainer over libmccfr, with the reference CLI's
// train, benchmark, eval and info modes, flags and defaults (poker_cuda/src/main.cpp:29-69,
// :98-139, :146-198, :295-398).
//
// Differences from the reference, all additive or fixes:
//   * --seed, --table-log2, --menu and --raise-cap configure the game beyond
//     the reference's fixed 8-action menu, 2^26-slot table and seed 42;
//   * benchmark mode times only the training launches: each trainer is built,
//     its hand table loaded, its scratch sized and one warm-up iteration run
//     (which captures the launch graph) before the clock starts — the
//     reference's timed region also held per-call setup (main.cpp:389-393);
//   * the hand table falls back to the project's generated table when the
//     default data/handranks.dat is absent (a path given with --handranks must
//     load, as in the reference);
//   * eval mode plays its hands on the GPU (one lane per hand, each hand with
//     its own Philox stream) instead of one std::mt19937 sequence on the CPU:
//     the same game and archetypes, results equal in distribution;
//   * one process drives one GPU (no MPI ranks); the reference's interactive
//     play and human modes are not provided;
//   * an unknown or incomplete option is an error (exit 2) rather than ignored.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mccfr/gpu_cfr_trainer.hpp"
#include "mccfr/mccfr.h"

namespace {

void print_usage(const char* prog) {
  std::printf(
      "Usage: %s [options]\n"
      "\n"
      "Training / benchmark flags:\n"
      "  --mode <train|benchmark|eval|info>  (default: train)\n"
      "  --players  <N>                      (default: 2)\n"
      "  --iters    <N>                      (default: 1000)\n"
      "  --batch    <N>                      (default: 65536)\n"
      "  --stack    <N>                      (default: 1000)\n"
      "  --sb       <N>                      (default: 10)\n"
      "  --bb       <N>                      (default: 20)\n"
      "  --save     <file>                   (default: strategy.bin)\n"
      "  --load     <file>                   load checkpoint before training\n"
      "  --handranks <file>                  (default: data/handranks.dat; 'generated' for the\n"
      "                                      project's own table, also used when the default is absent)\n"
      "  --no-cfrplus / --cfrplus            CFR+ clamping (default: on)\n"
      "  --no-lcfr / --lcfr                  Linear CFR weighting (default: on)\n"
      "\n"
      "Game flags:\n"
      "  --seed       <N>                    traversal stream seed (default: 42)\n"
      "  --table-log2 <K>                    info-set table of 2^K slots (default: 26)\n"
      "  --menu       <list>                 comma list of fold, check, call, checkcall, min,\n"
      "                                      pot, allin and pot fractions N/D\n"
      "                                      (default: fold,check,call,min,1/2,2/3,pot,allin)\n"
      "  --raise-cap  <N>                    aggressive actions per street, 0 = no cap (default: 0)\n"
      "\n"
      "Eval flags (--mode eval):\n"
      "  --strategy  <file>                  strategy .bin (or .ckpt) to evaluate (default: --save)\n"
      "  --opponent  <type>                  calling_station|nit|maniac|balanced|random\n"
      "                                      (default: calling_station)\n"
      "  --hands     <N>                     number of hands to play (default: 10000)\n"
      "  --players   <N>                     2-9 seats, the strategy in seat 0\n"
      "\n"
      "Benchmark flags:\n"
      "  --batch-sizes <list>                (default: 4096,16384,65536,262144)\n"
      "  --bench-iters <N>                   timed iterations per batch size (default: 10)\n",
      prog);
}

void print_gpu_info() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  std::printf("GPUs found: %d\n", n);
  for (int i = 0; i < n; ++i) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, i) != hipSuccess) continue;
    std::printf("  [%d] %s (%s)  CU=%d  VRAM=%.0fGB  BW=%.0fGB/s  L2=%.0fMB\n", i, p.name, p.gcnArchName,
                p.multiProcessorCount, p.totalGlobalMem / 1e9,
                2.0 * p.memoryClockRate * (p.memoryBusWidth / 8) / 1e6, p.l2CacheSize / 1e6);
  }
}

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  size_t a = 0;
  while (a <= s.size()) {
    size_t b = s.find(sep, a);
    if (b == std::string::npos) b = s.size();
    out.push_back(s.substr(a, b - a));
    a = b + 1;
  }
  return out;
}

// Parses a --menu list into cfg.menu / num_actions; false with a message on error.
bool parse_menu(const std::string& text, mccfr_game_cfg& cfg, std::string& why) {
  const std::vector<std::string> items = split(text, ',');
  if (items.empty() || items.size() > MCCFR_MAX_ACTIONS) {
    why = "--menu needs 1.." + std::to_string(MCCFR_MAX_ACTIONS) + " actions";
    return false;
  }
  std::memset(cfg.menu, 0, sizeof cfg.menu);
  for (size_t i = 0; i < items.size(); ++i) {
    const std::string& t = items[i];
    mccfr_action a{0, 0, 1, 0};
    if (t == "fold") a.kind = MCCFR_ACT_FOLD;
    else if (t == "check") a.kind = MCCFR_ACT_CHECK;
    else if (t == "call") a.kind = MCCFR_ACT_CALL;
    else if (t == "checkcall") a.kind = MCCFR_ACT_CHECK_CALL;
    else if (t == "min") a.kind = MCCFR_ACT_RAISE_MIN;
    else if (t == "allin") a.kind = MCCFR_ACT_ALL_IN;
    else if (t == "pot") a = mccfr_action{MCCFR_ACT_POT_FRAC, 1, 1, 0};
    else {
      const size_t slash = t.find('/');
      char* e1 = nullptr;
      char* e2 = nullptr;
      const long num = slash == std::string::npos ? -1 : std::strtol(t.c_str(), &e1, 10);
      const long den = slash == std::string::npos ? -1 : std::strtol(t.c_str() + slash + 1, &e2, 10);
      if (slash == std::string::npos || e1 != t.c_str() + slash || *e2 != '\0' || num < 1 || num > 255 || den < 1 ||
          den > 255) {
        why = "unknown --menu action '" + t + "'";
        return false;
      }
      a = mccfr_action{MCCFR_ACT_POT_FRAC, (uint8_t)num, (uint8_t)den, 0};
    }
    cfg.menu[i] = a;
  }
  cfg.num_actions = (int32_t)items.size();
  int bits = 1;
  while ((1 << bits) < cfg.num_actions) ++bits;
  cfg.history_bits = bits;
  return true;
}

bool load_table(GPUCFRTrainer& t, const std::string& path, bool given) {
  if (path == "generated") return t.use_generated_hand_table();
  if (t.load_hand_table(path.c_str())) return true;
  if (!given && t.use_generated_hand_table()) {
    std::fprintf(stderr, "NOTE: %s not found; using the generated hand table\n", path.c_str());
    return true;
  }
  std::fprintf(stderr, "ERROR: Could not load a compatible handranks.dat from %s\n", path.c_str());
  return false;
}

uint32_t strategy_entries(const std::string& path) {
  uint32_t n = 0;
  std::FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return 0;
  if (std::fseek(f, 16, SEEK_SET) != 0 || std::fread(&n, 4, 1, f) != 1) n = 0;
  std::fclose(f);
  return n;
}

// main.cpp:146-198: load a strategy (.bin, or a .ckpt converted through the
// trainer as strategy_load_auto does, strategy.cpp:112-156), play it against
// the archetype, print the reference's result lines.
int run_eval_mode(mccfr_game_cfg cfg, const std::string& path, const std::string& opp_name, int n_players,
                  long long hands, uint64_t seed, const std::string& hr_path, bool hr_given) {
  int opp = MCCFR_OPP_CALLING_STATION;  // main.cpp:176-181: unknown names fall back to calling_station
  if (opp_name == "nit") opp = MCCFR_OPP_NIT;
  else if (opp_name == "maniac") opp = MCCFR_OPP_MANIAC;
  else if (opp_name == "balanced") opp = MCCFR_OPP_BALANCED;
  else if (opp_name == "random") opp = MCCFR_OPP_RANDOM;
  std::printf("Eval: loading strategy from %s ...\n", path.c_str());
  uint64_t magic = 0;
  uint32_t ts = 0, na = 0;
  {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    const bool ok = f && std::fread(&magic, 8, 1, f) == 1 && std::fread(&ts, 4, 1, f) == 1 &&
                    std::fread(&na, 4, 1, f) == 1;
    if (f) std::fclose(f);
    if (!ok) {
      std::fprintf(stderr, "ERROR: Could not load strategy from %s\n", path.c_str());
      return 1;
    }
  }
  constexpr uint64_t kCkpt = 0x43465247505500ull, kV2 = 0x53545241540100ull;
  if ((magic != kCkpt && magic != kV2) || ts == 0 || (ts & (ts - 1)) != 0) {
    uint32_t t0 = 0, a0 = 0, e0 = 0;
    if (mccfr_load_strategy(path.c_str(), &t0, &a0, &e0, nullptr, nullptr, 0) != MCCFR_OK)
      std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(nullptr));
    std::fprintf(stderr, "ERROR: Could not load strategy from %s\n", path.c_str());
    return 1;
  }
  int log2 = 0;
  while ((1u << log2) < ts) ++log2;
  cfg.table_log2 = log2;  // the file fixes the table size
  GPUCFRTrainer bot(cfg, 4096);
  if (!load_table(bot, hr_path, hr_given)) return 1;
  if (bot.row_actions() != (int)na) {
    std::fprintf(stderr, "ERROR: %s holds %u actions per row, this game %d (see --menu)\n", path.c_str(), na,
                 bot.row_actions());
    return 1;
  }
  mccfr_eval_result res{};
  int rc;
  if (magic == kCkpt) {
    std::fprintf(stderr, "NOTE: %s is a checkpoint; converting it to a playable strategy in memory...\n",
                 path.c_str());
    if (!bot.load_checkpoint(path)) {
      std::fprintf(stderr, "ERROR: Could not load checkpoint from %s\n", path.c_str());
      return 1;
    }
    std::string out = path;
    if (out.size() >= 5 && out.compare(out.size() - 5, 5, ".ckpt") == 0)
      out.erase(out.size() - 5);
    else
      out += ".bin";
    if (bot.save_strategy(out)) {
      std::fprintf(stderr, "NOTE: Exported playable strategy to %s\n", out.c_str());
      std::printf("  Loaded %u info sets.\n", strategy_entries(out));
    }
    std::printf("Opponent: %s  |  Players: %d  |  Hands: %lld\n\n", opp_name.c_str(), n_players, hands);
    rc = mccfr_evaluate_strategy(bot.context(), 0, n_players, opp, hands, seed, nullptr, &res);
  } else {
    uint32_t t1 = 0, a1 = 0, n = 0;
    if (mccfr_load_strategy(path.c_str(), &t1, &a1, &n, nullptr, nullptr, 0) != MCCFR_OK) {
      std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(nullptr));
      return 1;
    }
    std::vector<uint32_t> keys(n);
    std::vector<float> probs((size_t)n * a1);
    if (mccfr_load_strategy(path.c_str(), &t1, &a1, &n, keys.data(), probs.data(), keys.size()) != MCCFR_OK) {
      std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(nullptr));
      return 1;
    }
    std::printf("  Loaded %u info sets.\n", n);
    std::printf("Opponent: %s  |  Players: %d  |  Hands: %lld\n\n", opp_name.c_str(), n_players, hands);
    rc = mccfr_evaluate_rows(bot.context(), &bot.config(), keys.data(), probs.data(), n, a1, n_players, opp, hands,
                             seed, nullptr, &res);
  }
  if (rc != MCCFR_OK) {
    std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(bot.context()));
    return 1;
  }
  std::printf("=== Eval Results (%d-player) ===\n", n_players);
  std::printf("  Hands played : %lld\n", (long long)res.hands_played);
  std::printf("  Net BB       : %.1f\n", res.net_bb);
  std::printf("  BB/100       : %+.2f\n", res.bb_per_100);
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  std::string mode = "train";
  int n_players = 2;
  long long iters = 1000;
  int batch_size = 65536;
  int stack_size = 1000, sb = 10, bb = 20;
  bool use_cfr_plus = true, use_linear_cfr = true;
  std::string save_path = "strategy.bin", load_path, hr_path = "data/handranks.dat";
  bool hr_given = false;
  long long seed = 42;
  int table_log2 = 26, raise_cap = 0;
  std::string menu, batch_sizes = "4096,16384,65536,262144";
  int bench_iters = 10;
  std::string strategy_path, opponent_str = "calling_station";
  long long eval_hands = 10000;

  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    const bool v = i + 1 < argc;
    if (!std::strcmp(a, "--mode") && v) mode = argv[++i];
    else if (!std::strcmp(a, "--players") && v) n_players = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--iters") && v) iters = std::atoll(argv[++i]);
    else if (!std::strcmp(a, "--batch") && v) batch_size = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--stack") && v) stack_size = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--sb") && v) sb = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--bb") && v) bb = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--save") && v) save_path = argv[++i];
    else if (!std::strcmp(a, "--load") && v) load_path = argv[++i];
    else if (!std::strcmp(a, "--handranks") && v) hr_path = argv[++i], hr_given = true;
    else if (!std::strcmp(a, "--no-cfrplus")) use_cfr_plus = false;
    else if (!std::strcmp(a, "--cfrplus")) use_cfr_plus = true;
    else if (!std::strcmp(a, "--no-lcfr")) use_linear_cfr = false;
    else if (!std::strcmp(a, "--lcfr")) use_linear_cfr = true;
    else if (!std::strcmp(a, "--seed") && v) seed = std::atoll(argv[++i]);
    else if (!std::strcmp(a, "--table-log2") && v) table_log2 = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--menu") && v) menu = argv[++i];
    else if (!std::strcmp(a, "--raise-cap") && v) raise_cap = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--batch-sizes") && v) batch_sizes = argv[++i];
    else if (!std::strcmp(a, "--bench-iters") && v) bench_iters = std::atoi(argv[++i]);
    else if (!std::strcmp(a, "--strategy") && v) strategy_path = argv[++i];
    else if (!std::strcmp(a, "--opponent") && v) opponent_str = argv[++i];
    else if (!std::strcmp(a, "--hands") && v) eval_hands = std::atoll(argv[++i]);
    else if (!std::strcmp(a, "--help")) {
      print_usage(argv[0]);
      return 0;
    } else {
      std::fprintf(stderr, "ERROR: unknown or incomplete option %s (see --help)\n", a);
      return 2;
    }
  }

  if (mode == "info") {
    print_gpu_info();
    return 0;
  }
  if (mode == "play" || mode == "human") {
    std::fprintf(stderr, "ERROR: --mode %s is an interactive tool of the reference CLI and is not part of poker_mccfr\n",
                 mode.c_str());
    return 2;
  }
  if (mode != "train" && mode != "benchmark" && mode != "eval") {
    std::fprintf(stderr, "ERROR: unknown --mode %s (train, benchmark, eval or info)\n", mode.c_str());
    return 2;
  }
  const int eval_players = n_players;
  if (mode == "eval") {
    if (n_players < 2 || n_players > 9) {
      std::fprintf(stderr, "ERROR: --players must be between 2 and 9 for eval mode\n");
      return 2;
    }
    if (eval_hands < 0) {
      std::fprintf(stderr, "ERROR: --hands must be >= 0\n");
      return 2;
    }
    if (n_players > 6) n_players = 6;  // the game the strategy was trained on holds at most 6 seats
  }

  // The game: the reference's, then the extra flags.
  mccfr_game_cfg cfg;
  mccfr_game_cfg_reference(&cfg, n_players, stack_size, sb, bb, use_cfr_plus ? 1 : 0, use_linear_cfr ? 1 : 0);
  cfg.seed = (uint64_t)seed;
  cfg.table_log2 = table_log2;
  cfg.max_raises = raise_cap;
  std::string why;
  if (!menu.empty() && !parse_menu(menu, cfg, why)) {
    std::fprintf(stderr, "ERROR: %s\n", why.c_str());
    return 2;
  }
  if (batch_size < 1) {
    std::fprintf(stderr, "ERROR: --batch must be positive\n");
    return 2;
  }
  std::vector<int> sizes;
  for (const std::string& s : split(batch_sizes, ',')) {
    const int b = std::atoi(s.c_str());
    if (b < 1) {
      std::fprintf(stderr, "ERROR: bad --batch-sizes entry '%s'\n", s.c_str());
      return 2;
    }
    sizes.push_back(b);
  }
  uint64_t table_bytes = 0;  // validates the game before any device work
  if (mccfr_tenant_footprint(&cfg, &table_bytes) != MCCFR_OK) {
    std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(nullptr));
    return 2;
  }

  std::printf("=== poker_mccfr (MI355X) ===\n");
  print_gpu_info();

  if (mode == "eval") return run_eval_mode(cfg, strategy_path.empty() ? save_path : strategy_path, opponent_str,
                                           eval_players, eval_hands, (uint64_t)seed, hr_path, hr_given);

  if (mode == "train") {
    GPUCFRTrainer trainer(cfg, (uint32_t)batch_size);
    if (!load_table(trainer, hr_path, hr_given)) return 1;
    if (!load_path.empty()) {
      std::printf("Loading checkpoint from %s ...\n", load_path.c_str());
      if (!trainer.load_checkpoint(load_path))
        std::fprintf(stderr, "WARNING: Could not load checkpoint %s\n", load_path.c_str());
    }
    std::printf("\nTraining: players=%d  iters=%lld  batch=%d  actions=%d  slots=2^%d (%.2f GB of tables)  seed=%llu\n",
                n_players, iters, batch_size, cfg.num_actions, cfg.table_log2, table_bytes / 1e9,
                (unsigned long long)cfg.seed);
    const std::string ckpt = save_path.empty() ? std::string() : save_path + ".ckpt";
    trainer.train(iters, batch_size, true, ckpt, 10000);
    std::printf("\nInfo sets active: %d\n", trainer.num_info_sets_active());
    if (trainer.save_checkpoint(ckpt)) std::printf("Checkpoint saved to %s\n", ckpt.c_str());
    if (trainer.save_strategy(save_path))
      std::printf("Strategy saved to %s  (%u info sets)\n", save_path.c_str(), strategy_entries(save_path));
    else
      std::fprintf(stderr, "ERROR: could not save strategy to %s\n", save_path.c_str());
    return 0;
  }

  // benchmark: setup and one warm-up iteration outside the clock
  std::printf("\nBenchmark: batch sizes vs throughput (%d timed iterations each, setup excluded)\n", bench_iters);
  for (int bs : sizes) {
    GPUCFRTrainer t2(cfg, (uint32_t)bs);
    if (!load_table(t2, hr_path, hr_given)) return 1;
    mccfr_ctx* ctx = t2.context();
    if (mccfr_train(ctx, 1, (uint32_t)bs) != MCCFR_OK) {
      std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(ctx));
      return 1;
    }
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = mccfr_train(ctx, bench_iters, (uint32_t)bs);  // returns after the device finishes
    const auto t1 = std::chrono::steady_clock::now();
    if (rc != MCCFR_OK) {
      std::fprintf(stderr, "ERROR: %s\n", mccfr_last_error(ctx));
      return 1;
    }
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    const long long hands = (long long)bench_iters * bs * n_players;
    std::printf("  batch=%6d  hands=%.0fM  time=%.2fs  speed=%.1fM/s\n", bs, hands / 1e6, sec, hands / sec / 1e6);
  }
  return 0;
}
