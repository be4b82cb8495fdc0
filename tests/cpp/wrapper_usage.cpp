// Compile-and-link check of the drop-in trainer class against the calls the
// reference CLI makes (poker_cuda/src/main.cpp:309-377, :383-397).  Built by
// tests/test_capi.py with g++ and run on the device by tests/test_gpu_io.py
// (argv[1] hand table or "-" for the generated one, argv[2] a scratch folder).
#include <cstdio>
#include <string>
#include <vector>

#include "mccfr/gpu_cfr_trainer.hpp"
#include "mccfr/strategy.hpp"

int main(int argc, char** argv) {
  const int n_players = 2, stack = 1000, sb = 10, bb = 20;
  GPUCFRTrainer trainer(n_players, stack, sb, bb, true, true);
  if (!trainer.load_hand_table(argc > 1 ? argv[1] : "data/handranks.dat") && !trainer.use_generated_hand_table())
    return 1;
  trainer.train(2, 4096, true, "", 10000);
  std::vector<float> reg, ssum;
  trainer.export_tables(reg, ssum);
  trainer.import_tables(reg, ssum);
  std::printf("Info sets active: %d\n", trainer.num_info_sets_active());
  HostStrategyTable strat = trainer.get_strategy();
  std::printf("strategy rows: %zu\n", strat.size());
  const std::string dir = argc > 2 ? argv[2] : "/tmp";
  const std::string ckpt = dir + "/wrapper_usage.ckpt";
  if (!trainer.save_checkpoint(ckpt) || !trainer.load_checkpoint(ckpt)) return 2;
  if (strat.empty()) return 3;
  // main.cpp:371-375: the strategy export after training
  const std::string bin = dir + "/wrapper_usage.bin";
  if (!strategy_save(strat, bin)) return 4;
  HostStrategyTable loaded;
  if (!strategy_load(loaded, bin) || loaded.size() != strat.size()) return 5;
  // main.cpp:170 (eval mode): a checkpoint converts through the trainer and is written next to it
  HostStrategyTable converted;
  if (!strategy_load_auto(converted, ckpt, n_players, stack, sb, bb) || converted.size() != strat.size()) return 6;
  for (const auto& kv : strat) {
    auto a = loaded.find(kv.first), b = converted.find(kv.first);
    if (a == loaded.end() || b == converted.end()) return 7;
    for (int i = 0; i < GPU_NUM_ACTIONS; ++i)
      if (a->second.probs[i] != kv.second.probs[i] || b->second.probs[i] != kv.second.probs[i]) return 8;
  }
  HostStrategyTable again;
  if (!strategy_load(again, dir + "/wrapper_usage") || again.size() != strat.size()) return 9;
  strategy_print_top(strat, strat.begin()->first, 3);
  return 0;
}
