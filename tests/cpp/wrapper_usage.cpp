// Compile-and-link check of the drop-in trainer class against the calls the
// reference CLI makes (poker_cuda/src/main.cpp:309-377, :383-397).  Built by
// tests/test_capi.py with g++; main() is only run on a GPU box.
#include <cstdio>
#include <string>
#include <vector>

#include "mccfr/gpu_cfr_trainer.hpp"

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
  const std::string ckpt = "/tmp/mccfr_wrapper_usage.ckpt";
  if (!trainer.save_checkpoint(ckpt) || !trainer.load_checkpoint(ckpt)) return 2;
  return strat.empty() ? 3 : 0;
}
