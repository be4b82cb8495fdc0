// gen_table: writes the 32,487,834-word rank-walk table (129,951,336 bytes,
// little-endian int32, the layout GPUCFRTrainer::load_hand_table reads,
// cfr_gpu.cu:1086-1104) at build time.  Same invocation as the reference's
// generator, tools/gen_lookup_table.cpp:63-66: `gen_table [output_path]`,
// default data/handranks.dat.  The reference's tool is a stub that asks for a
// download (:68-95); this one runs the project's generator (host code, no
// device needed) and checks the result the way every load does: word count,
// every state word inside the table, the royal flush ranking 7462.
//
// Exit status: 0 written (or already present and valid), 1 on any failure.
// An existing file is kept when it validates, and regenerated otherwise.
#include <mccfr/mccfr.h>

#include <cstdio>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr size_t kWords = 32487834;

bool write_all(const char* path, const std::vector<int32_t>& hr) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) {
    std::fprintf(stderr, "gen_table: cannot create %s\n", tmp.c_str());
    return false;
  }
  const size_t n = std::fwrite(hr.data(), sizeof(int32_t), hr.size(), f);
  const bool ok = n == hr.size() && std::fflush(f) == 0;
  if (std::fclose(f) != 0 || !ok) {
    std::fprintf(stderr, "gen_table: short write to %s\n", tmp.c_str());
    std::remove(tmp.c_str());
    return false;
  }
  // a reader never sees a partial table: the finished file replaces the old one
  if (std::rename(tmp.c_str(), path) != 0) {
    std::fprintf(stderr, "gen_table: cannot rename %s to %s\n", tmp.c_str(), path);
    std::remove(tmp.c_str());
    return false;
  }
  return true;
}

}  // namespace

int main(int argc, char** argv) {
  const char* out = argc > 1 ? argv[1] : "data/handranks.dat";
  if (mccfr_validate_hand_table_file(out) == MCCFR_OK) {
    std::printf("gen_table: %s is present and valid\n", out);
    return 0;
  }
  std::vector<int32_t> hr;
  try {
    hr.resize(kWords);
  } catch (...) {
    std::fprintf(stderr, "gen_table: no host memory for the table (%zu bytes)\n", kWords * sizeof(int32_t));
    return 1;
  }
  const unsigned hw = std::thread::hardware_concurrency();
  const int threads = hw == 0 ? 1 : (int)(hw > 16 ? 16 : hw);
  if (mccfr_generate_hand_table(hr.data(), hr.size(), threads) != MCCFR_OK) {
    std::fprintf(stderr, "gen_table: %s\n", mccfr_last_error(nullptr));
    return 1;
  }
  if (mccfr_validate_hand_table(hr.data(), hr.size()) != MCCFR_OK) {
    std::fprintf(stderr, "gen_table: generated table failed its check: %s\n", mccfr_last_error(nullptr));
    return 1;
  }
  if (!write_all(out, hr)) return 1;
  if (mccfr_validate_hand_table_file(out) != MCCFR_OK) {
    std::fprintf(stderr, "gen_table: %s does not read back: %s\n", out, mccfr_last_error(nullptr));
    return 1;
  }
  std::printf("gen_table: wrote %s (%zu bytes)\n", out, kWords * sizeof(int32_t));
  return 0;
}
