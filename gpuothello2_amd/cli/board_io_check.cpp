// board_io_check INPUT OUTPUT — runs the solver's board-file path without a
// GPU: reads, validates, sorts and de-duplicates INPUT as `solver` does, and
// writes each unique board with score 0 (and, with a third argument, move
// "ps").  Prints the phase times to stderr.  Used by the tests and for
// timing multi-million-line files.
#include <chrono>
#include <cstdio>
#include <vector>

#include "board_io.hpp"

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s INPUT OUTPUT [with-moves]\n", argv[0]);
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const auto boards = go2cli::load_boards(argv[1]);
  const auto t1 = std::chrono::steady_clock::now();
  std::vector<int8_t> scores(boards.size(), 0);
  std::vector<uint8_t> moves(boards.size(), GO2_MOVE_PASS);
  FILE* out = fopen(argv[2], "w");
  if (!out) go2cli::die(2, std::string("cannot open output ") + argv[2]);
  go2cli::write_results(out, boards, scores.data(), argc > 3 ? moves.data() : nullptr,
                        [](int) { return std::string("ps"); });
  fclose(out);
  const auto t2 = std::chrono::steady_clock::now();
  fprintf(stderr, "unique %zu, read+sort %.3f s, write %.3f s\n", boards.size(),
          std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
  return 0;
}
