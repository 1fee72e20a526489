// solver INPUT OUTPUT DEPTH — drop-in for the reference's solve mode
// (main.cu:123-185): reads n base81 boards, sorts and de-duplicates them,
// solves every board exactly on the GPU through libgo2, and writes one
// "<base81> <score>" line per unique board in sorted (player, opponent) order.
//
// Behaviour kept from the reference:
//   * input: first line n, then n lines whose first 16 bytes are a base81 board
//     (main.cu:49-63); a missing line prints "Failed to load boards" and exits 1;
//   * boards sorted and de-duplicated (main.cu:64-65);
//   * batches of at most 2,000,000 boards, each on its own stream, all launched
//     before polling (main.cu:130-155); each batch's upload is queued on its
//     stream from page-locked memory ahead of its solve, so batch k+1 copies
//     while batch k solves (GO2_CLI_BATCH overrides the batch size);
//   * stderr lines "n = ...", "start!", the status/elapsed/table line and
//     "total nodes: ..." (main.cu:129,144,158-162,172);
//   * usage text and exit code 1 for fewer than three arguments (main.cu:176-179).
// Undefined behaviour of the reference becomes a defined error (exit code 2):
// unreadable files, a non-numeric DEPTH or one outside 1..60, a board with
// more empties than DEPTH (or, solving exactly, than GO2_MAX_EMPTIES),
// malformed characters, overlapping discs, device errors.
// Exactly five arguments select think mode (main.cu:180-183): a depth-limited
// search of THINK_DEPTH moves scored by the evaluator in GO2_WEIGHTS, each
// output line "<base81> <score> <move>" (main.cu:114-117); more arguments
// solve exactly, as in the reference.
// Board files are read, sorted and written by the multi-threaded routines
// of board_io.hpp, with the reference's reading rules.
//
// Search knobs come from the environment: GO2_CONFIG=file.json (integer
// go2_config fields, as the autotuner writes them), then GO2_ORDERING=
// static|mobility|learned, GO2_BLOCK_SIZE, GO2_MOBILITY_MIN_EMPTIES,
// GO2_LEARNED_MIN_EMPTIES; the defaults are go2_config_default()'s.
// GO2_ORDERING=learned needs GO2_WEIGHTS=<raw file: 167265 int16 weights,
// 64 int16 biases, little endian> (scale 64 per disc).
#include <go2.h>

#include "board_io.hpp"

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

constexpr size_t kDefaultBatch = 2000000;  // main.cu:130

using go2cli::die;

// Integer fields of a flat JSON object {"name": value, ...}.
void apply_json_config(const char* path, go2_config* c) {
  File f(fopen(path, "r"));
  if (!f) die(2, std::string("cannot open GO2_CONFIG ") + path);
  std::string text;
  char buf[4096];
  size_t got;
  while ((got = fread(buf, 1, sizeof buf, f.get())) > 0) text.append(buf, got);
  struct Field {
    const char* name;
    int32_t* slot;
  } fields[] = {{"ordering", &c->ordering},
                {"mobility_min_empties", &c->mobility_min_empties},
                {"learned_min_empties", &c->learned_min_empties},
                {"block_size", &c->block_size},
                {"blocks_per_cu", &c->blocks_per_cu},
                {"positions_per_wave", &c->positions_per_wave},
                {"split_min_empties", &c->split_min_empties},
                {"lds_levels", &c->lds_levels},
                {"time_limit_ms", &c->time_limit_ms},
                {"split_cut_after", &c->split_cut_after},
                {"learned_key_eval", &c->learned_key_eval},
                {"learned_key_mobility", &c->learned_key_mobility},
                {"max_pops_per_trip", &c->max_pops_per_trip},
                {"tt_log2_buckets", &c->tt_log2_buckets},
                {"tt_min_empties", &c->tt_min_empties},
                {"graph_launch", &c->graph_launch},
                {"export_min_empties", &c->export_min_empties},
                {"export_low_empties", &c->export_low_empties},
                {"leaf_batch_min", &c->leaf_batch_min},
                {"score_batch_min", &c->score_batch_min},
                {"aspiration_width", &c->aspiration_width},
                {"pvs_min_empties", &c->pvs_min_empties}};
  size_t pos = 0;
  while ((pos = text.find('"', pos)) != std::string::npos) {
    const size_t end = text.find('"', pos + 1);
    if (end == std::string::npos) break;
    const std::string key = text.substr(pos + 1, end - pos - 1);
    const size_t colon = text.find(':', end);
    if (colon == std::string::npos) break;
    char* stop = nullptr;
    const long v = strtol(text.c_str() + colon + 1, &stop, 10);
    bool known = false;
    for (auto& fd : fields)
      if (key == fd.name) {
        *fd.slot = (int32_t)v;
        known = true;
      }
    if (!known) die(2, "GO2_CONFIG: unknown field " + key);
    pos = (size_t)(stop - text.c_str());
  }
}

// Column letter and row digit (bit 0 = A1); "ps" for a pass, "--" when the
// game is over (exact solve's GO2_BEST_MOVE column).
std::string move_text(int m) {
  if (m == GO2_MOVE_PASS) return "ps";
  if (m < 0 || m > 63) return "--";
  return std::string(1, (char)('A' + m % 8)) + (char)('1' + m / 8);
}

// Think mode writes moves as the reference's hand_to_s (main.cu:31-39): a
// board whose game is over is passed first and reported as a pass
// (thinker.cu:151-154,166-169), so it prints "ps" as well.
std::string think_move_text(int m) { return m == GO2_MOVE_NONE ? "ps" : move_text(m); }

int env_int(const char* name, int def) {
  const char* s = getenv(name);
  return s && *s ? atoi(s) : def;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s INPUT OUTPUT DEPTH [THINK_DEPTH]\n", argv[0]);
    return 1;
  }
  // Exactly five arguments select think mode (reference main.cu:175-185,
  // 69-121): a depth-limited search of THINK_DEPTH moves scored by the
  // evaluator in GO2_WEIGHTS, each line "<board> <score> <move>".  DEPTH is
  // read but, as in the reference, not used by think mode.
  const bool think = argc == 5;
  char* end = nullptr;
  const long depth = strtol(argv[3], &end, 10);
  if (!end || *end || depth < 1 || depth > 60) die(2, "DEPTH must be an integer in [1, 60]");
  long think_depth = 0;
  if (think) {
    think_depth = strtol(argv[4], &end, 10);
    if (!end || *end || think_depth < 1 || think_depth > 30) die(2, "THINK_DEPTH must be an integer in [1, 30]");
  }

  const auto boards = go2cli::load_boards(argv[1]);
  for (size_t i = 0; i < boards.size() && !think; ++i) {
    const int e = 64 - __builtin_popcountll(boards[i].player | boards[i].opponent);
    if (e > depth) die(2, "a board has " + std::to_string(e) + " empties, more than DEPTH");
    if (e > GO2_MAX_EMPTIES)
      die(2, "a board has " + std::to_string(e) + " empties; the exact solver takes at most " +
                 std::to_string(GO2_MAX_EMPTIES));
  }
  File out(fopen(argv[2], "w"));
  if (!out) die(2, std::string("cannot open output ") + argv[2]);

  go2_config cfg;
  go2_config_default(&cfg);
  cfg.time_limit_ms = 0;  // a command-line job runs to completion
  if (const char* path = getenv("GO2_CONFIG")) apply_json_config(path, &cfg);
  if (const char* o = getenv("GO2_ORDERING")) {
    if (!strcmp(o, "static")) cfg.ordering = GO2_ORDER_STATIC;
    else if (!strcmp(o, "mobility")) cfg.ordering = GO2_ORDER_MOBILITY;
    else if (!strcmp(o, "learned")) cfg.ordering = GO2_ORDER_LEARNED;
    else die(2, "GO2_ORDERING must be static, mobility or learned");
  }
  if (think) {
    cfg.think_depth = (int)think_depth;
    cfg.ordering = GO2_ORDER_LEARNED;
  }
  cfg.block_size = env_int("GO2_BLOCK_SIZE", cfg.block_size);
  cfg.mobility_min_empties = env_int("GO2_MOBILITY_MIN_EMPTIES", cfg.mobility_min_empties);
  cfg.learned_min_empties = env_int("GO2_LEARNED_MIN_EMPTIES", cfg.learned_min_empties);
  if (cfg.ordering == GO2_ORDER_LEARNED) {
    const char* wpath = getenv("GO2_WEIGHTS");
    if (!wpath) die(2, think ? "think mode needs GO2_WEIGHTS (the evaluator)" : "GO2_ORDERING=learned needs GO2_WEIGHTS");
    File wf(fopen(wpath, "rb"));
    if (!wf) die(2, std::string("cannot open GO2_WEIGHTS ") + wpath);
    int ni = 0, nw = 0, nb = 0;
    go2_pattern_info(&ni, &nw, &nb);
    std::vector<int16_t> w((size_t)nw), b((size_t)nb);
    if (fread(w.data(), 2, w.size(), wf.get()) != w.size() || fread(b.data(), 2, b.size(), wf.get()) != b.size())
      die(2, "GO2_WEIGHTS is shorter than the evaluator");
    if (go2_set_evaluator(w.data(), b.data(), 64.0f) != GO2_OK) die(2, go2_last_error());
  }

  // GO2_BEST_MOVE=1 adds a best move to each output line, written like the
  // reference's think-mode moves (main.cu:31-39,114-117).
  const bool with_moves = think || env_int("GO2_BEST_MOVE", 0) != 0;
  const size_t n = boards.size();
  const long batch_env = env_int("GO2_CLI_BATCH", 0);
  const size_t kBatch = batch_env > 0 ? (size_t)batch_env : kDefaultBatch;
  fprintf(stderr, "n = %zu\n", n);
  // Boards in page-locked memory, so that every batch's copy is queued
  // without waiting and overlaps the solves of the batches before it.
  go2_board* pinned = nullptr;
  if (n) {
    void* p = nullptr;
    if (go2_host_alloc(n * sizeof(go2_board), &p) != GO2_OK) die(2, go2_last_error());
    pinned = static_cast<go2_board*>(p);
    go2cli::parallel_ranges(n, go2cli::worker_count(n), [&](unsigned, size_t lo, size_t hi) {
      std::memcpy(pinned + lo, boards.data() + lo, (hi - lo) * sizeof(go2_board));
    });
  }
  std::vector<go2_task*> tasks;
  for (size_t i = 0; i < n; i += kBatch) {
    go2_task* t = go2_task_create(std::min(kBatch, n - i), think ? 60 : (int)depth, &cfg);
    if (!t) die(2, go2_last_error());
    tasks.push_back(t);
  }
  fprintf(stderr, "start!\n");
  const auto t0 = std::chrono::steady_clock::now();
  for (size_t k = 0; k < tasks.size(); ++k) {
    const size_t i = k * kBatch, cnt = std::min(kBatch, n - i);
    if (go2_task_upload_async(tasks[k], pinned + i, cnt) != GO2_OK) die(2, go2_last_error());
    if (go2_task_launch(tasks[k]) != GO2_OK) die(2, go2_last_error());
  }
  for (;;) {
    bool done = true;
    for (go2_task* t : tasks) {
      const int r = go2_task_query(t);
      if (r < 0) die(2, go2_last_error());
      if (r == 0) done = false;
    }
    if (done) break;
    std::this_thread::sleep_for(std::chrono::microseconds(200));
  }
  const auto t1 = std::chrono::steady_clock::now();
  std::vector<int8_t> scores(n);
  std::vector<uint8_t> moves(with_moves ? n : 0);
  uint64_t nodes = 0, tt_stores = 0, tt_hits = 0, tt_probes = 0;
  for (size_t k = 0; k < tasks.size(); ++k) {
    go2_stats st;
    if (go2_task_download(tasks[k], scores.data() + k * kBatch, &st) != GO2_OK) die(2, go2_last_error());
    if (with_moves && go2_task_download_moves(tasks[k], moves.data() + k * kBatch) != GO2_OK)
      die(2, go2_last_error());
    nodes += st.nodes;
    tt_stores += st.tt_stores;
    tt_hits += st.tt_hits;
    tt_probes += st.tt_probes;
    go2_task_destroy(tasks[k]);
  }
  const long ms = (long)std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
  // The table's update / hit / lookup counts (main.cu:158-162, table.cuh:26-40).
  fprintf(stderr,
          "hipSuccess (no error), elapsed: %ldms, table update count: %" PRIu64 ", table hit: %" PRIu64
          ", table find: %" PRIu64 "\n",
          ms, tt_stores, tt_hits, tt_probes);
  go2_host_free(pinned);
  go2cli::write_results(out.get(), boards, scores.data(), with_moves ? moves.data() : nullptr,
                        think ? think_move_text : move_text);
  fprintf(stderr, "total nodes: %" PRIu64 "\n", nodes);
  return 0;
}
