// C ABI of libgo2: argument checks, device buffers, streams, events and the
// base81 board text.  Kernels live in solver.hip / generator.hip; this file
// only orchestrates them.
//
// Counterparts in the reference:
//   go2_solve / go2_solve_device  <- solve()          main.cu:123-173 (minus file I/O)
//   go2_task_*                    <- BatchedTask      solver.cuh:8-25, solver.cu:313-348
//   go2_board_from_text/_to_text  <- toBoard/fromBoard to_board.cpp:15-60
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "internal.hpp"

namespace go2 {

static thread_local std::string g_last_error;
static thread_local int g_last_code = GO2_OK;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  g_last_code = code;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(GO2_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int resolve_config(const go2_config* cfg, go2_config* out) {
  go2_config c;
  if (cfg) {
    c = *cfg;
  } else {
    go2_config_default(&c);
  }
  if (c.ordering < GO2_ORDER_STATIC || c.ordering > GO2_ORDER_LEARNED)
    return fail(GO2_E_ARG, "config.ordering must be GO2_ORDER_STATIC, _MOBILITY or _LEARNED");
  if (c.block_size < 64 || c.block_size > 256 || c.block_size % 64)
    return fail(GO2_E_ARG, "config.block_size must be a multiple of 64 in [64, 256]");
  if (c.mobility_min_empties < 3) c.mobility_min_empties = 3;
  if (c.learned_min_empties < 3) c.learned_min_empties = 3;
  if (c.blocks_per_cu < 0) return fail(GO2_E_ARG, "config.blocks_per_cu must be >= 0");
  if (c.positions_per_wave < 0 || c.positions_per_wave > 64)
    return fail(GO2_E_ARG, "config.positions_per_wave must be in [0, 64]");
  if (c.split_min_empties < 0) return fail(GO2_E_ARG, "config.split_min_empties must be >= 0");
  if (c.lds_levels < 0 || c.lds_levels > 30) return fail(GO2_E_ARG, "config.lds_levels must be in [0, 30]");
  if (c.time_limit_ms < 0) return fail(GO2_E_ARG, "config.time_limit_ms must be >= 0");
  if (c.export_min_empties < 0) return fail(GO2_E_ARG, "config.export_min_empties must be >= 0");
  if (c.export_low_empties < 0) return fail(GO2_E_ARG, "config.export_low_empties must be >= 0");
  if (c.think_depth < 0 || c.think_depth > 30) return fail(GO2_E_ARG, "config.think_depth must be in [0, 30]");
  if (c.tt_log2_buckets < 0 || c.tt_log2_buckets > 32)
    return fail(GO2_E_ARG, "config.tt_log2_buckets must be in [0, 32]");
  if (c.leaf_batch_min < 0) return fail(GO2_E_ARG, "config.leaf_batch_min must be >= 0");
  if (c.score_batch_min < 0) return fail(GO2_E_ARG, "config.score_batch_min must be >= 0");
  if (c.aspiration_width < 0) return fail(GO2_E_ARG, "config.aspiration_width must be >= 0");
  if (c.pvs_min_empties < 0) return fail(GO2_E_ARG, "config.pvs_min_empties must be >= 0");
  if (c.split_cut_after < 0 || c.split_cut_after > 3)
    return fail(GO2_E_ARG, "config.split_cut_after must be in [0, 3]");
  *out = c;
  return GO2_OK;
}

// Host-array calls also need device copies of the boards and scores.
struct IoBufs {
  go2_board* boards = nullptr;
  int8_t* scores = nullptr;
  uint8_t* moves = nullptr;
  ~IoBufs() {
    if (boards) (void)hipFree(boards);
    if (scores) (void)hipFree(scores);
    if (moves) (void)hipFree(moves);
  }
  int alloc(size_t n) {
    hipError_t e;
    if ((e = hipMalloc(&boards, n * sizeof(go2_board))) != hipSuccess)
      return hip_fail(e, "hipMalloc(boards)");
    if ((e = hipMalloc(&scores, n)) != hipSuccess) return hip_fail(e, "hipMalloc(scores)");
    if ((e = hipMalloc(&moves, n)) != hipSuccess) return hip_fail(e, "hipMalloc(moves)");
    return GO2_OK;
  }
};

// DEPTH, the reference's maximum number of empties (main.cu:127): any value
// a game can have.  Boards with more empties than DEPTH, or (exact solve)
// than GO2_MAX_EMPTIES, are rejected per board by the kernel.
static int check_depth(int max_depth) {
  if (max_depth < 1 || max_depth > 60) return fail(GO2_E_ARG, "max_depth must be in [1, 60]");
  return GO2_OK;
}

// Largest number of empties among host boards.
static int max_empties(const go2_board* b, size_t n) {
  int m = 0;
  for (size_t i = 0; i < n; ++i) {
    const int e = 64 - __builtin_popcountll(b[i].player | b[i].opponent);
    if (e > m) m = e;
  }
  return m;
}

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  int create() {
    hipError_t e;
    if ((e = hipEventCreate(&a)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&b)) != hipSuccess) return hip_fail(e, "hipEventCreate");
    return GO2_OK;
  }
};

}  // namespace go2

using namespace go2;

struct go2_task {
  size_t n = 0;
  size_t count = 0;
  int max_depth = 0;
  int upload_max_e = 0;  // most empties among the uploaded boards: sizes the stack
  go2_config cfg{};
  hipStream_t stream = nullptr;
  Workspace ws;
  IoBufs bufs;
  Events ev;
  GraphCache graph;
  bool launched = false;
};

extern "C" {

int go2_version(void) { return 2; }

const char* go2_last_error(void) { return g_last_error.c_str(); }

int go2_set_device(int device) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return GO2_OK;
}

void go2_config_default(go2_config* cfg) {
  if (!cfg) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->ordering = GO2_ORDER_MOBILITY;
  cfg->mobility_min_empties = 7;
  cfg->learned_min_empties = 11;
  cfg->block_size = 256;
  cfg->blocks_per_cu = 0;
  cfg->positions_per_wave = 0;
  cfg->split_min_empties = 0;
  cfg->lds_levels = 0;
  cfg->time_limit_ms = 600000;
  cfg->aspiration_width = 8;
  cfg->pvs_min_empties = 8;
}

// base81: character i encodes squares 4i..4i+3 as 33 + d0 + 3 d1 + 9 d2 + 32 d3,
// digit 0 empty / 1 side to move / 2 other side (reference test.cpp:7-19).
int go2_board_from_text(const char* text, go2_board* out) {
  if (!text || !out) return fail(GO2_E_ARG, "null argument");
  uint64_t p = 0, o = 0;
  for (int i = 0; i < 16; ++i) {
    const int x = (unsigned char)text[i] - 33;
    const int hi = x >> 5, lo = x & 31;
    if (x < 0 || hi > 2 || lo > 26)
      return fail(GO2_E_PARSE, "character " + std::to_string(i) + " of board text is not base81");
    const int d[4] = {lo % 3, (lo / 3) % 3, lo / 9, hi};
    for (int j = 0; j < 4; ++j) {
      const uint64_t b = 1ULL << (4 * i + j);
      if (d[j] == 1) p |= b;
      if (d[j] == 2) o |= b;
    }
  }
  out->player = p;
  out->opponent = o;
  return GO2_OK;
}

void go2_board_to_text(const go2_board* board, char out[17]) {
  static const int w[4] = {1, 3, 9, 32};
  for (int i = 0; i < 16; ++i) {
    int v = 33;
    for (int j = 0; j < 4; ++j) {
      const int sq = 4 * i + j;
      v += (int)((board->player >> sq) & 1) * w[j] + 2 * (int)((board->opponent >> sq) & 1) * w[j];
    }
    out[i] = (char)v;
  }
  out[16] = '\0';
}

int go2_solve_device(const go2_board* d_boards, size_t n, int max_depth, int8_t* d_scores,
                     const go2_config* cfg, go2_stats* stats, void* hip_stream) {
  return go2_solve_device_moves(d_boards, n, max_depth, d_scores, nullptr, cfg, stats, hip_stream);
}

int go2_solve_device_moves(const go2_board* d_boards, size_t n, int max_depth, int8_t* d_scores,
                           uint8_t* d_moves, const go2_config* cfg, go2_stats* stats, void* hip_stream) {
  int rc;
  if (n && (!d_boards || !d_scores)) return fail(GO2_E_ARG, "null device array");
  go2_config c;
  if ((rc = resolve_config(cfg, &c)) != GO2_OK) return rc;
  if ((rc = check_depth(max_depth)) != GO2_OK) return rc;
  hipStream_t s = (hipStream_t)hip_stream;
  if (n == 0) {
    if (stats) *stats = go2_stats{};
    return GO2_OK;
  }
  Workspace ws;
  Events ev;
  if ((rc = ev.create()) != GO2_OK) return rc;
  if ((rc = launch_solve(d_boards, n, max_depth, d_scores, d_moves, c, ws, s, ev.a, ev.b)) != GO2_OK) return rc;
  return collect_stats(ws, s, ev.a, ev.b, stats);
}

int go2_solve_device_noalloc(const go2_board* d_boards, size_t n, int max_depth, int8_t* d_scores,
                             uint8_t* d_moves, const go2_config* cfg, go2_stats* stats, void* hip_stream) {
  int rc;
  if (n && (!d_boards || !d_scores)) return fail(GO2_E_ARG, "null device array");
  go2_config c;
  if ((rc = resolve_config(cfg, &c)) != GO2_OK) return rc;
  if ((rc = check_depth(max_depth)) != GO2_OK) return rc;
  if (c.tt_log2_buckets > 0)
    return fail(GO2_E_ARG, "this entry point has no transposition table (config.tt_log2_buckets must be 0)");
  hipStream_t s = (hipStream_t)hip_stream;
  if (n == 0) {
    if (stats) *stats = go2_stats{};
    return GO2_OK;
  }
  Workspace ws;  // fixed mode: nothing allocated, nothing freed
  if ((rc = noalloc_scratch(&ws.fixed, &ws.fixed_cap)) != GO2_OK) return rc;
  if ((rc = launch_solve(d_boards, n, max_depth, d_scores, d_moves, c, ws, s, nullptr, nullptr)) != GO2_OK)
    return rc;
  return collect_stats(ws, s, nullptr, nullptr, stats);
}

int go2_solve(const go2_board* boards, size_t n, int max_depth, int8_t* scores,
              const go2_config* cfg, go2_stats* stats) {
  return go2_solve_moves(boards, n, max_depth, scores, nullptr, cfg, stats);
}

int go2_solve_moves(const go2_board* boards, size_t n, int max_depth, int8_t* scores, uint8_t* moves,
                    const go2_config* cfg, go2_stats* stats) {
  if (n && (!boards || !scores)) return fail(GO2_E_ARG, "null host array");
  go2_task* t = go2_task_create(n, max_depth, cfg);
  if (!t) return g_last_code;
  int rc = go2_task_upload(t, boards, n);
  if (rc == GO2_OK) rc = go2_task_launch(t);
  if (rc == GO2_OK) rc = go2_task_download(t, scores, stats);
  if (rc == GO2_OK && moves) rc = go2_task_download_moves(t, moves);
  go2_task_destroy(t);
  return rc;
}

go2_task* go2_task_create(size_t n, int max_depth, const go2_config* cfg) {
  go2_task* t = new go2_task;
  t->n = n;
  t->max_depth = max_depth;
  if (resolve_config(cfg, &t->cfg) != GO2_OK || check_depth(max_depth) != GO2_OK) {
    delete t;
    return nullptr;
  }
  hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    hip_fail(e, "hipStreamCreate");
    delete t;
    return nullptr;
  }
  if (t->bufs.alloc(n ? n : 1) != GO2_OK || t->ev.create() != GO2_OK) {
    (void)hipStreamDestroy(t->stream);
    delete t;
    return nullptr;
  }
  return t;
}

int go2_task_upload(go2_task* t, const go2_board* boards, size_t count) {
  if (!t) return fail(GO2_E_ARG, "null task");
  if (count > t->n) return fail(GO2_E_ARG, "upload count exceeds the task size");
  if (count && !boards) return fail(GO2_E_ARG, "null boards");
  if (count) {
    hipError_t e = hipMemcpyAsync(t->bufs.boards, boards, count * sizeof(go2_board),
                                  hipMemcpyHostToDevice, t->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(boards)");
    if ((e = hipStreamSynchronize(t->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  }
  t->count = count;
  t->upload_max_e = max_empties(boards, count);
  t->launched = false;
  return GO2_OK;
}

int go2_task_upload_async(go2_task* t, const go2_board* boards, size_t count) {
  if (!t) return fail(GO2_E_ARG, "null task");
  if (count > t->n) return fail(GO2_E_ARG, "upload count exceeds the task size");
  if (count && !boards) return fail(GO2_E_ARG, "null boards");
  if (count) {
    const hipError_t e = hipMemcpyAsync(t->bufs.boards, boards, count * sizeof(go2_board),
                                        hipMemcpyHostToDevice, t->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(boards)");
  }
  t->count = count;
  t->upload_max_e = max_empties(boards, count);
  t->launched = false;
  return GO2_OK;
}

int go2_host_alloc(size_t bytes, void** out) {
  if (!out) return fail(GO2_E_ARG, "null output");
  *out = nullptr;
  const hipError_t e = hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault);
  if (e != hipSuccess) {
    *out = nullptr;
    return fail(GO2_E_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  return GO2_OK;
}

void go2_host_free(void* ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

int go2_task_launch(go2_task* t) {
  if (!t) return fail(GO2_E_ARG, "null task");
  int rc = GO2_OK;
  // The stack covers the deepest uploaded board (or DEPTH, if that is
  // smaller: deeper boards are then rejected).
  const int depth = t->upload_max_e < t->max_depth ? (t->upload_max_e > 1 ? t->upload_max_e : 1) : t->max_depth;
  if (t->count)
    rc = launch_solve(t->bufs.boards, t->count, depth, t->bufs.scores, t->bufs.moves, t->cfg, t->ws,
                      t->stream, t->ev.a, t->ev.b, t->cfg.graph_launch ? &t->graph : nullptr);
  if (rc == GO2_OK) t->launched = true;
  return rc;
}

int go2_task_query(go2_task* t) {
  if (!t) return fail(GO2_E_ARG, "null task");
  if (!t->launched) return fail(GO2_E_STATE, "task not launched");
  hipError_t e = hipStreamQuery(t->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return hip_fail(e, "hipStreamQuery");
}

int go2_task_wait(go2_task* t) {
  if (!t) return fail(GO2_E_ARG, "null task");
  if (!t->launched) return fail(GO2_E_STATE, "task not launched");
  hipError_t e = hipStreamSynchronize(t->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return GO2_OK;
}

int go2_task_download(go2_task* t, int8_t* scores, go2_stats* stats) {
  if (!t) return fail(GO2_E_ARG, "null task");
  if (!t->launched) return fail(GO2_E_STATE, "task not launched");
  if (t->count && !scores) return fail(GO2_E_ARG, "null scores");
  if (t->count == 0) {
    if (stats) *stats = go2_stats{};
    return GO2_OK;
  }
  int rc = collect_stats(t->ws, t->stream, t->ev.a, t->ev.b, stats);
  if (rc != GO2_OK) return rc;
  hipError_t e = hipMemcpyAsync(scores, t->bufs.scores, t->count, hipMemcpyDeviceToHost, t->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(scores)");
  if ((e = hipStreamSynchronize(t->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return GO2_OK;
}

int go2_think(const go2_board* boards, size_t n, int depth, int8_t* scores, uint8_t* moves,
              const go2_config* cfg, go2_stats* stats) {
  if (depth < 1 || depth > 30) return fail(GO2_E_ARG, "think depth must be in [1, 30]");
  go2_config c;
  int rc;
  if ((rc = resolve_config(cfg, &c)) != GO2_OK) return rc;
  c.think_depth = depth;
  c.ordering = GO2_ORDER_LEARNED;
  return go2_solve_moves(boards, n, 60, scores, moves, &c, stats);
}

int go2_task_download_moves(go2_task* t, uint8_t* moves) {
  if (!t) return fail(GO2_E_ARG, "null task");
  if (!t->launched) return fail(GO2_E_STATE, "task not launched");
  if (t->count && !moves) return fail(GO2_E_ARG, "null moves");
  if (t->count == 0) return GO2_OK;
  hipError_t e = hipStreamSynchronize(t->stream);
  if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  if ((e = hipMemcpyAsync(moves, t->bufs.moves, t->count, hipMemcpyDeviceToHost, t->stream)) != hipSuccess)
    return hip_fail(e, "hipMemcpyAsync(moves)");
  if ((e = hipStreamSynchronize(t->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return GO2_OK;
}

void go2_task_destroy(go2_task* t) {
  if (!t) return;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  t->graph.reset();
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
}

int go2_generate_device(uint64_t seed, uint32_t config_id, size_t n, int empties_min,
                        int empties_max, go2_board* d_out, void* hip_stream) {
  if (empties_min < 0 || empties_max > 60 || empties_min > empties_max)
    return fail(GO2_E_ARG, "need 0 <= empties_min <= empties_max <= 60");
  if (n && !d_out) return fail(GO2_E_ARG, "null output");
  return launch_generate(seed, config_id, n, empties_min, empties_max, d_out,
                         (hipStream_t)hip_stream);
}

int go2_generate(uint64_t seed, uint32_t config_id, size_t n, int empties_min, int empties_max,
                 go2_board* out) {
  if (n && !out) return fail(GO2_E_ARG, "null output");
  if (n == 0) return GO2_OK;
  go2_board* d = nullptr;
  hipError_t e = hipMalloc(&d, n * sizeof(go2_board));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(generate)");
  int rc = go2_generate_device(seed, config_id, n, empties_min, empties_max, d, nullptr);
  if (rc == GO2_OK) {
    e = hipMemcpy(out, d, n * sizeof(go2_board), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy(generate)");
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
