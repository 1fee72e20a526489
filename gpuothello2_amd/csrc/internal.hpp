// Internal declarations shared by the HIP sources of libgo2.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/go2.h"

namespace go2 {

// Device-side counters of one solve; reset by a fill kernel (never by a
// memset node, so the launch sequence stays capturable in a HIP graph).
struct SolveCounters {
  unsigned int next;   // work-queue head: next board index to hand out
  unsigned int error;  // 1 + index of the first invalid board, kTimedOut, or 0
  unsigned long long positions;
  unsigned long long nodes;
  unsigned long long score_steps;
  unsigned long long iterations;
  unsigned long long steals;
  unsigned long long lane_slots;
  unsigned long long tt_probes;
  unsigned long long tt_hits;
  unsigned long long tt_cuts;
  unsigned long long tt_stores;
  unsigned int tt_gen;  // launch generation 1..255 (advanced by the reset kernel, never zeroed)
  // Cross-wavefront help.  A wavefront counts itself active when it starts
  // (not at launch: a launched wavefront need not be resident yet) and while
  // it holds work; idle ones poll for exported siblings and leave once no
  // wavefront is active.
  unsigned int active_waves;
  unsigned int idle_waves;  // wavefronts polling for exported siblings
  unsigned long long exports;
  // Device clock (s_memrealtime, 100 MHz) at the reset and, atomically
  // maximised, when the last wavefront of the solve left.
  unsigned long long t_start;
  unsigned long long t_end;
};

constexpr unsigned int kTimedOut = 0xFFFFFFFFu;

// Records a failure for go2_last_error() and returns `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// Validates cfg (or fills defaults when null) into `out`.
int resolve_config(const go2_config* cfg, go2_config* out);

// Device memory of a solve that outlives one launch: counters and the
// per-lane search stacks kept in HBM.  Grown on demand, freed on destruction;
// or, in fixed mode (`fixed` set), carved from caller-provided device memory
// and never allocated or freed.
struct Workspace {
  uint8_t* fixed = nullptr;  // fixed mode: base of the region, fixed_cap bytes
  size_t fixed_cap = 0;
  SolveCounters* ctr = nullptr;
  void* frames = nullptr;
  size_t frames_bytes = 0;
  unsigned long long grid_waves = 0;  // of the last launch
  // transposition table: 2^tt_log2 buckets of 16 bytes; entries of older
  // launches are ignored through an 8-bit generation
  void* tt = nullptr;
  int tt_log2 = 0;
  unsigned int tt_gen = 0;
  // cross-wavefront help: slots and mailboxes
  void* help = nullptr;
  size_t help_bytes = 0;
  Workspace() = default;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
  ~Workspace();
  int ensure(size_t frame_bytes);
  int ensure_tt(int log2_buckets, hipStream_t stream);
  int ensure_help(size_t bytes, hipStream_t stream);
  // Fixed mode: counters, then help_bytes of help slots and mailboxes, then
  // frame_bytes of frames, within fixed_cap.
  int carve(size_t help_bytes, size_t frame_bytes);
};

// Device memory of the no-allocation entry point on the current device (a
// static device array, its address looked up once per device).
int noalloc_scratch(uint8_t** base, size_t* bytes);

// A launch sequence (counter reset + solve kernel) captured in a HIP graph,
// re-captured when any launch argument changes.
struct GraphCache {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned char key[512] = {0};
  size_t key_len = 0;
  GraphCache() = default;
  GraphCache(const GraphCache&) = delete;
  GraphCache& operator=(const GraphCache&) = delete;
  ~GraphCache();
  void reset();
};

// Enqueue counter reset + solve kernel on `stream` (directly, or as a replay
// of a captured graph when `graph` is non-null); events bracket the launch
// when non-null.
// d_moves (may be null) receives each board's best move: a square 0..63,
// GO2_MOVE_PASS or GO2_MOVE_NONE.
int launch_solve(const go2_board* d_boards, size_t n, int max_depth, int8_t* d_scores, uint8_t* d_moves,
                 const go2_config& cfg, Workspace& ws, hipStream_t stream, hipEvent_t ev0,
                 hipEvent_t ev1, GraphCache* graph = nullptr);

// Copy counters back (stream-ordered, then synchronises the stream) and fill
// stats; kernel_ms from the events when given, else from the device clock.
int collect_stats(Workspace& ws, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1,
                  go2_stats* stats);

int launch_generate(uint64_t seed, uint32_t config_id, size_t n, int emin, int emax,
                    go2_board* d_out, hipStream_t stream);

}  // namespace go2
