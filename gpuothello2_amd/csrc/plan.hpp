// Launch plan of one solve: stack levels, LDS bytes, grid, boards per
// wavefront and the help levels, computed on the host from the batch size,
// DEPTH, the configuration and the device's shape.  Plain C++ with no HIP
// dependency, so a CPU harness (tests/plan_check.cpp) checks the same code
// the library runs.
//
// Counterpart in the reference: the launch arithmetic of BatchedTask
// (solver.cu:313-333: grid ceil(n/2048) x 64 threads, DEPTH-9 upper frames,
// 10 shared-memory levels), redone for a resident persistent grid.
#pragma once
#include <cstddef>

#include "../../include/go2.h"

namespace go2 {

// Lowest stack level: nodes with three empties solve their children in
// straight-line code, so only nodes with four or more push frames.
constexpr int kFrameBase = 4;
constexpr size_t kLastTableBytes = 2048;   // last-square flip counts (LDS)
constexpr size_t kBase3TableBytes = 2048;  // base-3 table of the evaluator (LDS)
constexpr size_t kLeafMapBytesPerWave = 64; // leaf batch: worker slot -> requesting lane (LDS)
// The solve kernel is built for two wavefronts per SIMD (its launch bound:
// 256 VGPRs, no spills), so one CU holds 8 wavefronts and each workgroup
// may use its share of the CU's 160 KB of LDS for stack levels.
constexpr int kWavesPerSimd = 2;
constexpr size_t kLdsBytesPerCU = 160 * 1024;
constexpr size_t kFrameBytesPerLevel = 32;  // fb, mv, od (u64) + wd, sx (u32) per lane
constexpr unsigned int kHelpSlots = 4096;
constexpr size_t kMaxHelpLanes = size_t(1) << 19;  // slot tags hold global lane ids below 2^19
constexpr int kExtRec = 4;  // siblings of one node out with other wavefronts at once
constexpr size_t kHelpSlotBytes = 64;   // one exported sibling (HelpSlot)
constexpr size_t kCounterBytes = 256;   // the launch counters (SolveCounters), rounded up

struct StackPlan {
  bool think = false;
  int depth = 0;       // boards with more empties are rejected
  int maxe = 0;        // highest stack level
  int levels = 0;      // frame levels kFrameBase..maxe
  int lds_levels = 0;  // of which in LDS (the lowest ones)
  size_t lds_bytes = 0;
};

// max_depth is the caller's DEPTH (1..60).  The exact solve handles boards
// with up to GO2_MAX_EMPTIES empties (one bit per level in 32-bit masks);
// its stack is sized for min(DEPTH, GO2_MAX_EMPTIES).  Think mode searches
// think_depth moves from any position (stack levels relative to each
// board's horizon).
inline StackPlan plan_stack(int max_depth, const go2_config& cfg) {
  StackPlan s;
  s.think = cfg.think_depth > 0;
  if (s.think) {
    s.depth = 60;
    s.maxe = cfg.think_depth + 2;
  } else {
    s.depth = max_depth < GO2_MAX_EMPTIES ? max_depth : GO2_MAX_EMPTIES;
    s.maxe = s.depth < 3 ? 3 : s.depth;
  }
  s.levels = s.maxe + 1 - kFrameBase;
  if (s.levels < 1) s.levels = 1;
  const size_t bs = (size_t)cfg.block_size;
  const size_t fixed = kLastTableBytes + (cfg.ordering == GO2_ORDER_LEARNED || s.think ? kBase3TableBytes : 0) +
                       bs / 64 * kLeafMapBytesPerWave;
  if (cfg.lds_levels > 0) {
    s.lds_levels = cfg.lds_levels;
  } else {
    // as many levels as fit this workgroup's share of the CU's LDS
    const size_t blocks_per_cu = (size_t)(4 * kWavesPerSimd) / (bs / 64);
    const size_t share = kLdsBytesPerCU / (blocks_per_cu ? blocks_per_cu : 1);
    const size_t fit = share > fixed ? (share - fixed) / (bs * kFrameBytesPerLevel) : 0;
    s.lds_levels = fit < 1 ? 1 : (int)fit;
  }
  if (s.lds_levels > s.levels) s.lds_levels = s.levels;
  s.lds_bytes = (size_t)s.lds_levels * bs * kFrameBytesPerLevel + fixed;
  return s;
}

struct GridPlan {
  size_t grid = 0;         // workgroups
  size_t waves = 0;        // wavefronts
  size_t total_lanes = 0;  // grid * block size
  size_t g_levels = 0;     // stack levels in HBM
  int K = 1;               // boards one wavefront owns at once
  int split_min = 0;       // helpers take siblings at nodes with >= this many empties
  int ext_min = 3;         // cross-wavefront export levels (when help is on)
  int ext_lo = 3;
  bool help = false;       // cross-wavefront help on
};

// cus: compute units of the device; per_cu: workgroups of this kernel that
// stay resident on one CU (from the occupancy query, >= 1).
inline GridPlan plan_grid(const StackPlan& s, size_t n, const go2_config& cfg, int cus, int per_cu) {
  GridPlan g;
  if (cfg.blocks_per_cu > 0 && cfg.blocks_per_cu < per_cu) per_cu = cfg.blocks_per_cu;
  const size_t bs = (size_t)cfg.block_size;
  const size_t waves_per_block = bs / 64;
  g.grid = (size_t)cus * (size_t)per_cu;
  int split_min = cfg.split_min_empties > 0 ? cfg.split_min_empties : 8;
  if (split_min < 5) split_min = 5;
  if (split_min > s.maxe + 1) split_min = s.maxe + 1;  // no split levels at all
  if (s.think) split_min = 64;                         // one lane per board
  g.split_min = split_min;
  // Export levels: ext_min (default 12) while few wavefronts are idle, down
  // to ext_lo (default 9) when nearly all are; never below split_min.  No
  // node of this stack has more than maxe empties, so every export level
  // above maxe is the same as maxe + 1 (nothing to export): capping there
  // keeps the kernel's level shifts (level - 3) below 32 and every frame and
  // mailbox index in range.
  int ext_min = cfg.export_min_empties > 0
                    ? (cfg.export_min_empties > split_min ? cfg.export_min_empties : split_min)
                    : (split_min > 12 ? split_min : 12);
  if (ext_min > s.maxe + 1) ext_min = s.maxe + 1;
  int lo = cfg.export_low_empties > 0 ? cfg.export_low_empties : 9;
  if (lo < split_min) lo = split_min;
  const int ext_lo = lo < ext_min ? lo : ext_min;
  g.help = !s.think && cfg.export_min_empties <= 60 && ext_lo <= s.maxe &&
           (size_t)cus * (size_t)per_cu * bs < kMaxHelpLanes;
  if (g.help) {
    g.ext_min = ext_min;
    g.ext_lo = ext_lo;
  }
  // Without cross-wavefront help, never launch more wavefronts than boards;
  // with it, the whole resident grid runs and idle wavefronts take siblings.
  const size_t max_grid = (n + waves_per_block - 1) / waves_per_block;
  if (!g.help && g.grid > max_grid) g.grid = max_grid;
  if (g.grid < 1) g.grid = 1;
  g.waves = g.grid * waves_per_block;
  int K = cfg.positions_per_wave;
  if (K <= 0) K = (int)((n + g.waves - 1) / g.waves);
  if (K > 64) K = 64;
  if (K < 1) K = 1;
  g.K = K;
  g.total_lanes = g.grid * bs;
  g.g_levels = (size_t)(s.levels - s.lds_levels);
  return g;
}

inline size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device bytes of the HBM stack frames: fb, mv, od (u64) and wd, sx (u32),
// each [g_levels][total_lanes], each region 256-byte aligned.
inline size_t frame_bytes(const GridPlan& g) {
  const size_t cells = g.total_lanes * g.g_levels;
  return 3 * round256(cells * 8) + 2 * round256(cells * 4);
}

// Device bytes of cross-wavefront help: the slots, then four mailboxes (best
// value, reports, cut flag, reclaim records x kExtRec) per help level and lane.
inline size_t mailbox_bytes(const StackPlan& s, const GridPlan& g) {
  return round256((size_t)(s.maxe + 1 - g.ext_lo) * g.total_lanes * 4);
}
inline size_t help_bytes(const StackPlan& s, const GridPlan& g) {
  if (!g.help) return 0;
  return (size_t)kHelpSlots * kHelpSlotBytes + (3 + kExtRec) * mailbox_bytes(s, g);
}

// Everything a solve without a table keeps in device memory besides its
// boards and results: counters, help slots and mailboxes, HBM frames.
inline size_t scratch_bytes(const StackPlan& s, const GridPlan& g) {
  return kCounterBytes + help_bytes(s, g) + frame_bytes(g);
}

// The grid of plan_grid, shrunk until its scratch fits `capacity` bytes:
// fewer resident workgroups per CU first, then no cross-wavefront help.
// False when even one workgroup per CU without help does not fit.
inline bool plan_grid_within(const StackPlan& s, size_t n, go2_config cfg, int cus, int per_cu,
                             size_t capacity, GridPlan* out) {
  if (cfg.blocks_per_cu > 0 && cfg.blocks_per_cu < per_cu) per_cu = cfg.blocks_per_cu;
  for (;;) {
    const GridPlan g = plan_grid(s, n, cfg, cus, per_cu);
    if (scratch_bytes(s, g) <= capacity) {
      *out = g;
      return true;
    }
    if (per_cu > 1) {
      --per_cu;
    } else if (g.help) {
      cfg.export_min_empties = 61;  // > 60: no cross-wavefront help
    } else {
      return false;
    }
  }
}

}  // namespace go2
