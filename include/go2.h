/*
 * go2 — C ABI of the MI355X (gfx950) batched exact Othello endgame solver.
 *
 * Plain C types only (no torch, no HIP types in signatures).  Every call
 * returns GO2_OK (0) or a negative GO2_E* code; go2_last_error() gives the
 * message of the calling thread's last failure.  The library owns every
 * device buffer it allocates; callers own their host arrays.  A task handle
 * is used by one host thread at a time; many handles may run concurrently,
 * each on its own HIP stream.
 *
 * Board layout (all entry points): square = row*8 + col, bit 0 = A1,
 * bit 7 = H1, bit 63 = H8; `player` is the side to move.
 * Scores are exact disc differences in [-64, 64] from the side to move,
 * winner taking the empty squares (reference board.cuh:7-18).
 */
#ifndef GO2_H
#define GO2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GO2_OK 0
#define GO2_E_ARG -1     /* bad argument (null pointer, size, depth) */
#define GO2_E_BOARD -2   /* invalid board: overlapping discs, or more empties than DEPTH
                            (or, in the exact solve, than GO2_MAX_EMPTIES) */
#define GO2_E_DEVICE -3  /* HIP runtime failure */
#define GO2_E_NOMEM -4   /* allocation failure */
#define GO2_E_STATE -5   /* call out of order (e.g. download before launch) */
#define GO2_E_PARSE -6   /* malformed base81 text */
#define GO2_E_TIMEOUT -7 /* a solve ran past config.time_limit_ms and was abandoned */

/* Largest number of empty squares one board of an exact solve may have
 * (DEPTH itself may be anything in 1..60, as the reference's is). */
#define GO2_MAX_EMPTIES 32

/* Best-move codes besides squares 0..63 (bit 0 = A1), the reference's
 * hand::PASS / hand::NOMOVE (alphabeta.cuh:8-9). */
#define GO2_MOVE_PASS 64 /* the side to move has no legal move; the opponent has */
#define GO2_MOVE_NONE 65 /* game over (or an invalid board) */

/* Move ordering of the solver. */
#define GO2_ORDER_STATIC 0   /* corners, then plain squares, X-squares last */
#define GO2_ORDER_MOBILITY 1 /* fastest-first (fewest replies) above a cutoff, parity below */
#define GO2_ORDER_LEARNED 2  /* as MOBILITY, plus the pattern evaluator above a higher cutoff */

/* One input position.  Replaces AlphaBetaProblem (reference alphabeta.cuh:12-21);
 * the window is always the full (-64, 64) one the reference's CLI passes
 * (reference main.cu:140), so it is not part of the record. */
typedef struct go2_board {
  uint64_t player;
  uint64_t opponent;
} go2_board;

/* Launch and search knobs.  go2_config_default() fills a valid set; zero in
 * an "auto" field lets the library choose from the batch and the device. */
typedef struct go2_config {
  int32_t ordering;             /* GO2_ORDER_* */
  int32_t mobility_min_empties; /* fastest-first at nodes with >= this many empties */
  int32_t learned_min_empties;  /* evaluator ordering at nodes with >= this many empties */
  int32_t block_size;           /* lanes per workgroup, multiple of 64, 64..256 */
  int32_t blocks_per_cu;        /* resident workgroups per CU to launch; 0 = as many as fit */
  int32_t positions_per_wave;   /* boards a wavefront owns at once (1..64); 0 = auto from n.
                                   Lanes without a board of their own help the others. */
  int32_t split_min_empties;    /* idle lanes take sibling subtrees of nodes with >= this many
                                   empties (young-brothers-wait); 0 = auto, >60 = never */
  int32_t lds_levels;           /* stack levels (from 4 empties up) kept in LDS, the rest in HBM;
                                   0 = as many as fit a workgroup's LDS share (9 for 256 lanes) */
  int32_t time_limit_ms;        /* abandon a launch after this long (GO2_E_TIMEOUT); 0 = no limit */
  int32_t split_cut_after;      /* at nodes expected to cut, helpers wait for this many children
                                   (1..3) instead of one; 0 = auto */
  int32_t learned_key_eval;     /* GO2_ORDER_LEARNED move key = eval(child)*this/4096 (evaluator
                                   units) + replies*learned_key_mobility; 0 = auto.  At nodes
                                   with >= 15 empties eval(child) is one ply deeper: the
                                   largest -eval over the child's replies */
  int32_t learned_key_mobility;
  int32_t max_pops_per_trip;    /* finished nodes a lane returns from per loop trip; 0 = auto (2) */
  int32_t tt_log2_buckets;      /* transposition table of 2^n 16-byte buckets shared by the whole
                                   launch (64-bit keys hashed from both bitboards, lock-free);
                                   0 = no table */
  int32_t tt_min_empties;       /* probe/store nodes with >= this many empties; 0 = auto (12) */
  int32_t graph_launch;         /* tasks capture their launch (counter reset + solve kernel) in a
                                   HIP graph once and replay it on later launches */
  int32_t export_min_empties;   /* while some wavefront is idle, busy wavefronts hand siblings of
                                   nodes with >= this many empties to idle wavefronts; 0 = auto
                                   (12), > 60 = never */
  int32_t export_low_empties;   /* the export level falls linearly from export_min_empties to this
                                   one as the idle fraction of wavefronts approaches one; 0 = auto
                                   (9), never below split_min_empties */
  int32_t think_depth;          /* > 0: the reference's think mode instead of the exact solve: the
                                   position after this many moves is scored by the installed
                                   evaluator (go2_think) */
  int32_t leaf_batch_min;       /* nodes with three empties queue all their children, and a
                                   wavefront scores queued children together once this many
                                   (1..64) are waiting; 0 = auto (32), > 64 = each lane scores its
                                   own, one child per loop trip */
  int32_t score_batch_min;      /* nodes that order their moves queue every child, and a
                                   wavefront computes the queued ordering keys together once
                                   this many (1..64) are waiting; 0 = auto (64), > 64 = each lane
                                   scores its own, one child per loop trip */
  int32_t aspiration_width;     /* GO2_ORDER_LEARNED: each board (>= 10 empties) is first searched
                                   with the window evaluator estimate +- this many discs, and
                                   searched again with one bound open when its value falls outside
                                   (scores are exact either way); 0 = full window; default 8 */
  int32_t pvs_min_empties;      /* principal variation search: after its first child, a node with
                                   >= this many empties (>= 5) searches each child with a null
                                   window first and again with a full one when it fails high
                                   inside the window; 0 = off; default 8 */
  int32_t reserved[1];
} go2_config;

/* Counters of one solve.  Replaces the reference's `total` node counter
 * (reference solver.cu:308-309) with defined meanings. */
typedef struct go2_stats {
  uint64_t positions;   /* boards solved */
  uint64_t nodes;       /* positions visited: roots, pushed children, leaves scored in place */
  uint64_t score_steps; /* child evaluations spent only on move ordering */
  uint64_t iterations;  /* lane-iterations of the search loop (sum over lanes) */
  double kernel_ms;     /* device time of the solve kernel(s), HIP events on the solve stream */
  uint64_t steals;      /* subtrees handed from a busy lane to an idle one */
  uint64_t lane_slots;  /* lanes launched x loop trips of their wavefront (busy + idle) */
  uint64_t grid_waves;  /* wavefronts launched */
  uint64_t tt_probes;   /* transposition-table lookups */
  uint64_t tt_hits;     /* lookups that found the position */
  uint64_t tt_cuts;     /* hits that settled the child without searching it */
  uint64_t tt_stores;   /* entries written */
  uint64_t exports;     /* siblings handed to other wavefronts */
} go2_stats;

/* ---- library ---------------------------------------------------------------- */
/* 2: best moves (go2_solve_moves, go2_solve_device_moves,
 *    go2_task_download_moves), think mode (go2_think, config.think_depth). */
int go2_version(void);
const char *go2_last_error(void);
/* Select the HIP device for the calling thread's subsequent calls. */
int go2_set_device(int device);
void go2_config_default(go2_config *cfg);

/* ---- board text (replaces toBoard/fromBoard, reference to_board.hpp:6-9,
 *      to_board.cpp:15-60) ------------------------------------------------------- */
/* Decode the first 16 characters of `text`.  GO2_E_PARSE on a character that
 * is no base81 digit group (the reference decodes garbage silently). */
int go2_board_from_text(const char *text, go2_board *out);
/* Encode to 16 characters plus a terminating NUL. */
void go2_board_to_text(const go2_board *board, char out[17]);

/* ---- one-shot solve ------------------------------------------------------------ */
/* Host arrays in, host scores out; blocks until done.  max_depth is the
 * reference's DEPTH (maximum number of empties, reference README.md:48-50,
 * main.cu:127), 1..60: a board with more empties than max_depth or than
 * GO2_MAX_EMPTIES fails the call with GO2_E_BOARD (the reference overruns
 * its stack instead).  Replaces the solve path of reference main.cu:123-173
 * minus file I/O. */
int go2_solve(const go2_board *boards, size_t n, int max_depth, int8_t *scores,
              const go2_config *cfg, go2_stats *stats);

/* Device arrays in and out (inputs already resident in HBM).  Enqueued on
 * `hip_stream` (a hipStream_t, or NULL for the library's own stream) and
 * waited for before returning, so `stats` is complete on return. */
int go2_solve_device(const go2_board *d_boards, size_t n, int max_depth, int8_t *d_scores,
                     const go2_config *cfg, go2_stats *stats, void *hip_stream);

/* As go2_solve / go2_solve_device, plus one best move per board in `moves`:
 * a square whose child has the board's exact score (any one of them when
 * several do), GO2_MOVE_PASS or GO2_MOVE_NONE.  The reference returns a best
 * move only from think mode (thinker.cu:169, printed by main.cu:114-117);
 * this is the solve path's counterpart. */
int go2_solve_moves(const go2_board *boards, size_t n, int max_depth, int8_t *scores,
                    uint8_t *moves, const go2_config *cfg, go2_stats *stats);
int go2_solve_device_moves(const go2_board *d_boards, size_t n, int max_depth, int8_t *d_scores,
                           uint8_t *d_moves, const go2_config *cfg, go2_stats *stats,
                           void *hip_stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): go2_solve_device_moves without any allocation.  The
 * solve's counters, help slots, mailboxes and HBM stack frames live in one
 * static device array of GO2_NOALLOC_SCRATCH_BYTES that each device the
 * library runs on holds until the process ends; the caller passes boards,
 * scores and moves (may be NULL) in device memory.  The scratch does not grow
 * with n; when a DEPTH's stack does not fit it at full occupancy, the grid
 * shrinks (fewer workgroups per CU, then no cross-wavefront help), and a
 * DEPTH no grid fits fails with GO2_E_ARG.  No transposition table
 * (config->tt_log2_buckets must be 0); stats->kernel_ms comes from the device
 * clock.  Every call on one device shares that device's array and leaves its
 * contents behind (each call resets what must start at zero): calls that
 * overlap on one GPU race, so serialise them (one stream, or a lock).
 * Waits for the solve like go2_solve_device_moves. */
#define GO2_NOALLOC_SCRATCH_BYTES (128u << 20)
int go2_solve_device_noalloc(const go2_board *d_boards, size_t n, int max_depth, int8_t *d_scores,
                             uint8_t *d_moves, const go2_config *cfg, go2_stats *stats,
                             void *hip_stream);

/* Think mode (replaces BatchedThinkTask, reference thinker.cu:351-389 and
 * main.cu:69-121): depth-limited negamax over `depth` (1..30) moves from each
 * board (passes cost no depth; a game that ends first scores exactly), the
 * position after the depth-th move scored by the evaluator installed with
 * go2_set_evaluator (rounded to whole discs, clamped to [-64, 64]).  Boards
 * may have up to 60 empties.  scores[i] and moves[i] (may be NULL) as
 * go2_solve_moves.  The reference's own weights (value/value52) are not
 * shipped with it, so its think-mode numbers cannot be reproduced; the
 * search rule is the same. */
int go2_think(const go2_board *boards, size_t n, int depth, int8_t *scores, uint8_t *moves,
              const go2_config *cfg, go2_stats *stats);

/* ---- batched task (replaces BatchedTask, reference solver.cuh:8-25,
 *      solver.cu:313-348) ------------------------------------------------------------ */
typedef struct go2_task go2_task;
/* Allocate device buffers for n boards and a private stream (reference
 * solver.cu:313-323).  NULL on failure (see go2_last_error). */
go2_task *go2_task_create(size_t n, int max_depth, const go2_config *cfg);
/* Copy boards [0, count) into the task (the reference's `abp[j] = ...`,
 * main.cu:138-141).  count <= n. */
int go2_task_upload(go2_task *task, const go2_board *boards, size_t count);
/* As go2_task_upload, but only enqueued on the task's stream: the copy runs
 * while earlier work on other streams (another task's solve) still runs,
 * and a go2_task_launch after it starts once the copy is done.  `boards`
 * must stay valid and unchanged until the task has finished (go2_task_wait
 * or go2_task_query == 1); only memory from go2_host_alloc (pinned) is
 * copied without the caller waiting.  Replaces the host writes into the
 * task's managed input array (reference main.cu:138-141) for pipelines
 * that upload batch k+1 while batch k solves (main.cu:135-155). */
int go2_task_upload_async(go2_task *task, const go2_board *boards, size_t count);
/* Enqueue the solve asynchronously (reference solver.cu:330-333). */
int go2_task_launch(go2_task *task);
/* 1 when finished, 0 while running, negative on error (reference
 * solver.cu:335-337 — which reports an error as "ready"; this one does not). */
int go2_task_query(go2_task *task);
/* Block until finished. */
int go2_task_wait(go2_task *task);
/* Copy scores (count entries) and stats out (the reference's result[] and
 * *total, main.cu:165-171).  Waits for the task first. */
int go2_task_download(go2_task *task, int8_t *scores, go2_stats *stats);
/* Copy the best moves (count entries, codes as go2_solve_moves) of the last
 * launch out.  Waits for the task first. */
int go2_task_download_moves(go2_task *task, uint8_t *moves);
/* Free device buffers and the stream (reference solver.cu:339-348). */
void go2_task_destroy(go2_task *task);

/* Page-locked host memory for go2_task_upload_async (replaces the managed
 * input array the reference allocates per batch, solver.cu:317-319);
 * go2_host_free releases it.  GO2_E_NOMEM when the allocation fails. */
int go2_host_alloc(size_t bytes, void **out);
void go2_host_free(void *ptr);

/* ---- synthetic positions -------------------------------------------------------- */
/* Random legal playouts from the standard start (passing when forced) to a
 * fixed number of empties; a game that ends first is replayed with the next
 * attempt number.  Position i uses the seed-derived stream (seed, config_id,
 * i, attempt); its empties are drawn uniformly from [empties_min,
 * empties_max].  Runs on the GPU; writes n boards to host memory `out`.
 * No counterpart in the reference (its inputs come from files). */
int go2_generate(uint64_t seed, uint32_t config_id, size_t n, int empties_min, int empties_max,
                 go2_board *out);
/* Same, writing device memory. */
int go2_generate_device(uint64_t seed, uint32_t config_id, size_t n, int empties_min,
                        int empties_max, go2_board *d_out, void *hip_stream);

/* ---- learned move ordering: pattern evaluator ------------------------------------
 * A linear model over 46 ternary pattern instances (11 classes, int16 weight
 * per class configuration) plus one bias per number of empties, predicting
 * the exact disc difference for the side to move.  Replaces the reference's
 * float Evaluator (reference eval.cuh:5-17, eval.cu:12-142), whose weight
 * file it does not ship; here the model is trained by the kernels below on
 * scores the solver produces, and orders moves at shallow nodes of
 * GO2_ORDER_LEARNED solves (scores never depend on it). */
/* Sizes: instances per board (46), weights (167265), bias terms (64). */
int go2_pattern_info(int *n_instances, int *n_weights, int *n_bias);
/* Global weight index of every instance: d_feat[n*46] (uint32), and the
 * number of empties of every board: d_empties[n]. */
int go2_pattern_features_device(const go2_board *d_boards, size_t n, uint32_t *d_feat,
                                uint8_t *d_empties, void *hip_stream);
/* int32 evaluation sum_k w[f_k] + bias[empties] of every board. */
int go2_pattern_eval_device(const int16_t *d_w, const int16_t *d_bias, const go2_board *d_boards,
                            size_t n, int32_t *d_out, void *hip_stream);
/* SGD forward + gradient of the batch squared error against d_y (FP32 master
 * weights): adds 2*(pred-y) per occurrence into d_grad[n_weights+n_bias],
 * occurrence counts into d_count, and the sum of squared residuals into
 * *d_sse.  d_grad/d_count must start zeroed (go2_sgd_apply_device leaves them
 * zeroed). */
int go2_sgd_gradient_device(const float *d_w, const float *d_bias, const uint32_t *d_feat,
                            const uint8_t *d_empties, const float *d_y, size_t n, float *d_grad,
                            uint32_t *d_count, double *d_sse, void *hip_stream);
/* w -= lr * grad / count for every weight with count > 0; clears grad and count. */
int go2_sgd_apply_device(float *d_w, float *d_bias, float *d_grad, uint32_t *d_count, float lr,
                         void *hip_stream);
/* int16 copy of the model: round(w * scale), clamped to +-32767. */
int go2_quantize_weights_device(const float *d_w, const float *d_bias, float scale, int16_t *d_w16,
                                int16_t *d_bias16, void *hip_stream);
/* Install the int16 evaluator (host arrays of n_weights and n_bias entries;
 * `scale` = units per disc) used by GO2_ORDER_LEARNED solves and think mode
 * on the current device.  Without one, those fail with GO2_E_STATE.  The
 * tables are replaced in place: call it before launching such solves, not
 * while any of them runs on the device. */
int go2_set_evaluator(const int16_t *w, const int16_t *bias, float scale);

#ifdef __cplusplus
}
#endif
#endif /* GO2_H */
