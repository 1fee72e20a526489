/*
 * CPU oracle for the batched exact Othello endgame solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark link it.  The
 * product (the HIP solver behind include/go2.h) never calls it.
 *
 * It restates, in plain C, what the reference computes on the solve path:
 *   - final score            reference board.cuh:7-18
 *   - flip / mobility        reference board.cu:47-105
 *   - move iteration order   reference node.cuh:21-28 (corner first, then LSB)
 *   - fail-soft negamax with pass handling and the upper/lower ply split
 *                            reference solver.cu:54-283, node.cuh:56-74
 *   - base81 board text      reference test.cpp:7-19 (closed form), to_board.cpp:15-60
 *   - pattern index math     reference eval.cu:56-142 (base-3 cell weights, me=1, op=2)
 * plus checks that need no reference at all: a ray-walk flip, a
 * flip-based mobility and a no-pruning minimax.
 *
 * Bit layout: square = row*8 + col, bit 0 = A1, bit 7 = H1, bit 63 = H8.
 * P is the side to move, O the other side.
 */
#ifndef GO2_ORACLE_H
#define GO2_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- bitboard primitives ------------------------------------------------ */
int      orc_final_score(uint64_t P, uint64_t O);
uint64_t orc_flip(uint64_t P, uint64_t O, int sq);       /* line-mask method */
uint64_t orc_flip_walk(uint64_t P, uint64_t O, int sq);  /* brute-force ray walk */
uint64_t orc_mobility(uint64_t P, uint64_t O);           /* shift-fill method */
uint64_t orc_mobility_walk(uint64_t P, uint64_t O);      /* via orc_flip_walk */
uint64_t orc_next_bit(uint64_t untried);                 /* corner first, then LSB */

/* ---- base81 text --------------------------------------------------------- */
void orc_encode(uint64_t P, uint64_t O, char out[17]);
int  orc_decode(const char *s, uint64_t *P, uint64_t *O); /* 0 ok, -1 bad char */

/* ---- exact search --------------------------------------------------------- */
/* Reference method: fastest-first ordering in the top (depth-9) plies,
 * corner-first/LSB below, window (-64, 64), no table.  nodes may be NULL. */
int orc_solve_ref(uint64_t P, uint64_t O, int depth, uint64_t *nodes);
/* Plain negamax with no pruning at all (tiny inputs only). */
int orc_minimax(uint64_t P, uint64_t O, uint64_t *nodes);
/* Multithreaded driver over orc_solve_ref; returns elapsed seconds of the solve. */
double orc_solve_batch(const uint64_t *P, const uint64_t *O, size_t n, int depth,
                       int threads, int8_t *scores, uint64_t *nodes);

/* Optimal moves of (P, O): *optimal = the legal squares whose child solves
 * to the exact score; *kind = ORC_MOVES_PLAY, or ORC_MOVES_PASS (P has no
 * move, O has) / ORC_MOVES_GAME_OVER with *optimal = 0.  Returns the score. */
#define ORC_MOVES_PLAY 0
#define ORC_MOVES_PASS 1
#define ORC_MOVES_GAME_OVER 2
int  orc_best_moves(uint64_t P, uint64_t O, int depth, uint64_t *optimal, int *kind);
void orc_best_moves_batch(const uint64_t *P, const uint64_t *O, size_t n, int depth, int threads,
                          int8_t *scores, uint64_t *optimal, uint8_t *kinds);

/* Depth-limited search, the reference's think mode (thinker.cu:120-320):
 * negamax over `depth` (>= 1) disc placements, passes free, exact at game
 * end, the evaluator (orc_eval_int / scale, rounded half away from zero,
 * clamped to [-64, 64]) after the depth-th move.  Returns the value; the
 * optimal squares and kind as orc_best_moves. */
int  orc_think_leaf(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O);
int  orc_think(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O, int depth,
               uint64_t *optimal, int *kind);
/* The value alone, by fail-soft alpha-beta (fast enough for deeper tests). */
int  orc_think_value(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O, int depth);
void orc_think_batch(const int16_t *w, const int16_t *bias, float scale, const uint64_t *P, const uint64_t *O,
                     size_t n, int depth, int threads, int8_t *scores, uint64_t *optimal, uint8_t *kinds);

/* ---- synthetic positions ---------------------------------------------------- */
uint64_t orc_splitmix64(uint64_t *state);
uint64_t orc_position_key(uint64_t seed, uint32_t config, uint64_t index, uint32_t attempt);
int      orc_empties_for(uint64_t seed, uint32_t config, uint64_t index, int emin, int emax);
/* Random playout from the standard start to exactly `empties` empty squares.
 * Returns the number of attempts used (>=1), or -1 if none succeeded. */
int orc_generate(uint64_t seed, uint32_t config, uint64_t index, int empties,
                 uint64_t *P, uint64_t *O);

/* ---- pattern evaluator --------------------------------------------------------- */
#define ORC_N_INSTANCES 46
#define ORC_N_CLASSES 11
#define ORC_N_BIAS 64
int  orc_pattern_layout(int *cls, int *ncell, int *cells /* [46*10] */, int *class_offset /* [12] */);
void orc_pattern_indices(uint64_t P, uint64_t O, uint32_t *feat /* [46], global weight index */);
/* Index of a ternary pattern exactly as reference eval.cu:113-121 computes it:
 * cell j (j-th set bit of mask, from LSB) has weight 3^j, me=1, op=2. */
uint32_t orc_ref_pattern_index(uint64_t me, uint64_t op, uint64_t mask);
uint64_t orc_pext(uint64_t x, uint64_t mask);
uint64_t orc_flip_vertical(uint64_t x);
uint64_t orc_flip_diag(uint64_t x);
uint64_t orc_rot90(uint64_t x);
int32_t  orc_eval_int(const int16_t *w, const int16_t *bias, uint64_t P, uint64_t O);
float    orc_eval_float(const float *w, const float *bias, uint64_t P, uint64_t O);
/* Count-normalised SGD step (mirror of the GPU training kernels):
 *   pred_i = sum_k w[f_ik] + bias[e_i];  r_i = pred_i - y_i
 *   g_f = sum_{i,k: f_ik=f} 2 r_i;  n_f = occurrences
 *   w_f -= lr * g_f / n_f;  bias likewise.   Returns the batch mean squared error. */
double orc_sgd_step(float *w, float *bias, const uint32_t *feat, const uint8_t *empt,
                    const float *y, size_t n, float lr, float *grad_out, uint32_t *cnt_out);

#ifdef __cplusplus
}
#endif
#endif
