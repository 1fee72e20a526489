/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).  Nothing in the
 * product path links or calls this file; it exists to check the HIP solver
 * and to time the reference method on host cores.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static inline int popc(uint64_t x) { return __builtin_popcountll(x); }

/* ------------------------------------------------------------------------- */
/* final score: reference board.cuh:7-18.  The side with more discs also gets
 * every empty square; a tie is 0. */
int orc_final_score(uint64_t P, uint64_t O) {
  int p = popc(P), o = popc(O);
  if (p == o) return 0;
  if (p > o) return 64 - 2 * o;
  return 2 * p - 64;
}

/* ------------------------------------------------------------------------- */
/* Ray tables, built once by walking the board.  dir d = (dr, dc). */
static const int k_dr[8] = {0, 0, 1, -1, 1, 1, -1, -1};
static const int k_dc[8] = {1, -1, 0, 0, 1, -1, 1, -1};
static uint64_t g_ray[64][8];
static pthread_once_t g_ray_once = PTHREAD_ONCE_INIT;

static void build_rays(void) {
  for (int sq = 0; sq < 64; ++sq) {
    int r0 = sq >> 3, c0 = sq & 7;
    for (int d = 0; d < 8; ++d) {
      uint64_t m = 0;
      int r = r0 + k_dr[d], c = c0 + k_dc[d];
      while (r >= 0 && r < 8 && c >= 0 && c < 8) {
        m |= 1ULL << (r * 8 + c);
        r += k_dr[d];
        c += k_dc[d];
      }
      g_ray[sq][d] = m;
    }
  }
}

/* Flip set of a move, by line masks (restates reference board.cu:47-73:
 * for each ray the first non-opponent square decides; if it holds a player
 * disc, every square before it on the ray flips).  Rays pointing to higher
 * bit numbers take the lowest candidate bit, rays to lower numbers the
 * highest one.  Result 0 means the move is illegal. */
uint64_t orc_flip(uint64_t P, uint64_t O, int sq) {
  pthread_once(&g_ray_once, build_rays);
  uint64_t flipped = 0;
  for (int d = 0; d < 8; ++d) {
    uint64_t ray = g_ray[sq][d];
    uint64_t stop = ray & ~O;
    if (!stop) continue;
    int up = (k_dr[d] * 8 + k_dc[d]) > 0;
    uint64_t first = up ? (stop & (0 - stop)) : (1ULL << (63 - __builtin_clzll(stop)));
    if (!(first & P)) continue;
    uint64_t between = up ? (first - 1) : ~((first << 1) - 1);
    flipped |= ray & between;
  }
  return flipped;
}

/* Brute force: walk each ray square by square. */
uint64_t orc_flip_walk(uint64_t P, uint64_t O, int sq) {
  if (((P | O) >> sq) & 1) return 0;
  int r0 = sq >> 3, c0 = sq & 7;
  uint64_t flipped = 0;
  for (int d = 0; d < 8; ++d) {
    uint64_t run = 0;
    int r = r0 + k_dr[d], c = c0 + k_dc[d];
    while (r >= 0 && r < 8 && c >= 0 && c < 8) {
      uint64_t b = 1ULL << (r * 8 + c);
      if (O & b) {
        run |= b;
      } else {
        if ((P & b) && run) flipped |= run;
        break;
      }
      r += k_dr[d];
      c += k_dc[d];
    }
  }
  return flipped;
}

/* Legal moves by directional shift-fill (restates reference board.cu:75-97:
 * the same set, computed one direction at a time). */
static inline uint64_t shift_dir(uint64_t b, int d) {
  const uint64_t notA = 0xFEFEFEFEFEFEFEFEULL, notH = 0x7F7F7F7F7F7F7F7FULL;
  switch (d) {
    case 0: return (b << 1) & notA;  /* +col */
    case 1: return (b >> 1) & notH;  /* -col */
    case 2: return b << 8;           /* +row */
    case 3: return b >> 8;           /* -row */
    case 4: return (b << 9) & notA;  /* +row +col */
    case 5: return (b << 7) & notH;  /* +row -col */
    case 6: return (b >> 7) & notA;  /* -row +col */
    default: return (b >> 9) & notH; /* -row -col */
  }
}

uint64_t orc_mobility(uint64_t P, uint64_t O) {
  uint64_t empty = ~(P | O), moves = 0;
  for (int d = 0; d < 8; ++d) {
    uint64_t t = shift_dir(P, d) & O;
    for (int i = 0; i < 5; ++i) t |= shift_dir(t, d) & O;
    moves |= shift_dir(t, d);
  }
  return moves & empty;
}

uint64_t orc_mobility_walk(uint64_t P, uint64_t O) {
  uint64_t m = 0;
  for (int sq = 0; sq < 64; ++sq)
    if (orc_flip_walk(P, O, sq)) m |= 1ULL << sq;
  return m;
}

/* Candidate order of the reference's lower plies, node.cuh:21-28:
 * the four corners first, then ascending square index. */
uint64_t orc_next_bit(uint64_t untried) {
  const uint64_t corner = 0x8100000000000081ULL;
  uint64_t c = untried & corner;
  uint64_t src = c ? c : untried;
  return src & (0 - src);
}

/* ------------------------------------------------------------------------- */
/* base81: 16 characters, character i covers squares 4i..4i+3 with weights
 * 1, 3, 9, 32 and digits 0 empty / 1 player / 2 opponent, plus 33.
 * Closed form of reference test.cpp:7-19; decoder inverts it
 * (reference to_board.cpp:15-39). */
void orc_encode(uint64_t P, uint64_t O, char out[17]) {
  static const int w[4] = {1, 3, 9, 32};
  for (int i = 0; i < 16; ++i) {
    int v = 33;
    for (int j = 0; j < 4; ++j) {
      int sq = 4 * i + j;
      if ((P >> sq) & 1) v += w[j];
      if ((O >> sq) & 1) v += 2 * w[j];
    }
    out[i] = (char)v;
  }
  out[16] = '\0';
}

int orc_decode(const char *s, uint64_t *P, uint64_t *O) {
  uint64_t p = 0, o = 0;
  for (int i = 0; i < 16; ++i) {
    int x = (unsigned char)s[i] - 33;
    if (x < 0 || x > 90) return -1;
    int hi = x >> 5, lo = x & 31;
    if (hi > 2 || lo > 26) return -1;
    int dig[4] = {lo % 3, (lo / 3) % 3, lo / 9, hi};
    for (int j = 0; j < 4; ++j) {
      int sq = 4 * i + j;
      if (dig[j] == 1) p |= 1ULL << sq;
      if (dig[j] == 2) o |= 1ULL << sq;
    }
  }
  *P = p;
  *O = o;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Reference method search, restating reference solver.cu:54-283.
 *   upper plies (ply < depth-9): every legal move, sorted ascending by the
 *     mover-to-be's mobility count after the move (stable: ties keep the
 *     corner-first/LSB generation order), result starts at -64
 *     (solver.cu:75-93,129,143);
 *   lower plies: candidates in next-bit order, illegal squares skipped,
 *     result starts at -SHRT_MAX (node.cuh:65, solver.cu:244-270);
 *   fail-soft commit: result = max(result, v), alpha = max(alpha, result),
 *     cut when alpha >= beta (node.cuh:70-73, solver.cu:146-149,225,257);
 *   a side without moves passes with the window negated on the same ply;
 *     two passes in a row score the board (solver.cu:65-73,162-165,214-221,
 *     246-256). */
typedef struct {
  uint64_t nodes;
  int upper_plies;
} ref_ctx;

static int ref_search(ref_ctx *c, uint64_t P, uint64_t O, int alpha, int beta, int ply,
                      int passed) {
  c->nodes++;
  if (ply < c->upper_plies) {
    int sq[64], key[64], n = 0;
    uint64_t untried = ~(P | O);
    while (untried) {
      uint64_t bit = orc_next_bit(untried);
      untried ^= bit;
      int s = __builtin_ctzll(bit);
      uint64_t f = orc_flip(P, O, s);
      if (!f) continue;
      int k = popc(orc_mobility(O ^ f, (P ^ f) | bit));
      int j = n++;
      while (j > 0 && key[j - 1] > k) { /* stable insertion */
        key[j] = key[j - 1];
        sq[j] = sq[j - 1];
        --j;
      }
      key[j] = k;
      sq[j] = s;
    }
    if (n == 0) {
      if (passed) return orc_final_score(P, O);
      return -ref_search(c, O, P, -beta, -alpha, ply, 1);
    }
    int result = -64;
    for (int i = 0; i < n; ++i) {
      uint64_t bit = 1ULL << sq[i];
      uint64_t f = orc_flip(P, O, sq[i]);
      int v = -ref_search(c, O ^ f, (P ^ f) | bit, -beta, -alpha, ply + 1, 0);
      if (v > result) result = v;
      if (result > alpha) alpha = result;
      if (alpha >= beta) break;
    }
    return result;
  }
  int result = -SHRT_MAX, moved = 0;
  uint64_t untried = ~(P | O);
  while (untried && alpha < beta) {
    uint64_t bit = orc_next_bit(untried);
    untried ^= bit;
    uint64_t f = orc_flip(P, O, __builtin_ctzll(bit));
    if (!f) continue;
    moved = 1;
    int v = -ref_search(c, O ^ f, (P ^ f) | bit, -beta, -alpha, ply + 1, 0);
    if (v > result) result = v;
    if (result > alpha) alpha = result;
  }
  if (!moved && !untried) {
    if (passed) return orc_final_score(P, O);
    return -ref_search(c, O, P, -beta, -alpha, ply, 1);
  }
  return result;
}

int orc_solve_ref(uint64_t P, uint64_t O, int depth, uint64_t *nodes) {
  ref_ctx c = {0, depth - 9};
  int v = ref_search(&c, P, O, -64, 64, 0, 0);
  if (nodes) *nodes += c.nodes;
  return v;
}

/* Negamax with no pruning: the definition the search must agree with. */
static int minimax_rec(uint64_t P, uint64_t O, int passed, uint64_t *nodes) {
  ++*nodes;
  uint64_t m = orc_mobility(P, O);
  if (!m) {
    if (passed) return orc_final_score(P, O);
    return -minimax_rec(O, P, 1, nodes);
  }
  int best = -65;
  while (m) {
    uint64_t bit = m & (0 - m);
    m ^= bit;
    uint64_t f = orc_flip(P, O, __builtin_ctzll(bit));
    int v = -minimax_rec(O ^ f, (P ^ f) | bit, 0, nodes);
    if (v > best) best = v;
  }
  return best;
}

int orc_minimax(uint64_t P, uint64_t O, uint64_t *nodes) {
  uint64_t n = 0;
  int v = minimax_rec(P, O, 0, &n);
  if (nodes) *nodes += n;
  return v;
}

/* Multithreaded batch: each worker pulls indices from one atomic counter. */
typedef struct {
  const uint64_t *P, *O;
  size_t n;
  int depth;
  int8_t *scores;
  atomic_size_t next;
  atomic_ullong nodes;
} batch_job;

static void *batch_worker(void *arg) {
  batch_job *j = (batch_job *)arg;
  uint64_t local = 0;
  for (;;) {
    size_t i = atomic_fetch_add(&j->next, 1);
    if (i >= j->n) break;
    j->scores[i] = (int8_t)orc_solve_ref(j->P[i], j->O[i], j->depth, &local);
  }
  atomic_fetch_add(&j->nodes, local);
  return NULL;
}

double orc_solve_batch(const uint64_t *P, const uint64_t *O, size_t n, int depth, int threads,
                       int8_t *scores, uint64_t *nodes) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  batch_job j;
  j.P = P;
  j.O = O;
  j.n = n;
  j.depth = depth;
  j.scores = scores;
  atomic_init(&j.next, 0);
  atomic_init(&j.nodes, 0);
  pthread_t tid[256];
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, batch_worker, &j);
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (nodes) *nodes = atomic_load(&j.nodes);
  return (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
}

/* Optimal moves: every legal square whose child solves to the position's
 * exact score (the definition of an optimal move; any one of them is a valid
 * best move for a solver to return).  Returns the score. */
int orc_best_moves(uint64_t P, uint64_t O, int depth, uint64_t *optimal, int *kind) {
  const uint64_t m = orc_mobility(P, O);
  *optimal = 0;
  if (!m) {
    if (orc_mobility(O, P)) {
      *kind = ORC_MOVES_PASS;
      return -orc_solve_ref(O, P, depth, NULL);
    }
    *kind = ORC_MOVES_GAME_OVER;
    return orc_final_score(P, O);
  }
  *kind = ORC_MOVES_PLAY;
  int vals[64];
  int best = -65;
  for (uint64_t r = m; r; r &= r - 1) {
    const int sq = __builtin_ctzll(r);
    const uint64_t bit = 1ULL << sq;
    const uint64_t f = orc_flip(P, O, sq);
    vals[sq] = -orc_solve_ref(O ^ f, (P ^ f) | bit, depth > 0 ? depth - 1 : 0, NULL);
    if (vals[sq] > best) best = vals[sq];
  }
  for (uint64_t r = m; r; r &= r - 1)
    if (vals[__builtin_ctzll(r)] == best) *optimal |= r & (0 - r);
  return best;
}

typedef struct {
  const uint64_t *P, *O;
  size_t n;
  int depth;
  int8_t *scores;
  uint64_t *optimal;
  uint8_t *kinds;
  atomic_size_t next;
} moves_job;

static void *moves_worker(void *arg) {
  moves_job *j = (moves_job *)arg;
  for (;;) {
    size_t i = atomic_fetch_add(&j->next, 1);
    if (i >= j->n) break;
    int kind = 0;
    j->scores[i] = (int8_t)orc_best_moves(j->P[i], j->O[i], j->depth, &j->optimal[i], &kind);
    j->kinds[i] = (uint8_t)kind;
  }
  return NULL;
}

void orc_best_moves_batch(const uint64_t *P, const uint64_t *O, size_t n, int depth, int threads,
                          int8_t *scores, uint64_t *optimal, uint8_t *kinds) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  moves_job j;
  j.P = P;
  j.O = O;
  j.n = n;
  j.depth = depth;
  j.scores = scores;
  j.optimal = optimal;
  j.kinds = kinds;
  atomic_init(&j.next, 0);
  pthread_t tid[256];
  for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, moves_worker, &j);
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

/* Depth-limited search (the reference's think mode, thinker.cu:120-320):
 * negamax over `depth` disc placements; a pass does not use up depth
 * (thinker.cu:137-147,270-276); a game that ends first scores exactly; the
 * position after the depth-th move is scored by the evaluator, rounded to
 * whole discs half away from zero and clamped to [-64, 64]
 * (thinker.cu:305).  Plain recursion without pruning: it defines the value. */
int orc_think_leaf(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O) {
  float v = roundf((float)orc_eval_int(w, bias, P, O) / scale);
  if (v > 64.f) v = 64.f;
  if (v < -64.f) v = -64.f;
  return (int)v;
}

static int think_rec(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O, int depth) {
  if (depth == 0) return orc_think_leaf(w, bias, scale, P, O);
  const uint64_t m = orc_mobility(P, O);
  if (!m) {
    if (orc_mobility(O, P)) return -think_rec(w, bias, scale, O, P, depth);
    return orc_final_score(P, O);
  }
  int best = -65;
  for (uint64_t r = m; r; r &= r - 1) {
    const int sq = __builtin_ctzll(r);
    const uint64_t f = orc_flip(P, O, sq);
    const int v = -think_rec(w, bias, scale, O ^ f, (P ^ f) | (1ULL << sq), depth - 1);
    if (v > best) best = v;
  }
  return best;
}

/* The same value by fail-soft alpha-beta (window (-65, 65) at the root),
 * for depths the plain recursion cannot reach; equal to think_rec by the
 * alpha-beta theorem and checked against it by the tests. */
static int think_ab(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O, int depth,
                    int alpha, int beta) {
  if (depth == 0) return orc_think_leaf(w, bias, scale, P, O);
  const uint64_t m = orc_mobility(P, O);
  if (!m) {
    if (orc_mobility(O, P)) return -think_ab(w, bias, scale, O, P, depth, -beta, -alpha);
    return orc_final_score(P, O);
  }
  int best = -65;
  for (uint64_t r = m; r; r &= r - 1) {
    const int sq = __builtin_ctzll(r);
    const uint64_t f = orc_flip(P, O, sq);
    const int v = -think_ab(w, bias, scale, O ^ f, (P ^ f) | (1ULL << sq), depth - 1, -beta,
                            -(alpha > best ? alpha : best));
    if (v > best) best = v;
    if (best >= beta) break;
  }
  return best;
}

int orc_think_value(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O, int depth) {
  return think_ab(w, bias, scale, P, O, depth, -65, 65);
}

int orc_think(const int16_t *w, const int16_t *bias, float scale, uint64_t P, uint64_t O, int depth,
              uint64_t *optimal, int *kind) {
  const uint64_t m = orc_mobility(P, O);
  *optimal = 0;
  if (!m) {
    if (orc_mobility(O, P)) {
      *kind = ORC_MOVES_PASS;
      return -think_rec(w, bias, scale, O, P, depth);
    }
    *kind = ORC_MOVES_GAME_OVER;
    return orc_final_score(P, O);
  }
  *kind = ORC_MOVES_PLAY;
  int vals[64];
  int best = -65;
  for (uint64_t r = m; r; r &= r - 1) {
    const int sq = __builtin_ctzll(r);
    const uint64_t f = orc_flip(P, O, sq);
    vals[sq] = -think_rec(w, bias, scale, O ^ f, (P ^ f) | (1ULL << sq), depth - 1);
    if (vals[sq] > best) best = vals[sq];
  }
  for (uint64_t r = m; r; r &= r - 1)
    if (vals[__builtin_ctzll(r)] == best) *optimal |= r & (0 - r);
  return best;
}

typedef struct {
  const int16_t *w, *bias;
  float scale;
  const uint64_t *P, *O;
  size_t n;
  int depth;
  int8_t *scores;
  uint64_t *optimal;
  uint8_t *kinds;
  atomic_size_t next;
} think_job;

static void *think_worker(void *arg) {
  think_job *j = (think_job *)arg;
  for (;;) {
    size_t i = atomic_fetch_add(&j->next, 1);
    if (i >= j->n) break;
    int kind = 0;
    j->scores[i] = (int8_t)orc_think(j->w, j->bias, j->scale, j->P[i], j->O[i], j->depth, &j->optimal[i], &kind);
    j->kinds[i] = (uint8_t)kind;
  }
  return NULL;
}

void orc_think_batch(const int16_t *w, const int16_t *bias, float scale, const uint64_t *P, const uint64_t *O,
                     size_t n, int depth, int threads, int8_t *scores, uint64_t *optimal, uint8_t *kinds) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  think_job j;
  j.w = w;
  j.bias = bias;
  j.scale = scale;
  j.P = P;
  j.O = O;
  j.n = n;
  j.depth = depth;
  j.scores = scores;
  j.optimal = optimal;
  j.kinds = kinds;
  atomic_init(&j.next, 0);
  pthread_t tid[256];
  for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, think_worker, &j);
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
}

/* ------------------------------------------------------------------------- */
/* Synthetic positions: random legal moves from the standard start (black to
 * move: D5,E4 black; D4,E5 white), passing when forced, until exactly E
 * squares are empty; a game that ends first is retried with the next
 * attempt number.  The reference ships no generator; this one is the
 * project's definition, and the GPU generator must agree with it exactly. */
static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_splitmix64(uint64_t *state) {
  *state += 0x9E3779B97F4A7C15ULL;
  return mix64(*state);
}

uint64_t orc_position_key(uint64_t seed, uint32_t config, uint64_t index, uint32_t attempt) {
  uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ULL);
  k = mix64(k ^ ((uint64_t)config * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
  k = mix64(k ^ (index * 0x9E3779B97F4A7C15ULL + 0x8CB92BA72F3D8DD7ULL));
  k = mix64(k ^ ((uint64_t)attempt * 0xC2B2AE3D27D4EB4FULL + 0x165667B19E3779F9ULL));
  return k;
}

int orc_empties_for(uint64_t seed, uint32_t config, uint64_t index, int emin, int emax) {
  uint64_t k = mix64(orc_position_key(seed, config, index, 0xFFFFFFFFu));
  return emin + (int)(k % (uint64_t)(emax - emin + 1));
}

int orc_generate(uint64_t seed, uint32_t config, uint64_t index, int empties, uint64_t *Pout,
                 uint64_t *Oout) {
  for (uint32_t attempt = 0; attempt < 1000; ++attempt) {
    uint64_t s = orc_position_key(seed, config, index, attempt);
    uint64_t P = 0x0000000810000000ULL, O = 0x0000001008000000ULL;
    int ok = 1;
    while (64 - popc(P | O) > empties) {
      uint64_t m = orc_mobility(P, O);
      if (!m) {
        if (!orc_mobility(O, P)) {
          ok = 0;
          break;
        }
        uint64_t t = P;
        P = O;
        O = t;
        continue;
      }
      uint32_t k = (uint32_t)popc(m);
      uint64_t r = orc_splitmix64(&s);
      uint32_t pick = (uint32_t)(((r >> 32) * (uint64_t)k) >> 32);
      for (uint32_t i = 0; i < pick; ++i) m &= m - 1;
      uint64_t bit = m & (0 - m);
      uint64_t f = orc_flip(P, O, __builtin_ctzll(bit));
      uint64_t nP = O ^ f, nO = (P ^ f) | bit;
      P = nP;
      O = nO;
    }
    if (ok) {
      *Pout = P;
      *Oout = O;
      return (int)attempt + 1;
    }
  }
  return -1;
}

/* ------------------------------------------------------------------------- */
/* Pattern evaluator.  11 classes of ternary patterns, 46 instances (every
 * distinct image of a class under the 8 board symmetries).  Index of an
 * instance with ordered cells c_0..c_{n-1}: sum_k 3^k * s(c_k), s = 1 for a
 * side-to-move disc, 2 for an opponent disc, 0 empty — the base-3 convention
 * of reference eval.cu:113-121.  The class shapes are this project's. */
static const int k_class_cells[ORC_N_CLASSES][10] = {
    {0, 1, 2, 8, 9, 10, 16, 17, 18, -1},    /* corner 3x3 */
    {0, 1, 2, 3, 4, 8, 9, 10, 11, 12},      /* corner 2x5 */
    {0, 1, 2, 3, 4, 5, 6, 7, 9, 14},        /* edge + 2 X-squares */
    {8, 9, 10, 11, 12, 13, 14, 15, -1, -1}, /* second row */
    {16, 17, 18, 19, 20, 21, 22, 23, -1, -1},
    {24, 25, 26, 27, 28, 29, 30, 31, -1, -1},
    {0, 9, 18, 27, 36, 45, 54, 63, -1, -1}, /* main diagonal */
    {1, 10, 19, 28, 37, 46, 55, -1, -1, -1},
    {2, 11, 20, 29, 38, 47, -1, -1, -1, -1},
    {3, 12, 21, 30, 39, -1, -1, -1, -1, -1},
    {4, 13, 22, 31, -1, -1, -1, -1, -1, -1},
};

static int g_cls[ORC_N_INSTANCES], g_ncell[ORC_N_INSTANCES], g_cells[ORC_N_INSTANCES][10];
static int g_class_offset[ORC_N_CLASSES + 1];
static pthread_once_t g_pat_once = PTHREAD_ONCE_INIT;

static int sym_map(int s, int sq) {
  int r = sq >> 3, c = sq & 7, t;
  if (s & 4) { t = r; r = c; c = t; }
  if (s & 1) c = 7 - c;
  if (s & 2) r = 7 - r;
  return r * 8 + c;
}

static void build_patterns(void) {
  int n = 0, off = 0;
  for (int k = 0; k < ORC_N_CLASSES; ++k) {
    int len = 0;
    while (len < 10 && k_class_cells[k][len] >= 0) ++len;
    g_class_offset[k] = off;
    int p3 = 1;
    for (int i = 0; i < len; ++i) p3 *= 3;
    off += p3;
    uint64_t seen[8];
    int nseen = 0;
    for (int s = 0; s < 8; ++s) {
      uint64_t set = 0;
      int cells[10];
      for (int i = 0; i < len; ++i) {
        cells[i] = sym_map(s, k_class_cells[k][i]);
        set |= 1ULL << cells[i];
      }
      int dup = 0;
      for (int j = 0; j < nseen; ++j) dup |= seen[j] == set;
      if (dup) continue;
      seen[nseen++] = set;
      g_cls[n] = k;
      g_ncell[n] = len;
      for (int i = 0; i < 10; ++i) g_cells[n][i] = i < len ? cells[i] : -1;
      ++n;
    }
  }
  g_class_offset[ORC_N_CLASSES] = off;
}

int orc_pattern_layout(int *cls, int *ncell, int *cells, int *class_offset) {
  pthread_once(&g_pat_once, build_patterns);
  for (int i = 0; i < ORC_N_INSTANCES; ++i) {
    if (cls) cls[i] = g_cls[i];
    if (ncell) ncell[i] = g_ncell[i];
    if (cells)
      for (int j = 0; j < 10; ++j) cells[i * 10 + j] = g_cells[i][j];
  }
  if (class_offset)
    for (int k = 0; k <= ORC_N_CLASSES; ++k) class_offset[k] = g_class_offset[k];
  return g_class_offset[ORC_N_CLASSES];
}

void orc_pattern_indices(uint64_t P, uint64_t O, uint32_t *feat) {
  pthread_once(&g_pat_once, build_patterns);
  for (int i = 0; i < ORC_N_INSTANCES; ++i) {
    uint32_t idx = 0, p3 = 1;
    for (int k = 0; k < g_ncell[i]; ++k) {
      int c = g_cells[i][k];
      uint32_t s = (uint32_t)((P >> c) & 1) + 2u * (uint32_t)((O >> c) & 1);
      idx += s * p3;
      p3 *= 3;
    }
    feat[i] = (uint32_t)g_class_offset[g_cls[i]] + idx;
  }
}

/* Software bit extract, as reference eval.cu:56-70 defines it (bits of x
 * under mask, packed towards bit 0 in mask order). */
uint64_t orc_pext(uint64_t x, uint64_t mask) {
  uint64_t r = 0;
  int k = 0;
  for (int b = 0; b < 64; ++b)
    if ((mask >> b) & 1) r |= ((x >> b) & 1) << k++;
  return r;
}

uint32_t orc_ref_pattern_index(uint64_t me, uint64_t op, uint64_t mask) {
  uint64_t a = orc_pext(me, mask), b = orc_pext(op, mask);
  uint32_t ia = 0, ib = 0, p3 = 1;
  for (int j = 0; j < 64 && (a >> j || b >> j); ++j) {
    ia += (uint32_t)((a >> j) & 1) * p3;
    ib += (uint32_t)((b >> j) & 1) * p3;
    p3 *= 3;
  }
  return ia + 2 * ib;
}

/* Board symmetries used by reference eval.cu:72-111: vertical mirror
 * (row r -> 7-r), transpose across A1-H8 (r,c) -> (c,r), and their
 * composition rot90 = vertical(transpose). */
uint64_t orc_flip_vertical(uint64_t x) {
  uint64_t r = 0;
  for (int sq = 0; sq < 64; ++sq)
    if ((x >> sq) & 1) r |= 1ULL << ((7 - (sq >> 3)) * 8 + (sq & 7));
  return r;
}

uint64_t orc_flip_diag(uint64_t x) {
  uint64_t r = 0;
  for (int sq = 0; sq < 64; ++sq)
    if ((x >> sq) & 1) r |= 1ULL << ((sq & 7) * 8 + (sq >> 3));
  return r;
}

uint64_t orc_rot90(uint64_t x) { return orc_flip_vertical(orc_flip_diag(x)); }

int32_t orc_eval_int(const int16_t *w, const int16_t *bias, uint64_t P, uint64_t O) {
  uint32_t f[ORC_N_INSTANCES];
  orc_pattern_indices(P, O, f);
  int32_t s = 0;
  for (int i = 0; i < ORC_N_INSTANCES; ++i) s += w[f[i]];
  return s + bias[64 - popc(P | O)];
}

float orc_eval_float(const float *w, const float *bias, uint64_t P, uint64_t O) {
  uint32_t f[ORC_N_INSTANCES];
  orc_pattern_indices(P, O, f);
  float s = 0.f;
  for (int i = 0; i < ORC_N_INSTANCES; ++i) s += w[f[i]];
  return s + bias[64 - popc(P | O)];
}

double orc_sgd_step(float *w, float *bias, const uint32_t *feat, const uint8_t *empt,
                    const float *y, size_t n, float lr, float *grad_out, uint32_t *cnt_out) {
  int nw = orc_pattern_layout(NULL, NULL, NULL, NULL);
  float *g = (float *)calloc((size_t)nw + ORC_N_BIAS, sizeof(float));
  uint32_t *c = (uint32_t *)calloc((size_t)nw + ORC_N_BIAS, sizeof(uint32_t));
  double sse = 0;
  for (size_t i = 0; i < n; ++i) {
    float s = 0.f;
    for (int k = 0; k < ORC_N_INSTANCES; ++k) s += w[feat[i * ORC_N_INSTANCES + k]];
    s += bias[empt[i]];
    float r = s - y[i];
    sse += (double)r * r;
    for (int k = 0; k < ORC_N_INSTANCES; ++k) {
      uint32_t f = feat[i * ORC_N_INSTANCES + k];
      g[f] += 2.f * r;
      c[f] += 1;
    }
    g[nw + empt[i]] += 2.f * r;
    c[nw + empt[i]] += 1;
  }
  for (int f = 0; f < nw + ORC_N_BIAS; ++f) {
    if (!c[f]) continue;
    float step = lr * g[f] / (float)c[f];
    if (f < nw) w[f] -= step;
    else bias[f - nw] -= step;
  }
  if (grad_out) memcpy(grad_out, g, sizeof(float) * ((size_t)nw + ORC_N_BIAS));
  if (cnt_out) memcpy(cnt_out, c, sizeof(uint32_t) * ((size_t)nw + ORC_N_BIAS));
  free(g);
  free(c);
  return n ? sse / (double)n : 0.0;
}
