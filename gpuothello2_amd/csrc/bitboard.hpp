// Bitboard primitives for the exact solver (gfx950 device code; the same
// inline functions also compile for the host, where the CLI needs none of
// them but unit builds may).
//
// Layout: square = row*8 + col, bit 0 = A1, bit 7 = H1, bit 63 = H8.
// P = side to move, O = the other side.
//
// Counterparts in the reference:
//   go2_final_score  <- final_score     board.cuh:7-18
//   go2_flip         <- flip/flip_impl  board.cu:47-73
//   go2_mobility     <- mobility        board.cu:75-97
//   go2_last_score   <- (no counterpart: the reference walks the last ply
//                        through its generic lower-node loop, solver.cu:244-270)
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#ifndef GO2_HD
#define GO2_HD __host__ __device__ __forceinline__
#endif

namespace go2 {

constexpr uint64_t kCorners = 0x8100000000000081ULL;
constexpr uint64_t kXSquares = 0x4200000000004200ULL;  // B2, G2, B7, G7
constexpr uint64_t kNotA = 0xFEFEFEFEFEFEFEFEULL;
constexpr uint64_t kNotH = 0x7F7F7F7F7F7F7F7FULL;
constexpr uint64_t kInner = 0x7E7E7E7E7E7E7E7EULL;   // columns B..G

GO2_HD int popc64(uint64_t x) { return __builtin_popcountll(x); }
GO2_HD int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// Winner takes the empty squares; a tie scores 0.
GO2_HD int final_score(uint64_t P, uint64_t O) {
  int p = popc64(P), o = popc64(O);
  int s = (p > o) ? 64 - 2 * o : 2 * p - 64;
  return p == o ? 0 : s;
}

// Highest set bit of x, or 0 when x == 0.
GO2_HD uint64_t high_bit(uint64_t x) {
  return (0x8000000000000000ULL >> __builtin_clzll(x | 1)) & x;
}

// Discs flipped along one line through sq.  `line` holds the squares of that
// line (sq included or not).  Above sq the first non-opponent square is the
// lowest candidate bit; below sq it is the highest one.  A player disc there
// closes the run and every square between flips.
GO2_HD uint64_t flip_line(uint64_t P, uint64_t Om, uint64_t line, uint64_t above, uint64_t below) {
  uint64_t up = line & above;
  uint64_t stop_up = up & ~Om;
  uint64_t out_up = stop_up & (0 - stop_up) & P;
  uint64_t f = (out_up - (uint64_t)(out_up != 0)) & up;
  uint64_t dn = line & below;
  uint64_t out_dn = high_bit(dn & ~Om) & P;
  // squares of dn above out_dn (none when out_dn is 0).  Written without
  // -(out_dn << 1), which the compiler turns into a 64-bit multiply by -2.
  f |= dn & ~(out_dn | (out_dn - 1));
  return f;
}

// Flip set of the move at sq (0 => illegal).  The four lines through sq are
// built exactly from sq (row, column, and the two diagonals shifted by whole
// rows, so no line wraps across an edge); each is split at sq into the part
// above and the part below.
GO2_HD uint64_t flip(uint64_t P, uint64_t O, int sq) {
  const uint64_t above = 0xFFFFFFFFFFFFFFFEULL << sq;
  const uint64_t below = ~(0xFFFFFFFFFFFFFFFFULL << sq);
  const int r = sq >> 3, c = sq & 7;
  uint64_t f = flip_line(P, O, 0xFFULL << (sq & 56), above, below);
  f |= flip_line(P, O, 0x0101010101010101ULL << c, above, below);
  const int d = c - r;  // a1-h8 direction
  const uint64_t diag = d >= 0 ? 0x8040201008040201ULL >> (8 * d) : 0x8040201008040201ULL << (-8 * d);
  f |= flip_line(P, O, diag, above, below);
  const int a = c + r - 7;  // h1-a8 direction
  const uint64_t anti = a >= 0 ? 0x0102040810204080ULL << (8 * a) : 0x0102040810204080ULL >> (-8 * a);
  f |= flip_line(P, O, anti, above, below);
  return f;
}

// Legal moves in one direction pair (shift s left and right), filling the
// opponent run by doubling: 1, 2, then 4 and 6 squares.
GO2_HD uint64_t mob_dir(uint64_t P, uint64_t Om, int s) {
  uint64_t l = Om & (P << s);
  uint64_t r = Om & (P >> s);
  l |= Om & (l << s);
  r |= Om & (r >> s);
  const uint64_t pl = Om & (Om << s);
  const uint64_t pr = pl >> s;
  l |= pl & (l << (2 * s));
  r |= pr & (r >> (2 * s));
  l |= pl & (l << (2 * s));
  r |= pr & (r >> (2 * s));
  return (l << s) | (r >> s);
}

GO2_HD uint64_t mobility(uint64_t P, uint64_t O) {
  const uint64_t Oi = O & kInner;
  uint64_t m = mob_dir(P, Oi, 1) | mob_dir(P, O, 8) | mob_dir(P, Oi, 7) | mob_dir(P, Oi, 9);
  return m & ~(P | O);
}

// Exact score of a board with exactly one empty square `sq`, from P's view:
// P plays it if legal, else O plays it if legal, else the game is over.
GO2_HD int last_score(uint64_t P, uint64_t O, int sq) {
  const int p = popc64(P);
  const int n1 = popc64(flip(P, O, sq));
  if (n1) return 2 * (p + n1) + 2 - 64;
  const int n2 = popc64(flip(O, P, sq));
  if (n2) return 2 * (p - n2) - 64;
  return final_score(P, O);
}


// ---- last empty square by table ---------------------------------------------
// On a board whose only empty square is sq, the discs a side flips there
// depend, line by line, only on that side's discs along the line.  kLastCount
// [pos][p] = discs flipped along one 8-square line when the mover's discs on
// it form the bit pattern p and it plays at index pos; squares off the board
// are 0 in p and (being opponent-like but never followed by a mover disc)
// flip nothing.
GO2_HD int line_flip_count(int pos, int p) {
  int n = 0, run = 0;
  for (int i = pos + 1; i < 8; ++i) {
    if ((p >> i) & 1) {
      n += run;
      break;
    }
    ++run;
  }
  run = 0;
  for (int i = pos - 1; i >= 0; --i) {
    if ((p >> i) & 1) {
      n += run;
      break;
    }
    ++run;
  }
  return n;
}

// Gather column c of x into bits 0..7 (row order).
GO2_HD uint32_t column_bits(uint64_t x, int c) {
  uint64_t t = (x >> c) & 0x0101010101010101ULL;
  t |= t >> 7;
  t |= t >> 14;
  t |= t >> 28;
  return (uint32_t)(t & 0xFF);
}

// The four line patterns through sq of a bitboard, each in bits 0..7 (row
// and diagonals indexed by column, the column indexed by row), plus the
// patterns of which line squares exist on the board.
struct LinePatterns {
  uint32_t row, col, diag, anti;      // mover's discs
  uint32_t diag_on, anti_on;          // squares of the diagonals on the board
};

GO2_HD LinePatterns line_patterns(uint64_t x, int sq) {
  const int r = sq >> 3, c = sq & 7;
  const int d = c - r, a = c + r - 7;
  const uint64_t dm = d >= 0 ? 0x8040201008040201ULL >> (8 * d) : 0x8040201008040201ULL << (-8 * d);
  const uint64_t am = a >= 0 ? 0x0102040810204080ULL << (8 * a) : 0x0102040810204080ULL >> (-8 * a);
  auto fold = [](uint64_t y) {
    y |= y >> 32;
    y |= y >> 16;
    y |= y >> 8;
    return (uint32_t)(y & 0xFF);
  };
  LinePatterns L;
  L.row = (uint32_t)((x >> (sq & 56)) & 0xFF);
  L.col = column_bits(x, c);
  L.diag = fold(x & dm);
  L.anti = fold(x & am);
  L.diag_on = fold(dm);
  L.anti_on = fold(am);
  return L;
}

// Exact score, from P's view, of a board with exactly one empty square sq,
// using the kLastCount table `t` (8 x 256 bytes).
template <class T>
GO2_HD int last_score_tab(uint64_t P, uint64_t O, int sq, const T& t) {
  const int r = sq >> 3, c = sq & 7;
  const LinePatterns L = line_patterns(P, sq);
  const int n1 = t[c * 256 + L.row] + t[r * 256 + L.col] + t[c * 256 + L.diag] + t[c * 256 + L.anti];
  const int p = popc64(P);
  if (n1) return 2 * (p + n1) + 2 - 64;
  const uint32_t hole = ~(1u << c) & 0xFF, holer = ~(1u << r) & 0xFF;
  const int n2 = t[c * 256 + (~L.row & hole)] + t[r * 256 + (~L.col & holer)] +
                 t[c * 256 + (~L.diag & L.diag_on & hole)] + t[c * 256 + (~L.anti & L.anti_on & hole)];
  if (n2) return 2 * (p - n2) - 64;
  return final_score(P, O);
}

// Both moves of O after P, with two empty squares a < b, has passed (rare;
// kept out of line).  Returns the score for P, 65 in *best_sq when the game
// is over.
template <class T>
GO2_HD int solve2_pass(uint64_t P, uint64_t O, int a, int b, const T& t,
                                                  int& nodes, int* best_sq) {
  const uint64_t ba = 1ULL << a, bb = 1ULL << b;
  const uint64_t ga = flip(O, P, a), gb = flip(O, P, b);
  if (!(ga | gb)) {
    if (best_sq) *best_sq = 65;
    return final_score(P, O);
  }
  const int wa = ga ? -last_score_tab(P ^ ga, O ^ ga ^ ba, b, t) : -65;
  const int wb = gb ? -last_score_tab(P ^ gb, O ^ gb ^ bb, a, t) : -65;
  nodes += 1 + (ga != 0) + (gb != 0);
  if (best_sq) *best_sq = 64;
  return -(wa > wb ? wa : wb);
}

// The score of solve2 when P has a legal move on one of the two empty
// squares; kNoMove2 when P has none (the caller settles the pass itself, so
// that many lanes running this together never run the pass branch).
constexpr int kNoMove2 = 127;
template <class T>
GO2_HD int solve2_mover(uint64_t P, uint64_t O, const T& t, int& nodes) {
  const uint64_t e = ~(P | O);
  const int a = ctz64(e);
  const int b = 63 - __builtin_clzll(e & (e - 1));
  int best = -65;
  uint64_t f = flip(P, O, a);
  if (f) {
    best = -last_score_tab(O ^ f, P ^ f ^ (1ULL << a), b, t);
    ++nodes;
  }
  f = flip(P, O, b);
  if (f) {
    const int v = -last_score_tab(O ^ f, P ^ f ^ (1ULL << b), a, t);
    ++nodes;
    best = v > best ? v : best;
  }
  return best > -65 ? best : kNoMove2;
}

// Exact score for P to move with exactly two empty squares, in straight-line
// code: each move of P leaves one empty square, scored directly; if P has no
// move, O's moves (or the final count) decide.  `nodes` counts the children
// evaluated.  `best_sq` (when non-null) receives the move that reaches the
// score: a square, 64 for a pass, 65 when the game is over.
template <class T>
GO2_HD int solve2(uint64_t P, uint64_t O, const T& t, int& nodes, int* best_sq = nullptr) {
  const uint64_t e = ~(P | O);
  const int a = ctz64(e);
  const int b = 63 - __builtin_clzll(e & (e - 1));
  int best = -65;
  uint64_t f = flip(P, O, a);
  if (f) {
    best = -last_score_tab(O ^ f, P ^ f ^ (1ULL << a), b, t);
    ++nodes;
    if (best_sq) *best_sq = a;
  }
  f = flip(P, O, b);
  if (f) {
    const int v = -last_score_tab(O ^ f, P ^ f ^ (1ULL << b), a, t);
    ++nodes;
    if (v > best) {
      best = v;
      if (best_sq) *best_sq = b;
    }
  }
  if (best > -65) return best;
  return solve2_pass(P, O, a, b, t, nodes, best_sq);
}

}  // namespace go2
