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
  f |= (0 - (out_dn << 1)) & dn;
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

}  // namespace go2
