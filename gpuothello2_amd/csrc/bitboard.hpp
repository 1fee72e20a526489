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

// Any boolean function of three words in one gfx950 instruction
// (v_bitop3_b32) per 32-bit half.  TT is the truth table: bit i is the
// result for the input bits a = i >> 2, b = (i >> 1) & 1, c = i & 1 (so
// a & b & c is 0x80, a | (b & c) is 0xF8).  Host builds evaluate the table.
template <int TT>
GO2_HD uint32_t bop3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, TT);
#else
  uint32_t r = 0;
  for (int i = 0; i < 8; ++i)
    if ((TT >> i) & 1) r |= ((i & 4) ? a : ~a) & ((i & 2) ? b : ~b) & ((i & 1) ? c : ~c);
  return r;
#endif
}
template <int TT>
GO2_HD uint64_t bop3(uint64_t a, uint64_t b, uint64_t c) {
  const uint32_t hi = bop3<TT>((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32));
  const uint32_t lo = bop3<TT>((uint32_t)a, (uint32_t)b, (uint32_t)c);
  uint64_t r = ((uint64_t)hi << 32) | lo;
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(r));  // one 64-bit register pair, so later 64-bit shifts stay whole
#endif
  return r;
}
constexpr int kAnd3 = 0x80;      // a & b & c
constexpr int kAndAndNot = 0x40; // a & b & ~c
constexpr int kOrAnd = 0xF8;     // a | (b & c)
constexpr int kOrAndNot = 0xF4;  // a | (b & ~c)
constexpr int kAndNotOr = 0x10;  // a & ~(b | c)

// Discs flipped along one line through sq.  `up` and `dn` hold the squares
// of that line above and below sq.  Above sq the first non-opponent square
// is the lowest candidate bit; below sq it is the highest one.  A player
// disc there closes the run and every square between flips.
GO2_HD uint64_t flip_line(uint64_t P, uint64_t Om, uint64_t up, uint64_t dn) {
  const uint64_t stop_up = up & ~Om;
  const uint64_t out_up = bop3<kAnd3>(stop_up, 0 - stop_up, P);
  // squares of up below out_up: out_up - 1, cleared when out_up is 0 (then
  // out_up - 1 is all ones, the only case with its top bit set)
  const uint64_t t = out_up - 1;
  const uint64_t none = (uint64_t)(int64_t)((int32_t)(uint32_t)(t >> 32) >> 31);
  uint64_t f = bop3<kAndAndNot>(t, up, none);
  // highest non-opponent square below sq, kept when it is a player disc
  const uint64_t sd = dn & ~Om;
  const uint64_t out_dn = bop3<kAnd3>(0x8000000000000000ULL >> __builtin_clzll(sd | 1), sd, P);
  // squares of dn above out_dn (none when out_dn is 0)
  f = bop3<kOrAndNot>(f, dn, out_dn | (out_dn - 1));
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
  const uint64_t row = 0xFFULL << (sq & 56);
  const uint64_t col = 0x0101010101010101ULL << c;
  // the diagonals through sq, shifted by whole rows (no wrap across an edge):
  // right shift by max(8d, 0) then left shift by max(-8d, 0), no branch
  const int d = 8 * (c - r);      // a1-h8 direction
  const uint64_t diag = (0x8040201008040201ULL >> (d > 0 ? d : 0)) << (d < 0 ? -d : 0);
  const int a = 8 * (c + r - 7);  // h1-a8 direction
  const uint64_t anti = (0x0102040810204080ULL << (a > 0 ? a : 0)) >> (a < 0 ? -a : 0);
  uint64_t f = flip_line(P, O, row & above, row & below);
  f |= flip_line(P, O, col & above, col & below);
  f |= flip_line(P, O, diag & above, diag & below);
  f |= flip_line(P, O, anti & above, anti & below);
  return f;
}

// A 64-bit shift kept as one 64-bit value: without the (empty) operand
// barrier the compiler splits a shift whose halves feed separate 32-bit
// operations into a funnel shift per half (two instructions instead of one).
GO2_HD uint64_t shl_whole(uint64_t x, int s) {
  uint64_t y = x << s;
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(y));
#endif
  return y;
}
GO2_HD uint64_t shr_whole(uint64_t x, int s) {
  uint64_t y = x >> s;
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(y));
#endif
  return y;
}

// Legal moves in one direction pair (shift s left and right), filling the
// opponent run by doubling: 1, 2, then 4 and 6 squares.  Each "x |= m & y"
// step is one bitop3 per 32-bit half.
GO2_HD uint64_t mob_dir(uint64_t P, uint64_t Om, int s) {
  uint64_t l = Om & shl_whole(P, s);
  uint64_t r = Om & shr_whole(P, s);
  l = bop3<kOrAnd>(l, Om, shl_whole(l, s));
  r = bop3<kOrAnd>(r, Om, shr_whole(r, s));
  const uint64_t pl = Om & shl_whole(Om, s);
  const uint64_t pr = shr_whole(pl, s);
  l = bop3<kOrAnd>(l, pl, shl_whole(l, 2 * s));
  r = bop3<kOrAnd>(r, pr, shr_whole(r, 2 * s));
  l = bop3<kOrAnd>(l, pl, shl_whole(l, 2 * s));
  r = bop3<kOrAnd>(r, pr, shr_whole(r, 2 * s));
  return shl_whole(l, s) | shr_whole(r, s);
}

GO2_HD uint64_t mobility(uint64_t P, uint64_t O) {
  const uint64_t Oi = O & kInner;
  const uint64_t m = mob_dir(P, Oi, 1) | mob_dir(P, O, 8) | mob_dir(P, Oi, 7) | mob_dir(P, Oi, 9);
  return bop3<kAndNotOr>(m, P, O);
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
  const int d = 8 * (c - r), a = 8 * (c + r - 7);
  const uint64_t dm = (0x8040201008040201ULL >> (d > 0 ? d : 0)) << (d < 0 ? -d : 0);
  const uint64_t am = (0x0102040810204080ULL << (a > 0 ? a : 0)) >> (a < 0 ? -a : 0);
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
  // both sides' counts in one round of table reads (no branch between them)
  const uint32_t hole = ~(1u << c) & 0xFF, holer = ~(1u << r) & 0xFF;
  const int n1 = t[c * 256 + L.row] + t[r * 256 + L.col] + t[c * 256 + L.diag] + t[c * 256 + L.anti];
  const int n2 = t[c * 256 + (~L.row & hole)] + t[r * 256 + (~L.col & holer)] +
                 t[c * 256 + (~L.diag & L.diag_on & hole)] + t[c * 256 + (~L.anti & L.anti_on & hole)];
  const int p = popc64(P);
  return n1 ? 2 * (p + n1) + 2 - 64 : n2 ? 2 * (p - n2) - 64 : final_score(P, O);
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
