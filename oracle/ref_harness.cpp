// Test-only harness that exposes the reference's own host-compilable
// functions with a C ABI, so the oracle can be checked against them.
//
// Built by oracle/Makefile (target `ref`) together with the reference's
// sources where they lie (board.cu, to_board.cpp, eval_host.cpp); the
// result goes to oracle/_ref/ and is never shipped.  Nothing of the
// reference is copied here: this file only includes its headers.
#include <algorithm>
#include <climits>
#include <cstdint>
using std::max;
using std::min;

#include "board.cuh"      // reference board.cuh:4-19
#include "eval_host.hpp"  // reference eval_host.hpp:4
#include "node.cuh"       // reference node.cuh:5-74
#include "to_board.hpp"   // reference to_board.hpp:6-9

extern "C" {
uint64_t ref_flip(uint64_t p, uint64_t o, int pos) { return flip(p, o, pos); }
uint64_t ref_mobility(uint64_t p, uint64_t o) { return mobility(p, o); }
int ref_mobility_count(uint64_t p, uint64_t o) { return mobility_count(p, o); }
int ref_final_score(uint64_t p, uint64_t o) { return final_score(p, o); }
uint64_t ref_pext(uint64_t x, uint64_t m) { return pext(x, m); }

// Board text codec, reference to_board.cpp:15-60.
void ref_to_board(const char* s, uint64_t* p, uint64_t* o) {
  char buf[32] = {0};
  for (int i = 0; i < 16; ++i) buf[i] = s[i];
  Board b = toBoard(buf);
  *p = b.first;
  *o = b.second;
}
void ref_from_board(uint64_t p, uint64_t o, char* out17) { fromBoard(Board(p, o), out17); }

// Candidate iteration order of the lower plies, reference node.cuh:21-28:
// writes the squares in the order next_bit() yields them; returns the count.
int ref_iteration_order(uint64_t p, uint64_t o, int* squares) {
  MobilityGenerator mg(p, o);
  int n = 0;
  while (!mg.completed()) {
    ull bit = mg.next_bit();
    squares[n++] = __builtin_ctzll(bit);
  }
  return n;
}

// Board after a move / a pass through the reference's own generator.
void ref_move(uint64_t p, uint64_t o, int pos, uint64_t* np, uint64_t* no) {
  MobilityGenerator mg(p, o);
  ull f = flip(p, o, pos);
  MobilityGenerator c = mg.move(f, 1ULL << pos);
  *np = c.player_pos();
  *no = c.opponent_pos();
}
int ref_mg_score(uint64_t p, uint64_t o) { return MobilityGenerator(p, o).score(); }
}
