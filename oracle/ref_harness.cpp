// Test-only harness that exposes the reference's own host functions with a
// C ABI, so the oracle can be checked against them.
//
// Built by oracle/Makefile (target `ref`) together with the two reference
// sources that compile unchanged as host C++ (to_board.cpp, eval_host.cpp),
// where they lie; the result goes to oracle/_ref/ and is never shipped.
// Nothing of the reference is copied here: this file only includes its
// headers.  board.cu and node.cuh are CUDA sources (their functions carry
// __host__ __device__) and are not part of this build: the functions they
// hold are pinned by brute force and closed forms instead.
#include <cstdint>

#include "eval_host.hpp"  // reference eval_host.hpp:4
#include "to_board.hpp"   // reference to_board.hpp:6-9

extern "C" {
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
}
