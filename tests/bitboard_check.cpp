// Host build of the solver's bitboard primitives (gpuothello2_amd/csrc/
// bitboard.hpp, its host branch of every bitop3), checked against the
// oracle's square-by-square ray walks and against each other:
//   flip, mobility            == orc_flip_walk, orc_mobility_walk
//   last_score_tab            == last_score (flip-based)
//   two-round leaf (solve2_mover, then the swapped pass round as the solve
//   kernel's leaf batch runs it) == -solve2 of the child, with equal node counts
// Prints "ok" and exits 0, or names the first mismatch.
#include <cstdint>
#include <cstdio>
#include <random>

#include "bitboard.hpp"

extern "C" uint64_t orc_flip_walk(uint64_t P, uint64_t O, int sq);
extern "C" uint64_t orc_mobility_walk(uint64_t P, uint64_t O);

using namespace go2;

static int leaf_two_round(uint64_t P, uint64_t O, int sq, const uint8_t* t, int& nodes) {
  auto job = [&](bool swap) {
    const uint64_t f = flip(P, O, sq), bit = 1ULL << sq;
    const uint64_t cP = swap ? P ^ f ^ bit : O ^ f, cO = swap ? O ^ f : P ^ f ^ bit;
    ++nodes;
    const int r = solve2_mover(cP, cO, t, nodes);
    if (r != kNoMove2) return swap ? r : -r;
    return swap ? final_score(cP, cO) : 1000;
  };
  const int v = job(false);
  return v == 1000 ? job(true) : v;
}

int main() {
  uint8_t t[2048];
  for (int k = 0; k < 2048; ++k) t[k] = (uint8_t)line_flip_count(k >> 8, k & 255);
  std::mt19937_64 rng(20261204);
  long flips = 0, leaves = 0, passes = 0;
  for (int it = 0; it < 60000; ++it) {
    const uint64_t x = rng(), y = rng(), z = rng();
    uint64_t P = x & y, O = x & ~y;
    if (it % 3 == 1) P &= z;
    if (it % 3 == 2) O &= z | rng();
    if (mobility(P, O) != orc_mobility_walk(P, O)) {
      std::printf("mobility mismatch %016llx %016llx\n", (unsigned long long)P, (unsigned long long)O);
      return 1;
    }
    for (int sq = 0; sq < 64; ++sq) {
      if (((P | O) >> sq) & 1) continue;
      ++flips;
      if (flip(P, O, sq) != orc_flip_walk(P, O, sq)) {
        std::printf("flip mismatch %016llx %016llx %d\n", (unsigned long long)P, (unsigned long long)O, sq);
        return 1;
      }
    }
    // one empty square: the table count against the flip-based count
    const int sq1 = (int)(z & 63);
    const uint64_t P1 = x & ~(1ULL << sq1), O1 = ~x & ~(1ULL << sq1);
    if (last_score_tab(P1, O1, sq1, t) != last_score(P1, O1, sq1)) {
      std::printf("last_score_tab mismatch %016llx %d\n", (unsigned long long)P1, sq1);
      return 1;
    }
    // three empty squares: every move's two-round leaf value against solve2
    uint64_t e = 0;
    while (__builtin_popcountll(e) < 3) e |= 1ULL << (rng() & 63);
    const uint64_t P3 = y & ~e, O3 = ~y & ~e;
    for (uint64_t m = mobility(P3, O3); m; m &= m - 1) {
      const int sq = __builtin_ctzll(m);
      const uint64_t f = flip(P3, O3, sq);
      int n1 = 1, n2 = 0;
      const int want = -solve2(O3 ^ f, P3 ^ f ^ (1ULL << sq), t, n1);
      const int got = leaf_two_round(P3, O3, sq, t, n2);
      passes += mobility(O3 ^ f, P3 ^ f ^ (1ULL << sq)) == 0;
      ++leaves;
      if (got != want || n1 != n2) {
        std::printf("leaf mismatch %016llx %016llx %d: %d/%d nodes %d/%d\n", (unsigned long long)P3,
                    (unsigned long long)O3, sq, got, want, n2, n1);
        return 1;
      }
    }
  }
  std::printf("ok flips %ld leaves %ld passes %ld\n", flips, leaves, passes);
  return passes > 0 ? 0 : 1;
}
