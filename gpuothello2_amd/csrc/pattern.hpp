// Pattern evaluator: 11 classes of ternary patterns, 46 instances (every
// distinct image of a class under the 8 symmetries of the board), int16
// weights per (class, configuration) plus one bias per number of empties.
//
// Index of an instance with cells c_0..c_{n-1} (ascending squares of the
// class in its base orientation, mapped by the instance's symmetry):
//   sum_k 3^k * s(c_k),  s = 1 for a disc of the side to move, 2 for an
// opponent disc, 0 for empty — the base-3 convention of the reference
// evaluator (eval.cu:113-121: base3[pext(me)] + 2 * base3[pext(op)]).  The
// class shapes are this project's; the reference's 13 patterns and weight
// file are not available.
//
// GPU evaluation: the board is mapped through the instance's symmetry with
// delta swaps, the base-orientation cells are gathered with shifts/folds
// (no PEXT on the GPU), and base3 comes from a 1024-entry table.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace go2 {

constexpr int kPatClasses = 11;
constexpr int kPatInstances = 46;
constexpr int kPatBias = 64;
constexpr int kPatWeights = 167265;  // sum over classes of 3^cells
constexpr int kBase3Size = 1024;

// Offset of each class in the weight table (3^cells summed in class order).
constexpr int kClassOffset[kPatClasses + 1] = {0,      19683,  78732,  137781, 144342, 150903,
                                               157464, 164025, 166212, 166941, 167184, 167265};

// (class, symmetry) of the 46 instances, class-major, first symmetry that
// gives a new cell set (symmetry s: bit 2 transposes, bit 0 mirrors columns,
// bit 1 mirrors rows, applied in that order).
struct Instance {
  int cls;
  int sym;
};
constexpr Instance kInstances[kPatInstances] = {
    {0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 0}, {1, 1}, {1, 2}, {1, 3}, {1, 4}, {1, 5}, {1, 6}, {1, 7},
    {2, 0}, {2, 2}, {2, 4}, {2, 5}, {3, 0}, {3, 2}, {3, 4}, {3, 5}, {4, 0}, {4, 2}, {4, 4}, {4, 5},
    {5, 0}, {5, 2}, {5, 4}, {5, 5}, {6, 0}, {6, 1}, {7, 0}, {7, 1}, {7, 2}, {7, 3}, {8, 0}, {8, 1},
    {8, 2}, {8, 3}, {9, 0}, {9, 1}, {9, 2}, {9, 3}, {10, 0}, {10, 1}, {10, 2}, {10, 3}};

__host__ __device__ __forceinline__ uint64_t delta_swap(uint64_t x, uint64_t mask, int delta) {
  const uint64_t t = (x ^ (x >> delta)) & mask;
  return x ^ t ^ (t << delta);
}

// Bit (r, c) of the result = bit (r, 7 - c) of x.
__host__ __device__ __forceinline__ uint64_t mirror_cols(uint64_t x) {
  x = delta_swap(x, 0x5555555555555555ULL, 1);
  x = delta_swap(x, 0x3333333333333333ULL, 2);
  return delta_swap(x, 0x0F0F0F0F0F0F0F0FULL, 4);
}

// Bit (r, c) of the result = bit (7 - r, c) of x.
__host__ __device__ __forceinline__ uint64_t mirror_rows(uint64_t x) { return __builtin_bswap64(x); }

// Bit (r, c) of the result = bit (c, r) of x.
__host__ __device__ __forceinline__ uint64_t transpose(uint64_t x) {
  x = delta_swap(x, 0x00AA00AA00AA00AAULL, 7);
  x = delta_swap(x, 0x0000CCCC0000CCCCULL, 14);
  return delta_swap(x, 0x00000000F0F0F0F0ULL, 28);
}

// Board seen through symmetry s: bit q of the result is bit sym(s, q) of x.
__host__ __device__ __forceinline__ uint64_t sym_board(uint64_t x, int s) {
  if (s & 1) x = mirror_cols(x);
  if (s & 2) x = mirror_rows(x);
  if (s & 4) x = transpose(x);
  return x;
}

// OR of the eight bytes of x (each diagonal square sits in its own column).
__host__ __device__ __forceinline__ uint32_t fold_bytes(uint64_t x) {
  x |= x >> 32;
  x |= x >> 16;
  x |= x >> 8;
  return (uint32_t)(x & 0xFF);
}

// Cells of class k in its base orientation, gathered to bits 0.. in
// ascending square order.
template <int K>
__host__ __device__ __forceinline__ uint32_t gather(uint64_t x) {
  if constexpr (K == 0) {  // 3x3 corner: squares 0-2, 8-10, 16-18
    return (uint32_t)((x & 7) | ((x >> 5) & 0x38) | ((x >> 10) & 0x1C0));
  } else if constexpr (K == 1) {  // 2x5 corner: squares 0-4, 8-12
    return (uint32_t)((x & 0x1F) | ((x >> 3) & 0x3E0));
  } else if constexpr (K == 2) {  // edge + X-squares: 0-7, 9, 14
    return (uint32_t)((x & 0xFF) | ((x >> 1) & 0x100) | ((x >> 5) & 0x200));
  } else if constexpr (K <= 5) {  // rows 1..3
    return (uint32_t)((x >> (8 * (K - 2))) & 0xFF);
  } else {  // diagonals starting at column K-6 of row 0
    constexpr int d = K - 6;
    constexpr uint64_t m = d == 0 ? 0x8040201008040201ULL : (0x8040201008040201ULL << d) & ~(0xFFFFFFFFFFFFFFFFULL << (8 * (8 - d)));
    return fold_bytes(x & m) >> d;
  }
}

template <int K>
__host__ __device__ __forceinline__ uint32_t gather_k(uint64_t x) {
  return gather<K>(x);
}

__host__ __device__ __forceinline__ uint32_t gather_dyn(int k, uint64_t x) {
  switch (k) {
    case 0: return gather<0>(x);
    case 1: return gather<1>(x);
    case 2: return gather<2>(x);
    case 3: return gather<3>(x);
    case 4: return gather<4>(x);
    case 5: return gather<5>(x);
    case 6: return gather<6>(x);
    case 7: return gather<7>(x);
    case 8: return gather<8>(x);
    case 9: return gather<9>(x);
    default: return gather<10>(x);
  }
}

// Symmetries visited in Gray-code order, so that each step is one cheap
// transform of the boards seen so far: toggling the column mirror before a
// transpose equals mirroring rows after it, and vice versa.
constexpr int kGray[8] = {0, 1, 3, 2, 6, 7, 5, 4};
enum { kStepNone = 0, kStepCols = 1, kStepRows = 2, kStepTranspose = 3 };
constexpr int kGrayStep[8] = {kStepNone, kStepCols, kStepRows, kStepCols,
                              kStepTranspose, kStepRows, kStepCols, kStepRows};
// Classes with an instance under symmetry s (bit k = class k), and the
// instance number of (class k, symmetry s) in kInstances order.
constexpr int kSymClasses[8] = {0x7FF, 0x7C3, 0x7BF, 0x783, 0x3E, 0x3E, 0x2, 0x2};
constexpr int kInstanceOf[8][kPatClasses] = {
    {0, 4, 12, 16, 20, 24, 28, 30, 34, 38, 42}, {1, 5, -1, -1, -1, -1, 29, 31, 35, 39, 43},
    {2, 6, 13, 17, 21, 25, -1, 32, 36, 40, 44}, {3, 7, -1, -1, -1, -1, -1, 33, 37, 41, 45},
    {-1, 8, 14, 18, 22, 26, -1, -1, -1, -1, -1}, {-1, 9, 15, 19, 23, 27, -1, -1, -1, -1, -1},
    {-1, 10, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {-1, 11, -1, -1, -1, -1, -1, -1, -1, -1, -1}};

// Calls f(instance, global weight index) for the 46 instances of (P, O).
// t3[b] = sum_k bit_k(b) * 3^k.  A loop over the symmetries (uniform across
// lanes) with one board transform per step keeps code and registers small.
template <class T3, class F>
__host__ __device__ __forceinline__ void for_each_feature(uint64_t P, uint64_t O, const T3& t3, F&& f) {
  uint64_t xp = P, xo = O;
#pragma unroll 1
  for (int i = 0; i < 8; ++i) {
    const int st = kGrayStep[i];
    if (st == kStepCols) {
      xp = mirror_cols(xp);
      xo = mirror_cols(xo);
    } else if (st == kStepRows) {
      xp = mirror_rows(xp);
      xo = mirror_rows(xo);
    } else if (st == kStepTranspose) {
      xp = transpose(xp);
      xo = transpose(xo);
    }
    const int s = kGray[i];
    const int m = kSymClasses[s];
#define GO2_PAT_CLASS(K)                                                                        \
    if (m & (1 << K)) {                                                                        \
      const uint32_t a = gather<K>(xp), b = gather<K>(xo);                                     \
      f(kInstanceOf[s][K], (uint32_t)kClassOffset[K] + (uint32_t)t3[a] + 2u * (uint32_t)t3[b]); \
    }
    GO2_PAT_CLASS(0)
    GO2_PAT_CLASS(1)
    GO2_PAT_CLASS(2)
    GO2_PAT_CLASS(3)
    GO2_PAT_CLASS(4)
    GO2_PAT_CLASS(5)
    GO2_PAT_CLASS(6)
    GO2_PAT_CLASS(7)
    GO2_PAT_CLASS(8)
    GO2_PAT_CLASS(9)
    GO2_PAT_CLASS(10)
#undef GO2_PAT_CLASS
  }
}

}  // namespace go2
