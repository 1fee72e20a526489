// Batched exact endgame solver kernel for gfx950.
//
// What it replaces: alpha_beta_kernel and the Solver state machine of the
// reference (solver.cu:15-311), which runs one CUDA thread per board with
// 80-byte global-memory frames for the top DEPTH-9 plies and 24-byte
// shared-memory frames for the last 10.
//
// Design here (wave64, one lane per board):
//   * Persistent grid sized to residency; each lane pulls the next board
//     index from one device counter when its board is done, so lanes never
//     idle while boards remain (replaces the per-block stride queue,
//     solver.cu:169-177,289-292).
//   * The node being searched lives in VGPRs.  Only ancestors are spilled to
//     LDS, one frame per empties level, struct-of-arrays with the lane index
//     fastest so every frame access is bank-conflict free whatever depth each
//     lane is at.  A frame stores the flip set of the move taken (the parent
//     board is rebuilt from the child by two XORs), the untried-move set, the
//     packed move order, and alpha/beta/best/move/pass in one dword:
//     28 bytes per level, levels for 3..MAXE empties.
//   * Every loop iteration does the same heavy work on every active lane:
//     one flip and one mobility (or, one ply from the end, the exact score of
//     the last empty square).  Move ordering work is spread over iterations
//     ("scoring" steps: one child per iteration), so lanes of a wave that are
//     ordering and lanes that are descending run the same instructions.
//   * Exact fail-soft negamax, full window (-64, 64); passes are folded into
//     the child frame (board swapped, window kept, value not negated), which
//     is the same recurrence as solver.cu:65-73,162-165,214-221,246-256.
// Scores do not depend on move order, so any ordering gives the reference's
// scores bit for bit; ordering only changes node counts.
#include "bitboard.hpp"
#include "internal.hpp"

namespace go2 {

// 4x4 quadrants for parity ordering.
__constant__ uint64_t kQuad[4] = {0x000000000F0F0F0FULL, 0x00000000F0F0F0F0ULL,
                                  0x0F0F0F0F00000000ULL, 0xF0F0F0F000000000ULL};

// Static order: corners, then squares other than X-squares, X-squares last;
// lowest square first inside a class.
__device__ __forceinline__ int pick_plain(uint64_t m) {
  const uint64_t c = m & kCorners;
  const uint64_t nx = m & ~kXSquares;
  return ctz64(c ? c : (nx ? nx : m));
}

// As pick_plain, but among non-corner squares prefer those in a quadrant
// holding an odd number of empties (the side to move there gets the last
// move of that region).
__device__ __forceinline__ int pick_parity(uint64_t m, uint64_t empty) {
  uint64_t odd = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) odd |= (popc64(empty & kQuad[q]) & 1) ? kQuad[q] : 0;
  const uint64_t c = m & kCorners;
  const uint64_t po = m & odd & ~kXSquares;
  const uint64_t nx = m & ~kXSquares;
  return ctz64(c ? c : (po ? po : (nx ? nx : m)));
}

// Packed sorted move list: up to 9 entries of 7 bits (63 bits).  Keys and
// squares are kept in two such words; comparisons are done field-parallel.
constexpr uint64_t kLo7 = 0x0102040810204081ULL;  // bit 0 of each 7-bit field
constexpr uint64_t kHi7 = kLo7 << 6;                                         // bit 6 of each field
constexpr uint64_t kFull63 = 0x7FFFFFFFFFFFFFFFULL;
static_assert(__builtin_popcountll(kLo7) == 9, "nine 7-bit fields");

__device__ __forceinline__ uint32_t pack_frame(int alpha, int beta, int best, int sq, bool passed) {
  return (uint32_t)(uint8_t)alpha | ((uint32_t)(uint8_t)beta << 8) |
         ((uint32_t)(uint8_t)best << 16) | ((uint32_t)sq << 24) | ((uint32_t)passed << 30);
}

struct SolveArgs {
  const go2_board* boards;
  int8_t* scores;
  unsigned int n;
  int depth;   // DEPTH: boards with more empties are rejected
  int kff;     // fastest-first ordering at nodes with >= kff empties
  SolveCounters* ctr;
};

template <int MAXE, int ORDER>
__global__ void __launch_bounds__(256) solve_kernel(SolveArgs a) {
  extern __shared__ uint64_t lds[];
  constexpr int L = MAXE - 2;  // frame levels: nodes with 3..MAXE empties
  const int bs = blockDim.x, tid = threadIdx.x;
  uint64_t* fbA = lds;
  uint64_t* mvA = fbA + L * bs;
  uint64_t* odA = mvA + L * bs;
  uint32_t* wdA = reinterpret_cast<uint32_t*>(odA + L * bs);

  uint64_t P = 0, O = 0, moves = 0, order = 0, to_score = 0, keys = 0;
  int alpha = 0, beta = 0, best = 0, E = 0, Eroot = 0, nkeys = 0;
  bool scoring = false, negate = false, active = false;
  unsigned int idx = 0;
  unsigned long long c_pos = 0, c_nodes = 0, c_steps = 0, c_iters = 0;

  for (;;) {
    if (!active) {
      idx = atomicAdd(&a.ctr->next, 1u);
      if (idx >= a.n) break;
      const go2_board b = a.boards[idx];
      P = b.player;
      O = b.opponent;
      E = 64 - popc64(P | O);
      ++c_pos;
      ++c_nodes;
      if ((P & O) || E > a.depth) {
        atomicCAS(&a.ctr->error, 0u, idx + 1);
        a.scores[idx] = 0;
        continue;
      }
      if (E <= 1) {
        a.scores[idx] = (int8_t)(E == 0 ? final_score(P, O) : last_score(P, O, ctz64(~(P | O))));
        continue;
      }
      moves = mobility(P, O);
      negate = false;
      if (!moves) {
        const uint64_t om = mobility(O, P);
        if (!om) {
          a.scores[idx] = (int8_t)final_score(P, O);
          continue;
        }
        const uint64_t t = P;
        P = O;
        O = t;
        moves = om;
        negate = true;
      }
      Eroot = E;
      alpha = -64;
      beta = 64;
      best = -64;
      order = 0;
      scoring = ORDER >= 1 && E >= a.kff && (moves & (moves - 1));
      to_score = moves;
      keys = 0;
      nkeys = 0;
      active = true;
      continue;
    }

    ++c_iters;
    int sq;
    if (scoring) {
      sq = ctz64(to_score);
    } else if (order) {
      sq = (int)(order & 127) - 1;
      order >>= 7;
    } else {
      sq = ORDER == 0 ? pick_plain(moves) : pick_parity(moves, ~(P | O));
    }
    const uint64_t bit = 1ULL << sq;
    const uint64_t f = flip(P, O, sq);
    const uint64_t cP = O ^ f, cO = P ^ f ^ bit;
    const bool leaf = !scoring && E == 2;
    uint64_t cm = 0;
    int v = 0;
    if (leaf) {
      v = -last_score(cP, cO, ctz64(~(cP | cO)));
    } else {
      cm = mobility(cP, cO);
    }

    if (scoring) {
      // Fastest first: fewest replies, corners preferred, X-squares delayed.
      ++c_steps;
      const int k = min(popc64(cm), 15) * 4 + ((bit & kCorners) ? 0 : 2) + ((bit & kXSquares) ? 1 : 0);
      const uint64_t valid = (1ULL << (7 * nkeys)) - 1;
      const uint64_t le = ((((uint64_t)k * kLo7) | kHi7) - keys) & kHi7 & valid;
      const int c = popc64(le);  // entries with key <= k: insert after them (stable)
      const uint64_t low = (1ULL << (7 * c)) - 1;
      keys = (((keys & ~low) << 7) | (keys & low) | ((uint64_t)k << (7 * c))) & kFull63;
      order = (((order & ~low) << 7) | (order & low) | ((uint64_t)(sq + 1) << (7 * c))) & kFull63;
      nkeys = min(nkeys + 1, 9);
      to_score ^= bit;
      scoring = to_score != 0;
      continue;
    }

    moves ^= bit;
    ++c_nodes;
    if (!leaf) {
      bool passed = false;
      bool terminal = false;
      if (!cm) {
        const uint64_t om = mobility(cO, cP);
        if (!om) {
          v = -final_score(cP, cO);
          terminal = true;
        } else {
          cm = om;
          passed = true;
        }
      }
      if (!terminal) {
        const int at = (E - 3) * bs + tid;
        fbA[at] = f | bit;
        mvA[at] = moves;
        odA[at] = order;
        wdA[at] = pack_frame(alpha, beta, best, sq, passed);
        if (passed) {
          P = cO;
          O = cP;
        } else {
          P = cP;
          O = cO;
          const int t = alpha;
          alpha = -beta;
          beta = -t;
        }
        moves = cm;
        best = -64;
        --E;
        order = 0;
        scoring = ORDER >= 1 && E >= a.kff && (cm & (cm - 1));
        to_score = cm;
        keys = 0;
        nkeys = 0;
        continue;
      }
    }

    // A child value is known: commit it, then return from every finished node.
    best = max(best, v);
    alpha = max(alpha, best);
    while (alpha >= beta || !moves) {
      if (E == Eroot) {
        a.scores[idx] = (int8_t)(negate ? -best : best);
        active = false;
        break;
      }
      ++E;
      const int at = (E - 3) * bs + tid;
      const uint64_t fb = fbA[at];
      moves = mvA[at];
      order = odA[at];
      const uint32_t w = wdA[at];
      const int palpha = (int)(int8_t)(w & 0xFF);
      const int pbeta = (int)(int8_t)((w >> 8) & 0xFF);
      const int pbest = (int)(int8_t)((w >> 16) & 0xFF);
      const uint64_t pbit = 1ULL << ((w >> 24) & 63);
      int vv;
      if ((w >> 30) & 1) {  // child was a pass frame: same side to move as here
        const uint64_t nP = P ^ fb;
        O = O ^ fb ^ pbit;
        P = nP;
        vv = best;
      } else {
        const uint64_t nP = O ^ fb;
        O = P ^ fb ^ pbit;
        P = nP;
        vv = -best;
      }
      beta = pbeta;
      best = max(pbest, vv);
      alpha = max(palpha, best);
    }
  }

  atomicAdd(&a.ctr->positions, c_pos);
  atomicAdd(&a.ctr->nodes, c_nodes);
  atomicAdd(&a.ctr->score_steps, c_steps);
  atomicAdd(&a.ctr->iterations, c_iters);
}

__global__ void reset_counters_kernel(SolveCounters* c) {
  c->next = 0;
  c->error = 0;
  c->positions = 0;
  c->nodes = 0;
  c->score_steps = 0;
  c->iterations = 0;
}

// ---------------------------------------------------------------------------
// Host side: pick the instantiation whose stack covers DEPTH, size the grid
// to what stays resident, launch.

using KernelFn = void (*)(SolveArgs);

template <int MAXE>
static KernelFn pick_order(int ordering) {
  switch (ordering) {
    case GO2_ORDER_STATIC: return solve_kernel<MAXE, 0>;
    default: return solve_kernel<MAXE, 1>;
  }
}

static int stack_levels_for(int max_depth, KernelFn* fn, int ordering) {
  const int d = max_depth < 3 ? 3 : max_depth;
  if (d <= 12) { *fn = pick_order<12>(ordering); return 12; }
  if (d <= 16) { *fn = pick_order<16>(ordering); return 16; }
  if (d <= 20) { *fn = pick_order<20>(ordering); return 20; }
  if (d <= 24) { *fn = pick_order<24>(ordering); return 24; }
  if (d <= 28) { *fn = pick_order<28>(ordering); return 28; }
  *fn = pick_order<32>(ordering);
  return 32;
}

int launch_solve(const go2_board* d_boards, size_t n, int max_depth, int8_t* d_scores,
                 const go2_config& cfg, SolveCounters* d_ctr, hipStream_t stream, hipEvent_t ev0,
                 hipEvent_t ev1) {
  if (n > 0xFFFFFFF0ULL) return fail(GO2_E_ARG, "too many boards in one solve (max 2^32-16)");
  KernelFn fn = nullptr;
  const int maxe = stack_levels_for(max_depth, &fn, cfg.ordering);
  const int bs = cfg.block_size;
  const size_t lds = (size_t)(maxe - 2) * bs * 28;
  hipError_t e;
  if (lds > 65536) {
    e = hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(dynamic LDS)");
  }
  int dev = 0, cus = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return hip_fail(e, "hipGetDevice");
  if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return hip_fail(e, "hipDeviceGetAttribute(CUs)");
  int per_cu = 0;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, bs, lds)) != hipSuccess)
    return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  if (per_cu < 1) return fail(GO2_E_ARG, "block size / stack depth do not fit one CU");
  if (cfg.blocks_per_cu > 0 && cfg.blocks_per_cu < per_cu) per_cu = cfg.blocks_per_cu;
  size_t grid = (size_t)cus * per_cu;
  const size_t need = (n + bs - 1) / bs;
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;

  reset_counters_kernel<<<1, 1, 0, stream>>>(d_ctr);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "reset_counters launch");
  SolveArgs args{d_boards, d_scores, (unsigned int)n, max_depth, cfg.mobility_min_empties, d_ctr};
  if (ev0 && (e = hipEventRecord(ev0, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(bs), lds, stream, args);
  if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "solve kernel launch");
  if (ev1 && (e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return GO2_OK;
}

int collect_stats(SolveCounters* d_ctr, size_t n, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1,
                  go2_stats* stats) {
  SolveCounters h{};
  hipError_t e = hipMemcpyAsync(&h, d_ctr, sizeof(h), hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(counters)");
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  if (h.error) {
    return fail(GO2_E_BOARD, "board " + std::to_string(h.error - 1) +
                                 " is invalid (overlapping discs or more empties than DEPTH)");
  }
  if (stats) {
    *stats = go2_stats{};
    stats->positions = h.positions;
    stats->nodes = h.nodes;
    stats->score_steps = h.score_steps;
    stats->iterations = h.iterations;
    float ms = 0.f;
    if (ev0 && ev1 && hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess) stats->kernel_ms = ms;
  }
  (void)n;
  return GO2_OK;
}

}  // namespace go2
