// Batched exact endgame solver kernel for gfx950.
//
// What it replaces: alpha_beta_kernel and the Solver state machine of the
// reference (solver.cu:15-311), which runs one CUDA thread per board with
// 80-byte global-memory frames for the top DEPTH-9 plies and 24-byte
// shared-memory frames for the last 10, and a per-block stride queue.
//
// Design here (wave64):
//   * Persistent grid sized to what stays resident on every CU.  Each
//     wavefront owns up to K boards at a time (K = positions_per_wave, from
//     the batch size so that every wavefront gets work); owner lanes pull
//     board indices from one device counter with one atomic per wavefront.
//   * Lanes with no board of their own help: once a node with at least
//     split_min empties has searched its first child (young brothers wait),
//     an idle lane of the same wavefront takes the next untried sibling,
//     searches it with the node's current window and merges its fail-soft
//     value into the node (max), while the owner keeps searching the other
//     siblings.  The owner returns from the node only when every taken
//     sibling has reported.  A merged value with a window that was narrower
//     when the sibling started is still a correct fail-soft bound, so the
//     exact root score is unchanged (any order, any split gives the same
//     minimax value).  All of this stays inside one wavefront: hand-offs,
//     results and stop requests travel by ballot and readlane in
//     wavefront-uniform phases, and every frame is touched only by the lane
//     that owns it, so no atomics or fences are needed.  A node that is cut
//     while helpers still search its siblings stops them (and their own
//     helpers, one level per loop trip) instead of waiting them out.
//   * The node being searched lives in VGPRs.  Ancestors are spilled to a
//     stack, one frame per empties level: levels 4..(lds_levels+3) in LDS
//     (struct-of-arrays, lane index fastest, bank-conflict free; 4..12 with
//     the default 256-lane workgroups), higher levels in HBM (same layout
//     over all lanes of the grid; touched only by the rare nodes near the
//     root).  A frame holds the flip set of the
//     move taken (the parent board is rebuilt from the child by XORs), the
//     untried moves, the packed move order and alpha/beta/best/move/flags.
//     Frames at split levels add a split word: helpers still out, the best
//     value they reported, the node's entry alpha and best move (for the
//     table) and the siblings out with other wavefronts.
//   * Every loop trip does the same heavy work on every searching lane: one
//     flip and one mobility.  Work that would leave most lanes idle is
//     batched across the wavefront instead: a node with three empties
//     queues all its children and, once enough are queued, all 64 lanes
//     score queued two-empties children in straight-line code (the leaf
//     batch); a node that orders its moves queues all its children and the
//     lanes compute their ordering keys together and rank them by ballots
//     (the score batch).
//   * Aspiration windows at board roots (learned ordering) and null-window
//     searches after the first child from 8 empties, each re-searched when
//     the bound it returns falls inside the window.
//   * Across wavefronts: while some wavefront is idle, busy ones export
//     siblings of high nodes to help slots (more, and lower, as the idle
//     fraction grows); idle wavefronts claim them and report through
//     per-frame mailboxes; owners take back unclaimed siblings.  Wavefronts
//     count themselves active when they start and leave once none is.
//   * Optional lock-free transposition table (64-bit multiply-xorshift keys,
//     XOR-checked 16-byte buckets, launch generations), probed when children
//     are pushed.
//   * Best moves: values reaching a board's root carry the move and an
//     "exact" flag; the largest such key names an optimal move.
//   * THINK instantiation: the reference's think mode, a depth-limited
//     search scored by the pattern evaluator at its horizon.
//   * Exact fail-soft negamax, full window (-64, 64) at the root; passes
//     are folded into the child frame (board swapped, window kept, value not
//     negated), the recurrence of solver.cu:65-73,162-165,214-221,246-256.
//   * A wall-clock deadline (s_memrealtime) bounds every launch.
#include <cstring>
#include <mutex>

#include "bitboard.hpp"
#include "internal.hpp"
#include "pattern.hpp"
#include "plan.hpp"

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

template <int ORDER>
__device__ __forceinline__ int pick_static(uint64_t m, uint64_t empty) {
  if constexpr (ORDER == 0) {
    return pick_plain(m);
  } else {
    return pick_parity(m, empty);
  }
}

// Packed sorted move list: up to 9 entries of 7 bits (63 bits), entry =
// square + 1, first entry in the low bits.  Keys are kept the same way and
// compared field-parallel.
constexpr uint64_t kLo7 = 0x0102040810204081ULL;  // bit 0 of each 7-bit field
constexpr uint64_t kHi7 = kLo7 << 6;              // bit 6 of each field
constexpr uint64_t kFull63 = 0x7FFFFFFFFFFFFFFFULL;
static_assert(__builtin_popcountll(kLo7) == 9, "nine 7-bit fields");

// Frame word.  Every window bound and value in this search is even (final
// scores are even, the root window is (-64, 64) and only max / negation
// combine them), so alpha, beta and best are kept halved in 7 bits each
// (bits 0-6, 7-13, 14-20); then the move square (21-26), the
// child-was-a-pass flag (27), the number of finished children saturated at
// 3 (28-29) and the expected node type (30-31).
enum NodeType { kPV = 0, kCut = 1, kAll = 2 };

__device__ __forceinline__ uint32_t pack_frame(int alpha, int beta, int best, int sq, bool passed,
                                               int done, int type) {
  return ((uint32_t)(alpha >> 1) & 127u) | (((uint32_t)(beta >> 1) & 127u) << 7) |
         (((uint32_t)(best >> 1) & 127u) << 14) | ((uint32_t)sq << 21) | ((uint32_t)passed << 27) |
         ((uint32_t)done << 28) | ((uint32_t)type << 30);
}

// Even value held halved in the 7-bit field starting at `shift`.
__device__ __forceinline__ int frame_val(uint32_t w, int shift) {
  return ((int)(w << (25 - shift)) >> 25) * 2;
}

// Think mode's values need not be even (rounded evaluations) and it has no
// helpers, so its frame word keeps alpha + 64 (bits 0-7), beta + 64 (8-15)
// and best + 64 (low five bits in 16-20, high three in 28-30) instead of the
// split bookkeeping; the move square and pass flag stay at 21-27.
template <bool THINK>
__device__ __forceinline__ uint32_t frame_word(int alpha, int beta, int best, int sq, bool passed, int done,
                                               int type) {
  if constexpr (THINK) {
    const uint32_t b = (uint32_t)(best + 64);
    return (uint32_t)(alpha + 64) | ((uint32_t)(beta + 64) << 8) | ((b & 31u) << 16) | ((uint32_t)sq << 21) |
           ((uint32_t)passed << 27) | ((b >> 5) << 28);
  } else {
    return pack_frame(alpha, beta, best, sq, passed, done, type);
  }
}
template <bool THINK>
__device__ __forceinline__ int fw_alpha(uint32_t w) {
  if constexpr (THINK) return (int)(w & 255u) - 64;
  else return frame_val(w, 0);
}
template <bool THINK>
__device__ __forceinline__ int fw_beta(uint32_t w) {
  if constexpr (THINK) return (int)((w >> 8) & 255u) - 64;
  else return frame_val(w, 7);
}
template <bool THINK>
__device__ __forceinline__ int fw_best(uint32_t w) {
  if constexpr (THINK) return (int)(((w >> 16) & 31u) | (((w >> 28) & 7u) << 5)) - 64;
  else return frame_val(w, 14);
}

// Expected type of a child: the first child of a PV node is PV, its other
// children are CUT nodes; children of CUT nodes are ALL nodes and children
// of ALL nodes are CUT nodes.
__device__ __forceinline__ int child_type(int type, int done) {
  return type == kPV ? (done == 0 ? kPV : kCut) : (type == kCut ? kAll : kCut);
}

constexpr int kNoValue = -127;  // "no helper reported yet"

// A child value on its way into a node, with the move that produced it:
// ((value + 64) << 1 | exact) << 6 | square, where exact says the value beat
// the alpha the child was searched with; the largest key over a node's
// children names a move that reaches the node's final value.  Negative =
// none.
__device__ __forceinline__ int move_key(int v, bool exact, int sq) {
  return ((((v + 64) << 1) | (int)exact) << 6) | sq;
}
__device__ __forceinline__ int key_value(int k) { return (k >> 7) - 64; }
constexpr int kMovePass = 64, kMoveNone = 65;
// A leaf batch runs at the latest when its oldest job has waited this many trips.
constexpr int kLeafAge = 3;
// ... and a score batch when its oldest node has waited this many.
constexpr int kScoreAge = 3;
// Learned ordering scores a child by the evaluator one ply deeper (its
// replies) at nodes with at least this many empties: about ten evaluations
// per child instead of one, paid only near the roots of deep boards.
constexpr int kPlyMin = 15;
// Mobility ordering uses the capped fastest-first key with quadrant-parity
// ties at nodes with at most this many empties (above it, more than 8
// replies are common and the uncapped count orders better).
constexpr int kParityKeyMax = 10;
// The capped key (replies <= 8, seven classes per count) and the uncapped one
// (replies <= 15, four per count) stay below 64: the score batch ranks
// 6-bit keys.
static_assert(8 * 7 + 5 + 1 <= 63 && 15 * 4 + 3 <= 63, "ordering keys fit 6 bits");

struct SolveArgs {
  const go2_board* boards;
  int8_t* scores;
  uint8_t* moves;  // best move per board (square, kMovePass, kMoveNone), or null
  unsigned int n;
  int depth;          // DEPTH: boards with more empties are rejected
  int kff;            // fastest-first ordering at nodes with >= kff empties
  int split_min;      // helpers take siblings at nodes with >= split_min empties
  int cut_after;      // ... at expected CUT nodes only after this many children
  int lds_levels;     // frame levels 4..lds_levels+3 live in LDS
  int K;              // boards owned by one wavefront at a time
  unsigned int total_lanes;
  unsigned long long limit_ticks;  // time limit in s_memrealtime ticks (100 MHz); 0 = none
  // learned ordering (ORDER 2) at nodes with >= learned_min empties:
  // key = clamp(((eval(child) * key_eval_mul) >> 12) + replies * key_mob_mul + 32, 0, 63)
  const int16_t* ev_w;
  const int16_t* ev_bias;
  int learned_min;
  int key_eval_mul;
  int key_mob_mul;
  int max_pops;       // finished nodes a lane returns from per loop trip
  int leaf_batch;     // nodes with three empties queue their children and the wavefront evaluates
                      // them together once this many are queued (0 = each lane on its own)
  int score_batch;    // nodes that order their children queue them and the wavefront computes
                      // the keys together once this many are queued (0 = one child per trip)
  int pvs_min;        // children after the first of nodes with >= pvs_min empties are searched
                      // with a null window first (0 = full windows)
  int asp_width;      // ORDER 2: board roots with >= asp_min empties are searched with the window
  int asp_min;        // evaluator estimate +- asp_width (discs, even), re-searched once with one
                      // bound open when the value falls outside (0 = full window)
  int think_depth;    // think mode: moves before the evaluator scores a position
  float ev_scale;     // evaluator units per disc
  // transposition table (null = none): probed when pushing children with
  // >= tt_min empties, written when such nodes finish
  ulonglong2* tt;
  unsigned long long tt_mask;
  int tt_min;
  const unsigned int* tt_gen;  // launch generation (device, set by the reset kernel)
  // cross-wavefront help (null = off): while some wavefront is idle, busy
  // wavefronts export siblings to `slots`: of nodes with >= ext_min empties
  // when few wavefronts are idle, down to ext_lo empties when nearly all
  // are; idle wavefronts claim them and report through per-frame mailboxes
  // [E - ext_lo][total_lanes]
  struct HelpSlot* slots;
  unsigned int nslots;
  int ext_min;
  int ext_lo;
  int* mb_best;
  unsigned int* mb_done;
  unsigned int* mb_cut;
  // [E - ext_lo][kExtRec][total_lanes]: 1 + (slot index << 6 | move square)
  // of a sibling that may still be unclaimed, 0 = none
  unsigned int* mb_rec;
  SolveCounters* ctr;
  // HBM frames of levels >= lds_levels+4: [level - lds_levels - 4][total_lanes]
  uint64_t* g_fb;
  uint64_t* g_mv;
  uint64_t* g_od;
  uint32_t* g_wd;
  uint32_t* g_sx;
};
// The graph key holds kernel, grid, block, LDS bytes and these arguments.
static_assert(sizeof(SolveArgs) + 4 * sizeof(size_t) <= 512, "graph key too small");

// Split word of a frame (levels >= split_min): helpers still out on this
// node (bits 0-7) and the best value they reported, halved and offset by 64
// (bits 8-15; 0 = none reported).  Only the owning lane reads or writes it.
__device__ __forceinline__ int sx_pending(uint32_t x) { return (int)(x & 255u); }
__device__ __forceinline__ int sx_best(uint32_t x) {
  const int b = (int)((x >> 8) & 255u);
  return b ? (b - 64) * 2 : kNoValue;
}
__device__ __forceinline__ uint32_t sx_with_best(uint32_t x, int v) {
  const uint32_t b = (uint32_t)(v / 2 + 64);
  return ((x >> 8) & 255u) >= b ? x : ((x & ~0xFF00u) | (b << 8));
}
// Bits 16-22 of the split word: the node's alpha at entry, halved (for the
// bound type of its transposition-table entry).
__device__ __forceinline__ uint32_t sx_init(int alpha0) { return ((uint32_t)(alpha0 >> 1) & 127u) << 16; }
__device__ __forceinline__ int sx_alpha0(uint32_t x) { return ((int)(x << 9) >> 25) * 2; }

// ---- transposition table ------------------------------------------------------
// Key: a 64-bit hash of both bitboards, K(P, O) = fmix64(P ^ fmix64(O ^ c)),
// fmix64 being the bijective finalizer of MurmurHash3 (two multiply-xorshift
// rounds).  It costs four 64-bit multiplies and no memory access, where
// tabulated Zobrist keys (one random word per byte position and value) cost
// 16 dependent cache loads per key.  For a fixed O it is a bijection of P,
// and positions with different O differ by fmix64 outputs, so two positions
// of one search share a key with probability ~2^-64, as with Zobrist keys.
// Bucket: (key ^ data, data), so a torn or foreign entry fails the key check
// (lock-free, no CAS).  Data: lower + 64 (bits 0-7), upper + 64 (8-15), best
// move (16-21, 63 = none), generation (24-31), empties (32-37).
__device__ __forceinline__ uint64_t fmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xFF51AFD7ED558CCDULL;
  x ^= x >> 33;
  x *= 0xC4CEB9FE1A85EC53ULL;
  x ^= x >> 33;
  return x;
}
__device__ __forceinline__ uint64_t tt_key(uint64_t P, uint64_t O) {
  return fmix64(P ^ fmix64(O ^ 0x9E3779B97F4A7C15ULL));
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)x, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// HBM frame accesses.  Written as relaxed wavefront-scope atomics (plain
// global loads/stores in the ISA) so that the compiler cannot fold an
// LDS-or-HBM pair into one slower flat access.
template <class T>
__device__ __forceinline__ T hbm_ld(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
template <class T>
__device__ __forceinline__ void hbm_st(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// One exported sibling: board, window, expected type and the owner's frame.
struct HelpSlot {
  uint64_t P, O;
  int alpha, beta, ctype, owner_E;
  unsigned int owner;  // global lane id of the owner
  unsigned int state;  // kSlotEmpty / kSlotWriting / full (tag << 2 | kSlotFull) / kSlotTaken
  int sq;              // the sibling's move square
  unsigned int pad[5];  // 64 bytes: one slot per cache line
};
constexpr unsigned int kSlotEmpty = 0, kSlotWriting = 1, kSlotFull = 2, kSlotTaken = 3;
// A full slot carries its owner's tag (global lane id < 2^19, frame level
// and move square), so an owner taking back an unclaimed sibling can never
// cancel another item that reused the slot, its own included.
__device__ __forceinline__ unsigned int slot_full(unsigned int g, int e, int sq) {
  return ((((g << 5 | (unsigned int)e) << 6) | (unsigned int)sq) << 2) | kSlotFull;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)x, src);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(x >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// A queued child of a node with three empties: play sq on (P, O) and score
// the resulting two-empties position exactly; returns the value for the
// node's side.  The child's mover may have no move there: the first round
// (swap false) then returns kNeedPass and the node queues the same move
// again with swap true, which scores the child with the node's side to move
// (or the final count when neither side can move).  A whole batch of lanes
// thus never runs a pass branch that only a few of them need.
constexpr int kNeedPass = 126;
template <class T>
__device__ __forceinline__ int leaf_job(uint64_t P, uint64_t O, int sq, bool swap, const T& lastc, int& nodes) {
  const uint64_t f = flip(P, O, sq);
  const uint64_t bit = 1ULL << sq;
  const uint64_t cP = swap ? P ^ f ^ bit : O ^ f;  // side to move in the scored position
  const uint64_t cO = swap ? O ^ f : P ^ f ^ bit;
  ++nodes;  // the child, or (swap) its pass
  const int r = solve2_mover(cP, cO, lastc, nodes);
  if (r != kNoMove2) return swap ? r : -r;
  return swap ? final_score(cP, cO) : kNeedPass;
}

template <class T>
__device__ __forceinline__ T ld_agent(T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Split word bits 29-31: siblings of this node out with other wavefronts
// (up to kExtRec; not yet reported, or unclaimed and recorded so that the
// owner can take them back).
__device__ __forceinline__ int sx_ext(uint32_t x) { return (int)(x >> 29); }
constexpr uint32_t kSxExtOne = 1u << 29;

// Board of the parent of (P, O), from the frame word and flip set the parent
// stored when it moved to (P, O).
__device__ __forceinline__ void undo_move(uint64_t& P, uint64_t& O, uint64_t fb, uint32_t w) {
  const uint64_t bit = 1ULL << ((w >> 21) & 63);
  if ((w >> 27) & 1) {  // the child was a pass frame: same side to move
    const uint64_t nP = P ^ fb;
    O = O ^ fb ^ bit;
    P = nP;
  } else {
    const uint64_t nP = O ^ fb;
    O = P ^ fb ^ bit;
    P = nP;
  }
}

// THINK: the reference's think mode (thinker.cu:120-320) instead of the exact
// solve: the position after the think_depth-th move from a board is scored
// by the evaluator (rounded to whole discs); passes cost no depth.  One lane
// per board (no helpers, no table: stored values would depend on each
// board's horizon), learned ordering.
template <int ORDER, bool THINK = false>
// Two wavefronts per SIMD: 256 VGPRs hold the whole loop state without
// spills, and the two workgroups of 256 lanes per CU keep nine stack levels
// (4..12) in LDS; three per SIMD (168 VGPRs, spills, six LDS levels) ran no
// faster.
__global__ void __launch_bounds__(256, 2) solve_kernel(SolveArgs a) {
  static_assert(!THINK || ORDER == 2, "think mode needs the evaluator's tables");
  extern __shared__ uint64_t lds[];
  const int bs = blockDim.x, tid = threadIdx.x;
  const int lane = __lane_id();
  const unsigned int gtid = blockIdx.x * bs + tid;
  const int LL = a.lds_levels;
  uint64_t* l_fb = lds;
  uint64_t* l_mv = l_fb + LL * bs;
  uint64_t* l_od = l_mv + LL * bs;
  uint32_t* l_wd = reinterpret_cast<uint32_t*>(l_od + LL * bs);
  uint32_t* l_sx = l_wd + LL * bs;
  uint8_t* lastc = reinterpret_cast<uint8_t*>(l_sx + LL * bs);  // last-square flip counts
  for (int k = tid; k < 2048; k += bs) lastc[k] = (uint8_t)line_flip_count(k >> 8, k & 255);
  uint16_t* t3 = reinterpret_cast<uint16_t*>(lastc + 2048);  // base-3 table (ORDER 2)
  // leaf batch: worker slot -> requesting lane, 64 bytes per wavefront
  uint8_t* lmap = lastc + 2048 + (ORDER == 2 ? kBase3Size * 2 : 0) + (tid & ~63);
  if constexpr (ORDER != 2) __syncthreads();
  if constexpr (ORDER == 2) {
    for (int b = tid; b < kBase3Size; b += bs) {
      uint32_t v = 0, p = 1;
      for (int k = 0; k < 10; ++k) {
        v += ((b >> k) & 1) * p;
        p *= 3;
      }
      t3[b] = (uint16_t)v;
    }
    __syncthreads();
  }

  // Every frame slot is read and written only by the lane that owns it.
  // LDS levels use ds_* accesses and HBM levels global ones (no flat
  // addressing on the hot path).
  // Frames are indexed by E - ebase: ebase is 0 for the exact solve; in
  // think mode it is (horizon empties - 2) of the lane's board, so the
  // frames a search can push are levels 4..think_depth+2.
  int ebase = 0;
  auto lds_k = [&](int E) { return (E - ebase - kFrameBase) * bs + tid; };
  auto hbm_k = [&](int E) { return (size_t)(E - ebase - kFrameBase - LL) * a.total_lanes + gtid; };
  auto in_lds = [&](int E) { return E - ebase - kFrameBase < LL; };
  // Split words: only the owning lane reads or writes them, and every frame
  // a sibling is taken from is an ancestor of the owner's current node, so
  // the current node's word lives in a register (cur_sx, below) and memory
  // holds the ancestors' words.
  auto get_sx = [&](int E) -> uint32_t { return in_lds(E) ? l_sx[lds_k(E)] : hbm_ld(&a.g_sx[hbm_k(E)]); };
  auto set_sx = [&](int E, uint32_t v) {
    if (in_lds(E)) {
      l_sx[lds_k(E)] = v;
    } else {
      hbm_st(&a.g_sx[hbm_k(E)], v);
    }
  };
  auto get_fb = [&](int E) -> uint64_t { return in_lds(E) ? l_fb[lds_k(E)] : hbm_ld(&a.g_fb[hbm_k(E)]); };
  auto get_mv = [&](int E) -> uint64_t { return in_lds(E) ? l_mv[lds_k(E)] : hbm_ld(&a.g_mv[hbm_k(E)]); };
  auto get_od = [&](int E) -> uint64_t { return in_lds(E) ? l_od[lds_k(E)] : hbm_ld(&a.g_od[hbm_k(E)]); };
  auto get_wd = [&](int E) -> uint32_t { return in_lds(E) ? l_wd[lds_k(E)] : hbm_ld(&a.g_wd[hbm_k(E)]); };
  auto set_mv_od = [&](int E, uint64_t mv, uint64_t od) {
    if (in_lds(E)) {
      l_mv[lds_k(E)] = mv;
      l_od[lds_k(E)] = od;
    } else {
      hbm_st(&a.g_mv[hbm_k(E)], mv);
      hbm_st(&a.g_od[hbm_k(E)], od);
    }
  };
  auto push_frame = [&](int E, uint64_t fb, uint64_t mv, uint64_t od, uint32_t wd) {
    if (in_lds(E)) {
      const int k = lds_k(E);
      l_fb[k] = fb;
      l_mv[k] = mv;
      l_od[k] = od;
      l_wd[k] = wd;
    } else {
      const size_t k = hbm_k(E);
      hbm_st(&a.g_fb[k], fb);
      hbm_st(&a.g_mv[k], mv);
      hbm_st(&a.g_od[k], od);
      hbm_st(&a.g_wd[k], wd);
    }
  };

  // Lane state.
  // LEAFWAIT: at a node with three empties whose children are queued for
  // the wavefront's next leaf batch.
  // SCOREWAIT: at a node whose children are queued for the wavefront's next
  // score batch (their ordering keys).
  enum { IDLE = 0, SEARCH = 1, WAIT = 2, ABORT = 3, REPORT = 4, LEAFWAIT = 5, SCOREWAIT = 6 };
  int mode = IDLE;
  uint64_t P = 0, O = 0, moves = 0, order = 0, to_score = 0, keys = 0;
  int alpha = 0, beta = 0, best = 0, E = 0, Eroot = 0, nkeys = 0;
  uint32_t cur_sx = 0;  // split word of the current node (levels >= split_min)
  // split word of level e: the register for the current level, else memory
  auto sx_of = [&](int e) -> uint32_t { return e == E ? cur_sx : get_sx(e); };
  auto sx_put = [&](int e, uint32_t v) {
    if (e == E) {
      cur_sx = v;
    } else {
      set_sx(e, v);
    }
  };
  bool scoring = false, negate = false;
  bool asp = false;             // the board's root is searched with an aspiration window
  uint32_t nw_mask = 0;         // bit E-3: the child of frame E is searched with a null window
  bool full_next = false;       // the next child of the current node is a re-search: full window
  bool pchk = false;            // a child without moves: decide pass / game over next trip
  bool leaf_swap = false;       // LEAFWAIT: the queued children are passes (see leaf_job)
  int pchk_sq = 0;
  uint64_t pchk_f = 0;
  bool pop_pending = false;     // more finished nodes to return from next trip
  int ndone = 0, ntype = kPV;   // finished children (saturating at 3) and expected type
  int owner = -1, owner_E = 0;  // owner lane and frame level of a helper's task (-1: own board)
  bool owner_ext = false;       // the owner is in another wavefront (owner = its global lane id)
  uint32_t ext_mask = 0;        // bit E-3: frame E has siblings out with other wavefronts
  bool wave_idle = false;       // wavefront-uniform: counted out of active_waves
  bool demand = false;          // wavefront-uniform: some wavefront is idle
  int dem_e = 64;               // wavefront-uniform: lowest level to export from
  unsigned int poll_base = gtid - (unsigned int)lane;  // wavefront-uniform
  // For the table's entries: the current node's alpha at entry, halved
  // (bits 6-12, two's complement), and the move that gave its best value
  // (bits 0-5, 63 = none), in one register.
  int a0mv = 63;
  auto alpha0_of = [&]() { return ((a0mv << 19) >> 25) * 2; };
  auto bestmv_of = [&]() { return a0mv & 63; };
  unsigned int idx = 0;
  uint32_t split_mask = 0;      // bit E-3: this lane's frame at level E offers siblings
  int helpers_out = 0;          // helpers still working on this lane's frames
  int dead_level = 64;          // frame whose helpers must stop: E, 0 = all, 64 = none
  bool dead_new = false;        // dead_level changed; helpers not told yet
  bool report_aborted = false;
  // One register, three uses that never overlap: a board's owner keeps the
  // best move_key so far at the board's root (-1 = none); a helper keeps
  // (owner alpha + 64) << 6 | move square of its task while searching it,
  // then the task's move_key while reporting it.
  int mkey = -1;
  bool queue_done = false;      // wavefront-uniform
  // Per-lane statistics, 32-bit (a lane adds a few per trip) and flushed to
  // the 64-bit launch counters every 65,536 trips and at exit.
  unsigned int c_pos = 0, c_nodes = 0, c_steps = 0, c_iters = 0, c_steals = 0;
  unsigned int c_probes = 0, c_hits = 0, c_cuts = 0, c_stores = 0, c_exports = 0;
  auto flush_counters = [&]() {
    auto add = [](unsigned long long* dst, unsigned int& v) {
      if (v) atomicAdd(dst, (unsigned long long)v);
      v = 0;
    };
    add(&a.ctr->positions, c_pos);
    add(&a.ctr->nodes, c_nodes);
    add(&a.ctr->score_steps, c_steps);
    add(&a.ctr->iterations, c_iters);
    add(&a.ctr->steals, c_steals);
    add(&a.ctr->exports, c_exports);
    add(&a.ctr->tt_probes, c_probes);
    add(&a.ctr->tt_hits, c_hits);
    add(&a.ctr->tt_cuts, c_cuts);
    add(&a.ctr->tt_stores, c_stores);
  };
  auto mb_idx = [&](int e, unsigned int g) { return (size_t)(e - a.ext_lo) * a.total_lanes + g; };
  auto rec_idx = [&](int e, int k, unsigned int g) {
    return ((size_t)(e - a.ext_lo) * kExtRec + (size_t)k) * a.total_lanes + g;
  };

  // Ordering key of the child (cP, cO) with replies cm, reached by the move
  // `bit` from a node with E empties: fastest first (fewest replies),
  // corners preferred, X-squares delayed.  Where reply counts stay small
  // (learned ordering below learned_min empties, mobility ordering up to
  // kParityKeyMax empties) they are capped at 8 and each count is ordered by
  // corner / plain / X-square and then a move into a 4x4 quadrant that held
  // an odd number of empties (the quadrant now holds an even number).  From
  // learned_min empties the learned evaluator of the child joins the key;
  // in the exact search (not think mode) from kPlyMin empties the child's
  // value is the evaluator one ply deeper, the opponent's best reply.
  auto move_key_of = [&](uint64_t cP, uint64_t cO, uint64_t cm, uint64_t bit, int e) -> int {
    int k;
    if ((ORDER == 2 && e < a.learned_min) || (ORDER == 1 && e <= kParityKeyMax)) {
      const int bsq = ctz64(bit);
      const uint64_t quad = 0x000000000F0F0F0FULL << ((((bsq >> 2) & 1) << 2) | (((bsq >> 5) & 1) << 5));
      k = min(popc64(cm), 8) * 7 + ((bit & kCorners) ? 0 : ((bit & kXSquares) ? 5 : 3)) +
          (popc64(~(cP | cO) & quad) & 1);
    } else {
      k = min(popc64(cm), 15) * 4 + ((bit & kCorners) ? 0 : 2) + ((bit & kXSquares) ? 1 : 0);
    }
    if constexpr (ORDER == 2) {
      if (e >= a.learned_min) {
        // One evaluation site: the child itself, or each of its replies
        // (the opponent's value is the largest -eval over them).
        const bool ply = !THINK && e >= kPlyMin && cm;
        uint64_t rm = ply ? cm : 1ULL;
        int lo = 1 << 30;
        do {
          uint64_t xP = cP, xO = cO;
          int bi = e - 1;
          if (ply) {
            const int rs = ctz64(rm);
            const uint64_t rf = flip(cP, cO, rs);
            xP = cO ^ rf;
            xO = cP ^ rf ^ (1ULL << rs);
            bi = e - 2;
          }
          rm &= rm - 1;
          int g = a.ev_bias[bi];
          for_each_feature(xP, xO, t3, [&](int, uint32_t fi) { g += a.ev_w[fi]; });
          lo = min(lo, ply ? g : -g);
        } while (rm);
        const int ev = -lo;
        k = min(max(((ev * a.key_eval_mul) >> 12) + popc64(cm) * a.key_mob_mul + 32, 0), 63);
      }
    }
    return k;
  };

  // A node at level E with legal moves `cm` becomes the current node.
  auto enter_node = [&](uint64_t cm, int type) {
    moves = cm;
    best = -64;
    order = 0;
    scoring = ORDER >= 1 && E >= a.kff && (cm & (cm - 1));
    to_score = cm;
    keys = 0;
    nkeys = 0;
    ndone = 0;
    ntype = type;
    pchk = false;
    pop_pending = false;
    a0mv = (((alpha >> 1) & 127) << 6) | 63;
    if (E >= a.split_min) cur_sx = sx_init(alpha);
    if (!THINK && scoring && a.score_batch) mode = SCOREWAIT;
    if (!THINK && E == 3 && a.leaf_batch && cm) {
      // every child is queued at once and scored exactly (an exact value is
      // a valid fail-soft result whatever the window)
      mode = LEAFWAIT;
      leaf_swap = false;
      scoring = false;
    }
  };

  // This lane starts searching child (cP, cO) of frame oe of its owner, with
  // the owner's window (ca, cb) on that frame.
  auto start_child_task = [&](uint64_t cP, uint64_t cO, int ca, int cb, int ct, int own, int oe,
                              bool ext, int csq) {
    ++c_nodes;
    mkey = ((ca + 64) << 6) | csq;
    nw_mask = 0;
    full_next = false;
    E = oe - 1;
    Eroot = E;
    owner = own;
    owner_ext = ext;
    owner_E = oe;
    dead_level = 64;
    const uint64_t cm = mobility(cP, cO);
    if (cm) {
      P = cP;
      O = cO;
      alpha = -cb;
      beta = -ca;
      negate = true;
      mode = SEARCH;
      enter_node(cm, ct);
    } else {
      const uint64_t om = mobility(cO, cP);
      if (om) {
        P = cO;
        O = cP;
        alpha = ca;
        beta = cb;
        negate = false;
        mode = SEARCH;
        enter_node(om, ct);
      } else {
        const int rv = -final_score(cP, cO);
        mkey = move_key(rv, rv > ca, csq);
        report_aborted = false;
        mode = REPORT;
      }
    }
  };

  const unsigned int tt_gen = a.tt ? *a.tt_gen : 0u;
  unsigned int trip = 0;
  int leaf_age = 0;   // wavefront-uniform: trips since the last leaf batch
  int score_age = 0;  // wavefront-uniform: trips since the last score batch
  const unsigned long long deadline = __builtin_amdgcn_s_memrealtime() + a.limit_ticks;
  if (a.slots && lane == 0) atomicAdd(&a.ctr->active_waves, 1u);
  for (;;) {
    // ---- time limit (checked every 256 trips; uniform per wavefront) ----
    ++trip;
    if ((trip & 0xFFFF) == 0) flush_counters();
    if (a.limit_ticks && (trip & 255) == 0) {
      if (__builtin_amdgcn_s_memrealtime() > deadline) {
        if (lane == 0) atomicExch(&a.ctr->error, kTimedOut);
        break;
      }
    }

    // ---- 1. owner lanes fetch boards: one atomic per wavefront ----
    if (!queue_done) {
      const uint64_t want = __ballot(mode == IDLE && lane < a.K);
      if (want) {
        const int leader = __builtin_ctzll(want);
        const unsigned int cnt = (unsigned int)__builtin_popcountll(want);
        unsigned int base = 0;
        if (lane == leader) base = atomicAdd(&a.ctr->next, cnt);
        base = (unsigned int)__builtin_amdgcn_readlane((int)base, leader);
        if (base + cnt >= a.n) queue_done = true;
        if ((want >> lane) & 1) {
          const unsigned int i = base + (unsigned int)__builtin_popcountll(want & __lanemask_lt());
          if (i < a.n) {
            idx = i;
            owner_ext = false;
            const go2_board b = a.boards[i];
            P = b.player;
            O = b.opponent;
            E = 64 - popc64(P | O);
            ++c_pos;
            ++c_nodes;
            if ((P & O) || E > a.depth) {
              atomicCAS(&a.ctr->error, 0u, i + 1);
              a.scores[i] = 0;
              if (a.moves) a.moves[i] = (uint8_t)kMoveNone;
            } else if (E <= 1 && (!THINK || a.think_depth > E)) {
              const int sq = ctz64(~(P | O));
              a.scores[i] = (int8_t)(E == 0 ? final_score(P, O) : last_score(P, O, sq));
              if (a.moves)
                a.moves[i] = (uint8_t)(E == 0 ? kMoveNone
                                              : flip(P, O, sq) ? sq : flip(O, P, sq) ? kMovePass : kMoveNone);
            } else if (E == 2 && (!THINK || a.think_depth > 2)) {
              int nl = 0, mv = kMoveNone;
              a.scores[i] = (int8_t)solve2(P, O, lastc, nl, &mv);
              c_nodes += nl;
              if (a.moves) a.moves[i] = (uint8_t)mv;
            } else {
              uint64_t m = mobility(P, O);
              negate = false;
              bool done = false;
              if (!m) {
                const uint64_t om = mobility(O, P);
                if (!om) {
                  a.scores[i] = (int8_t)final_score(P, O);
                  if (a.moves) a.moves[i] = (uint8_t)kMoveNone;
                  done = true;
                } else {
                  const uint64_t t = P;
                  P = O;
                  O = t;
                  m = om;
                  negate = true;
                }
              }
              if (!done) {
                Eroot = E;
                alpha = -64;
                beta = 64;
                asp = false;
                if constexpr (ORDER == 2 && !THINK) {
                  if (a.asp_width > 0 && E >= a.asp_min) {
                    // window around the evaluator's estimate of the score
                    // (side to move, whole discs rounded to an even value)
                    int ev = a.ev_bias[E];
                    for_each_feature(P, O, t3, [&](int, uint32_t fi) { ev += a.ev_w[fi]; });
                    const int g = 2 * (int)floorf((float)ev / a.ev_scale * 0.5f + 0.5f);
                    alpha = max(-64, g - a.asp_width);
                    beta = min(64, g + a.asp_width);
                    if (alpha >= beta || (alpha == -64 && beta == 64)) {
                      alpha = -64;
                      beta = 64;
                    } else {
                      asp = true;
                    }
                  }
                }
                owner = -1;
                mkey = -1;
                nw_mask = 0;
                full_next = false;
                if constexpr (THINK) ebase = E - a.think_depth - 2;
                mode = SEARCH;
                enter_node(m, kPV);
              }
            }
          }
        }
      }
    } else if (__ballot(mode != IDLE) == 0) {
      if (!a.slots) break;
      // Whole wavefront idle: claim siblings other wavefronts exported, or
      // leave once no wavefront is active (an active one may still export;
      // an exported sibling always has its active owner waiting for it).
      if (!wave_idle) {
        wave_idle = true;
        if (lane == 0) {
          atomicAdd(&a.ctr->idle_waves, 1u);
          atomicSub(&a.ctr->active_waves, 1u);
        }
      }
      const unsigned int sidx = (poll_base + (unsigned int)lane) % a.nslots;
      poll_base += 64;
      bool won = false;
      const unsigned int sst = ld_agent(&a.slots[sidx].state);
      if ((sst & 3u) == kSlotFull) {
        unsigned int expect = sst;
        won = __hip_atomic_compare_exchange_strong(&a.slots[sidx].state, &expect, kSlotTaken, __ATOMIC_ACQUIRE,
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      if (__ballot(won)) {
        if (lane == 0) {
          atomicAdd(&a.ctr->active_waves, 1u);
          atomicSub(&a.ctr->idle_waves, 1u);
        }
        wave_idle = false;
        if (won) {
          HelpSlot* hs = &a.slots[sidx];
          const uint64_t cP = ld_agent(&hs->P), cO = ld_agent(&hs->O);
          const int ca = ld_agent(&hs->alpha), cb = ld_agent(&hs->beta);
          const int ct = ld_agent(&hs->ctype), oe = ld_agent(&hs->owner_E), csq = ld_agent(&hs->sq);
          const unsigned int og = ld_agent(&hs->owner);
          __hip_atomic_store(&hs->state, kSlotEmpty, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
          ++c_steals;
          start_child_task(cP, cO, ca, cb, ct, (int)og, oe, true, csq);
        }
      } else {
        unsigned int act = 0;
        if (lane == 0) act = ld_agent(&a.ctr->active_waves);
        act = (unsigned int)__builtin_amdgcn_readlane((int)act, 0);
        if (act == 0) {
          if (lane == 0) atomicSub(&a.ctr->idle_waves, 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(16);
      }
    }

    // ---- 2. helpers report to their owners, one at a time ----
    if (mode == REPORT && owner_ext) {
      const size_t m = mb_idx(owner_E, (unsigned int)owner);
      if (!report_aborted) atomicMax(&a.mb_best[m], mkey);
      __hip_atomic_fetch_add(&a.mb_done[m], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      mode = IDLE;
    }
    if (a.slots && (trip & 15) == 0) {
      // Owners collect siblings that came back from other wavefronts; helpers
      // of other wavefronts stop when their owner's node was cut.
      if (ext_mask) {
        uint32_t em = ext_mask;
        while (em) {
          const int e = __builtin_ctz(em) + 3;
          em &= em - 1;
          const size_t m = mb_idx(e, gtid);
          uint32_t sx = sx_of(e);
          const uint32_t sx0 = sx;
          const unsigned int got =
              __hip_atomic_exchange(&a.mb_done[m], 0u, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
          if (got) {
            sx -= got * (1u + kSxExtOne);
            helpers_out -= (int)got;
            const int mb = ld_agent(&a.mb_best[m]);
            if (mb >= 0) {
              sx = sx_with_best(sx, key_value(mb));
              if (e == Eroot && owner < 0) mkey = max(mkey, mb);
            }
          }
          // Waiting on this frame (or aborting): take back every sibling
          // nobody claimed.  Otherwise, after reports, drop the records of
          // siblings that were claimed (their slot no longer holds the tag).
          const bool waiting = mode == WAIT && E == e;
          const bool take = waiting || mode == ABORT;
          if (sx_ext(sx) && (take || got)) {
            for (int k = 0; k < kExtRec; ++k) {
              const size_t ri = rec_idx(e, k, gtid);
              const unsigned int r = ld_agent(&a.mb_rec[ri]);
              if (!r) continue;
              const unsigned int sl = r - 1u;
              unsigned int expect = slot_full(gtid, e, (int)(sl & 63));
              if (take) {
                if (__hip_atomic_compare_exchange_strong(&a.slots[sl >> 6].state, &expect, kSlotEmpty,
                                                         __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)) {
                  sx -= 1u + kSxExtOne;
                  --helpers_out;
                  if (waiting && alpha < beta) moves |= 1ULL << (sl & 63);
                }
                st_agent(&a.mb_rec[ri], 0u);
              } else if (ld_agent(&a.slots[sl >> 6].state) != expect) {
                st_agent(&a.mb_rec[ri], 0u);
              }
            }
            if (waiting && alpha < beta && moves) mode = SEARCH;
          }
          if (sx != sx0) sx_put(e, sx);
          if (!sx_ext(sx)) ext_mask &= ~(1u << (e - 3));
        }
      }
      if (owner_ext && (mode == SEARCH || mode == WAIT || mode == LEAFWAIT || mode == SCOREWAIT) &&
          ld_agent(&a.mb_cut[mb_idx(owner_E, (unsigned int)owner)])) {
        mode = ABORT;
        dead_level = 0;
        dead_new = true;
        split_mask = 0;
      }
      // Export level: ext_min while few wavefronts are idle, falling
      // linearly to ext_lo as the idle fraction approaches one.
      if (lane == 0) {
        const unsigned int iw = ld_agent(&a.ctr->idle_waves);
        const unsigned int gw = gridDim.x * (blockDim.x >> 6);
        dem_e = iw ? a.ext_min - (int)(((unsigned long long)(a.ext_min - a.ext_lo + 1) * iw) / (gw + 1u)) : 64;
        if (dem_e < a.ext_lo) dem_e = a.ext_lo;
      }
      dem_e = __builtin_amdgcn_readlane(dem_e, 0);
      demand = dem_e < 64;
    }
    uint64_t rm = __ballot(mode == REPORT && !owner_ext);
    while (rm) {
      const int r = __builtin_ctzll(rm);
      rm &= rm - 1;
      const int o = __builtin_amdgcn_readlane(owner, r);
      const int oe = __builtin_amdgcn_readlane(owner_E, r);
      const int key = __builtin_amdgcn_readlane(mkey, r);
      const int ab = __builtin_amdgcn_readlane((int)report_aborted, r);
      if (lane == o) {
        uint32_t sx = sx_of(oe) - 1u;
        if (!ab) {
          sx = sx_with_best(sx, key_value(key));
          if (oe == Eroot && owner < 0) mkey = max(mkey, key);
        }
        sx_put(oe, sx);
        --helpers_out;
      }
      if (lane == r) mode = IDLE;
    }

    // ---- 3. cut frames stop their helpers (cascades one level per trip) ----
    if (dead_new && ext_mask) {
      uint32_t em = ext_mask;
      while (em) {
        const int e = __builtin_ctz(em) + 3;
        em &= em - 1;
        if (dead_level == 0 || dead_level == e) st_agent(&a.mb_cut[mb_idx(e, gtid)], 1u);
      }
    }
    uint64_t dm = __ballot(dead_new);
    while (dm) {
      const int l = __builtin_ctzll(dm);
      dm &= dm - 1;
      const int dl = __builtin_amdgcn_readlane(dead_level, l);
      if (lane == l) dead_new = false;
      if ((mode == SEARCH || mode == WAIT || mode == LEAFWAIT || mode == SCOREWAIT) && !owner_ext && owner == l &&
          (dl == 0 || dl == owner_E)) {
        mode = ABORT;
        dead_level = 0;
        dead_new = true;
        split_mask = 0;
      }
    }

    // ---- 4. an idle lane takes a sibling from the lane with the highest offer ----
    const uint64_t idle = __ballot(mode == IDLE);
    if (idle && __ballot(split_mask != 0)) {
      uint32_t lv = split_mask;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) lv |= (uint32_t)__shfl_xor((int)lv, o);
      lv = (uint32_t)__builtin_amdgcn_readfirstlane((int)lv);
      const int e = 31 - __builtin_clz(lv) + 3;
      const int victim = __builtin_ctzll(__ballot((split_mask >> (e - 3)) & 1));
      const int thief = __builtin_ctzll(idle);
      int give = 0, ca = 0, cb = 0, ct = 0, csq = 0;
      uint64_t cP = 0, cO = 0;
      if (lane == victim) {
        uint64_t mv = get_mv(e), od = get_od(e);
        const uint32_t w = get_wd(e);
        const uint32_t sx = get_sx(e);
        const int fa = fw_alpha<THINK>(w), fbeta = fw_beta<THINK>(w);
        const int am = max(fa, sx_best(sx));
        if (mv == 0 || am >= fbeta) {
          split_mask &= ~(1u << (e - 3));
        } else {
          // Rebuild the board of the node at level e from the current node.
          uint64_t bp = P, bo = O;
          for (int l = E + 1; l <= e; ++l) undo_move(bp, bo, get_fb(l), get_wd(l));
          int sq;
          if (od) {
            sq = (int)(od & 127) - 1;
            od >>= 7;
          } else {
            sq = pick_static<ORDER>(mv, ~(bp | bo));
          }
          const uint64_t bit = 1ULL << sq;
          mv ^= bit;
          set_mv_od(e, mv, od);
          set_sx(e, sx + 1u);
          ++helpers_out;
          if (!mv) split_mask &= ~(1u << (e - 3));
          const uint64_t f = flip(bp, bo, sq);
          cP = bo ^ f;
          cO = bp ^ f ^ bit;
          ca = am;
          cb = fbeta;
          ct = child_type((int)(w >> 30), 1);
          csq = sq;
          give = 1;
        }
      }
      give = __builtin_amdgcn_readlane(give, victim);
      if (give) {
        cP = readlane64(cP, victim);
        cO = readlane64(cO, victim);
        ca = __builtin_amdgcn_readlane(ca, victim);
        cb = __builtin_amdgcn_readlane(cb, victim);
        ct = __builtin_amdgcn_readlane(ct, victim);
        csq = __builtin_amdgcn_readlane(csq, victim);
        if (lane == thief) {
          ++c_steals;
          start_child_task(cP, cO, ca, cb, ct, victim, e, false, csq);
        }
      }
    } else if (a.slots && demand && !idle && __ballot((split_mask >> (dem_e - 3)) != 0)) {
      // No idle lane here but an idle wavefront elsewhere: export the
      // highest offered sibling (>= dem_e empties) to a help slot.
      uint32_t lv = split_mask >> (dem_e - 3);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) lv |= (uint32_t)__shfl_xor((int)lv, o);
      lv = (uint32_t)__builtin_amdgcn_readfirstlane((int)lv);
      const int e = 31 - __builtin_clz(lv) + dem_e;
      const int victim = __builtin_ctzll(__ballot((split_mask >> (e - 3)) & 1));
      if (lane == victim) {
        uint64_t mv = get_mv(e), od = get_od(e);
        const uint32_t w = get_wd(e);
        const uint32_t sx = get_sx(e);
        const int fa = fw_alpha<THINK>(w), fbeta = fw_beta<THINK>(w);
        const int am = max(fa, sx_best(sx));
        // a free record for taking the sibling back (records of claimed
        // siblings are dropped when their reports arrive)
        const int cnt = sx_ext(sx);
        int rk = cnt == 0 ? 0 : kExtRec;
        for (int k = 0; cnt > 0 && cnt < kExtRec && k < kExtRec; ++k)
          if (rk == kExtRec && ld_agent(&a.mb_rec[rec_idx(e, k, gtid)]) == 0u) rk = k;
        if (mv == 0 || am >= fbeta) {
          split_mask &= ~(1u << (e - 3));
        } else if (rk < kExtRec) {
          // claim an empty slot first
          unsigned int sidx = (gtid * 2654435761u + trip) % a.nslots;
          bool got = false;
          for (int k = 0; k < 8 && !got; ++k) {
            unsigned int expect = kSlotEmpty;
            got = __hip_atomic_compare_exchange_strong(&a.slots[sidx].state, &expect, kSlotWriting,
                                                       __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            if (!got) sidx = (sidx + 1) % a.nslots;
          }
          if (got) {
            uint64_t bp = P, bo = O;
            for (int l = E + 1; l <= e; ++l) undo_move(bp, bo, get_fb(l), get_wd(l));
            int sq;
            if (od) {
              sq = (int)(od & 127) - 1;
              od >>= 7;
            } else {
              sq = pick_static<ORDER>(mv, ~(bp | bo));
            }
            const uint64_t bit = 1ULL << sq;
            mv ^= bit;
            set_mv_od(e, mv, od);
            if (cnt == 0) {
              // first sibling out from this node: fresh mailbox and records
              const size_t m = mb_idx(e, gtid);
              st_agent(&a.mb_best[m], kNoValue);
              st_agent(&a.mb_done[m], 0u);
              st_agent(&a.mb_cut[m], 0u);
              for (int k = 1; k < kExtRec; ++k) st_agent(&a.mb_rec[rec_idx(e, k, gtid)], 0u);
            }
            set_sx(e, sx + 1u + kSxExtOne);
            st_agent(&a.mb_rec[rec_idx(e, rk, gtid)], 1u + ((sidx << 6) | (unsigned int)sq));
            ++helpers_out;
            ext_mask |= 1u << (e - 3);
            if (!mv) split_mask &= ~(1u << (e - 3));
            const uint64_t f = flip(bp, bo, sq);
            HelpSlot* hs = &a.slots[sidx];
            st_agent(&hs->P, bo ^ f);
            st_agent(&hs->O, bp ^ f ^ bit);
            st_agent(&hs->alpha, am);
            st_agent(&hs->beta, fbeta);
            st_agent(&hs->ctype, child_type((int)(w >> 30), 1));
            st_agent(&hs->owner_E, e);
            st_agent(&hs->owner, gtid);
            st_agent(&hs->sq, sq);
            __hip_atomic_store(&hs->state, slot_full(gtid, e, sq), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            ++c_exports;
          }
        }
      }
    }

    // ---- 5. one search step ----
    // Every searching lane runs one flip and one mobility per trip.  A child
    // whose side has no move is settled next trip with the other side's
    // mobility in the same slot (pass or game over), so no trip pays for a
    // second mobility.
    bool commit = false;
    int v = 0;
    if (mode == SEARCH && !pop_pending) {
      ++c_iters;
      int sq;
      uint64_t f, mx, my;
      bool leaf = false, tleaf = false;
      if (pchk) {
        sq = pchk_sq;
        f = pchk_f;
      } else {
        if (scoring) {
          sq = ctz64(to_score);
        } else if (order) {
          sq = (int)(order & 127) - 1;
          order >>= 7;
        } else {
          sq = pick_static<ORDER>(moves, ~(P | O));
        }
        f = flip(P, O, sq);
        if constexpr (THINK) {
          tleaf = !scoring && E - 1 == ebase + 2;               // the child is at the horizon
          // exact below only when the horizon lies beyond the end: a
          // position reached by the depth-th move is scored by the evaluator
          // even when the board is full
          leaf = !scoring && E == 3 && ebase + 2 < 0 && !tleaf;
        } else {
          leaf = !scoring && E == 3;
        }
      }
      const uint64_t bit = 1ULL << sq;
      const uint64_t cP = O ^ f, cO = P ^ f ^ bit;
      mx = pchk ? cO : cP;
      my = pchk ? cP : cO;
      uint64_t cm = 0;
      if (THINK && tleaf) {
        if constexpr (ORDER == 2) {
          int ev = a.ev_bias[E - 1];
          for_each_feature(cP, cO, t3, [&](int, uint32_t fi) { ev += a.ev_w[fi]; });
          const float r = roundf((float)ev / a.ev_scale);
          v = -(int)fminf(fmaxf(r, -64.f), 64.f);
        }
      } else if (leaf) {
        // the child has two empty squares: its exact value in this trip
        int nl = 0;
        v = -solve2(cP, cO, lastc, nl);
        c_nodes += nl;
      } else {
        cm = mobility(mx, my);
      }

      bool push = false, passed = false;
      if (pchk) {
        pchk = false;
        if (!cm) {
          v = -final_score(cP, cO);  // neither side can move: game over
          commit = true;
        } else {
          push = true;
          passed = true;
        }
      } else if (scoring) {
        ++c_steps;
        const int k = move_key_of(cP, cO, cm, bit, E);
        const uint64_t valid = (1ULL << (7 * nkeys)) - 1;
        const uint64_t le = ((((uint64_t)k * kLo7) | kHi7) - keys) & kHi7 & valid;
        const int c = popc64(le);  // entries with key <= k: insert after them (stable)
        const uint64_t low = (1ULL << (7 * c)) - 1;
        keys = (((keys & ~low) << 7) | (keys & low) | ((uint64_t)k << (7 * c))) & kFull63;
        order = (((order & ~low) << 7) | (order & low) | ((uint64_t)(sq + 1) << (7 * c))) & kFull63;
        nkeys = min(nkeys + 1, 9);
        to_score ^= bit;
        scoring = to_score != 0;
      } else {
        moves ^= bit;
        ++c_nodes;
        if (leaf || tleaf) {
          commit = true;
        } else if (!cm) {
          pchk = true;
          pchk_sq = sq;
          pchk_f = f;
        } else {
          push = true;
        }
      }
      int ca = passed ? alpha : -beta, cb = passed ? beta : -alpha;  // child window
      bool nullw = false;
      if constexpr (!THINK) {
        // Principal variation search: after the first child, a child is
        // first searched with the null window (alpha, alpha + 2) (every
        // value is even); one that fails high inside (alpha, beta) is
        // searched again with its bound as alpha.
        if (push) {
          if (!passed && a.pvs_min && E >= a.pvs_min && ndone > 0 && cb - ca > 2 && !full_next) {
            ca = cb - 2;
            nullw = true;
          }
          full_next = false;
        }
      }
      bool research = false;
      if (push && !passed && a.tt && E - 1 >= a.tt_min) {
        // Probe the child: an entry may settle it or narrow its window.
        ++c_probes;
        const uint64_t h = tt_key(cP, cO);
        const ulonglong2 b = a.tt[h & a.tt_mask];
        if ((b.x ^ b.y) == h && (uint32_t)((b.y >> 24) & 255) == tt_gen) {
          ++c_hits;
          const int lo = (int)(b.y & 255) - 64, up = (int)((b.y >> 8) & 255) - 64;
          int tv = 0;
          bool cut = true;
          if (lo >= cb) {
            tv = lo;
          } else if (up <= ca) {
            tv = up;
          } else if (lo == up) {
            tv = lo;
          } else {
            cut = false;
            ca = max(ca, lo);
            cb = min(cb, up);
          }
          if (cut) {
            ++c_cuts;
            push = false;
            v = -tv;
            // an upper bound below a null window is only a bound on our side
            if (nullw && tv == up && lo != up && v < beta) {
              research = true;
            } else {
              commit = true;
            }
          }
        }
      }
      if (research) {
        // search this child again with the full window, next
        moves |= bit;
        order = ((order << 7) | (uint64_t)(sq + 1)) & kFull63;
        alpha = max(alpha, v - 2);
        full_next = true;
      }
      if (push) {
        if (nullw) nw_mask |= 1u << (E - 3);
        push_frame(E, f | bit, moves, order, frame_word<THINK>(alpha, beta, best, sq, passed, ndone, ntype));
        // Young-brothers-wait: siblings are offered after the first child
        // (after cut_after children at expected CUT nodes).  When nearly
        // every wavefront is idle, speculation costs little and the serial
        // chain of the last boards is what remains: offer after one child
        // at CUT nodes and at once at ALL nodes.
        const bool tail = dem_e <= a.ext_lo;
        const int thr = ntype == kCut ? (tail ? 1 : a.cut_after) : (tail && ntype == kAll ? 0 : 1);
        if (E >= a.split_min && moves && ndone >= thr) split_mask |= 1u << (E - 3);
        const int ct = child_type(ntype, ndone);
        if (E >= a.split_min)
          set_sx(E, a.tt && E >= a.tt_min ? (cur_sx & ~(63u << 23)) | ((uint32_t)bestmv_of() << 23) : cur_sx);
        if (passed) {
          P = cO;
          O = cP;
        } else {
          P = cP;
          O = cO;
        }
        alpha = ca;
        beta = cb;
        --E;
        enter_node(cm, ct);
      }
      if (commit) {
        if (E == Eroot && owner < 0) mkey = max(mkey, move_key(v, v > alpha, sq));
        if (v > best) a0mv = (a0mv & ~63) | sq;
        best = max(best, v);
        alpha = max(alpha, best);
        ndone = min(ndone + 1, 3);
      }
    } else if (mode == SEARCH) {
      // continue returning from finished nodes (pops are capped per trip)
      pop_pending = false;
      commit = true;
    } else if (mode == WAIT) {
      // Waiting for helpers of the current node; stop them once it is cut.
      const uint32_t sx = cur_sx;
      if (sx_pending(sx) == 0) {
        mode = SEARCH;
        commit = true;
        if (dead_level == E) dead_level = 64;
      } else if (dead_level != E && max(alpha, sx_best(sx)) >= beta) {
        dead_level = E;
        dead_new = true;
      }
    } else if (mode == ABORT) {
      if (helpers_out == 0) {
        report_aborted = true;
        mode = REPORT;
      }
    }

    // ---- 5b. leaf batch: the queued children of nodes with three empties ----
    // Each waiting lane has one job per legal move (at most three).  Once at
    // least leaf_batch jobs are queued, or the oldest has waited kLeafAge
    // trips, the whole wavefront evaluates up to 64 of them at once (whole
    // nodes only): worker w learns its requesting lane from an LDS map,
    // fetches the board by lane permute, and the requester reads its
    // children's values back the same way.
    if constexpr (!THINK) {
      const uint64_t lw = __ballot(mode == LEAFWAIT);
      if (lw) {
        ++leaf_age;
        const int cnt = mode == LEAFWAIT ? popc64(moves) : 0;  // 1..3
        const uint64_t lt = __lanemask_lt();
        const uint64_t c1 = __ballot(cnt & 1), c2 = __ballot(cnt & 2);
        const int jobs = popc64(c1) + 2 * popc64(c2);
        if (jobs >= a.leaf_batch || leaf_age > kLeafAge) {
          leaf_age = 0;
          const int off = popc64(c1 & lt) + 2 * popc64(c2 & lt);  // first job slot of this lane
          const bool sel = cnt > 0 && off + cnt <= 64;
          const uint64_t selm = __ballot(sel);
          const int nsel = (int)(64 - __builtin_clzll(selm));     // requesters are a lane prefix
          const int njobs = __builtin_amdgcn_readlane(off + cnt, nsel - 1);
          if (sel)
            for (int k = 0; k < cnt; ++k) lmap[off + k] = (uint8_t)lane;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          const int r = lane < njobs ? (int)lmap[lane] : lane;
          const uint64_t jP = shfl64(P, r), jO = shfl64(O, r), jm = shfl64(moves, r);
          const int k = lane - __shfl(off, r);
          const bool jswap = __shfl((int)leaf_swap, r) != 0;
          int jv = 0;
          if (lane < njobs) {
            uint64_t m = jm;
            for (int i = 0; i < k; ++i) m &= m - 1;
            int nl = 0;
            jv = leaf_job(jP, jO, ctz64(m), jswap, lastc, nl);
            c_nodes += (unsigned int)nl;
          }
          // results back to the requesters
          int vk[3];
#pragma unroll
          for (int i = 0; i < 3; ++i) vk[i] = __shfl(jv, min(off + i, 63));
          if (sel) {
            uint64_t m = moves, pend = 0;
            int nb = -64, got = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              if (i < cnt) {
                const int sq = ctz64(m);
                m &= m - 1;
                if (vk[i] == kNeedPass) {
                  pend |= 1ULL << sq;
                } else {
                  if (E == Eroot && owner < 0) mkey = max(mkey, move_key(vk[i], true, sq));
                  nb = max(nb, vk[i]);
                  ++got;
                }
              }
            }
            best = max(best, nb);
            alpha = max(alpha, best);
            ndone = min(ndone + got, 3);
            if (pend && alpha < beta) {
              // children whose mover passes: queued again for the next batch
              moves = pend;
              leaf_swap = true;
            } else {
              moves = 0;
              mode = SEARCH;
              commit = true;
            }
          }
        }
      }
    }

    // ---- 5c. score batch: the children of nodes that order their moves ----
    // Each waiting lane has one job per legal move (its to_score set).  A
    // worker plays its move, counts the replies and computes the key; the
    // node's jobs are ranked by (key, square) through lane permutes, and the
    // node's packed order list takes the first nine: the list the one-child-
    // per-trip scoring builds by stable insertion.
    if constexpr (!THINK) {
      const uint64_t swm = __ballot(mode == SCOREWAIT);
      if (swm) {
        ++score_age;
        const int cnt = mode == SCOREWAIT ? popc64(to_score) : 0;  // 2..32
        const uint64_t lt = __lanemask_lt();
        int off = 0, jobs = 0;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
          const uint64_t bb = __ballot((cnt >> b) & 1);
          off += popc64(bb & lt) << b;
          jobs += popc64(bb) << b;
        }
        if (jobs >= a.score_batch || score_age > kScoreAge) {
          score_age = 0;
          const bool sel = cnt > 0 && off + cnt <= 64;
          const uint64_t selm = __ballot(sel);
          const int njobs = __builtin_amdgcn_readlane(off + cnt, 63 - __builtin_clzll(selm));
          if (sel)
            for (int k = 0; k < cnt; ++k) lmap[off + k] = (uint8_t)lane;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          const bool worker = lane < njobs;
          const int r = worker ? (int)lmap[lane] : lane;
          const uint64_t jP = shfl64(P, r), jO = shfl64(O, r), jm = shfl64(to_score, r);
          const int jE = __shfl(E, r), joff = __shfl(off, r), jcnt = __shfl(cnt, r);
          const int jbeta = __shfl(beta, r);
          const int k = lane - joff;
          int key = 0, jsq = 0;
          if (worker) {
            uint64_t m = jm;
            for (int i = 0; i < k; ++i) m &= m - 1;
            jsq = ctz64(m);
            const uint64_t bit = 1ULL << jsq;
            const uint64_t f = flip(jP, jO, jsq);
            const uint64_t cP = jO ^ f, cO = jP ^ f ^ bit;
            key = move_key_of(cP, cO, mobility(cP, cO), bit, jE);
            ++c_steps;
            if (a.tt && jE - 1 >= a.tt_min) {
              // A child whose table entry already proves the node's cutoff
              // (an upper bound <= -beta) is searched first: its probe then
              // settles it and the node is cut after one step.
              const uint64_t h = tt_key(cP, cO);
              const ulonglong2 e = a.tt[h & a.tt_mask];
              if ((e.x ^ e.y) == h && (uint32_t)((e.y >> 24) & 255) == tt_gen &&
                  (int)((e.y >> 8) & 255) - 64 <= -jbeta)
                key = 0;
            }
          }
          // Rank of each job inside its node's segment of lanes [joff,
          // joff + jcnt): jobs with a smaller key, then equal keys on lower
          // lanes (lower squares).  Keys are 6-bit; six ballots from the top
          // bit down split the lanes into "key below mine" and "key equal".
          uint64_t lt = 0, eq = ~0ull;
#pragma unroll
          for (int b = 5; b >= 0; --b) {
            const uint64_t B = __ballot((key >> b) & 1);
            const bool one = (key >> b) & 1;
            lt |= one ? (eq & ~B) : 0ull;
            eq &= one ? B : ~B;
          }
          const uint64_t seg = (jcnt >= 64 ? ~0ull : ((1ull << jcnt) - 1)) << (joff & 63);
          const int rank = popc64(lt & seg) + popc64(eq & seg & __lanemask_lt());
          // Each node's first nine squares in rank order, through the same
          // LDS map (every lane has read its entry above): job slot joff +
          // rank gets square + 1, the node's lane reads its slots back.
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          if (worker && rank < 9) lmap[joff + rank] = (uint8_t)(jsq + 1);
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          uint64_t ord = 0;
          if (sel) {
#pragma unroll
            for (int i = 0; i < 9; ++i)
              if (i < cnt) ord |= (uint64_t)lmap[min(off + i, 63)] << (7 * i);
          }
          if (sel) {
            order = ord;
            to_score = 0;
            scoring = false;
            mode = SEARCH;
          }
        }
      }
    }

    // ---- 6. return from finished nodes, at most max_pops levels per trip ----
    if (commit) {
      for (int pops = 0;; ++pops) {
        if (E >= a.split_min) {
          const uint32_t sx = cur_sx;
          const int pd = sx_pending(sx);
          best = max(best, sx_best(sx));
          alpha = max(alpha, best);
          if (!(alpha >= beta || !moves)) break;
          if (pd != 0) {
            mode = WAIT;
            if (alpha >= beta) {
              dead_level = E;
              dead_new = true;
            }
            break;
          }
        } else if (!(alpha >= beta || !moves)) {
          break;
        }
        if (a.tt && E >= a.tt_min) {
          // Store the finished node: fail low -> upper bound, fail high ->
          // lower bound, inside the window -> exact.
          ++c_stores;
          const int lo = best > alpha0_of() ? best : -64, up = best < beta ? best : 64;
          const uint64_t d = (uint64_t)(lo + 64) | ((uint64_t)(up + 64) << 8) | ((uint64_t)bestmv_of() << 16) |
                             ((uint64_t)tt_gen << 24) | ((uint64_t)E << 32);
          const uint64_t h = tt_key(P, O);
          a.tt[h & a.tt_mask] = make_ulonglong2(h ^ d, d);
        }
        if (E == Eroot) {
          if (asp && owner < 0) {
            // Aspiration root: a value at or outside the window is only a
            // bound (fail-soft); the true value lies on its side, so one
            // search with that side open is exact.
            asp = false;
            const int a0 = alpha0_of();
            const bool low = best <= a0 && a0 > -64, high = best >= beta && beta < 64;
            if (low || high) {
              alpha = low ? -64 : best - 2;
              beta = low ? best + 2 : 64;
              mkey = -1;
              mode = SEARCH;
              enter_node(mobility(P, O), kPV);
              break;
            }
          }
          const int r = negate ? -best : best;
          if (owner < 0) {
            a.scores[idx] = (int8_t)r;
            if (a.moves)
              a.moves[idx] = (uint8_t)(negate ? kMovePass : mkey >= 0 ? (mkey & 63) : kMoveNone);
            mode = IDLE;
          } else {
            mkey = move_key(r, r > (mkey >> 6) - 64, mkey & 63);
            report_aborted = false;
            mode = REPORT;
          }
          break;
        }
        if (pops == a.max_pops) {
          pop_pending = true;
          break;
        }
        ++E;
        uint64_t fb;
        uint32_t w;
        if (in_lds(E)) {
          const int k = lds_k(E);
          fb = l_fb[k];
          moves = l_mv[k];
          order = l_od[k];
          w = l_wd[k];
        } else {
          const size_t k = hbm_k(E);
          fb = hbm_ld(&a.g_fb[k]);
          moves = hbm_ld(&a.g_mv[k]);
          order = hbm_ld(&a.g_od[k]);
          w = hbm_ld(&a.g_wd[k]);
        }
        if constexpr (!THINK) split_mask &= ~(1u << (E - 3));
        const int palpha = fw_alpha<THINK>(w);
        const int pbeta = fw_beta<THINK>(w);
        const int pbest = fw_best<THINK>(w);
        undo_move(P, O, fb, w);
        const int vv = ((w >> 27) & 1) ? best : -best;
        const int msq = (int)((w >> 21) & 63);
        bool again = false;
        if constexpr (!THINK) {
          if ((nw_mask >> (E - 3)) & 1) {
            nw_mask &= ~(1u << (E - 3));
            again = vv > palpha && vv < pbeta;  // a lower bound inside the window
          }
        }
        beta = pbeta;
        if (again) {
          best = pbest;
          alpha = vv - 2;
          moves |= 1ULL << msq;
          order = ((order << 7) | (uint64_t)(msq + 1)) & kFull63;
          full_next = true;
        } else {
          if (E == Eroot && owner < 0) mkey = max(mkey, move_key(vv, vv > palpha, msq));
          best = max(pbest, vv);
          alpha = max(palpha, best);
        }
        if (E >= a.split_min) {
          const uint32_t psx = get_sx(E);
          cur_sx = psx;
          a0mv = (((sx_alpha0(psx) >> 1) & 127) << 6) | (!again && vv > pbest ? msq : (int)((psx >> 23) & 63));
        }
        if constexpr (THINK) {
          ndone = 1;
          ntype = kAll;
        } else {
          ndone = min((int)((w >> 28) & 3) + 1, 3);
          ntype = (int)(w >> 30);
        }
        scoring = false;
      }
    }
  }

  flush_counters();
  if (lane == 0) {
    atomicAdd(&a.ctr->lane_slots, 64ull * trip);  // every lane takes part in every trip
    atomicMax(&a.ctr->t_end, (unsigned long long)__builtin_amdgcn_s_memrealtime());
  }
}

__global__ void reset_counters_kernel(SolveCounters* c) {
  c->t_start = __builtin_amdgcn_s_memrealtime();
  c->t_end = 0;
  c->active_waves = 0;
  c->idle_waves = 0;
  c->exports = 0;
  c->tt_gen = c->tt_gen % 255u + 1u;
  c->next = 0;
  c->error = 0;
  c->positions = 0;
  c->nodes = 0;
  c->score_steps = 0;
  c->iterations = 0;
  c->steals = 0;
  c->lane_slots = 0;
  c->tt_probes = 0;
  c->tt_hits = 0;
  c->tt_cuts = 0;
  c->tt_stores = 0;
}

// ---------------------------------------------------------------------------
// Host side: pick the instantiation for the ordering, size the grid to what
// stays resident (plan.hpp), lay out the HBM frames, launch.

using KernelFn = void (*)(SolveArgs);

static KernelFn pick_kernel(int ordering, bool think) {
  if (think) return solve_kernel<2, true>;
  switch (ordering) {
    case GO2_ORDER_STATIC: return solve_kernel<0>;
    case GO2_ORDER_MOBILITY: return solve_kernel<1>;
    default: return solve_kernel<2>;
  }
}

// Installed int16 evaluator per device (GO2_ORDER_LEARNED).
struct EvalTables {
  int16_t* w = nullptr;
  int16_t* bias = nullptr;
  float scale = 0.f;
};
static EvalTables g_eval[64];
// Guards the per-device evaluator tables while they are set up.
static std::mutex g_eval_mutex;

}  // namespace go2

extern "C" int go2_set_evaluator(const int16_t* w, const int16_t* bias, float scale) {
  using namespace go2;
  if (!w || !bias || !(scale > 0.f)) return fail(GO2_E_ARG, "null weights or non-positive scale");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(GO2_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_eval_mutex);
  EvalTables& t = g_eval[dev];
  if (!t.w && (e = hipMalloc(&t.w, kPatWeights * sizeof(int16_t))) != hipSuccess)
    return hip_fail(e, "hipMalloc(evaluator)");
  if (!t.bias && (e = hipMalloc(&t.bias, kPatBias * sizeof(int16_t))) != hipSuccess)
    return hip_fail(e, "hipMalloc(evaluator bias)");
  if ((e = hipMemcpy(t.w, w, kPatWeights * sizeof(int16_t), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "hipMemcpy(evaluator)");
  if ((e = hipMemcpy(t.bias, bias, kPatBias * sizeof(int16_t), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "hipMemcpy(evaluator bias)");
  t.scale = scale;
  return GO2_OK;
}

namespace go2 {

Workspace::~Workspace() {
  if (fixed) return;  // carved from memory the workspace does not own
  if (ctr) (void)hipFree(ctr);
  if (frames) (void)hipFree(frames);
  if (tt) (void)hipFree(tt);
  if (help) (void)hipFree(help);
}

static_assert(sizeof(SolveCounters) <= kCounterBytes, "counters outgrew their scratch slot");
static_assert(sizeof(HelpSlot) == kHelpSlotBytes, "help slot size");

int Workspace::carve(size_t help_b, size_t frame_b) {
  if (!fixed) return fail(GO2_E_STATE, "workspace is not in fixed mode");
  if (kCounterBytes + help_b + frame_b > fixed_cap) return fail(GO2_E_ARG, "solve does not fit the scratch");
  ctr = reinterpret_cast<SolveCounters*>(fixed);
  help = help_b ? fixed + kCounterBytes : nullptr;
  help_bytes = help_b;
  frames = fixed + kCounterBytes + help_b;
  frames_bytes = frame_b;
  return GO2_OK;
}

// ---- the no-allocation entry point's memory ------------------------------------
// One static device array: each device the library runs on holds its own
// copy until the process ends.  Host code never passes the array's name to
// a kernel (on the host it is only a stand-in for the symbol); it looks the
// device address up once per device and passes that.
static __device__ __align__(256) uint8_t scratch[GO2_NOALLOC_SCRATCH_BYTES];
static uint8_t* g_scratch[64];
static std::mutex g_scratch_mutex;

int noalloc_scratch(uint8_t** base, size_t* bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return fail(GO2_E_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess)
      return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(GO2_E_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch[dev] = reinterpret_cast<uint8_t*>(p);
  }
  *base = g_scratch[dev];
  *bytes = GO2_NOALLOC_SCRATCH_BYTES;
  return GO2_OK;
}

__global__ void clear_tt_kernel(ulonglong2* tt, unsigned long long n) {
  for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    tt[i] = make_ulonglong2(0, 0);
}

int Workspace::ensure_tt(int log2_buckets, hipStream_t stream) {
  hipError_t e;
  if (log2_buckets != tt_log2) {
    if (tt) (void)hipFree(tt);
    tt = nullptr;
    tt_log2 = 0;
    if ((e = hipMalloc(&tt, (size_t)16 << log2_buckets)) != hipSuccess)
      return hip_fail(e, "hipMalloc(transposition table)");
    tt_log2 = log2_buckets;
    tt_gen = 0;
  }
  // The reset kernel of the coming launch moves the device generation to
  // tt_gen % 255 + 1; the host mirrors it and clears the table (a fill
  // kernel, outside any captured graph) before generation 1 comes round.
  tt_gen = tt_gen % 255u + 1u;
  if (tt_gen == 1) {
    const unsigned long long n = 1ULL << tt_log2;
    hipLaunchKernelGGL(clear_tt_kernel, dim3(4096), dim3(256), 0, stream, (ulonglong2*)tt, n);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "clear table launch");
  }
  return GO2_OK;
}

int Workspace::ensure_help(size_t bytes, hipStream_t stream) {
  if (bytes > help_bytes) {
    if (help) (void)hipFree(help);
    help = nullptr;
    help_bytes = 0;
    hipError_t e = hipMalloc(&help, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(help slots)");
    help_bytes = bytes;
  }
  (void)stream;
  return GO2_OK;
}

__global__ void clear_slots_kernel(HelpSlot* s, unsigned int n) {
  for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    s[i].state = kSlotEmpty;
}

int Workspace::ensure(size_t frame_bytes) {
  hipError_t e;
  if (!ctr) {
    if ((e = hipMalloc(&ctr, sizeof(SolveCounters))) != hipSuccess) return hip_fail(e, "hipMalloc(counters)");
    if ((e = hipMemset(ctr, 0, sizeof(SolveCounters))) != hipSuccess) return hip_fail(e, "hipMemset(counters)");
  }
  if (frame_bytes > frames_bytes) {
    if (frames) (void)hipFree(frames);
    frames = nullptr;
    frames_bytes = 0;
    if ((e = hipMalloc(&frames, frame_bytes)) != hipSuccess) return hip_fail(e, "hipMalloc(stack frames)");
    frames_bytes = frame_bytes;
  }
  return GO2_OK;
}

int launch_solve(const go2_board* d_boards, size_t n, int max_depth, int8_t* d_scores, uint8_t* d_moves,
                 const go2_config& cfg, Workspace& ws, hipStream_t stream, hipEvent_t ev0,
                 hipEvent_t ev1, GraphCache* graph) {
  if (n > 0xFFFFFF00ULL) return fail(GO2_E_ARG, "too many boards in one solve (max 2^32-256)");
  const StackPlan sp = plan_stack(max_depth, cfg);
  const KernelFn fn = pick_kernel(cfg.ordering, sp.think);
  const int bs = cfg.block_size;
  const size_t lds = sp.lds_bytes;
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
  if (per_cu < 1) return fail(GO2_E_ARG, "block size / LDS levels do not fit one CU");
  GridPlan gp;
  int rc;
  if (ws.fixed) {
    // Memory the caller provides: no table, and a grid whose frames and help
    // fit (fewer workgroups per CU, then no cross-wavefront help).
    if (cfg.tt_log2_buckets > 0 && !sp.think)
      return fail(GO2_E_ARG, "this entry point has no transposition table (config.tt_log2_buckets must be 0)");
    if (!plan_grid_within(sp, n, cfg, cus, per_cu, ws.fixed_cap, &gp))
      return fail(GO2_E_ARG, "DEPTH " + std::to_string(max_depth) + " needs a search stack larger than the " +
                                 std::to_string(ws.fixed_cap >> 20) + " MiB this entry point holds");
    if ((rc = ws.carve(help_bytes(sp, gp), frame_bytes(gp))) != GO2_OK) return rc;
  } else {
    gp = plan_grid(sp, n, cfg, cus, per_cu);
    if ((rc = ws.ensure(frame_bytes(gp))) != GO2_OK) return rc;
    if (gp.help && (rc = ws.ensure_help(help_bytes(sp, gp), stream)) != GO2_OK) return rc;
  }
  const size_t grid = gp.grid, waves = gp.waves, total_lanes = gp.total_lanes, g_levels = gp.g_levels;
  const int lds_levels = sp.lds_levels;
  // frames: fb, mv, od (u64) | wd, sx (u32), each [level][total_lanes]
  uint8_t* base = reinterpret_cast<uint8_t*>(ws.frames);
  SolveArgs args{};
  args.boards = d_boards;
  args.scores = d_scores;
  args.moves = d_moves;
  args.n = (unsigned int)n;
  args.depth = sp.depth;
  args.think_depth = cfg.think_depth;
  args.kff = cfg.mobility_min_empties;
  args.split_min = gp.split_min;
  args.cut_after = cfg.split_cut_after > 0 ? cfg.split_cut_after : 3;
  args.lds_levels = lds_levels;
  args.K = gp.K;
  args.total_lanes = (unsigned int)total_lanes;
  args.limit_ticks = 0;
  args.max_pops = cfg.max_pops_per_trip > 0 ? cfg.max_pops_per_trip : 2;
  args.leaf_batch = sp.think || cfg.leaf_batch_min > 64 ? 0 : (cfg.leaf_batch_min > 0 ? cfg.leaf_batch_min : 32);
  args.score_batch = sp.think || cfg.score_batch_min > 64 ? 0 : (cfg.score_batch_min > 0 ? cfg.score_batch_min : 64);
  args.asp_width = sp.think || cfg.aspiration_width > 64 ? 0 : 2 * ((cfg.aspiration_width + 1) / 2);
  args.pvs_min = sp.think || cfg.pvs_min_empties <= 0 ? 0 : max(cfg.pvs_min_empties, 5);
  args.asp_min = 10;
  if (gp.help) {
    const unsigned int nslots = kHelpSlots;
    const size_t mb = mailbox_bytes(sp, gp);
    uint8_t* hb = reinterpret_cast<uint8_t*>(ws.help);
    args.slots = reinterpret_cast<HelpSlot*>(hb);
    args.nslots = nslots;
    args.ext_min = gp.ext_min;
    args.ext_lo = gp.ext_lo;
    size_t o = (size_t)nslots * sizeof(HelpSlot);
    args.mb_best = reinterpret_cast<int*>(hb + o);
    o += mb;
    args.mb_done = reinterpret_cast<unsigned int*>(hb + o);
    o += mb;
    args.mb_cut = reinterpret_cast<unsigned int*>(hb + o);
    o += mb;
    args.mb_rec = reinterpret_cast<unsigned int*>(hb + o);
  } else {
    args.ext_min = args.ext_lo = 3;
  }
  if (cfg.tt_log2_buckets > 0 && !sp.think) {
    if ((rc = ws.ensure_tt(cfg.tt_log2_buckets, stream)) != GO2_OK) return rc;
    args.tt = reinterpret_cast<ulonglong2*>(ws.tt);
    args.tt_mask = (1ULL << cfg.tt_log2_buckets) - 1;
    args.tt_gen = &ws.ctr->tt_gen;
    args.tt_min = max(max(cfg.tt_min_empties > 0 ? cfg.tt_min_empties : 12, gp.split_min), 4);
  }
  if (cfg.ordering == GO2_ORDER_LEARNED || sp.think) {
    const EvalTables& t = g_eval[dev];
    if (!t.w)
      return fail(GO2_E_STATE, sp.think ? "think mode needs go2_set_evaluator() on this device"
                                        : "GO2_ORDER_LEARNED needs go2_set_evaluator() on this device");
    args.ev_scale = t.scale;
    args.ev_w = t.w;
    args.ev_bias = t.bias;
    args.learned_min = cfg.learned_min_empties;
    // key units: by default one key step per four discs of the evaluator
    // (`scale` units per disc) and one per reply, the best of a sweep over
    // 16-empty positions with the 14-empty-trained table.
    args.key_eval_mul = cfg.learned_key_eval > 0 ? cfg.learned_key_eval : (int)(1024.f / t.scale);
    args.key_mob_mul = cfg.learned_key_mobility > 0 ? cfg.learned_key_mobility : 1;
  }
  args.ctr = ws.ctr;
  size_t off = 0;
  auto take = [&](size_t b) {
    uint8_t* p = base + off;
    off += (b + 255) & ~(size_t)255;
    return p;
  };
  args.g_fb = reinterpret_cast<uint64_t*>(take(total_lanes * g_levels * 8));
  args.g_mv = reinterpret_cast<uint64_t*>(take(total_lanes * g_levels * 8));
  args.g_od = reinterpret_cast<uint64_t*>(take(total_lanes * g_levels * 8));
  args.g_wd = reinterpret_cast<uint32_t*>(take(total_lanes * g_levels * 4));
  args.g_sx = reinterpret_cast<uint32_t*>(take(total_lanes * g_levels * 4));
  if (off != frame_bytes(gp)) return fail(GO2_E_STATE, "frame layout does not match its size");
  // s_memrealtime runs at 100 MHz; each wavefront adds its own start time.
  if (cfg.time_limit_ms > 0) args.limit_ticks = (unsigned long long)cfg.time_limit_ms * 100000ULL;
  ws.grid_waves = waves;

  auto enqueue = [&]() -> int {
    reset_counters_kernel<<<1, 1, 0, stream>>>(ws.ctr);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(le, "reset_counters launch");
    if (args.slots) {
      // a launch abandoned at its time limit may leave slots behind
      clear_slots_kernel<<<16, 256, 0, stream>>>(args.slots, args.nslots);
      if ((le = hipGetLastError()) != hipSuccess) return hip_fail(le, "clear slots launch");
    }
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(bs), lds, stream, args);
    if ((le = hipGetLastError()) != hipSuccess) return hip_fail(le, "solve kernel launch");
    return GO2_OK;
  };
  if (!graph) {
    if (ev0 && (e = hipEventRecord(ev0, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    if ((rc = enqueue()) != GO2_OK) return rc;
    if (ev1 && (e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
    return GO2_OK;
  }
  // Replay a captured graph of the same launch; capture it when the
  // kernel, shape or any argument changed.
  unsigned char key[512] = {0};
  size_t klen = 0;
  auto put = [&](const void* p, size_t b) {
    memcpy(key + klen, p, b);
    klen += b;
  };
  const void* fnp = (const void*)fn;
  put(&fnp, sizeof(fnp));
  put(&grid, sizeof(grid));
  put(&bs, sizeof(bs));
  put(&lds, sizeof(lds));
  put(&args, sizeof(args));
  if (!graph->exec || graph->key_len != klen || memcmp(graph->key, key, klen) != 0) {
    graph->reset();
    if ((e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return hip_fail(e, "hipStreamBeginCapture");
    rc = enqueue();
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(stream, &g);
    if (rc != GO2_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    graph->graph = g;
    if ((e = hipGraphInstantiate(&graph->exec, g, nullptr, nullptr, 0)) != hipSuccess)
      return hip_fail(e, "hipGraphInstantiate");
    memcpy(graph->key, key, klen);
    graph->key_len = klen;
  }
  if (ev0 && (e = hipEventRecord(ev0, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  if ((e = hipGraphLaunch(graph->exec, stream)) != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  if (ev1 && (e = hipEventRecord(ev1, stream)) != hipSuccess) return hip_fail(e, "hipEventRecord");
  return GO2_OK;
}

void GraphCache::reset() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  key_len = 0;
}

GraphCache::~GraphCache() { reset(); }

int collect_stats(Workspace& ws, hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1,
                  go2_stats* stats) {
  SolveCounters h{};
  hipError_t e = hipMemcpyAsync(&h, ws.ctr, sizeof(h), hipMemcpyDeviceToHost, stream);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(counters)");
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  if (h.error == kTimedOut) return fail(GO2_E_TIMEOUT, "solve exceeded its time limit");
  if (h.error) {
    return fail(GO2_E_BOARD, "board " + std::to_string(h.error - 1) +
                                 " is invalid (overlapping discs, or more empties than DEPTH or than " +
                                 std::to_string(GO2_MAX_EMPTIES) + ")");
  }
  if (stats) {
    *stats = go2_stats{};
    stats->positions = h.positions;
    stats->nodes = h.nodes;
    stats->score_steps = h.score_steps;
    stats->iterations = h.iterations;
    stats->steals = h.steals;
    stats->lane_slots = h.lane_slots;
    stats->tt_probes = h.tt_probes;
    stats->tt_hits = h.tt_hits;
    stats->tt_cuts = h.tt_cuts;
    stats->tt_stores = h.tt_stores;
    stats->exports = h.exports;
    stats->grid_waves = ws.grid_waves;
    float ms = 0.f;
    if (ev0 && ev1) {
      if (hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess) stats->kernel_ms = ms;
    } else if (h.t_end > h.t_start) {
      stats->kernel_ms = (double)(h.t_end - h.t_start) * 1e-5;  // 100 MHz ticks
    }
  }
  return GO2_OK;
}

}  // namespace go2
