// Synthetic positions on the GPU: one thread per position, a random legal
// playout from the standard start to a target number of empties (passing
// when forced; a game that ends first is replayed with the next attempt).
// The reference ships no generator; the stream definition is this
// project's and is shared bit for bit with the CPU oracle's generator.
#include "bitboard.hpp"
#include "internal.hpp"
#include "rng.hpp"

namespace go2 {

__global__ void generate_kernel(uint64_t seed, uint32_t config, size_t n, int emin, int emax,
                                go2_board* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int target = empties_for(seed, config, i, emin, emax);
  go2_board res{0, 0};
  for (uint32_t attempt = 0; attempt < 1000; ++attempt) {
    uint64_t s = position_key(seed, config, i, attempt);
    uint64_t P = 0x0000000810000000ULL, O = 0x0000001008000000ULL;
    bool ok = true;
    while (64 - popc64(P | O) > target) {
      uint64_t m = mobility(P, O);
      if (!m) {
        if (!mobility(O, P)) {
          ok = false;
          break;
        }
        const uint64_t t = P;
        P = O;
        O = t;
        continue;
      }
      const uint32_t k = (uint32_t)popc64(m);
      const uint64_t r = splitmix64(s);
      const uint32_t pick = (uint32_t)(((r >> 32) * (uint64_t)k) >> 32);
      for (uint32_t j = 0; j < pick; ++j) m &= m - 1;
      const uint64_t bit = m & (0 - m);
      const uint64_t f = flip(P, O, ctz64(bit));
      const uint64_t nP = O ^ f;
      O = (P ^ f) | bit;
      P = nP;
    }
    if (ok) {
      res.player = P;
      res.opponent = O;
      break;
    }
  }
  out[i] = res;
}

int launch_generate(uint64_t seed, uint32_t config_id, size_t n, int emin, int emax,
                    go2_board* d_out, hipStream_t stream) {
  if (n == 0) return GO2_OK;
  const int bs = 256;
  const size_t grid = (n + bs - 1) / bs;
  hipLaunchKernelGGL(generate_kernel, dim3((unsigned)grid), dim3(bs), 0, stream, seed, config_id, n,
                     emin, emax, d_out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "generate kernel launch");
  return GO2_OK;
}

}  // namespace go2
