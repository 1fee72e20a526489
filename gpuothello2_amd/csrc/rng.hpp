// Seeded streams for the synthetic position generator (SplitMix64 finaliser
// and step).  Shared by the device generator; the host never needs them.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#ifndef GO2_HD
#define GO2_HD __host__ __device__ __forceinline__
#endif

namespace go2 {

GO2_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

GO2_HD uint64_t splitmix64(uint64_t& state) {
  state += 0x9E3779B97F4A7C15ULL;
  return mix64(state);
}

// Key of (seed, config, position index, attempt): chained finaliser rounds.
GO2_HD uint64_t position_key(uint64_t seed, uint32_t config, uint64_t index, uint32_t attempt) {
  uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ULL);
  k = mix64(k ^ ((uint64_t)config * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
  k = mix64(k ^ (index * 0x9E3779B97F4A7C15ULL + 0x8CB92BA72F3D8DD7ULL));
  k = mix64(k ^ ((uint64_t)attempt * 0xC2B2AE3D27D4EB4FULL + 0x165667B19E3779F9ULL));
  return k;
}

GO2_HD int empties_for(uint64_t seed, uint32_t config, uint64_t index, int emin, int emax) {
  const uint64_t k = mix64(position_key(seed, config, index, 0xFFFFFFFFu));
  return emin + (int)(k % (uint64_t)(emax - emin + 1));
}

}  // namespace go2
