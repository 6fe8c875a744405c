// Counter-based random numbers for the device: Philox4x32-10 (Salmon,
// Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).
// Every draw is a pure function of (key, counter), so a particle's collision
// and emission outcomes depend only on (seed, particle id, step, species,
// stream tag), never on thread order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace scdat {

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                         uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return U4{c0, c1, c2, c3};
}

// (0,1) uniform from 32 bits: (u + 0.5) 2^-32, exact in double.
__device__ __forceinline__ double u01(uint32_t u) { return ((double)u + 0.5) * 2.3283064365386963e-10; }

// Stream tags (fourth counter word).
constexpr uint32_t kTagMccA = 0x4D43, kTagMccB = 0x4D44, kTagMccId = 0x4D45, kTagSee = 0x5345,
                   kTagSeeK = 0x5346, kTagSeeId = 0x5347, kTagPhoto = 0x5048, kTagPhotoK = 0x5049,
                   kTagPhotoK2 = 0x504A;

// Generated particle ids carry bit 31.
__device__ __forceinline__ int32_t gen_id(uint32_t bits) { return (int32_t)(0x80000000u | (bits & 0x7FFFFFFFu)); }

}  // namespace scdat
