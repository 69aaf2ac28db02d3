// Philox4x64-10 counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11), usable on host and
// device.  Streams used by the library:
//   lattice jitter : key {seed, 0x4C415454}, counter {atom, chain, 0, 0}
//   heat bath      : key {seed, 0x48454154}, counter {atom, trajectory, chain, 0}
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chmc {

struct u64x4 { uint64_t v[4]; };

__host__ __device__ inline void mulhilo64(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
    lo = a * b;
    hi = __umul64hi(a, b);
#else
    unsigned __int128 prod = (unsigned __int128)a * b;
    lo = (uint64_t)prod;
    hi = (uint64_t)(prod >> 64);
#endif
}

__host__ __device__ inline u64x4 philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                              uint64_t k0, uint64_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B97F4A7C15ULL;
            k1 += 0xBB67AE8584CAA73BULL;
        }
        uint64_t lo0, hi0, lo1, hi1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, lo0, hi0);
        mulhilo64(0xCA5A826395121157ULL, c2, lo1, hi1);
        uint64_t n0 = hi1 ^ c1 ^ k0;
        uint64_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    u64x4 r;
    r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
    return r;
}

// 53-bit uniforms: [0,1) and (0,1)
__host__ __device__ inline double u01_co(uint64_t w) { return (double)(w >> 11) * 0x1.0p-53; }
__host__ __device__ inline double u01_oo(uint64_t w) { return ((double)(w >> 11) + 0.5) * 0x1.0p-53; }

constexpr uint64_t kKeyLattice = 0x4C415454ULL;
constexpr uint64_t kKeyHeatbath = 0x48454154ULL;

}  // namespace chmc
