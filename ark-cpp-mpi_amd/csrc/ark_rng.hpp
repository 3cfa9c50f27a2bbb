// Counter-based random numbers: Philox4x32-10 (J. K. Salmon, M. A. Moraes,
// R. O. Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11).  A pure function of (counter, key), so every cell, mode and step
// draws its own numbers with no state to carry; the same code runs on the
// host and the device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace arkd {

struct U4 { uint32_t v[4]; };

__host__ __device__ inline void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

__host__ __device__ inline U4 philox4x32_10(U4 ctr, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;   // round multipliers
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;   // key schedule (Weyl) increments
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, ctr.v[0], hi0, lo0);
        mulhilo32(M1, ctr.v[2], hi1, lo1);
        U4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += W0;
        k1 += W1;
    }
    return ctr;
}

// Uniform in (0, 1]: 32 random bits, never 0, so log() below is finite.
__host__ __device__ inline double u01(uint32_t x) { return ((double)x + 1.0) * (1.0 / 4294967296.0); }

// Four standard normal deviates from one Philox draw (Box-Muller).
__host__ __device__ inline void gauss4(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       double g[4]) {
    U4 c{{c0, c1, c2, c3}};
    const U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double two_pi = 6.283185307179586476925286766559;
    for (int h = 0; h < 2; ++h) {
        const double rad = sqrt(-2.0 * log(u01(r.v[2 * h])));
        const double ang = two_pi * u01(r.v[2 * h + 1]);
        g[2 * h] = rad * cos(ang);
        g[2 * h + 1] = rad * sin(ang);
    }
}

}  // namespace arkd
