// Philox4x32-10 counter-based generator (J. K. Salmon, M. A. Moraes, R. O.
// Dror, D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011):
// ten rounds of two 32x32->64 multiplications with the published multipliers
// and Weyl key increments. Host and device; used for batch encryption on the
// device, where every ciphertext draws from its own counter.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define TFHE_AMD_HD __host__ __device__ inline
#else
#define TFHE_AMD_HD inline
#endif

namespace tfhe_amd {

struct Philox4 {
    uint32_t v[4];
};

TFHE_AMD_HD Philox4 philox4x32_10(Philox4 c, uint32_t k0, uint32_t k1) {
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += W0;
            k1 += W1;
        }
        const uint64_t p0 = (uint64_t)M0 * c.v[0];
        const uint64_t p1 = (uint64_t)M1 * c.v[2];
        const Philox4 n{{(uint32_t)(p1 >> 32) ^ c.v[1] ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1,
                         (uint32_t)p0}};
        c = n;
    }
    return c;
}

}  // namespace tfhe_amd
