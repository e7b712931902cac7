// Known-answer vectors of Philox4x32-10 (the Random123 distribution's
// kat_vectors for philox4x32_10) against tfhe_amd/csrc/philox.h.
#include <cstdio>

#include "../../tfhe_amd/csrc/philox.h"

using tfhe_amd::Philox4;

int main() {
    struct Kat {
        Philox4 ctr;
        unsigned k0, k1;
        Philox4 expect;
    } kats[] = {
        {{{0u, 0u, 0u, 0u}}, 0u, 0u, {{0x6627e8d5u, 0xe169c58du, 0xbc57ac4cu, 0x9b00dbd8u}}},
        {{{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}}, 0xffffffffu, 0xffffffffu,
         {{0x408f276du, 0x41c83b0eu, 0xa20bc7c6u, 0x6d5451fdu}}},
        {{{0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}}, 0xa4093822u, 0x299f31d0u,
         {{0xd16cfe09u, 0x94fdccebu, 0x5001e420u, 0x24126ea1u}}},
    };
    int bad = 0;
    for (const Kat& k : kats) {
        const Philox4 r = tfhe_amd::philox4x32_10(k.ctr, k.k0, k.k1);
        for (int i = 0; i < 4; i++) bad += r.v[i] != k.expect.v[i];
    }
    printf("%s\n", bad ? "MISMATCH" : "OK");
    return bad != 0;
}
