// Lane-by-lane CPU emulation of one key row of the two-wavefront N = 2048 blind
// rotation (brq_kernel in tfhe_amd/csrc/engine.hip: rotation from the shared
// accumulator with one mask and a sign bit-field per value, decomposition, the
// pair transform with fused butterflies of negacyclic_fft.h, multiply-accumulate,
// the handover of the other component's half, both inverse halves and the
// accumulator update), checked against the exact ACC + sum_{c,p} dec_p((X^abar - 1)
// ACC_c) * key[c l + p] mod 2^32 with a random accumulator and random key rows
// (l = 2, Bgbit = 10). Usage: brq_emu SEED ABAR -> "abar A mismatches M of 4096".
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include <cmath>
#include "../tfhe_amd/csrc/negacyclic_fft.h"
using namespace tfhe_amd;
static inline int32_t sbfe(int32_t v, int off, int w) { return (int32_t)((uint32_t)v << (32 - off - w)) >> (32 - w); }
static inline int32_t round_mod32(double v) { return (int32_t)(uint32_t)(long long)std::nearbyint(v); }
int main(int argc, char** argv) {
    const int N = 2048, L = 2, KPL = 4, Bgbit = 10;
    std::mt19937_64 rng(argc > 1 ? atoi(argv[1]) : 1);
    const int abar = argc > 2 ? atoi(argv[2]) : 777;
    // tables
    std::vector<c64> twq(FftCTPair::COUNT);
    build_twiddles_ct_pair(twq.data());
    // random key row: kpl x 2 components of N int32 (torus)
    std::vector<int32_t> key((size_t)KPL * 2 * N);
    for (auto& v : key) v = (int32_t)(uint32_t)rng();
    // key in the pair CT layout: [r][comp][h][slot][lane] via the emulated conversion
    std::vector<c64> bkq((size_t)KPL * 2 * 1024);
    for (int poly = 0; poly < KPL * 2; poly++) {
        const int32_t* p = &key[(size_t)poly * N];
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(FftCT::LDS_C64);
            c64 X[64][8];
            for (int lane = 0; lane < 64; lane++) {
                c64 x[8], ov[8];
                for (int a = 0; a < 8; a++) {
                    x[a] = c64{(double)p[512 * h + 64 * a + lane], (double)p[512 * h + 64 * a + lane + 1024]};
                    ov[a] = c64{(double)p[512 * (1 - h) + 64 * a + lane], (double)p[512 * (1 - h) + 64 * a + lane + 1024]};
                }
                if (h == 0) FftPair::pair_fwd<0>(x, ov); else FftPair::pair_fwd<1>(x, ov);
                for (int a = 0; a < 8; a++) X[lane][a] = x[a];
            }
            const ct64* tw = (const ct64*)(twq.data() + h * FftCT::COUNT);
            for (int l = 0; l < 64; l++) { if (h == 0) FftCTPair::fwd1<0>(X[l], l, lds.data()); else FftCTPair::fwd1<1>(X[l], l, lds.data()); }
            for (int l = 0; l < 64; l++) { ct64 w[4]; for (int k = 0; k < 4; k++) w[k] = tw[FftCT::FB + 64 * k + l]; FftCT::fwd2(X[l], l, lds.data(), w); }
            for (int l = 0; l < 64; l++) FftCT::fwd3(X[l], l, lds.data());
            for (int l = 0; l < 64; l++) { ct64 w[4]; for (int k = 0; k < 4; k++) w[k] = tw[FftCT::FC + 64 * k + l]; FftCT::fwd4(X[l], l, lds.data(), w); }
            for (int l = 0; l < 64; l++) for (int s = 0; s < 8; s++) bkq[(size_t)poly * 1024 + h * 512 + s * 64 + l] = X[l][s];
        }
    }
    // accumulator
    std::vector<int32_t> acc0(2 * N);
    for (auto& v : acc0) v = (int32_t)(uint32_t)rng();
    std::vector<int32_t> s_acc = acc0;  // [2][N]
    const uint32_t offset = (1u << (32 - Bgbit - 1)) + (1u << (32 - 2 * Bgbit - 1));
    const int32_t dhalf = 1 << (Bgbit - 1);
    // per wave registers
    c64 r0[2][64][8], r1[2][64][8];
    memset(r0, 0, sizeof r0); memset(r1, 0, sizeof r1);
    std::vector<c64> s_fft(2 * FftCT::LDS_C64);
    for (int c = 0; c < 2; c++) {
        uint32_t t[2][64][32];
        for (int h = 0; h < 2; h++)
            for (int lane = 0; lane < 64; lane++) {
                const int jo = 512 * h + lane, jp = 512 * (1 - h) + lane;
                const uint32_t xo = ((uint32_t)(jo - abar) & (2 * N - 1)) * 4u, xp = ((uint32_t)(jp - abar) & (2 * N - 1)) * 4u;
                for (int k = 0; k < 32; k++) {
                    const int g = k / 16, q = (k / 8) & 1, a = k % 8;
                    const uint32_t x = (g ? xp : xo) + 4u * (uint32_t)(1024 * q + 64 * a);
                    const uint32_t m = (uint32_t)sbfe((int32_t)x, 13, 1);
                    const uint32_t rv = (uint32_t)s_acc[c * N + ((x & (4u * N - 1)) >> 2)];
                    const uint32_t av = (uint32_t)s_acc[c * N + (g ? jp : jo) + 1024 * q + 64 * a];
                    t[h][lane][k] = ((rv ^ m) - m) - av + offset;
                }
            }
        for (int p = 0; p < L; p++) {
            const int shift = 32 - (p + 1) * Bgbit;
            for (int h = 0; h < 2; h++) {
                c64 X[64][8];
                c64* fb = s_fft.data() + h * FftCT::LDS_C64;
                for (int lane = 0; lane < 64; lane++) {
                    c64 x[8], ov[8];
                    auto dig = [&](uint32_t v) { return (double)((int32_t)((v >> shift) & ((1u << Bgbit) - 1)) - dhalf); };
                    for (int a = 0; a < 8; a++) {
                        x[a] = c64{dig(t[h][lane][a]), dig(t[h][lane][8 + a])};
                        ov[a] = c64{dig(t[h][lane][16 + a]), dig(t[h][lane][24 + a])};
                    }
                    if (h == 0) FftPair::pair_fwd<0>(x, ov); else FftPair::pair_fwd<1>(x, ov);
                    for (int a = 0; a < 8; a++) X[lane][a] = x[a];
                }
                const ct64* tw = (const ct64*)(twq.data() + h * FftCT::COUNT);
                for (int l = 0; l < 64; l++) { if (h == 0) FftCTPair::fwd1<0>(X[l], l, fb); else FftCTPair::fwd1<1>(X[l], l, fb); }
                for (int l = 0; l < 64; l++) { ct64 w[4]; for (int k = 0; k < 4; k++) w[k] = tw[FftCT::FB + 64 * k + l]; FftCT::fwd2(X[l], l, fb, w); }
                for (int l = 0; l < 64; l++) FftCT::fwd3(X[l], l, fb);
                for (int l = 0; l < 64; l++) { ct64 w[4]; for (int k = 0; k < 4; k++) w[k] = tw[FftCT::FC + 64 * k + l]; FftCT::fwd4(X[l], l, fb, w); }
                for (int l = 0; l < 64; l++) {
                    const c64* q = bkq.data() + ((size_t)(c * L + p)) * 2048 + h * 512 + l;
                    for (int s = 0; s < 8; s++) { cmac(r0[h][l][s], X[l][s], q[s * 64]); cmac(r1[h][l][s], X[l][s], q[1024 + s * 64]); }
                }
            }
        }
    }
    // handover, inverse, combine
    for (int h = 0; h < 2; h++) {
        c64* fb = s_fft.data() + h * FftCT::LDS_C64;
        c64 Y0[64][8], Y1[64][8];
        for (int l = 0; l < 64; l++)
            for (int s = 0; s < 8; s++) {
                if (h == 0) { Y0[l][s] = r0[0][l][s]; Y1[l][s] = r0[1][l][s]; }
                else { Y0[l][s] = r1[0][l][s]; Y1[l][s] = r1[1][l][s]; }
            }
        for (int hh = 0; hh < 2; hh++) {
            c64 (*X)[8] = hh ? Y1 : Y0;
            const ct64* tw = (const ct64*)(twq.data());
            for (int l = 0; l < 64; l++) FftCT::inv1(X[l], l, fb);
            for (int l = 0; l < 64; l++) { ct64 w[4]; for (int k = 0; k < 4; k++) w[k] = tw[FftCT::IB + 64 * k + l]; FftCT::inv2(X[l], l, fb, w); }
            for (int l = 0; l < 64; l++) FftCT::inv3(X[l], l, fb);
            for (int l = 0; l < 64; l++) {
                ct64 w[4]; c64 ut[8];
                for (int k = 0; k < 4; k++) w[k] = tw[FftCT::IA + 64 * k + l];
                for (int k = 0; k < 8; k++) ut[k] = twq[hh * FftCT::COUNT + FftCT::UT + 64 * k + l];
                FftCT::inv4(X[l], l, fb, w, ut);
            }
        }
        for (int l = 0; l < 64; l++)
            for (int a = 0; a < 8; a++) {
                const c64 lo = cadd(Y0[l][a], Y1[l][a]);
                const c64 dd = csub(Y0[l][a], Y1[l][a]);
                const c64 hi = c64{FftPair::SQ * (dd.re + dd.im), FftPair::SQ * (dd.im - dd.re)};
                int32_t* sa = &s_acc[h * N + l];
                sa[64 * a] = (int32_t)((uint32_t)sa[64 * a] + (uint32_t)round_mod32(lo.re));
                sa[1024 + 64 * a] = (int32_t)((uint32_t)sa[1024 + 64 * a] + (uint32_t)round_mod32(lo.im));
                sa[512 + 64 * a] = (int32_t)((uint32_t)sa[512 + 64 * a] + (uint32_t)round_mod32(hi.re));
                sa[1536 + 64 * a] = (int32_t)((uint32_t)sa[1536 + 64 * a] + (uint32_t)round_mod32(hi.im));
            }
    }
    // exact: ACC += sum_{c,p} dec_p(T_c) * key[c L + p][out]
    std::vector<uint32_t> ref(2 * N);
    for (int i = 0; i < 2 * N; i++) ref[i] = (uint32_t)acc0[i];
    for (int c = 0; c < 2; c++) {
        std::vector<int32_t> T(N);
        for (int j = 0; j < N; j++) {
            const int src = ((j - abar) % (2 * N) + 2 * N) % (2 * N);
            const uint32_t rv = src < N ? (uint32_t)acc0[c * N + src] : 0u - (uint32_t)acc0[c * N + src - N];
            T[j] = (int32_t)(rv - (uint32_t)acc0[c * N + j] + offset);
        }
        for (int p = 0; p < L; p++) {
            std::vector<int32_t> d(N);
            for (int j = 0; j < N; j++) d[j] = (int32_t)(((uint32_t)T[j] >> (32 - (p + 1) * Bgbit)) & ((1u << Bgbit) - 1)) - (1 << (Bgbit - 1));
            for (int o = 0; o < 2; o++) {
                const int32_t* kk = &key[((size_t)(c * L + p) * 2 + o) * N];
                for (int i = 0; i < N; i++) {
                    if (!d[i]) continue;
                    for (int j = 0; j < N; j++) {
                        const uint32_t v = (uint32_t)d[i] * (uint32_t)kk[j];
                        const int k = i + j;
                        if (k < N) ref[o * N + k] += v; else ref[o * N + k - N] -= v;
                    }
                }
            }
        }
    }
    int bad = 0;
    for (int i = 0; i < 2 * N; i++) if ((uint32_t)s_acc[i] != ref[i]) bad++;
    printf("abar %d mismatches %d of %d\n", abar, bad, 2 * N);
    return bad != 0;
}
