// Lane-by-lane CPU emulation of the wavefront negacyclic transform in
// tfhe_amd/csrc/negacyclic_fft.h (the same functions the GPU kernels call),
// checked against the exact negacyclic product mod 2^32.
// Usage: fft_emu N TRIALS SEED DIGIT_BITS   -> prints "maxerr <e> mismatches <m>"
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <random>
#include <string>
#include <vector>
#include <cmath>

#include "../tfhe_amd/csrc/negacyclic_fft.h"

using namespace tfhe_amd;

template <int M>
struct Emu {
    using F = Fft<M>;
    static constexpr int R = F::R;
    std::vector<c64> tw;
    Emu() : tw(TwiddleTable<M>::COUNT) { build_twiddles<M>(tw.data()); }
    const c64* tw1(int lane) const { return tw.data() + lane * (R + 8); }
    const c64* tw2(int lane) const { return tw.data() + lane * (R + 8) + R; }

    // poly (length 2M, as doubles) -> frequency registers x[64][R]
    void forward(const double* p, c64 (*x)[R]) {
        std::vector<c64> lds(F::LDS_C64);
        for (int b = 0; b < 64; b++)
            for (int a = 0; a < R; a++) x[b][a] = c64{p[64 * a + b], p[64 * a + b + M]};
        for (int l = 0; l < 64; l++) F::fwd1(x[l], l, lds.data(), tw1(l));
        for (int l = 0; l < 64; l++) F::fwd2(x[l], l, lds.data(), tw2(l)[1]);
        for (int l = 0; l < 64; l++) F::fwd3(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) F::fwd4(x[l], l, lds.data());
    }
    void inverse(c64 (*x)[R], double* p) {
        std::vector<c64> lds(F::LDS_C64);
        for (int l = 0; l < 64; l++) F::inv1(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) F::inv2(x[l], l, lds.data(), tw2(l)[1]);
        for (int l = 0; l < 64; l++) F::inv3(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) F::inv4(x[l], l, lds.data(), tw1(l));
        for (int b = 0; b < 64; b++)
            for (int a = 0; a < R; a++) {
                p[64 * a + b] = x[b][a].re;
                p[64 * a + b + M] = x[b][a].im;
            }
    }
};

// N = 2048 on a pair of wavefronts (FftPair): X[h][lane][slot]
struct EmuPair {
    using F = Fft<512>;
    using X = c64[2][64][8];
    std::vector<c64> tw;
    EmuPair() : tw(PAIR_TW_COUNT) { build_twiddles_pair(tw.data()); }
    const c64* tw1(int h, int lane) const { return tw.data() + (h * 64 + lane) * 16; }
    const c64* tw2(int h, int lane) const { return tw1(h, lane) + 8; }
    void forward(const double* p, c64 (*x)[8]) {  // x = X[2*64][8]
        c64 u[2][64][8];
        for (int h = 0; h < 2; h++)
            for (int b = 0; b < 64; b++)
                for (int a = 0; a < 8; a++) {
                    const int j = 512 * h + 64 * a + b;
                    u[h][b][a] = c64{p[j], p[j + 1024]};
                }
        for (int b = 0; b < 64; b++) {
            for (int a = 0; a < 8; a++) x[b][a] = u[0][b][a], x[64 + b][a] = u[1][b][a];
            FftPair::pair_fwd<0>(x[b], u[1][b]);
            FftPair::pair_fwd<1>(x[64 + b], u[0][b]);
        }
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(F::LDS_C64);
            c64 (*xh)[8] = x + 64 * h;
            for (int l = 0; l < 64; l++) {
                if (h == 0) FftPair::fwd1<0>(xh[l], l, lds.data(), tw1(0, l));
                else FftPair::fwd1<1>(xh[l], l, lds.data(), tw1(1, l));
            }
            for (int l = 0; l < 64; l++) F::fwd2(xh[l], l, lds.data(), tw2(h, l)[1]);
            for (int l = 0; l < 64; l++) F::fwd3(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) F::fwd4(xh[l], l, lds.data());
        }
    }
    void inverse(c64 (*x)[8], double* p) {
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(F::LDS_C64);
            c64 (*xh)[8] = x + 64 * h;
            for (int l = 0; l < 64; l++) F::inv1(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) F::inv2(xh[l], l, lds.data(), tw2(h, l)[1]);
            for (int l = 0; l < 64; l++) F::inv3(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) {
                if (h == 0) FftPair::inv4<0>(xh[l], l, lds.data(), tw1(0, l));
                else FftPair::inv4<1>(xh[l], l, lds.data(), tw1(1, l));
            }
        }
        for (int b = 0; b < 64; b++) {
            c64 y0[8], y1[8];
            for (int a = 0; a < 8; a++) y0[a] = x[b][a], y1[a] = x[64 + b][a];
            c64 u0[8], u1[8];
            for (int a = 0; a < 8; a++) u0[a] = y0[a], u1[a] = y1[a];
            FftPair::pair_inv<0>(u0, y1);
            FftPair::pair_inv<1>(u1, y0);
            for (int a = 0; a < 8; a++) {
                const int j0 = 64 * a + b, j1 = 512 + 64 * a + b;
                p[j0] = u0[a].re, p[j0 + 1024] = u0[a].im;
                p[j1] = u1[a].re, p[j1 + 1024] = u1[a].im;
            }
        }
    }
};


// N = 1024 on a pair of wavefronts, 256 points each (FftHalf): X[h*64 + lane][slot]
struct EmuHalf {
    using H = FftHalf;
    std::vector<c64> tw;
    EmuHalf() : tw(HALF_TW_COUNT) { build_twiddles_half(tw.data()); }
    const c64* t(int h, int lane) const { return tw.data() + (h * 64 + lane) * HALF_TW_PER_LANE; }
    void forward(const double* p, c64 (*x)[4]) {  // x = X[2*64][4]
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(H::LDS_C64);
            c64 (*xh)[4] = x + 64 * h;
            for (int b = 0; b < 64; b++)
                for (int a = 0; a < 4; a++) {
                    const int j = b + 64 * a;
                    const c64 lo{p[j], p[j + 512]}, hi{p[j + 256], p[j + 768]};
                    xh[b][a] = h == 0 ? H::pair_fwd<0>(lo, hi) : H::pair_fwd<1>(lo, hi);
                }
            for (int l = 0; l < 64; l++) {
                if (h == 0) H::fwd1<0>(xh[l], l, lds.data(), t(0, l));
                else H::fwd1<1>(xh[l], l, lds.data(), t(1, l));
            }
            for (int l = 0; l < 64; l++) H::fwd2(xh[l], l, lds.data(), t(h, l));
            for (int l = 0; l < 64; l++) H::fwd2w(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) H::fwd3(xh[l], l, lds.data(), t(h, l));
            for (int l = 0; l < 64; l++) H::fwd3w(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) H::fwd4(xh[l], l, lds.data());
        }
    }
    void inverse(c64 (*x)[4], double* p) {
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(H::LDS_C64);
            c64 (*xh)[4] = x + 64 * h;
            for (int l = 0; l < 64; l++) H::inv1(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) H::inv2(xh[l], l, lds.data(), t(h, l));
            for (int l = 0; l < 64; l++) H::inv2w(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) H::inv3(xh[l], l, lds.data(), t(h, l));
            for (int l = 0; l < 64; l++) H::inv3w(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) {
                if (h == 0) H::inv4<0>(xh[l], l, lds.data(), t(0, l));
                else H::inv4<1>(xh[l], l, lds.data(), t(1, l));
            }
        }
        for (int b = 0; b < 64; b++)
            for (int a = 0; a < 4; a++) {
                const c64 y0 = x[b][a], y1 = x[64 + b][a];
                const c64 lo = H::pair_inv_lo(y0, y1), hi = H::pair_inv_hi(y0, y1);
                const int j = b + 64 * a;
                p[j] = lo.re, p[j + 512] = lo.im;
                p[j + 256] = hi.re, p[j + 768] = hi.im;
            }
    }
};

// N = 1024 with fused butterflies (FftCT): X[lane][slot]
struct EmuCT {
    using F = FftCT;
    std::vector<c64> tw;
    EmuCT() : tw(FftCT::COUNT) { build_twiddles_ct(tw.data()); }
    // the per-lane table entries a lane reads: k-th entry of table `off`
    void lane_tab(int off, int lane, ct64* w, int n) const {
        for (int k = 0; k < n; k++) w[k] = ct64{tw[off + 64 * k + lane].re, tw[off + 64 * k + lane].im};
    }
    void forward(const double* p, c64 (*x)[8]) {
        std::vector<c64> lds(F::LDS_C64);
        for (int b = 0; b < 64; b++)
            for (int a = 0; a < 8; a++) x[b][a] = c64{p[64 * a + b], p[64 * a + b + 512]};
        for (int l = 0; l < 64; l++) F::fwd1(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) {
            ct64 w[4];
            lane_tab(FftCT::FB, l, w, 4);
            F::fwd2(x[l], l, lds.data(), w);
        }
        for (int l = 0; l < 64; l++) F::fwd3(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) {
            ct64 w[4];
            lane_tab(FftCT::FC, l, w, 4);
            F::fwd4(x[l], l, lds.data(), w);
        }
    }
    void inverse(c64 (*x)[8], double* p) {
        std::vector<c64> lds(F::LDS_C64);
        for (int l = 0; l < 64; l++) F::inv1(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) {
            ct64 w[4];
            lane_tab(FftCT::IB, l, w, 4);
            F::inv2(x[l], l, lds.data(), w);
        }
        for (int l = 0; l < 64; l++) F::inv3(x[l], l, lds.data());
        for (int l = 0; l < 64; l++) {
            ct64 w[4];
            c64 ut[8];
            lane_tab(FftCT::IA, l, w, 4);
            for (int a = 0; a < 8; a++) ut[a] = tw[FftCT::UT + 64 * a + l];
            F::inv4(x[l], l, lds.data(), w, ut);
        }
        for (int b = 0; b < 64; b++)
            for (int a = 0; a < 8; a++) {
                p[64 * a + b] = x[b][a].re;
                p[64 * a + b + 512] = x[b][a].im;
            }
    }
};

// N = 2048 on a pair of wavefronts with fused butterflies (FftCTPair): X[h*64 + lane][slot]
struct EmuCTPair {
    using F = FftCT;
    std::vector<c64> tw;
    EmuCTPair() : tw(FftCTPair::COUNT) { build_twiddles_ct_pair(tw.data()); }
    void lane_tab(int h, int off, int lane, ct64* w, int n) const {
        const c64* t = tw.data() + h * FftCT::COUNT;
        for (int k = 0; k < n; k++) w[k] = ct64{t[off + 64 * k + lane].re, t[off + 64 * k + lane].im};
    }
    void forward(const double* p, c64 (*x)[8]) {  // x = X[2*64][8]
        c64 u[2][64][8];
        for (int h = 0; h < 2; h++)
            for (int b = 0; b < 64; b++)
                for (int a = 0; a < 8; a++) {
                    const int j = 512 * h + 64 * a + b;
                    u[h][b][a] = c64{p[j], p[j + 1024]};
                }
        for (int b = 0; b < 64; b++) {
            for (int a = 0; a < 8; a++) x[b][a] = u[0][b][a], x[64 + b][a] = u[1][b][a];
            FftPair::pair_fwd<0>(x[b], u[1][b]);
            FftPair::pair_fwd<1>(x[64 + b], u[0][b]);
        }
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(F::LDS_C64);
            c64 (*xh)[8] = x + 64 * h;
            for (int l = 0; l < 64; l++) {
                if (h == 0) FftCTPair::fwd1<0>(xh[l], l, lds.data());
                else FftCTPair::fwd1<1>(xh[l], l, lds.data());
            }
            for (int l = 0; l < 64; l++) {
                ct64 w[4];
                lane_tab(h, FftCT::FB, l, w, 4);
                F::fwd2(xh[l], l, lds.data(), w);
            }
            for (int l = 0; l < 64; l++) F::fwd3(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) {
                ct64 w[4];
                lane_tab(h, FftCT::FC, l, w, 4);
                F::fwd4(xh[l], l, lds.data(), w);
            }
        }
    }
    void inverse(c64 (*x)[8], double* p) {
        for (int h = 0; h < 2; h++) {
            std::vector<c64> lds(F::LDS_C64);
            c64 (*xh)[8] = x + 64 * h;
            for (int l = 0; l < 64; l++) F::inv1(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) {
                ct64 w[4];
                lane_tab(h, FftCT::IB, l, w, 4);
                F::inv2(xh[l], l, lds.data(), w);
            }
            for (int l = 0; l < 64; l++) F::inv3(xh[l], l, lds.data());
            for (int l = 0; l < 64; l++) {
                ct64 w[4];
                c64 ut[8];
                lane_tab(h, FftCT::IA, l, w, 4);
                for (int a = 0; a < 8; a++) ut[a] = tw[h * FftCT::COUNT + FftCT::UT + 64 * a + l];
                F::inv4(xh[l], l, lds.data(), w, ut);
            }
        }
        for (int b = 0; b < 64; b++) {
            c64 y0[8], y1[8], u0[8], u1[8];
            for (int a = 0; a < 8; a++) y0[a] = x[b][a], y1[a] = x[64 + b][a];
            for (int a = 0; a < 8; a++) u0[a] = y0[a], u1[a] = y1[a];
            FftPair::pair_inv<0>(u0, y1);
            FftPair::pair_inv<1>(u1, y0);
            for (int a = 0; a < 8; a++) {
                const int j0 = 64 * a + b, j1 = 512 + 64 * a + b;
                p[j0] = u0[a].re, p[j0 + 1024] = u0[a].im;
                p[j1] = u1[a].re, p[j1 + 1024] = u1[a].im;
            }
        }
    }
};

template <class Em, int M, int SLOTS>
int run_with(Em& E, int trials, unsigned seed, int dbits);

template <int M>
int run(int trials, unsigned seed, int dbits) {
    Emu<M> E;
    return run_with<Emu<M>, M, M / 64>(E, trials, seed, dbits);
}

template <class Em, int M, int SLOTS>
int run_with(Em& E, int trials, unsigned seed, int dbits) {
    const int N = 2 * M;
    std::mt19937_64 rng(seed);
    double maxerr = 0;
    long mismatches = 0;
    constexpr int LANES = M / SLOTS;
    static c64 X[LANES][SLOTS], Y[LANES][SLOTS];
    std::vector<double> pd(N), pt(N), out(N);
    std::vector<int32_t> d(N), t(N);
    for (int tr = 0; tr < trials; tr++) {
        const int32_t half = 1 << (dbits - 1);
        for (int j = 0; j < N; j++) {
            d[j] = (int32_t)(rng() % (uint64_t)(2 * half)) - half;
            t[j] = (int32_t)(uint32_t)rng();
            pd[j] = d[j];
            pt[j] = t[j];
        }
        // special cases: a monomial and a constant
        if (tr == 0) { for (int j = 0; j < N; j++) { d[j] = (j == 5); pd[j] = d[j]; } }
        if (tr == 1) { for (int j = 0; j < N; j++) { t[j] = 1000 * (j == N - 1); pt[j] = t[j]; } }
        E.forward(pd.data(), X);
        E.forward(pt.data(), Y);
        for (int l = 0; l < LANES; l++)
            for (int s = 0; s < SLOTS; s++) X[l][s] = cmul(X[l][s], Y[l][s]);
        E.inverse(X, out.data());
        // exact product mod 2^32
        std::vector<uint32_t> ex(N, 0);
        for (int i = 0; i < N; i++) {
            if (!d[i]) continue;
            for (int j = 0; j < N; j++) {
                uint32_t v = (uint32_t)d[i] * (uint32_t)t[j];
                int k = i + j;
                if (k < N) ex[k] += v; else ex[k - N] -= v;
            }
        }
        for (int j = 0; j < N; j++) {
            double r = std::nearbyint(out[j]);
            maxerr = std::fmax(maxerr, std::fabs(out[j] - r));
            uint32_t got = (uint32_t)(int64_t)r;
            if (got != ex[j]) mismatches++;
        }
    }
    // round trip of a single polynomial
    double rt = 0;
    for (int j = 0; j < N; j++) pt[j] = (int32_t)(uint32_t)rng();
    E.forward(pt.data(), X);
    E.inverse(X, out.data());
    for (int j = 0; j < N; j++) rt = std::fmax(rt, std::fabs(out[j] - pt[j]));
    printf("maxerr %.6g mismatches %ld roundtrip %.6g\n", maxerr, mismatches, rt);
    return mismatches ? 1 : 0;
}

// Multiply-accumulate as the blind-rotation kernels do it: sum over r < kpl of
// F(d_r) F(t_r), one inverse, rounding; compared with the exact sum of the
// negacyclic products mod 2^32. worst = 0: digits uniform in [-2^(b-1), 2^(b-1))
// and torus values uniform (what a bootstrap sees: the key is uniformly random);
// worst = 1: every digit at its extreme and every key value at +-2^31 with signs
// aligned so that output coefficient 0 reaches kpl N 2^(b-1) 2^31 in magnitude.
// Prints the largest distance of an output to the nearest integer, the
// mismatches, and log2 of the largest exact output magnitude.
template <class Em, int M, int SLOTS>
int run_mac(Em& E, int kpl, int trials, unsigned seed, int dbits, int worst) {
    const int N = 2 * M;
    constexpr int LANES = M / SLOTS;
    std::mt19937_64 rng(seed);
    static c64 X[LANES][SLOTS], Y[LANES][SLOTS], A[LANES][SLOTS];
    std::vector<double> pd(N), pt(N), out(N);
    double maxerr = 0, maxmag = 0;
    long mismatches = 0;
    const int64_t half = (int64_t)1 << (dbits - 1);
    for (int tr = 0; tr < trials; tr++) {
        std::vector<std::vector<int64_t>> d(kpl, std::vector<int64_t>(N)), t(kpl, std::vector<int64_t>(N));
        for (int r = 0; r < kpl; r++)
            for (int j = 0; j < N; j++) {
                if (worst) {
                    // coefficient 0 = sum_i d_i t_{-i} with the wrap sign -1 for i > 0
                    const int64_t sgn = (rng() & 1) ? 1 : -1;
                    d[r][j] = sgn > 0 ? half - 1 : -half;
                    const int64_t tv = (j == 0) ? sgn : -sgn;  // sign of t_{-j} that aligns the product
                    t[r][(N - j) % N] = tv > 0 ? INT32_MAX : INT32_MIN;
                } else {
                    d[r][j] = (int64_t)(rng() % (uint64_t)(2 * half)) - half;
                    t[r][j] = (int32_t)(uint32_t)rng();
                }
            }
        for (int l = 0; l < LANES; l++)
            for (int s = 0; s < SLOTS; s++) A[l][s] = c64{0, 0};
        for (int r = 0; r < kpl; r++) {
            for (int j = 0; j < N; j++) pd[j] = (double)d[r][j], pt[j] = (double)t[r][j];
            E.forward(pd.data(), X);
            E.forward(pt.data(), Y);
            for (int l = 0; l < LANES; l++)
                for (int s = 0; s < SLOTS; s++) cmac(A[l][s], X[l][s], Y[l][s]);
        }
        E.inverse(A, out.data());
        for (int k = 0; k < N; k++) {
            __int128 ex = 0;
            for (int r = 0; r < kpl; r++)
                for (int i = 0; i < N; i++) {
                    const int j = k - i;
                    const __int128 p = (__int128)d[r][i] * t[r][j >= 0 ? j : j + N];
                    ex += j >= 0 ? p : -p;
                }
            maxmag = std::fmax(maxmag, std::fabs((double)ex));
            const double rr = std::nearbyint(out[k]);
            maxerr = std::fmax(maxerr, std::fabs(out[k] - rr));
            if ((uint32_t)(int64_t)rr != (uint32_t)(uint64_t)ex) mismatches++;
        }
    }
    printf("maxerr %.6g mismatches %ld log2max %.3f\n", maxerr, mismatches, std::log2(maxmag));
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "mac") {
        // fft_emu mac N KPL TRIALS SEED DIGIT_BITS WORST   (N = 1: N = 1024 fused; -2048: pair; 2: pair fused)
        const int N = atoi(argv[2]), kpl = atoi(argv[3]), trials = atoi(argv[4]), dbits = atoi(argv[6]);
        const unsigned seed = atoi(argv[5]);
        const int worst = atoi(argv[7]);
        if (N == 1) { EmuCT E; return run_mac<EmuCT, 512, 8>(E, kpl, trials, seed, dbits, worst); }
        if (N == 1024) { Emu<512> E; return run_mac<Emu<512>, 512, 8>(E, kpl, trials, seed, dbits, worst); }
        if (N == -2048) { EmuPair E; return run_mac<EmuPair, 1024, 8>(E, kpl, trials, seed, dbits, worst); }
        if (N == 2) { EmuCTPair E; return run_mac<EmuCTPair, 1024, 8>(E, kpl, trials, seed, dbits, worst); }
        if (N == -1024) { EmuHalf E; return run_mac<EmuHalf, 512, 4>(E, kpl, trials, seed, dbits, worst); }
        return 2;
    }
    const int N = argc > 1 ? atoi(argv[1]) : 1024;
    const int trials = argc > 2 ? atoi(argv[2]) : 4;
    const unsigned seed = argc > 3 ? atoi(argv[3]) : 1;
    const int dbits = argc > 4 ? atoi(argv[4]) : 7;
    if (N == 1024) return run<512>(trials, seed, dbits);
    if (N == 2048) return run<1024>(trials, seed, dbits);
    if (N == -2048) {  // pair-of-wavefronts transform
        EmuPair E;
        return run_with<EmuPair, 1024, 8>(E, trials, seed, dbits);
    }
    if (N == 2) {  // N = 2048, pair of wavefronts, fused butterflies
        EmuCTPair E;
        return run_with<EmuCTPair, 1024, 8>(E, trials, seed, dbits);
    }
    if (N == 1) {  // N = 1024, fused butterflies
        EmuCT E;
        return run_with<EmuCT, 512, 8>(E, trials, seed, dbits);
    }
    if (N == -1024) {  // pair of wavefronts, 256 points each
        EmuHalf E;
        return run_with<EmuHalf, 512, 4>(E, trials, seed, dbits);
    }
    fprintf(stderr, "N must be 1024 or 2048\n");
    return 2;
}
