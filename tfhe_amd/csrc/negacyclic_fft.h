// Negacyclic FP64 transform for polynomials mod X^N+1, laid out for one
// 64-lane wavefront.
//
// A real polynomial p of degree < N = 2M is folded to M complex points
//   z_j = (p_j + i p_{j+M}) * psi^j,   psi = exp(i*pi/N),   j < M
// and Z = DFT_M(z) (positive exponent) evaluates p at the M roots of X^N+1 of
// the form exp(i*pi*(4k+1)/N); the other M roots are their conjugates. The
// negacyclic product is then pointwise, as in the reference's
// IntPolynomial_ifft / TorusPolynomial_fft pair
// (fft_processors/spqlios/fft_processor_spqlios.cpp:38-115), whose element
// order, scaling and truncation are NOT reproduced: the order here is private
// to this library (the bootstrapping key is transformed by the same routine),
// the inverse is exactly scaled and results are rounded to nearest.
//
// Wavefront layout, M = 64 R (R = 8 for N = 1024, 16 for N = 2048):
//   L0  lane b, slot a      : z_{64a+b}                        (coefficient order)
//   stage 1  (in registers) : R-point DFT over a, times psi^b * w_M^{b k1}
//   T1  (LDS transpose)     : lane (b0 + 8c), slot (g, b1)  with b = b0 + 8 b1, k1 = cG+g
//   stage 2a (in registers) : 8-point DFT over b1, times w_64^{b0 k2a}
//   T2  (LDS transpose)     : lane (k2a + 8c), slot (g, b0)
//   stage 2b (in registers) : 8-point DFT over b0
// Register (lane, g*8+j) then holds frequency k1 + R*(k2a + 8*k2b) with
// k2b = bitrev3(j). Forward stages use decimation in frequency (natural in,
// bit-reversed out); the inverse runs the same steps backwards with
// decimation-in-time butterflies, so no explicit bit reversal is ever done.
//
// The functions below are split at every point where lanes exchange data
// through LDS, so the same code runs on the GPU (one wavefront, in order) and
// in a lane-by-lane CPU emulation used by the tests.
#pragma once

#include <stdint.h>

#include <cmath>

#if defined(__HIPCC__)
#define TA_HD __host__ __device__ __forceinline__
#else
#define TA_HD inline
#endif

namespace tfhe_amd {

struct __attribute__((aligned(16))) c64 {
    double re, im;
};

TA_HD c64 cadd(c64 a, c64 b) { return {a.re + b.re, a.im + b.im}; }
TA_HD c64 csub(c64 a, c64 b) { return {a.re - b.re, a.im - b.im}; }
TA_HD c64 cmul(c64 a, c64 b) { return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)}; }
// a * conj(b)
TA_HD c64 cmulc(c64 a, c64 b) { return {fma(a.re, b.re, a.im * b.im), fma(a.im, b.re, -(a.re * b.im))}; }
// cmac into a zero accumulator, the same roundings: {fma(-a.im, b.im, a.re b.re), fma(a.im, b.re, a.re b.im)}
TA_HD c64 cmul_first(c64 a, c64 b) { return {fma(-a.im, b.im, a.re * b.re), fma(a.im, b.re, a.re * b.im)}; }
// acc += a * b
TA_HD void cmac(c64& acc, c64 a, c64 b) {
    acc.re = fma(a.re, b.re, acc.re);
    acc.re = fma(-a.im, b.im, acc.re);
    acc.im = fma(a.re, b.im, acc.im);
    acc.im = fma(a.im, b.re, acc.im);
}

// cos(pi k / 32), k = 0..16 (sin(pi k/32) = cos(pi (16-k)/32))
constexpr double kCos32[17] = {
    1,
    0.995184726672196886231,
    0.980785280403230449119,
    0.956940335732208864931,
    0.923879532511286756101,
    0.881921264348355029715,
    0.831469612302545237081,
    0.773010453362736960797,
    0.707106781186547524382,
    0.634393284163645498203,
    0.555570233019602224757,
    0.471396736825997648572,
    0.382683432365089771696,
    0.290284677254462367618,
    0.19509032201612826787,
    0.0980171403295606019347,
    -2.50827880633416601173e-20};

constexpr double kC8 = 0.92387953251128675613;  // cos(pi/8)
constexpr double kS8 = 0.38268343236508977173;  // sin(pi/8)
constexpr double kH = 0.70710678118654752440;   // cos(pi/4)

// v * exp(S * 2*pi*i * m / L) for L in {8, 16}; m is a compile-time constant
// after unrolling, so the switch folds away.
template <int L, int S>
TA_HD c64 twc(c64 v, int m) {
    const int q = ((m * (16 / L)) % 16 + 16) % 16;
    double c, s;
    switch (q) {
        case 0: return v;
        case 4: return S > 0 ? c64{-v.im, v.re} : c64{v.im, -v.re};
        case 8: return c64{-v.re, -v.im};
        case 12: return S > 0 ? c64{v.im, -v.re} : c64{-v.im, v.re};
        case 1: c = kC8; s = kS8; break;
        case 2: c = kH; s = kH; break;
        case 3: c = kS8; s = kC8; break;
        case 5: c = -kS8; s = kC8; break;
        case 6: c = -kH; s = kH; break;
        case 7: c = -kC8; s = kS8; break;
        case 9: c = -kC8; s = -kS8; break;
        case 10: c = -kH; s = -kH; break;
        case 11: c = -kS8; s = -kC8; break;
        case 13: c = kS8; s = -kC8; break;
        case 14: c = kH; s = -kH; break;
        default: c = kC8; s = -kS8; break;  // 15
    }
    return cmul(v, c64{c, S > 0 ? s : -s});
}

// In-register DFT, natural order in, bit-reversed order out:
// x[j] <- sum_n x[n] exp(S 2 pi i n bitrev(j) / L)
template <int L, int S>
TA_HD void dif(c64* x) {
#pragma unroll
    for (int span = L / 2; span >= 1; span >>= 1) {
#pragma unroll
        for (int start = 0; start < L; start += 2 * span) {
#pragma unroll
            for (int j = 0; j < span; j++) {
                c64 a = x[start + j], b = x[start + j + span];
                x[start + j] = cadd(a, b);
                x[start + j + span] = twc<L, S>(csub(a, b), j * (L / (2 * span)));
            }
        }
    }
}

// In-register DFT, bit-reversed order in, natural order out.
template <int L, int S>
TA_HD void dit(c64* x) {
#pragma unroll
    for (int span = 1; span < L; span <<= 1) {
#pragma unroll
        for (int start = 0; start < L; start += 2 * span) {
#pragma unroll
            for (int j = 0; j < span; j++) {
                c64 a = x[start + j], b = twc<L, S>(x[start + j + span], j * (L / (2 * span)));
                x[start + j] = cadd(a, b);
                x[start + j + span] = csub(a, b);
            }
        }
    }
}

TA_HD constexpr int bitrev(int v, int bits) {
    int r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

template <int M>
struct Fft {
    static constexpr int R = M / 64;  // slots per lane
    static constexpr int G = R / 8;   // 8-point groups per lane in stage 2
    static constexpr int LOGR = (R == 8) ? 3 : 4;
    static constexpr int ROW = 72;    // LDS row stride in c64 (64 + 8 pad)
    static constexpr int LDS_C64 = R * ROW;
    static_assert(R == 8 || R == 16, "M must be 512 or 1024");

    TA_HD static int a1(int k1, int b) { return k1 * ROW + b; }
    TA_HD static int a2(int k1, int b0, int k2a) { return k1 * ROW + k2a * 9 + b0; }

    // psi^{64a} = exp(i pi 32 a / M) = cos/sin(pi k / 32) with k = 16 a / R: a compile-time constant
    TA_HD static constexpr c64 twist(int a) {
        return c64{kCos32[a * (16 / R)], kCos32[16 - a * (16 / R)]};
    }
    // ---- forward ----
    // x[a] = p_{64a+b} + i p_{64a+b+M} (untwisted); tw1[k1] = psi^b w_M^{b k1}
    // tw1[k1 * TS] = psi^b w_M^{b k1} (TS = 1: per-lane array; TS = 64: shared table column)
    template <int TS = 1>
    TA_HD static void fwd1(c64* x, int b, c64* lds, const c64* tw1) {
        // twiddles requested before the in-register DFT so their loads overlap it
        c64 t[R];
#pragma unroll
        for (int k = 0; k < R; k++) t[k] = tw1[k * TS];
#pragma unroll
        for (int a = 1; a < R; a++) x[a] = cmul(x[a], twist(a));
        dif<R, 1>(x);
#pragma unroll
        for (int j = 0; j < R; j++) {
            const int k1 = bitrev(j, LOGR);
            lds[a1(k1, b)] = cmul(x[j], t[k1]);
        }
    }
    // Powers w^0..w^7 of w = w_64^{b0}, b0 = lane & 7 (the stage-2a twiddles),
    // rebuilt from one base value each time instead of occupying 28 registers.
    TA_HD static void powers8(c64 w, c64* t) {
        t[0] = c64{1.0, 0.0};
        t[1] = w;
        t[2] = cmul(w, w);
        t[3] = cmul(t[2], w);
        t[4] = cmul(t[2], t[2]);
        t[5] = cmul(t[4], w);
        t[6] = cmul(t[4], t[2]);
        t[7] = cmul(t[4], t[3]);
    }
    TA_HD static void fwd2(c64* x, int lane, const c64* lds, c64 w) {
        c64 t[8];
        powers8(w, t);
        fwd2_pow(x, lane, lds, t);
    }
    TA_HD static void fwd2_pow(c64* x, int lane, const c64* lds, const c64* t) {
        const int b0 = lane & 7, c = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++) {
#pragma unroll
            for (int b1 = 0; b1 < 8; b1++) x[g * 8 + b1] = lds[a1(c * G + g, b0 + 8 * b1)];
            dif<8, 1>(x + g * 8);
#pragma unroll
            for (int j = 1; j < 8; j++) x[g * 8 + j] = cmul(x[g * 8 + j], t[bitrev(j, 3)]);
        }
    }
    TA_HD static void fwd3(const c64* x, int lane, c64* lds) {
        const int b0 = lane & 7, c = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int j = 0; j < 8; j++) lds[a2(c * G + g, b0, bitrev(j, 3))] = x[g * 8 + j];
    }
    TA_HD static void fwd4(c64* x, int lane, const c64* lds) {
        const int k2a = lane & 7, c = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++) {
#pragma unroll
            for (int b0 = 0; b0 < 8; b0++) x[g * 8 + b0] = lds[a2(c * G + g, b0, k2a)];
            dif<8, 1>(x + g * 8);
        }
    }

    // ---- inverse (exact inverse of the forward steps, times M) ----
    TA_HD static void inv1(c64* x, int lane, c64* lds) {
        const int k2a = lane & 7, c = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++) {
            dit<8, -1>(x + g * 8);
#pragma unroll
            for (int b0 = 0; b0 < 8; b0++) lds[a2(c * G + g, b0, k2a)] = x[g * 8 + b0];
        }
    }
    TA_HD static void inv2(c64* x, int lane, const c64* lds, c64 w) {
        c64 t[8];
        powers8(w, t);
        inv2_pow(x, lane, lds, t);
    }
    TA_HD static void inv2_pow(c64* x, int lane, const c64* lds, const c64* t) {
        const int b0 = lane & 7, c = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++) {
            x[g * 8] = lds[a2(c * G + g, b0, 0)];
#pragma unroll
            for (int j = 1; j < 8; j++) {
                const int k2a = bitrev(j, 3);
                x[g * 8 + j] = cmulc(lds[a2(c * G + g, b0, k2a)], t[k2a]);
            }
            dit<8, -1>(x + g * 8);
        }
    }
    TA_HD static void inv3(const c64* x, int lane, c64* lds) {
        const int b0 = lane & 7, c = lane >> 3;
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int b1 = 0; b1 < 8; b1++) lds[a1(c * G + g, b0 + 8 * b1)] = x[g * 8 + b1];
    }
    // result x[a] = p_{64a+b} + i p_{64a+b+M} (scaled by 1/M and untwisted by conj(psi^{64a}))
    template <int TS = 1>
    TA_HD static void inv4(c64* x, int b, const c64* lds, const c64* tw1) {
#pragma unroll
        for (int j = 0; j < R; j++) {
            const int k1 = bitrev(j, LOGR);
            x[j] = cmulc(lds[a1(k1, b)], tw1[k1 * TS]);
        }
        dit<R, -1>(x);
#pragma unroll
        for (int a = 0; a < R; a++) {
            const c64 t = twist(a);
            x[a] = cmul(x[a], c64{t.re / M, -t.im / M});
        }
    }
};

// exp(i pi m / 32) for any integer m (a compile-time constant after unrolling)
TA_HD constexpr double cos32(int m) {
    return (m = ((m % 64) + 64) % 64) <= 16 ? kCos32[m]
           : m <= 32                         ? -kCos32[32 - m]
           : m <= 48                         ? -kCos32[m - 32]
                                             : kCos32[64 - m];
}
TA_HD constexpr c64 e32(int m) { return c64{cos32(m), cos32(16 - m)}; }

// Negacyclic transform for N = 2048 on a PAIR of wavefronts. Folding as above
// gives u_j = p_j + i p_{j+1024} (j < 1024) and the wanted spectrum is
// X_m = sum_j u_j psi^j w^{jm} (psi = exp(i pi/2048), w = exp(2 pi i/1024)).
// Wavefront h of the pair holds u_j for j = 512h + 64a + b (lane b, slot a).
// One radix-2 step across the pair (through LDS) forms
//   y_h[j'] = u_{j'} + (-1)^h e^{i pi/4} u_{j'+512},       j' = 64a + b < 512,
// and then X_{2k+h} = sum_{j'} y_h[j'] psi^{(1+4h) j'} w_512^{j'k}: a 512-point
// DFT of y_h with the twist psi^{(1+4h) j'}, run by wavefront h alone with the
// Fft<512> stages (the twist split into the compile-time part psi^{(1+4h) 64a}
// and the per-lane part psi^{(1+4h) b}, which is folded into tw1). So
// wavefront h ends with the odd (h = 1) or even (h = 0) bins. The inverse runs
// the steps backwards; the pair step's 1/2 is folded into the final scale.
struct FftPair {
    using F = Fft<512>;
    static constexpr int R = 8;
    static constexpr double SQ = kH;  // e^{i pi/4} = (SQ, SQ)

    // y from own u (u) and the partner's u (o)
    template <int H>
    TA_HD static void pair_fwd(c64* x, const c64* o) {
#pragma unroll
        for (int a = 0; a < R; a++) {
            const c64 u = x[a], v = o[a];
            if (H == 0) {  // u_{j'} + e^{i pi/4} u_{j'+512}; v = u_{j'+512}
                x[a] = c64{u.re + SQ * (v.re - v.im), u.im + SQ * (v.re + v.im)};
            } else {  // u_{j'} - e^{i pi/4} u_{j'+512}; u = u_{j'+512}, v = u_{j'}
                x[a] = c64{v.re - SQ * (u.re - u.im), v.im - SQ * (u.re + u.im)};
            }
        }
    }
    // stage 1 of wavefront h: twist, 8-point DFT, per-lane twiddle, transpose write
    template <int H, int TS = 1>
    TA_HD static void fwd1(c64* x, int b, c64* lds, const c64* tw1) {
#pragma unroll
        for (int a = 1; a < R; a++) x[a] = cmul(x[a], e32((1 + 4 * H) * a));
        dif<R, 1>(x);
#pragma unroll
        for (int j = 0; j < R; j++) {
            const int k1 = bitrev(j, 3);
            lds[F::a1(k1, b)] = cmul(x[j], tw1[k1 * TS]);
        }
    }
    // last inverse stage of wavefront h: y_h / 2 (scaled by 1/1024)
    template <int H, int TS = 1>
    TA_HD static void inv4(c64* x, int b, const c64* lds, const c64* tw1) {
#pragma unroll
        for (int j = 0; j < R; j++) {
            const int k1 = bitrev(j, 3);
            x[j] = cmulc(lds[F::a1(k1, b)], tw1[k1 * TS]);
        }
        dit<R, -1>(x);
#pragma unroll
        for (int a = 0; a < R; a++) {
            const c64 t = e32((1 + 4 * H) * a);
            x[a] = cmul(x[a], c64{t.re / 1024.0, -t.im / 1024.0});
        }
    }
    // u from own y/2 and the partner's y/2: u_{j'} = y0 + y1 (h = 0),
    // u_{j'+512} = (y0 - y1) e^{-i pi/4} (h = 1)
    template <int H>
    TA_HD static void pair_inv(c64* x, const c64* o) {
#pragma unroll
        for (int a = 0; a < R; a++) {
            const c64 y = x[a], v = o[a];
            if (H == 0) {
                x[a] = cadd(y, v);
            } else {
                const c64 d = csub(v, y);  // y0 - y1
                x[a] = c64{SQ * (d.re + d.im), SQ * (d.im - d.re)};
            }
        }
    }
};

// Negacyclic transform for N = 1024 on a PAIR of wavefronts, 256 points each
// (four slots per lane), for the latency kernel. With u_j = p_j + i p_{j+512}
// (j < 512) and X_m = sum_j u_j psi^j w_512^{jm} (psi = exp(i pi/1024)),
// splitting j = j' + 256 t gives
//   y_h[j'] = u_{j'} + (-1)^h e^{i pi/4} u_{j'+256},            j' < 256,
//   X_{2K+h} = sum_{j'} y_h[j'] psi^{(1+4h) j'} w_256^{j' K},
// so wavefront h forms y_h from the whole polynomial (no exchange with its
// partner) and runs a 256-point DFT. With j' = b + 64a (lane b, slot a) and
// K = k1 + 4 k2:
//   stage 1  (in registers)   : twist psi^{(1+4h) 64a}, 4-point DFT over a,
//                               times tw1[k1] = psi^{(1+4h) b} w_256^{b k1};
//   then the 64-point DFT over b for each row r = k1, as three radix-4 steps
//   with b = g + 16 tau, g = g0 + 4 g1, k2 = mu + 4 (nu + 4 omega):
//   step A, lane (r, g)       : 4-point DFT over tau -> mu, times w_64^{g mu}
//   step B, lane (r, mu, g0)  : 4-point DFT over g1  -> nu, times w_16^{g0 nu}
//   step C, lane (r, mu, nu)  : 4-point DFT over g0  -> omega
// (lane numbers r*16 + g and r*16 + mu*4 + g0 / nu). Slot omega of lane
// (r, mu, nu) then holds X_{2K+h} with K = r + 4 mu + 16 nu + 64 omega. The
// inverse runs the steps backwards with negative exponents; the pair's
// combination u_{j'} = y0 + y1, u_{j'+256} = (y0 - y1) e^{-i pi/4} needs both
// halves. Each function either computes and writes the LDS buffer or reads it
// and computes, so the same code runs on the GPU (one wavefront, in order) and
// in the lane-by-lane CPU emulation of the tests.
struct FftHalf {
    static constexpr int R = 4;
    static constexpr int LDS_C64 = 256;
    static constexpr double SQ = kH;  // e^{i pi/4} = (SQ, SQ)
    // Element e = 64 d3 + 16 d2 + 4 d1 + d0 (base-4 digits) sits at
    //   64 d3 + 16 d2 + 4 (d1 ^ d2) + (d0 ^ d1).
    // Every step below holds three of the digits in the lane and one (its slot)
    // in a register; with this placement any 8 or 16 consecutive lanes of a
    // step touch distinct 16-byte bank groups, so no LDS access conflicts. The
    // bases are the slot-0 addresses; a slot adds 64 k1 (step 1) or is XORed in.
    TA_HD static int base1(int b) {  // lane (d0, d1, d2); slot d3: + 64 k1
        const int d0 = b & 3, d1 = (b >> 2) & 3, d2 = (b >> 4) & 3;
        return 16 * d2 + 4 * (d1 ^ d2) + (d0 ^ d1);
    }
    TA_HD static int base2(int lane) {  // lane (d0, d1, d3); slot t = d2: (base ^ 4 t) + 16 t
        const int d0 = lane & 3, d1 = (lane >> 2) & 3, d3 = lane >> 4;
        return 64 * d3 + 4 * d1 + (d0 ^ d1);
    }
    TA_HD static int base3(int lane) {  // lane (d0, d2, d3); slot g1 = d1: base ^ 5 g1
        const int d0 = lane & 3, d2 = (lane >> 2) & 3, d3 = lane >> 4;
        return 64 * d3 + 20 * d2 + d0;
    }
    TA_HD static int base4(int lane) {  // lane (d1, d2, d3); slot g0 = d0: base ^ g0
        const int d1 = lane & 3, d2 = (lane >> 2) & 3, d3 = lane >> 4;
        return 64 * d3 + 16 * d2 + 4 * (d1 ^ d2) + d1;
    }
    TA_HD static int at2(int b2, int t) { return (b2 ^ (4 * t)) + 16 * t; }

    // natural-order 4-point DFT in place, exponent sign S
    template <int S>
    TA_HD static void dft4(c64* x) {
        const c64 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
        const c64 t2 = cadd(x[1], x[3]), t3 = csub(x[1], x[3]);
        // i * t3 (S > 0) or -i * t3 (S < 0)
        const c64 it3 = S > 0 ? c64{-t3.im, t3.re} : c64{t3.im, -t3.re};
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = cadd(t1, it3);
        x[3] = csub(t1, it3);
    }
    // y_h from u_lo = u_{j'} and u_hi = u_{j'+256}
    template <int H>
    TA_HD static c64 pair_fwd(c64 lo, c64 hi) {
        const c64 r = c64{SQ * (hi.re - hi.im), SQ * (hi.re + hi.im)};  // e^{i pi/4} hi
        return H == 0 ? cadd(lo, r) : csub(lo, r);
    }
    // ---- forward; tw holds tw1[0..3], w_64^{g mu} (mu = 1..3), w_16^{g0 nu} (nu = 1..3) ----
    template <int H>
    TA_HD static void fwd1(c64* x, int b, c64* lds, const c64* tw) {
#pragma unroll
        for (int a = 1; a < R; a++) x[a] = cmul(x[a], e32(2 * (1 + 4 * H) * a));
        dft4<1>(x);
        c64* d = lds + base1(b);
#pragma unroll
        for (int k1 = 0; k1 < R; k1++) d[64 * k1] = cmul(x[k1], tw[k1]);
    }
    TA_HD static void fwd2(c64* x, int lane, const c64* lds, const c64* tw) {
        const int b2 = base2(lane);
#pragma unroll
        for (int t = 0; t < R; t++) x[t] = lds[at2(b2, t)];
        dft4<1>(x);
#pragma unroll
        for (int mu = 1; mu < R; mu++) x[mu] = cmul(x[mu], tw[3 + mu]);
    }
    TA_HD static void fwd2w(const c64* x, int lane, c64* lds) {
        const int b2 = base2(lane);
#pragma unroll
        for (int mu = 0; mu < R; mu++) lds[at2(b2, mu)] = x[mu];
    }
    TA_HD static void fwd3(c64* x, int lane, const c64* lds, const c64* tw) {
        const int b3 = base3(lane);
#pragma unroll
        for (int g1 = 0; g1 < R; g1++) x[g1] = lds[b3 ^ (5 * g1)];
        dft4<1>(x);
#pragma unroll
        for (int nu = 1; nu < R; nu++) x[nu] = cmul(x[nu], tw[6 + nu]);
    }
    TA_HD static void fwd3w(const c64* x, int lane, c64* lds) {
        const int b3 = base3(lane);
#pragma unroll
        for (int nu = 0; nu < R; nu++) lds[b3 ^ (5 * nu)] = x[nu];
    }
    TA_HD static void fwd4(c64* x, int lane, const c64* lds) {
        const int b4 = base4(lane);
#pragma unroll
        for (int g0 = 0; g0 < R; g0++) x[g0] = lds[b4 ^ g0];
        dft4<1>(x);
    }
    // ---- inverse (exact inverse of the forward steps, times 256) ----
    TA_HD static void inv1(c64* x, int lane, c64* lds) {
        dft4<-1>(x);
        const int b4 = base4(lane);
#pragma unroll
        for (int g0 = 0; g0 < R; g0++) lds[b4 ^ g0] = x[g0];
    }
    TA_HD static void inv2(c64* x, int lane, const c64* lds, const c64* tw) {
        const int b3 = base3(lane);
        x[0] = lds[b3];
#pragma unroll
        for (int nu = 1; nu < R; nu++) x[nu] = cmulc(lds[b3 ^ (5 * nu)], tw[6 + nu]);
        dft4<-1>(x);
    }
    TA_HD static void inv2w(const c64* x, int lane, c64* lds) {
        const int b3 = base3(lane);
#pragma unroll
        for (int g1 = 0; g1 < R; g1++) lds[b3 ^ (5 * g1)] = x[g1];
    }
    TA_HD static void inv3(c64* x, int lane, const c64* lds, const c64* tw) {
        const int b2 = base2(lane);
        x[0] = lds[b2];
#pragma unroll
        for (int mu = 1; mu < R; mu++) x[mu] = cmulc(lds[at2(b2, mu)], tw[3 + mu]);
        dft4<-1>(x);
    }
    TA_HD static void inv3w(const c64* x, int lane, c64* lds) {
        const int b2 = base2(lane);
#pragma unroll
        for (int t = 0; t < R; t++) lds[at2(b2, t)] = x[t];
    }
    // result: y_h[b + 64a] / 2 (the pair's 1/2 and the 1/256 folded in)
    template <int H>
    TA_HD static void inv4(c64* x, int b, const c64* lds, const c64* tw) {
        const c64* q = lds + base1(b);
#pragma unroll
        for (int k1 = 0; k1 < R; k1++) x[k1] = cmulc(q[64 * k1], tw[k1]);
        dft4<-1>(x);
#pragma unroll
        for (int a = 0; a < R; a++) {
            const c64 t = e32(2 * (1 + 4 * H) * a);
            x[a] = cmul(x[a], c64{t.re / 512.0, -t.im / 512.0});
        }
    }
    // u_{j'} = y0 + y1 and u_{j'+256} = (y0 - y1) e^{-i pi/4} from y0/2 and y1/2
    TA_HD static c64 pair_inv_lo(c64 y0, c64 y1) { return cadd(y0, y1); }
    TA_HD static c64 pair_inv_hi(c64 y0, c64 y1) {
        const c64 d = csub(y0, y1);
        return c64{SQ * (d.re + d.im), SQ * (d.im - d.re)};
    }
};

// Negacyclic transform for N = 1024 (M = 512) with fused butterflies, for the
// batched kernel. Same folding z_j = p_j + i p_{j+512} and same wavefront
// layouts and LDS transposes as Fft<512>, but the arithmetic is a negacyclic
// Cooley-Tukey split of U(X) = sum_j u_j X^j modulo X^512 - i: a block held
// modulo X^{2m} - rho splits into X^m - r and X^m + r (r^2 = rho) with the
// butterfly (x_j, x_{j+m}) -> (x_j + r x_{j+m}, x_j - r x_{j+m}). The twist by
// psi^j is part of the roots, so there is no separate twist, and every
// butterfly has its twiddle on an input, which fuses: with r = c (1 + i t),
//   v = (y.re - t y.im, y.im + t y.re)      (2 FMA)
//   x +- r y = (x.re +- c v.re, x.im +- c v.im)   (4 FMA)
// six FMAs per butterfly instead of a complex product and four additions. The
// variants i r, -i r and -r use the same (c, t) with the parts of v swapped or
// negated. Starting from rho = i = psi^512 (psi = exp(i pi/1024)), a block whose
// root is psi^e has children psi^{e/2} and psi^{e/2 + 1024}; after nine levels
// position P (9 bits, most significant = first split) holds U(psi^{1 + 4t})
// with t = bitrev9(P), i.e. the cyclic spectrum of u_j psi^j in bit-reversed
// order. The inverse is therefore a decimation-in-time cyclic inverse DFT
// (again twiddles on inputs, fused) from bit-reversed input to natural order,
// followed by the untwist psi^{-j} / 512.
//   forward  stage A (registers, slot a = position bits 8..6): levels 1-3,
//            compile-time roots; T1 (LDS): lane b0 + 8a, slot b1;
//            stage B: levels 4-6, roots of lane's a (table FB);
//            T2 (LDS): lane q + 8a, slot b0; stage C: levels 7-9 (table FC).
//            Register (lane q + 8a, slot p) holds position P = 64a + 8q + p.
//   inverse  stage C' (spans 1, 2, 4 over the slot, compile-time twiddles);
//            T2 backwards: lane p + 8a, slot q; stage B' (spans 8, 16, 32,
//            table IB, p = lane & 7); T1 backwards: lane b, slot a; stage A'
//            (spans 64, 128, 256, table IA, b = lane); untwist (table UT).
// Each function either computes and writes the LDS buffer or reads it and
// computes, so the same code runs on the GPU and in the CPU emulation.
struct __attribute__((aligned(16))) ct64 {
    double c, t;  // r = c (1 + i t)
};

// (x, y) <- (x + w y, x - w y) with w = r (V = 0), i r (1), -i r (2), -r (3)
template <int V>
TA_HD void ctb(c64& x, c64& y, double c, double t) {
    const double vr = fma(-t, y.im, y.re), vi = fma(t, y.re, y.im);
    double pr, pi;  // w y = c (pr, pi)
    if (V == 0) { pr = vr; pi = vi; }
    else if (V == 1) { pr = -vi; pi = vr; }
    else if (V == 2) { pr = vi; pi = -vr; }
    else { pr = -vr; pi = -vi; }
    const c64 a = x;
    x = c64{fma(c, pr, a.re), fma(c, pi, a.im)};
    y = c64{fma(-c, pr, a.re), fma(-c, pi, a.im)};
}
template <int V>
TA_HD void ctb(c64& x, c64& y, ct64 w) { ctb<V>(x, y, w.c, w.t); }
// (x, y) <- (x + y, x - y)
TA_HD void bf1(c64& x, c64& y) {
    const c64 a = x;
    x = cadd(a, y);
    y = csub(a, y);
}
// (x, y) <- (x - i y, x + i y)
TA_HD void bfmi(c64& x, c64& y) {
    const c64 a = x;
    x = c64{a.re + y.im, a.im - y.re};
    y = c64{a.re - y.im, a.im + y.re};
}

struct FftCT {
    static constexpr int R = 8;
    static constexpr int LDS_C64 = Fft<512>::LDS_C64;
    // compile-time roots (cos, tan) of pi/4, pi/8, pi/16, -3 pi/16
    static constexpr double C4 = 0.70710678118654752440, T4 = 1.0;
    static constexpr double C8 = 0.92387953251128675613, T8 = 0.41421356237309504880;
    static constexpr double C16 = 0.98078528040323044913, T16 = 0.19891236737965800691;
    static constexpr double C316 = 0.83146961230254523708, T316 = -0.66817863791929891999;
    // per-lane tables (ct64 / c64 [k][lane]), offsets in 16-byte entries
    static constexpr int FB = 0, FC = 4 * 64, IB = 8 * 64, IA = 12 * 64, UT = 16 * 64;
    static constexpr int COUNT = 24 * 64;

    TA_HD static int a1(int k1, int b) { return Fft<512>::a1(k1, b); }
    TA_HD static int a2(int k1, int b0, int k2a) { return Fft<512>::a2(k1, b0, k2a); }

    // levels 1-3 on slots 0..7 with roots psi^256; psi^128, i psi^128;
    // psi^64, i psi^64, psi^320 = i e^{-3 i pi/16}, psi^832 = -e^{-3 i pi/16}
    TA_HD static void stageA(c64* x) {
#pragma unroll
        for (int a = 0; a < 4; a++) ctb<0>(x[a], x[a + 4], C4, T4);
        ctb<0>(x[0], x[2], C8, T8);
        ctb<0>(x[1], x[3], C8, T8);
        ctb<1>(x[4], x[6], C8, T8);
        ctb<1>(x[5], x[7], C8, T8);
        ctb<0>(x[0], x[1], C16, T16);
        ctb<1>(x[2], x[3], C16, T16);
        ctb<1>(x[4], x[5], C316, T316);
        ctb<3>(x[6], x[7], C316, T316);
    }
    // levels with table roots w0 (level 1 of the stage), w1 (level 2), w2, w3 (level 3)
    TA_HD static void stage_tab(c64* x, const ct64* w) {
#pragma unroll
        for (int q = 0; q < 4; q++) ctb<0>(x[q], x[q + 4], w[0]);
        ctb<0>(x[0], x[2], w[1]);
        ctb<0>(x[1], x[3], w[1]);
        ctb<1>(x[4], x[6], w[1]);
        ctb<1>(x[5], x[7], w[1]);
        ctb<0>(x[0], x[1], w[2]);
        ctb<1>(x[2], x[3], w[2]);
        ctb<0>(x[4], x[5], w[3]);
        ctb<1>(x[6], x[7], w[3]);
    }
    // ---- forward: x[a] = p_{64a+b} + i p_{64a+b+512} at lane b ----
    TA_HD static void fwd1(c64* x, int b, c64* lds) {
        stageA(x);
#pragma unroll
        for (int a = 0; a < R; a++) lds[a1(a, b)] = x[a];
    }
    // reads in the order 0, 4, 1, 5, ...: the first butterflies pair q and q + 4
    TA_HD static void fwd2(c64* x, int lane, const c64* lds, const ct64* wb) {
        const int b0 = lane & 7, a = lane >> 3;
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int q = (k >> 1) + 4 * (k & 1);
            x[q] = lds[a1(a, b0 + 8 * q)];
        }
        stage_tab(x, wb);
    }
    TA_HD static void fwd3(const c64* x, int lane, c64* lds) {
        const int b0 = lane & 7, a = lane >> 3;
#pragma unroll
        for (int q = 0; q < R; q++) lds[a2(a, b0, q)] = x[q];
    }
    TA_HD static void fwd4(c64* x, int lane, const c64* lds, const ct64* wc) {
        const int q = lane & 7, a = lane >> 3;
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int p = (k >> 1) + 4 * (k & 1);
            x[p] = lds[a2(a, p, q)];
        }
        stage_tab(x, wc);
    }
    // ---- inverse (times 512 before the untwist, which divides it out) ----
    // spans 1, 2, 4 of the decimation-in-time inverse DFT: twiddles 1; 1, -i;
    // 1, e^{-i pi/4}, -i, -i e^{-i pi/4}
    TA_HD static void stageC_inv(c64* x) {
#pragma unroll
        for (int p = 0; p < R; p += 2) bf1(x[p], x[p + 1]);
        bf1(x[0], x[2]);
        bfmi(x[1], x[3]);
        bf1(x[4], x[6]);
        bfmi(x[5], x[7]);
        bf1(x[0], x[4]);
        ctb<0>(x[1], x[5], C4, -T4);
        bfmi(x[2], x[6]);
        ctb<2>(x[3], x[7], C4, -T4);
    }
    // spans m, 2m, 4m with table twiddles w0 (span m), w1 (2m), w2, w3 (4m: slots 0, 1; 2, 3 use -i w2, -i w3)
    TA_HD static void stage_inv_tab(c64* x, const ct64* w) {
#pragma unroll
        for (int q = 0; q < R; q += 2) ctb<0>(x[q], x[q + 1], w[0]);
        ctb<0>(x[0], x[2], w[1]);
        ctb<2>(x[1], x[3], w[1]);
        ctb<0>(x[4], x[6], w[1]);
        ctb<2>(x[5], x[7], w[1]);
        ctb<0>(x[0], x[4], w[2]);
        ctb<0>(x[1], x[5], w[3]);
        ctb<2>(x[2], x[6], w[2]);
        ctb<2>(x[3], x[7], w[3]);
    }
    TA_HD static void inv1(c64* x, int lane, c64* lds) {
        const int q = lane & 7, a = lane >> 3;
        stageC_inv(x);
#pragma unroll
        for (int p = 0; p < R; p++) lds[a2(a, p, q)] = x[p];
    }
    TA_HD static void inv2(c64* x, int lane, const c64* lds, const ct64* wb) {
        const int p = lane & 7, a = lane >> 3;
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = lds[a2(a, p, q)];
        stage_inv_tab(x, wb);
    }
    TA_HD static void inv3(const c64* x, int lane, c64* lds) {
        const int p = lane & 7, a = lane >> 3;
#pragma unroll
        for (int q = 0; q < R; q++) lds[a1(a, p + 8 * q)] = x[q];
    }
    // result x[a] = p_{64a+b} + i p_{64a+b+512} (ut[a] = psi^{-(64a+b)} / 512)
    TA_HD static void inv4(c64* x, int b, const c64* lds, const ct64* wa, const c64* ut) {
#pragma unroll
        for (int a = 0; a < R; a++) x[a] = lds[a1(a, b)];
        stage_inv_tab(x, wa);
#pragma unroll
        for (int a = 0; a < R; a++) x[a] = cmul(x[a], ut[a]);
    }
};

// Per-lane tables of the fused-butterfly transform of a polynomial modulo
// X^512 - psi^E0, [k][lane], built in long double and rounded once (psi =
// exp(i pi / (half)), a half turn being `half` units; a square root of psi^e is
// psi^{e/2} or psi^{e/2 + half}):
// FB (forward levels 4-6, lane b0 + 8a): psi^{E/2}, psi^{E/4}, psi^{E/8}, psi^{E/8+Q}
//    with psi^E the root of block a after three levels and Q a quarter of `half`
//    (psi^Q = e^{i pi/4});
// FC (levels 7-9, lane q + 8a): the same with the root of block (a, q) after six;
// IB (spans 8-32, p = lane & 7): w16^-p, w32^-p, w64^-p, e^{-i pi/4} w64^-p;
// IA (spans 64-256, b = lane): w128^-b, w256^-b, w512^-b, e^{-i pi/4} w512^-b
//    (w_L = exp(2 pi i / L)); UT [a][lane]: psi^{-(E0/512)(64a+b)} * scale as c64.
// FftCT (N = 1024): E0 = 512, half = 1024 (psi^E0 = i), scale 1/512.
inline void build_twiddles_ct_base(c64* out, int E0, int half, long double scale) {
    const long double pi = 3.141592653589793238462643383279502884L;
    // (cos, tan) of ang as two doubles in a c64 slot. At odd multiples of pi/2
    // (w = +-i, tan infinite) w is represented as 2^-60 (1 + i (+-2^60)): off by
    // 2^-60 in the real part, far below FP64 rounding, and in the fused
    // butterfly c v is an exact power-of-two scaling of the rounded v, so the
    // result keeps full precision.
    auto ct = [&](long double ang) {
        const long double c = cosl(ang), s = sinl(ang);
        if (fabsl(c) < 1e-12L) return c64{0x1p-60, (double)(s > 0 ? 0x1p60L : -0x1p60L)};
        return c64{(double)c, (double)(s / c)};
    };
    auto root = [&](int bits, int path) {  // exponent e of psi^e for a block after `bits` levels
        int e = E0;
        for (int k = bits - 1; k >= 0; k--) e = e / 2 + half * ((path >> k) & 1);
        return e;
    };
    const long double ps = pi / half;
    const int Q = half / 4, tw = E0 / 512;
    for (int lane = 0; lane < 64; lane++) {
        const int a = lane >> 3, q = lane & 7;
        const int E = root(3, a), E6 = root(6, 8 * a + q);
        const int fb[4] = {E / 2, E / 4, E / 8, E / 8 + Q}, fc[4] = {E6 / 2, E6 / 4, E6 / 8, E6 / 8 + Q};
        for (int k = 0; k < 4; k++) {
            out[FftCT::FB + 64 * k + lane] = ct(ps * fb[k]);
            out[FftCT::FC + 64 * k + lane] = ct(ps * fc[k]);
        }
        const int p = lane & 7, b = lane;
        out[FftCT::IB + 0 * 64 + lane] = ct(-2 * pi * p / 16);
        out[FftCT::IB + 1 * 64 + lane] = ct(-2 * pi * p / 32);
        out[FftCT::IB + 2 * 64 + lane] = ct(-2 * pi * p / 64);
        out[FftCT::IB + 3 * 64 + lane] = ct(-2 * pi * p / 64 - pi / 4);
        out[FftCT::IA + 0 * 64 + lane] = ct(-2 * pi * b / 128);
        out[FftCT::IA + 1 * 64 + lane] = ct(-2 * pi * b / 256);
        out[FftCT::IA + 2 * 64 + lane] = ct(-2 * pi * b / 512);
        out[FftCT::IA + 3 * 64 + lane] = ct(-2 * pi * b / 512 - pi / 4);
        for (int aa = 0; aa < 8; aa++) {
            const long double ang = -ps * (long double)((tw * (64 * aa + b)) % (2 * half));
            out[FftCT::UT + 64 * aa + lane] = c64{(double)(cosl(ang) * scale), (double)(sinl(ang) * scale)};
        }
    }
}
inline void build_twiddles_ct(c64* out) { build_twiddles_ct_base(out, 512, 1024, 1.0L / 512); }

// N = 2048 on a pair of wavefronts with the fused butterflies. With u_j = p_j +
// i p_{j+1024} (j < 1024) the folded polynomial is taken modulo X^1024 - i
// (psi = exp(i pi / 2048), i = psi^1024). The first level splits it into
// X^512 - zeta_h, zeta_0 = e^{i pi/4} = psi^512, zeta_1 = -zeta_0 = psi^2560:
//   y_h[j'] = u_{j'} + (-1)^h e^{i pi/4} u_{j'+512}   (FftPair::pair_fwd),
// and wavefront h runs the nine remaining levels on y_h, modulo X^512 - zeta_h,
// exactly as FftCT does modulo X^512 - i but with the roots of zeta_h: stage A
// with compile-time roots per h, then FftCT's stages with the h tables
// (build_twiddles_ct_base with E0 = 512 + 2048 h, half = 2048). The inverse is
// FftCT's cyclic inverse, whose untwist psi^{-(1+4h) j} / 1024 also halves y_h
// for the pair step u_{j'} = y0 + y1, u_{j'+512} = (y0 - y1) e^{-i pi/4}.
struct FftCTPair {
    static constexpr int R = 8;
    static constexpr int COUNT = 2 * FftCT::COUNT;  // [h][FftCT table layout]
    // (cos, tan) of pi/8, pi/16, -3 pi/16, pi/32, 5 pi/32, 3 pi/32, 7 pi/32
    static constexpr double C8 = FftCT::C8, T8 = FftCT::T8, C16 = FftCT::C16, T16 = FftCT::T16;
    static constexpr double C316 = FftCT::C316, T316 = FftCT::T316;
    static constexpr double C132 = 0.99518472667219692873, T132 = 0.098491403357164247767;
    static constexpr double C532 = 0.88192126434835504956, T532 = 0.53451113595079169372;
    static constexpr double C332 = 0.95694033573220882438, T332 = 0.30334668360734240444;
    static constexpr double C732 = 0.77301045336273699338, T732 = 0.8206787908286603539;

    // levels 2-4 of the whole transform (1-3 of wavefront h) on slots 0..7:
    // h = 0: roots psi^256 = e^{i pi/8}; e^{i pi/16}, i e^{i pi/16};
    //        e^{i pi/32}, i e^{i pi/32}, e^{9 i pi/32} = i e^{-7 i pi/32}, -e^{-7 i pi/32}
    // h = 1: e^{5 i pi/8} = i e^{i pi/8}; e^{5 i pi/16} = i e^{-3 i pi/16}, -e^{-3 i pi/16};
    //        e^{5 i pi/32}, i e^{5 i pi/32}, e^{13 i pi/32} = i e^{-3 i pi/32}, -e^{-3 i pi/32}
    template <int H>
    TA_HD static void stageA(c64* x) {
        if (H == 0) {
#pragma unroll
            for (int a = 0; a < 4; a++) ctb<0>(x[a], x[a + 4], C8, T8);
            ctb<0>(x[0], x[2], C16, T16);
            ctb<0>(x[1], x[3], C16, T16);
            ctb<1>(x[4], x[6], C16, T16);
            ctb<1>(x[5], x[7], C16, T16);
            ctb<0>(x[0], x[1], C132, T132);
            ctb<1>(x[2], x[3], C132, T132);
            ctb<1>(x[4], x[5], C732, -T732);
            ctb<3>(x[6], x[7], C732, -T732);
        } else {
#pragma unroll
            for (int a = 0; a < 4; a++) ctb<1>(x[a], x[a + 4], C8, T8);
            ctb<1>(x[0], x[2], C316, T316);
            ctb<1>(x[1], x[3], C316, T316);
            ctb<3>(x[4], x[6], C316, T316);
            ctb<3>(x[5], x[7], C316, T316);
            ctb<0>(x[0], x[1], C532, T532);
            ctb<1>(x[2], x[3], C532, T532);
            ctb<1>(x[4], x[5], C332, -T332);
            ctb<3>(x[6], x[7], C332, -T332);
        }
    }
    template <int H>
    TA_HD static void fwd1(c64* x, int b, c64* lds) {
        stageA<H>(x);
#pragma unroll
        for (int a = 0; a < R; a++) lds[FftCT::a1(a, b)] = x[a];
    }
};
inline void build_twiddles_ct_pair(c64* out) {
    for (int h = 0; h < 2; h++) build_twiddles_ct_base(out + h * FftCT::COUNT, 512 + 2048 * h, 2048, 1.0L / 1024);
}

// Twiddles of the half transform: [h][lane][10]: tw1_h[k1] = psi^{(1+4h) b} w_256^{b k1}
// (psi = exp(i pi/1024), b = lane), then w_64^{g mu} (g = lane & 15, mu = 1..3) and
// w_16^{g0 nu} (g0 = lane & 3, nu = 1..3).
constexpr int HALF_TW_PER_LANE = 10;
constexpr int HALF_TW_COUNT = 2 * 64 * HALF_TW_PER_LANE;
inline void build_twiddles_half(c64* out) {
    const long double pi = 3.141592653589793238462643383279502884L;
    auto ex = [&](long double ang) { return c64{(double)cosl(ang), (double)sinl(ang)}; };
    for (int h = 0; h < 2; h++)
        for (int b = 0; b < 64; b++) {
            c64* t = out + (h * 64 + b) * HALF_TW_PER_LANE;
            for (int k1 = 0; k1 < 4; k1++) {
                // exp(i pi ((1+4h) b + 8 b k1) / 1024)
                const long e = ((long)(1 + 4 * h) * b + 8L * b * k1) % 2048;
                t[k1] = ex(pi * (long double)e / 1024.0L);
            }
            for (int mu = 1; mu < 4; mu++) t[3 + mu] = ex(2.0L * pi * (long double)(((b & 15) * mu) % 64) / 64.0L);
            for (int nu = 1; nu < 4; nu++) t[6 + nu] = ex(2.0L * pi * (long double)(((b & 3) * nu) % 16) / 16.0L);
        }
}

// Twiddles of the pair transform: [h][lane][16]: tw1_h[k1] = psi^{(1+4h) b} w_512^{b k1}
// (psi = exp(i pi/2048)), then the 8 stage-2a twiddles w_64^{(b&7) k2a}.
constexpr int PAIR_TW_COUNT = 2 * 64 * 16;
inline void build_twiddles_pair(c64* out) {
    const long double pi = 3.141592653589793238462643383279502884L;
    auto ex = [&](long double ang) { return c64{(double)cosl(ang), (double)sinl(ang)}; };
    for (int h = 0; h < 2; h++)
        for (int b = 0; b < 64; b++) {
            c64* t = out + (h * 64 + b) * 16;
            for (int k1 = 0; k1 < 8; k1++) {
                // exp(i pi ((1+4h) b + 8 b k1) / 2048)
                const long e = ((long)(1 + 4 * h) * b + 8L * b * k1) % 4096;
                t[k1] = ex(pi * (long double)e / 2048.0L);
            }
            for (int k2a = 0; k2a < 8; k2a++) t[8 + k2a] = ex(2.0L * pi * (long double)(((b & 7) * k2a) % 64) / 64.0L);
        }
}

// Per-lane twiddle tables, built on the host in long double and rounded once.
// Layout: for each lane, R entries of tw1 then 8 entries of tw2; then the R
// uniform twist entries and R inverse-twist entries; then the tw1 entries again
// as [k1][lane] (TW1_OFF), so that a wavefront reads one k1 column with one
// coalesced 1 KiB load.
template <int M>
struct TwiddleTable {
    static constexpr int R = M / 64;
    static constexpr int PER_LANE = R + 8;
    static constexpr int TW1_OFF = 64 * PER_LANE + 2 * R;
    static constexpr int COUNT = TW1_OFF + R * 64;
};

template <int M>
inline void build_twiddles(c64* out) {
    constexpr int R = M / 64;
    const long double pi = 3.141592653589793238462643383279502884L;
    auto ex = [&](long double ang) { return c64{(double)cosl(ang), (double)sinl(ang)}; };
    for (int b = 0; b < 64; b++) {
        c64* t = out + b * (R + 8);
        for (int k1 = 0; k1 < R; k1++) {
            // psi^b * w_M^{b k1} = exp(i pi b (1 + 4 k1) / (2M))
            const long e = ((long)b * (1 + 4 * k1)) % (4 * M);
            t[k1] = ex(pi * (long double)e / (2.0L * M));
            out[TwiddleTable<M>::TW1_OFF + k1 * 64 + b] = t[k1];
        }
        const int b0 = b & 7;
        for (int k2a = 0; k2a < 8; k2a++) t[R + k2a] = ex(2.0L * pi * (long double)((b0 * k2a) % 64) / 64.0L);
    }
    c64* tw = out + 64 * (R + 8);
    for (int a = 0; a < R; a++) {
        // psi^{64a} = exp(i pi 64 a / (2M))
        c64 v = ex(pi * 64.0L * a / (2.0L * M));
        tw[a] = v;
        tw[R + a] = c64{v.re / M, -v.im / M};
    }
}

}  // namespace tfhe_amd
