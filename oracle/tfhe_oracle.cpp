// CPU oracle for TFHE gate bootstrapping — TEST INFRASTRUCTURE ONLY.
// Restates the reference algorithm (file:line citations are into tfhe/tfhe,
// src/). Used only by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg; never linked into the product library.
#include "tfhe_oracle.h"

#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

namespace {

// libstdc++ default_random_engine seeded from a seed_seq, as
// libtfhe/numeric-functions.cpp:11,16-19. Distribution objects are created
// exactly where the reference creates them, because std::normal_distribution
// caches its second polar-method value inside the object.
std::default_random_engine g_gen;

inline uint32_t U(int32_t x) { return (uint32_t)x; }
inline int32_t S(uint32_t x) { return (int32_t)x; }

int32_t uniform_torus() {
    // libtfhe/numeric-functions.cpp:12 (uniform_int_distribution<Torus32>(INT32_MIN, INT32_MAX))
    std::uniform_int_distribution<int32_t> d(INT32_MIN, INT32_MAX);
    return d(g_gen);
}

int32_t gaussian32(int32_t message, double sigma) {
    // libtfhe/numeric-functions.cpp:22-27: a fresh normal_distribution per call
    std::normal_distribution<double> d(0., sigma);
    double err = d(g_gen);
    return S(U(message) + U(orc_dtot32(err)));
}

int32_t key_bit() {
    std::uniform_int_distribution<int32_t> d(0, 1);
    return d(g_gen);
}

}  // namespace

extern "C" {

void orc_set_seed(const uint32_t* values, int32_t size) {
    std::seed_seq seeds(values, values + size);
    g_gen.seed(seeds);
}

// libtfhe/numeric-functions.cpp:32-34
int32_t orc_dtot32(double d) { return (int32_t)(int64_t)((d - (double)(int64_t)d) * 4294967296.0); }

// libtfhe/numeric-functions.cpp:59-65: round(phase * Msize / 2^32) in [0, Msize)
int32_t orc_modswitch_from_torus32(int32_t phase, int32_t Msize) {
    uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    uint64_t half = interv / 2;
    uint64_t phase64 = ((uint64_t)U(phase) << 32) + half;
    return (int32_t)(phase64 / interv);
}

// libtfhe/numeric-functions.cpp:71-76
int32_t orc_modswitch_to_torus32(int32_t mu, int32_t Msize) {
    uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    uint64_t phase64 = (uint64_t)(int64_t)mu * interv;
    return (int32_t)(uint32_t)(phase64 >> 32);
}

// libtfhe/lwe-functions.cpp:36-47 (gaussian first, then n uniform mask values)
void orc_lwe_encrypt(const int32_t* key, int32_t n, int32_t mu, double alpha, int32_t* out) {
    uint32_t b = U(gaussian32(mu, alpha));
    for (int32_t i = 0; i < n; i++) {
        out[i] = uniform_torus();
        b += U(out[i]) * U(key[i]);
    }
    out[n] = S(b);
}

// libtfhe/tfhe_gate_bootstrapping.cpp:154-160
void orc_boots_encrypt(const orc_params* P, const int32_t* key, int32_t bit, int32_t* out) {
    int32_t s8 = orc_modswitch_to_torus32(1, 8);
    orc_lwe_encrypt(key, P->n, bit ? s8 : -s8, P->lwe_stdev, out);
}

// libtfhe/lwe-functions.cpp:72-82
int32_t orc_lwe_phase(const int32_t* ct, const int32_t* key, int32_t n) {
    uint32_t axs = 0;
    for (int32_t i = 0; i < n; i++) axs += U(ct[i]) * U(key[i]);
    return S(U(ct[n]) - axs);
}

// Linear form applied before the bootstrap, libtfhe/boot-gates.cpp:34-319 and,
// for MUX (which = 0 / 1 selects the two halves), :342-353.
void orc_gate_linear(int32_t op, int32_t which, const int32_t* a, const int32_t* b, const int32_t* c, int32_t n,
                     int32_t* out) {
    const int32_t e8 = orc_modswitch_to_torus32(1, 8), e4 = orc_modswitch_to_torus32(1, 4);
    int32_t k0 = 0, ca = 0, cb = 0, cc = 0;
    switch (op) {
        case ORC_NAND: k0 = e8; ca = -1; cb = -1; break;
        case ORC_AND: k0 = -e8; ca = 1; cb = 1; break;
        case ORC_OR: k0 = e8; ca = 1; cb = 1; break;
        case ORC_XOR: k0 = e4; ca = 2; cb = 2; break;
        case ORC_XNOR: k0 = -e4; ca = -2; cb = -2; break;
        case ORC_NOR: k0 = -e8; ca = -1; cb = -1; break;
        case ORC_ANDNY: k0 = -e8; ca = -1; cb = 1; break;
        case ORC_ANDYN: k0 = -e8; ca = 1; cb = -1; break;
        case ORC_ORNY: k0 = e8; ca = -1; cb = 1; break;
        case ORC_ORYN: k0 = e8; ca = 1; cb = -1; break;
        case ORC_MUX:
            k0 = -e8;
            if (which == 0) { ca = 1; cb = 1; } else { ca = -1; cc = 1; }
            break;
        default: break;
    }
    for (int32_t i = 0; i <= n; i++) {
        uint32_t v = (i == n) ? U(k0) : 0u;
        v += U(ca) * U(a[i]) + U(cb) * U(b[i]);
        if (cc) v += U(cc) * U(c[i]);
        out[i] = S(v);
    }
}

// libtfhe/tgsw-functions.cpp:297-409 with constants from libtfhe/tgsw.cpp:7-28:
// truncating gadget decomposition, digits in [-Bg/2, Bg/2).
void orc_decompose(const orc_params* P, const int32_t* poly, int32_t* digits) {
    const int32_t N = P->N, l = P->l, Bgbit = P->Bgbit;
    const uint32_t Bg = 1u << Bgbit, half = Bg / 2, mask = Bg - 1;
    uint32_t offset = 0;
    for (int32_t i = 0; i < l; i++) offset += 1u << (32 - (i + 1) * Bgbit);
    offset *= half;
    for (int32_t j = 0; j < N; j++) {
        uint32_t u = U(poly[j]) + offset;
        for (int32_t p = 0; p < l; p++) {
            uint32_t t = (u >> (32 - (p + 1) * Bgbit)) & mask;
            digits[p * N + j] = S(t - half);
        }
    }
}

// libtfhe/toruspolynomial-functions.cpp:140-160 (out = X^a * in, 0 <= a < 2N)
void orc_mul_xai(int32_t a, const int32_t* in, int32_t N, int32_t* out) {
    for (int32_t j = 0; j < N; j++) {
        int32_t src = j - a;
        uint32_t sign = 0;
        while (src < 0) { src += N; sign ^= 1; }
        out[j] = sign ? S(0u - U(in[src])) : in[src];
    }
}

// libtfhe/toruspolynomial-functions.cpp:117-136 (out = (X^a - 1) * in)
void orc_mul_xai_minus_one(int32_t a, const int32_t* in, int32_t N, int32_t* out) {
    std::vector<int32_t> t(N);
    orc_mul_xai(a, in, N, t.data());
    for (int32_t j = 0; j < N; j++) out[j] = S(U(t[j]) - U(in[j]));
}

// Exact negacyclic product mod (X^N+1, 2^32): acc += ipoly * tpoly.
// Same result as libtfhe/multiplication.cpp:53-77 (torusPolynomialMultNaive).
void orc_negacyclic_mul_add(const int32_t* ipoly, const int32_t* tpoly, int32_t N, int32_t* acc) {
    uint32_t* r = (uint32_t*)acc;
    const uint32_t* b = (const uint32_t*)tpoly;
    for (int32_t t = 0; t < N; t++) {
        const uint32_t d = U(ipoly[t]);
        if (d == 0) continue;
        for (int32_t j = 0; j < t; j++) r[j] -= d * b[j - t + N];
        for (int32_t j = t; j < N; j++) r[j] += d * b[j - t];
    }
}

// External product, libtfhe/tgsw-fft-operations.cpp:89-114: decompose the k+1
// polynomials (row r = i*l + p), then acc_c = sum_r digit_r * bk_row[r][c].
void orc_extprod(const orc_params* P, const int32_t* bk_row, int32_t* acc) {
    const int32_t N = P->N, k = P->k, l = P->l, kpl = (k + 1) * l;
    std::vector<int32_t> dig((size_t)kpl * N);
    for (int32_t i = 0; i <= k; i++) orc_decompose(P, acc + (size_t)i * N, dig.data() + (size_t)i * l * N);
    std::vector<int32_t> out((size_t)(k + 1) * N, 0);
    for (int32_t r = 0; r < kpl; r++)
        for (int32_t c = 0; c <= k; c++)
            orc_negacyclic_mul_add(dig.data() + (size_t)r * N, bk_row + ((size_t)r * (k + 1) + c) * N, N,
                                   out.data() + (size_t)c * N);
    memcpy(acc, out.data(), sizeof(int32_t) * (k + 1) * N);
}

// libtfhe/lwe-bootstrapping-functions-fft.cpp:67-112
void orc_blind_rotate(const orc_params* P, const int32_t* bk, const int32_t* bara, int32_t* acc) {
    const int32_t N = P->N, k = P->k, kpl = (k + 1) * P->l;
    const size_t row = (size_t)kpl * (k + 1) * N;
    std::vector<int32_t> t((size_t)(k + 1) * N);
    for (int32_t i = 0; i < P->n; i++) {
        if (bara[i] == 0) continue;
        for (int32_t c = 0; c <= k; c++) orc_mul_xai_minus_one(bara[i], acc + (size_t)c * N, N, t.data() + (size_t)c * N);
        orc_extprod(P, bk + row * i, t.data());
        for (size_t j = 0; j < (size_t)(k + 1) * N; j++) acc[j] = S(U(acc[j]) + U(t[j]));
    }
}

// libtfhe/lwe-bootstrapping-functions-fft.cpp:127-156 and libtfhe/lwe.cpp:41-61
void orc_blind_rotate_and_extract(const orc_params* P, const int32_t* bk, const int32_t* v, int32_t barb,
                                  const int32_t* bara, int32_t* out) {
    const int32_t N = P->N, k = P->k;
    std::vector<int32_t> acc((size_t)(k + 1) * N, 0);
    if (barb != 0) orc_mul_xai(2 * N - barb, v, N, acc.data() + (size_t)k * N);
    else memcpy(acc.data() + (size_t)k * N, v, sizeof(int32_t) * N);
    orc_blind_rotate(P, bk, bara, acc.data());
    for (int32_t i = 0; i < k; i++) {
        const int32_t* a = acc.data() + (size_t)i * N;
        out[i * N] = a[0];
        for (int32_t j = 1; j < N; j++) out[i * N + j] = S(0u - U(a[N - j]));
    }
    out[k * N] = acc[(size_t)k * N];
}

// libtfhe/lwe-bootstrapping-functions-fft.cpp:169-200
void orc_bootstrap_woks(const orc_params* P, const int32_t* bk, int32_t mu, const int32_t* x, int32_t* out) {
    const int32_t N = P->N, n = P->n;
    std::vector<int32_t> bara(n), v(N, mu);
    int32_t barb = orc_modswitch_from_torus32(x[n], 2 * N);
    for (int32_t i = 0; i < n; i++) bara[i] = orc_modswitch_from_torus32(x[i], 2 * N);
    orc_blind_rotate_and_extract(P, bk, v.data(), barb, bara.data(), out);
}

// libtfhe/lwe-keyswitch-functions.cpp:114-129,228-238; ksk row (i,j,h) at ((i*t+j)*base+h)*(n+1)
void orc_keyswitch(const orc_params* P, const int32_t* ksk, const int32_t* in, int32_t* out) {
    const int32_t n = P->n, nin = P->k * P->N, t = P->ks_t, basebit = P->ks_basebit, base = 1 << basebit;
    const uint32_t prec_offset = 1u << (32 - (1 + basebit * t));
    const uint32_t mask = (uint32_t)base - 1;
    std::vector<uint32_t> r((size_t)n + 1, 0);
    r[n] = U(in[nin]);
    for (int32_t i = 0; i < nin; i++) {
        const uint32_t abar = U(in[i]) + prec_offset;
        for (int32_t j = 0; j < t; j++) {
            const uint32_t d = (abar >> (32 - (j + 1) * basebit)) & mask;
            if (d == 0) continue;
            const int32_t* row = ksk + ((size_t)(i * t + j) * base + d) * (n + 1);
            for (int32_t c = 0; c <= n; c++) r[c] -= U(row[c]);
        }
    }
    for (int32_t c = 0; c <= n; c++) out[c] = S(r[c]);
}

// libtfhe/boot-gates.cpp (bootstrap with MU = 1/8, then keyswitch; MUX per :329-370)
void orc_gate(const orc_params* P, const int32_t* bk, const int32_t* ksk, int32_t op, const int32_t* a,
              const int32_t* b, const int32_t* c, int32_t* out) {
    const int32_t n = P->n, nx = P->k * P->N;
    const int32_t mu = orc_modswitch_to_torus32(1, 8);
    std::vector<int32_t> lin((size_t)n + 1), u1((size_t)nx + 1), u2((size_t)nx + 1);
    orc_gate_linear(op, 0, a, b, c, n, lin.data());
    orc_bootstrap_woks(P, bk, mu, lin.data(), u1.data());
    if (op == ORC_MUX) {
        orc_gate_linear(op, 1, a, b, c, n, lin.data());
        orc_bootstrap_woks(P, bk, mu, lin.data(), u2.data());
        for (int32_t i = 0; i <= nx; i++) u1[i] = S(U(u1[i]) + U(u2[i]) + (i == nx ? U(mu) : 0u));
    }
    orc_keyswitch(P, ksk, u1.data(), out);
}

void orc_gates(const orc_params* P, const int32_t* bk, const int32_t* ksk, int32_t count, const int32_t* ops,
               const int32_t* in0, const int32_t* in1, const int32_t* in2, int32_t* out, int32_t threads) {
    const size_t w = (size_t)P->n + 1;
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    for (int32_t t = 0; t < threads; t++)
        th.emplace_back([=]() {
            for (int32_t g = t; g < count; g += threads)
                orc_gate(P, bk, ksk, ops[g], in0 + w * g, in1 + w * g, in2 ? in2 + w * g : in1 + w * g, out + w * g);
        });
    for (auto& x : th) x.join();
}

// Key generation in the reference's RNG order:
// libtfhe/tfhe_gate_bootstrapping.cpp:97-109 -> lweKeyGen (lwe-functions.cpp:21-27),
// tGswKeyGen (tlwe-functions.cpp:15-23), lweCreateKeySwitchKey
// (lwe-keyswitch-functions.cpp:163-215), then per i tGswSymEncryptInt
// (tgsw-functions.cpp:186-195 -> tlwe-functions.cpp:26-39 and tgsw-functions.cpp:114-123).
// The BK b polynomial is computed with the exact product (the reference uses
// its FFT product here, which truncates: expect |difference| <= 1 unit).
void orc_keygen(const orc_params* P, int32_t* lwe_key, int32_t* tlwe_key, int32_t* ksk, int32_t* bk) {
    const int32_t n = P->n, N = P->N, k = P->k, l = P->l, Bgbit = P->Bgbit;
    const int32_t t = P->ks_t, basebit = P->ks_basebit, base = 1 << basebit, nin = k * N;
    for (int32_t i = 0; i < n; i++) lwe_key[i] = key_bit();
    for (int32_t i = 0; i < k * N; i++) tlwe_key[i] = key_bit();
    // keyswitch key: recentred gaussian noises first, then masks row by row
    const int32_t sizeks = nin * t * (base - 1);
    std::vector<double> noise(sizeks);
    double err = 0;
    for (int32_t i = 0; i < sizeks; i++) {
        std::normal_distribution<double> d(0., P->lwe_stdev);
        noise[i] = d(g_gen);
        err += noise[i];
    }
    err = err / sizeks;
    for (int32_t i = 0; i < sizeks; i++) noise[i] -= err;
    int32_t index = 0;
    for (int32_t i = 0; i < nin; i++)
        for (int32_t j = 0; j < t; j++) {
            int32_t* row0 = ksk + ((size_t)(i * t + j) * base) * (n + 1);
            memset(row0, 0, sizeof(int32_t) * (n + 1));
            for (int32_t h = 1; h < base; h++) {
                int32_t* row = ksk + ((size_t)(i * t + j) * base + h) * (n + 1);
                uint32_t mess = (U(tlwe_key[i]) * (uint32_t)h) * (1u << (32 - (j + 1) * basebit));
                uint32_t b = mess + U(orc_dtot32(noise[index]));
                for (int32_t p = 0; p < n; p++) {
                    row[p] = uniform_torus();
                    b += U(row[p]) * U(lwe_key[p]);
                }
                row[n] = S(b);
                index++;
            }
        }
    // bootstrapping key: TGSW encryptions of the LWE key bits
    const int32_t kpl = (k + 1) * l;
    for (int32_t i = 0; i < n; i++)
        for (int32_t r = 0; r < kpl; r++) {
            int32_t* samp = bk + ((size_t)i * kpl + r) * (k + 1) * N;  // [k+1][N]
            int32_t* bpoly = samp + (size_t)k * N;
            for (int32_t j = 0; j < N; j++) bpoly[j] = gaussian32(0, P->bk_stdev);
            for (int32_t c = 0; c < k; c++) {
                int32_t* apoly = samp + (size_t)c * N;
                for (int32_t j = 0; j < N; j++) apoly[j] = uniform_torus();
                orc_negacyclic_mul_add(tlwe_key + (size_t)c * N, apoly, N, bpoly);
            }
            // + mu * h on coefficient 0 of component blk of row (blk, p)
            const int32_t blk = r / l, p = r % l;
            samp[(size_t)blk * N] = S(U(samp[(size_t)blk * N]) + U(lwe_key[i]) * (1u << (32 - (p + 1) * Bgbit)));
        }
}

uint64_t orc_fnv1a(const void* data, uint64_t bytes) {
    const uint8_t* p = (const uint8_t*)data;
    uint64_t h = 1469598103934665603ULL;
    for (uint64_t i = 0; i < bytes; i++) {
        h ^= p[i];
        h *= 1099511628211ULL;
    }
    return h;
}

}  // extern "C"
