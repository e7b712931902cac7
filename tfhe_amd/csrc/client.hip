// Client-side batch encryption and decryption on the device.
//
// Decryption restates bootsSymDecrypt (libtfhe/tfhe_gate_bootstrapping.cpp:162-165):
// phase = b - sum a_i s_i (lwePhase, libtfhe/lwe-functions.cpp:72-81), bit =
// phase > 0. Encryption restates bootsSymEncrypt (tfhe_gate_bootstrapping.cpp:
// 154-159) over lweSymEncrypt (lwe-functions.cpp:36-46): mu = +-1/8, b = mu +
// dtot32(e) + sum a_i s_i with a_i uniform on the torus and e ~ N(0, alpha_min)
// (gaussian32 / dtot32, libtfhe/numeric-functions.cpp:21-34). The randomness
// comes from Philox4x32-10 instead of the reference's global sequential
// generator, so device ciphertexts are reproducible from (seed, index) but are
// not the reference's ciphertexts for the same seed.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "engine.h"
#include "philox.h"

namespace tfhe_amd {
namespace {

// Rows of w = n + 1 int32 (a[0..n), b). Each wavefront keeps its slice of the
// key in registers (key[lane + 64 k], n <= 1024) and walks the rows four at a
// time (grid-stride), so 4 x ceil(n/64) row loads are in flight per step.
constexpr int DECRYPT_KS = 16;
__global__ void __launch_bounds__(256) decrypt_kernel(const int32_t* __restrict__ cts, int count, int n,
                                                      const int32_t* __restrict__ key, int32_t* __restrict__ bits,
                                                      int32_t* __restrict__ phases) {
    const int lane = threadIdx.x & 63;
    uint32_t kr[DECRYPT_KS];
#pragma unroll
    for (int k = 0; k < DECRYPT_KS; k++) {
        const int i = lane + 64 * k;
        kr[k] = i < n ? (uint32_t)key[i] : 0u;
    }
    const int stride = gridDim.x * 4 * 4;
    for (int g0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 4; g0 < count; g0 += stride) {
        uint32_t sum[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            sum[c] = 0;
            if (g0 + c < count) {
                const int32_t* a = cts + (size_t)(g0 + c) * (n + 1);
#pragma unroll
                for (int k = 0; k < DECRYPT_KS; k++) {
                    const int i = lane + 64 * k;
                    if (i < n) sum[c] += (uint32_t)a[i] * kr[k];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
            for (int d = 32; d >= 1; d >>= 1) sum[c] += (uint32_t)__shfl_xor((int)sum[c], d, 64);
        if (lane < 4 && g0 + lane < count) {
            const int g = g0 + lane;
            const uint32_t s = lane == 0 ? sum[0] : lane == 1 ? sum[1] : lane == 2 ? sum[2] : sum[3];
            const int32_t phase = (int32_t)((uint32_t)cts[(size_t)g * (n + 1) + n] - s);
            if (bits) bits[g] = phase > 0 ? 1 : 0;
            if (phases) phases[g] = phase;
        }
    }
}

// One wavefront per ciphertext. Ciphertext index q = first + g draws its mask
// from Philox blocks (q, j, 0, 0), j = 0 .. ceil(n/4)-1 (one 32-bit word per
// coefficient), and its noise from block (q, 0, 1, 0): two 53-bit uniforms
// for the Box-Muller transform.
__global__ void __launch_bounds__(256) encrypt_kernel(const int32_t* __restrict__ bits, int count, int n,
                                                      const int32_t* __restrict__ key, double alpha, uint32_t k0,
                                                      uint32_t k1, uint64_t first, int32_t mu,
                                                      int32_t* __restrict__ cts) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= count) return;
    const uint64_t q = first + (uint64_t)g;
    const uint32_t q0 = (uint32_t)q, q1 = (uint32_t)(q >> 32);
    int32_t* out = cts + (size_t)g * (n + 1);
    uint32_t s = 0;
    for (int j = lane; 4 * j < n; j += 64) {
        const Philox4 r = philox4x32_10(Philox4{{q0, (uint32_t)j, q1, 0u}}, k0, k1);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int i = 4 * j + t;
            if (i < n) {
                out[i] = (int32_t)r.v[t];
                s += r.v[t] * (uint32_t)key[i];
            }
        }
    }
    for (int d = 32; d >= 1; d >>= 1) s += (uint32_t)__shfl_xor((int)s, d, 64);
    if (lane == 0) {
        const Philox4 r = philox4x32_10(Philox4{{q0, 0u, q1, 1u}}, k0, k1);
        // uniforms in (0, 1] and [0, 1) with 53 bits each
        const double u1 = ((double)((((uint64_t)r.v[0] << 32) | r.v[1]) >> 11) + 1.0) * 0x1p-53;
        const double u2 = (double)((((uint64_t)r.v[2] << 32) | r.v[3]) >> 11) * 0x1p-53;
        const double e = alpha * sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
        // dtot32: the fractional part of e scaled to 2^32, truncated toward zero
        const int32_t et = (int32_t)(int64_t)((e - (double)(int64_t)e) * 4294967296.0);
        const int32_t m = bits[g] ? mu : -mu;
        out[n] = (int32_t)((uint32_t)m + (uint32_t)et + s);
    }
}

// ------------------------------------------------------- key generation --
//
// Keys drawn on the device from Philox4x32-10 keyed by a 64-bit seed; value
// `idx` of stream `st` is word 0 of block (idx lo, idx hi, st, 0), a Gaussian
// of stream `st` comes from the Box-Muller transform of the whole block. The
// procedure follows the reference's key generation
// (libtfhe/tfhe_gate_bootstrapping.cpp:97-109): LWE and TLWE secret bits, the
// keyswitch key rows ks[i][j][h] = LWE_s(h * t_i / base^(j+1)) for h >= 1
// (libtfhe/lwe-keyswitch-functions.cpp:163-215) and the TGSW encryptions of the
// LWE key bits (libtfhe/tgsw-functions.cpp:114-140), with the exact negacyclic
// product for the TLWE b polynomials. The reference's generator order is not
// reproduced (new_random_gate_bootstrapping_secret_keyset does that on the host).

__device__ inline uint32_t rnd32(uint64_t idx, uint32_t st, uint32_t k0, uint32_t k1) {
    return philox4x32_10(Philox4{{(uint32_t)idx, (uint32_t)(idx >> 32), st, 0u}}, k0, k1).v[0];
}

__device__ inline int32_t gauss_torus(uint64_t idx, uint32_t st, double sigma, uint32_t k0, uint32_t k1) {
    const Philox4 r = philox4x32_10(Philox4{{(uint32_t)idx, (uint32_t)(idx >> 32), st, 1u}}, k0, k1);
    const double u1 = ((double)((((uint64_t)r.v[0] << 32) | r.v[1]) >> 11) + 1.0) * 0x1p-53;
    const double u2 = (double)((((uint64_t)r.v[2] << 32) | r.v[3]) >> 11) * 0x1p-53;
    const double e = sigma * sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    return (int32_t)(int64_t)((e - (double)(int64_t)e) * 4294967296.0);  // dtot32
}

constexpr uint32_t ST_LWE_KEY = 1, ST_TLWE_KEY = 2, ST_KS_MASK = 3, ST_KS_NOISE = 4, ST_BK_MASK = 5, ST_BK_NOISE = 6;

__global__ void __launch_bounds__(256) keygen_bits_kernel(int n, int nk, uint32_t k0, uint32_t k1,
                                                          int32_t* __restrict__ lwe, int32_t* __restrict__ tlwe) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) lwe[idx] = (int32_t)(rnd32(idx, ST_LWE_KEY, k0, k1) & 1u);
    if (idx < nk) tlwe[idx] = (int32_t)(rnd32(idx, ST_TLWE_KEY, k0, k1) & 1u);
}

// one wavefront per keyswitch row (i, j, h), layout [nin][t][base][n+1]; h = 0 rows are zero
__global__ void __launch_bounds__(256) keygen_ksk_kernel(int nin, int t, int basebit, int n, double alpha,
                                                         const int32_t* __restrict__ lwe,
                                                         const int32_t* __restrict__ tlwe, uint32_t k0, uint32_t k1,
                                                         int32_t* __restrict__ ksk) {
    const int lane = threadIdx.x & 63;
    const int base = 1 << basebit;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= (int64_t)nin * t * base) return;
    const int h = (int)(row % base), j = (int)((row / base) % t), i = (int)(row / ((int64_t)base * t));
    int32_t* out = ksk + row * (n + 1);
    if (h == 0) {
        for (int p = lane; p <= n; p += 64) out[p] = 0;
        return;
    }
    uint32_t s = 0;
    for (int p = lane; p < n; p += 64) {
        const uint32_t a = rnd32((uint64_t)row * n + p, ST_KS_MASK, k0, k1);
        out[p] = (int32_t)a;
        s += a * (uint32_t)lwe[p];
    }
    for (int d = 32; d >= 1; d >>= 1) s += (uint32_t)__shfl_xor((int)s, d, 64);
    if (lane == 0) {
        const uint32_t msg = ((uint32_t)tlwe[i] * (uint32_t)h) << (32 - (j + 1) * basebit);
        out[n] = (int32_t)(msg + (uint32_t)gauss_torus((uint64_t)row, ST_KS_NOISE, alpha, k0, k1) + s);
    }
}

// one workgroup per TGSW row (i, r): masks uniform, b = sum_c a_c * s_c + e
// (exact negacyclic products with the binary TLWE key), then mu * h_p added to
// coefficient 0 of component r / l (libtfhe/tgsw-functions.cpp:114-140).
__global__ void __launch_bounds__(1024) keygen_bk_kernel(int N, int k, int l, int Bgbit, double alpha,
                                                         const int32_t* __restrict__ lwe,
                                                         const int32_t* __restrict__ tlwe, uint32_t k0, uint32_t k1,
                                                         int32_t* __restrict__ bk) {
    extern __shared__ int32_t sm[];  // a[k][N], s[k][N]
    const int kpl = (k + 1) * l;
    const int64_t row = blockIdx.x;  // i * kpl + r
    const int i = (int)(row / kpl), r = (int)(row % kpl);
    int32_t* poly = bk + row * (k + 1) * N;
    for (int q = threadIdx.x; q < k * N; q += blockDim.x) {
        const int32_t a = (int32_t)rnd32((uint64_t)row * k * N + q, ST_BK_MASK, k0, k1);
        sm[q] = a;
        sm[k * N + q] = tlwe[q];
        poly[q] = a;
    }
    __syncthreads();
    for (int jj = threadIdx.x; jj < N; jj += blockDim.x) {
        uint32_t b = (uint32_t)gauss_torus((uint64_t)row * N + jj, ST_BK_NOISE, alpha, k0, k1);
        for (int c = 0; c < k; c++) {
            const int32_t* a = sm + c * N;
            const int32_t* sc = sm + k * N + c * N;
            // (a * s)_jj = sum_{t <= jj} s_t a_{jj-t} - sum_{t > jj} s_t a_{jj-t+N}
            for (int tt = 0; tt < N; tt++) {
                const uint32_t prod = (uint32_t)sc[tt] * (uint32_t)a[(jj - tt) & (N - 1)];
                b += (tt <= jj) ? prod : 0u - prod;
            }
        }
        poly[k * N + jj] = (int32_t)b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int blk = r / l, p = r % l;
        int32_t* c0 = poly + (size_t)blk * N;
        c0[0] = (int32_t)((uint32_t)c0[0] + (uint32_t)lwe[i] * (1u << (32 - (p + 1) * Bgbit)));
    }
}

// Per LWE secret key: its device copy and the first ciphertext index that no
// encryption under this key has used yet.
struct DevKey {
    int32_t* d = nullptr;
    uint64_t next_index = 0;
};

struct KeyCache {
    std::mutex mu;
    std::map<const LweKey*, DevKey> dev;
};

KeyCache& key_cache() {
    static KeyCache c;
    return c;
}

// Device copy of the key. With `index` non-null it also reserves `count`
// ciphertext indices: *index = TFHE_AMD_NEXT_INDEX takes the next unused ones;
// an explicit *index is kept, and the key's next unused index moves past it.
const int32_t* device_lwe_key(const LweKey* k, int n, uint64_t* index = nullptr, uint64_t count = 0) {
    KeyCache& C = key_cache();
    std::lock_guard<std::mutex> g(C.mu);
    auto it = C.dev.find(k);
    if (it == C.dev.end()) {
        int32_t* d = nullptr;
        if (hipMalloc(&d, sizeof(int32_t) * n) != hipSuccess ||
            hipMemcpy(d, k->key, sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d);
            set_error("device copy of the LWE secret key failed");
            return nullptr;
        }
        it = C.dev.emplace(k, DevKey{d, 0}).first;
    }
    DevKey& e = it->second;
    if (index) {
        if (*index == TFHE_AMD_NEXT_INDEX) *index = e.next_index;
        if (*index > TFHE_AMD_NEXT_INDEX - 1 - count) {
            set_error("encrypt_device: ciphertext index range overflows");
            return nullptr;
        }
        e.next_index = std::max(e.next_index, *index + count);
    }
    return e.d;
}

}  // namespace

void engine_forget_secret_key(const LweKey* k) {
    KeyCache& C = key_cache();
    std::lock_guard<std::mutex> g(C.mu);
    auto it = C.dev.find(k);
    if (it == C.dev.end()) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(it->second.d);
    C.dev.erase(it);
}

}  // namespace tfhe_amd

using namespace tfhe_amd;

TFHE_AMD_EXPORT int32_t tfhe_amd_encrypt_device(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                                const int32_t* d_bits, int32_t* d_cts, uint64_t seed,
                                                uint64_t first_index, void* stream) {
    if (!key || !key->lwe_key || !key->params || count < 0 || (count && (!d_bits || !d_cts)))
        return set_error("encrypt_device: bad arguments");
    if (!count) return 0;
    const int n = key->lwe_key->params->n;
    const int32_t* dkey = device_lwe_key(key->lwe_key, n, &first_index, (uint64_t)count);
    if (!dkey) return -1;
    const int32_t mu = (int32_t)(1u << 29);  // modSwitchToTorus32(1, 8)
    const double alpha = key->params->in_out_params->alpha_min;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)engine_stream();
    hipLaunchKernelGGL(encrypt_kernel, dim3((count + 3) / 4), dim3(256), 0, s, d_bits, count, n, dkey, alpha,
                       (uint32_t)seed, (uint32_t)(seed >> 32), first_index, mu, d_cts);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : set_error("encrypt_device: launch failed: %s", hipGetErrorString(e));
}

TFHE_AMD_EXPORT int32_t tfhe_amd_decrypt_device(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                                const int32_t* d_cts, int32_t* d_bits, int32_t* d_phases,
                                                void* stream) {
    if (!key || !key->lwe_key || count < 0 || (count && (!d_cts || (!d_bits && !d_phases))))
        return set_error("decrypt_device: bad arguments");
    if (!count) return 0;
    const int n = key->lwe_key->params->n;
    const int32_t* dkey = device_lwe_key(key->lwe_key, n);
    if (!dkey) return -1;
    hipStream_t s = stream ? (hipStream_t)stream : (hipStream_t)engine_stream();
    if (n > 64 * DECRYPT_KS) return set_error("decrypt_device: n = %d above %d", n, 64 * DECRYPT_KS);
    // up to 16 workgroups (64 wavefronts) per CU's worth of rows, at least one
    const int blocks = std::max(1, std::min((count + 15) / 16, 256 * 16));
    hipLaunchKernelGGL(decrypt_kernel, dim3(blocks), dim3(256), 0, s, d_cts, count, n, dkey, d_bits, d_phases);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : set_error("decrypt_device: launch failed: %s", hipGetErrorString(e));
}

TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* tfhe_amd_keygen_device(const TFheGateBootstrappingParameterSet* P,
                                                                        uint64_t seed) {
    if (!P) {
        set_error("keygen_device: null parameters");
        return nullptr;
    }
    const TGswParams* gp = P->tgsw_params;
    const TLweParams* tp = gp->tlwe_params;
    const int n = P->in_out_params->n, N = tp->N, k = tp->k, l = gp->l, kpl = gp->kpl;
    const int t = P->ks_t, basebit = P->ks_basebit, base = 1 << basebit, nin = k * N;
    const size_t ksk_words = (size_t)nin * t * base * (n + 1), bk_words = (size_t)n * kpl * (k + 1) * N;
    hipStream_t s = (hipStream_t)engine_stream();
    if (!s) return nullptr;
    int32_t *d_lwe = nullptr, *d_tlwe = nullptr, *d_ksk = nullptr, *d_bk = nullptr;
    std::vector<int32_t> lwe(n), tlwe(nin), ksk(ksk_words), bk(bk_words);
    auto done = [&](bool ok) {
        (void)hipFree(d_lwe);
        (void)hipFree(d_tlwe);
        (void)hipFree(d_ksk);
        (void)hipFree(d_bk);
        return ok;
    };
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    bool ok = hipMalloc(&d_lwe, sizeof(int32_t) * n) == hipSuccess &&
              hipMalloc(&d_tlwe, sizeof(int32_t) * nin) == hipSuccess &&
              hipMalloc(&d_ksk, sizeof(int32_t) * ksk_words) == hipSuccess &&
              hipMalloc(&d_bk, sizeof(int32_t) * bk_words) == hipSuccess;
    if (!ok) {
        done(false);
        set_error("keygen_device: device allocation failed");
        return nullptr;
    }
    const int m = n > nin ? n : nin;
    hipLaunchKernelGGL(keygen_bits_kernel, dim3((m + 255) / 256), dim3(256), 0, s, n, nin, k0, k1, d_lwe, d_tlwe);
    const int64_t rows = (int64_t)nin * t * base;
    hipLaunchKernelGGL(keygen_ksk_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, nin, t, basebit, n,
                       P->in_out_params->alpha_min, d_lwe, d_tlwe, k0, k1, d_ksk);
    hipLaunchKernelGGL(keygen_bk_kernel, dim3((unsigned)(n * kpl)), dim3(1024), sizeof(int32_t) * 2 * k * N, s, N, k, l,
                       gp->Bgbit, tp->alpha_min, d_lwe, d_tlwe, k0, k1, d_bk);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(lwe.data(), d_lwe, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(tlwe.data(), d_tlwe, sizeof(int32_t) * nin, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(ksk.data(), d_ksk, sizeof(int32_t) * ksk_words, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipMemcpyAsync(bk.data(), d_bk, sizeof(int32_t) * bk_words, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess;
    done(ok);
    if (!ok) {
        set_error("keygen_device: key generation failed on the device");
        return nullptr;
    }
    const double ks_var = P->in_out_params->alpha_min * P->in_out_params->alpha_min;
    const double bk_var = tp->alpha_min * tp->alpha_min;
    return host_secret_from_arrays(P, bk.data(), ksk.data(), ks_var, bk_var, lwe.data(), tlwe.data());
}
