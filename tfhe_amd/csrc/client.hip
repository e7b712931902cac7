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

#include <cmath>
#include <map>
#include <mutex>

#include "engine.h"
#include "philox.h"

namespace tfhe_amd {
namespace {

// One wavefront per ciphertext; rows of w = n + 1 int32 (a[0..n), b).
__global__ void __launch_bounds__(256) decrypt_kernel(const int32_t* __restrict__ cts, int count, int n,
                                                      const int32_t* __restrict__ key, int32_t* __restrict__ bits,
                                                      int32_t* __restrict__ phases) {
    const int lane = threadIdx.x & 63;
    const int g = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (g >= count) return;
    const int32_t* a = cts + (size_t)g * (n + 1);
    uint32_t s = 0;
    for (int i = lane; i < n; i += 64) s += (uint32_t)a[i] * (uint32_t)key[i];
    for (int d = 32; d >= 1; d >>= 1) s += (uint32_t)__shfl_xor((int)s, d, 64);
    if (lane == 0) {
        const int32_t phase = (int32_t)((uint32_t)a[n] - s);
        if (bits) bits[g] = phase > 0 ? 1 : 0;
        if (phases) phases[g] = phase;
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

struct KeyCache {
    std::mutex mu;
    std::map<const LweKey*, int32_t*> dev;
};

KeyCache& key_cache() {
    static KeyCache c;
    return c;
}

const int32_t* device_lwe_key(const LweKey* k, int n) {
    KeyCache& C = key_cache();
    std::lock_guard<std::mutex> g(C.mu);
    auto it = C.dev.find(k);
    if (it != C.dev.end()) return it->second;
    int32_t* d = nullptr;
    if (hipMalloc(&d, sizeof(int32_t) * n) != hipSuccess ||
        hipMemcpy(d, k->key, sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        set_error("device copy of the LWE secret key failed");
        return nullptr;
    }
    C.dev[k] = d;
    return d;
}

}  // namespace

void engine_forget_secret_key(const LweKey* k) {
    KeyCache& C = key_cache();
    std::lock_guard<std::mutex> g(C.mu);
    auto it = C.dev.find(k);
    if (it == C.dev.end()) return;
    (void)hipDeviceSynchronize();
    (void)hipFree(it->second);
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
    const int32_t* dkey = device_lwe_key(key->lwe_key, n);
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
    hipLaunchKernelGGL(decrypt_kernel, dim3((count + 3) / 4), dim3(256), 0, s, d_cts, count, n, dkey, d_bits,
                       d_phases);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : set_error("decrypt_device: launch failed: %s", hipGetErrorString(e));
}
