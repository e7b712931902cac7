// Internal interface between the host library (host.cpp) and the GPU engine
// (engine.hip). Not part of the public C-ABI.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "../../include/tfhe_amd.h"

namespace tfhe_amd {

// Gate codes beyond the public TFHE_AMD_* ones, used by the subsystem entry points.
constexpr int OP_IDENTITY = 11;  // bootstrap the input sample as is (tfhe_bootstrap_FFT)

struct HostParams {
    int32_t n, N, k, l, Bgbit, ks_t, ks_basebit;
};

HostParams host_params(const TFheGateBootstrappingParameterSet* p);
HostParams host_params_fft(const LweBootstrappingKeyFFT* bf);

// Records an error message (thread-local) and returns -1.
int set_error(const char* fmt, ...);
[[noreturn]] void die(const char* what);

// What the library keeps beside a transform-domain key it made (the public
// struct has the reference's layout): the coefficient-domain key it was built
// from and the handle of its device copy (null until uploaded). Dies on a
// pointer this library did not make.
const LweBootstrappingKey* bkfft_coef(const LweBootstrappingKeyFFT* bf);
tfhe_amd_device_key*& bkfft_dkey(const LweBootstrappingKeyFFT* bf);
// The transform-domain key whose bkFFT array holds `rows`, and the index of
// `rows` in it; nullptr if no key made by this library holds it.
const LweBootstrappingKeyFFT* bkfft_of_rows(const TGswSampleFFT* rows, int32_t* row);
// Device key of a transform-domain key; uploads on first use. Returns nullptr on error.
tfhe_amd_device_key* engine_key(const LweBootstrappingKeyFFT* bf);

// Gates on host ciphertexts (stages through pinned memory). in2 may be null.
// nz (with keyswitch, if not null): per gate, the number of key rows the keyswitch added.
int engine_gates_host(const LweBootstrappingKeyFFT* bf, int32_t count, const int32_t* ops,
                      LweSample* const* out, const LweSample* const* in0, const LweSample* const* in1,
                      const LweSample* const* in2, int32_t mu, bool keyswitch, int32_t* nz = nullptr);

// Subsystem entry points on host buffers.
// acc: (k+1)*N int32 in/out, bara: n entries, rows [row0, row0+n) of the key.
int engine_blind_rotate_host(tfhe_amd_device_key* dk, int32_t row0, int32_t n, const int32_t* bara, int32_t* acc,
                             int32_t* ext_out);
int engine_keyswitch_host(tfhe_amd_device_key* dk, const int32_t* in, int32_t* out, int32_t* nz = nullptr);

// Drops the host key's device copy. A copy pinned by a server stays allocated
// (and in use by that server) until its last unpin.
void engine_forget_host_key(const LweBootstrappingKeyFFT* bkfft);
// Identity of one upload (a re-uploaded key gets a new one), and pins that keep a
// device key alive after its host key releases it or is deleted.
uint64_t engine_key_gen(const tfhe_amd_device_key* dk);
tfhe_amd_device_key* engine_key_pinned(const LweBootstrappingKeyFFT* bf);
void engine_key_unpin(tfhe_amd_device_key* dk);
// Frees the device copy of an LWE secret key (client-side device encryption).
void engine_forget_secret_key(const LweKey* k);

// For the level scheduler: device scratch a batch of `count` gates needs, the
// engine lock, its stream, and a batch launch with caller-provided scratch
// (call with engine_mutex() held; `capturing` disables event timing; `shared`:
// other launches are running on the device, so automatic mode does not pick the
// latency kernel, which takes a whole CU per bootstrap).
size_t engine_batch_scratch_bytes(const tfhe_amd_device_key* dk, int count);
std::mutex& engine_mutex();
void* engine_stream();
int engine_launch_batch_locked(tfhe_amd_device_key* dk, int count, const int32_t* d_op, const int32_t* d_in0,
                               const int32_t* d_in1, const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                               char* scr, void* stream, bool capturing,
                               bool shared = false);
// Several clients' batches (segments) in one blind-rotation launch, each segment
// with its own key (same parameters, N = 1024), then each segment's keyswitch.
// d_idx holds [5][total] (op, in0, in1, in2, out) with segment g at
// [sum of earlier counts, + counts[g]). At most 32 segments; call with
// engine_mutex() held.
bool engine_can_segment(const tfhe_amd_device_key* dk);
size_t engine_segments_scratch_bytes(tfhe_amd_device_key* const* dks, const int32_t* counts, int nseg);
int engine_launch_segments_locked(tfhe_amd_device_key* const* dks, const int32_t* counts, int nseg,
                                  const int32_t* d_idx, int32_t* d_pool, char* scr, void* stream);
int engine_param_n(const tfhe_amd_device_key* dk);
// True when two device keys have the same parameters (ciphertext width, scratch sizes).
bool engine_same_params(const tfhe_amd_device_key* a, const tfhe_amd_device_key* b);
size_t engine_key_bytes(const tfhe_amd_device_key* dk);

// Host key-set constructors shared by the gate API and the key-file readers
// (host.cpp). Arrays: bootstrapping key [n][kpl][k+1][N], keyswitch key
// [k*N][t][base][n+1] (a, then b), every sample given the stated variance.
TFheGateBootstrappingParameterSet* host_make_params(int32_t n, double lwe_alpha_min, double lwe_alpha_max, int32_t N,
                                                    int32_t k, double tlwe_alpha_min, double tlwe_alpha_max,
                                                    int32_t bk_l, int32_t bk_Bgbit, int32_t ks_t, int32_t ks_basebit);
TFheGateBootstrappingCloudKeySet* host_cloud_from_arrays(const TFheGateBootstrappingParameterSet* P, const int32_t* bkc,
                                                         const int32_t* ksk, double ks_var, double bk_var);
TFheGateBootstrappingSecretKeySet* host_secret_from_arrays(const TFheGateBootstrappingParameterSet* P,
                                                           const int32_t* bkc, const int32_t* ksk, double ks_var,
                                                           double bk_var, const int32_t* lwe_key,
                                                           const int32_t* tlwe_key);

// Key sets around a coefficient-domain bootstrapping key already built (it and
// its keyswitch key become the key set's), for the key-file readers; lwe_key
// [n], tlwe_key [k N].
TFheGateBootstrappingCloudKeySet* host_cloud_from_bk(const TFheGateBootstrappingParameterSet* P,
                                                     LweBootstrappingKey* bk);
TFheGateBootstrappingSecretKeySet* host_secret_from_bk(const TFheGateBootstrappingParameterSet* P,
                                                       LweBootstrappingKey* bk, const int32_t* lwe_key,
                                                       const int32_t* tlwe_key);
// Parameters that live until the process ends (parameters read from files).
const LweParams* host_keep(LweParams* p);
const TLweParams* host_keep(TLweParams* p);
const TGswParams* host_keep(TGswParams* p);

// Levelisation of a gate netlist (schedule 0 = as soon as possible, 1 = as
// late as possible): gate order sorted by level and the gates per level.
// Validates gate codes, wire ranges, single drivers and acyclicity.
int engine_levelize(int32_t n_wires, int32_t n_gates, const int32_t* op, const int32_t* in0, const int32_t* in1,
                    const int32_t* in2, const int32_t* out, int schedule, std::vector<int32_t>& order,
                    std::vector<int32_t>& level_count);

}  // namespace tfhe_amd
