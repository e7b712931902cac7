/* tfhe_amd — MI355X gate-bootstrapping engine with the tfhe/tfhe C gate API.
 *
 * Drop-in boundary: the declarations below keep the names, argument meaning
 * and struct layouts of the reference's public headers (src/include/ in
 * tfhe/tfhe; each group cites the header it replaces), so C/C++ code written
 * against libtfhe compiles and links against libtfhe_amd.so unchanged. The
 * bootstrapping and keyswitch run on the GPU; key generation, encryption and
 * decryption stay on the host, as in the reference.
 *
 * Error behaviour: like the reference, the void gate API aborts the process
 * on a fatal error (after printing the reason to stderr, as die_dramatically
 * does in libtfhe/tfhe_gate_bootstrapping.cpp:11-15). The added batch and
 * device entry points return a status code instead (0 = success) and record
 * the reason for tfhe_amd_last_error().
 */
#ifndef TFHE_AMD_H
#define TFHE_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
#define TFHE_AMD_EXPORT extern "C" __attribute__((visibility("default")))
#else
#define TFHE_AMD_EXPORT __attribute__((visibility("default")))
#endif

/* ---- types: include/tfhe_core.h:37 and the struct headers ------------- */
typedef int32_t Torus32;

typedef struct LweParams LweParams;
typedef struct LweKey LweKey;
typedef struct LweSample LweSample;
typedef struct LweKeySwitchKey LweKeySwitchKey;
typedef struct TLweParams TLweParams;
typedef struct TLweKey TLweKey;
typedef struct TLweSample TLweSample;
typedef struct TLweSampleFFT TLweSampleFFT;
typedef struct TGswParams TGswParams;
typedef struct TGswKey TGswKey;
typedef struct TGswSample TGswSample;
typedef struct TGswSampleFFT TGswSampleFFT;
typedef struct LweBootstrappingKey LweBootstrappingKey;
typedef struct LweBootstrappingKeyFFT LweBootstrappingKeyFFT;
typedef struct IntPolynomial IntPolynomial;
typedef struct TorusPolynomial TorusPolynomial;
typedef struct TFheGateBootstrappingParameterSet TFheGateBootstrappingParameterSet;
typedef struct TFheGateBootstrappingCloudKeySet TFheGateBootstrappingCloudKeySet;
typedef struct TFheGateBootstrappingSecretKeySet TFheGateBootstrappingSecretKeySet;

/* opaque device-resident key (bootstrapping key in FP64 transform form + keyswitch key) */
typedef struct tfhe_amd_device_key tfhe_amd_device_key;

/* include/lweparams.h:13-27 */
struct LweParams { int32_t n; double alpha_min; double alpha_max; };
/* include/tlwe.h:10-28 */
struct TLweParams { int32_t N; int32_t k; double alpha_min; double alpha_max; LweParams extracted_lweparams; };
/* include/tgsw.h:10-32 */
struct TGswParams {
    int32_t l; int32_t Bgbit; int32_t Bg; int32_t halfBg; uint32_t maskMod;
    const TLweParams* tlwe_params; int32_t kpl; Torus32* h; uint32_t offset;
};
/* include/tfhe_gate_bootstrapping_structures.h:9-24 */
struct TFheGateBootstrappingParameterSet {
    int32_t ks_t; int32_t ks_basebit; const LweParams* in_out_params; const TGswParams* tgsw_params;
};
/* include/lwekey.h:10-20 */
struct LweKey { const LweParams* params; int32_t* key; };
/* include/lwesamples.h:9-20 */
struct LweSample { Torus32* a; Torus32 b; double current_variance; };
/* include/polynomials.h:10-34 */
struct IntPolynomial { int32_t N; int32_t* coefs; };
struct TorusPolynomial { int32_t N; Torus32* coefsT; };
/* include/tlwe.h:30-63 */
struct TLweKey { const TLweParams* params; IntPolynomial* key; };
struct TLweSample { TorusPolynomial* a; TorusPolynomial* b; double current_variance; int32_t k; };
/* include/tgsw.h:35-76 */
struct TGswKey { const TGswParams* params; const TLweParams* tlwe_params; IntPolynomial* key; TLweKey tlwe_key; };
struct TGswSample { TLweSample* all_sample; TLweSample** bloc_sample; int32_t k; int32_t l; };
/* include/tgsw.h:78-96, the same layout. There is no host transform-domain copy
 * in this library (all_samples/sample are NULL): the rows of a bootstrapping key
 * live on the device, and an array of them is recognised by its address (the
 * bkFFT array of a LweBootstrappingKeyFFT made by this library). */
struct TGswSampleFFT {
    TLweSampleFFT* all_samples; TLweSampleFFT** sample; int32_t k; int32_t l;
};
/* include/lwekeyswitch.h:11-28 (ks[i][j][h] = ks0_raw[(i*t+j)*base+h]) */
struct LweKeySwitchKey {
    int32_t n; int32_t t; int32_t basebit; int32_t base; const LweParams* out_params;
    LweSample* ks0_raw; LweSample** ks1_raw; LweSample*** ks;
};
/* include/lwebootstrappingkey.h:10-33 */
struct LweBootstrappingKey {
    const LweParams* in_out_params; const TGswParams* bk_params; const TLweParams* accum_params;
    const LweParams* extract_params; TGswSample* bk; LweKeySwitchKey* ks;
};
/* include/lwebootstrappingkey.h:36-59, the same layout. Made only by this
 * library (new_LweBootstrappingKeyFFT, the key-set constructors and readers),
 * which keeps beside it the coefficient-domain key it was built from and the
 * handle of its HBM-resident copy (uploaded on first use or by
 * tfhe_amd_cloud_key_upload). */
struct LweBootstrappingKeyFFT {
    const LweParams* in_out_params; const TGswParams* bk_params; const TLweParams* accum_params;
    const LweParams* extract_params; const TGswSampleFFT* bkFFT; const LweKeySwitchKey* ks;
};
/* include/tfhe_gate_bootstrapping_structures.h:26-62 */
struct TFheGateBootstrappingCloudKeySet {
    const TFheGateBootstrappingParameterSet* params; const LweBootstrappingKey* bk; const LweBootstrappingKeyFFT* bkFFT;
};
struct TFheGateBootstrappingSecretKeySet {
    const TFheGateBootstrappingParameterSet* params; const LweKey* lwe_key; const TGswKey* tgsw_key;
    TFheGateBootstrappingCloudKeySet cloud;
};

/* ---- include/tfhe_gate_bootstrapping_functions.h:14-87 ----------------- */
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_default_gate_bootstrapping_parameters(int32_t minimum_lambda);
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_random_gate_bootstrapping_secret_keyset(const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT void delete_gate_bootstrapping_parameters(TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT void delete_gate_bootstrapping_secret_keyset(TFheGateBootstrappingSecretKeySet* keyset);
TFHE_AMD_EXPORT void delete_gate_bootstrapping_cloud_keyset(TFheGateBootstrappingCloudKeySet* keyset);
TFHE_AMD_EXPORT LweSample* new_gate_bootstrapping_ciphertext(const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT LweSample* new_gate_bootstrapping_ciphertext_array(int32_t nbelems, const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT void delete_gate_bootstrapping_ciphertext(LweSample* sample);
TFHE_AMD_EXPORT void delete_gate_bootstrapping_ciphertext_array(int32_t nbelems, LweSample* samples);
TFHE_AMD_EXPORT void bootsSymEncrypt(LweSample* result, int32_t message, const TFheGateBootstrappingSecretKeySet* params);
TFHE_AMD_EXPORT int32_t bootsSymDecrypt(const LweSample* sample, const TFheGateBootstrappingSecretKeySet* params);
TFHE_AMD_EXPORT void bootsCONSTANT(LweSample* result, int32_t value, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsNAND(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsOR(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsAND(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsXOR(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsXNOR(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsNOT(LweSample* result, const LweSample* ca, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsCOPY(LweSample* result, const LweSample* ca, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsNOR(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsANDNY(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsANDYN(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsORNY(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsORYN(LweSample* result, const LweSample* ca, const LweSample* cb, const TFheGateBootstrappingCloudKeySet* bk);
TFHE_AMD_EXPORT void bootsMUX(LweSample* result, const LweSample* a, const LweSample* b, const LweSample* c, const TFheGateBootstrappingCloudKeySet* bk);

/* ---- include/tfhe.h:34-45 (seed + bootstrapping subsystem) ------------- */
TFHE_AMD_EXPORT void tfhe_random_generator_setSeed(uint32_t* values, int32_t size);
TFHE_AMD_EXPORT void tfhe_blindRotate_FFT(TLweSample* accum, const TGswSampleFFT* bk, const int32_t* bara, const int32_t n, const TGswParams* bk_params);
TFHE_AMD_EXPORT void tfhe_blindRotateAndExtract_FFT(LweSample* result, const TorusPolynomial* v, const TGswSampleFFT* bk, const int32_t barb, const int32_t* bara, const int32_t n, const TGswParams* bk_params);
TFHE_AMD_EXPORT void tfhe_bootstrap_woKS_FFT(LweSample* result, const LweBootstrappingKeyFFT* bk, Torus32 mu, const LweSample* x);
TFHE_AMD_EXPORT void tfhe_bootstrap_FFT(LweSample* result, const LweBootstrappingKeyFFT* bk, Torus32 mu, const LweSample* x);

/* ---- include/lwe-functions.h:17-79 (host LWE ops + device keyswitch) ---- */
TFHE_AMD_EXPORT void lweKeyGen(LweKey* result);
TFHE_AMD_EXPORT void lweSymEncrypt(LweSample* result, Torus32 message, double alpha, const LweKey* key);
TFHE_AMD_EXPORT Torus32 lwePhase(const LweSample* sample, const LweKey* key);
TFHE_AMD_EXPORT void lweNoiselessTrivial(LweSample* result, Torus32 mu, const LweParams* params);
TFHE_AMD_EXPORT void lweKeySwitch(LweSample* result, const LweKeySwitchKey* ks, const LweSample* sample);

/* ---- include/numeric_functions.h:24-52 ---------------------------------- */
TFHE_AMD_EXPORT Torus32 dtot32(double d);
TFHE_AMD_EXPORT double t32tod(Torus32 x);
TFHE_AMD_EXPORT int32_t modSwitchFromTorus32(Torus32 phase, int32_t Msize);
TFHE_AMD_EXPORT Torus32 modSwitchToTorus32(int32_t mu, int32_t Msize);

/* ---- allocators used by callers of the subsystem entry points
 *      (include/lwesamples.h, include/lweparams.h, include/tlwe.h, include/polynomials.h) */
TFHE_AMD_EXPORT LweParams* new_LweParams(int32_t n, double alpha_min, double alpha_max);
TFHE_AMD_EXPORT void delete_LweParams(LweParams* obj);
TFHE_AMD_EXPORT TLweParams* new_TLweParams(int32_t N, int32_t k, double alpha_min, double alpha_max);
TFHE_AMD_EXPORT void delete_TLweParams(TLweParams* obj);
TFHE_AMD_EXPORT TGswParams* new_TGswParams(int32_t l, int32_t Bgbit, const TLweParams* tlwe_params);
TFHE_AMD_EXPORT void delete_TGswParams(TGswParams* obj);
TFHE_AMD_EXPORT LweSample* new_LweSample(const LweParams* params);
TFHE_AMD_EXPORT LweSample* new_LweSample_array(int32_t nbelts, const LweParams* params);
TFHE_AMD_EXPORT void delete_LweSample(LweSample* obj);
TFHE_AMD_EXPORT void delete_LweSample_array(int32_t nbelts, LweSample* obj);
TFHE_AMD_EXPORT TLweSample* new_TLweSample(const TLweParams* params);
TFHE_AMD_EXPORT void delete_TLweSample(TLweSample* obj);
TFHE_AMD_EXPORT TorusPolynomial* new_TorusPolynomial(int32_t N);
TFHE_AMD_EXPORT void delete_TorusPolynomial(TorusPolynomial* obj);

/* ======================= additions (no reference counterpart) =========== */

/* Gate codes for bootsGate_batch / the device API. */
enum {
    TFHE_AMD_NAND = 0, TFHE_AMD_AND = 1, TFHE_AMD_OR = 2, TFHE_AMD_XOR = 3, TFHE_AMD_XNOR = 4, TFHE_AMD_NOR = 5,
    TFHE_AMD_ANDNY = 6, TFHE_AMD_ANDYN = 7, TFHE_AMD_ORNY = 8, TFHE_AMD_ORYN = 9, TFHE_AMD_MUX = 10
};

/* Parameter set from explicit values (the reference builds these only from
 * C++ constructors, libtfhe/tfhe_gate_bootstrapping.cpp:51-76). */
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_gate_bootstrapping_parameters(
    int32_t n, int32_t N, int32_t k, int32_t bk_l, int32_t bk_Bgbit, int32_t ks_t, int32_t ks_basebit,
    double ks_stdev, double bk_stdev, double max_stdev);

/* Cloud key from coefficient-domain arrays (e.g. keys generated by another
 * implementation): bk int32 [n][(k+1)l][k+1][N], ksk int32 [k*N][t][base][n+1]. */
TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_gate_bootstrapping_cloud_keyset_from_arrays(
    const TFheGateBootstrappingParameterSet* params, const int32_t* bk, const int32_t* ksk);

/* Batched gates: count independent gates, gate_op[g] in TFHE_AMD_*; in2 is
 * only read for MUX and may be NULL otherwise. All inputs are read before any
 * output is written, so outputs may alias inputs. Returns 0 on success. */
TFHE_AMD_EXPORT int32_t bootsGate_batch(int32_t count, const int32_t* gate_op, LweSample* const* out,
                                        const LweSample* const* in0, const LweSample* const* in1,
                                        const LweSample* const* in2, const TFheGateBootstrappingCloudKeySet* ck);

/* ---- bulk host helpers --------------------------------------------------- */
/* bootsSymEncrypt of count bits into int32 [count][n+1] (a[0..n), b). */
TFHE_AMD_EXPORT int32_t tfhe_amd_encrypt_bits(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                              const int32_t* bits, int32_t* out);
/* lwePhase of count samples int32 [count][n+1]. */
TFHE_AMD_EXPORT int32_t tfhe_amd_phases(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                        const int32_t* cts, int32_t* phases);
/* Key material in flat layouts: lwe_key [n], tlwe_key [k*N],
 * bk [n][(k+1)l][k+1][N], ksk [k*N][t][base][n+1]; NULL pointers are skipped. */
TFHE_AMD_EXPORT int32_t tfhe_amd_export_keys(const TFheGateBootstrappingSecretKeySet* key, int32_t* lwe_key,
                                             int32_t* tlwe_key, int32_t* bk, int32_t* ksk);
/* &key->cloud, for callers without the struct layout. */
TFHE_AMD_EXPORT const TFheGateBootstrappingCloudKeySet* tfhe_amd_cloud_of(const TFheGateBootstrappingSecretKeySet* key);

/* ---- device-resident API (keys and ciphertexts stay in HBM) ------------- */
/* Uploads (or returns the already uploaded) device key of a cloud key set. */
TFHE_AMD_EXPORT int32_t tfhe_amd_cloud_key_upload(const TFheGateBootstrappingCloudKeySet* ck);
/* Frees the device copy; the host key stays valid and is re-uploaded on next use
 * (a netlist captured into a graph on the old copy is re-captured on its next
 * run). A copy that a server registered (tfhe_amd_server_add_client) stays
 * allocated for that server until tfhe_amd_server_destroy. */
TFHE_AMD_EXPORT void tfhe_amd_cloud_key_release(const TFheGateBootstrappingCloudKeySet* ck);
/* Bytes of HBM the device key of ck occupies (0 if not uploaded). */
TFHE_AMD_EXPORT size_t tfhe_amd_cloud_key_device_bytes(const TFheGateBootstrappingCloudKeySet* ck);

/* Device ciphertext pool layout: int32 [slots][n+1] (a[0..n), b).
 * Gate g reads pool[in0[g]], pool[in1[g]] (and pool[in2[g]] for MUX) and
 * writes pool[out[g]]. All index arrays and the pool are device pointers;
 * stream is a hipStream_t (NULL = the library's stream). Asynchronous.
 * Each stream passed to these calls (and to tfhe_amd_keyswitch_device) gets a
 * device workspace of its own, up to 8 streams, so calls on different streams
 * overlap on the device; further streams share the library stream's
 * workspace, and a call whose workspace was last used from another stream
 * waits for that use on the device (never on the host). A workspace is kept
 * until the process ends; the first call on a stream (or one with a larger
 * count than before) allocates it, so make that call before capturing the
 * stream into a graph. */
TFHE_AMD_EXPORT int32_t tfhe_amd_gates_device(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                              const int32_t* d_op, const int32_t* d_in0, const int32_t* d_in1,
                                              const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                                              void* stream);
/* Same gates and outputs as tfhe_amd_gates_device, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * No counterpart in the reference. Calls no hipMalloc / hipFree: its
 * workspace is a static 32 MiB device array per GPU, held until the process
 * ends, which limits a call to about 2,700 gates with the 110-bit parameters
 * (larger counts return -1 with the limit in tfhe_amd_last_error). The cloud
 * key must already be on the device (tfhe_amd_cloud_key_upload, or any
 * earlier call with it); otherwise -1. All calls on one GPU share the array:
 * calls that overlap on the device (on different streams, or from several
 * threads without ordering) race and give wrong results. */
TFHE_AMD_EXPORT int32_t tfhe_amd_gates_device_noalloc(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                                      const int32_t* d_op, const int32_t* d_in0, const int32_t* d_in1,
                                                      const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                                                      void* stream);
/* Same, bootstrapping only (no keyswitch): writes extracted samples int32
 * [count][k*N+1] to d_ext. For MUX gates both halves are summed with +1/8. */
TFHE_AMD_EXPORT int32_t tfhe_amd_gates_device_woks(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                                   const int32_t* d_op, const int32_t* d_in0, const int32_t* d_in1,
                                                   const int32_t* d_in2, const int32_t* d_pool, int32_t* d_ext,
                                                   void* stream);
/* Keyswitch count extracted samples int32 [count][k*N+1] into [count][n+1]. */
TFHE_AMD_EXPORT int32_t tfhe_amd_keyswitch_device(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                                  const int32_t* d_in, int32_t* d_out, void* stream);

/* ---- netlists (level scheduler; no reference counterpart) -------------
 * A netlist is a set of bootstrapped gates (codes TFHE_AMD_NAND..MUX) over
 * wires 0..n_wires-1; wire w is slot w of a device ciphertext pool. Gate g
 * reads in0[g], in1[g] (and in2[g] for MUX; in2 may be NULL otherwise) and
 * writes out[g]; each wire is written by at most one gate, wires written by no
 * gate are inputs. Gates are levelised (schedule 0: as soon as possible,
 * 1: as late as possible, except that gates whose outputs nothing reads keep
 * their earliest level; 2: balanced, the as-soon-as-possible depth with each
 * level holding the gates whose as-late-as-possible level it is, filled with
 * ready gates up to the smallest width base * 2^k that holds those, where base
 * is the current GPU's SIMD count (4 per CU; 1024 without a GPU) and a MUX
 * counts twice; a value of 64 or more: balanced with that base); each level is
 * one batch launch, and with use_graph the whole sequence is captured once
 * into a hipGraph and replayed (the graph is re-captured when the pool or
 * stream changes). Other schedule values are an error. */
typedef struct tfhe_amd_netlist tfhe_amd_netlist;
TFHE_AMD_EXPORT tfhe_amd_netlist* tfhe_amd_netlist_create(const TFheGateBootstrappingCloudKeySet* ck, int32_t n_wires,
                                                          int32_t n_gates, const int32_t* op, const int32_t* in0,
                                                          const int32_t* in1, const int32_t* in2, const int32_t* out,
                                                          int32_t schedule);
/* Host-only levelisation (no GPU): returns the number of levels (negative on
 * error: bad gate code, wire out of range, wire driven twice, cycle) and the
 * 1-based level of every gate. */
TFHE_AMD_EXPORT int32_t tfhe_amd_netlist_levelize(int32_t n_wires, int32_t n_gates, const int32_t* op,
                                                  const int32_t* in0, const int32_t* in1, const int32_t* in2,
                                                  const int32_t* out, int32_t schedule, int32_t* level_of_gate);
/* Number of levels; widths (if not NULL) receives the gates per level. */
TFHE_AMD_EXPORT int32_t tfhe_amd_netlist_levels(const tfhe_amd_netlist* nl, int32_t* widths);
/* Asynchronous on stream (NULL = the library's stream). */
TFHE_AMD_EXPORT int32_t tfhe_amd_netlist_run(tfhe_amd_netlist* nl, int32_t* d_pool, void* stream, int32_t use_graph);
TFHE_AMD_EXPORT void tfhe_amd_netlist_destroy(tfhe_amd_netlist* nl);

/* ---- keys and key material as separate objects: include/lwekey.h:69-78,
 *      include/tlwe.h:159-172, include/tgsw.h:118-177, include/lwekeyswitch.h:50-55,
 *      include/lwebootstrappingkey.h:82-115. Parameters are not owned (as in the
 *      reference); new_LweBootstrappingKey also builds its keyswitch key (input
 *      dimension k N), and new_LweBootstrappingKeyFFT makes the handle that the
 *      gates and tfhe_bootstrap_FFT take (the device key is built from bk, which
 *      must outlive it, on first use). */
TFHE_AMD_EXPORT LweKey* new_LweKey(const LweParams* params);
TFHE_AMD_EXPORT void delete_LweKey(LweKey* key);
TFHE_AMD_EXPORT TLweKey* new_TLweKey(const TLweParams* params);
TFHE_AMD_EXPORT void delete_TLweKey(TLweKey* obj);
TFHE_AMD_EXPORT TGswKey* new_TGswKey(const TGswParams* params);
TFHE_AMD_EXPORT void delete_TGswKey(TGswKey* obj);
TFHE_AMD_EXPORT TGswSample* new_TGswSample(const TGswParams* params);
TFHE_AMD_EXPORT void delete_TGswSample(TGswSample* obj);
TFHE_AMD_EXPORT LweKeySwitchKey* new_LweKeySwitchKey(int32_t n, int32_t t, int32_t basebit, const LweParams* out_params);
TFHE_AMD_EXPORT void delete_LweKeySwitchKey(LweKeySwitchKey* obj);
TFHE_AMD_EXPORT LweBootstrappingKey* new_LweBootstrappingKey(const int32_t ks_t, const int32_t ks_basebit,
                                                             const LweParams* in_out_params, const TGswParams* bk_params);
TFHE_AMD_EXPORT void delete_LweBootstrappingKey(LweBootstrappingKey* obj);
TFHE_AMD_EXPORT LweBootstrappingKeyFFT* new_LweBootstrappingKeyFFT(const LweBootstrappingKey* bk);
TFHE_AMD_EXPORT void delete_LweBootstrappingKeyFFT(LweBootstrappingKeyFFT* obj);

/* ---- Key, parameter, sample and ciphertext files: include/tfhe_io.h:30-496 --
 * The reference's wire format (libtfhe/tfhe_io.cpp): text parameter
 * sections plus native-order binary key material, so files written by either
 * library load in the other. Constructors (new_*_from*) return NULL on
 * malformed input (reason in tfhe_amd_last_error; parsed sizes are checked
 * before allocating); importers into caller-allocated objects and writers
 * abort on bad input or arguments, as the reference. Parameter objects that
 * readers of keys create live until the process ends, as the reference's
 * garbage-collected ones; the parameter objects new_*Params_from* return are
 * the caller's. The std::ostream / std::istream forms follow below for C++. */
TFHE_AMD_EXPORT void export_lweParams_toFile(FILE* F, const LweParams* params);
TFHE_AMD_EXPORT LweParams* new_lweParams_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_lweSample_toFile(FILE* F, const LweSample* sample, const LweParams* params);
TFHE_AMD_EXPORT void import_lweSample_fromFile(FILE* F, LweSample* sample, const LweParams* params);
TFHE_AMD_EXPORT void export_lweKey_toFile(FILE* F, const LweKey* key);
TFHE_AMD_EXPORT LweKey* new_lweKey_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tLweParams_toFile(FILE* F, const TLweParams* params);
TFHE_AMD_EXPORT TLweParams* new_tLweParams_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tlweSample_toFile(FILE* F, const TLweSample* sample, const TLweParams* params);
TFHE_AMD_EXPORT void import_tlweSample_fromFile(FILE* F, TLweSample* sample, const TLweParams* params);
TFHE_AMD_EXPORT void export_tlweKey_toFile(FILE* F, const TLweKey* key);
TFHE_AMD_EXPORT TLweKey* new_tlweKey_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tGswParams_toFile(FILE* F, const TGswParams* params);
TFHE_AMD_EXPORT TGswParams* new_tGswParams_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tgswSample_toFile(FILE* F, const TGswSample* sample, const TGswParams* params);
TFHE_AMD_EXPORT void import_tgswSample_fromFile(FILE* F, TGswSample* sample, const TGswParams* params);
TFHE_AMD_EXPORT void export_tgswKey_toFile(FILE* F, const TGswKey* key);
TFHE_AMD_EXPORT TGswKey* new_tgswKey_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_lweKeySwitchKey_toFile(FILE* F, const LweKeySwitchKey* ks);
TFHE_AMD_EXPORT LweKeySwitchKey* new_lweKeySwitchKey_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_lweBootstrappingKey_toFile(FILE* F, const LweBootstrappingKey* bk);
TFHE_AMD_EXPORT LweBootstrappingKey* new_lweBootstrappingKey_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tfheGateBootstrappingParameterSet_toFile(FILE* F,
                                                                     const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tfheGateBootstrappingCloudKeySet_toFile(FILE* F, const TFheGateBootstrappingCloudKeySet* key);
TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_tfheGateBootstrappingSecretKeySet_toFile(FILE* F,
                                                                     const TFheGateBootstrappingSecretKeySet* key);
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromFile(FILE* F);
TFHE_AMD_EXPORT void export_gate_bootstrapping_ciphertext_toFile(FILE* F, const LweSample* sample,
                                                                 const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT void import_gate_bootstrapping_ciphertext_fromFile(FILE* F, LweSample* sample,
                                                                   const TFheGateBootstrappingParameterSet* params);

#ifdef __cplusplus
/* the same over C++ streams (include/tfhe_io.h:44-496; C linkage, as the
 * reference's EXPORT gives them) */
#include <iosfwd>
TFHE_AMD_EXPORT void export_lweParams_toStream(std::ostream& F, const LweParams* params);
TFHE_AMD_EXPORT LweParams* new_lweParams_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_lweSample_toStream(std::ostream& F, const LweSample* sample, const LweParams* params);
TFHE_AMD_EXPORT void import_lweSample_fromStream(std::istream& F, LweSample* sample, const LweParams* params);
TFHE_AMD_EXPORT void export_lweKey_toStream(std::ostream& F, const LweKey* key);
TFHE_AMD_EXPORT LweKey* new_lweKey_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tLweParams_toStream(std::ostream& F, const TLweParams* params);
TFHE_AMD_EXPORT TLweParams* new_tLweParams_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tlweSample_toStream(std::ostream& F, const TLweSample* sample, const TLweParams* params);
TFHE_AMD_EXPORT void import_tlweSample_fromStream(std::istream& F, TLweSample* sample, const TLweParams* params);
TFHE_AMD_EXPORT void export_tlweKey_toStream(std::ostream& F, const TLweKey* key);
TFHE_AMD_EXPORT TLweKey* new_tlweKey_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tGswParams_toStream(std::ostream& F, const TGswParams* params);
TFHE_AMD_EXPORT TGswParams* new_tGswParams_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tgswSample_toStream(std::ostream& F, const TGswSample* sample, const TGswParams* params);
TFHE_AMD_EXPORT void import_tgswSample_fromStream(std::istream& F, TGswSample* sample, const TGswParams* params);
TFHE_AMD_EXPORT void export_tgswKey_toStream(std::ostream& F, const TGswKey* key);
TFHE_AMD_EXPORT TGswKey* new_tgswKey_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_lweKeySwitchKey_toStream(std::ostream& F, const LweKeySwitchKey* ks);
TFHE_AMD_EXPORT LweKeySwitchKey* new_lweKeySwitchKey_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_lweBootstrappingKey_toStream(std::ostream& F, const LweBootstrappingKey* bk);
TFHE_AMD_EXPORT LweBootstrappingKey* new_lweBootstrappingKey_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tfheGateBootstrappingParameterSet_toStream(std::ostream& F,
                                                                       const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tfheGateBootstrappingCloudKeySet_toStream(std::ostream& F,
                                                                      const TFheGateBootstrappingCloudKeySet* key);
TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_tfheGateBootstrappingSecretKeySet_toStream(std::ostream& F,
                                                                       const TFheGateBootstrappingSecretKeySet* key);
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromStream(std::istream& F);
TFHE_AMD_EXPORT void export_gate_bootstrapping_ciphertext_toStream(std::ostream& F, const LweSample* sample,
                                                                   const TFheGateBootstrappingParameterSet* params);
TFHE_AMD_EXPORT void import_gate_bootstrapping_ciphertext_fromStream(std::istream& F, LweSample* sample,
                                                                     const TFheGateBootstrappingParameterSet* params);
#endif

/* Client-side batch encryption / decryption on the device, over contiguous
 * ciphertext rows [count][n+1]. decrypt: bits[g] = (phase > 0) as
 * bootsSymDecrypt (libtfhe/tfhe_gate_bootstrapping.cpp:162-165), phases[g] =
 * b - <a, s> as lwePhase (libtfhe/lwe-functions.cpp:72-81); either output may
 * be NULL. encrypt: bootsSymEncrypt (tfhe_gate_bootstrapping.cpp:154-159) of
 * bits[g] with noise stdev alpha_min, randomness from Philox4x32-10 keyed by
 * seed with ciphertext index first_index + g as counter (reproducible from
 * (seed, index); not the reference generator's stream). A (seed, index) pair
 * must never be used twice under one key: two ciphertexts with the same mask
 * and noise reveal whether their bits differ. first_index =
 * TFHE_AMD_NEXT_INDEX takes the key's next unused indices (the library keeps,
 * per key, one past the highest index any call has used, explicit ones
 * included), which is what callers without their own index bookkeeping
 * should pass. The LWE key is copied to the device on first use and freed
 * with the key set. Asynchronous on stream (NULL = the library's stream). */
#define TFHE_AMD_NEXT_INDEX UINT64_MAX
TFHE_AMD_EXPORT int32_t tfhe_amd_encrypt_device(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                                const int32_t* d_bits, int32_t* d_cts, uint64_t seed,
                                                uint64_t first_index, void* stream);
TFHE_AMD_EXPORT int32_t tfhe_amd_decrypt_device(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                                const int32_t* d_cts, int32_t* d_bits, int32_t* d_phases,
                                                void* stream);

/* Key generation on the device: a full secret key set (LWE and TLWE keys,
 * keyswitch key, bootstrapping key) drawn from Philox4x32-10 keyed by seed,
 * following the reference's procedure (libtfhe/tfhe_gate_bootstrapping.cpp:
 * 97-109) but not its generator order, so these are not the keys
 * new_random_gate_bootstrapping_secret_keyset gives for any seed. Takes
 * milliseconds instead of the host's sequential generator; free with
 * delete_gate_bootstrapping_secret_keyset. */
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* tfhe_amd_keygen_device(const TFheGateBootstrappingParameterSet* params,
                                                                        uint64_t seed);

/* Serving loop (no counterpart in the reference; its callers are circuits
 * like test/test-long-run.cpp:61-185 submitted by many clients). One
 * scheduler thread admits requests in arrival order as their wires fit in a
 * device wire arena of wire_budget_bytes, batches the next dependency level of
 * every admitted request of a client into one launch (requests that arrive
 * while others run join the next launch), spreads clients over n_streams
 * streams with up to `depth` launches in flight per client, and copies each
 * request's outputs back when its last level has finished. Keys stay resident
 * (their HBM is counted in key_bytes); all clients of a server share one
 * parameter set. 0 for n_streams / depth / max_batch picks 8 / 2 / 4096. */
typedef struct tfhe_amd_server tfhe_amd_server;
typedef struct {
    int64_t submitted, completed, failed, gates, launches, max_launch_gates;
    int32_t queued, active, clients, n_streams;
    size_t wire_bytes_budget, wire_bytes_in_use, wire_bytes_peak, key_bytes;
} tfhe_amd_server_stats;
TFHE_AMD_EXPORT tfhe_amd_server* tfhe_amd_server_create(int32_t n_streams, int32_t depth, int32_t max_batch,
                                                        size_t wire_budget_bytes);
/* Registers a client's cloud key (uploaded now if needed); returns the client id.
 * The server holds the device copy until it is destroyed, so the key set may be
 * released or deleted meanwhile without affecting the server. */
TFHE_AMD_EXPORT int32_t tfhe_amd_server_add_client(tfhe_amd_server* srv, const TFheGateBootstrappingCloudKeySet* ck);
/* Submits a netlist (same conventions as tfhe_amd_netlist_create) whose
 * undriven wires input_wires[0..n_inputs) hold input_cts[n_inputs][n+1]
 * (copied now). output_cts[n_outputs][n+1] receives the ciphertexts of
 * output_wires when the request completes and must stay valid until then.
 * Returns a request id >= 0, or -1 (bad netlist, wire read but neither input
 * nor driven, request larger than the wire budget). */
TFHE_AMD_EXPORT int64_t tfhe_amd_server_submit(tfhe_amd_server* srv, int32_t client, int32_t n_wires, int32_t n_gates,
                                               const int32_t* op, const int32_t* in0, const int32_t* in1,
                                               const int32_t* in2, const int32_t* out, int32_t n_inputs,
                                               const int32_t* input_wires, const int32_t* input_cts,
                                               int32_t n_outputs, const int32_t* output_wires, int32_t* output_cts);
/* Waits for a request (timeout_ms < 0: no limit). 0 = done (latency_ms =
 * submit to outputs written; the id is then retired), 1 = timed out, -1 = failed. */
TFHE_AMD_EXPORT int32_t tfhe_amd_server_wait(tfhe_amd_server* srv, int64_t request, int32_t timeout_ms,
                                             double* latency_ms);
TFHE_AMD_EXPORT int32_t tfhe_amd_server_get_stats(tfhe_amd_server* srv, tfhe_amd_server_stats* stats);
/* Packing (no reference counterpart; for loaded servers): when at least two
 * clients have gates ready and together at least min_gates, one launch carries
 * the next level of each of them (up to 32 clients and max_batch gates), every
 * client's gates with its own key in the batched blind-rotation kernel, on a
 * stream of its own; while packing is on, a client whose newest launch was
 * packed waits for the next packed launch. 0 (the default) turns it off.
 * Outputs are bit-identical either way. N = 1024 keys only. */
TFHE_AMD_EXPORT int32_t tfhe_amd_server_set_packing(tfhe_amd_server* srv, int32_t min_gates);
/* Packed launches made so far. */
TFHE_AMD_EXPORT int64_t tfhe_amd_server_packed_launches(tfhe_amd_server* srv);
/* Finishes every submitted request, then frees the server. */
TFHE_AMD_EXPORT void tfhe_amd_server_destroy(tfhe_amd_server* srv);

/* Blind-rotation kernel choice. mode: 0 = automatic, 1 = one wavefront per
 * bootstrap, 2 = one wavefront per bootstrap with the key and twiddles read
 * through the vector memory path (the batched kernel, on the fused-butterfly
 * transform: the batch's whole device rounds on workgroups of eight such
 * bootstraps held together by a barrier per key row, the rest on workgroups of
 * one), 9 = the batched kernel on workgroups of one only, 7 = the same on the
 * radix-8 transform (9 and 7 kept for comparison), 3 = two wavefronts per
 * bootstrap, 5 = 4l wavefronts per bootstrap (lowest latency; modes 2, 3, 5, 7,
 * 9 for N = 1024); for N = 2048:
 * 4 = the pair-transform kernel (four wavefronts per bootstrap, two bootstraps
 * per CU) on the fused-butterfly transform, 6 = the two-wavefront kernel (four
 * bootstraps per CU, same transform), 8 = the pair kernel on the radix-8
 * stages (kept for comparison; modes 6 and 8 run mode 2 for N = 1024). All
 * modes give bit-identical results. In automatic mode, for N = 1024, a key the
 * autotuner has measured takes its per-size choice (below) unless crossovers
 * have been set here since; otherwise batches of at most latency_max gates use
 * mode 5, then up to small_batch_max mode 3, larger ones mode 2 (a negative
 * value keeps the current crossover; setting either one makes the crossovers
 * apply instead of the per-size choice until the next autotune); for N = 2048,
 * a batch takes the kernel (4 or 6) the autotuner measured the faster at the
 * smallest of its ladder sizes that holds the batch (until it runs: mode 4 up to
 * 512 gates), batches beyond the ladder mode 6. */
TFHE_AMD_EXPORT void tfhe_amd_set_launch_config(int32_t br_mode, int32_t latency_max, int32_t small_batch_max);
TFHE_AMD_EXPORT int32_t tfhe_amd_get_launch_config(int32_t* br_mode, int32_t* latency_max, int32_t* small_batch_max);
/* The blind-rotation kernel (mode number as above, never 0) that a batch of
 * `count` gates under this key runs with the current settings, for a launch
 * with nothing else on the device; uploads the key if needed; -1 on error.
 * No counterpart in the reference: for tests and tuning reports. */
TFHE_AMD_EXPORT int32_t tfhe_amd_kernel_for(const TFheGateBootstrappingCloudKeySet* ck, int32_t count);
/* Autotuner: for an N = 1024 key, times modes 5, 3 and 2 at batch sizes 64,
 * 128, 256, 384, 512, 640, 768, 896, 1024, 1280, 1536, 2048, 3072, 4096
 * (milliseconds per batch including the keyswitch, into ms[3][14] in that mode
 * order, if not NULL), records the fastest mode per size for automatic mode (a
 * batch takes the choice of the smallest of these sizes that holds it, mode 2
 * beyond them) and sets the crossovers: latency_max to the largest size up to
 * which mode 5 is the fastest at every measured point, small_batch_max to the
 * largest size up to which mode 3 then beats mode 2 at every further point
 * (they apply instead of the per-size choice once tfhe_amd_set_launch_config
 * sets crossovers). For an N = 2048 key, times modes 4 and 6 at the same sizes
 * (ms rows 0 and 1; row 2 zero), records the faster one per size for automatic
 * mode, and returns in small_batch_max the largest size up to which mode 4 is
 * the faster at every point (latency_max 0).
 * Synchronous; allocates and frees its own buffers. */
TFHE_AMD_EXPORT int32_t tfhe_amd_autotune(const TFheGateBootstrappingCloudKeySet* ck, int32_t* latency_max,
                                          int32_t* small_batch_max, float* ms);
/* Keyswitch kernels: 0 (default) = exact int8 matrix-core product of the
 * one-hot digit matrix with the key's signed byte planes (keys with t = 8,
 * basebit = 2); 1 = lanes-over-gates integer kernels (any t, basebit). Both
 * give identical results. */
TFHE_AMD_EXPORT void tfhe_amd_set_keyswitch_path(int32_t path);

/* Per-launch timing: when enabled, every device gate batch records HIP
 * events on its stream around the blind-rotation kernel and around the
 * keyswitch kernels. profile_read waits for them and returns the summed
 * milliseconds and the number of batches since the last read. */
TFHE_AMD_EXPORT void tfhe_amd_profile_enable(int32_t on);
TFHE_AMD_EXPORT int32_t tfhe_amd_profile_read(double* br_ms, double* ks_ms, int32_t* launches);

/* Device memory helpers for callers without a HIP runtime of their own. */
/* Asynchronous device-to-device copy with a plain copy kernel (16-byte
 * aligned sizes and pointers), e.g. to move ciphertexts between pools. */
TFHE_AMD_EXPORT int32_t tfhe_amd_copy_device(void* dst, const void* src, size_t bytes, void* stream);
TFHE_AMD_EXPORT void* tfhe_amd_malloc(size_t bytes);
TFHE_AMD_EXPORT int32_t tfhe_amd_free(void* ptr);
/* Host <-> device copies, ordered after the work queued on the library's
 * stream; they have completed when the call returns. */
TFHE_AMD_EXPORT int32_t tfhe_amd_memcpy_h2d(void* dst, const void* src, size_t bytes);
TFHE_AMD_EXPORT int32_t tfhe_amd_memcpy_d2h(void* dst, const void* src, size_t bytes);
TFHE_AMD_EXPORT int32_t tfhe_amd_synchronize(void);
TFHE_AMD_EXPORT void* tfhe_amd_stream(void);
TFHE_AMD_EXPORT int32_t tfhe_amd_device_count(void);
TFHE_AMD_EXPORT const char* tfhe_amd_last_error(void);

#endif
