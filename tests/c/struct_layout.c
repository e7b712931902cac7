/* Prints the size and field offsets of every public struct of the gate API.
 * Built twice by tests/test_library_cpu.py: as C against include/tfhe_amd.h,
 * and (where the reference sources are present) as C++ against the
 * reference's own headers (-DREFERENCE_HEADERS), so the two outputs can be
 * compared line by line: a program compiled against the reference's headers
 * then finds every field where it expects it. */
#include <stddef.h>
#include <stdio.h>
#ifdef REFERENCE_HEADERS
#include "tfhe.h"
#include "tfhe_gate_bootstrapping_structures.h"
#else
#include "tfhe_amd.h"
#endif

#define S(T) printf("%s size %zu\n", #T, sizeof(T))
#define F(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f))

int main(void) {
    S(LweParams); F(LweParams, n); F(LweParams, alpha_min); F(LweParams, alpha_max);
    S(TLweParams); F(TLweParams, N); F(TLweParams, k); F(TLweParams, alpha_min); F(TLweParams, alpha_max);
    F(TLweParams, extracted_lweparams);
    S(TGswParams); F(TGswParams, l); F(TGswParams, Bgbit); F(TGswParams, Bg); F(TGswParams, halfBg);
    F(TGswParams, maskMod); F(TGswParams, tlwe_params); F(TGswParams, kpl); F(TGswParams, h); F(TGswParams, offset);
    S(LweKey); F(LweKey, params); F(LweKey, key);
    S(LweSample); F(LweSample, a); F(LweSample, b); F(LweSample, current_variance);
    S(IntPolynomial); F(IntPolynomial, N); F(IntPolynomial, coefs);
    S(TorusPolynomial); F(TorusPolynomial, N); F(TorusPolynomial, coefsT);
    S(TLweKey); F(TLweKey, params); F(TLweKey, key);
    S(TLweSample); F(TLweSample, a); F(TLweSample, b); F(TLweSample, current_variance); F(TLweSample, k);
    S(TGswKey); F(TGswKey, params); F(TGswKey, tlwe_params); F(TGswKey, key); F(TGswKey, tlwe_key);
    S(TGswSample); F(TGswSample, all_sample); F(TGswSample, bloc_sample); F(TGswSample, k); F(TGswSample, l);
    S(TGswSampleFFT); F(TGswSampleFFT, all_samples); F(TGswSampleFFT, sample); F(TGswSampleFFT, k);
    F(TGswSampleFFT, l);
    S(LweKeySwitchKey); F(LweKeySwitchKey, n); F(LweKeySwitchKey, t); F(LweKeySwitchKey, basebit);
    F(LweKeySwitchKey, base); F(LweKeySwitchKey, out_params); F(LweKeySwitchKey, ks0_raw);
    F(LweKeySwitchKey, ks1_raw); F(LweKeySwitchKey, ks);
    S(LweBootstrappingKey); F(LweBootstrappingKey, in_out_params); F(LweBootstrappingKey, bk_params);
    F(LweBootstrappingKey, accum_params); F(LweBootstrappingKey, extract_params); F(LweBootstrappingKey, bk);
    F(LweBootstrappingKey, ks);
    S(LweBootstrappingKeyFFT); F(LweBootstrappingKeyFFT, in_out_params); F(LweBootstrappingKeyFFT, bk_params);
    F(LweBootstrappingKeyFFT, accum_params); F(LweBootstrappingKeyFFT, extract_params);
    F(LweBootstrappingKeyFFT, bkFFT); F(LweBootstrappingKeyFFT, ks);
    S(TFheGateBootstrappingParameterSet); F(TFheGateBootstrappingParameterSet, ks_t);
    F(TFheGateBootstrappingParameterSet, ks_basebit); F(TFheGateBootstrappingParameterSet, in_out_params);
    F(TFheGateBootstrappingParameterSet, tgsw_params);
    S(TFheGateBootstrappingCloudKeySet); F(TFheGateBootstrappingCloudKeySet, params);
    F(TFheGateBootstrappingCloudKeySet, bk); F(TFheGateBootstrappingCloudKeySet, bkFFT);
    S(TFheGateBootstrappingSecretKeySet); F(TFheGateBootstrappingSecretKeySet, params);
    F(TFheGateBootstrappingSecretKeySet, lwe_key); F(TFheGateBootstrappingSecretKeySet, tgsw_key);
    F(TFheGateBootstrappingSecretKeySet, cloud);
    return 0;
}
