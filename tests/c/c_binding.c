/* A C program using the gate API exactly as C code written against the
 * reference does (cf. the reference's test/test-c-binding.c), compiled with a
 * C compiler against include/tfhe_amd.h and linked to libtfhe_amd.so.
 * "host" checks key generation / encryption / the non-bootstrapped gates;
 * "gpu" also bootstraps every gate through the one-gate-per-call API. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "tfhe_amd.h"

static int fails = 0;
#define CHECK(c) do { if (!(c)) { fprintf(stderr, "FAILED: %s (line %d)\n", #c, __LINE__); fails++; } } while (0)

int main(int argc, char** argv) {
    int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    uint32_t seed[] = {314, 1592, 657};
    tfhe_random_generator_setSeed(seed, 3);
    TFheGateBootstrappingParameterSet* params = new_default_gate_bootstrapping_parameters(110);
    CHECK(params->in_out_params->n == 630 && params->tgsw_params->l == 3 && params->tgsw_params->Bgbit == 7);
    CHECK(params->tgsw_params->tlwe_params->N == 1024 && params->ks_t == 8 && params->ks_basebit == 2);
    TFheGateBootstrappingSecretKeySet* key = new_random_gate_bootstrapping_secret_keyset(params);
    const TFheGateBootstrappingCloudKeySet* bk = &key->cloud;
    LweSample* ct = new_gate_bootstrapping_ciphertext_array(4, params);
    for (int b = 0; b < 2; b++) {
        bootsSymEncrypt(&ct[0], b, key);
        CHECK(bootsSymDecrypt(&ct[0], key) == b);
        bootsNOT(&ct[1], &ct[0], bk);
        CHECK(bootsSymDecrypt(&ct[1], key) == 1 - b);
        bootsCOPY(&ct[2], &ct[0], bk);
        CHECK(bootsSymDecrypt(&ct[2], key) == b);
        bootsCONSTANT(&ct[3], b, bk);
        CHECK(bootsSymDecrypt(&ct[3], key) == b);
    }
    if (gpu) {
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
                for (int c = 0; c < 2; c++) {
                    bootsSymEncrypt(&ct[0], a, key);
                    bootsSymEncrypt(&ct[1], b, key);
                    bootsSymEncrypt(&ct[2], c, key);
                    bootsNAND(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == !(a && b));
                    bootsAND(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (a && b));
                    bootsOR(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (a || b));
                    bootsXOR(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (a ^ b));
                    bootsXNOR(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == !(a ^ b));
                    bootsNOR(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == !(a || b));
                    bootsANDNY(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (!a && b));
                    bootsANDYN(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (a && !b));
                    bootsORNY(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (!a || b));
                    bootsORYN(&ct[3], &ct[0], &ct[1], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (a || !b));
                    bootsMUX(&ct[3], &ct[0], &ct[1], &ct[2], bk); CHECK(bootsSymDecrypt(&ct[3], key) == (a ? b : c));
                }
    }
    delete_gate_bootstrapping_ciphertext_array(4, ct);
    delete_gate_bootstrapping_secret_keyset(key);
    delete_gate_bootstrapping_parameters(params);
    printf("%s: %d failures\n", gpu ? "gpu" : "host", fails);
    return fails ? 1 : 0;
}
