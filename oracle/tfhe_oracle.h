/* CPU oracle for TFHE gate bootstrapping — TEST INFRASTRUCTURE ONLY.
 *
 * A plain restatement of the reference algorithm (tfhe/tfhe, libtfhe/) used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker. It is never linked into, loaded by or called from the product
 * library (tfhe_amd/).
 *
 * All torus arithmetic is done in uint32_t so that wrap-around is defined.
 * The external product is the EXACT negacyclic integer product (mod 2^32), not
 * an FFT: the reference computes it with FP64 transforms and truncation
 * (fft_processor_spqlios.cpp:114), so reference outputs differ from this one
 * by the transform's rounding only.
 *
 * Parity pins (what checks each function; see tests/test_oracle_cpu.py):
 *   keygen / encrypt (RNG replay)     : bit-exact vs reference keys (oracle/_ref)
 *                                        except BK b, which the reference builds
 *                                        with an FFT product (|diff| <= 1 unit)
 *   modswitch, gate constants          : closed forms + reference truth tables
 *   decompose                          : recomposition bound + brute force
 *   negacyclic product                 : brute-force double sum on tiny N,
 *                                        X^a shift identities
 *   blind rotation                     : noiseless-BK closed form X^{-barb+sum bara_i s_i} v
 *   keyswitch                          : bit-exact vs reference lweKeySwitch
 *   full gates                         : decrypted truth tables; phase vs the
 *                                        reference spqlios output on the same keys
 */
#ifndef TFHE_ORACLE_H
#define TFHE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t n;          /* LWE dimension */
    int32_t N;          /* ring degree */
    int32_t k;          /* TLWE mask size */
    int32_t l;          /* gadget length */
    int32_t Bgbit;      /* log2 gadget base */
    int32_t ks_t;       /* keyswitch length */
    int32_t ks_basebit; /* log2 keyswitch base */
    double lwe_stdev;   /* LWE / keyswitch noise stdev */
    double bk_stdev;    /* bootstrapping key noise stdev */
} orc_params;

/* gate codes (shared with the product library's bootsGate_batch) */
enum { ORC_NAND = 0, ORC_AND, ORC_OR, ORC_XOR, ORC_XNOR, ORC_NOR, ORC_ANDNY, ORC_ANDYN, ORC_ORNY, ORC_ORYN, ORC_MUX };

void orc_set_seed(const uint32_t* values, int32_t size);
/* ksk layout [k*N][t][2^basebit][n+1] (a..., b); bk layout [n][(k+1)l][k+1][N] */
void orc_keygen(const orc_params* P, int32_t* lwe_key, int32_t* tlwe_key, int32_t* ksk, int32_t* bk);
void orc_lwe_encrypt(const int32_t* key, int32_t n, int32_t mu, double alpha, int32_t* out);
void orc_boots_encrypt(const orc_params* P, const int32_t* key, int32_t bit, int32_t* out);
int32_t orc_lwe_phase(const int32_t* ct, const int32_t* key, int32_t n);
int32_t orc_dtot32(double d);
int32_t orc_modswitch_from_torus32(int32_t phase, int32_t Msize);
int32_t orc_modswitch_to_torus32(int32_t mu, int32_t Msize);
void orc_gate_linear(int32_t op, int32_t which, const int32_t* a, const int32_t* b, const int32_t* c, int32_t n, int32_t* out);
void orc_decompose(const orc_params* P, const int32_t* poly, int32_t* digits);
void orc_mul_xai(int32_t a, const int32_t* in, int32_t N, int32_t* out);
void orc_mul_xai_minus_one(int32_t a, const int32_t* in, int32_t N, int32_t* out);
void orc_negacyclic_mul_add(const int32_t* ipoly, const int32_t* tpoly, int32_t N, int32_t* acc);
void orc_extprod(const orc_params* P, const int32_t* bk_row, int32_t* acc);
void orc_blind_rotate(const orc_params* P, const int32_t* bk, const int32_t* bara, int32_t* acc);
void orc_blind_rotate_and_extract(const orc_params* P, const int32_t* bk, const int32_t* v, int32_t barb,
                                  const int32_t* bara, int32_t* out);
void orc_bootstrap_woks(const orc_params* P, const int32_t* bk, int32_t mu, const int32_t* x, int32_t* out);
void orc_keyswitch(const orc_params* P, const int32_t* ksk, const int32_t* in, int32_t* out);
void orc_gate(const orc_params* P, const int32_t* bk, const int32_t* ksk, int32_t op, const int32_t* a,
              const int32_t* b, const int32_t* c, int32_t* out);
/* count gates on up to `threads` host threads; in* are [count][n+1] (in2 may be NULL) */
void orc_gates(const orc_params* P, const int32_t* bk, const int32_t* ksk, int32_t count, const int32_t* ops,
               const int32_t* in0, const int32_t* in1, const int32_t* in2, int32_t* out, int32_t threads);
uint64_t orc_fnv1a(const void* data, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif
