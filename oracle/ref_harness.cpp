// Test infrastructure only (never linked into the product).
//
// A driver around the reference library built from its own sources by
// oracle/Makefile (`make -C oracle ref`). It uses only the reference's public
// C API and struct layouts:
//   keygen  : new_random_gate_bootstrapping_secret_keyset
//             (libtfhe/tfhe_gate_bootstrapping.cpp:97-109)
//   gates   : boots* (libtfhe/boot-gates.cpp:34-370)
//   keyswitch: lweKeySwitch (libtfhe/lwe-keyswitch-functions.cpp:228-238)
//
// Commands (all binary outputs are little-endian int32 unless stated):
//   keys  SEED OUTDIR           dump lwe_key, tlwe_key, bk (coefficient domain), ksk
//   gates SEED COUNT OPS OUTFILE  encrypt random bits (global RNG, in order) and run gates
//   ks    SEED COUNT OUTFILE    keyswitch random extracted-dimension samples (input, output,
//                               the output's current_variance as a double)
//   bench SEED COUNT THREADS    time bootsNAND on fresh encryptions; prints one JSON line
//   files SEED DIR COUNT        write params.gb, secret.key, cloud.key, lwekey.bin, COUNT
//                               ciphertexts cts.bin (bits in bits.txt) and nand.bin = NAND(ct0, ct1)
//                               with the reference's file functions (tfhe_io.cpp)
//   readfiles SEED DIR COUNT    read those files with the reference's readers; print the
//                               decrypted bits, NAND(ct0, ct1) decrypted, and key hashes
//   smallkeys SEED DIR COUNT    key set at the 110-bit parameters except n = 16 (the
//                               reference's transforms are fixed at N = 1024): every object
//                               its file functions write, one file each (tfhe_io.h:30-496),
//                               and COUNT gates (all eleven types in turn) on fresh
//                               encryptions: gates.bin = per gate op, 3 bits, 3 input
//                               samples, the output sample (n+1 int32 each) and its phase
//   extprod SEED OUTFILE COUNT  on the smallkeys key set of the same seed: COUNT one-row
//                               blind rotations tfhe_blindRotate_FFT(acc, &bkFFT[i], &a, 1)
//                               (lwe-bootstrapping-functions-fft.cpp:88-112: acc += row i (x)
//                               ((X^a - 1) acc), the external product on the reference's
//                               transform); per case i, a, the input and the output
//                               accumulators ((k+1) N int32 each)
#include <tfhe.h>
#include <tfhe_io.h>
#include <lwesamples.h>
#include <lwekey.h>
#include <lweparams.h>
#include <tlwe.h>
#include <tgsw.h>
#include <polynomials.h>
#include <lwekeyswitch.h>
#include <lwebootstrappingkey.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

// Gate codes shared with the oracle and the product library.
enum { G_NAND = 0, G_AND, G_OR, G_XOR, G_XNOR, G_NOR, G_ANDNY, G_ANDYN, G_ORNY, G_ORYN, G_MUX };

static TFheGateBootstrappingSecretKeySet* make_keys(uint32_t seed) {
    uint32_t s[1] = {seed};
    tfhe_random_generator_setSeed(s, 1);
    TFheGateBootstrappingParameterSet* params = new_default_gate_bootstrapping_parameters(110);
    return new_random_gate_bootstrapping_secret_keyset(params);
}

static void write_all(FILE* f, const void* p, size_t bytes) {
    if (fwrite(p, 1, bytes, f) != bytes) { perror("fwrite"); exit(2); }
}

static void dump_keys(const TFheGateBootstrappingSecretKeySet* key, const std::string& dir) {
    const TFheGateBootstrappingParameterSet* P = key->params;
    const int n = P->in_out_params->n;
    const TGswParams* gp = P->tgsw_params;
    const int N = gp->tlwe_params->N, k = gp->tlwe_params->k, kpl = gp->kpl;
    FILE* f = fopen((dir + "/lwe_key.bin").c_str(), "wb");
    write_all(f, key->lwe_key->key, sizeof(int32_t) * n);
    fclose(f);
    f = fopen((dir + "/tlwe_key.bin").c_str(), "wb");
    for (int i = 0; i < k; i++) write_all(f, key->tgsw_key->tlwe_key.key[i].coefs, sizeof(int32_t) * N);
    fclose(f);
    const LweBootstrappingKey* bk = key->cloud.bk;
    f = fopen((dir + "/bk.bin").c_str(), "wb");
    for (int i = 0; i < n; i++)
        for (int r = 0; r < kpl; r++)
            for (int c = 0; c <= k; c++)
                write_all(f, bk->bk[i].all_sample[r].a[c].coefsT, sizeof(int32_t) * N);
    fclose(f);
    const LweKeySwitchKey* ks = bk->ks;
    f = fopen((dir + "/ksk.bin").c_str(), "wb");
    const int rows = ks->n * ks->t * ks->base;
    for (int r = 0; r < rows; r++) {
        write_all(f, ks->ks0_raw[r].a, sizeof(int32_t) * n);
        write_all(f, &ks->ks0_raw[r].b, sizeof(int32_t));
    }
    fclose(f);
    f = fopen((dir + "/params.txt").c_str(), "w");
    fprintf(f, "n=%d N=%d k=%d l=%d Bgbit=%d ks_t=%d ks_basebit=%d\n", n, N, k, gp->l, gp->Bgbit, P->ks_t, P->ks_basebit);
    fclose(f);
}

static void run_gate(int op, LweSample* out, const LweSample* a, const LweSample* b, const LweSample* c,
                     const TFheGateBootstrappingCloudKeySet* ck) {
    switch (op) {
        case G_NAND: bootsNAND(out, a, b, ck); break;
        case G_AND: bootsAND(out, a, b, ck); break;
        case G_OR: bootsOR(out, a, b, ck); break;
        case G_XOR: bootsXOR(out, a, b, ck); break;
        case G_XNOR: bootsXNOR(out, a, b, ck); break;
        case G_NOR: bootsNOR(out, a, b, ck); break;
        case G_ANDNY: bootsANDNY(out, a, b, ck); break;
        case G_ANDYN: bootsANDYN(out, a, b, ck); break;
        case G_ORNY: bootsORNY(out, a, b, ck); break;
        case G_ORYN: bootsORYN(out, a, b, ck); break;
        case G_MUX: bootsMUX(out, a, b, c, ck); break;
        default: fprintf(stderr, "bad op %d\n", op); exit(2);
    }
}

static void put_sample(FILE* f, const LweSample* s, int n) {
    write_all(f, s->a, sizeof(int32_t) * n);
    write_all(f, &s->b, sizeof(int32_t));
}

static uint64_t fnv(const void* p, size_t bytes, uint64_t h = 1469598103934665603ULL) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < bytes; i++) h = (h ^ c[i]) * 1099511628211ULL;
    return h;
}

static FILE* open_or_die(const std::string& path, const char* mode) {
    FILE* f = fopen(path.c_str(), mode);
    if (!f) { perror(path.c_str()); exit(2); }
    return f;
}

// A key set at explicit parameters (the reference's C++ parameter-set constructor,
// include/tfhe_gate_bootstrapping_structures.h).
static TFheGateBootstrappingSecretKeySet* make_keys_with(uint32_t seed, int n, int N, int l, int Bgbit, int t,
                                                         int basebit, double a_lwe, double a_bk) {
    uint32_t s[1] = {seed};
    tfhe_random_generator_setSeed(s, 1);
    LweParams* lp = new_LweParams(n, a_lwe, 0.012467);
    TLweParams* tp = new_TLweParams(N, 1, a_bk, 0.012467);
    TGswParams* gp = new_TGswParams(l, Bgbit, tp);
    auto* P = new TFheGateBootstrappingParameterSet(t, basebit, lp, gp);
    return new_random_gate_bootstrapping_secret_keyset(P);
}

static int write_small_keys(uint32_t seed, const std::string& dir, int count) {
    TFheGateBootstrappingSecretKeySet* key = make_keys_with(seed, 16, 1024, 3, 7, 8, 2, 1. / 32768, 1. / 33554432);
    const TFheGateBootstrappingParameterSet* P = key->params;
    const int n = P->in_out_params->n;
    const TGswParams* gp = P->tgsw_params;
    const TLweParams* tp = gp->tlwe_params;
    const LweBootstrappingKey* bk = key->cloud.bk;
    std::mt19937_64 bitgen(0x5eed0000ULL + seed);
    auto out = [&](const char* name) { return open_or_die(dir + "/" + name, "wb"); };
    FILE* f;
    f = out("secret.key"); export_tfheGateBootstrappingSecretKeySet_toFile(f, key); fclose(f);
    f = out("cloud.key"); export_tfheGateBootstrappingCloudKeySet_toFile(f, &key->cloud); fclose(f);
    f = out("params.gb"); export_tfheGateBootstrappingParameterSet_toFile(f, P); fclose(f);
    f = out("lweparams.txt"); export_lweParams_toFile(f, P->in_out_params); fclose(f);
    f = out("lwekey.bin"); export_lweKey_toFile(f, key->lwe_key); fclose(f);
    f = out("tlweparams.txt"); export_tLweParams_toFile(f, tp); fclose(f);
    f = out("tlwekey.bin"); export_tlweKey_toFile(f, &key->tgsw_key->tlwe_key); fclose(f);
    f = out("tlwesample.bin"); export_tlweSample_toFile(f, &bk->bk[0].all_sample[1], tp); fclose(f);
    f = out("tgswparams.txt"); export_tGswParams_toFile(f, gp); fclose(f);
    f = out("tgswkey.bin"); export_tgswKey_toFile(f, key->tgsw_key); fclose(f);
    f = out("tgswsample.bin"); export_tgswSample_toFile(f, &bk->bk[1], gp); fclose(f);
    f = out("kskey.bin"); export_lweKeySwitchKey_toFile(f, bk->ks); fclose(f);
    f = out("bk.bin"); export_lweBootstrappingKey_toFile(f, bk); fclose(f);
    {
        LweSample* s1 = new_gate_bootstrapping_ciphertext(P);
        bootsSymEncrypt(s1, 1, key);
        f = out("lwesample.bin"); export_lweSample_toFile(f, s1, P->in_out_params); fclose(f);
        delete_gate_bootstrapping_ciphertext(s1);
    }
    f = open_or_die(dir + "/gates.bin", "wb");
    int32_t hdr[2] = {count, n};
    write_all(f, hdr, sizeof hdr);
    LweSample* ct = new_gate_bootstrapping_ciphertext_array(4, P);
    for (int g = 0; g < count; g++) {
        const int op = g % 11;
        int32_t rec[4] = {op, 0, 0, 0};
        for (int j = 0; j < 3; j++) {
            rec[1 + j] = (int32_t)(bitgen() & 1);
            bootsSymEncrypt(&ct[j], rec[1 + j], key);
        }
        run_gate(op, &ct[3], &ct[0], &ct[1], &ct[2], &key->cloud);
        write_all(f, rec, sizeof rec);
        for (int j = 0; j < 4; j++) put_sample(f, &ct[j], n);
        const int32_t ph = lwePhase(&ct[3], key->lwe_key);
        write_all(f, &ph, sizeof ph);
    }
    fclose(f);
    return 0;
}

static int write_extprod(uint32_t seed, const std::string& path, int count) {
    TFheGateBootstrappingSecretKeySet* key = make_keys_with(seed, 16, 1024, 3, 7, 8, 2, 1. / 32768, 1. / 33554432);
    const TFheGateBootstrappingParameterSet* P = key->params;
    const TGswParams* gp = P->tgsw_params;
    const int N = gp->tlwe_params->N, k = gp->tlwe_params->k, n = P->in_out_params->n;
    std::mt19937_64 gen(0xe7e7ULL + seed);
    TLweSample* acc = new_TLweSample(gp->tlwe_params);
    FILE* f = open_or_die(path, "wb");
    int32_t hdr[3] = {count, k, N};
    write_all(f, hdr, sizeof hdr);
    for (int c = 0; c < count; c++) {
        const int32_t row = c % n;
        int32_t a = (int32_t)(1 + gen() % (2 * N - 1));
        if (c == 1) a = N;  // a pure negation
        for (int q = 0; q <= k; q++)
            for (int j = 0; j < N; j++) acc->a[q].coefsT[j] = (int32_t)(uint32_t)gen();
        int32_t rec[2] = {row, a};
        write_all(f, rec, sizeof rec);
        for (int q = 0; q <= k; q++) write_all(f, acc->a[q].coefsT, sizeof(int32_t) * N);
        tfhe_blindRotate_FFT(acc, key->cloud.bkFFT->bkFFT + row, &a, 1, gp);
        for (int q = 0; q <= k; q++) write_all(f, acc->a[q].coefsT, sizeof(int32_t) * N);
    }
    fclose(f);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: ref_harness keys|gates|ks|bench|files|readfiles|components|smallkeys SEED ...\n");
        return 2;
    }
    const std::string cmd = argv[1];
    const uint32_t seed = (uint32_t)strtoul(argv[2], nullptr, 10);
    if (cmd == "smallkeys" && argc >= 5) return write_small_keys(seed, argv[3], atoi(argv[4]));
    if (cmd == "extprod" && argc >= 5) return write_extprod(seed, argv[3], atoi(argv[4]));
    TFheGateBootstrappingSecretKeySet* key = make_keys(seed);
    const TFheGateBootstrappingParameterSet* P = key->params;
    const int n = P->in_out_params->n;
    // input bits and gate codes come from a separate generator so that the
    // library's global generator is consumed only by key generation and encryption
    std::mt19937_64 bitgen(0x5eed0000ULL + seed);

    if (cmd == "files" && argc >= 5) {
        const std::string dir = argv[3];
        const int count = atoi(argv[4]);
        FILE* f = open_or_die(dir + "/params.gb", "wb");
        export_tfheGateBootstrappingParameterSet_toFile(f, P);
        fclose(f);
        f = open_or_die(dir + "/secret.key", "wb");
        export_tfheGateBootstrappingSecretKeySet_toFile(f, key);
        fclose(f);
        f = open_or_die(dir + "/cloud.key", "wb");
        export_tfheGateBootstrappingCloudKeySet_toFile(f, &key->cloud);
        fclose(f);
        f = open_or_die(dir + "/lwekey.bin", "wb");
        export_lweKey_toFile(f, key->lwe_key);
        fclose(f);
        LweSample* ct = new_gate_bootstrapping_ciphertext_array(count, P);
        f = open_or_die(dir + "/cts.bin", "wb");
        FILE* fb = open_or_die(dir + "/bits.txt", "w");
        for (int g = 0; g < count; g++) {
            const int32_t bit = (int32_t)(bitgen() & 1);
            bootsSymEncrypt(&ct[g], bit, key);
            export_gate_bootstrapping_ciphertext_toFile(f, &ct[g], P);
            fprintf(fb, "%d", bit);
        }
        fclose(f);
        fclose(fb);
        LweSample* r = new_gate_bootstrapping_ciphertext(P);
        bootsNAND(r, &ct[0], &ct[1], &key->cloud);
        f = open_or_die(dir + "/nand.bin", "wb");
        export_gate_bootstrapping_ciphertext_toFile(f, r, P);
        fclose(f);
        return 0;
    }
    if (cmd == "readfiles" && argc >= 5) {
        const std::string dir = argv[3];
        const int count = atoi(argv[4]);
        FILE* f = open_or_die(dir + "/params.gb", "rb");
        TFheGateBootstrappingParameterSet* p2 = new_tfheGateBootstrappingParameterSet_fromFile(f);
        fclose(f);
        f = open_or_die(dir + "/secret.key", "rb");
        TFheGateBootstrappingSecretKeySet* sk = new_tfheGateBootstrappingSecretKeySet_fromFile(f);
        fclose(f);
        f = open_or_die(dir + "/cloud.key", "rb");
        TFheGateBootstrappingCloudKeySet* ck = new_tfheGateBootstrappingCloudKeySet_fromFile(f);
        fclose(f);
        LweSample* ct = new_gate_bootstrapping_ciphertext_array(count, p2);
        f = open_or_die(dir + "/cts.bin", "rb");
        printf("bits ");
        for (int g = 0; g < count; g++) {
            import_gate_bootstrapping_ciphertext_fromFile(f, &ct[g], p2);
            printf("%d", bootsSymDecrypt(&ct[g], sk));
        }
        fclose(f);
        LweSample* r = new_gate_bootstrapping_ciphertext(p2);
        bootsNAND(r, &ct[0], &ct[1], ck);
        printf("\nnand %d\n", bootsSymDecrypt(r, sk));
        printf("params %d %d %d %d %d %d %d %.10g %.10g %.10g %.10g\n", p2->in_out_params->n,
               p2->tgsw_params->tlwe_params->N, p2->tgsw_params->tlwe_params->k, p2->tgsw_params->l,
               p2->tgsw_params->Bgbit, p2->ks_t, p2->ks_basebit, p2->in_out_params->alpha_min,
               p2->in_out_params->alpha_max, p2->tgsw_params->tlwe_params->alpha_min,
               p2->tgsw_params->tlwe_params->alpha_max);
        // key hashes over the cloud key read from cloud.key
        const LweKeySwitchKey* ks = ck->bk->ks;
        const int n2 = p2->in_out_params->n;
        uint64_t hk = 1469598103934665603ULL;
        for (int i = 0; i < ks->n * ks->t * ks->base; i++) {
            hk = fnv(ks->ks0_raw[i].a, sizeof(int32_t) * n2, hk);
            hk = fnv(&ks->ks0_raw[i].b, sizeof(int32_t), hk);
        }
        uint64_t hb = 1469598103934665603ULL;
        const int kpl = p2->tgsw_params->kpl, N2 = p2->tgsw_params->tlwe_params->N, k2 = p2->tgsw_params->tlwe_params->k;
        for (int i = 0; i < n2; i++)
            for (int j = 0; j < kpl; j++)
                for (int c = 0; c <= k2; c++) hb = fnv(ck->bk->bk[i].all_sample[j].a[c].coefsT, sizeof(int32_t) * N2, hb);
        printf("ks_hash %llu\nbk_hash %llu\n", (unsigned long long)hk, (unsigned long long)hb);
        return 0;
    }
    if (cmd == "keys" && argc >= 4) {
        dump_keys(key, argv[3]);
        return 0;
    }
    if (cmd == "gates" && argc >= 6) {
        const int count = atoi(argv[3]);
        const std::string ops = argv[4];  // comma-free list of op codes as letters 'a'+code, or "rand6"
        FILE* f = fopen(argv[5], "wb");
        int32_t hdr[2] = {count, n};
        write_all(f, hdr, sizeof hdr);
        LweSample* ct = new_gate_bootstrapping_ciphertext_array(4, P);
        for (int g = 0; g < count; g++) {
            int op;
            if (ops == "rand6") op = (int)(bitgen() % 6);
            else op = ops[g % ops.size()] - 'a';
            int32_t bits[3];
            for (int j = 0; j < 3; j++) bits[j] = (int32_t)(bitgen() & 1);
            const int nin = (op == G_MUX) ? 3 : 2;
            for (int j = 0; j < nin; j++) bootsSymEncrypt(&ct[j], bits[j], key);
            if (nin == 2) lweNoiselessTrivial(&ct[2], 0, P->in_out_params);
            auto t0 = std::chrono::steady_clock::now();
            run_gate(op, &ct[3], &ct[0], &ct[1], &ct[2], &key->cloud);
            auto t1 = std::chrono::steady_clock::now();
            int32_t rec[4] = {op, bits[0], bits[1], bits[2]};
            write_all(f, rec, sizeof rec);
            for (int j = 0; j < 4; j++) put_sample(f, &ct[j], n);
            int32_t ph = lwePhase(&ct[3], key->lwe_key);
            write_all(f, &ph, sizeof ph);
            int64_t ns = std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
            write_all(f, &ns, sizeof ns);
        }
        fclose(f);
        return 0;
    }
    if (cmd == "ks" && argc >= 5) {
        const int count = atoi(argv[3]);
        const LweKeySwitchKey* ks = key->cloud.bk->ks;
        const LweParams* ext = &P->tgsw_params->tlwe_params->extracted_lweparams;
        LweSample* in = new_LweSample(ext);
        LweSample* out = new_LweSample(P->in_out_params);
        FILE* f = fopen(argv[4], "wb");
        int32_t hdr[3] = {count, ext->n, n};
        write_all(f, hdr, sizeof hdr);
        for (int g = 0; g < count; g++) {
            for (int i = 0; i < ext->n; i++) in->a[i] = (int32_t)(uint32_t)bitgen();
            in->b = (int32_t)(uint32_t)bitgen();
            lweKeySwitch(out, ks, in);
            put_sample(f, in, ext->n);
            put_sample(f, out, n);
            write_all(f, &out->current_variance, sizeof(double));
        }
        fclose(f);
        return 0;
    }
    if (cmd == "bench" && argc >= 5) {
        const int count = atoi(argv[3]);
        const int threads = std::max(1, atoi(argv[4]));
        std::vector<LweSample*> ins;
        for (int g = 0; g < 2 * count; g++) {
            LweSample* s = new_gate_bootstrapping_ciphertext(P);
            bootsSymEncrypt(s, (int32_t)(bitgen() & 1), key);
            ins.push_back(s);
        }
        std::vector<double> lat(count);
        auto work = [&](int tid) {
            LweSample* out = new_gate_bootstrapping_ciphertext(P);
            for (int g = tid; g < count; g += threads) {
                auto t0 = std::chrono::steady_clock::now();
                bootsNAND(out, ins[2 * g], ins[2 * g + 1], &key->cloud);
                auto t1 = std::chrono::steady_clock::now();
                lat[g] = std::chrono::duration<double>(t1 - t0).count();
            }
            delete_gate_bootstrapping_ciphertext(out);
        };
        auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++) th.emplace_back(work, t);
        for (auto& t : th) t.join();
        double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<double> s = lat;
        std::sort(s.begin(), s.end());
        printf("{\"gates\": %d, \"threads\": %d, \"wall_s\": %.6f, \"gates_per_s\": %.3f, "
               "\"p50_ms\": %.4f, \"p99_ms\": %.4f}\n",
               count, threads, wall, count / wall, 1e3 * s[count / 2], 1e3 * s[(size_t)(0.99 * (count - 1))]);
        return 0;
    }
    fprintf(stderr, "bad command\n");
    return 2;
}
