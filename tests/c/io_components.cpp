// Test program for the file and stream functions of include/tfhe_amd.h (the
// reference's include/tfhe_io.h:30-496), run by tests/test_io_cpu.py.
//
//   io_components GOLD OUT
//
// GOLD holds files the reference wrote for one key set. Every object is read
// with the FILE* reader and written back with the stream writer, and read with
// the stream reader and written back with the FILE* writer; both copies must
// equal the reference's bytes (checked here for the small files, and by the
// caller, against the committed digests, for the large ones written to OUT).
// Prints "ok <what>" per check; exits non-zero at the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "tfhe_amd.h"

static std::string G, O;

static std::string slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        exit(2);
    }
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}
static void expect(bool ok, const std::string& what) {
    if (!ok) {
        fprintf(stderr, "MISMATCH %s (%s)\n", what.c_str(), tfhe_amd_last_error());
        exit(1);
    }
    printf("ok %s\n", what.c_str());
}
static FILE* fopen_or_die(const std::string& path, const char* mode) {
    FILE* f = fopen(path.c_str(), mode);
    if (!f) {
        perror(path.c_str());
        exit(2);
    }
    return f;
}

// For an object read one way and written the other: `read_file` reads GOLD/name
// with FILE*, `write_stream` writes it to a stream; `read_stream` reads the
// bytes from a stream, `write_file` writes the object to OUT/name. Both results
// must equal GOLD/name.
template <class T, class RF, class WS, class RS, class WF>
static T* both_ways(const std::string& name, RF read_file, WS write_stream, RS read_stream, WF write_file) {
    const std::string ref = slurp(G + "/" + name);
    FILE* f = fopen_or_die(G + "/" + name, "rb");
    T* a = read_file(f);
    fclose(f);
    expect(a != nullptr, name + " read from FILE");
    std::ostringstream os;
    write_stream(os, a);
    expect(os.str() == ref, name + " FILE reader -> stream writer, same bytes");
    std::istringstream is(ref);
    T* b = read_stream(is);
    expect(b != nullptr, name + " read from stream");
    f = fopen_or_die(O + "/" + name, "wb");
    write_file(f, b);
    fclose(f);
    expect(slurp(O + "/" + name) == ref, name + " stream reader -> FILE writer, same bytes");
    return b;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    G = argv[1];
    O = argv[2];

    LweParams* lp = both_ways<LweParams>(
        "lweparams.txt", new_lweParams_fromFile, export_lweParams_toStream, new_lweParams_fromStream,
        export_lweParams_toFile);
    TLweParams* tp = both_ways<TLweParams>(
        "tlweparams.txt", new_tLweParams_fromFile, export_tLweParams_toStream, new_tLweParams_fromStream,
        export_tLweParams_toFile);
    TGswParams* gp = both_ways<TGswParams>(
        "tgswparams.txt", new_tGswParams_fromFile, export_tGswParams_toStream, new_tGswParams_fromStream,
        export_tGswParams_toFile);
    TFheGateBootstrappingParameterSet* P = both_ways<TFheGateBootstrappingParameterSet>(
        "params.gb", new_tfheGateBootstrappingParameterSet_fromFile, export_tfheGateBootstrappingParameterSet_toStream,
        new_tfheGateBootstrappingParameterSet_fromStream, export_tfheGateBootstrappingParameterSet_toFile);
    expect(lp->n == P->in_out_params->n && tp->N == P->tgsw_params->tlwe_params->N && gp->l == P->tgsw_params->l &&
               gp->kpl == P->tgsw_params->kpl,
           "parameters agree with the gate parameter set");

    LweKey* lk = both_ways<LweKey>("lwekey.bin", new_lweKey_fromFile, export_lweKey_toStream, new_lweKey_fromStream,
                                   export_lweKey_toFile);
    TLweKey* tk = both_ways<TLweKey>("tlwekey.bin", new_tlweKey_fromFile, export_tlweKey_toStream,
                                     new_tlweKey_fromStream, export_tlweKey_toFile);
    TGswKey* gk = both_ways<TGswKey>("tgswkey.bin", new_tgswKey_fromFile, export_tgswKey_toStream,
                                     new_tgswKey_fromStream, export_tgswKey_toFile);
    expect(memcmp(tk->key[0].coefs, gk->key[0].coefs, sizeof(int32_t) * tp->N) == 0,
           "TLWE key and TGSW key hold the same polynomial");

    // samples: imported into objects allocated with the parameters read above
    auto sample_both = [&](const std::string& name, auto alloc, auto imp_f, auto exp_s, auto imp_s, auto exp_f,
                           auto params) {
        const std::string ref = slurp(G + "/" + name);
        auto* a = alloc(params);
        FILE* f = fopen_or_die(G + "/" + name, "rb");
        imp_f(f, a, params);
        fclose(f);
        std::ostringstream os;
        exp_s(os, a, params);
        expect(os.str() == ref, name + " FILE importer -> stream writer, same bytes");
        auto* b = alloc(params);
        std::istringstream is(ref);
        imp_s(is, b, params);
        f = fopen_or_die(O + "/" + name, "wb");
        exp_f(f, b, params);
        fclose(f);
        expect(slurp(O + "/" + name) == ref, name + " stream importer -> FILE writer, same bytes");
    };
    sample_both("lwesample.bin", new_LweSample, import_lweSample_fromFile, export_lweSample_toStream,
                import_lweSample_fromStream, export_lweSample_toFile, (const LweParams*)lp);
    sample_both("tlwesample.bin", new_TLweSample, import_tlweSample_fromFile, export_tlweSample_toStream,
                import_tlweSample_fromStream, export_tlweSample_toFile, (const TLweParams*)tp);
    sample_both("tgswsample.bin", new_TGswSample, import_tgswSample_fromFile, export_tgswSample_toStream,
                import_tgswSample_fromStream, export_tgswSample_toFile, (const TGswParams*)gp);

    // the secret key set, then its parts written as separate objects into OUT
    const std::string sref = slurp(G + "/secret.key");
    std::istringstream sis(sref);
    TFheGateBootstrappingSecretKeySet* sk = new_tfheGateBootstrappingSecretKeySet_fromStream(sis);
    expect(sk != nullptr, "secret.key read from stream");
    expect(memcmp(sk->lwe_key->key, lk->key, sizeof(int32_t) * lp->n) == 0, "LWE key of secret.key = lwekey.bin");
    expect(memcmp(sk->tgsw_key->key[0].coefs, gk->key[0].coefs, sizeof(int32_t) * tp->N) == 0,
           "TGSW key of secret.key = tgswkey.bin");
    FILE* f = fopen_or_die(O + "/secret.key", "wb");
    export_tfheGateBootstrappingSecretKeySet_toFile(f, sk);
    fclose(f);
    expect(slurp(O + "/secret.key") == sref, "secret.key stream reader -> FILE writer, same bytes");
    {
        std::ofstream o(O + "/cloud.key", std::ios::binary);
        export_tfheGateBootstrappingCloudKeySet_toStream(o, &sk->cloud);
    }
    f = fopen_or_die(O + "/kskey.bin", "wb");
    export_lweKeySwitchKey_toFile(f, sk->cloud.bk->ks);
    fclose(f);
    {
        std::ofstream o(O + "/bk.bin", std::ios::binary);
        export_lweBootstrappingKey_toStream(o, sk->cloud.bk);
    }
    // and read back the other way: stream <-> FILE
    {
        std::ifstream i(O + "/kskey.bin", std::ios::binary);
        LweKeySwitchKey* ks = new_lweKeySwitchKey_fromStream(i);
        expect(ks != nullptr, "kskey.bin read from stream");
        f = fopen_or_die(O + "/kskey2.bin", "wb");
        export_lweKeySwitchKey_toFile(f, ks);
        fclose(f);
        expect(slurp(O + "/kskey2.bin") == slurp(O + "/kskey.bin"), "kskey.bin rewritten, same bytes");
        delete_LweKeySwitchKey(ks);
    }
    {
        f = fopen_or_die(O + "/bk.bin", "rb");
        LweBootstrappingKey* bk = new_lweBootstrappingKey_fromFile(f);
        fclose(f);
        expect(bk != nullptr, "bk.bin read from FILE");
        std::ofstream o(O + "/bk2.bin", std::ios::binary);
        export_lweBootstrappingKey_toStream(o, bk);
        o.close();
        expect(slurp(O + "/bk2.bin") == slurp(O + "/bk.bin"), "bk.bin rewritten, same bytes");
        LweBootstrappingKeyFFT* bf = new_LweBootstrappingKeyFFT(bk);
        expect(bf != nullptr && bf->ks == bk->ks && bf->in_out_params == bk->in_out_params && bf->bkFFT != nullptr,
               "transform-domain handle over a loaded bootstrapping key");
        delete_LweBootstrappingKeyFFT(bf);
        delete_LweBootstrappingKey(bk);
    }
    {
        f = fopen_or_die(O + "/cloud.key", "rb");
        TFheGateBootstrappingCloudKeySet* ck = new_tfheGateBootstrappingCloudKeySet_fromFile(f);
        fclose(f);
        expect(ck != nullptr, "cloud.key read from FILE");
        f = fopen_or_die(O + "/cloud2.key", "wb");
        export_tfheGateBootstrappingCloudKeySet_toFile(f, ck);
        fclose(f);
        expect(slurp(O + "/cloud2.key") == slurp(O + "/cloud.key"), "cloud.key rewritten, same bytes");
        expect(sref.compare(0, slurp(O + "/cloud.key").size(), slurp(O + "/cloud.key")) == 0,
               "cloud.key is the leading part of secret.key");
        delete_gate_bootstrapping_cloud_keyset(ck);
    }
    // malformed input to a stream reader: NULL and a reason, no abort
    {
        std::istringstream bad("-----BEGIN TLWEPARAMS-----\nN: 1000\nk: 1\nalpha_min: 0.1\nalpha_max: 0.1\n"
                               "-----END TLWEPARAMS-----\n");
        expect(new_tLweParams_fromStream(bad) == nullptr && strstr(tfhe_amd_last_error(), "power of two"),
               "stream reader rejects N = 1000");
        std::istringstream cut(sref.substr(0, sref.size() / 2));
        expect(new_tfheGateBootstrappingSecretKeySet_fromStream(cut) == nullptr &&
                   strstr(tfhe_amd_last_error(), "truncated"),
               "stream reader rejects a truncated secret key");
    }
    delete_gate_bootstrapping_secret_keyset(sk);
    delete_LweKey(lk);
    delete_TLweKey(tk);
    delete_TGswKey(gk);
    delete_LweParams(lp);
    delete_TLweParams(tp);
    delete_TGswParams(gp);
    printf("done\n");
    return 0;
}
