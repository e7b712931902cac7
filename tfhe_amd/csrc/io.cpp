// Key, parameter and ciphertext files in the reference's wire format
// (libtfhe/tfhe_io.cpp, libtfhe/tfhe_generic_streams.cpp), so key sets and
// ciphertexts written by either library load in the other.
//
// Format: parameter sections are text blocks
//     -----BEGIN <TITLE>-----
//     <name>: <value>            (names in byte order; integers "%10d"-style,
//     -----END <TITLE>-----       doubles "%.8f" -- the reference rounds the
//                                 noise parameters to 8 decimals on export)
// and key material is binary, native byte order, each object led by its type
// id (include/tfhe_generic_streams.h:15-30): LWE sample 42 (a[n], b, double
// variance), LWE key 43 (n ints), TGSW key 169 (k*N ints), keyswitch key 200
// (max variance, then [k*N][t][base] rows of a[n], b), bootstrapping key 201
// (max variance, then [n][kpl][k+1][N] coefficients).
// Files: gate parameters = GATEBOOTSPARAMS, LWEPARAMS, TLWEPARAMS, TGSWPARAMS
// (tfhe_io.cpp:1018-1046); cloud key set = gate parameters, LWEKSPARAMS,
// keyswitch key, bootstrapping key (:1087-1111, :937-944); secret key set =
// cloud key set followed by the LWE key and the TGSW key (:1148-1158).
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "engine.h"

namespace tfhe_amd {
namespace {

constexpr int32_t UID_LWE_SAMPLE = 42;
constexpr int32_t UID_LWE_KEY = 43;
constexpr int32_t UID_TGSW_KEY = 169;
constexpr int32_t UID_KS_KEY = 200;
constexpr int32_t UID_BK_KEY = 201;

// ---- text sections

struct Section {
    std::string title;
    std::map<std::string, std::string> props;  // written in key order, as a std::map
    void set_int(const char* name, int64_t v) {
        char buf[64];
        snprintf(buf, sizeof buf, "%10" PRId64, v);
        props[name] = buf;
    }
    void set_double(const char* name, double v) {
        char buf[64];
        snprintf(buf, sizeof buf, "%.8f", v);
        props[name] = buf;
    }
    bool has(const char* name) const { return props.count(name) != 0; }
    int64_t get_int(const char* name) const { return strtoll(props.at(name).c_str(), nullptr, 10); }
    double get_double(const char* name) const { return (double)strtold(props.at(name).c_str(), nullptr); }
};

void write_section(FILE* F, const Section& s) {
    fprintf(F, "-----BEGIN %s-----\n", s.title.c_str());
    for (const auto& kv : s.props) fprintf(F, "%s: %s\n", kv.first.c_str(), kv.second.c_str());
    fprintf(F, "-----END %s-----\n", s.title.c_str());
}

bool read_line(FILE* F, std::string& line) {
    line.clear();
    int c;
    bool any = false;
    while ((c = fgetc(F)) != EOF) {
        any = true;
        if (c == '\r') continue;
        if (c == '\n') return true;
        line.push_back((char)c);
    }
    return any;
}

// Reads the next section; lines before its BEGIN marker are skipped, as the
// reference does (tfhe_generic_streams.cpp:119-158).
bool read_section(FILE* F, const char* title, Section& s) {
    std::string line, end;
    bool started = false;
    while (read_line(F, line)) {
        const size_t n = line.size();
        if (n >= 16 && !line.compare(0, 11, "-----BEGIN ") && !line.compare(n - 5, 5, "-----")) {
            s.title = line.substr(11, n - 16);
            end = "-----END " + s.title + "-----";
            started = true;
            continue;
        }
        if (!started) continue;
        if (line == end) {
            if (s.title != title) {
                set_error("key file: expected section %s, found %s", title, s.title.c_str());
                return false;
            }
            return true;
        }
        const size_t pos = line.find(": ");
        if (pos == std::string::npos) continue;
        s.props[line.substr(0, pos)] = line.substr(pos + 2);
    }
    set_error("key file: section %s missing or truncated", title);
    return false;
}

bool need(const Section& s, std::initializer_list<const char*> names) {
    for (const char* nm : names)
        if (!s.has(nm)) {
            set_error("key file: section %s lacks %s", s.title.c_str(), nm);
            return false;
        }
    return true;
}

// ---- binary blocks

bool rd(FILE* F, void* p, size_t bytes) {
    if (fread(p, 1, bytes, F) == bytes) return true;
    set_error("key file: truncated binary block");
    return false;
}
bool rd_uid(FILE* F, int32_t want) {
    int32_t uid = -1;
    if (!rd(F, &uid, sizeof uid)) return false;
    if (uid != want) {
        set_error("key file: object type %d where %d was expected", uid, want);
        return false;
    }
    return true;
}
void wr(FILE* F, const void* p, size_t bytes) { fwrite(p, 1, bytes, F); }

// ---- parameter sections

Section lwe_params_section(const LweParams* p) {
    Section s;
    s.title = "LWEPARAMS";
    s.set_int("n", p->n);
    s.set_double("alpha_min", p->alpha_min);
    s.set_double("alpha_max", p->alpha_max);
    return s;
}

void write_gate_params(FILE* F, const TFheGateBootstrappingParameterSet* P) {
    Section g;
    g.title = "GATEBOOTSPARAMS";
    g.set_int("ks_t", P->ks_t);
    g.set_int("ks_basebit", P->ks_basebit);
    write_section(F, g);
    write_section(F, lwe_params_section(P->in_out_params));
    const TLweParams* tp = P->tgsw_params->tlwe_params;
    Section t;
    t.title = "TLWEPARAMS";
    t.set_int("N", tp->N);
    t.set_int("k", tp->k);
    t.set_double("alpha_min", tp->alpha_min);
    t.set_double("alpha_max", tp->alpha_max);
    write_section(F, t);
    Section w;
    w.title = "TGSWPARAMS";
    w.set_int("l", P->tgsw_params->l);
    w.set_int("Bgbit", P->tgsw_params->Bgbit);
    write_section(F, w);
}

TFheGateBootstrappingParameterSet* read_gate_params(FILE* F) {
    Section g, l, t, w;
    if (!read_section(F, "GATEBOOTSPARAMS", g) || !need(g, {"ks_t", "ks_basebit"})) return nullptr;
    if (!read_section(F, "LWEPARAMS", l) || !need(l, {"n", "alpha_min", "alpha_max"})) return nullptr;
    if (!read_section(F, "TLWEPARAMS", t) || !need(t, {"N", "k", "alpha_min", "alpha_max"})) return nullptr;
    if (!read_section(F, "TGSWPARAMS", w) || !need(w, {"l", "Bgbit"})) return nullptr;
    return host_make_params((int32_t)l.get_int("n"), l.get_double("alpha_min"), l.get_double("alpha_max"),
                            (int32_t)t.get_int("N"), (int32_t)t.get_int("k"), t.get_double("alpha_min"),
                            t.get_double("alpha_max"), (int32_t)w.get_int("l"), (int32_t)w.get_int("Bgbit"),
                            (int32_t)g.get_int("ks_t"), (int32_t)g.get_double("ks_basebit"));
}

// ---- keys

struct KeyArrays {
    std::vector<int32_t> ksk, bk;
    double ks_var = 0, bk_var = 0;
};

// LWEKSPARAMS section, keyswitch key, bootstrapping key (tfhe_io.cpp:937-944)
void write_bootstrapping_key(FILE* F, const LweBootstrappingKey* bk) {
    const LweKeySwitchKey* ks = bk->ks;
    Section s;
    s.title = "LWEKSPARAMS";
    s.set_int("n", ks->n);
    s.set_int("t", ks->t);
    s.set_int("basebit", ks->basebit);
    write_section(F, s);
    const int n = ks->out_params->n;
    const size_t rows = (size_t)ks->n * ks->t * ks->base;
    double var = -1;
    for (size_t r = 0; r < rows; r++) var = ks->ks0_raw[r].current_variance > var ? ks->ks0_raw[r].current_variance : var;
    wr(F, &UID_KS_KEY, sizeof(int32_t));
    wr(F, &var, sizeof var);
    for (size_t r = 0; r < rows; r++) {
        wr(F, ks->ks0_raw[r].a, sizeof(int32_t) * n);
        wr(F, &ks->ks0_raw[r].b, sizeof(int32_t));
    }
    const int nn = bk->in_out_params->n, kpl = bk->bk_params->kpl, k = bk->accum_params->k, N = bk->accum_params->N;
    double bvar = -1;
    for (int i = 0; i < nn; i++)
        for (int j = 0; j < kpl; j++)
            bvar = bk->bk[i].all_sample[j].current_variance > bvar ? bk->bk[i].all_sample[j].current_variance : bvar;
    wr(F, &UID_BK_KEY, sizeof(int32_t));
    wr(F, &bvar, sizeof bvar);
    for (int i = 0; i < nn; i++)
        for (int j = 0; j < kpl; j++)
            for (int c = 0; c <= k; c++) wr(F, bk->bk[i].all_sample[j].a[c].coefsT, sizeof(int32_t) * N);
}

bool read_bootstrapping_key(FILE* F, const TFheGateBootstrappingParameterSet* P, KeyArrays& K) {
    Section s;
    if (!read_section(F, "LWEKSPARAMS", s) || !need(s, {"n", "t", "basebit"})) return false;
    const TLweParams* tp = P->tgsw_params->tlwe_params;
    const int n = P->in_out_params->n, N = tp->N, k = tp->k, kpl = P->tgsw_params->kpl;
    if (s.get_int("n") != (int64_t)N * k || s.get_int("t") != P->ks_t || s.get_int("basebit") != P->ks_basebit) {
        set_error("key file: keyswitch parameters do not match the gate parameters");
        return false;
    }
    const size_t rows = (size_t)N * k * P->ks_t * (1 << P->ks_basebit);
    K.ksk.resize(rows * (n + 1));
    if (!rd_uid(F, UID_KS_KEY) || !rd(F, &K.ks_var, sizeof(double)) ||
        !rd(F, K.ksk.data(), sizeof(int32_t) * K.ksk.size()))
        return false;
    K.bk.resize((size_t)n * kpl * (k + 1) * N);
    return rd_uid(F, UID_BK_KEY) && rd(F, &K.bk_var, sizeof(double)) &&
           rd(F, K.bk.data(), sizeof(int32_t) * K.bk.size());
}

FILE* check_file(FILE* F, const char* what) {
    if (!F) set_error("%s: null FILE", what);
    return F;
}

}  // namespace
}  // namespace tfhe_amd

using namespace tfhe_amd;

// ------------------------------------------------ gate-level files (C API) --

TFHE_AMD_EXPORT void export_tfheGateBootstrappingParameterSet_toFile(FILE* F,
                                                                     const TFheGateBootstrappingParameterSet* params) {
    if (!check_file(F, "export parameters") || !params) die("export_tfheGateBootstrappingParameterSet_toFile: bad arguments");
    write_gate_params(F, params);
}

TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromFile(FILE* F) {
    if (!check_file(F, "read parameters")) return nullptr;
    return read_gate_params(F);
}

TFHE_AMD_EXPORT void export_tfheGateBootstrappingCloudKeySet_toFile(FILE* F, const TFheGateBootstrappingCloudKeySet* key) {
    if (!check_file(F, "export cloud key") || !key) die("export_tfheGateBootstrappingCloudKeySet_toFile: bad arguments");
    write_gate_params(F, key->params);
    write_bootstrapping_key(F, key->bk);
}

TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromFile(FILE* F) {
    if (!check_file(F, "read cloud key")) return nullptr;
    TFheGateBootstrappingParameterSet* P = read_gate_params(F);
    if (!P) return nullptr;
    KeyArrays K;
    if (!read_bootstrapping_key(F, P, K)) return nullptr;
    return host_cloud_from_arrays(P, K.bk.data(), K.ksk.data(), K.ks_var, K.bk_var);
}

TFHE_AMD_EXPORT void export_tfheGateBootstrappingSecretKeySet_toFile(FILE* F,
                                                                     const TFheGateBootstrappingSecretKeySet* key) {
    if (!check_file(F, "export secret key") || !key) die("export_tfheGateBootstrappingSecretKeySet_toFile: bad arguments");
    write_gate_params(F, key->params);
    write_bootstrapping_key(F, key->cloud.bk);
    const int n = key->params->in_out_params->n;
    wr(F, &UID_LWE_KEY, sizeof(int32_t));
    wr(F, key->lwe_key->key, sizeof(int32_t) * n);
    const TLweParams* tp = key->params->tgsw_params->tlwe_params;
    wr(F, &UID_TGSW_KEY, sizeof(int32_t));
    for (int i = 0; i < tp->k; i++) wr(F, key->tgsw_key->key[i].coefs, sizeof(int32_t) * tp->N);
}

TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromFile(FILE* F) {
    if (!check_file(F, "read secret key")) return nullptr;
    TFheGateBootstrappingParameterSet* P = read_gate_params(F);
    if (!P) return nullptr;
    KeyArrays K;
    if (!read_bootstrapping_key(F, P, K)) return nullptr;
    const TLweParams* tp = P->tgsw_params->tlwe_params;
    std::vector<int32_t> lwe(P->in_out_params->n), tlwe((size_t)tp->k * tp->N);
    if (!rd_uid(F, UID_LWE_KEY) || !rd(F, lwe.data(), sizeof(int32_t) * lwe.size()) || !rd_uid(F, UID_TGSW_KEY) ||
        !rd(F, tlwe.data(), sizeof(int32_t) * tlwe.size()))
        return nullptr;
    return host_secret_from_arrays(P, K.bk.data(), K.ksk.data(), K.ks_var, K.bk_var, lwe.data(), tlwe.data());
}

// LWE samples: type id, a[n], b, variance (tfhe_io.cpp:90-108)
TFHE_AMD_EXPORT void export_lweSample_toFile(FILE* F, const LweSample* sample, const LweParams* params) {
    if (!check_file(F, "export sample") || !sample || !params) die("export_lweSample_toFile: bad arguments");
    wr(F, &UID_LWE_SAMPLE, sizeof(int32_t));
    wr(F, sample->a, sizeof(int32_t) * params->n);
    wr(F, &sample->b, sizeof(int32_t));
    wr(F, &sample->current_variance, sizeof(double));
}

TFHE_AMD_EXPORT void import_lweSample_fromFile(FILE* F, LweSample* sample, const LweParams* params) {
    if (!check_file(F, "import sample") || !sample || !params) die("import_lweSample_fromFile: bad arguments");
    if (!rd_uid(F, UID_LWE_SAMPLE) || !rd(F, sample->a, sizeof(int32_t) * params->n) ||
        !rd(F, &sample->b, sizeof(int32_t)) || !rd(F, &sample->current_variance, sizeof(double)))
        die(tfhe_amd_last_error());
}

TFHE_AMD_EXPORT void export_gate_bootstrapping_ciphertext_toFile(FILE* F, const LweSample* sample,
                                                                 const TFheGateBootstrappingParameterSet* params) {
    export_lweSample_toFile(F, sample, params ? params->in_out_params : nullptr);
}

TFHE_AMD_EXPORT void import_gate_bootstrapping_ciphertext_fromFile(FILE* F, LweSample* sample,
                                                                   const TFheGateBootstrappingParameterSet* params) {
    import_lweSample_fromFile(F, sample, params ? params->in_out_params : nullptr);
}

// LWE parameters and LWE keys on their own (tfhe_io.cpp:35-60, 154-205)
TFHE_AMD_EXPORT void export_lweParams_toFile(FILE* F, const LweParams* params) {
    if (!check_file(F, "export LWE parameters") || !params) die("export_lweParams_toFile: bad arguments");
    write_section(F, lwe_params_section(params));
}

TFHE_AMD_EXPORT LweParams* new_lweParams_fromFile(FILE* F) {
    Section s;
    if (!check_file(F, "read LWE parameters") || !read_section(F, "LWEPARAMS", s) ||
        !need(s, {"n", "alpha_min", "alpha_max"}))
        return nullptr;
    return new_LweParams((int32_t)s.get_int("n"), s.get_double("alpha_min"), s.get_double("alpha_max"));
}

TFHE_AMD_EXPORT void export_lweKey_toFile(FILE* F, const LweKey* key) {
    if (!check_file(F, "export LWE key") || !key) die("export_lweKey_toFile: bad arguments");
    write_section(F, lwe_params_section(key->params));
    wr(F, &UID_LWE_KEY, sizeof(int32_t));
    wr(F, key->key, sizeof(int32_t) * key->params->n);
}

// The key owns its parameters; free with delete_LweKey.
TFHE_AMD_EXPORT LweKey* new_lweKey_fromFile(FILE* F) {
    LweParams* p = new_lweParams_fromFile(F);
    if (!p) return nullptr;
    auto* key = new LweKey{p, new int32_t[p->n]};
    if (!rd_uid(F, UID_LWE_KEY) || !rd(F, key->key, sizeof(int32_t) * p->n)) {
        delete[] key->key;
        delete key;
        delete p;
        return nullptr;
    }
    return key;
}

TFHE_AMD_EXPORT void delete_LweKey(LweKey* key) {
    if (!key) return;
    delete[] key->key;
    delete key->params;
    delete key;
}
