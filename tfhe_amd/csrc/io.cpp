// Key, parameter, sample and ciphertext files in the reference's wire format
// (libtfhe/tfhe_io.cpp, libtfhe/tfhe_generic_streams.cpp), so objects written
// by either library load in the other, through C FILE* or C++ std::iostream.
//
// Format: parameter sections are text blocks
//     -----BEGIN <TITLE>-----
//     <name>: <value>            (names in byte order; integers "%10d"-style,
//     -----END <TITLE>-----       doubles "%.8f" -- the reference rounds the
//                                 noise parameters to 8 decimals on export)
// and key material is binary, native byte order, each object led by its type
// id (include/tfhe_generic_streams.h:15-30): LWE sample 42 (a[n], b, double
// variance), LWE key 43 (n ints), TLWE sample 84 ((k+1) polynomials of N,
// variance), TLWE key 85 (k N ints), TGSW sample 168 ((k+1) l TLWE samples),
// TGSW key 169 (k N ints), keyswitch key 200 (max variance, then
// [k N][t][base] rows of a[n], b), bootstrapping key 201 (max variance, then
// [n][kpl][k+1][N] coefficients).
// Composite objects (tfhe_io.cpp): LWE key = LWEPARAMS + key; TLWE key =
// TLWEPARAMS + key; TGSW parameters = TLWEPARAMS + TGSWPARAMS; TGSW key =
// TGSW parameters + key; keyswitch key = LWEPARAMS (output) + LWEKSPARAMS +
// content (:819-854); bootstrapping key = LWEPARAMS + TGSW parameters +
// LWEKSPARAMS + keyswitch content + bootstrapping content (:937-944); gate
// parameters = GATEBOOTSPARAMS + LWEPARAMS + TGSW parameters (:1018-1046); cloud
// key set = gate parameters + LWEKSPARAMS + the two contents (:1087-1111);
// secret key set = cloud key set + LWE key content + TGSW key content
// (:1148-1158).
//
// Readers check every size they parse against the limits below before
// allocating; malformed input makes the constructors return NULL (reason in
// tfhe_amd_last_error) and the importers into caller-allocated objects abort,
// as the reference does.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <istream>
#include <map>
#include <new>
#include <ostream>
#include <string>
#include <vector>

#include "engine.h"

namespace tfhe_amd {
namespace {

constexpr int32_t UID_LWE_SAMPLE = 42;
constexpr int32_t UID_LWE_KEY = 43;
constexpr int32_t UID_TLWE_SAMPLE = 84;
constexpr int32_t UID_TLWE_KEY = 85;
constexpr int32_t UID_TGSW_SAMPLE = 168;
constexpr int32_t UID_TGSW_KEY = 169;
constexpr int32_t UID_KS_KEY = 200;
constexpr int32_t UID_BK_KEY = 201;

// Limits on parsed sizes (far above any parameter set in use; they keep a
// corrupted header from asking for absurd allocations or overflowing sizes).
constexpr int64_t MAX_LWE_N = 1 << 16;
constexpr int64_t MAX_POLY_N = 1 << 16;
constexpr int64_t MAX_K = 8;
constexpr size_t MAX_WORDS = (size_t)1 << 31;  // 8 GiB of int32 in one object

// ---- byte sinks and sources over FILE* or std::ostream / std::istream

struct Out {
    virtual void write(const void* p, size_t bytes) = 0;
    virtual ~Out() = default;
};
struct In {
    virtual bool read(void* p, size_t bytes) = 0;  // false when short
    virtual int get() = 0;                         // next byte or EOF
    virtual ~In() = default;
};
struct FileOut : Out {
    FILE* f;
    explicit FileOut(FILE* F) : f(F) {}
    void write(const void* p, size_t bytes) override { fwrite(p, 1, bytes, f); }
};
struct FileIn : In {
    FILE* f;
    explicit FileIn(FILE* F) : f(F) {}
    bool read(void* p, size_t bytes) override { return fread(p, 1, bytes, f) == bytes; }
    int get() override { return fgetc(f); }
};
struct StreamOut : Out {
    std::ostream& o;
    explicit StreamOut(std::ostream& O) : o(O) {}
    void write(const void* p, size_t bytes) override { o.write((const char*)p, (std::streamsize)bytes); }
};
struct StreamIn : In {
    std::istream& i;
    explicit StreamIn(std::istream& I) : i(I) {}
    bool read(void* p, size_t bytes) override {
        i.read((char*)p, (std::streamsize)bytes);
        return (size_t)i.gcount() == bytes;
    }
    int get() override {
        const int c = i.get();
        return i ? c : EOF;
    }
};

// Parse failures unwind to the exported entry point, which turns them into
// NULL + tfhe_amd_last_error (constructors) or an abort (importers).
struct ParseError {};
[[noreturn]] void fail(const char* fmt, const char* a = "", const char* b = "") {
    set_error(fmt, a, b);
    throw ParseError{};
}

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
    const std::string& at(const char* name) const {
        auto it = props.find(name);
        if (it == props.end()) fail("key file: section %s lacks %s", title.c_str(), name);
        return it->second;
    }
    int64_t get_int(const char* name, int64_t lo, int64_t hi) const {
        const std::string& v = at(name);
        char* end = nullptr;
        const long long x = strtoll(v.c_str(), &end, 10);
        if (end == v.c_str() || x < lo || x > hi) fail("key file: %s = %s out of range", name, v.c_str());
        return x;
    }
    double get_double(const char* name) const { return (double)strtold(at(name).c_str(), nullptr); }
};

void write_section(Out& F, const Section& s) {
    std::string t = "-----BEGIN " + s.title + "-----\n";
    for (const auto& kv : s.props) t += kv.first + ": " + kv.second + "\n";
    t += "-----END " + s.title + "-----\n";
    F.write(t.data(), t.size());
}

bool read_line(In& F, std::string& line) {
    line.clear();
    int c;
    bool any = false;
    while ((c = F.get()) != EOF) {
        any = true;
        if (c == '\r') continue;
        if (c == '\n') return true;
        line.push_back((char)c);
        if (line.size() > 4096) fail("key file: text line too long");
    }
    return any;
}

// Reads the next section, which must be `title`; lines before its BEGIN marker
// are skipped, as the reference does (tfhe_generic_streams.cpp:119-158).
Section read_section(In& F, const char* title) {
    Section s;
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
            if (s.title != title) fail("key file: expected section %s, found %s", title, s.title.c_str());
            return s;
        }
        const size_t pos = line.find(": ");
        if (pos == std::string::npos) continue;
        s.props[line.substr(0, pos)] = line.substr(pos + 2);
    }
    fail("key file: section %s missing or truncated", title);
}

// ---- binary blocks

void rd(In& F, void* p, size_t bytes) {
    if (!F.read(p, bytes)) fail("key file: truncated binary block");
}
void rd_uid(In& F, int32_t want) {
    int32_t uid = -1;
    rd(F, &uid, sizeof uid);
    if (uid != want) {
        char a[16], b[16];
        snprintf(a, sizeof a, "%d", uid);
        snprintf(b, sizeof b, "%d", want);
        fail("key file: object type %s where %s was expected", a, b);
    }
}
void wr(Out& F, const void* p, size_t bytes) { F.write(p, bytes); }
void wr_uid(Out& F, int32_t uid) { wr(F, &uid, sizeof uid); }

// ---- parameters (tfhe_io.cpp:35-60, 243-296, 480-548, 731-755, 1014-1046)

void write_lwe_params(Out& F, const LweParams* p) {
    Section s;
    s.title = "LWEPARAMS";
    s.set_int("n", p->n);
    s.set_double("alpha_min", p->alpha_min);
    s.set_double("alpha_max", p->alpha_max);
    write_section(F, s);
}
const LweParams* read_lwe_params(In& F) {
    const Section s = read_section(F, "LWEPARAMS");
    return host_keep(new_LweParams((int32_t)s.get_int("n", 1, MAX_LWE_N), s.get_double("alpha_min"),
                                   s.get_double("alpha_max")));
}
void write_tlwe_params(Out& F, const TLweParams* p) {
    Section s;
    s.title = "TLWEPARAMS";
    s.set_int("N", p->N);
    s.set_int("k", p->k);
    s.set_double("alpha_min", p->alpha_min);
    s.set_double("alpha_max", p->alpha_max);
    write_section(F, s);
}
const TLweParams* read_tlwe_params(In& F) {
    const Section s = read_section(F, "TLWEPARAMS");
    const int64_t N = s.get_int("N", 1, MAX_POLY_N), k = s.get_int("k", 1, MAX_K);
    if (N & (N - 1)) fail("key file: N is not a power of two");
    return host_keep(new_TLweParams((int32_t)N, (int32_t)k, s.get_double("alpha_min"), s.get_double("alpha_max")));
}
void write_tgsw_params(Out& F, const TGswParams* p) {
    write_tlwe_params(F, p->tlwe_params);
    Section s;
    s.title = "TGSWPARAMS";
    s.set_int("l", p->l);
    s.set_int("Bgbit", p->Bgbit);
    write_section(F, s);
}
const TGswParams* read_tgsw_params(In& F) {
    const TLweParams* tp = read_tlwe_params(F);
    const Section s = read_section(F, "TGSWPARAMS");
    const int64_t Bgbit = s.get_int("Bgbit", 1, 31), l = s.get_int("l", 1, 32 / Bgbit);
    return host_keep(new_TGswParams((int32_t)l, (int32_t)Bgbit, tp));
}
struct KsParams {
    int32_t n, t, basebit;
};
void write_ks_params(Out& F, const LweKeySwitchKey* ks) {
    Section s;
    s.title = "LWEKSPARAMS";
    s.set_int("n", ks->n);
    s.set_int("t", ks->t);
    s.set_int("basebit", ks->basebit);
    write_section(F, s);
}
KsParams read_ks_params(In& F) {
    const Section s = read_section(F, "LWEKSPARAMS");
    KsParams k;
    k.n = (int32_t)s.get_int("n", 1, MAX_POLY_N * MAX_K);
    k.basebit = (int32_t)s.get_int("basebit", 1, 16);
    k.t = (int32_t)s.get_int("t", 1, 32 / k.basebit);
    return k;
}
void write_gate_params(Out& F, const TFheGateBootstrappingParameterSet* P) {
    Section g;
    g.title = "GATEBOOTSPARAMS";
    g.set_int("ks_t", P->ks_t);
    g.set_int("ks_basebit", P->ks_basebit);
    write_section(F, g);
    write_lwe_params(F, P->in_out_params);
    write_tgsw_params(F, P->tgsw_params);
}
TFheGateBootstrappingParameterSet* read_gate_params(In& F) {
    const Section g = read_section(F, "GATEBOOTSPARAMS");
    const int32_t ks_t = (int32_t)g.get_int("ks_t", 1, 32);
    // the reference reads ks_basebit through its double accessor (tfhe_io.cpp:1027)
    const double bb = g.get_double("ks_basebit");
    if (!(bb >= 1 && bb <= 16)) fail("key file: ks_basebit out of range");
    const LweParams* lp = read_lwe_params(F);
    const TGswParams* gp = read_tgsw_params(F);
    auto* P = new TFheGateBootstrappingParameterSet{ks_t, (int32_t)bb, lp, gp};
    return P;
}

// ---- samples and keys (tfhe_io.cpp:90-224, 303-462, 555-713)

void write_lwe_sample(Out& F, const LweSample* s, const LweParams* p) {
    wr_uid(F, UID_LWE_SAMPLE);
    wr(F, s->a, sizeof(int32_t) * p->n);
    wr(F, &s->b, sizeof(int32_t));
    wr(F, &s->current_variance, sizeof(double));
}
void read_lwe_sample(In& F, LweSample* s, const LweParams* p) {
    rd_uid(F, UID_LWE_SAMPLE);
    rd(F, s->a, sizeof(int32_t) * p->n);
    rd(F, &s->b, sizeof(int32_t));
    rd(F, &s->current_variance, sizeof(double));
}
void write_tlwe_sample(Out& F, const TLweSample* s, const TLweParams* p) {
    wr_uid(F, UID_TLWE_SAMPLE);
    for (int i = 0; i <= p->k; i++) wr(F, s->a[i].coefsT, sizeof(int32_t) * p->N);
    wr(F, &s->current_variance, sizeof(double));
}
void read_tlwe_sample(In& F, TLweSample* s, const TLweParams* p) {
    rd_uid(F, UID_TLWE_SAMPLE);
    for (int i = 0; i <= p->k; i++) rd(F, s->a[i].coefsT, sizeof(int32_t) * p->N);
    rd(F, &s->current_variance, sizeof(double));
}
void write_tgsw_sample(Out& F, const TGswSample* s, const TGswParams* p) {
    wr_uid(F, UID_TGSW_SAMPLE);
    for (int r = 0; r < p->kpl; r++) write_tlwe_sample(F, &s->all_sample[r], p->tlwe_params);
}
void read_tgsw_sample(In& F, TGswSample* s, const TGswParams* p) {
    rd_uid(F, UID_TGSW_SAMPLE);
    for (int r = 0; r < p->kpl; r++) read_tlwe_sample(F, &s->all_sample[r], p->tlwe_params);
}
void write_poly_key(Out& F, int32_t uid, const IntPolynomial* key, int k, int N) {
    wr_uid(F, uid);
    for (int i = 0; i < k; i++) wr(F, key[i].coefs, sizeof(int32_t) * N);
}
void read_poly_key(In& F, int32_t uid, IntPolynomial* key, int k, int N) {
    rd_uid(F, uid);
    for (int i = 0; i < k; i++) rd(F, key[i].coefs, sizeof(int32_t) * N);
}

// ---- keyswitch and bootstrapping keys (tfhe_io.cpp:757-930)

void write_ks_content(Out& F, const LweKeySwitchKey* ks) {
    const int n = ks->out_params->n;
    const size_t rows = (size_t)ks->n * ks->t * ks->base;
    double var = -1;
    for (size_t r = 0; r < rows; r++) var = ks->ks0_raw[r].current_variance > var ? ks->ks0_raw[r].current_variance : var;
    wr_uid(F, UID_KS_KEY);
    wr(F, &var, sizeof var);
    for (size_t r = 0; r < rows; r++) {
        wr(F, ks->ks0_raw[r].a, sizeof(int32_t) * n);
        wr(F, &ks->ks0_raw[r].b, sizeof(int32_t));
    }
}
void read_ks_content(In& F, LweKeySwitchKey* ks) {
    const int n = ks->out_params->n;
    const size_t rows = (size_t)ks->n * ks->t * ks->base;
    rd_uid(F, UID_KS_KEY);
    double var = 0;
    rd(F, &var, sizeof var);
    for (size_t r = 0; r < rows; r++) {
        rd(F, ks->ks0_raw[r].a, sizeof(int32_t) * n);
        rd(F, &ks->ks0_raw[r].b, sizeof(int32_t));
        ks->ks0_raw[r].current_variance = var;
    }
}
void write_bk_content(Out& F, const LweBootstrappingKey* bk) {
    const int n = bk->in_out_params->n, kpl = bk->bk_params->kpl, k = bk->accum_params->k, N = bk->accum_params->N;
    double var = -1;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < kpl; j++)
            var = bk->bk[i].all_sample[j].current_variance > var ? bk->bk[i].all_sample[j].current_variance : var;
    wr_uid(F, UID_BK_KEY);
    wr(F, &var, sizeof var);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < kpl; j++)
            for (int c = 0; c <= k; c++) wr(F, bk->bk[i].all_sample[j].a[c].coefsT, sizeof(int32_t) * N);
}
void read_bk_content(In& F, LweBootstrappingKey* bk) {
    const int n = bk->in_out_params->n, kpl = bk->bk_params->kpl, k = bk->accum_params->k, N = bk->accum_params->N;
    rd_uid(F, UID_BK_KEY);
    double var = 0;
    rd(F, &var, sizeof var);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < kpl; j++) {
            TLweSample& s = bk->bk[i].all_sample[j];
            for (int c = 0; c <= k; c++) rd(F, s.a[c].coefsT, sizeof(int32_t) * N);
            s.current_variance = var;
        }
}
void check_words(size_t words) {
    if (words > MAX_WORDS) fail("key file: object larger than the reader's limit");
}
size_t ks_words(int32_t nin, const KsParams& k, int n_out) {
    return (size_t)nin * (size_t)k.t * ((size_t)1 << k.basebit) * ((size_t)n_out + 1);
}
void write_bootstrapping_key(Out& F, const LweBootstrappingKey* bk, bool with_params) {
    if (with_params) {
        write_lwe_params(F, bk->in_out_params);
        write_tgsw_params(F, bk->bk_params);
    }
    write_ks_params(F, bk->ks);
    write_ks_content(F, bk->ks);
    write_bk_content(F, bk);
}
// reads into a new key; the partial key is freed on a parse failure
LweBootstrappingKey* read_bootstrapping_key(In& F, const LweParams* in_out, const TGswParams* gp) {
    const KsParams k = read_ks_params(F);
    const TLweParams* tp = gp->tlwe_params;
    if (k.n != tp->N * tp->k) fail("key file: keyswitch input dimension is not k N");
    check_words(ks_words(k.n, k, in_out->n));
    check_words((size_t)in_out->n * gp->kpl * (tp->k + 1) * tp->N);
    LweBootstrappingKey* bk = new_LweBootstrappingKey(k.t, k.basebit, in_out, gp);
    try {
        read_ks_content(F, bk->ks);
        read_bk_content(F, bk);
    } catch (...) {
        delete_LweBootstrappingKey(bk);
        throw;
    }
    return bk;
}

// Runs fn, turning parse failures and allocation failures into nullptr.
template <class Fn>
auto guarded(Fn fn) -> decltype(fn()) {
    try {
        return fn();
    } catch (const ParseError&) {
        return nullptr;
    } catch (const std::bad_alloc&) {
        set_error("key file: out of memory for the object it describes");
        return nullptr;
    }
}
// For importers into caller-allocated objects: abort on failure, as the reference.
template <class Fn>
void guarded_or_die(Fn fn) {
    try {
        fn();
    } catch (const ParseError&) {
        die(tfhe_amd_last_error());
    }
}

void need(bool ok, const char* what) {
    if (!ok) die(what);
}

// ---- key sets

TFheGateBootstrappingCloudKeySet* read_cloud(In& F) {
    TFheGateBootstrappingParameterSet* P = read_gate_params(F);
    LweBootstrappingKey* bk = read_bootstrapping_key(F, P->in_out_params, P->tgsw_params);
    return host_cloud_from_bk(P, bk);
}
TFheGateBootstrappingSecretKeySet* read_secret(In& F) {
    TFheGateBootstrappingParameterSet* P = read_gate_params(F);
    LweBootstrappingKey* bk = read_bootstrapping_key(F, P->in_out_params, P->tgsw_params);
    const TLweParams* tp = P->tgsw_params->tlwe_params;
    std::vector<int32_t> lwe(P->in_out_params->n), tlwe((size_t)tp->k * tp->N);
    try {
        rd_uid(F, UID_LWE_KEY);
        rd(F, lwe.data(), sizeof(int32_t) * lwe.size());
        rd_uid(F, UID_TGSW_KEY);
        rd(F, tlwe.data(), sizeof(int32_t) * tlwe.size());
    } catch (...) {
        delete_LweBootstrappingKey(bk);
        throw;
    }
    return host_secret_from_bk(P, bk, lwe.data(), tlwe.data());
}
void write_secret(Out& F, const TFheGateBootstrappingSecretKeySet* key) {
    write_gate_params(F, key->params);
    write_bootstrapping_key(F, key->cloud.bk, false);
    const int n = key->params->in_out_params->n;
    wr_uid(F, UID_LWE_KEY);
    wr(F, key->lwe_key->key, sizeof(int32_t) * n);
    const TLweParams* tp = key->params->tgsw_params->tlwe_params;
    write_poly_key(F, UID_TGSW_KEY, key->tgsw_key->key, tp->k, tp->N);
}

}  // namespace
}  // namespace tfhe_amd

using namespace tfhe_amd;

// Every reference file function exists twice: over a C FILE* and over a C++
// std::ostream / std::istream; both forms go through the same Out / In code.

// ---- LWE parameters, samples, keys (include/tfhe_io.h:30-110)

static void lwe_params_out(Out& F, const LweParams* p) {
    need(p != nullptr, "export_lweParams: null parameters");
    write_lwe_params(F, p);
}
TFHE_AMD_EXPORT void export_lweParams_toFile(FILE* f, const LweParams* p) {
    need(f != nullptr, "export_lweParams_toFile: null FILE");
    FileOut F(f);
    lwe_params_out(F, p);
}
TFHE_AMD_EXPORT void export_lweParams_toStream(std::ostream& o, const LweParams* p) {
    StreamOut F(o);
    lwe_params_out(F, p);
}
static LweParams* lwe_params_in(In& F) {
    // the caller may delete_LweParams the result: give it a copy of the kept one
    return guarded([&]() -> LweParams* {
        const LweParams* p = read_lwe_params(F);
        return new_LweParams(p->n, p->alpha_min, p->alpha_max);
    });
}
TFHE_AMD_EXPORT LweParams* new_lweParams_fromFile(FILE* f) {
    if (!f) return set_error("new_lweParams_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return lwe_params_in(F);
}
TFHE_AMD_EXPORT LweParams* new_lweParams_fromStream(std::istream& i) {
    StreamIn F(i);
    return lwe_params_in(F);
}

TFHE_AMD_EXPORT void export_lweSample_toFile(FILE* f, const LweSample* s, const LweParams* p) {
    need(f && s && p, "export_lweSample_toFile: bad arguments");
    FileOut F(f);
    write_lwe_sample(F, s, p);
}
TFHE_AMD_EXPORT void export_lweSample_toStream(std::ostream& o, const LweSample* s, const LweParams* p) {
    need(s && p, "export_lweSample_toStream: bad arguments");
    StreamOut F(o);
    write_lwe_sample(F, s, p);
}
TFHE_AMD_EXPORT void import_lweSample_fromFile(FILE* f, LweSample* s, const LweParams* p) {
    need(f && s && p, "import_lweSample_fromFile: bad arguments");
    FileIn F(f);
    guarded_or_die([&] { read_lwe_sample(F, s, p); });
}
TFHE_AMD_EXPORT void import_lweSample_fromStream(std::istream& i, LweSample* s, const LweParams* p) {
    need(s && p, "import_lweSample_fromStream: bad arguments");
    StreamIn F(i);
    guarded_or_die([&] { read_lwe_sample(F, s, p); });
}

static void lwe_key_out(Out& F, const LweKey* key) {
    need(key != nullptr, "export_lweKey: null key");
    write_lwe_params(F, key->params);
    wr_uid(F, UID_LWE_KEY);
    wr(F, key->key, sizeof(int32_t) * key->params->n);
}
static LweKey* lwe_key_in(In& F) {
    return guarded([&]() -> LweKey* {
        const LweParams* p = read_lwe_params(F);
        LweKey* key = new_LweKey(p);
        try {
            rd_uid(F, UID_LWE_KEY);
            rd(F, key->key, sizeof(int32_t) * p->n);
        } catch (...) {
            delete_LweKey(key);
            throw;
        }
        return key;
    });
}
TFHE_AMD_EXPORT void export_lweKey_toFile(FILE* f, const LweKey* key) {
    need(f != nullptr, "export_lweKey_toFile: null FILE");
    FileOut F(f);
    lwe_key_out(F, key);
}
TFHE_AMD_EXPORT void export_lweKey_toStream(std::ostream& o, const LweKey* key) {
    StreamOut F(o);
    lwe_key_out(F, key);
}
TFHE_AMD_EXPORT LweKey* new_lweKey_fromFile(FILE* f) {
    if (!f) return set_error("new_lweKey_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return lwe_key_in(F);
}
TFHE_AMD_EXPORT LweKey* new_lweKey_fromStream(std::istream& i) {
    StreamIn F(i);
    return lwe_key_in(F);
}

// ---- TLWE parameters, samples, keys (include/tfhe_io.h:126-207)

static void tlwe_params_out(Out& F, const TLweParams* p) {
    need(p != nullptr, "export_tLweParams: null parameters");
    write_tlwe_params(F, p);
}
static TLweParams* tlwe_params_in(In& F) {
    return guarded([&]() -> TLweParams* {
        const TLweParams* p = read_tlwe_params(F);
        return new_TLweParams(p->N, p->k, p->alpha_min, p->alpha_max);
    });
}
TFHE_AMD_EXPORT void export_tLweParams_toFile(FILE* f, const TLweParams* p) {
    need(f != nullptr, "export_tLweParams_toFile: null FILE");
    FileOut F(f);
    tlwe_params_out(F, p);
}
TFHE_AMD_EXPORT void export_tLweParams_toStream(std::ostream& o, const TLweParams* p) {
    StreamOut F(o);
    tlwe_params_out(F, p);
}
TFHE_AMD_EXPORT TLweParams* new_tLweParams_fromFile(FILE* f) {
    if (!f) return set_error("new_tLweParams_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return tlwe_params_in(F);
}
TFHE_AMD_EXPORT TLweParams* new_tLweParams_fromStream(std::istream& i) {
    StreamIn F(i);
    return tlwe_params_in(F);
}

TFHE_AMD_EXPORT void export_tlweSample_toFile(FILE* f, const TLweSample* s, const TLweParams* p) {
    need(f && s && p, "export_tlweSample_toFile: bad arguments");
    FileOut F(f);
    write_tlwe_sample(F, s, p);
}
TFHE_AMD_EXPORT void export_tlweSample_toStream(std::ostream& o, const TLweSample* s, const TLweParams* p) {
    need(s && p, "export_tlweSample_toStream: bad arguments");
    StreamOut F(o);
    write_tlwe_sample(F, s, p);
}
TFHE_AMD_EXPORT void import_tlweSample_fromFile(FILE* f, TLweSample* s, const TLweParams* p) {
    need(f && s && p, "import_tlweSample_fromFile: bad arguments");
    FileIn F(f);
    guarded_or_die([&] { read_tlwe_sample(F, s, p); });
}
TFHE_AMD_EXPORT void import_tlweSample_fromStream(std::istream& i, TLweSample* s, const TLweParams* p) {
    need(s && p, "import_tlweSample_fromStream: bad arguments");
    StreamIn F(i);
    guarded_or_die([&] { read_tlwe_sample(F, s, p); });
}

static void tlwe_key_out(Out& F, const TLweKey* key) {
    need(key != nullptr, "export_tlweKey: null key");
    write_tlwe_params(F, key->params);
    write_poly_key(F, UID_TLWE_KEY, key->key, key->params->k, key->params->N);
}
static TLweKey* tlwe_key_in(In& F) {
    return guarded([&]() -> TLweKey* {
        const TLweParams* p = read_tlwe_params(F);
        check_words((size_t)p->k * p->N);
        TLweKey* key = new_TLweKey(p);
        try {
            read_poly_key(F, UID_TLWE_KEY, key->key, p->k, p->N);
        } catch (...) {
            delete_TLweKey(key);
            throw;
        }
        return key;
    });
}
TFHE_AMD_EXPORT void export_tlweKey_toFile(FILE* f, const TLweKey* key) {
    need(f != nullptr, "export_tlweKey_toFile: null FILE");
    FileOut F(f);
    tlwe_key_out(F, key);
}
TFHE_AMD_EXPORT void export_tlweKey_toStream(std::ostream& o, const TLweKey* key) {
    StreamOut F(o);
    tlwe_key_out(F, key);
}
TFHE_AMD_EXPORT TLweKey* new_tlweKey_fromFile(FILE* f) {
    if (!f) return set_error("new_tlweKey_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return tlwe_key_in(F);
}
TFHE_AMD_EXPORT TLweKey* new_tlweKey_fromStream(std::istream& i) {
    StreamIn F(i);
    return tlwe_key_in(F);
}

// ---- TGSW parameters, samples, keys (include/tfhe_io.h:225-308)

static void tgsw_params_out(Out& F, const TGswParams* p) {
    need(p != nullptr, "export_tGswParams: null parameters");
    write_tgsw_params(F, p);
}
static TGswParams* tgsw_params_in(In& F) {
    // the TLWE parameters it points at are kept for the process (as the
    // reference's garbage collector); the TGSW parameters are the caller's
    return guarded([&]() -> TGswParams* {
        const TGswParams* p = read_tgsw_params(F);
        return new_TGswParams(p->l, p->Bgbit, p->tlwe_params);
    });
}
TFHE_AMD_EXPORT void export_tGswParams_toFile(FILE* f, const TGswParams* p) {
    need(f != nullptr, "export_tGswParams_toFile: null FILE");
    FileOut F(f);
    tgsw_params_out(F, p);
}
TFHE_AMD_EXPORT void export_tGswParams_toStream(std::ostream& o, const TGswParams* p) {
    StreamOut F(o);
    tgsw_params_out(F, p);
}
TFHE_AMD_EXPORT TGswParams* new_tGswParams_fromFile(FILE* f) {
    if (!f) return set_error("new_tGswParams_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return tgsw_params_in(F);
}
TFHE_AMD_EXPORT TGswParams* new_tGswParams_fromStream(std::istream& i) {
    StreamIn F(i);
    return tgsw_params_in(F);
}

TFHE_AMD_EXPORT void export_tgswSample_toFile(FILE* f, const TGswSample* s, const TGswParams* p) {
    need(f && s && p, "export_tgswSample_toFile: bad arguments");
    FileOut F(f);
    write_tgsw_sample(F, s, p);
}
TFHE_AMD_EXPORT void export_tgswSample_toStream(std::ostream& o, const TGswSample* s, const TGswParams* p) {
    need(s && p, "export_tgswSample_toStream: bad arguments");
    StreamOut F(o);
    write_tgsw_sample(F, s, p);
}
TFHE_AMD_EXPORT void import_tgswSample_fromFile(FILE* f, TGswSample* s, const TGswParams* p) {
    need(f && s && p, "import_tgswSample_fromFile: bad arguments");
    FileIn F(f);
    guarded_or_die([&] { read_tgsw_sample(F, s, p); });
}
TFHE_AMD_EXPORT void import_tgswSample_fromStream(std::istream& i, TGswSample* s, const TGswParams* p) {
    need(s && p, "import_tgswSample_fromStream: bad arguments");
    StreamIn F(i);
    guarded_or_die([&] { read_tgsw_sample(F, s, p); });
}

static void tgsw_key_out(Out& F, const TGswKey* key) {
    need(key != nullptr, "export_tgswKey: null key");
    write_tgsw_params(F, key->params);
    write_poly_key(F, UID_TGSW_KEY, key->key, key->tlwe_params->k, key->tlwe_params->N);
}
static TGswKey* tgsw_key_in(In& F) {
    return guarded([&]() -> TGswKey* {
        const TGswParams* p = read_tgsw_params(F);
        check_words((size_t)p->tlwe_params->k * p->tlwe_params->N);
        TGswKey* key = new_TGswKey(p);
        try {
            read_poly_key(F, UID_TGSW_KEY, key->key, p->tlwe_params->k, p->tlwe_params->N);
        } catch (...) {
            delete_TGswKey(key);
            throw;
        }
        return key;
    });
}
TFHE_AMD_EXPORT void export_tgswKey_toFile(FILE* f, const TGswKey* key) {
    need(f != nullptr, "export_tgswKey_toFile: null FILE");
    FileOut F(f);
    tgsw_key_out(F, key);
}
TFHE_AMD_EXPORT void export_tgswKey_toStream(std::ostream& o, const TGswKey* key) {
    StreamOut F(o);
    tgsw_key_out(F, key);
}
TFHE_AMD_EXPORT TGswKey* new_tgswKey_fromFile(FILE* f) {
    if (!f) return set_error("new_tgswKey_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return tgsw_key_in(F);
}
TFHE_AMD_EXPORT TGswKey* new_tgswKey_fromStream(std::istream& i) {
    StreamIn F(i);
    return tgsw_key_in(F);
}

// ---- keyswitch and bootstrapping keys (include/tfhe_io.h:319-368)

static void ks_key_out(Out& F, const LweKeySwitchKey* ks) {
    need(ks != nullptr, "export_lweKeySwitchKey: null key");
    write_lwe_params(F, ks->out_params);
    write_ks_params(F, ks);
    write_ks_content(F, ks);
}
static LweKeySwitchKey* ks_key_in(In& F) {
    return guarded([&]() -> LweKeySwitchKey* {
        const LweParams* out = read_lwe_params(F);
        const KsParams k = read_ks_params(F);
        check_words(ks_words(k.n, k, out->n));
        LweKeySwitchKey* ks = new_LweKeySwitchKey(k.n, k.t, k.basebit, out);
        try {
            read_ks_content(F, ks);
        } catch (...) {
            delete_LweKeySwitchKey(ks);
            throw;
        }
        return ks;
    });
}
TFHE_AMD_EXPORT void export_lweKeySwitchKey_toFile(FILE* f, const LweKeySwitchKey* ks) {
    need(f != nullptr, "export_lweKeySwitchKey_toFile: null FILE");
    FileOut F(f);
    ks_key_out(F, ks);
}
TFHE_AMD_EXPORT void export_lweKeySwitchKey_toStream(std::ostream& o, const LweKeySwitchKey* ks) {
    StreamOut F(o);
    ks_key_out(F, ks);
}
TFHE_AMD_EXPORT LweKeySwitchKey* new_lweKeySwitchKey_fromFile(FILE* f) {
    if (!f) return set_error("new_lweKeySwitchKey_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return ks_key_in(F);
}
TFHE_AMD_EXPORT LweKeySwitchKey* new_lweKeySwitchKey_fromStream(std::istream& i) {
    StreamIn F(i);
    return ks_key_in(F);
}

static LweBootstrappingKey* bk_key_in(In& F) {
    return guarded([&]() -> LweBootstrappingKey* {
        const LweParams* in_out = read_lwe_params(F);
        const TGswParams* gp = read_tgsw_params(F);
        return read_bootstrapping_key(F, in_out, gp);
    });
}
TFHE_AMD_EXPORT void export_lweBootstrappingKey_toFile(FILE* f, const LweBootstrappingKey* bk) {
    need(f && bk, "export_lweBootstrappingKey_toFile: bad arguments");
    FileOut F(f);
    write_bootstrapping_key(F, bk, true);
}
TFHE_AMD_EXPORT void export_lweBootstrappingKey_toStream(std::ostream& o, const LweBootstrappingKey* bk) {
    need(bk != nullptr, "export_lweBootstrappingKey_toStream: null key");
    StreamOut F(o);
    write_bootstrapping_key(F, bk, true);
}
TFHE_AMD_EXPORT LweBootstrappingKey* new_lweBootstrappingKey_fromFile(FILE* f) {
    if (!f) return set_error("new_lweBootstrappingKey_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return bk_key_in(F);
}
TFHE_AMD_EXPORT LweBootstrappingKey* new_lweBootstrappingKey_fromStream(std::istream& i) {
    StreamIn F(i);
    return bk_key_in(F);
}

// ---- gate-level files (include/tfhe_io.h:379-496)

TFHE_AMD_EXPORT void export_tfheGateBootstrappingParameterSet_toFile(FILE* f,
                                                                     const TFheGateBootstrappingParameterSet* params) {
    need(f && params, "export_tfheGateBootstrappingParameterSet_toFile: bad arguments");
    FileOut F(f);
    write_gate_params(F, params);
}
TFHE_AMD_EXPORT void export_tfheGateBootstrappingParameterSet_toStream(std::ostream& o,
                                                                       const TFheGateBootstrappingParameterSet* params) {
    need(params != nullptr, "export_tfheGateBootstrappingParameterSet_toStream: null parameters");
    StreamOut F(o);
    write_gate_params(F, params);
}
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromFile(FILE* f) {
    if (!f) return set_error("new_tfheGateBootstrappingParameterSet_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return guarded([&] { return read_gate_params(F); });
}
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_tfheGateBootstrappingParameterSet_fromStream(std::istream& i) {
    StreamIn F(i);
    return guarded([&] { return read_gate_params(F); });
}

TFHE_AMD_EXPORT void export_tfheGateBootstrappingCloudKeySet_toFile(FILE* f, const TFheGateBootstrappingCloudKeySet* key) {
    need(f && key, "export_tfheGateBootstrappingCloudKeySet_toFile: bad arguments");
    FileOut F(f);
    write_gate_params(F, key->params);
    write_bootstrapping_key(F, key->bk, false);
}
TFHE_AMD_EXPORT void export_tfheGateBootstrappingCloudKeySet_toStream(std::ostream& o,
                                                                      const TFheGateBootstrappingCloudKeySet* key) {
    need(key != nullptr, "export_tfheGateBootstrappingCloudKeySet_toStream: null key");
    StreamOut F(o);
    write_gate_params(F, key->params);
    write_bootstrapping_key(F, key->bk, false);
}
TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromFile(FILE* f) {
    if (!f) return set_error("new_tfheGateBootstrappingCloudKeySet_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return guarded([&] { return read_cloud(F); });
}
TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_tfheGateBootstrappingCloudKeySet_fromStream(std::istream& i) {
    StreamIn F(i);
    return guarded([&] { return read_cloud(F); });
}

TFHE_AMD_EXPORT void export_tfheGateBootstrappingSecretKeySet_toFile(FILE* f,
                                                                     const TFheGateBootstrappingSecretKeySet* key) {
    need(f && key, "export_tfheGateBootstrappingSecretKeySet_toFile: bad arguments");
    FileOut F(f);
    write_secret(F, key);
}
TFHE_AMD_EXPORT void export_tfheGateBootstrappingSecretKeySet_toStream(std::ostream& o,
                                                                       const TFheGateBootstrappingSecretKeySet* key) {
    need(key != nullptr, "export_tfheGateBootstrappingSecretKeySet_toStream: null key");
    StreamOut F(o);
    write_secret(F, key);
}
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromFile(FILE* f) {
    if (!f) return set_error("new_tfheGateBootstrappingSecretKeySet_fromFile: null FILE"), nullptr;
    FileIn F(f);
    return guarded([&] { return read_secret(F); });
}
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_tfheGateBootstrappingSecretKeySet_fromStream(std::istream& i) {
    StreamIn F(i);
    return guarded([&] { return read_secret(F); });
}

TFHE_AMD_EXPORT void export_gate_bootstrapping_ciphertext_toFile(FILE* F, const LweSample* sample,
                                                                 const TFheGateBootstrappingParameterSet* params) {
    export_lweSample_toFile(F, sample, params ? params->in_out_params : nullptr);
}
TFHE_AMD_EXPORT void export_gate_bootstrapping_ciphertext_toStream(std::ostream& o, const LweSample* sample,
                                                                   const TFheGateBootstrappingParameterSet* params) {
    export_lweSample_toStream(o, sample, params ? params->in_out_params : nullptr);
}
TFHE_AMD_EXPORT void import_gate_bootstrapping_ciphertext_fromFile(FILE* F, LweSample* sample,
                                                                   const TFheGateBootstrappingParameterSet* params) {
    import_lweSample_fromFile(F, sample, params ? params->in_out_params : nullptr);
}
TFHE_AMD_EXPORT void import_gate_bootstrapping_ciphertext_fromStream(std::istream& i, LweSample* sample,
                                                                     const TFheGateBootstrappingParameterSet* params) {
    import_lweSample_fromStream(i, sample, params ? params->in_out_params : nullptr);
}
