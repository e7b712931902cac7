// Host side of tfhe_amd: parameter sets, key generation, encryption and the
// reference gate API (include/tfhe_gate_bootstrapping_functions.h). Every
// bootstrap and keyswitch is handed to the GPU engine (engine.hip); there is
// no CPU implementation of the bootstrapping path in this library.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "engine.h"

namespace tfhe_amd {

namespace {
thread_local std::string g_err;

// The reference keeps one global libstdc++ default_random_engine, seeded via
// seed_seq (libtfhe/numeric-functions.cpp:11,16-19). Key generation and
// encryption below draw from it in the reference's order, so a given seed
// produces the reference's keys and ciphertexts.
std::default_random_engine& rng() {
    static std::default_random_engine g;
    return g;
}

inline uint32_t U(int32_t x) { return (uint32_t)x; }
inline int32_t S(uint32_t x) { return (int32_t)x; }

int32_t draw_torus() {
    std::uniform_int_distribution<int32_t> d(INT32_MIN, INT32_MAX);
    return d(rng());
}
int32_t draw_bit() {
    std::uniform_int_distribution<int32_t> d(0, 1);
    return d(rng());
}
// one fresh normal_distribution per draw, as the reference does (the
// distribution object caches the second polar-method value)
double draw_normal(double sigma) {
    std::normal_distribution<double> d(0., sigma);
    return d(rng());
}

// Parameter objects referenced by parameter sets live until process exit,
// as with the reference's TfheGarbageCollector.
std::mutex g_param_mu;
std::vector<std::unique_ptr<LweParams>> g_lwe_params;
std::vector<std::unique_ptr<TLweParams>> g_tlwe_params;
std::vector<TGswParams*> g_tgsw_params;
}  // namespace

int set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

void die(const char* what) {
    fprintf(stderr, "tfhe_amd: %s\n", what);
    abort();
}

HostParams host_params(const TFheGateBootstrappingParameterSet* p) {
    HostParams h;
    h.n = p->in_out_params->n;
    h.N = p->tgsw_params->tlwe_params->N;
    h.k = p->tgsw_params->tlwe_params->k;
    h.l = p->tgsw_params->l;
    h.Bgbit = p->tgsw_params->Bgbit;
    h.ks_t = p->ks_t;
    h.ks_basebit = p->ks_basebit;
    return h;
}

HostParams host_params_fft(const LweBootstrappingKeyFFT* bf) {
    HostParams h;
    h.n = bf->in_out_params->n;
    h.N = bf->accum_params->N;
    h.k = bf->accum_params->k;
    h.l = bf->bk_params->l;
    h.Bgbit = bf->bk_params->Bgbit;
    h.ks_t = bf->ks->t;
    h.ks_basebit = bf->ks->basebit;
    return h;
}

namespace {

// ---------------------------------------------------------- key storage --

// Coefficient-domain bootstrapping key, contiguous [n][kpl][k+1][N] with the
// reference's struct views on top (libtfhe/tgsw-functions.cpp:24-35).
struct BkOwner {
    LweBootstrappingKey pub;  // must stay first: delete casts back from &pub
    std::vector<int32_t> coefs;
    std::vector<TorusPolynomial> polys;
    std::vector<TLweSample> samples;
    std::vector<TLweSample*> blocs;
    std::vector<TGswSample> gsw;
};

// Keyswitch key: ks[i][j][h] = ks0_raw[(i*t+j)*base+h] (libtfhe/lwekeyswitch.cpp:3-18).
struct KsOwner {
    LweKeySwitchKey pub;  // must stay first
    std::vector<int32_t> a;
    std::vector<LweSample> rows;
    std::vector<LweSample*> l1;
    std::vector<LweSample**> l0;
};

struct BkFftOwner {
    LweBootstrappingKeyFFT pub;  // must stay first
    uint64_t magic = kBkFftMagic;
    std::vector<TGswSampleFFT> rows;
    const LweBootstrappingKey* bk_coef = nullptr;
    tfhe_amd_device_key* dkey = nullptr;
    static constexpr uint64_t kBkFftMagic = 0x7466686542464654ull;
};

// every live transform-domain key this library made, for finding the key of a
// row array (tfhe_blindRotate_FFT gets the rows, not the key)
std::mutex g_bkfft_mu;
std::vector<BkFftOwner*> g_bkfft;

BkFftOwner* bkfft_owner(const LweBootstrappingKeyFFT* bf) {
    auto* o = reinterpret_cast<BkFftOwner*>(const_cast<LweBootstrappingKeyFFT*>(bf));
    if (!bf || o->magic != BkFftOwner::kBkFftMagic) die("LweBootstrappingKeyFFT not made by tfhe_amd");
    return o;
}

KsOwner* new_ks(int32_t nin, int32_t t, int32_t basebit, const LweParams* out_params) {
    auto* o = new KsOwner();
    const int base = 1 << basebit, n = out_params->n;
    const size_t rows = (size_t)nin * t * base;
    o->a.assign(rows * n, 0);
    o->rows.resize(rows);
    for (size_t r = 0; r < rows; r++) o->rows[r] = LweSample{o->a.data() + r * n, 0, 0.0};
    o->l1.resize((size_t)nin * t);
    for (size_t p = 0; p < o->l1.size(); p++) o->l1[p] = o->rows.data() + base * p;
    o->l0.resize(nin);
    for (int p = 0; p < nin; p++) o->l0[p] = o->l1.data() + (size_t)t * p;
    o->pub = LweKeySwitchKey{nin, t, basebit, base, out_params, o->rows.data(), o->l1.data(), o->l0.data()};
    return o;
}

BkOwner* new_bk(const LweParams* in_out, const TGswParams* gp, LweKeySwitchKey* ks) {
    auto* o = new BkOwner();
    const TLweParams* tp = gp->tlwe_params;
    const int n = in_out->n, N = tp->N, k = tp->k, l = gp->l, kpl = gp->kpl;
    o->coefs.assign((size_t)n * kpl * (k + 1) * N, 0);
    o->polys.resize((size_t)n * kpl * (k + 1));
    for (size_t q = 0; q < o->polys.size(); q++) o->polys[q] = TorusPolynomial{N, o->coefs.data() + q * N};
    o->samples.resize((size_t)n * kpl);
    for (size_t q = 0; q < o->samples.size(); q++)
        o->samples[q] = TLweSample{o->polys.data() + q * (k + 1), o->polys.data() + q * (k + 1) + k, 0.0, k};
    o->blocs.resize((size_t)n * (k + 1));
    o->gsw.resize(n);
    for (int i = 0; i < n; i++) {
        for (int b = 0; b <= k; b++) o->blocs[(size_t)i * (k + 1) + b] = o->samples.data() + (size_t)i * kpl + b * l;
        o->gsw[i] = TGswSample{o->samples.data() + (size_t)i * kpl, o->blocs.data() + (size_t)i * (k + 1), k, l};
    }
    o->pub = LweBootstrappingKey{in_out, gp, tp, &tp->extracted_lweparams, o->gsw.data(), ks};
    return o;
}
BkOwner* new_bk(const TFheGateBootstrappingParameterSet* P, LweKeySwitchKey* ks) {
    return new_bk(P->in_out_params, P->tgsw_params, ks);
}

BkFftOwner* new_bkfft(const LweBootstrappingKey* bk) {
    auto* o = new BkFftOwner();
    const int n = bk->in_out_params->n;
    o->rows.resize(n);
    o->pub = LweBootstrappingKeyFFT{bk->in_out_params, bk->bk_params, bk->accum_params, bk->extract_params,
                                    o->rows.data(), bk->ks};
    o->bk_coef = bk;
    for (int i = 0; i < n; i++) o->rows[i] = TGswSampleFFT{nullptr, nullptr, bk->accum_params->k, bk->bk_params->l};
    std::lock_guard<std::mutex> g(g_bkfft_mu);
    g_bkfft.push_back(o);
    return o;
}

void unregister_ks(const LweBootstrappingKeyFFT* bf);

void delete_bkfft(const LweBootstrappingKeyFFT* bf) {
    if (!bf) return;
    BkFftOwner* o = bkfft_owner(bf);
    unregister_ks(bf);
    engine_forget_host_key(bf);
    {
        std::lock_guard<std::mutex> g(g_bkfft_mu);
        g_bkfft.erase(std::remove(g_bkfft.begin(), g_bkfft.end(), o), g_bkfft.end());
    }
    o->magic = 0;
    delete o;
}

void delete_bk(const LweBootstrappingKey* bk) {
    if (!bk) return;
    delete reinterpret_cast<KsOwner*>(bk->ks);
    delete reinterpret_cast<BkOwner*>(const_cast<LweBootstrappingKey*>(bk));
}

struct SecretOwner {
    TFheGateBootstrappingSecretKeySet pub;  // must stay first
    LweKey lwe_key;
    std::vector<int32_t> lwe_bits;
    TGswKey tgsw_key;
    std::vector<int32_t> tlwe_bits;
    std::vector<IntPolynomial> tlwe_polys;
};

struct CloudOwner {
    TFheGateBootstrappingCloudKeySet pub;  // must stay first
};

// keyswitch key -> transform-domain key that owns it (for lweKeySwitch)
std::mutex g_ks_mu;
std::vector<std::pair<const LweKeySwitchKey*, const LweBootstrappingKeyFFT*>> g_ks_owner;

void register_ks(const LweKeySwitchKey* ks, const LweBootstrappingKeyFFT* bf) {
    std::lock_guard<std::mutex> g(g_ks_mu);
    g_ks_owner.emplace_back(ks, bf);
}
void unregister_ks(const LweBootstrappingKeyFFT* bf) {
    std::lock_guard<std::mutex> g(g_ks_mu);
    for (size_t i = 0; i < g_ks_owner.size();)
        if (g_ks_owner[i].second == bf) g_ks_owner.erase(g_ks_owner.begin() + i);
        else i++;
}
const LweBootstrappingKeyFFT* owner_of_ks(const LweKeySwitchKey* ks) {
    std::lock_guard<std::mutex> g(g_ks_mu);
    for (auto& e : g_ks_owner)
        if (e.first == ks) return e.second;
    return nullptr;
}

// exact negacyclic b += s * a for a binary key polynomial s
void add_binary_product(const int32_t* s, const int32_t* a, int N, int32_t* b) {
    uint32_t* r = (uint32_t*)b;
    const uint32_t* x = (const uint32_t*)a;
    for (int t = 0; t < N; t++) {
        if (!s[t]) continue;
        for (int j = 0; j < t; j++) r[j] -= x[j - t + N];
        for (int j = t; j < N; j++) r[j] += x[j - t];
    }
}

int32_t to_torus(double d) { return (int32_t)(int64_t)((d - (double)(int64_t)d) * 4294967296.0); }

}  // namespace

const LweBootstrappingKey* bkfft_coef(const LweBootstrappingKeyFFT* bf) { return bkfft_owner(bf)->bk_coef; }
tfhe_amd_device_key*& bkfft_dkey(const LweBootstrappingKeyFFT* bf) { return bkfft_owner(bf)->dkey; }

const LweBootstrappingKeyFFT* bkfft_of_rows(const TGswSampleFFT* rows, int32_t* row) {
    if (!rows) return nullptr;
    std::lock_guard<std::mutex> g(g_bkfft_mu);
    for (BkFftOwner* o : g_bkfft) {
        const TGswSampleFFT* b = o->rows.data();
        if (rows >= b && rows < b + o->rows.size()) {
            *row = (int32_t)(rows - b);
            return &o->pub;
        }
    }
    return nullptr;
}

}  // namespace tfhe_amd

using namespace tfhe_amd;

// ------------------------------------------------------------ numerics ----

// libtfhe/numeric-functions.cpp:16-19
TFHE_AMD_EXPORT void tfhe_random_generator_setSeed(uint32_t* values, int32_t size) {
    std::seed_seq seeds(values, values + size);
    rng().seed(seeds);
}

TFHE_AMD_EXPORT Torus32 dtot32(double d) { return to_torus(d); }
TFHE_AMD_EXPORT double t32tod(Torus32 x) { return double(x) / 4294967296.0; }

// round(phase * Msize / 2^32) mod Msize, libtfhe/numeric-functions.cpp:59-65
TFHE_AMD_EXPORT int32_t modSwitchFromTorus32(Torus32 phase, int32_t Msize) {
    const uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    const uint64_t phase64 = ((uint64_t)U(phase) << 32) + interv / 2;
    return (int32_t)(phase64 / interv);
}

// mu / Msize as a torus value, libtfhe/numeric-functions.cpp:71-76
TFHE_AMD_EXPORT Torus32 modSwitchToTorus32(int32_t mu, int32_t Msize) {
    const uint64_t interv = ((UINT64_C(1) << 63) / (uint64_t)Msize) * 2;
    return (Torus32)(uint32_t)(((uint64_t)(int64_t)mu * interv) >> 32);
}

TFHE_AMD_EXPORT const char* tfhe_amd_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------- parameters ----

TFHE_AMD_EXPORT LweParams* new_LweParams(int32_t n, double alpha_min, double alpha_max) {
    return new LweParams{n, alpha_min, alpha_max};
}
TFHE_AMD_EXPORT void delete_LweParams(LweParams* obj) { delete obj; }

TFHE_AMD_EXPORT TLweParams* new_TLweParams(int32_t N, int32_t k, double alpha_min, double alpha_max) {
    return new TLweParams{N, k, alpha_min, alpha_max, LweParams{N * k, alpha_min, alpha_max}};
}
TFHE_AMD_EXPORT void delete_TLweParams(TLweParams* obj) { delete obj; }

// h_i = 2^(32-(i+1)Bgbit) and offset, libtfhe/tgsw.cpp:7-28
TFHE_AMD_EXPORT TGswParams* new_TGswParams(int32_t l, int32_t Bgbit, const TLweParams* tlwe_params) {
    auto* p = new TGswParams;
    p->l = l;
    p->Bgbit = Bgbit;
    p->Bg = 1 << Bgbit;
    p->halfBg = p->Bg / 2;
    p->maskMod = (uint32_t)p->Bg - 1;
    p->tlwe_params = tlwe_params;
    p->kpl = (tlwe_params->k + 1) * l;
    p->h = new Torus32[l];
    uint32_t off = 0;
    for (int i = 0; i < l; i++) {
        p->h[i] = S(1u << (32 - (i + 1) * Bgbit));
        off += 1u << (32 - (i + 1) * Bgbit);
    }
    p->offset = off * (uint32_t)p->halfBg;
    return p;
}
TFHE_AMD_EXPORT void delete_TGswParams(TGswParams* obj) {
    if (!obj) return;
    delete[] obj->h;
    delete obj;
}

namespace tfhe_amd {
TFheGateBootstrappingParameterSet* host_make_params(int32_t n, double lwe_alpha_min, double lwe_alpha_max, int32_t N,
                                                    int32_t k, double tlwe_alpha_min, double tlwe_alpha_max,
                                                    int32_t bk_l, int32_t bk_Bgbit, int32_t ks_t, int32_t ks_basebit) {
    if (n <= 0 || N <= 0 || (N & (N - 1)) || k <= 0 || bk_l <= 0 || bk_Bgbit <= 0 || bk_l * bk_Bgbit > 32 ||
        ks_t <= 0 || ks_basebit <= 0 || ks_t * ks_basebit > 31) {
        set_error("invalid gate bootstrapping parameters");
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_param_mu);
    g_lwe_params.emplace_back(new_LweParams(n, lwe_alpha_min, lwe_alpha_max));
    g_tlwe_params.emplace_back(new_TLweParams(N, k, tlwe_alpha_min, tlwe_alpha_max));
    TGswParams* gp = new_TGswParams(bk_l, bk_Bgbit, g_tlwe_params.back().get());
    g_tgsw_params.push_back(gp);
    return new TFheGateBootstrappingParameterSet{ks_t, ks_basebit, g_lwe_params.back().get(), gp};
}
}  // namespace tfhe_amd

namespace tfhe_amd {
// Parameter objects read from files live until the process ends, as the
// reference's TfheGarbageCollector keeps them (libtfhe/tfhe_garbage_collector.cpp).
const LweParams* host_keep(LweParams* p) {
    std::lock_guard<std::mutex> g(g_param_mu);
    g_lwe_params.emplace_back(p);
    return p;
}
const TLweParams* host_keep(TLweParams* p) {
    std::lock_guard<std::mutex> g(g_param_mu);
    g_tlwe_params.emplace_back(p);
    return p;
}
const TGswParams* host_keep(TGswParams* p) {
    std::lock_guard<std::mutex> g(g_param_mu);
    g_tgsw_params.push_back(p);
    return p;
}
}  // namespace tfhe_amd

TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_gate_bootstrapping_parameters(
    int32_t n, int32_t N, int32_t k, int32_t bk_l, int32_t bk_Bgbit, int32_t ks_t, int32_t ks_basebit,
    double ks_stdev, double bk_stdev, double max_stdev) {
    return host_make_params(n, ks_stdev, max_stdev, N, k, bk_stdev, max_stdev, bk_l, bk_Bgbit, ks_t, ks_basebit);
}

// presets, libtfhe/tfhe_gate_bootstrapping.cpp:23-90
TFHE_AMD_EXPORT TFheGateBootstrappingParameterSet* new_default_gate_bootstrapping_parameters(int32_t minimum_lambda) {
    if (minimum_lambda > 128) die("Sorry, for now, the parameters are only implemented for 80bit and 128bit of security!");
    if (minimum_lambda > 80)
        return new_gate_bootstrapping_parameters(630, 1024, 1, 3, 7, 8, 2, std::pow(2., -15), std::pow(2., -25), 0.012467);
    if (minimum_lambda > 0)
        return new_gate_bootstrapping_parameters(500, 1024, 1, 2, 10, 8, 2, 2.44e-5, 7.18e-9, 0.012467);
    die("the requested security parameter must be positive (currently, 80 and 128-bits are supported)");
}

TFHE_AMD_EXPORT void delete_gate_bootstrapping_parameters(TFheGateBootstrappingParameterSet* params) { delete params; }

// ------------------------------------------------------------- samples ----

TFHE_AMD_EXPORT LweSample* new_LweSample(const LweParams* params) {
    return new LweSample{new Torus32[params->n](), 0, 0.0};
}
TFHE_AMD_EXPORT LweSample* new_LweSample_array(int32_t nbelts, const LweParams* params) {
    auto* s = new LweSample[nbelts];
    for (int i = 0; i < nbelts; i++) s[i] = LweSample{new Torus32[params->n](), 0, 0.0};
    return s;
}
TFHE_AMD_EXPORT void delete_LweSample(LweSample* obj) {
    if (!obj) return;
    delete[] obj->a;
    delete obj;
}
TFHE_AMD_EXPORT void delete_LweSample_array(int32_t nbelts, LweSample* obj) {
    if (!obj) return;
    for (int i = 0; i < nbelts; i++) delete[] obj[i].a;
    delete[] obj;
}
TFHE_AMD_EXPORT TorusPolynomial* new_TorusPolynomial(int32_t N) { return new TorusPolynomial{N, new Torus32[N]()}; }
TFHE_AMD_EXPORT void delete_TorusPolynomial(TorusPolynomial* obj) {
    if (!obj) return;
    delete[] obj->coefsT;
    delete obj;
}
TFHE_AMD_EXPORT TLweSample* new_TLweSample(const TLweParams* params) {
    const int k = params->k, N = params->N;
    auto* a = new TorusPolynomial[k + 1];
    for (int i = 0; i <= k; i++) a[i] = TorusPolynomial{N, new Torus32[N]()};
    return new TLweSample{a, a + k, 0.0, k};
}
TFHE_AMD_EXPORT void delete_TLweSample(TLweSample* obj) {
    if (!obj) return;
    for (int i = 0; i <= obj->k; i++) delete[] obj->a[i].coefsT;
    delete[] obj->a;
    delete obj;
}

// ---- keys and key material as separate objects (include/lwekey.h:69-78,
//      include/tlwe.h:159-172, include/tgsw.h:118-177, include/lwekeyswitch.h:50-55,
//      include/lwebootstrappingkey.h:82-115). Parameters are not owned.
TFHE_AMD_EXPORT LweKey* new_LweKey(const LweParams* params) { return new LweKey{params, new int32_t[params->n]()}; }
TFHE_AMD_EXPORT void delete_LweKey(LweKey* key) {
    if (!key) return;
    delete[] key->key;
    delete key;
}
TFHE_AMD_EXPORT TLweKey* new_TLweKey(const TLweParams* params) {
    auto* polys = new IntPolynomial[params->k];
    for (int i = 0; i < params->k; i++) polys[i] = IntPolynomial{params->N, new int32_t[params->N]()};
    return new TLweKey{params, polys};
}
TFHE_AMD_EXPORT void delete_TLweKey(TLweKey* obj) {
    if (!obj) return;
    for (int i = 0; i < obj->params->k; i++) delete[] obj->key[i].coefs;
    delete[] obj->key;
    delete obj;
}
// the TGSW key is its TLWE key (libtfhe/tgsw.cpp: key aliases tlwe_key.key)
TFHE_AMD_EXPORT TGswKey* new_TGswKey(const TGswParams* params) {
    TLweKey* t = new_TLweKey(params->tlwe_params);
    auto* k = new TGswKey{params, params->tlwe_params, t->key, *t};
    delete t;  // the polynomials now belong to k
    return k;
}
TFHE_AMD_EXPORT void delete_TGswKey(TGswKey* obj) {
    if (!obj) return;
    for (int i = 0; i < obj->tlwe_params->k; i++) delete[] obj->key[i].coefs;
    delete[] obj->key;
    delete obj;
}
// rows all_sample[(k+1) l], block b = rows b l .. b l + l - 1 (libtfhe/tgsw.cpp)
TFHE_AMD_EXPORT TGswSample* new_TGswSample(const TGswParams* params) {
    const int k = params->tlwe_params->k, l = params->l, kpl = params->kpl;
    auto* rows = new TLweSample[kpl];
    for (int r = 0; r < kpl; r++) {
        TLweSample* t = new_TLweSample(params->tlwe_params);
        rows[r] = *t;
        delete t;
    }
    auto* blocs = new TLweSample*[k + 1];
    for (int b = 0; b <= k; b++) blocs[b] = rows + b * l;
    return new TGswSample{rows, blocs, k, l};
}
TFHE_AMD_EXPORT void delete_TGswSample(TGswSample* obj) {
    if (!obj) return;
    const int kpl = (obj->k + 1) * obj->l;
    for (int r = 0; r < kpl; r++) {
        for (int c = 0; c <= obj->k; c++) delete[] obj->all_sample[r].a[c].coefsT;
        delete[] obj->all_sample[r].a;
    }
    delete[] obj->all_sample;
    delete[] obj->bloc_sample;
    delete obj;
}
TFHE_AMD_EXPORT LweKeySwitchKey* new_LweKeySwitchKey(int32_t n, int32_t t, int32_t basebit, const LweParams* out_params) {
    return &new_ks(n, t, basebit, out_params)->pub;
}
TFHE_AMD_EXPORT void delete_LweKeySwitchKey(LweKeySwitchKey* obj) { delete reinterpret_cast<KsOwner*>(obj); }
// with its keyswitch key from the extracted dimension k N (libtfhe/lwebootstrappingkey.cpp)
TFHE_AMD_EXPORT LweBootstrappingKey* new_LweBootstrappingKey(const int32_t ks_t, const int32_t ks_basebit,
                                                             const LweParams* in_out_params, const TGswParams* bk_params) {
    const TLweParams* tp = bk_params->tlwe_params;
    KsOwner* ks = new_ks(tp->N * tp->k, ks_t, ks_basebit, in_out_params);
    return &new_bk(in_out_params, bk_params, &ks->pub)->pub;
}
TFHE_AMD_EXPORT void delete_LweBootstrappingKey(LweBootstrappingKey* obj) { delete_bk(obj); }
// the transform-domain view used by the gates and tfhe_bootstrap_FFT: here a
// handle on the device key, built from bk (which must outlive it) on first use
TFHE_AMD_EXPORT LweBootstrappingKeyFFT* new_LweBootstrappingKeyFFT(const LweBootstrappingKey* bk) {
    LweBootstrappingKeyFFT* bf = &new_bkfft(bk)->pub;
    register_ks(bk->ks, bf);
    return bf;
}
TFHE_AMD_EXPORT void delete_LweBootstrappingKeyFFT(LweBootstrappingKeyFFT* obj) { delete_bkfft(obj); }

TFHE_AMD_EXPORT LweSample* new_gate_bootstrapping_ciphertext(const TFheGateBootstrappingParameterSet* params) {
    return new_LweSample(params->in_out_params);
}
TFHE_AMD_EXPORT LweSample* new_gate_bootstrapping_ciphertext_array(int32_t nbelems, const TFheGateBootstrappingParameterSet* params) {
    return new_LweSample_array(nbelems, params->in_out_params);
}
TFHE_AMD_EXPORT void delete_gate_bootstrapping_ciphertext(LweSample* sample) { delete_LweSample(sample); }
TFHE_AMD_EXPORT void delete_gate_bootstrapping_ciphertext_array(int32_t nbelems, LweSample* samples) {
    delete_LweSample_array(nbelems, samples);
}

// ----------------------------------------------------------- LWE on host ----

TFHE_AMD_EXPORT void lweKeyGen(LweKey* result) {
    for (int i = 0; i < result->params->n; i++) result->key[i] = draw_bit();
}

// b = m + e + <a, s>: one gaussian first, then n uniform mask values
// (libtfhe/lwe-functions.cpp:36-47)
TFHE_AMD_EXPORT void lweSymEncrypt(LweSample* result, Torus32 message, double alpha, const LweKey* key) {
    const int n = key->params->n;
    uint32_t b = U(message) + U(to_torus(draw_normal(alpha)));
    for (int i = 0; i < n; i++) {
        result->a[i] = draw_torus();
        b += U(result->a[i]) * U(key->key[i]);
    }
    result->b = S(b);
    result->current_variance = alpha * alpha;
}

TFHE_AMD_EXPORT Torus32 lwePhase(const LweSample* sample, const LweKey* key) {
    uint32_t axs = 0;
    for (int i = 0; i < key->params->n; i++) axs += U(sample->a[i]) * U(key->key[i]);
    return S(U(sample->b) - axs);
}

TFHE_AMD_EXPORT void lweNoiselessTrivial(LweSample* result, Torus32 mu, const LweParams* params) {
    memset(result->a, 0, sizeof(Torus32) * params->n);
    result->b = mu;
    result->current_variance = 0.;
}

// --------------------------------------------------------------- keysets ----

static TFheGateBootstrappingCloudKeySet make_cloud(const TFheGateBootstrappingParameterSet* P, const LweBootstrappingKey* bk) {
    const LweBootstrappingKeyFFT* bf = &new_bkfft(bk)->pub;
    register_ks(bk->ks, bf);
    return TFheGateBootstrappingCloudKeySet{P, bk, bf};
}

// Key generation in the reference's RNG order (libtfhe/tfhe_gate_bootstrapping.cpp:97-109):
// LWE key, TLWE key, keyswitch key (lwe-keyswitch-functions.cpp:163-215), then
// the TGSW encryptions of the LWE key bits (lwe-bootstrapping-functions.cpp:185-216).
// The TLWE b polynomials use the exact product (the reference's FFT product
// truncates, so its b coefficients can differ by one unit).
TFHE_AMD_EXPORT TFheGateBootstrappingSecretKeySet* new_random_gate_bootstrapping_secret_keyset(
    const TFheGateBootstrappingParameterSet* P) {
    const TGswParams* gp = P->tgsw_params;
    const TLweParams* tp = gp->tlwe_params;
    const int n = P->in_out_params->n, N = tp->N, k = tp->k, l = gp->l, kpl = gp->kpl;
    const int t = P->ks_t, basebit = P->ks_basebit, base = 1 << basebit, nin = k * N;
    auto* so = new SecretOwner();
    so->lwe_bits.resize(n);
    so->lwe_key = LweKey{P->in_out_params, so->lwe_bits.data()};
    lweKeyGen(&so->lwe_key);
    so->tlwe_bits.resize((size_t)k * N);
    so->tlwe_polys.resize(k);
    for (int i = 0; i < k; i++) so->tlwe_polys[i] = IntPolynomial{N, so->tlwe_bits.data() + (size_t)i * N};
    so->tgsw_key = TGswKey{gp, tp, so->tlwe_polys.data(), TLweKey{tp, so->tlwe_polys.data()}};
    for (int i = 0; i < k * N; i++) so->tlwe_bits[i] = draw_bit();

    // keyswitch key from the extracted key (the TLWE key bits) to the LWE key
    KsOwner* ks = new_ks(nin, t, basebit, P->in_out_params);
    {
        const double alpha = P->in_out_params->alpha_min;
        const int sizeks = nin * t * (base - 1);
        std::vector<double> noise(sizeks);
        double err = 0;
        for (int i = 0; i < sizeks; i++) {
            noise[i] = draw_normal(alpha);
            err += noise[i];
        }
        err /= sizeks;
        for (int i = 0; i < sizeks; i++) noise[i] -= err;
        int idx = 0;
        for (int i = 0; i < nin; i++)
            for (int j = 0; j < t; j++)
                for (int h = 1; h < base; h++) {
                    LweSample& row = ks->pub.ks[i][j][h];
                    const uint32_t mess = (U(so->tlwe_bits[i]) * (uint32_t)h) * (1u << (32 - (j + 1) * basebit));
                    uint32_t b = mess + U(to_torus(noise[idx++]));
                    for (int p = 0; p < n; p++) {
                        row.a[p] = draw_torus();
                        b += U(row.a[p]) * U(so->lwe_bits[p]);
                    }
                    row.b = S(b);
                    row.current_variance = alpha * alpha;
                }
    }
    // bootstrapping key: TGSW(s_i) rows = TLWE(0) + s_i * h_p on component blk
    BkOwner* bk = new_bk(P, &ks->pub);
    const double balpha = tp->alpha_min;
    for (int i = 0; i < n; i++)
        for (int r = 0; r < kpl; r++) {
            TLweSample& smp = bk->pub.bk[i].all_sample[r];
            int32_t* bpoly = smp.b->coefsT;
            for (int j = 0; j < N; j++) bpoly[j] = S(U(to_torus(draw_normal(balpha))));
            for (int c = 0; c < k; c++) {
                int32_t* apoly = smp.a[c].coefsT;
                for (int j = 0; j < N; j++) apoly[j] = draw_torus();
                add_binary_product(so->tlwe_bits.data() + (size_t)c * N, apoly, N, bpoly);
            }
            const int blk = r / l, p = r % l;
            smp.a[blk].coefsT[0] = S(U(smp.a[blk].coefsT[0]) + U(so->lwe_bits[i]) * U(gp->h[p]));
            smp.current_variance = balpha * balpha;
        }
    TFheGateBootstrappingCloudKeySet cloud = make_cloud(P, &bk->pub);
    so->pub = TFheGateBootstrappingSecretKeySet{P, &so->lwe_key, &so->tgsw_key, cloud};
    return &so->pub;
}

static const LweBootstrappingKey* bk_from_arrays(const TFheGateBootstrappingParameterSet* P, const int32_t* bkc,
                                                 const int32_t* ksk, double ks_var, double bk_var) {
    const TGswParams* gp = P->tgsw_params;
    const TLweParams* tp = gp->tlwe_params;
    const int n = P->in_out_params->n, N = tp->N, k = tp->k, kpl = gp->kpl;
    const int t = P->ks_t, base = 1 << P->ks_basebit, nin = k * N;
    KsOwner* ks = new_ks(nin, t, P->ks_basebit, P->in_out_params);
    for (size_t r = 0; r < (size_t)nin * t * base; r++) {
        memcpy(ks->rows[r].a, ksk + r * (n + 1), sizeof(int32_t) * n);
        ks->rows[r].b = ksk[r * (n + 1) + n];
        ks->rows[r].current_variance = ks_var;
    }
    BkOwner* bk = new_bk(P, &ks->pub);
    memcpy(bk->coefs.data(), bkc, sizeof(int32_t) * (size_t)n * kpl * (k + 1) * N);
    for (auto& smp : bk->samples) smp.current_variance = bk_var;
    return &bk->pub;
}

namespace tfhe_amd {
TFheGateBootstrappingCloudKeySet* host_cloud_from_arrays(const TFheGateBootstrappingParameterSet* P, const int32_t* bkc,
                                                         const int32_t* ksk, double ks_var, double bk_var) {
    auto* co = new CloudOwner();
    co->pub = make_cloud(P, bk_from_arrays(P, bkc, ksk, ks_var, bk_var));
    return &co->pub;
}

TFheGateBootstrappingCloudKeySet* host_cloud_from_bk(const TFheGateBootstrappingParameterSet* P,
                                                     LweBootstrappingKey* bk) {
    auto* co = new CloudOwner();
    co->pub = make_cloud(P, bk);
    return &co->pub;
}

TFheGateBootstrappingSecretKeySet* host_secret_from_bk(const TFheGateBootstrappingParameterSet* P,
                                                       LweBootstrappingKey* bk, const int32_t* lwe_key,
                                                       const int32_t* tlwe_key) {
    const TGswParams* gp = P->tgsw_params;
    const TLweParams* tp = gp->tlwe_params;
    const int n = P->in_out_params->n, N = tp->N, k = tp->k;
    auto* so = new SecretOwner();
    so->lwe_bits.assign(lwe_key, lwe_key + n);
    so->lwe_key = LweKey{P->in_out_params, so->lwe_bits.data()};
    so->tlwe_bits.assign(tlwe_key, tlwe_key + (size_t)k * N);
    so->tlwe_polys.resize(k);
    for (int i = 0; i < k; i++) so->tlwe_polys[i] = IntPolynomial{N, so->tlwe_bits.data() + (size_t)i * N};
    so->tgsw_key = TGswKey{gp, tp, so->tlwe_polys.data(), TLweKey{tp, so->tlwe_polys.data()}};
    so->pub = TFheGateBootstrappingSecretKeySet{P, &so->lwe_key, &so->tgsw_key, make_cloud(P, bk)};
    return &so->pub;
}

TFheGateBootstrappingSecretKeySet* host_secret_from_arrays(const TFheGateBootstrappingParameterSet* P,
                                                           const int32_t* bkc, const int32_t* ksk, double ks_var,
                                                           double bk_var, const int32_t* lwe_key,
                                                           const int32_t* tlwe_key) {
    return host_secret_from_bk(P, const_cast<LweBootstrappingKey*>(bk_from_arrays(P, bkc, ksk, ks_var, bk_var)),
                               lwe_key, tlwe_key);
}
}  // namespace tfhe_amd

TFHE_AMD_EXPORT TFheGateBootstrappingCloudKeySet* new_gate_bootstrapping_cloud_keyset_from_arrays(
    const TFheGateBootstrappingParameterSet* P, const int32_t* bkc, const int32_t* ksk) {
    if (!P || !bkc || !ksk) {
        set_error("null argument");
        return nullptr;
    }
    return host_cloud_from_arrays(P, bkc, ksk, P->in_out_params->alpha_min * P->in_out_params->alpha_min,
                                  P->tgsw_params->tlwe_params->alpha_min * P->tgsw_params->tlwe_params->alpha_min);
}

TFHE_AMD_EXPORT void delete_gate_bootstrapping_secret_keyset(TFheGateBootstrappingSecretKeySet* keyset) {
    if (!keyset) return;
    engine_forget_secret_key(keyset->lwe_key);
    delete_bkfft(keyset->cloud.bkFFT);
    delete_bk(keyset->cloud.bk);
    delete reinterpret_cast<SecretOwner*>(keyset);
}

TFHE_AMD_EXPORT void delete_gate_bootstrapping_cloud_keyset(TFheGateBootstrappingCloudKeySet* keyset) {
    if (!keyset) return;
    delete_bkfft(keyset->bkFFT);
    delete_bk(keyset->bk);
    delete reinterpret_cast<CloudOwner*>(keyset);
}

// --------------------------------------------------- encrypt / decrypt ----

// libtfhe/tfhe_gate_bootstrapping.cpp:154-165
TFHE_AMD_EXPORT void bootsSymEncrypt(LweSample* result, int32_t message, const TFheGateBootstrappingSecretKeySet* key) {
    const Torus32 s8 = modSwitchToTorus32(1, 8);
    lweSymEncrypt(result, message ? s8 : -s8, key->params->in_out_params->alpha_min, key->lwe_key);
}

TFHE_AMD_EXPORT int32_t bootsSymDecrypt(const LweSample* sample, const TFheGateBootstrappingSecretKeySet* key) {
    return lwePhase(sample, key->lwe_key) > 0 ? 1 : 0;
}

// ----------------------------------------------------------------- gates ----

namespace {
// current_variance of a keyswitched sample, as the reference keeps it: the
// result starts as a noiseless trivial sample (variance 0) and every nonzero
// digit subtracts one key row, adding that row's variance
// (lwe-keyswitch-functions.cpp:114-129,228-238; lwe-functions.cpp:140-145).
// The rows carry one variance (key generation and the key files give every
// row the same), so a count of nz rows gives nz additions of it, in order:
// ks_variances(bf)[nz] for nz = 0 .. k*N*t.
std::vector<double> ks_variances(const LweBootstrappingKeyFFT* bf) {
    const LweKeySwitchKey* ks = bf->ks;
    const double r = ks->ks0_raw[ks->base > 1 ? 1 : 0].current_variance;
    std::vector<double> v((size_t)ks->n * ks->t + 1);
    v[0] = 0.;
    for (size_t c = 1; c < v.size(); c++) v[c] = v[c - 1] + r;
    return v;
}

void one_gate(int op, LweSample* result, const LweSample* a, const LweSample* b, const LweSample* c,
              const TFheGateBootstrappingCloudKeySet* ck) {
    if (!ck) die("null cloud key");
    int32_t o = op, nz = 0;
    LweSample* outs[1] = {result};
    const LweSample* i0[1] = {a};
    const LweSample* i1[1] = {b};
    const LweSample* i2[1] = {c};
    if (engine_gates_host(ck->bkFFT, 1, &o, outs, i0, i1, i2, modSwitchToTorus32(1, 8), true, &nz)) die(g_err.c_str());
    result->current_variance = ks_variances(ck->bkFFT)[nz];
}
}  // namespace

TFHE_AMD_EXPORT void bootsNAND(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_NAND, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_OR, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsAND(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_AND, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsXOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_XOR, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsXNOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_XNOR, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsNOR(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_NOR, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsANDNY(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_ANDNY, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsANDYN(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_ANDYN, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsORNY(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_ORNY, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsORYN(LweSample* r, const LweSample* a, const LweSample* b, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_ORYN, r, a, b, nullptr, k); }
TFHE_AMD_EXPORT void bootsMUX(LweSample* r, const LweSample* a, const LweSample* b, const LweSample* c, const TFheGateBootstrappingCloudKeySet* k) { one_gate(TFHE_AMD_MUX, r, a, b, c, k); }

// libtfhe/boot-gates.cpp:164-189 (no bootstrap)
TFHE_AMD_EXPORT void bootsNOT(LweSample* result, const LweSample* ca, const TFheGateBootstrappingCloudKeySet* bk) {
    const int n = bk->params->in_out_params->n;
    for (int i = 0; i < n; i++) result->a[i] = S(0u - U(ca->a[i]));
    result->b = S(0u - U(ca->b));
    result->current_variance = ca->current_variance;
}
TFHE_AMD_EXPORT void bootsCOPY(LweSample* result, const LweSample* ca, const TFheGateBootstrappingCloudKeySet* bk) {
    const int n = bk->params->in_out_params->n;
    if (result != ca) memcpy(result->a, ca->a, sizeof(Torus32) * n);
    result->b = ca->b;
    result->current_variance = ca->current_variance;
}
TFHE_AMD_EXPORT void bootsCONSTANT(LweSample* result, int32_t value, const TFheGateBootstrappingCloudKeySet* bk) {
    const Torus32 mu = modSwitchToTorus32(1, 8);
    lweNoiselessTrivial(result, value ? mu : -mu, bk->params->in_out_params);
}

TFHE_AMD_EXPORT int32_t bootsGate_batch(int32_t count, const int32_t* gate_op, LweSample* const* out,
                                        const LweSample* const* in0, const LweSample* const* in1,
                                        const LweSample* const* in2, const TFheGateBootstrappingCloudKeySet* ck) {
    if (!ck) return set_error("null cloud key");
    for (int i = 0; i < count; i++)
        if (gate_op[i] < TFHE_AMD_NAND || gate_op[i] > TFHE_AMD_MUX) return set_error("gate %d: bad gate code %d", i, gate_op[i]);
    std::vector<int32_t> nz(count > 0 ? count : 0);
    if (engine_gates_host(ck->bkFFT, count, gate_op, out, in0, in1, in2, modSwitchToTorus32(1, 8), true, nz.data()))
        return -1;
    const std::vector<double> var = ks_variances(ck->bkFFT);
    // outputs named twice: the later gate wins, as for the ciphertexts
    for (int i = 0; i < count; i++) out[i]->current_variance = var[nz[i]];
    return 0;
}

// ------------------------------------------- bootstrapping subsystem ----

// libtfhe/lwe-bootstrapping-functions-fft.cpp:213-225
TFHE_AMD_EXPORT void tfhe_bootstrap_FFT(LweSample* result, const LweBootstrappingKeyFFT* bk, Torus32 mu, const LweSample* x) {
    int32_t op = OP_IDENTITY;
    LweSample* o[1] = {result};
    const LweSample* i0[1] = {x};
    int32_t nz = 0;
    if (engine_gates_host(bk, 1, &op, o, i0, i0, i0, mu, true, &nz)) die(g_err.c_str());
    result->current_variance = ks_variances(bk)[nz];
}

// libtfhe/lwe-bootstrapping-functions-fft.cpp:169-200 (result has dimension k*N)
TFHE_AMD_EXPORT void tfhe_bootstrap_woKS_FFT(LweSample* result, const LweBootstrappingKeyFFT* bk, Torus32 mu, const LweSample* x) {
    int32_t op = OP_IDENTITY;
    LweSample* o[1] = {result};
    const LweSample* i0[1] = {x};
    if (engine_gates_host(bk, 1, &op, o, i0, i0, i0, mu, false)) die(g_err.c_str());
    result->current_variance = 0.;
}

static tfhe_amd_device_key* key_of_rows(const TGswSampleFFT* bk, int32_t* row) {
    const LweBootstrappingKeyFFT* bf = bkfft_of_rows(bk, row);
    if (!bf) die("TGswSampleFFT does not belong to a tfhe_amd key");
    tfhe_amd_device_key* dk = engine_key(bf);
    if (!dk) die(g_err.c_str());
    return dk;
}

// libtfhe/lwe-bootstrapping-functions-fft.cpp:88-112
TFHE_AMD_EXPORT void tfhe_blindRotate_FFT(TLweSample* accum, const TGswSampleFFT* bk, const int32_t* bara, const int32_t n,
                                          const TGswParams* bk_params) {
    int32_t row = 0;
    tfhe_amd_device_key* dk = key_of_rows(bk, &row);
    const int N = bk_params->tlwe_params->N, k = bk_params->tlwe_params->k;
    std::vector<int32_t> acc((size_t)(k + 1) * N);
    for (int c = 0; c <= k; c++) memcpy(&acc[(size_t)c * N], accum->a[c].coefsT, sizeof(int32_t) * N);
    if (engine_blind_rotate_host(dk, row, n, bara, acc.data(), nullptr)) die(g_err.c_str());
    for (int c = 0; c <= k; c++) memcpy(accum->a[c].coefsT, &acc[(size_t)c * N], sizeof(int32_t) * N);
}

// libtfhe/lwe-bootstrapping-functions-fft.cpp:127-156
TFHE_AMD_EXPORT void tfhe_blindRotateAndExtract_FFT(LweSample* result, const TorusPolynomial* v, const TGswSampleFFT* bk,
                                                    const int32_t barb, const int32_t* bara, const int32_t n,
                                                    const TGswParams* bk_params) {
    int32_t row = 0;
    tfhe_amd_device_key* dk = key_of_rows(bk, &row);
    const int N = bk_params->tlwe_params->N, k = bk_params->tlwe_params->k;
    std::vector<int32_t> acc((size_t)(k + 1) * N, 0), ext((size_t)k * N + 1);
    // test vector X^{2N-barb} v in the b component
    int32_t* tv = &acc[(size_t)k * N];
    const int a = barb ? 2 * N - barb : 0;
    for (int j = 0; j < N; j++) {
        int idx = ((j - a) % (2 * N) + 2 * N) % (2 * N);
        tv[j] = idx < N ? v->coefsT[idx] : S(0u - U(v->coefsT[idx - N]));
    }
    if (engine_blind_rotate_host(dk, row, n, bara, acc.data(), ext.data())) die(g_err.c_str());
    memcpy(result->a, ext.data(), sizeof(int32_t) * k * N);
    result->b = ext[(size_t)k * N];
    result->current_variance = 0.;
}

// libtfhe/lwe-keyswitch-functions.cpp:228-238
TFHE_AMD_EXPORT void lweKeySwitch(LweSample* result, const LweKeySwitchKey* ks, const LweSample* sample) {
    const LweBootstrappingKeyFFT* bf = owner_of_ks(ks);
    if (!bf) die("lweKeySwitch: keyswitch key does not belong to a tfhe_amd key set");
    tfhe_amd_device_key* dk = engine_key(bf);
    if (!dk) die(g_err.c_str());
    const int nin = ks->n, n = ks->out_params->n;
    std::vector<int32_t> in((size_t)nin + 1), out((size_t)n + 1);
    memcpy(in.data(), sample->a, sizeof(int32_t) * nin);
    in[nin] = sample->b;
    int32_t nz = 0;
    if (engine_keyswitch_host(dk, in.data(), out.data(), &nz)) die(g_err.c_str());
    memcpy(result->a, out.data(), sizeof(int32_t) * n);
    result->b = out[n];
    result->current_variance = ks_variances(bf)[nz];
}

// --------------------------------------------- bulk helpers (no reference counterpart) ----

// Encrypt count bits with bootsSymEncrypt semantics into int32 [count][n+1].
TFHE_AMD_EXPORT int32_t tfhe_amd_encrypt_bits(const TFheGateBootstrappingSecretKeySet* key, int32_t count,
                                              const int32_t* bits, int32_t* out) {
    if (!key || count < 0 || (count && (!bits || !out))) return set_error("bad arguments");
    const int n = key->params->in_out_params->n;
    LweSample s{nullptr, 0, 0.0};
    for (int g = 0; g < count; g++) {
        s.a = out + (size_t)g * (n + 1);
        bootsSymEncrypt(&s, bits[g], key);
        out[(size_t)g * (n + 1) + n] = s.b;
    }
    return 0;
}

// Phases b - <a, s> of count samples int32 [count][n+1].
TFHE_AMD_EXPORT int32_t tfhe_amd_phases(const TFheGateBootstrappingSecretKeySet* key, int32_t count, const int32_t* cts,
                                        int32_t* phases) {
    if (!key || count < 0) return set_error("bad arguments");
    const int n = key->params->in_out_params->n;
    const int32_t* s = key->lwe_key->key;
    for (int g = 0; g < count; g++) {
        const int32_t* c = cts + (size_t)g * (n + 1);
        uint32_t axs = 0;
        for (int i = 0; i < n; i++) axs += U(c[i]) * U(s[i]);
        phases[g] = S(U(c[n]) - axs);
    }
    return 0;
}

// Copies the key material out in flat layouts: lwe_key [n], tlwe_key [k*N],
// bk [n][(k+1)l][k+1][N], ksk [k*N][t][base][n+1]. Any pointer may be NULL.
TFHE_AMD_EXPORT int32_t tfhe_amd_export_keys(const TFheGateBootstrappingSecretKeySet* key, int32_t* lwe_key,
                                             int32_t* tlwe_key, int32_t* bk, int32_t* ksk) {
    if (!key) return set_error("null key");
    const TFheGateBootstrappingParameterSet* P = key->params;
    const int n = P->in_out_params->n, N = P->tgsw_params->tlwe_params->N, k = P->tgsw_params->tlwe_params->k;
    const int kpl = P->tgsw_params->kpl;
    if (lwe_key) memcpy(lwe_key, key->lwe_key->key, sizeof(int32_t) * n);
    if (tlwe_key)
        for (int i = 0; i < k; i++) memcpy(tlwe_key + (size_t)i * N, key->tgsw_key->tlwe_key.key[i].coefs, sizeof(int32_t) * N);
    const LweBootstrappingKey* b = key->cloud.bk;
    if (bk)
        for (int i = 0; i < n; i++)
            for (int r = 0; r < kpl; r++)
                for (int c = 0; c <= k; c++)
                    memcpy(bk + (((size_t)i * kpl + r) * (k + 1) + c) * N, b->bk[i].all_sample[r].a[c].coefsT,
                           sizeof(int32_t) * N);
    if (ksk) {
        const LweKeySwitchKey* ks = b->ks;
        const size_t rows = (size_t)ks->n * ks->t * ks->base;
        for (size_t r = 0; r < rows; r++) {
            memcpy(ksk + r * (n + 1), ks->ks0_raw[r].a, sizeof(int32_t) * n);
            ksk[r * (n + 1) + n] = ks->ks0_raw[r].b;
        }
    }
    return 0;
}

// Cloud key set embedded in a secret key set.
TFHE_AMD_EXPORT const TFheGateBootstrappingCloudKeySet* tfhe_amd_cloud_of(const TFheGateBootstrappingSecretKeySet* key) {
    return key ? &key->cloud : nullptr;
}
