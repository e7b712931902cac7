// GPU engine for TFHE gate bootstrapping on MI355X (gfx950).
//
// Kernels (blind rotation = gate linear form + modulus switch, test-vector
// rotation, n external products, sample extraction: tfhe_bootstrap_woKS_FFT /
// tfhe_blindRotateAndExtract_FFT / tfhe_blindRotate_FFT,
// libtfhe/lwe-bootstrapping-functions-fft.cpp:67-200):
//   br_kernel          blind rotation, one bootstrap per 64-lane wavefront, the
//                      accumulator in LDS (the explicit-accumulator entry points)
//   br4_kernel         blind rotation for large batches: one bootstrap per
//                      one-wavefront workgroup, key read from global memory
//   br4_seg_kernel     the same over several clients' batches in one launch,
//                      each segment with its own key (serving)
//   br2_kernel         blind rotation, two wavefronts per bootstrap (one per TLWE
//                      component), partial sums handed over through LDS
//   brl_kernel         blind rotation for one gate: 4l wavefronts per bootstrap on
//                      half transforms
//   brp_kernel         blind rotation for N = 2048: four wavefronts per bootstrap,
//                      1024-point transforms on pairs of wavefronts
//   bk_convert_*       coefficient-domain bootstrapping key -> FP64 transform
//                      domain in each kernel's layout (replaces
//                      init_LweBootstrappingKeyFFT,
//                      libtfhe/lwe-bootstrapping-functions-fft.cpp:23-52)
//   ks_prep_kernel     sums the two halves of MUX gates and transposes the
//                      extracted masks for the keyswitch
//   ks_planes_kernel,  exact keyswitch as an int8 matrix-core product over byte
//   ks_mma_kernel      planes of the key (lweKeySwitch,
//                      lwe-keyswitch-functions.cpp:114-129,228-238), sliced over
//                      the input coefficients for small batches
//   ks_kernel,         exact int32 keyswitch with lanes over gates, key rows
//   ks_reduce_kernel   staged in LDS (the integer path)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <atomic>
#include <type_traits>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "negacyclic_fft.h"

using tfhe_amd::c64;

struct tfhe_amd_device_key {
    int device;
    tfhe_amd::HostParams P;
    int M, kpl, npad;
    c64* bkf;      // [n][kpl][k+1][M]     FP64 transform of every TGSW row polynomial
    int32_t* ksk;  // [k*N][t][base-1][npad]  non-trivial keyswitch rows, coefficient b at column n
    c64* tw;       // TwiddleTable<M>::COUNT
    c64* bkp;      // M = 1024 only: [n][kpl][k+1][2][512], the pair-of-wavefronts transform (FftPair)
    c64* twp;      // M = 1024 only: PAIR_TW_COUNT
    c64* bkh;      // M = 512 only: [n][kpl][k+1][2][256], the half transform (FftHalf) for the latency kernel
    c64* twh;      // M = 512 only: HALF_TW_COUNT
    int8_t* kp;    // t = 8, basebit = 2: keyswitch key as signed byte planes [4][k*N][npad][32] (int8 matrix cores)
    c64* bkc;      // M = 512 only: [n][kpl][k+1][M] in the fused-butterfly transform's order (FftCT)
    c64* twc;      // M = 512 only: FftCT::COUNT per-lane tables
    c64* bkq;      // M = 1024 only: like bkp, in the pair transform with fused butterflies (FftCTPair)
    c64* twq;      // M = 1024 only: FftCTPair::COUNT per-lane tables
    size_t bytes;
    uint64_t gen;  // unique per upload: a re-uploaded key never compares equal to a freed one
    int pins;      // holders besides the host key (servers); the memory is freed when the last one lets go
    uint32_t pair_sizes;  // M = 1024, automatic mode: bit i set when the four-wavefront pair kernel
                          // beats the two-wavefront kernel at the autotuner's ladder size i; a batch
                          // takes the choice of the smallest ladder size that holds it, batches
                          // beyond the ladder the two-wavefront kernel
    uint8_t br_choice[16];  // M = 512, automatic mode after the autotuner: the fastest kernel (mode 5,
                            // 3 or 2) at ladder size i; 0 before tuning (the crossovers apply)
    bool orphaned; // the host key has forgotten it (release or delete) while it was pinned
};

namespace tfhe_amd {

#define HIP_TRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) return set_error("%s: %s", #call, hipGetErrorString(e_));       \
    } while (0)

// ---------------------------------------------------------------- kernels --

struct BrArgs {
    const int32_t* op;
    const int32_t* in0;
    const int32_t* in1;
    const int32_t* in2;
    const int32_t* pool;
    int32_t count;              // gates; jobs [0, count) are their first bootstraps
    const int32_t* mux_gate;    // gate of job count + k (second bootstrap of the k-th MUX gate)
    const int32_t* n_jobs;      // count + number of MUX gates (device)
    int32_t mu;
    const int32_t* bara_in;  // explicit mode (one job): n modulus-switched coefficients
    const int32_t* acc_in;   // explicit mode: initial accumulator [(k+1)][N]
    int32_t* acc_out;        // explicit mode: final accumulator
    int32_t* ext;            // extracted samples [2*count][kN+1]
    const c64* bkf;          // key rows, already offset to the first row used
    const c64* tw;
    const c64* twc;  // FftCT tables (kernels on the fused-butterfly transform)
    int32_t n;
    int32_t Bgbit;
    uint32_t offset;  // gadget offset, libtfhe/tgsw.cpp:21-27
    int32_t split_round;    // > 0 when the several-bootstrap workgroups ran the first
    int32_t split_partial;  //   round_split(n_jobs, split_round, split_partial) jobs
};

// Orders LDS accesses between the lanes of one wavefront. The blind-rotation
// block is a single wavefront, whose LDS instructions execute in program
// order, so only the compiler needs fencing; a workgroup barrier here would
// also force every outstanding global load (the key prefetch) to complete.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Zero, opaque to the optimiser (an empty asm that "modifies" it): loads at
// p + opaque_zero() are not hoisted out of loops. The offset, not the pointer,
// goes through the asm, so the pointer keeps the global address space the
// compiler infers from the kernel argument and the loads stay global_load (a
// flat load would also count against the LDS wait counter, so every LDS wait
// after it would wait for the load as well).
__device__ __forceinline__ int opaque_zero() {
    int z = 0;
    asm volatile("" : "+s"(z));
    return z;
}

// a + b + c mod 2^32 in one instruction (c wave-uniform); the compiler splits the
// sum into two adds when it sees -x + uniform
__device__ __forceinline__ uint32_t add3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// round to nearest, result mod 2^32 (valid for |v| < 2^51)
__device__ __forceinline__ int32_t round_mod32(double v) {
    const double t = v + 6755399441055744.0;  // 1.5 * 2^52
    return (int32_t)(uint32_t)(unsigned long long)__double_as_longlong(t);
}

// Linear form applied to the gate inputs before the bootstrap
// (libtfhe/boot-gates.cpp:34-319; MUX halves :342-353).
__device__ __forceinline__ void gate_coeffs(int op, int half, int32_t& k0, int32_t& ca, int32_t& cb, int32_t& cc) {
    const int32_t e8 = 1 << 29, e4 = 1 << 30;
    cc = 0;
    switch (op) {
        case TFHE_AMD_NAND: k0 = e8; ca = -1; cb = -1; break;
        case TFHE_AMD_AND: k0 = -e8; ca = 1; cb = 1; break;
        case TFHE_AMD_OR: k0 = e8; ca = 1; cb = 1; break;
        case TFHE_AMD_XOR: k0 = e4; ca = 2; cb = 2; break;
        case TFHE_AMD_XNOR: k0 = -e4; ca = -2; cb = -2; break;
        case TFHE_AMD_NOR: k0 = -e8; ca = -1; cb = -1; break;
        case TFHE_AMD_ANDNY: k0 = -e8; ca = -1; cb = 1; break;
        case TFHE_AMD_ANDYN: k0 = -e8; ca = 1; cb = -1; break;
        case TFHE_AMD_ORNY: k0 = e8; ca = -1; cb = 1; break;
        case TFHE_AMD_ORYN: k0 = e8; ca = 1; cb = -1; break;
        case TFHE_AMD_MUX:
            k0 = -e8;
            if (half == 0) { ca = 1; cb = 1; } else { ca = -1; cb = 0; cc = 1; }
            break;
        default: k0 = 0; ca = 1; cb = 0; break;  // OP_IDENTITY
    }
}

template <int M, int L>
__global__ __launch_bounds__(64, M == 512 ? 2 : 1) void br_kernel(BrArgs A) {
    using F = Fft<M>;
    constexpr int R = F::R, N = 2 * M, KPL = 2 * L;
    constexpr int LOG2N2 = (M == 512) ? 11 : 12;  // log2(2N)
    __shared__ int32_t s_acc[2 * N];
    __shared__ c64 s_fft[F::LDS_C64];
    __shared__ int16_t s_bara[1024];
    __shared__ int32_t s_barb;

    const int lane = threadIdx.x;
    const int job = blockIdx.x;
    const int n = A.n;

    if (A.bara_in == nullptr) {
        if (job >= *A.n_jobs) return;
        int gate = job, half = 0;
        if (job >= A.count) { gate = A.mux_gate[job - A.count]; half = 1; }
        const int op = A.op[gate];
        int32_t k0, ca, cb, cc;
        gate_coeffs(op, half, k0, ca, cb, cc);
        const int32_t* a = A.pool + (size_t)A.in0[gate] * (n + 1);
        const int32_t* b = (op == OP_IDENTITY) ? a : A.pool + (size_t)A.in1[gate] * (n + 1);
        const int32_t* c = (op == TFHE_AMD_MUX) ? A.pool + (size_t)A.in2[gate] * (n + 1) : a;
        for (int i = lane; i <= n; i += 64) {
            uint32_t v = (uint32_t)ca * (uint32_t)a[i] + (uint32_t)cb * (uint32_t)b[i] + (uint32_t)cc * (uint32_t)c[i];
            if (i == n) v += (uint32_t)k0;
            // modSwitchFromTorus32(v, 2N), libtfhe/numeric-functions.cpp:59-65
            const uint32_t bar = (v + (1u << (31 - LOG2N2))) >> (32 - LOG2N2);
            if (i < n) s_bara[i] = (int16_t)bar;
            else s_barb = (int32_t)bar;
        }
        __syncthreads();
        const int barb = s_barb;
        // ACC = (0, X^{2N-barb} * [mu...mu]), lwe-bootstrapping-functions-fft.cpp:146-148
        for (int j = lane; j < N; j += 64) {
            s_acc[j] = 0;
            s_acc[N + j] = (((j + barb) & (2 * N - 1)) < N) ? A.mu : -A.mu;
        }
    } else {
        for (int i = lane; i < n; i += 64) s_bara[i] = (int16_t)A.bara_in[i];
        for (int j = lane; j < 2 * N; j += 64) s_acc[j] = A.acc_in[j];
    }
    __syncthreads();

    c64 tw1[R];
#pragma unroll
    for (int k = 0; k < R; k++) tw1[k] = A.tw[lane * (R + 8) + k];
    const c64 tw2 = A.tw[lane * (R + 8) + R + 1];
    const uint32_t offset = A.offset;
    const int Bgbit = A.Bgbit;
    const int32_t dhalf = 1 << (Bgbit - 1);

    // Key rows of the current digit polynomial, loaded one digit ahead of use
    // so the loads are in flight during the transform.
    c64 bk0[R], bk1[R];
    auto load_bk = [&](int row, int r) {
        const c64* q = A.bkf + ((size_t)row * KPL + r) * 2 * M + lane;
#pragma unroll
        for (int s = 0; s < R; s++) {
            bk0[s] = q[s * 64];
            bk1[s] = q[M + s * 64];
        }
    };

    for (int i = 0; i < n; i++) {
        const int abar = s_bara[i];
        if (abar == 0) continue;  // lwe-bootstrapping-functions-fft.cpp:101
        c64 r0[R], r1[R];
#pragma unroll
        for (int s = 0; s < R; s++) r0[s] = r1[s] = c64{0.0, 0.0};
        uint32_t tv[2 * R];
#pragma unroll 1
        for (int r = 0; r < KPL; r++) {
            const int c = r >= L ? 1 : 0, p = r - c * L;
            if (p == 0) {
                // T = (X^abar - 1) * ACC_c (+ gadget offset) at this lane's 2R coefficients
                const int32_t* acc = s_acc + c * N;
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int s = 0; s < R; s++) {
                        const int j = 64 * s + lane + h * M;
                        const int idx = (j - abar) & (2 * N - 1);
                        const uint32_t v = (uint32_t)acc[idx & (N - 1)];
                        const uint32_t rot = (idx & N) ? 0u - v : v;
                        tv[h * R + s] = rot - (uint32_t)acc[j] + offset;
                    }
            }
            // digit p of the truncating gadget decomposition, tgsw-functions.cpp:337-348
            const int shift = 32 - (p + 1) * Bgbit;
            c64 x[R];
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int32_t d0 = (int32_t)__builtin_amdgcn_ubfe(tv[s], shift, Bgbit) - dhalf;
                const int32_t d1 = (int32_t)__builtin_amdgcn_ubfe(tv[R + s], shift, Bgbit) - dhalf;
                x[s] = c64{(double)d0, (double)d1};
            }
            F::fwd1(x, lane, s_fft, tw1);
            wave_sync();
            F::fwd2(x, lane, s_fft, tw2);
            wave_sync();
            F::fwd3(x, lane, s_fft);
            wave_sync();
            F::fwd4(x, lane, s_fft);
            wave_sync();
            load_bk(i, r);
#pragma unroll
            for (int s = 0; s < R; s++) {
                cmac(r0[s], x[s], bk0[s]);
                cmac(r1[s], x[s], bk1[s]);
            }
        }
        // inverse transforms and ACC += result (libtfhe/lwe-bootstrapping-functions-fft.cpp:67-76)
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            c64* x = cc ? r1 : r0;
            F::inv1(x, lane, s_fft);
            wave_sync();
            F::inv2(x, lane, s_fft, tw2);
            wave_sync();
            F::inv3(x, lane, s_fft);
            wave_sync();
            F::inv4(x, lane, s_fft, tw1);
            int32_t* acc = s_acc + cc * N;
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int j = 64 * s + lane;
                acc[j] = (int32_t)((uint32_t)acc[j] + (uint32_t)round_mod32(x[s].re));
                acc[j + M] = (int32_t)((uint32_t)acc[j + M] + (uint32_t)round_mod32(x[s].im));
            }
            wave_sync();
        }
    }

    if (A.acc_out) {
        for (int j = lane; j < 2 * N; j += 64) A.acc_out[j] = s_acc[j];
    }
    if (A.ext) {
        // tLweExtractLweSampleIndex(.., 0, ..), libtfhe/lwe.cpp:41-56
        int32_t* e = A.ext + (size_t)job * (N + 1);
        for (int m = lane; m < N; m += 64) e[m] = (m == 0) ? s_acc[0] : (int32_t)(0u - (uint32_t)s_acc[N - m]);
        if (lane == 0) e[N] = s_acc[N];
    }
}

// Job list for the blind-rotation kernels: jobs [0, count) are the first bootstrap
// of every gate; jobs [count, count + n_mux) the second bootstrap of the MUX
// gates, in gate order. second[g] = job index of gate g's second bootstrap or -1.
// One block; count is split into contiguous chunks, one per thread.
__global__ __launch_bounds__(1024) void mux_scan_kernel(const int32_t* __restrict__ op, int count,
                                                        int32_t* __restrict__ mux_gate, int32_t* __restrict__ second,
                                                        int32_t* __restrict__ n_jobs) {
    __shared__ int32_t s_cnt[1024];
    const int t = threadIdx.x;
    const int per = (count + 1023) / 1024;
    const int g0 = min(count, t * per), g1 = min(count, g0 + per);
    int c = 0;
    for (int g = g0; g < g1; g++) c += (op[g] == TFHE_AMD_MUX);
    s_cnt[t] = c;
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int d = 1; d < 1024; d <<= 1) {
        const int v = t >= d ? s_cnt[t - d] : 0;
        __syncthreads();
        s_cnt[t] += v;
        __syncthreads();
    }
    int rank = s_cnt[t] - c;
    for (int g = g0; g < g1; g++) {
        if (op[g] == TFHE_AMD_MUX) {
            mux_gate[rank] = g;
            second[g] = count + rank;
            rank++;
        } else {
            second[g] = -1;
        }
    }
    if (t == 1023) *n_jobs = count + s_cnt[1023];
}

// Blind rotation with one bootstrap per one-wavefront workgroup (M = 512, i.e.
// N = 1024). Bootstraps share nothing, so each is its own workgroup: the
// dispatcher refills a CU as soon as any one of them ends, and a batch of 512
// gates still spreads over all 256 CUs. The transposes of the transform go
// through the workgroup's LDS, and the transformed key and the twiddles come
// straight from global memory (cached; the 4096 wavefronts of a batch walk the
// key rows nearly together). The accumulator lives in registers in the
// transform's coefficient layout; the negacyclic rotation is staged through
// the wavefront's transform buffer, so every rotated read is a plain load.
struct Br4Lds {
    // radix-8 form: one transform buffer, also holding (v, -v, v) for the rotation (3N int32)
    static constexpr int BUF_C64 = (Fft<512>::LDS_C64 > 3 * 1024 / 4) ? Fft<512>::LDS_C64 : 3 * 1024 / 4;
    // fused-butterfly form: two transform buffers (the two inverse transforms of a
    // row run interleaved), the first also holding the rotation's 2N words
    static constexpr int CT_C64 = FftCT::LDS_C64;
    static_assert(CT_C64 * 16 >= 2 * 1024 * 4, "rotation buffer");
};

// CT: the fused-butterfly transform (FftCT, key bkc, tables twc; LDS fb and fb2
// of Br4Lds::CT_C64 entries) instead of Fft<512> (LDS fb of Br4Lds::BUF_C64).
// The calling wavefront (lane `lane`) runs job `job`. ROW_SYNC (fused-butterfly form only): a
// workgroup barrier after every key row, which holds the workgroup's wavefronts
// (bootstraps) together so that they read each key row at about the same time.
template <int L, bool CT = false, bool ROW_SYNC = false>
__device__ __forceinline__ void br4_body(BrArgs A, const int32_t* __restrict__ mux_gate,
                                         const int32_t* __restrict__ n_jobs, c64* __restrict__ fb,
                                         int16_t* __restrict__ bara, c64* __restrict__ fb2, const int job,
                                         const int lane) {
    constexpr int M = 512;
    using F = Fft<M>;
    using G = FftCT;
    constexpr int R = F::R, N = 2 * M, KPL = 2 * L;
    constexpr int LOG2N2 = 11;

    if (job >= *n_jobs) return;
    const int n = A.n;
    int32_t* fbi = (int32_t*)fb;

    // accumulator: acc[c][h*R + s] = ACC_c[64 s + lane + h M]
    uint32_t acc[2][2 * R];
    {
        const int gate = job < A.count ? job : mux_gate[job - A.count];
        const int half = job < A.count ? 0 : 1;
        const int op = A.op[gate];
        int32_t k0, ca, cb, cc;
        gate_coeffs(op, half, k0, ca, cb, cc);
        const int32_t* a = A.pool + (size_t)A.in0[gate] * (n + 1);
        const int32_t* b = (op == OP_IDENTITY) ? a : A.pool + (size_t)A.in1[gate] * (n + 1);
        const int32_t* c = (op == TFHE_AMD_MUX) ? A.pool + (size_t)A.in2[gate] * (n + 1) : a;
        uint32_t barb = 0;
        for (int i = lane; i <= n; i += 64) {
            uint32_t v = (uint32_t)ca * (uint32_t)a[i] + (uint32_t)cb * (uint32_t)b[i] + (uint32_t)cc * (uint32_t)c[i];
            if (i == n) v += (uint32_t)k0;
            // modSwitchFromTorus32(v, 2N), libtfhe/numeric-functions.cpp:59-65
            const uint32_t bar = (v + (1u << (31 - LOG2N2))) >> (32 - LOG2N2);
            if (i < n) bara[i] = (int16_t)bar;
            else barb = bar;
        }
        // the lane that handled i == n holds barb; broadcast it
        barb = __shfl(barb, n & 63);
        // ACC = (0, X^{2N-barb} * [mu...mu]), lwe-bootstrapping-functions-fft.cpp:146-148
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int j = 64 * s + lane + h * M;
                acc[0][h * R + s] = 0;
                acc[1][h * R + s] = (((j + barb) & (2 * N - 1)) < N) ? (uint32_t)A.mu : 0u - (uint32_t)A.mu;
            }
    }

    // stage-1 twiddles psi^b w_M^{b k1}, read per transform from the table's [k1][lane] copy
    // (8 KiB, cached) through the vector memory path; the stage-2a twiddles are powers of one
    // per-lane value
    const c64* tw1 = A.tw + TwiddleTable<M>::TW1_OFF + lane;  // tw1[k1 * 64]
    const c64 tw2 = CT ? c64{0.0, 0.0} : A.tw[lane * (R + 8) + R + 1];
    // FftCT: this lane's column of every table, entry k at [k * 64]; the base
    // pointer is passed through an empty asm at every use so the compiler
    // reloads the (cached) tables where they are needed instead of hoisting
    // ~96 registers of them out of the row loop
    const ct64* twc_base = (const ct64*)A.twc;
    // the forward tables stay in registers for the whole kernel (32 VGPRs)
    ct64 fwb[4], fwc[4];
    if constexpr (CT) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            fwb[k] = twc_base[G::FB + 64 * k + lane];
            fwc[k] = twc_base[G::FC + 64 * k + lane];
        }
    }
    const uint32_t offset = A.offset;
    const int Bgbit = A.Bgbit;
    // the top bit of every digit field: with it flipped, the signed field is the balanced digit
    uint32_t flip = 0;
    for (int p = 0; p < L; p++) flip |= 1u << (31 - p * Bgbit);
    wave_sync();  // bara written by the whole wavefront

    for (int i = 0; i < n; i++) {
        const int abar = bara[i];
        c64 r0[R], r1[R];
        if constexpr (CT) {
            // digit p of component c: extraction, forward transform, products with key
            // digit polynomial (i, c l + p) added into (r0, r1) (written, for the row's
            // first digit, with the roundings of an addition into zero)
            auto digit = [&](const int c, const int p, const uint32_t* tv, auto first_tag) {
                constexpr bool FIRST = decltype(first_tag)::value;
                const int r = c * L + p;
                // digit p of the truncating gadget decomposition, tgsw-functions.cpp:337-348:
                // ((T + offset) >> shift & (Bg - 1)) - Bg/2 = the signed field of the flipped word
                const int shift = 32 - (p + 1) * Bgbit;
                c64 x[R];
#pragma unroll
                for (int s = 0; s < R; s++) {
                    const int32_t d0 = __builtin_amdgcn_sbfe((int32_t)tv[s], shift, Bgbit);
                    const int32_t d1 = __builtin_amdgcn_sbfe((int32_t)tv[R + s], shift, Bgbit);
                    x[s] = c64{(double)d0, (double)d1};
                }
                wave_sync();  // rotated reads or the previous transform done with the buffer
                // key digit polynomial (i, r): component 0 at q[s * 64], component 1 at
                // q[M + s * 64]. Component 0 is requested before the transform, half of
                // component 1 before its last stage, the other half as the component-0
                // products free their registers.
                const c64* q = A.bkf + ((size_t)i * KPL + r) * 2 * M + lane;
                c64 kv[R], kw[R / 2];
#pragma unroll
                for (int s = 0; s < R; s++) kv[s] = q[s * 64];
                G::fwd1(x, lane, fb);
                wave_sync();
                G::fwd2(x, lane, fb, fwb);
                wave_sync();
                G::fwd3(x, lane, fb);
#pragma unroll
                for (int s = 0; s < R / 2; s++) kw[s] = q[M + s * 64];
                wave_sync();
                G::fwd4(x, lane, fb, fwc);
#pragma unroll
                for (int s = 0; s < R; s++) {
                    if (FIRST) r0[s] = cmul_first(x[s], kv[s]);
                    else cmac(r0[s], x[s], kv[s]);
                    if (s < R / 2) kv[s] = q[M + (s + R / 2) * 64];
                }
#pragma unroll
                for (int s = 0; s < R; s++) {
                    const c64 k = s < R / 2 ? kw[s] : kv[s - R / 2];
                    if (FIRST) r1[s] = cmul_first(x[s], k);
                    else cmac(r1[s], x[s], k);
                }
            };
#pragma unroll
            for (int c = 0; c < 2; c++) {
                // T = (X^abar - 1) * ACC_c + gadget offset. With E = (ACC_c, -ACC_c, ACC_c),
                // (X^abar ACC_c)[j] = E[j + 2N - abar]; for abar <= N every read falls in
                // E[N, 3N) = (-ACC_c, ACC_c), otherwise in E[0, 2N) = (ACC_c, -ACC_c), so
                // the buffer holds those 2N words: one base address, no masks or selects.
                uint32_t tv[2 * R];
                {
                    const bool low = abar <= N;
                    const int vofs = low ? N : 0, nofs = low ? 0 : N;
                    wave_sync();  // the previous transform is done with the buffer
                    uint32_t nv[2 * R];  // -ACC_c, for the buffer and for T (one three-input add)
#pragma unroll
                    for (int q = 0; q < 2 * R; q++) {
                        const int j = 64 * (q % R) + lane + (q / R) * M;
                        nv[q] = 0u - acc[c][q];
                        fbi[j + nofs] = (int32_t)nv[q];
                        fbi[j + vofs] = (int32_t)acc[c][q];
                    }
                    wave_sync();
                    const int32_t* er = fbi + lane + (low ? N : 2 * N) - abar;
#pragma unroll
                    for (int q = 0; q < 2 * R; q++) tv[q] = (uint32_t)er[64 * (q % R) + (q / R) * M];
#pragma unroll
                    for (int q = 0; q < 2 * R; q++) tv[q] = add3_u32(tv[q], nv[q], offset) ^ flip;
                }
                int p0 = 0;
                if (c == 0) {
                    digit(0, 0, tv, std::true_type{});
                    p0 = 1;
                }
#pragma unroll 1
                for (int p = p0; p < L; p++) digit(c, p, tv, std::false_type{});
            }
            // the inverse tables, loaded once per key row for both components (at an
            // opaque offset, so not hoisted out of the row loop)
            ct64 iwb[4], iwa[4];
            c64 iut[R];
            const ct64* twc = twc_base + opaque_zero() + lane;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                iwb[k] = twc[G::IB + 64 * k];
                iwa[k] = twc[G::IA + 64 * k];
            }
#pragma unroll
            for (int k = 0; k < R; k++) iut[k] = ((const c64*)twc)[G::UT + 64 * k];
            // inverse transforms and ACC += result (lwe-bootstrapping-functions-fft.cpp:67-76),
            // the two components interleaved stage by stage, component 1 in the second buffer
            wave_sync();
            G::inv1(r0, lane, fb);
            G::inv1(r1, lane, fb2);
            wave_sync();
            G::inv2(r0, lane, fb, iwb);
            G::inv2(r1, lane, fb2, iwb);
            wave_sync();
            G::inv3(r0, lane, fb);
            G::inv3(r1, lane, fb2);
            wave_sync();
            G::inv4(r0, lane, fb, iwa, iut);
            G::inv4(r1, lane, fb2, iwa, iut);
#pragma unroll
            for (int s = 0; s < R; s++) {
                acc[0][s] += (uint32_t)round_mod32(r0[s].re);
                acc[0][R + s] += (uint32_t)round_mod32(r0[s].im);
                acc[1][s] += (uint32_t)round_mod32(r1[s].re);
                acc[1][R + s] += (uint32_t)round_mod32(r1[s].im);
            }
            // (no memory is shared between the wavefronts: the barrier only aligns them)
            if constexpr (ROW_SYNC) __builtin_amdgcn_s_barrier();
        } else {
#pragma unroll
            for (int s = 0; s < R; s++) r0[s] = r1[s] = c64{0.0, 0.0};
            uint32_t tv[2 * R];
#pragma unroll 1
            for (int r = 0; r < KPL; r++) {
                const int c = r >= L ? 1 : 0, p = r - c * L;
                if (p == 0) {
                    // T = (X^abar - 1) * ACC_c + gadget offset. The wave's buffer holds
                    // E = (ACC_c, -ACC_c, ACC_c), so (X^abar ACC_c)[j] = E[j + 2N - abar] with
                    // 0 < j + 2N - abar < 3N: one base address, constant offsets, no masks.
                    wave_sync();  // the previous transform is done with the buffer
#pragma unroll
                    for (int q = 0; q < 2 * R; q++) {
                        const int j = 64 * (q % R) + lane + (q / R) * M;
                        const uint32_t v = c ? acc[1][q] : acc[0][q];
                        fbi[j] = (int32_t)v;
                        fbi[j + N] = (int32_t)(0u - v);
                        fbi[j + 2 * N] = (int32_t)v;
                    }
                    wave_sync();
                    const int32_t* er = fbi + lane + 2 * N - abar;
#pragma unroll
                    for (int q = 0; q < 2 * R; q++) tv[q] = (uint32_t)er[64 * (q % R) + (q / R) * M];
#pragma unroll
                    for (int q = 0; q < 2 * R; q++) tv[q] = (tv[q] - (c ? acc[1][q] : acc[0][q]) + offset) ^ flip;
                }
                const int shift = 32 - (p + 1) * Bgbit;
                c64 x[R];
#pragma unroll
                for (int s = 0; s < R; s++) {
                    const int32_t d0 = __builtin_amdgcn_sbfe((int32_t)tv[s], shift, Bgbit);
                    const int32_t d1 = __builtin_amdgcn_sbfe((int32_t)tv[R + s], shift, Bgbit);
                    x[s] = c64{(double)d0, (double)d1};
                }
                wave_sync();  // rotated reads (p = 0) or the previous transform done with the buffer
                const c64* q = A.bkf + ((size_t)i * KPL + r) * 2 * M + lane;
                c64 kv[R];
                F::template fwd1<64>(x, lane, fb, tw1);
                wave_sync();
                F::fwd2(x, lane, fb, tw2);
                wave_sync();
                F::fwd3(x, lane, fb);
#pragma unroll
                for (int s = 0; s < R; s++) kv[s] = q[s * 64];
                wave_sync();
                F::fwd4(x, lane, fb);
#pragma unroll
                for (int s = 0; s < R; s++) {
                    cmac(r0[s], x[s], kv[s]);
                    kv[s] = q[M + s * 64];
                }
#pragma unroll
                for (int s = 0; s < R; s++) cmac(r1[s], x[s], kv[s]);
            }
            // inverse transforms and ACC += result (lwe-bootstrapping-functions-fft.cpp:67-76)
#pragma unroll
            for (int cc = 0; cc < 2; cc++) {
                c64* x = cc ? r1 : r0;
                wave_sync();
                F::inv1(x, lane, fb);
                wave_sync();
                F::inv2(x, lane, fb, tw2);
                wave_sync();
                F::inv3(x, lane, fb);
                wave_sync();
                F::template inv4<64>(x, lane, fb, tw1);
#pragma unroll
                for (int s = 0; s < R; s++) {
                    acc[cc][s] += (uint32_t)round_mod32(x[s].re);
                    acc[cc][R + s] += (uint32_t)round_mod32(x[s].im);
                }
            }
        }
    }
    // tLweExtractLweSampleIndex(.., 0, ..), libtfhe/lwe.cpp:41-56: a'_0 = a_0, a'_m = -a_{N-m}
    int32_t* e = A.ext + (size_t)job * (N + 1);
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int s = 0; s < R; s++) {
            const int j = 64 * s + lane + h * M;
            const uint32_t v = acc[0][h * R + s];
            e[j == 0 ? 0 : N - j] = (int32_t)(j == 0 ? v : 0u - v);
        }
    if (lane == 0) e[N] = (int32_t)acc[1][0];
}

// Jobs [0, split) of a batch of n jobs that the several-bootstrap workgroups of
// br4w_kernel or br2_kernel<L, 4> run, for device rounds of `round` jobs: all of them when
// n is whole rounds; otherwise every whole round but the last, which runs with the
// remainder on one-bootstrap workgroups, when that leaves at least BR4W_MIN_ROUNDS rounds
// to the former, else none. The one-bootstrap workgroups
// start the remainder's bootstraps as soon as bootstraps of the round before end, where a
// kernel boundary holds them until the slowest CU is done: that costs about a tenth of a
// round, each round on br4w_kernel saves a twenty-fifth (4,470 jobs: 25.9 ms split after
// two rounds, 24.4 ms after one, 23.7 ms all on one-bootstrap workgroups). A batch of
// less than one round with more than `partial` jobs (partial > 0) runs there as one
// partial round.
constexpr int BR4W_MIN_ROUNDS = 3;
__device__ __forceinline__ int round_split(int n, int round, int partial) {
    const int k = n / round;
    if (n == k * round) return n;
    if (k == 0 && partial > 0 && n > partial) return n;
    return k - 1 >= BR4W_MIN_ROUNDS ? (k - 1) * round : 0;
}

// Workgroup x runs job x, or job split + x after br4w_kernel ran jobs [0, split)
// (A.split_round > 0).
template <int L, bool CT = false>
__global__ __launch_bounds__(64, 8) void br4_kernel(BrArgs A, const int32_t* __restrict__ mux_gate,
                                                   const int32_t* __restrict__ n_jobs) {
    __shared__ int16_t bara[1024];
    const int job = (A.split_round > 0 ? round_split(*n_jobs, A.split_round, A.split_partial) : 0) + (int)blockIdx.x;
    if constexpr (CT) {
        __shared__ c64 fb[Br4Lds::CT_C64], fb2[Br4Lds::CT_C64];
        br4_body<L, CT>(A, mux_gate, n_jobs, fb, bara, fb2, job, (int)threadIdx.x);
    } else {
        __shared__ c64 fb[Br4Lds::BUF_C64];
        br4_body<L, CT>(A, mux_gate, n_jobs, fb, bara, nullptr, job, (int)threadIdx.x);
    }
}

// The batched kernel's rounds that fill the device (jobs [0, round_split(n_jobs,
// round)), `round` = resident workgroups x W): W bootstraps per workgroup (one per
// wavefront, each with its own 20 KB of LDS; W >= 5 leaves room for one workgroup
// per CU), workgroup x running jobs W x + w, with a workgroup barrier after every
// key row. Held together, the bootstraps of a CU read each key row at about the
// same time, so most of their key reads hit the CU's L1 (with eight: a quarter of
// the one-wavefront workgroups' L2 requests, half their fabric reads) and the device
// holds a higher clock under the load (2340 against 2010 MHz): 2.5-4% more gates/s
// over whole rounds of eight, and a batch of 1536 gates, one round of six per CU,
// 21% more than on one-bootstrap workgroups (which leave it three quarters of a
// round). Partial rounds stay on br4_kernel, which spreads them over every CU.
constexpr int BR4W_MIN = 5, BR4W_MAX = 8;
template <int L, int W>
__global__ __launch_bounds__(64 * W, 1) void br4w_kernel(BrArgs A, const int32_t* __restrict__ mux_gate,
                                                        const int32_t* __restrict__ n_jobs, int round) {
    static_assert(W >= BR4W_MIN && W <= BR4W_MAX, "bootstraps per workgroup");
    __shared__ int16_t bara[W][1024];
    __shared__ c64 fb[W][Br4Lds::CT_C64], fb2[W][Br4Lds::CT_C64];
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int job = (int)blockIdx.x * W + w;
    // uniform per workgroup except in a partial round's last workgroup, whose wavefronts
    // past the job count end (the barrier then waits for the others only)
    if (job >= round_split(*n_jobs, round, A.split_partial)) return;
    br4_body<L, true, true>(A, mux_gate, n_jobs, fb[w], bara[w], fb2[w], job, (int)threadIdx.x & 63);
}

// The bootstraps per workgroup of br4w_kernel for a batch of `count` gates on `cus` CUs:
// the bootstraps per CU the batch needs when that is 5 to 8 (one round, whole when the
// batch has no MUX gates), else BR4W_MAX. Several rounds of fewer than eight lose to the
// one-bootstrap workgroups (3072 gates as two rounds of six: 192.0k against 199.1k
// gates/s; one round of 5 / 6 / 7 per CU: +10% / +24% / +11%).
int br4w_waves(int count, int cus) {
    const int q = (count + cus - 1) / cus;
    return q >= BR4W_MIN && q <= BR4W_MAX ? q : BR4W_MAX;
}

template <int L>
void launch_br4w(int w, int blocks, hipStream_t s, const BrArgs& A, const int32_t* mux_gate, const int32_t* n_jobs,
                 int round) {
    switch (w) {
    case 5: hipLaunchKernelGGL((br4w_kernel<L, 5>), dim3(blocks), dim3(64 * 5), 0, s, A, mux_gate, n_jobs, round); break;
    case 6: hipLaunchKernelGGL((br4w_kernel<L, 6>), dim3(blocks), dim3(64 * 6), 0, s, A, mux_gate, n_jobs, round); break;
    case 7: hipLaunchKernelGGL((br4w_kernel<L, 7>), dim3(blocks), dim3(64 * 7), 0, s, A, mux_gate, n_jobs, round); break;
    default: hipLaunchKernelGGL((br4w_kernel<L, 8>), dim3(blocks), dim3(64 * 8), 0, s, A, mux_gate, n_jobs, round); break;
    }
}

template <int L>
hipError_t br4w_occupancy(int w, int* per_cu) {
    switch (w) {
    case 5: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, br4w_kernel<L, 5>, 64 * 5, 0);
    case 6: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, br4w_kernel<L, 6>, 64 * 6, 0);
    case 7: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, br4w_kernel<L, 7>, 64 * 7, 0);
    default: return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, br4w_kernel<L, 8>, 64 * 8, 0);
    }
}

// One launch over several clients' batches (segments), each with its own
// bootstrapping key and job list: workgroup (x, y) runs job x of segment y.
// The segments share the parameters, the twiddles and the ciphertext pool;
// segment y's gates are entries [off, off + count) of the launch's index arrays.
constexpr int SEG_MAX = 32;
struct BrSeg {
    const c64* bkf;
    int32_t* ext;
    const int32_t* mux_gate;
    const int32_t* n_jobs;
    int32_t off, count;
};
struct BrSegTable {
    BrArgs base;
    BrSeg seg[SEG_MAX];
};

template <int L>
__global__ __launch_bounds__(64, 8) void br4_seg_kernel(BrSegTable T) {
    // segments run the batched kernel's body on the fused-butterfly transform
    // (their keys' bkc and the tables in T.base.twc)
    const BrSeg& g = T.seg[blockIdx.y];
    BrArgs A = T.base;
    A.bkf = g.bkf;
    A.ext = g.ext;
    A.count = g.count;
    A.op += g.off;
    A.in0 += g.off;
    A.in1 += g.off;
    A.in2 += g.off;
    __shared__ c64 fb[Br4Lds::CT_C64], fb2[Br4Lds::CT_C64];
    __shared__ int16_t bara[1024];
    br4_body<L, true>(A, g.mux_gate, g.n_jobs, fb, bara, fb2, (int)blockIdx.x, (int)threadIdx.x);
}

// largest LWE dimension the latency kernel's LDS table holds
constexpr int BRL_NMAX = 1024;
// largest LWE dimension for which the autotuner compares the N = 1024 kernels
constexpr int BR4_ANY_NMAX = 768;

// Blind rotation for latency with the transforms split across pairs of
// wavefronts (M = 512, FftHalf): one bootstrap per workgroup of 4l
// wavefronts, wavefront w = h + 2c + 4p owning half h of the transform of
// digit polynomial p of accumulator component c. Per key row every wavefront
// decomposes its component (the accumulator and its rotation live in LDS), forms
// its half y_h from the whole digit polynomial, runs its 256-point transform
// and multiplies the result by its half of key row (c*l + p), held in
// registers and requested a row ahead. Wavefronts (c, 0, h), one per SIMD,
// sum the 2l products of output component c for half h, run the inverse
// 256-point transform, exchange halves with their partner and update their
// half of ACC_c. The rotated accumulator is formed once per row by the whole
// workgroup. Four workgroup barriers per row; the forward work is spread
// evenly over the SIMDs (three half transforms each for l = 3).
template <int L>
__global__ __launch_bounds__(4 * L * 64, 1) void brl_kernel(BrArgs A, const c64* __restrict__ bkh,
                                                           const c64* __restrict__ twh,
                                                           const int32_t* __restrict__ mux_gate,
                                                           const int32_t* __restrict__ n_jobs) {
    using H = FftHalf;
    constexpr int R = H::R, N = 1024, KPL = 2 * L, W = 4 * L;
    constexpr int LOG2N2 = 11;
    __shared__ c64 s_fft[W][H::LDS_C64];  // transform buffer, then the product for output 0
    __shared__ c64 s_par[W][256];         // product for output 1
    __shared__ c64 s_inv[4][H::LDS_C64];  // inverse transforms of wavefronts (c, 0, h)
    __shared__ int32_t s_acc[2][N];
    __shared__ uint32_t s_tv[2][N];  // (X^abar - 1) * ACC + gadget offset, digit tops flipped
    __shared__ int16_t s_bara[BRL_NMAX];
    __shared__ int32_t s_barb;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int w = tid >> 6;
    const int h = w & 1, c = (w >> 1) & 1, p = w >> 2;
    const int job = blockIdx.x;
    if (job >= *n_jobs) return;  // uniform per workgroup
    const int n = A.n;
    c64* fb = s_fft[w];

    if (w == 0) {
        const int gate = job < A.count ? job : mux_gate[job - A.count];
        const int half = job < A.count ? 0 : 1;
        const int op = A.op[gate];
        int32_t k0, ca, cb, cc;
        gate_coeffs(op, half, k0, ca, cb, cc);
        const int32_t* a = A.pool + (size_t)A.in0[gate] * (n + 1);
        const int32_t* b = (op == OP_IDENTITY) ? a : A.pool + (size_t)A.in1[gate] * (n + 1);
        const int32_t* cp = (op == TFHE_AMD_MUX) ? A.pool + (size_t)A.in2[gate] * (n + 1) : a;
        for (int i = lane; i <= n; i += 64) {
            uint32_t v = (uint32_t)ca * (uint32_t)a[i] + (uint32_t)cb * (uint32_t)b[i] + (uint32_t)cc * (uint32_t)cp[i];
            if (i == n) v += (uint32_t)k0;
            // modSwitchFromTorus32(v, 2N), libtfhe/numeric-functions.cpp:59-65
            const uint32_t bar = (v + (1u << (31 - LOG2N2))) >> (32 - LOG2N2);
            if (i < n) s_bara[i] = (int16_t)bar;
            else s_barb = (int32_t)bar;
        }
    }
    // this lane's twiddles: tw1[0..3], w_64^{g mu}, w_16^{g0 nu}
    c64 tw[HALF_TW_PER_LANE];
#pragma unroll
    for (int k = 0; k < HALF_TW_PER_LANE; k++) tw[k] = twh[(h * 64 + lane) * HALF_TW_PER_LANE + k];
    const uint32_t offset = A.offset;
    const int Bgbit = A.Bgbit;
    // the top bit of every digit field: with it flipped, the signed field is the balanced digit
    uint32_t flip = 0;
    for (int pp = 0; pp < L; pp++) flip |= 1u << (31 - pp * Bgbit);
    __syncthreads();
    // ACC = (0, X^{2N-barb} * [mu...mu]), lwe-bootstrapping-functions-fft.cpp:146-148.
    // Wavefront (c, 0, h) owns coefficients lane + 64 a + 256 h (+ 512) of ACC_c in
    // registers (acc[a], acc[R + a]); s_acc is the copy the rotation reads.
    const int barb = s_barb;
    auto acc_init = [&](int cc, int j) { return cc == 0 ? 0 : ((((j + barb) & (2 * N - 1)) < N) ? A.mu : -A.mu); };
    for (int j = tid; j < N; j += W * 64) {
        s_acc[0][j] = acc_init(0, j);
        s_acc[1][j] = acc_init(1, j);
    }
    uint32_t acc[2 * R];
#pragma unroll
    for (int a = 0; a < R; a++) {
        const int j = lane + 64 * a + 256 * h;
        acc[a] = (uint32_t)acc_init(c, j);
        acc[R + a] = (uint32_t)acc_init(c, j + 512);
    }
    __syncthreads();

    // rows whose rotation is zero are skipped (lwe-bootstrapping-functions-fft.cpp:101;
    // uniform per workgroup)
    auto next_row = [&](int from) {
        while (from < n && s_bara[from] == 0) from++;
        return from;
    };
    // half h of key row (i, c*l + p) for both output components, into registers;
    // row i + 1's is requested as soon as row i's products are formed
    c64 k0[R], k1[R];
    auto load_key = [&](int row) {
        const c64* q = bkh + ((((size_t)row * KPL + c * L + p) * 2) * 2 + h) * 256 + lane;
#pragma unroll
        for (int s = 0; s < R; s++) {
            k0[s] = q[s * 64];
            k1[s] = q[512 + s * 64];
        }
    };
    int i = next_row(0);
    if (i < n) load_key(i);
    while (i < n) {
        const int abar = s_bara[i];
        const int inext = next_row(i + 1);
        // T = (X^abar - 1) * ACC + gadget offset, both components, once for the workgroup
        {
            constexpr int PER = (2 * N + W * 64 - 1) / (W * 64);
            uint32_t rv[PER], av[PER];
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int j2 = tid + u * W * 64, cc = (j2 >> 10) & 1, j = j2 & (N - 1);
                rv[u] = (uint32_t)s_acc[cc][(j - abar) & (N - 1)];
                av[u] = (uint32_t)s_acc[cc][j];
            }
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int j2 = tid + u * W * 64, cc = (j2 >> 10) & 1, j = j2 & (N - 1);
                if (j2 < 2 * N) s_tv[cc][j] = (((((j - abar) & (2 * N - 1)) & N) ? 0u - rv[u] : rv[u]) - av[u] + offset) ^ flip;
            }
        }
        __syncthreads();
        // digit p of component c (tgsw-functions.cpp:337-348), coefficient
        // j = lane + 64 a + 256 q in tv[q * R + a]
        uint32_t tv[4 * R];
#pragma unroll
        for (int k = 0; k < 4 * R; k++) tv[k] = s_tv[c][64 * k + lane];
        const int shift = 32 - (p + 1) * Bgbit;
        c64 x[R];
#pragma unroll
        for (int a = 0; a < R; a++) {
            // u_{j'} = d_{j'} + i d_{j'+512}, u_{j'+256} = d_{j'+256} + i d_{j'+768}, j' = lane + 64 a;
            // y_h = u_{j'} +- e^{i pi/4} u_{j'+256}: the integer parts of e^{i pi/4} u_{j'+256}
            // are formed exactly in int32, then one FMA per component
            int32_t d[4];
#pragma unroll
            for (int q = 0; q < 4; q++) d[q] = __builtin_amdgcn_sbfe((int32_t)tv[q * R + a], shift, Bgbit);
            const double sq = h == 0 ? H::SQ : -H::SQ;
            x[a] = c64{fma(sq, (double)(d[1] - d[3]), (double)d[0]), fma(sq, (double)(d[1] + d[3]), (double)d[2])};
        }
        if (h == 0) H::fwd1<0>(x, lane, fb, tw);
        else H::fwd1<1>(x, lane, fb, tw);
        wave_sync();
        H::fwd2(x, lane, fb, tw);
        wave_sync();
        H::fwd2w(x, lane, fb);
        wave_sync();
        H::fwd3(x, lane, fb, tw);
        wave_sync();
        H::fwd3w(x, lane, fb);
        wave_sync();
        H::fwd4(x, lane, fb);
        wave_sync();
#pragma unroll
        for (int s = 0; s < R; s++) {
            fb[s * 64 + lane] = cmul(x[s], k0[s]);
            s_par[w][s * 64 + lane] = cmul(x[s], k1[s]);
        }
        if (inext < n) load_key(inext);
        __syncthreads();  // all products written
        if (p == 0) {
            // output component c, half h: the 2l products of the wavefronts (c', p', h)
            c64 y[R];
#pragma unroll
            for (int s = 0; s < R; s++) y[s] = c64{0.0, 0.0};
#pragma unroll
            for (int v = 0; v < 2 * L; v++) {
                const int wv = h + 2 * (v & 1) + 4 * (v >> 1);
                const c64* src = c ? s_par[wv] : s_fft[wv];
#pragma unroll
                for (int s = 0; s < R; s++) y[s] = cadd(y[s], src[s * 64 + lane]);
            }
            c64* ib = s_inv[w];
            H::inv1(y, lane, ib);
            wave_sync();
            H::inv2(y, lane, ib, tw);
            wave_sync();
            H::inv2w(y, lane, ib);
            wave_sync();
            H::inv3(y, lane, ib, tw);
            wave_sync();
            H::inv3w(y, lane, ib);
            wave_sync();
            if (h == 0) H::inv4<0>(y, lane, ib, tw);
            else H::inv4<1>(y, lane, ib, tw);
            wave_sync();
#pragma unroll
            for (int s = 0; s < R; s++) ib[s * 64 + lane] = y[s];
            __syncthreads();  // both halves of each component written
            // ACC_c += result (lwe-bootstrapping-functions-fft.cpp:67-76): half h holds
            // coefficients j' + 256 h and j' + 256 h + 512, j' = lane + 64 a
            const c64* o = s_inv[w ^ 1];
            int32_t* sa = s_acc[c];
#pragma unroll
            for (int a = 0; a < R; a++) {
                const c64 yo = o[a * 64 + lane];
                const c64 u = h == 0 ? H::pair_inv_lo(y[a], yo) : H::pair_inv_hi(yo, y[a]);
                const int j = lane + 64 * a + 256 * h;
                acc[a] += (uint32_t)round_mod32(u.re);
                acc[R + a] += (uint32_t)round_mod32(u.im);
                sa[j] = (int32_t)acc[a];
                sa[j + 512] = (int32_t)acc[R + a];
            }
        } else {
            __syncthreads();
        }
        __syncthreads();  // ACC updated for the next row
        i = inext;
    }
    if (w != 0) return;
    // tLweExtractLweSampleIndex(.., 0, ..), libtfhe/lwe.cpp:41-56
    int32_t* e = A.ext + (size_t)job * (N + 1);
    for (int j = lane; j < N; j += 64) e[j == 0 ? 0 : N - j] = j == 0 ? s_acc[0][0] : (int32_t)(0u - (uint32_t)s_acc[0][j]);
    if (lane == 0) e[N] = s_acc[1][0];
}

// Half-transform copy of every bootstrapping-key polynomial (N = 1024): one
// 128-thread workgroup per polynomial, output [poly][h][slot][lane].
__global__ __launch_bounds__(128) void bk_convert_half_kernel(const int32_t* __restrict__ bk, c64* __restrict__ bkh,
                                                              const c64* __restrict__ twh) {
    using H = FftHalf;
    constexpr int R = H::R;
    __shared__ c64 s_fft[2][H::LDS_C64];
    const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * 1024;
    c64* fb = s_fft[h];
    c64 tw[HALF_TW_PER_LANE], x[R];
#pragma unroll
    for (int k = 0; k < HALF_TW_PER_LANE; k++) tw[k] = twh[(h * 64 + lane) * HALF_TW_PER_LANE + k];
#pragma unroll
    for (int a = 0; a < R; a++) {
        const int j = lane + 64 * a;
        const c64 lo{(double)p[j], (double)p[j + 512]}, hi{(double)p[j + 256], (double)p[j + 768]};
        x[a] = h == 0 ? H::pair_fwd<0>(lo, hi) : H::pair_fwd<1>(lo, hi);
    }
    if (h == 0) H::fwd1<0>(x, lane, fb, tw);
    else H::fwd1<1>(x, lane, fb, tw);
    wave_sync();
    H::fwd2(x, lane, fb, tw);
    wave_sync();
    H::fwd2w(x, lane, fb);
    wave_sync();
    H::fwd3(x, lane, fb, tw);
    wave_sync();
    H::fwd3w(x, lane, fb);
    wave_sync();
    H::fwd4(x, lane, fb);
    c64* out = bkh + (poly * 2 + h) * 256;
#pragma unroll
    for (int s = 0; s < R; s++) out[s * 64 + lane] = x[s];
}

// largest LWE dimension the two-wavefront kernel's LDS table holds
constexpr int BR2_NMAX = 1024;

// Blind rotation with two wavefronts per bootstrap on the fused-butterfly
// transform (M = 512): wavefront c owns TLWE component c of the accumulator (in
// registers). Per key row it rotates and decomposes its own component (through
// its 2N-word buffer, as br4_kernel), transforms its l digit polynomials and
// multiply-accumulates them against key rows (c l + p) of both output
// components, read straight from global memory; then it hands the partial sum
// of the other component to its partner through LDS, takes the partner's partial
// sum of its own, inverse-transforms and updates its component. Two workgroup
// barriers per row; four bootstraps per CU. B = 1: one bootstrap per 128-thread
// workgroup, workgroup x running job x (or job split + x after the B = 4 kernel
// ran jobs [0, split), A.split_round > 0). B = 4: whole device rounds (jobs
// [0, round_split(n_jobs, round))), four bootstraps per 512-thread workgroup,
// pair b of workgroup x running job 4 x + b with its own LDS (one workgroup per
// CU): the barriers then hold the CU's four bootstraps together, so their key
// reads share the CU's L1 (one whole round of 1,024 two-input gates: 176.6k ->
// 188.7k gates/s; two per workgroup 177.5k).
template <int L, int B = 1>
__global__ __launch_bounds__(128 * B, 2 / (B > 1 ? 2 : 1)) void br2_kernel(BrArgs A, const int32_t* __restrict__ mux_gate,
                                                                          const int32_t* __restrict__ n_jobs, int round) {
    constexpr int M = 512;
    using G = FftCT;
    constexpr int R = 8, N = 2 * M, KPL = 2 * L, LOG2N2 = 11;
    __shared__ c64 s_fft_b[B][2][G::LDS_C64];  // transform and rotation buffer of wavefront c
    __shared__ c64 s_swap_b[B][2][M];          // s_swap[d]: partial sum of output d from wavefront 1 - d
    __shared__ int16_t s_bara_b[B][BR2_NMAX];
    __shared__ int32_t s_barb_b[B];
    const int pb = B == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);  // this bootstrap's pair
    auto& s_fft = s_fft_b[pb];
    auto& s_swap = s_swap_b[pb];
    auto& s_bara = s_bara_b[pb];
    int32_t& s_barb = s_barb_b[pb];

    const int tid = B == 1 ? (int)threadIdx.x : (int)threadIdx.x & 127;
    const int lane = tid & 63;
    const int c = __builtin_amdgcn_readfirstlane(tid >> 6);
    int job;
    if constexpr (B == 1) {
        job = (A.split_round > 0 ? round_split(*n_jobs, A.split_round, A.split_partial) : 0) + (int)blockIdx.x;
    } else {
        job = (int)blockIdx.x * B + pb;
        // uniform per workgroup except in a partial round's last workgroup
        if (job >= round_split(*n_jobs, round, A.split_partial)) return;
    }
    // uniform per pair; its wavefronts end, and the workgroup's barriers then wait for the
    // other pairs only
    if (job >= *n_jobs) return;
    const int n = A.n;
    c64* fb = s_fft[c];
    int32_t* fbi = (int32_t*)fb;

    if (c == 1) {
        // gate linear form and modulus switch (wavefront 1, while wavefront 0 loads its tables)
        const int gate = job < A.count ? job : mux_gate[job - A.count];
        const int half = job < A.count ? 0 : 1;
        const int op = A.op[gate];
        int32_t k0, ca, cb, cc;
        gate_coeffs(op, half, k0, ca, cb, cc);
        const int32_t* a = A.pool + (size_t)A.in0[gate] * (n + 1);
        const int32_t* b = (op == OP_IDENTITY) ? a : A.pool + (size_t)A.in1[gate] * (n + 1);
        const int32_t* cp = (op == TFHE_AMD_MUX) ? A.pool + (size_t)A.in2[gate] * (n + 1) : a;
        for (int i = lane; i <= n; i += 64) {
            uint32_t v =
                (uint32_t)ca * (uint32_t)a[i] + (uint32_t)cb * (uint32_t)b[i] + (uint32_t)cc * (uint32_t)cp[i];
            if (i == n) v += (uint32_t)k0;
            // modSwitchFromTorus32(v, 2N), libtfhe/numeric-functions.cpp:59-65
            const uint32_t bar = (v + (1u << (31 - LOG2N2))) >> (32 - LOG2N2);
            if (i < n) s_bara[i] = (int16_t)bar;
            else s_barb = (int32_t)bar;
        }
    }
    // the forward tables stay in registers for the whole kernel (32 VGPRs)
    const ct64* twc_base = (const ct64*)A.twc;
    ct64 fwb[4], fwc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        fwb[k] = twc_base[G::FB + 64 * k + lane];
        fwc[k] = twc_base[G::FC + 64 * k + lane];
    }
    const uint32_t offset = A.offset;
    const int Bgbit = A.Bgbit;
    // the top bit of every digit field: with it flipped, the signed field is the balanced digit
    uint32_t flip = 0;
    for (int p = 0; p < L; p++) flip |= 1u << (31 - p * Bgbit);
    __syncthreads();
    // ACC_c in registers: acc[h*R + s] = ACC_c[64 s + lane + h M];
    // ACC = (0, X^{2N-barb} * [mu...mu]), lwe-bootstrapping-functions-fft.cpp:146-148
    uint32_t acc[2 * R];
    {
        const int barb = s_barb;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int j = 64 * s + lane + h * M;
                acc[h * R + s] =
                    c == 0 ? 0u : ((((j + barb) & (2 * N - 1)) < N) ? (uint32_t)A.mu : 0u - (uint32_t)A.mu);
            }
    }

    for (int i = 0; i < n; i++) {
        const int abar = s_bara[i];
        c64 r0[R], r1[R];  // partial spectra of output components 0 and 1
        // T = (X^abar - 1) * ACC_c + gadget offset, through this wavefront's buffer: for
        // abar <= N every rotated read E[j + 2N - abar] of E = (ACC_c, -ACC_c, ACC_c) falls
        // in E[N, 3N) = (-ACC_c, ACC_c), otherwise in E[0, 2N) = (ACC_c, -ACC_c)
        uint32_t tv[2 * R];
        {
            const bool low = abar <= N;
            const int vofs = low ? N : 0, nofs = low ? 0 : N;
            wave_sync();  // the previous inverse transform is done with the buffer
            uint32_t nv[2 * R];  // -ACC_c, for the buffer and for T (one three-input add)
#pragma unroll
            for (int q = 0; q < 2 * R; q++) {
                const int j = 64 * (q % R) + lane + (q / R) * M;
                nv[q] = 0u - acc[q];
                fbi[j + nofs] = (int32_t)nv[q];
                fbi[j + vofs] = (int32_t)acc[q];
            }
            wave_sync();
            const int32_t* er = fbi + lane + (low ? N : 2 * N) - abar;
#pragma unroll
            for (int q = 0; q < 2 * R; q++) tv[q] = (uint32_t)er[64 * (q % R) + (q / R) * M];
#pragma unroll
            for (int q = 0; q < 2 * R; q++) tv[q] = add3_u32(tv[q], nv[q], offset) ^ flip;
        }
        // digit p of this component: extraction (tgsw-functions.cpp:337-348, the signed field
        // of the flipped word), forward transform, products with key digit polynomial
        // (i, c l + p) of both output components (written, for the first digit)
        auto digit = [&](const int p, auto first_tag) {
            constexpr bool FIRST = decltype(first_tag)::value;
            const int shift = 32 - (p + 1) * Bgbit;
            c64 x[R];
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int32_t d0 = __builtin_amdgcn_sbfe((int32_t)tv[s], shift, Bgbit);
                const int32_t d1 = __builtin_amdgcn_sbfe((int32_t)tv[R + s], shift, Bgbit);
                x[s] = c64{(double)d0, (double)d1};
            }
            wave_sync();  // rotated reads or the previous transform done with the buffer
            const c64* q = A.bkf + ((size_t)i * KPL + c * L + p) * 2 * M + lane;
            c64 kv[R], kw[R / 2];
#pragma unroll
            for (int s = 0; s < R; s++) kv[s] = q[s * 64];
            G::fwd1(x, lane, fb);
            wave_sync();
            G::fwd2(x, lane, fb, fwb);
            wave_sync();
            G::fwd3(x, lane, fb);
#pragma unroll
            for (int s = 0; s < R / 2; s++) kw[s] = q[M + s * 64];
            wave_sync();
            G::fwd4(x, lane, fb, fwc);
#pragma unroll
            for (int s = 0; s < R; s++) {
                if (FIRST) r0[s] = cmul_first(x[s], kv[s]);
                else cmac(r0[s], x[s], kv[s]);
                if (s < R / 2) kv[s] = q[M + (s + R / 2) * 64];
            }
#pragma unroll
            for (int s = 0; s < R; s++) {
                const c64 k = s < R / 2 ? kw[s] : kv[s - R / 2];
                if (FIRST) r1[s] = cmul_first(x[s], k);
                else cmac(r1[s], x[s], k);
            }
        };
        digit(0, std::true_type{});
#pragma unroll 1
        for (int p = 1; p < L; p++) digit(p, std::false_type{});
        // hand the other component's partial sum to the partner, take this one's
        __syncthreads();  // the partner has read the previous row's hand-over
        c64 x[R];
        if (c == 0) {
#pragma unroll
            for (int s = 0; s < R; s++) s_swap[1][s * 64 + lane] = r1[s];
        } else {
#pragma unroll
            for (int s = 0; s < R; s++) s_swap[0][s * 64 + lane] = r0[s];
        }
        __syncthreads();  // both hand-overs written
        if (c == 0) {
#pragma unroll
            for (int s = 0; s < R; s++) x[s] = cadd(r0[s], s_swap[0][s * 64 + lane]);
        } else {
#pragma unroll
            for (int s = 0; s < R; s++) x[s] = cadd(r1[s], s_swap[1][s * 64 + lane]);
        }
        // inverse transform of component c and ACC_c += result
        // (lwe-bootstrapping-functions-fft.cpp:67-76); the inverse tables are loaded
        // once per row at an opaque offset, so they are not hoisted out of the loop
        const ct64* twc = twc_base + opaque_zero() + lane;
        ct64 iwb[4], iwa[4];
        c64 iut[R];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            iwb[k] = twc[G::IB + 64 * k];
            iwa[k] = twc[G::IA + 64 * k];
        }
#pragma unroll
        for (int k = 0; k < R; k++) iut[k] = ((const c64*)twc)[G::UT + 64 * k];
        wave_sync();
        G::inv1(x, lane, fb);
        wave_sync();
        G::inv2(x, lane, fb, iwb);
        wave_sync();
        G::inv3(x, lane, fb);
        wave_sync();
        G::inv4(x, lane, fb, iwa, iut);
#pragma unroll
        for (int s = 0; s < R; s++) {
            acc[s] += (uint32_t)round_mod32(x[s].re);
            acc[R + s] += (uint32_t)round_mod32(x[s].im);
        }
    }
    // tLweExtractLweSampleIndex(.., 0, ..), libtfhe/lwe.cpp:41-56: a'_0 = a_0, a'_m = -a_{N-m}, b' = b_0
    int32_t* e = A.ext + (size_t)job * (N + 1);
    if (c == 0) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int s = 0; s < R; s++) {
                const int j = 64 * s + lane + h * M;
                const uint32_t v = acc[h * R + s];
                e[j == 0 ? 0 : N - j] = (int32_t)(j == 0 ? v : 0u - v);
            }
    } else if (lane == 0) {
        e[N] = (int32_t)acc[0];
    }
}

// largest LWE dimension the pair kernel's LDS table holds
constexpr int BRP_NMAX = 768;

// Blind rotation for N = 2048 with one bootstrap per 256-thread workgroup of
// four wavefronts (two workgroups per CU). Wavefront (c, h): accumulator
// component c, half h of the pair-of-wavefronts transform (FftPair); it owns
// the quarters m = h and h + 2 of ACC_c (coefficients 512 m + 64 a + lane,
// a < 8). At the start of a key row each wavefront also holds its partner's
// quarters, so it writes the whole of ACC_c into its own LDS buffer and the
// rotation, the gadget decomposition and the transform's radix-2 step across
// the pair (y_h from all four quarters) need no exchange inside the pair. Each
// wavefront then runs its 512-point transform and multiply-accumulates its
// bins against its half of key rows (c*l + p), read straight from global
// memory (cached); the component partners swap partial sums, each pair
// inverse-transforms its own component and exchanges the two halves once, and
// both wavefronts of the pair form the whole update (the partner's old
// quarters come through s_xacc). Four workgroup barriers per key row; the two
// workgroups on a CU never wait for each other.
// CT: the transform with fused butterflies (FftCTPair; key in its order, twp
// its tables: the forward ones staged in LDS where s_tw1 was, the inverse ones
// read once per key row) instead of the radix-8 stages of Fft<512>.
template <int L, bool CT = false>
__global__ __launch_bounds__(256, 2) void brp_kernel(BrArgs A, const c64* __restrict__ bkp, const c64* __restrict__ twp,
                                                    const int32_t* __restrict__ mux_gate,
                                                    const int32_t* __restrict__ n_jobs) {
    using F = Fft<512>;
    using P = FftPair;
    using G = FftCT;
    constexpr int R = 8, N = 2048, KPL = 2 * L, LOG2N2 = 12;
    constexpr int NMAX = BRP_NMAX;
    __shared__ c64 s_fft[4][F::LDS_C64];
    __shared__ int32_t s_xacc[4][2 * R * 64];  // wavefront w's old quarters of ACC_c, for its partner
    // radix-8: [h][k1 * 64 + lane] stage-1 twiddles; CT: [h][k * 64 + lane] the FB (k < 4)
    // and FC (k >= 4) entries of wavefront h's tables, as ct64
    __shared__ c64 s_tw1[2][R * 64];
    __shared__ int16_t s_bara[NMAX];
    __shared__ int32_t s_barb;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // wave-uniform. The fused form takes it as a scalar: its c/h offsets then sit in SGPRs and
    // it needs no spill within 256 VGPRs (the radix-8 form spills less with the vector value)
    const int w = CT ? __builtin_amdgcn_readfirstlane(tid >> 6) : tid >> 6;
    const int c = w >> 1, h = w & 1;
    const int job = blockIdx.x;
    if (job >= *n_jobs) return;  // whole block idle (uniform)
    const int n = A.n;
    int16_t* bara = s_bara;
    c64* fb = s_fft[w];
    int32_t* fbi = (int32_t*)fb;
    int32_t* const acc_out = s_xacc[w];
    const int32_t* const acc_in = s_xacc[w ^ 1];

    // gate linear form and modulus switch, one wavefront per bootstrap
    if (c == 1 && h == 0) {
        {
            const int gate = job < A.count ? job : mux_gate[job - A.count];
            const int half = job < A.count ? 0 : 1;
            const int op = A.op[gate];
            int32_t k0, ca, cb, cc;
            gate_coeffs(op, half, k0, ca, cb, cc);
            const int32_t* a = A.pool + (size_t)A.in0[gate] * (n + 1);
            const int32_t* b = (op == OP_IDENTITY) ? a : A.pool + (size_t)A.in1[gate] * (n + 1);
            const int32_t* cp = (op == TFHE_AMD_MUX) ? A.pool + (size_t)A.in2[gate] * (n + 1) : a;
            for (int i = lane; i <= n; i += 64) {
                uint32_t v =
                    (uint32_t)ca * (uint32_t)a[i] + (uint32_t)cb * (uint32_t)b[i] + (uint32_t)cc * (uint32_t)cp[i];
                if (i == n) v += (uint32_t)k0;
                // modSwitchFromTorus32(v, 2N), libtfhe/numeric-functions.cpp:59-65
                const uint32_t bar = (v + (1u << (31 - LOG2N2))) >> (32 - LOG2N2);
                if (i < n) bara[i] = (int16_t)bar;
                else s_barb = (int32_t)bar;
            }
        }
    }
    for (int k = tid; k < 2 * R * 64; k += 256) {
        const int hh = k >> 9, k1 = (k >> 6) & 7, l = k & 63;
        if constexpr (CT)  // FB entries 0..3, then FC entries 0..3
            s_tw1[hh][k1 * 64 + l] = twp[hh * G::COUNT + (k1 < 4 ? G::FB : G::FC - 4 * 64) + k1 * 64 + l];
        else
            s_tw1[hh][k1 * 64 + l] = twp[(hh * 64 + l) * 16 + k1];
    }
    const c64* tw1 = s_tw1[h] + lane;
    const c64 tw2 = CT ? c64{0.0, 0.0} : twp[(h * 64 + lane) * 16 + 9];
    const ct64* const ctw = (const ct64*)s_tw1[h] + lane;  // CT: FB at [k * 64], FC at [(4 + k) * 64]
    const c64* const twq_base = twp + h * G::COUNT;         // CT: this wavefront's tables in global memory
    const uint32_t offset = A.offset;
    const int Bgbit = A.Bgbit;
    const int32_t dhalf = 1 << (Bgbit - 1);
    __syncthreads();

    // own quarters acc[q * R + a] = ACC_c[512 (h + 2q) + 64 a + lane]; pacc: the partner's
    // quarters (1 - h + 2q), valid from the end of one key row to the start of the next.
    uint32_t acc[2 * R], pacc[2 * R];
#pragma unroll
    for (int k = 0; k < 2 * R; k++) acc[k] = pacc[k] = 0;
    if (c == 1) {
        // ACC_1 = X^{2N-barb} * [mu...mu], lwe-bootstrapping-functions-fft.cpp:146-148
        const int barb = s_barb;
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int a = 0; a < R; a++) {
                const int jo = 512 * (h + 2 * q) + 64 * a + lane, jp = 512 * (1 - h + 2 * q) + 64 * a + lane;
                acc[q * R + a] = (((jo + barb) & (2 * N - 1)) < N) ? (uint32_t)A.mu : 0u - (uint32_t)A.mu;
                pacc[q * R + a] = (((jp + barb) & (2 * N - 1)) < N) ? (uint32_t)A.mu : 0u - (uint32_t)A.mu;
            }
    }

    for (int i = 0; i < n; i++) {
        const int abar = bara[i];
        c64 p0[R], p1[R];  // set by the first digit's product
        // digits 1..L-1 of the own and the partner's quarters, two 16-bit fields per word;
        // the next one to use is always dig[0]
        uint32_t dig[L - 1][2][R];
        // one gadget digit: decomposition, transform, product with the key read from global
        // memory. FIRST (digit 0) also rotates ACC_c and decomposes every digit.
        auto digit = [&](const int p, auto first) {
            constexpr bool FIRST = decltype(first)::value;
            if constexpr (FIRST) __syncthreads();  // the partner has read this wavefront's buffer
            else wave_sync();                      // the previous digit's transform is done with it
            // x: u_{j'} = d_{j'} + i d_{j'+1024} of the own quarters, ov: of the partner's
            c64 x[R], ov[R];
            if constexpr (FIRST) {
                // T = (X^abar - 1) * ACC_c + gadget offset, from this wavefront's copy of ACC_c
                int32_t* const fo = fbi + 512 * h + lane;
                int32_t* const fp = fbi + 512 * (1 - h) + lane;
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int a = 0; a < R; a++) {
                        fo[1024 * q + 64 * a] = (int32_t)acc[q * R + a];
                        fp[1024 * q + 64 * a] = (int32_t)pacc[q * R + a];
                    }
                wave_sync();
                // rotated positions x = base + compile-time offset, base = (own or partner's
                // quarter 0 position - abar) mod 2N, in bytes: (X^abar ACC_c)[j] = +-ACC_c[x mod N]
                // with the sign from bit log2(N) of x, which x < 3N keeps (x mod 2N has the same
                // bits below it), so the read needs one mask and the sign one bit-field extract
                const uint32_t bo4 = ((uint32_t)(512 * h + lane - abar) & (2 * N - 1)) * 4u;
                const uint32_t bp4 = ((uint32_t)(512 * (1 - h) + lane - abar) & (2 * N - 1)) * 4u;
                const char* fbc = (const char*)fbi;
                uint32_t to[2 * R], tp[2 * R];
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int a = 0; a < R; a++) {
                        const uint32_t xo = bo4 + 4u * (1024 * q + 64 * a), xp = bp4 + 4u * (1024 * q + 64 * a);
                        to[q * R + a] = *(const uint32_t*)(fbc + (xo & (4u * N - 1)));
                        tp[q * R + a] = *(const uint32_t*)(fbc + (xp & (4u * N - 1)));
                    }
#pragma unroll
                for (int q = 0; q < 2; q++)
#pragma unroll
                    for (int a = 0; a < R; a++) {
                        const uint32_t xo = bo4 + 4u * (1024 * q + 64 * a), xp = bp4 + 4u * (1024 * q + 64 * a);
                        // 0 or -1: the sign of the wrapped coefficient
                        const uint32_t mo = (uint32_t)__builtin_amdgcn_sbfe((int32_t)xo, 2 + LOG2N2 - 1, 1);
                        const uint32_t mp = (uint32_t)__builtin_amdgcn_sbfe((int32_t)xp, 2 + LOG2N2 - 1, 1);
                        to[q * R + a] = ((to[q * R + a] ^ mo) - mo) - acc[q * R + a] + offset;
                        tp[q * R + a] = ((tp[q * R + a] ^ mp) - mp) - pacc[q * R + a] + offset;
                    }
                wave_sync();  // the reads are done before the transform reuses the buffer
                // the truncating gadget decomposition, tgsw-functions.cpp:337-348
#pragma unroll
                for (int pp = 1; pp < L; pp++) {
                    const int sh = 32 - (pp + 1) * Bgbit;
#pragma unroll
                    for (int a = 0; a < R; a++) {
                        dig[pp - 1][0][a] = __builtin_amdgcn_ubfe(to[a], sh, Bgbit) |
                                            (__builtin_amdgcn_ubfe(to[R + a], sh, Bgbit) << 16);
                        dig[pp - 1][1][a] = __builtin_amdgcn_ubfe(tp[a], sh, Bgbit) |
                                            (__builtin_amdgcn_ubfe(tp[R + a], sh, Bgbit) << 16);
                    }
                }
                const int sh = 32 - Bgbit;
#pragma unroll
                for (int a = 0; a < R; a++) {
                    x[a] = c64{(double)((int32_t)__builtin_amdgcn_ubfe(to[a], sh, Bgbit) - dhalf),
                               (double)((int32_t)__builtin_amdgcn_ubfe(to[R + a], sh, Bgbit) - dhalf)};
                    ov[a] = c64{(double)((int32_t)__builtin_amdgcn_ubfe(tp[a], sh, Bgbit) - dhalf),
                                (double)((int32_t)__builtin_amdgcn_ubfe(tp[R + a], sh, Bgbit) - dhalf)};
                }
            } else {
#pragma unroll
                for (int a = 0; a < R; a++) {
                    const uint32_t vo = dig[0][0][a], vp = dig[0][1][a];
                    x[a] = c64{(double)((int32_t)(vo & 0xffffu) - dhalf), (double)((int32_t)(vo >> 16) - dhalf)};
                    ov[a] = c64{(double)((int32_t)(vp & 0xffffu) - dhalf), (double)((int32_t)(vp >> 16) - dhalf)};
                }
#pragma unroll
                for (int pp = 0; pp + 1 < L - 1; pp++)
#pragma unroll
                    for (int a = 0; a < R; a++) {
                        dig[pp][0][a] = dig[pp + 1][0][a];
                        dig[pp][1][a] = dig[pp + 1][1][a];
                    }
            }
            if constexpr (CT) {
                if (h == 0) {
                    P::pair_fwd<0>(x, ov);
                    FftCTPair::fwd1<0>(x, lane, fb);
                } else {
                    P::pair_fwd<1>(x, ov);
                    FftCTPair::fwd1<1>(x, lane, fb);
                }
                wave_sync();
                {
                    ct64 w[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) w[k] = ctw[k * 64];
                    G::fwd2(x, lane, fb, w);
                }
                wave_sync();
                G::fwd3(x, lane, fb);
            } else {
                if (h == 0) {
                    P::pair_fwd<0>(x, ov);
                    P::template fwd1<0, 64>(x, lane, fb, tw1);
                } else {
                    P::pair_fwd<1>(x, ov);
                    P::template fwd1<1, 64>(x, lane, fb, tw1);
                }
                wave_sync();
                F::fwd2(x, lane, fb, tw2);
                wave_sync();
                F::fwd3(x, lane, fb);
            }
            // key row (i, c L + p), half h: output component 0 at bk[s * 64], 1 at bk[1024 + s * 64]
            const c64* bk = bkp + ((size_t)i * KPL + c * L + p) * 2048 + h * 512 + lane;
            c64 kv[R];
#pragma unroll
            for (int s = 0; s < R; s++) kv[s] = bk[s * 64];
            wave_sync();
            if constexpr (CT) {
                ct64 w[4];
#pragma unroll
                for (int k = 0; k < 4; k++) w[k] = ctw[(4 + k) * 64];
                G::fwd4(x, lane, fb, w);
            } else {
                F::fwd4(x, lane, fb);
            }
#pragma unroll
            for (int s = 0; s < R; s++) {
                if constexpr (FIRST) p0[s] = cmul_first(x[s], kv[s]);
                else cmac(p0[s], x[s], kv[s]);
                kv[s] = bk[1024 + s * 64];
            }
#pragma unroll
            for (int s = 0; s < R; s++) {
                if constexpr (FIRST) p1[s] = cmul_first(x[s], kv[s]);
                else cmac(p1[s], x[s], kv[s]);
            }
        };
        digit(0, std::true_type{});
        if constexpr (L == 2) {
            digit(1, std::false_type{});
        } else {
#pragma unroll 1
            for (int p = 1; p < L; p++) digit(p, std::false_type{});
        }
        // swap partial sums with the other component's wavefront of the same half
        {
            wave_sync();  // the last transform is done with the buffer
            if (c) {
#pragma unroll
                for (int s = 0; s < R; s++) fb[s * 64 + lane] = p0[s];
            } else {
#pragma unroll
                for (int s = 0; s < R; s++) fb[s * 64 + lane] = p1[s];
            }
        }
        __syncthreads();  // partials written
        c64 x[R];
        {
#pragma unroll
            for (int k = 0; k < 2 * R; k++) acc_out[k * 64 + lane] = (int32_t)acc[k];
            const c64* other = s_fft[w ^ 2];
            if (c) {
#pragma unroll
                for (int s = 0; s < R; s++) x[s] = cadd(p1[s], other[s * 64 + lane]);
            } else {
#pragma unroll
                for (int s = 0; s < R; s++) x[s] = cadd(p0[s], other[s * 64 + lane]);
            }
        }
        __syncthreads();
        if constexpr (CT) {
            // this wavefront's inverse tables, read once per key row (opaque: kept out of
            // the row loop, where they would hold registers for the whole row)
            const c64* tq = twq_base + opaque_zero() + lane;
            ct64 w[4];
            G::inv1(x, lane, fb);
#pragma unroll
            for (int k = 0; k < 4; k++) w[k] = ((const ct64*)tq)[G::IB + 64 * k];
            wave_sync();
            G::inv2(x, lane, fb, w);
            wave_sync();
            G::inv3(x, lane, fb);
            c64 ut[R];
#pragma unroll
            for (int k = 0; k < 4; k++) w[k] = ((const ct64*)tq)[G::IA + 64 * k];
#pragma unroll
            for (int k = 0; k < R; k++) ut[k] = tq[G::UT + 64 * k];
            wave_sync();
            G::inv4(x, lane, fb, w, ut);
            wave_sync();
#pragma unroll
            for (int s = 0; s < R; s++) fb[s * 64 + lane] = x[s];
        } else {
            F::inv1(x, lane, fb);
            wave_sync();
            F::inv2(x, lane, fb, tw2);
            wave_sync();
            F::inv3(x, lane, fb);
            wave_sync();
            if (h == 0) P::template inv4<0, 64>(x, lane, fb, tw1);
            else P::template inv4<1, 64>(x, lane, fb, tw1);
            wave_sync();
#pragma unroll
            for (int s = 0; s < R; s++) fb[s * 64 + lane] = x[s];
        }
        __syncthreads();
        {
            // y_0 / 2 and y_1 / 2 -> u_{j'} = y0 + y1 (quarters 0, 2), u_{j'+512} =
            // (y0 - y1) e^{-i pi/4} (quarters 1, 3); ACC_c += result
            // (lwe-bootstrapping-functions-fft.cpp:67-76) for the own and the partner's quarters.
            // h is the same on the whole wavefront: one branch, no per-value selects.
            const c64* o = s_fft[w ^ 1];
            auto combine = [&](auto hc) {
                constexpr int H = decltype(hc)::value;
#pragma unroll
                for (int a = 0; a < R; a++) {
                    const c64 ya = x[a], yb = o[a * 64 + lane];
                    const c64 y0 = H == 0 ? ya : yb, y1 = H == 0 ? yb : ya;
                    const c64 u0 = cadd(y0, y1);
                    const c64 dd = csub(y0, y1);
                    const c64 u1 = c64{P::SQ * (dd.re + dd.im), P::SQ * (dd.im - dd.re)};
                    const c64 uo = H == 0 ? u0 : u1, up = H == 0 ? u1 : u0;
                    acc[a] += (uint32_t)round_mod32(uo.re);
                    acc[R + a] += (uint32_t)round_mod32(uo.im);
                    pacc[a] = (uint32_t)acc_in[a * 64 + lane] + (uint32_t)round_mod32(up.re);
                    pacc[R + a] = (uint32_t)acc_in[(R + a) * 64 + lane] + (uint32_t)round_mod32(up.im);
                }
            };
            if (h == 0) combine(std::integral_constant<int, 0>{});
            else combine(std::integral_constant<int, 1>{});
        }
    }
    // tLweExtractLweSampleIndex(.., 0, ..), libtfhe/lwe.cpp:41-56
    int32_t* e = A.ext + (size_t)job * (N + 1);
    if (c == 0) {
#pragma unroll
        for (int q = 0; q < 2; q++)
#pragma unroll
            for (int a = 0; a < R; a++) {
                const int j = 512 * (h + 2 * q) + 64 * a + lane;
                const uint32_t v = acc[q * R + a];
                e[j == 0 ? 0 : N - j] = (int32_t)(j == 0 ? v : 0u - v);
            }
    } else if (h == 0 && lane == 0) {
        e[N] = (int32_t)acc[0];
    }
}

// Blind rotation for N = 2048 with one bootstrap per pair of wavefronts (four
// per CU, twice the bootstraps resident of brp_kernel) on the fused-butterfly
// pair transform (FftCTPair; key and tables as for brp_kernel<L, true>).
// Wavefront h runs half h of every transform, for both accumulator components:
// per key row and input component it rotates and decomposes the whole of ACC_c
// (its own quarters h, h + 2 and its partner's, which the transform's radix-2
// step across the pair needs) from the accumulator in LDS, transforms each digit
// polynomial and multiply-accumulates the bins of both output components against
// its half of the key rows; then wavefront h takes over output component h (the
// pair swaps the halves of the other component once through LDS),
// inverse-transforms both of its halves and adds the result to ACC_h. Seven
// workgroup barriers per key row. A workgroup holds BRQ_BOOTSTRAPS bootstraps
// (job 2 x + b for pair b of workgroup x, each with its own LDS): the barriers
// keep the two together, so that they read each key row at about the same time
// and the second one's reads mostly hit the CU's L1 (8192-gate launch 141.5 ->
// 126.4 ms against one bootstrap per workgroup; 135.6 ms with four).
constexpr int BRQ_BOOTSTRAPS = 2;
template <int L>
__global__ __launch_bounds__(128 * BRQ_BOOTSTRAPS, 4 / BRQ_BOOTSTRAPS) void brq_kernel(BrArgs A, const c64* __restrict__ bkq,
                                                                  const c64* __restrict__ twq,
                                                                  const int32_t* __restrict__ mux_gate,
                                                                  const int32_t* __restrict__ n_jobs) {
    using G = FftCT;
    using P = FftPair;
    constexpr int R = 8, N = 2048, KPL = 2 * L, LOG2N2 = 12;
    constexpr int NMAX = BRP_NMAX;
    __shared__ c64 s_fft_b[BRQ_BOOTSTRAPS][2][G::LDS_C64];  // transform buffer of wavefront h; the halves handed over at a row's end
    __shared__ int32_t s_acc_b[BRQ_BOOTSTRAPS][2][N];       // ACC_0, ACC_1
    __shared__ int16_t s_bara_b[BRQ_BOOTSTRAPS][NMAX];
    __shared__ int32_t s_barb_b[BRQ_BOOTSTRAPS];
    const int pb = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 7);  // this bootstrap's pair
    auto& s_fft = s_fft_b[pb];
    auto& s_acc = s_acc_b[pb];
    auto& s_bara = s_bara_b[pb];
    int32_t& s_barb = s_barb_b[pb];

    const int tid = threadIdx.x & 127;
    const int lane = tid & 63;
    const int h = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int job = (int)blockIdx.x * BRQ_BOOTSTRAPS + pb;
    // uniform per pair; the pair's wavefronts end, and the workgroup's barriers then wait
    // for the other pair only
    if (job >= *n_jobs) return;
    const int n = A.n;
    c64* fb = s_fft[h];

    // gate linear form and modulus switch (wavefront 1, while wavefront 0 loads its tables)
    if (h == 1) {
        const int gate = job < A.count ? job : mux_gate[job - A.count];
        const int half = job < A.count ? 0 : 1;
        const int op = A.op[gate];
        int32_t k0, ca, cb, cc;
        gate_coeffs(op, half, k0, ca, cb, cc);
        const int32_t* a = A.pool + (size_t)A.in0[gate] * (n + 1);
        const int32_t* b = (op == OP_IDENTITY) ? a : A.pool + (size_t)A.in1[gate] * (n + 1);
        const int32_t* cp = (op == TFHE_AMD_MUX) ? A.pool + (size_t)A.in2[gate] * (n + 1) : a;
        for (int i = lane; i <= n; i += 64) {
            uint32_t v =
                (uint32_t)ca * (uint32_t)a[i] + (uint32_t)cb * (uint32_t)b[i] + (uint32_t)cc * (uint32_t)cp[i];
            if (i == n) v += (uint32_t)k0;
            // modSwitchFromTorus32(v, 2N), libtfhe/numeric-functions.cpp:59-65
            const uint32_t bar = (v + (1u << (31 - LOG2N2))) >> (32 - LOG2N2);
            if (i < n) s_bara[i] = (int16_t)bar;
            else s_barb = (int32_t)bar;
        }
    }
    // this wavefront's forward tables, in registers for the whole kernel
    const ct64* twl = (const ct64*)(twq + h * G::COUNT) + lane;
    ct64 fwb[4], fwc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        fwb[k] = twl[G::FB + 64 * k];
        fwc[k] = twl[G::FC + 64 * k];
    }
    const uint32_t offset = A.offset;
    const int Bgbit = A.Bgbit;
    const int32_t dhalf = 1 << (Bgbit - 1);
    __syncthreads();
    {
        // ACC = (0, X^{2N-barb} * [mu...mu]), lwe-bootstrapping-functions-fft.cpp:146-148
        const int barb = s_barb;
        for (int j = tid; j < N; j += 128) {
            s_acc[0][j] = 0;
            s_acc[1][j] = (((j + barb) & (2 * N - 1)) < N) ? A.mu : -A.mu;
        }
    }
    __syncthreads();

    // coefficient positions: own quarters h, h + 2 from jo = 512 h + lane, the partner's
    // 1 - h, 3 - h from jp = 512 (1 - h) + lane, plus 1024 q + 64 a (compile-time)
    const int jo = 512 * h + lane;
    for (int i = 0; i < n; i++) {
        const int abar = s_bara[i];
        c64 r0[R], r1[R];
#pragma unroll
        for (int s = 0; s < R; s++) r0[s] = r1[s] = c64{0.0, 0.0};
#pragma unroll 1
        for (int c = 0; c < 2; c++) {
            // T = (X^abar - 1) * ACC_c + gadget offset: slot k = 8 (2 g + q) + a
            // holds coefficient j = (g ? jp : jo) + 1024 q + 64 a. In bytes, x = 4 ((j0 - abar)
            // mod 2N) + 4 (1024 q + 64 a) for the slot's base j0: the rotated coefficient is
            // +-ACC_c[x / 4 mod N], its sign bit log2(N) of x / 4 (x / 4 < 3N)
            uint32_t t[4 * R];
            {
                // each wavefront forms T for its own quarters (slots k < 2R) and hands them to
                // its partner through its transform buffer, idle at this point, taking the
                // partner's (slots 2R + k: the partner's own quarters are this one's jp)
                const char* ac = (const char*)s_acc[c];
                const int z = opaque_zero();
                const uint32_t xo = ((uint32_t)(jo + z - abar) & (2 * N - 1)) * 4u;
                const int32_t* ao = s_acc[c] + jo;
                uint32_t* tb = (uint32_t*)fb;
#pragma unroll
                for (int k = 0; k < 2 * R; k++) {
                    const int q = k / R, a = k % R;
                    const uint32_t x = xo + 4u * (uint32_t)(1024 * q + 64 * a);
                    const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)x, 2 + LOG2N2 - 1, 1);
                    const uint32_t rv = *(const uint32_t*)(ac + (x & (4u * N - 1)));
                    const uint32_t av = (uint32_t)ao[1024 * q + 64 * a];
                    t[k] = ((rv ^ m) - m) - av + offset;
                    tb[k * 64 + lane] = t[k];
                }
                __syncthreads();  // both halves of T written
                const uint32_t* tp = (const uint32_t*)s_fft[h ^ 1];
#pragma unroll
                for (int k = 0; k < 2 * R; k++) t[2 * R + k] = tp[k * 64 + lane];
                __syncthreads();  // the partner has read this wavefront's half
            }
#pragma unroll 1
            for (int p = 0; p < L; p++) {
                // digit p (tgsw-functions.cpp:337-348): field p minus Bg/2. (With a signed field
                // extract of the flipped word instead, this kernel's compiled form converted the
                // field to double as unsigned.)
                // x: u_{j'} = d_{j'} + i d_{j'+1024} of the own quarters, ov: of the partner's
                const int shift = 32 - (p + 1) * Bgbit;
                auto dig = [&](uint32_t v) { return (double)((int32_t)__builtin_amdgcn_ubfe(v, shift, Bgbit) - dhalf); };
                c64 x[R], ov[R];
#pragma unroll
                for (int a = 0; a < R; a++) {
                    x[a] = c64{dig(t[a]), dig(t[R + a])};
                    ov[a] = c64{dig(t[2 * R + a]), dig(t[3 * R + a])};
                }
                wave_sync();  // the previous transform is done with the buffer
                if (h == 0) {
                    P::pair_fwd<0>(x, ov);
                    FftCTPair::fwd1<0>(x, lane, fb);
                } else {
                    P::pair_fwd<1>(x, ov);
                    FftCTPair::fwd1<1>(x, lane, fb);
                }
                wave_sync();
                G::fwd2(x, lane, fb, fwb);
                wave_sync();
                G::fwd3(x, lane, fb);
                // key row (i, c L + p), half h: output component 0 at q[s * 64], 1 at q[1024 + s * 64]
                const c64* q = bkq + ((size_t)i * KPL + c * L + p) * 2048 + h * 512 + lane;
                c64 kv[R];
#pragma unroll
                for (int s = 0; s < R; s++) kv[s] = q[s * 64];
                wave_sync();
                G::fwd4(x, lane, fb, fwc);
#pragma unroll
                for (int s = 0; s < R; s++) {
                    cmac(r0[s], x[s], kv[s]);
                    kv[s] = q[1024 + s * 64];
                }
#pragma unroll
                for (int s = 0; s < R; s++) cmac(r1[s], x[s], kv[s]);
            }
        }
        // wavefront h keeps output component h: it hands its half of the other component
        // to the partner and takes the partner's half of its own
        __syncthreads();  // both wavefronts are done with their transform buffers
        {
            c64* pb = s_fft[h ^ 1];
            if (h == 0) {
#pragma unroll
                for (int s = 0; s < R; s++) pb[s * 64 + lane] = r1[s];
            } else {
#pragma unroll
                for (int s = 0; s < R; s++) pb[s * 64 + lane] = r0[s];
            }
        }
        __syncthreads();  // the halves handed over
        {
            // component c = h: its bins of half h (own, in registers) and of half 1 - h (from
            // the partner, in this wavefront's buffer); inverse transform of each half with
            // that half's untwist (tables read once per row at an opaque offset, so they are not
            // hoisted out of the row loop), then u_{j'} = y0 + y1 (quarters 0, 2) and
            // u_{j'+512} = (y0 - y1) e^{-i pi/4} (quarters 1, 3), ACC_c += result
            // (lwe-bootstrapping-functions-fft.cpp:67-76)
            c64 y0[R], y1[R];
            if (h == 0) {
#pragma unroll
                for (int s = 0; s < R; s++) {
                    y0[s] = r0[s];
                    y1[s] = fb[s * 64 + lane];
                }
            } else {
#pragma unroll
                for (int s = 0; s < R; s++) {
                    y0[s] = fb[s * 64 + lane];
                    y1[s] = r1[s];
                }
            }
            const int z = opaque_zero();
            const ct64* tb = (const ct64*)twq + z + lane;  // half 0's tables; half 1's at + COUNT
            ct64 wb[4], wa[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                wb[k] = tb[G::IB + 64 * k];
                wa[k] = tb[G::IA + 64 * k];
            }
#pragma unroll
            for (int hh = 0; hh < 2; hh++) {
                c64* x = hh ? y1 : y0;
                c64 ut[R];
#pragma unroll
                for (int k = 0; k < R; k++) ut[k] = twq[hh * G::COUNT + G::UT + 64 * k + z + lane];
                wave_sync();
                G::inv1(x, lane, fb);
                wave_sync();
                G::inv2(x, lane, fb, wb);
                wave_sync();
                G::inv3(x, lane, fb);
                wave_sync();
                G::inv4(x, lane, fb, wa, ut);
            }
            int32_t* sa = s_acc[h] + lane;
#pragma unroll
            for (int a = 0; a < R; a++) {
                const c64 lo = cadd(y0[a], y1[a]);
                const c64 dd = csub(y0[a], y1[a]);
                const c64 hi = c64{P::SQ * (dd.re + dd.im), P::SQ * (dd.im - dd.re)};
                sa[64 * a] = (int32_t)((uint32_t)sa[64 * a] + (uint32_t)round_mod32(lo.re));
                sa[1024 + 64 * a] = (int32_t)((uint32_t)sa[1024 + 64 * a] + (uint32_t)round_mod32(lo.im));
                sa[512 + 64 * a] = (int32_t)((uint32_t)sa[512 + 64 * a] + (uint32_t)round_mod32(hi.re));
                sa[1536 + 64 * a] = (int32_t)((uint32_t)sa[1536 + 64 * a] + (uint32_t)round_mod32(hi.im));
            }
        }
        __syncthreads();  // ACC updated for the next row
    }
    // tLweExtractLweSampleIndex(.., 0, ..), libtfhe/lwe.cpp:41-56
    int32_t* e = A.ext + (size_t)job * (N + 1);
    for (int j = tid; j < N; j += 128) e[j == 0 ? 0 : N - j] = j == 0 ? s_acc[0][0] : (int32_t)(0u - (uint32_t)s_acc[0][j]);
    if (tid == 0) e[N] = s_acc[1][0];
}

// Pair-of-wavefronts transform of every bootstrapping-key polynomial (N = 2048):
// one 128-thread workgroup per polynomial, output [poly][h][slot][lane].
__global__ __launch_bounds__(128) void bk_convert_pair_kernel(const int32_t* __restrict__ bk, c64* __restrict__ bkp,
                                                              const c64* __restrict__ twp) {
    using F = Fft<512>;
    using P = FftPair;
    constexpr int R = 8;
    __shared__ c64 s_fft[2][F::LDS_C64];
    const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * 2048;
    c64* fb = s_fft[h];
    c64 x[R], tw1[R];
#pragma unroll
    for (int a = 0; a < R; a++) {
        x[a] = c64{(double)p[512 * h + 64 * a + lane], (double)p[512 * h + 64 * a + lane + 1024]};
        fb[a * 64 + lane] = x[a];
    }
    __syncthreads();
    c64 ov[R];
#pragma unroll
    for (int a = 0; a < R; a++) ov[a] = s_fft[h ^ 1][a * 64 + lane];
    if (h == 0) P::pair_fwd<0>(x, ov);
    else P::pair_fwd<1>(x, ov);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; k++) tw1[k] = twp[(h * 64 + lane) * 16 + k];
    const c64 tw2 = twp[(h * 64 + lane) * 16 + 9];
    if (h == 0) P::fwd1<0>(x, lane, fb, tw1);
    else P::fwd1<1>(x, lane, fb, tw1);
    wave_sync();
    F::fwd2(x, lane, fb, tw2);
    wave_sync();
    F::fwd3(x, lane, fb);
    wave_sync();
    F::fwd4(x, lane, fb);
    c64* out = bkp + poly * 1024 + h * 512;
#pragma unroll
    for (int s = 0; s < R; s++) out[s * 64 + lane] = x[s];
}

// The same on the pair transform with fused butterflies (FftCTPair): output
// [poly][h][slot][lane] in that transform's order, twq its [h] tables.
__global__ __launch_bounds__(128) void bk_convert_pair_ct_kernel(const int32_t* __restrict__ bk, c64* __restrict__ bkq,
                                                                 const c64* __restrict__ twq) {
    using G = FftCT;
    using P = FftPair;
    constexpr int R = 8;
    __shared__ c64 s_fft[2][G::LDS_C64];
    const int lane = threadIdx.x & 63, h = threadIdx.x >> 6;
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * 2048;
    c64* fb = s_fft[h];
    const ct64* tw = (const ct64*)(twq + h * G::COUNT) + lane;
    c64 x[R];
#pragma unroll
    for (int a = 0; a < R; a++) {
        x[a] = c64{(double)p[512 * h + 64 * a + lane], (double)p[512 * h + 64 * a + lane + 1024]};
        fb[a * 64 + lane] = x[a];
    }
    __syncthreads();
    c64 ov[R];
#pragma unroll
    for (int a = 0; a < R; a++) ov[a] = s_fft[h ^ 1][a * 64 + lane];
    if (h == 0) P::pair_fwd<0>(x, ov);
    else P::pair_fwd<1>(x, ov);
    __syncthreads();
    ct64 wb[4], wc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        wb[k] = tw[G::FB + 64 * k];
        wc[k] = tw[G::FC + 64 * k];
    }
    if (h == 0) FftCTPair::fwd1<0>(x, lane, fb);
    else FftCTPair::fwd1<1>(x, lane, fb);
    wave_sync();
    G::fwd2(x, lane, fb, wb);
    wave_sync();
    G::fwd3(x, lane, fb);
    wave_sync();
    G::fwd4(x, lane, fb, wc);
    c64* out = bkq + poly * 1024 + h * 512;
#pragma unroll
    for (int s = 0; s < R; s++) out[s * 64 + lane] = x[s];
}

template <int M>
__global__ __launch_bounds__(64) void bk_convert_kernel(const int32_t* __restrict__ bk, c64* __restrict__ bkf,
                                                        const c64* __restrict__ tw) {
    using F = Fft<M>;
    constexpr int R = F::R;
    __shared__ c64 s_fft[F::LDS_C64];
    const int lane = threadIdx.x;
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * 2 * M;
    c64 x[R], tw1[R];
#pragma unroll
    for (int k = 0; k < R; k++) tw1[k] = tw[lane * (R + 8) + k];
    const c64 tw2 = tw[lane * (R + 8) + R + 1];
#pragma unroll
    for (int s = 0; s < R; s++) x[s] = c64{(double)p[64 * s + lane], (double)p[64 * s + lane + M]};
    F::fwd1(x, lane, s_fft, tw1);
    __syncthreads();
    F::fwd2(x, lane, s_fft, tw2);
    __syncthreads();
    F::fwd3(x, lane, s_fft);
    __syncthreads();
    F::fwd4(x, lane, s_fft);
#pragma unroll
    for (int s = 0; s < R; s++) bkf[poly * M + s * 64 + lane] = x[s];
}

// Bootstrapping-key polynomials into the fused-butterfly transform's order
// (N = 1024): one wavefront per polynomial, output [poly][slot][lane].
__global__ __launch_bounds__(64) void bk_convert_ct_kernel(const int32_t* __restrict__ bk, c64* __restrict__ bkc,
                                                           const c64* __restrict__ twc) {
    using G = FftCT;
    __shared__ c64 s_fft[G::LDS_C64];
    const int lane = threadIdx.x;
    const size_t poly = blockIdx.x;
    const int32_t* p = bk + poly * 1024;
    const ct64* tw = (const ct64*)twc + lane;
    c64 x[8];
    ct64 wb[4], wc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        wb[k] = tw[G::FB + 64 * k];
        wc[k] = tw[G::FC + 64 * k];
    }
#pragma unroll
    for (int s = 0; s < 8; s++) x[s] = c64{(double)p[64 * s + lane], (double)p[64 * s + lane + 512]};
    G::fwd1(x, lane, s_fft);
    __syncthreads();
    G::fwd2(x, lane, s_fft, wb);
    __syncthreads();
    G::fwd3(x, lane, s_fft);
    __syncthreads();
    G::fwd4(x, lane, s_fft, wc);
#pragma unroll
    for (int s = 0; s < 8; s++) bkc[poly * 512 + s * 64 + lane] = x[s];
}

// ext: [2*count][nin+1]; op may be null (no MUX). Writes abar[i][cpad] = a_i + prec_offset
// (masks transposed so the keyswitch reads one word per gate per i) and binit[g] = b.
// nz (if not null): nz[g] = the number of nonzero keyswitch digits of gate g (t digits of
// basebit bits per coefficient), i.e. of key rows the keyswitch adds
// (libtfhe/lwe-keyswitch-functions.cpp:114-129: one lweSubTo per nonzero digit).
__global__ __launch_bounds__(256) void ks_prep_kernel(const int32_t* __restrict__ ext,
                                                      const int32_t* __restrict__ second, int nin, int32_t mu,
                                                      uint32_t prec_offset, uint32_t* __restrict__ abar, int cpad,
                                                      int32_t* __restrict__ binit, int t, int basebit,
                                                      int32_t* __restrict__ nz) {
    __shared__ int s_nz[4];
    const int g = blockIdx.x;
    const int j2 = second ? second[g] : -1;
    const bool mux = j2 >= 0;
    const int32_t* e0 = ext + (size_t)g * (nin + 1);
    const int32_t* e1 = ext + (size_t)(mux ? j2 : g) * (nin + 1);
    const uint32_t mask = (1u << basebit) - 1;
    int cnt = 0;
    for (int i = threadIdx.x; i <= nin; i += blockDim.x) {
        uint32_t v = (uint32_t)e0[i];
        // libtfhe/boot-gates.cpp:358-361: (0, 1/8) + u1 + u2
        if (mux) v += (uint32_t)e1[i] + (i == nin ? (uint32_t)mu : 0u);
        if (i < nin) {
            const uint32_t ab = v + prec_offset;
            abar[(size_t)i * cpad + g] = ab;
            if (nz)
                for (int j = 0; j < t; j++) cnt += ((ab >> (32 - (j + 1) * basebit)) & mask) != 0;
        } else {
            binit[g] = (int32_t)v;
        }
    }
    if (nz) {  // block sum: wavefront sums, then the four partials
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if ((threadIdx.x & 63) == 0) s_nz[threadIdx.x >> 6] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) nz[g] = s_nz[0] + s_nz[1] + s_nz[2] + s_nz[3];
    }
}

__global__ __launch_bounds__(256) void sum_halves_kernel(const int32_t* ext, const int32_t* second, int nin,
                                                         int32_t mu, int32_t* out) {
    const int g = blockIdx.x;
    const int j2 = second[g];
    const bool mux = j2 >= 0;
    const int32_t* e0 = ext + (size_t)g * (nin + 1);
    const int32_t* e1 = ext + (size_t)(mux ? j2 : g) * (nin + 1);
    int32_t* o = out + (size_t)g * (nin + 1);
    for (int i = threadIdx.x; i <= nin; i += blockDim.x) {
        uint32_t v = (uint32_t)e0[i];
        if (mux) v += (uint32_t)e1[i] + (i == nin ? (uint32_t)mu : 0u);
        o[i] = (int32_t)v;
    }
}

// Plain device-to-device copy (ciphertext pools); 16-byte vector accesses.
__global__ __launch_bounds__(256) void copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Keyswitch with lanes over gates: a block covers WAVES*64 gates, one chunk of
// 64 output coefficients and a range of input coefficients i. For every
// (i, j) the block stages the three non-trivial key rows of its chunk in LDS
// (next to a zero row for digit 0); each lane then subtracts row[digit] of its
// own gate with broadcast 16-byte reads, so the key is read once per block
// and each gate-coefficient costs one integer subtraction. Partial sums go to
// part[split][gate][npad]; ks_reduce_kernel adds the splits.
template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void ks_kernel(const int32_t* __restrict__ ksk,
                                                        const uint32_t* __restrict__ abar, int cpad, int nin, int t,
                                                        int npad, int i_per_split, int32_t* __restrict__ part) {
    __shared__ int4 tab[2][4][16];
    const int tid = threadIdx.x;
    const int g = blockIdx.y * (WAVES * 64) + tid;  // < cpad by construction
    const int chunk = blockIdx.x;
    const int i0 = blockIdx.z * i_per_split;
    const int i1 = min(nin, i0 + i_per_split);
    if (tid < 32) tab[tid >> 4][0][tid & 15] = make_int4(0, 0, 0, 0);
    int4 acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = make_int4(0, 0, 0, 0);
    int step = 0;
    for (int i = i0; i < i1; i++) {
        const uint32_t a = abar[(size_t)i * cpad + g];
        for (int j = 0; j < t; j++, step++) {
            const int buf = step & 1;
            if (tid < 48) {
                const int h = tid >> 4, q = tid & 15;
                tab[buf][1 + h][q] =
                    *(const int4*)(ksk + ((size_t)(i * t + j) * 3 + h) * npad + chunk * 64 + q * 4);
            }
            __syncthreads();
            const int d = (a >> (30 - 2 * j)) & 3;
            const int4* row = tab[buf][d];
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int4 v = row[q];
                acc[q].x = (int32_t)((uint32_t)acc[q].x - (uint32_t)v.x);
                acc[q].y = (int32_t)((uint32_t)acc[q].y - (uint32_t)v.y);
                acc[q].z = (int32_t)((uint32_t)acc[q].z - (uint32_t)v.z);
                acc[q].w = (int32_t)((uint32_t)acc[q].w - (uint32_t)v.w);
            }
        }
    }
    int4* dst = (int4*)(part + ((size_t)blockIdx.z * cpad + g) * npad + chunk * 64);
#pragma unroll
    for (int q = 0; q < 16; q++) dst[q] = acc[q];
}

// out[gate] = (0, ..., 0, b) + sum over splits of the partial keyswitch sums
__global__ __launch_bounds__(256) void ks_reduce_kernel(const int32_t* __restrict__ part, int splits, int cpad,
                                                        int npad, const int32_t* __restrict__ binit, int n,
                                                        int32_t* __restrict__ out, const int32_t* __restrict__ out_idx) {
    const int g = blockIdx.x;
    int32_t* o = out + (size_t)(out_idx ? out_idx[g] : g) * (n + 1);
    for (int c = threadIdx.x; c <= n; c += blockDim.x) {
        uint32_t v = (c == n) ? (uint32_t)binit[g] : 0u;
        for (int sp = 0; sp < splits; sp++) v += (uint32_t)part[((size_t)sp * cpad + g) * npad + c];
        o[c] = (int32_t)v;
    }
}

// ------------------------------------------------ keyswitch on matrix cores --
//
// The keyswitch out = (0, ..., 0, b) - sum_{i<kN, j<t} K[i][j][d_ij] is the
// product OUT = B - D x K of the one-hot digit matrix D[g][(i,j,h)] = [d_gij = h]
// with the key rows K[(i,j,h)][c] (h = 0 rows are zero). The key is split into
// four signed byte planes, K = sum_b 2^{8b} s_b (mod 2^32), s_b in [-128, 127]
// (balanced base-256 digits, the carry out of the top byte dropped), so every
// plane product runs on the int8 matrix cores and is exact in int32
// (|sum| <= kN t 128 < 2^31); the planes are recombined mod 2^32.
// Layout for v_mfma_i32_16x16x64_i8 (lane l: matrix row l%16, k = 16(l/16) + 0..15):
// the 64 k of one instruction are two input coefficients i, each with t = 8
// digits x 4 values h, so a lane's 16 A-bytes are the one-hot words of four
// digits and its 16 B-bytes are kp[b][i][c][16 (l/16 & 1) + 0..15].

typedef int v4i __attribute__((ext_vector_type(4)));

// gate tiles of 16 per wavefront in the matrix-core keyswitch: 8 above 64 gates (each key
// byte plane is read by half as many gate groups as with 4: 4096 gates 0.42 -> 0.37 ms),
// 4 up to 64 gates (no more padded gate rows than one group of 64)
constexpr int KS_MMA_GT_SMALL = 4, KS_MMA_GT_LARGE = 8;

// 0: matrix cores when the key has byte planes; 1: always the lanes-over-gates kernels
std::atomic<int> g_ks_path{0};

// one thread per (i, c, j): the 4 values h of key row (i, j) at column c, for the 4 planes
__global__ __launch_bounds__(256) void ks_planes_kernel(const int32_t* __restrict__ ksk, int nin, int npad,
                                                        int8_t* __restrict__ kp) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)nin * npad * 8) return;
    const int j = (int)(idx & 7);
    const int c = (int)((idx >> 3) % npad);
    const int i = (int)(idx / ((size_t)8 * npad));
    const size_t plane = (size_t)nin * npad * 32;
    int8_t* dst = kp + ((size_t)i * npad + c) * 32 + j * 4;
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const uint32_t v = h ? (uint32_t)ksk[((size_t)(i * 8 + j) * 3 + (h - 1)) * npad + c] : 0u;
        int carry = 0;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            int byte = (int)((v >> (8 * b)) & 255u) + carry;
            carry = byte >= 128;
            if (carry) byte -= 256;
            dst[b * plane + h] = (int8_t)byte;
        }
    }
}

// A workgroup = 4 wavefronts; wavefront w owns 16 output columns (group
// blockIdx.x*4 + w) of GT*16 gates (GT tiles of 16, from blockIdx.y*GT*16)
// over input coefficients [blockIdx.z * ips, +ips). abar is the digit source
// a_i + prec_offset, [kN][cpad] with cpad a multiple of 128. With one slice
// (gridDim.z == 1) the output rows are written; with several, each slice
// subtracts its partial sum from rows already set to (0, ..., 0, b) — integer
// atomics, exact mod 2^32 in any order.
template <int GT>
__global__ __launch_bounds__(256, 2) void ks_mma_kernel(const int8_t* __restrict__ kp, const uint32_t* __restrict__ abar,
                                                     int cpad, int nin, int npad, int count,
                                                     const int32_t* __restrict__ binit, int n,
                                                     int32_t* __restrict__ out, const int32_t* __restrict__ out_idx,
                                                     int ips) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int q = lane >> 4, r16 = lane & 15;
    const int col = (blockIdx.x * 4 + w) * 16 + r16;
    const int g0 = blockIdx.y * GT * 16;
    // one-hot words of the four 2-bit digits in a byte: s_onehot[x][jj] = 1 << 8 d_jj
    __shared__ v4i s_onehot[256];
    {
        const uint32_t x = threadIdx.x;
        v4i o;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) o[jj] = 1 << (((x >> (6 - 2 * jj)) & 3u) << 3);
        s_onehot[x] = o;
    }
    __syncthreads();
    v4i acc[GT][4];
#pragma unroll
    for (int gt = 0; gt < GT; gt++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[gt][b] = v4i{0, 0, 0, 0};
    const size_t plane = (size_t)nin * npad * 32;
    const int8_t* kb = kp + (size_t)col * 32 + 16 * (q & 1);
    const uint32_t* ab = abar + g0 + r16;
    const int sh = (q & 1) ? 16 : 24;  // digits j = 4(q&1) .. 4(q&1)+3 sit in this byte of a_i
    // operands of the next pair of input coefficients are loaded while the
    // current pair's products run
    v4i bv[4];
    uint32_t av8[GT];
    auto load = [&](int i0) {
        const int i = i0 + (q >> 1);
#pragma unroll
        for (int b = 0; b < 4; b++) bv[b] = *(const v4i*)(kb + b * plane + (size_t)i * npad * 32);
#pragma unroll
        for (int gt = 0; gt < GT; gt++) av8[gt] = ab[(size_t)i * cpad + gt * 16];
    };
    const int ibeg = blockIdx.z * ips, iend = min(nin, ibeg + ips);
    load(ibeg);
    for (int i0 = ibeg; i0 < iend; i0 += 2) {
        v4i b_cur[4];
        uint32_t a_cur[GT];
#pragma unroll
        for (int b = 0; b < 4; b++) b_cur[b] = bv[b];
#pragma unroll
        for (int gt = 0; gt < GT; gt++) a_cur[gt] = av8[gt];
        if (i0 + 2 < iend) load(i0 + 2);
#pragma unroll
        for (int gt = 0; gt < GT; gt++) {
            const v4i av = s_onehot[(a_cur[gt] >> sh) & 255u];
#pragma unroll
            for (int b = 0; b < 4; b++)
                acc[gt][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, b_cur[b], acc[gt][b], 0, 0, 0);
        }
    }
    if (col > n) return;
#pragma unroll
    for (int gt = 0; gt < GT; gt++)
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int gate = g0 + gt * 16 + 4 * q + rr;
            if (gate >= count) continue;
            const uint32_t v = (uint32_t)acc[gt][0][rr] + ((uint32_t)acc[gt][1][rr] << 8) +
                               ((uint32_t)acc[gt][2][rr] << 16) + ((uint32_t)acc[gt][3][rr] << 24);
            int32_t* dst = out + (size_t)(out_idx ? out_idx[gate] : gate) * (n + 1) + col;
            if (gridDim.z == 1) *dst = (int32_t)((col == n ? (uint32_t)binit[gate] : 0u) - v);
            else atomicSub((unsigned int*)dst, v);
        }
}

// Output rows (0, ..., 0, b) for the sliced matrix-core keyswitch.
__global__ __launch_bounds__(256) void ks_out_init_kernel(const int32_t* __restrict__ binit, int n,
                                                          int32_t* __restrict__ out, const int32_t* __restrict__ out_idx) {
    const int gate = blockIdx.x;
    int32_t* row = out + (size_t)(out_idx ? out_idx[gate] : gate) * (n + 1);
    for (int c = threadIdx.x; c <= n; c += 256) row[c] = c == n ? binit[gate] : 0;
}

// ---------------------------------------------------------------- runtime --

// Workspace of the entry point that allocates nothing
// (tfhe_amd_gates_device_noalloc): a static device array, one copy per GPU,
// held until the process ends. Host code never passes `scratch` itself to a
// kernel (in host code it is only a stand-in for the symbol): the address of
// the current device's copy comes from hipGetSymbolAddress at set-up.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)32 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

// A device workspace, the stream that used it last and an event recorded after
// that use.
struct Scratch {
    void* d = nullptr;
    size_t cap = 0;
    hipStream_t last = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
};

// caller streams that get a workspace of their own; further streams share the
// library stream's (ordered, not overlapped)
constexpr int MAX_STREAM_WORKSPACES = 8;

struct Ctx {
    std::mutex mu;
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    Scratch scratch;                                        // the library stream's workspace
    std::vector<std::pair<hipStream_t, Scratch>> stream_ws;  // caller streams' own workspaces
    void* h_pinned = nullptr;
    size_t h_cap = 0;
    int br_mode = 0;
    int brl_max = 256;        // automatic mode: batches up to this many gates use the latency kernel (4l wavefronts per bootstrap)
    int br_small_max = 1024;  // automatic mode: batches up to this many gates use two wavefronts per bootstrap
    bool crossovers_set = false;  // set_launch_config gave crossovers: they apply instead of the per-size choice
    char* noalloc_scr = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress at set-up)
    int round_jobs[5][2] = {{-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}, {-1, -1}};  // bootstraps in one whole
                            // round of br2_kernel<l, 4> [0][l == 3] and br4w_kernel<l, W> [W - 4][l == 3],
                            // once queried
    int cus = -1;           // CUs of the device, once queried
    // optional per-launch timing with HIP events on the launch stream
    bool prof = false;
    struct Ev { hipEvent_t e[3]; };
    std::vector<Ev> ev;
};

Ctx& ctx() {
    static Ctx c;
    return c;
}

int whole_round(Ctx& C, int w, int l);

int ensure_ready(Ctx& C) {
    if (C.ready) return 0;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (ndev < 1) return set_error("no GPU device visible");
    HIP_TRY(hipGetDevice(&C.device));
    HIP_TRY(hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking));
    // the device address of this GPU's copy of the static workspace (the device is current)
    HIP_TRY(hipGetSymbolAddress((void**)&C.noalloc_scr, HIP_SYMBOL(scratch)));
    if (!C.noalloc_scr) return set_error("hipGetSymbolAddress(scratch) returned a null address");
    // the round sizes of the several-bootstrap workgroups (occupancy queries), here rather
    // than at a first launch that may be made under stream capture
    for (int l = 2; l <= 3; l++)
        for (int w = 4; w <= BR4W_MAX; w++) whole_round(C, w, l);
    C.ready = true;
    return 0;
}

// The workspace for work about to be queued on stream s: the library stream's
// own, or the caller stream's (the first MAX_STREAM_WORKSPACES distinct streams;
// later ones share the library stream's). It is grown to at least `bytes` (a
// regrowth waits for the whole device). A workspace whose last use was on
// another stream (the shared one, or a stream handle reused after its stream
// was destroyed) is first waited for on the device, so calls never overlap in
// one workspace, while calls on streams with workspaces of their own overlap
// freely. The host does not block; no wait goes into a stream under graph
// capture.
int ensure_scratch(Ctx& C, size_t bytes, hipStream_t s, Scratch** out) {
    Scratch* w = &C.scratch;
    if (s != C.stream) {
        for (auto& e : C.stream_ws)
            if (e.first == s) w = &e.second;
        if (w == &C.scratch && (int)C.stream_ws.size() < MAX_STREAM_WORKSPACES) {
            C.stream_ws.reserve(MAX_STREAM_WORKSPACES);  // workspaces never move
            C.stream_ws.emplace_back(s, Scratch{});
            w = &C.stream_ws.back().second;
        }
    }
    if (w->cap < bytes) {
        if (w->d) {
            // work queued on any stream may still use the old buffer
            HIP_TRY(hipDeviceSynchronize());
            HIP_TRY(hipFree(w->d));
            w->d = nullptr;
            w->cap = 0;
            w->busy = false;
        }
        size_t cap = bytes + bytes / 2 + (1 << 20);
        HIP_TRY(hipMalloc(&w->d, cap));
        w->cap = cap;
    }
    if (w->busy && w->last != s) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(s, &cs));
        if (cs == hipStreamCaptureStatusNone) HIP_TRY(hipStreamWaitEvent(s, w->ev, 0));
    }
    *out = w;
    return 0;
}

// Records that work queued on stream s uses workspace w (after its launches).
int scratch_done(Scratch* w, hipStream_t s) {
    if (!w->ev) HIP_TRY(hipEventCreateWithFlags(&w->ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(w->ev, s));
    w->last = s;
    w->busy = true;
    return 0;
}

int ensure_pinned(Ctx& C, size_t bytes) {
    if (C.h_cap >= bytes) return 0;
    if (C.h_pinned) {
        HIP_TRY(hipStreamSynchronize(C.stream));
        HIP_TRY(hipHostFree(C.h_pinned));
        C.h_pinned = nullptr;
        C.h_cap = 0;
    }
    size_t cap = bytes + bytes / 2 + (1 << 20);
    HIP_TRY(hipHostMalloc(&C.h_pinned, cap, hipHostMallocDefault));
    C.h_cap = cap;
    return 0;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

uint32_t gadget_offset(const HostParams& P) {
    uint32_t o = 0;
    for (int i = 0; i < P.l; i++) o += 1u << (32 - (i + 1) * P.Bgbit);
    return o * (1u << (P.Bgbit - 1));
}

int check_supported(const HostParams& P) {
    if (P.k != 1) return set_error("only k = 1 is supported (got k = %d)", P.k);
    if (P.N != 1024 && P.N != 2048) return set_error("only N = 1024 or 2048 is supported (got %d)", P.N);
    if (!((P.N == 1024 && P.l == 3) || (P.N == 2048 && P.l == 2) || (P.N == 1024 && P.l == 2) || (P.N == 2048 && P.l == 3)))
        return set_error("unsupported gadget length l = %d", P.l);
    if (P.n > 1024 || P.n < 1) return set_error("LWE dimension n = %d out of range [1, 1024]", P.n);
    if (P.ks_basebit != 2) return set_error("keyswitch basebit must be 2 (got %d)", P.ks_basebit);
    if (P.ks_t * P.ks_basebit > 31) return set_error("keyswitch t*basebit must be < 32");
    if (P.Bgbit * P.l > 32 || P.Bgbit < 1) return set_error("bad gadget base");
    return 0;
}

template <int M>
void launch_br_t(const BrArgs& A, int blocks, int l, hipStream_t s) {
    if (l == 3) hipLaunchKernelGGL((br_kernel<M, 3>), dim3(blocks), dim3(64), 0, s, A);
    else hipLaunchKernelGGL((br_kernel<M, 2>), dim3(blocks), dim3(64), 0, s, A);
}

void launch_br(const tfhe_amd_device_key* dk, const BrArgs& A, int blocks, hipStream_t s) {
    if (dk->M == 512) launch_br_t<512>(A, blocks, dk->P.l, s);
    else launch_br_t<1024>(A, blocks, dk->P.l, s);
}

BrArgs base_args(const tfhe_amd_device_key* dk) {
    BrArgs A;
    memset(&A, 0, sizeof A);
    A.bkf = dk->bkf;
    A.tw = dk->tw;
    A.n = dk->P.n;
    A.Bgbit = dk->P.Bgbit;
    A.offset = gadget_offset(dk->P);
    A.mu = 1 << 29;
    return A;
}

struct KsPlan {
    int waves, tiles, chunks, splits, cpad;
};

KsPlan ks_plan(const HostParams& P, int count) {
    KsPlan k;
    k.waves = count >= 256 ? 4 : 1;
    const int nt = 64 * k.waves;
    k.tiles = (count + nt - 1) / nt;
    k.cpad = k.tiles * nt;
    k.chunks = (P.n + 1 + 63) / 64;
    const int blocks = k.tiles * k.chunks;
    k.splits = 1;
    // enough wavefronts to fill the device (about 8 per CU on 256 CUs)
    while (blocks * k.splits * k.waves < 2048 && k.splits < 64) k.splits *= 2;
    return k;
}

// Keyswitch scratch of one path (mma: the matrix-core kernels; else the
// lanes-over-gates ones), or of either (the larger) when the path may change.
size_t ks_scratch_bytes_path(const HostParams& P, int count, bool mma) {
    if (mma) {
        const size_t cpad_m = (size_t)(count + 127) / 128 * 128;
        return align256(sizeof(uint32_t) * (size_t)P.k * P.N * cpad_m) + align256(sizeof(int32_t) * cpad_m);
    }
    const KsPlan k = ks_plan(P, count);
    const int npad = ((P.n + 1 + 63) / 64) * 64;
    return align256(sizeof(uint32_t) * (size_t)P.k * P.N * k.cpad) + align256(sizeof(int32_t) * (size_t)k.cpad) +
           align256(sizeof(int32_t) * (size_t)k.splits * k.cpad * npad);
}

size_t ks_scratch_bytes(const HostParams& P, int count) {
    return std::max(ks_scratch_bytes_path(P, count, true), ks_scratch_bytes_path(P, count, false));
}

// The keyswitch path a call takes now: 0 = matrix cores, 1 = lanes over gates.
int ks_path_now(const tfhe_amd_device_key* dk) {
    return (dk->kp && g_ks_path.load(std::memory_order_relaxed) == 0) ? 0 : 1;
}

// Keyswitch of ext rows (sum of halves for MUX) into out rows, using the
// scratch region starting at `scr` (ks_scratch_bytes(count) bytes).
// path: -1 = the current setting (ks_path_now), else the path to take (0 only
// for keys with byte planes).
// nz (if not null): per gate, the number of nonzero digits (key rows added).
int run_keyswitch(const tfhe_amd_device_key* dk, const int32_t* ext, const int32_t* second, int count, int32_t* out,
                  const int32_t* out_idx, char* scr, hipStream_t s, int path = -1, int32_t* nz = nullptr) {
    const HostParams& P = dk->P;
    const int nin = P.k * P.N;
    const uint32_t prec_offset0 = 1u << (32 - (1 + P.ks_basebit * P.ks_t));
    if (path < 0) path = ks_path_now(dk);
    if (path == 0 && dk->kp) {
        // matrix-core path: prep (abar [kN][cpad], b) then one product kernel
        const int cpad = (count + 127) / 128 * 128;
        uint32_t* abar = (uint32_t*)scr;
        int32_t* binit = (int32_t*)(scr + align256(sizeof(uint32_t) * (size_t)nin * cpad));
        hipLaunchKernelGGL(ks_prep_kernel, dim3(count), dim3(256), 0, s, ext, second, nin, 1 << 29, prec_offset0, abar,
                           cpad, binit, P.ks_t, P.ks_basebit, nz);
        // small batches leave most CUs idle: slice the input coefficients over
        // up to 32 workgroups per tile (an even number of coefficients each). Tiles of 128
        // gates: up to 1024 workgroups, at least 128 coefficients per slice (4096 gates:
        // four slices, 0.374 -> 0.331 ms; 2048: 0.206 -> 0.189; 256: 0.078 -> 0.056; 1024
        // gates keep eight, sixteen ran 0.117 -> 0.125); tiles of 64: up to 512, at least
        // 16 (one gate: 32 slices)
        const int gt = count > 64 ? KS_MMA_GT_LARGE : KS_MMA_GT_SMALL;
        const int gx = dk->npad / 64, gy = (count + 16 * gt - 1) / (16 * gt);
        const int target = gt == KS_MMA_GT_LARGE ? 1024 : 512, min_coefs = gt == KS_MMA_GT_LARGE ? 128 : 16;
        int slices = 1;
        while (gx * gy * slices < target && slices < 32 && nin / (2 * slices) >= min_coefs) slices *= 2;
        const int ips = (nin / slices + 1) & ~1;
        if (slices > 1) hipLaunchKernelGGL(ks_out_init_kernel, dim3(count), dim3(256), 0, s, binit, P.n, out, out_idx);
        if (gt == KS_MMA_GT_LARGE)
            hipLaunchKernelGGL(ks_mma_kernel<KS_MMA_GT_LARGE>, dim3(gx, gy, slices), dim3(256), 0, s, dk->kp, abar,
                               cpad, nin, dk->npad, count, binit, P.n, out, out_idx, ips);
        else
            hipLaunchKernelGGL(ks_mma_kernel<KS_MMA_GT_SMALL>, dim3(gx, gy, slices), dim3(256), 0, s, dk->kp, abar,
                               cpad, nin, dk->npad, count, binit, P.n, out, out_idx, ips);
        HIP_TRY(hipGetLastError());
        return 0;
    }
    const KsPlan k = ks_plan(P, count);
    uint32_t* abar = (uint32_t*)scr;
    char* p = scr + align256(sizeof(uint32_t) * (size_t)nin * k.cpad);
    int32_t* binit = (int32_t*)p;
    int32_t* part = (int32_t*)(p + align256(sizeof(int32_t) * (size_t)k.cpad));
    const uint32_t prec_offset = 1u << (32 - (1 + P.ks_basebit * P.ks_t));
    hipLaunchKernelGGL(ks_prep_kernel, dim3(count), dim3(256), 0, s, ext, second, nin, 1 << 29, prec_offset, abar,
                       k.cpad, binit, P.ks_t, P.ks_basebit, nz);
    const int ips = (nin + k.splits - 1) / k.splits;
    dim3 grid(k.chunks, k.tiles, k.splits);
    if (k.waves == 4)
        hipLaunchKernelGGL(ks_kernel<4>, grid, dim3(256), 0, s, dk->ksk, abar, k.cpad, nin, P.ks_t, dk->npad, ips, part);
    else
        hipLaunchKernelGGL(ks_kernel<1>, grid, dim3(64), 0, s, dk->ksk, abar, k.cpad, nin, P.ks_t, dk->npad, ips, part);
    hipLaunchKernelGGL(ks_reduce_kernel, dim3(count), dim3(256), 0, s, part, k.splits, k.cpad, dk->npad, binit, P.n,
                       out, out_idx);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Per-batch bootstrap workspace: extracted samples of every job, then the
// MUX job list (mux_gate[count], second[count], n_jobs).
size_t ext_rows_bytes(const HostParams& P, int count) {
    return align256(sizeof(int32_t) * 2 * (size_t)count * (P.k * P.N + 1));
}
size_t ext_bytes(const HostParams& P, int count) {
    return ext_rows_bytes(P, count) + 2 * align256(sizeof(int32_t) * (size_t)count) + 256;
}

// Core: gates on device-resident data. `scr` must hold ext_bytes(count) +
// ks_scratch_bytes(count). d_ext_out (optional) receives the summed extracted
// samples instead of running the keyswitch.
// the autotuner's ladder of batch sizes (gates)
constexpr int TUNE_SIZES[] = {64, 128, 256, 384, 512, 640, 768, 896, 1024, 1280, 1536, 2048, 3072, 4096};
constexpr int TUNE_NS = sizeof(TUNE_SIZES) / sizeof(TUNE_SIZES[0]);
static_assert(TUNE_NS <= 16, "per-size choice tables");

// N = 1024, automatic mode: the autotuner's choice for a batch of `count` gates
// (that of the smallest ladder size that holds it; the batched kernel beyond the
// ladder), or 0 when the key has not been tuned or explicit crossovers are set
int tuned_choice(const Ctx& C, const tfhe_amd_device_key* dk, int count) {
    if (C.crossovers_set || !dk->br_choice[0]) return 0;
    for (int si = 0; si < TUNE_NS; si++)
        if (count <= TUNE_SIZES[si]) return dk->br_choice[si];
    return 2;
}

// N = 2048, automatic mode: the pair kernel when it was the faster at the smallest
// ladder size that holds the batch; beyond the ladder the two-wavefront kernel
bool pair_kernel_chosen(const tfhe_amd_device_key* dk, int count) {
    for (int si = 0; si < TUNE_NS; si++)
        if (count <= TUNE_SIZES[si]) return (dk->pair_sizes >> si) & 1u;
    return false;
}

// CUs of the device (0 if the query fails).
int device_cus(Ctx& C) {
    if (C.cus < 0) {
        int cus = 0;
        C.cus = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C.device) == hipSuccess && cus > 0
                    ? cus : 0;
    }
    return C.cus;
}

// Bootstraps that one whole round of the several-bootstrap workgroups holds on the
// device (resident workgroups x bootstraps per workgroup; 0 if a query fails):
// w = 4 br2_kernel<l, 4> (four two-wavefront bootstraps per workgroup), w = 5..8
// br4w_kernel<l, w>.
int whole_round(Ctx& C, int w, int l) {
    int& r = C.round_jobs[w - 4][l == 3];
    if (r < 0) {
        int per_cu = 0;
        hipError_t e1;
        if (w == 4)
            e1 = l == 3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, br2_kernel<3, 4>, 512, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, br2_kernel<2, 4>, 512, 0);
        else
            e1 = l == 3 ? br4w_occupancy<3>(w, &per_cu) : br4w_occupancy<2>(w, &per_cu);
        const int cus = device_cus(C);
        r = (e1 == hipSuccess && per_cu > 0 && cus > 0) ? per_cu * cus * w : 0;
    }
    return r;
}

// The blind-rotation kernel (mode) a batch of `count` gates runs with: the forced mode, or
// in automatic mode the autotuner's per-size choice or the crossovers (shared: other launches
// are on the device, so not the latency kernel), mapped to a kernel the key supports.

int resolve_mode(const Ctx& C, const tfhe_amd_device_key* dk, int count, bool shared) {
    int mode = C.br_mode;
    if (dk->M != 512) {
        // N = 2048: 1 one-wavefront, 4 pair (four wavefronts), 6 two-wavefront, 8 pair on the
        // radix-8 stages; otherwise automatic: the autotuner's (or the default) per-size choice
        if (mode == 1 || dk->P.n > BRP_NMAX || !dk->bkp) return 1;
        if (mode == 8 || !dk->bkq) return 8;
        if (mode != 4 && mode != 6) mode = pair_kernel_chosen(dk, count) ? 4 : 6;
        return mode;
    }
    if (mode == 8 || mode == 6 || mode == 4) mode = 2;
    if (mode == 0) {
        mode = tuned_choice(C, dk, count);
        if (mode == 0) {
            if (count <= C.brl_max && dk->P.n <= BRL_NMAX) mode = 5;
            else mode = (count <= C.br_small_max && dk->P.n <= BR2_NMAX) ? 3 : 2;
        }
        // beside other launches the latency kernel (one bootstrap per CU) would hold CUs
        // that the others' bootstraps could share; the two-wavefront kernel fits four per CU
        if (shared && mode == 5) mode = 3;
    }
    if (mode == 5 && (dk->P.n > BRL_NMAX || !dk->bkh)) mode = 3;
    if (mode == 3 && (dk->P.n > BR2_NMAX || !dk->bkc)) mode = 2;
    if ((mode == 2 || mode == 9) && !dk->bkc) mode = 7;
    return mode;
}

int gates_device_locked(Ctx& C, tfhe_amd_device_key* dk, int count, const int32_t* d_op, const int32_t* d_in0,
                        const int32_t* d_in1, const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                        int32_t* d_ext_out, int32_t mu, char* scr, hipStream_t s, bool allow_prof = true,
                        int ks_path = -1, int32_t* d_nz = nullptr, bool shared = false) {
    if (count <= 0) return count == 0 ? 0 : set_error("negative gate count");
    const HostParams& P = dk->P;
    const size_t eb = ext_bytes(P, count);
    int32_t* ext = (int32_t*)scr;
    char* jp = scr + ext_rows_bytes(P, count);
    int32_t* mux_gate = (int32_t*)jp;
    int32_t* second = (int32_t*)(jp + align256(sizeof(int32_t) * (size_t)count));
    int32_t* n_jobs = (int32_t*)(jp + 2 * align256(sizeof(int32_t) * (size_t)count));
    hipLaunchKernelGGL(mux_scan_kernel, dim3(1), dim3(1024), 0, s, d_op, count, mux_gate, second, n_jobs);
    BrArgs A = base_args(dk);
    A.mux_gate = mux_gate;
    A.n_jobs = n_jobs;
    A.op = d_op;
    A.in0 = d_in0;
    A.in1 = d_in1;
    A.in2 = d_in2;
    A.pool = d_pool;
    A.count = count;
    A.mu = mu;
    A.ext = ext;
    Ctx::Ev ev{};
    const bool prof = C.prof && allow_prof;
    if (prof) {
        for (auto& e : ev.e) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventRecord(ev.e[0], s));
    }
    // kernel choice: 1 = one wavefront per bootstrap (accumulator in LDS), 2 = the batched kernel (whole
    // rounds on eight-bootstrap workgroups, the rest on one-wavefront workgroups; 9: all on the latter),
    // 3 = two wavefronts per bootstrap, 5 = 4l wavefronts per bootstrap (half transforms); 0 = automatic
    // (5 for the smallest batches, 3 while the batch leaves CUs idle, 2 for large
    // batches). N = 2048 uses the pair-transform kernel (4: fused butterflies, 8: radix-8 stages)
    // unless mode 1 is forced.
    const int mode = resolve_mode(C, dk, count, shared);
    if (mode == 5) {
        const int blocks = 2 * count;  // one bootstrap per workgroup, at most 2*count jobs
        if (dk->P.l == 3) hipLaunchKernelGGL(brl_kernel<3>, dim3(blocks), dim3(768), 0, s, A, dk->bkh, dk->twh, mux_gate, n_jobs);
        else hipLaunchKernelGGL(brl_kernel<2>, dim3(blocks), dim3(512), 0, s, A, dk->bkh, dk->twh, mux_gate, n_jobs);
    } else if (mode == 4) {
        const int blocks = 2 * count;  // one bootstrap per workgroup, at most 2*count jobs
        if (dk->P.l == 2) hipLaunchKernelGGL((brp_kernel<2, true>), dim3(blocks), dim3(256), 0, s, A, dk->bkq, dk->twq, mux_gate, n_jobs);
        else hipLaunchKernelGGL((brp_kernel<3, true>), dim3(blocks), dim3(256), 0, s, A, dk->bkq, dk->twq, mux_gate, n_jobs);
    } else if (mode == 6) {
        // BRQ_BOOTSTRAPS bootstraps per workgroup, at most 2*count jobs
        const int blocks = (2 * count + BRQ_BOOTSTRAPS - 1) / BRQ_BOOTSTRAPS;
        if (dk->P.l == 2) hipLaunchKernelGGL(brq_kernel<2>, dim3(blocks), dim3(128 * BRQ_BOOTSTRAPS), 0, s, A, dk->bkq, dk->twq, mux_gate, n_jobs);
        else hipLaunchKernelGGL(brq_kernel<3>, dim3(blocks), dim3(128 * BRQ_BOOTSTRAPS), 0, s, A, dk->bkq, dk->twq, mux_gate, n_jobs);
    } else if (mode == 8) {
        const int blocks = 2 * count;  // the same kernel on the radix-8 stages (for comparison)
        if (dk->P.l == 2) hipLaunchKernelGGL(brp_kernel<2>, dim3(blocks), dim3(256), 0, s, A, dk->bkp, dk->twp, mux_gate, n_jobs);
        else hipLaunchKernelGGL(brp_kernel<3>, dim3(blocks), dim3(256), 0, s, A, dk->bkp, dk->twp, mux_gate, n_jobs);
    } else if (mode == 2 || mode == 9) {
        A.bkf = dk->bkc;  // the fused-butterfly transform's key and tables
        A.twc = dk->twc;
        // jobs [0, round_split(n_jobs, round)) (on the device, which knows the job count):
        // whole rounds of the device on the row-synchronised workgroups of w bootstraps; the
        // rest (at most 2*count jobs in all) on one-bootstrap workgroups (mode 9: every job
        // on these). Workgroups past the split or the job count leave at once.
        const int w = device_cus(C) > 0 ? br4w_waves(count, device_cus(C)) : BR4W_MAX;
        const int round = mode == 2 ? whole_round(C, w, dk->P.l) : 0;
        const int jmax = round > 0 ? 2 * count / round * round : 0;
        if (jmax > 0) {
            // a partial round on these when it gives every CU w - 1 or more bootstraps
            A.split_round = round;
            A.split_partial = round - round / w;
            if (dk->P.l == 3) launch_br4w<3>(w, jmax / w, s, A, mux_gate, n_jobs, round);
            else launch_br4w<2>(w, jmax / w, s, A, mux_gate, n_jobs, round);
        }
        const int blocks = 2 * count;
        if (dk->P.l == 3) hipLaunchKernelGGL((br4_kernel<3, true>), dim3(blocks), dim3(64), 0, s, A, mux_gate, n_jobs);
        else hipLaunchKernelGGL((br4_kernel<2, true>), dim3(blocks), dim3(64), 0, s, A, mux_gate, n_jobs);
    } else if (mode == 7) {
        const int blocks = 2 * count;  // the same kernel on the radix-8 transform Fft<512> (for comparison)
        if (dk->P.l == 3) hipLaunchKernelGGL(br4_kernel<3>, dim3(blocks), dim3(64), 0, s, A, mux_gate, n_jobs);
        else hipLaunchKernelGGL(br4_kernel<2>, dim3(blocks), dim3(64), 0, s, A, mux_gate, n_jobs);
    } else if (mode == 3) {
        A.bkf = dk->bkc;  // the fused-butterfly transform's key and tables
        A.twc = dk->twc;
        // whole device rounds on four-bootstrap workgroups, the rest on one-bootstrap ones
        // (split on the device as for mode 2; at most 2*count jobs in all)
        const int round = whole_round(C, 4, dk->P.l);
        const int jmax = round > 0 ? 2 * count / round * round : 0;
        if (jmax > 0) {
            // a partial round on these when it leaves at most an eighth of the slots empty
            // (one round of 1000 gates +3.9%; 900 even, 800 -3.3%)
            A.split_round = round;
            A.split_partial = round - round / 8;
            if (dk->P.l == 3) hipLaunchKernelGGL((br2_kernel<3, 4>), dim3(jmax / 4), dim3(512), 0, s, A, mux_gate, n_jobs, round);
            else hipLaunchKernelGGL((br2_kernel<2, 4>), dim3(jmax / 4), dim3(512), 0, s, A, mux_gate, n_jobs, round);
        }
        const int blocks = 2 * count;
        if (dk->P.l == 3) hipLaunchKernelGGL((br2_kernel<3, 1>), dim3(blocks), dim3(128), 0, s, A, mux_gate, n_jobs, 0);
        else hipLaunchKernelGGL((br2_kernel<2, 1>), dim3(blocks), dim3(128), 0, s, A, mux_gate, n_jobs, 0);
    } else {
        launch_br(dk, A, 2 * count, s);
    }
    HIP_TRY(hipGetLastError());
    if (prof) HIP_TRY(hipEventRecord(ev.e[1], s));
    int rc = 0;
    if (d_ext_out) {
        hipLaunchKernelGGL(sum_halves_kernel, dim3(count), dim3(256), 0, s, ext, second, P.k * P.N, mu, d_ext_out);
        HIP_TRY(hipGetLastError());
    } else {
        rc = run_keyswitch(dk, ext, second, count, d_pool, d_out, scr + eb, s, ks_path, d_nz);
    }
    if (prof) {
        HIP_TRY(hipEventRecord(ev.e[2], s));
        C.ev.push_back(ev);
    }
    return rc;
}

}  // namespace

// ------------------------------------------- internal entry points (scheduler) --

size_t engine_batch_scratch_bytes(const tfhe_amd_device_key* dk, int count) {
    return ext_bytes(dk->P, count) + ks_scratch_bytes(dk->P, count);
}

std::mutex& engine_mutex() { return ctx().mu; }

void* engine_stream() {
    Ctx& C = ctx();
    return ensure_ready(C) ? nullptr : (void*)C.stream;
}

int engine_launch_batch_locked(tfhe_amd_device_key* dk, int count, const int32_t* d_op, const int32_t* d_in0,
                               const int32_t* d_in1, const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                               char* scr, void* stream, bool capturing, bool shared) {
    return gates_device_locked(ctx(), dk, count, d_op, d_in0, d_in1, d_in2, d_out, d_pool, nullptr, 1 << 29, scr,
                               (hipStream_t)stream, !capturing, -1, nullptr, shared);
}

bool engine_can_segment(const tfhe_amd_device_key* dk) {
    return dk && dk->M == 512 && dk->bkc && dk->twc && (dk->P.l == 2 || dk->P.l == 3);
}

size_t engine_segments_scratch_bytes(tfhe_amd_device_key* const* dks, const int32_t* counts, int nseg) {
    // the segments' extracted samples and job lists, then one keyswitch area reused by
    // every segment in turn (its size is not monotone in the count, so take the largest)
    size_t b = 0, ks = 0;
    for (int g = 0; g < nseg; g++) {
        b += ext_bytes(dks[g]->P, counts[g]);
        ks = std::max(ks, ks_scratch_bytes(dks[g]->P, counts[g]));
    }
    return b + ks;
}

int engine_launch_segments_locked(tfhe_amd_device_key* const* dks, const int32_t* counts, int nseg,
                                  const int32_t* d_idx, int32_t* d_pool, char* scr, void* stream) {
    if (nseg <= 0 || nseg > SEG_MAX) return set_error("segments: 1..%d segments per launch", SEG_MAX);
    int total = 0, maxc = 0;
    for (int g = 0; g < nseg; g++) {
        if (!engine_can_segment(dks[g]) || !engine_same_params(dks[0], dks[g]))
            return set_error("segments: every key must be an N = 1024 key with the same parameters");
        if (counts[g] <= 0) return set_error("segments: empty segment");
        total += counts[g];
        maxc = std::max(maxc, (int)counts[g]);
    }
    const HostParams& P = dks[0]->P;
    hipStream_t s = (hipStream_t)stream;
    // index arrays [5][total]: op, in0, in1, in2, out; segment g is [off_g, off_g + count_g)
    BrSegTable T;
    memset(&T, 0, sizeof T);
    T.base = base_args(dks[0]);
    T.base.twc = dks[0]->twc;  // the tables are the same for every key
    T.base.op = d_idx;
    T.base.in0 = d_idx + total;
    T.base.in1 = d_idx + 2 * (size_t)total;
    T.base.in2 = d_idx + 3 * (size_t)total;
    T.base.pool = d_pool;
    int32_t* second[SEG_MAX];
    char* p = scr;
    for (int g = 0, off = 0; g < nseg; off += counts[g], g++) {
        const int c = counts[g];
        int32_t* ext = (int32_t*)p;
        char* jp = p + ext_rows_bytes(P, c);
        int32_t* mux_gate = (int32_t*)jp;
        second[g] = (int32_t*)(jp + align256(sizeof(int32_t) * (size_t)c));
        int32_t* n_jobs = (int32_t*)(jp + 2 * align256(sizeof(int32_t) * (size_t)c));
        p += ext_bytes(P, c);
        hipLaunchKernelGGL(mux_scan_kernel, dim3(1), dim3(1024), 0, s, d_idx + off, c, mux_gate, second[g], n_jobs);
        T.seg[g] = BrSeg{dks[g]->bkc, ext, mux_gate, n_jobs, off, c};
    }
    // at most 2*count jobs per segment; workgroups past a segment's job count leave at once
    if (P.l == 3) hipLaunchKernelGGL(br4_seg_kernel<3>, dim3(2 * maxc, nseg), dim3(64), 0, s, T);
    else hipLaunchKernelGGL(br4_seg_kernel<2>, dim3(2 * maxc, nseg), dim3(64), 0, s, T);
    HIP_TRY(hipGetLastError());
    // each segment's keyswitch with its own key, one after another in the shared scratch
    for (int g = 0; g < nseg; g++)
        if (int rc = run_keyswitch(dks[g], T.seg[g].ext, second[g], counts[g], d_pool, d_idx + 4 * (size_t)total + T.seg[g].off,
                                   p, s))
            return rc;
    return 0;
}

int engine_param_n(const tfhe_amd_device_key* dk) { return dk->P.n; }

bool engine_same_params(const tfhe_amd_device_key* a, const tfhe_amd_device_key* b) {
    return !memcmp(&a->P, &b->P, sizeof(HostParams));
}

size_t engine_key_bytes(const tfhe_amd_device_key* dk) { return dk->bytes; }

// ----------------------------------------------------------- key upload --

static int upload_key(Ctx& C, const LweBootstrappingKeyFFT* bf, tfhe_amd_device_key** out) {
    const HostParams P = host_params_fft(bf);
    if (int rc = check_supported(P)) return rc;
    const LweBootstrappingKey* bk = bkfft_coef(bf);
    if (!bk || !bk->bk || !bk->ks) return set_error("cloud key has no coefficient-domain bootstrapping key");
    const int N = P.N, k = P.k, kpl = (k + 1) * P.l, n = P.n, M = N / 2;
    const int base = 1 << P.ks_basebit, t = P.ks_t, nin = k * N;
    const int npad = ((n + 1 + 63) / 64) * 64;
    auto* dk = new tfhe_amd_device_key();
    static uint64_t next_gen = 0;
    dk->device = C.device;
    dk->gen = ++next_gen;  // under C.mu
    dk->pins = 0;
    dk->orphaned = false;
    dk->pair_sizes = 0x1f;  // until the autotuner measures it: the pair kernel up to 512 gates
    memset(dk->br_choice, 0, sizeof(dk->br_choice));
    dk->P = P;
    dk->M = M;
    dk->kpl = kpl;
    dk->npad = npad;
    const size_t npoly = (size_t)n * kpl * (k + 1);
    const size_t bkf_bytes = npoly * M * sizeof(c64);
    const size_t ksk_rows = (size_t)nin * t * (base - 1);
    const size_t ksk_bytes = ksk_rows * npad * sizeof(int32_t);
    const int twc = (M == 512) ? TwiddleTable<512>::COUNT : TwiddleTable<1024>::COUNT;
    dk->bytes = bkf_bytes + ksk_bytes + twc * sizeof(c64);
    dk->bkp = nullptr;
    dk->twp = nullptr;
    dk->bkh = nullptr;
    dk->twh = nullptr;
    dk->kp = nullptr;
    dk->bkc = nullptr;
    dk->twc = nullptr;
    dk->bkq = nullptr;
    dk->twq = nullptr;
    auto fail = [&](int rc) {
        (void)hipFree(dk->bkf);
        (void)hipFree(dk->ksk);
        (void)hipFree(dk->tw);
        (void)hipFree(dk->bkp);
        (void)hipFree(dk->twp);
        (void)hipFree(dk->bkh);
        (void)hipFree(dk->twh);
        (void)hipFree(dk->kp);
        (void)hipFree(dk->bkc);
        (void)hipFree(dk->twc);
        (void)hipFree(dk->bkq);
        (void)hipFree(dk->twq);
        delete dk;
        return rc;
    };
    const size_t kp_bytes = (size_t)4 * nin * npad * 32;
    if (P.ks_t == 8 && P.ks_basebit == 2 && npad % 64 == 0 && nin % 2 == 0) {
        if (hipMalloc(&dk->kp, kp_bytes) != hipSuccess)
            return fail(set_error("hipMalloc(keyswitch byte planes, %zu bytes) failed", kp_bytes));
        dk->bytes += kp_bytes;
    }
    if (M == 1024 && k == 1 && n <= BRP_NMAX) {
        if (hipMalloc(&dk->bkp, bkf_bytes) != hipSuccess || hipMalloc(&dk->twp, PAIR_TW_COUNT * sizeof(c64)) != hipSuccess)
            return fail(set_error("hipMalloc(pair-transform key, %zu bytes) failed", bkf_bytes));
        dk->bytes += bkf_bytes + PAIR_TW_COUNT * sizeof(c64);
        if (hipMalloc(&dk->bkq, bkf_bytes) != hipSuccess || hipMalloc(&dk->twq, FftCTPair::COUNT * sizeof(c64)) != hipSuccess)
            return fail(set_error("hipMalloc(fused pair-transform key, %zu bytes) failed", bkf_bytes));
        dk->bytes += bkf_bytes + FftCTPair::COUNT * sizeof(c64);
    }
    if (M == 512 && k == 1 && n <= BRL_NMAX) {
        if (hipMalloc(&dk->bkh, bkf_bytes) != hipSuccess || hipMalloc(&dk->twh, HALF_TW_COUNT * sizeof(c64)) != hipSuccess)
            return fail(set_error("hipMalloc(half-transform key, %zu bytes) failed", bkf_bytes));
        dk->bytes += bkf_bytes + HALF_TW_COUNT * sizeof(c64);
    }
    if (M == 512 && k == 1) {
        if (hipMalloc(&dk->bkc, bkf_bytes) != hipSuccess || hipMalloc(&dk->twc, FftCT::COUNT * sizeof(c64)) != hipSuccess)
            return fail(set_error("hipMalloc(fused-transform key, %zu bytes) failed", bkf_bytes));
        dk->bytes += bkf_bytes + FftCT::COUNT * sizeof(c64);
    }
    if (hipMalloc(&dk->bkf, bkf_bytes) != hipSuccess) return fail(set_error("hipMalloc(bootstrapping key, %zu bytes) failed", bkf_bytes));
    if (hipMalloc(&dk->ksk, ksk_bytes) != hipSuccess) return fail(set_error("hipMalloc(keyswitch key, %zu bytes) failed", ksk_bytes));
    if (hipMalloc(&dk->tw, twc * sizeof(c64)) != hipSuccess) return fail(set_error("hipMalloc(twiddles) failed"));
    // twiddles
    std::vector<c64> tw(twc);
    if (M == 512) build_twiddles<512>(tw.data());
    else build_twiddles<1024>(tw.data());
    // coefficient-domain BK, [n][kpl][k+1][N], staged through one host buffer
    std::vector<int32_t> coef(npoly * N);
    for (int i = 0; i < n; i++)
        for (int r = 0; r < kpl; r++)
            for (int c = 0; c <= k; c++)
                memcpy(&coef[(((size_t)i * kpl + r) * (k + 1) + c) * N], bk->bk[i].all_sample[r].a[c].coefsT,
                       sizeof(int32_t) * N);
    // keyswitch key rows h >= 1, [nin][t][base-1][npad]
    std::vector<int32_t> kh(ksk_rows * npad, 0);
    for (int i = 0; i < nin; i++)
        for (int j = 0; j < t; j++)
            for (int h = 1; h < base; h++) {
                const LweSample& s = bk->ks->ks0_raw[((size_t)i * t + j) * base + h];
                int32_t* dst = &kh[(((size_t)i * t + j) * (base - 1) + (h - 1)) * npad];
                memcpy(dst, s.a, sizeof(int32_t) * n);
                dst[n] = s.b;
            }
    int32_t* d_coef = nullptr;
    if (hipMalloc(&d_coef, coef.size() * sizeof(int32_t)) != hipSuccess) return fail(set_error("hipMalloc(staging) failed"));
    hipError_t e = hipMemcpyAsync(d_coef, coef.data(), coef.size() * sizeof(int32_t), hipMemcpyHostToDevice, C.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dk->tw, tw.data(), twc * sizeof(c64), hipMemcpyHostToDevice, C.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dk->ksk, kh.data(), ksk_bytes, hipMemcpyHostToDevice, C.stream);
    std::vector<c64> twp;
    if (dk->twp) {
        twp.resize(PAIR_TW_COUNT);
        build_twiddles_pair(twp.data());
        if (e == hipSuccess) e = hipMemcpyAsync(dk->twp, twp.data(), PAIR_TW_COUNT * sizeof(c64), hipMemcpyHostToDevice, C.stream);
    }
    std::vector<c64> twq;
    if (dk->twq) {
        twq.resize(FftCTPair::COUNT);
        build_twiddles_ct_pair(twq.data());
        if (e == hipSuccess) e = hipMemcpyAsync(dk->twq, twq.data(), FftCTPair::COUNT * sizeof(c64), hipMemcpyHostToDevice, C.stream);
    }
    std::vector<c64> twct;
    if (dk->twc) {
        twct.resize(FftCT::COUNT);
        build_twiddles_ct(twct.data());
        if (e == hipSuccess) e = hipMemcpyAsync(dk->twc, twct.data(), FftCT::COUNT * sizeof(c64), hipMemcpyHostToDevice, C.stream);
    }
    std::vector<c64> twh;
    if (dk->twh) {
        twh.resize(HALF_TW_COUNT);
        build_twiddles_half(twh.data());
        if (e == hipSuccess) e = hipMemcpyAsync(dk->twh, twh.data(), HALF_TW_COUNT * sizeof(c64), hipMemcpyHostToDevice, C.stream);
    }
    if (e == hipSuccess) {
        if (M == 512) hipLaunchKernelGGL(bk_convert_kernel<512>, dim3((unsigned)npoly), dim3(64), 0, C.stream, d_coef, dk->bkf, dk->tw);
        else hipLaunchKernelGGL(bk_convert_kernel<1024>, dim3((unsigned)npoly), dim3(64), 0, C.stream, d_coef, dk->bkf, dk->tw);
        e = hipGetLastError();
        if (e == hipSuccess && dk->bkp) {
            hipLaunchKernelGGL(bk_convert_pair_kernel, dim3((unsigned)npoly), dim3(128), 0, C.stream, d_coef, dk->bkp, dk->twp);
            e = hipGetLastError();
        }
        if (e == hipSuccess && dk->bkq) {
            hipLaunchKernelGGL(bk_convert_pair_ct_kernel, dim3((unsigned)npoly), dim3(128), 0, C.stream, d_coef, dk->bkq, dk->twq);
            e = hipGetLastError();
        }
        if (e == hipSuccess && dk->bkc) {
            hipLaunchKernelGGL(bk_convert_ct_kernel, dim3((unsigned)npoly), dim3(64), 0, C.stream, d_coef, dk->bkc, dk->twc);
            e = hipGetLastError();
        }
        if (e == hipSuccess && dk->bkh) {
            hipLaunchKernelGGL(bk_convert_half_kernel, dim3((unsigned)npoly), dim3(128), 0, C.stream, d_coef, dk->bkh, dk->twh);
            e = hipGetLastError();
        }
        if (e == hipSuccess && dk->kp) {
            const size_t threads = (size_t)nin * npad * 8;
            hipLaunchKernelGGL(ks_planes_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, C.stream,
                               dk->ksk, nin, npad, dk->kp);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(C.stream);
    (void)hipFree(d_coef);
    if (e != hipSuccess) return fail(set_error("key upload failed: %s", hipGetErrorString(e)));
    *out = dk;
    return 0;
}

static tfhe_amd_device_key* engine_key_impl(const LweBootstrappingKeyFFT* bf0, bool pin);

tfhe_amd_device_key* engine_key(const LweBootstrappingKeyFFT* bf0) { return engine_key_impl(bf0, false); }

tfhe_amd_device_key* engine_key_pinned(const LweBootstrappingKeyFFT* bf0) { return engine_key_impl(bf0, true); }

static tfhe_amd_device_key* engine_key_impl(const LweBootstrappingKeyFFT* bf0, bool pin) {
    if (!bf0) {
        set_error("null bootstrapping key");
        return nullptr;
    }
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    if (ensure_ready(C)) return nullptr;
    LweBootstrappingKeyFFT* bf = const_cast<LweBootstrappingKeyFFT*>(bf0);
    tfhe_amd_device_key*& slot = bkfft_dkey(bf);
    if (!slot) {
        tfhe_amd_device_key* dk = nullptr;
        if (upload_key(C, bf, &dk)) return nullptr;
        slot = dk;
    }
    if (pin) slot->pins++;  // under the same lock as the lookup: no release in between
    return slot;
}

static void free_key(Ctx& C, tfhe_amd_device_key* dk) {
    if (!dk) return;
    // launches queued on any stream (the caller's, a server's) may still read the key
    if (C.ready) (void)hipDeviceSynchronize();
    (void)hipFree(dk->bkf);
    (void)hipFree(dk->ksk);
    (void)hipFree(dk->tw);
    (void)hipFree(dk->bkp);
    (void)hipFree(dk->twp);
    (void)hipFree(dk->bkh);
    (void)hipFree(dk->twh);
    (void)hipFree(dk->kp);
    (void)hipFree(dk->bkc);
    (void)hipFree(dk->twc);
    (void)hipFree(dk->bkq);
    (void)hipFree(dk->twq);
    delete dk;
}

void engine_forget_host_key(const LweBootstrappingKeyFFT* bkfft) {
    if (!bkfft || !bkfft_dkey(bkfft)) return;
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    tfhe_amd_device_key*& slot = bkfft_dkey(bkfft);
    tfhe_amd_device_key* dk = slot;
    slot = nullptr;  // the next use through this host key uploads a fresh copy
    if (dk->pins > 0) dk->orphaned = true;  // a server still runs on it; the last unpin frees it
    else free_key(C, dk);
}

uint64_t engine_key_gen(const tfhe_amd_device_key* dk) { return dk->gen; }

void engine_key_unpin(tfhe_amd_device_key* dk) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    if (--dk->pins == 0 && dk->orphaned) free_key(C, dk);
}

// ------------------------------------------------------ host-buffer paths --

// Runs fn(begin, end) over [0, n) on up to 8 host threads for large n (the
// per-sample copies between LweSample arrays and the pinned staging buffer).
template <class Fn>
static void host_parallel(int n, Fn fn) {
    const int hw = (int)std::thread::hardware_concurrency();
    const int nt = n < 1024 ? 1 : (hw < 2 ? 1 : (hw < 8 ? hw : 8));
    if (nt == 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    const int per = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        const int b = t * per, e = b + per < n ? b + per : n;
        if (b < e) th.emplace_back(fn, b, e);
    }
    for (auto& x : th) x.join();
}

int engine_gates_host(const LweBootstrappingKeyFFT* bf, int32_t count, const int32_t* ops,
                      LweSample* const* out, const LweSample* const* in0, const LweSample* const* in1,
                      const LweSample* const* in2, int32_t mu, bool keyswitch, int32_t* nz) {
    if (count == 0) return 0;
    if (count < 0 || !ops || !out || !in0) return set_error("bad arguments to gate batch");
    tfhe_amd_device_key* dk = engine_key(bf);
    if (!dk) return -1;
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    const HostParams& P = dk->P;
    const int w = P.n + 1, nx = P.k * P.N + 1;
    for (int i = 0; i < count; i++) {
        const int op = ops[i];
        if (op < 0 || op > OP_IDENTITY) return set_error("gate %d: unknown gate code %d", i, op);
        if (!out[i] || !in0[i] || (op != OP_IDENTITY && !(in1 && in1[i])) || (op == TFHE_AMD_MUX && !(in2 && in2[i])))
            return set_error("gate %d: missing ciphertext", i);
    }
    // deduplicate input ciphertexts into pool slots
    std::unordered_map<const LweSample*, int32_t> slot;
    std::vector<const LweSample*> uniq;
    auto slot_of = [&](const LweSample* p) {
        auto it = slot.find(p);
        if (it != slot.end()) return it->second;
        int32_t s = (int32_t)uniq.size();
        slot.emplace(p, s);
        uniq.push_back(p);
        return s;
    };
    std::vector<int32_t> idx(5 * (size_t)count);
    for (int i = 0; i < count; i++) {
        idx[i] = ops[i];
        idx[count + i] = slot_of(in0[i]);
        idx[2 * count + i] = ops[i] == OP_IDENTITY ? idx[count + i] : slot_of(in1[i]);
        idx[3 * count + i] = ops[i] == TFHE_AMD_MUX ? slot_of(in2[i]) : idx[count + i];
    }
    const int nin_slots = (int)uniq.size();
    for (int i = 0; i < count; i++) idx[4 * count + i] = nin_slots + i;
    const size_t idx_bytes = align256(idx.size() * sizeof(int32_t));
    const size_t pool_in_bytes = (size_t)nin_slots * w * sizeof(int32_t);
    const size_t out_row = keyswitch ? (size_t)w : (size_t)nx;
    const size_t out_bytes = (size_t)count * out_row * sizeof(int32_t);
    const size_t pool_bytes = align256(pool_in_bytes + (keyswitch ? out_bytes : 0));
    const size_t extout_bytes = keyswitch ? 0 : align256(out_bytes);
    const size_t nz_bytes = (keyswitch && nz) ? align256(sizeof(int32_t) * (size_t)count) : 0;
    if (int rc = ensure_pinned(C, idx_bytes + pool_bytes + extout_bytes + nz_bytes)) return rc;
    char* h = (char*)C.h_pinned;
    memcpy(h, idx.data(), idx.size() * sizeof(int32_t));
    int32_t* hp = (int32_t*)(h + idx_bytes);
    host_parallel(nin_slots, [&](int b, int e) {
        for (int s = b; s < e; s++) {
            memcpy(hp + (size_t)s * w, uniq[s]->a, sizeof(int32_t) * P.n);
            hp[(size_t)s * w + P.n] = uniq[s]->b;
        }
    });
    // device layout: [idx][pool][ext out][nonzero digit counts] then the engine's own scratch
    const size_t dev_bytes = idx_bytes + pool_bytes + extout_bytes + nz_bytes;
    const size_t work = ext_bytes(P, count) + ks_scratch_bytes(P, count);
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, dev_bytes + work, C.stream, &ws)) return rc;
    // the engine's working buffers follow the staged region
    char* d = (char*)ws->d;
    HIP_TRY(hipMemcpyAsync(d, h, idx_bytes + pool_in_bytes, hipMemcpyHostToDevice, C.stream));
    const int32_t* d_idx = (const int32_t*)d;
    int32_t* d_pool = (int32_t*)(d + idx_bytes);
    int32_t* d_extout = keyswitch ? nullptr : (int32_t*)(d + idx_bytes + pool_bytes);
    int32_t* d_nz = nz_bytes ? (int32_t*)(d + idx_bytes + pool_bytes + extout_bytes) : nullptr;
    int rc = gates_device_locked(C, dk, count, d_idx, d_idx + count, d_idx + 2 * count, d_idx + 3 * count,
                                 d_idx + 4 * count, d_pool, d_extout, mu, d + dev_bytes, C.stream, true, -1, d_nz);
    if (rc) return rc;
    const void* src = keyswitch ? (const void*)(d_pool + (size_t)nin_slots * w) : (const void*)d_extout;
    HIP_TRY(hipMemcpyAsync(h + idx_bytes, src, out_bytes, hipMemcpyDeviceToHost, C.stream));
    char* h_nz = h + idx_bytes + pool_bytes + extout_bytes;
    if (d_nz) HIP_TRY(hipMemcpyAsync(h_nz, d_nz, sizeof(int32_t) * (size_t)count, hipMemcpyDeviceToHost, C.stream));
    HIP_TRY(hipStreamSynchronize(C.stream));
    if (d_nz) memcpy(nz, h_nz, sizeof(int32_t) * (size_t)count);
    ws->busy = false;  // the workspace is idle again
    const int32_t* ho = (const int32_t*)(h + idx_bytes);
    auto scatter = [&](int b, int e) {
        for (int i = b; i < e; i++) {
            const int32_t* row = ho + (size_t)i * out_row;
            LweSample* o = out[i];
            memcpy(o->a, row, sizeof(int32_t) * (out_row - 1));
            o->b = row[out_row - 1];
        }
    };
    // outputs named twice keep the serial order (the later gate wins), so only
    // distinct outputs are written from several threads
    bool distinct = count >= 1024;
    if (distinct) {
        std::vector<const LweSample*> sorted(out, out + count);
        std::sort(sorted.begin(), sorted.end());
        distinct = std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end();
    }
    if (distinct) host_parallel(count, scatter);
    else scatter(0, count);
    return 0;
}

int engine_blind_rotate_host(tfhe_amd_device_key* dk, int32_t row0, int32_t n, const int32_t* bara, int32_t* acc,
                             int32_t* ext_out) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    const HostParams& P = dk->P;
    if (n < 0 || n > 1024 || row0 < 0 || row0 + n > P.n) return set_error("blind rotate: bad key row range");
    const int N = P.N, nx = P.k * N + 1;
    const size_t acc_bytes = align256(sizeof(int32_t) * 2 * N);
    const size_t bara_bytes = align256(sizeof(int32_t) * 1024);
    const size_t ext_b = align256(sizeof(int32_t) * nx);
    const size_t total = 2 * acc_bytes + bara_bytes + ext_b;
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, total, C.stream, &ws)) return rc;
    if (int rc = ensure_pinned(C, total)) return rc;
    char* h = (char*)C.h_pinned;
    memcpy(h, acc, sizeof(int32_t) * 2 * N);
    if (n) memcpy(h + acc_bytes, bara, sizeof(int32_t) * n);
    char* d = (char*)ws->d;
    HIP_TRY(hipMemcpyAsync(d, h, acc_bytes + bara_bytes, hipMemcpyHostToDevice, C.stream));
    BrArgs A = base_args(dk);
    A.n = n;
    A.bkf = dk->bkf + (size_t)row0 * dk->kpl * (P.k + 1) * dk->M;
    A.acc_in = (const int32_t*)d;
    A.bara_in = (const int32_t*)(d + acc_bytes);
    A.acc_out = (int32_t*)(d + acc_bytes + bara_bytes);
    A.ext = ext_out ? (int32_t*)(d + 2 * acc_bytes + bara_bytes) : nullptr;
    launch_br(dk, A, 1, C.stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(h, d + acc_bytes + bara_bytes, acc_bytes + ext_b, hipMemcpyDeviceToHost, C.stream));
    HIP_TRY(hipStreamSynchronize(C.stream));
    ws->busy = false;
    memcpy(acc, h, sizeof(int32_t) * 2 * N);
    if (ext_out) memcpy(ext_out, h + acc_bytes, sizeof(int32_t) * nx);
    return 0;
}

int engine_keyswitch_host(tfhe_amd_device_key* dk, const int32_t* in, int32_t* out, int32_t* nz) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    const HostParams& P = dk->P;
    const int nx = P.k * P.N + 1, w = P.n + 1;
    // device: [in][out + the nonzero digit count][keyswitch scratch]
    const size_t in_b = align256(sizeof(int32_t) * 2 * nx), out_b = align256(sizeof(int32_t) * (w + 1));
    const size_t work = ks_scratch_bytes(P, 1);
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, in_b + out_b + work, C.stream, &ws)) return rc;
    if (int rc = ensure_pinned(C, in_b + out_b)) return rc;
    char* h = (char*)C.h_pinned;
    memcpy(h, in, sizeof(int32_t) * nx);
    char* d = (char*)ws->d;
    HIP_TRY(hipMemcpyAsync(d, h, sizeof(int32_t) * nx, hipMemcpyHostToDevice, C.stream));
    int32_t* d_out = (int32_t*)(d + in_b);
    if (int rc = run_keyswitch(dk, (const int32_t*)d, nullptr, 1, d_out, nullptr, d + in_b + out_b, C.stream, -1,
                               d_out + w))
        return rc;
    HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(int32_t) * (w + 1), hipMemcpyDeviceToHost, C.stream));
    HIP_TRY(hipStreamSynchronize(C.stream));
    ws->busy = false;
    memcpy(out, h, sizeof(int32_t) * w);
    if (nz) *nz = ((const int32_t*)h)[w];
    return 0;
}

}  // namespace tfhe_amd

// ------------------------------------------------------------- C-ABI --------
using namespace tfhe_amd;

TFHE_AMD_EXPORT int32_t tfhe_amd_cloud_key_upload(const TFheGateBootstrappingCloudKeySet* ck) {
    if (!ck) return set_error("null cloud key");
    return engine_key(ck->bkFFT) ? 0 : -1;
}

TFHE_AMD_EXPORT void tfhe_amd_cloud_key_release(const TFheGateBootstrappingCloudKeySet* ck) {
    if (ck) engine_forget_host_key(ck->bkFFT);
}

TFHE_AMD_EXPORT size_t tfhe_amd_cloud_key_device_bytes(const TFheGateBootstrappingCloudKeySet* ck) {
    return (ck && ck->bkFFT && bkfft_dkey(ck->bkFFT)) ? bkfft_dkey(ck->bkFFT)->bytes : 0;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_gates_device(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                              const int32_t* d_op, const int32_t* d_in0, const int32_t* d_in1,
                                              const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                                              void* stream) {
    if (!ck) return set_error("null cloud key");
    tfhe_amd_device_key* dk = engine_key(ck->bkFFT);
    if (!dk) return -1;
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    if (count < 0) return set_error("negative gate count");
    const hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, ext_bytes(dk->P, count) + ks_scratch_bytes(dk->P, count), s, &ws)) return rc;
    if (int rc = gates_device_locked(C, dk, count, d_op, d_in0, d_in1, d_in2, d_out, d_pool, nullptr, 1 << 29,
                                     (char*)ws->d, s))
        return rc;
    return scratch_done(ws, s);
}

TFHE_AMD_EXPORT int32_t tfhe_amd_gates_device_noalloc(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                                      const int32_t* d_op, const int32_t* d_in0, const int32_t* d_in1,
                                                      const int32_t* d_in2, const int32_t* d_out, int32_t* d_pool,
                                                      void* stream) {
    if (!ck || !ck->bkFFT) return set_error("gates_device_noalloc: null cloud key");
    if (count < 0) return set_error("gates_device_noalloc: negative gate count");
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    // an upload allocates: the key must already be on the device (read under the
    // lock a release takes)
    tfhe_amd_device_key* dk = bkfft_dkey(ck->bkFFT);
    if (!C.ready || !dk)
        return set_error("gates_device_noalloc: the cloud key is not on the device (call tfhe_amd_cloud_key_upload first)");
    if (count == 0) return 0;
    // the keyswitch path is fixed for this call, so the scratch it needs is known
    const int path = ks_path_now(dk);
    const size_t need = ext_bytes(dk->P, count) + ks_scratch_bytes_path(dk->P, count, path == 0);
    if (need > NOALLOC_SCRATCH_BYTES) {
        int most = count;
        while (most > 0 && ext_bytes(dk->P, most) + ks_scratch_bytes_path(dk->P, most, path == 0) > NOALLOC_SCRATCH_BYTES)
            most = most * 7 / 8;
        return set_error("gates_device_noalloc: %d gates need %zu bytes of workspace, above the static %zu bytes "
                         "(about %d gates fit with this key)", count, need, NOALLOC_SCRATCH_BYTES, most);
    }
    const hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    return gates_device_locked(C, dk, count, d_op, d_in0, d_in1, d_in2, d_out, d_pool, nullptr, 1 << 29, C.noalloc_scr,
                               s, false, path);
}

TFHE_AMD_EXPORT int32_t tfhe_amd_gates_device_woks(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                                   const int32_t* d_op, const int32_t* d_in0, const int32_t* d_in1,
                                                   const int32_t* d_in2, const int32_t* d_pool, int32_t* d_ext,
                                                   void* stream) {
    if (!ck) return set_error("null cloud key");
    tfhe_amd_device_key* dk = engine_key(ck->bkFFT);
    if (!dk) return -1;
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    if (count < 0) return set_error("negative gate count");
    const hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, ext_bytes(dk->P, count) + ks_scratch_bytes(dk->P, count), s, &ws)) return rc;
    if (int rc = gates_device_locked(C, dk, count, d_op, d_in0, d_in1, d_in2, nullptr, const_cast<int32_t*>(d_pool),
                                     d_ext, 1 << 29, (char*)ws->d, s))
        return rc;
    return scratch_done(ws, s);
}

TFHE_AMD_EXPORT int32_t tfhe_amd_keyswitch_device(const TFheGateBootstrappingCloudKeySet* ck, int32_t count,
                                                  const int32_t* d_in, int32_t* d_out, void* stream) {
    if (!ck) return set_error("null cloud key");
    tfhe_amd_device_key* dk = engine_key(ck->bkFFT);
    if (!dk) return -1;
    if (count <= 0) return count == 0 ? 0 : set_error("negative count");
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    // run_keyswitch reads ext rows [2*count][..]; with op == null only the first count are used
    const hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, ks_scratch_bytes(dk->P, count), s, &ws)) return rc;
    if (int rc = run_keyswitch(dk, d_in, nullptr, count, d_out, nullptr, (char*)ws->d, s)) return rc;
    return scratch_done(ws, s);
}

TFHE_AMD_EXPORT void tfhe_amd_set_launch_config(int32_t br_mode, int32_t latency_max, int32_t small_batch_max) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    C.br_mode = br_mode;
    if (latency_max >= 0) C.brl_max = latency_max;
    if (small_batch_max >= 0) C.br_small_max = small_batch_max;
    if (latency_max >= 0 || small_batch_max >= 0) C.crossovers_set = true;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_kernel_for(const TFheGateBootstrappingCloudKeySet* ck, int32_t count) {
    if (!ck || !ck->bkFFT) return set_error("kernel_for: null cloud key");
    if (count <= 0) return set_error("kernel_for: the gate count must be positive");
    tfhe_amd_device_key* dk = engine_key(ck->bkFFT);
    if (!dk) return -1;
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    return resolve_mode(C, dk, count, false);
}

TFHE_AMD_EXPORT int32_t tfhe_amd_get_launch_config(int32_t* br_mode, int32_t* latency_max, int32_t* small_batch_max) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    if (br_mode) *br_mode = C.br_mode;
    if (latency_max) *latency_max = C.brl_max;
    if (small_batch_max) *small_batch_max = C.br_small_max;
    return 0;
}

// Times the three N = 1024 blind-rotation kernels (each with its keyswitch) on
// pseudo-random ciphertexts over a ladder of batch sizes and sets the automatic
// mode's crossovers: the latency kernel up to the largest size up to
// which it is the fastest at every measured point, then the two-wavefront
// kernel up to the largest size up to which it beats the four-gate kernel at
// every further point.
TFHE_AMD_EXPORT int32_t tfhe_amd_autotune(const TFheGateBootstrappingCloudKeySet* ck, int32_t* latency_max,
                                          int32_t* small_batch_max, float* ms) {
    if (!ck) return set_error("null cloud key");
    tfhe_amd_device_key* dk = engine_key(ck->bkFFT);
    if (!dk) return -1;
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    const int* sizes = TUNE_SIZES;
    constexpr int NS = TUNE_NS;
    // N = 1024: the latency, two-wavefront and batched kernels (modes 5, 3, 2);
    // N = 2048: the four-wavefront pair kernel and the two-wavefront kernel (modes 4, 6)
    const bool pbig = dk->M == 1024;
    static const int modes512[3] = {5, 3, 2}, modes1024[2] = {4, 6};
    const int nm = pbig ? 2 : 3;
    const int* modes = pbig ? modes1024 : modes512;
    if (pbig ? (dk->P.n > BRP_NMAX || !dk->bkq) : (dk->M != 512 || dk->P.n > BR4_ANY_NMAX)) {
        if (!pbig) {
            C.brl_max = 0;
            C.br_small_max = 0;
        }
        if (latency_max) *latency_max = 0;
        if (small_batch_max) *small_batch_max = 0;
        return 0;
    }
    const int maxB = sizes[NS - 1], w = dk->P.n + 1;
    int32_t *d_pool = nullptr, *d_idx = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<int32_t> idx(4 * (size_t)maxB);
    for (int i = 0; i < maxB; i++) {
        idx[i] = TFHE_AMD_NAND;
        idx[maxB + i] = i;
        idx[2 * maxB + i] = maxB + i;
        idx[3 * maxB + i] = 2 * maxB + i;
    }
    auto cleanup = [&](int rc) {
        (void)hipStreamSynchronize(C.stream);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(d_pool);
        (void)hipFree(d_idx);
        return rc;
    };
    if (hipMalloc(&d_pool, sizeof(int32_t) * 3 * (size_t)maxB * w) != hipSuccess ||
        hipMalloc(&d_idx, sizeof(int32_t) * idx.size()) != hipSuccess)
        return cleanup(set_error("autotune: device allocation failed"));
    // pseudo-random ciphertexts: every key row is used (a zero mask would let
    // the kernels that skip rows with a zero rotation look faster than they are)
    std::vector<int32_t> pool(3 * (size_t)maxB * w);
    uint64_t st = 0x9e3779b97f4a7c15ULL;
    for (auto& v : pool) {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        v = (int32_t)(st >> 32);
    }
    if (hipMemcpy(d_pool, pool.data(), sizeof(int32_t) * pool.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpyAsync(d_idx, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice, C.stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return cleanup(set_error("autotune: set-up failed"));
    Scratch* ws = nullptr;
    if (int rc = ensure_scratch(C, ext_bytes(dk->P, maxB) + ks_scratch_bytes(dk->P, maxB), C.stream, &ws))
        return cleanup(rc);
    const int saved = C.br_mode;
    float best[3][NS] = {};
    int rc = 0;
    for (int m = 0; m < nm && !rc; m++) {
        C.br_mode = modes[m];
        for (int si = 0; si < NS && !rc; si++) {
            const int B = sizes[si];
            best[m][si] = 1e30f;
            for (int rep = 0; rep < 4 && !rc; rep++) {
                rc = (int)(hipEventRecord(e0, C.stream) != hipSuccess);
                if (!rc)
                    rc = gates_device_locked(C, dk, B, d_idx, d_idx + maxB, d_idx + 2 * maxB, d_idx + maxB,
                                             d_idx + 3 * maxB, d_pool, nullptr, 1 << 29, (char*)ws->d,
                                             C.stream, false);
                float t = 0;
                if (!rc && (hipEventRecord(e1, C.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&t, e0, e1) != hipSuccess))
                    rc = set_error("autotune: event timing failed");
                if (!rc && rep) best[m][si] = std::min(best[m][si], t);  // first run warms up
            }
        }
    }
    C.br_mode = saved;
    if (rc) return cleanup(rc < 0 ? rc : set_error("autotune: event record failed"));
    if (pbig) {
        // per ladder size, the faster kernel; the crossover reported is the largest size up to
        // which the pair kernel is the faster at every point
        uint32_t mask = 0;
        for (int si = 0; si < NS; si++)
            if (best[0][si] <= best[1][si]) mask |= 1u << si;
        dk->pair_sizes = mask;
        int pm = 0;
        for (int si = 0; si < NS && ((mask >> si) & 1u); si++) pm = sizes[si];
        if (latency_max) *latency_max = 0;
        if (small_batch_max) *small_batch_max = pm;
        if (ms)
            for (int m = 0; m < 3; m++)
                for (int k = 0; k < NS; k++) ms[m * NS + k] = m < nm ? best[m][k] : 0.0f;
        return cleanup(0);
    }
    // per ladder size, the fastest kernel (automatic mode); the crossovers reported are the
    // largest size up to which the latency kernel is the fastest at every point and, after
    // it, the largest size up to which the two-wavefront kernel beats the batched one
    for (int k = 0; k < NS; k++) {
        const int m = (best[0][k] <= best[1][k] && best[0][k] <= best[2][k]) ? 0 : (best[1][k] < best[2][k] ? 1 : 2);
        dk->br_choice[k] = (uint8_t)modes[m];
    }
    int si = 0, lat = 0, small = 0;
    for (; si < NS && best[0][si] <= best[1][si] && best[0][si] <= best[2][si]; si++) lat = sizes[si];
    small = lat;
    for (; si < NS && best[1][si] < best[2][si]; si++) small = sizes[si];
    C.brl_max = lat;
    C.br_small_max = small;
    C.crossovers_set = false;
    if (latency_max) *latency_max = lat;
    if (small_batch_max) *small_batch_max = small;
    if (ms)
        for (int m = 0; m < 3; m++)
            for (int k = 0; k < NS; k++) ms[m * NS + k] = best[m][k];
    return cleanup(0);
}

TFHE_AMD_EXPORT void tfhe_amd_set_keyswitch_path(int32_t path) { g_ks_path.store(path == 1 ? 1 : 0); }

TFHE_AMD_EXPORT void tfhe_amd_profile_enable(int32_t on) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    C.prof = on != 0;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_profile_read(double* br_ms, double* ks_ms, int32_t* launches) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    double br = 0, ks = 0;
    int rc = 0;
    for (auto& e : C.ev) {
        float a = 0, b = 0;
        if (hipEventSynchronize(e.e[2]) != hipSuccess || hipEventElapsedTime(&a, e.e[0], e.e[1]) != hipSuccess ||
            hipEventElapsedTime(&b, e.e[1], e.e[2]) != hipSuccess)
            rc = set_error("event timing failed");
        br += a;
        ks += b;
        for (auto x : e.e) (void)hipEventDestroy(x);
    }
    if (br_ms) *br_ms = br;
    if (ks_ms) *ks_ms = ks;
    if (launches) *launches = (int32_t)C.ev.size();
    C.ev.clear();
    return rc;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_copy_device(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return set_error("copy_device: 16-byte multiples only");
    Ctx& C = ctx();
    {
        std::lock_guard<std::mutex> g(C.mu);
        if (int rc = ensure_ready(C)) return rc;
    }
    hipStream_t s = stream ? (hipStream_t)stream : C.stream;
    hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, s, (const int4*)src, (int4*)dst, bytes / 16);
    HIP_TRY(hipGetLastError());
    return 0;
}

TFHE_AMD_EXPORT void* tfhe_amd_malloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        set_error("hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_free(void* ptr) {
    HIP_TRY(hipFree(ptr));
    return 0;
}

// Both copies are ordered after the work queued on the library's stream and
// complete before returning.
TFHE_AMD_EXPORT int32_t tfhe_amd_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    Ctx& C = ctx();
    hipStream_t s;
    {
        std::lock_guard<std::mutex> g(C.mu);
        if (int rc = ensure_ready(C)) return rc;
        s = C.stream;
    }
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    Ctx& C = ctx();
    hipStream_t s;
    {
        std::lock_guard<std::mutex> g(C.mu);
        if (int rc = ensure_ready(C)) return rc;
        s = C.stream;
    }
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return 0;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_synchronize(void) {
    Ctx& C = ctx();
    {
        std::lock_guard<std::mutex> g(C.mu);
        if (int rc = ensure_ready(C)) return rc;
    }
    HIP_TRY(hipStreamSynchronize(C.stream));
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

TFHE_AMD_EXPORT void* tfhe_amd_stream(void) {
    Ctx& C = ctx();
    std::lock_guard<std::mutex> g(C.mu);
    if (ensure_ready(C)) return nullptr;
    return (void*)C.stream;
}

TFHE_AMD_EXPORT int32_t tfhe_amd_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
