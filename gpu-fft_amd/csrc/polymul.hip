// Fused batched polynomial multiplication (additive API; the reference needs
// two GPU_FFNT calls and a caller-side rounding for the same product,
// example/test_ffnt_gpu_R_R.cu:137-162).
//
// A plan owns its root tables (generated on the GPU in FP64 in the
// reference's FFNT layouts) and the scratch buffer of the split kernels.
//   GPUFFT_PREC_F64        double in, double out (unrounded, like GPU_FFNT)
//   GPUFFT_PREC_F32_SPLIT  int32 in, int64 out: the operands are split into
//                          balanced base-2^w limbs, every limb product runs in
//                          FP32 and is rounded, and the rounded limb products
//                          are recombined exactly in int64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "fused.hpp"
#include "status.hpp"

using namespace gfa;

struct gpufft_polymul_plan_s {
    int device;
    int ring_power;  // log2 N
    int n;           // log2 L, L = N/2
    int precision;
    int max_batch;
    int wa, ka, wb, kb;
    int ca;          // FP32 split: a split as a - 2^ca (0: none), plus the limb 2^ca J b
    void* jspec;     // FP32 split with ca: forward FFNT spectrum of J (complex float, L)
    int S, SB;       // sub-problem split
    int rb;          // radix bits per LDS step of the product kernel
    int e;           // log2 points per thread of the product kernel (2^(S-e) threads)
    int cmb;         // FP64, SB > 0: 1 = top stages in the product kernel, 0 = separate pass (default)
    void* root;      // forward FFNT root table (L entries)
    void* iroot;
    void* twist;
    void* untwist;
    void* wc;        // stage root constants and their conjugates (32 each)
    void* wci;
    void* scratch;
    size_t scratch_bytes;
    int* arrive;     // per-product arrival counters (max_batch), kept at 0 between launches
};

namespace {

constexpr int F64_MAX_S = 13;  // 2^13 doubles complex = 128 KiB of LDS
constexpr int F32_MAX_S = 14;  // 2^14 floats complex = 128 KiB of LDS
constexpr int MAX_SB = 3;
#ifndef GFA_E64
#define GFA_E64 4
#endif



template <typename T> using ProdKernel = void (*)(ProdArgs<T>);

// Zero fill by a kernel rather than hipMemsetAsync, so that a caller may
// capture the call into a HIP graph.
__global__ void zero_kernel(double* p, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void fill_ones(float* p, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 1.0f;
}

__global__ void zero_ints(int* p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Whether the product kernel exists for a sub-problem of 2^S points, 2^E
// points per thread and radix-2^RB steps: the thread's groups of V loaded
// positions must hold the 2^(S mod RB) partners of the register top step.
template <typename T, int S, int E, int PL, int RB> constexpr bool prod_ok()
{
    constexpr int HB = S % RB;
    constexpr int V = in_vec<T, PL, E, HB>();
    return RB <= E && S > E && ((1 << E) / V) / (1 << HB) >= 1;
}

// 2^E points per thread: 2^(S-E) threads per workgroup
template <typename T, int PL, int RB, int E> ProdKernel<T> prod_kernel_rb(int S, int SB, int cmb)
{
    constexpr int SMAX = sizeof(T) == 8 ? F64_MAX_S : F32_MAX_S;
    if constexpr (sizeof(T) == 8 && PL == PL_FOLD) {
        if (cmb && S == SMAX) {
            if (SB == 1) return &product_sub_kernel<T, SMAX, 1, E, PL, RB, true>;
            if (SB == 2) return &product_sub_kernel<T, SMAX, 2, E, PL, RB, true>;
            if (SB == 3) return &product_sub_kernel<T, SMAX, 3, E, PL, RB, true>;
        }
    }
#define GFA_PK(SS, BB)                                                 \
    if constexpr (prod_ok<T, SS, E, PL, RB>())                         \
        if (S == SS && SB == BB) return &product_sub_kernel<T, SS, BB, E, PL, RB>;
    if (SB > 0) {
        GFA_PK(SMAX, 1)
        GFA_PK(SMAX, 2)
        GFA_PK(SMAX, 3)
    } else {
        GFA_PK(6, 0)
        GFA_PK(7, 0)
        GFA_PK(8, 0)
        GFA_PK(9, 0)
        GFA_PK(10, 0)
        GFA_PK(11, 0)
        GFA_PK(12, 0)
        GFA_PK(13, 0)
        if constexpr (SMAX >= 14) {
            GFA_PK(14, 0)
        }
    }
#undef GFA_PK
    return nullptr;
}

// radix bits per LDS step: 3 or 4 (double), 3..5 (float); points per thread
// 2^e: 16 (double), 32 or 16 (float); both chosen per plan
template <typename T, int PL, int E> ProdKernel<T> prod_kernel_e(int S, int SB, int rb, int cmb)
{
    switch (rb) {
        case 3: return prod_kernel_rb<T, PL, 3, E>(S, SB, cmb);
        case 4: return prod_kernel_rb<T, PL, 4, E>(S, SB, cmb);
        case 5:
            if constexpr (sizeof(T) == 4 && E >= 5) return prod_kernel_rb<T, PL, 5, E>(S, SB, cmb);
            return nullptr;
        default: return nullptr;
    }
}

template <typename T, int PL> ProdKernel<T> prod_kernel(int S, int SB, int rb, int cmb, int e)
{
    if constexpr (sizeof(T) == 8) {
        return e == GFA_E64 ? prod_kernel_e<T, PL, GFA_E64>(S, SB, rb, cmb) : nullptr;
    } else {
        if (e == 5) return prod_kernel_e<T, PL, 5>(S, SB, rb, cmb);
        if (e == 4) return prod_kernel_e<T, PL, 4>(S, SB, rb, cmb);
        return nullptr;
    }
}

template <typename T> int make_tables(gpufft_polymul_plan_s* pl, hipStream_t st)
{
    size_t bytes = sizeof(cplx<T>) << pl->n;
    void** dst[4] = {&pl->root, &pl->iroot, &pl->twist, &pl->untwist};
    for (void** d : dst) {
        hipError_t e = hipMalloc(d, bytes);
        if (e != hipSuccess) return fail_hip(e, "gpufft_polymul_plan_create: hipMalloc");
    }
    for (void** d : {&pl->wc, &pl->wci}) {
        hipError_t e = hipMalloc(d, sizeof(cplx<T>) * 32);
        if (e != hipSuccess) return fail_hip(e, "gpufft_polymul_plan_create: hipMalloc");
    }
    int (*gen)(void*, int, int, int, void*) =
        sizeof(T) == 8 ? gpufft_generate_table_f64 : gpufft_generate_table_f32;
    int s;
    if ((s = gen(pl->root, GPUFFT_TABLE_FFNT_ROOT, pl->n, 0, st))) return s;
    if ((s = gen(pl->iroot, GPUFFT_TABLE_FFNT_ROOT, pl->n, 1, st))) return s;
    if ((s = gen(pl->twist, GPUFFT_TABLE_FFNT_TWIST, pl->n, 0, st))) return s;
    if ((s = gen(pl->untwist, GPUFFT_TABLE_FFNT_TWIST, pl->n, 1, st))) return s;
    if ((s = gen(pl->wc, GPUFFT_TABLE_STAGE_ROOTS, 5, 0, st))) return s;
    if ((s = gen(pl->wci, GPUFFT_TABLE_STAGE_ROOTS, 5, 1, st))) return s;
    return GPUFFT_OK;
}

void release(gpufft_polymul_plan_s* pl)
{
    if (!pl) return;
    void* ptrs[9] = {pl->root, pl->iroot, pl->twist, pl->untwist, pl->wc, pl->wci, pl->scratch,
                     pl->arrive, pl->jspec};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete pl;
}

template <typename T, int PL>
int launch_products(const gpufft_polymul_plan_s* pl, const void* a, const void* b, T* out_direct,
                    int batch, int cmb, hipStream_t st, const char* who)
{
    ProdArgs<T> pa{};
    pa.n = pl->n;
    pa.S = pl->S;
    pa.batch = batch;
    pa.nega = 0;
    pa.tab = (const cplx<T>*)pl->root;
    pa.itab = (const cplx<T>*)pl->iroot;
    pa.twist = (const cplx<T>*)pl->twist;
    pa.untwist = (const cplx<T>*)pl->untwist;
    pa.wc = (const cplx<T>*)pl->wc;
    pa.wci = (const cplx<T>*)pl->wci;
    pa.a = a;
    pa.b = b;
    pa.in_stride = 2LL << pl->n;
    pa.scratch = (cplx<T>*)pl->scratch;
    pa.out = out_direct;
    pa.out_stride = 2LL << pl->n;
    pa.scale = mk<T>((T)(1.0 / (double)(1LL << pl->n)), (T)0);
    pa.wa = pl->wa;
    pa.ka = pl->ka;
    pa.ca = pl->ca;
    pa.jspec = (const cplx<T>*)pl->jspec;
    pa.wb = pl->wb;
    pa.kb = pl->kb;
    auto offset = [](int w, int k) {
        long long o = 0;
        for (int j = 0; j + 1 < k; j++) o += (1LL << (w - 1)) << (w * j);
        return o;
    };
    pa.oa = pl->ka > 1 ? offset(pl->wa, pl->ka) : 0;
    pa.ob = pl->kb > 1 ? offset(pl->wb, pl->kb) : 0;
    pa.arrive = pl->arrive;
    ProdKernel<T> k = prod_kernel<T, PL>(pl->S, pl->SB, pl->rb, cmb, pl->e);
    if (!k) return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": no kernel for this size");
    const int threads = (1 << pl->S) >> pl->e;
    const long long groups = ((long long)batch + 7) / 8;
    const long long blocks = groups * 8 * (1LL << pl->SB);
    hipLaunchKernelGGL(k, dim3((unsigned)blocks), dim3(threads), 0, st, pa);
    GFA_CHECK_LAUNCH(who);
    return GPUFFT_OK;
}

int check_plan(const gpufft_polymul_plan_s* pl, int precision, int batch, const char* who)
{
    if (!pl) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (pl->precision != precision)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": plan precision mismatch");
    if (batch < 1) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch must be >= 1");
    return GPUFFT_OK;
}

// FP64 products of a validated plan (or of the no-allocation entry point's
// stack plan), in launches of at most max_batch products.
int run_f64(const gpufft_polymul_plan_s* plan, const double* a, const double* b, double* out,
            int batch, int phases, hipStream_t st, const char* who)
{
    const long long N = 2LL << plan->n;
    for (int b0 = 0; b0 < batch; b0 += plan->max_batch) {
        const int nb = std::min(plan->max_batch, batch - b0);
        const double* pa = a + b0 * N;
        const double* pb = b + b0 * N;
        double* po = out + b0 * N;
        // in-kernel combine: one launch does both phases (PHASE_TRANSFORM
        // runs it, PHASE_COMBINE alone does nothing)
        const int cmb = plan->cmb && plan->SB > 0;
        if (phases & GPUFFT_PHASE_TRANSFORM)
            if (int s = launch_products<double, PL_FOLD>(plan, pa, pb, po, nb, cmb, st, who)) return s;
        if (plan->SB > 0 && !cmb && (phases & GPUFFT_PHASE_COMBINE)) {
            // top SB inverse stages + scale + untwist + unfold
            F64CombineArgs ca{};
            ca.n = plan->n;
            ca.S = plan->S;
            ca.batch = nb;
            ca.scratch = (const cplx<double>*)plan->scratch;
            ca.itab = (const cplx<double>*)plan->iroot;
            ca.untwist = (const cplx<double>*)plan->untwist;
            ca.scale = 1.0 / (double)(1LL << plan->n);
            ca.out = po;
            ca.out_stride = N;
            // 512 columns per workgroup, every workgroup inside one product
            if (plan->S < 9) return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": tile too small");
            const unsigned blocks = (unsigned)((long long)nb << (plan->S - 9));
            switch (plan->SB) {
                case 1: hipLaunchKernelGGL(f64_combine_kernel<1>, dim3(blocks), dim3(256), 0, st, ca); break;
                case 2: hipLaunchKernelGGL(f64_combine_kernel<2>, dim3(blocks), dim3(256), 0, st, ca); break;
                case 3: hipLaunchKernelGGL(f64_combine_kernel<3>, dim3(blocks), dim3(256), 0, st, ca); break;
                default: return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": split too deep");
            }
            GFA_CHECK_LAUNCH(who);
        }
    }
    return ok();
}

// ---- products without allocation ----
// Static device scratch of gpufft_polymul_f64_noalloc: the stage root
// constants (64 complex doubles in the first 1 KiB, written by every call)
// and the sub-problem results of products split over several workgroups.
// Each device holds its own copy from module load until the process ends;
// calls that overlap in time on one device share it and race.
constexpr size_t NOALLOC_BYTES = (size_t)64 << 20;
constexpr size_t NOALLOC_CONST_BYTES = 1024;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device

// Device address of this device's `scratch`, looked up once per device with
// the device current (the host-side name is only a stand-in).
int noalloc_scratch(int dev, void** out, const char* who)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": device index out of range");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, who);
        }
        noalloc_ptr[dev] = p;
    }
    *out = noalloc_ptr[dev];
    return GPUFFT_OK;
}

}  // namespace

extern "C" {

int gpufft_polymul_plan_create(gpufft_polymul_plan* out, int ring_power, int precision,
                               int max_batch, int a_limb_bits, int a_limbs, int b_limb_bits,
                               int b_limbs, void* stream)
{
    const char* who = "gpufft_polymul_plan_create";
    if (!out) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": out is NULL");
    *out = nullptr;
    if (precision != GPUFFT_PREC_F64 && precision != GPUFFT_PREC_F32_SPLIT)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": unknown precision");
    const int E = precision == GPUFFT_PREC_F64 ? 4 : 5;
    const int smax = precision == GPUFFT_PREC_F64 ? F64_MAX_S : F32_MAX_S;
    const int n = ring_power - 1;
    if (n < E + 2 || n > smax + MAX_SB)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT,
                    std::string(who) + ": ring_power out of the fused range");
    if (max_batch < 1) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": max_batch < 1");
    if (precision == GPUFFT_PREC_F32_SPLIT) {
        if (a_limbs < 1 || b_limbs < 1 || a_limbs * b_limbs > 64 ||
            (a_limbs > 1 && (a_limb_bits < 1 || a_limb_bits > 16)) ||
            (b_limbs > 1 && (b_limb_bits < 1 || b_limb_bits > 16)) ||
            (a_limbs > 1 && (a_limbs - 1) * a_limb_bits > 30) ||
            (b_limbs > 1 && (b_limbs - 1) * b_limb_bits > 30))
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": invalid limb split");
    }
    gpufft_polymul_plan_s* pl = new gpufft_polymul_plan_s();
    (void)hipGetDevice(&pl->device);
    pl->ring_power = ring_power;
    pl->n = n;
    pl->precision = precision;
    pl->max_batch = max_batch;
    if (precision == GPUFFT_PREC_F32_SPLIT) {
        pl->wa = a_limbs > 1 ? a_limb_bits : 0;
        pl->ka = a_limbs;
        pl->wb = b_limbs > 1 ? b_limb_bits : 0;
        pl->kb = b_limbs;
    } else {
        pl->ka = pl->kb = 1;
    }
    pl->S = std::min(n, smax);
    pl->SB = n - pl->S;
    pl->rb = 4;
    // FP32: 1024 threads of 16 points per full 2^14-point tile (measured
    // faster than 512 threads of 32 at N = 2^16: 0.38 against 0.45 ms per
    // batch of 256), 512 threads of 32 points below the full tile
    // FP32: 16 points per thread (1024 threads) for full tiles of products
    // in one or two sub-problems; with more, the top-row gather's partner
    // loads spill at 128 VGPRs and 32 points per thread (256 VGPRs) is faster
    pl->e = precision == GPUFFT_PREC_F64 ? GFA_E64 : (pl->S == F32_MAX_S && pl->SB <= 1 ? 4 : 5);
    pl->cmb = 0;  // separate pass: measured faster at N = 2^16, batch 256
    hipStream_t st = (hipStream_t)stream;
    int s = precision == GPUFFT_PREC_F64 ? make_tables<double>(pl, st) : make_tables<float>(pl, st);
    if (s) {
        release(pl);
        return s;
    }
    const bool need_scratch = precision == GPUFFT_PREC_F32_SPLIT || pl->SB > 0;
    if (need_scratch) {
        size_t elem = precision == GPUFFT_PREC_F64 ? 16 : 8;
        // FP32 split: room for the optional J limb (gpufft_polymul_plan_set_a_center)
        const size_t limbs = precision == GPUFFT_PREC_F32_SPLIT ? (size_t)(pl->ka + 1) * pl->kb : 1;
        pl->scratch_bytes = (size_t)max_batch * limbs * (elem << n);
        hipError_t e = hipMalloc(&pl->scratch, pl->scratch_bytes);
        if (e != hipSuccess) {
            release(pl);
            return fail_hip(e, "gpufft_polymul_plan_create: hipMalloc scratch");
        }
    }
    if (precision == GPUFFT_PREC_F64 && pl->SB > 0) {
        hipError_t e = hipMalloc((void**)&pl->arrive, sizeof(int) * (size_t)max_batch);
        if (e != hipSuccess) {
            release(pl);
            return fail_hip(e, "gpufft_polymul_plan_create: hipMalloc counters");
        }
        hipLaunchKernelGGL(zero_ints, dim3((unsigned)((max_batch + 255) / 256)), dim3(256), 0, st,
                           pl->arrive, max_batch);
    }
    // tables and counters are ready before the plan is used on any stream
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        release(pl);
        return fail_hip(e, "gpufft_polymul_plan_create: table generation");
    }
    *out = pl;
    return ok();
}

int gpufft_polymul_plan_destroy(gpufft_polymul_plan plan)
{
    release(plan);
    return GPUFFT_OK;
}

int gpufft_polymul_f64(gpufft_polymul_plan plan, const double* a, const double* b, double* out,
                       int batch, void* stream)
{
    return gpufft_polymul_f64_phases(plan, a, b, out, batch, GPUFFT_PHASE_ALL, stream);
}

int gpufft_polymul_f64_phases(gpufft_polymul_plan plan, const double* a, const double* b,
                              double* out, int batch, int phases, void* stream)
{
    const char* who = "gpufft_polymul_f64";
    if (int s = check_plan(plan, GPUFFT_PREC_F64, batch, who)) return s;
    if (!a || !b || !out) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    return run_f64(plan, a, b, out, batch, phases, (hipStream_t)stream, who);
}

int gpufft_polymul_i32(gpufft_polymul_plan plan, const int32_t* a, const int32_t* b, int64_t* out,
                       int batch, double* err_stats, void* stream)
{
    return gpufft_polymul_i32_phases(plan, a, b, out, batch, err_stats, GPUFFT_PHASE_ALL, stream);
}

int gpufft_polymul_i32_phases(gpufft_polymul_plan plan, const int32_t* a, const int32_t* b,
                              int64_t* out, int batch, double* err_stats, int phases, void* stream)
{
    const char* who = "gpufft_polymul_i32";
    if (int s = check_plan(plan, GPUFFT_PREC_F32_SPLIT, batch, who)) return s;
    if (!a || !b || !out) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    hipStream_t st = (hipStream_t)stream;
    const long long N = 2LL << plan->n;
    if (err_stats && (phases & GPUFFT_PHASE_COMBINE)) {
        const long long cnt = 2LL * batch;
        hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, st,
                           err_stats, cnt);
        GFA_CHECK_LAUNCH(who);
    }
    for (int b0 = 0; b0 < batch; b0 += plan->max_batch) {
        const int nb = std::min(plan->max_batch, batch - b0);
        if (phases & GPUFFT_PHASE_TRANSFORM)
            if (int s = launch_products<float, PL_FOLD_LIMB>(plan, a + b0 * N, b + b0 * N, nullptr,
                                                             nb, 0, st, who))
                return s;
        if (!(phases & GPUFFT_PHASE_COMBINE)) continue;
        const long long nthreads_total = (long long)nb << plan->S;
        if ((nthreads_total + 255) / 256 > 0x7fffffffLL)
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch too large");
        const unsigned blocks = (unsigned)((nthreads_total + 255) / 256);
#define GFA_COMB(BB)                                                                       \
    case BB: {                                                                             \
        CombineArgs<BB> ca{};                                                              \
        ca.n = plan->n;                                                                    \
        ca.S = plan->S;                                                                    \
        ca.batch = nb;                                                                     \
        ca.nprod = (plan->ka + (plan->ca ? 1 : 0)) * plan->kb;                             \
        ca.ka = plan->ka;                                                                  \
        ca.ca = plan->ca;                                                                  \
        ca.kb = plan->kb;                                                                  \
        ca.wa = plan->wa;                                                                  \
        ca.wb = plan->wb;                                                                  \
        ca.scratch = (const cplx<float>*)plan->scratch;                                    \
        ca.itab = (const cplx<float>*)plan->iroot;                                         \
        ca.untwist = (const cplx<float>*)plan->untwist;                                    \
        ca.scale = (float)(1.0 / (double)(1LL << plan->n));                                \
        ca.out = (long long*)(out + b0 * N);                                               \
        ca.out_stride = N;                                                                 \
        ca.err = err_stats ? err_stats + 2 * b0 : nullptr;                                 \
        hipLaunchKernelGGL(limb_combine_kernel<BB>, dim3(blocks), dim3(256), 0, st, ca);   \
        break;                                                                             \
    }
        switch (plan->SB) {
            GFA_COMB(0)
            GFA_COMB(1)
            GFA_COMB(2)
            GFA_COMB(3)
            default: return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": split too deep");
        }
#undef GFA_COMB
        GFA_CHECK_LAUNCH(who);
    }
    return ok();
}

size_t gpufft_polymul_plan_scratch_bytes(gpufft_polymul_plan plan)
{
    return plan ? plan->scratch_bytes : 0;
}

int gpufft_polymul_plan_set_radix_bits(gpufft_polymul_plan plan, int radix_bits)
{
    const char* who = "gpufft_polymul_plan_set_radix_bits";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    const int hi = plan->precision == GPUFFT_PREC_F64 ? 4 : 5;
    if (radix_bits < 3 || radix_bits > hi)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": radix_bits out of range");
    if (plan->precision == GPUFFT_PREC_F32_SPLIT &&
        !prod_kernel<float, PL_FOLD_LIMB>(plan->S, plan->SB, radix_bits, 0, plan->e))
        return fail(GPUFFT_ERR_UNSUPPORTED,
                    std::string(who) + ": no kernel for this radix at the plan's points per thread");
    plan->rb = radix_bits;
    return ok();
}

int gpufft_polymul_plan_set_thread_points(gpufft_polymul_plan plan, int log2_points)
{
    const char* who = "gpufft_polymul_plan_set_thread_points";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    const bool okv = plan->precision == GPUFFT_PREC_F64 ? log2_points == GFA_E64
                                                        : (log2_points == 4 || log2_points == 5);
    if (!okv) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": unsupported points per thread");
    if (!prod_kernel<float, PL_FOLD_LIMB>(plan->S, plan->SB, plan->rb, 0, log2_points) &&
        plan->precision == GPUFFT_PREC_F32_SPLIT)
        return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": no kernel for this size and radix");
    plan->e = log2_points;
    return ok();
}

int gpufft_polymul_plan_get_thread_points(gpufft_polymul_plan plan)
{
    return plan ? plan->e : -1;
}

int gpufft_polymul_plan_get_radix_bits(gpufft_polymul_plan plan)
{
    return plan ? plan->rb : -1;
}

int gpufft_polymul_plan_set_combine(gpufft_polymul_plan plan, int in_kernel)
{
    const char* who = "gpufft_polymul_plan_set_combine";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (in_kernel != 0 && in_kernel != 1)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": in_kernel must be 0 or 1");
    // only FP64 plans with more than one workgroup per product have the choice
    plan->cmb = plan->precision == GPUFFT_PREC_F64 && plan->SB > 0 ? in_kernel : 0;
    return ok();
}

int gpufft_polymul_plan_get_combine(gpufft_polymul_plan plan)
{
    return plan ? plan->cmb : -1;
}

int gpufft_polymul_plan_set_a_center(gpufft_polymul_plan plan, int center_log2)
{
    const char* who = "gpufft_polymul_plan_set_a_center";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (plan->precision != GPUFFT_PREC_F32_SPLIT)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": only GPUFFT_PREC_F32_SPLIT plans");
    if (center_log2 < 0 || center_log2 > 30)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": center_log2 must be 0..30");
    if (center_log2 > 0 && !plan->jspec) {
        // spectrum of J (all ones), once per plan: the GPU_FFNT forward of the
        // plan's own tables, so it is the transform the product kernel applies
        const long long N = 2LL << plan->n, L = 1LL << plan->n;
        hipError_t e = hipMalloc(&plan->jspec, sizeof(cplx<float>) * (size_t)L);
        if (e != hipSuccess) return fail_hip(e, "gpufft_polymul_plan_set_a_center: hipMalloc");
        float* ones = nullptr;
        e = hipMalloc((void**)&ones, sizeof(float) * (size_t)N);
        if (e != hipSuccess) {
            (void)hipFree(plan->jspec);
            plan->jspec = nullptr;
            return fail_hip(e, "gpufft_polymul_plan_set_a_center: hipMalloc");
        }
        hipLaunchKernelGGL(fill_ones, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, 0, ones, N);
        gpufft_config cfg{};
        cfg.n_power = plan->ring_power;
        cfg.fft_type = GPUFFT_FORWARD;
        cfg.stream = nullptr;
        int st = gpufft_ffnt_f32(ones, plan->jspec, plan->twist, plan->root, &cfg, 1, 0);
        hipError_t e2 = hipDeviceSynchronize();
        (void)hipFree(ones);
        if (st) return st;
        if (e2 != hipSuccess) return fail_hip(e2, "gpufft_polymul_plan_set_a_center: J spectrum");
    }
    plan->ca = center_log2;
    return ok();
}

int gpufft_polymul_plan_get_a_center(gpufft_polymul_plan plan)
{
    return plan ? plan->ca : -1;
}

int gpufft_polymul_f64_noalloc(const double* a, const double* b, double* out, int ring_power,
                               int batch, const void* twist, const void* untwist,
                               const void* root_table, const void* inv_root_table, void* stream)
{
    const char* who = "gpufft_polymul_f64_noalloc";
    const int n = ring_power - 1;
    if (n < 4 + 2 || n > F64_MAX_S + MAX_SB)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": ring_power must be 7..17");
    if (batch < 1) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch must be >= 1");
    if (!a || !b || !out || !twist || !untwist || !root_table || !inv_root_table)
        return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_hip(e, who);
    void* base = nullptr;
    if (int s = noalloc_scratch(dev, &base, who)) return s;
    hipStream_t st = (hipStream_t)stream;
    gpufft_polymul_plan_s pl{};
    pl.device = dev;
    pl.ring_power = ring_power;
    pl.n = n;
    pl.precision = GPUFFT_PREC_F64;
    pl.ka = pl.kb = 1;
    pl.S = std::min(n, F64_MAX_S);
    pl.SB = n - pl.S;
    pl.rb = 4;
    pl.e = GFA_E64;
    pl.cmb = 0;
    pl.root = const_cast<void*>(root_table);
    pl.iroot = const_cast<void*>(inv_root_table);
    pl.twist = const_cast<void*>(twist);
    pl.untwist = const_cast<void*>(untwist);
    cplx<double>* consts = (cplx<double>*)base;
    pl.wc = consts;
    pl.wci = consts + 32;
    // the sub-problem results of one launch must fit the rest of the array:
    // the batch runs in launches of at most that many products
    const size_t per_product = pl.SB > 0 ? (sizeof(cplx<double>) << n) : 0;
    const size_t room = NOALLOC_BYTES - NOALLOC_CONST_BYTES;
    pl.max_batch = per_product ? (int)std::min<size_t>(room / per_product, (size_t)batch) : batch;
    pl.scratch = pl.SB > 0 ? (void*)((unsigned char*)base + NOALLOC_CONST_BYTES) : nullptr;
    pl.scratch_bytes = pl.SB > 0 ? room : 0;
    // stage root constants: rewritten by every call (a kernel, so the call
    // can be captured into a graph), ahead of the product kernels on `st`
    if (int s = gpufft_generate_table_f64(pl.wc, GPUFFT_TABLE_STAGE_ROOTS, 5, 0, st)) return s;
    if (int s = gpufft_generate_table_f64(pl.wci, GPUFFT_TABLE_STAGE_ROOTS, 5, 1, st)) return s;
    return run_f64(&pl, a, b, out, batch, GPUFFT_PHASE_ALL, st, who);
}

}  // extern "C"
