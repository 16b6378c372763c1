// Batched products of uneven ring sizes in one persistent launch (FP32 limb
// split, int32 in / int64 out).  No counterpart in the reference, which runs
// one GPU_FFT / GPU_FFNT call pair per ring size.
//
// Every product is cut into 2^SB sub-problems of 2^S points (S = min(log2 L,
// 14), L the complex transform length) as in fused.hpp; the work list holds
// one item per sub-problem, largest products first.  A persistent grid of
// workgroups takes items from a device counter until the list is exhausted
// (every workgroup reaches that exit).  An item with SB = 0 is a whole
// product: it rounds each limb product and accumulates it into the output.
// An item with SB > 0 writes its limb sub-products to scratch, publishes them
// (agent-scope release), and increments the product's arrival counter; the
// workgroup that arrives last (acquire) runs the top SB inverse stages, the
// output transform and the rounding for the whole product.  Nothing waits
// on another workgroup.
//
// Product kinds:
//   NEGACYCLIC  c = a b mod X^N + 1 (FFNT: L = N/2, fold + twist; reference
//               GPU_FFNT, src/lib/fft.cu:1013-1384)
//   LINEAR      c = a b, 2N - 1 coefficients (zero-padded cyclic C_C product
//               of length L = 2N; reference example/test_fft_gpu_C_C.cu)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "fused.hpp"
#include "status.hpp"

using namespace gfa;

namespace {

constexpr int KIND_NEGA = 0;
constexpr int KIND_LINEAR = 1;
constexpr int BT = 512;     // threads of the persistent workgroup
constexpr int SMAX = 14;    // 2^14 float complex = 128 KiB of LDS
constexpr int SMIN = 6;
constexpr int SBMAX = 4;

struct BProd {
    int kind;
    int n;        // log2 L
    int S, SB;
    int N;        // coefficients per operand
    int tabset;   // index of the table set for (kind, n)
    long long a_off, b_off, out_off;
    long long scratch_off;  // cplx<float> offset (SB > 0)
};

struct BTables {
    const cplx<float>* root;
    const cplx<float>* iroot;
    const cplx<float>* twist;    // NEGACYCLIC only
    const cplx<float>* untwist;  // NEGACYCLIC only
};

struct BArgs {
    const BProd* prods;
    const int2* items;  // (product, q)
    const BTables* tabs;
    int n_items;
    int* next_item;     // work counter
    int* arrive;        // per product
    const int* a;
    const int* b;
    long long* out;
    cplx<float>* scratch;
    const cplx<float>* wc;
    const cplx<float>* wci;
    int wa, ka, wb, kb;
    long long oa, ob;
    double* err;        // optional [products]
};

__global__ void zero_ints(int* p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void zero_doubles(double* p, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__device__ __forceinline__ void err_update(double* err, int prod, float e)
{
    if (!err) return;
    double d = e;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) d = fmax(d, __shfl_xor(d, off, 64));
    if ((threadIdx.x & 63) == 0)
        atomicMax(reinterpret_cast<unsigned long long*>(err + prod),
                  (unsigned long long)__double_as_longlong(d));
}

// Input at position P of the complex transform of one operand limb.
template <int KIND>
__device__ __forceinline__ void item_input(const int* x, int N, int L, int w, int limb, int k,
                                           long long o, int P, float& re, float& im)
{
    if constexpr (KIND == KIND_NEGA) {
        re = (float)limb_digit(x[P], w, limb, k, o);
        im = (float)limb_digit(x[P + L], w, limb, k, o);
    } else {
        re = P < N ? (float)limb_digit(x[P], w, limb, k, o) : 0.f;
        im = 0.f;
    }
}

template <int KIND, int S, int SB, int E>
__device__ __forceinline__ void item_load(cplx<float>* lds, const ProdArgs<float>& pa, const int* x,
                                          int N, int w, int limb, int k, long long o,
                                          const cplx<float>* C, int tid)
{
    constexpr int NT = (1 << S) >> E;
    constexpr int Q = 1 << SB;
    const int L = 1 << pa.n;
    if (tid >= NT) return;
#pragma unroll 2
    for (int i = 0; i < (1 << E); i++) {
        const int c = tid + i * NT;
        float ax = 0.f, ay = 0.f;
#pragma unroll
        for (int j = 0; j < Q; j++) {
            float re, im;
            item_input<KIND>(x, N, L, w, limb, k, o, (j << S) | c, re, im);
            ax = ax + C[j].x * re - C[j].y * im;
            ay = ay + C[j].x * im + C[j].y * re;
        }
        cplx<float> y = mk<float>(ax, ay);
        if constexpr (KIND == KIND_NEGA) y = pa.twist[c] * y;
        lds[swz(c)] = y;
    }
}

// Whole product or one sub-problem of it.
template <int KIND, int S, int SB, int E, int RB>
__device__ void run_item(cplx<float>* lds, const BArgs& ba, const BProd& pr, const BTables& tb,
                         int q, int tid, int* sh_last)
{
    constexpr int EPT = 1 << E;
    constexpr int NT = (1 << S) >> E;
    constexpr int Q = 1 << SB;
    constexpr int NST = (S + RB - 1) / RB;
    const bool active = tid < NT;
    const int L = 1 << pr.n;

    ProdArgs<float> pa{};
    pa.n = pr.n;
    pa.S = S;
    pa.nega = 0;
    pa.tab = tb.root;
    pa.itab = tb.iroot;
    pa.twist = tb.twist;
    PassArgs<float> pf{};
    pf.n = pr.n;
    pf.K = S;
    pf.tab = tb.root;
    PassArgs<float> pinv = pf;
    pinv.tab = tb.iroot;
    TileGeo g;
    g.hi_pos = q << S;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = make_rsrc(tb.root);
    TileGeo gi = g;
    gi.tab = make_rsrc(tb.iroot);

    // coefficients of row q of the top network (twist factored for FFNT)
    cplx<float> C[Q];
    {
        cplx<float> wq[SB > 0 ? SB : 1];
        top_row_twiddles<float, SB>(pa, q, wq);
#pragma unroll
        for (int j = 0; j < Q; j++) {
            cplx<float> e[Q];
#pragma unroll
            for (int m = 0; m < Q; m++) e[m] = mk<float>(0.f, 0.f);
            e[j] = KIND == KIND_NEGA ? tb.twist[j << S] : mk<float>(1.f, 0.f);
            C[j] = top_forward_row<float, SB>(e, wq);
        }
    }
    const int* xa = ba.a + pr.a_off;
    const int* xb = ba.b + pr.b_off;
    const int nprod = ba.ka * ba.kb;
    const float scale = 1.0f / (float)L;
    float emax = 0.f;

    cplx<float> breg[EPT];
    for (int jb = 0; jb < ba.kb; jb++) {
        int t = tid;
        asm volatile("" : "+v"(t));
        __syncthreads();
        item_load<KIND, S, SB, E>(lds, pa, xb, pr.N, ba.wb, jb, ba.kb, ba.ob, C, t);
        run_steps_ct_active<float, MODE_STD_FWD, E, NT, S, RB, 0, NST - 1>(lds, pf, g, t, ba.wc, active);
        if (active) {
#pragma unroll
            for (int kk = 0; kk < (1 << (E - RB)); kk++) {
                const int rbase = (t + kk * NT) << RB;
                cplx<float> v[1 << RB];
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) v[j] = lds[swz(rbase | j)];
                group_stages_derived<float, MODE_STD_FWD, RB>(v, pf, g, rbase, 0, 0, ba.wc);
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) breg[(kk << RB) + j] = v[j];
            }
        }
        for (int ia = 0; ia < ba.ka; ia++) {
            int u = tid;
            asm volatile("" : "+v"(u));
            __syncthreads();
            item_load<KIND, S, SB, E>(lds, pa, xa, pr.N, ba.wa, ia, ba.ka, ba.oa, C, u);
            run_steps_ct_active<float, MODE_STD_FWD, E, NT, S, RB, 0, NST - 1>(lds, pf, g, u, ba.wc,
                                                                           active);
            if (active) {
#pragma unroll
                for (int kk = 0; kk < (1 << (E - RB)); kk++) {
                    const int rbase = (u + kk * NT) << RB;
                    cplx<float> v[1 << RB];
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) v[j] = lds[swz(rbase | j)];
                    group_stages_derived<float, MODE_STD_FWD, RB>(v, pf, g, rbase, 0, 0, ba.wc);
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) v[j] = v[j] * breg[(kk << RB) + j];
                    group_stages_derived<float, MODE_STD_INV, RB>(v, pinv, gi, rbase, 0, 0, ba.wci);
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) lds[swz(rbase | j)] = v[j];
                }
            }
            run_steps_ct_active<float, MODE_STD_INV, E, NT, S, RB, 1, NST>(lds, pinv, gi, u, ba.wci,
                                                                         active);
            const int limb = ia * ba.kb + jb;
            if constexpr (SB == 0) {
                // the whole product: round this limb product and accumulate
                if (active) {
                    const long long wgt = 1LL << (ba.wa * ia + ba.wb * jb);
                    long long* o = ba.out + pr.out_off;
#pragma unroll 4
                    for (int i = 0; i < EPT; i++) {
                        const int c = u + i * NT;
                        cplx<float> z = lds[swz(c)];
                        z = mk<float>(z.x * scale, z.y * scale);
                        if constexpr (KIND == KIND_NEGA) {
                            z = z * tb.untwist[c];
                            const float r0 = rintf(z.x), r1 = rintf(z.y);
                            emax = fmaxf(emax, fmaxf(fabsf(z.x - r0), fabsf(z.y - r1)));
                            const long long v0 = (long long)r0 * wgt, v1 = (long long)r1 * wgt;
                            if (limb == 0) {
                                o[c] = v0;
                                o[c + L] = v1;
                            } else {
                                o[c] += v0;
                                o[c + L] += v1;
                            }
                        } else {
                            const float r0 = rintf(z.x);
                            emax = fmaxf(emax, fabsf(z.x - r0));
                            const long long v0 = (long long)r0 * wgt;
                            o[c] = limb == 0 ? v0 : o[c] + v0;
                        }
                    }
                }
            } else {
                if (active) {
                    cplx<float>* dst = ba.scratch + pr.scratch_off + (long long)limb * L +
                                       ((long long)q << S);
#pragma unroll 4
                    for (int i = 0; i < EPT; i++) {
                        const int c = u + i * NT;
                        stc_nt(dst + c, lds[swz(c)]);
                    }
                }
            }
        }
    }
    if constexpr (SB == 0) {
        err_update(ba.err, (int)(&pr - ba.prods), emax);
    } else {
        // publish this sub-problem; the last of the product's Q workgroups
        // combines them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            const int prod = (int)(&pr - ba.prods);
            const int prev = __hip_atomic_fetch_add(ba.arrive + prod, 1, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
            *sh_last = (prev == Q - 1) ? 1 : 0;
        }
        __syncthreads();
        if (*sh_last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            long long* o = ba.out + pr.out_off;
            float em = 0.f;
            for (int c = tid; c < (1 << S); c += BT) {
                long long acc[Q][2];
#pragma unroll
                for (int j = 0; j < Q; j++) acc[j][0] = acc[j][1] = 0;
                for (int l = 0; l < nprod; l++) {
                    const cplx<float>* src = ba.scratch + pr.scratch_off + (long long)l * L;
                    cplx<float> y[Q];
#pragma unroll
                    for (int j = 0; j < Q; j++) y[j] = src[(j << S) + c];
#pragma unroll
                    for (int i = 0; i < SB; i++) {  // Gentleman-Sande, bits S .. n-1
#pragma unroll
                        for (int j0 = 0; j0 < Q; j0++) {
                            if (j0 & (1 << i)) continue;
                            const int j1 = j0 | (1 << i);
                            const cplx<float> w = tb.iroot[j0 >> (i + 1)];
                            cplx<float> U = y[j0], V = y[j1];
                            y[j0] = U + V;
                            y[j1] = (U - V) * w;
                        }
                    }
                    const long long wgt = 1LL << (ba.wa * (l / ba.kb) + ba.wb * (l % ba.kb));
#pragma unroll
                    for (int j = 0; j < Q; j++) {
                        const int P = (j << S) | c;
                        cplx<float> z = mk<float>(y[j].x * scale, y[j].y * scale);
                        if constexpr (KIND == KIND_NEGA) {
                            z = z * tb.untwist[P];
                            const float r0 = rintf(z.x), r1 = rintf(z.y);
                            em = fmaxf(em, fmaxf(fabsf(z.x - r0), fabsf(z.y - r1)));
                            acc[j][0] += (long long)r0 * wgt;
                            acc[j][1] += (long long)r1 * wgt;
                        } else {
                            const float r0 = rintf(z.x);
                            em = fmaxf(em, fabsf(z.x - r0));
                            acc[j][0] += (long long)r0 * wgt;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < Q; j++) {
                    const int P = (j << S) | c;
                    o[P] = acc[j][0];
                    if constexpr (KIND == KIND_NEGA) o[P + L] = acc[j][1];
                }
            }
            err_update(ba.err, (int)(&pr - ba.prods), em);
        }
    }
}

template <int KIND, int S, int RB>
__device__ __forceinline__ void dispatch_sb(cplx<float>* lds, const BArgs& ba, const BProd& pr,
                                            const BTables& tb, int q, int tid, int* sh)
{
    constexpr int E = S >= 14 ? 5 : (S == 13 ? 4 : 3);  // NT <= 512
    constexpr int R = RB < E ? RB : E;
    if constexpr (S == SMAX) {
        switch (pr.SB) {
            case 0: run_item<KIND, S, 0, E, R>(lds, ba, pr, tb, q, tid, sh); break;
            case 1: run_item<KIND, S, 1, E, R>(lds, ba, pr, tb, q, tid, sh); break;
            case 2: run_item<KIND, S, 2, E, R>(lds, ba, pr, tb, q, tid, sh); break;
            case 3: run_item<KIND, S, 3, E, R>(lds, ba, pr, tb, q, tid, sh); break;
            case 4: run_item<KIND, S, 4, E, R>(lds, ba, pr, tb, q, tid, sh); break;
            default: break;
        }
    } else {
        run_item<KIND, S, 0, E, R>(lds, ba, pr, tb, q, tid, sh);
    }
}

template <int KIND, int RB>
__device__ __forceinline__ void dispatch_s(cplx<float>* lds, const BArgs& ba, const BProd& pr,
                                           const BTables& tb, int q, int tid, int* sh)
{
    switch (pr.S) {
        case 6: dispatch_sb<KIND, 6, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 7: dispatch_sb<KIND, 7, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 8: dispatch_sb<KIND, 8, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 9: dispatch_sb<KIND, 9, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 10: dispatch_sb<KIND, 10, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 11: dispatch_sb<KIND, 11, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 12: dispatch_sb<KIND, 12, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 13: dispatch_sb<KIND, 13, RB>(lds, ba, pr, tb, q, tid, sh); break;
        case 14: dispatch_sb<KIND, 14, RB>(lds, ba, pr, tb, q, tid, sh); break;
        default: break;
    }
}

template <int RB>
__global__ __launch_bounds__(BT) void batched_kernel(BArgs ba)
{
    __shared__ cplx<float> lds[1 << SMAX];
    __shared__ int sh_item;
    __shared__ int sh_last;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) sh_item = atomicAdd(ba.next_item, 1);
        __syncthreads();
        const int it = sh_item;
        __syncthreads();
        if (it >= ba.n_items) break;  // every workgroup ends here
        const int2 item = ba.items[it];
        const BProd& pr = ba.prods[item.x];
        const BTables tb = ba.tabs[pr.tabset];
        if (pr.kind == KIND_NEGA)
            dispatch_s<KIND_NEGA, RB>(lds, ba, pr, tb, item.y, tid, &sh_last);
        else
            dispatch_s<KIND_LINEAR, RB>(lds, ba, pr, tb, item.y, tid, &sh_last);
        __syncthreads();
    }
}

}  // namespace

struct gpufft_batched_plan_s {
    int count;
    int n_items;
    int wa, ka, wb, kb;
    int rb;
    int grid;
    BProd* d_prods;
    int2* d_items;
    BTables* d_tabs;
    int* d_counters;  // [0] work counter, [1..count] arrival counters
    cplx<float>* scratch;
    void* wc;
    void* wci;
    std::vector<void*> tables;
    std::vector<BProd> prods;
    hipGraphExec_t graph;
    const void *g_a, *g_b, *g_out, *g_err;
    void* g_stream;
};

namespace {

void release(gpufft_batched_plan_s* p)
{
    if (!p) return;
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    for (void* t : p->tables) (void)hipFree(t);
    void* bufs[7] = {p->d_prods, p->d_items, p->d_tabs, p->d_counters, p->scratch, p->wc, p->wci};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

template <typename V> int upload(V** dst, const std::vector<V>& src, hipStream_t st)
{
    hipError_t e = hipMalloc((void**)dst, sizeof(V) * std::max<size_t>(1, src.size()));
    if (e != hipSuccess) return fail_hip(e, "gpufft_batched_plan_create: hipMalloc");
    if (!src.empty()) {
        e = hipMemcpyAsync(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return fail_hip(e, "gpufft_batched_plan_create: hipMemcpyAsync");
    }
    return GPUFFT_OK;
}

int launch(gpufft_batched_plan_s* p, const int* a, const int* b, long long* out, double* err,
           hipStream_t st)
{
    const char* who = "gpufft_batched_execute";
    hipLaunchKernelGGL(zero_ints, dim3((p->count + 1 + 255) / 256), dim3(256), 0, st, p->d_counters,
                       p->count + 1);
    GFA_CHECK_LAUNCH(who);
    if (err) {
        hipLaunchKernelGGL(zero_doubles, dim3((p->count + 255) / 256), dim3(256), 0, st, err, p->count);
        GFA_CHECK_LAUNCH(who);
    }
    BArgs ba{};
    ba.prods = p->d_prods;
    ba.items = p->d_items;
    ba.tabs = p->d_tabs;
    ba.n_items = p->n_items;
    ba.next_item = p->d_counters;
    ba.arrive = p->d_counters + 1;
    ba.a = a;
    ba.b = b;
    ba.out = out;
    ba.scratch = p->scratch;
    ba.wc = (const cplx<float>*)p->wc;
    ba.wci = (const cplx<float>*)p->wci;
    ba.wa = p->wa;
    ba.ka = p->ka;
    ba.wb = p->wb;
    ba.kb = p->kb;
    auto offset = [](int w, int k) {
        long long o = 0;
        for (int j = 0; j + 1 < k; j++) o += (1LL << (w - 1)) << (w * j);
        return o;
    };
    ba.oa = p->ka > 1 ? offset(p->wa, p->ka) : 0;
    ba.ob = p->kb > 1 ? offset(p->wb, p->kb) : 0;
    ba.err = err;
    switch (p->rb) {
        case 3: hipLaunchKernelGGL(batched_kernel<3>, dim3(p->grid), dim3(BT), 0, st, ba); break;
        case 4: hipLaunchKernelGGL(batched_kernel<4>, dim3(p->grid), dim3(BT), 0, st, ba); break;
        default: hipLaunchKernelGGL(batched_kernel<5>, dim3(p->grid), dim3(BT), 0, st, ba); break;
    }
    GFA_CHECK_LAUNCH(who);
    return GPUFFT_OK;
}

}  // namespace

extern "C" {

int gpufft_batched_plan_create(gpufft_batched_plan* out, const gpufft_product_desc* descs, int count,
                               int a_limb_bits, int a_limbs, int b_limb_bits, int b_limbs,
                               void* stream)
{
    const char* who = "gpufft_batched_plan_create";
    if (!out || !descs) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL argument");
    *out = nullptr;
    if (count < 1) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": count < 1");
    if (a_limbs < 1 || b_limbs < 1 || a_limbs * b_limbs > 64 ||
        (a_limbs > 1 && (a_limb_bits < 1 || a_limb_bits > 16)) ||
        (b_limbs > 1 && (b_limb_bits < 1 || b_limb_bits > 16)))
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": invalid limb split");
    hipStream_t st = (hipStream_t)stream;
    gpufft_batched_plan_s* p = new gpufft_batched_plan_s();
    p->count = count;
    p->wa = a_limbs > 1 ? a_limb_bits : 0;
    p->ka = a_limbs;
    p->wb = b_limbs > 1 ? b_limb_bits : 0;
    p->kb = b_limbs;
    p->rb = 5;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // persistent grid: one 128 KiB-LDS workgroup per CU is resident at a time
    p->grid = cus;

    std::map<std::pair<int, int>, int> tabset;
    std::vector<BTables> tabs;
    long long scratch = 0;
    const int nprod = p->ka * p->kb;
    for (int i = 0; i < count; i++) {
        const gpufft_product_desc& d = descs[i];
        BProd pr{};
        if (d.kind != GPUFFT_PRODUCT_NEGACYCLIC && d.kind != GPUFFT_PRODUCT_LINEAR) {
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": unknown product kind");
        }
        pr.kind = d.kind == GPUFFT_PRODUCT_NEGACYCLIC ? KIND_NEGA : KIND_LINEAR;
        pr.N = 1 << d.ring_power;
        pr.n = pr.kind == KIND_NEGA ? d.ring_power - 1 : d.ring_power + 1;
        if (d.ring_power < 2 || pr.n < SMIN || pr.n > SMAX + SBMAX) {
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": ring_power out of range");
        }
        if (d.a_offset < 0 || d.b_offset < 0 || d.out_offset < 0) {
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": negative offset");
        }
        pr.S = std::min(pr.n, SMAX);
        pr.SB = pr.n - pr.S;
        pr.a_off = d.a_offset;
        pr.b_off = d.b_offset;
        pr.out_off = d.out_offset;
        auto key = std::make_pair(pr.kind, pr.n);
        auto f = tabset.find(key);
        if (f == tabset.end()) {
            BTables t{};
            void* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
            const int nb = pr.kind == KIND_NEGA ? 4 : 2;
            for (int k = 0; k < nb; k++) {
                hipError_t e = hipMalloc(&bufs[k], sizeof(cplx<float>) << pr.n);
                if (e != hipSuccess) {
                    release(p);
                    return fail_hip(e, who);
                }
                p->tables.push_back(bufs[k]);
            }
            int s = 0;
            if (pr.kind == KIND_NEGA) {
                s |= gpufft_generate_table_f32(bufs[0], GPUFFT_TABLE_FFNT_ROOT, pr.n, 0, st);
                s |= gpufft_generate_table_f32(bufs[1], GPUFFT_TABLE_FFNT_ROOT, pr.n, 1, st);
                s |= gpufft_generate_table_f32(bufs[2], GPUFFT_TABLE_FFNT_TWIST, pr.n, 0, st);
                s |= gpufft_generate_table_f32(bufs[3], GPUFFT_TABLE_FFNT_TWIST, pr.n, 1, st);
            } else {
                // cyclic length L = 2^n: FFT<T>(L/2) tables (L/2 entries)
                s |= gpufft_generate_table_f32(bufs[0], GPUFFT_TABLE_FFT, pr.n - 1, 0, st);
                s |= gpufft_generate_table_f32(bufs[1], GPUFFT_TABLE_FFT, pr.n - 1, 1, st);
            }
            if (s) {
                release(p);
                return s;
            }
            t.root = (const cplx<float>*)bufs[0];
            t.iroot = (const cplx<float>*)bufs[1];
            t.twist = (const cplx<float>*)bufs[2];
            t.untwist = (const cplx<float>*)bufs[3];
            f = tabset.emplace(key, (int)tabs.size()).first;
            tabs.push_back(t);
        }
        pr.tabset = f->second;
        if (pr.SB > 0) {
            pr.scratch_off = scratch;
            scratch += (long long)nprod << pr.n;
        }
        p->prods.push_back(pr);
    }
    // items, largest products first (cost ~ L log L per sub-problem)
    std::vector<int> order(count);
    for (int i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return p->prods[x].S > p->prods[y].S; });
    std::vector<int2> items;
    for (int i : order)
        for (int q = 0; q < (1 << p->prods[i].SB); q++) items.push_back(make_int2(i, q));
    p->n_items = (int)items.size();
    int s;
    if ((s = upload(&p->d_prods, p->prods, st)) || (s = upload(&p->d_items, items, st)) ||
        (s = upload(&p->d_tabs, tabs, st))) {
        release(p);
        return s;
    }
    hipError_t e = hipMalloc((void**)&p->d_counters, sizeof(int) * (count + 1));
    if (e == hipSuccess && scratch > 0) e = hipMalloc((void**)&p->scratch, sizeof(cplx<float>) * scratch);
    if (e == hipSuccess) e = hipMalloc(&p->wc, sizeof(cplx<float>) * 32);
    if (e == hipSuccess) e = hipMalloc(&p->wci, sizeof(cplx<float>) * 32);
    if (e != hipSuccess) {
        release(p);
        return fail_hip(e, who);
    }
    if ((s = gpufft_generate_table_f32(p->wc, GPUFFT_TABLE_STAGE_ROOTS, 5, 0, st)) ||
        (s = gpufft_generate_table_f32(p->wci, GPUFFT_TABLE_STAGE_ROOTS, 5, 1, st))) {
        release(p);
        return s;
    }
    e = hipStreamSynchronize(st);  // host vectors above are freed on return
    if (e != hipSuccess) {
        release(p);
        return fail_hip(e, who);
    }
    *out = p;
    return ok();
}

int gpufft_batched_plan_destroy(gpufft_batched_plan plan)
{
    release(plan);
    return GPUFFT_OK;
}

int gpufft_batched_plan_set_tuning(gpufft_batched_plan plan, int radix_bits, int grid)
{
    const char* who = "gpufft_batched_plan_set_tuning";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (radix_bits < 3 || radix_bits > 5 || grid < 1)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": invalid tuning");
    plan->rb = radix_bits;
    plan->grid = grid;
    if (plan->graph) {
        (void)hipGraphExecDestroy(plan->graph);
        plan->graph = nullptr;
    }
    return ok();
}

int gpufft_batched_plan_info(gpufft_batched_plan plan, int* n_items, long long* scratch_elems,
                             int* radix_bits, int* grid)
{
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, "gpufft_batched_plan_info: plan is NULL");
    if (n_items) *n_items = plan->n_items;
    long long s = 0;
    for (const BProd& pr : plan->prods)
        if (pr.SB > 0) s += (long long)plan->ka * plan->kb << pr.n;
    if (scratch_elems) *scratch_elems = s;
    if (radix_bits) *radix_bits = plan->rb;
    if (grid) *grid = plan->grid;
    return ok();
}

int gpufft_batched_execute(gpufft_batched_plan plan, const int32_t* a, const int32_t* b,
                           int64_t* out, double* err_stats, void* stream)
{
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, "gpufft_batched_execute: plan is NULL");
    if (!a || !b || !out) return fail(GPUFFT_ERR_NULL_POINTER, "gpufft_batched_execute: NULL buffer");
    if (int s = launch(plan, a, b, (long long*)out, err_stats, (hipStream_t)stream)) return s;
    return ok();
}

int gpufft_batched_graph_launch(gpufft_batched_plan plan, const int32_t* a, const int32_t* b,
                                int64_t* out, double* err_stats, void* stream)
{
    const char* who = "gpufft_batched_graph_launch";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (!a || !b || !out) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    hipStream_t st = (hipStream_t)stream;
    if (!plan->graph || plan->g_a != a || plan->g_b != b || plan->g_out != out ||
        plan->g_err != err_stats || plan->g_stream != stream) {
        if (plan->graph) {
            (void)hipGraphExecDestroy(plan->graph);
            plan->graph = nullptr;
        }
        hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return fail_hip(e, who);
        int s = launch(plan, a, b, (long long*)out, err_stats, st);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(st, &g);
        if (s) {
            if (g) (void)hipGraphDestroy(g);
            return s;
        }
        if (e != hipSuccess) return fail_hip(e, who);
        e = hipGraphInstantiate(&plan->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail_hip(e, who);
        plan->g_a = a;
        plan->g_b = b;
        plan->g_out = out;
        plan->g_err = err_stats;
        plan->g_stream = stream;
    }
    hipError_t e = hipGraphLaunch(plan->graph, st);
    if (e != hipSuccess) return fail_hip(e, who);
    return ok();
}

}  // extern "C"
