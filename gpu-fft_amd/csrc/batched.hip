// Batched products of uneven ring sizes in one persistent launch, int32 in /
// int64 out, in one of two precisions: FP32 with a and b split into limbs, or
// FP64 (one limb each).  No counterpart in the reference, which runs one
// GPU_FFT / GPU_FFNT call pair per ring size.
//
// Both product kinds are FFNT products (fold + twist, L complex points;
// reference GPU_FFNT, src/lib/fft.cu:1013-1384):
//   NEGACYCLIC  c = a b mod X^N + 1, ring N, L = N/2.
//   LINEAR      c = a b (2N - 1 coefficients, then a 0), computed as the
//               negacyclic product of ring 2N of the zero-padded operands:
//               deg(a b) < 2N, so nothing wraps.  L = N, and the fold's upper
//               half is the padding, so the imaginary inputs are 0.  This
//               stands for the reference's zero-padded cyclic C_C product of
//               length 2N (example/test_fft_gpu_C_C.cu) at half its
//               transform length.
//
// Every product is cut into 2^SB sub-problems of 2^S points, S = min(log2 L,
// T), as in fused.hpp; T is the plan's LDS tile: FP32 14 (128 KiB, one
// workgroup per CU) or 13 (64 KiB, two per CU), FP64 13 (128 KiB).  The work
// list holds one item per sub-problem.  The
// list is split into 8 queues, one per XCD: the hardware deals the workgroups
// of the persistent grid round-robin to the 8 XCDs, and workgroup b serves
// queue b mod 8 first.  All sub-problems of a product sit in one queue, so
// the workgroups that re-read its operands share an L2.  Products are dealt
// to the queues largest first, each to the least loaded queue.  A workgroup
// whose queue is exhausted serves the other queues in turn and exits once all
// 8 are exhausted (every counter only grows, so every workgroup gets there).
//
// Limb products: each limb of a and of b is gathered and forward-transformed
// once per item.  Limb products of equal weight 2^(wa ia + wb jb) (with
// wa = wb: a0 b1 and a1 b0, ...) are summed in the frequency domain and share
// one inverse transform; their magnitudes add, which the error statistics
// see.  Spectra and partial sums that do not stay in registers live in a
// small per-workgroup stash.  For a and b in 2 limbs each that is 4 forward
// and 3 inverse transforms, against 10 when each limb product is formed on
// its own (b limb, then every a limb again, then the inverse).
//
// An item with SB = 0 is a whole product: it rounds each group and
// accumulates it into the output.  An item with SB > 0 writes its group
// sub-products to scratch.  The top SB inverse stages, the rounding and the
// recombination of a split product then run in one of two places:
//   * a second kernel over all split products (the default);
//   * the persistent kernel itself: each sub-problem publishes its scratch
//     (agent-scope release) and increments the product's arrival counter, and
//     the workgroup that arrives last (acquire) combines the product.  Nothing
//     waits on another workgroup.
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
constexpr int SMAX = 14;    // largest tile: 2^14 float complex = 128 KiB of LDS
constexpr int SMAX64 = 13;  // FP64: 2^13 double complex = 128 KiB
constexpr int SMIN = 5;
constexpr int SBMAX = 4;
constexpr int NQ = 8;       // work queues, one per XCD
constexpr int CT = 256;     // threads of a combine workgroup

struct BProd {
    int kind;
    int n;        // log2 L
    int S, SB;
    int N;        // coefficients per operand
    int tabset;   // index of the table set for n
    long long a_off, b_off, out_off;
    long long scratch_off;  // cplx<T> offset (SB > 0)
};

// FFNT tables of one L, cplx<float> or cplx<double> as the plan's precision
struct BTables {
    const void* root;
    const void* iroot;
    const void* twist;
    const void* untwist;
};

template <typename T> __device__ __forceinline__ const cplx<T>* tabp(const void* p)
{
    return static_cast<const cplx<T>*>(p);
}

// Order of the limb products of one product.  Limb products of equal
// weight wa ia + wb jb form a group: they are summed in the frequency domain
// and share one inverse transform.  Index k = ia kb + jb.
constexpr int MAXP = 64;
struct BSched {
    int G;               // groups = inverse transforms per product
    int gshift[MAXP];    // weight exponent of group g
    int gslot[MAXP];     // stash slot of group g's partial sum (-1: none)
    int gdone[MAXP];     // a limb after which group g is complete
    int cgroup[MAXP];    // group of limb product k
    int cfirst[MAXP];    // 1: k is the first contribution to its group's slot
    int dgroup[MAXP];    // per a limb: group summed in registers and inverted
                         // right after it (-1: none)
    int dload[MAXP];     // per a limb: 1 if that group has a partial sum in its slot
};

struct BArgs {
    const BProd* prods;
    const int2* items;  // (product, q), queue after queue
    const BTables* tabs;
    int qbeg[NQ];       // first item of each queue
    int qlen[NQ];       // items of each queue
    int* next_item;     // NQ queue counters
    int* arrive;        // per product (in-kernel combine)
    int* exited;        // workgroups that have left the persistent kernel
    int count;          // products
    const int* split;   // the products with SB > 0 (combine kernel)
    int cmb;            // 1: split products are combined in the persistent kernel
    const int* a;
    const int* b;
    long long* out;
    void* scratch;       // cplx<T> sub-results of split products
    const BSched* sched;
    void* stash;         // cplx<T>: stash_slots x 2^SMAX per workgroup of the persistent grid
    int stash_slots;     // b-limb spectra 1 .. kb-1, then group partial sums
    const void* wc;      // stage root constants, cplx<T>
    const void* wci;
    int wa, ka, wb, kb;
    long long oa, ob;
    double* err;        // optional [products]
    int out_pair;       // 1: every product's output starts 16-byte aligned (paired int64 stores)
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

// Largest e over the wave, folded into err[prod]; every lane of the wave
// takes part.
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

// Global-memory views of the pointers the item code works through: the
// pointers reach it inside structures, where the compiler no longer knows
// their address space and would use flat instructions (which also count
// against the LDS wait counter).
// (The host pass of the compiler parses this code too but never runs it; it
// gets plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T> using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T> using gptr = T*;
#endif
template <typename T> __device__ __forceinline__ gptr<T> to_global(T* p)
{
    return (gptr<T>)p;
}

// Digit `limb` of an operand coefficient as T (one limb: the coefficient
// itself, exact in FP32 below 2^24 and in FP64 for every int32).
template <typename T>
__device__ __forceinline__ T item_digit(int v, int w, int limb, int k, int o)
{
    if (k == 1) return (T)v;
    return (T)limb_digit(v, w, limb, k, o, 0);
}

// Twist factors of the 4 consecutive positions c0 .. c0+3 (16-byte loads).
template <typename T>
__device__ __forceinline__ void twist4(gptr<const cplx<T>> twist, int c0, cplx<T>* t)
{
    if constexpr (sizeof(T) == 4) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const f32x4 t0 = *(gptr<const f32x4>)(twist + c0);
        const f32x4 t1 = *(gptr<const f32x4>)(twist + c0 + 2);
        t[0] = mk<float>(t0[0], t0[1]);
        t[1] = mk<float>(t0[2], t0[3]);
        t[2] = mk<float>(t1[0], t1[1]);
        t[3] = mk<float>(t1[2], t1[3]);
    } else {
#pragma unroll
        for (int v = 0; v < 4; v++) t[v] = twist[c0 + v];
    }
}

// Sub-problem q of one limb of one operand into LDS: for each of the
// workgroup's 2^S positions c, row q of the top SB forward stages over the
// folded inputs at j 2^S + c (Q = 2^SB partners), times the twist
// (coefficients C as in top_row_coeffs).  LINEAR products have no upper fold
// half.  A lane takes 4 consecutive positions with one 16-byte load per
// partner (operand offsets are multiples of 4 elements), U such groups per
// round with all their loads issued together.
// FP64 (TWL): C and the twist factors come from LDS (C written by the
// item, twist[c] = tws[c >> 7] tws[NHI + (c & 127)]), which keeps the 2^SB
// double coefficients and the twist loads out of the registers and the
// vector memory path.
template <typename T, int KIND, int S, int Q, int E, bool TWL = false>
__device__ __forceinline__ void item_load(cplx<T>* lds, gptr<const cplx<T>> twist,
                                          gptr<const int> x, int L, int w_in, int limb_in,
                                          int k_in, int o_in, const cplx<T>* C, int tid,
                                          const cplx<T>* tws = nullptr)
{
    // the split parameters are uniform: scalar copies keep the digit
    // extraction free of per-lane branches
    const int w = __builtin_amdgcn_readfirstlane(w_in);
    const int limb = __builtin_amdgcn_readfirstlane(limb_in);
    const int k = __builtin_amdgcn_readfirstlane(k_in);
    const int o = __builtin_amdgcn_readfirstlane(o_in);
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    constexpr int NT = (1 << S) >> E;
    constexpr int NV = (1 << E) / 4;  // groups of 4 positions per lane
    constexpr int U0 = Q >= 4 ? 1 : (Q == 2 ? 2 : 4);
    constexpr int U = U0 < NV ? U0 : NV;
    if (tid >= NT) return;
    for (int i0 = 0; i0 < NV; i0 += U) {
        i32x4 xr[U][Q], xi[U][Q];
        cplx<T> tw[U][4];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c0 = (tid + (i0 + u) * NT) * 4;
            if constexpr (!TWL) twist4<T>(twist, c0, tw[u]);
#pragma unroll
            for (int j = 0; j < Q; j++) {
                const int P = (j << S) | c0;
                xr[u][j] = *(gptr<const i32x4>)(x + P);
                if constexpr (KIND == KIND_NEGA) xi[u][j] = *(gptr<const i32x4>)(x + P + L);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int c0 = (tid + (i0 + u) * NT) * 4;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T ax = T(0), ay = T(0);
#pragma unroll
                for (int j = 0; j < Q; j++) {
                    const T re = item_digit<T>(xr[u][j][v], w, limb, k, o);
                    if constexpr (KIND == KIND_NEGA) {
                        const T im = item_digit<T>(xi[u][j][v], w, limb, k, o);
                        ax = ax + C[j].x * re - C[j].y * im;
                        ay = ay + C[j].x * im + C[j].y * re;
                    } else {
                        ax = ax + C[j].x * re;
                        ay = ay + C[j].y * re;
                    }
                }
                if constexpr (TWL)
                    lds[swz(c0 + v)] = twist_of<T, S>(tws, c0 + v) * mk<T>(ax, ay);
                else
                    lds[swz(c0 + v)] = tw[u][v] * mk<T>(ax, ay);
            }
        }
    }
}

// FP64 full tiles (S = the tile, S mod RB = 1): item_load with the short
// top forward step (bit S-1) in registers.  A lane's blocks i and i + NV/2
// of 4 positions are 2^(S-1) apart, the two halves of that step; partners
// are accumulated 4 at a time to bound the registers of the loads.  C and
// the twist factors come from LDS as in item_load<TWL>.
template <typename T, int KIND, int S, int Q, int E>
__device__ __forceinline__ void item_load_top(cplx<T>* lds, gptr<const int> x, int L,
                                              const cplx<T>* C, const cplx<T>* tws,
                                              const ProdArgs<T>& pa, int q, int tid)
{
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    constexpr int NT = (1 << S) >> E;
    constexpr int NV = (1 << E) / 4;  // blocks of 4 positions per lane
    constexpr int NH = NV / 2;        // block i pairs with block i + NH
    constexpr int JC = Q < 4 ? Q : 4;  // partners per load round
    static_assert(NH >= 1, "two blocks per lane");
    if (tid >= NT) return;
#pragma unroll 1
    for (int i = 0; i < NH; i++) {
        int c0[2];
        c0[0] = (tid + i * NT) * 4;
        c0[1] = (tid + (i + NH) * NT) * 4;
        T ax[2][4], ay[2][4];
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int v = 0; v < 4; v++) ax[h][v] = ay[h][v] = T(0);
#pragma unroll
        for (int j0 = 0; j0 < Q; j0 += JC) {
            i32x4 xr[2][JC], xi[2][JC];
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int jj = 0; jj < JC; jj++) {
                    const int P = ((j0 + jj) << S) | c0[h];
                    xr[h][jj] = *(gptr<const i32x4>)(x + P);
                    if constexpr (KIND == KIND_NEGA) xi[h][jj] = *(gptr<const i32x4>)(x + P + L);
                }
#pragma unroll
            for (int jj = 0; jj < JC; jj++) {
                const cplx<T> cj = C[j0 + jj];
#pragma unroll
                for (int h = 0; h < 2; h++)
#pragma unroll
                    for (int v = 0; v < 4; v++) {
                        const T re = (T)xr[h][jj][v];
                        if constexpr (KIND == KIND_NEGA) {
                            const T im = (T)xi[h][jj][v];
                            ax[h][v] = ax[h][v] + cj.x * re - cj.y * im;
                            ay[h][v] = ay[h][v] + cj.x * im + cj.y * re;
                        } else {
                            ax[h][v] = ax[h][v] + cj.x * re;
                            ay[h][v] = ay[h][v] + cj.y * re;
                        }
                    }
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            cplx<T> z[2];
#pragma unroll
            for (int h = 0; h < 2; h++) z[h] = twist_of<T, S>(tws, c0[h] + v) * mk<T>(ax[h][v], ay[h][v]);
            top_ct<T, S, 1>(z, pa, q);
#pragma unroll
            for (int h = 0; h < 2; h++) lds[swz(c0[h] + v)] = z[h];
        }
    }
}

// Streaming load and store of one element through a global pointer.
template <typename T> __device__ __forceinline__ cplx<T> ldg_nt(gptr<const cplx<T>> p)
{
    typedef T vec2 __attribute__((ext_vector_type(2)));
    const vec2 w = __builtin_nontemporal_load((gptr<const vec2>)p);
    return mk<T>(w.x, w.y);
}

template <typename T> __device__ __forceinline__ void stg_nt(gptr<cplx<T>> p, cplx<T> v)
{
    typedef T vec2 __attribute__((ext_vector_type(2)));
    vec2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (gptr<vec2>)p);
}

// Rounding to the nearest integer (ties to even) and the distance to it.
template <typename T> __device__ __forceinline__ T rnd(T x)
{
    if constexpr (sizeof(T) == 4)
        return rintf(x);
    else
        return rint(x);
}

template <typename T> __device__ __forceinline__ float rdist(cplx<T> z, T r0, T r1)
{
    const T d0 = z.x - r0, d1 = z.y - r1;
    const T a0 = d0 < T(0) ? -d0 : d0, a1 = d1 < T(0) ? -d1 : d1;
    return (float)(a0 > a1 ? a0 : a1);
}

// The top SB inverse stages of position c < 2^S of a split product (S the
// plan's tile) over all its limb-group sub-products in scratch, then the
// scale 1/L, the untwist, the rounding of each group and the int64
// recombination; writes the unfolded outputs out[P], out[P + L] for
// P = j 2^S + c.  Returns the largest distance of a group output to its
// rounding.
// UL: the untwist factor of P = j 2^S + c is uhi[4 j + ((c - cbase) >> 7)]
// ulo[c & 127] from LDS (the combine workgroup's 512 positions from cbase),
// else untwist[P] from the table.
template <typename T, int SB, bool UL = false>
__device__ __forceinline__ float combine_pos(const BArgs& ba, const BProd& pr, const BTables& tb,
                                             int c, const cplx<T>* uhi = nullptr,
                                             const cplx<T>* ulo = nullptr, int cbase = 0)
{
    constexpr int Q = 1 << SB;
    constexpr bool ONE = sizeof(T) == 8;  // one group: outputs stored as they are rounded
    const int S = pr.S;
    const int L = 1 << pr.n;
    const T scale = T(1) / (T)L;
    const int ngrp = ONE ? 1 : ba.sched->G;
    const gptr<long long> o = to_global(ba.out + pr.out_off);
    long long acc[ONE ? 1 : Q][2];
#pragma unroll
    for (int j = 0; j < (ONE ? 1 : Q); j++) acc[j][0] = acc[j][1] = 0;
    float em = 0.f;
    const gptr<const cplx<T>> untwist = to_global(tabp<T>(tb.untwist));
    const cplx<T>* iroot = tabp<T>(tb.iroot);
    for (int l = 0; l < ngrp; l++) {
        const gptr<const cplx<T>> src =
            to_global(static_cast<const cplx<T>*>(ba.scratch) + pr.scratch_off + (long long)l * L);
        cplx<T> y[Q];
#pragma unroll
        for (int j = 0; j < Q; j++) y[j] = ldg_nt<T>(src + (j << S) + c);
#pragma unroll
        for (int i = 0; i < SB; i++) {  // Gentleman-Sande, bits S .. n-1
#pragma unroll
            for (int j0 = 0; j0 < Q; j0++) {
                if (j0 & (1 << i)) continue;
                const int j1 = j0 | (1 << i);
                const cplx<T> w = iroot[j0 >> (i + 1)];
                const cplx<T> U = y[j0], V = y[j1];
                y[j0] = U + V;
                y[j1] = (U - V) * w;
            }
        }
        const long long wgt = ONE ? 1 : 1LL << ba.sched->gshift[l];
#pragma unroll
        for (int j = 0; j < Q; j++) {
            const int P = (j << S) | c;
            cplx<T> ut;
            if constexpr (UL)
                ut = uhi[4 * j + ((c - cbase) >> 7)] * ulo[c & 127];
            else
                ut = untwist[P];
            const cplx<T> z = mk<T>(y[j].x * scale, y[j].y * scale) * ut;
            const T r0 = rnd(z.x), r1 = rnd(z.y);
            em = fmaxf(em, rdist(z, r0, r1));
            if constexpr (ONE) {
                __builtin_nontemporal_store((long long)r0, o + P);
                __builtin_nontemporal_store((long long)r1, o + P + L);
            } else {
                acc[j][0] += (long long)r0 * wgt;
                acc[j][1] += (long long)r1 * wgt;
            }
        }
    }
    if constexpr (ONE) return em;
#pragma unroll
    for (int j = 0; j < (ONE ? 1 : Q); j++) {
        const int P = (j << S) | c;
        __builtin_nontemporal_store(acc[j][0], o + P);
        __builtin_nontemporal_store(acc[j][1], o + P + L);
    }
    return em;
}

// Elements c and c + 1 (c even) of a cplx<T> array: one 16-byte load for
// float, two for double.
template <typename T>
__device__ __forceinline__ void load_pair_nt(gptr<const cplx<T>> p, cplx<T>& z0, cplx<T>& z1)
{
    if constexpr (sizeof(T) == 4) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const f32x4 w = __builtin_nontemporal_load((gptr<const f32x4>)p);
        z0 = mk<float>(w[0], w[1]);
        z1 = mk<float>(w[2], w[3]);
    } else {
        z0 = ldg_nt<T>(p);
        z1 = ldg_nt<T>(p + 1);
    }
}

template <typename T>
__device__ __forceinline__ void load_pair(gptr<const cplx<T>> p, cplx<T>& z0, cplx<T>& z1)
{
    if constexpr (sizeof(T) == 4) {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const f32x4 w = *(gptr<const f32x4>)p;
        z0 = mk<float>(w[0], w[1]);
        z1 = mk<float>(w[2], w[3]);
    } else {
        z0 = p[0];
        z1 = p[1];
    }
}

// Outputs of positions P, P + 1 (P even): r[v][0] to o[P + v] and r[v][1] to
// o[P + v + L]; one 16-byte store per pair when the output is 16-byte
// aligned (8-byte stores of every other element leave half-written lines to
// the memory system), else two 8-byte stores.
__device__ __forceinline__ void store_pair(gptr<long long> o, int P, int L, const long long (*r)[2],
                                           int paired)
{
    if (paired) {
        typedef long long i64x2 __attribute__((ext_vector_type(2)));
        i64x2 re, im;
        re.x = r[0][0];
        re.y = r[1][0];
        im.x = r[0][1];
        im.y = r[1][1];
        __builtin_nontemporal_store(re, (gptr<i64x2>)(o + P));
        __builtin_nontemporal_store(im, (gptr<i64x2>)(o + P + L));
    } else {
#pragma unroll
        for (int v = 0; v < 2; v++) {
            __builtin_nontemporal_store(r[v][0], o + P + v);
            __builtin_nontemporal_store(r[v][1], o + P + v + L);
        }
    }
}

// combine_pos for the two adjacent positions c, c + 1 (c even): paired
// loads of every row, group and untwist entry.
template <typename T, int SB, bool UL = false>
__device__ __forceinline__ float combine_pos2(const BArgs& ba, const BProd& pr, const BTables& tb,
                                              int c, const cplx<T>* uhi = nullptr,
                                              const cplx<T>* ulo = nullptr, int cbase = 0)
{
    constexpr int Q = 1 << SB;
    constexpr bool ONE = sizeof(T) == 8;  // one group: outputs stored as they are rounded
    const int S = pr.S;
    const int L = 1 << pr.n;
    const T scale = T(1) / (T)L;
    const int ngrp = ONE ? 1 : ba.sched->G;
    const gptr<long long> o = to_global(ba.out + pr.out_off);
    long long acc[ONE ? 1 : Q][2][2];  // [row][position][re, im]
#pragma unroll
    for (int j = 0; j < (ONE ? 1 : Q); j++) acc[j][0][0] = acc[j][0][1] = acc[j][1][0] = acc[j][1][1] = 0;
    float em = 0.f;
    const gptr<const cplx<T>> untwist = to_global(tabp<T>(tb.untwist));
    const cplx<T>* iroot = tabp<T>(tb.iroot);
    for (int l = 0; l < ngrp; l++) {
        const gptr<const cplx<T>> src =
            to_global(static_cast<const cplx<T>*>(ba.scratch) + pr.scratch_off + (long long)l * L);
        cplx<T> y[2][Q];
#pragma unroll
        for (int j = 0; j < Q; j++) load_pair_nt<T>(src + (j << S) + c, y[0][j], y[1][j]);
#pragma unroll
        for (int i = 0; i < SB; i++) {  // Gentleman-Sande, bits S .. n-1
#pragma unroll
            for (int j0 = 0; j0 < Q; j0++) {
                if (j0 & (1 << i)) continue;
                const int j1 = j0 | (1 << i);
                const cplx<T> w = iroot[j0 >> (i + 1)];
#pragma unroll
                for (int v = 0; v < 2; v++) {
                    const cplx<T> U = y[v][j0], V = y[v][j1];
                    y[v][j0] = U + V;
                    y[v][j1] = (U - V) * w;
                }
            }
        }
        const long long wgt = ONE ? 1 : 1LL << ba.sched->gshift[l];
#pragma unroll
        for (int j = 0; j < Q; j++) {
            const int P = (j << S) | c;
            cplx<T> ut[2];
            if constexpr (UL) {
                const cplx<T> h = uhi[4 * j + ((c - cbase) >> 7)];
                ut[0] = h * ulo[c & 127];
                ut[1] = h * ulo[(c & 127) + 1];
            } else {
                load_pair<T>(untwist + P, ut[0], ut[1]);
            }
            long long rr[2][2];
#pragma unroll
            for (int v = 0; v < 2; v++) {
                const cplx<T> z = mk<T>(y[v][j].x * scale, y[v][j].y * scale) * ut[v];
                const T r0 = rnd(z.x), r1 = rnd(z.y);
                em = fmaxf(em, rdist(z, r0, r1));
                if constexpr (ONE) {
                    rr[v][0] = (long long)r0;
                    rr[v][1] = (long long)r1;
                } else {
                    acc[j][v][0] += (long long)r0 * wgt;
                    acc[j][v][1] += (long long)r1 * wgt;
                }
            }
            if constexpr (ONE) store_pair(o, P, L, rr, ba.out_pair);
        }
    }
    if constexpr (ONE) return em;
#pragma unroll
    for (int j = 0; j < (ONE ? 1 : Q); j++) store_pair(o, (j << S) | c, L, acc[j], ba.out_pair);
    return em;
}

template <typename T>
__device__ __forceinline__ float combine_pos_rt(const BArgs& ba, const BProd& pr,
                                                const BTables& tb, int c)
{
    switch (pr.SB) {
        case 1: return combine_pos<T, 1>(ba, pr, tb, c);
        case 2: return combine_pos<T, 2>(ba, pr, tb, c);
        case 3: return combine_pos<T, 3>(ba, pr, tb, c);
        case 4: return combine_pos<T, 4>(ba, pr, tb, c);
        default: return 0.f;
    }
}

// Copy of a workgroup-uniform value held in scalar registers.
template <typename V> __device__ __forceinline__ V uniform_copy(const V& x)
{
    static_assert(sizeof(V) % 4 == 0, "whole dwords");
    unsigned w[sizeof(V) / 4];
    __builtin_memcpy(w, &x, sizeof(V));
#pragma unroll
    for (unsigned i = 0; i < sizeof(V) / 4; i++) w[i] = __builtin_amdgcn_readfirstlane(w[i]);
    V r;
    __builtin_memcpy(&r, w, sizeof(V));
    return r;
}

// Stash element e (index in the thread's register layout) of slot s, for
// thread t: for a fixed e consecutive threads touch consecutive elements.
// Each thread only reads back what it wrote itself.
template <int NT, typename T>
__device__ __forceinline__ gptr<cplx<T>> stash_at(gptr<cplx<T>> base, int s, int e, int t)
{
    return base + ((long long)s << SMAX) + e * NT + t;
}

// Whole product or one sub-problem of it, in three phases:
//   1. each b limb: gather, forward steps; the spectrum of limb 0 stays in
//      registers, the others go to the workgroup's stash;
//   2. each a limb: gather, forward steps, then in registers the products
//      with every b-limb spectrum, summed by weight group (sc.cgroup): the
//      group that completes with this limb (sc.dgroup) is summed in registers
//      and inverse-transformed at once, the others are summed in stash slots;
//   3. after a limb, every other group it completed is inverse-transformed
//      from its slot.
// An inverse transform ends in emit(): rounding and accumulation into the
// output (SB = 0) or the store of the sub-result to scratch (SB > 0).
template <typename T, int KIND, int S, int E, int RB, int TILE>
__device__ __forceinline__ void run_item(cplx<T>* lds, const BArgs& ba, const BProd& pr,
                                         const BTables& tb, int prod, int q, int tid, int* sh_last)
{
    constexpr int EPT = 1 << E;
    constexpr int NT = (1 << S) >> E;
    constexpr int QMAX = S == TILE ? 1 << SBMAX : 1;  // only S = TILE items are split
    const int SB = S == TILE ? pr.SB : 0;
    const int Q = 1 << SB;
    constexpr int NST = (S + RB - 1) / RB;
    constexpr int NG = 1 << (E - RB);  // register groups of a thread
    const bool active = tid < NT;
    const int L = 1 << pr.n;
    const BSched& sc = *ba.sched;

    const cplx<T>* root = tabp<T>(tb.root);
    const cplx<T>* iroot = tabp<T>(tb.iroot);
    const cplx<T>* wc = static_cast<const cplx<T>*>(ba.wc);
    const cplx<T>* wci = static_cast<const cplx<T>*>(ba.wci);
    ProdArgs<T> pa{};
    pa.n = pr.n;
    pa.S = S;
    pa.nega = 0;
    pa.tab = root;
    pa.itab = iroot;
    pa.twist = tabp<T>(tb.twist);
    PassArgs<T> pf{};
    pf.n = pr.n;
    pf.K = S;
    pf.tab = root;
    PassArgs<T> pinv = pf;
    pinv.tab = iroot;
    TileGeo g;
    g.hi_pos = q << S;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = make_rsrc(root);
    TileGeo gi = g;
    gi.tab = make_rsrc(iroot);

    // coefficients of row q of the top network, twist factored
    auto coeffs = [&](cplx<T>* C) {
        if constexpr (QMAX == 1) {
            top_row_coeffs<T, 0>(pa, q, C);
        } else {
            switch (SB) {
                case 0: top_row_coeffs<T, 0>(pa, q, C); break;
                case 1: top_row_coeffs<T, 1>(pa, q, C); break;
                case 2: top_row_coeffs<T, 2>(pa, q, C); break;
                case 3: top_row_coeffs<T, 3>(pa, q, C); break;
                default: top_row_coeffs<T, 4>(pa, q, C); break;
            }
        }
    };
    // FP32: in registers, once per item.  FP64: in LDS beside the tile
    // (2^SB double coefficients would not fit the registers beside the b
    // spectrum), with the split twist tables; the phase-1 barrier publishes
    // them.
    constexpr bool TWL = sizeof(T) == 8;
    constexpr bool ONE = sizeof(T) == 8;  // one limb per operand, one group
    // FP64 full tiles: the short top step (S mod RB = 1 bit) in registers, in
    // the gather (forward) and in emit (inverse), one LDS round trip fewer
    // per transform
    constexpr int HBT = (ONE && S == TILE) ? S % RB : 0;
    static_assert(HBT <= 1, "one-bit register top step");
    constexpr int FS0 = HBT ? 1 : 0;           // first forward step through LDS
    constexpr int IEND = HBT ? NST - 1 : NST;  // end of the inverse steps through LDS
    cplx<T> C_item[TWL ? 1 : QMAX];
    cplx<T>* cl = lds + (1 << TILE);
    cplx<T>* tws = cl + (1 << SBMAX);
    if constexpr (TWL) {
        if (tid < Q) {  // thread j: C[j]
            cplx<T> cj;
            if constexpr (QMAX == 1) {
                cj = top_row_coeff<T, 0>(pa, q, tid);
            } else {
                switch (SB) {
                    case 0: cj = top_row_coeff<T, 0>(pa, q, tid); break;
                    case 1: cj = top_row_coeff<T, 1>(pa, q, tid); break;
                    case 2: cj = top_row_coeff<T, 2>(pa, q, tid); break;
                    case 3: cj = top_row_coeff<T, 3>(pa, q, tid); break;
                    default: cj = top_row_coeff<T, 4>(pa, q, tid); break;
                }
            }
            cl[tid] = cj;
        }
        fill_twist_split<T, S>(tws, pa.twist, tid, BT);
    } else {
        coeffs(C_item);
    }
    const gptr<const int> xa = to_global(ba.a + pr.a_off);
    const gptr<const int> xb = to_global(ba.b + pr.b_off);
    const gptr<const cplx<T>> twist = to_global(tabp<T>(tb.twist));
    const gptr<const cplx<T>> untwist = to_global(tabp<T>(tb.untwist));
    const gptr<cplx<T>> stash = to_global(static_cast<cplx<T>*>(ba.stash) +
                                          (((long long)blockIdx.x * ba.stash_slots) << SMAX));
    const T scale = T(1) / (T)L;
    float emax = 0.f;
    bool first_out = true;

    // gather with a compile-time partner count
    auto gather = [&](gptr<const int> x, int w, int limb, int k, int o, int t) {
        const cplx<T>* C = TWL ? cl : C_item;
        if constexpr (HBT) {
            switch (SB) {
                case 0: item_load_top<T, KIND, S, 1, E>(lds, x, L, C, tws, pa, q, t); break;
                case 1: item_load_top<T, KIND, S, 2, E>(lds, x, L, C, tws, pa, q, t); break;
                case 2: item_load_top<T, KIND, S, 4, E>(lds, x, L, C, tws, pa, q, t); break;
                case 3: item_load_top<T, KIND, S, 8, E>(lds, x, L, C, tws, pa, q, t); break;
                default: item_load_top<T, KIND, S, 16, E>(lds, x, L, C, tws, pa, q, t); break;
            }
            return;
        }
        if constexpr (QMAX == 1) {
            item_load<T, KIND, S, 1, E, TWL>(lds, twist, x, L, w, limb, k, o, C, t, tws);
        } else {
            switch (SB) {
                case 0: item_load<T, KIND, S, 1, E, TWL>(lds, twist, x, L, w, limb, k, o, C, t, tws); break;
                case 1: item_load<T, KIND, S, 2, E, TWL>(lds, twist, x, L, w, limb, k, o, C, t, tws); break;
                case 2: item_load<T, KIND, S, 4, E, TWL>(lds, twist, x, L, w, limb, k, o, C, t, tws); break;
                case 3: item_load<T, KIND, S, 8, E, TWL>(lds, twist, x, L, w, limb, k, o, C, t, tws); break;
                default: item_load<T, KIND, S, 16, E, TWL>(lds, twist, x, L, w, limb, k, o, C, t, tws); break;
            }
        }
    };

    // the inverse of group grp is in LDS
    auto emit = [&](int grp, int u) {
        if constexpr (HBT) {
            // element pairs c, c + 2^(S-1): the top inverse step, then the
            // output (SB = 0) or the scratch sub-result
            constexpr int NI = EPT / 2;
            if (!active) return;
            if (SB == 0) {
                const gptr<long long> o = to_global(ba.out + pr.out_off);
#pragma unroll 2
                for (int i = 0; i < NI; i++) {
                    const int c = u + i * NT;
                    cplx<T> z[2];
#pragma unroll
                    for (int m = 0; m < 2; m++) z[m] = lds[swz(c + (m << (S - 1)))];
                    top_gs<T, S, 1>(z, pa, q);
#pragma unroll
                    for (int m = 0; m < 2; m++) {
                        const int cm = c + (m << (S - 1));
                        const cplx<T> y = mk<T>(z[m].x * scale, z[m].y * scale) * untwist[cm];
                        const T r0 = rnd(y.x), r1 = rnd(y.y);
                        emax = fmaxf(emax, rdist(y, r0, r1));
                        o[cm] = (long long)r0;
                        o[cm + L] = (long long)r1;
                    }
                }
            } else {
                const gptr<cplx<T>> dst = to_global(static_cast<cplx<T>*>(ba.scratch) +
                                                    pr.scratch_off + ((long long)q << S));
#pragma unroll 2
                for (int i = 0; i < NI; i++) {
                    const int c = u + i * NT;
                    cplx<T> z[2];
#pragma unroll
                    for (int m = 0; m < 2; m++) z[m] = lds[swz(c + (m << (S - 1)))];
                    top_gs<T, S, 1>(z, pa, q);
#pragma unroll
                    for (int m = 0; m < 2; m++) stg_nt<T>(dst + c + (m << (S - 1)), z[m]);
                }
            }
            return;
        }
        if (SB == 0) {
            if (active) {
                const long long wgt = ONE ? 1 : 1LL << __builtin_amdgcn_readfirstlane(sc.gshift[grp]);
                const gptr<long long> o = to_global(ba.out + pr.out_off);
#pragma unroll 4
                for (int i = 0; i < EPT; i++) {
                    const int c = u + i * NT;
                    cplx<T> z = lds[swz(c)];
                    z = mk<T>(z.x * scale, z.y * scale) * untwist[c];
                    const T r0 = rnd(z.x), r1 = rnd(z.y);
                    emax = fmaxf(emax, rdist(z, r0, r1));
                    const long long v0 = (long long)r0 * wgt, v1 = (long long)r1 * wgt;
                    if (ONE || first_out) {
                        o[c] = v0;
                        o[c + L] = v1;
                    } else {
                        o[c] += v0;
                        o[c + L] += v1;
                    }
                }
            }
            first_out = false;
        } else {
            if (active) {
                const gptr<cplx<T>> dst = to_global(static_cast<cplx<T>*>(ba.scratch) +
                                                    pr.scratch_off + (long long)grp * L +
                                                    ((long long)q << S));
#pragma unroll 4
                for (int i = 0; i < EPT; i++) {
                    const int c = u + i * NT;
                    stg_nt<T>(dst + c, lds[swz(c)]);
                }
            }
        }
    };

    // phase 1: b-limb spectra
    cplx<T> breg[EPT];
    const int kb = ONE ? 1 : ba.kb;
    for (int jb = 0; jb < kb; jb++) {
        int t = tid;
        asm volatile("" : "+v"(t));
        __syncthreads();
        gather(xb, ba.wb, jb, ba.kb, (int)ba.ob, t);
        run_steps_ct_active<T, MODE_STD_FWD, E, NT, S, RB, FS0, NST - 1>(lds, pf, g, t, wc, active);
        if (active) {
#pragma unroll
            for (int kk = 0; kk < NG; kk++) {
                const int rbase = (t + kk * NT) << RB;
                cplx<T> v[1 << RB];
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) v[j] = lds[swz(rbase | j)];
                group_stages_derived<T, MODE_STD_FWD, RB>(v, pf, g, rbase, 0, 0, wc);
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) {
                    const int e = (kk << RB) + j;
                    if (ONE || jb == 0)
                        breg[e] = v[j];
                    else
                        *stash_at<NT>(stash, jb - 1, e, t) = v[j];
                }
            }
        }
    }

    // one limb each (FP64): the a spectrum times the b spectrum in registers,
    // then the inverse
    if constexpr (ONE) {
        int u = tid;
        asm volatile("" : "+v"(u));
        __syncthreads();
        gather(xa, 0, 0, 1, 0, u);
        run_steps_ct_active<T, MODE_STD_FWD, E, NT, S, RB, FS0, NST - 1>(lds, pf, g, u, wc, active);
        if (active) {
#pragma unroll
            for (int kk = 0; kk < NG; kk++) {
                const int rbase = (u + kk * NT) << RB;
                cplx<T> v[1 << RB];
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) v[j] = lds[swz(rbase | j)];
                group_stages_derived<T, MODE_STD_FWD, RB>(v, pf, g, rbase, 0, 0, wc);
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) v[j] = v[j] * breg[(kk << RB) + j];
                group_stages_derived<T, MODE_STD_INV, RB>(v, pinv, gi, rbase, 0, 0, wci);
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) lds[swz(rbase | j)] = v[j];
            }
        }
        run_steps_ct_active<T, MODE_STD_INV, E, NT, S, RB, 1, IEND>(lds, pinv, gi, u, wci, active);
        emit(0, u);
    }

    // phases 2 and 3
    for (int ia = 0; ia < (ONE ? 0 : ba.ka); ia++) {
        int u = tid;
        asm volatile("" : "+v"(u));
        __syncthreads();
        gather(xa, ba.wa, ia, ba.ka, (int)ba.oa, u);
        run_steps_ct_active<T, MODE_STD_FWD, E, NT, S, RB, FS0, NST - 1>(lds, pf, g, u, wc,
                                                                       active);
        const int dg = __builtin_amdgcn_readfirstlane(sc.dgroup[ia]);
        const bool dl = __builtin_amdgcn_readfirstlane(sc.dload[ia]) != 0;
        const int dslot = dl ? __builtin_amdgcn_readfirstlane(sc.gslot[dg]) : 0;
        if (active) {
#pragma unroll
            for (int kk = 0; kk < NG; kk++) {
                const int rbase = (u + kk * NT) << RB;
                cplx<T> v[1 << RB];
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) v[j] = lds[swz(rbase | j)];
                group_stages_derived<T, MODE_STD_FWD, RB>(v, pf, g, rbase, 0, 0, wc);
                // products with the b-limb spectra, CH elements at a time;
                // each round issues all its stash loads before its stores
                constexpr int CH = (1 << RB) < 8 ? (1 << RB) : 8;
#pragma unroll
                for (int j0 = 0; j0 < (1 << RB); j0 += CH) {
                    const int e0 = (kk << RB) + j0;
                    cplx<T> w[CH];
                    if (dl) {
                        const gptr<cplx<T>> ds = stash_at<NT>(stash, dslot, e0, u);
#pragma unroll
                        for (int jj = 0; jj < CH; jj++) w[jj] = ds[jj * NT];
                    } else {
#pragma unroll
                        for (int jj = 0; jj < CH; jj++) w[jj] = mk<T>(T(0), T(0));
                    }
                    for (int jb = 0; jb < ba.kb; jb++) {
                        const int k = ia * ba.kb + jb;
                        const int cg = __builtin_amdgcn_readfirstlane(sc.cgroup[k]);
                        const bool regs = cg == dg;
                        const bool first = __builtin_amdgcn_readfirstlane(sc.cfirst[k]) != 0;
                        const int cslot = regs ? 0 : __builtin_amdgcn_readfirstlane(sc.gslot[cg]);
                        const gptr<cplx<T>> slot = stash_at<NT>(stash, cslot, e0, u);
                        cplx<T> bj[CH], c[CH];
                        if (jb == 0) {
#pragma unroll
                            for (int jj = 0; jj < CH; jj++) bj[jj] = breg[e0 + jj];
                        } else {
                            const gptr<cplx<T>> bsp = stash_at<NT>(stash, jb - 1, e0, u);
#pragma unroll
                            for (int jj = 0; jj < CH; jj++) bj[jj] = bsp[jj * NT];
                        }
#pragma unroll
                        for (int jj = 0; jj < CH; jj++) c[jj] = v[j0 + jj] * bj[jj];
                        if (regs) {
#pragma unroll
                            for (int jj = 0; jj < CH; jj++) w[jj] = w[jj] + c[jj];
                        } else if (first) {
#pragma unroll
                            for (int jj = 0; jj < CH; jj++) slot[jj * NT] = c[jj];
                        } else {
                            cplx<T> old[CH];
#pragma unroll
                            for (int jj = 0; jj < CH; jj++) old[jj] = slot[jj * NT];
#pragma unroll
                            for (int jj = 0; jj < CH; jj++) slot[jj * NT] = old[jj] + c[jj];
                        }
                    }
#pragma unroll
                    for (int jj = 0; jj < CH; jj++) v[j0 + jj] = w[jj];
                }
                if (dg >= 0) {
                    group_stages_derived<T, MODE_STD_INV, RB>(v, pinv, gi, rbase, 0, 0, wci);
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) lds[swz(rbase | j)] = v[j];
                }
            }
        }
        if (dg >= 0) {
            run_steps_ct_active<T, MODE_STD_INV, E, NT, S, RB, 1, IEND>(lds, pinv, gi, u, wci,
                                                                         active);
            emit(dg, u);
        }
        const int ngrp = __builtin_amdgcn_readfirstlane(sc.G);
        for (int grp = 0; grp < ngrp; grp++) {
            if (__builtin_amdgcn_readfirstlane(sc.gdone[grp]) != ia || grp == dg) continue;
            const int gs = __builtin_amdgcn_readfirstlane(sc.gslot[grp]);
            __syncthreads();
            if (active) {
#pragma unroll
                for (int kk = 0; kk < NG; kk++) {
                    const int rbase = (u + kk * NT) << RB;
                    cplx<T> v[1 << RB];
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++)
                        v[j] = *stash_at<NT>(stash, gs, (kk << RB) + j, u);
                    group_stages_derived<T, MODE_STD_INV, RB>(v, pinv, gi, rbase, 0, 0, wci);
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) lds[swz(rbase | j)] = v[j];
                }
            }
            run_steps_ct_active<T, MODE_STD_INV, E, NT, S, RB, 1, IEND>(lds, pinv, gi, u, wci,
                                                                         active);
            emit(grp, u);
        }
    }
    if (SB == 0) {
        err_update(ba.err, prod, emax);
    } else if (ba.cmb) {
        // publish this sub-problem; the last of the product's Q workgroups
        // combines them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            const int prev = __hip_atomic_fetch_add(ba.arrive + prod, 1, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
            *sh_last = (prev == Q - 1) ? 1 : 0;
        }
        __syncthreads();
        if (*sh_last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            float em = 0.f;
            // 2^S / BT iterations for every thread
            for (int c = tid; c < (1 << S); c += BT) em = fmaxf(em, combine_pos_rt<T>(ba, pr, tb, c));
            err_update(ba.err, prod, em);
        }
    }
}

// The kernel's argument structure in the kernel-argument segment (it is
// the kernel's only argument, at offset 0).  Calls take it from there: a
// reference to the by-value argument would make every wave keep a private
// copy of it, written and reloaded through the vector memory path.
typedef __attribute__((address_space(4))) const BArgs* kargs_t;

__device__ __forceinline__ kargs_t kernel_args()
{
    return (kargs_t)__builtin_amdgcn_kernarg_segment_ptr();
}

// A workgroup-uniform structure from constant memory (scalar loads).
template <typename V> __device__ __forceinline__ V load_uniform(__attribute__((address_space(4))) const V* p)
{
    static_assert(sizeof(V) % 4 == 0, "whole dwords");
    typedef __attribute__((address_space(4))) const unsigned* cu32;
    const cu32 w = (cu32)p;
    unsigned v[sizeof(V) / 4];
#pragma unroll
    for (unsigned i = 0; i < sizeof(V) / 4; i++) v[i] = w[i];
    V r;
    __builtin_memcpy(&r, v, sizeof(V));
    return r;
}

// Small items (S <= 12) as calls, which keeps the kernel's code size and
// compile time in bounds; the large ones inlined.  The callee takes only
// scalars: the arguments from the kernel-argument segment, the item's
// product and table records from memory by index, each made a scalar copy
// (a buffer load through a table pointer held in vector registers would
// become a loop over lanes).
template <typename T, int KIND, int S, int E, int RB, int TILE>
__device__ __noinline__ void run_item_call(cplx<T>* lds, kargs_t kba, int prod_in, int q_in,
                                           int tid, int* sh)
{
    const BArgs ba = load_uniform(kba);
    const int prod = __builtin_amdgcn_readfirstlane(prod_in);
    const int q = __builtin_amdgcn_readfirstlane(q_in);
    const BProd pr = uniform_copy(ba.prods[prod]);
    const BTables tb = uniform_copy(ba.tabs[pr.tabset]);
    run_item<T, KIND, S, E, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh);
}

template <typename T, int KIND, int S, int RB, int TILE>
__device__ __forceinline__ void dispatch_sb(cplx<T>* lds, const BArgs& ba, const BProd& pr,
                                            const BTables& tb, int prod, int q, int tid, int* sh)
{
    if constexpr (S <= TILE) {
        constexpr int E = S >= 14 ? 5 : (S == 13 ? 4 : 3);  // NT <= 512
        constexpr int R = RB < E ? RB : E;
        // FP64: every item inlined (the FP32 kernel's larger code keeps the
        // small items as calls)
        if constexpr (S >= TILE - 1 || sizeof(T) == 8)
            run_item<T, KIND, S, E, R, TILE>(lds, ba, pr, tb, prod, q, tid, sh);
        else
            run_item_call<T, KIND, S, E, R, TILE>(lds, kernel_args(), prod, q, tid, sh);
    }
}

template <typename T, int KIND, int RB, int TILE>
__device__ __forceinline__ void dispatch_s(cplx<T>* lds, const BArgs& ba, const BProd& pr,
                                           const BTables& tb, int prod, int q, int tid, int* sh)
{
    switch (pr.S) {
        case 5: dispatch_sb<T, KIND, 5, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 6: dispatch_sb<T, KIND, 6, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 7: dispatch_sb<T, KIND, 7, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 8: dispatch_sb<T, KIND, 8, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 9: dispatch_sb<T, KIND, 9, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 10: dispatch_sb<T, KIND, 10, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 11: dispatch_sb<T, KIND, 11, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 12: dispatch_sb<T, KIND, 12, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 13: dispatch_sb<T, KIND, 13, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        case 14: dispatch_sb<T, KIND, 14, RB, TILE>(lds, ba, pr, tb, prod, q, tid, sh); break;
        default: break;
    }
}

// FP32: TILE = 14, one 128 KiB workgroup per CU; TILE = 13, 64 KiB
// workgroups, compiled for 4 waves per SIMD so that two of them share a CU.
// FP64: TILE = 13, one 128 KiB workgroup per CU.
template <typename T, int RB, int TILE>
__global__ __launch_bounds__(BT)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && TILE == 13 ? 4 : 1)))
void batched_kernel(BArgs ba)
{
    // FP64: the tile, then the item's top-row coefficients (2^SBMAX) and
    // split twist tables (at most 2^(TILE-7) + 128 entries)
    constexpr int EXTRA = sizeof(T) == 8 ? (1 << SBMAX) + (1 << (TILE - 7)) + 128 : 0;
    __shared__ cplx<T> lds[(1 << TILE) + EXTRA];
    __shared__ int sh_item;
    __shared__ int sh_last;
    const int tid = threadIdx.x;
    const int home = (int)(blockIdx.x % NQ);
    for (int r = 0; r < NQ; r++) {
        const int x = (home + r) % NQ;
        for (;;) {
            if (tid == 0) sh_item = atomicAdd(ba.next_item + x, 1);
            __syncthreads();
            const int it = sh_item;
            __syncthreads();
            if (it >= ba.qlen[x]) break;  // queue x is exhausted
            // the item's records are read from memory, which the compiler
            // cannot prove uniform: scalar copies keep the table resources in
            // scalar registers (a buffer load through a resource in vector
            // registers becomes a loop over lanes)
            const int2 it2 = ba.items[ba.qbeg[x] + it];
            const int2 item = make_int2(__builtin_amdgcn_readfirstlane(it2.x),
                                        __builtin_amdgcn_readfirstlane(it2.y));
            const BProd pr = uniform_copy(ba.prods[item.x]);
            const BTables tb = uniform_copy(ba.tabs[pr.tabset]);
            if (pr.kind == KIND_NEGA)
                dispatch_s<T, KIND_NEGA, RB, TILE>(lds, ba, pr, tb, item.x, item.y, tid, &sh_last);
            else
                dispatch_s<T, KIND_LINEAR, RB, TILE>(lds, ba, pr, tb, item.x, item.y, tid, &sh_last);
            __syncthreads();
        }
    }
    // The last workgroup to leave clears the queue, arrival and exit
    // counters for the next launch (every other workgroup has made its last
    // atomic by then), so a launch needs no separate clearing kernel.
    __syncthreads();
    if (tid == 0) sh_last = atomicAdd(ba.exited, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (sh_last) {
        for (int i = tid; i < NQ; i += BT) atomicExch(ba.next_item + i, 0);
        for (int i = tid; i < ba.count; i += BT) atomicExch(ba.arrive + i, 0);
        if (tid == 0) atomicExch(ba.exited, 0);
    }
}

// Positions cbase + t and cbase + CT + t of one split product (t the
// thread, cbase the workgroup's first of 2 CT positions).
template <typename T, int SB, bool UL>
__device__ __forceinline__ float combine_two(const BArgs& ba, const BProd& pr, const BTables& tb,
                                             const cplx<T>* uhi, const cplx<T>* ulo, int cbase)
{
    const int c = cbase + threadIdx.x;
    return fmaxf(combine_pos<T, SB, UL>(ba, pr, tb, c, uhi, ulo, cbase),
                 combine_pos<T, SB, UL>(ba, pr, tb, c + CT, uhi, ulo, cbase));
}

// The 2 CT positions from cbase < 2^S (S the tile) of one split product
// (blockIdx.x: the product's index in the split list, blockIdx.y: cbase).
// FP64: the untwist factors come from LDS, untwist[P] = untwist[P & ~127]
// untwist[P & 127] (exp(-2 pi i P / 2N) is multiplicative in P; the FP32
// form keeps the direct table entries, whose rounding its limb error
// statistics were measured with).
template <typename T> __global__ __launch_bounds__(CT) void batched_combine_kernel(BArgs ba)
{
    constexpr bool UL = sizeof(T) == 8;
    __shared__ cplx<T> ulo[UL ? 128 : 1];
    __shared__ cplx<T> uhi[UL ? 4 << SBMAX : 1];
    const int prod = __builtin_amdgcn_readfirstlane(ba.split[blockIdx.x]);
    const BProd pr = uniform_copy(ba.prods[prod]);
    const BTables tb = uniform_copy(ba.tabs[pr.tabset]);
    const int cbase = blockIdx.y * CT * 2;  // this workgroup's 512 positions (4 blocks of 128)
    const int c = cbase + threadIdx.x * 2;
    if constexpr (UL) {
        const cplx<T>* ut = tabp<T>(tb.untwist);
        const int t = threadIdx.x;
        if (t < 128) ulo[t] = ut[t];
        if (t < (4 << pr.SB)) uhi[t] = ut[((t >> 2) << pr.S) | (cbase + ((t & 3) << 7))];
        __syncthreads();
    }
    float em;
    if constexpr (sizeof(T) == 8) {
        // FP64: positions cbase + t and cbase + CT + t, so that the lanes of
        // every 16-byte scratch load and 8-byte output store touch
        // consecutive elements (pairs of adjacent positions per thread put
        // the lanes of each load 32 bytes apart)
        switch (pr.SB) {
            case 1: em = combine_two<T, 1, UL>(ba, pr, tb, uhi, ulo, cbase); break;
            case 2: em = combine_two<T, 2, UL>(ba, pr, tb, uhi, ulo, cbase); break;
            case 3: em = combine_two<T, 3, UL>(ba, pr, tb, uhi, ulo, cbase); break;
            default: em = combine_two<T, 4, UL>(ba, pr, tb, uhi, ulo, cbase); break;
        }
    } else {
        // FP32: a pair of adjacent positions is one 16-byte load per row and
        // one 16-byte store per output pair, consecutive across the lanes
        switch (pr.SB) {
            case 1: em = combine_pos2<T, 1, UL>(ba, pr, tb, c, uhi, ulo, cbase); break;
            case 2: em = combine_pos2<T, 2, UL>(ba, pr, tb, c, uhi, ulo, cbase); break;
            case 3: em = combine_pos2<T, 3, UL>(ba, pr, tb, c, uhi, ulo, cbase); break;
            default:  // 2^4 rows: one position at a time (registers), consecutive per lane
                em = combine_two<T, 4, UL>(ba, pr, tb, uhi, ulo, cbase);
                break;
        }
    }
    err_update(ba.err, prod, em);
}

}  // namespace

struct gpufft_batched_plan_s {
    int prec;         // GPUFFT_PREC_F32_SPLIT (limbs) or GPUFFT_PREC_F64 (one limb each)
    int count;
    int n_items;
    int n_split;      // products with SB > 0
    int wa, ka, wb, kb;
    int rb;
    int grid;
    int cmb;          // 1: combine split products in the persistent kernel
    int tile;         // log2 of the LDS tile: FP32 14 (128 KiB) or 13 (64 KiB, two per CU); FP64 13
    int max_n;        // largest log2 L of the list
    int out_even;     // every out_offset is even
    long long scratch_elems;  // allocated scratch (cplx<T>)
    int qbeg[NQ], qlen[NQ];
    BProd* d_prods;
    int2* d_items;
    BTables* d_tabs;
    int* d_counters;  // [0, NQ) queue counters, [NQ, NQ + count) arrival counters, then the exit counter
    int* d_split;
    BSched sched;
    BSched* d_sched;
    int stash_slots;
    int stash_grid;   // workgroups the stash is sized for
    void* stash;
    void* scratch;
    void* wc;
    void* wci;
    std::vector<void*> tables;
    std::vector<BProd> prods;
    hipGraphExec_t graph;
    const void *g_a, *g_b, *g_out, *g_err;
    void* g_stream;
};

namespace {

size_t elem_bytes(const gpufft_batched_plan_s* p)
{
    return p->prec == GPUFFT_PREC_F64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
}

void release(gpufft_batched_plan_s* p)
{
    if (!p) return;
    if (p->graph) (void)hipGraphExecDestroy(p->graph);
    for (void* t : p->tables) (void)hipFree(t);
    void* bufs[10] = {p->d_prods, p->d_items, p->d_tabs, p->d_counters, p->d_split, p->d_sched,
                      p->stash, p->scratch, p->wc, p->wci};
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

// Groups of equal weight, the order of their inverse transforms and the
// stash slots of their partial sums (a slot is reused once its group has
// been inverted, by a group whose first contribution comes with a later a
// limb).  Returns the number of group slots.
int build_sched(int wa, int ka, int wb, int kb, BSched& sc)
{
    std::memset(&sc, 0, sizeof(sc));
    std::map<int, int> gid;
    std::vector<int> gfirst;  // a limb of each group's first contribution
    for (int ia = 0; ia < ka; ia++)
        for (int jb = 0; jb < kb; jb++) {
            const int w = wa * ia + wb * jb;
            auto f = gid.find(w);
            if (f == gid.end()) {
                f = gid.emplace(w, sc.G).first;
                sc.gshift[sc.G] = w;
                gfirst.push_back(ia);
                sc.cfirst[ia * kb + jb] = 1;
                sc.G++;
            }
            sc.cgroup[ia * kb + jb] = f->second;
            sc.gdone[f->second] = ia;  // the last a limb seen so far
        }
    for (int ia = 0; ia < ka; ia++) {
        // prefer a group that already has a partial sum: it is read anyway
        sc.dgroup[ia] = -1;
        for (int g = 0; g < sc.G; g++)
            if (sc.gdone[g] == ia && (sc.dgroup[ia] < 0 || (gfirst[g] < ia && gfirst[sc.dgroup[ia]] == ia)))
                sc.dgroup[ia] = g;
        sc.dload[ia] = sc.dgroup[ia] >= 0 && gfirst[sc.dgroup[ia]] < ia ? 1 : 0;
    }
    // slot intervals [gfirst, gdone], groups in order of first contribution
    std::vector<int> slot_free_after;  // gdone of the slot's last group
    for (int g = 0; g < sc.G; g++) {
        const bool direct = sc.dgroup[sc.gdone[g]] == g;
        sc.gslot[g] = -1;
        if (direct && gfirst[g] == sc.gdone[g]) continue;  // summed in registers only
        int s = -1;
        for (int k = 0; k < (int)slot_free_after.size() && s < 0; k++)
            if (slot_free_after[k] < gfirst[g]) s = k;
        if (s < 0) {
            s = (int)slot_free_after.size();
            slot_free_after.push_back(0);
        }
        slot_free_after[s] = sc.gdone[g];
        sc.gslot[g] = s;
    }
    // a group's contributions before its completing limb go to its slot; the
    // first of them is a store
    for (int k = 0; k < ka * kb; k++) sc.cfirst[k] = 0;
    std::vector<int> seen(sc.G, 0);
    for (int ia = 0; ia < ka; ia++)
        for (int jb = 0; jb < kb; jb++) {
            const int g = sc.cgroup[ia * kb + jb];
            const bool to_regs = sc.dgroup[ia] == g;
            if (!to_regs && !seen[g]) sc.cfirst[ia * kb + jb] = 1;
            if (!to_regs) seen[g] = 1;
        }
    return (int)slot_free_after.size();
}

// Sub-problems, work queues, split list and scratch for LDS tile 2^tile:
// each product is cut into 2^SB sub-problems of 2^S points, S = min(n, tile).
// Products go largest first (cost ~ L log L), each to the least loaded queue;
// a queue lists its products' sub-problems in that order.  Synchronises the
// stream (the host vectors are freed on return).
int build_layout(gpufft_batched_plan_s* p, int tile, hipStream_t st)
{
    const char* who = "gpufft_batched_plan: layout";
    const int count = p->count;
    long long scratch = 0;
    for (BProd& pr : p->prods) {
        pr.S = std::min(pr.n, tile);
        pr.SB = pr.n - pr.S;
        pr.scratch_off = 0;
        if (pr.SB > 0) {
            pr.scratch_off = scratch;
            scratch += (long long)p->sched.G << pr.n;
        }
    }
    auto cost = [&](int i) { return (long long)p->prods[i].n << p->prods[i].n; };
    std::vector<int> order(count);
    for (int i = 0; i < count; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
    std::vector<int> queue_of(count);
    long long load[NQ] = {};
    for (int i : order) {
        const int x = (int)(std::min_element(load, load + NQ) - load);
        queue_of[i] = x;
        load[x] += cost(i);
    }
    std::vector<int2> items;
    std::vector<int> split;
    for (int x = 0; x < NQ; x++) {
        p->qbeg[x] = (int)items.size();
        for (int i : order)
            if (queue_of[i] == x)
                for (int q = 0; q < (1 << p->prods[i].SB); q++) items.push_back(make_int2(i, q));
        p->qlen[x] = (int)items.size() - p->qbeg[x];
    }
    for (int i : order)
        if (p->prods[i].SB > 0) split.push_back(i);
    void* old[3] = {p->d_prods, p->d_items, p->d_split};
    for (void* b : old)
        if (b) (void)hipFree(b);
    p->d_prods = nullptr;
    p->d_items = nullptr;
    p->d_split = nullptr;
    int s;
    if ((s = upload(&p->d_prods, p->prods, st)) || (s = upload(&p->d_items, items, st)) ||
        (s = upload(&p->d_split, split, st)))
        return s;
    if (scratch > p->scratch_elems) {
        if (p->scratch) (void)hipFree(p->scratch);
        p->scratch = nullptr;
        p->scratch_elems = 0;
        hipError_t e = hipMalloc(&p->scratch, elem_bytes(p) * scratch);
        if (e != hipSuccess) return fail_hip(e, who);
        p->scratch_elems = scratch;
    }
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail_hip(e, who);
    p->n_items = (int)items.size();
    p->n_split = (int)split.size();
    p->tile = tile;
    return GPUFFT_OK;
}

int alloc_stash(gpufft_batched_plan_s* p, int grid)
{
    if (p->stash_slots == 0 || grid <= p->stash_grid) return GPUFFT_OK;
    // the new stash first: on failure the plan keeps its old stash (and any
    // captured graph that points to it) unchanged and usable
    void* fresh = nullptr;
    hipError_t e = hipMalloc(&fresh, elem_bytes(p) * ((size_t)grid * p->stash_slots << SMAX));
    if (e != hipSuccess) {
        (void)hipGetLastError();  // the failure is reported here, not by the next launch
        return fail_hip(e, "gpufft_batched_plan: stash");
    }
    if (p->graph) {  // captured with the old stash's address
        (void)hipGraphExecDestroy(p->graph);
        p->graph = nullptr;
    }
    if (p->stash) (void)hipFree(p->stash);
    p->stash = fresh;
    p->stash_grid = grid;
    return GPUFFT_OK;
}

int launch(gpufft_batched_plan_s* p, const int* a, const int* b, long long* out, double* err,
           hipStream_t st)
{
    const char* who = "gpufft_batched_execute";
    // the counters are zero: cleared at plan creation and by the end of every launch
    if (err) {
        hipLaunchKernelGGL(zero_doubles, dim3((p->count + 255) / 256), dim3(256), 0, st, err, p->count);
        GFA_CHECK_LAUNCH(who);
    }
    BArgs ba{};
    ba.prods = p->d_prods;
    ba.items = p->d_items;
    ba.tabs = p->d_tabs;
    for (int x = 0; x < NQ; x++) {
        ba.qbeg[x] = p->qbeg[x];
        ba.qlen[x] = p->qlen[x];
    }
    ba.next_item = p->d_counters;
    ba.arrive = p->d_counters + NQ;
    ba.exited = p->d_counters + NQ + p->count;
    ba.count = p->count;
    ba.split = p->d_split;
    ba.cmb = p->cmb;
    ba.a = a;
    ba.b = b;
    ba.out = out;
    ba.scratch = p->scratch;
    ba.sched = p->d_sched;
    ba.stash = p->stash;
    ba.stash_slots = p->stash_slots;
    ba.wc = p->wc;
    ba.wci = p->wci;
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
    ba.out_pair = p->out_even && ((uintptr_t)out & 15) == 0;
    const dim3 grid(p->grid), block(BT);
    if (p->prec == GPUFFT_PREC_F64) {
        if (p->rb == 3)
            hipLaunchKernelGGL((batched_kernel<double, 3, 13>), grid, block, 0, st, ba);
        else
            hipLaunchKernelGGL((batched_kernel<double, 4, 13>), grid, block, 0, st, ba);
    } else if (p->tile == 13) {
        switch (p->rb) {
            case 3: hipLaunchKernelGGL((batched_kernel<float, 3, 13>), grid, block, 0, st, ba); break;
            case 4: hipLaunchKernelGGL((batched_kernel<float, 4, 13>), grid, block, 0, st, ba); break;
            default: hipLaunchKernelGGL((batched_kernel<float, 5, 13>), grid, block, 0, st, ba); break;
        }
    } else {
        switch (p->rb) {
            case 3: hipLaunchKernelGGL((batched_kernel<float, 3, 14>), grid, block, 0, st, ba); break;
            case 4: hipLaunchKernelGGL((batched_kernel<float, 4, 14>), grid, block, 0, st, ba); break;
            default: hipLaunchKernelGGL((batched_kernel<float, 5, 14>), grid, block, 0, st, ba); break;
        }
    }
    GFA_CHECK_LAUNCH(who);
    if (!p->cmb && p->n_split > 0) {
        const dim3 cgrid(p->n_split, (1 << p->tile) / (2 * CT));
        if (p->prec == GPUFFT_PREC_F64)
            hipLaunchKernelGGL(batched_combine_kernel<double>, cgrid, dim3(CT), 0, st, ba);
        else
            hipLaunchKernelGGL(batched_combine_kernel<float>, cgrid, dim3(CT), 0, st, ba);
        GFA_CHECK_LAUNCH(who);
    }
    return GPUFFT_OK;
}

int create_plan(gpufft_batched_plan* out, const gpufft_product_desc* descs, int count,
                int a_limb_bits, int a_limbs, int b_limb_bits, int b_limbs, int prec,
                void* stream, const char* who)
{
    if (!out || !descs) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL argument");
    *out = nullptr;
    if (count < 1) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": count < 1");
    if (a_limbs < 1 || b_limbs < 1 || a_limbs * b_limbs > MAXP ||
        (a_limbs > 1 && (a_limb_bits < 1 || a_limb_bits > 16)) ||
        (b_limbs > 1 && (b_limb_bits < 1 || b_limb_bits > 16)) ||
        (a_limbs > 1 && (a_limbs - 1) * a_limb_bits > 30) ||
        (b_limbs > 1 && (b_limbs - 1) * b_limb_bits > 30))
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": invalid limb split");
    hipStream_t st = (hipStream_t)stream;
    const bool f64 = prec == GPUFFT_PREC_F64;
    const int tmax = f64 ? SMAX64 : SMAX;  // default (and FP64's only) tile
    gpufft_batched_plan_s* p = new gpufft_batched_plan_s();
    p->prec = prec;
    p->count = count;
    p->wa = a_limbs > 1 ? a_limb_bits : 0;
    p->ka = a_limbs;
    p->wb = b_limbs > 1 ? b_limb_bits : 0;
    p->kb = b_limbs;
    p->rb = f64 ? 4 : 5;
    p->cmb = 0;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // persistent grid: one 128 KiB-LDS workgroup per CU is resident at a time
    p->grid = cus;

    std::map<int, int> tabset;  // n -> table set (both kinds use the FFNT tables of L = 2^n)
    std::vector<BTables> tabs;
    // group slots after the kb - 1 stashed b-limb spectra
    const int gslots = build_sched(p->wa, p->ka, p->wb, p->kb, p->sched);
    for (int g = 0; g < p->sched.G; g++)
        if (p->sched.gslot[g] >= 0) p->sched.gslot[g] += p->kb - 1;
    p->stash_slots = p->kb - 1 + gslots;
    for (int i = 0; i < count; i++) {
        const gpufft_product_desc& d = descs[i];
        BProd pr{};
        if (d.kind != GPUFFT_PRODUCT_NEGACYCLIC && d.kind != GPUFFT_PRODUCT_LINEAR) {
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": unknown product kind");
        }
        pr.kind = d.kind == GPUFFT_PRODUCT_NEGACYCLIC ? KIND_NEGA : KIND_LINEAR;
        // NEGACYCLIC: FFNT of ring N (L = N/2); LINEAR: FFNT of ring 2N (L = N)
        pr.n = pr.kind == KIND_NEGA ? d.ring_power - 1 : d.ring_power;
        if (pr.n < SMIN || pr.n > tmax + SBMAX) {
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": ring_power out of range");
        }
        if (d.a_offset < 0 || d.b_offset < 0 || d.out_offset < 0) {
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": negative offset");
        }
        if ((d.a_offset | d.b_offset) & 3) {  // 16-byte operand loads
            release(p);
            return fail(GPUFFT_ERR_INVALID_ARGUMENT,
                        std::string(who) + ": a_offset and b_offset must be multiples of 4");
        }
        pr.a_off = d.a_offset;
        pr.b_off = d.b_offset;
        pr.N = 1 << d.ring_power;
        pr.out_off = d.out_offset;
        if (i == 0) p->out_even = 1;
        if (d.out_offset & 1) p->out_even = 0;
        auto f = tabset.find(pr.n);
        if (f == tabset.end()) {
            BTables t{};
            void* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
            for (int k = 0; k < 4; k++) {
                hipError_t e = hipMalloc(&bufs[k], elem_bytes(p) << pr.n);
                if (e != hipSuccess) {
                    release(p);
                    return fail_hip(e, who);
                }
                p->tables.push_back(bufs[k]);
            }
            auto gen = f64 ? gpufft_generate_table_f64 : gpufft_generate_table_f32;
            int s = 0;
            s |= gen(bufs[0], GPUFFT_TABLE_FFNT_ROOT, pr.n, 0, st);
            s |= gen(bufs[1], GPUFFT_TABLE_FFNT_ROOT, pr.n, 1, st);
            s |= gen(bufs[2], GPUFFT_TABLE_FFNT_TWIST, pr.n, 0, st);
            s |= gen(bufs[3], GPUFFT_TABLE_FFNT_TWIST, pr.n, 1, st);
            if (s) {
                release(p);
                return s;
            }
            t.root = bufs[0];
            t.iroot = bufs[1];
            t.twist = bufs[2];
            t.untwist = bufs[3];
            f = tabset.emplace(pr.n, (int)tabs.size()).first;
            tabs.push_back(t);
        }
        pr.tabset = f->second;
        p->max_n = std::max(p->max_n, pr.n);
        p->prods.push_back(pr);
    }
    const std::vector<BSched> schedv(1, p->sched);
    int s;
    if ((s = upload(&p->d_tabs, tabs, st)) || (s = upload(&p->d_sched, schedv, st)) ||
        (s = build_layout(p, tmax, st)) || (s = alloc_stash(p, p->grid))) {
        release(p);
        return s;
    }
    hipError_t e = hipMalloc((void**)&p->d_counters, sizeof(int) * (count + NQ + 1));
    if (e == hipSuccess) e = hipMalloc(&p->wc, elem_bytes(p) * 32);
    if (e == hipSuccess) e = hipMalloc(&p->wci, elem_bytes(p) * 32);
    if (e != hipSuccess) {
        release(p);
        return fail_hip(e, who);
    }
    hipLaunchKernelGGL(zero_ints, dim3((count + NQ + 1 + 255) / 256), dim3(256), 0, st,
                       p->d_counters, count + NQ + 1);
    e = hipGetLastError();
    if (e != hipSuccess) {
        release(p);
        return fail_hip(e, who);
    }
    auto gen = f64 ? gpufft_generate_table_f64 : gpufft_generate_table_f32;
    if ((s = gen(p->wc, GPUFFT_TABLE_STAGE_ROOTS, 5, 0, st)) ||
        (s = gen(p->wci, GPUFFT_TABLE_STAGE_ROOTS, 5, 1, st))) {
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

}  // namespace

extern "C" {

int gpufft_batched_plan_create(gpufft_batched_plan* out, const gpufft_product_desc* descs, int count,
                               int a_limb_bits, int a_limbs, int b_limb_bits, int b_limbs,
                               void* stream)
{
    return create_plan(out, descs, count, a_limb_bits, a_limbs, b_limb_bits, b_limbs,
                       GPUFFT_PREC_F32_SPLIT, stream, "gpufft_batched_plan_create");
}

int gpufft_batched_plan_create_f64(gpufft_batched_plan* out, const gpufft_product_desc* descs,
                                   int count, void* stream)
{
    return create_plan(out, descs, count, 0, 1, 0, 1, GPUFFT_PREC_F64, stream,
                       "gpufft_batched_plan_create_f64");
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
    if (radix_bits < 3 || radix_bits > (plan->prec == GPUFFT_PREC_F64 ? 4 : 5) || grid < 1)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": invalid tuning");
    if (int s = alloc_stash(plan, grid)) return s;
    plan->rb = radix_bits;
    plan->grid = grid;
    if (plan->graph) {
        (void)hipGraphExecDestroy(plan->graph);
        plan->graph = nullptr;
    }
    return ok();
}

int gpufft_batched_plan_set_combine(gpufft_batched_plan plan, int in_kernel)
{
    const char* who = "gpufft_batched_plan_set_combine";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (in_kernel != 0 && in_kernel != 1)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": in_kernel must be 0 or 1");
    plan->cmb = in_kernel;
    if (plan->graph) {
        (void)hipGraphExecDestroy(plan->graph);
        plan->graph = nullptr;
    }
    return ok();
}

int gpufft_batched_plan_set_tile(gpufft_batched_plan plan, int log2_tile)
{
    const char* who = "gpufft_batched_plan_set_tile";
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": plan is NULL");
    if (plan->prec == GPUFFT_PREC_F64 && log2_tile != SMAX64)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": FP64 plans have log2_tile 13");
    if (log2_tile != 13 && log2_tile != 14)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": log2_tile must be 13 or 14");
    if (plan->max_n - log2_tile > SBMAX)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT,
                    std::string(who) + ": the largest product needs more than 2^4 sub-problems");
    if (log2_tile == plan->tile) return ok();
    if (plan->graph) {
        (void)hipGraphExecDestroy(plan->graph);
        plan->graph = nullptr;
    }
    if (int s = build_layout(plan, log2_tile, nullptr)) return s;
    return ok();
}

int gpufft_batched_plan_get_tile(gpufft_batched_plan plan, int* log2_tile)
{
    const char* who = "gpufft_batched_plan_get_tile";
    if (!plan || !log2_tile) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL argument");
    *log2_tile = plan->tile;
    return ok();
}

int gpufft_batched_plan_get_combine(gpufft_batched_plan plan, int* in_kernel)
{
    const char* who = "gpufft_batched_plan_get_combine";
    if (!plan || !in_kernel) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL argument");
    *in_kernel = plan->cmb;
    return ok();
}

int gpufft_batched_plan_info(gpufft_batched_plan plan, int* n_items, long long* scratch_elems,
                             int* radix_bits, int* grid)
{
    if (!plan) return fail(GPUFFT_ERR_NULL_POINTER, "gpufft_batched_plan_info: plan is NULL");
    if (n_items) *n_items = plan->n_items;
    long long s = 0;
    for (const BProd& pr : plan->prods)
        if (pr.SB > 0) s += (long long)plan->sched.G << pr.n;
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
    if (!stream)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT,
                    std::string(who) + ": graph capture needs a stream other than the default stream");
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
