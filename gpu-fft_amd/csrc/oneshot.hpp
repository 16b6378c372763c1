// Transforms of 2^n points, n = S + SB (S: the LDS tile, 2^13 double /
// 2^14 float points; SB = 1..3), for the out-of-place drop-in calls (R_R and
// FFNT): the forward in one launch, the inverse in two.
//
// A transform is cut into Q = 2^SB sub-problems of 2^S points, one workgroup
// each.  A forward workgroup reads the whole transform (the Q workgroups of a
// transform get block ids congruent mod 8, so they run on one XCD, close in
// time, and the repeated reads are served by its L2) and evaluates, for its
// own row only, the Q-point network that couples the sub-problems:
// the forward (Cooley-Tukey, natural in, bit-reversed out) couplings are the
// first SB stages (bits n-1 .. S), between positions j 2^S + c, with
// coefficients uniform over the workgroup (the fused products' gather,
// fused.hpp); sub-problem q owns the output block [q 2^S, (q+1) 2^S).
// The other S stages of a sub-problem are radix-2^RB steps through LDS with
// twiddles derived from one table load per step (in-step roots tab[t], see
// group_stages_derived); the short top step runs in registers in the gather.
// HBM sees every input once and every output once.
//
// The inverse (Gentleman-Sande, bit-reversed in, natural out) is two
// launches: the stages inside each contiguous sub-problem (its bits 0 .. S-1)
// in place, then the SB coupling stages with the output transform as a
// streaming pass.  (A one-launch inverse, partitioned by the low bits of the
// output index so that its coupling stages come first, writes every Q-th
// output, and those partial-line stores cost more than the second pass.)
// Stage order and twiddle indices are
// those of the reference kernels (ForwardCore / InverseCore,
// src/lib/fft.cu:11-260, and the FFNT cores, fft.cu:1013-1384), so results
// match the pass chain up to rounding.
#pragma once

#include "fused.hpp"

namespace gfa {

template <typename T> struct OneArgs {
    int n;       // log2 transform length
    int batch;   // transforms (products) in the launch
    int nega;    // X^N+1 table layout (base offsets)
    const cplx<T>* tab;      // root table of this direction (reference layout)
    const cplx<T>* twist;    // PL_FOLD: FFNT twist table (2^n entries)
    const void* src;
    const void* src2;        // second operand (LD_MUL)
    long long src_stride;    // elements between transforms
    void* dst;
    long long dst_stride;
    const cplx<T>* untwist;  // ST_UNFOLD: untwist table (2^n entries)
    int do_scale;
    cplx<T> scale;
};

// XCD-aware block order: the Q workgroups of a transform get block ids
// congruent mod 8 (consecutive ids go to consecutive XCDs).
template <int Q> __device__ __forceinline__ void one_block(int& row, long long& t)
{
    const int x = blockIdx.x;
    row = (x >> 3) % Q;
    t = (long long)((x >> 3) / Q) * 8 + (x & 7);
}

template <typename T, int S, int SB, int E, int RB, int PL>
__global__ __launch_bounds__((1 << S) >> E) void fwd_sub_kernel(OneArgs<T> o)
{
    constexpr int TILE = 1 << S;
    constexpr int EPT = 1 << E;
    constexpr int NT = TILE / EPT;
    constexpr int Q = 1 << SB;
    constexpr bool TWIST = PL != PL_REAL;
    constexpr int NST = (S + RB - 1) / RB;
    constexpr int HB = S % RB;
    constexpr int FS0 = HB ? 1 : 0;
    __shared__ cplx<T> lds[TILE];
    __shared__ cplx<T> tws[TWIST ? TwistSplit<S>::NHI + TwistSplit<S>::NLO : 1];
    const int tid = threadIdx.x;
    int q;
    long long t;
    one_block<Q>(q, t);
    if (t >= o.batch) return;

    ProdArgs<T> a{};
    a.n = o.n;
    a.S = S;
    a.nega = o.nega;
    a.tab = o.tab;
    a.twist = o.twist;
    PassArgs<T> pf{};
    pf.n = o.n;
    pf.K = S;
    pf.nega = o.nega;
    pf.tab = o.tab;
    TileGeo g;
    g.hi_pos = q << S;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = make_rsrc(o.tab);
    if constexpr (TWIST) fill_twist_split<T, S>(tws, o.twist, tid, NT);  // read after a barrier
    cplx<T> C[Q];
    top_row_coeffs<T, SB, TWIST>(a, q, C);
    const rsrc_t xa = make_rsrc(reinterpret_cast<const T*>(o.src) + t * o.src_stride);
    __syncthreads();
    load_sub<T, S, SB, E, PL, HB>(lds, a, xa, tws, 0, 0, 1, 0, 0, false, C, q, tid);
    run_steps_ct<T, MODE_STD_FWD, E, NT, S, RB, FS0, NST, true, true>(lds, pf, g, tid, o.tab);
    cplx<T>* dst = reinterpret_cast<cplx<T>*>(o.dst) + t * o.dst_stride + ((long long)q << S);
#pragma unroll 4
    for (int i = 0; i < EPT; i++) {
        const int c = tid + i * NT;
        stc_nt(dst + c, lds[swz(c)]);
    }
}

// Inverse, first launch: the Gentleman-Sande stages on bits 0 .. S-1 of
// sub-problem q (the contiguous block [q 2^S, (q+1) 2^S) of a bit-reversed
// spectrum, so its stages need nothing outside it), in place in the source
// buffer (the reference's inverse also works in place in temp,
// fft.cu:1456-1510); with LD_MUL the block is the product of src and src2
// and goes to src.  The short top step (HB bits) runs in registers before
// the stores.
template <typename T, int S, int E, int RB, int LK>
__global__ __launch_bounds__((1 << S) >> E) void inv_local_kernel(OneArgs<T> o)
{
    constexpr int TILE = 1 << S;
    constexpr int EPT = 1 << E;
    constexpr int NT = TILE / EPT;
    constexpr int NST = (S + RB - 1) / RB;
    constexpr int HB = S % RB;
    constexpr int IEND = HB ? NST - 1 : NST;
    __shared__ cplx<T> lds[TILE];
    const int tid = threadIdx.x;
    const int Q = 1 << (o.n - S);
    const int q = blockIdx.x & (Q - 1);
    const long long t = blockIdx.x >> (o.n - S);
    if (t >= o.batch) return;
    cplx<T>* x = reinterpret_cast<cplx<T>*>(const_cast<void*>(o.src)) + t * o.src_stride +
                 ((long long)q << S);
    const cplx<T>* y = LK == LD_MUL ? reinterpret_cast<const cplx<T>*>(o.src2) + t * o.src_stride +
                                          ((long long)q << S)
                                    : nullptr;
#pragma unroll 4
    for (int i = 0; i < EPT; i++) {
        const int c = tid + i * NT;
        cplx<T> v = ldc(x + c);
        if constexpr (LK == LD_MUL) v = v * ldc(y + c);
        lds[swz(c)] = v;
    }
    PassArgs<T> pi{};
    pi.n = o.n;
    pi.K = S;
    pi.nega = o.nega;
    pi.tab = o.tab;
    TileGeo g;
    g.hi_pos = q << S;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = make_rsrc(o.tab);
    run_steps_ct<T, MODE_STD_INV, E, NT, S, RB, 0, IEND, true, true>(lds, pi, g, tid, o.tab);
    ProdArgs<T> a{};
    a.n = o.n;
    a.S = S;
    a.nega = o.nega;
    a.itab = o.tab;
    constexpr int NP = 1 << HB;
    constexpr int NI = EPT / NP;
#pragma unroll 2
    for (int i = 0; i < NI; i++) {
        const int c = tid + i * NT;
        cplx<T> z[NP];
#pragma unroll
        for (int m = 0; m < NP; m++) z[m] = lds[swz(c + (m << (S - HB)))];
        if constexpr (HB > 0) top_gs<T, S, HB>(z, a, q);
#pragma unroll
        for (int m = 0; m < NP; m++) stc(x + c + (m << (S - HB)), z[m]);
    }
}

// Inverse, second launch (SB = 2..4): the top SB Gentleman-Sande stages (bits S .. n-1)
// from the 2^SB rows j 2^S + c of the first launch's result, twiddle
// tab[base_b + (j0 >> (i+1))] (uniform: the column bits are below the stage
// bit), then the scale and the output: ST_REAL Re(z mod_inverse) (reference
// InverseCore R_R store, fft.cu:499-514), ST_UNFOLD z mod_inverse untwist[p]
// unfolded to Re at p and Im at p + L (InverseFFNTCoreUnique,
// fft.cu:1366-1383), with untwist[p] = untwist[p & ~127] untwist[p & 127]
// from LDS.  A workgroup takes 512 consecutive columns of one transform
// (S >= 9), each thread the columns c and c + 256, so that the lanes of
// every load and every store touch consecutive elements.
template <typename T, int SB, int SK>
__global__ __launch_bounds__(256) void inv_top_kernel(OneArgs<T> o, int S)
{
    constexpr int Q = 1 << SB;
    constexpr int UTN = SK == ST_UNFOLD ? (1 << (17 - 7)) + 128 : 1;  // n <= 17
    __shared__ cplx<T> ut[UTN];
    const int tid = threadIdx.x;
    const int n = o.n;
    if constexpr (SK == ST_UNFOLD) {
        const int nhi = 1 << (n - 7);
        for (int i = tid; i < nhi + 128; i += 256)
            ut[i] = i < nhi ? o.untwist[i << 7] : o.untwist[i - nhi];
        __syncthreads();
    }
    const long long g0 = (long long)blockIdx.x * 512;
    const long long t = g0 >> S;
    const int c = (int)(g0 & ((1LL << S) - 1)) + tid;
    if (t >= o.batch) return;
    const long long L = 1LL << n;
    const cplx<T>* src = reinterpret_cast<const cplx<T>*>(o.src) + t * o.src_stride;
    cplx<T> w[SB][Q / 2];
#pragma unroll
    for (int i = 0; i < SB; i++) {
        const int base = o.nega ? (1 << (n - 1 - (S + i))) : 0;
#pragma unroll
        for (int m = 0; m < Q / 2; m++)
            if ((m << (i + 1)) < Q) w[i][m] = o.tab[base + m];
    }
    cplx<T> y[2][Q];
#pragma unroll
    for (int j = 0; j < Q; j++) {
        y[0][j] = ldc_nt(src + ((long long)j << S) + c);
        y[1][j] = ldc_nt(src + ((long long)j << S) + c + 256);
    }
#pragma unroll
    for (int v = 0; v < 2; v++) {
#pragma unroll
        for (int i = 0; i < SB; i++) {
#pragma unroll
            for (int j0 = 0; j0 < Q; j0++) {
                if (j0 & (1 << i)) continue;
                const int j1 = j0 | (1 << i);
                const cplx<T> U = y[v][j0], V = y[v][j1];
                y[v][j0] = U + V;
                y[v][j1] = (U - V) * w[i][j0 >> (i + 1)];
            }
        }
    }
    T* out = reinterpret_cast<T*>(o.dst) + t * o.dst_stride;
#pragma unroll
    for (int j = 0; j < Q; j++) {
#pragma unroll
        for (int v = 0; v < 2; v++) {
            const int P = (j << S) | (c + 256 * v);
            cplx<T> z = y[v][j] * o.scale;
            if constexpr (SK == ST_UNFOLD) z = z * (ut[P >> 7] * ut[(1 << (n - 7)) + (P & 127)]);
            __builtin_nontemporal_store(z.x, out + P);
            if constexpr (SK == ST_UNFOLD) __builtin_nontemporal_store(z.y, out + P + L);
        }
    }
}

}  // namespace gfa
