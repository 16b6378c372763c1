// One-launch transforms for the out-of-place drop-in calls (R_R and FFNT,
// both directions) of length 2^n with n = S + SB, S the LDS tile (2^13
// double / 2^14 float points) and SB = 1..3.
//
// A transform is cut into Q = 2^SB sub-problems of 2^S points, one workgroup
// each; a workgroup reads the whole transform (the Q workgroups of a
// transform get block ids congruent mod 8, so they run on one XCD, close in
// time, and the repeated reads are served by its L2) and evaluates the Q-point
// network that couples the sub-problems in registers, for its own row only:
//   forward (Cooley-Tukey, natural in, bit-reversed out): the couplings are
//     the first SB stages (bits n-1 .. S), between positions j 2^S + c, and
//     sub-problem q owns the output block [q 2^S, (q+1) 2^S); the row
//     coefficients are uniform over the workgroup;
//   inverse (Gentleman-Sande, bit-reversed in, natural out): the couplings
//     are the first SB stages (bits 0 .. SB-1), between the adjacent
//     positions (u << SB) | v, v < Q, and sub-problem r owns the outputs
//     j = (u << SB) | r, u < 2^S (every Q-th output); the stage twiddles
//     depend on u (one table load per distinct twiddle).
// The remaining S stages of a sub-problem are the LDS radix steps of the pass
// engine.  HBM sees every input once and every output once, where the pass
// chain of the same transform reads and writes the whole transform twice.
// Stage order and twiddle indices are those of the reference kernels
// (ForwardCore / InverseCore, src/lib/fft.cu:11-260, and the FFNT cores,
// fft.cu:1013-1384), so results match the pass chain up to rounding.
// In-place calls (C_C, special) keep the pass chain: a workgroup here reads
// positions that its siblings write.
#pragma once

#include "engine.hpp"

namespace gfa {

// Row q of the top SB forward stages as Q coefficients: y_q[c] = sum_j C[j] x_j[c],
// x_j[c] the input at position j 2^S + c.  Stage i (bit S+i, top first)
// keeps z[j] + z[j + 2^i] * wq[i], wq[i] = +-tab[base + (q >> (i+1))] (the
// sign is bit i of q).
template <typename T, int SB>
__device__ __forceinline__ void oneshot_row_coeffs(const cplx<T>* tab, int n, int S, int nega,
                                                   int q, cplx<T>* C)
{
    constexpr int Q = 1 << SB;
    cplx<T> wq[SB];
#pragma unroll
    for (int i = 0; i < SB; i++) {
        const int b = S + i;
        const int base = nega ? (1 << (n - 1 - b)) : 0;
        cplx<T> w = tab[base + (q >> (i + 1))];
        if ((q >> i) & 1) w = mk<T>(-w.x, -w.y);
        wq[i] = w;
    }
#pragma unroll
    for (int j = 0; j < Q; j++) {
        cplx<T> z[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) z[k] = mk<T>(T(j == k), T(0));
#pragma unroll
        for (int i = SB - 1; i >= 0; i--) {
#pragma unroll
            for (int k = 0; k < (1 << i); k++) z[k] = z[k] + z[k + (1 << i)] * wq[i];
        }
        C[j] = z[0];
    }
}

// io: the call's input/output description with io.n = n (fold/unfold
// offsets, output twiddle positions); stages use the same table.
template <typename T, int E, int SB, int LK, int SK>
__global__ __launch_bounds__(512) void oneshot_forward_kernel(PassArgs<T> io)
{
    constexpr int NT = 512;
    constexpr int EPT = 1 << E;
    constexpr int TILE = NT * EPT;
    constexpr int S = E + 9;
    constexpr int Q = 1 << SB;
    __shared__ cplx<T> lds[TILE];
    const int tid = threadIdx.x;
    const int x = blockIdx.x;
    const int q = (x >> 3) % Q;
    const long long t = (long long)((x >> 3) / Q) * 8 + (x & 7);
    if (t >= io.batch) return;

    cplx<T> C[Q];
    oneshot_row_coeffs<T, SB>(io.tab, io.n, S, io.nega, q, C);
#pragma unroll 2
    for (int i = 0; i < EPT; i++) {
        const int c = tid + i * NT;
        cplx<T> xv[Q];
#pragma unroll
        for (int j = 0; j < Q; j++) xv[j] = load_elem<T, LK>(io, t, (j << S) | c);
        cplx<T> y = C[0] * xv[0];
#pragma unroll
        for (int j = 1; j < Q; j++) y = y + C[j] * xv[j];
        lds[swz(c)] = y;
    }
    PassArgs<T> st = io;
    st.blo = 0;
    st.K = S;
    st.logC = 0;
    TileGeo g;
    g.hi_pos = q << S;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = make_rsrc(io.tab);
    run_steps<T, MODE_STD_FWD, E, NT>(lds, st, g, tid, 0, num_steps(S, E));
#pragma unroll 4
    for (int i = 0; i < EPT; i++) {
        const int c = tid + i * NT;
        store_elem<T, SK>(io, t, (q << S) | c, lds[swz(c)]);
    }
}

template <typename T, int E, int SB, int LK, int SK>
__global__ __launch_bounds__(512) void oneshot_inverse_kernel(PassArgs<T> io)
{
    constexpr int NT = 512;
    constexpr int EPT = 1 << E;
    constexpr int TILE = NT * EPT;
    constexpr int S = E + 9;
    constexpr int Q = 1 << SB;
    __shared__ cplx<T> lds[TILE];
    const int tid = threadIdx.x;
    const int x = blockIdx.x;
    const int r = (x >> 3) % Q;
    const long long t = (long long)((x >> 3) / Q) * 8 + (x & 7);
    if (t >= io.batch) return;
    const rsrc_t tab = make_rsrc(io.tab);
    const int n = io.n;

#pragma unroll 2
    for (int i = 0; i < EPT; i++) {
        const int u = tid + i * NT;
        cplx<T> z[Q];
#pragma unroll
        for (int v = 0; v < Q; v++) z[v] = load_elem<T, LK>(io, t, (u << SB) | v);
        // stages on bits 0 .. SB-1 of p = (u << SB) | v, twiddle
        // tab[base_b + (p >> (b+1))] (reference InverseCore, fft.cu:154-260)
#pragma unroll
        for (int b = 0; b < SB; b++) {
            const int base = io.nega ? (1 << (n - 1 - b)) : 0;
#pragma unroll
            for (int v0 = 0; v0 < Q; v0++) {
                if (v0 & (1 << b)) continue;
                const int v1 = v0 | (1 << b);
                const cplx<T> w = bldc<T>(tab, base + (((u << SB) | v0) >> (b + 1)));
                const cplx<T> U = z[v0], V = z[v1];
                z[v0] = U + V;
                z[v1] = (U - V) * w;
            }
        }
        cplx<T> y = z[0];
#pragma unroll
        for (int v = 1; v < Q; v++)
            if (r == v) y = z[v];
        lds[swz(u)] = y;
    }
    // the remaining stages (bits SB .. n-1 of p = bits 0 .. S-1 of u) are
    // those of a transform of 2^S points: p >> (b+1) = u >> (b-SB+1) and
    // base 2^(n-1-b) = 2^(S-1-(b-SB))
    PassArgs<T> st = io;
    st.n = S;
    st.blo = 0;
    st.K = S;
    st.logC = 0;
    TileGeo g;
    g.hi_pos = 0;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = tab;
    run_steps<T, MODE_STD_INV, E, NT>(lds, st, g, tid, 0, num_steps(S, E));
#pragma unroll 4
    for (int i = 0; i < EPT; i++) {
        const int u = tid + i * NT;
        store_elem<T, SK>(io, t, (u << SB) | r, lds[swz(u)]);
    }
}

}  // namespace gfa
