// Fused polynomial-product kernels for gfx950.
//
// A product of two length-2^n transforms (FFNT folding: n = log2(N) - 1) is
// computed by 2^SB workgroups per product.  Workgroup q owns the sub-problem
// of positions [q 2^S, (q+1) 2^S), S = n - SB: after the top SB Cooley-Tukey
// stages (bits n-1 .. S) the remaining S stages, the pointwise product and
// the first S Gentleman-Sande inverse stages of those positions depend on
// nothing else.  So the workgroup
//   1. gathers, for each of its 2^S positions c, the 2^SB inputs
//      j 2^S + c (fold + twist fused) and evaluates row q of the top
//      SB-stage network in registers (the other workgroups of the product
//      read the same inputs, which the XCD-aware block order keeps in L2),
//   2. runs the S forward stages of operand b in LDS and keeps the result in
//      registers, then does the same for operand a (or each limb of a),
//   3. multiplies in registers, inverse-transforms in LDS and writes the
//      2^S partial results to a scratch buffer.
// A second, streaming pass applies the top SB inverse stages, the scale, the
// untwist and the unfold (FP64: f64_combine_kernel; FP32 limbs:
// limb_combine_kernel, which also rounds each limb product and recombines
// them in int64).  With SB = 0 an FP64 product is one workgroup and writes the
// output directly.  The scratch is written and read with streaming hints: it
// is used once, and must not displace the operand lines that the sibling
// workgroups of a product re-read from L2.
//
// The stage/twiddle semantics are those of the reference GPU_FFNT
// (src/lib/fft.cu:1013-1384) and GPU_FFT (fft.cu:11-260): bit-reversed
// spectra, table index (p >> (b+1)) (+2^(n-1-b) for X^N+1).
#pragma once

#include "engine.hpp"

namespace gfa {


enum ProdLoad {
    PL_FOLD = 0,       // T inputs [batch][2^(n+1)]: fold + twist
    PL_FOLD_LIMB = 1,  // int32 inputs: balanced base-2^w digit, fold + twist
    PL_REAL = 2,       // T inputs [batch][2^n]: real part, no twist (R_R forward)
};

template <typename T> struct ProdArgs {
    int n;      // log2 of the complex transform length L
    int S;      // log2 of the sub-problem length
    int batch;  // products in this launch
    int nega;
    const cplx<T>* tab;    // forward root table (reference layout)
    const cplx<T>* itab;   // inverse root table
    const cplx<T>* twist;  // FFNT twist table (L entries)
    const cplx<T>* untwist;
    const cplx<T>* wc;     // stage root constants wc[2^s + t] (32 entries)
    const cplx<T>* wci;    // their conjugates, for the inverse table
    const void* a;
    const void* b;
    long long in_stride;   // input elements per product
    cplx<T>* scratch;      // [batch][limb products][L] partial results (SB > 0)
    T* out;                // direct output (SB == 0, PL_FOLD)
    long long out_stride;
    cplx<T> scale;
    int wa, ka, wb, kb;    // limb split of a and b (PL_FOLD_LIMB)
    long long oa, ob;      // digit offsets: sum_{k < K-1} 2^(w-1) 2^(w k)
    int ca;                // PL_FOLD_LIMB: a is split as a - 2^ca (0: as is); the
                           // product then gains the limb 2^ca J b (J = all ones),
                           // run after the ka digit limbs
    const cplx<T>* jspec;  // forward FFNT spectrum of J (L entries, bit-reversed order)
    int* arrive;           // per-product arrival counters (in-kernel combine)
};

// Balanced base-2^w digit i of x = v - center, v an int32 (digits d_0 ..
// d_{k-1}, the top one keeps the rest): with o = sum_{j < k-1} 2^(w-1) 2^(w j),
// the unsigned base-2^w digits u_j of x + o give d_j = u_j - 2^(w-1) for
// j < k-1, and the top digit is floor((x + o) / 2^s), s = w (k-1).  In 32-bit
// arithmetic, exact for every int32 v when s <= 30 and center <= 2^30 (the
// plans check both): the digits below the top are bits of v - center + o
// mod 2^32, and the top digit is (v >> s) + (((v mod 2^s) - center + o) >> s),
// whose second sum lies in (-2^31, 2^31).
__device__ __forceinline__ float limb_digit(int v, int w, int i, int k, int o, int center)
{
    if (k == 1) return (float)((long long)v - center);
    const int s = w * (k - 1);
    if (i == k - 1) return (float)((v >> s) + (((v & ((1 << s) - 1)) - center + o) >> s));
    const unsigned u = (unsigned)v - (unsigned)center + (unsigned)o;
    return (float)((int)((u >> (w * i)) & ((1u << w) - 1)) - (1 << (w - 1)));
}

// Positions per load of an operand: 16-byte loads of 2 doubles (FP64 form);
// 16-byte loads of 4 int32 (limb form) when the thread's 2^(E-2) groups of
// 4 still hold the 2^HB partners of the top step, else 8-byte loads of 2.
template <typename T, int PL, int E, int HB> constexpr int in_vec()
{
    return (sizeof(T) == 8 && PL != PL_FOLD_LIMB) ? 2 : ((E - 2) >= HB ? 4 : 2);
}

// Folded inputs of V consecutive positions P .. P+V-1 of this product's
// operand, as T: re[v] = x_{P+v}, im[v] = x_{P+v+L}; two vector buffer loads
// (P and L are multiples of V).  Wide lanes cut the address and tag work per
// byte, which bounds the gather of the top network (each input is read by
// every workgroup of its product, each limb's digits by every limb pass).
template <typename T, int PL, int V>
__device__ __forceinline__ void prod_input_vec(rsrc_t x, int L, int w, int limb, int k, long long o,
                                               int center, int P, T* re, T* im)
{
    if constexpr (PL == PL_FOLD || PL == PL_REAL) {
        // V consecutive T of 16 bytes (2 doubles, 4 floats) or 8 bytes (2 floats)
        static_assert(V * sizeof(T) == 16 || V * sizeof(T) == 8, "16- or 8-byte loads");
        if constexpr (V * sizeof(T) == 16) {
            const u32x4 r = __builtin_amdgcn_raw_buffer_load_b128(x, (unsigned)P * (unsigned)sizeof(T), 0, 0);
            __builtin_memcpy(re, &r, 16);
            if constexpr (PL == PL_FOLD) {
                const u32x4 i = __builtin_amdgcn_raw_buffer_load_b128(
                    x, (unsigned)(P + L) * (unsigned)sizeof(T), 0, 0);
                __builtin_memcpy(im, &i, 16);
            }
        } else {
            const u32x2 r = __builtin_amdgcn_raw_buffer_load_b64(x, (unsigned)P * (unsigned)sizeof(T), 0, 0);
            __builtin_memcpy(re, &r, 8);
            if constexpr (PL == PL_FOLD) {
                const u32x2 i = __builtin_amdgcn_raw_buffer_load_b64(
                    x, (unsigned)(P + L) * (unsigned)sizeof(T), 0, 0);
                __builtin_memcpy(im, &i, 8);
            }
        }
        if constexpr (PL == PL_REAL) {
#pragma unroll
            for (int v = 0; v < V; v++) im[v] = T(0);
        }
    } else if constexpr (V == 4) {
        const u32x4 r = __builtin_amdgcn_raw_buffer_load_b128(x, (unsigned)P * 4u, 0, 0);
        const u32x4 i = __builtin_amdgcn_raw_buffer_load_b128(x, (unsigned)(P + L) * 4u, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            re[v] = (T)limb_digit((int)r[v], w, limb, k, (int)o, center);
            im[v] = (T)limb_digit((int)i[v], w, limb, k, (int)o, center);
        }
    } else {
        const u32x2 r = __builtin_amdgcn_raw_buffer_load_b64(x, (unsigned)P * 4u, 0, 0);
        const u32x2 i = __builtin_amdgcn_raw_buffer_load_b64(x, (unsigned)(P + L) * 4u, 0, 0);
#pragma unroll
        for (int v = 0; v < 2; v++) {
            re[v] = (T)limb_digit((int)r[v], w, limb, k, (int)o, center);
            im[v] = (T)limb_digit((int)i[v], w, limb, k, (int)o, center);
        }
    }
}

// Row q of the top SB forward stages (bits n-1 .. S) of zz[j] = input at
// position j 2^S + c.  Only the entries whose processed bits equal q's
// survive each stage, and their twiddle, tab[base + (q >> (i+1))], is
// uniform, so stage i (bit S+i, top first) is z[j] += z[j + 2^i] * wq[i]
// with wq[i] = +-tab[...] (sign: bit i of q), j < 2^i.
template <typename T, int SB>
__device__ __forceinline__ void top_row_twiddles(const ProdArgs<T>& a, int q, cplx<T>* wq)
{
#pragma unroll
    for (int i = 0; i < SB; i++) {
        const int b = a.S + i;
        const int base = a.nega ? (1 << (a.n - 1 - b)) : 0;
        cplx<T> w = a.tab[base + (q >> (i + 1))];
        if ((q >> i) & 1) w = mk<T>(-w.x, -w.y);
        wq[i] = w;
    }
}

template <typename T, int SB>
__device__ __forceinline__ cplx<T> top_forward_row(cplx<T>* zz, const cplx<T>* wq)
{
#pragma unroll
    for (int i = SB - 1; i >= 0; i--) {
#pragma unroll
        for (int j = 0; j < (1 << i); j++) zz[j] = zz[j] + zz[j + (1 << i)] * wq[i];
    }
    return zz[0];
}

// Row q of the top network is linear in the inputs and the twist factors
// as twist[j 2^S + c] = twist[c] * twist[j 2^S] (up to rounding), so
//     y_q[c] = twist[c] * sum_j C[j] z_j,  C[j] = row_q(e_j) * twist[j 2^S],
// with uniform coefficients C computed once per workgroup.
template <typename T, int SB, bool TWIST = true>
__device__ __forceinline__ void top_row_coeffs(const ProdArgs<T>& a, int q, cplx<T>* C)
{
    constexpr int Q = 1 << SB;
    cplx<T> wq[SB > 0 ? SB : 1];
    top_row_twiddles<T, SB>(a, q, wq);
#pragma unroll
    for (int j = 0; j < Q; j++) {
        cplx<T> e[Q];
#pragma unroll
        for (int k = 0; k < Q; k++) e[k] = mk<T>(T(0), T(0));
        e[j] = TWIST ? a.twist[j << a.S] : mk<T>(T(1), T(0));
        C[j] = top_forward_row<T, SB>(e, wq);
    }
}

// Coefficient C[j] of top_row_coeffs alone (one thread per j).
template <typename T, int SB>
__device__ __forceinline__ cplx<T> top_row_coeff(const ProdArgs<T>& a, int q, int j)
{
    constexpr int Q = 1 << SB;
    cplx<T> wq[SB > 0 ? SB : 1];
    top_row_twiddles<T, SB>(a, q, wq);
    const cplx<T> tj = a.twist[j << a.S];
    cplx<T> e[Q];
#pragma unroll
    for (int k = 0; k < Q; k++) e[k] = k == j ? tj : mk<T>(T(0), T(0));
    return top_forward_row<T, SB>(e, wq);
}

// Twist factors twist[c], c < 2^S, as twist[c & ~(2^TWB-1)] * twist[c & (2^TWB-1)]
// (exp(2 pi i c / 2N) is multiplicative in c; both factors are entries of the
// FFNT twist table), from two small LDS tables: the gather then moves only
// operand bytes through the vector memory path.
template <int S> struct TwistSplit {
    static constexpr int TWB = S > 7 ? 7 : S;
    static constexpr int NLO = 1 << TWB;
    static constexpr int NHI = 1 << (S - TWB);
};

template <typename T, int S>
__device__ __forceinline__ void fill_twist_split(cplx<T>* tws, const cplx<T>* twist, int tid, int nthreads)
{
    using TS = TwistSplit<S>;
    for (int i = tid; i < TS::NHI + TS::NLO; i += nthreads)
        tws[i] = i < TS::NHI ? twist[i << TS::TWB] : twist[i - TS::NHI];
}

template <typename T, int S>
__device__ __forceinline__ cplx<T> twist_of(const cplx<T>* tws, int c)
{
    using TS = TwistSplit<S>;
    return tws[c >> TS::TWB] * tws[TS::NHI + (c & (TS::NLO - 1))];
}

// The short top step of a sub-problem (its HB = S mod RB highest bits, a
// step that would otherwise cost a full LDS round trip) runs in registers:
// forward in the gather, inverse in the final store.  A thread holds the
// 2^HB positions c + m 2^(S-HB), m < 2^HB, bit k of m being position bit
// S-HB+k, and the stage twiddles tab[base_b + ((q 2^S + c) >> (b+1))] =
// tab[base_b + (q << (HB-1-k)) + (m >> (k+1))] (b = S-HB+k) are uniform over
// the workgroup (scalar loads).
template <typename T, int S, int HB>
__device__ __forceinline__ cplx<T> top_twiddle(const ProdArgs<T>& a, const cplx<T>* tab, int q,
                                               int k, int t)
{
    const int b = S - HB + k;
    const int base = a.nega ? (1 << (a.n - 1 - b)) : 0;
    return tab[base + (q << (HB - 1 - k)) + t];
}

template <typename T, int S, int HB>
__device__ __forceinline__ void top_ct(cplx<T>* y, const ProdArgs<T>& a, int q)
{
#pragma unroll
    for (int k = HB - 1; k >= 0; k--) {
#pragma unroll
        for (int m = 0; m < (1 << HB); m++) {
            if (m & (1 << k)) continue;
            const cplx<T> w = top_twiddle<T, S, HB>(a, a.tab, q, k, m >> (k + 1));
            const cplx<T> U = y[m], Vw = y[m | (1 << k)] * w;
            y[m] = U + Vw;
            y[m | (1 << k)] = U - Vw;
        }
    }
}

template <typename T, int S, int HB>
__device__ __forceinline__ void top_gs(cplx<T>* y, const ProdArgs<T>& a, int q)
{
#pragma unroll
    for (int k = 0; k < HB; k++) {
#pragma unroll
        for (int m = 0; m < (1 << HB); m++) {
            if (m & (1 << k)) continue;
            const cplx<T> w = top_twiddle<T, S, HB>(a, a.itab, q, k, m >> (k + 1));
            const cplx<T> U = y[m], V = y[m | (1 << k)];
            y[m] = U + V;
            y[m | (1 << k)] = (U - V) * w;
        }
    }
}

// Gathers the sub-problem's input into LDS: fold + twist + row q of the top
// network for every position c < 2^S, then the HB-bit top step (top_ct).
// ones: the operand is the all-ones polynomial J (no loads); center: the
// limb form splits x - center.
template <typename T, int S, int SB, int E, int PL, int HB>
__device__ __forceinline__ void load_sub(cplx<T>* lds, const ProdArgs<T>& a, rsrc_t x,
                                         const cplx<T>* tws, int w, int limb, int k, long long o,
                                         int center, bool ones, const cplx<T>* C, int q, int tid)
{
    constexpr int NT = (1 << S) >> E;
    constexpr int Q = 1 << SB;
    constexpr int V = in_vec<T, PL, E, HB>();
    constexpr int NG = (1 << E) / V;  // groups of V consecutive positions per thread
    constexpr int NP = 1 << HB;       // groups joined by the top step
    constexpr int NGH = NG / NP;      // partner m of group g: g + m NGH (c + m 2^(S-HB))
    static_assert(NGH >= 1, "the top step must fit the thread's groups");
    const int L = 1 << a.n;
#pragma unroll 1
    for (int g0 = 0; g0 < NGH; g0++) {
        cplx<T> y[V][NP];
#pragma unroll
        for (int m = 0; m < NP; m++) {
            const int c0 = (tid + (g0 + m * NGH) * NT) * V;
            T re[Q][V], im[Q][V];
            if (ones) {
#pragma unroll
                for (int j = 0; j < Q; j++)
#pragma unroll
                    for (int v = 0; v < V; v++) re[j][v] = im[j][v] = T(1);
            } else {
#pragma unroll
                for (int j = 0; j < Q; j++)
                    prod_input_vec<T, PL, V>(x, L, w, limb, k, o, center, (j << S) | c0, re[j],
                                             im[j]);
            }
#pragma unroll
            for (int v = 0; v < V; v++) {
                T ax = T(0), ay = T(0);
#pragma unroll
                for (int j = 0; j < Q; j++) {
                    if constexpr (PL == PL_REAL) {  // imaginary parts are 0, no twist
                        ax = ax + C[j].x * re[j][v];
                        ay = ay + C[j].y * re[j][v];
                    } else {
                        ax = ax + C[j].x * re[j][v] - C[j].y * im[j][v];
                        ay = ay + C[j].x * im[j][v] + C[j].y * re[j][v];
                    }
                }
                if constexpr (PL == PL_REAL)
                    y[v][m] = mk<T>(ax, ay);
                else
                    y[v][m] = twist_of<T, S>(tws, c0 + v) * mk<T>(ax, ay);
            }
        }
#pragma unroll
        for (int v = 0; v < V; v++) {
            if constexpr (HB > 0) top_ct<T, S, HB>(y[v], a, q);
#pragma unroll
            for (int m = 0; m < NP; m++)
                lds[swz((tid + (g0 + m * NGH) * NT) * V + v)] = y[v][m];
        }
    }
}

// Last forward step (bits [0, RB)) of the thread's 2^(E-RB) groups, into
// v[k 2^RB + j] = position ((tid + k NT) << RB) | j.
template <typename T, int E, int RB, int NT, bool TW = false>
__device__ __forceinline__ void last_forward_step(const cplx<T>* lds, const PassArgs<T>& pf,
                                                  const TileGeo& g, int tid, cplx<T>* v,
                                                  const cplx<T>* wc)
{
#pragma unroll
    for (int k = 0; k < (1 << (E - RB)); k++) {
        const int rbase = (tid + k * NT) << RB;
        cplx<T> w[1 << RB];
#pragma unroll
        for (int j = 0; j < (1 << RB); j++) w[j] = lds[swz(rbase | j)];
        group_stages_derived<T, MODE_STD_FWD, RB, TW>(w, pf, g, rbase, 0, 0, wc);
#pragma unroll
        for (int j = 0; j < (1 << RB); j++) v[(k << RB) + j] = w[j];
    }
}

// Top SB Gentleman-Sande stages (bits S .. n-1) of one whole product from
// its 2^SB sub-problem results in scratch, then scale, untwist and unfold:
// out[P] = Re, out[P + L] = Im (reference InverseFFNTCoreUnique,
// src/lib/fft.cu:1298-1384).  Run by the last workgroup of the product.
template <typename T, int S, int SB, int NT>
__device__ __forceinline__ void combine_product(const ProdArgs<T>& a, long long p, int tid)
{
    constexpr int Q = 1 << SB;
    const int L = 1 << a.n;
    const rsrc_t src = make_rsrc(a.scratch + p * (long long)L);
    const rsrc_t ut = make_rsrc(a.untwist);
    T* o = a.out + p * a.out_stride;
    // stage i, pair (j0, j0 + 2^i): twiddle itab[(j0 2^S + c) >> (S + i + 1)] = itab[j0 >> (i+1)]
    cplx<T> w[Q / 2];
#pragma unroll
    for (int m = 0; m < Q / 2; m++) w[m] = a.itab[m];
#pragma unroll 2
    for (int c = tid; c < (1 << S); c += NT) {
        cplx<T> y[Q];
#pragma unroll
        for (int j = 0; j < Q; j++) y[j] = bldc<T>(src, (j << S) | c);
#pragma unroll
        for (int i = 0; i < SB; i++) {
#pragma unroll
            for (int j0 = 0; j0 < Q; j0++) {
                if (j0 & (1 << i)) continue;
                const int j1 = j0 | (1 << i);
                const cplx<T> U = y[j0], V = y[j1];
                y[j0] = U + V;
                y[j1] = (U - V) * w[j0 >> (i + 1)];
            }
        }
#pragma unroll
        for (int j = 0; j < Q; j++) {
            const int P = (j << S) | c;
            const cplx<T> z = (y[j] * a.scale) * bldc<T>(ut, P);
            o[P] = z.x;
            o[P + L] = z.y;
        }
    }
}

// CMB (FP64, SB > 0): the combine runs inside this kernel.  Each workgroup
// publishes its sub-problem result (agent-scope release) and counts itself
// in arrive[p]; the workgroup that arrives last acquires, combines the whole
// product and resets arrive[p] to 0 for the next launch.  No workgroup waits
// for another.
template <typename T, int S, int SB, int E, int PL, int RB, bool CMB = false>
__global__ __launch_bounds__((1 << S) >> E) void product_sub_kernel(ProdArgs<T> a)
{
    constexpr int TILE = 1 << S;
    constexpr int EPT = 1 << E;
    constexpr int NT = TILE / EPT;
    constexpr int Q = 1 << SB;
    __shared__ cplx<T> lds[TILE];
    __shared__ cplx<T> tws[TwistSplit<S>::NHI + TwistSplit<S>::NLO];

    const int tid = threadIdx.x;
    // XCD-aware order: consecutive workgroups go to consecutive XCDs, so the
    // Q workgroups of one product are given block ids congruent mod 8 and
    // dispatched in consecutive rounds, i.e. on one XCD, close in time.
    const int x = blockIdx.x;
    const int xcd = x & 7;
    const int slot = x >> 3;
    const int q = slot % Q;
    const long long p = (long long)(slot / Q) * 8 + xcd;
    if (p >= a.batch) return;

    PassArgs<T> pf{};
    pf.n = a.n;
    pf.K = S;
    pf.nega = a.nega;
    pf.tab = a.tab;
    PassArgs<T> pinv = pf;
    pinv.tab = a.itab;
    TileGeo g;
    g.hi_pos = q << S;
    g.blo = 0;
    g.K = S;
    g.logC = 0;
    g.tab = make_rsrc(a.tab);
    TileGeo gi = g;
    gi.tab = make_rsrc(a.itab);
    constexpr int NST = (S + RB - 1) / RB;
    const int nla = a.ka + (a.ca ? 1 : 0);  // a limbs: ka digits (+ J)
    const int nprod = nla * a.kb;
    const long long L = 1LL << a.n;

    const rsrc_t xa = make_rsrc(reinterpret_cast<const char*>(a.a) +
                                p * a.in_stride * (PL == PL_FOLD ? (long long)sizeof(T) : 4LL));
    const rsrc_t xb = make_rsrc(reinterpret_cast<const char*>(a.b) +
                                p * a.in_stride * (PL == PL_FOLD ? (long long)sizeof(T) : 4LL));
    fill_twist_split<T, S>(tws, a.twist, tid, NT);  // visible after the first barrier below
    // compile-time single limb for the FP64 form
    const int KA = PL == PL_FOLD ? 1 : nla;
    const int KB = PL == PL_FOLD ? 1 : a.kb;

    cplx<T> C[Q];
    top_row_coeffs<T, SB>(a, q, C);
    // the short top step (HB bits) runs in registers: in the gather (forward)
    // and in the final store (inverse)
    constexpr int HB = S % RB;
    constexpr int FS0 = HB ? 1 : 0;           // first forward step through LDS
    constexpr int IEND = HB ? NST - 1 : NST;  // end of the inverse steps through LDS

    cplx<T> breg[EPT];
    for (int jb = 0; jb < KB; jb++) {
        // an opaque copy of tid per iteration keeps the address arithmetic of
        // the loop body inside the loop (hoisted, it would hold registers)
        int t = tid;
        asm volatile("" : "+v"(t));
        __syncthreads();
        load_sub<T, S, SB, E, PL, HB>(lds, a, xb, tws, a.wb, jb, KB, a.ob, 0, false, C, q, t);
        run_steps_ct<T, MODE_STD_FWD, E, NT, S, RB, FS0, NST - 1, true>(lds, pf, g, t, a.wc);
        last_forward_step<T, E, RB, NT>(lds, pf, g, t, breg, a.wc);
        for (int ia = 0; ia < KA; ia++) {
            int u = tid;
            asm volatile("" : "+v"(u));
            __syncthreads();
            // the J limb (ia == ka) needs no transform: its spectrum is a.jspec
            const bool jl = PL != PL_FOLD && ia >= a.ka;
            if (!jl) {
                load_sub<T, S, SB, E, PL, HB>(lds, a, xa, tws, a.wa, ia, a.ka, a.oa,
                                              a.ca ? (1 << a.ca) : 0, false, C, q, u);
                run_steps_ct<T, MODE_STD_FWD, E, NT, S, RB, FS0, NST - 1, true>(lds, pf, g, u,
                                                                               a.wc);
            }
            // last forward step, product, first inverse step: in registers,
            // one radix-2^RB group at a time
#pragma unroll
            for (int k = 0; k < (1 << (E - RB)); k++) {
                const int rbase = (u + k * NT) << RB;
                cplx<T> v[1 << RB];
                if (jl) {
                    const rsrc_t js = make_rsrc(a.jspec);
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) v[j] = bldc<T>(js, g.hi_pos | rbase | j);
                } else {
#pragma unroll
                    for (int j = 0; j < (1 << RB); j++) v[j] = lds[swz(rbase | j)];
                    group_stages_derived<T, MODE_STD_FWD, RB>(v, pf, g, rbase, 0, 0, a.wc);
                }
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) v[j] = v[j] * breg[(k << RB) + j];
                group_stages_derived<T, MODE_STD_INV, RB>(v, pinv, gi, rbase, 0, 0, a.wci);
#pragma unroll
                for (int j = 0; j < (1 << RB); j++) lds[swz(rbase | j)] = v[j];
            }
            run_steps_ct<T, MODE_STD_INV, E, NT, S, RB, 1, IEND, true>(lds, pinv, gi, u, a.wci);
            // element i of this thread: position u + i NT; the top step joins
            // elements i + m EPT/2^HB (positions c + m 2^(S-HB))
            constexpr int NP = 1 << HB;
            constexpr int NI = EPT / NP;
            if constexpr (SB == 0 && PL == PL_FOLD) {
                // whole product in this workgroup: scale, untwist, unfold
                T* o = a.out + p * a.out_stride;
                const rsrc_t utw = make_rsrc(a.untwist);
#pragma unroll 2
                for (int i = 0; i < NI; i++) {
                    const int c = u + i * NT;
                    cplx<T> z[NP];
#pragma unroll
                    for (int m = 0; m < NP; m++) z[m] = lds[swz(c + (m << (S - HB)))];
                    if constexpr (HB > 0) top_gs<T, S, HB>(z, a, q);
#pragma unroll
                    for (int m = 0; m < NP; m++) {
                        const int cm = c + (m << (S - HB));
                        const cplx<T> r = (z[m] * a.scale) * bldc<T>(utw, cm);
                        o[cm] = r.x;
                        o[cm + L] = r.y;
                    }
                }
            } else {
                cplx<T>* dst = a.scratch + (p * nprod + ia * KB + jb) * L + ((long long)q << S);
#pragma unroll 2
                for (int i = 0; i < NI; i++) {
                    const int c = u + i * NT;
                    cplx<T> z[NP];
#pragma unroll
                    for (int m = 0; m < NP; m++) z[m] = lds[swz(c + (m << (S - HB)))];
                    if constexpr (HB > 0) top_gs<T, S, HB>(z, a, q);
#pragma unroll
                    for (int m = 0; m < NP; m++) stc_nt(dst + c + (m << (S - HB)), z[m]);
                }
            }
        }
    }
    if constexpr (CMB && SB > 0 && PL == PL_FOLD) {
        __shared__ int sh_last;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (tid == 0) {
            const int prev = __hip_atomic_fetch_add(a.arrive + p, 1, __ATOMIC_ACQ_REL,
                                                    __HIP_MEMORY_SCOPE_AGENT);
            sh_last = prev == Q - 1;
        }
        __syncthreads();
        if (sh_last) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            combine_product<T, S, SB, NT>(a, p, tid);
            if (tid == 0)
                __hip_atomic_store(a.arrive + p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    (void)Q;
}

// FP32 limb products -> int64: top SB inverse stages, scale, untwist, unfold,
// round every limb product and add it at weight 2^(wa*ia + wb*jb).
// One thread per (product, sub-problem position c).  err (optional) receives
// per product the largest |x - rint(x)| (as double bits) and the sum of
// squares over all limb outputs.
// FP64 products: top SB Gentleman-Sande stages (bits S .. n-1), scale,
// untwist and unfold (reference InverseFFNTCoreUnique, src/lib/fft.cu:
// 1298-1384).  A workgroup takes 512 consecutive columns of one product
// (S >= 9), each thread the columns c and c + 256: the lanes of every load
// read consecutive 16-byte scratch entries and the lanes of every store
// write consecutive doubles.  untwist[P] = untwist[P & ~127] *
// untwist[P & 127] from LDS (the table is exp(-2 pi i P / 2N),
// multiplicative in P).
struct F64CombineArgs {
    int n, S, batch;
    const cplx<double>* scratch;  // [batch][2^n]
    const cplx<double>* itab;
    const cplx<double>* untwist;
    double scale;
    double* out;                  // [batch][out_stride]
    long long out_stride;
};

constexpr int F64C_MAX_N = 16;  // 2^(16-7) entries in the LDS high-part table

template <int SB>
__global__ __launch_bounds__(256) void f64_combine_kernel(F64CombineArgs a)
{
    constexpr int Q = 1 << SB;
    __shared__ cplx<double> ut[128 + (1 << (F64C_MAX_N - 7))];
    const int tid = threadIdx.x;
    const int nhi = 1 << (a.n - 7);
    for (int i = tid; i < 128 + nhi; i += 256)
        ut[i] = i < 128 ? a.untwist[i] : a.untwist[(i - 128) << 7];
    __syncthreads();
    const long long g0 = (long long)blockIdx.x * 512;
    const long long p = g0 >> a.S;
    const int c = (int)(g0 & ((1LL << a.S) - 1)) + tid;
    if (p >= a.batch) return;
    const long long L = 1LL << a.n;
    const cplx<double>* src = a.scratch + p * L;
    cplx<double> w[Q > 1 ? Q / 2 : 1];
#pragma unroll
    for (int m = 0; m < Q / 2; m++) w[m] = a.itab[m];
    cplx<double> y[2][Q];
#pragma unroll
    for (int j = 0; j < Q; j++) {
        y[0][j] = ldc_nt(src + ((long long)j << a.S) + c);
        y[1][j] = ldc_nt(src + ((long long)j << a.S) + c + 256);
    }
    double* o = a.out + p * a.out_stride;
#pragma unroll
    for (int v = 0; v < 2; v++) {
#pragma unroll
        for (int i = 0; i < SB; i++) {
#pragma unroll
            for (int j0 = 0; j0 < Q; j0++) {
                if (j0 & (1 << i)) continue;
                const int j1 = j0 | (1 << i);
                const cplx<double> U = y[v][j0], V = y[v][j1];
                y[v][j0] = U + V;
                y[v][j1] = (U - V) * w[j0 >> (i + 1)];
            }
        }
    }
#pragma unroll
    for (int j = 0; j < Q; j++) {
#pragma unroll
        for (int v = 0; v < 2; v++) {
            const int P = (j << a.S) | (c + 256 * v);
            const cplx<double> z = mk<double>(y[v][j].x * a.scale, y[v][j].y * a.scale) *
                                   (ut[128 + (P >> 7)] * ut[P & 127]);
            __builtin_nontemporal_store(z.x, o + P);
            __builtin_nontemporal_store(z.y, o + P + L);
        }
    }
}

template <int SB> struct CombineArgs {
    int n, S, batch, nprod, ka, kb, wa, wb, ca;
    const cplx<float>* scratch;
    const cplx<float>* itab;
    const cplx<float>* untwist;
    float scale;
    long long* out;
    long long out_stride;
    double* err;
};

template <int SB>
__global__ __launch_bounds__(256) void limb_combine_kernel(CombineArgs<SB> a)
{
    constexpr int Q = 1 << SB;
    const long long gid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long p = gid >> a.S;
    const int c = (int)(gid & ((1LL << a.S) - 1));
    if (p >= a.batch) return;
    const long long L = 1LL << a.n;
    long long acc[Q][2];
#pragma unroll
    for (int j = 0; j < Q; j++) acc[j][0] = acc[j][1] = 0;
    float emax = 0.f;
    float esq = 0.f;
    for (int l = 0; l < a.nprod; l++) {
        const cplx<float>* src = a.scratch + (p * a.nprod + l) * L;
        cplx<float> y[Q];
#pragma unroll
        for (int j = 0; j < Q; j++) y[j] = ldc_nt(src + ((long long)j << a.S) + c);
#pragma unroll
        for (int i = 0; i < SB; i++) {  // Gentleman-Sande, bits S .. n-1
#pragma unroll
            for (int j0 = 0; j0 < Q; j0++) {
                if (j0 & (1 << i)) continue;
                const int j1 = j0 | (1 << i);
                const cplx<float> w = ldc(a.itab + (j0 >> (i + 1)));
                cplx<float> U = y[j0], V = y[j1];
                y[j0] = U + V;
                y[j1] = (U - V) * w;
            }
        }
        const int ia = l / a.kb, jb = l % a.kb;  // ia == ka: the J limb, weight 2^ca
        const int shift = (ia < a.ka ? a.wa * ia : a.ca) + a.wb * jb;
#pragma unroll
        for (int j = 0; j < Q; j++) {
            const int P = (j << a.S) | c;
            cplx<float> z = mk<float>(y[j].x * a.scale, y[j].y * a.scale) * ldc(a.untwist + P);
            const float r0 = rintf(z.x), r1 = rintf(z.y);
            const float e0 = fabsf(z.x - r0), e1 = fabsf(z.y - r1);
            emax = fmaxf(emax, fmaxf(e0, e1));
            esq += e0 * e0 + e1 * e1;
            acc[j][0] += (long long)r0 * (1LL << shift);
            acc[j][1] += (long long)r1 * (1LL << shift);
        }
    }
    long long* o = a.out + p * a.out_stride;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const int P = (j << a.S) | c;
        __builtin_nontemporal_store(acc[j][0], o + P);
        __builtin_nontemporal_store(acc[j][1], o + P + L);
    }
    if (a.err) {
        // one wavefront holds 64 consecutive positions of one product
        double dmax = emax, dsq = esq;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
            dsq += __shfl_xor(dsq, off, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMax(reinterpret_cast<unsigned long long*>(a.err + 2 * p),
                      (unsigned long long)__double_as_longlong(dmax));
            atomicAdd(a.err + 2 * p + 1, dsq);
        }
    }
}

}  // namespace gfa
