// Generic LDS-tiled radix pass for gfx950.
//
// A transform of length L = 2^n is a sequence of radix-2 stages, each one
// pairing positions p and p ^ 2^b for one position bit b:
//   MODE_STD_FWD  Cooley-Tukey, bits n-1 .. 0, twiddle tab[base + (p >> (b+1))]
//                 (ForwardCore, reference src/lib/fft.cu:11-152; base = 2^(n-1-b)
//                 for X^N+1, 0 for X^N-1, fft.cu:67-74)
//   MODE_STD_INV  Gentleman-Sande, bits 0 .. n-1, same twiddle index
//                 (InverseCore, fft.cu:154-260)
//   MODE_SPE_FWD  Cooley-Tukey, bits 0 .. n-1, twiddle tab[2^b + (p & (2^b-1))]
//                 (Special_ForwardCore, fft.cu:856-926)
//   MODE_SPE_INV  Gentleman-Sande, bits n-1 .. 0, same twiddle index
//                 (Special_InverseCore, fft.cu:727-854)
//
// One launch ("pass") applies the stages of K consecutive bits [blo, blo+K).
// A workgroup owns a tile of 2^LOGT elements: 2^K "rows" (the pass bits) by
// 2^logC "columns" (consecutive positions below blo, so that every global
// row segment is a contiguous run), or, when the pass covers the whole
// transform, 2^logBT whole transforms.  The tile is read from HBM once with
// consecutive lanes on consecutive addresses, kept in LDS (XOR-swizzled in
// 16-element blocks against bank conflicts) while each thread performs
// radix-2^E steps in registers, and written back once.  The first pass can
// fuse the operation's input transform (real promotion, pointwise product,
// FFNT fold+twist) and the last pass its output transform (scale, real part,
// FFNT untwist+unfold), so a single-pass transform touches HBM once each way.
#pragma once

#include "cplx.hpp"

namespace gfa {

enum Mode { MODE_STD_FWD = 0, MODE_STD_INV = 1, MODE_SPE_FWD = 2, MODE_SPE_INV = 3 };
enum LoadKind {
    LD_CPLX = 0,
    LD_REAL = 1,
    LD_MUL = 2,
    LD_FOLD = 3,
    LD_SPLIT_HI = 4,  // T = float: double complex source, its float rounding
    LD_SPLIT_LO = 5,  // T = float: double complex source, the float of the rest
    LD_RR_PACK = 6,   // R_R inverse by half-length packing (see load_elem)
};
enum StoreKind {
    ST_CPLX = 0,
    ST_REAL = 1,
    ST_UNFOLD = 2,
    ST_SUM_F64 = 3,  // T = float: double complex dst = (double)aux[p] + (double)z
};

__host__ __device__ constexpr bool mode_is_ct(int m) { return m == MODE_STD_FWD || m == MODE_SPE_FWD; }
__host__ __device__ constexpr bool mode_descending(int m) { return m == MODE_STD_FWD || m == MODE_SPE_INV; }
__host__ __device__ constexpr bool mode_special(int m) { return m == MODE_SPE_FWD || m == MODE_SPE_INV; }

template <typename T> struct PassArgs {
    int n;       // log2 transform length
    int blo;     // lowest bit of this pass
    int K;       // number of bits (stages) of this pass
    int logC;    // log2 columns per tile (0 when the pass reaches bit 0)
    int logBT;   // log2 transforms per tile (only when K == n)
    int batch;   // transforms in the launch
    int nega;    // X^N+1 twiddle base (standard modes only)
    int do_scale;
    int nt_in, nt_out;   // streaming (non-temporal) source loads / destination stores
    cplx<T> scale;
    const cplx<T>* tab;  // root-of-unity table (reference layout)
    // input
    const void* src;
    const void* src2;      // second operand (LD_MUL)
    const cplx<T>* ltab;   // twist table (LD_FOLD)
    long long src_stride;  // elements between transforms
    // output
    void* dst;
    const cplx<T>* stab;  // untwist table (ST_UNFOLD); first summand (ST_SUM_F64)
    long long dst_stride;
};

// Bank-conflict swizzle: permutes each aligned block of 16 elements by the
// XOR of the next two 4-bit fields of the index.
__device__ __forceinline__ int swz(int i) { return i ^ (((i >> 4) ^ (i >> 8)) & 15); }

// Source and destination accesses of a pass: streaming (non-temporal) when
// the buffer is read or written once by this transform (first-pass source,
// last-pass destination), so that it does not displace cached tables and
// intermediate data; cached otherwise.
template <typename U> __device__ __forceinline__ U ld_s(const U* p, bool nt)
{
    return nt ? __builtin_nontemporal_load(p) : *p;
}
template <typename U> __device__ __forceinline__ void st_s(U* p, U v, bool nt)
{
    if (nt)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}
template <typename T> __device__ __forceinline__ cplx<T> ldc_s(const cplx<T>* p, bool nt)
{
    return nt ? ldc_nt(p) : ldc(p);
}
template <typename T> __device__ __forceinline__ void stc_s(cplx<T>* p, cplx<T> v, bool nt)
{
    if (nt)
        stc_nt(p, v);
    else
        stc(p, v);
}

template <typename T, int LK>
__device__ __forceinline__ cplx<T> load_elem(const PassArgs<T>& a, long long b, int p)
{
    const bool nt = a.nt_in;
    if constexpr (LK == LD_CPLX) {
        return ldc_s(reinterpret_cast<const cplx<T>*>(a.src) + b * a.src_stride + p, nt);
    } else if constexpr (LK == LD_REAL) {
        return mk<T>(ld_s(reinterpret_cast<const T*>(a.src) + b * a.src_stride + p, nt), T(0));
    } else if constexpr (LK == LD_MUL) {
        cplx<T> u = ldc_s(reinterpret_cast<const cplx<T>*>(a.src) + b * a.src_stride + p, nt);
        cplx<T> v = ldc_s(reinterpret_cast<const cplx<T>*>(a.src2) + b * a.src_stride + p, nt);
        return u * v;
    } else if constexpr (LK == LD_FOLD) {  // z = (f[p] + i f[p+L]) * twist[p]
        const T* f = reinterpret_cast<const T*>(a.src) + b * a.src_stride;
        int L = 1 << a.n;
        return mk<T>(ld_s(f + p, nt), ld_s(f + p + L, nt)) * ldc(a.ltab + p);
    } else if constexpr (LK == LD_RR_PACK) {
        // Half-length form of the X^N-1 R_R inverse.  The source holds full
        // N-point spectra C in bit-reversed order (C_k at 2 bitrev(k), C_{k+M}
        // next to it, M = N/2 = 2^n); for a spectrum of real data the N-point
        // inverse D_j = sum_k C_k w^-jk has D_2m + i D_2m+1 = (M-point inverse
        // of Z)_m with Z_k = (C_k + C_{k+M}) + i w^-k (C_k - C_{k+M}), and
        // w^-k = itab[p] at bit-reversed position p = bitrev(k).  src2
        // (optional) is the second operand of the product.
        const cplx<T>* X = reinterpret_cast<const cplx<T>*>(a.src) + b * a.src_stride + 2 * p;
        cplx<T> x0 = ldc_s(X, nt), x1 = ldc_s(X + 1, nt);
        if (a.src2) {
            const cplx<T>* Y = reinterpret_cast<const cplx<T>*>(a.src2) + b * a.src_stride + 2 * p;
            x0 = x0 * ldc_s(Y, nt);
            x1 = x1 * ldc_s(Y + 1, nt);
        }
        const cplx<T> s0 = x0 + x1, d = (x0 - x1) * ldc(a.ltab + p);
        return mk<T>(s0.x - d.y, s0.y + d.x);
    } else {  // LD_SPLIT_HI / LD_SPLIT_LO: x = hi + lo, hi = float(x), lo = float(x - hi)
        const cplx<double> x =
            ldc_s(reinterpret_cast<const cplx<double>*>(a.src) + b * a.src_stride + p, nt);
        const T hx = (T)x.x, hy = (T)x.y;
        if constexpr (LK == LD_SPLIT_HI) return mk<T>(hx, hy);
        return mk<T>((T)(x.x - (double)hx), (T)(x.y - (double)hy));
    }
}

template <typename T, int SK>
__device__ __forceinline__ void store_elem(const PassArgs<T>& a, long long b, int p, cplx<T> z)
{
    const bool nt = a.nt_out;
    if (a.do_scale) z = z * a.scale;
    if constexpr (SK == ST_CPLX) {
        stc_s(reinterpret_cast<cplx<T>*>(a.dst) + b * a.dst_stride + p, z, nt);
    } else if constexpr (SK == ST_REAL) {
        st_s(reinterpret_cast<T*>(a.dst) + b * a.dst_stride + p, z.x, nt);
    } else if constexpr (SK == ST_UNFOLD) {  // (z * untwist[p]) -> out[p] = re, out[p+L] = im
        z = z * ldc(a.stab + p);
        T* o = reinterpret_cast<T*>(a.dst) + b * a.dst_stride;
        int L = 1 << a.n;
        st_s(o + p, z.x, nt);
        st_s(o + p + L, z.y, nt);
    } else {  // ST_SUM_F64
        const cplx<T> h = ldc_s(a.stab + b * a.dst_stride + p, nt);
        stc_s(reinterpret_cast<cplx<double>*>(a.dst) + b * a.dst_stride + p,
              mk<double>((double)h.x + (double)z.x, (double)h.y + (double)z.y), nt);
    }
}

// Position (within its transform) of tile row r, column col.
struct TileGeo {
    int hi_pos;   // (hi << (blo+K)) | c0
    int blo, K, logC;
    rsrc_t tab;   // root-of-unity table
    __device__ __forceinline__ int pos(int r, int col) const
    {
        return hi_pos | ((r & ((1 << K) - 1)) << blo) | col;
    }
};

template <typename T, int MODE>
__device__ __forceinline__ int twiddle_index(int n, int nega, int b, int p)
{
    if constexpr (mode_special(MODE)) {
        return (1 << b) + (p & ((1 << b) - 1));
    } else {
        int base = nega ? (1 << (n - 1 - b)) : 0;
        return base + (p >> (b + 1));
    }
}

// Runs the RHO stages of one group held in registers: v[j] sits at tile row
// rbase | (j << q), column col.  Standard twiddles depend only on the group
// bits above the stage bit i, special ones only on the bits below it, so
// each distinct twiddle is loaded once.
template <typename T, int MODE, int RHO>
__device__ __forceinline__ void group_stages(cplx<T>* v, const PassArgs<T>& a, const TileGeo& g,
                                             int rbase, int col, int q)
{
    constexpr int GS = 1 << RHO;
#pragma unroll
    for (int st = 0; st < RHO; st++) {
        const int i = mode_descending(MODE) ? (RHO - 1 - st) : st;
        const int bpos = g.blo + q + i;  // position bit of this stage
#pragma unroll
        for (int outer = 0; outer < (mode_special(MODE) ? (1 << i) : (GS >> (i + 1))); outer++) {
            const int j_first = mode_special(MODE) ? outer : (outer << (i + 1));
            const int p0 = g.pos(rbase | (j_first << q), col);
            const cplx<T> w = bldc<T>(g.tab, twiddle_index<T, MODE>(a.n, a.nega, bpos, p0));
#pragma unroll
            for (int inner = 0; inner < (mode_special(MODE) ? (GS >> (i + 1)) : (1 << i)); inner++) {
                const int j0 = mode_special(MODE) ? ((inner << (i + 1)) | outer)
                                                  : ((outer << (i + 1)) | inner);
                const int j1 = j0 | (1 << i);
                cplx<T> U = v[j0], V = v[j1];
                if constexpr (mode_is_ct(MODE)) {
                    cplx<T> Vw = V * w;
                    v[j0] = U + Vw;
                    v[j1] = U - Vw;
                } else {
                    v[j0] = U + V;
                    v[j1] = (U - V) * w;
                }
            }
        }
    }
}

// Standard-mode stages with derived twiddles.  Inside a step over bits
// [q, q+RHO) every twiddle of the stage with in-group bit i is
//     table[base + ((P >> (q+RHO)) << s) + t] = table[base + ((P >> (q+RHO)) << s)]
//                                             * exp(2 pi i bitrev_s(t) / 2^(s+1)),
// s = RHO-1-i, t < 2^s the group bits above i (the reference tables hold
// bit-reversed powers, fft_cpu.cu:200-211 and 434-446, so an index t below
// a multiple of 2^s only adds a root of order 2^(s+1)).  One table load per
// stage; wc[2^s + t] holds the root constants (s < 5), or, with TW, the
// root table itself does at wc[t]: every reference table (FFT<T>, FFNT<T>,
// forward or inverse, either reduction) has tab[t] = exp(+-2 pi i bitrev_s(t)
// / 2^(s+1)) for t < 2^s.  The values differ from direct table reads by
// rounding only.
template <typename T, int MODE, int RHO, bool TW = false>
__device__ __forceinline__ void group_stages_derived(cplx<T>* v, const PassArgs<T>& a,
                                                     const TileGeo& g, int rbase, int col, int q,
                                                     const cplx<T>* __restrict__ wc)
{
    static_assert(!mode_special(MODE), "derived twiddles are for the standard modes");
#pragma unroll
    for (int st = 0; st < RHO; st++) {
        const int i = mode_descending(MODE) ? (RHO - 1 - st) : st;
        const int s = RHO - 1 - i;
        const int bpos = g.blo + q + i;
        const int p0 = g.pos(rbase, col);
        const cplx<T> base = bldc<T>(g.tab, twiddle_index<T, MODE>(a.n, a.nega, bpos, p0));
#pragma unroll
        for (int t = 0; t < (1 << s); t++) {
            cplx<T> w = base;
            if (t == 1 && s == 1) {
                // times exp(+-i pi/2), exact: forward chains use the roots,
                // inverse chains (Gentleman-Sande) their conjugates
                w = mode_is_ct(MODE) ? mk<T>(-base.y, base.x) : mk<T>(base.y, -base.x);
            } else if (t > 0) {
                w = base * wc[TW ? t : (1 << s) + t];
            }
#pragma unroll
            for (int lo = 0; lo < (1 << i); lo++) {
                const int j0 = (t << (i + 1)) | lo;
                const int j1 = j0 | (1 << i);
                cplx<T> U = v[j0], V = v[j1];
                if constexpr (mode_is_ct(MODE)) {
                    cplx<T> Vw = V * w;
                    v[j0] = U + Vw;
                    v[j1] = U - Vw;
                } else {
                    v[j0] = U + V;
                    v[j1] = (U - V) * w;
                }
            }
        }
    }
}

// Group geometry of thread `tid`, k-th group, for a step over tile-row bits
// [q, q+RHO): the group's rows are rbase | (j << q).
template <int RHO, int NT>
__device__ __forceinline__ void group_of(int tid, int k, int logC, int q, int& rbase, int& col)
{
    const int gamma = tid + k * NT;
    col = gamma & ((1 << logC) - 1);
    const int rest = gamma >> logC;
    const int rlo = rest & ((1 << q) - 1);
    const int rhi = rest >> q;
    rbase = (rhi << (q + RHO)) | rlo;
}

// One radix-2^RHO step over tile-row bits [q, q+RHO): every thread takes
// EPT/2^RHO groups of 2^RHO elements from LDS, runs the RHO stages in
// registers and writes them back.
template <typename T, int MODE, int E, int RHO, int NT, bool DRV = false, bool TW = false>
__device__ __forceinline__ void radix_step(cplx<T>* lds, const PassArgs<T>& a, const TileGeo& g,
                                           int q, int tid, const cplx<T>* wc = nullptr)
{
    constexpr int GS = 1 << RHO;
    constexpr int GPT = (1 << E) / GS;
    // with derived twiddles (fused kernels, tight on registers) the groups
    // of a thread are processed one after the other
    constexpr int UNROLL = DRV ? 1 : GPT;
#pragma unroll UNROLL
    for (int k = 0; k < GPT; k++) {
        int rbase, col;
        group_of<RHO, NT>(tid, k, g.logC, q, rbase, col);
        cplx<T> v[GS];
#pragma unroll
        for (int j = 0; j < GS; j++) v[j] = lds[swz(((rbase | (j << q)) << g.logC) | col)];
        if constexpr (DRV)
            group_stages_derived<T, MODE, RHO, TW>(v, a, g, rbase, col, q, wc);
        else
            group_stages<T, MODE, RHO>(v, a, g, rbase, col, q);
#pragma unroll
        for (int j = 0; j < GS; j++) lds[swz(((rbase | (j << q)) << g.logC) | col)] = v[j];
    }
}

// Bit range of step s (0-based, in execution order) of a pass of K bits made
// of radix-2^E steps: the one short step sits at the top end of the range,
// so a descending chain runs it first and an ascending chain last, and the
// last step of a descending chain covers the same bits [0, E) as the first
// step of the ascending chain that inverts it.
__host__ __device__ __forceinline__ void step_range(bool descending, int K, int E, int s, int& q,
                                                    int& rho)
{
    const int full = K / E;
    const int rem = K - full * E;
    const int nsteps = full + (rem ? 1 : 0);
    const int as = descending ? (nsteps - 1 - s) : s;  // index in ascending order
    q = as * E;
    rho = (as < full) ? E : rem;
}

__host__ __device__ __forceinline__ int num_steps(int K, int E) { return (K + E - 1) / E; }

template <typename T, int MODE, int E, int NT, bool DRV = false>
__device__ __forceinline__ void radix_step_rt(cplx<T>* lds, const PassArgs<T>& a, const TileGeo& g,
                                              int q, int rho, int tid, const cplx<T>* wc)
{
    switch (rho) {
        case 1: radix_step<T, MODE, E, 1, NT, DRV>(lds, a, g, q, tid, wc); break;
        case 2: if constexpr (E >= 2) radix_step<T, MODE, E, 2, NT, DRV>(lds, a, g, q, tid, wc); break;
        case 3: if constexpr (E >= 3) radix_step<T, MODE, E, 3, NT, DRV>(lds, a, g, q, tid, wc); break;
        case 4: if constexpr (E >= 4) radix_step<T, MODE, E, 4, NT, DRV>(lds, a, g, q, tid, wc); break;
        case 5: if constexpr (E >= 5) radix_step<T, MODE, E, 5, NT, DRV>(lds, a, g, q, tid, wc); break;
        default: break;
    }
}

// Runs steps [s_begin, s_end) of the pass, each between two barriers; the
// steps are radix-2^RB (RB <= E; a thread holds 2^E elements).
template <typename T, int MODE, int E, int NT, bool DRV = false, int RB = E>
__device__ __forceinline__ void run_steps(cplx<T>* lds, const PassArgs<T>& a, const TileGeo& g,
                                          int tid, int s_begin, int s_end,
                                          const cplx<T>* wc = nullptr)
{
    for (int s = s_begin; s < s_end; s++) {
        int q, rho;
        step_range(mode_descending(MODE), a.K, RB, s, q, rho);
        __syncthreads();
        radix_step_rt<T, MODE, E, NT, DRV>(lds, a, g, q, rho, tid, wc);
    }
    __syncthreads();
}

// Steps [S, SEND) of a pass of K bits with radix-2^RB steps, unrolled at
// compile time so that every step's bit range is a constant (address and
// twiddle-index arithmetic then folds).
template <typename T, int MODE, int E, int NT, int K, int RB, int S, int SEND, bool DRV,
          bool TW = false>
__device__ __forceinline__ void run_steps_ct(cplx<T>* lds, const PassArgs<T>& a, const TileGeo& g,
                                             int tid, const cplx<T>* wc)
{
    if constexpr (S < SEND) {
        constexpr int full = K / RB;
        constexpr int rem = K - full * RB;
        constexpr int nst = full + (rem ? 1 : 0);
        constexpr int as = mode_descending(MODE) ? (nst - 1 - S) : S;
        constexpr int q = as * RB;
        constexpr int rho = (as < full) ? RB : rem;
        __syncthreads();
        radix_step<T, MODE, E, rho, NT, DRV, TW>(lds, a, g, q, tid, wc);
        run_steps_ct<T, MODE, E, NT, K, RB, S + 1, SEND, DRV, TW>(lds, a, g, tid, wc);
    } else {
        __syncthreads();
    }
}

// run_steps_ct for workgroups larger than the tile's thread count: every
// thread takes the barriers, only the first NT threads compute.
template <typename T, int MODE, int E, int NT, int K, int RB, int S, int SEND>
__device__ __forceinline__ void run_steps_ct_active(cplx<T>* lds, const PassArgs<T>& a,
                                                    const TileGeo& g, int tid, const cplx<T>* wc,
                                                    bool active)
{
    if constexpr (S < SEND) {
        constexpr int full = K / RB;
        constexpr int rem = K - full * RB;
        constexpr int nst = full + (rem ? 1 : 0);
        constexpr int as = mode_descending(MODE) ? (nst - 1 - S) : S;
        constexpr int q = as * RB;
        constexpr int rho = (as < full) ? RB : rem;
        __syncthreads();
        if (active) radix_step<T, MODE, E, rho, NT, true>(lds, a, g, q, tid, wc);
        run_steps_ct_active<T, MODE, E, NT, K, RB, S + 1, SEND>(lds, a, g, tid, wc, active);
    } else {
        __syncthreads();
    }
}

template <typename T, int MODE, int LOGT, int E, int LK, int SK>
__global__ __launch_bounds__((1 << LOGT) >> E) void pass_kernel(PassArgs<T> a)
{
    constexpr int TILE = 1 << LOGT;
    constexpr int EPT = 1 << E;
    constexpr int NT = TILE / EPT;
    __shared__ cplx<T> lds[TILE];

    const int tid = threadIdx.x;
    const int n = a.n;
    const int logC = a.logC;
    const int bhi = a.blo + a.K;
    TileGeo g;
    g.blo = a.blo;
    g.K = a.K;
    g.logC = logC;
    g.tab = make_rsrc(a.tab);

    // tile -> (transform block, hi bits above the pass, column block)
    const int ncb_log = a.blo - logC;
    const int nhi_log = n - bhi;
    long long tile = blockIdx.x;
    const int cb = (int)(tile & ((1LL << ncb_log) - 1));
    tile >>= ncb_log;
    const int hi = (int)(tile & ((1LL << nhi_log) - 1));
    tile >>= nhi_log;
    const long long bt0 = tile << a.logBT;  // first transform of this tile
    g.hi_pos = (hi << bhi) | (cb << logC);
    const int C = 1 << logC;
    const int Kmask = (1 << a.K) - 1;

    // ---- load: consecutive lanes on consecutive positions
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        int idx = tid + i * NT;
        int col = idx & (C - 1);
        int rext = idx >> logC;
        long long b = bt0 + (rext >> a.K);
        int p = g.pos(rext & Kmask, col);
        cplx<T> v = mk<T>(T(0), T(0));
        if (b < a.batch) v = load_elem<T, LK>(a, b, p);
        lds[swz(idx)] = v;
    }

    run_steps<T, MODE, E, NT>(lds, a, g, tid, 0, num_steps(a.K, E));

    // ---- store
#pragma unroll
    for (int i = 0; i < EPT; i++) {
        int idx = tid + i * NT;
        int col = idx & (C - 1);
        int rext = idx >> logC;
        long long b = bt0 + (rext >> a.K);
        int p = g.pos(rext & Kmask, col);
        if (b < a.batch) store_elem<T, SK>(a, b, p, lds[swz(idx)]);
    }
}

}  // namespace gfa
