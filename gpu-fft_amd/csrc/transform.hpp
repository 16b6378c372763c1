// Host-side planning and launch of a full transform as a chain of passes.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "engine.hpp"
#include "oneshot.hpp"
#include "status.hpp"

namespace gfa {

// Tile geometry per element type: 2^LOGT elements of 16 bytes (double) or
// 8 bytes (float) = 64 KiB of LDS, 16 elements per thread, 256 / 512 threads.
template <typename T> struct TileCfg;
template <> struct TileCfg<double> {
    static constexpr int LOGT = 12;
    static constexpr int E = 4;
};
template <> struct TileCfg<float> {
    static constexpr int LOGT = 13;
    static constexpr int E = 4;
};

struct PassPlan {
    int blo, K, logC, logBT;
};

// Splits the n position bits into passes of at most LOGT bits.  The pass
// that reaches bit 0 takes LOGT bits (contiguous tiles); the others take at
// most LOGT-4 bits so that every tile row is at least 16 contiguous elements.
// Descending modes run the top bits first, ascending modes bit 0 first.
inline std::vector<PassPlan> plan_passes(int n, int LOGT, bool descending)
{
    std::vector<PassPlan> p;
    if (n <= LOGT) {
        p.push_back({0, n, 0, LOGT - n});
        return p;
    }
    int rest = n - LOGT;
    int maxk = LOGT - 4;
    int cnt = (rest + maxk - 1) / maxk;
    std::vector<int> ks;
    for (int i = 0; i < cnt; i++) ks.push_back(rest / cnt + (i < rest % cnt ? 1 : 0));
    // ascending order of bits: [0,LOGT), then the chunks upward
    std::vector<PassPlan> asc;
    asc.push_back({0, LOGT, 0, 0});
    int b = LOGT;
    for (int k : ks) {
        asc.push_back({b, k, LOGT - k, 0});
        b += k;
    }
    if (descending) {
        for (int i = (int)asc.size() - 1; i >= 0; i--) p.push_back(asc[i]);
    } else {
        p = asc;
    }
    return p;
}

// What the first pass reads and the last pass writes.
template <typename T> struct TransformIO {
    int lk;                 // LoadKind of the first pass
    const void* src;
    const void* src2;
    const cplx<T>* ltab;
    long long src_stride;
    int sk;                 // StoreKind of the last pass
    void* dst;
    const cplx<T>* stab;
    long long dst_stride;
    cplx<T>* work;          // complex [batch][2^n] buffer for intermediate passes
    bool do_scale;
    cplx<T> scale;
};

template <typename T>
using PassFn = void (*)(PassArgs<T>);

template <typename T, int MODE, int LK, int SK, int BIG> inline PassFn<T> pass_fn()
{
    return &pass_kernel<T, MODE, TileCfg<T>::LOGT + BIG, TileCfg<T>::E, LK, SK>;
}

// Kernel for (mode, load kind, store kind); nullptr for combinations no
// operation uses.  BIG = 1: tile of 2^(LOGT+1) elements (twice the LDS and
// threads), so that a transform of exactly 2^(LOGT+1) points is one pass.
template <typename T, int BIG> PassFn<T> select_pass(int mode, int lk, int sk)
{
#define GFA_SEL(M, L, S) \
    if (mode == M && lk == L && sk == S) return pass_fn<T, M, L, S, BIG>();
    GFA_SEL(MODE_STD_FWD, LD_CPLX, ST_CPLX)
    GFA_SEL(MODE_STD_FWD, LD_REAL, ST_CPLX)
    GFA_SEL(MODE_STD_FWD, LD_FOLD, ST_CPLX)
    GFA_SEL(MODE_STD_INV, LD_CPLX, ST_CPLX)
    GFA_SEL(MODE_STD_INV, LD_MUL, ST_CPLX)
    GFA_SEL(MODE_STD_INV, LD_CPLX, ST_REAL)
    GFA_SEL(MODE_STD_INV, LD_MUL, ST_REAL)
    GFA_SEL(MODE_STD_INV, LD_CPLX, ST_UNFOLD)
    GFA_SEL(MODE_STD_INV, LD_RR_PACK, ST_CPLX)
    GFA_SEL(MODE_STD_INV, LD_MUL, ST_UNFOLD)
    GFA_SEL(MODE_SPE_FWD, LD_CPLX, ST_CPLX)
    GFA_SEL(MODE_SPE_INV, LD_CPLX, ST_CPLX)
    if constexpr (sizeof(T) == 4) {
        GFA_SEL(MODE_SPE_FWD, LD_SPLIT_HI, ST_CPLX)
        GFA_SEL(MODE_SPE_FWD, LD_SPLIT_LO, ST_CPLX)
        GFA_SEL(MODE_SPE_FWD, LD_SPLIT_LO, ST_SUM_F64)
        GFA_SEL(MODE_SPE_FWD, LD_CPLX, ST_SUM_F64)
        GFA_SEL(MODE_SPE_INV, LD_SPLIT_HI, ST_CPLX)
        GFA_SEL(MODE_SPE_INV, LD_SPLIT_LO, ST_CPLX)
        GFA_SEL(MODE_SPE_INV, LD_SPLIT_LO, ST_SUM_F64)
        GFA_SEL(MODE_SPE_INV, LD_CPLX, ST_SUM_F64)
    }
#undef GFA_SEL
    return nullptr;
}

// Launches every pass of one transform of length 2^n over `batch` transforms.
template <typename T>
int run_transform(int mode, int n, int batch, int nega, const cplx<T>* tab,
                  const TransformIO<T>& io, hipStream_t stream, const char* who)
{
    constexpr int LOGT0 = TileCfg<T>::LOGT;
    // double, n = LOGT0 + 1: one pass over a double-size tile (128 KiB of LDS,
    // 512 threads) instead of two passes (for float the doubled tile would
    // need 1024 threads, whose 128-register budget spills)
    const bool big = sizeof(T) == 8 && n == LOGT0 + 1;
    const int LOGT = big ? LOGT0 + 1 : LOGT0;
    const int NT = (1 << LOGT) >> TileCfg<T>::E;
    std::vector<PassPlan> plan = plan_passes(n, LOGT, mode_descending(mode));
    const long long L = 1LL << n;
    for (size_t i = 0; i < plan.size(); i++) {
        const bool first = (i == 0), last = (i + 1 == plan.size());
        const PassPlan& pp = plan[i];
        PassArgs<T> a{};
        a.n = n;
        a.blo = pp.blo;
        a.K = pp.K;
        a.logC = pp.logC;
        a.logBT = pp.logBT;
        a.batch = batch;
        a.nega = nega;
        a.tab = tab;
        int lk = LD_CPLX, sk = ST_CPLX;
        a.nt_in = first ? 1 : 0;
        a.nt_out = last ? 1 : 0;
        if (first) {
            lk = io.lk;
            a.src = io.src;
            a.src2 = io.src2;
            a.ltab = io.ltab;
            a.src_stride = io.src_stride;
        } else {
            a.src = io.work;
            a.src_stride = L;
        }
        if (last) {
            sk = io.sk;
            a.dst = io.dst;
            a.stab = io.stab;
            a.dst_stride = io.dst_stride;
            a.do_scale = io.do_scale ? 1 : 0;
            a.scale = io.scale;
        } else {
            a.dst = io.work;
            a.dst_stride = L;
        }
        PassFn<T> fn = select_pass<T, 0>(mode, lk, sk);
        if constexpr (sizeof(T) == 8)
            if (big) fn = select_pass<T, 1>(mode, lk, sk);
        if (!fn) return fail(GPUFFT_ERR_UNSUPPORTED, std::string(who) + ": no kernel for this pass");
        long long tiles = ((long long)batch << n) >> LOGT;
        if (pp.logBT > 0) tiles = ((long long)batch + (1LL << pp.logBT) - 1) >> pp.logBT;
        if (tiles <= 0) tiles = 1;
        if (tiles > 0x7fffffffLL) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch too large");
        hipLaunchKernelGGL(fn, dim3((unsigned)tiles), dim3(NT), 0, stream, a);
        GFA_CHECK_LAUNCH(who);
    }
    return ok();
}

// Out-of-place calls (R_R, FFNT) of 2^n points, n = S + SB with S the LDS
// tile (13 double, 14 float) and SB = 1..3 (oneshot.hpp): the forward is one
// launch of Q = 2^SB workgroups per transform; the inverse is a launch of
// the in-sub-problem stages (in place in the source) and a streaming launch
// of the top SB stages with the output transform.
template <typename T> struct OneCfg;
// The forward's workgroups read the whole transform, so it takes the largest
// tile (one workgroup per CU); the inverse's first launch reads only its own
// sub-problem and takes half of it, so that two workgroups share a CU
// (measured at N = 2^16, batch 256, FP64: inverse 0.146 -> 0.135 ms; the
// forward at the smaller tile went 0.150 -> 0.161 ms).
template <> struct OneCfg<double> {
    static constexpr int S = 13, E = 4, RB = 4;   // forward: 512 threads, 16 points each
    static constexpr int SI = 12, EI = 4;         // inverse: 256 threads, 16 points each
};
template <> struct OneCfg<float> {
    static constexpr int S = 14, E = 5, RB = 4;   // forward: 512 threads, 32 points each
    static constexpr int SI = 13, EI = 5;         // inverse: 256 threads, 32 points each
};

template <typename T> using OneFn = void (*)(OneArgs<T>);
template <typename T> using TopFn = void (*)(OneArgs<T>, int);

template <typename T, int SB> OneFn<T> select_forward_sb(int lk)
{
    constexpr int S = OneCfg<T>::S, E = OneCfg<T>::E, RB = OneCfg<T>::RB;
    if (lk == LD_REAL) return &fwd_sub_kernel<T, S, SB, E, RB, PL_REAL>;
    if (lk == LD_FOLD) return &fwd_sub_kernel<T, S, SB, E, RB, PL_FOLD>;
    return nullptr;
}

template <typename T, int SB> TopFn<T> select_top_sb(int sk)  // SB = 2..4
{
    if (sk == ST_REAL) return &inv_top_kernel<T, SB, ST_REAL>;
    if (sk == ST_UNFOLD) return &inv_top_kernel<T, SB, ST_UNFOLD>;
    return nullptr;
}

// Returns true and launches when this path covers the transform (status in
// *status); false leaves the pass chain to the caller.  The source must not
// overlap the destination; the inverse uses the source as its work buffer.
template <typename T>
bool run_oneshot(int mode, int n, int batch, int nega, const cplx<T>* tab,
                 const TransformIO<T>& io, hipStream_t stream, const char* who, int* status)
{
    constexpr int S = OneCfg<T>::S, E = OneCfg<T>::E, RB = OneCfg<T>::RB;
    const int SB = n - S;
    if (SB < 1 || SB > 3) return false;
    OneArgs<T> a{};
    a.n = n;
    a.batch = batch;
    a.nega = nega;
    a.tab = tab;
    a.twist = io.ltab;
    a.src = io.src;
    a.src2 = io.src2;
    a.src_stride = io.src_stride;
    a.dst = io.dst;
    a.dst_stride = io.dst_stride;
    a.untwist = io.stab;
    a.do_scale = io.do_scale ? 1 : 0;
    a.scale = io.do_scale ? io.scale : mk<T>(T(1), T(0));
    if (mode == MODE_STD_FWD && io.sk == ST_CPLX && (io.lk == LD_REAL || io.lk == LD_FOLD)) {
        OneFn<T> fn = SB == 1 ? select_forward_sb<T, 1>(io.lk)
                    : SB == 2 ? select_forward_sb<T, 2>(io.lk)
                              : select_forward_sb<T, 3>(io.lk);
        const long long blocks = (((long long)batch + 7) / 8) * 8 * (1LL << SB);
        if (blocks > 0x7fffffffLL) {
            *status = fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch too large");
            return true;
        }
        hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3((1 << S) >> E), 0, stream, a);
    } else if (mode == MODE_STD_INV && (io.lk == LD_CPLX || io.lk == LD_MUL) &&
               (io.sk == ST_REAL || io.sk == ST_UNFOLD)) {
        constexpr int SI = OneCfg<T>::SI, EI = OneCfg<T>::EI;
        const int SBI = n - SI;  // 2..4
        const long long blocks1 = (long long)batch << SBI;
        static_assert(SI >= 9, "the second inverse launch takes 512 columns per workgroup");
        const long long blocks2 = (long long)batch << (SI - 9);
        if (blocks1 > 0x7fffffffLL || blocks2 > 0x7fffffffLL) {
            *status = fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch too large");
            return true;
        }
        OneFn<T> f1 = io.lk == LD_MUL ? &inv_local_kernel<T, SI, EI, RB, LD_MUL>
                                      : &inv_local_kernel<T, SI, EI, RB, LD_CPLX>;
        hipLaunchKernelGGL(f1, dim3((unsigned)blocks1), dim3((1 << SI) >> EI), 0, stream, a);
        TopFn<T> f2 = SBI == 2 ? select_top_sb<T, 2>(io.sk)
                    : SBI == 3 ? select_top_sb<T, 3>(io.sk)
                               : select_top_sb<T, 4>(io.sk);
        hipLaunchKernelGGL(f2, dim3((unsigned)blocks2), dim3(256), 0, stream, a, SI);
    } else {
        return false;
    }
    hipError_t e = hipGetLastError();
    *status = e != hipSuccess ? fail_hip(e, who) : ok();
    return true;
}

}  // namespace gfa
