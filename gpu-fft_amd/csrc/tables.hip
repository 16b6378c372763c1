// Root-of-unity tables generated on the GPU in FP64, in the reference's
// layouts (the host generators are FFT<T> / FFNT<T>, src/lib/fft_cpu.cu:93-129,
// 200-226, 234-299, 434-470; the special-FFT layout is the one
// src/lib/fft.cu:782,908 index).  Entry k is exp(+-2 pi i e(k) / M) with an
// exact rational angle e(k)/M, evaluated by sincospi in double and, for the
// float tables, rounded once to float.
#include <hip/hip_runtime.h>

#include "status.hpp"
#include "cplx.hpp"

using namespace gfa;

namespace {

__device__ __forceinline__ int brev(int x, int bits)
{
    return bits == 0 ? 0 : (int)(__brev((unsigned)x) >> (32 - bits));
}

__device__ __forceinline__ long long powmod5(long long e, long long m)
{
    long long r = 1 % m, b = 5 % m;
    while (e) {
        if (e & 1) r = (r * b) % m;
        b = (b * b) % m;
        e >>= 1;
    }
    return r;
}

template <typename T>
__global__ void table_kernel(cplx<T>* out, int kind, int count, int logm, int inverse)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    long long num;  // angle = 2*pi*num / den
    long long den;
    switch (kind) {
        case GPUFFT_TABLE_FFT:  // FFT<T>(count).ReverseRootTable(): omega_{2n}^{bitrev(k)}
            num = brev(k, logm);
            den = 2LL * count;
            break;
        case GPUFFT_TABLE_FFNT_ROOT:  // FFNT<T>(2L).ReverseRootTable_ffnt(): omega_L^{bitrev_{logL-1}(k)}
            num = brev(k, logm - 1 < 0 ? 0 : logm - 1);
            den = count;
            break;
        case GPUFFT_TABLE_FFNT_TWIST:  // twist_table_ffnt(): omega_{4L}^k
            num = k;
            den = 4LL * count;
            break;
        case GPUFFT_TABLE_STAGE_ROOTS: {  // [2^s + t] = exp(2 pi i bitrev_s(t) / 2^(s+1))
            if (k == 0) {
                num = 0;
                den = 1;
            } else {
                const int s = 31 - __clz(k);
                num = brev(k - (1 << s), s);
                den = 2LL << s;
            }
            break;
        }
        default: {  // GPUFFT_TABLE_SPECIAL: T[h+j] = zeta_{8h}^{5^j mod 8h}
            if (k == 0) {
                num = 0;
                den = 1;
            } else {
                long long h = 1LL << (31 - __clz(k));
                long long j = k - h;
                den = 8 * h;
                num = powmod5(j, den);
            }
        }
    }
    double s, c;
    sincospi(2.0 * (double)num / (double)den, &s, &c);
    if (inverse) s = -s;
    out[k] = mk<T>((T)c, (T)s);
}

template <typename T> int gen(void* dst, int kind, int n_power, int inverse, void* stream)
{
    if (!dst) return fail(GPUFFT_ERR_NULL_POINTER, "gpufft_generate_table: dst is NULL");
    if (n_power < 0 || n_power > 26)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, "gpufft_generate_table: n_power out of range");
    if (kind < GPUFFT_TABLE_FFT || kind > GPUFFT_TABLE_STAGE_ROOTS)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, "gpufft_generate_table: unknown kind");
    const int count = 1 << n_power;
    hipLaunchKernelGGL(table_kernel<T>, dim3((count + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, (cplx<T>*)dst, kind, count, n_power, inverse);
    GFA_CHECK_LAUNCH("gpufft_generate_table");
    return ok();
}

}  // namespace

extern "C" {

int gpufft_generate_table_f64(void* dst, int kind, int n_power, int inverse, void* stream)
{
    return gen<double>(dst, kind, n_power, inverse, stream);
}

int gpufft_generate_table_f32(void* dst, int kind, int n_power, int inverse, void* stream)
{
    return gen<float>(dst, kind, n_power, inverse, stream);
}

}  // extern "C"
