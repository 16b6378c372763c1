// Device-side complex arithmetic for the gfx950 kernels.
//
// The storage layout is the reference's COMPLEX<T>: interleaved (re, im) of
// 8 bytes (float) or 16 bytes (double) (src/include/gpufft/complex.cuh:14-16,30),
// so caller buffers and root tables are read and written unchanged.
// The arithmetic follows complex.cuh:65-139 (textbook 4-mul product); hipcc
// contracts a*b+c into FMA, which changes results only in the last bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gfa {

template <typename T> struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T> __host__ __device__ __forceinline__ cplx<T> mk(T re, T im)
{
    cplx<T> z;
    z.x = re;
    z.y = im;
    return z;
}
template <typename T> __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T> __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T> __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b)
{
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// 16-byte / 8-byte vector loads and stores of one complex element.
template <typename T> __device__ __forceinline__ cplx<T> ldc(const cplx<T>* p)
{
    return *p;
}
template <typename T> __device__ __forceinline__ void stc(cplx<T>* p, cplx<T> v)
{
    *p = v;
}

}  // namespace gfa
