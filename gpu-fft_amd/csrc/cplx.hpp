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

// Buffer resources: a 64-bit base in scalar registers plus a 32-bit byte
// offset per lane, so that every access costs one VGPR of address instead of
// two, and an access beyond num_records bytes reads 0 / is dropped.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr int RSRC_FLAGS = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* base, unsigned bytes = 0x7fffffffu)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             RSRC_FLAGS);
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

template <typename T> __device__ __forceinline__ cplx<T> bld(rsrc_t r, unsigned off);
template <> __device__ __forceinline__ cplx<double> bld<double>(rsrc_t r, unsigned off)
{
    u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    cplx<double> z;
    __builtin_memcpy(&z, &v, 16);
    return z;
}
template <> __device__ __forceinline__ cplx<float> bld<float>(rsrc_t r, unsigned off)
{
    u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    cplx<float> z;
    __builtin_memcpy(&z, &v, 8);
    return z;
}
// complex element `idx` of a buffer of cplx<T>
template <typename T> __device__ __forceinline__ cplx<T> bldc(rsrc_t r, int idx)
{
    return bld<T>(r, (unsigned)idx * (unsigned)sizeof(cplx<T>));
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

// Streaming load (non-temporal hint): data read once.
template <typename T> __device__ __forceinline__ cplx<T> ldc_nt(const cplx<T>* p)
{
    typedef T vec2 __attribute__((ext_vector_type(2)));
    const vec2 w = __builtin_nontemporal_load(reinterpret_cast<const vec2*>(p));
    return mk<T>(w.x, w.y);
}

// Streaming store (non-temporal hint): data written once and read back by a
// later launch, kept from displacing the lines the current launch re-reads.
template <typename T> __device__ __forceinline__ void stc_nt(cplx<T>* p, cplx<T> v)
{
    typedef T vec2 __attribute__((ext_vector_type(2)));
    vec2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<vec2*>(p));
}

}  // namespace gfa
