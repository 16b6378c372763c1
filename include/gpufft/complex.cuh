// COMPLEX<T>: the complex element type of the GPU-FFT interface, for HIP.
//
// Same names, layout and operations as the reference header
// (src/include/gpufft/complex.cuh:14-318): storage is HIP's float2/double2
// (interleaved re, im; 8 or 16 bytes), usable in host code and in device
// code compiled by hipcc.  Two deliberate differences, both bug fixes:
// `z * scalar` scales both parts (the reference writes real*s into the
// imaginary part, complex.cuh:153-159), and the compound assignments are
// non-const members that modify *this.
#ifndef GPUFFT_AMD_COMPLEX_CUH
#define GPUFFT_AMD_COMPLEX_CUH

#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <iostream>
#include <type_traits>
#include <vector>

template <typename T>
using fp2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;

typedef float Float32;
typedef double Float64;

template <typename T>
using Float = typename std::conditional<std::is_same<T, float>::value, Float32, Float64>::type;

namespace complex_arithmetic {

template <typename T1> struct ComplexOperations {
    fp2<T1> data;

    __host__ __device__ __forceinline__ T1& real() { return data.x; }
    __host__ __device__ __forceinline__ T1& imag() { return data.y; }
    __host__ __device__ __forceinline__ const T1& real() const { return data.x; }
    __host__ __device__ __forceinline__ const T1& imag() const { return data.y; }

    template <typename T2> __host__ __device__ __forceinline__ ComplexOperations(T2 r, T2 i)
    {
        data.x = static_cast<T1>(r);
        data.y = static_cast<T1>(i);
    }
    template <typename T2> __host__ __device__ __forceinline__ ComplexOperations(T2 r)
    {
        data.x = static_cast<T1>(r);
        data.y = static_cast<T1>(0);
    }
    __host__ __device__ __forceinline__ ComplexOperations()
    {
        data.x = static_cast<T1>(0);
        data.y = static_cast<T1>(0);
    }

    __host__ __device__ __forceinline__ ComplexOperations operator+(const ComplexOperations& o) const
    {
        return ComplexOperations(real() + o.real(), imag() + o.imag());
    }
    __host__ __device__ __forceinline__ ComplexOperations operator-(const ComplexOperations& o) const
    {
        return ComplexOperations(real() - o.real(), imag() - o.imag());
    }
    // textbook 4-multiply product (reference complex.cuh:133-139)
    __host__ __device__ __forceinline__ ComplexOperations operator*(const ComplexOperations& o) const
    {
        return ComplexOperations(real() * o.real() - imag() * o.imag(),
                                 real() * o.imag() + imag() * o.real());
    }
    __host__ __device__ __forceinline__ ComplexOperations operator/(const ComplexOperations& o) const
    {
        T1 den = o.real() * o.real() + o.imag() * o.imag();
        return ComplexOperations((real() * o.real() + imag() * o.imag()) / den,
                                 (imag() * o.real() - real() * o.imag()) / den);
    }
    template <typename T2>
    __host__ __device__ __forceinline__ ComplexOperations operator+(const T2& s) const
    {
        return ComplexOperations(real() + static_cast<T1>(s), imag());
    }
    template <typename T2>
    __host__ __device__ __forceinline__ ComplexOperations operator-(const T2& s) const
    {
        return ComplexOperations(real() - static_cast<T1>(s), imag());
    }
    template <typename T2>
    __host__ __device__ __forceinline__ ComplexOperations operator*(const T2& s) const
    {
        return ComplexOperations(real() * static_cast<T1>(s), imag() * static_cast<T1>(s));
    }
    template <typename T2>
    __host__ __device__ __forceinline__ ComplexOperations operator/(const T2& s) const
    {
        return ComplexOperations(real() / static_cast<T1>(s), imag() / static_cast<T1>(s));
    }
    __host__ __device__ __forceinline__ ComplexOperations& operator+=(const ComplexOperations& o)
    {
        return *this = *this + o;
    }
    __host__ __device__ __forceinline__ ComplexOperations& operator-=(const ComplexOperations& o)
    {
        return *this = *this - o;
    }
    __host__ __device__ __forceinline__ ComplexOperations& operator*=(const ComplexOperations& o)
    {
        return *this = *this * o;
    }
    __host__ __device__ __forceinline__ ComplexOperations& operator/=(const ComplexOperations& o)
    {
        return *this = *this / o;
    }

    __host__ __device__ __forceinline__ ComplexOperations conjugate() const
    {
        return ComplexOperations(real(), -imag());
    }
    __host__ __device__ __forceinline__ ComplexOperations negate() const
    {
        return ComplexOperations(-real(), -imag());
    }
    __host__ __device__ __forceinline__ T1 square() const { return real() * real() + imag() * imag(); }
    __host__ __device__ __forceinline__ ComplexOperations inverse() const
    {
        T1 sq = square();
        return ComplexOperations(real() / sq, -imag() / sq);
    }
    // integer power by square-and-multiply (reference complex.cuh:244-291)
    template <typename I> __host__ __device__ __forceinline__ ComplexOperations pow_int(I e) const
    {
        ComplexOperations r(static_cast<T1>(1), static_cast<T1>(0));
        ComplexOperations b = *this;
        while (e) {
            if (e & 1) r = r * b;
            b = b * b;
            e >>= 1;
        }
        return r;
    }
    __host__ __device__ __forceinline__ ComplexOperations exp(int& e) const { return pow_int(e); }
    __host__ __device__ __forceinline__ ComplexOperations exp(unsigned long long& e) const
    {
        return pow_int(e);
    }

    friend std::ostream& operator<<(std::ostream& out, const ComplexOperations& c)
    {
        out << c.real();
        if (c.imag() >= 0)
            out << " + " << c.imag() << "i";
        else
            out << " - " << -c.imag() << "i";
        return out;
    }
};

// e^z on the host (reference complex.cuh:305-311)
template <typename T> ComplexOperations<T> exp(const ComplexOperations<T>& z)
{
    T e = std::exp(z.real());
    return ComplexOperations<T>(e * std::cos(z.imag()), e * std::sin(z.imag()));
}

}  // namespace complex_arithmetic

template <typename T> using COMPLEX = complex_arithmetic::ComplexOperations<T>;

typedef COMPLEX<Float32> Complex32;
typedef COMPLEX<Float64> Complex64;

static_assert(sizeof(Complex32) == 8 && sizeof(Complex64) == 16, "interleaved complex layout");

#endif  // GPUFFT_AMD_COMPLEX_CUH
