// Host side of the GPU-FFT interface: root-table generators, the CPU FFT
// path and the integer reference products (declared in
// include/gpufft/fft_cpu.cuh; reference src/lib/fft_cpu.cu).
//
// The tables are computed with the same sequence of rounded operations as the
// reference generators (angle i * (2 pi / 2n) rounded in T, cos/sin in T,
// inverse by complex division), so FFT<T>/FFNT<T> tables are the reference's
// tables bit for bit under the same libm.  The CPU transforms are the same
// radix-2 algorithms (same order, same twiddles); FFT<T>::ifft scales both
// parts of the result (the reference's scalar product writes real*s into the
// imaginary part, complex.cuh:153-159).
#include <cmath>
#include <vector>

#include "../../include/gpufft/fft_cpu.cuh"

namespace gpufft {

int bitreverse(int index, int n_power)
{
    int r = 0;
    for (int i = 0; i < n_power; i++) {
        r = (r << 1) | (index & 1);
        index >>= 1;
    }
    return r;
}

namespace {

int log2_int(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

// exp(i * theta) as the reference's Release build evaluates it: its g++ -O3
// host compile turns std::cos(x), std::sin(x) into one glibc sincos(x)
// (CMakeLists.txt:31-33), which can differ from separate calls in the last
// bit; call it explicitly.
inline void sin_cos(double x, double* s, double* c) { ::sincos(x, s, c); }
inline void sin_cos(float x, float* s, float* c) { ::sincosf(x, s, c); }

template <typename T> COMPLEX<T> unit(T theta)
{
    T s, c;
    sin_cos(theta, &s, &c);
    return COMPLEX<T>(c, s);
}

// Cooley-Tukey over a length-2^logm vector, natural in, bit-reversed out;
// group i of a stage uses tab[bitrev_bits(i)].
template <typename T>
void ct_loop(std::vector<COMPLEX<T>>& x, int len, const std::vector<COMPLEX<T>>& tab, int bits)
{
    int t = len;
    for (int m = 1; m < len; m <<= 1) {
        t >>= 1;
        for (int i = 0; i < m; i++) {
            const int j1 = 2 * i * t;
            const COMPLEX<T> S = tab[bitreverse(i, bits)];
            for (int j = j1; j < j1 + t; j++) {
                COMPLEX<T> U = x[j], V = x[j + t] * S;
                x[j] = U + V;
                x[j + t] = U - V;
            }
        }
    }
}

// Gentleman-Sande, bit-reversed in, natural out.
template <typename T>
void gs_loop(std::vector<COMPLEX<T>>& x, int len, const std::vector<COMPLEX<T>>& tab, int bits)
{
    int t = 1;
    for (int m = len; m > 1; m >>= 1) {
        int j1 = 0;
        const int h = m >> 1;
        for (int i = 0; i < h; i++) {
            const COMPLEX<T> S = tab[bitreverse(i, bits)];
            for (int j = j1; j < j1 + t; j++) {
                COMPLEX<T> U = x[j], V = x[j + t];
                x[j] = U + V;
                x[j + t] = (U - V) * S;
            }
            j1 += t << 1;
        }
        t <<= 1;
    }
}

}  // namespace

template <typename T>
std::vector<T> schoolbook_poly_multiplication(std::vector<T> a, std::vector<T> b, T modulus, int size)
{
    std::vector<T> r(2 * size, 0);
    for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) r[i + j] += a[i] * b[j];
    for (int i = 0; i < size; i++) r[i] = (r[i] - r[i + size]) % modulus;
    return r;
}

template <typename T>
std::vector<T> schoolbook_poly_multiplication_without_reduction(std::vector<T> a, std::vector<T> b,
                                                                T modulus, int size)
{
    std::vector<T> r(2 * size, 0);
    for (int i = 0; i < size; i++)
        for (int j = 0; j < size; j++) r[i + j] += a[i] * b[j];
    for (int i = 0; i < 2 * size; i++) r[i] = r[i] % modulus;
    return r;
}

template std::vector<int> schoolbook_poly_multiplication(std::vector<int>, std::vector<int>, int, int);
template std::vector<unsigned long long> schoolbook_poly_multiplication(
    std::vector<unsigned long long>, std::vector<unsigned long long>, unsigned long long, int);
template std::vector<int> schoolbook_poly_multiplication_without_reduction(std::vector<int>,
                                                                           std::vector<int>, int, int);
template std::vector<unsigned long long> schoolbook_poly_multiplication_without_reduction(
    std::vector<unsigned long long>, std::vector<unsigned long long>, unsigned long long, int);

// ---------------------------------------------------------------- FFT<T>
template <typename T> FFT<T>::FFT(int size)
{
    n = size;
    logn = log2_int(n);
    x = COMPLEX<T>(1.0, 0.0);
    max_size = 2 * n;
    // 2 * pi / (2n): exact scaling of the rounded pi
    root = COMPLEX<T>(static_cast<T>(2.0) * static_cast<T>(M_PI) / static_cast<T>(max_size));
    n_inverse = static_cast<Float<T>>(1.0 / max_size);
    GenerateRootTable();
    GenerateInverseRootTable();
}

template <typename T> void FFT<T>::GenerateRootTable()
{
    root_tables.clear();
    for (int i = 0; i < n; i++) root_tables.push_back(unit<T>(static_cast<T>(i) * root.real()));
}

template <typename T> void FFT<T>::GenerateInverseRootTable()
{
    inverse_root_tables.clear();
    const COMPLEX<T> one(1.0);
    for (int i = 0; i < n; i++) inverse_root_tables.push_back(one / root_tables[i]);
}

template <typename T> void FFT<T>::fft(std::vector<COMPLEX<T>>& input)
{
    ct_loop<T>(input, max_size, root_tables, logn);
}

template <typename T> void FFT<T>::ifft(std::vector<COMPLEX<T>>& input)
{
    gs_loop<T>(input, max_size, inverse_root_tables, logn);
    for (int i = 0; i < max_size; i++) input[i] = input[i] * n_inverse;
}

template <typename T> std::vector<COMPLEX<T>> FFT<T>::ReverseRootTable()
{
    std::vector<COMPLEX<T>> r;
    for (int i = 0; i < n; i++) r.push_back(root_tables[bitreverse(i, logn)]);
    return r;
}

template <typename T> std::vector<COMPLEX<T>> FFT<T>::InverseReverseRootTable()
{
    std::vector<COMPLEX<T>> r;
    for (int i = 0; i < n; i++) r.push_back(inverse_root_tables[bitreverse(i, logn)]);
    return r;
}

template class FFT<Float32>;
template class FFT<Float64>;

// ---------------------------------------------------------------- FFNT<T>
template <typename T> FFNT<T>::FFNT(int size)
{
    n = size;
    logn = log2_int(n);
    x = COMPLEX<T>(1.0, 0.0);
    max_size = 2 * n;
    root = COMPLEX<T>(static_cast<T>(2.0 * M_PI / static_cast<T>(max_size)));
    n_inverse = static_cast<T>(1.0 / max_size);
    GenerateRootTable();
    GenerateInverseRootTable();
    GenerateRootTableTwist();
    GenerateInverseRootTableTwist();
}

template <typename T> void FFNT<T>::GenerateRootTable()
{
    root_tables_new.clear();
    // angle step 2 pi / (n/2), computed in double and stored in T
    const T step = static_cast<T>(2.0 * M_PI / static_cast<T>(n >> 1));
    for (int i = 0; i < (n >> 1); i++) root_tables_new.push_back(unit<T>(static_cast<T>(i) * step));
}

template <typename T> void FFNT<T>::GenerateInverseRootTable()
{
    inverse_root_tables_new.clear();
    const COMPLEX<T> one(1.0);
    for (int i = 0; i < (n >> 1); i++) inverse_root_tables_new.push_back(one / root_tables_new[i]);
}

template <typename T> void FFNT<T>::GenerateRootTableTwist()
{
    root_tables_twist.clear();
    const T step = static_cast<T>(2.0 * M_PI / static_cast<T>(2 * n));
    for (int i = 0; i < (n >> 1); i++) root_tables_twist.push_back(unit<T>(static_cast<T>(i) * step));
}

template <typename T> void FFNT<T>::GenerateInverseRootTableTwist()
{
    inverse_root_tables_twist.clear();
    const COMPLEX<T> one(1.0);
    for (int i = 0; i < (n >> 1); i++) inverse_root_tables_twist.push_back(one / root_tables_twist[i]);
}

// fold + twist, two forward FFTs of length n/2, pointwise product, inverse,
// untwist, scale by 1/(n/2), round, unfold (reference fft_cpu.cu:301-432)
template <typename T>
std::vector<unsigned long long> FFNT<T>::MULT(std::vector<unsigned long long>& input1,
                                              std::vector<unsigned long long>& input2)
{
    const int h = n >> 1;
    const int bits = log2_int(h) - 1;
    std::vector<COMPLEX<T>> a(h), b(h);
    for (int i = 0; i < h; i++) {
        a[i] = COMPLEX<T>(static_cast<T>(input1[i]), static_cast<T>(input1[i + h])) * root_tables_twist[i];
        b[i] = COMPLEX<T>(static_cast<T>(input2[i]), static_cast<T>(input2[i + h])) * root_tables_twist[i];
    }
    ct_loop<T>(a, h, root_tables_new, bits);
    ct_loop<T>(b, h, root_tables_new, bits);
    for (int i = 0; i < h; i++) a[i] = a[i] * b[i];
    gs_loop<T>(a, h, inverse_root_tables_new, bits);
    const T ninv = static_cast<T>(1.0 / h);
    std::vector<unsigned long long> r(n);
    for (int i = 0; i < h; i++) {
        COMPLEX<T> z = a[i] * inverse_root_tables_twist[i];
        z = z * COMPLEX<T>(ninv);
        // negative values wrap modulo 2^64, as two's complement
        r[i] = static_cast<unsigned long long>(static_cast<long long>(std::round(z.real())));
        r[i + h] = static_cast<unsigned long long>(static_cast<long long>(std::round(z.imag())));
    }
    return r;
}

template <typename T> std::vector<COMPLEX<T>> FFNT<T>::ReverseRootTable_ffnt()
{
    std::vector<COMPLEX<T>> r;
    const int bits = log2_int(n >> 1) - 1;
    for (int i = 0; i < (n >> 1); i++) r.push_back(root_tables_new[bitreverse(i, bits)]);
    return r;
}

template <typename T> std::vector<COMPLEX<T>> FFNT<T>::InverseReverseRootTable_ffnt()
{
    std::vector<COMPLEX<T>> r;
    const int bits = log2_int(n >> 1) - 1;
    for (int i = 0; i < (n >> 1); i++) r.push_back(inverse_root_tables_new[bitreverse(i, bits)]);
    return r;
}

template <typename T> std::vector<COMPLEX<T>> FFNT<T>::twist_table_ffnt()
{
    return root_tables_twist;
}

template <typename T> std::vector<COMPLEX<T>> FFNT<T>::untwist_table_ffnt()
{
    return inverse_root_tables_twist;
}

template class FFNT<Float32>;
template class FFNT<Float64>;

}  // namespace gpufft
