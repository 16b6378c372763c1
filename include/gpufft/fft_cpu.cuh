// Host-side table generators, CPU FFT path and integer reference products
// of the GPU-FFT interface (reference src/include/gpufft/fft_cpu.cuh:14-105,
// implemented in gpu-fft_amd/csrc/host_fft_cpu.cpp, linked from
// libgpufft_amd.so).  Same class names, public members and table layouts.
#ifndef GPUFFT_AMD_FFT_CPU_CUH
#define GPUFFT_AMD_FFT_CPU_CUH

#include <vector>

#include "complex.cuh"

namespace gpufft {

int bitreverse(int index, int n_power);

enum ReductionPolynomial { X_N_plus, X_N_minus };  // X^n + 1, X^n - 1

// Linear product reduced mod X^size + 1, then each coefficient % modulus
// (reference fft_cpu.cu:23-50).
template <typename T>
std::vector<T> schoolbook_poly_multiplication(std::vector<T> a, std::vector<T> b, T modulus,
                                              int size);
// Linear product, each of the 2*size coefficients % modulus (fft_cpu.cu:52-73).
template <typename T>
std::vector<T> schoolbook_poly_multiplication_without_reduction(std::vector<T> a,
                                                                std::vector<T> b, T modulus,
                                                                int size);

template <typename T> class FFT {
  public:
    int n;
    int logn;
    COMPLEX<T> x;
    int max_size;
    COMPLEX<T> root;
    std::vector<COMPLEX<T>> root_tables;
    std::vector<COMPLEX<T>> inverse_root_tables;
    Float<T> n_inverse;

    explicit FFT(int size);

  private:
    void GenerateRootTable();
    void GenerateInverseRootTable();

  public:
    // in-place transforms of a length-2n vector (the CPU FFT path)
    void fft(std::vector<COMPLEX<T>>& input);
    void ifft(std::vector<COMPLEX<T>>& input);
    std::vector<COMPLEX<T>> ReverseRootTable();
    std::vector<COMPLEX<T>> InverseReverseRootTable();
};

template <typename T> class FFNT {
  public:
    int n;
    int logn;
    COMPLEX<T> x;
    int max_size;
    COMPLEX<T> root;
    std::vector<COMPLEX<T>> root_tables_new;
    std::vector<COMPLEX<T>> inverse_root_tables_new;
    std::vector<COMPLEX<T>> root_tables_twist;
    std::vector<COMPLEX<T>> inverse_root_tables_twist;
    T n_inverse;

    explicit FFNT(int size);

  private:
    void GenerateRootTable();
    void GenerateInverseRootTable();
    void GenerateRootTableTwist();
    void GenerateInverseRootTableTwist();

  public:
    // negacyclic product mod X^n + 1, rounded (the CPU FFNT path)
    std::vector<unsigned long long> MULT(std::vector<unsigned long long>& input1,
                                         std::vector<unsigned long long>& input2);
    std::vector<COMPLEX<T>> ReverseRootTable_ffnt();
    std::vector<COMPLEX<T>> InverseReverseRootTable_ffnt();
    std::vector<COMPLEX<T>> twist_table_ffnt();
    std::vector<COMPLEX<T>> untwist_table_ffnt();
};

}  // namespace gpufft

#endif  // GPUFFT_AMD_FFT_CPU_CUH
