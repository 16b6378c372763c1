// GPU entry points of the GPU-FFT interface on MI355X (gfx950).
//
// Same names, templates, enums (same numeric values), configuration fields
// and exception behaviour as the reference header
// (src/include/gpufft/fft.cuh:27-273); the only source-visible change is that
// fft_configuration::stream is a hipStream_t.  Each template is an inline
// wrapper over the C ABI of include/gpufft_amd.h, which validates on the host
// and enqueues hand-written HIP kernels on cfg.stream.
//   std::invalid_argument  unsupported n_power / bad argument (before any launch)
//   gpufft::CudaException  launch failure (hipGetLastError after a launch)
#ifndef GPUFFT_AMD_FFT_CUH
#define GPUFFT_AMD_FFT_CUH

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <type_traits>

#include "../gpufft_amd.h"
#include "common_fft.cuh"
#include "complex.cuh"
#include "fft_cpu.cuh"

namespace gpufft {

typedef unsigned location_t;

enum type { FORWARD, INVERSE };

template <typename T> struct fft_configuration {
    int n_power;
    type fft_type;
    ReductionPolynomial reduction_poly;
    bool zero_padding;  // accepted and ignored, as in the reference
    COMPLEX<T> mod_inverse;
    hipStream_t stream;
};

// Launch-shape record of the reference (fft.cuh:53-67).  Kept for source
// compatibility; this library plans its own launches.
struct KernelConfig {
    int griddim_x;
    int griddim_y;
    int blockdim_x;
    int blockdim_y;
    size_t shared_memory;
    int shared_index;
    int logm;
    int k;
    int outer_iteration_count;
    bool not_last_kernel;
};

namespace detail {

template <typename T> inline gpufft_config to_c(const fft_configuration<T>& cfg)
{
    gpufft_config c;
    c.n_power = cfg.n_power;
    c.fft_type = cfg.fft_type == FORWARD ? GPUFFT_FORWARD : GPUFFT_INVERSE;
    c.reduction_poly = cfg.reduction_poly == X_N_plus ? GPUFFT_X_N_PLUS : GPUFFT_X_N_MINUS;
    c.zero_padding = cfg.zero_padding ? 1 : 0;
    c.mod_inverse_re = static_cast<double>(cfg.mod_inverse.real());
    c.mod_inverse_im = static_cast<double>(cfg.mod_inverse.imag());
    c.stream = reinterpret_cast<void*>(cfg.stream);
    return c;
}

inline void raise(int status, const char* file, int line)
{
    if (status == GPUFFT_OK) return;
    if (status == GPUFFT_ERR_HIP)
        throw CudaException(file, line, static_cast<hipError_t>(gpufft_last_hip_error()));
    throw std::invalid_argument(gpufft_last_error_message());
}

template <typename T> constexpr bool is_f32() { return std::is_same<T, float>::value; }

}  // namespace detail

// COMPLEX<T> to COMPLEX<T> (reference fft.cuh:121-124)
template <typename T>
inline void GPU_FFT(COMPLEX<T>* device_inout, COMPLEX<T>* root_of_unity_table,
                    fft_configuration<T> cfg, int batch_size, bool multiplication)
{
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "float/double");
    gpufft_config c = detail::to_c(cfg);
    int st = detail::is_f32<T>()
                 ? gpufft_fft_c2c_f32(device_inout, root_of_unity_table, &c, batch_size, multiplication)
                 : gpufft_fft_c2c_f64(device_inout, root_of_unity_table, &c, batch_size, multiplication);
    detail::raise(st, __FILE__, __LINE__);
}

// FFT: real to COMPLEX<T>, IFFT: COMPLEX<T> to real (reference fft.cuh:126-132)
template <typename T>
inline void GPU_FFT(T* device_fix_inout, COMPLEX<T>* device_Complex64_inout,
                    COMPLEX<T>* root_of_unity_table, fft_configuration<T> cfg, int batch_size,
                    bool multiplication)
{
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "float/double");
    gpufft_config c = detail::to_c(cfg);
    int st = detail::is_f32<T>()
                 ? gpufft_fft_r2r_f32(device_fix_inout, device_Complex64_inout, root_of_unity_table, &c,
                                      batch_size, multiplication)
                 : gpufft_fft_r2r_f64(device_fix_inout, device_Complex64_inout, root_of_unity_table, &c,
                                      batch_size, multiplication);
    detail::raise(st, __FILE__, __LINE__);
}

// Special FFT for CKKS encoding (reference fft.cuh:134-197)
template <typename T>
inline void GPU_Special_FFT(COMPLEX<T>* device_inout, COMPLEX<T>* root_of_unity_table,
                            fft_configuration<T> cfg, int batch_size)
{
    gpufft_config c = detail::to_c(cfg);
    int st = detail::is_f32<T>()
                 ? gpufft_special_fft_f32(device_inout, root_of_unity_table, &c, batch_size)
                 : gpufft_special_fft_f64(device_inout, root_of_unity_table, &c, batch_size);
    detail::raise(st, __FILE__, __LINE__);
}

// FFNT negacyclic transform (reference fft.cuh:199-273); pass the untwist
// table as twist_table for INVERSE.
template <typename T>
inline void GPU_FFNT(T* device_inout, COMPLEX<T>* device_temp, COMPLEX<T>* twist_table,
                     COMPLEX<T>* root_of_unity_table, fft_configuration<T> cfg, int batch_size,
                     bool multiplication)
{
    gpufft_config c = detail::to_c(cfg);
    int st = detail::is_f32<T>()
                 ? gpufft_ffnt_f32(device_inout, device_temp, twist_table, root_of_unity_table, &c,
                                   batch_size, multiplication)
                 : gpufft_ffnt_f64(device_inout, device_temp, twist_table, root_of_unity_table, &c,
                                   batch_size, multiplication);
    detail::raise(st, __FILE__, __LINE__);
}

}  // namespace gpufft

#endif  // GPUFFT_AMD_FFT_CUH
