// Error convention of the GPU-FFT interface, for HIP.
//
// Same names as the reference (src/include/gpufft/common_fft.cuh:19-51):
// gpufft::CudaException carries file, line and the runtime error code (here a
// hipError_t) and GPUFFT_CUDA_CHECK throws it.  GPUFFT_HIP_CHECK is the same
// macro under the HIP name.
#ifndef GPUFFT_AMD_COMMON_FFT_CUH
#define GPUFFT_AMD_COMMON_FFT_CUH

#include <hip/hip_runtime.h>

#include <cassert>
#include <exception>
#include <iostream>
#include <string>

namespace gpufft {

class CudaException : public std::exception {
  public:
    CudaException(const std::string& file, int line, hipError_t error)
        : file_(file), line_(line), error_(error),
          msg_("HIP error in " + file_ + " at line " + std::to_string(line_) + ": " +
               hipGetErrorString(error_))
    {
    }
    const char* what() const noexcept override { return msg_.c_str(); }
    hipError_t error() const noexcept { return error_; }

  private:
    std::string file_;
    int line_;
    hipError_t error_;
    std::string msg_;
};

using HipException = CudaException;

inline void customAssert(bool condition, const std::string& message)
{
    if (!condition) {
        std::cerr << "Custom assertion failed: " << message << std::endl;
        assert(condition);
    }
}

}  // namespace gpufft

#define GPUFFT_HIP_CHECK(err)                                            \
    do {                                                                 \
        hipError_t gpufft_error_ = (err);                                \
        if (gpufft_error_ != hipSuccess)                                 \
            throw ::gpufft::CudaException(__FILE__, __LINE__, gpufft_error_); \
    } while (0)
#define GPUFFT_CUDA_CHECK(err) GPUFFT_HIP_CHECK(err)

#endif  // GPUFFT_AMD_COMMON_FFT_CUH
