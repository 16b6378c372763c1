// Per-thread error state of the C ABI.
#include "status.hpp"

namespace {
thread_local std::string g_msg;
thread_local int g_hip = 0;
}  // namespace

namespace gfa {

int fail(int status, const std::string& msg)
{
    g_msg = msg;
    return status;
}

int fail_hip(hipError_t e, const char* where)
{
    g_hip = (int)e;
    g_msg = std::string(where) + ": " + hipGetErrorString(e);
    return GPUFFT_ERR_HIP;
}

int ok()
{
    return GPUFFT_OK;
}

}  // namespace gfa

extern "C" {

const char* gpufft_last_error_message(void)
{
    return g_msg.c_str();
}

int gpufft_last_hip_error(void)
{
    return g_hip;
}

}  // extern "C"
