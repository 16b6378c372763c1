// Drop-in entry points: GPU_FFT (C_C and R_R), GPU_FFNT and GPU_Special_FFT
// of the reference (src/lib/fft.cu:517-725, 928-1007, 1386-1515), as C ABI.
// Each call validates on the host, then enqueues one launch per pass on
// cfg->stream (one launch in total up to 2^12 double / 2^13 float points).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "status.hpp"
#include "transform.hpp"

using namespace gfa;

namespace {

int check_cfg(const gpufft_config* cfg, int lo, int hi, const char* who)
{
    if (!cfg) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": cfg is NULL");
    // same message as the reference (fft.cu:522-525)
    if (cfg->n_power < lo || cfg->n_power > hi)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, "Invalid n_power range!");
    if (cfg->fft_type != GPUFFT_FORWARD && cfg->fft_type != GPUFFT_INVERSE)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": invalid fft_type");
    return GPUFFT_OK;
}

int check_batch(int batch, const char* who)
{
    if (batch < 1) return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": batch must be >= 1");
    return GPUFFT_OK;
}

// Out-of-place transform (source and destination buffers differ): one
// launch where the one-launch path covers the size, else the pass chain.
template <typename T>
int transform_oop(int mode, int n, int batch, int nega, const cplx<T>* tab, const TransformIO<T>& io,
                  hipStream_t st, const char* who)
{
    int s = GPUFFT_OK;
    if (run_oneshot<T>(mode, n, batch, nega, tab, io, st, who, &s)) return s;
    return run_transform<T>(mode, n, batch, nega, tab, io, st, who);
}

template <typename T>
int c2c(void* inout, const void* table, const gpufft_config* cfg, int batch, int mult,
        const char* who)
{
    if (int s = check_cfg(cfg, GPUFFT_MIN_N_POWER, GPUFFT_MAX_N_POWER, who)) return s;
    if (int s = check_batch(batch, who)) return s;
    if (!inout || !table) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const int n = cfg->n_power;
    const long long L = 1LL << n;
    const int nega = cfg->reduction_poly == GPUFFT_X_N_PLUS;
    hipStream_t st = (hipStream_t)cfg->stream;
    cplx<T>* x = (cplx<T>*)inout;
    TransformIO<T> io{};
    io.src = x;
    io.src_stride = L;
    io.dst = x;
    io.dst_stride = L;
    io.work = x;
    if (cfg->fft_type == GPUFFT_FORWARD) {
        io.lk = LD_CPLX;
        io.sk = ST_CPLX;
        return run_transform<T>(MODE_STD_FWD, n, batch, nega, (const cplx<T>*)table, io, st, who);
    }
    io.lk = mult ? LD_MUL : LD_CPLX;
    io.src2 = x + (long long)batch * L;
    io.sk = ST_CPLX;
    io.do_scale = true;
    io.scale = mk<T>((T)cfg->mod_inverse_re, (T)cfg->mod_inverse_im);
    return run_transform<T>(MODE_STD_INV, n, batch, nega, (const cplx<T>*)table, io, st, who);
}

template <typename T>
int r2r(void* real, void* temp, const void* table, const gpufft_config* cfg, int batch, int mult,
        const char* who)
{
    if (int s = check_cfg(cfg, GPUFFT_MIN_N_POWER, GPUFFT_MAX_N_POWER, who)) return s;
    if (int s = check_batch(batch, who)) return s;
    if (!real || !temp || !table) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const int n = cfg->n_power;
    const long long L = 1LL << n;
    const int nega = cfg->reduction_poly == GPUFFT_X_N_PLUS;
    hipStream_t st = (hipStream_t)cfg->stream;
    cplx<T>* t = (cplx<T>*)temp;
    TransformIO<T> io{};
    io.work = t;
    if (cfg->fft_type == GPUFFT_FORWARD) {
        io.lk = LD_REAL;
        io.src = real;
        io.src_stride = L;
        io.sk = ST_CPLX;
        io.dst = t;
        io.dst_stride = L;
        return transform_oop<T>(MODE_STD_FWD, n, batch, nega, (const cplx<T>*)table, io, st, who);
    }
    io.lk = mult ? LD_MUL : LD_CPLX;
    io.src = t;
    io.src2 = t + (long long)batch * L;
    io.src_stride = L;
    io.sk = ST_REAL;
    io.dst = real;
    io.dst_stride = L;
    io.do_scale = true;
    io.scale = mk<T>((T)cfg->mod_inverse_re, (T)cfg->mod_inverse_im);
    return transform_oop<T>(MODE_STD_INV, n, batch, nega, (const cplx<T>*)table, io, st, who);
}

// X^N-1 R_R inverse by half-length packing, for spectra of real data only.
// For k = bitrev(r), C_k and C_{k+M} (M = N/2) sit at temp[2r], temp[2r+1];
// when temp is the spectrum of real data (Hermitian: C_{N-k} = conj(C_k)),
// the N-point inverse D is real and D_2m + i D_2m+1 is the M-point inverse
// of Z_k = (C_k + C_{k+M}) + i w^-k (C_k - C_{k+M}) (LD_RR_PACK), so one
// transform of half the length gives the real output Re(mod_inverse) D_j,
// which equals the reference's Re(mod_inverse * D_j) (fft.cu:499-506) for
// such temp.  For any other temp (or a complex mod_inverse) the result
// differs from the reference's; gpufft_fft_r2r_* is the general form.
// X^N+1 and n < 3 run the general inverse.
template <typename T>
int r2r_inverse_hermitian(void* real, void* temp, const void* table, const gpufft_config* cfg,
                          int batch, int mult, const char* who)
{
    if (int s = check_cfg(cfg, GPUFFT_MIN_N_POWER, GPUFFT_MAX_N_POWER, who)) return s;
    if (cfg->fft_type != GPUFFT_INVERSE)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": fft_type must be INVERSE");
    const int n = cfg->n_power;
    const int nega = cfg->reduction_poly == GPUFFT_X_N_PLUS;
    if (nega || n < 3) return r2r<T>(real, temp, table, cfg, batch, mult, who);
    if (int s = check_batch(batch, who)) return s;
    if (!real || !temp || !table) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const long long L = 1LL << n;
    const long long M = L / 2;
    cplx<T>* t = (cplx<T>*)temp;
    TransformIO<T> io{};
    io.lk = LD_RR_PACK;
    io.src = t;
    io.src2 = mult ? t + (long long)batch * L : nullptr;
    io.ltab = (const cplx<T>*)table;
    io.src_stride = L;
    io.sk = ST_CPLX;
    io.dst = real;
    io.dst_stride = M;
    io.work = (cplx<T>*)real;  // the real buffer holds the intermediate passes
    io.do_scale = true;
    io.scale = mk<T>((T)cfg->mod_inverse_re, (T)0);
    return run_transform<T>(MODE_STD_INV, n - 1, batch, 0, (const cplx<T>*)table, io,
                            (hipStream_t)cfg->stream, who);
}

template <typename T>
int ffnt(void* inout, void* temp, const void* twist, const void* table, const gpufft_config* cfg,
         int batch, int mult, const char* who)
{
    if (int s = check_cfg(cfg, GPUFFT_MIN_N_POWER_FFNT, GPUFFT_MAX_N_POWER, who)) return s;
    if (int s = check_batch(batch, who)) return s;
    if (!inout || !temp || !twist || !table)
        return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const int n = cfg->n_power - 1;  // length of the complex transform
    const long long L = 1LL << n;
    hipStream_t st = (hipStream_t)cfg->stream;
    cplx<T>* t = (cplx<T>*)temp;
    TransformIO<T> io{};
    io.work = t;
    if (cfg->fft_type == GPUFFT_FORWARD) {
        io.lk = LD_FOLD;
        io.src = inout;
        io.ltab = (const cplx<T>*)twist;
        io.src_stride = 2 * L;
        io.sk = ST_CPLX;
        io.dst = t;
        io.dst_stride = L;
        return transform_oop<T>(MODE_STD_FWD, n, batch, 0, (const cplx<T>*)table, io, st, who);
    }
    io.lk = mult ? LD_MUL : LD_CPLX;
    io.src = t;
    io.src2 = t + (long long)batch * L;
    io.src_stride = L;
    io.sk = ST_UNFOLD;
    io.dst = inout;
    io.stab = (const cplx<T>*)twist;
    io.dst_stride = 2 * L;
    io.do_scale = true;
    io.scale = mk<T>((T)cfg->mod_inverse_re, (T)cfg->mod_inverse_im);
    return transform_oop<T>(MODE_STD_INV, n, batch, 0, (const cplx<T>*)table, io, st, who);
}

template <typename T>
int special(void* inout, const void* table, const gpufft_config* cfg, int batch, const char* who)
{
    if (int s = check_cfg(cfg, GPUFFT_MIN_N_POWER, GPUFFT_MAX_N_POWER, who)) return s;
    if (int s = check_batch(batch, who)) return s;
    if (!inout || !table) return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const int n = cfg->n_power;
    const long long L = 1LL << n;
    hipStream_t st = (hipStream_t)cfg->stream;
    cplx<T>* x = (cplx<T>*)inout;
    TransformIO<T> io{};
    io.lk = LD_CPLX;
    io.src = x;
    io.src_stride = L;
    io.sk = ST_CPLX;
    io.dst = x;
    io.dst_stride = L;
    io.work = x;
    if (cfg->fft_type == GPUFFT_FORWARD)
        return run_transform<T>(MODE_SPE_FWD, n, batch, 0, (const cplx<T>*)table, io, st, who);
    io.do_scale = true;
    io.scale = mk<T>((T)cfg->mod_inverse_re, (T)cfg->mod_inverse_im);
    return run_transform<T>(MODE_SPE_INV, n, batch, 0, (const cplx<T>*)table, io, st, who);
}

// Special FFT of double-complex slots in FP32 arithmetic: x = hi + lo with
// hi = float(x), lo = float(x - hi); T(x) = T(hi) + T(lo) by linearity, each
// transform in float, the two results added in double.  work: float complex
// [2][batch][l] (T(hi), then the intermediate passes of T(lo)).
int special_split(void* inout, const void* table, const gpufft_config* cfg, int batch, void* work,
                  const char* who)
{
    if (int s = check_cfg(cfg, GPUFFT_MIN_N_POWER, GPUFFT_MAX_N_POWER, who)) return s;
    if (int s = check_batch(batch, who)) return s;
    if (!inout || !table || !work)
        return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const int n = cfg->n_power;
    const long long L = 1LL << n;
    hipStream_t st = (hipStream_t)cfg->stream;
    const int mode = cfg->fft_type == GPUFFT_FORWARD ? MODE_SPE_FWD : MODE_SPE_INV;
    cplx<float>* whi = (cplx<float>*)work;
    cplx<float>* wlo = whi + (long long)batch * L;
    TransformIO<float> io{};
    io.src = inout;
    io.src_stride = L;
    io.do_scale = cfg->fft_type == GPUFFT_INVERSE;
    io.scale = mk<float>((float)cfg->mod_inverse_re, (float)cfg->mod_inverse_im);
    // T(hi) -> whi
    io.lk = LD_SPLIT_HI;
    io.sk = ST_CPLX;
    io.dst = whi;
    io.dst_stride = L;
    io.work = whi;
    if (int s = run_transform<float>(mode, n, batch, 0, (const cplx<float>*)table, io, st, who))
        return s;
    // T(lo), intermediates in wlo, last pass: inout = whi + T(lo) in double
    io.lk = LD_SPLIT_LO;
    io.sk = ST_SUM_F64;
    io.dst = inout;
    io.stab = whi;
    io.work = wlo;
    return run_transform<float>(mode, n, batch, 0, (const cplx<float>*)table, io, st, who);
}

// CKKS slot <-> coefficient steps around the special FFT.  l = 2^n slots,
// ring N = 2l: coefficient i pairs with slot-transform entry bitrev(i) (the
// special inverse leaves its output bit-reversed, the special forward wants
// its input bit-reversed; fft.cu:727-926), real parts give m[0..l), imaginary
// parts m[l..2l).
__global__ void ckks_round_kernel(const cplx<double>* u, long long* m, int n, long long total,
                                  double delta)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const long long l = 1LL << n;
    const long long b = t >> n;
    const int i = (int)(t & (l - 1));
    const int br = (int)(__brev((unsigned)i) >> (32 - n));
    const cplx<double> z = u[b * l + br];
    m[2 * b * l + i] = llround(z.x * delta);
    m[2 * b * l + i + l] = llround(z.y * delta);
}

__global__ void ckks_unround_kernel(const long long* m, cplx<double>* v, int n, long long total,
                                    double delta)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const long long l = 1LL << n;
    const long long b = t >> n;
    const int i = (int)(t & (l - 1));
    const int br = (int)(__brev((unsigned)i) >> (32 - n));
    v[b * l + br] = mk<double>((double)m[2 * b * l + i] / delta, (double)m[2 * b * l + i + l] / delta);
}

int check_ckks(int n_power, double delta, int batch, const char* who)
{
    if (n_power < GPUFFT_MIN_N_POWER || n_power > GPUFFT_MAX_N_POWER)
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, "Invalid n_power range!");
    if (!(delta > 0.0) || !std::isfinite(delta))
        return fail(GPUFFT_ERR_INVALID_ARGUMENT, std::string(who) + ": scale must be finite and > 0");
    return check_batch(batch, who);
}

int ckks_encode(const void* slots, int64_t* coeffs, const void* inv_table, int n_power,
                double delta, int batch, void* work, void* stream, const char* who)
{
    if (int s = check_ckks(n_power, delta, batch, who)) return s;
    if (!slots || !coeffs || !inv_table || !work)
        return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const long long L = 1LL << n_power;
    hipStream_t st = (hipStream_t)stream;
    TransformIO<double> io{};
    io.lk = LD_CPLX;
    io.src = slots;
    io.src_stride = L;
    io.sk = ST_CPLX;
    io.dst = work;
    io.dst_stride = L;
    io.work = (cplx<double>*)work;
    io.do_scale = true;
    io.scale = mk<double>(1.0 / (double)L, 0.0);
    if (int s = run_transform<double>(MODE_SPE_INV, n_power, batch, 0, (const cplx<double>*)inv_table,
                                      io, st, who))
        return s;
    const long long total = (long long)batch << n_power;
    hipLaunchKernelGGL(ckks_round_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       (const cplx<double>*)work, (long long*)coeffs, n_power, total, delta);
    GFA_CHECK_LAUNCH(who);
    return ok();
}

int ckks_decode(const int64_t* coeffs, void* slots, const void* table, int n_power, double delta,
                int batch, void* stream, const char* who)
{
    if (int s = check_ckks(n_power, delta, batch, who)) return s;
    if (!slots || !coeffs || !table)
        return fail(GPUFFT_ERR_NULL_POINTER, std::string(who) + ": NULL buffer");
    const long long L = 1LL << n_power;
    hipStream_t st = (hipStream_t)stream;
    const long long total = (long long)batch << n_power;
    hipLaunchKernelGGL(ckks_unround_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       (const long long*)coeffs, (cplx<double>*)slots, n_power, total, delta);
    GFA_CHECK_LAUNCH(who);
    cplx<double>* x = (cplx<double>*)slots;
    TransformIO<double> io{};
    io.lk = LD_CPLX;
    io.src = x;
    io.src_stride = L;
    io.sk = ST_CPLX;
    io.dst = x;
    io.dst_stride = L;
    io.work = x;
    return run_transform<double>(MODE_SPE_FWD, n_power, batch, 0, (const cplx<double>*)table, io, st,
                                 who);
}

}  // namespace

extern "C" {

int gpufft_ckks_encode_f64(const void* slots, int64_t* coeffs, const void* inv_table, int n_power,
                           double scale, int batch, void* work, void* stream)
{
    return ckks_encode(slots, coeffs, inv_table, n_power, scale, batch, work, stream,
                       "gpufft_ckks_encode_f64");
}

int gpufft_ckks_decode_f64(const int64_t* coeffs, void* slots, const void* table, int n_power,
                           double scale, int batch, void* stream)
{
    return ckks_decode(coeffs, slots, table, n_power, scale, batch, stream, "gpufft_ckks_decode_f64");
}

int gpufft_special_fft_split_f32(void* inout, const void* root_table, const gpufft_config* cfg,
                                 int batch, void* work)
{
    return special_split(inout, root_table, cfg, batch, work, "gpufft_special_fft_split_f32");
}

int gpufft_fft_c2c_f32(void* inout, const void* t, const gpufft_config* cfg, int batch, int mult)
{
    return c2c<float>(inout, t, cfg, batch, mult, "gpufft_fft_c2c_f32");
}
int gpufft_fft_c2c_f64(void* inout, const void* t, const gpufft_config* cfg, int batch, int mult)
{
    return c2c<double>(inout, t, cfg, batch, mult, "gpufft_fft_c2c_f64");
}
int gpufft_fft_r2r_f32(void* real, void* temp, const void* t, const gpufft_config* cfg, int batch,
                       int mult)
{
    return r2r<float>(real, temp, t, cfg, batch, mult, "gpufft_fft_r2r_f32");
}
int gpufft_fft_r2r_f64(void* real, void* temp, const void* t, const gpufft_config* cfg, int batch,
                       int mult)
{
    return r2r<double>(real, temp, t, cfg, batch, mult, "gpufft_fft_r2r_f64");
}
int gpufft_fft_r2r_inverse_hermitian_f32(void* real, void* temp, const void* t,
                                         const gpufft_config* cfg, int batch, int mult)
{
    return r2r_inverse_hermitian<float>(real, temp, t, cfg, batch, mult,
                                        "gpufft_fft_r2r_inverse_hermitian_f32");
}
int gpufft_fft_r2r_inverse_hermitian_f64(void* real, void* temp, const void* t,
                                         const gpufft_config* cfg, int batch, int mult)
{
    return r2r_inverse_hermitian<double>(real, temp, t, cfg, batch, mult,
                                         "gpufft_fft_r2r_inverse_hermitian_f64");
}
int gpufft_ffnt_f32(void* inout, void* temp, const void* tw, const void* t, const gpufft_config* cfg,
                    int batch, int mult)
{
    return ffnt<float>(inout, temp, tw, t, cfg, batch, mult, "gpufft_ffnt_f32");
}
int gpufft_ffnt_f64(void* inout, void* temp, const void* tw, const void* t, const gpufft_config* cfg,
                    int batch, int mult)
{
    return ffnt<double>(inout, temp, tw, t, cfg, batch, mult, "gpufft_ffnt_f64");
}
int gpufft_special_fft_f32(void* inout, const void* t, const gpufft_config* cfg, int batch)
{
    return special<float>(inout, t, cfg, batch, "gpufft_special_fft_f32");
}
int gpufft_special_fft_f64(void* inout, const void* t, const gpufft_config* cfg, int batch)
{
    return special<double>(inout, t, cfg, batch, "gpufft_special_fft_f64");
}

}  // extern "C"
