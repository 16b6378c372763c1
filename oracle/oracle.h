/*
 * CPU oracle for the GPU-FFT polynomial-multiplication path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else. No product code path calls into it.
 *
 * It restates, in plain C, the algorithm of the reference library
 * (Alisah-Ozcan/GPU-FFT):
 *   - host table generation and the CPU FFT path of src/lib/fft_cpu.cu,
 *   - the observable semantics of the GPU entry points of src/lib/fft.cu
 *     (the stage order, butterfly and twiddle indexing, fused product and
 *     scaling, whatever launch split the reference used),
 *   - exact integer products (schoolbook, as src/lib/fft_cpu.cu:23-73, and an
 *     exact NTT over the prime 2^64-2^32+1 for large sizes).
 * Complex arrays are interleaved (re, im) pairs, the layout of the
 * reference's COMPLEX<T> (src/include/gpufft/complex.cuh:14-16,30).
 *
 * Every function names the reference lines it follows; how each is pinned
 * (golden vectors, brute force, closed forms) is written in the tests.
 */
#ifndef GPUFFT_ORACLE_H
#define GPUFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fft_cpu.cu:11-21 */
int orc_bitreverse(int index, int n_power);

/* ---- tables: FFT<T>(n) (fft_cpu.cu:93-129,200-226) ----
 * root[i] = exp(i*2*pi*i/(2n)), inv[i] = 1/root[i] by complex division,
 * rev[i] = root[bitrev_log2(n)(i)], inv_rev likewise.  n entries each. */
void orc_fft_tables_f64(int n, double *root, double *inv, double *rev, double *inv_rev);
void orc_fft_tables_f32(int n, float *root, float *inv, float *rev, float *inv_rev);

/* ---- tables: FFNT<T>(N) (fft_cpu.cu:234-299,434-470) ----
 * N/2 entries each: reversed root table, reversed inverse root table,
 * twist table, untwist table. */
void orc_ffnt_tables_f64(int N, double *rev_root, double *inv_rev_root, double *twist, double *untwist);
void orc_ffnt_tables_f32(int N, float *rev_root, float *inv_rev_root, float *twist, float *untwist);

/* ---- special-FFT table (layout derived from fft.cu:782,805,824,908 and the
 * docstring fft.cuh:134-180): T[h+j] = exp(+-2*pi*i*(5^j mod 8h)/(8h)),
 * h = 1,2,...,l/2, j < h; T[0] = 1 (never read).  l entries. */
void orc_special_table_f64(int l, int inverse, double *tab);

/* ---- CPU FFT path ---- */
/* FFT<double>::fft on a length-2n vector (fft_cpu.cu:131-161) */
void orc_cpu_fft_f64(int n, const double *root, double *data);
/* FFT<double>::ifft (fft_cpu.cu:163-198); both parts are scaled by 1/(2n). */
void orc_cpu_ifft_f64(int n, const double *inv_root, double *data);
/* FFNT<double>::MULT (fft_cpu.cu:301-432) with signed 64-bit coefficients. */
void orc_cpu_ffnt_mult_f64(int N, const int64_t *a, const int64_t *b, int64_t *out);
/* Same, for `count` independent products (stride N), on `threads` threads. */
void orc_cpu_ffnt_mult_batch_f64(int N, int count, const int64_t *a, const int64_t *b,
                                 int64_t *out, int threads);
/* The plain-method CPU product of example/test_fft_cpu.cu:65-142 (FFT<double>(n)
 * fft of both zero-padded operands, product, ifft, round): `count` linear
 * products of n-coefficient operands (stride n), 2n outputs each (stride 2n),
 * on `threads` threads. */
void orc_cpu_fft_mult_batch_f64(int n, int count, const int64_t *a, const int64_t *b,
                                int64_t *out, int threads);

/* ---- observable semantics of the GPU entry points ----
 * f64 and f32 variants; for f32 the arithmetic is done in float, as on
 * the device.  `table` is the caller's table in the reference layout. */
#define ORC_DECL(T, S)                                                                   \
    void orc_c2c_forward_##S(T *data, int n_power, int batch, const T *table, int nega); \
    void orc_c2c_inverse_##S(T *data, int n_power, int batch, const T *table, int nega,  \
                             int mult, T scale_re, T scale_im);                          \
    void orc_r2r_forward_##S(const T *real, T *temp, int n_power, int batch,             \
                             const T *table, int nega);                                  \
    void orc_r2r_inverse_##S(T *real, T *temp, int n_power, int batch, const T *table,   \
                             int nega, int mult, T scale_re, T scale_im);                \
    void orc_ffnt_forward_##S(const T *in, T *temp, int n_power, int batch,              \
                              const T *twist, const T *table);                           \
    void orc_ffnt_inverse_##S(T *out, T *temp, int n_power, int batch,                   \
                              const T *untwist, const T *table, int mult, T scale_re,    \
                              T scale_im);                                               \
    void orc_special_forward_##S(T *data, int n_power, int batch, const T *table);       \
    void orc_special_inverse_##S(T *data, int n_power, int batch, const T *table,        \
                                 T scale_re, T scale_im);
ORC_DECL(double, f64)
ORC_DECL(float, f32)
#undef ORC_DECL

/* ---- CKKS encode / decode around the special FFT ----
 * The reference documents these caller steps (src/include/gpufft/fft.cuh:
 * 134-180, the HEAAN-style fftSpecialInv / fftSpecial pair) without code.
 * l = 2^n_power slots, ring N = 2l, scale delta:
 *   encode: u = special_inverse(z) (bit-reversed order, scaled 1/l);
 *           m[i] = llround(Re u[bitrev(i)] * delta), m[i + l] = llround(Im ... * delta)
 *   decode: v[bitrev(i)] = (m[i] / delta, m[i + l] / delta); z = special_forward(v)
 * so decode(m)_j = m(zeta_{2N}^(5^j mod 2N)) / delta. */
void orc_ckks_encode_f64(const double *slots, int n_power, int batch, const double *inv_table,
                         double delta, int64_t *coeffs);
void orc_ckks_decode_f64(const int64_t *coeffs, int n_power, int batch, const double *table,
                         double delta, double *slots);

/* ---- exact integer products ---- */
/* linear product, out has 2n entries (last one 0): fft_cpu.cu:52-73 without % */
void orc_schoolbook_linear(int n, const int64_t *a, const int64_t *b, int64_t *out);
/* negacyclic product mod X^n+1: fft_cpu.cu:23-50 without % */
void orc_schoolbook_negacyclic(int n, const int64_t *a, const int64_t *b, int64_t *out);
/* cyclic product mod X^n-1 */
void orc_schoolbook_cyclic(int n, const int64_t *a, const int64_t *b, int64_t *out);
/* exact products for n a power of two (|result| < 2^62), by NTT mod 2^64-2^32+1 */
void orc_ntt_negacyclic(int n, int count, const int64_t *a, const int64_t *b, int64_t *out);
void orc_ntt_cyclic(int n, int count, const int64_t *a, const int64_t *b, int64_t *out);

/* ---- seeded synthetic inputs (counter-based splitmix64) ----
 * value_i = lo + floor(mix64(seed + (i+1)*0x9E3779B97F4A7C15) * span / 2^64),
 * span = hi - lo (values in [lo, hi)). */
uint64_t orc_mix64(uint64_t z);
void orc_fill_uniform_i64(uint64_t seed, int64_t lo, int64_t hi, int64_t count, int64_t *out);

#ifdef __cplusplus
}
#endif
#endif
