/*
 * gpufft_amd — C ABI of the MI355X (gfx950) GPU-FFT polynomial-multiplication
 * library.
 *
 * Plain pointers, sizes and a POD configuration; no C++ types, no exceptions.
 * Every entry point returns a status (GPUFFT_OK on success).  All work is
 * enqueued asynchronously on cfg->stream (NULL = the default stream); nothing
 * synchronises.  Device buffers and root tables are caller-owned device
 * memory in the layouts of the reference library (Alisah-Ozcan/GPU-FFT):
 * complex values are interleaved (re, im) pairs of float (8 bytes) or
 * double (16 bytes), as COMPLEX<T> in src/include/gpufft/complex.cuh:14-16.
 *
 * The C++ templates of include/gpufft/fft.cuh (the reference's public
 * interface) are inline wrappers over these functions.
 */
#ifndef GPUFFT_AMD_H
#define GPUFFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define GPUFFT_OK 0
#define GPUFFT_ERR_INVALID_ARGUMENT 1 /* maps to std::invalid_argument */
#define GPUFFT_ERR_HIP 2              /* a HIP runtime error; see gpufft_last_hip_error() */
#define GPUFFT_ERR_NULL_POINTER 3     /* maps to std::invalid_argument */
#define GPUFFT_ERR_UNSUPPORTED 4      /* maps to std::invalid_argument */

/* ---- configuration ----
 * Mirrors fft_configuration<T> (reference src/include/gpufft/fft.cuh:43-51)
 * with the enum values of fft.cuh:37-41 and fft_cpu.cuh:16-20. */
#define GPUFFT_FORWARD 0
#define GPUFFT_INVERSE 1
#define GPUFFT_X_N_PLUS 0  /* X^N + 1 (negacyclic) */
#define GPUFFT_X_N_MINUS 1 /* X^N - 1 (cyclic) */

typedef struct gpufft_config {
    int n_power;        /* log2 transform length (FFNT: log2 ring size) */
    int fft_type;       /* GPUFFT_FORWARD / GPUFFT_INVERSE */
    int reduction_poly; /* GPUFFT_X_N_PLUS / GPUFFT_X_N_MINUS */
    int zero_padding;   /* accepted and ignored, as in the reference */
    double mod_inverse_re, mod_inverse_im; /* complex scale of INVERSE */
    void* stream;       /* hipStream_t */
} gpufft_config;

/* ---- supported ranges of n_power (the reference accepts 12..24 for
 * GPU_FFT/GPU_FFNT and 11..24 for GPU_Special_FFT; this library extends the
 * lower end and keeps the upper one) ---- */
#define GPUFFT_MIN_N_POWER 1
#define GPUFFT_MIN_N_POWER_FFNT 2
#define GPUFFT_MAX_N_POWER 24

/* ---- drop-in transforms ---- */

/* Complex-to-complex transform, in place on inout[batch][2^n_power].
 * FORWARD: Cooley-Tukey, natural order in, bit-reversed order out, twiddle
 *   table in the FFT<T>::ReverseRootTable() layout (X^N-1: 2^(n_power-1)
 *   entries; X^N+1: 2^n_power entries).
 * INVERSE: Gentleman-Sande, bit-reversed in, natural out, times mod_inverse.
 *   With multiplication != 0 the input is inout[b] * inout[b + batch] and
 *   the result goes to inout[b], b < batch.
 * Replaces gpufft::GPU_FFT(COMPLEX<T>*, COMPLEX<T>*, fft_configuration<T>,
 * int, bool), reference src/include/gpufft/fft.cuh:121-124, src/lib/fft.cu:517-598. */
int gpufft_fft_c2c_f32(void* inout, const void* root_table, const gpufft_config* cfg,
                       int batch, int multiplication);
int gpufft_fft_c2c_f64(void* inout, const void* root_table, const gpufft_config* cfg,
                       int batch, int multiplication);

/* Real transform through a complex buffer.
 * FORWARD: real[batch][2^n] in, full bit-reversed complex spectrum out in
 *   temp[batch][2^n]; real is not modified.
 * INVERSE: temp[batch][2^n] (times temp[b + batch] with multiplication) in,
 *   real[b][i] = Re(x[i] * mod_inverse) out; temp is used as scratch.
 * Replaces gpufft::GPU_FFT(T*, COMPLEX<T>*, COMPLEX<T>*, fft_configuration<T>,
 * int, bool), reference fft.cuh:126-132, fft.cu:600-725. */
int gpufft_fft_r2r_f32(void* real_inout, void* temp, const void* root_table,
                       const gpufft_config* cfg, int batch, int multiplication);
int gpufft_fft_r2r_f64(void* real_inout, void* temp, const void* root_table,
                       const gpufft_config* cfg, int batch, int multiplication);

/* The R_R INVERSE for the common case where temp holds spectra of real data
 * (as the R_R FORWARD writes them; Hermitian: C[N-k] = conj(C[k])), so the
 * product of two such spectra is one too.  For X^N-1 (n_power >= 3) it runs
 * a half-length transform of the packed spectrum, half the flops and less
 * HBM traffic than the general inverse, and writes
 * real[b][i] = Re(mod_inverse) * D[i]; under the precondition that equals
 * the general result Re(mod_inverse * D[i]) (D is then real).  For any other
 * temp, or a mod_inverse with an imaginary part, the result differs from the
 * reference's: use gpufft_fft_r2r_* then.  The real buffer also holds the
 * intermediate passes; temp is read only.  X^N+1 and n_power < 3 run the
 * general inverse.  fft_type must be GPUFFT_INVERSE.
 * No counterpart in the reference (whose GPU_FFT R_R inverse, fft.cu:600-725,
 * always runs the full-length complex inverse); for callers that feed the
 * R_R inverse only with R_R forward spectra. */
int gpufft_fft_r2r_inverse_hermitian_f32(void* real_inout, void* temp, const void* root_table,
                                         const gpufft_config* cfg, int batch, int multiplication);
int gpufft_fft_r2r_inverse_hermitian_f64(void* real_inout, void* temp, const void* root_table,
                                         const gpufft_config* cfg, int batch, int multiplication);

/* FFNT negacyclic transform (ring size N = 2^n_power, length-N/2 complex FFT).
 * FORWARD: real inout[batch][N] -> fold z_j = f_j + i f_{j+N/2}, times
 *   twist_table[j], forward FFT -> temp[batch][N/2] (bit-reversed).
 * INVERSE: temp (times temp[b + batch] with multiplication) -> inverse FFT,
 *   times mod_inverse, times twist_table[j] (pass the untwist table here),
 *   unfold -> inout[b][j] = Re, inout[b][j + N/2] = Im (not rounded).
 * Replaces gpufft::GPU_FFNT, reference fft.cuh:269-273, fft.cu:1386-1515. */
int gpufft_ffnt_f32(void* inout, void* temp, const void* twist_table, const void* root_table,
                    const gpufft_config* cfg, int batch, int multiplication);
int gpufft_ffnt_f64(void* inout, void* temp, const void* twist_table, const void* root_table,
                    const gpufft_config* cfg, int batch, int multiplication);

/* Special (CKKS canonical-embedding) FFT, in place on inout[batch][2^n_power].
 * FORWARD: bit-reversed in, natural out, w_j = sum_k z_k zeta^{(5^j mod 4l) k}.
 * INVERSE: natural in, bit-reversed out, times mod_inverse.
 * Table layout: T[h + j] = exp(+-2 pi i (5^j mod 8h) / 8h), h = 1..l/2, j < h.
 * Replaces gpufft::GPU_Special_FFT, reference fft.cuh:194-197, fft.cu:928-1007. */
int gpufft_special_fft_f32(void* inout, const void* root_table, const gpufft_config* cfg,
                           int batch);
int gpufft_special_fft_f64(void* inout, const void* root_table, const gpufft_config* cfg,
                           int batch);

/* The same special FFT on double-complex slots, evaluated in FP32: each slot
 * is split x = hi + lo (hi = float(x), lo = float(x - hi)), both halves are
 * transformed with the float table (same layout) and the two results are
 * added in double, so the 2^30-scaled CKKS slots are not truncated to float
 * on input.  work: device buffer of 2 * batch * 2^n_power float complex.
 * No counterpart in the reference (which runs GPU_Special_FFT<float> on float
 * data); the FP64 entry point above is the drop-in. */
int gpufft_special_fft_split_f32(void* inout, const void* root_table, const gpufft_config* cfg,
                                 int batch, void* work);

/* ---- root-of-unity tables generated on the GPU (FP64 arithmetic) ----
 * Fills dst[0 .. 2^n_power) on `stream` in the reference layouts:
 *   GPUFFT_TABLE_FFT         FFT<T>(2^n_power).ReverseRootTable()
 *                            (reference src/lib/fft_cpu.cu:93-118,200-211);
 *                            inverse != 0: InverseReverseRootTable()
 *   GPUFFT_TABLE_FFNT_ROOT   FFNT<T>(2^(n_power+1)).ReverseRootTable_ffnt()
 *                            (fft_cpu.cu:251-263,434-446); inverse: _Inverse_
 *   GPUFFT_TABLE_FFNT_TWIST  FFNT<T>(2^(n_power+1)).twist_table_ffnt()
 *                            (fft_cpu.cu:276-288,462-465); inverse: untwist
 *   GPUFFT_TABLE_SPECIAL     special-FFT table of 2^n_power slots,
 *                            T[h+j] = exp(2 pi i (5^j mod 8h)/8h) (layout
 *                            indexed at fft.cu:782,908); inverse: conjugate
 *   GPUFFT_TABLE_STAGE_ROOTS [2^s + t] = exp(2 pi i bitrev_s(t) / 2^(s+1)),
 *                            the in-step root constants of the fused kernels
 * The reference builds these on the host with libm (same layout and values
 * to within an ulp); this entry point has no counterpart there. */
#define GPUFFT_TABLE_FFT 0
#define GPUFFT_TABLE_FFNT_ROOT 1
#define GPUFFT_TABLE_FFNT_TWIST 2
#define GPUFFT_TABLE_SPECIAL 3
#define GPUFFT_TABLE_STAGE_ROOTS 4
int gpufft_generate_table_f64(void* dst, int kind, int n_power, int inverse, void* stream);
int gpufft_generate_table_f32(void* dst, int kind, int n_power, int inverse, void* stream);

/* ---- fused batched negacyclic products (no counterpart in the reference,
 * which multiplies with GPU_FFNT FORWARD + GPU_FFNT INVERSE(multiplication)
 * and rounds on the caller side, example/test_ffnt_gpu_R_R.cu:137-172) ----
 * A plan fixes the ring size N = 2^ring_power, the precision and the largest
 * batch of one launch; it owns its root tables (generated on the GPU) and
 * its scratch buffer (allocated once at creation).  Execution allocates
 * nothing; calls on one plan must not overlap in time (they share scratch).
 *   GPUFFT_PREC_F64       a, b, out: double [batch][N]; out is the unrounded
 *                         product mod X^N+1 (same values as the two-call
 *                         GPU_FFNT flow); ring_power 7..17.
 *   GPUFFT_PREC_F32_SPLIT a, b: int32 [batch][N]; out: int64 [batch][N], the
 *                         product mod X^N+1.  a is split into a_limbs
 *                         balanced base-2^a_limb_bits digits (the top one
 *                         keeps the rest), b into b_limbs digits (the top
 *                         digit must start at bit 30 or below:
 *                         (limbs - 1) * limb_bits <= 30); each limb
 *                         product is an FP32 FFNT product rounded to an
 *                         integer, and the limbs are recombined in int64.
 *                         err_stats (optional, double [batch][2]) receives
 *                         per product the largest distance of a limb output
 *                         to its rounded value and the sum of squares of
 *                         those distances: a largest distance below 0.5
 *                         means every rounding was to the exact integer.
 *                         ring_power 8..18. */
#define GPUFFT_PREC_F64 0
#define GPUFFT_PREC_F32_SPLIT 1
typedef struct gpufft_polymul_plan_s* gpufft_polymul_plan;
int gpufft_polymul_plan_create(gpufft_polymul_plan* out, int ring_power, int precision,
                               int max_batch, int a_limb_bits, int a_limbs, int b_limb_bits,
                               int b_limbs, void* stream);
int gpufft_polymul_plan_destroy(gpufft_polymul_plan plan);
size_t gpufft_polymul_plan_scratch_bytes(gpufft_polymul_plan plan);
/* Radix of the LDS steps of the product kernel, 2^radix_bits: 3 or 4 for
 * GPUFFT_PREC_F64, 3..5 for GPUFFT_PREC_F32_SPLIT (default 4).  Results do
 * not depend on it beyond rounding; the autotuner (gpufft_amd.autotune)
 * sets it per ring size and device. */
int gpufft_polymul_plan_set_radix_bits(gpufft_polymul_plan plan, int radix_bits);
int gpufft_polymul_plan_get_radix_bits(gpufft_polymul_plan plan);
/* Points per thread of the product kernel, 2^log2_points: 4 for
 * GPUFFT_PREC_F64 (512 threads per 2^13-point sub-problem); 5 (512
 * threads) or 4 (1024 threads) per 2^14-point sub-problem for
 * GPUFFT_PREC_F32_SPLIT (default 4 for ring_power 15 and 16, 5 otherwise).
 * Results do not depend on it beyond rounding.
 * Fails with GPUFFT_ERR_UNSUPPORTED where no kernel exists for the plan's
 * size and radix (16 points per thread allow radix bits 3 and 4 only). */
int gpufft_polymul_plan_set_thread_points(gpufft_polymul_plan plan, int log2_points);
int gpufft_polymul_plan_get_thread_points(gpufft_polymul_plan plan);
/* GPUFFT_PREC_F64 plans whose product spans several workgroups (ring_power
 * >= 15): in_kernel = 1 runs the last inverse stages inside the product
 * kernel, in the workgroup that finishes a product last (one launch); 0
 * (default) runs them as a second, streaming launch.  On MI355X the second
 * launch is faster: the in-kernel form's device-scope release/acquire write
 * back and invalidate L2, which the product kernel's input reads rely on.
 * Other plans ignore the setting (get returns 0).  Same results either way. */
int gpufft_polymul_plan_set_combine(gpufft_polymul_plan plan, int in_kernel);
int gpufft_polymul_plan_get_combine(gpufft_polymul_plan plan);
/* GPUFFT_PREC_F32_SPLIT plans: split a - 2^center_log2 instead of a (0, the
 * default: a as is), and add the exact correction 2^center_log2 * (J * b),
 * J the all-ones polynomial, as one more limb product (no operand loads).
 * For a in [0, 2^20) with 4-bit limbs, center 19 makes 5 limbs enough (every
 * digit in [-8, 8]) where a as is needs 6: one limb product less per call.
 * Results are the same exact products; only the split changes. */
int gpufft_polymul_plan_set_a_center(gpufft_polymul_plan plan, int center_log2);
int gpufft_polymul_plan_get_a_center(gpufft_polymul_plan plan);
int gpufft_polymul_f64(gpufft_polymul_plan plan, const double* a, const double* b, double* out,
                       int batch, void* stream);
int gpufft_polymul_i32(gpufft_polymul_plan plan, const int32_t* a, const int32_t* b, int64_t* out,
                       int batch, double* err_stats, void* stream);
/* The same products as two separately enqueued phases, for callers that
 * schedule or time them apart: GPUFFT_PHASE_TRANSFORM runs the forward
 * transforms, products and inverse sub-transforms into the plan's scratch,
 * GPUFFT_PHASE_COMBINE the last inverse stages and the output (for plans
 * whose whole product fits one workgroup, and for FP64 plans with the
 * in-kernel combine, the first phase writes the output and the second does
 * nothing).  A combine must follow the transform of the same inputs on the
 * same plan, with batch <= max_batch. */
#define GPUFFT_PHASE_TRANSFORM 1
#define GPUFFT_PHASE_COMBINE 2
#define GPUFFT_PHASE_ALL 3
int gpufft_polymul_f64_phases(gpufft_polymul_plan plan, const double* a, const double* b,
                              double* out, int batch, int phases, void* stream);
int gpufft_polymul_i32_phases(gpufft_polymul_plan plan, const int32_t* a, const int32_t* b,
                              int64_t* out, int batch, double* err_stats, int phases,
                              void* stream);

/* The FP64 product of gpufft_polymul_f64 without a plan and without any
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The caller passes the inputs, the
 * output and the four FFNT tables of ring N = 2^ring_power (N/2 entries
 * each, the layouts GPU_FFNT takes: twist, untwist, root and inverse root,
 * e.g. from gpufft_generate_table_f64); the library uses a static device
 * array of 64 MiB (per device, held until the process ends) for the stage
 * constants and the sub-problem results, so the batch runs in launches of
 * at most 63 MiB / (8 N) products for ring_power >= 15.  ring_power 7..17
 * (else GPUFFT_ERR_INVALID_ARGUMENT); a, b, out: double [batch][N]; out is
 * the unrounded product mod X^N+1, equal to gpufft_polymul_f64's.  Calls
 * that overlap in time on one device share the array and race; calls on one
 * stream are ordered.  Only kernels are enqueued (no memset, no
 * allocation), so a call can be captured into a HIP graph. */
int gpufft_polymul_f64_noalloc(const double* a, const double* b, double* out, int ring_power,
                               int batch, const void* twist, const void* untwist,
                               const void* root_table, const void* inv_root_table, void* stream);

/* ---- CKKS encode / decode around the special FFT (no counterpart in the
 * reference, which documents these caller steps around GPU_Special_FFT in
 * src/include/gpufft/fft.cuh:134-180 without implementing them) ----
 * l = 2^n_power slots (n_power 1..24), ring N = 2l, scale = Delta > 0.
 *   encode: slots (complex double [batch][l], read only) -> coeffs (int64
 *           [batch][2l]): u = special inverse FFT of the slots with
 *           inv_table (GPUFFT_TABLE_SPECIAL, inverse) and mod_inverse 1/l;
 *           coeffs[i] = llround(Re u[bitrev(i)] * scale),
 *           coeffs[i + l] = llround(Im u[bitrev(i)] * scale).
 *           work: complex double [batch][l].  |u * scale| must stay below
 *           2^63 (not checked).
 *   decode: coeffs -> slots: slots[bitrev(i)] = (coeffs[i] / scale,
 *           coeffs[i + l] / scale), then the special forward FFT in place
 *           with table (GPUFFT_TABLE_SPECIAL, forward), so that
 *           slots[j] = m(zeta_2N^(5^j mod 2N)) / scale. */
int gpufft_ckks_encode_f64(const void* slots, int64_t* coeffs, const void* inv_table, int n_power,
                           double scale, int batch, void* work, void* stream);
int gpufft_ckks_decode_f64(const int64_t* coeffs, void* slots, const void* table, int n_power,
                           double scale, int batch, void* stream);

/* ---- batched products of uneven ring sizes, one persistent launch ----
 * (no counterpart in the reference, which needs one GPU_FFT/GPU_FFNT call
 * pair per ring size).  Each product of the list is
 *   GPUFFT_PRODUCT_NEGACYCLIC  out[0..N)  = a * b mod X^N + 1
 *   GPUFFT_PRODUCT_LINEAR      out[0..2N) = a * b (2N-1 coefficients, then 0)
 * with N = 2^ring_power coefficients per operand (NEGACYCLIC: ring_power
 * 6..19, LINEAR: 5..18) at the given element offsets of the int32 inputs a,
 * b (multiples of 4: operands are read 16 bytes at a time) and the int64
 * output.  A LINEAR product is computed as the negacyclic
 * product of ring 2N of the zero-padded operands.  Arithmetic is FP32 with
 * the operands split into balanced limbs (as gpufft_polymul_i32; the top limb
 * must start at bit 30 or below: (limbs - 1) * limb_bits <= 30) and exact
 * int64 recombination.  Limb products of equal weight are summed before their
 * inverse transform.  The plan (created once; it allocates its tables,
 * work list and scratch) runs every product in one persistent kernel, whose
 * workgroups take sub-problems from 8 per-XCD work queues.  The top inverse
 * stages of products split over several workgroups run in a second, streaming
 * kernel (default) or in the persistent kernel itself, by the last workgroup
 * of the product to finish (gpufft_batched_plan_set_combine(plan, 1)).
 * gpufft_batched_execute enqueues the launches on `stream`;
 * gpufft_batched_graph_launch captures them into a HIP graph on first use
 * per buffer set and replays the graph; it needs a stream other than the
 * default stream, which cannot be captured.  err_stats
 * (optional, double [count]) receives each product's largest distance of a
 * limb output to its rounding (< 0.5: every limb was rounded exactly).
 * Calls on one plan must not overlap. */
#define GPUFFT_PRODUCT_NEGACYCLIC 0
#define GPUFFT_PRODUCT_LINEAR 1
typedef struct gpufft_product_desc {
    int kind;
    int ring_power;
    long long a_offset, b_offset, out_offset;
} gpufft_product_desc;
typedef struct gpufft_batched_plan_s* gpufft_batched_plan;
int gpufft_batched_plan_create(gpufft_batched_plan* out, const gpufft_product_desc* descs, int count,
                               int a_limb_bits, int a_limbs, int b_limb_bits, int b_limbs,
                               void* stream);
/* The same list in FP64, one limb per operand: exact rounded products while
 * the coefficients and the transform's rounding stay well inside FP64
 * (err_stats shows the margin; rings to 2^17 with 9-bit operands leave
 * errors near 1e-6).  NEGACYCLIC ring_power 6..18, LINEAR 5..17; LDS tile
 * 2^13 (128 KiB); radix_bits 3 or 4.  Everything else as above.  (No
 * counterpart in the reference.) */
int gpufft_batched_plan_create_f64(gpufft_batched_plan* out, const gpufft_product_desc* descs,
                                   int count, void* stream);
int gpufft_batched_plan_destroy(gpufft_batched_plan plan);
/* radix of the LDS steps (2^3..2^5; FP64 plans 2^3..2^4) and persistent
 * grid size (workgroups); chosen by the autotuner (gpufft_amd.autotune) */
int gpufft_batched_plan_set_tuning(gpufft_batched_plan plan, int radix_bits, int grid);
/* where split products are combined: 0 = separate streaming kernel (default),
 * 1 = in the persistent kernel (one launch); chosen by the autotuner */
int gpufft_batched_plan_set_combine(gpufft_batched_plan plan, int in_kernel);
int gpufft_batched_plan_get_combine(gpufft_batched_plan plan, int* in_kernel);
/* LDS tile of a workgroup, log2 of its points: 14 (128 KiB, one workgroup
 * of 8 waves per CU; default) or 13 (64 KiB, two per CU; products above
 * the tile are split into more sub-problems, at most 2^4).  Rebuilds the
 * work list and synchronises; chosen by the autotuner */
int gpufft_batched_plan_set_tile(gpufft_batched_plan plan, int log2_tile);
int gpufft_batched_plan_get_tile(gpufft_batched_plan plan, int* log2_tile);
int gpufft_batched_plan_info(gpufft_batched_plan plan, int* n_items, long long* scratch_elems,
                             int* radix_bits, int* grid);
int gpufft_batched_execute(gpufft_batched_plan plan, const int32_t* a, const int32_t* b,
                           int64_t* out, double* err_stats, void* stream);
int gpufft_batched_graph_launch(gpufft_batched_plan plan, const int32_t* a, const int32_t* b,
                                int64_t* out, double* err_stats, void* stream);

/* ---- errors ---- */
/* Human-readable description of the last failure on the calling thread. */
const char* gpufft_last_error_message(void);
/* hipError_t value of the last GPUFFT_ERR_HIP on the calling thread. */
int gpufft_last_hip_error(void);

#ifdef __cplusplus
}
#endif
#endif
