/*
 * CPU oracle for the GPU-FFT polynomial-multiplication path.
 * TEST INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's cpu_baseline).
 * See oracle.h for the contract; each function cites the reference lines
 * (Alisah-Ozcan/GPU-FFT) it restates.
 *
 * Built with -ffp-contract=off so that products and sums round exactly as
 * written, the way the reference's host code (g++, no FMA) rounds them.
 * cos and sin of one angle are evaluated with glibc's sincos(): the
 * reference's Release build (g++ -O3 on its host code, CMakeLists.txt:31-33)
 * fuses std::cos(x), std::sin(x) into that call, whose result differs from
 * separate cos()/sin() in the last bit for some x.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* fft_cpu.cu:11-21 */
int orc_bitreverse(int index, int n_power)
{
    int r = 0;
    for (int i = 0; i < n_power; i++) {
        r = (r << 1) | (index & 1);
        index >>= 1;
    }
    return r;
}

static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

/* ------------------------------------------------------------------------ */
/* complex helpers that round the way complex.cuh:65-180 rounds             */
/* ------------------------------------------------------------------------ */
#define DEF_CPLX(T, S)                                                              \
    typedef struct { T re, im; } c_##S;                                             \
    static inline c_##S mk_##S(T re, T im) { c_##S z; z.re = re; z.im = im; return z; } \
    static inline c_##S add_##S(c_##S a, c_##S b) { return mk_##S(a.re + b.re, a.im + b.im); } \
    static inline c_##S sub_##S(c_##S a, c_##S b) { return mk_##S(a.re - b.re, a.im - b.im); } \
    /* complex.cuh:133-139 */                                                       \
    static inline c_##S mul_##S(c_##S a, c_##S b)                                   \
    {                                                                               \
        return mk_##S(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);        \
    }                                                                               \
    /* complex.cuh:172-180 */                                                       \
    static inline c_##S div_##S(c_##S a, c_##S b)                                   \
    {                                                                               \
        T den = b.re * b.re + b.im * b.im;                                          \
        return mk_##S((a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den); \
    }
DEF_CPLX(double, f64)
DEF_CPLX(float, f32)

/* ------------------------------------------------------------------------ */
/* tables                                                                   */
/* ------------------------------------------------------------------------ */

/* FFT<T>::FFT + GenerateRootTable + GenerateInverseRootTable
 * (fft_cpu.cu:93-129), ReverseRootTable / InverseReverseRootTable
 * (fft_cpu.cu:200-226).  root = (2)*(pi)/(2n) as complex ops, exact for
 * power-of-two n; element = exp((0,1)*(i,0)*(root,0)) = (cos th, sin th)
 * with th = i*root rounded in T (complex.cuh:305-311). */
static inline void sc_f64(double x, double *s, double *c) { sincos(x, s, c); }
static inline void sc_f32(float x, float *s, float *c) { sincosf(x, s, c); }

#define DEF_FFT_TABLES(T, S, COS, SIN)                                              \
    void orc_fft_tables_##S(int n, T *root, T *inv, T *rev, T *inv_rev)             \
    {                                                                               \
        int lg = ilog2(n);                                                          \
        T max_size = (T)(2 * n);                                                    \
        T r = ((T)2.0 * (T)M_PI) / max_size;                                        \
        c_##S one = mk_##S((T)1, (T)0);                                             \
        c_##S *rt = (c_##S *)malloc(sizeof(c_##S) * (size_t)n);                     \
        c_##S *it = (c_##S *)malloc(sizeof(c_##S) * (size_t)n);                     \
        for (int i = 0; i < n; i++) {                                               \
            T th = (T)i * r, sv, cv;                                                \
            sc_##S(th, &sv, &cv);                                                   \
            rt[i] = mk_##S(cv, sv);                                                 \
            it[i] = div_##S(one, rt[i]);                                            \
        }                                                                           \
        for (int i = 0; i < n; i++) {                                               \
            if (root) { root[2 * i] = rt[i].re; root[2 * i + 1] = rt[i].im; }       \
            if (inv) { inv[2 * i] = it[i].re; inv[2 * i + 1] = it[i].im; }          \
            int b = orc_bitreverse(i, lg);                                          \
            if (rev) { rev[2 * i] = rt[b].re; rev[2 * i + 1] = rt[b].im; }          \
            if (inv_rev) { inv_rev[2 * i] = it[b].re; inv_rev[2 * i + 1] = it[b].im; } \
        }                                                                           \
        free(rt);                                                                   \
        free(it);                                                                   \
    }
DEF_FFT_TABLES(double, f64, cos, sin)
DEF_FFT_TABLES(float, f32, cosf, sinf)

/* FFNT<T>::FFNT (fft_cpu.cu:234-299) and the table accessors
 * (fft_cpu.cu:434-470).  root_new = 2*pi/(N/2) and the twist angle step
 * 2*pi/(2N) are computed in double and then stored in T, as the reference
 * does with `T root_new = 2.0 * M_PI / static_cast<T>(...)`.  The reversed
 * tables use bitreverse over log2(N/2)-1 bits, so entries >= N/4 repeat. */
#define DEF_FFNT_TABLES(T, S, COS, SIN)                                             \
    void orc_ffnt_tables_##S(int N, T *rev_root, T *inv_rev_root, T *twist, T *untwist) \
    {                                                                               \
        int L = N >> 1;                                                             \
        int lgn = ilog2(L);                                                         \
        c_##S one = mk_##S((T)1, (T)0);                                             \
        T rn = (T)(2.0 * M_PI / (double)(T)L);                                      \
        T rt = (T)(2.0 * M_PI / (double)(T)(2 * N));                                \
        c_##S *root = (c_##S *)malloc(sizeof(c_##S) * (size_t)L);                   \
        c_##S *iroot = (c_##S *)malloc(sizeof(c_##S) * (size_t)L);                  \
        for (int i = 0; i < L; i++) {                                               \
            T th = (T)i * rn, sv, cv;                                               \
            sc_##S(th, &sv, &cv);                                                   \
            root[i] = mk_##S(cv, sv);                                               \
            iroot[i] = div_##S(one, root[i]);                                       \
            T tw = (T)i * rt;                                                       \
            sc_##S(tw, &sv, &cv);                                                   \
            c_##S w = mk_##S(cv, sv);                                               \
            c_##S uw = div_##S(one, w);                                             \
            if (twist) { twist[2 * i] = w.re; twist[2 * i + 1] = w.im; }            \
            if (untwist) { untwist[2 * i] = uw.re; untwist[2 * i + 1] = uw.im; }    \
        }                                                                           \
        for (int i = 0; i < L; i++) {                                               \
            int b = orc_bitreverse(i, lgn - 1);                                     \
            if (rev_root) { rev_root[2 * i] = root[b].re; rev_root[2 * i + 1] = root[b].im; } \
            if (inv_rev_root) {                                                     \
                inv_rev_root[2 * i] = iroot[b].re;                                  \
                inv_rev_root[2 * i + 1] = iroot[b].im;                              \
            }                                                                       \
        }                                                                           \
        free(root);                                                                 \
        free(iroot);                                                                \
    }
DEF_FFNT_TABLES(double, f64, cos, sin)
DEF_FFNT_TABLES(float, f32, cosf, sinf)

/* Special-FFT table.  The reference ships no generator; the layout is the
 * one its kernels index (fft.cu:782,805,824,908: T[h + (w & (h-1))]) and its
 * docstring describes (fft.cuh:134-180: Psi[(5^j mod 4m)*(l/m)] with
 * Psi[k] = exp(pi*i*k/(2l)), m = 2h):  T[h+j] = exp(2*pi*i*(5^j mod 8h)/(8h)).
 * The inverse table is the conjugate. */
void orc_special_table_f64(int l, int inverse, double *tab)
{
    tab[0] = 1.0;
    tab[1] = 0.0;
    for (int h = 1; h < l; h <<= 1) {
        long long mod = 8LL * h;
        long long p5 = 1;
        for (int j = 0; j < h; j++) {
            /* angle = 2 pi r / (8h), evaluated in long double and rounded
             * once to double */
            long double ang = 2.0L * 3.141592653589793238462643383279502884L * (long double)p5 /
                              (long double)mod;
            double c = (double)cosl(ang), s = (double)sinl(ang);
            tab[2 * (h + j)] = c;
            tab[2 * (h + j) + 1] = inverse ? -s : s;
            p5 = (p5 * 5) % mod;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* CPU FFT path                                                             */
/* ------------------------------------------------------------------------ */

/* FFT<T>::fft, fft_cpu.cu:131-161; `root` is FFT(n).root_tables. */
void orc_cpu_fft_f64(int n, const double *root, double *data)
{
    const c_f64 *rt = (const c_f64 *)root;
    c_f64 *x = (c_f64 *)data;
    int logn = ilog2(n);
    int max_size = 2 * n;
    int t = max_size, m = 1;
    while (m < max_size) {
        t >>= 1;
        for (int i = 0; i < m; i++) {
            int j1 = 2 * i * t;
            c_f64 S = rt[orc_bitreverse(i, logn)];
            for (int j = j1; j < j1 + t; j++) {
                c_f64 U = x[j], V = x[j + t];
                x[j] = add_f64(U, mul_f64(V, S));
                x[j + t] = sub_f64(U, mul_f64(V, S));
            }
        }
        m <<= 1;
    }
}

/* FFT<T>::ifft, fft_cpu.cu:163-198.  The reference's scalar product
 * (complex.cuh:153-159) writes real*s into the imaginary part; only the real
 * part of its result is ever consumed, and here both parts are scaled. */
void orc_cpu_ifft_f64(int n, const double *inv_root, double *data)
{
    const c_f64 *it = (const c_f64 *)inv_root;
    c_f64 *x = (c_f64 *)data;
    int logn = ilog2(n);
    int max_size = 2 * n;
    int t = 1, m = max_size;
    while (m > 1) {
        int j1 = 0, h = m >> 1;
        for (int i = 0; i < h; i++) {
            c_f64 S = it[orc_bitreverse(i, logn)];
            for (int j = j1; j < j1 + t; j++) {
                c_f64 U = x[j], V = x[j + t];
                x[j] = add_f64(U, V);
                x[j + t] = mul_f64(sub_f64(U, V), S);
            }
            j1 += (t << 1);
        }
        t <<= 1;
        m >>= 1;
    }
    double ninv = 1.0 / (double)max_size;
    for (int i = 0; i < max_size; i++) {
        x[i].re *= ninv;
        x[i].im *= ninv;
    }
}

/* FFNT<T>::MULT, fft_cpu.cu:301-432: fold + twist, two forward FFTs of
 * length N/2 with bitreverse over log2(N/2)-1 bits, pointwise product, GS
 * inverse, untwist then scale by 1/(N/2), round, unfold.  The reference
 * takes unsigned long long (negative values wrap); this takes int64. */
void orc_cpu_ffnt_mult_f64(int N, const int64_t *a, const int64_t *b, int64_t *out)
{
    int L = N >> 1;
    int lgn = ilog2(L);
    double rn = 2.0 * M_PI / (double)L;
    double rtw = 2.0 * M_PI / (double)(2 * N);
    c_f64 one = mk_f64(1.0, 0.0);
    c_f64 *root = (c_f64 *)malloc(sizeof(c_f64) * (size_t)L);
    c_f64 *iroot = (c_f64 *)malloc(sizeof(c_f64) * (size_t)L);
    c_f64 *tw = (c_f64 *)malloc(sizeof(c_f64) * (size_t)L);
    c_f64 *utw = (c_f64 *)malloc(sizeof(c_f64) * (size_t)L);
    c_f64 *x1 = (c_f64 *)malloc(sizeof(c_f64) * (size_t)L);
    c_f64 *x2 = (c_f64 *)malloc(sizeof(c_f64) * (size_t)L);
    for (int i = 0; i < L; i++) {
        double th = (double)i * rn, sv, cv;
        sincos(th, &sv, &cv);
        root[i] = mk_f64(cv, sv);
        iroot[i] = div_f64(one, root[i]);
        double tt = (double)i * rtw;
        sincos(tt, &sv, &cv);
        tw[i] = mk_f64(cv, sv);
        utw[i] = div_f64(one, tw[i]);
    }
    for (int i = 0; i < L; i++) {
        x1[i] = mul_f64(mk_f64((double)a[i], (double)a[i + L]), tw[i]);
        x2[i] = mul_f64(mk_f64((double)b[i], (double)b[i + L]), tw[i]);
    }
    c_f64 *xs[2] = {x1, x2};
    for (int v = 0; v < 2; v++) {
        c_f64 *x = xs[v];
        int t = L, m = 1;
        while (m < L) {
            t >>= 1;
            for (int i = 0; i < m; i++) {
                int j1 = 2 * i * t;
                c_f64 S = root[orc_bitreverse(i, lgn - 1)];
                for (int j = j1; j < j1 + t; j++) {
                    c_f64 U = x[j], V = x[j + t];
                    x[j] = add_f64(U, mul_f64(V, S));
                    x[j + t] = sub_f64(U, mul_f64(V, S));
                }
            }
            m <<= 1;
        }
    }
    for (int i = 0; i < L; i++) x1[i] = mul_f64(x1[i], x2[i]);
    {
        int t = 1, m = L;
        while (m > 1) {
            int j1 = 0, h = m >> 1;
            for (int i = 0; i < h; i++) {
                c_f64 S = iroot[orc_bitreverse(i, lgn - 1)];
                for (int j = j1; j < j1 + t; j++) {
                    c_f64 U = x1[j], V = x1[j + t];
                    x1[j] = add_f64(U, V);
                    x1[j + t] = mul_f64(sub_f64(U, V), S);
                }
                j1 += (t << 1);
            }
            t <<= 1;
            m >>= 1;
        }
    }
    double ninv = 1.0 / (double)L;
    for (int i = 0; i < L; i++) {
        c_f64 z = mul_f64(x1[i], utw[i]);
        z = mul_f64(z, mk_f64(ninv, 0.0));
        out[i] = (int64_t)round(z.re);
        out[i + L] = (int64_t)round(z.im);
    }
    free(root); free(iroot); free(tw); free(utw); free(x1); free(x2);
}

typedef struct {
    int N, begin, end;
    const int64_t *a, *b;
    int64_t *out;
} mult_job;

static void *mult_worker(void *p)
{
    mult_job *j = (mult_job *)p;
    for (int i = j->begin; i < j->end; i++)
        orc_cpu_ffnt_mult_f64(j->N, j->a + (size_t)i * j->N, j->b + (size_t)i * j->N,
                              j->out + (size_t)i * j->N);
    return NULL;
}

void orc_cpu_ffnt_mult_batch_f64(int N, int count, const int64_t *a, const int64_t *b,
                                 int64_t *out, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > count) threads = count;
    pthread_t th[256];
    mult_job jobs[256];
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; t++) {
        jobs[t].N = N;
        jobs[t].begin = (int)((long long)count * t / threads);
        jobs[t].end = (int)((long long)count * (t + 1) / threads);
        jobs[t].a = a; jobs[t].b = b; jobs[t].out = out;
        pthread_create(&th[t], NULL, mult_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
}

/* The reference's CPU plain-method product (example/test_fft_cpu.cu:65-142):
 * a, b of n coefficients zero-padded to 2n, FFT<double>(n)::fft of both,
 * pointwise product, ifft, round the real parts: the linear product, 2n
 * outputs (the last one 0).  `root` and `inv` are FFT(n).root_tables and
 * inverse_root_tables (fft_cpu.cu:93-129); work holds 2 x 2n complex. */
static void cpu_fft_mult_one(int n, const double *root, const double *inv, const int64_t *a,
                             const int64_t *b, int64_t *out, c_f64 *work)
{
    c_f64 *x = work, *y = work + 2 * n;
    for (int i = 0; i < 2 * n; i++) {
        x[i] = mk_f64(i < n ? (double)a[i] : 0.0, 0.0);
        y[i] = mk_f64(i < n ? (double)b[i] : 0.0, 0.0);
    }
    orc_cpu_fft_f64(n, root, (double *)x);
    orc_cpu_fft_f64(n, root, (double *)y);
    for (int i = 0; i < 2 * n; i++) x[i] = mul_f64(x[i], y[i]);
    orc_cpu_ifft_f64(n, inv, (double *)x);
    for (int i = 0; i < 2 * n; i++) out[i] = (int64_t)round(x[i].re);
}

typedef struct {
    int n, begin, end;
    const double *root, *inv;
    const int64_t *a, *b;
    int64_t *out;
} fft_mult_job;

static void *fft_mult_worker(void *p)
{
    fft_mult_job *j = (fft_mult_job *)p;
    c_f64 *work = (c_f64 *)malloc(sizeof(c_f64) * 4 * (size_t)j->n);
    for (int i = j->begin; i < j->end; i++)
        cpu_fft_mult_one(j->n, j->root, j->inv, j->a + (size_t)i * j->n, j->b + (size_t)i * j->n,
                         j->out + (size_t)i * 2 * j->n, work);
    free(work);
    return NULL;
}

void orc_cpu_fft_mult_batch_f64(int n, int count, const int64_t *a, const int64_t *b,
                                int64_t *out, int threads)
{
    /* tables once per call, as the reference builds FFT<T>(n) once for a batch */
    double *root = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *inv = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *rev = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *inv_rev = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    orc_fft_tables_f64(n, root, inv, rev, inv_rev);
    if (threads < 1) threads = 1;
    if (threads > count) threads = count;
    if (threads > 256) threads = 256;
    pthread_t th[256];
    fft_mult_job jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t].n = n;
        jobs[t].begin = (int)((long long)count * t / threads);
        jobs[t].end = (int)((long long)count * (t + 1) / threads);
        jobs[t].root = root; jobs[t].inv = inv;
        jobs[t].a = a; jobs[t].b = b; jobs[t].out = out;
        pthread_create(&th[t], NULL, fft_mult_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(root); free(inv); free(rev); free(inv_rev);
}

/* ------------------------------------------------------------------------ */
/* Observable semantics of the GPU entry points (fft.cu)                    */
/* ------------------------------------------------------------------------ */
/*
 * Cooley-Tukey forward (ForwardCore, fft.cu:11-152, with CooleyTukeyUnit
 * fft.cuh:69-78): natural order in, bit-reversed order out.  Stage s has
 * m = 2^s groups of half-size t = L/(2m); group g uses
 * table[g] (X^N-1) or table[m + g] (X^N+1), fft.cu:67-74.
 * Gentleman-Sande inverse (InverseCore, fft.cu:154-260, GentlemanSandeUnit
 * fft.cuh:80-89): the same stages in reverse order, U+V, (U-V)*w.
 */
#define DEF_GPU_SEM(T, S)                                                           \
    static void ct_fwd_##S(c_##S *x, int n_power, const c_##S *tab, int nega)      \
    {                                                                               \
        int L = 1 << n_power;                                                       \
        for (int s = 0; s < n_power; s++) {                                         \
            int m = 1 << s, t = L >> (s + 1);                                       \
            for (int g = 0; g < m; g++) {                                           \
                c_##S w = tab[(nega ? m : 0) + g];                                  \
                for (int j = 2 * g * t; j < 2 * g * t + t; j++) {                   \
                    c_##S U = x[j], V = mul_##S(x[j + t], w);                       \
                    x[j] = add_##S(U, V);                                           \
                    x[j + t] = sub_##S(U, V);                                       \
                }                                                                   \
            }                                                                       \
        }                                                                           \
    }                                                                               \
    static void gs_inv_##S(c_##S *x, int n_power, const c_##S *tab, int nega)      \
    {                                                                               \
        int L = 1 << n_power;                                                       \
        for (int s = n_power - 1; s >= 0; s--) {                                    \
            int m = 1 << s, t = L >> (s + 1);                                       \
            for (int g = 0; g < m; g++) {                                           \
                c_##S w = tab[(nega ? m : 0) + g];                                  \
                for (int j = 2 * g * t; j < 2 * g * t + t; j++) {                   \
                    c_##S U = x[j], V = x[j + t];                                   \
                    x[j] = add_##S(U, V);                                           \
                    x[j + t] = mul_##S(sub_##S(U, V), w);                           \
                }                                                                   \
            }                                                                       \
        }                                                                           \
    }                                                                               \
    /* GPU_FFT complex-to-complex FORWARD (fft.cu:531-553) */                       \
    void orc_c2c_forward_##S(T *data, int n_power, int batch, const T *table, int nega) \
    {                                                                               \
        size_t L = (size_t)1 << n_power;                                            \
        for (int b = 0; b < batch; b++)                                             \
            ct_fwd_##S((c_##S *)data + b * L, n_power, (const c_##S *)table, nega); \
    }                                                                               \
    /* GPU_FFT complex-to-complex INVERSE (fft.cu:554-593): optional product    \
     * X[b]*X[b+batch] on load (fft.cu:194-205), scale by mod_inverse as a      \
     * complex product on the last launch (fft.cu:246-253). */                   \
    void orc_c2c_inverse_##S(T *data, int n_power, int batch, const T *table, int nega, \
                             int mult, T scale_re, T scale_im)                      \
    {                                                                               \
        size_t L = (size_t)1 << n_power;                                            \
        c_##S *x = (c_##S *)data;                                                   \
        c_##S sc = mk_##S(scale_re, scale_im);                                      \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *p = x + b * L;                                                   \
            if (mult)                                                               \
                for (size_t i = 0; i < L; i++) p[i] = mul_##S(p[i], p[i + batch * L]); \
            gs_inv_##S(p, n_power, (const c_##S *)table, nega);                     \
            for (size_t i = 0; i < L; i++) p[i] = mul_##S(p[i], sc);                \
        }                                                                           \
    }                                                                               \
    /* GPU_FFT real-to-complex FORWARD (fft.cu:614-654, R->C promotion at       \
     * fft.cu:303-305): real [batch][L] in, full complex spectrum out. */         \
    void orc_r2r_forward_##S(const T *real, T *temp, int n_power, int batch,        \
                             const T *table, int nega)                              \
    {                                                                               \
        size_t L = (size_t)1 << n_power;                                            \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *p = (c_##S *)temp + b * L;                                       \
            for (size_t i = 0; i < L; i++) p[i] = mk_##S(real[b * L + i], (T)0);    \
            ct_fwd_##S(p, n_power, (const c_##S *)table, nega);                     \
        }                                                                           \
    }                                                                               \
    /* GPU_FFT complex-to-real INVERSE (fft.cu:655-720, store at 499-514):      \
     * real[b][i] = Re(x[i] * mod_inverse).  temp is used as scratch. */          \
    void orc_r2r_inverse_##S(T *real, T *temp, int n_power, int batch, const T *table, \
                             int nega, int mult, T scale_re, T scale_im)            \
    {                                                                               \
        size_t L = (size_t)1 << n_power;                                            \
        c_##S *x = (c_##S *)temp;                                                   \
        c_##S sc = mk_##S(scale_re, scale_im);                                      \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *p = x + b * L;                                                   \
            if (mult)                                                               \
                for (size_t i = 0; i < L; i++) p[i] = mul_##S(p[i], p[i + batch * L]); \
            gs_inv_##S(p, n_power, (const c_##S *)table, nega);                     \
            for (size_t i = 0; i < L; i++) real[b * L + i] = mul_##S(p[i], sc).re;  \
        }                                                                           \
    }                                                                               \
    /* GPU_FFNT FORWARD (fft.cu:1403-1455; fold + twist at fft.cu:1057-1069):   \
     * in [batch][2^n_power] real, temp [batch][2^(n_power-1)] complex. */       \
    void orc_ffnt_forward_##S(const T *in, T *temp, int n_power, int batch,         \
                              const T *twist, const T *table)                       \
    {                                                                               \
        size_t N = (size_t)1 << n_power, L = N >> 1;                                \
        const c_##S *tw = (const c_##S *)twist;                                     \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *p = (c_##S *)temp + b * L;                                       \
            const T *f = in + b * N;                                                \
            for (size_t i = 0; i < L; i++) p[i] = mul_##S(mk_##S(f[i], f[i + L]), tw[i]); \
            ct_fwd_##S(p, n_power - 1, (const c_##S *)table, 0);                    \
        }                                                                           \
    }                                                                               \
    /* GPU_FFNT INVERSE (fft.cu:1456-1510; product fft.cu:1252-1263; scale,     \
     * untwist, unfold fft.cu:1366-1383). */                                      \
    void orc_ffnt_inverse_##S(T *out, T *temp, int n_power, int batch,              \
                              const T *untwist, const T *table, int mult, T scale_re, \
                              T scale_im)                                           \
    {                                                                               \
        size_t N = (size_t)1 << n_power, L = N >> 1;                                \
        const c_##S *utw = (const c_##S *)untwist;                                  \
        c_##S sc = mk_##S(scale_re, scale_im);                                      \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *p = (c_##S *)temp + b * L;                                       \
            if (mult)                                                               \
                for (size_t i = 0; i < L; i++) p[i] = mul_##S(p[i], p[i + batch * L]); \
            gs_inv_##S(p, n_power - 1, (const c_##S *)table, 0);                    \
            for (size_t i = 0; i < L; i++) {                                        \
                c_##S z = mul_##S(mul_##S(p[i], sc), utw[i]);                       \
                out[b * N + i] = z.re;                                              \
                out[b * N + i + L] = z.im;                                          \
            }                                                                       \
        }                                                                           \
    }                                                                               \
    /* GPU_Special_FFT FORWARD (Special_ForwardCore fft.cu:856-926): stride      \
     * ascending, CT with twiddle T[h + (p & (h-1))] (fft.cu:908). */            \
    void orc_special_forward_##S(T *data, int n_power, int batch, const T *table)   \
    {                                                                               \
        size_t L = (size_t)1 << n_power;                                            \
        const c_##S *tab = (const c_##S *)table;                                    \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *x = (c_##S *)data + b * L;                                       \
            for (size_t h = 1; h < L; h <<= 1)                                      \
                for (size_t p = 0; p < L; p++) {                                    \
                    if (p & h) continue;                                            \
                    c_##S U = x[p], V = mul_##S(x[p + h], tab[h + (p & (h - 1))]);  \
                    x[p] = add_##S(U, V);                                           \
                    x[p + h] = sub_##S(U, V);                                       \
                }                                                                   \
        }                                                                           \
    }                                                                               \
    /* GPU_Special_FFT INVERSE (Special_InverseCore fft.cu:727-854): stride      \
     * descending, GS with twiddle T[h + (p & (h-1))] (fft.cu:782), then the     \
     * complex scale on the last launch (fft.cu:846-853). */                     \
    void orc_special_inverse_##S(T *data, int n_power, int batch, const T *table,   \
                                 T scale_re, T scale_im)                            \
    {                                                                               \
        size_t L = (size_t)1 << n_power;                                            \
        const c_##S *tab = (const c_##S *)table;                                    \
        c_##S sc = mk_##S(scale_re, scale_im);                                      \
        for (int b = 0; b < batch; b++) {                                           \
            c_##S *x = (c_##S *)data + b * L;                                       \
            for (size_t h = L >> 1; h >= 1; h >>= 1)                                \
                for (size_t p = 0; p < L; p++) {                                    \
                    if (p & h) continue;                                            \
                    c_##S U = x[p], V = x[p + h];                                   \
                    x[p] = add_##S(U, V);                                           \
                    x[p + h] = mul_##S(sub_##S(U, V), tab[h + (p & (h - 1))]);      \
                }                                                                   \
            for (size_t i = 0; i < L; i++) x[i] = mul_##S(x[i], sc);                \
        }                                                                           \
    }
DEF_GPU_SEM(double, f64)
DEF_GPU_SEM(float, f32)

/* CKKS encode / decode: the steps a caller of GPU_Special_FFT performs around
 * it (fft.cuh:134-180); see oracle.h. */
void orc_ckks_encode_f64(const double *slots, int n_power, int batch, const double *inv_table,
                         double delta, int64_t *coeffs)
{
    size_t l = (size_t)1 << n_power;
    double *u = (double *)malloc(sizeof(double) * 2 * l * (size_t)batch);
    memcpy(u, slots, sizeof(double) * 2 * l * (size_t)batch);
    orc_special_inverse_f64(u, n_power, batch, inv_table, 1.0 / (double)l, 0.0);
    for (int b = 0; b < batch; b++)
        for (size_t i = 0; i < l; i++) {
            const double *z = u + 2 * ((size_t)b * l + (size_t)orc_bitreverse((int)i, n_power));
            coeffs[(size_t)b * 2 * l + i] = llround(z[0] * delta);
            coeffs[(size_t)b * 2 * l + i + l] = llround(z[1] * delta);
        }
    free(u);
}

void orc_ckks_decode_f64(const int64_t *coeffs, int n_power, int batch, const double *table,
                         double delta, double *slots)
{
    size_t l = (size_t)1 << n_power;
    for (int b = 0; b < batch; b++)
        for (size_t i = 0; i < l; i++) {
            double *v = slots + 2 * ((size_t)b * l + (size_t)orc_bitreverse((int)i, n_power));
            v[0] = (double)coeffs[(size_t)b * 2 * l + i] / delta;
            v[1] = (double)coeffs[(size_t)b * 2 * l + i + l] / delta;
        }
    orc_special_forward_f64(slots, n_power, batch, table);
}

/* ------------------------------------------------------------------------ */
/* exact integer products                                                   */
/* ------------------------------------------------------------------------ */

/* schoolbook_poly_multiplication_without_reduction, fft_cpu.cu:52-73,
 * with exact int64 accumulation instead of `% modulus`. */
void orc_schoolbook_linear(int n, const int64_t *a, const int64_t *b, int64_t *out)
{
    memset(out, 0, sizeof(int64_t) * 2 * (size_t)n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) out[i + j] += a[i] * b[j];
}

/* schoolbook_poly_multiplication, fft_cpu.cu:23-50: linear product, then
 * out[i] = lin[i] - lin[i+n] (reduction mod X^n+1), without `% modulus`. */
void orc_schoolbook_negacyclic(int n, const int64_t *a, const int64_t *b, int64_t *out)
{
    int64_t *lin = (int64_t *)calloc(2 * (size_t)n, sizeof(int64_t));
    orc_schoolbook_linear(n, a, b, lin);
    for (int i = 0; i < n; i++) out[i] = lin[i] - lin[i + n];
    free(lin);
}

void orc_schoolbook_cyclic(int n, const int64_t *a, const int64_t *b, int64_t *out)
{
    int64_t *lin = (int64_t *)calloc(2 * (size_t)n, sizeof(int64_t));
    orc_schoolbook_linear(n, a, b, lin);
    for (int i = 0; i < n; i++) out[i] = lin[i] + lin[i + n];
    free(lin);
}

/* NTT over P = 2^64 - 2^32 + 1 (multiplicative generator 7). */
#define GP 0xFFFFFFFF00000001ULL
#define GEPS 0xFFFFFFFFULL
typedef unsigned __int128 u128;

static inline uint64_t gred(u128 x)
{
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t hh = hi >> 32, hl = hi & GEPS;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GEPS;
    uint64_t t1 = hl * GEPS;
    uint64_t r = t0 + t1;
    if (r < t0) r += GEPS;
    if (r >= GP) r -= GP;
    return r;
}
static inline uint64_t gmul(uint64_t a, uint64_t b) { return gred((u128)a * b); }
static inline uint64_t gadd(uint64_t a, uint64_t b)
{
    u128 s = (u128)a + b;
    if (s >= GP) s -= GP;
    return (uint64_t)s;
}
static inline uint64_t gsub(uint64_t a, uint64_t b) { return a >= b ? a - b : a + (GP - b); }
static uint64_t gpow(uint64_t b, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = gmul(r, b);
        b = gmul(b, b);
        e >>= 1;
    }
    return r;
}
static inline uint64_t to_g(int64_t v) { return v >= 0 ? (uint64_t)v : GP - (uint64_t)(-v); }
static inline int64_t from_g(uint64_t v) { return v > GP / 2 ? -(int64_t)(GP - v) : (int64_t)v; }

/* in-place cyclic NTT of length n (natural order in and out) */
static void gntt(uint64_t *x, int n, int inverse)
{
    int lg = ilog2(n);
    for (int i = 0; i < n; i++) {
        int j = orc_bitreverse(i, lg);
        if (i < j) { uint64_t t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        uint64_t w = gpow(7, (GP - 1) / (uint64_t)len);
        if (inverse) w = gpow(w, GP - 2);
        int half = len >> 1;
        uint64_t *ws = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)half);
        ws[0] = 1;
        for (int k = 1; k < half; k++) ws[k] = gmul(ws[k - 1], w);
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < half; k++) {
                uint64_t u = x[i + k], v = gmul(x[i + k + half], ws[k]);
                x[i + k] = gadd(u, v);
                x[i + k + half] = gsub(u, v);
            }
        free(ws);
    }
    if (inverse) {
        uint64_t ninv = gpow((uint64_t)n, GP - 2);
        for (int i = 0; i < n; i++) x[i] = gmul(x[i], ninv);
    }
}

static void ntt_product(int n, const int64_t *a, const int64_t *b, int64_t *out, int nega)
{
    uint64_t *x = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint64_t *y = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint64_t psi = nega ? gpow(7, (GP - 1) / (2 * (uint64_t)n)) : 1;
    uint64_t pw = 1;
    for (int i = 0; i < n; i++) {
        x[i] = gmul(to_g(a[i]), pw);
        y[i] = gmul(to_g(b[i]), pw);
        pw = gmul(pw, psi);
    }
    gntt(x, n, 0);
    gntt(y, n, 0);
    for (int i = 0; i < n; i++) x[i] = gmul(x[i], y[i]);
    gntt(x, n, 1);
    uint64_t ipsi = nega ? gpow(psi, GP - 2) : 1;
    pw = 1;
    for (int i = 0; i < n; i++) {
        out[i] = from_g(gmul(x[i], pw));
        pw = gmul(pw, ipsi);
    }
    free(x);
    free(y);
}

typedef struct {
    int n, nega, begin, end;
    const int64_t *a, *b;
    int64_t *out;
} ntt_job;

static void *ntt_worker(void *p)
{
    ntt_job *j = (ntt_job *)p;
    for (int i = j->begin; i < j->end; i++)
        ntt_product(j->n, j->a + (size_t)i * j->n, j->b + (size_t)i * j->n,
                    j->out + (size_t)i * j->n, j->nega);
    return NULL;
}

static void ntt_batch(int n, int count, const int64_t *a, const int64_t *b, int64_t *out,
                      int nega)
{
    int threads = count < 8 ? count : 8;
    if (threads < 1) return;
    pthread_t th[8];
    ntt_job jobs[8];
    for (int t = 0; t < threads; t++) {
        jobs[t].n = n;
        jobs[t].nega = nega;
        jobs[t].begin = (int)((long long)count * t / threads);
        jobs[t].end = (int)((long long)count * (t + 1) / threads);
        jobs[t].a = a; jobs[t].b = b; jobs[t].out = out;
        pthread_create(&th[t], NULL, ntt_worker, &jobs[t]);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
}

void orc_ntt_negacyclic(int n, int count, const int64_t *a, const int64_t *b, int64_t *out)
{
    ntt_batch(n, count, a, b, out, 1);
}

void orc_ntt_cyclic(int n, int count, const int64_t *a, const int64_t *b, int64_t *out)
{
    ntt_batch(n, count, a, b, out, 0);
}

/* ------------------------------------------------------------------------ */
/* seeded inputs                                                            */
/* ------------------------------------------------------------------------ */
uint64_t orc_mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void orc_fill_uniform_i64(uint64_t seed, int64_t lo, int64_t hi, int64_t count, int64_t *out)
{
    uint64_t span = (uint64_t)(hi - lo);
    for (int64_t i = 0; i < count; i++) {
        uint64_t r = orc_mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL);
        out[i] = lo + (int64_t)(uint64_t)(((u128)r * span) >> 64);
    }
}
