// Drop-in examples through the C++ interface (include/gpufft/fft.cuh), in the
// flow of the reference's example programs (example/test_fft_gpu_C_C.cu,
// test_fft_gpu_R_R.cu, test_ffnt_gpu_R_R.cu, test_fft_cpu.cu,
// test_ffnt_cpu.cu), with seeded inputs and exact checks: every rounded
// coefficient must equal the schoolbook product (not only modulo 7).
//   dropin_examples <case> [logn] [batch]
// cases: cc rr ffnt special cpu_fft cpu_ffnt errors (the last three need no GPU)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpufft/fft.cuh"

using namespace gpufft;
typedef double TestDataType;

static int fails = 0;
#define EXPECT(c)                                                   \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            fails++;                                                \
        }                                                           \
    } while (0)

template <typename V> V* to_device(const std::vector<V>& h)
{
    V* d = nullptr;
    GPUFFT_HIP_CHECK(hipMalloc(&d, h.size() * sizeof(V)));
    GPUFFT_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
    return d;
}

template <typename V> std::vector<V> to_host(const V* d, size_t count)
{
    std::vector<V> h(count);
    GPUFFT_HIP_CHECK(hipMemcpy(h.data(), d, count * sizeof(V), hipMemcpyDeviceToHost));
    return h;
}

static const unsigned long long Q = 1ULL << 62;  // large modulus: exact products

// C_C cyclic product of zero-padded inputs (example/test_fft_gpu_C_C.cu:53-203)
static void case_cc(int logn, int batch)
{
    const int n = 1 << logn;
    std::mt19937_64 gen(1234 + logn);
    std::uniform_int_distribution<int> dis(0, 15);
    std::vector<std::vector<unsigned long long>> A(batch, std::vector<unsigned long long>(n)),
        B(batch, std::vector<unsigned long long>(n));
    std::vector<COMPLEX<TestDataType>> host(2 * batch * 2 * n);
    for (int j = 0; j < batch; j++)
        for (int i = 0; i < n; i++) {
            A[j][i] = dis(gen);
            B[j][i] = dis(gen);
            host[j * 2 * n + i] = COMPLEX<TestDataType>((double)A[j][i]);
            host[(j + batch) * 2 * n + i] = COMPLEX<TestDataType>((double)B[j][i]);
        }
    FFT<TestDataType> gen_t(n);
    auto* d = to_device(host);
    auto* rt = to_device(gen_t.ReverseRootTable());
    auto* irt = to_device(gen_t.InverseReverseRootTable());
    fft_configuration<TestDataType> cf{};
    cf.n_power = logn + 1;
    cf.fft_type = FORWARD;
    cf.reduction_poly = ReductionPolynomial::X_N_minus;
    cf.stream = 0;
    GPU_FFT(d, rt, cf, batch * 2, false);
    fft_configuration<TestDataType> ci{};
    ci.n_power = logn + 1;
    ci.fft_type = INVERSE;
    ci.reduction_poly = ReductionPolynomial::X_N_minus;
    ci.mod_inverse = COMPLEX<TestDataType>(gen_t.n_inverse, 0.0);
    ci.stream = 0;
    GPU_FFT(d, irt, ci, batch, true);
    auto out = to_host(d, (size_t)batch * 2 * n);
    for (int j = 0; j < batch; j++) {
        auto ref = schoolbook_poly_multiplication_without_reduction(A[j], B[j], Q, n);
        for (int i = 0; i < 2 * n; i++)
            EXPECT((unsigned long long)std::llround(out[j * 2 * n + i].real()) == ref[i]);
    }
    (void)hipFree(d);
    (void)hipFree(rt);
    (void)hipFree(irt);
}

// R_R cyclic product (example/test_fft_gpu_R_R.cu:53-190)
static void case_rr(int logn, int batch)
{
    const int n = 1 << logn;
    std::mt19937_64 gen(99 + logn);
    std::uniform_int_distribution<int> dis(0, 255);
    std::vector<std::vector<unsigned long long>> A(batch, std::vector<unsigned long long>(n)),
        B(batch, std::vector<unsigned long long>(n));
    std::vector<TestDataType> host(2 * batch * 2 * n, 0.0);
    for (int j = 0; j < batch; j++)
        for (int i = 0; i < n; i++) {
            A[j][i] = dis(gen);
            B[j][i] = dis(gen);
            host[j * 2 * n + i] = (double)A[j][i];
            host[(j + batch) * 2 * n + i] = (double)B[j][i];
        }
    FFT<TestDataType> gen_t(n);
    auto* d = to_device(host);
    COMPLEX<TestDataType>* temp = nullptr;
    GPUFFT_HIP_CHECK(hipMalloc(&temp, sizeof(COMPLEX<TestDataType>) * 2 * batch * 2 * n));
    auto* rt = to_device(gen_t.ReverseRootTable());
    auto* irt = to_device(gen_t.InverseReverseRootTable());
    fft_configuration<TestDataType> cf{};
    cf.n_power = logn + 1;
    cf.fft_type = FORWARD;
    cf.reduction_poly = ReductionPolynomial::X_N_minus;
    GPU_FFT(d, temp, rt, cf, batch * 2, false);
    fft_configuration<TestDataType> ci{};
    ci.n_power = logn + 1;
    ci.fft_type = INVERSE;
    ci.reduction_poly = ReductionPolynomial::X_N_minus;
    ci.mod_inverse = COMPLEX<TestDataType>(gen_t.n_inverse, 0.0);
    GPU_FFT(d, temp, irt, ci, batch, true);
    auto out = to_host(d, (size_t)batch * 2 * n);
    for (int j = 0; j < batch; j++) {
        auto ref = schoolbook_poly_multiplication_without_reduction(A[j], B[j], Q, n);
        for (int i = 0; i < 2 * n; i++)
            EXPECT((unsigned long long)std::llround(out[j * 2 * n + i]) == ref[i]);
    }
    (void)hipFree(d);
    (void)hipFree(temp);
    (void)hipFree(rt);
    (void)hipFree(irt);
}

// FFNT negacyclic product (example/test_ffnt_gpu_R_R.cu:53-181), batch > 1
static void case_ffnt(int logn, int batch)
{
    const int n = 1 << logn;
    std::mt19937_64 gen(7 + logn);
    std::uniform_int_distribution<int> dis(0, 1023);
    std::vector<std::vector<long long>> A(batch, std::vector<long long>(n)),
        B(batch, std::vector<long long>(n));
    std::vector<TestDataType> host(2 * batch * n);
    for (int j = 0; j < batch; j++)
        for (int i = 0; i < n; i++) {
            A[j][i] = dis(gen);
            B[j][i] = dis(gen) - 512;
            host[j * n + i] = (double)A[j][i];
            host[(j + batch) * n + i] = (double)B[j][i];
        }
    FFNT<TestDataType> gen_t(n);
    auto* d = to_device(host);
    COMPLEX<TestDataType>* temp = nullptr;
    GPUFFT_HIP_CHECK(hipMalloc(&temp, sizeof(COMPLEX<TestDataType>) * 2 * batch * (n >> 1)));
    auto* rt = to_device(gen_t.ReverseRootTable_ffnt());
    auto* irt = to_device(gen_t.InverseReverseRootTable_ffnt());
    auto* tw = to_device(gen_t.twist_table_ffnt());
    auto* utw = to_device(gen_t.untwist_table_ffnt());
    fft_configuration<TestDataType> cf{};
    cf.n_power = logn;
    cf.fft_type = FORWARD;
    GPU_FFNT(d, temp, tw, rt, cf, batch * 2, false);
    fft_configuration<TestDataType> ci{};
    ci.n_power = logn;
    ci.fft_type = INVERSE;
    ci.mod_inverse = COMPLEX<TestDataType>(1.0 / (n >> 1), 0.0);
    GPU_FFNT(d, temp, utw, irt, ci, batch, true);
    auto out = to_host(d, (size_t)batch * n);
    for (int j = 0; j < batch; j++) {
        std::vector<long long> lin(2 * n, 0);
        for (int x = 0; x < n; x++)
            for (int y = 0; y < n; y++) lin[x + y] += A[j][x] * B[j][y];
        for (int i = 0; i < n; i++) EXPECT(std::llround(out[j * n + i]) == lin[i] - lin[i + n]);
    }
    (void)hipFree(d);
    (void)hipFree(temp);
    (void)hipFree(rt);
    (void)hipFree(irt);
    (void)hipFree(tw);
    (void)hipFree(utw);
}

// special FFT: encode (INVERSE) then decode (FORWARD) returns the slots
static void case_special(int logn, int batch)
{
    const int l = 1 << logn;
    std::vector<COMPLEX<TestDataType>> T(l), Ti(l), z(batch * l);
    T[0] = Ti[0] = COMPLEX<TestDataType>(1.0);
    for (int h = 1; h < l; h <<= 1) {
        long long p = 1;
        for (int j = 0; j < h; j++) {
            double ang = 2.0 * M_PI * (double)p / (8.0 * h);
            T[h + j] = COMPLEX<TestDataType>(std::cos(ang), std::sin(ang));
            Ti[h + j] = T[h + j].conjugate();
            p = (p * 5) % (8LL * h);
        }
    }
    std::mt19937_64 gen(5);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto& v : z) v = COMPLEX<TestDataType>(u(gen), u(gen));
    auto* d = to_device(z);
    auto* dt = to_device(T);
    auto* dti = to_device(Ti);
    fft_configuration<TestDataType> ci{};
    ci.n_power = logn;
    ci.fft_type = INVERSE;
    ci.mod_inverse = COMPLEX<TestDataType>(1.0 / l, 0.0);
    GPU_Special_FFT(d, dti, ci, batch);
    fft_configuration<TestDataType> cf{};
    cf.n_power = logn;
    cf.fft_type = FORWARD;
    GPU_Special_FFT(d, dt, cf, batch);
    auto out = to_host(d, z.size());
    double err = 0;
    for (size_t i = 0; i < z.size(); i++)
        err = std::fmax(err, std::fabs(out[i].real() - z[i].real()) + std::fabs(out[i].imag() - z[i].imag()));
    EXPECT(err < 1e-12);
    (void)hipFree(d);
    (void)hipFree(dt);
    (void)hipFree(dti);
}

// CPU FFT path (example/test_fft_cpu.cu:53-154)
static void case_cpu_fft(int logn)
{
    const int n = 1 << logn;
    std::mt19937_64 gen(3);
    std::uniform_int_distribution<int> dis(0, 6);
    std::vector<unsigned long long> A(n), B(n);
    std::vector<COMPLEX<TestDataType>> a(2 * n), b(2 * n);
    for (int i = 0; i < n; i++) {
        A[i] = dis(gen);
        B[i] = dis(gen);
        a[i] = COMPLEX<TestDataType>((double)A[i]);
        b[i] = COMPLEX<TestDataType>((double)B[i]);
    }
    FFT<TestDataType> f(n);
    f.fft(a);
    f.fft(b);
    for (int i = 0; i < 2 * n; i++) a[i] = a[i] * b[i];
    f.ifft(a);
    auto ref = schoolbook_poly_multiplication_without_reduction(A, B, Q, n);
    for (int i = 0; i < 2 * n; i++) EXPECT((unsigned long long)std::llround(a[i].real()) == ref[i]);
}

// CPU FFNT path (example/test_ffnt_cpu.cu:25-84)
static void case_cpu_ffnt(int logn)
{
    const int n = 1 << logn;
    std::mt19937_64 gen(4);
    std::uniform_int_distribution<int> dis(0, 6);
    std::vector<unsigned long long> A(n), B(n);
    for (int i = 0; i < n; i++) {
        A[i] = dis(gen);
        B[i] = dis(gen);
    }
    FFNT<TestDataType> f(n);
    auto got = f.MULT(A, B);
    std::vector<long long> lin(2 * n, 0);
    for (int x = 0; x < n; x++)
        for (int y = 0; y < n; y++) lin[x + y] += (long long)(A[x] * B[y]);
    for (int i = 0; i < n; i++) EXPECT((long long)got[i] == lin[i] - lin[i + n]);
    // the reference's own check (mod 7 after wrapping)
    auto ref7 = schoolbook_poly_multiplication(A, B, 7ULL, n);
    for (int i = 0; i < n; i++) EXPECT(got[i] % 7 == ref7[i]);
}

// argument errors are thrown before any launch (fft.cu:522-525)
static void case_errors()
{
    fft_configuration<TestDataType> c{};
    c.n_power = 25;
    c.fft_type = FORWARD;
    bool thrown = false;
    try {
        GPU_FFT<TestDataType>(reinterpret_cast<COMPLEX<TestDataType>*>(0x1000),
                              reinterpret_cast<COMPLEX<TestDataType>*>(0x2000), c, 1, false);
    } catch (const std::invalid_argument& e) {
        thrown = std::string(e.what()) == "Invalid n_power range!";
    }
    EXPECT(thrown);
    thrown = false;
    c.n_power = 1;
    try {
        GPU_FFNT<TestDataType>(reinterpret_cast<TestDataType*>(0x1000),
                               reinterpret_cast<COMPLEX<TestDataType>*>(0x2000),
                               reinterpret_cast<COMPLEX<TestDataType>*>(0x3000),
                               reinterpret_cast<COMPLEX<TestDataType>*>(0x4000), c, 1, false);
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    EXPECT(thrown);
    // tables of the host generators: sizes and a closed-form spot check
    FFT<TestDataType> f(8);
    auto rev = f.ReverseRootTable();
    EXPECT(rev.size() == 8);
    EXPECT(std::fabs(rev[1].real() - 0.0) < 1e-15 && std::fabs(rev[1].imag() - 1.0) < 1e-15);  // w_16^4 = i
    FFNT<TestDataType> g(16);
    EXPECT(g.twist_table_ffnt().size() == 8 && g.ReverseRootTable_ffnt().size() == 8);
}

// writes the host generators' tables (double and float) to `path`, in the
// order FFT<T>(n).Reverse, InverseReverse, FFNT<T>(2n).Reverse_ffnt,
// InverseReverse_ffnt, twist, untwist (T = double, then float)
template <typename T> static void dump(FILE* f, int n)
{
    FFT<T> a(n);
    FFNT<T> b(2 * n);
    std::vector<std::vector<COMPLEX<T>>> ts = {a.ReverseRootTable(), a.InverseReverseRootTable(),
                                               b.ReverseRootTable_ffnt(),
                                               b.InverseReverseRootTable_ffnt(),
                                               b.twist_table_ffnt(), b.untwist_table_ffnt()};
    for (auto& t : ts) std::fwrite(t.data(), sizeof(COMPLEX<T>), t.size(), f);
}

int main(int argc, char** argv)
{
    if (argc == 4 && std::string(argv[1]) == "tables") {
        FILE* f = std::fopen(argv[3], "wb");
        if (!f) return 2;
        dump<double>(f, std::atoi(argv[2]));
        dump<float>(f, std::atoi(argv[2]));
        std::fclose(f);
        return 0;
    }
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s case [logn] [batch]\n", argv[0]);
        return 2;
    }
    std::string c = argv[1];
    int logn = argc > 2 ? std::atoi(argv[2]) : 12;
    int batch = argc > 3 ? std::atoi(argv[3]) : 2;
    try {
        if (c == "cc") case_cc(logn, batch);
        else if (c == "rr") case_rr(logn, batch);
        else if (c == "ffnt") case_ffnt(logn, batch);
        else if (c == "special") case_special(logn, batch);
        else if (c == "cpu_fft") case_cpu_fft(logn);
        else if (c == "cpu_ffnt") case_cpu_ffnt(logn);
        else if (c == "errors") case_errors();
        else {
            std::fprintf(stderr, "unknown case %s\n", c.c_str());
            return 2;
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 3;
    }
    if (fails) return 1;
    std::printf("PASS %s logn=%d batch=%d\n", c.c_str(), logn, batch);
    return 0;
}
