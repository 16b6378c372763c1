"""Pins the CPU oracle (oracle/oracle.c) before any GPU result is compared
with it.  The reference ships no golden vectors, no known-answer tests and no
fixtures for this path, and its CPU code cannot be built here (it includes
cuda_runtime.h, which this image does not have), so every oracle function is
checked against what the mathematics fixes, chosen so that a plausible
mistake (a dropped term, a wrong sign, an un-reversed index, a transposed
operand) fails: numpy's FFT, numpy's convolution, brute-force polynomial
evaluation at roots of unity, closed forms of the tables, round trips, and
the committed golden fixtures of tests/golden/ (regression of the oracle)."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "oracle_golden.npz")


def br(n):
    return np.array([O.bitreverse(i, n) for i in range(1 << n)], dtype=np.int64)


def test_bitreverse_brute_force():
    for n in range(0, 12):
        for i in range(0, 1 << n, max(1, (1 << n) // 97)):
            s = format(i, "0%db" % n) if n else ""
            assert O.bitreverse(i, n) == (int(s[::-1], 2) if n else 0)


@pytest.mark.parametrize("n", [1, 2, 8, 1024])
def test_fft_tables_closed_form(n):
    t = O.fft_tables(n)
    k = np.arange(n)
    lg = n.bit_length() - 1
    ref = np.exp(2j * np.pi * br(lg) / (2 * n)) if n > 1 else np.ones(1)
    assert np.max(np.abs(t["rev"] - ref)) < 4e-16
    assert np.max(np.abs(t["inv_rev"] - np.conj(ref))) < 4e-16
    assert np.max(np.abs(t["root"] - np.exp(2j * np.pi * k / (2 * n)))) < 4e-16
    t32 = O.fft_tables(n, "f32")
    assert np.max(np.abs(t32["rev"] - ref)) < 1e-6  # float angle + cosf/sinf


@pytest.mark.parametrize("N", [4, 16, 4096])
def test_ffnt_tables_closed_form(N):
    t = O.ffnt_tables(N)
    L = N // 2
    lg = L.bit_length() - 1
    k = np.arange(L)
    ref = np.exp(2j * np.pi * br(lg - 1)[k % (L // 2)] / L)  # entries >= L/2 repeat
    assert np.max(np.abs(t["rev_root"] - ref)) < 4e-16
    assert np.max(np.abs(t["inv_rev_root"] - np.conj(ref))) < 4e-16
    assert np.max(np.abs(t["twist"] - np.exp(2j * np.pi * k / (2 * N)))) < 4e-16
    assert np.max(np.abs(t["untwist"] - np.exp(-2j * np.pi * k / (2 * N)))) < 4e-16


@pytest.mark.parametrize("l", [2, 8, 256])
def test_special_table_matches_docstring_psi_indexing(l):
    """fft.cuh:134-180: twiddle Psi[(5^j mod 4m) * (l/m)], Psi[k] = exp(pi i k/(2l)),
    for sub-array size m and j < m/2; the kernels read it at T[m/2 + j]."""
    T = O.special_table(l)
    for m in (2 ** e for e in range(1, l.bit_length())):
        for j in range(m // 2):
            k = (pow(5, j, 4 * m) * (l // m)) % (4 * l)
            assert abs(T[m // 2 + j] - np.exp(1j * np.pi * k / (2 * l))) < 8e-16
    assert np.max(np.abs(O.special_table(l, True) - np.conj(T))) == 0


@pytest.mark.parametrize("n_power", [1, 3, 7, 11])
def test_c2c_forward_is_bitreversed_positive_dft(n_power):
    """X[bitrev(k)] = sum_j x_j w^{+jk} (cyclic) and sum_j x_j psi^j w^{jk} (negacyclic)."""
    L = 1 << n_power
    rng = np.random.default_rng(n_power)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    cyc = O.c2c_forward(x, n_power, 1, O.fft_tables(max(1, L // 2))["rev"], 0)
    assert np.max(np.abs(cyc - (L * np.fft.ifft(x))[br(n_power)])) < 1e-12 * L
    psi = np.exp(1j * np.pi / L)
    neg = O.c2c_forward(x, n_power, 1, O.fft_tables(L)["rev"], 1)
    ref = (L * np.fft.ifft(x * psi ** np.arange(L)))[br(n_power)]
    assert np.max(np.abs(neg - ref)) < 1e-12 * L


def test_wrong_table_layout_is_caught():
    """A plausible mistake planted in the oracle's input (natural instead of
    bit-reversed root table) breaks the DFT identity of the test above."""
    n_power, L = 6, 64
    x = np.random.default_rng(1).standard_normal(L).astype(complex)
    bad = O.c2c_forward(x, n_power, 1, O.fft_tables(L // 2)["root"], 0)
    assert np.max(np.abs(bad - (L * np.fft.ifft(x))[br(n_power)])) > 1.0


@pytest.mark.parametrize("n_power", [2, 5, 10])
@pytest.mark.parametrize("nega", [0, 1])
def test_c2c_inverse_round_trip_and_product(n_power, nega):
    L = 1 << n_power
    t = O.fft_tables(L if nega else max(1, L // 2))
    rng = np.random.default_rng(10 + n_power)
    a = rng.integers(-50, 50, L)
    b = rng.integers(-50, 50, L)
    f = O.c2c_forward(np.concatenate([a, b]).astype(complex), n_power, 2, t["rev"], nega)
    back = O.c2c_inverse(f, n_power, 2, t["inv_rev"], nega, 0, complex(1.0 / L))
    assert np.max(np.abs(back - np.concatenate([a, b]))) < 1e-9
    p = O.c2c_inverse(f, n_power, 1, t["inv_rev"], nega, 1, complex(1.0 / L))[:L]
    exact = O.schoolbook_negacyclic(a, b) if nega else O.schoolbook_cyclic(a, b)
    assert np.array_equal(np.round(p.real).astype(np.int64), exact)


@pytest.mark.parametrize("n_power", [2, 6, 12])
def test_r2r_matches_c2c_on_real_data(n_power):
    L = 1 << n_power
    t = O.fft_tables(max(1, L // 2))
    x = np.random.default_rng(3).standard_normal(2 * L)
    f = O.r2r_forward(x, n_power, 2, t["rev"], 0)
    assert np.array_equal(f, O.c2c_forward(x.astype(complex), n_power, 2, t["rev"], 0))
    r = O.r2r_inverse(f, n_power, 1, t["inv_rev"], 0, 1, complex(1.0 / L))
    c = O.c2c_inverse(f, n_power, 1, t["inv_rev"], 0, 1, complex(1.0 / L))[:L]
    assert np.array_equal(r, c.real)


@pytest.mark.parametrize("n", [1, 4, 64, 1024])
def test_cpu_fft_mult_is_linear_product(n):
    """The reference's CPU plain-method product (FFT::fft / ifft of the
    zero-padded operands) equals numpy's convolution, batch-threaded or not."""
    rng = np.random.default_rng(60 + n)
    a = rng.integers(-1024, 1024, (3, n))
    b = rng.integers(-1024, 1024, (3, n))
    for threads in (1, 2):
        got = O.cpu_fft_mult(a, b, threads=threads)
        for i in range(3):
            ref = np.zeros(2 * n, np.int64)
            ref[:2 * n - 1] = np.convolve(a[i], b[i])
            assert np.array_equal(got[i], ref)


@pytest.mark.parametrize("n_power", [1, 3, 8])
@pytest.mark.parametrize("mult", [0, 1])
def test_r2r_inverse_general_temp(n_power, mult):
    """R_R inverse on arbitrary (non-Hermitian) temp with a complex
    mod_inverse: real_j = Re(mi * D_j), D_j = sum_k C_k w^{-jk} with
    C_k = temp[bitrev(k)] (times the second operand with mult); a dropped
    Im(mi) or Im(D) term fails here."""
    L = 1 << n_power
    t = O.fft_tables(max(1, L // 2))
    rng = np.random.default_rng(40 + n_power + 7 * mult)
    temp = rng.standard_normal(2 * L) + 1j * rng.standard_normal(2 * L)
    mi = complex(0.3, -0.7)
    nb = 1 if mult else 2
    got = O.r2r_inverse(temp, n_power, nb, t["inv_rev"], 0, mult, mi)
    for b in range(nb):
        C = temp[b * L:(b + 1) * L] * (temp[L:2 * L] if mult else 1)
        D = np.fft.fft(C[br(n_power)])
        ref = (mi * D).real
        assert np.max(np.abs(got[b * L:(b + 1) * L] - ref)) < 1e-12 * L
        assert np.max(np.abs(ref - mi.real * D.real)) > 1e-3  # the Im parts matter here


@pytest.mark.parametrize("n_power", [2, 4, 9, 13])
def test_ffnt_is_negacyclic_product(n_power):
    N = 1 << n_power
    L = N // 2
    t = O.ffnt_tables(N)
    rng = np.random.default_rng(20 + n_power)
    a = rng.integers(0, 1 << 20, N)
    b = rng.integers(-128, 129, N)
    f = O.ffnt_forward(np.concatenate([a, b]).astype(float), n_power, 2, t["twist"], t["rev_root"])
    # forward = cyclic DFT of the folded, twisted input, bit-reversed
    z = (a[:L] + 1j * a[L:]) * np.exp(2j * np.pi * np.arange(L) / (2 * N))
    assert np.max(np.abs(f[:L] - (L * np.fft.ifft(z))[br(n_power - 1)])) < 1e-9 * np.max(np.abs(f[:L]))
    out = O.ffnt_inverse(f, n_power, 1, t["untwist"], t["inv_rev_root"], 1, complex(1.0 / L))
    assert np.array_equal(np.round(out).astype(np.int64), O.schoolbook_negacyclic(a, b))


@pytest.mark.parametrize("n_power", [1, 2, 5, 9])
def test_special_forward_is_canonical_embedding(n_power):
    """forward(bitrev(z))_j = sum_k z_k zeta^{(5^j mod 4l) k}, zeta = exp(2 pi i / 4l)
    (evaluation of the slot polynomial at the primitive 4l-th roots 5^j)."""
    l = 1 << n_power
    M = 4 * l
    rng = np.random.default_rng(30 + n_power)
    z = rng.standard_normal(l) + 1j * rng.standard_normal(l)
    w = O.special_forward(z[br(n_power)], n_power, 1, O.special_table(l))
    zeta = np.exp(2j * np.pi / M)
    ref = np.array([np.sum(z * zeta ** ((pow(5, j, M) * np.arange(l)) % M)) for j in range(l)])
    assert np.max(np.abs(w - ref)) < 1e-12 * l
    back = O.special_inverse(w, n_power, 1, O.special_table(l, True), complex(1.0 / l))
    assert np.max(np.abs(back - z[br(n_power)])) < 1e-12


def test_schoolbook_against_numpy_convolve():
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 64):
        a, b = rng.integers(-1000, 1000, n), rng.integers(-1000, 1000, n)
        lin = np.convolve(a, b)
        assert np.array_equal(O.schoolbook_linear(a, b)[: 2 * n - 1], lin)
        padded = np.concatenate([lin, [0]])
        assert np.array_equal(O.schoolbook_negacyclic(a, b), padded[:n] - padded[n:])
        assert np.array_equal(O.schoolbook_cyclic(a, b), padded[:n] + padded[n:])


@pytest.mark.parametrize("n", [1, 2, 16, 512])
def test_ntt_exact_products(n):
    rng = np.random.default_rng(n)
    a = rng.integers(-(1 << 20), 1 << 20, (3, n))
    b = rng.integers(-(1 << 20), 1 << 20, (3, n))
    neg, cyc = O.ntt_negacyclic(a, b), O.ntt_cyclic(a, b)
    for i in range(3):
        assert np.array_equal(neg[i], O.schoolbook_negacyclic(a[i], b[i]))
        assert np.array_equal(cyc[i], O.schoolbook_cyclic(a[i], b[i]))


def test_cpu_fft_path():
    """FFT<double>::fft / ifft: the same transform as the GPU semantics, and a
    linear product through them is exact."""
    n = 256
    t = O.fft_tables(n)
    rng = np.random.default_rng(5)
    a = np.zeros(2 * n, complex)
    b = np.zeros(2 * n, complex)
    a[:n] = rng.integers(0, 7, n)
    b[:n] = rng.integers(0, 7, n)
    fa, fb = O.cpu_fft(a.copy(), t["root"]), O.cpu_fft(b.copy(), t["root"])
    assert np.array_equal(fa, O.c2c_forward(a, 9, 1, t["rev"], 0))
    p = O.cpu_ifft(fa * fb, t["inv"])
    assert np.array_equal(np.round(p.real).astype(np.int64),
                          O.schoolbook_linear(a[:n].real.astype(np.int64), b[:n].real.astype(np.int64)))
    assert np.max(np.abs(O.cpu_ifft(fa.copy(), t["inv"]) - a)) < 1e-12


def test_cpu_ffnt_mult():
    N = 1024
    rng = np.random.default_rng(6)
    a = rng.integers(0, 1 << 20, (4, N))
    b = rng.integers(-128, 129, (4, N))
    got = O.cpu_ffnt_mult(a, b, threads=2)
    assert np.array_equal(got, O.ntt_negacyclic(a, b))


def test_fill_uniform():
    x = O.fill_uniform(123, -128, 129, 200000)
    assert x.min() == -128 and x.max() == 128
    assert abs(x.mean()) < 1.0
    assert np.array_equal(x[:100], O.fill_uniform(123, -128, 129, 100))
    assert not np.array_equal(x[:100], O.fill_uniform(124, -128, 129, 100))


def test_golden_fixtures_regenerate_identically():
    """The committed fixtures (tests/golden/make_golden.py) are reproduced bit
    for bit by the current oracle, and their products are exact."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fresh = mg.build()
    with np.load(GOLDEN, allow_pickle=False) as g:
        assert sorted(g.files) == sorted(fresh)
        for k in g.files:
            assert g[k].dtype == fresh[k].dtype and np.array_equal(g[k], fresh[k]), k
        assert np.array_equal(g["c1_product"], O.schoolbook_linear(g["c1_a"], g["c1_b"]))
        assert np.array_equal(g["ffnt_product"], O.schoolbook_negacyclic(g["ffnt_a"], g["ffnt_b"]))


@pytest.mark.parametrize("n_power", [1, 2, 5, 8])
def test_ckks_oracle_is_canonical_embedding(n_power):
    """The oracle's CKKS decode is the canonical embedding at the roots
    zeta_2N^(5^j mod 2N) (brute force), and its encode inverts it exactly on
    integer coefficient vectors."""
    l, delta = 1 << n_power, 2.0 ** 20
    N = 2 * l
    rng = np.random.default_rng(900 + n_power)
    m = rng.integers(-2 ** 35, 2 ** 35, size=2 * l)
    z = O.ckks_decode(m, n_power, 1, O.special_table(l), delta)
    e = np.array([pow(5, j, 2 * N) for j in range(l)])
    ref = np.exp(2j * np.pi * np.outer(e, np.arange(2 * l)) / (2 * N)) @ m.astype(float) / delta
    assert np.max(np.abs(z - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.array_equal(O.ckks_encode(ref, n_power, 1, O.special_table(l, True), delta), m)
    # a planted mistake (natural instead of bit-reversed slot order) is caught
    zb = ref[[O.bitreverse(i, n_power) for i in range(l)]] if l > 2 else ref[::-1]
    assert not np.array_equal(O.ckks_encode(zb, n_power, 1, O.special_table(l, True), delta), m)
