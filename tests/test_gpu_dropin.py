"""GPU parity of the drop-in entry points against the oracle.

Each test runs the HIP library through its C ABI (gpufft_amd ctypes binding)
and compares with the oracle's restatement of the reference semantics
(oracle/oracle.c), on seeded inputs, with the reference's table layouts:
floating-point outputs within helpers.TOL, rounded integer products
bit-exact against exact integer products.
"""
import numpy as np
import pytest

import oracle as O
from helpers import TOL, cdtype, rdtype, rel_err, rng_complex, to_dev, to_host

pytestmark = pytest.mark.gpu

PRECS = ["f64", "f32"]
# 1-pass and multi-pass sizes for both tile widths (2^12 double, 2^13 float)
SIZES = [1, 2, 3, 5, 8, 11, 12, 13, 14, 16, 17, 20]


def _g():
    import gpufft_amd as G
    return G


def _tables(n_power, nega, prec):
    t = O.fft_tables((1 << n_power) if nega else (1 << (n_power - 1)) if n_power > 0 else 1, prec)
    return t


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("nega", [0, 1])
@pytest.mark.parametrize("n_power", SIZES)
def test_c2c_forward_inverse(prec, nega, n_power):
    G = _g()
    rng = np.random.default_rng(1000 + n_power * 4 + nega * 2 + (prec == "f32"))
    batch = 3
    L = 1 << n_power
    tabs = _tables(n_power, nega, prec)
    x = rng_complex(rng, 2 * batch * L, prec)
    d = to_dev(x)
    rp = G.X_N_PLUS if nega else G.X_N_MINUS
    G.GPU_FFT(d, to_dev(tabs["rev"]), G.config(n_power, G.FORWARD, rp), 2 * batch)
    got = to_host(d)
    ref = O.c2c_forward(x, n_power, 2 * batch, tabs["rev"], nega, prec)
    assert rel_err(got, ref) < TOL[prec]
    # inverse with pointwise product, on the oracle's forward output
    d2 = to_dev(ref)
    scale = 1.0 / L
    G.GPU_FFT(d2, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, rp, scale), batch, True)
    got2 = to_host(d2)
    ref2 = O.c2c_inverse(ref, n_power, batch, tabs["inv_rev"], nega, 1, complex(scale), prec)
    assert rel_err(got2[: batch * L], ref2[: batch * L]) < TOL[prec]
    # operand B is untouched
    assert np.array_equal(got2[batch * L:], ref[batch * L:])


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_power", [1, 4, 12, 13, 15])
def test_c2c_inverse_no_mult(prec, n_power):
    G = _g()
    rng = np.random.default_rng(77 + n_power)
    batch, L = 2, 1 << n_power
    tabs = _tables(n_power, 0, prec)
    x = rng_complex(rng, batch * L, prec)
    d = to_dev(x)
    G.GPU_FFT(d, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, G.X_N_MINUS, 0.5 - 0.25j),
              batch, False)
    ref = O.c2c_inverse(x, n_power, batch, tabs["inv_rev"], 0, 0, complex(0.5, -0.25), prec)
    assert rel_err(to_host(d), ref) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("nega", [0, 1])
@pytest.mark.parametrize("n_power", [2, 7, 12, 13, 14, 16, 18])
def test_r2r_forward_inverse(prec, nega, n_power):
    G = _g()
    import torch
    rng = np.random.default_rng(2000 + n_power * 2 + nega)
    batch, L = 2, 1 << n_power
    tabs = _tables(n_power, nega, prec)
    real = rng.standard_normal(2 * batch * L).astype(rdtype(prec))
    dr = to_dev(real)
    dt = torch.zeros(2 * batch * L, dtype=torch.complex128 if prec == "f64" else torch.complex64,
                     device="cuda")
    rp = G.X_N_PLUS if nega else G.X_N_MINUS
    G.GPU_FFT_R(dr, dt, to_dev(tabs["rev"]), G.config(n_power, G.FORWARD, rp), 2 * batch)
    ref = O.r2r_forward(real, n_power, 2 * batch, tabs["rev"], nega, prec)
    assert rel_err(to_host(dt), ref) < TOL[prec]
    assert np.array_equal(to_host(dr), real)  # the real buffer is left alone
    dt2 = to_dev(ref)
    G.GPU_FFT_R(dr, dt2, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, rp, 1.0 / L), batch,
                True)
    ref2 = O.r2r_inverse(ref, n_power, batch, tabs["inv_rev"], nega, 1, complex(1.0 / L), prec)
    assert rel_err(to_host(dr)[: batch * L], ref2) < TOL[prec]
    # inverse without product, complex mod_inverse (the output is Re(mod_inverse * D)),
    # all 2 batch spectra: the round trip returns the input scaled by Re(mod_inverse) L
    dt3 = to_dev(ref)
    mi = complex(1.0 / L, 0.37 / L)
    G.GPU_FFT_R(dr, dt3, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, rp, mi), 2 * batch)
    ref3 = O.r2r_inverse(ref, n_power, 2 * batch, tabs["inv_rev"], nega, 0, mi, prec)
    assert rel_err(to_host(dr), ref3) < TOL[prec]
    assert rel_err(to_host(dr), real) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_power", [2, 3, 8, 12, 13, 14, 15, 16, 17, 19])
def test_ffnt_forward_inverse(prec, n_power):
    G = _g()
    import torch
    rng = np.random.default_rng(3000 + n_power)
    batch, N = 2, 1 << n_power
    L = N // 2
    t = O.ffnt_tables(N, prec)
    real = rng.standard_normal(2 * batch * N).astype(rdtype(prec))
    dr = to_dev(real)
    dt = torch.zeros(2 * batch * L, dtype=torch.complex128 if prec == "f64" else torch.complex64,
                     device="cuda")
    G.GPU_FFNT(dr, dt, to_dev(t["twist"]), to_dev(t["rev_root"]), G.config(n_power, G.FORWARD),
               2 * batch)
    ref = O.ffnt_forward(real, n_power, 2 * batch, t["twist"], t["rev_root"], prec)
    assert rel_err(to_host(dt), ref) < TOL[prec]
    dt2 = to_dev(ref)
    G.GPU_FFNT(dr, dt2, to_dev(t["untwist"]), to_dev(t["inv_rev_root"]),
               G.config(n_power, G.INVERSE, mod_inverse=1.0 / L), batch, True)
    ref2 = O.ffnt_inverse(ref, n_power, batch, t["untwist"], t["inv_rev_root"], 1, complex(1.0 / L),
                          prec)
    assert rel_err(to_host(dr)[: batch * N], ref2) < TOL[prec]


# sizes on the one-launch forward / two-launch inverse: 2^(S+1)..2^(S+3)
# points, S = 13 (double) or 14 (float); R_R transforms 2^n_power points,
# FFNT 2^(n_power-1)
ONESHOT = [("f64", "r2r", 14), ("f64", "r2r", 16), ("f64", "ffnt", 15), ("f64", "ffnt", 17),
           ("f32", "r2r", 15), ("f32", "r2r", 17), ("f32", "ffnt", 16), ("f32", "ffnt", 18)]


@pytest.mark.parametrize("prec,kind,n_power", ONESHOT)
def test_oneshot_odd_batches(prec, kind, n_power):
    """Batches of 13 and 21 transforms on the one-launch paths: their grids
    deal the workgroups of a transform 8 transforms at a time (block ids
    congruent mod 8), so a batch that is not a multiple of 8 leaves a partial
    group, and one above 8 spans several; forward and inverse with product
    against the oracle."""
    G = _g()
    import torch
    rng = np.random.default_rng(4000 + n_power * 4 + (kind == "ffnt") * 2 + (prec == "f32"))
    for batch in (13, 21):
        if kind == "r2r":
            N = L = 1 << n_power
            tabs = _tables(n_power, 0, prec)
            ftab, itab = tabs["rev"], tabs["inv_rev"]
        else:
            N = 1 << n_power
            L = N // 2
            t = O.ffnt_tables(N, prec)
            ftab, itab = t["rev_root"], t["inv_rev_root"]
        real = rng.standard_normal(2 * batch * N).astype(rdtype(prec))
        dr = to_dev(real)
        dt = torch.zeros(2 * batch * L, dtype=torch.complex128 if prec == "f64" else torch.complex64,
                         device="cuda")
        if kind == "r2r":
            G.GPU_FFT_R(dr, dt, to_dev(ftab), G.config(n_power, G.FORWARD, G.X_N_MINUS), 2 * batch)
            ref = O.r2r_forward(real, n_power, 2 * batch, ftab, 0, prec)
        else:
            G.GPU_FFNT(dr, dt, to_dev(t["twist"]), to_dev(ftab), G.config(n_power, G.FORWARD),
                       2 * batch)
            ref = O.ffnt_forward(real, n_power, 2 * batch, t["twist"], ftab, prec)
        assert rel_err(to_host(dt), ref) < TOL[prec], batch
        dt2 = to_dev(ref)
        if kind == "r2r":
            G.GPU_FFT_R(dr, dt2, to_dev(itab), G.config(n_power, G.INVERSE, G.X_N_MINUS, 1.0 / L),
                        batch, True)
            ref2 = O.r2r_inverse(ref, n_power, batch, itab, 0, 1, complex(1.0 / L), prec)
        else:
            G.GPU_FFNT(dr, dt2, to_dev(t["untwist"]), to_dev(itab),
                       G.config(n_power, G.INVERSE, mod_inverse=1.0 / L), batch, True)
            ref2 = O.ffnt_inverse(ref, n_power, batch, t["untwist"], itab, 1, complex(1.0 / L), prec)
        assert rel_err(to_host(dr)[: batch * N], ref2) < TOL[prec], batch
        # the second half of the real buffer (b) is left alone by the inverse
        assert np.array_equal(to_host(dr)[batch * N:], real[batch * N:]), batch


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_power", [1, 2, 5, 11, 12, 13, 14, 15, 17])
def test_special_forward_inverse(prec, n_power):
    G = _g()
    rng = np.random.default_rng(4000 + n_power)
    batch, l = 3, 1 << n_power
    T = O.special_table(l, False, prec)
    Ti = O.special_table(l, True, prec)
    z = rng_complex(rng, batch * l, prec)
    d = to_dev(z)
    G.GPU_Special_FFT(d, to_dev(T), G.config(n_power, G.FORWARD), batch)
    ref = O.special_forward(z, n_power, batch, T, prec)
    assert rel_err(to_host(d), ref) < TOL[prec] * (4 if prec == "f32" else 1)
    d2 = to_dev(ref)
    G.GPU_Special_FFT(d2, to_dev(Ti), G.config(n_power, G.INVERSE, mod_inverse=1.0 / l), batch)
    ref2 = O.special_inverse(ref, n_power, batch, Ti, complex(1.0 / l), prec)
    assert rel_err(to_host(d2), ref2) < TOL[prec] * (4 if prec == "f32" else 1)
    # encode -> decode round trip returns the input
    assert rel_err(to_host(d2), z) < TOL[prec] * (8 if prec == "f32" else 100)


# ---- integer products through the drop-in calls, rounded, bit-exact ----
def _cc_product(G, a, b, n, prec):
    """The reference example flow (example/test_fft_gpu_C_C.cu:53-183)."""
    batch = a.shape[0]
    L = 2 * n
    n_power = L.bit_length() - 1
    t = O.fft_tables(n, prec)
    host = np.zeros((2 * batch, L), cdtype(prec))
    host[:batch, :n] = a
    host[batch:, :n] = b
    d = to_dev(host.reshape(-1))
    G.GPU_FFT(d, to_dev(t["rev"]), G.config(n_power, G.FORWARD, G.X_N_MINUS), 2 * batch)
    G.GPU_FFT(d, to_dev(t["inv_rev"]), G.config(n_power, G.INVERSE, G.X_N_MINUS, t["n_inverse"]),
              batch, True)
    out = to_host(d).reshape(2 * batch, L)[:batch].real
    return np.round(out).astype(np.int64)


@pytest.mark.parametrize("logn", [1, 5, 11, 12])
def test_c2c_linear_product_exact(logn):
    """Config-1 flow: C_C cyclic product of zero-padded inputs in [0,16)."""
    G = _g()
    n = 1 << logn
    batch = 2
    a = np.stack([O.fill_uniform(0x6770756666740001 + 2 * i, 0, 16, n) for i in range(batch)])
    b = np.stack([O.fill_uniform(0x6770756666740001 + 2 * i + 1, 0, 16, n) for i in range(batch)])
    got = _cc_product(G, a, b, n, "f64")
    for i in range(batch):
        assert np.array_equal(got[i], O.schoolbook_linear(a[i], b[i]))


@pytest.mark.parametrize("n_power", [4, 12, 16])
def test_ffnt_product_exact(n_power):
    """Reference example/test_ffnt_gpu_R_R.cu flow, batch > 1, signed inputs."""
    G = _g()
    import torch
    N = 1 << n_power
    L = N // 2
    batch = 3
    a = O.fill_uniform(0x6770756666740003, 0, 1 << 20, batch * N).reshape(batch, N)
    b = O.fill_uniform(0x6770756666740013, -128, 129, batch * N).reshape(batch, N)
    t = O.ffnt_tables(N)
    dr = to_dev(np.concatenate([a, b]).astype(np.float64).reshape(-1))
    dt = torch.zeros(2 * batch * L, dtype=torch.complex128, device="cuda")
    G.GPU_FFNT(dr, dt, to_dev(t["twist"]), to_dev(t["rev_root"]), G.config(n_power, G.FORWARD),
               2 * batch)
    G.GPU_FFNT(dr, dt, to_dev(t["untwist"]), to_dev(t["inv_rev_root"]),
               G.config(n_power, G.INVERSE, mod_inverse=1.0 / L), batch, True)
    got = np.round(to_host(dr).reshape(2 * batch, N)[:batch]).astype(np.int64)
    exact = O.ntt_negacyclic(a, b)
    assert np.array_equal(got, exact)


@pytest.mark.parametrize("n_power", [3, 12, 16])
def test_r2r_cyclic_product_exact(n_power):
    """Reference example/test_fft_gpu_R_R.cu flow (zero-padded cyclic)."""
    G = _g()
    import torch
    L = 1 << n_power
    n = L // 2
    batch = 2
    a = O.fill_uniform(0x6770756666740002, -1024, 1024, batch * n).reshape(batch, n)
    b = O.fill_uniform(0x6770756666740012, -1024, 1024, batch * n).reshape(batch, n)
    host = np.zeros((2 * batch, L))
    host[:batch, :n] = a
    host[batch:, :n] = b
    t = O.fft_tables(n)
    dr = to_dev(host.reshape(-1))
    dt = torch.zeros(2 * batch * L, dtype=torch.complex128, device="cuda")
    G.GPU_FFT_R(dr, dt, to_dev(t["rev"]), G.config(n_power, G.FORWARD, G.X_N_MINUS), 2 * batch)
    G.GPU_FFT_R(dr, dt, to_dev(t["inv_rev"]),
                G.config(n_power, G.INVERSE, G.X_N_MINUS, t["n_inverse"]), batch, True)
    got = np.round(to_host(dr).reshape(2 * batch, L)[:batch]).astype(np.int64)
    for i in range(batch):
        assert np.array_equal(got[i], O.schoolbook_linear(a[i], b[i]))


@pytest.mark.parametrize("n_power", [3, 13])
def test_c2c_negacyclic_product_exact(n_power):
    """X_N_plus through GPU_FFT: FFT(N) tables, mod_inverse = 1/N."""
    G = _g()
    N = 1 << n_power
    batch = 2
    a = O.fill_uniform(11, -500, 500, batch * N).reshape(batch, N)
    b = O.fill_uniform(12, -500, 500, batch * N).reshape(batch, N)
    t = O.fft_tables(N)
    d = to_dev(np.concatenate([a, b]).astype(np.complex128).reshape(-1))
    G.GPU_FFT(d, to_dev(t["rev"]), G.config(n_power, G.FORWARD, G.X_N_PLUS), 2 * batch)
    G.GPU_FFT(d, to_dev(t["inv_rev"]), G.config(n_power, G.INVERSE, G.X_N_PLUS, 1.0 / N), batch, True)
    got = np.round(to_host(d).reshape(2 * batch, N)[:batch].real).astype(np.int64)
    assert np.array_equal(got, O.ntt_negacyclic(a, b))


# ---- the top of the drop-in range (n_power 24: 2^24 points, four passes;
# float splits into three passes from n_power 23) ----
@pytest.mark.parametrize("prec,n_power", [("f64", 24), ("f32", 23), ("f32", 24)])
@pytest.mark.parametrize("kind", ["c2c", "c2c_nega", "r2r", "ffnt", "special"])
def test_dropin_max_size(kind, prec, n_power):
    """Every drop-in entry point at the top of the reference's range
    (n_power 24, fft.cu:522-525), batch 1, FP64 and FP32, against the
    oracle."""
    G = _g()
    import torch
    rng = np.random.default_rng(7700 + len(kind) + n_power)
    L = 1 << n_power
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    tol = TOL[prec]
    if kind in ("c2c", "c2c_nega"):
        nega = kind == "c2c_nega"
        tabs = _tables(n_power, nega, prec)
        x = rng_complex(rng, L, prec)
        d = to_dev(x)
        rp = G.X_N_PLUS if nega else G.X_N_MINUS
        G.GPU_FFT(d, to_dev(tabs["rev"]), G.config(n_power, G.FORWARD, rp), 1)
        ref = O.c2c_forward(x, n_power, 1, tabs["rev"], nega, prec)
        assert rel_err(to_host(d), ref) < tol
        G.GPU_FFT(d, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, rp, 1.0 / L), 1)
        assert rel_err(to_host(d), x) < 100 * tol
    elif kind == "r2r":
        tabs = _tables(n_power, 0, prec)
        real = rng.standard_normal(L).astype(rdtype(prec))
        dr = to_dev(real)
        dt = torch.zeros(L, dtype=cdt, device="cuda")
        G.GPU_FFT_R(dr, dt, to_dev(tabs["rev"]), G.config(n_power, G.FORWARD), 1)
        ref = O.r2r_forward(real, n_power, 1, tabs["rev"], 0, prec)
        assert rel_err(to_host(dt), ref) < tol
        G.GPU_FFT_R(dr, dt, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, mod_inverse=1.0 / L), 1)
        assert rel_err(to_host(dr), real) < 100 * tol
    elif kind == "ffnt":
        N, Lh = L, L // 2
        t = O.ffnt_tables(N, prec)
        real = rng.standard_normal(N).astype(rdtype(prec))
        dr = to_dev(real)
        dt = torch.zeros(Lh, dtype=cdt, device="cuda")
        G.GPU_FFNT(dr, dt, to_dev(t["twist"]), to_dev(t["rev_root"]), G.config(n_power, G.FORWARD), 1)
        ref = O.ffnt_forward(real, n_power, 1, t["twist"], t["rev_root"], prec)
        assert rel_err(to_host(dt), ref) < tol
        G.GPU_FFNT(dr, dt, to_dev(t["untwist"]), to_dev(t["inv_rev_root"]),
                   G.config(n_power, G.INVERSE, mod_inverse=1.0 / Lh), 1)
        assert rel_err(to_host(dr), real) < 100 * tol
    else:
        T, Ti = O.special_table(L, False, prec), O.special_table(L, True, prec)
        z = rng_complex(rng, L, prec)
        d = to_dev(z)
        G.GPU_Special_FFT(d, to_dev(T), G.config(n_power, G.FORWARD), 1)
        ref = O.special_forward(z, n_power, 1, T, prec)
        assert rel_err(to_host(d), ref) < 4 * tol
        G.GPU_Special_FFT(d, to_dev(Ti), G.config(n_power, G.INVERSE, mod_inverse=1.0 / L), 1)
        assert rel_err(to_host(d), z) < 100 * tol


# ---- R_R inverse on arbitrary temp (reference fft.cu:499-506: the output is
# Re(mod_inverse * D) for any temp, D the complex GS inverse) ----
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("mult", [0, 1])
@pytest.mark.parametrize("n_power", [3, 12, 16])
def test_r2r_inverse_non_hermitian(prec, mult, n_power):
    """temp holds arbitrary complex values (not the spectrum of real data)
    and mod_inverse is complex: the drop-in R_R inverse must still give the
    reference's Re(mod_inverse * D_j)."""
    G = _g()
    rng = np.random.default_rng(8100 + n_power * 4 + mult * 2 + (prec == "f32"))
    batch, L = 2, 1 << n_power
    tabs = _tables(n_power, 0, prec)
    temp = rng_complex(rng, 2 * batch * L, prec)
    mi = complex(1.0 / L, -0.61 / L)
    nb = batch if mult else 2 * batch
    dr = to_dev(np.zeros(2 * batch * L, rdtype(prec)))
    dt = to_dev(temp)
    G.GPU_FFT_R(dr, dt, to_dev(tabs["inv_rev"]), G.config(n_power, G.INVERSE, G.X_N_MINUS, mi),
                nb, bool(mult))
    ref = O.r2r_inverse(temp, n_power, nb, tabs["inv_rev"], 0, mult, mi, prec)
    assert rel_err(to_host(dr)[: nb * L], ref) < TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_power", [2, 3, 12, 16, 17])
def test_r2r_inverse_hermitian_entry(prec, n_power):
    """The opt-in packed inverse equals the general one on spectra of real
    data (its precondition), with and without the product."""
    G = _g()
    import torch
    rng = np.random.default_rng(8300 + n_power * 2 + (prec == "f32"))
    batch, L = 2, 1 << n_power
    tabs = _tables(n_power, 0, prec)
    real = rng.standard_normal(2 * batch * L).astype(rdtype(prec))
    spec = O.r2r_forward(real, n_power, 2 * batch, tabs["rev"], 0, prec)
    for mult, nb in ((1, batch), (0, 2 * batch)):
        dr = torch.zeros(2 * batch * L, dtype=torch.float64 if prec == "f64" else torch.float32,
                         device="cuda")
        G.GPU_FFT_R_inverse_hermitian(dr, to_dev(spec), to_dev(tabs["inv_rev"]),
                                      G.config(n_power, G.INVERSE, G.X_N_MINUS, 1.0 / L), nb,
                                      bool(mult))
        ref = O.r2r_inverse(spec, n_power, nb, tabs["inv_rev"], 0, mult, complex(1.0 / L), prec)
        assert rel_err(to_host(dr)[: nb * L], ref) < TOL[prec]
