"""CPU oracle for the GPU-FFT polynomial-multiplication path (numpy front end).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, as the checker. Nothing in the product
(gpu-fft_amd/) imports it.

It wraps liboracle.so (oracle.c, plain C built with gcc), which restates the
reference's algorithm; each C function cites the reference file:line it
follows. Complex arrays are numpy complex128/complex64, whose memory layout is
the reference's interleaved COMPLEX<T> (src/include/gpufft/complex.cuh:14-16).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_F = ctypes.c_float


def _declare(L):
    L.orc_bitreverse.argtypes = [_I, _I]
    L.orc_bitreverse.restype = _I
    for s in ("f64", "f32"):
        getattr(L, "orc_fft_tables_" + s).argtypes = [_I, _P, _P, _P, _P]
        getattr(L, "orc_ffnt_tables_" + s).argtypes = [_I, _P, _P, _P, _P]
    L.orc_special_table_f64.argtypes = [_I, _I, _P]
    L.orc_cpu_fft_f64.argtypes = [_I, _P, _P]
    L.orc_cpu_ifft_f64.argtypes = [_I, _P, _P]
    L.orc_cpu_ffnt_mult_f64.argtypes = [_I, _P, _P, _P]
    L.orc_cpu_ffnt_mult_batch_f64.argtypes = [_I, _I, _P, _P, _P, _I]
    L.orc_cpu_fft_mult_batch_f64.argtypes = [_I, _I, _P, _P, _P, _I]
    for s, T in (("f64", _D), ("f32", _F)):
        getattr(L, "orc_c2c_forward_" + s).argtypes = [_P, _I, _I, _P, _I]
        getattr(L, "orc_c2c_inverse_" + s).argtypes = [_P, _I, _I, _P, _I, _I, T, T]
        getattr(L, "orc_r2r_forward_" + s).argtypes = [_P, _P, _I, _I, _P, _I]
        getattr(L, "orc_r2r_inverse_" + s).argtypes = [_P, _P, _I, _I, _P, _I, _I, T, T]
        getattr(L, "orc_ffnt_forward_" + s).argtypes = [_P, _P, _I, _I, _P, _P]
        getattr(L, "orc_ffnt_inverse_" + s).argtypes = [_P, _P, _I, _I, _P, _P, _I, T, T]
        getattr(L, "orc_special_forward_" + s).argtypes = [_P, _I, _I, _P]
        getattr(L, "orc_special_inverse_" + s).argtypes = [_P, _I, _I, _P, T, T]
    for name in ("orc_schoolbook_linear", "orc_schoolbook_negacyclic", "orc_schoolbook_cyclic"):
        getattr(L, name).argtypes = [_I, _P, _P, _P]
    L.orc_ckks_encode_f64.argtypes = [_P, _I, _I, _P, ctypes.c_double, _P]
    L.orc_ckks_decode_f64.argtypes = [_P, _I, _I, _P, ctypes.c_double, _P]
    L.orc_ntt_negacyclic.argtypes = [_I, _I, _P, _P, _P]
    L.orc_ntt_cyclic.argtypes = [_I, _I, _P, _P, _P]
    L.orc_mix64.argtypes = [ctypes.c_uint64]
    L.orc_mix64.restype = ctypes.c_uint64
    L.orc_fill_uniform_i64.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                       ctypes.c_int64, _P]


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be contiguous"
    return a.ctypes.data


def _cdt(prec):
    return np.complex128 if prec == "f64" else np.complex64


def _rdt(prec):
    return np.float64 if prec == "f64" else np.float32


def bitreverse(index, n_power):
    return lib().orc_bitreverse(index, n_power)


# ---------------------------------------------------------------- tables
def fft_tables(n, prec="f64"):
    """FFT<T>(n) tables: dict(root, inv, rev, inv_rev, n_inverse)
    (reference src/lib/fft_cpu.cu:93-129,200-226)."""
    d = {k: np.zeros(n, _cdt(prec)) for k in ("root", "inv", "rev", "inv_rev")}
    getattr(lib(), "orc_fft_tables_" + prec)(n, _ptr(d["root"]), _ptr(d["inv"]),
                                             _ptr(d["rev"]), _ptr(d["inv_rev"]))
    d["n_inverse"] = 1.0 / (2 * n)
    return d


def ffnt_tables(N, prec="f64"):
    """FFNT<T>(N) tables: dict(rev_root, inv_rev_root, twist, untwist), N/2 each
    (reference src/lib/fft_cpu.cu:234-299,434-470)."""
    L = N // 2
    d = {k: np.zeros(L, _cdt(prec)) for k in ("rev_root", "inv_rev_root", "twist", "untwist")}
    getattr(lib(), "orc_ffnt_tables_" + prec)(N, _ptr(d["rev_root"]), _ptr(d["inv_rev_root"]),
                                              _ptr(d["twist"]), _ptr(d["untwist"]))
    return d


def special_table(l, inverse=False, prec="f64"):
    t = np.zeros(l, np.complex128)
    lib().orc_special_table_f64(l, 1 if inverse else 0, _ptr(t))
    return t if prec == "f64" else t.astype(np.complex64)


# ---------------------------------------------------------------- CPU path
def cpu_fft(data, root):
    """FFT<double>::fft in place on a length-2n complex128 vector."""
    n = len(root)
    assert data.dtype == np.complex128 and len(data) == 2 * n
    lib().orc_cpu_fft_f64(n, _ptr(root), _ptr(data))
    return data


def cpu_ifft(data, inv_root):
    n = len(inv_root)
    assert data.dtype == np.complex128 and len(data) == 2 * n
    lib().orc_cpu_ifft_f64(n, _ptr(inv_root), _ptr(data))
    return data


def cpu_fft_mult(a, b, threads=1):
    """The reference's CPU plain-method product (example/test_fft_cpu.cu:65-142)
    for rows of a, b (int64, shape [count, n]): linear products, [count, 2n]."""
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    if a.ndim == 1:
        a, b = a[None], b[None]
    out = np.zeros((a.shape[0], 2 * a.shape[1]), np.int64)
    lib().orc_cpu_fft_mult_batch_f64(a.shape[1], a.shape[0], _ptr(a), _ptr(b), _ptr(out), threads)
    return out


def cpu_ffnt_mult(a, b, threads=1):
    """FFNT<double>::MULT for rows of a, b (int64, shape [count, N])."""
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    if a.ndim == 1:
        a, b = a[None], b[None]
    out = np.zeros_like(a)
    lib().orc_cpu_ffnt_mult_batch_f64(a.shape[1], a.shape[0], _ptr(a), _ptr(b), _ptr(out), threads)
    return out


# ---------------------------------------------------------------- GPU semantics
def c2c_forward(data, n_power, batch, table, nega, prec="f64"):
    data = np.ascontiguousarray(data, _cdt(prec)).copy()
    getattr(lib(), "orc_c2c_forward_" + prec)(_ptr(data), n_power, batch,
                                              _ptr(np.ascontiguousarray(table, _cdt(prec))), int(nega))
    return data


def c2c_inverse(data, n_power, batch, table, nega, mult, scale, prec="f64"):
    data = np.ascontiguousarray(data, _cdt(prec)).copy()
    getattr(lib(), "orc_c2c_inverse_" + prec)(_ptr(data), n_power, batch,
                                              _ptr(np.ascontiguousarray(table, _cdt(prec))),
                                              int(nega), int(mult), scale.real, scale.imag)
    return data


def r2r_forward(real, n_power, batch, table, nega, prec="f64"):
    real = np.ascontiguousarray(real, _rdt(prec))
    temp = np.zeros(batch << n_power, _cdt(prec))
    getattr(lib(), "orc_r2r_forward_" + prec)(_ptr(real), _ptr(temp), n_power, batch,
                                              _ptr(np.ascontiguousarray(table, _cdt(prec))), int(nega))
    return temp


def r2r_inverse(temp, n_power, batch, table, nega, mult, scale, prec="f64"):
    temp = np.ascontiguousarray(temp, _cdt(prec)).copy()
    real = np.zeros(batch << n_power, _rdt(prec))
    getattr(lib(), "orc_r2r_inverse_" + prec)(_ptr(real), _ptr(temp), n_power, batch,
                                              _ptr(np.ascontiguousarray(table, _cdt(prec))),
                                              int(nega), int(mult), scale.real, scale.imag)
    return real


def ffnt_forward(real, n_power, batch, twist, table, prec="f64"):
    real = np.ascontiguousarray(real, _rdt(prec))
    temp = np.zeros(batch << (n_power - 1), _cdt(prec))
    getattr(lib(), "orc_ffnt_forward_" + prec)(_ptr(real), _ptr(temp), n_power, batch,
                                               _ptr(np.ascontiguousarray(twist, _cdt(prec))),
                                               _ptr(np.ascontiguousarray(table, _cdt(prec))))
    return temp


def ffnt_inverse(temp, n_power, batch, untwist, table, mult, scale, prec="f64"):
    temp = np.ascontiguousarray(temp, _cdt(prec)).copy()
    out = np.zeros(batch << n_power, _rdt(prec))
    getattr(lib(), "orc_ffnt_inverse_" + prec)(_ptr(out), _ptr(temp), n_power, batch,
                                               _ptr(np.ascontiguousarray(untwist, _cdt(prec))),
                                               _ptr(np.ascontiguousarray(table, _cdt(prec))),
                                               int(mult), scale.real, scale.imag)
    return out


def special_forward(data, n_power, batch, table, prec="f64"):
    data = np.ascontiguousarray(data, _cdt(prec)).copy()
    getattr(lib(), "orc_special_forward_" + prec)(_ptr(data), n_power, batch,
                                                  _ptr(np.ascontiguousarray(table, _cdt(prec))))
    return data


def ckks_encode(slots, n_power, batch, inv_table, delta):
    """int64 [batch][2l] coefficients of the slots (oracle.h: CKKS encode)."""
    z = np.ascontiguousarray(slots, np.complex128)
    out = np.empty(batch * (2 << n_power), np.int64)
    lib().orc_ckks_encode_f64(_ptr(z), n_power, batch,
                              _ptr(np.ascontiguousarray(inv_table, np.complex128)),
                              ctypes.c_double(delta), _ptr(out))
    return out


def ckks_decode(coeffs, n_power, batch, table, delta):
    """complex [batch][l] slots of int64 coefficients (oracle.h: CKKS decode)."""
    m = np.ascontiguousarray(coeffs, np.int64)
    out = np.empty(batch << n_power, np.complex128)
    lib().orc_ckks_decode_f64(_ptr(m), n_power, batch,
                              _ptr(np.ascontiguousarray(table, np.complex128)),
                              ctypes.c_double(delta), _ptr(out))
    return out


def special_inverse(data, n_power, batch, table, scale, prec="f64"):
    data = np.ascontiguousarray(data, _cdt(prec)).copy()
    getattr(lib(), "orc_special_inverse_" + prec)(_ptr(data), n_power, batch,
                                                  _ptr(np.ascontiguousarray(table, _cdt(prec))),
                                                  scale.real, scale.imag)
    return data


# ---------------------------------------------------------------- exact products
def _pair(a, b):
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    assert a.shape == b.shape
    return a, b


def schoolbook_linear(a, b):
    a, b = _pair(a, b)
    out = np.zeros(2 * len(a), np.int64)
    lib().orc_schoolbook_linear(len(a), _ptr(a), _ptr(b), _ptr(out))
    return out


def schoolbook_negacyclic(a, b):
    a, b = _pair(a, b)
    out = np.zeros(len(a), np.int64)
    lib().orc_schoolbook_negacyclic(len(a), _ptr(a), _ptr(b), _ptr(out))
    return out


def schoolbook_cyclic(a, b):
    a, b = _pair(a, b)
    out = np.zeros(len(a), np.int64)
    lib().orc_schoolbook_cyclic(len(a), _ptr(a), _ptr(b), _ptr(out))
    return out


def ntt_negacyclic(a, b):
    """Exact products mod X^n+1 for rows of a, b ([count, n] int64)."""
    a, b = _pair(a, b)
    one = a.ndim == 1
    if one:
        a, b = a[None], b[None]
    out = np.zeros_like(a)
    lib().orc_ntt_negacyclic(a.shape[1], a.shape[0], _ptr(a), _ptr(b), _ptr(out))
    return out[0] if one else out


def ntt_cyclic(a, b):
    a, b = _pair(a, b)
    one = a.ndim == 1
    if one:
        a, b = a[None], b[None]
    out = np.zeros_like(a)
    lib().orc_ntt_cyclic(a.shape[1], a.shape[0], _ptr(a), _ptr(b), _ptr(out))
    return out[0] if one else out


# ---------------------------------------------------------------- inputs
def fill_uniform(seed, lo, hi, count):
    out = np.zeros(count, np.int64)
    lib().orc_fill_uniform_i64(seed, lo, hi, count, _ptr(out))
    return out
