"""Python binding of libgpufft_amd.so (the MI355X GPU-FFT C ABI) via ctypes.

This is the host-side mirror of the reference's GPU entry points
(Alisah-Ozcan/GPU-FFT, src/include/gpufft/fft.cuh): the same operations,
argument meaning and errors (`ValueError` where the reference throws
std::invalid_argument, `HipError` where it throws gpufft::CudaException).
Device memory comes from torch tensors (torch is only the allocator and the
stream provider here); every call goes through the compiled C ABI.

There is no CPU fallback: if the shared library is missing, importing the
binding raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# always the in-tree build: tests and benchmarks must run the library built
# from this source tree
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libgpufft_amd.so")

FORWARD, INVERSE = 0, 1
X_N_PLUS, X_N_MINUS = 0, 1

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_HIP = 2
ERR_NULL_POINTER = 3
ERR_UNSUPPORTED = 4

MIN_N_POWER = 1
MIN_N_POWER_FFNT = 2
MAX_N_POWER = 24


class HipError(RuntimeError):
    """A HIP runtime error reported by the library (gpufft::CudaException)."""


class Config(ctypes.Structure):
    """gpufft_config (include/gpufft_amd.h) = fft_configuration<T>."""
    _fields_ = [
        ("n_power", ctypes.c_int),
        ("fft_type", ctypes.c_int),
        ("reduction_poly", ctypes.c_int),
        ("zero_padding", ctypes.c_int),
        ("mod_inverse_re", ctypes.c_double),
        ("mod_inverse_im", ctypes.c_double),
        ("stream", ctypes.c_void_p),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libgpufft_amd.so not built: run `make -C gpu-fft_amd` "
            "(or __graft_entry__.build()) first; there is no fallback path")
    return ctypes.CDLL(LIB_PATH)


_lib = _load()
_P = ctypes.c_void_p
_I = ctypes.c_int
_CP = ctypes.POINTER(Config)

_SIGS = {
    "gpufft_fft_c2c_f32": [_P, _P, _CP, _I, _I],
    "gpufft_fft_c2c_f64": [_P, _P, _CP, _I, _I],
    "gpufft_fft_r2r_f32": [_P, _P, _P, _CP, _I, _I],
    "gpufft_fft_r2r_f64": [_P, _P, _P, _CP, _I, _I],
    "gpufft_fft_r2r_inverse_hermitian_f32": [_P, _P, _P, _CP, _I, _I],
    "gpufft_fft_r2r_inverse_hermitian_f64": [_P, _P, _P, _CP, _I, _I],
    "gpufft_ffnt_f32": [_P, _P, _P, _P, _CP, _I, _I],
    "gpufft_ffnt_f64": [_P, _P, _P, _P, _CP, _I, _I],
    "gpufft_special_fft_f32": [_P, _P, _CP, _I],
    "gpufft_special_fft_f64": [_P, _P, _CP, _I],
    "gpufft_special_fft_split_f32": [_P, _P, _CP, _I, _P],
    "gpufft_generate_table_f64": [_P, _I, _I, _I, _P],
    "gpufft_generate_table_f32": [_P, _I, _I, _I, _P],
    "gpufft_polymul_plan_create": [_P, _I, _I, _I, _I, _I, _I, _I, _P],
    "gpufft_polymul_plan_destroy": [_P],
    "gpufft_polymul_f64": [_P, _P, _P, _P, _I, _P],
    "gpufft_polymul_i32": [_P, _P, _P, _P, _I, _P, _P],
    "gpufft_polymul_f64_phases": [_P, _P, _P, _P, _I, _I, _P],
    "gpufft_polymul_f64_noalloc": [_P, _P, _P, _I, _I, _P, _P, _P, _P, _P],
    "gpufft_polymul_plan_set_radix_bits": [_P, _I],
    "gpufft_batched_plan_create": [_P, _P, _I, _I, _I, _I, _I, _P],
    "gpufft_batched_plan_create_f64": [_P, _P, _I, _P],
    "gpufft_batched_plan_destroy": [_P],
    "gpufft_batched_plan_set_tuning": [_P, _I, _I],
    "gpufft_batched_plan_set_combine": [_P, _I],
    "gpufft_batched_plan_get_combine": [_P, _P],
    "gpufft_batched_plan_set_tile": [_P, _I],
    "gpufft_batched_plan_get_tile": [_P, _P],
    "gpufft_batched_plan_info": [_P, _P, _P, _P, _P],
    "gpufft_batched_execute": [_P, _P, _P, _P, _P, _P],
    "gpufft_batched_graph_launch": [_P, _P, _P, _P, _P, _P],
    "gpufft_polymul_plan_get_radix_bits": [_P],
    "gpufft_polymul_plan_set_thread_points": [_P, _I],
    "gpufft_polymul_plan_get_thread_points": [_P],
    "gpufft_polymul_plan_set_combine": [_P, _I],
    "gpufft_polymul_plan_get_combine": [_P],
    "gpufft_polymul_plan_set_a_center": [_P, _I],
    "gpufft_polymul_plan_get_a_center": [_P],
    "gpufft_polymul_i32_phases": [_P, _P, _P, _P, _I, _P, _I, _P],
    "gpufft_ckks_encode_f64": [_P, _P, _P, _I, ctypes.c_double, _I, _P, _P],
    "gpufft_ckks_decode_f64": [_P, _P, _P, _I, ctypes.c_double, _I, _P],
}


def _declare():
    for name, args in _SIGS.items():
        f = getattr(_lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib.gpufft_last_error_message.argtypes = []
    _lib.gpufft_last_error_message.restype = ctypes.c_char_p
    _lib.gpufft_last_hip_error.argtypes = []
    _lib.gpufft_last_hip_error.restype = ctypes.c_int
    _lib.gpufft_polymul_plan_scratch_bytes.argtypes = [_P]
    _lib.gpufft_polymul_plan_scratch_bytes.restype = ctypes.c_size_t


_declare()


def lib():
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def check(status):
    """Raise the reference's exception type for a non-zero status."""
    if status == OK:
        return
    msg = _lib.gpufft_last_error_message().decode()
    if status == ERR_HIP:
        raise HipError(msg)
    raise ValueError(msg)


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def _prec(t):
    """Element precision of a tensor argument; raw pointers mean double."""
    if t is None or isinstance(t, int):
        return "f64"
    import torch
    if t.dtype in (torch.complex64, torch.float32, torch.int32):
        return "f32"
    return "f64"


def config(n_power, fft_type, reduction_poly=X_N_MINUS, mod_inverse=1.0, stream=None,
           zero_padding=False):
    c = Config()
    c.n_power = n_power
    c.fft_type = fft_type
    c.reduction_poly = reduction_poly
    c.zero_padding = 1 if zero_padding else 0
    mi = complex(mod_inverse)
    c.mod_inverse_re = mi.real
    c.mod_inverse_im = mi.imag
    c.stream = _stream(stream)
    return c


# ---- drop-in operations (names of the reference's GPU entry points) ----
def GPU_FFT(inout, root_table, cfg, batch_size, multiplication=False):
    """Complex-to-complex GPU_FFT (reference fft.cuh:121-124)."""
    fn = getattr(_lib, "gpufft_fft_c2c_" + _prec(inout))
    check(fn(_ptr(inout), _ptr(root_table), ctypes.byref(cfg), batch_size, int(multiplication)))


def GPU_FFT_R(real_inout, temp, root_table, cfg, batch_size, multiplication=False):
    """Real-to-complex / complex-to-real GPU_FFT (reference fft.cuh:126-132)."""
    fn = getattr(_lib, "gpufft_fft_r2r_" + _prec(temp))
    check(fn(_ptr(real_inout), _ptr(temp), _ptr(root_table), ctypes.byref(cfg), batch_size,
             int(multiplication)))


def GPU_FFT_R_inverse_hermitian(real_inout, temp, root_table, cfg, batch_size,
                                multiplication=False):
    """R_R inverse for spectra of real data only (half-length packed form for
    X^N-1); equals GPU_FFT_R's INVERSE when temp is Hermitian and mod_inverse
    is real."""
    fn = getattr(_lib, "gpufft_fft_r2r_inverse_hermitian_" + _prec(temp))
    check(fn(_ptr(real_inout), _ptr(temp), _ptr(root_table), ctypes.byref(cfg), batch_size,
             int(multiplication)))


def GPU_FFNT(inout, temp, twist_table, root_table, cfg, batch_size, multiplication=False):
    """GPU_FFNT (reference fft.cuh:269-273); pass the untwist table for INVERSE."""
    fn = getattr(_lib, "gpufft_ffnt_" + _prec(temp))
    check(fn(_ptr(inout), _ptr(temp), _ptr(twist_table), _ptr(root_table), ctypes.byref(cfg),
             batch_size, int(multiplication)))


def special_fft_split(inout, root_table, cfg, batch_size, work):
    """Special FFT of complex128 slots in FP32 arithmetic (hi/lo split)."""
    check(_lib.gpufft_special_fft_split_f32(_ptr(inout), _ptr(root_table), ctypes.byref(cfg),
                                            batch_size, _ptr(work)))


def GPU_Special_FFT(inout, root_table, cfg, batch_size):
    """GPU_Special_FFT (reference fft.cuh:194-197)."""
    fn = getattr(_lib, "gpufft_special_fft_" + _prec(inout))
    check(fn(_ptr(inout), _ptr(root_table), ctypes.byref(cfg), batch_size))


# ---- GPU table generation (reference layouts) ----
TABLE_FFT, TABLE_FFNT_ROOT, TABLE_FFNT_TWIST, TABLE_SPECIAL = 0, 1, 2, 3


def ckks_encode(slots, coeffs, inv_table, n_power, scale, batch, work, stream=None):
    """CKKS encode: complex slots [batch][2^n_power] -> int64 coefficients
    [batch][2^(n_power+1)] (special inverse FFT, bit reversal, x scale, round)."""
    check(_lib.gpufft_ckks_encode_f64(_ptr(slots), _ptr(coeffs), _ptr(inv_table), n_power,
                                      float(scale), batch, _ptr(work), _stream(stream)))


def ckks_decode(coeffs, slots, table, n_power, scale, batch, stream=None):
    """CKKS decode: int64 coefficients -> complex slots (/ scale, bit reversal,
    special forward FFT)."""
    check(_lib.gpufft_ckks_decode_f64(_ptr(coeffs), _ptr(slots), _ptr(table), n_power,
                                      float(scale), batch, _stream(stream)))


def generate_table(kind, n_power, inverse=False, prec="f64", stream=None):
    """Root-of-unity table of 2^n_power complex entries, built on the GPU."""
    import torch
    dt = torch.complex128 if prec == "f64" else torch.complex64
    out = torch.empty(1 << n_power, dtype=dt, device="cuda")
    fn = getattr(_lib, "gpufft_generate_table_" + prec)
    check(fn(out.data_ptr(), kind, n_power, int(inverse), _stream(stream)))
    return out


# ---- fused batched negacyclic products ----
PREC_F64, PREC_F32_SPLIT = 0, 1
PHASE_TRANSFORM, PHASE_COMBINE, PHASE_ALL = 1, 2, 3


class PolymulPlan:
    """Plan of gpufft_polymul_* for ring size N = 2^ring_power (mod X^N+1)."""

    def __init__(self, ring_power, precision=PREC_F64, max_batch=256, a_limb_bits=0, a_limbs=1,
                 b_limb_bits=0, b_limbs=1, stream=None):
        h = ctypes.c_void_p()
        check(_lib.gpufft_polymul_plan_create(ctypes.byref(h), ring_power, precision, max_batch,
                                              a_limb_bits, a_limbs, b_limb_bits, b_limbs,
                                              _stream(stream)))
        self.handle = h
        self.ring_power = ring_power
        self.precision = precision

    @property
    def radix_bits(self):
        return _lib.gpufft_polymul_plan_get_radix_bits(self.handle)

    @radix_bits.setter
    def radix_bits(self, rb):
        check(_lib.gpufft_polymul_plan_set_radix_bits(self.handle, rb))

    @property
    def thread_points(self):
        """log2 of the points per thread of the product kernel: 4 for FP64
        (512 threads per 2^13-point sub-problem); 5 or 4 for the FP32 split
        (512 or 1024 threads per 2^14-point sub-problem)."""
        return _lib.gpufft_polymul_plan_get_thread_points(self.handle)

    @thread_points.setter
    def thread_points(self, e):
        check(_lib.gpufft_polymul_plan_set_thread_points(self.handle, int(e)))

    @property
    def combine_in_kernel(self):
        """FP64 plans of ring_power >= 15: last inverse stages in the product
        kernel (True) or in a second launch (False, default)."""
        return bool(_lib.gpufft_polymul_plan_get_combine(self.handle))

    @combine_in_kernel.setter
    def combine_in_kernel(self, v):
        check(_lib.gpufft_polymul_plan_set_combine(self.handle, 1 if v else 0))

    @property
    def a_center_log2(self):
        """FP32-split plans: a is split as a - 2^c (0: as is), plus the exact
        correction limb 2^c (J b)."""
        return _lib.gpufft_polymul_plan_get_a_center(self.handle)

    @a_center_log2.setter
    def a_center_log2(self, c):
        check(_lib.gpufft_polymul_plan_set_a_center(self.handle, c))

    @property
    def scratch_bytes(self):
        return _lib.gpufft_polymul_plan_scratch_bytes(self.handle)

    def multiply_f64(self, a, b, out, batch, stream=None, phases=3):
        check(_lib.gpufft_polymul_f64_phases(self.handle, _ptr(a), _ptr(b), _ptr(out), batch,
                                             phases, _stream(stream)))

    def multiply_i32(self, a, b, out, batch, err_stats=None, stream=None, phases=3):
        check(_lib.gpufft_polymul_i32_phases(self.handle, _ptr(a), _ptr(b), _ptr(out), batch,
                                             _ptr(err_stats), phases, _stream(stream)))

    def close(self):
        if self.handle:
            _lib.gpufft_polymul_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- batched products of uneven ring sizes ----
PRODUCT_NEGACYCLIC, PRODUCT_LINEAR = 0, 1


class ProductDesc(ctypes.Structure):
    """gpufft_product_desc (include/gpufft_amd.h)."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("ring_power", ctypes.c_int),
        ("a_offset", ctypes.c_longlong),
        ("b_offset", ctypes.c_longlong),
        ("out_offset", ctypes.c_longlong),
    ]


def output_length(kind, ring_power):
    """int64 output elements of one product."""
    return (1 << ring_power) * (1 if kind == PRODUCT_NEGACYCLIC else 2)


def pack_products(kinds, ring_powers):
    """Descriptors for products stored back to back: returns (descs, a/b
    element count, output element count)."""
    descs = (ProductDesc * len(kinds))()
    io = oo = 0
    for i, (k, r) in enumerate(zip(kinds, ring_powers)):
        descs[i].kind, descs[i].ring_power = int(k), int(r)
        descs[i].a_offset = descs[i].b_offset = io
        descs[i].out_offset = oo
        io += 1 << int(r)
        oo += output_length(int(k), int(r))
    return descs, io, oo


class BatchedPlan:
    """gpufft_batched_plan: uneven negacyclic / linear products, one launch.
    precision PREC_F32_SPLIT (limbs as given) or PREC_F64 (one limb each;
    the limb arguments are then ignored)."""

    def __init__(self, descs, a_limb_bits=5, a_limbs=2, b_limb_bits=5, b_limbs=2, stream=None,
                 precision=PREC_F32_SPLIT):
        h = ctypes.c_void_p()
        if precision == PREC_F64:
            check(_lib.gpufft_batched_plan_create_f64(ctypes.byref(h), ctypes.byref(descs),
                                                      len(descs), _stream(stream)))
        elif precision == PREC_F32_SPLIT:
            check(_lib.gpufft_batched_plan_create(ctypes.byref(h), ctypes.byref(descs), len(descs),
                                                  a_limb_bits, a_limbs, b_limb_bits, b_limbs,
                                                  _stream(stream)))
        else:
            raise ValueError("precision must be PREC_F32_SPLIT or PREC_F64")
        self.handle = h
        self.count = len(descs)
        self.precision = precision

    def set_tuning(self, radix_bits, grid):
        check(_lib.gpufft_batched_plan_set_tuning(self.handle, radix_bits, grid))

    @property
    def combine_in_kernel(self):
        """True: split products are combined inside the persistent kernel
        (one launch); False: by a second, streaming kernel (default)."""
        v = ctypes.c_int()
        check(_lib.gpufft_batched_plan_get_combine(self.handle, ctypes.byref(v)))
        return bool(v.value)

    @combine_in_kernel.setter
    def combine_in_kernel(self, on):
        check(_lib.gpufft_batched_plan_set_combine(self.handle, 1 if on else 0))

    @property
    def tile(self):
        """log2 of the LDS tile (points per workgroup): 14 or 13."""
        v = ctypes.c_int()
        check(_lib.gpufft_batched_plan_get_tile(self.handle, ctypes.byref(v)))
        return v.value

    @tile.setter
    def tile(self, log2_tile):
        check(_lib.gpufft_batched_plan_set_tile(self.handle, int(log2_tile)))

    def info(self):
        n, s, rb, g = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int()
        check(_lib.gpufft_batched_plan_info(self.handle, ctypes.byref(n), ctypes.byref(s),
                                            ctypes.byref(rb), ctypes.byref(g)))
        return {"items": n.value, "scratch_elems": s.value, "radix_bits": rb.value,
                "grid": g.value}

    def execute(self, a, b, out, err_stats=None, stream=None):
        check(_lib.gpufft_batched_execute(self.handle, _ptr(a), _ptr(b), _ptr(out),
                                          _ptr(err_stats), _stream(stream)))

    def graph_launch(self, a, b, out, err_stats=None, stream=None):
        check(_lib.gpufft_batched_graph_launch(self.handle, _ptr(a), _ptr(b), _ptr(out),
                                               _ptr(err_stats), _stream(stream)))

    def close(self):
        if self.handle:
            _lib.gpufft_batched_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
