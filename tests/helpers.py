"""Shared helpers of the test suite: device transfer and comparisons."""
import numpy as np

SEED_BASE = 0x6770756666740000


def to_dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_err(got, ref):
    """max |got - ref| / max |ref| (max-norm relative error)."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    scale = np.max(np.abs(ref)) if ref.size else 1.0
    if scale == 0:
        scale = 1.0
    return float(np.max(np.abs(got.astype(np.complex128) - ref.astype(np.complex128))) / scale)


# tolerances of the floating-point transforms (max-norm, relative to the
# largest reference value): fused multiply-add and a different summation
# order against the oracle's plain multiply-then-add.
TOL = {"f64": 1e-12, "f32": 2e-5}


def cdtype(prec):
    return np.complex128 if prec == "f64" else np.complex64


def rdtype(prec):
    return np.float64 if prec == "f64" else np.float32


def rng_complex(rng, n, prec):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(cdtype(prec))
