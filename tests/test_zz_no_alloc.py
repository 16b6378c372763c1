"""GPU test of gpufft_polymul_f64_noalloc, the FP64 fused product that
allocates nothing (static device scratch): called through the C ABI with the
caller's FFNT tables, its outputs must equal gpufft_polymul_f64's (same
kernels, same table values) on two calls in a row, the second with a smaller
input, and round to the exact products."""
import ctypes

import numpy as np
import pytest

import oracle as O
from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _lib():
    import gpufft_amd as G
    lib = G.lib()
    f = lib.gpufft_polymul_f64_noalloc
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
    f.restype = ctypes.c_int
    return G, f


def _tables(G, ring_power, stream):
    n = ring_power - 1
    return [G.generate_table(k, n, inv, stream=stream) for k, inv in
            ((G.TABLE_FFNT_TWIST, False), (G.TABLE_FFNT_TWIST, True),
             (G.TABLE_FFNT_ROOT, False), (G.TABLE_FFNT_ROOT, True))]


def _case(G, f, ring_power, batch, seed, stream):
    import torch
    N = 1 << ring_power
    a = O.fill_uniform(seed, 0, 1 << 20, batch * N)
    b = O.fill_uniform(seed + 1, -128, 129, batch * N)
    da = to_dev(a.astype(np.float64))
    db = to_dev(b.astype(np.float64))
    ref = torch.empty(batch * N, dtype=torch.float64, device="cuda")
    plan = G.PolymulPlan(ring_power, G.PREC_F64, max_batch=batch, stream=stream)
    plan.multiply_f64(da, db, ref, batch, stream=stream)
    tw, utw, rt, irt = _tables(G, ring_power, stream)
    out = torch.full((batch * N,), -1.0, dtype=torch.float64, device="cuda")
    st = f(da.data_ptr(), db.data_ptr(), out.data_ptr(), ring_power, batch, tw.data_ptr(),
           utw.data_ptr(), rt.data_ptr(), irt.data_ptr(), stream)
    G.check(st)
    got, want = to_host(out), to_host(ref)
    plan.close()
    assert np.array_equal(got, want), float(np.max(np.abs(got - want)))
    # rounded products are the exact ones (first and last product of the batch)
    for p in (0, batch - 1):
        exact = O.ntt_negacyclic(a[p * N:(p + 1) * N], b[p * N:(p + 1) * N])
        assert np.array_equal(np.round(got[p * N:(p + 1) * N]).astype(np.int64), exact)


def test_noalloc_matches_plan_two_calls():
    """First call: N = 2^16, batch 150 (launches of 127 and 23 products
    through the 64 MiB scratch); second call: N = 2^15, batch 3; then
    N = 2^12 (one workgroup per product, no scratch)."""
    import torch
    G, f = _lib()
    s = torch.cuda.current_stream().cuda_stream
    print("noalloc: ring 2^16 batch 150", flush=True)
    _case(G, f, 16, 150, 0x6770756666740101, s)
    print("noalloc: ring 2^15 batch 3", flush=True)
    _case(G, f, 15, 3, 0x6770756666740103, s)
    print("noalloc: ring 2^12 batch 5", flush=True)
    _case(G, f, 12, 5, 0x6770756666740105, s)


def test_noalloc_rejects_out_of_range():
    import torch
    G, f = _lib()
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = x.data_ptr()
    s = torch.cuda.current_stream().cuda_stream
    for rp in (6, 18):
        assert f(p, p, p, rp, 1, p, p, p, p, s) == G.ERR_INVALID_ARGUMENT
        assert "ring_power" in G.lib().gpufft_last_error_message().decode()
    assert f(p, p, p, 12, 0, p, p, p, p, s) == G.ERR_INVALID_ARGUMENT
    assert f(p, None, p, 12, 1, p, p, p, p, s) == G.ERR_NULL_POINTER
