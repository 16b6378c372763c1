"""GPU parity of the fused batched negacyclic products (gpufft_polymul_*) and of
the GPU-generated root tables.

Products are compared bit-exactly (after rounding, for the FP64 form) with
the oracle's exact integer product (NTT over 2^64-2^32+1, itself pinned to the
schoolbook product in test_oracle.py); the FP64 form is also compared, before
rounding, with the two-call GPU_FFNT flow of the reference.
"""
import numpy as np
import pytest

import oracle as O
from helpers import SEED_BASE, to_dev, to_host

pytestmark = pytest.mark.gpu


def _g():
    import gpufft_amd as G
    return G


def _operands(N, batch, seed, a_bits=20, b_lo=-128, b_hi=129):
    a = O.fill_uniform(seed, 0, 1 << a_bits, batch * N).reshape(batch, N)
    b = O.fill_uniform(seed + 0x100, b_lo, b_hi, batch * N).reshape(batch, N)
    return a, b


# ---- tables ----
@pytest.mark.parametrize("n_power", [1, 2, 6, 12, 15])
def test_gpu_tables_match_reference_layout(n_power):
    G = _g()
    n = 1 << n_power
    ft = O.fft_tables(n)
    assert np.max(np.abs(to_host(G.generate_table(G.TABLE_FFT, n_power)) - ft["rev"])) < 5e-16
    assert np.max(np.abs(to_host(G.generate_table(G.TABLE_FFT, n_power, True)) - ft["inv_rev"])) < 5e-16
    nt = O.ffnt_tables(2 * n)
    for kind, inv, key in ((G.TABLE_FFNT_ROOT, False, "rev_root"), (G.TABLE_FFNT_ROOT, True, "inv_rev_root"),
                           (G.TABLE_FFNT_TWIST, False, "twist"), (G.TABLE_FFNT_TWIST, True, "untwist")):
        got = to_host(G.generate_table(kind, n_power, inv))
        assert np.max(np.abs(got - nt[key])) < 5e-16, key
    for inv in (False, True):
        got = to_host(G.generate_table(G.TABLE_SPECIAL, n_power, inv))
        assert np.max(np.abs(got - O.special_table(n, inv))) < 5e-16
    # float tables: the double values rounded once
    got32 = to_host(G.generate_table(G.TABLE_FFNT_TWIST, n_power, False, prec="f32"))
    assert np.array_equal(got32, nt["twist"].astype(np.complex64))


# ---- FP64 fused product ----
# rings below 2^15 are one workgroup per product (no combine step); from
# 2^15 the product spans 2^(ring_power-14) workgroups and its last inverse
# stages run in the product kernel ("in_kernel") or in a second launch
# ("pass")
F64_CASES = ([(r, "single") for r in (7, 8, 12, 14)] +
             [(r, c) for r in (15, 16, 17) for c in ("in_kernel", "pass")])


@pytest.mark.parametrize("rb", [3, 4])
@pytest.mark.parametrize("ring_power,combine", F64_CASES)
def test_polymul_f64_exact(ring_power, combine, rb):
    G = _g()
    import torch
    N = 1 << ring_power
    batch = 11  # not a multiple of the 8-XCD block grouping
    a, b = _operands(N, batch, SEED_BASE + 3 + ring_power)
    plan = G.PolymulPlan(ring_power, G.PREC_F64, max_batch=batch)
    plan.radix_bits = rb
    plan.combine_in_kernel = combine == "in_kernel"
    assert plan.combine_in_kernel == (combine == "in_kernel")
    out = torch.empty(batch * N, dtype=torch.float64, device="cuda")
    plan.multiply_f64(to_dev(a.astype(np.float64).reshape(-1)), to_dev(b.astype(np.float64).reshape(-1)),
                      out, batch)
    got = to_host(out).reshape(batch, N)
    exact = O.ntt_negacyclic(a, b)
    assert np.array_equal(np.round(got).astype(np.int64), exact)
    # same unrounded values as the oracle's two-call GPU_FFNT flow, to FP64 rounding
    t = O.ffnt_tables(N)
    L = N // 2
    fwd = O.ffnt_forward(np.concatenate([a, b]).astype(np.float64).reshape(-1), ring_power, 2 * batch,
                         t["twist"], t["rev_root"])
    ref = O.ffnt_inverse(fwd, ring_power, batch, t["untwist"], t["inv_rev_root"], 1, complex(1.0 / L))
    ref = ref.reshape(batch, N)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_polymul_f64_linear_through_negacyclic():
    """A linear product whose degree stays below the ring (operands of degree
    < 2^12 in the ring mod X^(2^13) + 1, the C1 shape of the reference's
    example C_C product) is the negacyclic one: the fused FP64 product
    equals the schoolbook convolution, with zeros above degree 2^13 - 2."""
    G = _g()
    import torch
    n, N = 1 << 12, 1 << 13
    a = O.fill_uniform(SEED_BASE + 91, 0, 16, n)
    b = O.fill_uniform(SEED_BASE + 92, 0, 16, n)
    pa = np.zeros(N, np.float64)
    pb = np.zeros(N, np.float64)
    pa[:n], pb[:n] = a, b
    plan = G.PolymulPlan(13, G.PREC_F64, max_batch=1)
    out = torch.empty(N, dtype=torch.float64, device="cuda")
    plan.multiply_f64(to_dev(pa), to_dev(pb), out, 1)
    got = np.round(to_host(out)).astype(np.int64)
    plan.close()
    ref = np.convolve(a.astype(np.int64), b.astype(np.int64))
    assert np.array_equal(got[:2 * n - 1], ref)
    assert not got[2 * n - 1:].any()


@pytest.mark.parametrize("in_kernel", [True, False])
def test_polymul_f64_chunks_and_reuse(in_kernel):
    """batch larger than max_batch runs in chunks; a second call on the same
    plan with a smaller batch gives the same products (the in-kernel combine's
    arrival counters are back at zero after every launch)."""
    G = _g()
    import torch
    rp, N = 15, 1 << 15
    a, b = _operands(N, 20, SEED_BASE + 77)
    plan = G.PolymulPlan(rp, G.PREC_F64, max_batch=8)
    plan.combine_in_kernel = in_kernel
    da, db = to_dev(a.astype(np.float64).reshape(-1)), to_dev(b.astype(np.float64).reshape(-1))
    out = torch.empty(20 * N, dtype=torch.float64, device="cuda")
    plan.multiply_f64(da, db, out, 20)
    exact = O.ntt_negacyclic(a, b)
    assert np.array_equal(np.round(to_host(out)).astype(np.int64).reshape(20, N), exact)
    out2 = torch.empty(3 * N, dtype=torch.float64, device="cuda")
    plan.multiply_f64(da, db, out2, 3)
    assert np.array_equal(np.round(to_host(out2)).astype(np.int64).reshape(3, N), exact[:3])


def test_polymul_f64_combine_forms_agree():
    """in-kernel combine and the separate combine pass give the same unrounded
    values to FP64 rounding, call after call."""
    G = _g()
    import torch
    rp, N, batch = 16, 1 << 16, 9
    a, b = _operands(N, batch, SEED_BASE + 91)
    da, db = to_dev(a.astype(np.float64).reshape(-1)), to_dev(b.astype(np.float64).reshape(-1))
    plan = G.PolymulPlan(rp, G.PREC_F64, max_batch=batch)
    outs = []
    for in_kernel in (True, False, True):
        plan.combine_in_kernel = in_kernel
        o = torch.empty(batch * N, dtype=torch.float64, device="cuda")
        plan.multiply_f64(da, db, o, batch)
        outs.append(to_host(o))
    scale = np.max(np.abs(outs[1]))
    assert np.max(np.abs(outs[0] - outs[1])) <= 1e-13 * scale
    assert np.array_equal(outs[0], outs[2])


# ---- FP32-split fused product ----
SPLITS = {
    # (a_limb_bits, a_limbs, b_limb_bits, b_limbs[, a center log2])
    "a5x5": (5, 5, 0, 1),
    "a4x6": (4, 6, 0, 1),
    "a7x3_b4x2": (7, 3, 4, 2),
    "a4x5c19": (4, 5, 0, 1, 19),
    "a6x4c19_b4x2": (6, 4, 4, 2, 19),
}


# splits at the edge of the FP32 margin at this ring size (a in [0, 2^20),
# b in [-2^7, 2^7]): the kernel may report a largest distance of 0.5
BEYOND_BOUND = {(17, "a5x5"), (17, "a7x3_b4x2")}


# (log2 points per thread, radix bits): 512 threads per 2^14-point
# sub-problem with radix 8, 16 or 32 steps, or 1024 threads with radix 8 or 16
THREAD_RADIX = [(5, 3), (5, 4), (5, 5), (4, 3), (4, 4)]


@pytest.mark.parametrize("e,rb", THREAD_RADIX)
@pytest.mark.parametrize("split", sorted(SPLITS))
@pytest.mark.parametrize("ring_power", [8, 12, 13, 14, 15, 16, 17])
def test_polymul_f32_split_exact(ring_power, split, e, rb):
    G = _g()
    import torch
    N = 1 << ring_power
    batch = 5
    a, b = _operands(N, batch, SEED_BASE + 33 + ring_power)
    wa, ka, wb, kb = SPLITS[split][:4]
    plan = G.PolymulPlan(ring_power, G.PREC_F32_SPLIT, batch, wa, ka, wb, kb)
    plan.radix_bits = min(rb, 4)
    plan.thread_points = e
    plan.radix_bits = rb
    assert (plan.thread_points, plan.radix_bits) == (e, rb)
    if len(SPLITS[split]) > 4:
        plan.a_center_log2 = SPLITS[split][4]
        assert plan.a_center_log2 == SPLITS[split][4]
    out = torch.empty(batch * N, dtype=torch.int64, device="cuda")
    err = torch.empty(batch * 2, dtype=torch.float64, device="cuda")
    plan.multiply_i32(to_dev(a.astype(np.int32).reshape(-1)), to_dev(b.astype(np.int32).reshape(-1)),
                      out, batch, err)
    got = to_host(out).reshape(batch, N)
    e = to_host(err).reshape(batch, 2)
    exact = O.ntt_negacyclic(a, b)
    # the guarantee: a product whose largest distance to the rounded value is
    # below 0.5 is exact; only splits at the edge of the FP32 margin may report
    # 0.5 or more (then the product is flagged, not promised)
    if (ring_power, split) not in BEYOND_BOUND:
        assert np.all(e[:, 0] < 0.5), e
    for i in range(batch):
        if e[i, 0] < 0.5:
            assert np.array_equal(got[i], exact[i]), i


def test_polymul_f32_unsplit_small_coefficients():
    """No split: exact while the products stay small (here |c| < 2^14)."""
    G = _g()
    import torch
    rp, N, batch = 12, 1 << 12, 4
    a, b = _operands(N, batch, SEED_BASE + 5, a_bits=3, b_lo=-4, b_hi=4)
    plan = G.PolymulPlan(rp, G.PREC_F32_SPLIT, batch, 0, 1, 0, 1)
    out = torch.empty(batch * N, dtype=torch.int64, device="cuda")
    plan.multiply_i32(to_dev(a.astype(np.int32).reshape(-1)), to_dev(b.astype(np.int32).reshape(-1)),
                      out, batch)
    assert np.array_equal(to_host(out).reshape(batch, N), O.ntt_negacyclic(a, b))
