"""GPU parity of the persistent batched product of uneven ring sizes
(gpufft_batched_*), bit-exact against the oracle's exact NTT products, for
both product kinds and both precisions (FP32 limbs, FP64), with the plain
launch, the HIP-graph replay (several replays with eager work and
synchronisation between them) and every tuning candidate."""
import numpy as np
import pytest

import oracle as O
from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu

LIMBS = dict(a_limb_bits=5, a_limbs=2, b_limb_bits=5, b_limbs=2)


def _workload(kinds, rps, seed):
    import gpufft_amd as G
    descs, n_in, n_out = G.pack_products(kinds, rps)
    a = O.fill_uniform(seed, -256, 256, n_in)
    b = O.fill_uniform(seed + 1, -256, 256, n_in)
    exact = np.zeros(n_out, np.int64)
    for d in descs:
        N = 1 << d.ring_power
        x = a[d.a_offset:d.a_offset + N]
        y = b[d.b_offset:d.b_offset + N]
        if d.kind == G.PRODUCT_NEGACYCLIC:
            exact[d.out_offset:d.out_offset + N] = O.ntt_negacyclic(x, y)
        else:
            z = np.zeros(2 * N, np.int64)
            xp, yp = z.copy(), z.copy()
            xp[:N], yp[:N] = x, y
            exact[d.out_offset:d.out_offset + 2 * N] = O.ntt_cyclic(xp, yp)
    return descs, a, b, exact


MIX_KINDS = [0, 1] * 9 + [0, 0, 1, 1, 0, 1]
MIX_RP = [10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 7, 5,
          16, 17, 14, 9, 15, 12]


def test_batched_exact_mixed():
    import torch

    import gpufft_amd as G
    descs, a, b, exact = _workload(MIX_KINDS, MIX_RP, 0x6770756666740004)
    plan = G.BatchedPlan(descs, **LIMBS)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
    plan.execute(da, db, out, err)
    got = to_host(out)
    e = to_host(err)
    assert np.all(e < 0.5), e
    assert np.array_equal(got, exact)
    info = plan.info()
    assert info["items"] > len(descs)  # some products were split


@pytest.mark.parametrize("in_kernel", [False, True])
@pytest.mark.parametrize("rb,grid", [(3, 256), (4, 256), (4, 512), (3, 16)])
def test_batched_f64_exact_mixed(rb, grid, in_kernel):
    """The FP64 plan (one limb per operand, 2^13-point tile, split products
    up to 2^4 sub-problems) on the mixed list: bit-exact against the NTT for
    every radix, grid and combine placement."""
    import torch

    import gpufft_amd as G
    descs, a, b, exact = _workload(MIX_KINDS, MIX_RP, 0x6770756666740005)
    plan = G.BatchedPlan(descs, stream=None, precision=G.PREC_F64)
    assert plan.tile == 13 and plan.info()["radix_bits"] == 4
    plan.set_tuning(rb, grid)
    plan.combine_in_kernel = in_kernel
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    out = torch.full((len(exact),), -3, dtype=torch.int64, device="cuda")
    err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
    plan.execute(da, db, out, err)
    e = to_host(err)
    assert np.all(e < 1e-3), e  # FP64: far from the rounding boundary
    assert np.array_equal(to_host(out), exact)


def test_batched_f64_range_edges_and_limits():
    """FP64 plan: smallest and largest ring powers of both kinds (NEGACYCLIC
    6 and 18, LINEAR 5 and 17), full-range int32 coefficients of a small
    ring (exact in FP64), graph replays; radix 2^5 and the 2^14 tile are
    refused."""
    import torch

    import gpufft_amd as G
    kinds, rps = [0, 1, 0, 1], [6, 5, 18, 17]
    descs, a, b, exact = _workload(kinds, rps, 0x67707566667400c5)
    plan = G.BatchedPlan(descs, precision=G.PREC_F64)
    with pytest.raises(ValueError):
        plan.set_tuning(5, 256)
    with pytest.raises(ValueError):
        plan.tile = 14
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    s = torch.cuda.Stream()
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    for rep in range(3):
        with torch.cuda.stream(s):
            out.fill_(-1)
            plan.graph_launch(da, db, out, stream=s)
        s.synchronize()
        assert np.array_equal(to_host(out), exact), rep
        _ = (da.float() * 2).sum().item()
    # a ring of 64 coefficients with a up to 2^26 and b up to 2^14: products
    # up to 2^46, exact in FP64 (an FP32 plan would need limbs)
    descs, n_in, n_out = G.pack_products([G.PRODUCT_NEGACYCLIC, G.PRODUCT_LINEAR], [6, 6])
    rng = np.random.default_rng(0x67707566667400c6)
    x = rng.integers(-(1 << 26), 1 << 26, n_in)
    y = rng.integers(-(1 << 14), 1 << 14, n_in)
    want = np.zeros(n_out, np.int64)
    for d in descs:
        N = 1 << d.ring_power
        p = np.convolve(x[d.a_offset:d.a_offset + N], y[d.b_offset:d.b_offset + N])
        if d.kind == G.PRODUCT_NEGACYCLIC:
            w = p[:N].copy()
            w[:N - 1] -= p[N:]
            want[d.out_offset:d.out_offset + N] = w
        else:
            want[d.out_offset:d.out_offset + 2 * N - 1] = p
    plan = G.BatchedPlan(descs, precision=G.PREC_F64)
    out = torch.full((n_out,), 9, dtype=torch.int64, device="cuda")
    err = torch.zeros(2, dtype=torch.float64, device="cuda")
    plan.execute(to_dev(x.astype(np.int32)), to_dev(y.astype(np.int32)), out, err)
    assert np.all(to_host(err) < 0.1), to_host(err)
    assert np.array_equal(to_host(out), want)


def test_batched_graph_replays():
    import torch

    import gpufft_amd as G
    kinds, rps = [0, 1, 0, 1, 0], [16, 15, 13, 10, 11]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740044)
    plan = G.BatchedPlan(descs, **LIMBS)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    for rep in range(3):
        with torch.cuda.stream(s):
            out.fill_(-7)  # eager work between replays
            plan.graph_launch(da, db, out, stream=s)
        s.synchronize()
        assert np.array_equal(to_host(out), exact), rep
        _ = (da.float() * 2).sum().item()  # unrelated eager work + sync
    # eager launch after graph replays agrees too
    out.zero_()
    plan.execute(da, db, out)
    assert np.array_equal(to_host(out), exact)


@pytest.mark.parametrize("in_kernel", [False, True])
@pytest.mark.parametrize("rb", [3, 4, 5])
@pytest.mark.parametrize("grid", [16, 256])
def test_batched_tuning_candidates_agree(rb, grid, in_kernel):
    import torch

    import gpufft_amd as G
    kinds, rps = [1, 0, 0, 1, 1], [16, 17, 12, 11, 18]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740084)
    plan = G.BatchedPlan(descs, **LIMBS)
    plan.set_tuning(rb, grid)
    plan.combine_in_kernel = in_kernel
    assert plan.combine_in_kernel == in_kernel
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    plan.execute(to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32)), out)
    assert np.array_equal(to_host(out), exact)


def test_batched_range_edges():
    """Smallest and largest ring powers of both kinds (NEGACYCLIC 6 and 19,
    LINEAR 5 and 18) in one plan, both combine placements."""
    import torch

    import gpufft_amd as G
    kinds, rps = [0, 1, 0, 1], [6, 5, 19, 18]
    descs, a, b, exact = _workload(kinds, rps, 0x67707566667400c4)
    plan = G.BatchedPlan(descs, **LIMBS)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    for in_kernel in (False, True):
        plan.combine_in_kernel = in_kernel
        out = torch.full((len(exact),), -1, dtype=torch.int64, device="cuda")
        err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
        plan.execute(da, db, out, err)
        e = to_host(err)
        assert np.all(e < 0.5), e
        assert np.array_equal(to_host(out), exact), in_kernel


@pytest.mark.parametrize("split", [(4, 3, 4, 3), (4, 3, 5, 2), (5, 2, 4, 3), (9, 1, 5, 2),
                                   (5, 2, 9, 1), (3, 4, 3, 4)])
def test_batched_limb_schedules(split):
    """Limb splits other than 2 x 2: equal weights summed before one inverse
    (3 x 3 limbs: groups of 1, 2, 3, 2, 1 products, partial sums in stash
    slots), unequal weights inverted one by one, an unsplit operand."""
    import torch

    import gpufft_amd as G
    kinds, rps = [0, 1, 0, 1, 0], [15, 14, 12, 9, 7]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740104)
    plan = G.BatchedPlan(descs, *split)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    for in_kernel in (False, True):
        plan.combine_in_kernel = in_kernel
        out = torch.full((len(exact),), -1, dtype=torch.int64, device="cuda")
        err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
        plan.execute(da, db, out, err)
        e = to_host(err)
        assert np.all(e < 0.5), (split, e)
        assert np.array_equal(to_host(out), exact), (split, in_kernel)


@pytest.mark.parametrize("rb", [3, 4, 5])
def test_batched_tile_13(rb):
    """The 64 KiB tile (two workgroups per CU; products above 2^13 points
    split into more sub-problems): exact for both combine placements, and
    switching the tile back and forth rebuilds the work list correctly."""
    import torch

    import gpufft_amd as G
    kinds, rps = [1, 0, 0, 1, 1, 0], [17, 17, 14, 13, 9, 12]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740144)
    plan = G.BatchedPlan(descs, **LIMBS)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    items14 = plan.info()["items"]
    plan.tile = 13
    assert plan.tile == 13 and plan.info()["items"] > items14
    plan.set_tuning(rb, 512)
    for in_kernel in (False, True):
        plan.combine_in_kernel = in_kernel
        out = torch.full((len(exact),), -1, dtype=torch.int64, device="cuda")
        err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
        plan.execute(da, db, out, err)
        assert np.all(to_host(err) < 0.5)
        assert np.array_equal(to_host(out), exact), in_kernel
    plan.tile = 14
    assert plan.info()["items"] == items14
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="default stream"):
        plan.graph_launch(da, db, out)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.graph_launch(da, db, out, stream=s)
    s.synchronize()
    assert np.array_equal(to_host(out), exact)


def test_batched_tile_refused_for_large_products():
    """A ring that would need 2^5 sub-problems at the 2^13 tile keeps 2^14."""
    import gpufft_amd as G
    descs, _, _, _ = _workload([0], [19], 0x6770756666740184)
    plan = G.BatchedPlan(descs, **LIMBS)
    with pytest.raises(ValueError, match="sub-problems"):
        plan.tile = 13
    assert plan.tile == 14


@pytest.mark.parametrize("trial", range(4))
def test_batched_random_lists(trial):
    """Seeded random lists (kinds, ring sizes, limb split, tile, radix, grid,
    combine placement), exact against the NTT products."""
    import torch

    import gpufft_amd as G
    rng = np.random.default_rng(0x6770756666740200 + trial)
    splits = [(5, 2, 5, 2), (4, 3, 4, 3), (4, 3, 5, 2), (6, 2, 3, 3)]
    cnt = int(rng.integers(3, 10))
    kinds = [int(k) for k in rng.integers(0, 2, cnt)]
    rps = [int(rng.integers(6 if k == 0 else 5, 15)) for k in kinds]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740300 + trial)
    plan = G.BatchedPlan(descs, *splits[trial])
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    for tile in (14, 13):
        plan.tile = tile
        plan.set_tuning(int(rng.integers(3, 6)), int(rng.choice([32, 256, 512])))
        plan.combine_in_kernel = bool(rng.integers(0, 2))
        out = torch.full((len(exact),), -5, dtype=torch.int64, device="cuda")
        err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
        plan.execute(da, db, out, err)
        assert np.all(to_host(err) < 0.5)
        assert np.array_equal(to_host(out), exact), (tile, kinds, rps)


@pytest.mark.parametrize("trial", range(4))
def test_batched_f64_random_lists(trial):
    """FP64 plans on seeded random lists (kinds, ring sizes up to NEGACYCLIC
    2^18 / LINEAR 2^17, radix, grid, combine placement), each executed twice
    with a retune between: exact against the NTT products every time."""
    import torch

    import gpufft_amd as G
    rng = np.random.default_rng(0x6770756666740400 + trial)
    cnt = int(rng.integers(3, 12))
    kinds = [int(k) for k in rng.integers(0, 2, cnt)]
    rps = [int(rng.integers(6 if k == 0 else 5, 19 if k == 0 else 18)) for k in kinds]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740500 + trial)
    plan = G.BatchedPlan(descs, precision=G.PREC_F64)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    for rep in range(2):
        plan.set_tuning(int(rng.integers(3, 5)), int(rng.choice([8, 64, 256, 512])))
        plan.combine_in_kernel = bool(rng.integers(0, 2))
        out = torch.full((len(exact),), -5, dtype=torch.int64, device="cuda")
        err = torch.zeros(len(descs), dtype=torch.float64, device="cuda")
        plan.execute(da, db, out, err)
        assert np.all(to_host(err) < 1e-3)
        assert np.array_equal(to_host(out), exact), (rep, kinds, rps)


def test_batched_autotune_db(tmp_path):
    """tune_batched times radix x grid x combine placement x tile, stores the
    winner in the tuning database, and a second plan gets it applied without
    timing; the tuned plan stays exact."""
    import torch

    import gpufft_amd as G
    from gpufft_amd import autotune as A
    kinds, rps = [1, 0, 1, 0], [16, 15, 11, 13]
    descs, a, b, exact = _workload(kinds, rps, 0x67707566667402c4)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    dbf = str(tmp_path / "tune.json")
    plan = G.BatchedPlan(descs, **LIMBS)
    with torch.cuda.stream(s):
        cand, timings = A.tune_batched(plan, lambda: plan.execute(da, db, out, stream=s), s,
                                       key="test_batched_db", reps=1, db_file=dbf)
    s.synchronize()
    assert len(cand) == 4 and len(timings) == 24
    assert {c[3] for c in timings} == {13, 14}
    plan2 = G.BatchedPlan(descs, **LIMBS)
    with torch.cuda.stream(s):
        cand2, timings2 = A.tune_batched(plan2, lambda: plan2.execute(da, db, out, stream=s), s,
                                         key="test_batched_db", reps=1, db_file=dbf)
        out.zero_()
        plan2.execute(da, db, out, stream=s)
    s.synchronize()
    assert tuple(cand2) == tuple(cand) and timings2 == {}
    info = plan2.info()
    assert (info["radix_bits"], info["grid"]) == (cand[0], cand[1])
    assert plan2.combine_in_kernel == bool(cand[2]) and plan2.tile == cand[3]
    assert np.array_equal(to_host(out), exact)


def test_batched_f64_autotune():
    """tune_batched on an FP64 plan times radix 3/4 x two grids x combine
    placement at the one tile 13; the tuned plan stays exact."""
    import torch

    import gpufft_amd as G
    from gpufft_amd import autotune as A
    kinds, rps = [1, 0, 1, 0], [17, 15, 11, 14]
    descs, a, b, exact = _workload(kinds, rps, 0x67707566667402c5)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    plan = G.BatchedPlan(descs, precision=G.PREC_F64)
    with torch.cuda.stream(s):
        cand, timings = A.tune_batched(plan, lambda: plan.execute(da, db, out, stream=s), s,
                                       key="test_batched_f64", reps=1, use_db=False)
        out.zero_()
        plan.execute(da, db, out, stream=s)
    s.synchronize()
    assert len(timings) == 8 and {c[3] for c in timings} == {13}
    assert {c[0] for c in timings} == {3, 4}
    assert np.array_equal(to_host(out), exact)


def test_batched_failed_retune_keeps_plan_usable():
    """A grid whose stash cannot be allocated is refused; the plan keeps its
    old stash, tuning and captured graph, and still gives exact products."""
    import torch

    import gpufft_amd as G
    kinds, rps = [0, 1, 0], [15, 12, 11]
    descs, a, b, exact = _workload(kinds, rps, 0x6770756666740045)
    plan = G.BatchedPlan(descs, **LIMBS)
    da, db = to_dev(a.astype(np.int32)), to_dev(b.astype(np.int32))
    out = torch.zeros(len(exact), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.graph_launch(da, db, out, stream=s)
    s.synchronize()
    assert np.array_equal(to_host(out), exact)
    before = plan.info()
    with pytest.raises(G.HipError):
        plan.set_tuning(4, 1 << 28)  # a stash of about 10^14 bytes
    assert plan.info() == before
    out.zero_()
    with torch.cuda.stream(s):
        plan.graph_launch(da, db, out, stream=s)
    s.synchronize()
    assert np.array_equal(to_host(out), exact)
    out.zero_()
    plan.execute(da, db, out)
    assert np.array_equal(to_host(out), exact)


def _exact_products(descs, a, b, threads=16):
    """Exact products of every descriptor by the oracle's NTT, in a thread
    pool (each oracle call releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    import gpufft_amd as G

    def one(d):
        N = 1 << d.ring_power
        x = a[d.a_offset:d.a_offset + N].astype(np.int64)
        y = b[d.b_offset:d.b_offset + N].astype(np.int64)
        if d.kind == G.PRODUCT_NEGACYCLIC:
            return O.ntt_negacyclic(x, y)
        xp = np.zeros(2 * N, np.int64)
        yp = np.zeros(2 * N, np.int64)
        xp[:N], yp[:N] = x, y
        return O.ntt_cyclic(xp, yp)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, list(descs)))


def test_batched_configured_size_4096():
    """The batched uneven configuration at its full size: 4096 products, ring
    sizes 2^e with e uniform in {10..17} (seeded), even index linear, odd
    negacyclic, coefficients uniform in [-2^8, 2^8) (the seeds of bench.py
    --batched-uneven), one persistent launch replayed as a HIP graph twice,
    every product bit-exact against the oracle's NTT."""
    import torch

    import gpufft_amd as G
    from gpufft_amd import synthetic as S
    count = 4096
    e = S.uniform(S.SEED_BASE + 4, 10, 18, count)
    kinds = [G.PRODUCT_LINEAR if i % 2 == 0 else G.PRODUCT_NEGACYCLIC for i in range(count)]
    descs, n_in, n_out = G.pack_products(kinds, e)
    a = S.uniform(S.SEED_BASE + 4 + 0x100, -256, 256, n_in).astype(np.int32)
    b = S.uniform(S.SEED_BASE + 4 + 0x200, -256, 256, n_in).astype(np.int32)
    print("batched 4096: exact products on the host", flush=True)
    exact = _exact_products(descs, a, b)
    da, db = to_dev(a), to_dev(b)
    out = torch.zeros(n_out, dtype=torch.int64, device="cuda")
    err = torch.zeros(count, dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    for name, plan in (("f32 split", G.BatchedPlan(descs, **LIMBS)),
                       ("f64", G.BatchedPlan(descs, precision=G.PREC_F64))):
        info = plan.info()
        assert info["items"] > count  # the large products are split
        for rep in range(2):
            print("batched 4096:", name, "graph replay", rep, flush=True)
            with torch.cuda.stream(s):
                out.fill_(-1)
                plan.graph_launch(da, db, out, err, stream=s)
            s.synchronize()
            got = to_host(out)
            e_host = to_host(err)
            assert np.all(e_host < 0.5), (name, float(e_host.max()))
            bad = [i for i, d in enumerate(descs)
                   if not np.array_equal(got[d.out_offset:d.out_offset + len(exact[i])], exact[i])]
            assert not bad, (name, rep, len(bad), bad[:8])
        plan.close()
