"""CPU checks of the C ABI boundary: the library loads, exports every symbol
include/gpufft_amd.h declares, and validates arguments on the host before any
HIP call (so these run without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "gpufft_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpufft_\w+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "gpufft_fft_c2c_f64" in names and "gpufft_ffnt_f32" in names


def test_library_exports_every_declared_symbol():
    import gpufft_amd as G
    lib = ctypes.CDLL(G.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def _cfg(G, n_power, fft_type=0):
    return G.config(n_power, fft_type, stream=0)


@pytest.mark.parametrize("n_power", [0, -3, 25, 40])
def test_invalid_n_power_raises_like_reference(n_power):
    """fft.cu:522-525 throws std::invalid_argument("Invalid n_power range!")."""
    import gpufft_amd as G
    with pytest.raises(ValueError, match="Invalid n_power range!"):
        G.GPU_FFT(0x1000, 0x2000, _cfg(G, n_power), 1)
    with pytest.raises(ValueError, match="Invalid n_power range!"):
        G.GPU_Special_FFT(0x1000, 0x2000, _cfg(G, n_power), 1)


def test_ffnt_lower_limit():
    import gpufft_amd as G
    with pytest.raises(ValueError, match="Invalid n_power range!"):
        G.check(G.lib().gpufft_ffnt_f64(0x1000, 0x2000, 0x3000, 0x4000,
                                         ctypes.byref(_cfg(G, 1)), 1, 0))


def test_null_and_batch_validation():
    import gpufft_amd as G
    with pytest.raises(ValueError, match="NULL"):
        G.GPU_FFT(None, 0x2000, _cfg(G, 12), 1)
    with pytest.raises(ValueError, match="batch"):
        G.GPU_FFT(0x1000, 0x2000, _cfg(G, 12), 0)
    with pytest.raises(ValueError, match="fft_type"):
        G.GPU_FFT(0x1000, 0x2000, _cfg(G, 12, 7), 1)
    st = G.lib().gpufft_fft_c2c_f64(0x1000, 0x2000, None, 1, 0)
    assert st == G.ERR_NULL_POINTER


def test_polymul_plan_validation_on_host():
    """Plan arguments are checked before any device call."""
    import gpufft_amd as G
    for args in ((3, G.PREC_F64), (30, G.PREC_F64), (12, 7)):
        with pytest.raises(ValueError):
            G.PolymulPlan(*args, stream=0)
    with pytest.raises(ValueError, match="limb"):
        G.PolymulPlan(12, G.PREC_F32_SPLIT, 4, 40, 2, 0, 1, stream=0)
    with pytest.raises(ValueError, match="limb"):  # top digit above bit 30
        G.PolymulPlan(12, G.PREC_F32_SPLIT, 4, 16, 3, 0, 1, stream=0)
    with pytest.raises(ValueError, match="max_batch"):
        G.PolymulPlan(12, G.PREC_F64, 0, stream=0)
    assert G.lib().gpufft_polymul_plan_set_combine(None, 1) == G.ERR_NULL_POINTER
    assert G.lib().gpufft_polymul_plan_get_combine(None) == -1
    assert G.lib().gpufft_polymul_plan_set_a_center(None, 19) == G.ERR_NULL_POINTER
    assert G.lib().gpufft_polymul_plan_get_a_center(None) == -1
    assert G.lib().gpufft_batched_plan_set_combine(None, 1) == G.ERR_NULL_POINTER
    assert G.lib().gpufft_batched_plan_get_combine(None, None) == G.ERR_NULL_POINTER
    assert G.lib().gpufft_batched_plan_set_tile(None, 13) == G.ERR_NULL_POINTER
    assert G.lib().gpufft_batched_plan_get_tile(None, None) == G.ERR_NULL_POINTER


def test_table_generation_validation_on_host():
    import gpufft_amd as G
    with pytest.raises(ValueError):
        G.check(G.lib().gpufft_generate_table_f64(0x1000, 9, 4, 0, None))
    with pytest.raises(ValueError):
        G.check(G.lib().gpufft_generate_table_f64(None, 0, 4, 0, None))


def test_batched_plan_validation_on_host():
    """Ring powers outside NEGACYCLIC 6..19 / LINEAR 5..18 and unknown kinds
    are refused before any device allocation."""
    import gpufft_amd as G
    for kind, rp in ((G.PRODUCT_NEGACYCLIC, 5), (G.PRODUCT_NEGACYCLIC, 20),
                     (G.PRODUCT_LINEAR, 4), (G.PRODUCT_LINEAR, 19), (7, 12)):
        descs, _, _ = G.pack_products([kind], [rp])
        with pytest.raises(ValueError):
            G.BatchedPlan(descs, stream=0)
    descs, _, _ = G.pack_products([G.PRODUCT_NEGACYCLIC], [12])
    for split in ((16, 3, 5, 2), (5, 2, 11, 4), (5, 0, 5, 2), (17, 2, 5, 2)):
        with pytest.raises(ValueError, match="limb split"):
            G.BatchedPlan(descs, *split, stream=0)
    descs[0].b_offset = 2
    with pytest.raises(ValueError, match="multiples of 4"):
        G.BatchedPlan(descs, stream=0)
    # FP64 plans: NEGACYCLIC 6..18, LINEAR 5..17 (2^13-point tile, at most
    # 2^4 sub-problems)
    for kind, rp in ((G.PRODUCT_NEGACYCLIC, 5), (G.PRODUCT_NEGACYCLIC, 19),
                     (G.PRODUCT_LINEAR, 4), (G.PRODUCT_LINEAR, 18)):
        descs, _, _ = G.pack_products([kind], [rp])
        with pytest.raises(ValueError, match="ring_power"):
            G.BatchedPlan(descs, stream=0, precision=G.PREC_F64)
    with pytest.raises(ValueError, match="precision"):
        G.BatchedPlan(descs, stream=0, precision=7)
    assert G.lib().gpufft_batched_plan_create_f64(None, None, 1, None) == G.ERR_NULL_POINTER
