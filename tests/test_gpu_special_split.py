"""Special (CKKS) FFT in FP32 arithmetic with hi/lo split slots, against the
oracle's FP64 special FFT (oracle/oracle.c orc_special_*, itself pinned to the
brute-force canonical embedding in test_oracle.py).

Workload shape of the CKKS encode/decode case: slots uniform on the unit disk
scaled by 2^30, l = 2^15 slots (ring 2^16).  The bound asserted is the FP32
transform error, relative to the largest output value; the split keeps the
inputs exact (a plain float copy of a 2^30-scaled slot already loses up to
ulp(2^30)/2 = 64)."""
import numpy as np
import pytest

import oracle as O
from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _slots(l, batch, seed):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.random(l * batch))
    th = 2 * np.pi * rng.random(l * batch)
    return (r * np.exp(1j * th) * 2.0 ** 30).astype(np.complex128)


@pytest.mark.parametrize("n_power", [3, 11, 13, 15])
def test_special_split_encode_decode(n_power):
    import torch

    import gpufft_amd as G
    l, batch = 1 << n_power, 6
    z = _slots(l, batch, 500 + n_power)
    T64, Ti64 = O.special_table(l), O.special_table(l, True)
    T32, Ti32 = T64.astype(np.complex64), Ti64.astype(np.complex64)
    work = torch.empty(2 * batch * l, dtype=torch.complex64, device="cuda")
    d = to_dev(z)
    G.special_fft_split(d, to_dev(Ti32), G.config(n_power, G.INVERSE, mod_inverse=1.0 / l), batch, work)
    enc = to_host(d)
    ref_enc = O.special_inverse(z, n_power, batch, Ti64, complex(1.0 / l))
    e_enc = np.max(np.abs(enc - ref_enc)) / np.max(np.abs(ref_enc))
    G.special_fft_split(d, to_dev(T32), G.config(n_power, G.FORWARD), batch, work)
    dec = to_host(d)
    ref_dec = O.special_forward(ref_enc, n_power, batch, T64)
    e_dec = np.max(np.abs(dec - ref_dec)) / np.max(np.abs(ref_dec))
    e_rt = np.max(np.abs(dec - z)) / np.max(np.abs(z))
    print("special split n=%d: encode %.3g decode %.3g round trip %.3g (max abs %.3g)"
          % (n_power, e_enc, e_dec, e_rt, np.max(np.abs(dec - z))))
    assert e_enc < 2e-6 and e_dec < 2e-6 and e_rt < 4e-6
    # the same transform on a float copy of the slots is worse or equal
    dz = to_dev(z.astype(np.complex64))
    G.GPU_Special_FFT(dz, to_dev(Ti32), G.config(n_power, G.INVERSE, mod_inverse=1.0 / l), batch)
    e_plain = np.max(np.abs(to_host(dz) - ref_enc)) / np.max(np.abs(ref_enc))
    assert e_enc <= e_plain * 1.5 + 1e-9
