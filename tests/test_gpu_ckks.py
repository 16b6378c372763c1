"""CKKS encode / decode around the special FFT (gpufft_ckks_*), against the
oracle (oracle.c orc_ckks_*, pinned below to the brute-force canonical
embedding) and by exact round trips.

Pins, all on integers where they can be:
  - decode(m) is the canonical embedding m(zeta_2N^(5^j mod 2N)) / Delta,
    evaluated by brute force (small l);
  - encode(decode(m)) == m exactly for random integer m (the embedding is
    invertible and the FP64 error stays far below 1/2);
  - GPU encode equals the oracle's encode, except where the oracle's
    pre-rounding value lies within 1e-6 of a rounding tie (+-1 allowed there).
Workload shape of the CKKS case: slots on the unit disk, Delta = 2^30,
l = 2^15 slots (ring 2^16)."""
import numpy as np
import pytest

import oracle as O
from helpers import to_dev, to_host

pytestmark = pytest.mark.gpu


def _g():
    import gpufft_amd as G
    return G


def _embed(m, n_power, delta):
    l = 1 << n_power
    N = 2 * l
    e = np.array([pow(5, j, 2 * N) for j in range(l)])
    k = np.arange(2 * l)
    return np.exp(2j * np.pi * np.outer(e, k) / (2 * N)) @ m.astype(np.float64) / delta


@pytest.mark.parametrize("n_power", [1, 3, 6, 9])
def test_decode_is_canonical_embedding_and_encode_inverts_it(n_power):
    import torch
    G = _g()
    l, batch, delta = 1 << n_power, 3, 2.0 ** 24
    rng = np.random.default_rng(40 + n_power)
    m = rng.integers(-2 ** 40, 2 ** 40, size=(batch, 2 * l))
    T, Ti = to_dev(O.special_table(l)), to_dev(O.special_table(l, True))
    slots = torch.empty(batch * l, dtype=torch.complex128, device="cuda")
    G.ckks_decode(to_dev(m.reshape(-1)), slots, T, n_power, delta, batch)
    z = to_host(slots).reshape(batch, l)
    for b in range(batch):
        ref = _embed(m[b], n_power, delta)
        assert np.max(np.abs(z[b] - ref)) <= 1e-12 * np.max(np.abs(ref))
    coeffs = torch.empty(batch * 2 * l, dtype=torch.int64, device="cuda")
    work = torch.empty(batch * l, dtype=torch.complex128, device="cuda")
    G.ckks_encode(slots, coeffs, Ti, n_power, delta, batch, work)
    assert np.array_equal(to_host(coeffs).reshape(batch, 2 * l), m)


@pytest.mark.parametrize("n_power", [4, 11, 13, 15])
def test_encode_matches_oracle_and_round_trips(n_power):
    import torch
    G = _g()
    l, batch, delta = 1 << n_power, 4, 2.0 ** 30
    rng = np.random.default_rng(70 + n_power)
    r = np.sqrt(rng.random(batch * l))
    th = 2 * np.pi * rng.random(batch * l)
    z = r * np.exp(1j * th)
    T64, Ti64 = O.special_table(l), O.special_table(l, True)
    coeffs = torch.empty(batch * 2 * l, dtype=torch.int64, device="cuda")
    work = torch.empty(batch * l, dtype=torch.complex128, device="cuda")
    G.ckks_encode(to_dev(z), coeffs, to_dev(Ti64), n_power, delta, batch, work)
    got = to_host(coeffs)
    ref = O.ckks_encode(z, n_power, batch, Ti64, delta)
    # pre-rounding values of the oracle, to find the near-ties
    u = O.special_inverse(z, n_power, batch, Ti64, complex(1.0 / l)).reshape(batch, l)
    br = np.array([O.bitreverse(i, n_power) for i in range(l)])
    pre = np.concatenate([u[:, br].real, u[:, br].imag], axis=1).reshape(-1) * delta
    near_tie = np.abs(np.abs(pre - np.floor(pre)) - 0.5) < 1e-6
    assert np.array_equal(got[~near_tie], ref[~near_tie])
    assert np.all(np.abs(got - ref) <= 1)
    # decode(encode(z)) = z up to the rounding of the coefficients
    slots = torch.empty(batch * l, dtype=torch.complex128, device="cuda")
    G.ckks_decode(coeffs, slots, to_dev(T64), n_power, delta, batch)
    back = to_host(slots)
    # each coefficient moves by <= 1/2 / delta; a slot sums 2l of them
    assert np.max(np.abs(back - z)) <= 2 * l * 0.5 / delta
    assert np.max(np.abs(back - O.ckks_decode(got, n_power, batch, T64, delta))) < 1e-12


def test_ckks_validation():
    G = _g()
    import torch
    x = torch.empty(8, dtype=torch.complex128, device="cuda")
    c = torch.empty(16, dtype=torch.int64, device="cuda")
    with pytest.raises(ValueError, match="scale"):
        G.ckks_encode(x, c, x, 3, 0.0, 1, x)
    with pytest.raises(ValueError, match="Invalid n_power range!"):
        G.ckks_decode(c, x, x, 0, 1.0, 1)
    with pytest.raises(ValueError, match="NULL"):
        G.ckks_encode(x, c, x, 3, 1.0, 1, None)
