"""Writes tests/golden/oracle_golden.npz: small seeded inputs and the oracle's
outputs for them (test fixtures; the reference ships none for this path).

Run: python tests/golden/make_golden.py   (rewrites the .npz)

Contents (seeds: splitmix64 counter streams, 0x6770756666740000 + config id):
  c1_a, c1_b, c1_product   config-1 shape: 4096 coefficients in [0,16), linear
                           product (schoolbook, 8192 entries)
  ffnt_a, ffnt_b, ffnt_product  negacyclic N=4096, a in [0,2^20), b in [-2^7,2^7]
  fft16_rev, fft16_inv_rev      FFT<double>(16) reversed tables
  ffnt32_*                      FFNT<double>(32) tables
  special16, special16_inv      special-FFT tables, 16 slots
  c2c_x, c2c_cyc, c2c_neg       forward transforms of a 32-point complex input
  ffnt_fwd                      FFNT forward spectrum of ffnt_a[:64]/ffnt_b[:64] (N=64)
  special_z, special_fwd        special forward of a 16-slot input
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402

SEED = 0x6770756666740000


def build():
    d = {}
    d["c1_a"] = O.fill_uniform(SEED + 1, 0, 16, 4096).astype(np.int64)
    d["c1_b"] = O.fill_uniform(SEED + 1 + 0x100, 0, 16, 4096).astype(np.int64)
    d["c1_product"] = O.schoolbook_linear(d["c1_a"], d["c1_b"])
    d["ffnt_a"] = O.fill_uniform(SEED + 3, 0, 1 << 20, 4096)
    d["ffnt_b"] = O.fill_uniform(SEED + 3 + 0x100, -128, 129, 4096)
    d["ffnt_product"] = O.schoolbook_negacyclic(d["ffnt_a"], d["ffnt_b"])
    t = O.fft_tables(16)
    d["fft16_rev"], d["fft16_inv_rev"] = t["rev"], t["inv_rev"]
    f = O.ffnt_tables(32)
    for k in ("rev_root", "inv_rev_root", "twist", "untwist"):
        d["ffnt32_" + k] = f[k]
    d["special16"], d["special16_inv"] = O.special_table(16), O.special_table(16, True)
    x = (O.fill_uniform(SEED + 9, -1000, 1000, 32) + 1j * O.fill_uniform(SEED + 10, -1000, 1000, 32))
    d["c2c_x"] = x
    d["c2c_cyc"] = O.c2c_forward(x, 5, 1, O.fft_tables(16)["rev"], 0)
    d["c2c_neg"] = O.c2c_forward(x, 5, 1, O.fft_tables(32)["rev"], 1)
    t64 = O.ffnt_tables(64)
    real = np.concatenate([d["ffnt_a"][:64], d["ffnt_b"][:64]]).astype(np.float64)
    d["ffnt_fwd"] = O.ffnt_forward(real, 6, 2, t64["twist"], t64["rev_root"])
    z = x[:16] / 1000.0
    d["special_z"] = z
    d["special_fwd"] = O.special_forward(z, 4, 1, d["special16"])
    return d


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "oracle_golden.npz"), **build())
    print("wrote", os.path.join(HERE, "oracle_golden.npz"))
