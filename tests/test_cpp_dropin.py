"""The C++ drop-in interface (include/gpufft/*.cuh) through example programs
written in the flow of the reference's own examples (tests/cpp/), plus the
host table generators compared bit-for-bit with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "bin", "dropin_examples")


def run(*args, timeout=120):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/bin/dropin_examples not built (run __graft_entry__.build())")
    r = subprocess.run([BIN] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
    return r.stdout


@pytest.mark.parametrize("logn", [4, 10, 12])
def test_cpu_fft_path(logn):
    run("cpu_fft", logn)


@pytest.mark.parametrize("logn", [4, 10, 12])
def test_cpu_ffnt_path(logn):
    run("cpu_ffnt", logn)


def test_host_errors_before_launch():
    run("errors")


@pytest.mark.parametrize("n", [1, 2, 64, 4096])
def test_host_tables_bitwise_equal_oracle(n, tmp_path):
    path = tmp_path / "t.bin"
    r = subprocess.run([BIN, "tables", str(n), str(path)], capture_output=True, timeout=60)
    assert r.returncode == 0
    raw = path.read_bytes()
    off = 0
    for prec, dt in (("f64", np.complex128), ("f32", np.complex64)):
        ft = O.fft_tables(n, prec)
        nt = O.ffnt_tables(2 * n, prec)
        for ref in (ft["rev"], ft["inv_rev"], nt["rev_root"], nt["inv_rev_root"], nt["twist"],
                    nt["untwist"]):
            got = np.frombuffer(raw, dtype=dt, count=len(ref), offset=off)
            off += got.nbytes
            assert np.array_equal(got.view(np.uint8), ref.astype(dt).view(np.uint8)), prec
    assert off == len(raw)


@pytest.mark.gpu
@pytest.mark.parametrize("case,logn,batch", [
    ("cc", 1, 1), ("cc", 11, 2), ("cc", 12, 3), ("cc", 15, 1),
    ("rr", 3, 2), ("rr", 12, 2), ("rr", 15, 1),
    ("ffnt", 4, 3), ("ffnt", 12, 4), ("ffnt", 16, 2), ("ffnt", 17, 1),
    ("special", 2, 2), ("special", 11, 3), ("special", 15, 2),
])
def test_gpu_dropin_examples(case, logn, batch):
    run(case, logn, batch)
