"""The CMake package (gpu-fft_amd/cmake/GPUFFTConfig.cmake) lets a downstream
project do what the reference's README asks of its users (README.md:91-105):
find_package(GPUFFT CONFIG REQUIRED) and link GPUFFT::fft.  A consumer that
includes gpufft/fft.cuh is configured, built and run (host code only, so no
GPU is needed), from the source tree and from an installed prefix."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "gpu-fft_amd")

CONSUMER_CMAKE = """
cmake_minimum_required(VERSION 3.21)
project(consumer LANGUAGES CXX)
set(CMAKE_CXX_STANDARD 17)
find_package(GPUFFT 1.0 CONFIG REQUIRED)
add_executable(consumer main.cpp)
target_link_libraries(consumer PRIVATE GPUFFT::fft)
set_target_properties(consumer PROPERTIES BUILD_RPATH "${GPUFFT_LIB_DIR}")
"""

CONSUMER_MAIN = r"""
#include <cstdio>
#include "gpufft/fft.cuh"
int main()
{
    gpufft::FFT<double> host(8);  // host tables of the reference's FFT<T>
    std::vector<COMPLEX<double>> t = host.ReverseRootTable();
    gpufft::fft_configuration<double> cfg{};
    cfg.n_power = 3;
    cfg.fft_type = gpufft::FORWARD;
    cfg.reduction_poly = gpufft::ReductionPolynomial::X_N_minus;
    std::printf("table %zu %.3f\n", t.size(), t[0].real());
    // out-of-range n_power is refused on the host, before any device call
    try {
        cfg.n_power = 40;
        gpufft::GPU_FFT<double>(nullptr, nullptr, cfg, 1, false);
    } catch (const std::invalid_argument& e) {
        std::printf("caught %s\n", e.what());
        return 0;
    }
    return 1;
}
"""


def _tools():
    if not shutil.which("cmake"):
        pytest.skip("cmake not installed")
    if not os.path.exists(os.path.join(PKG, "lib", "libgpufft_amd.so")):
        pytest.skip("library not built")


def _consume(tmp_path, gpufft_dir):
    src = tmp_path / "src"
    src.mkdir()
    (src / "CMakeLists.txt").write_text(CONSUMER_CMAKE)
    (src / "main.cpp").write_text(CONSUMER_MAIN)
    build = tmp_path / "build"
    env = dict(os.environ)
    subprocess.run(["cmake", "-S", str(src), "-B", str(build), "-DGPUFFT_DIR=%s" % gpufft_dir,
                    "-DCMAKE_PREFIX_PATH=/opt/rocm", "-DGPU_TARGETS=gfx950"],
                   check=True, capture_output=True, env=env, timeout=300)
    subprocess.run(["cmake", "--build", str(build)], check=True, capture_output=True, env=env,
                   timeout=300)
    r = subprocess.run([str(build / "consumer")], capture_output=True, text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "table 8 1.000" in r.stdout
    assert "caught Invalid n_power range!" in r.stdout


def test_find_package_from_source_tree(tmp_path):
    _tools()
    _consume(tmp_path, os.path.join(PKG, "cmake"))


def test_find_package_from_install_prefix(tmp_path):
    _tools()
    prefix = tmp_path / "prefix"
    subprocess.run(["make", "-C", PKG, "install", "PREFIX=%s" % prefix], check=True,
                   capture_output=True, timeout=300)
    assert (prefix / "lib" / "cmake" / "GPUFFT-1.0" / "GPUFFTConfig.cmake").exists()
    _consume(tmp_path, str(prefix / "lib" / "cmake" / "GPUFFT-1.0"))
