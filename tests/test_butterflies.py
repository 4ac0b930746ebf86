"""The DFT butterflies of the kernels (dfft_math.hpp, compiled for the host
here) against a direct long-double DFT: every radix the compile-time and
runtime schedules use, in double and float."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

RADICES = (2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16)


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("bfly") / "butterfly_check")
    src = os.path.join(ROOT, "tests", "butterfly_check.cpp")
    inc = os.path.join(ROOT, "daggerfft.jl_amd", "csrc")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", inc, src, "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.split("\n"):
        if line.strip():
            r, prec, err = line.split()
            res[(int(r), prec)] = float(err)
    return res


@pytest.mark.parametrize("radix", RADICES)
def test_butterfly_matches_dft(results, radix):
    assert results[(radix, "f64")] < 1e-15, results[(radix, "f64")]
    assert results[(radix, "f32")] < 3e-7, results[(radix, "f32")]
