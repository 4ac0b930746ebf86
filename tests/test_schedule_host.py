"""The planner (host code of libdaggerfft_hip) schedules every axis pass of
every measured transform -- the configurations, the 64-transform batch and
each shape of the serving trace -- both for a plan and for a grouped launch,
with and without the shipped schedule cache (wisdom/gfx950.txt).  No GPU
needed: scheduling only reads the catalog and the cache."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

LIBDIR = os.path.join(ROOT, "daggerfft.jl_amd")
WISDOM = os.path.join(LIBDIR, "wisdom", "gfx950.txt")


def _shapes():
    from daggerfft import serve as S
    from daggerfft.batch import batch_config_sizes
    rows = [(0, (512, 512, 512), 1), (0, (128, 128, 128), 1), (1, (256, 256, 256), 1), (0, (384, 240, 210), 1)]
    rows += [(0, s, p) for s, p in batch_config_sizes(64, 7)]
    rows += sorted({(q.kind, q.shape, q.precision) for q in S.serving_trace()})
    return "".join("%d %d %d %d %d\n" % (k, s[0], s[1], s[2], p) for k, s, p in rows)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.exists(os.path.join(LIBDIR, "libdaggerfft_hip.so")):
        pytest.skip("needs g++ and the built library")
    exe = str(tmp_path_factory.mktemp("sched") / "schedule_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(LIBDIR, "csrc"), os.path.join(ROOT, "tests", "schedule_check.cpp"),
                    "-o", exe, "-L", LIBDIR, "-ldaggerfft_hip", "-Wl,-rpath," + LIBDIR], check=True)
    return exe


@pytest.mark.parametrize("wisdom", [True, False])
def test_every_measured_pass_schedules(checker, wisdom):
    args = [checker] + ([WISDOM] if wisdom else [])
    env = dict(os.environ, DFFT_WISDOM="off")
    r = subprocess.run(args, input=_shapes(), capture_output=True, text=True, env=env)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]
    assert "bad 0" in r.stdout


def test_bluestein_choices_on_the_host(tmp_path):
    """Convolution passes are chosen exactly for lengths with a prime factor
    above 7 -- Rader for C2C primes with 7-smooth n-1, Bluestein otherwise --
    with a valid convolution length and schedule (tests/bluestein_check.cpp)."""
    if shutil.which("g++") is None or not os.path.exists(os.path.join(LIBDIR, "libdaggerfft_hip.so")):
        pytest.skip("needs g++ and the built library")
    exe = str(tmp_path / "bluestein_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(LIBDIR, "csrc"), os.path.join(ROOT, "tests", "bluestein_check.cpp"),
                    "-o", exe, "-L", LIBDIR, "-ldaggerfft_hip", "-Wl,-rpath," + LIBDIR], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:]
