"""The schedule cache's wisdom files on the host (no GPU calls): the shipped
file loads, dfft_tune_save writes a stamp with the HIP version the library was built with, a
file stamped for another HIP build or architecture is refused with
DFFT_ERR_WISDOM and leaves the cache untouched, unstamped hand-written files
load, and a file that fails to parse loads nothing (all or nothing).

Mirrors the reference's wisdom import/export (src/fft.jl:113-124,162): FFTW
likewise refuses wisdom of another FFTW build."""
import os

import pytest

from conftest import ROOT

SHIPPED = os.path.join(ROOT, "daggerfft.jl_amd", "wisdom", "gfx950.txt")


@pytest.fixture
def D():
    import daggerfft as D
    D.tune_clear()
    yield D
    D.tune_clear()


def _entries(path):
    return [ln for ln in open(path) if ln.strip() and not ln.startswith("#")]


def test_shipped_wisdom_is_stamped_and_loads(D):
    first = open(SHIPPED).readline()
    assert "hip=" in first and "arch=gfx950" in first
    D.tune_load(SHIPPED)
    assert D.tune_entries() == len(set(ln.split()[0] for ln in _entries(SHIPPED)))


def test_save_round_trip_with_stamp(D, tmp_path):
    D.tune_load(SHIPPED)
    n = D.tune_entries()
    out = str(tmp_path / "w.txt")
    D.tune_save(out)
    first = open(out).readline()
    # the stamp names the HIP version of the library's build, the same as the shipped file's
    assert first.split("hip=")[1].split()[0] == open(SHIPPED).readline().split("hip=")[1].split()[0]
    assert sorted(_entries(out)) == sorted(set(_entries(SHIPPED)))
    D.tune_clear()
    assert D.tune_entries() == 0
    D.tune_load(out)
    assert D.tune_entries() == n


def test_foreign_stamp_refused(D, tmp_path):
    from daggerfft import _lib as L
    lines = open(SHIPPED).read().splitlines()
    for bad in ("hip=60342134", "arch=gfx942"):
        stamp = lines[0].replace("hip=70226015", bad) if bad.startswith("hip") else lines[0].replace(
            "arch=gfx950", bad)
        p = tmp_path / ("foreign_%s.txt" % bad.split("=")[0])
        p.write_text("\n".join([stamp] + lines[1:]) + "\n")
        D.tune_clear()
        if bad.startswith("arch"):
            # the architecture is known only with a device; without one the
            # stamp cannot be compared and the file loads
            D.tune_load(str(p))
            continue
        with pytest.raises(D.DFFTError) as ei:
            D.tune_load(str(p))
        assert ei.value.status == L.DFFT_ERR_WISDOM
        assert b"hip=60342134" in L.lib().dfft_last_error()
        assert D.tune_entries() == 0


def test_unstamped_override_and_all_or_nothing(D, tmp_path):
    p = tmp_path / "override.txt"
    p.write_text("1/0/512/0/27 16,16,2 0 32 0 2 97\n")
    D.tune_load(str(p))
    assert D.tune_entries() == 1
    D.tune_clear()
    bad = tmp_path / "bad.txt"
    bad.write_text("1/0/512/0/27 16,16,2 0 32 0 2 97\n1/0/256/0/24 11,5 0 16 0 2 1\n")  # radix 11
    with pytest.raises(D.DFFTError):
        D.tune_load(str(bad))
    assert D.tune_entries() == 0
