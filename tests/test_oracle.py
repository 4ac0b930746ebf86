"""Pins the CPU oracle before it is trusted as the checker.

The reference ships no golden vectors or fixtures (its only checks are
commented out, benchmarks/fftmpi.jl:80-100), so the oracle is pinned by:
  * tests/golden/golden.npz: brute-force mpmath sums of FFTW's definitions on
    tiny shapes, inputs from an independent pure-Python SplitMix64;
  * scipy.fft (pocketfft), an independent library, at the config shapes;
  * invariants: Parseval, linearity, shift theorem, delta/constant inputs,
    Hermitian symmetry, REDFT01 o REDFT10 = 2n, round trips;
  * planted mistakes (wrong sign, missing normalisation, wrong DCT offset)
    that these checks must catch.
"""
import hashlib
import os

import numpy as np
import pytest
import scipy.fft as sf

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "golden.npz")


@pytest.fixture(scope="module")
def gold():
    h = hashlib.sha256(open(GOLD, "rb").read()).hexdigest()
    man = open(os.path.join(HERE, "golden", "MANIFEST.sha256")).read().split()[0]
    assert h == man
    return np.load(GOLD, allow_pickle=False)


# ---------------------------------------------------------------- generator
def test_splitmix64_known_answer():
    # SplitMix64 from state 0: first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4
    assert O.lib().dfo_splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert O.lib().dfo_splitmix64(0, 1) == 0x6E789E6AA1B965F4


def test_generator_matches_pure_python(gold):
    assert np.array_equal(gold["splitmix_seed0_first"], [O.uniform(0, i) for i in range(4)])
    a = O.fill(101, (4, 3, 2), np.complex128)
    assert np.array_equal(a, gold["c2c_4x3x2_in"])
    r = O.fill(201, (6, 4, 3), np.float64)
    assert np.array_equal(r, gold["r2c_6x4x3_in"])
    f = O.fill(5, (9,), np.float32)
    assert np.array_equal(f, O.fill(5, (9,), np.float64).astype(np.float32))


# ---------------------------------------------------------------- golden vectors
@pytest.mark.parametrize("name", ["c2c_4x3x2", "c2c_8x6x5", "c2c_7x5x3"])
def test_c2c_golden(gold, name):
    a = gold[name + "_in"]
    assert O.rel_l2(O.fftn(a), gold[name + "_fwd"]) < 1e-15
    assert O.rel_l2(O.fftn(O.fftn(a), inverse=True), gold[name + "_rt"]) < 1e-15
    assert O.rel_l2(O.fft_pencil(a, chunk=2), gold[name + "_fwd"]) < 1e-15
    assert O.rel_l2(O.fft_slab(a, chunk=1), gold[name + "_fwd"]) < 1e-15


@pytest.mark.parametrize("name", ["r2c_6x4x3", "r2c_5x4x2"])
def test_r2c_golden(gold, name):
    a = gold[name + "_in"]
    F = O.rfftn(a)
    assert O.rel_l2(F, gold[name + "_fwd"]) < 1e-15
    assert O.rel_l2(O.irfftn(F, a.shape[0]), a) < 1e-15


@pytest.mark.parametrize("kind", ["REDFT10", "REDFT01", "REDFT00", "REDFT11", "RODFT00", "RODFT10", "RODFT01",
                                  "RODFT11", "DHT"])
def test_r2r_golden(gold, kind):
    a = gold["r2r_%s_6x5_in" % kind]
    k = getattr(O, kind)
    assert O.rel_l2(O.r2rn(a, k), gold["r2r_%s_6x5" % kind]) < 1e-15


# ---------------------------------------------------------------- scipy at config shapes
@pytest.mark.parametrize("shape", [(128, 128, 128), (384, 240, 210), (96, 160, 192), (48, 32, 256)])
def test_c2c_vs_scipy(shape):
    a = O.fill(1, shape, np.complex128)
    assert O.rel_l2(O.fftn(a), sf.fftn(a)) < 1e-14
    assert O.rel_l2(O.fftn(a, inverse=True), sf.ifftn(a)) < 1e-14


@pytest.mark.parametrize("shape", [(11, 13, 17), (1009, 3, 2), (2, 3, 4093), (22, 26, 34), (97, 1, 5)])
def test_c2c_prime_lengths_vs_scipy(shape):
    """Lengths with prime factors above 7 (the library's Bluestein passes):
    the oracle's recursive FFT ends in direct DFTs of the prime factors."""
    a = O.fill(3, shape, np.complex128)
    assert O.rel_l2(O.fftn(a), sf.fftn(a)) < 1e-14
    assert O.rel_l2(O.fftn(a, inverse=True), sf.ifftn(a)) < 1e-14
    r = O.fill(4, shape, np.float64)
    ref = np.asfortranarray(np.transpose(sf.rfftn(np.transpose(r), axes=(0, 1, 2))))
    assert O.rel_l2(O.rfftn(r), ref) < 1e-14
    assert O.rel_l2(O.r2rn(r, O.REDFT10), sf.dctn(r, type=2)) < 1e-14
    assert O.rel_l2(O.r2rn(r, O.REDFT01), sf.dctn(r, type=3)) < 1e-14


def test_rfft_256_cubed_vs_scipy():
    a = O.fill(2, (256, 256, 256), np.float64)
    F = O.rfftn(a)
    ref = np.asfortranarray(np.transpose(sf.rfftn(np.transpose(a), axes=(0, 1, 2))))  # halve Julia dim 1
    assert O.rel_l2(F, ref) < 1e-14
    assert O.rel_l2(O.irfftn(F, 256), a) < 1e-14


@pytest.mark.parametrize("shape", [(256, 256), (128, 128, 128), (7, 5, 3)])
def test_dct_vs_scipy(shape):
    a = O.fill(4, shape, np.float64)
    assert O.rel_l2(O.r2rn(a, O.REDFT10), sf.dctn(a, type=2)) < 1e-14
    assert O.rel_l2(O.r2rn(a, O.REDFT01), sf.dctn(a, type=3)) < 1e-14


@pytest.mark.parametrize("kind,ref", [("REDFT00", ("dct", 1)), ("REDFT11", ("dct", 4)), ("RODFT00", ("dst", 1)),
                                      ("RODFT10", ("dst", 2)), ("RODFT01", ("dst", 3)), ("RODFT11", ("dst", 4))])
@pytest.mark.parametrize("shape", [(9, 6, 5), (7, 5, 9), (12, 2, 3), (11, 13, 2), (97, 3, 2)])
def test_r2r_kinds_vs_scipy(kind, ref, shape):
    """The remaining FFTW kinds against pocketfft's DCT/DST types 1-4
    (norm=None uses FFTW's unnormalised definitions)."""
    a = O.fill(16, shape, np.float64)
    fn = getattr(sf, ref[0] + "n")
    assert O.rel_l2(O.r2rn(a, getattr(O, kind)), fn(a, type=ref[1])) < 1e-14


def test_dht_vs_fft():
    """DHT: cas kernel = Re - Im of the forward DFT, per dim."""
    for n in (1, 2, 7, 12):
        x = O.fill(17, (n,), np.float64)
        F = np.fft.fft(x)
        assert O.rel_l2(O.r2rn(x, O.DHT), F.real - F.imag) < 1e-14


@pytest.mark.parametrize("dims", [(1,), (2,), (3,), (2, 3), (3, 1, 2)])
def test_dims_subsets_vs_scipy(dims):
    a = O.fill(3, (12, 10, 6), np.complex128)
    axes = [d - 1 for d in dims]
    assert O.rel_l2(O.fftn(a, dims=dims), sf.fftn(a, axes=axes)) < 1e-15
    r = O.fill(3, (12, 10, 6), np.float64)
    F = O.rfftn(r, dims=dims)
    ref = sf.fftn(r, axes=axes)
    sl = [slice(None)] * 3
    sl[dims[0] - 1] = slice(0, r.shape[dims[0] - 1] // 2 + 1)
    assert O.rel_l2(F, ref[tuple(sl)]) < 1e-15


# ---------------------------------------------------------------- invariants
def test_invariants_c2c():
    shape = (16, 12, 10)
    N = int(np.prod(shape))
    a = O.fill(7, shape, np.complex128)
    b = O.fill(8, shape, np.complex128)
    A, B = O.fftn(a), O.fftn(b)
    # Parseval
    assert abs(np.vdot(A, A).real / N - np.vdot(a, a).real) / np.vdot(a, a).real < 1e-14
    # linearity
    assert O.rel_l2(O.fftn(2.5 * a - 1j * b), 2.5 * A - 1j * B) < 1e-14
    # shift theorem along dim 2
    s = np.roll(a, 3, axis=1)
    k = np.arange(shape[1]).reshape(1, -1, 1)
    assert O.rel_l2(O.fftn(s), A * np.exp(-2j * np.pi * 3 * k / shape[1])) < 1e-13
    # delta -> ones, constant -> N delta
    d = np.zeros(shape, np.complex128, order="F")
    d[0, 0, 0] = 1
    assert np.allclose(O.fftn(d), 1, atol=1e-15)
    c = np.ones(shape, np.complex128, order="F")
    C = O.fftn(c)
    assert abs(C[0, 0, 0] - N) < 1e-10 and np.abs(C).ravel()[1:].max() < 1e-10
    # plane wave -> single bin
    i, j, kk = np.meshgrid(np.arange(16), np.arange(12), np.arange(10), indexing="ij")
    w = np.asfortranarray(np.exp(2j * np.pi * (3 * i / 16 + 5 * j / 12 + 7 * kk / 10)))
    W = O.fftn(w)
    assert abs(W[3, 5, 7] - N) < 1e-9
    W[3, 5, 7] = 0
    assert np.abs(W).max() < 1e-9


def test_invariants_real():
    shape = (10, 8, 6)
    a = O.fill(9, shape, np.float64)
    F = O.fftn(a.astype(np.complex128))
    # Hermitian symmetry of the full spectrum; rfft keeps bins 0..n/2 of dim 1
    idx = tuple((-np.arange(n)) % n for n in shape)
    assert O.rel_l2(F, np.conj(F[np.ix_(*idx)])) < 1e-15
    assert O.rel_l2(O.rfftn(a), F[: shape[0] // 2 + 1]) < 1e-15
    # REDFT01 o REDFT10 = 2n per dim
    for n in (1, 2, 5, 8):
        x = O.fill(10, (n,), np.float64)
        y = O.r2rn(O.r2rn(x, O.REDFT10), O.REDFT01)
        assert O.rel_l2(y, 2 * n * x) < 1e-14
    for k in O.R2R_KINDS:
        x = O.fill(11, (6, 5), np.float64)
        inv = {O.REDFT10: O.REDFT01, O.REDFT01: O.REDFT10, O.RODFT10: O.RODFT01, O.RODFT01: O.RODFT10}.get(k, k)
        y = O.r2rn(O.r2rn(x, k), inv)
        scale = O.r2r_logical_size(k, 6) * O.r2r_logical_size(k, 5)
        assert O.rel_l2(y, scale * x) < 1e-13, k


# ---------------------------------------------------------------- restatement structure
@pytest.mark.parametrize("chunk", [1, 3, 4, 16])
def test_chunked_pipeline_equals_plain(chunk):
    a = O.fill(12, (16, 12, 10), np.complex128)
    ref = O.fftn(a)
    assert O.rel_l2(O.fft_pencil(a, chunk=chunk), ref) < 1e-15
    assert O.rel_l2(O.fft_slab(a, chunk=chunk), ref) < 1e-15
    assert O.rel_l2(O.ifft_pencil(ref, chunk=chunk), a) < 1e-15
    assert O.rel_l2(O.ifft_slab(ref, chunk=chunk), a) < 1e-15


def test_box_maps_match_numpy_slicing():
    src = O.fill(13, (9, 7, 5), np.complex128)
    buf = np.zeros(3 + 4 * 3 * 2, np.complex128)
    O.pack(buf, src, (2, 1, 3), (4, 3, 2), 3)
    assert np.array_equal(buf[3:].reshape((4, 3, 2), order="F"), src[2:6, 1:4, 3:5])
    dst = np.zeros((6, 6, 6), np.complex128, order="F")
    O.unpack(dst, buf, (1, 2, 0), (4, 3, 2), 3)
    assert np.array_equal(dst[1:5, 2:5, 0:2], src[2:6, 1:4, 3:5])
    d2 = np.zeros((5, 5, 5), np.complex128, order="F")
    O.local_copy(d2, (0, 1, 2), src, (3, 2, 1), (4, 3, 3))
    assert np.array_equal(d2[0:4, 1:4, 2:5], src[3:7, 2:5, 1:4])
    for perm in [(1, 0, 2), (2, 0, 1), (1, 2, 0)]:
        assert np.array_equal(O.permute(src, perm), np.transpose(src, perm))


# ---------------------------------------------------------------- planted mistakes
def test_checks_catch_planted_mistakes():
    a = O.fill(14, (8, 6, 4), np.complex128)
    ref = sf.fftn(a)
    # wrong sign: the backward transform in place of the forward one
    wrong = np.array(a, order="F", copy=True)
    for ax in range(3):
        O.c2c_axis(wrong, ax, +1, 1.0)
    assert O.rel_l2(wrong, ref) > 0.1
    # missing inverse normalisation
    inv = np.array(ref, order="F", copy=True)
    for ax in range(3):
        O.c2c_axis(inv, ax, +1, 1.0)
    assert O.rel_l2(inv, a) > 0.1
    # DCT-II computed with the DCT-IV half-sample offset
    x = O.fill(15, (8,), np.float64)
    assert O.rel_l2(O.r2rn(x, O.REDFT11), sf.dct(x, type=2)) > 0.1
    # a dropped axis
    one = np.array(a, order="F", copy=True)
    O.c2c_axis(one, 0, -1, 1.0)
    O.c2c_axis(one, 1, -1, 1.0)
    assert O.rel_l2(one, ref) > 0.1


# ---------------------------------------------------------------- spot bins
@pytest.mark.parametrize("name", ["c2c_4x3x2", "c2c_8x6x5", "c2c_7x5x3"])
def test_spot_bins_golden(gold, name):
    """Every bin of the golden tiny transforms (mpmath brute force), by the
    long-double separable summation; the wrong sign fails."""
    a = gold[name + "_in"]
    ref = gold[name + "_fwd"]
    bins = [(i, j, k) for k in range(a.shape[2]) for j in range(a.shape[1]) for i in range(a.shape[0])]
    got = O.spot_bins(a, bins)
    assert O.rel_l2(got, np.array([ref[b] for b in bins])) < 1e-15
    assert O.rel_l2(O.spot_bins(a, bins, sign=+1), np.array([ref[b] for b in bins])) > 1e-2


def test_spot_bins_match_oracle_fft_mixed_sizes():
    """Spot bins at 36x20x21 (factors 2, 3, 5, 7) against the oracle's FFT,
    both signs, including the Nyquist and the last bins."""
    a = O.fill(17, (36, 20, 21), np.complex128)
    bins = [(0, 0, 0), (18, 10, 0), (35, 19, 20), (1, 0, 0), (0, 1, 0), (0, 0, 1), (7, 13, 5)]
    F = O.fftn(a)
    Fi = O.fftn(a, inverse=True) * a.size
    assert O.rel_l2(O.spot_bins(a, bins), np.array([F[b] for b in bins])) < 1e-14
    assert O.rel_l2(O.spot_bins(a, bins, sign=+1), np.array([Fi[b] for b in bins])) < 1e-14
