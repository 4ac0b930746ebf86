"""C2C parity of the HIP path against the oracle (FFTW definitions restated in
oracle/dfft_oracle.c), through the C ABI via the Python mirror.

Tolerance: relative L2 <= 1e-12 for ComplexF64 (the stated FP64 target);
ComplexF32 errors are measured against the FP64 oracle and bounded by 1e-5."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL64 = 1e-12
TOL32 = 1e-5


def _dev(a):
    import daggerfft as D
    return D.DeviceArray.from_host(a)


def test_fill_bit_exact(gpu):
    import daggerfft as D
    for dt in (np.complex128, np.complex64, np.float64, np.float32):
        for shape in ((17, 5, 3), (1001, 999, 1), (3,)):
            A = D.rand(shape, dt, seed=7)
            ref = O.fill(7, shape, dt)
            assert np.array_equal(A.to_host(), ref), (dt, shape)


def test_fill_unaligned_pointer(gpu):
    """dfft_fill_uniform at an address that is not 16-byte aligned (one
    scalar into a buffer) takes the one-scalar-per-lane kernel; the values
    are the same as at an aligned address."""
    import ctypes
    import torch
    from daggerfft import _lib as L
    for dt, prec, cplx in ((np.complex128, 1, 1), (np.complex64, 0, 1), (np.float64, 1, 0), (np.float32, 0, 0)):
        n = 100003
        es = np.dtype(dt).itemsize // (2 if cplx else 1)
        buf = torch.zeros((n + 4) * np.dtype(dt).itemsize, dtype=torch.uint8, device="cuda")
        L.check(L.lib().dfft_fill_uniform(ctypes.c_void_p(buf.data_ptr() + es), n, cplx, prec, ctypes.c_uint64(9),
                                          None), "dfft_fill_uniform")
        torch.cuda.synchronize()
        got = buf.cpu().numpy()[es:es + n * np.dtype(dt).itemsize].view(dt)
        assert np.array_equal(got, O.fill(9, (n,), dt)), dt


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 12, 10), (7, 5, 3), (64, 48, 30), (128, 128, 128),
                                   (2, 3, 5), (1, 4, 6), (32, 1, 9)])
@pytest.mark.parametrize("decomp", ["pencil", "slab"])
def test_c2c_roundtrip_vs_oracle(gpu, shape, decomp):
    import daggerfft as D
    dec = D.Pencil() if decomp == "pencil" else D.Slab()
    A = D.rand(shape, np.complex128, seed=1)
    a = O.fill(1, shape, np.complex128)
    F = D.fft(A, decomp=dec, dims=(1, 2, 3))
    ref = O.fftn(a)
    e = O.rel_l2(F.to_host(), ref)
    print("shape", shape, decomp, "fwd rel-L2", e)
    assert e <= TOL64
    B = D.ifft(F, decomp=dec, dims=(1, 2, 3))
    e2 = O.rel_l2(B.to_host(), O.fftn(ref, inverse=True))
    e3 = O.rel_l2(B.to_host(), a)
    print("inv rel-L2", e2, "round trip", e3)
    assert e2 <= TOL64 and e3 <= TOL64


@pytest.mark.parametrize("dims", [(1,), (2,), (3,), (1, 2), (2, 3), (3, 1, 2), (2, 1)])
def test_c2c_dims_subsets(gpu, dims):
    import daggerfft as D
    shape = (24, 20, 14)
    a = O.fill(3, shape, np.complex128)
    F = D.fft(_dev(a), dims=dims)
    assert O.rel_l2(F.to_host(), O.fftn(a, dims=dims)) <= TOL64
    B = D.ifft(F, dims=dims)
    assert O.rel_l2(B.to_host(), a) <= TOL64


@pytest.mark.parametrize("shape", [(256,), (384, 240), (96, 160, 192)])
def test_c2c_rank_and_mixed_radix(gpu, shape):
    import daggerfft as D
    a = O.fill(5, shape, np.complex128)
    F = D.fft(_dev(a))
    assert O.rel_l2(F.to_host(), O.fftn(a)) <= TOL64


def test_c2c_mixed_radix_384x240x210(gpu):
    import daggerfft as D
    shape = (384, 240, 210)
    A = D.rand(shape, np.complex128, seed=6)
    a = O.fill(6, shape, np.complex128)
    F = D.fft(A, decomp=D.Slab())
    ref = O.fftn(a)
    e = O.rel_l2(F.to_host(), ref)
    print("384x240x210 fwd rel-L2", e)
    assert e <= TOL64
    B = D.ifft(F, decomp=D.Slab())
    assert O.rel_l2(B.to_host(), a) <= TOL64


@pytest.mark.parametrize("shape", [(64, 64, 64), (48, 96, 160), (210, 7, 5)])
def test_c2c_f32_error_vs_f64_oracle(gpu, shape):
    import daggerfft as D
    A = D.rand(shape, np.complex64, seed=9)
    a = O.fill(9, shape, np.complex64).astype(np.complex128)
    F = D.fft(A)
    e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a))
    print("f32", shape, "rel-L2 vs f64 oracle", e)
    assert e <= TOL32
    B = D.ifft(F)
    assert O.rel_l2(B.to_host().astype(np.complex128), a) <= TOL32


def test_c2c_inplace_and_graph(gpu):
    import daggerfft as D
    from daggerfft import _lib as L
    shape = (32, 24, 20)
    a = O.fill(11, shape, np.complex128)
    A = _dev(a)
    p = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    p.execute(A, A)
    assert O.rel_l2(A.to_host(), O.fftn(a)) <= TOL64
    pg = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3), flags=L.FLAG_GRAPH)
    X = _dev(a)
    Y = X.similar()
    import torch
    for _ in range(3):  # replays, with eager work between them
        pg.execute(X, Y)
        torch.cuda.synchronize()
        assert O.rel_l2(Y.to_host(), O.fftn(a)) <= TOL64
        Y.tensor.zero_()


@pytest.mark.parametrize("shape", [(64, 48, 40), (30, 14, 9), (128, 128, 128)])
def test_slab_plane_pipeline_matches_pencil(gpu, shape):
    """Slab plans with the plane-pipeline flag software-pipeline the 2D stage
    over chunks of z-planes in grouped launches; results equal the per-pass
    Pencil schedule and the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    a = O.fill(17, shape, np.complex128)
    A = _dev(a)
    ps = D.Plan(shape, L.C2C, -1, L.F64, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    assert ps.schedule()[0] == "plane_pipeline"
    pp = D.Plan(shape, L.C2C, -1, L.F64, D.Pencil(), (1, 2, 3))
    assert pp.schedule()[0] == "passes"
    Fs = ps.execute(A, A.similar())
    Fp = pp.execute(A, A.similar())
    assert O.rel_l2(Fs.to_host(), Fp.to_host()) < 1e-14  # grouped vs catalog kernels: rounding only
    assert O.rel_l2(Fs.to_host(), O.fftn(a)) <= TOL64
    pi = D.Plan(shape, L.C2C, +1, L.F64, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    assert pi.schedule()[0] == "plane_pipeline"
    B = pi.execute(Fs, Fs.similar())
    assert O.rel_l2(B.to_host(), a) <= TOL64
    # f32 through the same grouped kernel
    a32 = O.fill(18, shape, np.complex64)
    p32 = D.Plan(shape, L.C2C, -1, L.F32, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    F32 = p32.execute(_dev(a32), D.DeviceArray(shape, np.complex64))
    assert O.rel_l2(F32.to_host().astype(np.complex128), O.fftn(a32.astype(np.complex128))) <= TOL32


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape,planes", [((64, 48, 40), 1), ((30, 14, 9), 3), ((32, 24, 20), 7)])
def test_slab_plane_pipeline_many_chunks(gpu, monkeypatch, shape, planes, prec):
    """The plane pipeline with chunks of a few z-planes (DFFT_SLAB_CHUNK_BYTES
    set to `planes` planes of the first pass's output), so every chunked pass
    runs at nonzero plane offsets: C2C forward and inverse, R2C and C2R,
    against the oracle."""
    import daggerfft as D
    from daggerfft import _lib as L
    P = L.F64 if prec == "f64" else L.F32
    cdt = np.complex128 if prec == "f64" else np.complex64
    rdt = np.float64 if prec == "f64" else np.float32
    tol = TOL64 if prec == "f64" else TOL32
    es = 16 if prec == "f64" else 8
    nh = shape[0] // 2 + 1
    monkeypatch.setenv("DFFT_SLAB_CHUNK_BYTES", str(planes * shape[0] * shape[1] * es))

    def chk(y, ref):
        e = O.rel_l2(np.asarray(y, dtype=np.complex128 if np.iscomplexobj(y) else np.float64), ref)
        print(shape, planes, prec, "rel-L2", e)
        assert e <= tol

    a = O.fill(31, shape, cdt)
    a64 = a.astype(np.complex128)
    pf = D.Plan(shape, L.C2C, -1, P, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    pi = D.Plan(shape, L.C2C, +1, P, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    assert pf.schedule()[0] == "plane_pipeline" and pi.schedule()[0] == "plane_pipeline"
    assert pf.schedule()[1] > 4 and pi.schedule()[1] > 4  # several chunks
    F = pf.execute(_dev(a), D.DeviceArray(shape, cdt))
    chk(F.to_host(), O.fftn(a64))
    B = pi.execute(_dev(a), D.DeviceArray(shape, cdt))
    chk(B.to_host(), O.fftn(a64, inverse=True))

    r = O.fill(32, shape, rdt)
    r64 = r.astype(np.float64)
    pr = D.Plan(shape, L.R2C, -1, P, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    assert pr.schedule()[0] == "plane_pipeline"
    Fr = pr.execute(_dev(r), D.DeviceArray((nh,) + shape[1:], cdt))
    ref = O.rfftn(r64)
    chk(Fr.to_host(), ref)
    pc = D.Plan(shape, L.C2R, +1, P, D.Slab(), (1, 2, 3), flags=L.FLAG_PLANE_PIPELINE)
    assert pc.schedule()[0] == "plane_pipeline"
    spec = np.asfortranarray(ref.astype(cdt))
    Br = pc.execute(_dev(spec), D.DeviceArray(shape, rdt))
    chk(Br.to_host(), O.irfftn(spec.astype(np.complex128), shape[0]))


def test_first_stage_load_order(gpu):
    """The first stage's loads of a catalog C2C pass are all issued at once,
    except along ComplexF64 lines whose stride is a power of two of 8 KiB or
    more (AxisPass::paired_loads, shown by dfft_plan_describe).  A plan whose
    passes take both orders matches the oracle forward and back; ComplexF32
    takes the all-at-once order on every pass."""
    import daggerfft as D
    from daggerfft import _lib as L
    expect = {
        ((512, 512, 512), L.F64): {0: 0, 1: 1, 2: 1},      # strides 16 B, 8 KiB, 4 MiB
        ((512, 512, 512), L.F32): {0: 0, 1: 0, 2: 0},      # ComplexF32: always all at once
        ((384, 240, 210), L.F64): {0: 0, 1: 0, 2: 0},      # 6 KiB, 1.4 MiB: not powers of two
    }
    for (shape, prec), want in expect.items():
        p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        got = {q["axis"]: q["paired_loads"] for q in p.passes()}
        assert got == want, (shape, prec, got)
    for prec, dt, tol in ((L.F64, np.complex128, TOL64), (L.F32, np.complex64, TOL32)):
        shape = (32, 32, 48)  # axis 3: 1024 elements apart (16 KiB in ComplexF64), axis 2: 32
        p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        orders = {q["axis"]: q["paired_loads"] for q in p.passes()}
        assert orders == {0: 0, 1: 0, 2: 1 if prec == L.F64 else 0}, orders
        a = O.fill(61, shape, dt)
        F = p.execute(D.DeviceArray.from_host(a), D.DeviceArray(shape, dt))
        e = O.rel_l2(F.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128)))
        B = D.Plan(shape, L.C2C, +1, prec, D.Pencil(), (1, 2, 3)).execute(F, F.similar())
        e2 = O.rel_l2(B.to_host().astype(np.complex128), a.astype(np.complex128))
        assert e <= tol and e2 <= tol, (prec, e, e2)


@pytest.mark.parametrize("prec", [1, 0], ids=["f64", "f32"])
def test_inplace_partial_tiles(gpu, prec):
    """In place, with tiles that run past the array's last line: the first
    stage of a C2C pass, and the copy-in of a Rader pass, read line 0 of the
    pass (or element 0) for the lines past the end rather than zeros -- values
    that are never stored, even while line 0 is being overwritten by its own
    tile.  Forward then inverse in place, against the oracle, on shapes whose
    passes end in partial tiles (catalog lengths and a Rader length, 257)."""
    import daggerfft as D
    from daggerfft import _lib as L
    dt = np.complex128 if prec else np.complex64
    tol = TOL64 if prec else TOL32
    for shape in ((96, 33, 7), (7, 96, 33), (257, 5, 3), (5, 257, 3)):
        print("in place, partial tiles", shape, "f64" if prec else "f32", flush=True)
        p = D.Plan(shape, L.C2C, -1, prec, D.Pencil(), (1, 2, 3))
        n = int(np.prod(shape))
        partial = [q["axis"] for q in p.passes() if q["blocks"] * q["TC"] * q["L"] != n]
        assert partial, (shape, p.passes())
        a = O.fill(71, shape, dt)
        A = D.DeviceArray.from_host(a)
        p.execute(A, A)
        e = O.rel_l2(A.to_host().astype(np.complex128), O.fftn(a.astype(np.complex128)))
        D.Plan(shape, L.C2C, +1, prec, D.Pencil(), (1, 2, 3)).execute(A, A)
        e2 = O.rel_l2(A.to_host().astype(np.complex128), a.astype(np.complex128))
        assert e <= tol and e2 <= tol, (shape, e, e2)
