"""CPU oracle for the DaggerFFT 3D transform path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  The product
library never calls it.

Arrays are numpy arrays in Julia (column-major, Fortran) order: an array of
Julia size ``(n1, n2, n3)`` is ``np.asfortranarray`` with that shape.  ``dims``
are 1-based, as in the reference.

What is restated, with the reference lines it follows:

* ``apply_fft!`` on one chunk: an in-place batched 1D transform along one
  axis (``src/fft.jl:126-174``); the arithmetic (FFTW's definitions) is in
  ``dfft_oracle.c``.
* ``fft!`` / ``ifft!`` Pencil: three stages in the order ``dims[1], dims[2],
  dims[3]`` (inverse: ``dims[3], dims[2], dims[1]``) with a chunk
  redistribution between stages (``src/fft.jl:650-724``), blockings
  ``(N,c,c) -> (c,N,c) -> (c,c,N)`` as in ``benchmarks/fftmpi.jl:35-41``.
* ``fft!`` / ``ifft!`` Slab: a 2D stage over ``(dims[1], dims[2])`` with
  ``transforms[1]`` then a 1D stage over ``dims[3]`` with ``transforms[3]``;
  ``transforms[2]`` is ignored (``src/fft.jl:726-776``).  Blockings
  ``(N,N,c) -> (c,N,N)`` (``benchmarks/fftmpi.jl:18-24``).
* the redistribution's box overlaps and copies (``compute_overlap``
  ``src/fft.jl:176-178``; ``local_copy!`` ``src/fft.jl:216-231``;
  ``pack_data!``/``unpack_data!`` ``src/fft.jl:180-214``).
* the README's high-level ``fft/ifft/rfft/irfft`` with ``R2R`` kinds
  (``README.md:29-67``), implemented from FFTW's documented definitions.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libdfft_oracle.so")
_lib = None
_lock = threading.Lock()

# FFTW fftw_r2r_kind numbering (FFTW's api/fftw3.h enum order)
R2HC, HC2R, DHT, REDFT00, REDFT01, REDFT10, REDFT11, RODFT00, RODFT01, RODFT10, RODFT11 = range(11)
R2R_KINDS = (DHT, REDFT00, REDFT01, REDFT10, REDFT11, RODFT00, RODFT01, RODFT10, RODFT11)


def build():
    """Compile the C restatement (gcc + OpenMP) into oracle/libdfft_oracle.so."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            L = ctypes.CDLL(_LIB_PATH)
            P = ctypes.c_void_p
            I64 = ctypes.c_int64
            I64P = ctypes.POINTER(ctypes.c_int64)
            L.dfo_splitmix64.restype = ctypes.c_uint64
            L.dfo_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
            L.dfo_uniform.restype = ctypes.c_double
            L.dfo_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
            for f in ("dfo_fill_f64", "dfo_fill_c128", "dfo_fill_f32", "dfo_fill_c64"):
                getattr(L, f).argtypes = [ctypes.c_uint64, I64, P]
                getattr(L, f).restype = None
            L.dfo_fft1d.argtypes = [P, P, I64, ctypes.c_int]
            L.dfo_dft_naive.argtypes = [P, P, I64, ctypes.c_int]
            L.dfo_c2c_axis.argtypes = [P, I64P, ctypes.c_int, ctypes.c_int, ctypes.c_double]
            L.dfo_r2c_axis.argtypes = [P, P, I64P, ctypes.c_int]
            L.dfo_c2r_axis.argtypes = [P, P, I64P, ctypes.c_int, ctypes.c_double]
            L.dfo_r2r_axis.argtypes = [P, I64P, ctypes.c_int, ctypes.c_int, ctypes.c_double]
            L.dfo_pack.argtypes = [P, P, I64P, I64P, I64P, I64, I64]
            L.dfo_unpack.argtypes = [P, P, I64P, I64P, I64P, I64, I64]
            L.dfo_local_copy.argtypes = [P, I64P, I64P, P, I64P, I64P, I64P, I64]
            L.dfo_permute.argtypes = [P, P, I64P, ctypes.POINTER(ctypes.c_int), I64]
            L.dfo_spot_bins.argtypes = [P, I64P, I64P, I64, ctypes.c_int, P]
            L.dfo_spot_bins.restype = None
            L.dfo_num_threads.restype = ctypes.c_int
            L.dfo_set_num_threads.argtypes = [ctypes.c_int]
            _lib = L
    return _lib


def _shape3(shape):
    s = list(shape) + [1] * (3 - len(shape))
    return (ctypes.c_int64 * 3)(*s)


def _i64x3(v):
    v = list(v) + [0] * (3 - len(v))
    return (ctypes.c_int64 * 3)(*v)


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def num_threads():
    return lib().dfo_num_threads()


def set_num_threads(n):
    lib().dfo_set_num_threads(int(n))


# ----------------------------------------------------------------------------
# inputs: SplitMix64 counter generator; complex element e -> (u(2e), u(2e+1))
# ----------------------------------------------------------------------------
def uniform(seed, i):
    return lib().dfo_uniform(seed, i)


def fill(seed, shape, dtype):
    """Seeded uniform[0,1) array of Julia size `shape` (Fortran order)."""
    n = int(np.prod(shape)) if len(shape) else 1
    dtype = np.dtype(dtype)
    flat = np.empty(n, dtype=dtype)
    f = {np.dtype(np.float64): lib().dfo_fill_f64, np.dtype(np.complex128): lib().dfo_fill_c128,
         np.dtype(np.float32): lib().dfo_fill_f32, np.dtype(np.complex64): lib().dfo_fill_c64}[dtype]
    f(seed, n, _ptr(flat))
    return flat.reshape(shape, order="F")


# ----------------------------------------------------------------------------
# per-axis transforms (one chunk's apply_fft!, src/fft.jl:168-174)
# ----------------------------------------------------------------------------
def _f(a, dtype):
    return np.require(a, dtype=dtype, requirements=["F_CONTIGUOUS", "WRITEABLE", "ALIGNED"])


def c2c_axis(a, axis, sign, scale=1.0):
    """In-place batched C2C along 0-based `axis` (forward sign=-1)."""
    assert a.dtype == np.complex128 and a.flags.f_contiguous and a.ndim <= 3
    rc = lib().dfo_c2c_axis(_ptr(a), _shape3(a.shape), axis, sign, scale)
    assert rc == 0
    return a


def r2c_axis(a, axis):
    a = _f(a, np.float64)
    oshape = list(a.shape)
    oshape[axis] = oshape[axis] // 2 + 1
    out = np.empty(oshape, dtype=np.complex128, order="F")
    assert lib().dfo_r2c_axis(_ptr(a), _ptr(out), _shape3(a.shape), axis) == 0
    return out


def c2r_axis(a, axis, n, scale=1.0):
    a = _f(a, np.complex128)
    oshape = list(a.shape)
    assert oshape[axis] == n // 2 + 1
    oshape[axis] = n
    out = np.empty(oshape, dtype=np.float64, order="F")
    assert lib().dfo_c2r_axis(_ptr(a), _ptr(out), _shape3(oshape), axis, scale) == 0
    return out


def r2r_axis(a, axis, kind, scale=1.0):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    rc = lib().dfo_r2r_axis(_ptr(a), _shape3(a.shape), axis, kind, scale)
    assert rc == 0, "unsupported r2r kind %r" % kind
    return a


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    lib().dfo_fft1d(_ptr(x), _ptr(y), x.size, sign)
    return y


def dft_naive(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(x)
    lib().dfo_dft_naive(_ptr(x), _ptr(y), x.size, sign)
    return y


# ----------------------------------------------------------------------------
# box maps (src/fft.jl:180-231), 0-based lo / counts
# ----------------------------------------------------------------------------
def pack(buffer, src, lo, cnt, offset=0):
    assert src.flags.f_contiguous and buffer.dtype == src.dtype
    lib().dfo_pack(_ptr(buffer), _ptr(src), _shape3(src.shape), _i64x3(lo), _i64x3(cnt),
                   offset, src.dtype.itemsize)
    return buffer


def unpack(dst, buffer, lo, cnt, offset=0):
    assert dst.flags.f_contiguous and buffer.dtype == dst.dtype
    lib().dfo_unpack(_ptr(dst), _ptr(buffer), _shape3(dst.shape), _i64x3(lo), _i64x3(cnt),
                     offset, dst.dtype.itemsize)
    return dst


def local_copy(dst, dlo, src, slo, cnt):
    assert dst.flags.f_contiguous and src.flags.f_contiguous and dst.dtype == src.dtype
    lib().dfo_local_copy(_ptr(dst), _shape3(dst.shape), _i64x3(dlo), _ptr(src), _shape3(src.shape),
                         _i64x3(slo), _i64x3(cnt), dst.dtype.itemsize)
    return dst


def permute(a, perm):
    """out[idx[perm[0]], idx[perm[1]], idx[perm[2]]] = a[idx] (0-based perm)."""
    assert a.flags.f_contiguous and a.ndim == 3
    oshape = tuple(a.shape[p] for p in perm)
    out = np.empty(oshape, dtype=a.dtype, order="F")
    lib().dfo_permute(_ptr(out), _ptr(a), _shape3(a.shape), (ctypes.c_int * 3)(*perm), a.dtype.itemsize)
    return out


# ----------------------------------------------------------------------------
# Blocked arrays and the redistribution (create_workspace local branch +
# flat_redistribute! LOCAL_COPIES; src/fft.jl:253-305, 453-462)
# ----------------------------------------------------------------------------
class Blocked:
    """A single-process stand-in for a DArray with Blocks(b1,b2,b3): one
    Fortran-ordered numpy chunk per block, domains as 0-based half-open boxes."""

    def __init__(self, shape, blocks, dtype):
        self.shape = tuple(shape)
        self.blocks = tuple(min(b, n) for b, n in zip(blocks, shape))
        self.dtype = np.dtype(dtype)
        grids = [range(0, n, b) for n, b in zip(self.shape, self.blocks)]
        self.domains = []
        self.chunks = []
        for k0 in grids[2]:
            for j0 in grids[1]:
                for i0 in grids[0]:
                    lo = (i0, j0, k0)
                    hi = tuple(min(l + b, n) for l, b, n in zip(lo, self.blocks, self.shape))
                    self.domains.append((lo, hi))
                    self.chunks.append(np.zeros(tuple(h - l for l, h in zip(lo, hi)), dtype=self.dtype, order="F"))

    @classmethod
    def distribute(cls, A, blocks):
        B = cls(A.shape, blocks, A.dtype)
        full = np.asfortranarray(A)
        for (lo, hi), ch in zip(B.domains, B.chunks):
            cnt = tuple(h - l for l, h in zip(lo, hi))
            local_copy(ch, (0, 0, 0), full, lo, cnt)
        return B

    def collect(self):
        out = np.zeros(self.shape, dtype=self.dtype, order="F")
        for (lo, hi), ch in zip(self.domains, self.chunks):
            cnt = tuple(h - l for l, h in zip(lo, hi))
            local_copy(out, lo, ch, (0, 0, 0), cnt)
        return out


def redistribute(dst: Blocked, src: Blocked):
    """Every overlap of a source chunk with a destination chunk is copied
    box-to-box (compute_overlap + local_copy!)."""
    for (slo, shi), sch in zip(src.domains, src.chunks):
        for (dlo, dhi), dch in zip(dst.domains, dst.chunks):
            lo = tuple(max(a, b) for a, b in zip(slo, dlo))
            hi = tuple(min(a, b) for a, b in zip(shi, dhi))
            if any(h <= l for l, h in zip(lo, hi)):
                continue
            cnt = tuple(h - l for l, h in zip(lo, hi))
            local_copy(dch, tuple(l - d for l, d in zip(lo, dlo)),
                       sch, tuple(l - s for l, s in zip(lo, slo)), cnt)


def _apply(chunk, transform, dims):
    """apply_fft!(chunk, transform, dims) with FFTW.jl semantics."""
    if isinstance(dims, int):
        dims = (dims,)
    for d in dims:
        ax = d - 1
        if transform == "FFT!":
            c2c_axis(chunk, ax, -1, 1.0)
        elif transform == "IFFT!":
            c2c_axis(chunk, ax, +1, 1.0 / chunk.shape[ax])
        else:
            raise ValueError("Unknown transform type")


def fft_pencil(A, transforms=("FFT!", "FFT!", "FFT!"), dims=(1, 2, 3), chunk=None):
    """fft! Pencil (src/fft.jl:651-688): 3 stages, 2 redistributions."""
    N = A.shape
    c = chunk or max(N)
    DA = Blocked.distribute(np.asarray(A, dtype=np.complex128), (N[0], c, c))
    DB = Blocked(N, (c, N[1], c), np.complex128)
    DC = Blocked(N, (c, c, N[2]), np.complex128)
    for ch in DA.chunks:
        _apply(ch, transforms[0], dims[0])
    redistribute(DB, DA)
    for ch in DB.chunks:
        _apply(ch, transforms[1], dims[1])
    redistribute(DC, DB)
    for ch in DC.chunks:
        _apply(ch, transforms[2], dims[2])
    return DC.collect()


def ifft_pencil(C, transforms=("IFFT!", "IFFT!", "IFFT!"), dims=(1, 2, 3), chunk=None):
    """ifft! Pencil (src/fft.jl:691-724): dims[3] on z-pencils, then dims[2], then dims[1]."""
    N = C.shape
    c = chunk or max(N)
    DA = Blocked.distribute(np.asarray(C, dtype=np.complex128), (c, c, N[2]))
    DB = Blocked(N, (c, N[1], c), np.complex128)
    DC = Blocked(N, (N[0], c, c), np.complex128)
    for ch in DA.chunks:
        _apply(ch, transforms[2], dims[2])
    redistribute(DB, DA)
    for ch in DB.chunks:
        _apply(ch, transforms[1], dims[1])
    redistribute(DC, DB)
    for ch in DC.chunks:
        _apply(ch, transforms[0], dims[0])
    return DC.collect()


def fft_slab(A, transforms=("FFT!", "FFT!", "FFT!"), dims=(1, 2, 3), chunk=None):
    """fft! Slab (src/fft.jl:727-750): 2D stage over (dims[1],dims[2]) with
    transforms[1], redistribute, 1D stage over dims[3] with transforms[3]."""
    N = A.shape
    c = chunk or max(N)
    DA = Blocked.distribute(np.asarray(A, dtype=np.complex128), (N[0], N[1], c))
    DB = Blocked(N, (c, N[1], N[2]), np.complex128)
    for ch in DA.chunks:
        _apply(ch, transforms[0], (dims[0], dims[1]))
    redistribute(DB, DA)
    for ch in DB.chunks:
        _apply(ch, transforms[2], dims[2])
    return DB.collect()


def ifft_slab(B, transforms=("IFFT!", "IFFT!", "IFFT!"), dims=(1, 2, 3), chunk=None):
    """ifft! Slab (src/fft.jl:753-776)."""
    N = B.shape
    c = chunk or max(N)
    DB = Blocked.distribute(np.asarray(B, dtype=np.complex128), (c, N[1], N[2]))
    DA = Blocked(N, (N[0], N[1], c), np.complex128)
    for ch in DB.chunks:
        _apply(ch, transforms[2], dims[2])
    redistribute(DA, DB)
    for ch in DA.chunks:
        _apply(ch, transforms[0], (dims[0], dims[1]))
    return DA.collect()


# ----------------------------------------------------------------------------
# README-level API semantics (README.md:29-67), FFTW conventions
# ----------------------------------------------------------------------------
def fftn(A, dims=None, inverse=False):
    """fft(A; dims) / ifft(A; dims): per-axis C2C in the reference's stage
    order (forward dims[1..], inverse reversed), ifft scaled by 1/n per dim."""
    a = np.array(A, dtype=np.complex128, order="F", copy=True)
    dims = tuple(dims) if dims is not None else tuple(range(1, a.ndim + 1))
    order = reversed(dims) if inverse else dims
    for d in order:
        ax = d - 1
        if inverse:
            c2c_axis(a, ax, +1, 1.0 / a.shape[ax])
        else:
            c2c_axis(a, ax, -1, 1.0)
    return a


def rfftn(A, dims=None):
    """rfft(A; dims): r2c along dims[1] (halved to n/2+1), then C2C forward along the rest."""
    a = np.asfortranarray(A, dtype=np.float64)
    dims = tuple(dims) if dims is not None else tuple(range(1, a.ndim + 1))
    out = r2c_axis(a, dims[0] - 1)
    for d in dims[1:]:
        c2c_axis(out, d - 1, -1, 1.0)
    return out


def irfftn(F, n, dims=None):
    """irfft(F, n; dims): C2C backward along dims[end..2], then c2r of length n
    along dims[1]; the whole transform is scaled by 1/prod(n_d)."""
    a = np.array(F, dtype=np.complex128, order="F", copy=True)
    dims = tuple(dims) if dims is not None else tuple(range(1, a.ndim + 1))
    for d in reversed(dims[1:]):
        ax = d - 1
        c2c_axis(a, ax, +1, 1.0 / a.shape[ax])
    return c2r_axis(a, dims[0] - 1, n, 1.0 / n)


def r2rn(A, kinds, dims=None, scale=1.0):
    """FFTW r2r with one kind per transformed dim (unnormalised), times `scale`."""
    a = np.array(A, dtype=np.float64, order="F", copy=True)
    dims = tuple(dims) if dims is not None else tuple(range(1, a.ndim + 1))
    if isinstance(kinds, int):
        kinds = (kinds,) * len(dims)
    for i, (d, k) in enumerate(zip(dims, kinds)):
        r2r_axis(a, d - 1, k, scale if i == 0 else 1.0)
    return a


def r2r_logical_size(kind, n):
    """FFTW's logical DFT size N for an r2r kind; the inverse pair composes to N*I."""
    if kind == REDFT00:
        return 2 * (n - 1)
    if kind == RODFT00:
        return 2 * (n + 1)
    if kind == DHT:
        return n
    return 2 * n


def spot_bins(a, bins, sign=-1):
    """Bins (k0, k1, k2) of the unnormalised 3D DFT of complex `a` (Julia
    size n0 x n1 x n2) by direct separable summation in long double
    (dfo_spot_bins): the reference value at sizes where a full oracle
    transform is too slow."""
    a = np.asfortranarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(np.asarray(bins, dtype=np.int64).reshape(-1, 3))
    out = np.empty(len(b), dtype=np.complex128)
    lib().dfo_spot_bins(_ptr(a), _shape3(a.shape), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(b),
                        int(sign), _ptr(out))
    return out


def rel_l2(y, ref):
    y = np.asarray(y)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((y - ref).ravel())
    return float(num / den) if den > 0 else float(num)
