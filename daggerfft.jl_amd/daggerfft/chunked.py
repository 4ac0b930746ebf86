"""Chunked device arrays on one device and the redistribution between them.

A ``DArray`` here is the reference's DArray with ``Blocks(b1, b2, b3)``
chunks, every chunk on this device: one column-major device array per block,
in the reference's chunk order (dim 1 fastest).  Domains are 0-based
half-open ``range`` triples (the reference's 1-based ``UnitRange``s shifted
by one).

  compute_overlap       src/fft.jl:176-178 (box intersection)
  create_workspace      src/fft.jl:233-343, the branch where source and
                        destination chunks share a rank: every overlap is a
                        local CoalescedPattern (src/fft.jl:301-305)
  redistribute          flat_redistribute!'s LOCAL_COPIES step
                        (src/fft.jl:453-462): each pattern moved by
                        dfft_box_copy, the device form of local_copy!
                        (src/fft.jl:216-231, src/fftgpu.jl:482-495)
  fft_ / ifft_          the Pencil and Slab stage drivers over chunked arrays
                        (src/fft.jl:651-776): per-chunk transforms along each
                        stage's dims (apply_fft!, src/fft.jl:168-174), in
                        place, with the redistributions between stages

On one device a single chunk per array is the fast path (the plain
``daggerfft.fft_``); this module serves callers whose arrays are already
blocked.
"""
from __future__ import annotations

import collections
import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, lib

CoalescedPattern = collections.namedtuple(
    "CoalescedPattern",
    "src_idx dst_idx overlap src_indices dst_indices buffer_offset buffer_size target_node")
CoalescedPattern.__doc__ = ("One box overlap of a source and a destination chunk (src/fft.jl:66-75); "
                            "indices are 0-based ranges, *_indices local to the chunk.")


class ArrayDomain:
    """The global index box of a chunk: ``indexes`` = three 0-based ranges."""

    def __init__(self, indexes):
        self.indexes = tuple(range(r.start, r.stop) for r in indexes)

    @property
    def shape(self):
        return tuple(len(r) for r in self.indexes)

    def __repr__(self):
        return "ArrayDomain(%s)" % ", ".join("%d:%d" % (r.start, r.stop) for r in self.indexes)


def compute_overlap(src_domain, dst_domain):
    """Per-dim intersection of two domains (src/fft.jl:176-178); an empty
    range in any dim means no overlap."""
    return tuple(range(max(a.start, b.start), max(max(a.start, b.start), min(a.stop, b.stop)))
                 for a, b in zip(src_domain.indexes, dst_domain.indexes))


def block_domains(shape, blocks):
    """The chunk domains of Blocks(b1, b2, b3) over `shape`, dim 1 fastest
    (the reference's chunk order); edge blocks are cut at the array's end."""
    shape = tuple(int(n) for n in shape)
    blocks = tuple(max(1, min(int(b), n)) for b, n in zip(blocks, shape))
    starts = [range(0, n, b) for n, b in zip(shape, blocks)]
    return [ArrayDomain(range(l, min(l + b, n)) for l, b, n in zip((i0, j0, k0), blocks, shape))
            for k0 in starts[2] for j0 in starts[1] for i0 in starts[0]]


class DArray:
    """Blocks(b1, b2, b3) of an n1 x n2 x n3 array, one device chunk per block."""

    def __init__(self, shape, blocks, dtype, device=None):
        from . import DeviceArray
        self.shape = tuple(int(n) for n in shape)
        if len(self.shape) != 3 or len(blocks) != 3:
            raise ValueError("DArray: rank-3 shape and blocks")
        self.blocks = tuple(max(1, min(int(b), n)) for b, n in zip(blocks, self.shape))
        self.dtype = np.dtype(dtype)
        self.subdomains = block_domains(self.shape, self.blocks)
        self.chunks = [DeviceArray(dom.shape, self.dtype, device=device) for dom in self.subdomains]

    @property
    def precision(self):
        return self.chunks[0].precision

    @classmethod
    def distribute(cls, A, blocks, stream=None):
        """distribute(A, Blocks(...)): A is a DeviceArray (or a host array)."""
        from . import DeviceArray
        if not isinstance(A, DeviceArray):
            A = DeviceArray.from_host(A)
        D = cls(A.shape, blocks, A.dtype)
        for dom, ch in zip(D.subdomains, D.chunks):
            _box_copy(ch, (0, 0, 0), A, tuple(r.start for r in dom.indexes), dom.shape, stream)
        return D

    def collect(self, stream=None):
        """The whole array as one DeviceArray."""
        from . import DeviceArray
        out = DeviceArray(self.shape, self.dtype)
        for dom, ch in zip(self.subdomains, self.chunks):
            _box_copy(out, tuple(r.start for r in dom.indexes), ch, (0, 0, 0), dom.shape, stream)
        return out


def _box_copy(dst, dlo, src, slo, cnt, stream=None):
    from . import _stream_handle
    if any(c == 0 for c in cnt):
        return
    check(lib().dfft_box_copy(ctypes.c_void_p(dst.ptr), L.i64(dst.shape), L.i64(dlo), ctypes.c_void_p(src.ptr),
                              L.i64(src.shape), L.i64(slo), L.i64(cnt), dst.dtype.itemsize,
                              _stream_handle(stream)), "dfft_box_copy")


class Workspace:
    """FFTWorkspace of a one-device redistribution: only local patterns
    (src/fft.jl:301-305); the send/receive peer maps of the reference are
    empty because every chunk lives on this device."""

    def __init__(self, src, dst, local_patterns):
        self.src_shape = tuple(src.shape)
        self.dst_shape = tuple(dst.shape)
        self.local_patterns = list(local_patterns)
        self.send_peers = {}
        self.recv_peers = {}


def create_workspace(src, dst, use_hierarchical=True, ranks_per_node=0):
    """create_workspace(src, dst) (src/fft.jl:233-343): every pair of a
    destination and a source chunk with a non-empty overlap, in the
    reference's loop order (destination chunks outer, source chunks inner)."""
    if tuple(src.shape) != tuple(dst.shape) or src.dtype != dst.dtype:
        raise ValueError("create_workspace: arrays must share shape and element type")
    pats = []
    for dst_idx, dst_dom in enumerate(dst.subdomains):
        for src_idx, src_dom in enumerate(src.subdomains):
            ov = compute_overlap(src_dom, dst_dom)
            if any(len(r) == 0 for r in ov):
                continue
            src_ind = tuple(range(o.start - s.start, o.stop - s.start) for o, s in zip(ov, src_dom.indexes))
            dst_ind = tuple(range(o.start - d.start, o.stop - d.start) for o, d in zip(ov, dst_dom.indexes))
            size = int(np.prod([len(r) for r in ov]))
            pats.append(CoalescedPattern(src_idx, dst_idx, ov, src_ind, dst_ind, 0, size, -1))
    return Workspace(src, dst, pats)


create_gpu_workspace = create_workspace


def redistribute(dst, src, ws, stream=None):
    """coalesced_redistribute!(dst, src, ws) on one device: every local
    pattern is one box copy (flat_redistribute! LOCAL_COPIES, src/fft.jl:453-462)."""
    for p in ws.local_patterns:
        _box_copy(dst.chunks[p.dst_idx], tuple(r.start for r in p.dst_indices), src.chunks[p.src_idx],
                  tuple(r.start for r in p.src_indices), tuple(len(r) for r in p.overlap), stream)


def _apply(chunk, transform, dims, stream=None):
    """apply_fft!(chunk, transform, dims) in place (src/fft.jl:168-174):
    FFT!/IFFT! as C2C (IFFT! scaled by 1/n per dim, FFTW.jl plan_ifft!),
    R2R(kinds) as the r2r kinds of those dims."""
    from . import FFT, FFT_, IFFT, IFFT_, R2R, Pencil, _cached_plan
    dims = tuple(dims) if isinstance(dims, (tuple, list)) else (int(dims),)
    if isinstance(transform, R2R):
        kinds = transform.kinds if len(transform.kinds) == len(dims) else transform.kinds[:1] * len(dims)
        p = _cached_plan(chunk.shape, L.R2R, -1, chunk.precision, Pencil(), dims, kinds)
    elif isinstance(transform, (FFT, FFT_)):
        p = _cached_plan(chunk.shape, L.C2C, -1, chunk.precision, Pencil(), dims)
    elif isinstance(transform, (IFFT, IFFT_)):
        p = _cached_plan(chunk.shape, L.C2C, +1, chunk.precision, Pencil(), dims)
    else:
        raise ValueError("Unknown transform type")
    p.execute(chunk, chunk, stream=stream)


def _stage(D, transform, dims, stream):
    for ch in D.chunks:
        _apply(ch, transform, dims, stream)


def fft_(*args, stream=None):
    """fft!(C, A, B, ws_AB, ws_BC, transforms, dims) [Pencil] or
    fft!(B, A, ws_AB, transforms, dims, Slab()) over chunked arrays
    (src/fft.jl:651-688, 727-750).  Mutates the stage arrays as the
    reference does; returns C (Pencil) or B (Slab)."""
    from . import Slab
    if len(args) == 7:
        C, A, B, ws_ab, ws_bc, transforms, dims = args
        _stage(A, transforms[0], dims[0], stream)
        redistribute(B, A, ws_ab, stream)
        _stage(B, transforms[1], dims[1], stream)
        redistribute(C, B, ws_bc, stream)
        _stage(C, transforms[2], dims[2], stream)
        return C
    if len(args) == 6 and isinstance(args[5], Slab):
        B, A, ws_ab, transforms, dims, _ = args
        _stage(A, transforms[0], (dims[0], dims[1]), stream)
        redistribute(B, A, ws_ab, stream)
        _stage(B, transforms[2], dims[2], stream)
        return B
    raise TypeError("fft_: expected (C, A, B, ws_AB, ws_BC, transforms, dims) or (B, A, ws, transforms, dims, Slab())")


def ifft_(*args, stream=None):
    """ifft!(C, A, B, ws_AB, ws_BC, transforms, dims) [Pencil: dims[3] on A,
    then dims[2] on B, then dims[1] on C] or ifft!(A, B, ws_BA, transforms,
    dims, Slab()) [dims[3] on B, then (dims[1], dims[2]) on A]
    (src/fft.jl:691-724, 753-776)."""
    from . import Slab
    if len(args) == 7:
        C, A, B, ws_ab, ws_bc, transforms, dims = args
        _stage(A, transforms[2], dims[2], stream)
        redistribute(B, A, ws_ab, stream)
        _stage(B, transforms[1], dims[1], stream)
        redistribute(C, B, ws_bc, stream)
        _stage(C, transforms[0], dims[0], stream)
        return C
    if len(args) == 6 and isinstance(args[5], Slab):
        A, B, ws_ba, transforms, dims, _ = args
        _stage(B, transforms[2], dims[2], stream)
        redistribute(A, B, ws_ba, stream)
        _stage(A, transforms[0], (dims[0], dims[1]), stream)
        return A
    raise TypeError("ifft_: expected (C, A, B, ws_AB, ws_BC, transforms, dims) or (A, B, ws, transforms, dims, Slab())")
