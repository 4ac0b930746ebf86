"""daggerfft -- Python host mirror of the DaggerFFT.jl transform API over
libdaggerfft_hip (MI355X / gfx950).

The Julia API this mirrors (README.md:29-67 and the exported bang API,
src/DaggerFFT.jl:32-37):

    F = fft(A; decomp=Pencil(), dims=(1,2,3))             -> fft(A, decomp=Pencil(), dims=(1, 2, 3))
    A = ifft(F; decomp=Pencil(), dims=(1,2,3))            -> ifft(F, ...)
    F = rfft(A; decomp=Pencil(), dims=(1,2,3))            -> rfft(A, ...)
    A = irfft(F, size(A,1); decomp=Pencil(), dims=...)    -> irfft(F, d, ...)
    fft(A; decomp=Slab(), transforms=(R2R((FFTW.REDFT10, FFTW.REDFT10)),), dims=(1,2))
    fft!(C, A, B, ws_AB, ws_BC, transforms, dims)          -> fft_(C, A, B, ws_AB, ws_BC, transforms, dims)
    fft!(B, A, ws, transforms, dims, Slab())               -> fft_(B, A, ws, transforms, dims, Slab())

Arrays are ``DeviceArray`` objects: column-major (Julia layout) device
buffers whose ``shape`` is the Julia size and ``dims`` are 1-based.  Device
memory and streams come from PyTorch (ROCm build); all arithmetic runs in the
library's HIP kernels.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _lib as L
from ._lib import DFFTError, check, lib

__all__ = [
    "DeviceArray", "Pencil", "Slab", "FFT", "IFFT", "RFFT", "IRFFT", "FFT_", "IFFT_", "RFFT_", "IRFFT_",
    "R2R", "DHT", "REDFT00", "REDFT01", "REDFT10", "REDFT11", "RODFT00", "RODFT01", "RODFT10", "RODFT11",
    "fft", "ifft", "rfft", "irfft", "fft_", "ifft_", "plan_fft", "plan_ifft", "plan_rfft", "plan_irfft",
    "plan_r2r", "Plan", "create_workspace", "cleanup_workspace", "DFFTError", "rand", "version",
    "DArray", "compute_overlap",
]

# ---------------------------------------------------------------------------
# tag types (src/fft.jl:1-24)
# ---------------------------------------------------------------------------
R2HC, HC2R, DHT, REDFT00, REDFT01, REDFT10, REDFT11, RODFT00, RODFT01, RODFT10, RODFT11 = range(11)
R2R_SUPPORTED_KINDS = (DHT, REDFT00, REDFT01, REDFT10, REDFT11, RODFT00, RODFT01, RODFT10, RODFT11)


class Decomposition:
    code = L.PENCIL

    def __repr__(self):
        return type(self).__name__ + "()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)


class Pencil(Decomposition):
    code = L.PENCIL


class Slab(Decomposition):
    code = L.SLAB


class _Tag:
    def __repr__(self):
        return type(self).__name__ + "()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)


class FFT(_Tag): pass
class IFFT(_Tag): pass
class RFFT(_Tag): pass
class IRFFT(_Tag): pass
class FFT_(_Tag): pass     # FFT!
class IFFT_(_Tag): pass    # IFFT!
class RFFT_(_Tag): pass    # RFFT!
class IRFFT_(_Tag): pass   # IRFFT!


class R2R(_Tag):
    """R2R((kind1, kind2, ...)): one FFTW r2r kind per transformed dim (README.md:59-66)."""

    def __init__(self, kinds):
        self.kinds = tuple(int(k) for k in (kinds if isinstance(kinds, (tuple, list)) else (kinds,)))
        for k in self.kinds:
            if k not in R2R_SUPPORTED_KINDS:
                raise ValueError("unsupported r2r kind %r" % k)

    def __repr__(self):
        return "R2R(%r)" % (self.kinds,)

    def __eq__(self, other):
        return isinstance(other, R2R) and other.kinds == self.kinds

    def __hash__(self):
        return hash(("R2R", self.kinds))


def version():
    a, b = ctypes.c_int(), ctypes.c_int()
    lib().dfft_version(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


# ---------------------------------------------------------------------------
# device arrays
# ---------------------------------------------------------------------------
_NP2PREC = {np.dtype(np.complex128): (L.F64, True), np.dtype(np.complex64): (L.F32, True),
            np.dtype(np.float64): (L.F64, False), np.dtype(np.float32): (L.F32, False)}


def _torch():
    import torch
    return torch


def _torch_dtype(dt):
    torch = _torch()
    return {np.dtype(np.complex128): torch.complex128, np.dtype(np.complex64): torch.complex64,
            np.dtype(np.float64): torch.float64, np.dtype(np.float32): torch.float32}[np.dtype(dt)]


class DeviceArray:
    """A column-major (Julia-layout) array in device memory.

    ``shape`` is the Julia size (n1, n2, n3), dim 1 fastest.  The storage is a
    flat torch tensor on the current ROCm device."""

    def __init__(self, shape, dtype, tensor=None, device=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _NP2PREC:
            raise TypeError("unsupported dtype %s" % self.dtype)
        torch = _torch()
        n = int(np.prod(self.shape)) if self.shape else 1
        if tensor is None:
            tensor = torch.empty(n, dtype=_torch_dtype(self.dtype), device=device or "cuda")
        if tensor.numel() != n or not tensor.is_contiguous() or tensor.dtype != _torch_dtype(self.dtype):
            raise ValueError("tensor does not match shape/dtype")
        self.tensor = tensor

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size(self):
        return int(np.prod(self.shape))

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    @property
    def ptr(self):
        return self.tensor.data_ptr()

    @property
    def precision(self):
        return _NP2PREC[self.dtype][0]

    @property
    def is_complex(self):
        return _NP2PREC[self.dtype][1]

    @classmethod
    def from_host(cls, a, device=None):
        a = np.asfortranarray(a)
        torch = _torch()
        flat = np.ascontiguousarray(a.T).reshape(-1)  # column-major order, flattened
        t = torch.from_numpy(flat.copy()).to(device or "cuda")
        return cls(a.shape, a.dtype, t)

    def to_host(self):
        flat = self.tensor.detach().cpu().numpy()
        return flat.reshape(self.shape[::-1]).T

    def similar(self, shape=None, dtype=None):
        return DeviceArray(self.shape if shape is None else shape, self.dtype if dtype is None else dtype,
                           device=self.tensor.device)

    def __repr__(self):
        return "DeviceArray(%s, %s)" % ("x".join(map(str, self.shape)), self.dtype)


def _stream_handle(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class _NoStream:
    def __enter__(self):
        return None

    def __exit__(self, *exc):
        return False


def _on_stream(stream):
    """Context in which torch allocations and conversions run on `stream`
    (torch's current stream when None), so they are ordered with the
    library's launches on that stream."""
    return _torch().cuda.stream(stream) if stream is not None else _NoStream()


def _splitmix_u(seed, i):
    """u(seed, i) of the SplitMix64 counter generator, on the host."""
    M = (1 << 64) - 1
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    z ^= z >> 31
    return (z >> 11) * 2.0 ** -53


def rand(shape, dtype=np.complex128, seed=1, device=None, stream=None):
    """Seeded uniform[0,1) array from the project's SplitMix64 counter generator
    (complex element e = (u(2e), u(2e+1)); real element e = u(e)), filled on device."""
    A = DeviceArray(shape, dtype, device=device)
    check(lib().dfft_fill_uniform(ctypes.c_void_p(A.ptr), A.size, 1 if A.is_complex else 0, A.precision,
                                  ctypes.c_uint64(seed), _stream_handle(stream)), "dfft_fill_uniform")
    return A


# ---------------------------------------------------------------------------
# plans (plan_transform / gpu_plan_transform, src/fft.jl:126-154, src/fftgpu.jl:85-105)
# ---------------------------------------------------------------------------
class Plan:
    """An immutable transform plan: a list of HIP axis passes (AbstractFFTs-style:
    ``p * A`` or ``p.mul(out, A)``)."""

    def __init__(self, shape, kind, sign, precision, decomp=Pencil(), dims=None, r2r_kinds=None, flags=0):
        self.shape = tuple(int(s) for s in shape)
        self.kind = kind
        self.sign = sign
        self.precision = precision
        self.decomp = decomp
        rank = len(self.shape)
        dims = tuple(range(1, rank + 1)) if dims is None else tuple(dims)
        self.dims = dims
        self.flags = flags
        h = ctypes.c_void_p()
        kinds = L.i32(r2r_kinds) if r2r_kinds is not None else None
        st = lib().dfft_plan_create(ctypes.byref(h), rank, L.i64(self.shape), kind, sign, kinds, precision,
                                    decomp.code, L.i32([d - 1 for d in dims]), len(dims), flags)
        check(st, "dfft_plan_create")
        self._h = h
        ws, ib, ob = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        check(lib().dfft_plan_workspace_size(h, ctypes.byref(ws)))
        check(lib().dfft_plan_io_size(h, ctypes.byref(ib), ctypes.byref(ob)))
        self.workspace_bytes, self.in_bytes, self.out_bytes = ws.value, ib.value, ob.value

    @property
    def handle(self):
        return self._h

    def passes(self):
        import json
        n = ctypes.c_int()
        check(lib().dfft_plan_num_passes(self._h, ctypes.byref(n)))
        out = []
        buf = ctypes.create_string_buffer(1024)
        for i in range(n.value):
            check(lib().dfft_plan_describe(self._h, i, buf, 1024))
            out.append(json.loads(buf.value.decode()))
        return out

    def schedule(self):
        """('passes' | 'plane_pipeline', kernel launches per execution)."""
        k, n = ctypes.c_int(), ctypes.c_int()
        check(lib().dfft_plan_schedule(self._h, ctypes.byref(k), ctypes.byref(n)))
        return ("plane_pipeline" if k.value == 1 else "passes"), n.value

    def export_wisdom(self, path):
        """Write this plan's pass schedules as a stamped wisdom file: a process
        that loads it (tune_load) and creates the same plan gets the same
        kernels and launch geometry."""
        check(lib().dfft_plan_export_wisdom(self._h, str(path).encode()), "dfft_plan_export_wisdom")

    def algorithmic_bytes(self):
        b = ctypes.c_double()
        check(lib().dfft_plan_algorithmic_bytes(self._h, ctypes.byref(b)))
        return b.value

    def execute(self, inp, out, work=None, stream=None):
        """Enqueue the plan on `stream` (default: torch's current stream).

        A plan that needs a workspace (C2R without DFFT_FLAG_DESTROY_INPUT)
        gets a fresh one per call from torch's stream-ordered allocator when
        `work` is None, so concurrent executions of one cached plan on
        different streams or threads never share scratch memory."""
        if inp.nbytes != self.in_bytes or out.nbytes != self.out_bytes:
            raise ValueError("array sizes do not match the plan (%d/%d vs %d/%d bytes)"
                             % (inp.nbytes, out.nbytes, self.in_bytes, self.out_bytes))
        if work is None and self.workspace_bytes:
            with _on_stream(stream):
                work = _torch().empty(self.workspace_bytes, dtype=_torch().uint8, device=out.tensor.device)
        wptr = None
        if work is not None:
            wptr = work.data_ptr() if hasattr(work, "data_ptr") else int(work)
        check(lib().dfft_execute(self._h, ctypes.c_void_p(inp.ptr), ctypes.c_void_p(out.ptr),
                                 ctypes.c_void_p(wptr) if wptr else None, _stream_handle(stream)), "dfft_execute")
        return out

    def execute_pass(self, i, inp, out, work=None, stream=None):
        """Enqueue axis pass `i` alone (passes 0..n-1 in order == execute).
        Without `work`, a plan that needs a workspace uses its own one, so the
        passes of one execution see the same intermediate buffer."""
        wptr = None
        if work is not None:
            wptr = work.data_ptr() if hasattr(work, "data_ptr") else int(work)
        check(lib().dfft_execute_pass(self._h, int(i), ctypes.c_void_p(inp.ptr), ctypes.c_void_p(out.ptr),
                                      ctypes.c_void_p(wptr) if wptr else None, _stream_handle(stream)),
              "dfft_execute_pass")
        return out

    def mul(self, out, inp):
        return self.execute(inp, out)

    def destroy(self):
        if getattr(self, "_h", None):
            try:
                lib().dfft_plan_destroy(self._h)
            finally:
                self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def autotune(shape, kind=L.C2C, sign=-1, precision=L.F64, decomp=Pencil(), dims=None, r2r_kinds=None, level=1,
             flags=0, stream=None):
    """Time candidate schedules of every axis pass of this transform on the
    device and record the fastest in the library's schedule cache; plans
    created afterwards use it.  Cached plans of the same shape are dropped."""
    dims = tuple(range(1, len(shape) + 1)) if dims is None else tuple(dims)
    kinds = L.i32(r2r_kinds) if r2r_kinds is not None else None
    _torch().cuda.init()
    check(lib().dfft_autotune(len(shape), L.i64(shape), kind, sign, kinds, precision, decomp.code,
                              L.i32([d - 1 for d in dims]), len(dims), flags, level, _stream_handle(stream)),
          "dfft_autotune")
    clear_plan_cache()


def tune_save(path):
    """Write the schedule cache as a wisdom file stamped with the HIP runtime
    version and GPU architecture (export_wisdom, src/fft.jl:122-124)."""
    check(lib().dfft_tune_save(str(path).encode()), "dfft_tune_save")


def tune_autosave(path):
    """Save the cache to `path` whenever tuning adds to it (None: off), as the
    reference saves its wisdom after every new plan (src/fft.jl:162)."""
    check(lib().dfft_tune_autosave(str(path).encode() if path else None), "dfft_tune_autosave")


def tune_load(path):
    check(lib().dfft_tune_load(str(path).encode()), "dfft_tune_load")
    clear_plan_cache()


def tune_clear():
    check(lib().dfft_tune_clear())
    clear_plan_cache()


def tune_entries():
    n = ctypes.c_int()
    check(lib().dfft_tune_entries(ctypes.byref(n)))
    return n.value


# GPU_PLAN_CACHE (src/fftgpu.jl:83,107-116), guarded by a lock (the reference's is not)
_PLAN_CACHE = {}
_PLAN_LOCK = threading.Lock()


def _cached_plan(shape, kind, sign, precision, decomp, dims, r2r_kinds=None, flags=0):
    key = (tuple(shape), kind, sign, precision, decomp.code, tuple(dims), tuple(r2r_kinds or ()), flags)
    with _PLAN_LOCK:
        p = _PLAN_CACHE.get(key)
        if p is None:
            p = Plan(shape, kind, sign, precision, decomp, dims, r2r_kinds, flags)
            _PLAN_CACHE[key] = p
        return p


def clear_plan_cache():
    with _PLAN_LOCK:
        for p in _PLAN_CACHE.values():
            p.destroy()
        _PLAN_CACHE.clear()


def _dims(A, dims):
    return tuple(range(1, A.ndim + 1)) if dims is None else tuple(int(d) for d in (dims if isinstance(dims, (tuple, list)) else (dims,)))


def _complex_of(dt):
    return np.dtype(np.complex128) if np.dtype(dt) in (np.dtype(np.float64), np.dtype(np.complex128)) else np.dtype(np.complex64)


def _real_of(dt):
    return np.dtype(np.float64) if np.dtype(dt) in (np.dtype(np.float64), np.dtype(np.complex128)) else np.dtype(np.float32)


def _r2r_kinds(transforms, dims, decomp):
    """Per-dim r2r kinds from a transforms tuple.  One R2R covering every dim
    (README.md:59-66) or one entry per stage; Slab applies transforms[1] to its
    2D stage and ignores transforms[2] (src/fft.jl:735,746)."""
    if transforms is None:
        return None
    ts = tuple(transforms) if isinstance(transforms, (tuple, list)) else (transforms,)
    if not any(isinstance(t, R2R) for t in ts):
        return None
    if len(ts) == 1:
        kinds = ts[0].kinds
        if len(kinds) == 1:
            kinds = kinds * len(dims)
        if len(kinds) != len(dims):
            raise ValueError("R2R needs one kind per transformed dim")
        return kinds
    if isinstance(decomp, Slab) and len(ts) == 3:
        k01 = ts[0].kinds if len(ts[0].kinds) == 2 else ts[0].kinds * 2
        return tuple(k01[:2]) + (ts[2].kinds[0],)
    return tuple(t.kinds[0] for t in ts[:len(dims)])


# ---------------------------------------------------------------------------
# README-level API
# ---------------------------------------------------------------------------
def plan_fft(A, dims=None, decomp=Pencil(), flags=0):
    return _cached_plan(A.shape, L.C2C, -1, A.precision, decomp, _dims(A, dims), None, flags)


def plan_ifft(A, dims=None, decomp=Pencil(), flags=0):
    return _cached_plan(A.shape, L.C2C, +1, A.precision, decomp, _dims(A, dims), None, flags)


def plan_rfft(A, dims=None, decomp=Pencil(), flags=0):
    return _cached_plan(A.shape, L.R2C, -1, A.precision, decomp, _dims(A, dims), None, flags)


def plan_irfft(F, d, dims=None, decomp=Pencil(), flags=0):
    dims = _dims(F, dims)
    shape = list(F.shape)
    if shape[dims[0] - 1] != d // 2 + 1:
        raise ValueError("irfft: size(F, %d) must be d÷2+1" % dims[0])
    shape[dims[0] - 1] = d
    return _cached_plan(tuple(shape), L.C2R, +1, F.precision, decomp, dims, None, flags)


def plan_r2r(A, kinds, dims=None, decomp=Pencil(), sign=-1, flags=0):
    return _cached_plan(A.shape, L.R2R, sign, A.precision, decomp, _dims(A, dims), tuple(kinds), flags)


def fft(A, decomp=Pencil(), dims=None, transforms=None, stream=None):
    """F = fft(A; decomp, dims[, transforms=(R2R(kinds),)]) -- unnormalised forward."""
    dims = _dims(A, dims)
    kinds = _r2r_kinds(transforms, dims, decomp)
    with _on_stream(stream):
        if kinds is not None:
            if A.is_complex:
                raise TypeError("R2R transforms need a real array")
            p = plan_r2r(A, kinds, dims, decomp, sign=-1)
            return p.execute(A, A.similar(), stream=stream)
        if not A.is_complex:
            A = _to_complex(A)
        p = plan_fft(A, dims, decomp)
        return p.execute(A, A.similar(), stream=stream)


def ifft(F, decomp=Pencil(), dims=None, transforms=None, stream=None):
    """A = ifft(F; decomp, dims) -- inverse scaled by 1/n per transformed dim.
    With R2R kinds the result is scaled by 1/prod(logical r2r size), so an
    REDFT01 ifft of an REDFT10 fft reconstructs the input."""
    dims = _dims(F, dims)
    kinds = _r2r_kinds(transforms, dims, decomp)
    with _on_stream(stream):
        if kinds is not None:
            p = plan_r2r(F, kinds, dims, decomp, sign=+1, flags=L.FLAG_NORMALIZE_R2R)
            return p.execute(F, F.similar(), stream=stream)
        if not F.is_complex:
            F = _to_complex(F)
        p = plan_ifft(F, dims, decomp)
        return p.execute(F, F.similar(), stream=stream)


def rfft(A, decomp=Pencil(), dims=None, stream=None):
    """F = rfft(A; decomp, dims): size(F, dims[1]) = size(A, dims[1])÷2+1."""
    if A.is_complex:
        raise TypeError("rfft needs a real array")
    dims = _dims(A, dims)
    p = plan_rfft(A, dims, decomp)
    shape = list(A.shape)
    shape[dims[0] - 1] = shape[dims[0] - 1] // 2 + 1
    with _on_stream(stream):
        out = DeviceArray(shape, _complex_of(A.dtype), device=A.tensor.device)
        return p.execute(A, out, stream=stream)


def irfft(F, d, decomp=Pencil(), dims=None, stream=None):
    """A = irfft(F, d; decomp, dims): real output with size(A, dims[1]) = d, scaled 1/N."""
    if not F.is_complex:
        raise TypeError("irfft needs a complex array")
    dims = _dims(F, dims)
    p = plan_irfft(F, d, dims, decomp)
    with _on_stream(stream):
        out = DeviceArray(p.shape, _real_of(F.dtype), device=F.tensor.device)
        return p.execute(F, out, stream=stream)


def _to_complex(A):
    """Complex copy of a real array, made on torch's current stream (callers
    enter the execution stream first)."""
    t = A.tensor.to(_torch_dtype(_complex_of(A.dtype)))
    return DeviceArray(A.shape, _complex_of(A.dtype), t)


# ---------------------------------------------------------------------------
# bang API over single-chunk arrays (src/fft.jl:651-776, src/fftgpu.jl:497-619)
# ---------------------------------------------------------------------------
class Workspace:
    """create_workspace(src, dst): on one device every chunk overlap is local
    and covers the whole block (src/fft.jl:301-305), so the redistribution is
    the identity in global index space and runs inside the next axis pass."""

    def __init__(self, src, dst):
        if tuple(src.shape) != tuple(dst.shape) or src.dtype != dst.dtype:
            raise ValueError("workspace arrays must share shape and element type")
        self.shape = tuple(src.shape)
        self.dtype = src.dtype
        self.local_patterns = [((0, 0, 0), self.shape)]


def create_workspace(src, dst, use_hierarchical=True, ranks_per_node=0):
    if isinstance(src, chunked.DArray) or isinstance(dst, chunked.DArray):
        return chunked.create_workspace(src, dst, use_hierarchical, ranks_per_node)
    return Workspace(src, dst)


create_gpu_workspace = create_workspace


def cleanup_workspace(ws):
    """cleanup_workspace! empties the global plan cache too (src/fft.jl:778-781)."""
    clear_plan_cache()


gpu_cleanup_workspace = cleanup_workspace


def _stage_list(transforms, dims, decomp, inverse):
    """The (transform, dims) of each stage in execution order.

    Pencil fft! runs transforms[i] along dims[i] for i = 1, 2, 3 and ifft!
    runs them in the order 3, 2, 1 (src/fft.jl:651-724).  Slab runs
    transforms[1] over (dims[1], dims[2]) and transforms[3] over dims[3],
    ignoring transforms[2] (src/fft.jl:727-776).  A single transform (not a
    tuple) applies to every stage; a one-element tuple holding an R2R with one
    kind per dim is the README form (README.md:59-66)."""
    ts = tuple(transforms) if isinstance(transforms, (tuple, list)) else None
    dims = tuple(int(d) for d in dims)
    slab = isinstance(decomp, Slab)
    groups = [(dims[0], dims[1]), (dims[2],)] if slab else [(d,) for d in dims]
    if ts is None:
        per = [transforms] * len(groups)
    elif len(ts) == 1 and isinstance(ts[0], R2R) and len(ts[0].kinds) == len(dims) > 1:
        ks, per, i = ts[0].kinds, [], 0
        for g in groups:
            per.append(R2R(ks[i:i + len(g)]))
            i += len(g)
    elif len(ts) == 1:
        per = [ts[0]] * len(groups)
    else:
        if len(ts) < len(dims):
            raise ValueError("transforms needs one entry per stage (%d given, %d dims)" % (len(ts), len(dims)))
        per = [ts[0], ts[2]] if slab else list(ts[:len(dims)])
    stages = list(zip(per, groups))
    return stages[::-1] if inverse else stages


def _stage_op(t, ndims):
    """('c2c', sign, None) or ('r2r', -1, kinds) for one stage (plan_transform,
    src/fft.jl:126-144: FFT/FFT! forward, IFFT/IFFT! inverse scaled by 1/n per
    dim).  RFFT!/IRFFT! change the element type, so an in-place stage cannot
    hold them (the reference's plan_rfft! does not exist in FFTW.jl): they
    raise, as the reference's call would."""
    if isinstance(t, R2R):
        kinds = t.kinds if len(t.kinds) == ndims else t.kinds[:1] * ndims
        return "r2r", -1, tuple(kinds)
    if isinstance(t, (FFT, FFT_)):
        return "c2c", -1, None
    if isinstance(t, (IFFT, IFFT_)):
        return "c2c", +1, None
    if isinstance(t, (RFFT, RFFT_, IRFFT, IRFFT_)):
        raise ValueError("%r changes the element type and cannot run as an in-place stage; "
                         "use rfft/irfft" % (t,))
    raise ValueError("Unknown transform type %r" % (t,))


def _run_stages(src, out, transforms, dims, decomp, inverse, stream):
    """Run the bang API's stages from `src` into `out` on one device block.

    Stages of one kind and direction form one plan (one launch per axis).
    Mixed tuples, e.g. (FFT!(), IFFT!(), FFT!()), run one plan per stage:
    the first stage out of place from `src`, the rest in place on `out`."""
    stages = [(_stage_op(t, len(g)), g) for t, g in _stage_list(transforms, dims, decomp, inverse)]
    kinds = {(op, sign) for (op, sign, _), _ in stages}
    if len(kinds) == 1:
        (op, sign), = kinds
        fwd = stages[::-1] if inverse else stages  # the plan takes dims in forward stage order
        all_dims = tuple(d for _, g in fwd for d in g)
        if op == "c2c":
            if not src.is_complex:
                raise TypeError("FFT!/IFFT! stages need a complex array")
            p = (plan_ifft if (sign > 0) else plan_fft)(src, all_dims, decomp)
        else:
            if src.is_complex:
                raise TypeError("R2R stages need a real array")
            ks = tuple(k for (_, _, kk), _ in fwd for k in kk)
            p = plan_r2r(src, ks, all_dims, decomp, sign=+1 if inverse else -1)
        return p.execute(src, out, stream=stream)
    cur = src
    for (op, sign, kk), g in stages:
        if op == "c2c":
            if not src.is_complex:
                raise TypeError("FFT!/IFFT! stages need a complex array")
            p = _cached_plan(src.shape, L.C2C, sign, src.precision, Pencil(), g)
        else:
            if src.is_complex:
                raise TypeError("R2R stages need a real array")
            p = _cached_plan(src.shape, L.R2R, -1, src.precision, Pencil(), g, kk)
        p.execute(cur, out, stream=stream)
        cur = out
    return out


def fft_(*args, stream=None):
    """fft!(C, A, B, ws_AB, ws_BC, transforms, dims) [Pencil] or
    fft!(B, A, ws, transforms, dims, Slab()).  Stage i applies transforms[i]
    along dims[i] (src/fft.jl:651-688, 727-750).  Returns the output array; the
    stage array A is the input (the reference also leaves stage results in
    A/B, which this single-chunk version does not).  Chunked arrays
    (chunked.DArray) take the reference's per-chunk stages and box-copy
    redistributions (daggerfft.chunked.fft_)."""
    if args and isinstance(args[0], chunked.DArray):
        return chunked.fft_(*args, stream=stream)
    if len(args) == 7:
        C, A, B, ws_ab, ws_bc, transforms, dims = args
        return _run_stages(A, C, transforms, tuple(dims), Pencil(), False, stream)
    if len(args) == 6 and isinstance(args[5], Slab):
        B, A, ws, transforms, dims, decomp = args
        return _run_stages(A, B, transforms, tuple(dims), decomp, False, stream)
    raise TypeError("fft_: expected (C, A, B, ws_AB, ws_BC, transforms, dims) or (B, A, ws, transforms, dims, Slab())")


def ifft_(*args, stream=None):
    """ifft!(C, A, B, ws_AB, ws_BC, transforms, dims) [Pencil, input A, output
    C: transforms[3] along dims[3], then [2], then [1]] or ifft!(A, B, ws,
    transforms, dims, Slab()) [input B, output A] (src/fft.jl:691-724,
    753-776); chunked arrays as in fft_."""
    if args and isinstance(args[0], chunked.DArray):
        return chunked.ifft_(*args, stream=stream)
    if len(args) == 7:
        C, A, B, ws_ab, ws_bc, transforms, dims = args
        return _run_stages(A, C, transforms, tuple(dims), Pencil(), True, stream)
    if len(args) == 6 and isinstance(args[5], Slab):
        A, B, ws, transforms, dims, decomp = args
        return _run_stages(B, A, transforms, tuple(dims), decomp, True, stream)
    raise TypeError("ifft_: expected (C, A, B, ws_AB, ws_BC, transforms, dims) or (A, B, ws, transforms, dims, Slab())")


gpu_fft_ = fft_
gpu_ifft_ = ifft_


# chunked arrays (the reference's Blocks-distributed DArray on one device)
from . import chunked  # noqa: E402
from .chunked import DArray, compute_overlap  # noqa: E402
