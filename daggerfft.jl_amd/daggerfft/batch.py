"""Batched launch: many independent 3D transforms per call (dfft_batch_* of
include/daggerfft_hip.h) -- grouped launches, one per axis-pass level and
precision, the same tiles in one cooperative launch, or each transform's own
plan on four forked streams (the library picks unless a schedule flag
says).  No reference counterpart; the
reference runs one cuFFT plan per chunk (src/fftgpu.jl:118-124)."""
from __future__ import annotations

import ctypes

from . import _lib as L
from ._lib import check, lib


class Batch:
    """descs: list of (shape, kind, sign, precision) with kind L.C2C or L.R2C."""

    def __init__(self, descs, flags=0):
        arr = (L.dfft_batch_desc * len(descs))()
        for i, (shape, kind, sign, prec) in enumerate(descs):
            arr[i].n[0], arr[i].n[1], arr[i].n[2] = [int(s) for s in shape]
            arr[i].kind, arr[i].sign, arr[i].precision, arr[i].flags = kind, sign, prec, 0
        self.descs = list(descs)
        h = ctypes.c_void_p()
        check(lib().dfft_batch_create(ctypes.byref(h), len(descs), arr, flags), "dfft_batch_create")
        self._h = h

    @property
    def tickets(self):
        n = ctypes.c_int64()
        check(lib().dfft_batch_num_items(self._h, ctypes.byref(n)))
        return n.value

    def algorithmic_bytes(self):
        b = ctypes.c_double()
        check(lib().dfft_batch_algorithmic_bytes(self._h, ctypes.byref(b)))
        return b.value

    def execute(self, ins, outs, stream=None):
        from . import _stream_handle
        n = len(self.descs)
        pi = (ctypes.c_void_p * n)(*[a.ptr for a in ins])
        po = (ctypes.c_void_p * n)(*[a.ptr for a in outs])
        check(lib().dfft_batch_execute(self._h, pi, po, _stream_handle(stream)), "dfft_batch_execute")
        return outs

    @property
    def launches(self):
        n = ctypes.c_int()
        check(lib().dfft_batch_num_launches(self._h, ctypes.byref(n)))
        return n.value

    def destroy(self):
        if getattr(self, "_h", None):
            lib().dfft_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def batch_config_sizes(count=64, seed=7):
    """Axis lengths of the mixed batch: floor(8 u(seed, 3t+a)) indexes
    {32,48,64,96,128,160,192,256}; even t ComplexF32, odd t ComplexF64."""
    from . import _splitmix_u
    choices = (32, 48, 64, 96, 128, 160, 192, 256)
    out = []
    for t in range(count):
        shape = tuple(choices[int(8 * _splitmix_u(seed, 3 * t + a))] for a in range(3))
        out.append((shape, L.F32 if t % 2 == 0 else L.F64))
    return out
