"""Writes tests/golden/golden.npz: tiny known-answer vectors for the oracle.

Independent of the oracle's code: inputs come from a pure-Python SplitMix64
(Steele, Lea, Flood 2014; counter form u(seed, i) = mix64(seed + (i+1)*gamma)
>> 11 * 2^-53), outputs from brute-force O(n^2) sums in mpmath at 40 digits,
one axis at a time (FFTW's multi-dimensional transforms are separable), using
FFTW's documented definitions.  Run: python tests/golden/make_golden.py
"""
import hashlib
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
M = (1 << 64) - 1
HERE = os.path.dirname(os.path.abspath(__file__))


def u(seed, i):
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    z ^= z >> 31
    return (z >> 11) * 2.0 ** -53


def fill(seed, shape, cplx):
    n = int(np.prod(shape))
    if cplx:
        flat = np.array([complex(u(seed, 2 * e), u(seed, 2 * e + 1)) for e in range(n)])
    else:
        flat = np.array([u(seed, e) for e in range(n)])
    return flat.reshape(shape, order="F")


def along(a, axis, f):
    """apply the 1D map f (list of mpc -> list) along `axis` of an object array"""
    a = np.moveaxis(a, axis, 0)
    out = None
    for idx in np.ndindex(a.shape[1:]):
        line = f([a[(j,) + idx] for j in range(a.shape[0])])
        if out is None:
            out = np.empty((len(line),) + a.shape[1:], dtype=object)
        for j, v in enumerate(line):
            out[(j,) + idx] = v
    return np.moveaxis(out, 0, axis)


def dft(x, sign):
    n = len(x)
    return [mp.fsum(x[j] * mp.expjpi(sign * mp.mpf(2 * j * k) / n) for j in range(n)) for k in range(n)]


def cosm(num, den):
    return mp.cospi(mp.mpf(num) / den)


def sinm(num, den):
    return mp.sinpi(mp.mpf(num) / den)


def r2r(x, kind):
    n = len(x)
    y = []
    for k in range(n):
        if kind == "REDFT10":
            y.append(2 * mp.fsum(x[j] * cosm((2 * j + 1) * k, 2 * n) for j in range(n)))
        elif kind == "REDFT01":
            y.append(x[0] + 2 * mp.fsum(x[j] * cosm(j * (2 * k + 1), 2 * n) for j in range(1, n)))
        elif kind == "REDFT00":
            y.append(x[0] + (-1) ** k * x[n - 1] + 2 * mp.fsum(x[j] * cosm(j * k, n - 1) for j in range(1, n - 1)))
        elif kind == "REDFT11":
            y.append(2 * mp.fsum(x[j] * cosm((2 * j + 1) * (2 * k + 1), 4 * n) for j in range(n)))
        elif kind == "RODFT00":
            y.append(2 * mp.fsum(x[j] * sinm((j + 1) * (k + 1), n + 1) for j in range(n)))
        elif kind == "RODFT10":
            y.append(2 * mp.fsum(x[j] * sinm((2 * j + 1) * (k + 1), 2 * n) for j in range(n)))
        elif kind == "RODFT01":
            y.append((-1) ** k * x[n - 1] + 2 * mp.fsum(x[j] * sinm((j + 1) * (2 * k + 1), 2 * n) for j in range(n - 1)))
        elif kind == "RODFT11":
            y.append(2 * mp.fsum(x[j] * sinm((2 * j + 1) * (2 * k + 1), 4 * n) for j in range(n)))
        elif kind == "DHT":
            y.append(mp.fsum(x[j] * (cosm(2 * j * k, n) + sinm(2 * j * k, n)) for j in range(n)))
    return y


def to_obj(a):
    o = np.empty(a.shape, dtype=object)
    for idx in np.ndindex(a.shape):
        o[idx] = mp.mpc(complex(a[idx])) if np.iscomplexobj(a) else mp.mpf(float(a[idx]))
    return o


def to_np(o, cplx):
    a = np.empty(o.shape, dtype=np.complex128 if cplx else np.float64)
    for idx in np.ndindex(o.shape):
        a[idx] = complex(o[idx]) if cplx else float(mp.re(o[idx]))
    return np.asfortranarray(a)


def main():
    out = {}
    # 3D C2C forward / inverse (normalised 1/n per dim)
    for name, shape, seed in [("c2c_4x3x2", (4, 3, 2), 101), ("c2c_8x6x5", (8, 6, 5), 102), ("c2c_7x5x3", (7, 5, 3), 103)]:
        a = fill(seed, shape, True)
        o = to_obj(a)
        for ax in range(3):
            o = along(o, ax, lambda x: dft(x, -1))
        out[name + "_in"] = a
        out[name + "_fwd"] = to_np(o, True)
        for ax in (2, 1, 0):
            o = along(o, ax, lambda x, n=shape[ax]: [v / n for v in dft(x, +1)])
        out[name + "_rt"] = to_np(o, True)
    # rfft: r2c along dim 1 (first n/2+1 bins of the full DFT), then C2C along dims 2, 3
    for name, shape, seed in [("r2c_6x4x3", (6, 4, 3), 201), ("r2c_5x4x2", (5, 4, 2), 202)]:
        a = fill(seed, shape, False)
        o = to_obj(a)
        o = along(o, 0, lambda x: dft(x, -1)[: len(x) // 2 + 1])
        for ax in (1, 2):
            o = along(o, ax, lambda x: dft(x, -1))
        out[name + "_in"] = a
        out[name + "_fwd"] = to_np(o, True)
    # r2r kinds along both dims of a 2D array (and a 1D line of 7)
    for kind in ("REDFT10", "REDFT01", "REDFT00", "REDFT11", "RODFT00", "RODFT10", "RODFT01", "RODFT11", "DHT"):
        a = fill(300, (6, 5), False)
        o = to_obj(a)
        for ax in range(2):
            o = along(o, ax, lambda x: r2r(x, kind))
        out["r2r_%s_6x5_in" % kind] = a
        out["r2r_%s_6x5" % kind] = to_np(o, False)
    out["splitmix_seed0_first"] = np.array([u(0, i) for i in range(4)])
    path = os.path.join(HERE, "golden.npz")
    np.savez(path, **out)
    h = hashlib.sha256(open(path, "rb").read()).hexdigest()
    with open(os.path.join(HERE, "MANIFEST.sha256"), "w") as f:
        f.write("%s  golden.npz\n" % h)
    print("wrote", path, len(out), "arrays", h)


if __name__ == "__main__":
    main()
