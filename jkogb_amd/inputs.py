"""Seeded synthetic inputs of the benchmark configurations (Y, J) and the
reference's file formats.

File formats (src/shared/binio.F90:87-325, src/D/bio.F90, src/Z/bio.F90):
FN.Y is N x N FP64 (or complex128, real/imag interleaved) column-major,
FN.J is N int64 of +1/-1; FN.YU and FN.ZZ are N x N, FN.SY is N FP64.
All are raw native-endian streams with no header.
"""
import numpy as np


def gaussian(n, seed, cplx=False):
    """Y_ij ~ N(0,1) (complex: real and imaginary parts ~ N(0,1/2))."""
    rng = np.random.default_rng(seed)
    if cplx:
        y = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5)
    else:
        y = rng.standard_normal((n, n))
    return np.asfortranarray(y), rng


def signs_random(n, rng):
    """J_i = +1 or -1, each with probability 1/2."""
    return np.where(rng.random(n) < 0.5, 1, -1).astype(np.int64)


def signs_balanced(n, rng):
    """n//2 entries +1, the rest -1, randomly permuted."""
    j = np.array([1] * (n // 2) + [-1] * (n - n // 2), dtype=np.int64)
    rng.shuffle(j)
    return j


def make(config, n=None, seed=0):
    """Inputs of the five configurations; n overrides the size (small tests).

    c1: N=64 real Gaussian, J random signs.
    c2: N=1024 real, Y = G*D, D = diag(10**U(-4,4)), J balanced.
    c3: N=2048 complex Gaussian, J balanced.
    c4: one instance of the batched config: N given, Gaussian, J random.
    c5: N=4096 real, column scales 10**U(-6,6), J balanced.
    """
    if config == "c1":
        n = n or 64
        y, rng = gaussian(n, seed)
        return y, signs_random(n, rng)
    if config == "c2":
        n = n or 1024
        y, rng = gaussian(n, seed)
        y = np.asfortranarray(y * 10.0 ** rng.uniform(-4, 4, n)[None, :])
        return y, signs_balanced(n, rng)
    if config == "c3":
        n = n or 2048
        y, rng = gaussian(n, seed, cplx=True)
        return y, signs_balanced(n, rng)
    if config == "c4":
        y, rng = gaussian(n, seed)
        return y, signs_random(n, rng)
    if config == "c5":
        n = n or 4096
        y, rng = gaussian(n, seed)
        y = np.asfortranarray(y * 10.0 ** rng.uniform(-6, 6, n)[None, :])
        return y, signs_balanced(n, rng)
    raise ValueError(config)


def write_yj(fn, y, j):
    np.asfortranarray(y).reshape(-1, order="F").tofile(fn + ".Y")
    np.asarray(j, dtype=np.int64).tofile(fn + ".J")


def read_yj(fn, n, cplx=False):
    dt = np.complex128 if cplx else np.float64
    y = np.fromfile(fn + ".Y", dtype=dt, count=n * n)
    j = np.fromfile(fn + ".J", dtype=np.int64, count=n)
    if y.size != n * n or j.size != n:
        raise IOError("short input files for N=%d" % n)
    return y.reshape(n, n, order="F"), j


def read_uzs(fn, n, cplx=False):
    dt = np.complex128 if cplx else np.float64
    yu = np.fromfile(fn + ".YU", dtype=dt).reshape(n, n, order="F")
    zz = np.fromfile(fn + ".ZZ", dtype=dt).reshape(n, n, order="F")
    sy = np.fromfile(fn + ".SY", dtype=np.float64)
    return yu, zz, sy


def write_uzs(fn, yu, zz, sy):
    np.asfortranarray(yu).reshape(-1, order="F").tofile(fn + ".YU")
    np.asfortranarray(zz).reshape(-1, order="F").tofile(fn + ".ZZ")
    np.asarray(sy, dtype=np.float64).tofile(fn + ".SY")
