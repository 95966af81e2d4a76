"""Test matrices with a prescribed hyperbolic spectrum, in the drop-in's file format.

The role of the reference's jkdatgen (src/var/old/jkdatgen/jkdatgen.F90:1-274):
eigenvalues lambda are drawn as GENLAM draws them (src/var/old/jkdatgen/genlam.F90:1-69:
LAPACK's DLARND with the seeds of SEEDIX, seedix.F90:1-20, values with |lambda| <= EPS
redrawn, then scaled by SCAL) or read from a text file (TXTLAM, txtlam.F90:1-52), and a
factor G of a symmetric matrix with those eigenvalues is generated on the GPU
(jkg_dgen: G = Q diag(sqrt|lambda|) H, J = sign(lambda)), so the hyperbolic singular
values of (G, J) are sqrt|lambda|.  As JPART (jpart.F90:1-79) does, the columns are
ordered with the J = +1 ones first, each sign class in drawing order.

Output: FN.Y (N x N FP64, column-major), FN.J (N int64) -- the input of djk.exe --
and FN.LAM (N FP64: lambda in the column order).  The reference writes HDF5 and
factors in extended precision; neither is reproduced (same lambda, other G bits).

Usage: python -m jkogb_amd.jkdatgen LAMGEN SEEDIX N FN [IDIST EPS SCAL | LAMFILE]
  LAMGEN 1: lambda from the text file LAMFILE (one per line);
  LAMGEN 2: random, IDIST 1 uniform(0,1), 2 uniform(-1,1), 3 normal(0,1).
"""
import math
import sys

import numpy as np

# SEEDIX, src/var/old/jkdatgen/seedix.F90:7-11
ISEEDS = {1: (2741, 1223, 1987, 541), 2: (3281, 123, 456, 333)}


def dlaran(iseed):
    """LAPACK DLARAN: the 48-bit multiplicative congruential generator in four 12-bit
    limbs (multiplier limbs 494, 322, 2508, 2549); updates iseed (a list) in place and
    returns a uniform (0, 1) number."""
    m1, m2, m3, m4, ipw2 = 494, 322, 2508, 2549, 4096
    r = 1.0 / ipw2
    while True:
        it4 = iseed[3] * m4
        it3 = it4 // ipw2
        it4 -= ipw2 * it3
        it3 += iseed[2] * m4 + iseed[3] * m3
        it2 = it3 // ipw2
        it3 -= ipw2 * it2
        it2 += iseed[1] * m4 + iseed[2] * m3 + iseed[3] * m2
        it1 = it2 // ipw2
        it2 -= ipw2 * it1
        it1 += iseed[0] * m4 + iseed[1] * m3 + iseed[2] * m2 + iseed[3] * m1
        it1 %= ipw2
        iseed[:] = [it1, it2, it3, it4]
        out = r * (float(it1) + r * (float(it2) + r * (float(it3) + r * float(it4))))
        if out != 1.0:
            return out


def dlarnd(idist, iseed):
    """LAPACK DLARND: 1 uniform (0,1), 2 uniform (-1,1), 3 normal (0,1) by Box-Muller."""
    t1 = dlaran(iseed)
    if idist == 1:
        return t1
    if idist == 2:
        return 2.0 * t1 - 1.0
    if idist == 3:
        t2 = dlaran(iseed)
        return math.sqrt(-2.0 * math.log(t1)) * math.cos(6.2831853071795864769252867663 * t2)
    return 0.0


def genlam(n, seedix, idist, eps, scal):
    """GENLAM: n draws of DLARND(idist) with |lambda| > eps (smaller ones redrawn), times scal."""
    if seedix not in ISEEDS:
        raise ValueError("SEEDIX must be 1 or 2")
    if idist not in (1, 2, 3) or not eps > 0.0 or scal == 0.0:
        raise ValueError("IDIST in 1..3, EPS > 0, SCAL != 0")
    iseed = list(ISEEDS[seedix])
    lam = []
    while len(lam) < n:
        v = dlarnd(idist, iseed)
        if abs(v) > eps:
            lam.append(v)
    lam = np.array(lam, dtype=np.float64)
    return lam * scal if scal != 1.0 else lam


def partition(lam):
    """Column order of JPART: the positive eigenvalues first, each class in drawing order."""
    lam = np.asarray(lam, dtype=np.float64)
    return np.concatenate([lam[lam > 0], lam[lam < 0]])


def generate(ctx, lam, rounds=0, tmax=0.5, seed=1):
    """(G, J) on the GPU (jkg_dgen) for the eigenvalues lam, in JPART's column order."""
    lam = partition(lam)
    G, J = ctx.dgen(lam, rounds=rounds, tmax=tmax, seed=seed)
    return G, J, lam


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) not in (5, 7):
        print(__doc__.split("Usage: ")[1], file=sys.stderr)
        return 2
    lamgen, seedix, n, fn = int(argv[0]), int(argv[1]), int(argv[2]), argv[3]
    if lamgen == 1:
        lam = np.loadtxt(argv[4], dtype=np.float64).reshape(-1)[:n]
        if lam.size != n or np.any(lam == 0.0):
            print("LAMFILE must hold N nonzero values", file=sys.stderr)
            return 1
    elif lamgen == 2:
        lam = genlam(n, seedix, int(argv[4]), float(argv[5]), float(argv[6]))
    else:
        print("LAMGEN must be 1 or 2", file=sys.stderr)
        return 1
    from jkogb_amd import inputs, jkg
    with jkg.Context(0) as ctx:
        G, J, lam = generate(ctx, lam, seed=seedix)
    inputs.write_yj(fn, G, J)
    lam.tofile(fn + ".LAM")
    print("N=%d: %d positive and %d negative eigenvalues -> %s.Y, %s.J, %s.LAM"
          % (n, int(np.sum(J > 0)), int(np.sum(J < 0)), fn, fn, fn))
    return 0


if __name__ == "__main__":
    sys.exit(main())
