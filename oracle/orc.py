"""ctypes front-end of the CPU oracle (oracle/libjkoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, never by the product package.  It drives the C
restatement of the reference's DSTEP_LOOP (src/D/dstep.F90:472-604,
src/Z/zstep.F90:482-615) the way src/D/djk.F90:23-121 does: JSTEP_LEN,
TRU, DSTEP_LOOP, then YU = U, ZZ = J Z^T J, SY = diag(A).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libjkoracle.so")


class AW(C.Structure):
    """The AW record of src/shared/atypes.F90:18-24."""
    _fields_ = [("w", C.c_double), ("p", C.c_int32), ("q", C.c_int32),
                ("b", C.c_int32), ("i", C.c_int32)]


AW_DTYPE = np.dtype([("w", "<f8"), ("p", "<i4"), ("q", "<i4"), ("b", "<i4"), ("i", "<i4")])
STEPREC_DTYPE = np.dtype([("step", "<i4"), ("oldlen", "<i4"), ("newlen", "<i4"), ("trig", "<i4"),
                          ("hyp", "<i4"), ("pad", "<i4"), ("tw", "<f8")])

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "lib"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def pfx(fused):
    return "orcf_" if fused else "orc_"


def fn(name, fused=False, restype=None, argtypes=None):
    f = getattr(lib(), (pfx(fused) if not name.startswith("orc_") else "") + name)
    if restype is not None:
        f.restype = restype
    if argtypes is not None:
        f.argtypes = argtypes
    return f


def hypot(x, y):
    f = lib().orc_hypot
    f.restype = C.c_double
    f.argtypes = [C.c_double, C.c_double]
    return f(x, y)


def jstep_len(n, n2):
    return lib().orc_jstep_len(n, n2)


def tru(J, id_tru=1):
    n = len(J)
    nn = n * (n - 1) // 2
    J32 = np.ascontiguousarray(J, dtype=np.int32)
    P = np.zeros(nn, np.int32)
    Q = np.zeros(nn, np.int32)
    PN = np.zeros(2, np.int32)
    tt = lib().orc_tru(id_tru, n, _p(J32), nn, _p(P), _p(Q), _p(PN))
    return tt, P, Q, PN


def dhsvd2(a, j, th1=0.75, fused=False):
    """2x2 HSVD of a (2x2 array); returns (A_out(2x2), U, Z, info) column-major arrays."""
    A = np.asfortranarray(np.array(a, dtype=np.float64)).reshape(4, order="F").copy()
    U = np.zeros(4)
    Z = np.zeros(4)
    Jv = np.array(j, dtype=np.int32)
    info = C.c_int32(0)
    f = getattr(lib(), pfx(fused) + "dhsvd2")
    f.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_void_p]
    f(_p(A), _p(Jv), _p(U), _p(Z), th1, C.byref(info))
    r = lambda v: v.reshape(2, 2, order="F")
    return r(A), r(U), r(Z), info.value


def zhsvd2(a, j, th1=0.75, fused=False):
    A = np.asfortranarray(np.array(a, dtype=np.complex128)).reshape(4, order="F").copy()
    U = np.zeros(4, np.complex128)
    Z = np.zeros(4, np.complex128)
    Jv = np.array(j, dtype=np.int32)
    info = C.c_int32(0)
    f = getattr(lib(), pfx(fused) + "zhsvd2")
    f.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_void_p]
    f(_p(A), _p(Jv), _p(U), _p(Z), th1, C.byref(info))
    r = lambda v: v.reshape(2, 2, order="F")
    return r(A), r(U), r(Z), info.value


def step_build(A, J, n2, P, Q, tt, id_ncp=2, fused=False, s=0):
    """One DSTEP_BUILD/ZSTEP_BUILD: returns (info, sorted AW list[:it], sl, step[:sl])."""
    cplx = np.iscomplexobj(A)
    A = np.asfortranarray(A, dtype=np.complex128 if cplx else np.float64)
    n = A.shape[0]
    nn = len(P)
    dz = np.zeros(2 * nn + 2, AW_DTYPE)
    step = np.zeros(max(n2, 1), np.int32)
    sl = C.c_int32(0)
    J32 = np.ascontiguousarray(J, dtype=np.int32)
    f = getattr(lib(), pfx(fused) + ("zstep_build" if cplx else "dstep_build"))
    f.restype = C.c_int64
    f.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    info = f(s, n, _p(A), n, _p(J32), nn, _p(P), _p(Q), id_ncp, len(dz), _p(dz), tt, n2,
             C.byref(sl), _p(step))
    it = int(np.sum(dz["w"][:nn] == dz["w"][:nn]))
    return info, dz, sl.value, step[:sl.value].copy(), it


def solve(Y, J, n2=0, id_ncp=0, id_tru=1, fused=False, max_steps=-1, max_trace=None):
    """The djk/zjk pipeline on the CPU.  Returns dict(YU, ZZ, SY, info, trace, tt, pn)."""
    cplx = np.iscomplexobj(Y)
    dt = np.complex128 if cplx else np.float64
    A = np.array(Y, dtype=dt, order="F", copy=True)
    n = A.shape[0]
    n2 = jstep_len(n, n2)
    nn = n * (n - 1) // 2
    if id_ncp == 0:
        id_ncp = 2
    tt, P, Q, PN = tru(J, id_tru)
    if tt < 0:
        raise ValueError("R%TRU failed: INFO=%d" % tt)
    U = np.zeros((n, n), dt, order="F")
    Z = np.zeros((n, n), dt, order="F")
    S = np.zeros(max(n, 3))
    dz = np.zeros(2 * nn + 2, AW_DTYPE)
    step = np.zeros(max(n2, 1), np.int32)
    W = np.zeros(6 * max(n2, 1), dt)
    if max_trace is None:
        max_trace = 64 * n + 64
    trace = np.zeros(max_trace, STEPREC_DTYPE)
    J32 = np.ascontiguousarray(J, dtype=np.int32)
    f = getattr(lib(), pfx(fused) + ("zstep_loop" if cplx else "dstep_loop"))
    f.restype = C.c_int32
    f.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                  C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    info = f(n, _p(U), n, _p(A), n, _p(Z), n, _p(J32), _p(S), nn, _p(P), _p(Q), id_ncp, len(dz),
             _p(dz), n2, _p(step), tt, _p(W), _p(trace), max_trace, max_steps)
    nsteps = info + 1 if info >= 0 else None
    return dict(YU=U, ZZ=Z, SY=S[:n].copy(), info=info,
                trace=trace[:min(nsteps if nsteps else max_trace, max_trace)], tt=tt, pn=PN, A=A)


def solve_cyclic(Y, J, fused=False, max_steps=-1, max_trace=None, order=1):
    """The mixed solver's cyclic sweep baselines on the CPU (dcyc_loop, no convergence
    tolerance): order 1 round robin, 2 row-cyclic and 3 column-cyclic pointwise (the
    reference's sweeps, src/old/jk1.f90:101-146).  Returns dict(YU, ZZ, SY, info, trace, A)."""
    A = np.array(Y, dtype=np.float64, order="F", copy=True)
    n = A.shape[0]
    U = np.zeros((n, n), order="F")
    Z = np.zeros((n, n), order="F")
    S = np.zeros(max(n, 3))
    W = np.zeros(6 * max(n // 2, 1))
    if max_trace is None:
        max_trace = 64 * n + 64
    trace = np.zeros(max_trace, STEPREC_DTYPE)
    J32 = np.ascontiguousarray(J, dtype=np.int32)
    f = getattr(lib(), pfx(fused) + "dcyc_loop")
    f.restype = C.c_int32
    f.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
    info = f(n, _p(U), n, _p(A), n, _p(Z), n, _p(J32), _p(S), _p(W), _p(trace), max_trace, max_steps, order)
    nsteps = info + 1 if info >= 0 else max_trace
    return dict(YU=U, ZZ=Z, SY=S[:n].copy(), info=info, trace=trace[:min(nsteps, max_trace)], A=A)


def pair_weights(A, J, fused=False):
    """Weight of every pair (DMAGF2T for J_p == J_q, DMAGF2H otherwise, with
    the hyperbolic -inf dropped to NaN as DSTEP_BUILD does), indexed by the
    tie-break order of AW_CMP: k = m(m-1)/2 + (m - p), m = N - (q - p)."""
    cplx = np.iscomplexobj(A)
    A = np.asfortranarray(A, dtype=np.complex128 if cplx else np.float64)
    n = A.shape[0]
    J32 = np.ascontiguousarray(J, dtype=np.int32)
    ft = getattr(lib(), pfx(fused) + ("zmagf2t" if cplx else "dmagf2t"))
    fh = getattr(lib(), pfx(fused) + ("zmagf2h" if cplx else "dmagf2h"))
    for f in (ft, fh):
        f.restype = C.c_double
        f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    w = np.zeros(n * (n - 1) // 2)
    pa, ja = _p(A), _p(J32)
    for m in range(1, n):
        for p in range(m, 0, -1):
            q = p + n - m
            k = m * (m - 1) // 2 + (m - p)
            if J32[p - 1] == J32[q - 1]:
                w[k] = ft(n, p, q, pa, n, ja)
            else:
                v = fh(n, p, q, pa, n, ja)
                w[k] = np.nan if v < -np.finfo(np.float64).max else v
    return w
