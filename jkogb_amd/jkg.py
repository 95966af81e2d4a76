"""ctypes front-end of libjkg.so (the C ABI of include/jkg.h).

The reference's interfaces on the hot path, in Python:
  Context.dsolve      -- djk.exe FN N N_2 [ID_NCP [ID_TRU]] between reading
                         FN.Y/FN.J and writing FN.YU/FN.ZZ/FN.SY
                         (src/D/djk.F90:27-121)
  Context.srt / .ncp  -- the DPROC strategy procedures ASRT / ANCP
                         (src/D/dtypes.F90:5-30)
  tru / jstep_len     -- ATRU (src/shared/jstep.F90:9-120), JSTEP_LEN (:124-133)
  Context.dhsvd2, .dweights, .dbuild, .dstep
                      -- DHSVD2, DMAGF2T/DMAGF2H, DSTEP_BUILD, DSTEP_EXEC
There is no CPU fallback: a missing libjkg.so raises at import.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjkg.so")

if not os.path.exists(LIB_PATH):
    raise ImportError("libjkg.so is not built (run __graft_entry__.build() or make -C jkogb_amd/csrc): %s"
                      % LIB_PATH)
_lib = C.CDLL(LIB_PATH)

JKG_OK = 0
JKG_EHIP = -1000
JKG_ENOMEM = -1001
JKG_ENOTSUP = -1002
JKG_EJ = -1003

AW_DTYPE = np.dtype([("w", "<f8"), ("p", "<i4"), ("q", "<i4"), ("b", "<i4"), ("i", "<i4")])
STEPREC_DTYPE = np.dtype([("step", "<i4"), ("oldlen", "<i4"), ("newlen", "<i4"), ("trig", "<i4"),
                          ("hyp", "<i4"), ("it", "<i4"), ("tw", "<f8"), ("build_s", "<f4"), ("transf_s", "<f4")])


PROGRESS = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int32)


class Opts(C.Structure):
    _fields_ = [("id_ncp", C.c_int32), ("id_tru", C.c_int32), ("fused", C.c_int32), ("max_steps", C.c_int32),
                ("steps_per_sync", C.c_int32), ("trace_cap", C.c_int32), ("trace", C.c_void_p),
                ("progress", PROGRESS), ("user", C.c_void_p)]


def _hook(o, progress):
    """Installs progress(step, done) as the solve's progress hook; returns the ctypes
    callback, which the caller keeps alive for the call."""
    if progress is None:
        return None
    cb = PROGRESS(lambda user, step, done: progress(int(step), bool(done)))
    o.progress = cb
    return cb


class Result(C.Structure):
    _fields_ = [("info", C.c_int32), ("tt", C.c_int32), ("npos", C.c_int32), ("nneg", C.c_int32),
                ("n2", C.c_int32), ("steps", C.c_int32), ("seconds", C.c_double), ("sum_sl", C.c_int64),
                ("sum_it", C.c_int64), ("sum_b2", C.c_int64), ("sum_b4", C.c_int64), ("sum_abod", C.c_int64)]


class MixedOpts(C.Structure):
    """jkg_mixed_opts: options of the mixed solver (no reference counterpart)."""
    _fields_ = [("weights", C.c_int32), ("fp32_warm", C.c_int32), ("switch_tol", C.c_double),
                ("key_bits", C.c_int32), ("controller", C.c_int32), ("ctl", C.c_void_p),
                ("threshold", C.c_double), ("tol", C.c_double), ("ctl_trace", C.c_void_p),
                ("ctl_trace_cap", C.c_int32), ("ordering", C.c_int32), ("warm", C.c_void_p),
                ("ctl_tau_max", C.c_double), ("ctl_handoff", C.c_float), ("reserved", C.c_int32)]


class WarmInfo(C.Structure):
    """jkg_warm_info: what the FP32 warm start of the complex mixed solver did."""
    _fields_ = [("steps32", C.c_int32), ("warm_end", C.c_int32), ("ns_u", C.c_int32), ("ns_z", C.c_int32),
                ("fallback", C.c_int32), ("reserved", C.c_int32), ("off_rel", C.c_double), ("res_u", C.c_double),
                ("res_z", C.c_double), ("seconds32", C.c_double), ("seconds_switch", C.c_double)]


class BatchOpts(C.Structure):
    """jkg_batch_opts."""
    _fields_ = [("fused", C.c_int32), ("max_steps", C.c_int32), ("grid", C.c_int32), ("variant", C.c_int32),
                ("tol", C.c_double), ("steps", C.c_int32), ("tuned_grid", C.c_int32), ("tuned_variant", C.c_int32),
                ("reserved", C.c_int32), ("seconds", C.c_double), ("tune_seconds", C.c_double)]


CTL_NPARAM = 5378
CTL_REC = 20
CTL_FEATURES = ["step/16N", "log10(TW/TW0)/20", "log10(TW/TWprev)/2", "SL/N2", "NL/SL", "IT/NN", "walk depth",
                "trig share", "log10(wmax/|Y|^2)/32", "log2(N)/16", "N2 fraction", "log10 tau (norm)",
                "log10(TW/|Y|^2)/32", "smoothed rate", "SL reached N2", "1"]


W_PAIR, W_GRAM = 0, 1

STAGES = ["begin", "weights", "sort", "select", "pivots", "rotate", "end"]


def _counts(r):
    return dict(sum_sl=r.sum_sl, sum_it=r.sum_it, sum_b2=r.sum_b2, sum_b4=r.sum_b4, sum_abod=r.sum_abod)


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_vp, _i, _d = C.c_void_p, C.c_int, C.c_double
_create = _sig("jkg_create", _i, [C.POINTER(C.c_void_p), _i])
_destroy = _sig("jkg_destroy", None, [_vp])
_strerror = _sig("jkg_strerror", C.c_char_p, [_i])
_last_hip = _sig("jkg_last_hip_error", _i, [])
_dsolve = _sig("jkg_dsolve", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(Opts), C.POINTER(Result)])
_zsolve = _sig("jkg_zsolve", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(Opts), C.POINTER(Result)])
_dsolve_from = _sig("jkg_dsolve_from", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, C.POINTER(Opts),
                                            C.POINTER(Result)])
_zsolve_from = _sig("jkg_zsolve_from", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, C.POINTER(Opts),
                                            C.POINTER(Result)])
_dsolve_noalloc = _sig("jkg_dsolve_noalloc", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(Opts),
                                                  C.POINTER(Result)])
_srt = _sig("jkg_srt", _i, [_vp, _i, _i, _vp, C.POINTER(C.c_int64)])
_ncp = _sig("jkg_ncp", _i, [_vp, _i, _i, _i, _vp, _i, C.POINTER(C.c_int32), _vp, C.POINTER(C.c_int64)])
_ncp3 = _sig("jkg_ncp3", _i, [_vp, _i, _i, _i, _vp, _i, C.POINTER(C.c_int32), _vp, C.POINTER(C.c_int64)])
_dgen = _sig("jkg_dgen", _i, [_vp, _i, _vp, _i, _d, C.c_uint64, _vp, _vp])
_set_select = _sig("jkg_set_select", _i, [_vp, _i, _i])
_set_frames = _sig("jkg_set_frames", _i, [_vp, _i, _i, _i, _vp, _vp])
_tru = _sig("jkg_tru", _i, [_i, _i, _vp, _i, _vp, _vp, _vp])
_jstep_len = _sig("jkg_jstep_len", _i, [_i, _i])
_dhsvd2 = _sig("jkg_dhsvd2", _i, [_vp, _i, _vp, _vp, _d, _i, _vp, _vp, _vp, _vp])
_zhsvd2 = _sig("jkg_zhsvd2", _i, [_vp, _i, _vp, _vp, _d, _i, _vp, _vp, _vp, _vp])
_dweights = _sig("jkg_dweights", _i, [_vp, _i, _vp, _vp, _i, _vp])
_zweights = _sig("jkg_zweights", _i, [_vp, _i, _vp, _vp, _i, _vp])
_dbuild = _sig("jkg_dbuild", _i, [_vp, _i, _i, _vp, _vp, _i, C.POINTER(C.c_int32), _vp, C.POINTER(C.c_int32), _vp])
_dstep = _sig("jkg_dstep", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp])
_dprofile = _sig("jkg_dprofile", _i, [_vp, _i, _i, _vp, _vp, _i, _i, _vp, C.POINTER(Result)])
_mixed_default = _sig("jkg_mixed_opts_default", None, [C.POINTER(MixedOpts)])
_dsolve_mixed = _sig("jkg_dsolve_mixed", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(Opts),
                                              C.POINTER(MixedOpts), C.POINTER(Result)])
_dweights_mixed = _sig("jkg_dweights_mixed", _i, [_vp, _i, _vp, _vp, _i, _vp])
_dprofile_mixed = _sig("jkg_dprofile_mixed", _i, [_vp, _i, _i, _vp, _vp, _i, _i, C.POINTER(MixedOpts), _vp,
                                                  C.POINTER(Result)])
_zsolve_mixed = _sig("jkg_zsolve_mixed", _i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, C.POINTER(Opts),
                                              C.POINTER(MixedOpts), C.POINTER(Result)])
_zweights_mixed = _sig("jkg_zweights_mixed", _i, [_vp, _i, _vp, _vp, _i, _vp])
_zprofile = _sig("jkg_zprofile", _i, [_vp, _i, _i, _vp, _vp, _i, _i, _vp, C.POINTER(Result)])
_zprofile_mixed = _sig("jkg_zprofile_mixed", _i, [_vp, _i, _i, _vp, _vp, _i, _i, C.POINTER(MixedOpts), _vp,
                                                  C.POINTER(Result)])
_batch_default = _sig("jkg_batch_opts_default", None, [C.POINTER(BatchOpts)])
_dsolve_batched = _sig("jkg_dsolve_batched", _i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                  C.POINTER(BatchOpts)])
_ctl_grad = _sig("jkg_ctl_grad", _i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp])
_ctl_fit = _sig("jkg_ctl_fit", _i, [_vp, _i, _vp, _vp, _vp, _i, _i, C.c_float, C.c_float, C.c_uint64, _vp, _vp])
_pair_index = _sig("jkg_pair_index", C.c_uint32, [_i, _i, _i])
_key_weight = _sig("jkg_key_weight", _d, [C.c_uint64])

EXPORTS = ["jkg_create", "jkg_destroy", "jkg_strerror", "jkg_last_hip_error", "jkg_opts_default", "jkg_dsolve",
           "jkg_zsolve", "jkg_srt", "jkg_ncp", "jkg_tru", "jkg_jstep_len", "jkg_dhsvd2", "jkg_zhsvd2",
           "jkg_dweights", "jkg_zweights", "jkg_dbuild", "jkg_dstep", "jkg_dprofile", "jkg_pair_index",
           "jkg_key_weight", "jkg_mixed_opts_default", "jkg_dsolve_mixed", "jkg_dweights_mixed",
           "jkg_dprofile_mixed", "jkg_ctl_grad", "jkg_ctl_fit", "jkg_batch_opts_default", "jkg_dsolve_batched",
           "jkg_zsolve_mixed", "jkg_zweights_mixed", "jkg_zprofile", "jkg_zprofile_mixed", "jkg_ncp3", "jkg_dgen",
           "jkg_set_select", "jkg_set_frames", "jkg_dsolve_noalloc", "jkg_dsolve_from", "jkg_zsolve_from"]


def mixed_opts(**kw):
    """jkg_mixed_opts with the library's defaults, then the given fields."""
    m = MixedOpts()
    _mixed_default(C.byref(m))
    for k, v in kw.items():
        if not hasattr(m, k):
            raise KeyError(k)
        setattr(m, k, v)
    return m


class JKGError(RuntimeError):
    def __init__(self, code, where):
        msg = "%s: %s (%d)" % (where, _strerror(code).decode(), code)
        if code == JKG_EHIP:
            msg += " [hipError %d]" % _last_hip()
        super().__init__(msg)
        self.code = code


def _chk(rc, where):
    if rc != JKG_OK:
        raise JKGError(rc, where)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def jstep_len(n, n2):
    return _jstep_len(n, n2)


def tru(J, id_tru=1):
    n = len(J)
    nn = n * (n - 1) // 2
    J32 = np.ascontiguousarray(J, dtype=np.int32)
    P = np.zeros(max(nn, 1), np.int32)
    Q = np.zeros(max(nn, 1), np.int32)
    PN = np.zeros(2, np.int32)
    tt = _tru(id_tru, n, _p(J32), nn, _p(P), _p(Q), _p(PN))
    return tt, P[:nn], Q[:nn], PN


def pair_index(n, p, q):
    return _pair_index(n, p, q)


def key_weight(key):
    return _key_weight(int(key))


def keys_to_weights(keys):
    """Vectorised jkg_key_weight: the weights the sort keys stand for."""
    keys = np.asarray(keys, dtype=np.uint64)
    ordv = ~keys
    neg = (ordv & np.uint64(1 << 63)) == 0
    u = np.where(neg, ~ordv, ordv & np.uint64((1 << 63) - 1))
    w = u.view(np.float64).copy()
    w[keys == np.uint64(0xFFFFFFFFFFFFFFFF)] = np.nan
    return w


class Context:
    """One HIP stream and the device buffers of the last problem size."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _chk(_create(C.byref(h), device), "jkg_create")
        self._h = h

    def close(self):
        if self._h:
            _destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- solve
    def dsolve(self, Y, J, n2=0, id_ncp=0, id_tru=0, fused=False, max_steps=-1, trace_cap=None,
               steps_per_sync=0, progress=None):
        """HSVD of a real or complex Y (N x N): dict(YU, ZZ, SY, info, trace, tt, pn, n2, seconds).
        progress(step, done), if given, is called at every host poll of the device."""
        cplx = np.iscomplexobj(Y)
        dt = np.complex128 if cplx else np.float64
        Y = np.asfortranarray(Y, dtype=dt)
        n = Y.shape[0]
        if Y.shape != (n, n):
            raise ValueError("Y must be square")
        J64 = np.ascontiguousarray(J, dtype=np.int64)
        YU = np.zeros((n, n), dt, order="F")
        ZZ = np.zeros((n, n), dt, order="F")
        SY = np.zeros(n)
        if trace_cap is None:
            trace_cap = 64 * n + 64 if max_steps < 0 else max_steps + 1
        trace = np.zeros(max(trace_cap, 1), STEPREC_DTYPE)
        o = Opts(id_ncp, id_tru, 1 if fused else 0, max_steps, steps_per_sync, trace_cap, _p(trace))
        keep = _hook(o, progress)
        r = Result()
        f = _zsolve if cplx else _dsolve
        _chk(f(self._h, n, n2, _p(Y), _p(J64), _p(YU), _p(ZZ), _p(SY), C.byref(o), C.byref(r)),
             "jkg_zsolve" if cplx else "jkg_dsolve")
        return dict(YU=YU, ZZ=ZZ, SY=SY, info=r.info, trace=trace[:min(r.steps, trace_cap)], tt=r.tt,
                    pn=(r.npos, r.nneg), n2=r.n2, seconds=r.seconds, steps=r.steps, **_counts(r))

    def dsolve_from(self, A, U, Z, J, step0=0, n2=0, id_ncp=0, id_tru=0, fused=False, max_steps=-1,
                    trace_cap=None, steps_per_sync=0, progress=None):
        """The loop resumed from its state (jkg_dsolve_from / jkg_zsolve_from): A, U, Z (N x N, the
        loop's own, as a previous call returned them; Y, I, I and step0 = 0 to start) and the next
        step's index.  Returns dsolve's dict plus the state after the last step (A, U, Z) and
        ended (False: stopped by max_steps; resume with step0 = info)."""
        cplx = np.iscomplexobj(A)
        dt = np.complex128 if cplx else np.float64
        A, U, Z = (np.array(x, dtype=dt, order="F", copy=True) for x in (A, U, Z))
        n = A.shape[0]
        if A.shape != (n, n) or U.shape != (n, n) or Z.shape != (n, n):
            raise ValueError("A, U and Z must be N x N")
        J64 = np.ascontiguousarray(J, dtype=np.int64)
        YU = np.zeros((n, n), dt, order="F")
        ZZ = np.zeros((n, n), dt, order="F")
        SY = np.zeros(n)
        if trace_cap is None:
            trace_cap = 64 * n + 64 if max_steps < 0 else max_steps + 1
        trace = np.zeros(max(trace_cap, 1), STEPREC_DTYPE)
        o = Opts(id_ncp, id_tru, 1 if fused else 0, max_steps, steps_per_sync, trace_cap, _p(trace))
        keep = _hook(o, progress)
        r = Result()
        f = _zsolve_from if cplx else _dsolve_from
        _chk(f(self._h, n, n2, _p(A), _p(U), _p(Z), _p(J64), step0, _p(YU), _p(ZZ), _p(SY), C.byref(o),
               C.byref(r)), "jkg_zsolve_from" if cplx else "jkg_dsolve_from")
        ended = max_steps < 0 or r.info < step0 + max_steps
        return dict(A=A, U=U, Z=Z, YU=YU, ZZ=ZZ, SY=SY, info=r.info, ended=ended,
                    trace=trace[:min(r.steps, trace_cap)], tt=r.tt, pn=(r.npos, r.nneg), n2=r.n2,
                    seconds=r.seconds, steps=r.steps, **_counts(r))

    def dsolve_mixed(self, Y, J, n2=0, fused=False, max_steps=-1, trace_cap=None, steps_per_sync=0,
                     ctl_params=None, ctl_table=None, ctl_trace=False, progress=None, **mixed):
        """The mixed solver (jkg_dsolve_mixed, or jkg_zsolve_mixed for a complex Y): same outputs
        as dsolve, agreeing to rounding.  ctl_params (JKG_CTL_NPARAM floats): the MLP controller;
        ctl_table (8 floats): a phase table; ctl_trace: also return the per-step controller
        records as out["ctl"]."""
        cplx = np.iscomplexobj(Y)
        dt = np.complex128 if cplx else np.float64
        Y = np.asfortranarray(Y, dtype=dt)
        n = Y.shape[0]
        if Y.shape != (n, n):
            raise ValueError("Y must be square")
        J64 = np.ascontiguousarray(J, dtype=np.int64)
        YU = np.zeros((n, n), dt, order="F")
        ZZ = np.zeros((n, n), dt, order="F")
        SY = np.zeros(n)
        if trace_cap is None:
            trace_cap = 64 * n + 64 if max_steps < 0 else max_steps + 1
        trace = np.zeros(max(trace_cap, 1), STEPREC_DTYPE)
        o = Opts(0, 0, 1 if fused else 0, max_steps, steps_per_sync, trace_cap, _p(trace))
        m = mixed_opts(**mixed)
        keep = [_hook(o, progress)]
        if ctl_params is not None:
            cp = np.ascontiguousarray(ctl_params, dtype=np.float32)
            if cp.size != CTL_NPARAM:
                raise ValueError("ctl_params needs %d floats" % CTL_NPARAM)
            keep.append(cp)
            m.controller, m.ctl = 1, cp.ctypes.data
        elif ctl_table is not None:
            ct = np.ascontiguousarray(ctl_table, dtype=np.float32)
            if ct.size != 8:
                raise ValueError("ctl_table needs 8 floats")
            keep.append(ct)
            m.controller, m.ctl = 2, ct.ctypes.data
        ctr = None
        if ctl_trace:
            ctr = np.zeros((max(trace_cap, 1), CTL_REC), np.float32)
            m.ctl_trace, m.ctl_trace_cap = ctr.ctypes.data, ctr.shape[0]
        wi = None
        if m.fp32_warm:
            wi = WarmInfo()
            m.warm = C.addressof(wi)
        r = Result()
        f = _zsolve_mixed if cplx else _dsolve_mixed
        _chk(f(self._h, n, n2, _p(Y), _p(J64), _p(YU), _p(ZZ), _p(SY), C.byref(o), C.byref(m), C.byref(r)),
             "jkg_zsolve_mixed" if cplx else "jkg_dsolve_mixed")
        out = dict(YU=YU, ZZ=ZZ, SY=SY, info=r.info, trace=trace[:min(r.steps, trace_cap)], tt=r.tt,
                   pn=(r.npos, r.nneg), n2=r.n2, seconds=r.seconds, steps=r.steps, **_counts(r))
        if ctr is not None:
            out["ctl"] = ctr[:min(r.steps, ctr.shape[0])].copy()
        if wi is not None:
            out["warm"] = {f: getattr(wi, f) for f, _ in WarmInfo._fields_ if f != "reserved"}
        return out

    def dsolve_noalloc(self, n, d_y, J, d_yu, d_zz, d_sy, n2=0, fused=False, max_steps=-1, trace_cap=None):
        """jkg_dsolve_noalloc: dsolve on device buffers (addresses of n x n column-major FP64 arrays
        d_y, d_yu, d_zz and n FP64 d_sy on this context's device), with no device allocation; N <= 256.
        Returns dict(info, trace, tt, pn, n2, seconds, steps)."""
        J64 = np.ascontiguousarray(J, dtype=np.int64)
        if trace_cap is None:
            trace_cap = 64 * n + 64 if max_steps < 0 else max_steps + 1
        trace = np.zeros(max(trace_cap, 1), STEPREC_DTYPE)
        o = Opts(0, 0, 1 if fused else 0, max_steps, 0, trace_cap, _p(trace))
        r = Result()
        _chk(_dsolve_noalloc(self._h, n, n2, C.c_void_p(d_y), _p(J64), C.c_void_p(d_yu), C.c_void_p(d_zz),
                             C.c_void_p(d_sy), C.byref(o), C.byref(r)), "jkg_dsolve_noalloc")
        return dict(info=r.info, trace=trace[:min(r.steps, trace_cap)], tt=r.tt, pn=(r.npos, r.nneg), n2=r.n2,
                    seconds=r.seconds, steps=r.steps, **_counts(r))

    # ---------------------------------------------------------------- batched solver
    def dsolve_batched(self, Ys, Js, n2s=None, fused=False, max_steps=-1, grid=0, variant=-1, tol=0.0):
        """Independent real HSVDs (djk semantics each) advanced together (jkg_dsolve_batched).
        Returns (list of dict(YU, ZZ, SY, info) per instance, dict of the run: seconds, steps,
        grid, variant, tune_seconds)."""
        ns = np.array([np.asarray(Y).shape[0] for Y in Ys], np.int32)
        for Y, J, n in zip(Ys, Js, ns):
            if np.asarray(Y).shape != (n, n) or len(J) != n:
                raise ValueError("instance shapes")
        y = np.concatenate([np.asfortranarray(Y, dtype=np.float64).ravel(order="F") for Y in Ys])
        j = np.concatenate([np.asarray(J, dtype=np.int64) for J in Js])
        n2a = None if n2s is None else np.ascontiguousarray(n2s, dtype=np.int32)
        yu = np.zeros_like(y)
        zz = np.zeros_like(y)
        sy = np.zeros(int(ns.sum()))
        info = np.zeros(len(Ys), np.int32)
        o = BatchOpts()
        _batch_default(C.byref(o))
        o.fused, o.max_steps, o.grid, o.variant, o.tol = (1 if fused else 0), max_steps, grid, variant, tol
        _chk(_dsolve_batched(self._h, len(Ys), _p(ns), None if n2a is None else _p(n2a), _p(y), _p(j), _p(yu),
                             _p(zz), _p(sy), _p(info), C.byref(o)), "jkg_dsolve_batched")
        out, mo, so = [], 0, 0
        for i, n in enumerate(ns):
            n = int(n)
            out.append(dict(YU=yu[mo:mo + n * n].reshape((n, n), order="F"),
                            ZZ=zz[mo:mo + n * n].reshape((n, n), order="F"), SY=sy[so:so + n].copy(),
                            info=int(info[i])))
            mo += n * n
            so += n
        return out, dict(seconds=o.seconds, steps=o.steps, grid=o.tuned_grid, variant=o.tuned_variant,
                         tune_seconds=o.tune_seconds)

    # ---------------------------------------------------------------- controller training
    def ctl_grad(self, X, T, params, W=None):
        """Loss, gradient and outputs of the controller MLP over the samples (library kernels)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        P = np.ascontiguousarray(params, dtype=np.float32)
        ns = X.shape[0]
        Wv = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
        g = np.zeros(CTL_NPARAM, np.float32)
        y = np.zeros((ns, 2), np.float32)
        loss = np.zeros(1, np.float32)
        _chk(_ctl_grad(self._h, ns, _p(X), _p(T), None if Wv is None else _p(Wv), _p(P), _p(g), _p(loss), _p(y)),
             "jkg_ctl_grad")
        return float(loss[0]), g, y

    def ctl_fit(self, X, T, params, W=None, epochs=50, batch=256, lr=1e-3, weight_decay=0.0, seed=1):
        """Trains the controller MLP on the GPU; returns (params, per-epoch full-set loss)."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        P = np.array(params, dtype=np.float32, copy=True)
        Wv = None if W is None else np.ascontiguousarray(W, dtype=np.float32)
        loss = np.zeros(epochs, np.float32)
        _chk(_ctl_fit(self._h, X.shape[0], _p(X), _p(T), None if Wv is None else _p(Wv), epochs, batch, lr,
                      weight_decay, seed, _p(P), _p(loss)), "jkg_ctl_fit")
        return P, loss

    def dprofile_mixed(self, Y, J, n2=0, steps=4, fused=False, **mixed):
        """As dprofile, for the steps of the mixed solver."""
        cplx = np.iscomplexobj(Y)
        Y = np.asfortranarray(Y, dtype=np.complex128 if cplx else np.float64)
        n = Y.shape[0]
        ms = np.zeros(len(STAGES))
        r = Result()
        m = mixed_opts(**mixed)
        f = _zprofile_mixed if cplx else _dprofile_mixed
        _chk(f(self._h, n, n2, _p(Y), _p(np.ascontiguousarray(J, dtype=np.int64)), 1 if fused else 0, steps,
               C.byref(m), _p(ms), C.byref(r)), "jkg_zprofile_mixed" if cplx else "jkg_dprofile_mixed")
        return dict(stage_ms=dict(zip(STAGES, ms.tolist())), steps=r.steps, n2=r.n2, seconds=r.seconds,
                    **_counts(r))

    def dweights_mixed(self, A, J, fused=False):
        """Sort keys of the mixed solver's pair weights, indexed by pair_index."""
        cplx = np.iscomplexobj(A)
        A = np.asfortranarray(A, dtype=np.complex128 if cplx else np.float64)
        n = A.shape[0]
        keys = np.zeros(n * (n - 1) // 2, np.uint64)
        f = _zweights_mixed if cplx else _dweights_mixed
        _chk(f(self._h, n, _p(A), _p(np.ascontiguousarray(J, dtype=np.int64)), 1 if fused else 0, _p(keys)),
             "jkg_zweights_mixed" if cplx else "jkg_dweights_mixed")
        return keys

    def dprofile(self, Y, J, n2=0, steps=4, fused=False):
        """Per-stage device time (ms summed over `steps` steps from Y) and work counts."""
        cplx = np.iscomplexobj(Y)
        Y = np.asfortranarray(Y, dtype=np.complex128 if cplx else np.float64)
        n = Y.shape[0]
        ms = np.zeros(len(STAGES))
        r = Result()
        f = _zprofile if cplx else _dprofile
        _chk(f(self._h, n, n2, _p(Y), _p(np.ascontiguousarray(J, dtype=np.int64)), 1 if fused else 0, steps,
               _p(ms), C.byref(r)), "jkg_zprofile" if cplx else "jkg_dprofile")
        return dict(stage_ms=dict(zip(STAGES, ms.tolist())), steps=r.steps, n2=r.n2, seconds=r.seconds,
                    **_counts(r))

    def set_frames(self, every=1, size=64, max_frames=256):
        """Record size x size pictures of |A| every `every` steps of later solves (jkg_set_frames);
        after a solve, frames() returns them as (count, size, size) uint8.  every=0 turns it off."""
        if every <= 0:
            _chk(_set_frames(self._h, 0, 0, 0, None, None), "jkg_set_frames")
            self._frames = None
            return
        buf = np.zeros((max_frames, size, size), np.uint8)
        cnt = np.zeros(1, np.int32)
        _chk(_set_frames(self._h, every, size, max_frames, _p(buf), _p(cnt)), "jkg_set_frames")
        self._frames = (buf, cnt)

    def frames(self):
        """The frames of the last solve (see set_frames)."""
        buf, cnt = getattr(self, "_frames", None) or (np.zeros((0, 0, 0), np.uint8), np.zeros(1, np.int32))
        return buf[:int(cnt[0])].copy()

    def set_select(self, rounds=0, limit=65536):
        """The greedy walk's schedule (jkg_set_select); the pivots do not depend on it."""
        _chk(_set_select(self._h, rounds, limit), "jkg_set_select")

    # ---------------------------------------------------------------- test matrices
    def dgen(self, lam, rounds=0, tmax=0.5, seed=1):
        """(G, J) with hyperbolic singular values sqrt|lam| (jkg_dgen): G = Q diag(sqrt|lam|) H."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        n = lam.size
        G = np.zeros((n, n), order="F")
        J = np.zeros(n, np.int64)
        _chk(_dgen(self._h, n, _p(lam), rounds, tmax, seed, _p(G), _p(J)), "jkg_dgen")
        return G, J

    # ---------------------------------------------------------------- plug-ins
    def srt(self, dz, nn=None):
        """ASRT: sorts dz[:nn] (AW records) in place by AW_CMP."""
        nn = len(dz) if nn is None else nn
        info = C.c_int64(0)
        _chk(_srt(self._h, nn, len(dz), _p(dz), C.byref(info)), "jkg_srt")
        return info.value

    def ncp(self, n, dz, n2, nn=None, id_ncp=2):
        """ANCP: disjoint pivots of the sorted dz[:nn] -- the greedy walk (id_ncp 1 or 2) or the best
        total weight over all start offsets (id_ncp 3, AW_NCP3); returns 1-based positions."""
        nn = len(dz) if nn is None else nn
        sl = C.c_int32(0)
        step = np.zeros(max(n2, 1), np.int32)
        info = C.c_int64(0)
        f, name = (_ncp3, "jkg_ncp3") if id_ncp == 3 else (_ncp, "jkg_ncp")
        _chk(f(self._h, n, nn, len(dz), _p(dz), n2, C.byref(sl), _p(step), C.byref(info)), name)
        return step[:sl.value].copy()

    # ---------------------------------------------------------------- step level
    def dhsvd2(self, A, J, th1=0.75, fused=False):
        """DHSVD2 of a batch: A (count, 4) column-major 2x2 blocks, J (count, 2).
        Returns U (count,4), Z (count,4), D (count,2), INFO (count)."""
        cplx = np.iscomplexobj(A)
        dt = np.complex128 if cplx else np.float64
        A = np.ascontiguousarray(A, dtype=dt).reshape(-1, 4)
        J = np.ascontiguousarray(J, dtype=np.int32).reshape(-1, 2)
        cnt = A.shape[0]
        U = np.zeros((cnt, 4), dt)
        Z = np.zeros((cnt, 4), dt)
        D = np.zeros((cnt, 2))
        info = np.zeros(cnt, np.int32)
        f = _zhsvd2 if cplx else _dhsvd2
        _chk(f(self._h, cnt, _p(A), _p(J), th1, 1 if fused else 0, _p(U), _p(Z), _p(D), _p(info)), "jkg_dhsvd2")
        return U, Z, D, info

    def dweights(self, A, J, fused=False):
        """Sort keys of every pair weight, indexed by pair_index."""
        cplx = np.iscomplexobj(A)
        A = np.asfortranarray(A, dtype=np.complex128 if cplx else np.float64)
        n = A.shape[0]
        keys = np.zeros(n * (n - 1) // 2, np.uint64)
        f = _zweights if cplx else _dweights
        _chk(f(self._h, n, _p(A), _p(np.ascontiguousarray(J, dtype=np.int64)), 1 if fused else 0, _p(keys)),
             "jkg_dweights")
        return keys

    def dbuild(self, A, J, n2, fused=False):
        """DSTEP_BUILD: (sorted finite candidates as AW records, chosen 1-based positions)."""
        A = np.asfortranarray(A, dtype=np.float64)
        n = A.shape[0]
        nn = n * (n - 1) // 2
        srt = np.zeros(nn, AW_DTYPE)
        step = np.zeros(max(n // 2, 1), np.int32)
        it = C.c_int32(0)
        sl = C.c_int32(0)
        _chk(_dbuild(self._h, n, n2, _p(A), _p(np.ascontiguousarray(J, dtype=np.int64)), 1 if fused else 0,
                     C.byref(it), _p(srt), C.byref(sl), _p(step)), "jkg_dbuild")
        return srt[:it.value].copy(), step[:sl.value].copy()

    def dstep(self, A, J, U, Z, n2, fused=False):
        """One DSTEP_EXEC on copies of (A, U, Z); returns (A, U, Z, trace record)."""
        A = np.array(A, dtype=np.float64, order="F", copy=True)
        U = np.array(U, dtype=np.float64, order="F", copy=True)
        Z = np.array(Z, dtype=np.float64, order="F", copy=True)
        n = A.shape[0]
        rec = np.zeros(1, STEPREC_DTYPE)
        _chk(_dstep(self._h, n, n2, _p(A), _p(np.ascontiguousarray(J, dtype=np.int64)), _p(U), _p(Z),
                    1 if fused else 0, _p(rec)), "jkg_dstep")
        return A, U, Z, rec[0]
