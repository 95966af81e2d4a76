"""Complete solves at the configurations' full sizes, through the C ABI.

  C2  N = 1024 real, column scales 10**U(-4,4), the plain solver (the reference's arithmetic)
  C3  N = 2048 complex, the FP32 warm start (FP32 steps until the relative off-norm is below
      1e-5, Newton-Schulz polar factors, FP64 steps)
  C5  N = 4096 real, column scales 10**U(-6,6), the trained step controller and the FP32
      warm start

Each result is checked by properties that do not depend on the trajectory:
  - reconstruction ||Y - YU diag(SY) ZZ||_F <= 64 N eps ||Y||_F;
  - orthogonality ||YU^H YU - I||_2 <= 64 N eps (spectral norm: the backward-error measure);
  - J-unitarity ||ZZ J ZZ^H - J||_2 <= 64 N eps ||ZZ||_2^2 (the hyperbolic factor may be
    ill-conditioned, so relative to its own norm);
  - the singular values against sqrt|eig(Y J Y^H)| of each J class, within that reference's own
    error bound N eps (sigma + ||Y||_2^2 / sigma), for the eigenvalues above eig's own error
    N eps ||Y||_2^2 (eigvalsh in FP64 on the GPU);
and, for C3 and C5, index by index against the plain solver's run of the same input
(tests/golden/make_plain_sy.py), the reference's arithmetic at a size the reference itself
needs days for:
  - C3: every value within the flat 64 N eps;
  - C5: every value within 64 N eps except where the reference's two builds themselves are
    further apart: the plain runs with FMAD unfused and fused (both committed) differ by up
    to 3069 N eps on sigma_min = 1.7e-7 and by more than 64 N eps on 17 values, so a result
    that follows another trajectory cannot be held to 64 N eps of one of them there.  The
    bound is max(64 N eps, the largest relative distance between the two builds), checked
    on each value, with the median held to 1 N eps and the count above 64 N eps reported.
    (The data do fix those values more tightly: one rounding of each entry of Y moves the
    exact sigma_min by at most 62 N eps to first order, kappa = |u|^T |Y| |z| / sigma.  The
    method does not: the mixed solver run on Y with every entry moved by one ulp lands 939
    N eps from its own run on Y; bench.py --c5-noise measures both.)
"""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def spectral_sym(t):
    """Largest |eigenvalue| of a Hermitian device matrix."""
    import torch
    return float(torch.linalg.eigvalsh(t).abs().max())


def properties(Y, J, out):
    import torch
    dev = torch.device("cuda")
    n = Y.shape[0]
    y = torch.from_numpy(np.asfortranarray(Y)).to(dev)
    u = torch.from_numpy(np.asfortranarray(out["YU"])).to(dev)
    zz = torch.from_numpy(np.asfortranarray(out["ZZ"])).to(dev)
    s = torch.from_numpy(out["SY"]).to(dev)
    jd = torch.from_numpy(np.asarray(J, dtype=np.float64)).to(dev).to(y.dtype)
    eye = torch.eye(n, dtype=y.dtype, device=dev)
    rec = float(torch.linalg.norm(y - (u * s[None, :].to(y.dtype)) @ zz) / torch.linalg.norm(y))
    orth = spectral_sym(u.mH @ u - eye)
    zn2 = spectral_sym(zz @ zz.mH)
    junit = spectral_sym((zz * jd[None, :]) @ zz.mH - torch.diag(jd)) / zn2
    # sqrt|eig(Y J Y^H)| per sign class, where eig resolves them (|lambda| above its own error
    # N eps ||Y||_2^2; below it the sign of an eigenvalue is not determined), largest first
    ev = torch.linalg.eigvalsh((y * jd[None, :]) @ y.mH).cpu().numpy()
    y2 = spectral_sym(y.mH @ y)
    del y, u, zz, s, eye
    torch.cuda.empty_cache()
    worst = 0.0
    sy = out["SY"]
    for cls in (1, -1):
        a = np.sort(np.abs(sy[np.asarray(J) == cls]))[::-1]
        e = np.sqrt(np.sort(np.abs(ev[ev > 0] if cls > 0 else ev[ev < 0]))[::-1])
        m = int(np.sum(e * e > n * EPS * y2))
        assert m >= 1 and a.size >= m
        worst = max(worst, float(np.max(np.abs(a[:m] - e[:m]) / (n * EPS * (e[:m] + y2 / e[:m])))))
    return rec, orth, junit, worst


def check_properties(Y, J, out):
    n = Y.shape[0]
    rec, orth, junit, eig = properties(Y, J, out)
    print("N=%d rec %.3g orth %.3g junit %.3g (64 N eps = %.3g); eig distance %.3g of its bound"
          % (n, rec, orth, junit, 64 * n * EPS, eig))
    assert rec <= 64 * n * EPS, rec
    assert orth <= 64 * n * EPS, orth
    assert junit <= 64 * n * EPS, junit
    assert eig <= 1.0, eig


def rel_n_eps(sy, ref):
    n = ref.size
    return np.abs(sy - ref) / np.abs(ref) / (n * EPS)


@pytest.mark.timeout(300)
def test_c2_plain_full_size(ctx):
    from jkogb_amd import inputs
    Y, J = inputs.make("c2", seed=0)
    out = ctx.dsolve(Y, J, 512)
    # the reference's termination (not the step guard): INFO steps with transforms and a last one without
    assert out["info"] >= 0 and out["steps"] == out["info"] + 1, (out["info"], out["steps"])
    assert out["pn"] == (512, 512)
    check_properties(Y, J, out)


@pytest.mark.timeout(300)
def test_c3_warm_start_full_size(ctx):
    from jkogb_amd import inputs
    Y, J = inputs.make("c3", seed=0)
    ref = np.load(os.path.join(GOLDEN, "solve_c3_n2048_plain_sy.npy"))
    out = ctx.dsolve_mixed(Y, J, 1024, fp32_warm=1, key_bits=32, trace_cap=1)
    w = out["warm"]
    assert out["info"] >= 0
    assert w["fallback"] == 0 and w["steps32"] > 0 and w["warm_end"] in (1, 2, 3), w
    r = rel_n_eps(out["SY"], ref)
    print("C3 warm: %d steps (%d FP32), SY vs plain: max %.3f, median %.4f N eps"
          % (out["steps"], w["steps32"], r.max(), np.median(r)))
    assert r.max() <= 64, r.max()
    check_properties(Y, J, out)


def c5_reference():
    unf = np.load(os.path.join(GOLDEN, "solve_c5_n4096_plain_sy.npy"))
    fma = np.load(os.path.join(GOLDEN, "solve_c5_n4096_plain_fma_sy.npy"))
    return unf, fma, float(rel_n_eps(fma, unf).max())


@pytest.mark.timeout(400)
@pytest.mark.parametrize("variant", ["learned", "warm", "warm_learned"])
def test_c5_full_size(ctx, variant):
    from jkogb_amd import inputs
    Y, J = inputs.make("c5", seed=0)
    ref, fma, spread = c5_reference()
    ctl = dict(ctl_params=np.load(os.path.join(ROOT, "jkogb_amd", "ctl_params.npy")))
    kw = {"learned": ctl, "warm": dict(fp32_warm=1, key_bits=32),
          "warm_learned": dict(fp32_warm=1, key_bits=32, **ctl)}[variant]
    out = ctx.dsolve_mixed(Y, J, 2048, trace_cap=1, **kw)
    assert out["info"] >= 0
    r = rel_n_eps(out["SY"], ref)
    rf = rel_n_eps(out["SY"], fma)
    print("C5 %s: %d steps, %.1f s; SY vs plain: max %.1f, median %.4f N eps, %d above 64 N eps; vs the fused "
          "plain run: max %.1f, %d above 64 N eps; builds' spread %.1f N eps"
          % (variant, out["steps"], out["seconds"], r.max(), np.median(r), int(np.sum(r > 64)), rf.max(),
             int(np.sum(rf > 64)), spread))
    assert np.median(r) <= 1.0
    assert r.max() <= max(64.0, spread), (r.max(), spread)
    check_properties(Y, J, out)
