#!/usr/bin/env python3
"""Benchmark of the HSVD step at N = 4096 on one MI355X: the mixed solver
against the plain FP64 solver on the same input.

One "step" is one pass of the hot path (pair weights -> sort -> greedy
select -> 2x2 HSVD -> rotations) on the N = 4096 real input of the
learned-controller configuration (Gaussian Y, column scales log-uniform in
[1e-6, 1e6], J half +1 / half -1, N_2 = 2048).  The timed regions start with
Y already in HBM and cover K consecutive steps from the start of the solve
(device events around the graph launches), once per solver:
  plain  jkg_dsolve: the reference's arithmetic (bit-exact with djk.exe)
  mixed  jkg_dsolve_mixed: Gram-form hyperbolic weights on the FP64 matrix
         cores and the FP64 convergence test (tol = DBL_EPSILON)

Printed as one JSON line:
  value        algorithmic HBM GB/s of the mixed solver's step (bytes below)
  ms_per_step  the mixed solver's device time per step
  plain        the same for the plain solver; step_speedup = plain / mixed
  roofline     the stage that takes longest in the mixed step, against its own bound:
               the weights stage against the FP64 matrix cores (the Gram product's
               2 npl nmi N flops), the others against HBM (algorithmic bytes below);
               traffic: HBM bytes per step measured by rocprofv3 FETCH_SIZE and
               WRITE_SIZE, scaled by a plain copy's readings (the summary in
               profiles/hbm_traffic_n4096_step_stages.md)
  rooflines    each stage against its bound (weights: FP64 MFMA; rotate, select,
               sort: HBM)
  c5_full      complete solves of the same N = 4096 input (the headline wall time):
               the mixed solver, the trained controller, and the controller after the
               FP32 warm start by default (also the warm start alone with
               --c5-variants mixed,learned,warm_learned,warm): device time, steps,
               the singular values against the plain (reference-arithmetic) run in
               tests/golden/solve_c5_n4096_plain_sy.npy in units of N eps, the
               residuals and the componentwise condition of the singular values
               (kappa / N: how far the exact values move with one rounding of each
               entry of Y); the reference's two builds' spread; with --c5-noise also
               the mixed solver on Y moved by one ulp per entry (the method's own
               spread); the plain solver's wall time is its measured per-step time
               times the 52 000 steps of that run, a labelled lower bound (it does
               not meet the reference's termination at this size)
  full_solve   complete HSVDs at N = 1024 (C2 input): plain, mixed, the FP32 warm
               start, and the trained step controller (jkogb_amd/ctl_params.npy):
               device time, steps, and the singular values' largest
               relative difference from the plain solver in units of N eps
  complex      the complex configuration (N=2048, Re/Im ~ N(0, 1/2), J balanced,
               N_2=1024): per-step device time of the plain, FP64 mixed and FP32
               warm-start steps (20 steps from the start each); with --complex-full
               also complete solves of the FP64 mixed solver and of the warm start
               (FP32 steps until the relative off-norm is below 1e-5, then FP64),
               with the warm start's SY difference from the FP64 mixed solver's
  batched      (with --batched) 4096 independent HSVDs, N uniform in [16, 256], Gaussian Y,
               random J, N_2 = N/2, advanced together by one persistent launch
               per step (jkg_dsolve_batched, tuned grid and variant, tol = eps):
               device time, against single mixed solves of a sample of the
               instances (projected to all 4096), and the sample's singular
               values against sqrt|eig(Y J Y^T)|
  cpu_baseline the reference (oracle/_ref/plain/djk.exe, built from the
               reference's own sources) on the same input, host cores: a bounded
               sample of its steps priced like the plain step, the CPU model, and
               its complete run of the N = 64 configuration beside the GPU's

Algorithmic bytes of one step (N, NN = N(N-1)/2, IT finite weights, SL
pivots, b2/b4 pivots with a row/column transform):
  (a) weights  8 N^2 (read A) + 8 NN (one key per pair, stored at the pair's index)
  (b) sort     24 NN (one pass reading and writing an 8-byte key and a 4-byte value: the
               reference figure; the sort itself makes eight passes over shrinking records)
  (c) select   12 IT (walk the sorted values once)
  (d) 2x2      112 SL
  (e) rotate   16 N^2 (A read + written once) + 32 N (b2 + b4) (U, Z columns)
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np
import torch  # before the library: torch carries its own HIP runtime under the soname libjkg.so links

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def step_bytes(n, it, sl, b2, b4):
    nn = n * (n - 1) // 2
    a = 8.0 * n * n + 8.0 * nn
    b = 24.0 * nn
    c = 12.0 * it
    d = 112.0 * sl
    e = 16.0 * n * n + 32.0 * n * (b2 + b4)
    return a + b + c + d + e, e


def cpu_reference_sample(y, j, n2, budget_s, threads):
    """Runs the reference djk.exe on the same input for at most budget_s and
    returns the per-step seconds of the steps it completed (its own CSV)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "plain", "djk.exe")
    if not os.path.exists(exe):
        return None
    from jkogb_amd import inputs
    n = y.shape[0]
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "c5")
        inputs.write_yj(fn, y, j)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        p = subprocess.Popen([exe, fn, str(n), str(n2)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env,
                             text=True)
        t0 = time.time()
        lines = []
        try:
            import selectors
            sel = selectors.DefaultSelector()
            sel.register(p.stderr, selectors.EVENT_READ)
            buf = ""
            while time.time() - t0 < budget_s and p.poll() is None:
                for key, _ in sel.select(timeout=1.0):
                    chunk = os.read(key.fileobj.fileno(), 65536).decode()
                    buf += chunk
                    while "\n" in buf:
                        line, buf = buf.split("\n", 1)
                        lines.append(line)
        finally:
            if p.poll() is None:
                p.kill()
            p.wait()
    steps = []
    for line in lines:
        f = [x.strip() for x in line.split(",")]
        if len(f) == 8 and re.match(r"^-?\d+$", f[0]):
            steps.append((int(f[0]), int(f[1]), float(f[2]) + float(f[3])))
    return steps


def timed_steps(ctx, solve, y, j, n2, steps, warmup, **kw):
    """K steps of one solver from the start (after W warmup steps that capture the step graph)."""
    solve(y, j, n2, max_steps=warmup, **kw)
    r = solve(y, j, n2, max_steps=steps, **kw)
    k = max(r["steps"], 1)
    ms = 1e3 * r["seconds"] / k
    b, _ = step_bytes(y.shape[0], r["sum_it"] / k, r["sum_sl"] / k, r["sum_b2"] / k, r["sum_b4"] / k)
    return dict(steps=k, ms_per_step=ms, gbs=b / (ms * 1e-3) / 1e9, bytes=b)


C5_GOLDEN = os.path.join(ROOT, "tests", "golden", "solve_c5_n4096_plain_sy.npy")
C5_GOLDEN_STEPS = 52000  # the plain (reference-arithmetic) run behind C5_GOLDEN: tests/golden/make_plain_sy.py


def accuracy(y, j, r, ref, eps):
    """SY against the reference-arithmetic run (per-index relative error in units of N eps) and the
    residuals ||Y - YU diag(SY) ZZ||_F / ||Y||_F, ||YU^T YU - I||_F, ||ZZ J ZZ^T - J||_F (FP64 on the GPU)."""
    import torch
    n = y.shape[0]
    out = {}
    if ref is not None:
        rel = np.abs(r["SY"] - ref) / np.abs(ref) / (n * eps)
        out.update({"sy_max_N_eps": round(float(rel.max()), 3), "sy_median_N_eps": round(float(np.median(rel)), 4),
                    "sy_over_64_N_eps": int(np.sum(rel > 64)),
                    "sy_worst": [round(float(ref[i]), 9) for i in np.argsort(-rel)[:3]]})
    dev = torch.device("cuda")
    Y = torch.from_numpy(np.asfortranarray(y)).to(dev)
    U = torch.from_numpy(np.asfortranarray(r["YU"])).to(dev)
    ZZ = torch.from_numpy(np.asfortranarray(r["ZZ"])).to(dev)
    S = torch.from_numpy(r["SY"]).to(dev)
    Jd = torch.from_numpy(np.asarray(j, dtype=np.float64)).to(dev)
    I = torch.eye(n, dtype=torch.float64, device=dev)
    out["res_rec"] = float(torch.linalg.norm(Y - (U * S[None, :]) @ ZZ) / torch.linalg.norm(Y))
    out["res_orth"] = float(torch.linalg.norm(U.T @ U - I))
    out["res_junit"] = float(torch.linalg.norm((ZZ * Jd[None, :]) @ ZZ.T - torch.diag(Jd)))
    # componentwise relative condition of each sigma_i under |dY| <= eps |Y|: first order,
    # |d sigma_i| / sigma_i <= eps (|u_i|^T |Y| |z_i|) / sigma_i with Z = J ZZ^T J (Y Z = U S), in units
    # of N eps: how far the exact values of the FP64 data move with one rounding of each entry
    Z = (Jd[:, None] * ZZ.T) * Jd[None, :]
    kap = ((U.abs() * (Y.abs() @ Z.abs())).sum(0) / S.abs()).cpu().numpy() / n
    imin = int(np.argmin(np.abs(r["SY"])))
    out["kappa_over_N"] = {"at_sigma_min": round(float(kap[imin]), 2), "max": round(float(kap.max()), 2),
                           "median": round(float(np.median(kap)), 4)}
    del Y, U, ZZ, S, I, Z
    torch.cuda.empty_cache()
    return out


def c5_full_run(ctx, y, j, n2, fused, plain_ms, log, variants, noise=False):
    """Complete solves of the N=4096 input (the headline wall time): the mixed solver, its FP32 warm
    start, the trained controller, and both together.  Accuracy against the plain solver's run of the
    same input (the reference's arithmetic; it never meets the reference's own termination at this size,
    so the committed run stops after 52 000 steps, NORMF2D 5.8e-75) and residuals.  The plain solver's
    wall time is the measured per-step time times those 52 000 steps (a lower bound: it does not end)."""
    n = y.shape[0]
    eps = np.finfo(float).eps
    ref = np.load(C5_GOLDEN) if os.path.exists(C5_GOLDEN) else None
    ctl = np.load(os.path.join(ROOT, "jkogb_amd", "ctl_params.npy"))
    last = [0]

    def prog(step, done):
        if done or step - last[0] >= 5000:
            last[0] = step
            log("c5 solve: step %d%s" % (step, " (done)" if done else ""))

    plain_s = plain_ms * 1e-3 * C5_GOLDEN_STEPS
    out = {"workload": "N=4096 real, column scales 10**U(-6,6), J 2048/2048, N_2=2048, complete solves",
           "plain_projected": {"steps": C5_GOLDEN_STEPS, "seconds": round(plain_s, 1),
                               "note": "the plain solver does not meet the reference's termination at this size; "
                                       "%d of its measured %.3f ms steps (the committed reference-arithmetic run "
                                       "took 712 s for them on one MI355X); speed-ups below are against this lower "
                                       "bound" % (C5_GOLDEN_STEPS, plain_ms)},
           "accuracy_reference": "tests/golden/solve_c5_n4096_plain_sy.npy (plain solver, 52000 steps)"
           if ref is not None else None,
           "bound_64_N_eps": 64}
    fma_ref = os.path.join(ROOT, "tests", "golden", "solve_c5_n4096_plain_fma_sy.npy")
    if ref is not None and os.path.exists(fma_ref):
        spread = np.abs(np.load(fma_ref) - ref) / np.abs(ref) / (n * eps)
        out["reference_builds_spread"] = {"what": "the plain solver with FMAD fused against unfused (the reference's "
                                                  "two builds): per-index relative difference in N eps",
                                          "max_N_eps": round(float(spread.max()), 1),
                                          "over_64_N_eps": int(np.sum(spread > 64))}
    for name, kw in variants:
        kw = dict(kw)
        if kw.pop("learned", False):
            kw["ctl_params"] = ctl
        last[0] = 0
        r = ctx.dsolve_mixed(y, j, n2, fused=fused, trace_cap=1, progress=prog, **kw)
        rec = {"steps": r["steps"], "seconds": round(r["seconds"], 3),
               "speedup_vs_plain": round(plain_s / r["seconds"], 2)}
        if "warm" in r:
            w = r["warm"]
            rec["warm"] = {"fp32_steps": w["steps32"], "fp32_seconds": round(w["seconds32"], 3),
                           "switch_seconds": round(w["seconds_switch"], 4), "off_rel_at_switch": w["off_rel"],
                           "newton_schulz": [w["ns_u"], w["ns_z"]], "fallback": w["fallback"]}
        rec.update(accuracy(y, j, r, ref, eps))
        out[name] = rec
        if name == "mixed" and noise:
            # the same solver on Y moved to a neighbouring double in every entry (relative change <= eps):
            # the spread of the method's own trajectory, against which a bound on the distance to one
            # run of the reference can be read
            rng = np.random.default_rng(7)
            yp = np.where(rng.random(y.shape) < 0.5, np.nextafter(y, np.inf), np.nextafter(y, -np.inf))
            rp = ctx.dsolve_mixed(yp, j, n2, fused=fused, trace_cap=1, progress=prog)
            rel = np.abs(rp["SY"] - r["SY"]) / np.abs(r["SY"]) / (n * eps)
            out["one_ulp_noise"] = {"what": "mixed solver on Y with every entry moved by one ulp, against the mixed "
                                            "solver on Y: per-index relative difference in N eps",
                                    "steps": rp["steps"], "max_N_eps": round(float(rel.max()), 1),
                                    "median_N_eps": round(float(np.median(rel)), 4),
                                    "over_64_N_eps": int(np.sum(rel > 64))}
            log("c5 one-ulp noise: max %.1f N eps" % rel.max())
        log("c5 %s: %d steps, %.1f s, max %.1f N eps" % (name, r["steps"], r["seconds"], rec.get("sy_max_N_eps", -1)))
    return out


def complex_run(ctx, fused, log, steps=20, full=True):
    """The complex configuration: per-step times and complete solves of the FP64 mixed solver
    and of its FP32 warm start."""
    from jkogb_amd import inputs
    y, j = inputs.make("c3", seed=0)
    n = y.shape[0]
    eps = np.finfo(float).eps

    def per_step(solve, **kw):
        solve(y, j, n // 2, max_steps=2, fused=fused, **kw)
        r = solve(y, j, n // 2, max_steps=steps, fused=fused, **kw)
        return 1e3 * r["seconds"] / r["steps"]

    plain_ms = per_step(ctx.dsolve)
    mixed_ms = per_step(ctx.dsolve_mixed)
    # the FP32 phase alone: its own device time (the step-limited run ends with the hand-over)
    ctx.dsolve_mixed(y, j, n // 2, max_steps=2, fused=fused, fp32_warm=1, key_bits=32, switch_tol=1e-300)
    r32 = ctx.dsolve_mixed(y, j, n // 2, max_steps=steps, fused=fused, fp32_warm=1, key_bits=32, switch_tol=1e-300)
    warm_ms = 1e3 * r32["warm"]["seconds32"] / max(r32["warm"]["steps32"], 1)
    last = [0]

    def prog(step, done):
        if done or step - last[0] >= 5000:
            last[0] = step
            log("complex solve: step %d%s" % (step, " (done)" if done else ""))

    if not full:
        m = w = None
    else:
        m = ctx.dsolve_mixed(y, j, n // 2, fused=fused, trace_cap=1, progress=prog)
        last[0] = 0
        w = ctx.dsolve_mixed(y, j, n // 2, fused=fused, trace_cap=1, progress=prog, fp32_warm=1, key_bits=32)
    out = {"workload": "complex N=2048, Re/Im ~ N(0,1/2), J 1024/1024 permuted, N_2=1024",
           "ms_per_step": {"plain": round(plain_ms, 4), "mixed_fp64": round(mixed_ms, 4),
                           "warm_fp32_phase": round(warm_ms, 4)},
           "step_speedup_mixed": round(plain_ms / mixed_ms, 3),
           "note": "ms_per_step: the first %d steps of each solver (warm_fp32_phase: the FP32 phase's own device "
                   "time); a complete plain solve at this size (~22k steps) is not part of the default run" % steps}
    if m is None:
        return out
    wi = w["warm"]
    out.update({"mixed": {"steps": m["steps"], "seconds": round(m["seconds"], 3)},
            "warm": {"steps": w["steps"], "seconds": round(w["seconds"], 3), "fp32_steps": wi["steps32"],
                     "fp32_seconds": round(wi["seconds32"], 3), "switch_seconds": round(wi["seconds_switch"], 4),
                     "off_rel_at_switch": wi["off_rel"], "newton_schulz": [wi["ns_u"], wi["ns_z"]],
                     "fallback": wi["fallback"],
                     "sy_err_vs_mixed_N_eps": round(float(np.max(np.abs(w["SY"] - m["SY"]) / np.abs(m["SY"])))
                                                    / (n * eps), 4)},
                "speedup_warm_vs_mixed": round(m["seconds"] / w["seconds"], 3)})
    return out


def batched_run(ctx, count, fused, sample=8):
    """The batched configuration: count instances, N uniform in [16, 256], solved together with the
    convergence test (tol = eps); a sample of them solved one at a time by the mixed solver (the same
    convergence test), and the batch's singular values of the sample against sqrt|eig(Y J Y^T)|."""
    from jkogb_amd import inputs
    rng = np.random.default_rng(4)
    ns = rng.integers(16, 257, count)
    ys, js = [], []
    for i, n in enumerate(ns):
        y, j = inputs.make("c4", n=int(n), seed=10000 + i)
        ys.append(y)
        js.append(j)
    eps = np.finfo(float).eps
    outs, run = ctx.dsolve_batched(ys, js, fused=fused, tol=eps)
    infos = np.array([o["info"] for o in outs])
    t1, worst, vs_eig = 0.0, 0.0, 0.0
    for i in rng.choice(count, sample, replace=False):
        n, y, j = int(ns[i]), ys[i], js[i]
        r = ctx.dsolve_mixed(y, j, n // 2, fused=fused)
        t1 += r["seconds"]
        worst = max(worst, float(np.max(np.abs(outs[i]["SY"] - r["SY"]) / np.abs(r["SY"]))) / (n * eps))
        # sqrt|eig(Y J Y^T)| in float64, per J class; its own error is about N eps (sigma + |Y|^2 / sigma)
        ev = np.linalg.eigvalsh(y @ np.diag(j.astype(float)) @ y.T)
        y2 = np.linalg.norm(y, 2) ** 2
        for cls in (1, -1):
            a = np.sort(np.abs(outs[i]["SY"][j == cls]))
            e = np.sort(np.sqrt(np.abs(ev[ev > 0] if cls > 0 else ev[ev < 0])))
            vs_eig = max(vs_eig, float(np.max(np.abs(a - e) / (n * eps * (e + y2 / e)))))
    single = t1 * count / sample
    return {"workload": "%d independent HSVDs, N uniform in [16,256] (sum N^2 = %d), Gaussian Y, random J, "
                        "N_2=N/2, one persistent launch per step, tol=eps" % (count, int(np.sum(ns.astype(np.int64) ** 2))),
            "seconds": round(run["seconds"], 4), "steps": run["steps"], "grid": run["grid"],
            "variant": run["variant"], "tune_seconds": round(run["tune_seconds"], 3),
            "instance_steps": {"min": int(infos.min()), "median": int(np.median(infos)), "max": int(infos.max())},
            "single_projected_seconds": round(single, 2), "speedup_vs_single": round(single / run["seconds"], 2),
            "sy_err_vs_single_N_eps_sampled": round(worst, 4),
            "sy_vs_eig_in_eig_error_units": round(vs_eig, 4),
            "note": "single_projected: %d sampled instances solved one at a time by jkg_dsolve_mixed (the same "
                    "convergence test), scaled to %d; sy_err_vs_single: the batch's largest relative difference "
                    "from those solves; sy_vs_eig: its largest difference from sqrt|eig(Y J Y^T)| (float64, per J "
                    "class) in units of that reference's own error N eps (sigma + |Y|^2/sigma)" % (sample, count)}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_n4096_step_stages.md")


def measured_traffic():
    """HBM bytes per step of each stage, from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary in
    profiles/ (its json block), or None if the summary is absent."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)["bytes_per_step"]
    except (OSError, IndexError, KeyError, ValueError):
        return None


def stage_profile(prof):
    ps = max(prof["steps"], 1)
    return {s: v / ps for s, v in prof["stage_ms"].items()}, ps


def cpu_c1_full(threads):
    """The reference djk.exe run to convergence on the N=64 configuration (seed 0), its own timer."""
    exe = os.path.join(ROOT, "oracle", "_ref", "plain", "djk.exe")
    if not os.path.exists(exe):
        return None
    from jkogb_amd import inputs
    y, j = inputs.make("c1", n=64, seed=0)
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "c1")
        inputs.write_yj(fn, y, j)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        p = subprocess.run([exe, fn, "64", "32"], capture_output=True, text=True, env=env, timeout=120)
    m = re.search(r"Executed\s+(\d+) steps with transformations in\s+([0-9.]+) s", p.stdout)
    return (int(m.group(1)), float(m.group(2))) if m else None


def dominant_roofline(n, j, pm, prof_m, traffic):
    """The roofline of the stage that takes the longest per mixed step, against its own bound:
    weights -> FP64 matrix cores (the Gram product's 2 npl nmi N flops), the others -> HBM with
    their algorithmic bytes (sort 24 NN, select 12 IT, rotate 16 N^2 + 32 N (b2 + b4))."""
    nn = n * (n - 1) // 2
    npl = int(np.sum(np.asarray(j) > 0))
    nmi = n - npl
    stage_bytes = {"sort": 24.0 * nn, "select": 12.0 * prof_m["sum_it"],
                   "rotate": step_bytes(n, 0, 0, prof_m["sum_b2"], prof_m["sum_b4"])[1]}
    gram_flops = 2.0 * npl * nmi * n
    dom = max(pm, key=lambda k: pm[k])
    ms = pm[dom]
    if dom == "weights":
        ach = gram_flops / (ms * 1e-3) / 1e12
        return dom, {"bound": "mfma", "kernel": "weights stage of the mixed step (the Gram product O(:,+)^T O(:,-) on "
                                                "the FP64 matrix cores, the column norms and the per-pair weights)",
                     "achieved": round(ach, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": (traffic or {}).get(dom),
                     "flops_per_launch": gram_flops, "ms_per_launch": round(ms, 4)}, stage_bytes, gram_flops
    ach = stage_bytes.get(dom, 0.0) / (ms * 1e-3) / 1e9
    return dom, {"bound": "hbm", "kernel": "%s stage of the mixed step" % dom, "achieved": round(ach, 2),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5),
                 "traffic": (traffic or {}).get(dom), "bytes_per_launch": int(stage_bytes.get(dom, 0.0)),
                 "ms_per_launch": round(ms, 4)}, stage_bytes, gram_flops


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--fused", action="store_true", help="FMAD = fma (default: the reference's unfused build)")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the complete N=1024 solves")
    ap.add_argument("--no-c5-full", action="store_true", help="skip the complete N=4096 solves (minutes)")
    ap.add_argument("--c5-noise", action="store_true",
                    help="also the mixed solver on Y moved by one ulp per entry (the method's own spread)")
    ap.add_argument("--c5-variants", default="mixed,learned,warm_learned",
                    help="complete N=4096 solves: any of mixed, warm, learned, warm_learned")
    ap.add_argument("--no-complex", action="store_true")
    ap.add_argument("--complex-full", action="store_true", help="also complete complex N=2048 solves")
    ap.add_argument("--batched", action="store_true", help="also the batched configuration (4096 instances)")
    ap.add_argument("--batch-count", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the step runs on one GPU (its sharding over several GPUs is a design, not built); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    torch.cuda.init()
    from jkogb_amd import inputs, jkg
    n = args.n
    y, j = inputs.make("c5", n=n, seed=0)
    n2 = n // 2
    ctx = jkg.Context(0)
    t0 = time.time()
    log = lambda msg: print("bench: %s (%.0f s)" % (msg, time.time() - t0), file=sys.stderr, flush=True)
    mixed = timed_steps(ctx, ctx.dsolve_mixed, y, j, n2, args.steps, args.warmup, fused=args.fused)
    plain = timed_steps(ctx, ctx.dsolve, y, j, n2, args.steps, args.warmup, fused=args.fused)
    wall = time.time() - t0
    log("timed steps: mixed %.3f ms, plain %.3f ms" % (mixed["ms_per_step"], plain["ms_per_step"]))
    # the FP32 phase of the warm start: its own device time per step (a tolerance it does not reach)
    ctx.dsolve_mixed(y, j, n2, max_steps=2, fused=args.fused, fp32_warm=1, key_bits=32, switch_tol=1e-300)
    r32 = ctx.dsolve_mixed(y, j, n2, max_steps=args.steps, fused=args.fused, fp32_warm=1, key_bits=32,
                           switch_tol=1e-300)
    warm_ms = 1e3 * r32["warm"]["seconds32"] / max(r32["warm"]["steps32"], 1)

    # per-stage device times (each stage launched and timed with HIP events on the library's stream)
    pm, psm = stage_profile(ctx.dprofile_mixed(y, j, n2, steps=args.profile_steps, fused=args.fused))
    pp, _ = stage_profile(ctx.dprofile(y, j, n2, steps=args.profile_steps, fused=args.fused))
    prof_m = ctx.dprofile_mixed(y, j, n2, steps=1, fused=args.fused)
    traffic = measured_traffic()
    dom, roof, stage_bytes, gram_flops = dominant_roofline(n, j, pm, prof_m, traffic)
    rot_gbs = stage_bytes["rotate"] / (pm["rotate"] * 1e-3) / 1e9
    sel_gbs = stage_bytes["select"] / (pm["select"] * 1e-3) / 1e9
    sort_gbs = stage_bytes["sort"] / (pm["sort"] * 1e-3) / 1e9
    log("stage profiles")

    c5 = None
    if not args.no_c5_full:
        table = {"mixed": {}, "warm": {"fp32_warm": 1, "key_bits": 32}, "learned": {"learned": True},
                 "warm_learned": {"fp32_warm": 1, "key_bits": 32, "learned": True}}
        variants = [(v, table[v]) for v in args.c5_variants.split(",") if v in table]
        c5 = c5_full_run(ctx, y, j, n2, args.fused, plain["ms_per_step"], log, variants, noise=args.c5_noise)
    full = None
    if not args.no_full:
        yc, jc = inputs.make("c2", n=1024, seed=0)
        rp = ctx.dsolve(yc, jc, 512, fused=args.fused)
        rf = ctx.dsolve(yc, jc, 512, fused=not args.fused)
        rm = ctx.dsolve_mixed(yc, jc, 512, fused=args.fused)
        rw = ctx.dsolve_mixed(yc, jc, 512, fused=args.fused, fp32_warm=1, key_bits=32)
        ctl = np.load(os.path.join(ROOT, "jkogb_amd", "ctl_params.npy"))
        rl = ctx.dsolve_mixed(yc, jc, 512, fused=args.fused, ctl_params=ctl)
        ne = 1024 * np.finfo(float).eps
        err = lambda r: round(float(np.max(np.abs(r["SY"] - rp["SY"]) / np.abs(rp["SY"])) / ne), 4)
        full = {"workload": "C2: N=1024 real, columns 10**U(-4,4), J balanced, N_2=512",
                "plain": {"steps": rp["steps"], "seconds": round(rp["seconds"], 4)},
                "mixed": {"steps": rm["steps"], "seconds": round(rm["seconds"], 4), "sy_err_N_eps": err(rm)},
                "warm": {"steps": rw["steps"], "seconds": round(rw["seconds"], 4), "sy_err_N_eps": err(rw),
                         "fp32_steps": rw["warm"]["steps32"]},
                "learned": {"steps": rl["steps"], "seconds": round(rl["seconds"], 4), "sy_err_N_eps": err(rl)},
                "speedup_mixed": round(rp["seconds"] / rm["seconds"], 3),
                "speedup_warm": round(rp["seconds"] / rw["seconds"], 3),
                "speedup_learned": round(rp["seconds"] / rl["seconds"], 3),
                "reference_build_spread_N_eps": err(rf),
                "note": "sy_err: largest relative SY difference against the plain (reference-exact) solver; "
                        "reference_build_spread: the same between the reference's unfused and fused builds"}
        log("full solves at N=1024")
    cplx = None
    if not args.no_complex:
        cplx = complex_run(ctx, args.fused, log, full=args.complex_full)
    batched = None
    if args.batched:
        batched = batched_run(ctx, args.batch_count, args.fused)
        log("batched: %d instances in %.3f s" % (args.batch_count, batched["seconds"]))
    gpu_c1 = None
    if not args.no_cpu:
        y1, j1 = inputs.make("c1", n=64, seed=0)
        r1 = ctx.dsolve(y1, j1, 32, fused=args.fused)
        gpu_c1 = (r1["info"], r1["seconds"])
    ctx.close()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        steps = cpu_reference_sample(y, j, n2, args.cpu_budget, threads)
        c1 = cpu_c1_full(threads)
        cpu_model = ""
        try:
            cpu_model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
        except (OSError, IndexError):
            pass
        if steps:
            per = float(np.mean([s[2] for s in steps]))
            cpu = {"value": round(plain["bytes"] / per / 1e9, 4), "unit": "GB/s", "cores": threads,
                   "kind": "reference", "cpu_model": cpu_model,
                   "sample": "the reference djk.exe (flang build of its own sources) on the same N=4096 input, "
                             "%d completed step(s) within %.0f s; %.3f s per step (its BUILDs+TRANSFs CSV), "
                             "priced with the plain step's algorithmic bytes" % (len(steps), args.cpu_budget, per),
                   "s_per_step": round(per, 4),
                   "c5_projected_s": round(per * C5_GOLDEN_STEPS, 0)}
        else:
            cpu = {"value": None, "unit": "GB/s", "cores": threads, "kind": "reference", "cpu_model": cpu_model,
                   "sample": "no step of the reference completed within %.0f s at N=4096" % args.cpu_budget}
        if c1:
            cpu["c1_full"] = {"workload": "N=64 real Gaussian, random J, N_2=32 (configuration 1), run to the end",
                              "reference_steps": c1[0], "reference_s": c1[1],
                              "gpu_plain_steps": gpu_c1[0] if gpu_c1 else None,
                              "gpu_plain_s": round(gpu_c1[1], 4) if gpu_c1 else None}
        log("cpu baseline")

    out = {
        "metric": "HSVD wall time + HBM GB/s at N=4096, mixed/learned vs plain FP64, same accuracy",
        "value": round(mixed["gbs"], 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": mixed["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "N=4096 real Gaussian Y, column scales log-uniform in [1e-6,1e6], J 2048 (+1) / "
                               "2048 (-1) permuted, N_2=2048; mixed solver (Gram-form FP64 weights, tol=eps, %s "
                               "FMAD) vs the plain FP64 solver" % ("fused" if args.fused else "unfused"),
                   "n": n, "n2": n2, "solver": "mixed-fp64"},
        "plain": {"ms_per_step": round(plain["ms_per_step"], 4), "gbs": round(plain["gbs"], 3),
                  "steps": plain["steps"]},
        "warm_fp32_ms_per_step": round(warm_ms, 4),
        "step_speedup": round(plain["ms_per_step"] / mixed["ms_per_step"], 3),
        "roofline": roof,
        "rooflines": [
            {"bound": "mfma", "kernel": "weights stage (k_colnorm_off + k_gemm_tn + k_trig_weights + "
                                        "k_hyp_weights_gram)",
             "achieved": round(gram_flops / (pm["weights"] * 1e-3) / 1e12, 3), "peak": FP64_PEAK_TFLOPS,
             "unit": "TFLOP/s", "frac": round(gram_flops / (pm["weights"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
             "flops_per_launch": gram_flops, "ms_per_launch": round(pm["weights"], 4)},
            {"bound": "hbm", "kernel": "k_rotate", "achieved": round(rot_gbs, 1), "peak": HBM_PEAK_GBS,
             "unit": "GB/s", "frac": round(rot_gbs / HBM_PEAK_GBS, 4), "traffic": (traffic or {}).get("rotate"),
             "bytes_per_launch": int(stage_bytes["rotate"]), "ms_per_launch": round(pm["rotate"], 4)},
            {"bound": "hbm", "kernel": "select stage (k_sel_walk rounds: locally dominant rounds over dense "
                                       "super-chunks, wave-0 resolution of sparse ones; k_sel_count + "
                                       "k_sel_write filters between rounds)",
             "achieved": round(sel_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(sel_gbs / HBM_PEAK_GBS, 5), "traffic": (traffic or {}).get("select"),
             "bytes_per_launch": int(stage_bytes["select"]), "ms_per_launch": round(pm["select"], 4)},
            {"bound": "hbm", "kernel": "sort stage (8 passes of k_sort_count, k_sort_scan, k_sort_scatter over "
                                       "records of 12, 8 and 4 bytes)",
             "achieved": round(sort_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(sort_gbs / HBM_PEAK_GBS, 5), "traffic": (traffic or {}).get("sort"),
             "bytes_per_launch": int(stage_bytes["sort"]), "ms_per_launch": round(pm["sort"], 4)},
        ],
        "stage_ms_per_step": {"mixed": {s: round(v, 4) for s, v in pm.items()},
                              "plain": {s: round(v, 4) for s, v in pp.items()}},
        "cpu_baseline": cpu,
        "c5_full": c5,
        "full_solve": full,
        "complex": cplx,
        "batched": batched,
        "host_wall_s": round(wall, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
