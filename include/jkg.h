/*
 * jkg.h -- C ABI of the MI355X J-Kogbetliantz hyperbolic SVD (libjkg.so).
 *
 * Y = YU * diag(SY) * ZZ with YU unitary and ZZ = J Z^H J, Z J-unitary
 * (Z^H J Z = J), J = diag(+-1): the decomposition computed by the reference's
 * djk.exe / zjk.exe (venovako/JKogb src/D/djk.F90, src/Z/zjk.F90).  Every
 * entry point is plain C: pointers, sizes and int status codes.  Matrices are
 * column-major, as the reference's FN.Y/FN.YU/FN.ZZ files; complex data are
 * interleaved (re, im) pairs of doubles (Fortran COMPLEX(KIND=8)).
 *
 * Return values: JKG_OK (0), a negative JKG_E* code, or (for argument
 * checks) minus the 1-based position of the offending argument, as the
 * reference's INFO convention does.
 */
#ifndef JKG_H
#define JKG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JKG_OK 0
#define JKG_EHIP (-1000)     /* a HIP runtime call failed (see jkg_last_hip_error) */
#define JKG_ENOMEM (-1001)   /* device allocation failed */
#define JKG_ENOTSUP (-1002)  /* an option or size this build does not support (e.g. N beyond the limits below) */
#define JKG_EJ (-1003)       /* a J entry outside {-1,+1} (the reference: R%TRU INFO=-2) */
#define JKG_ENOCONV (-1004)  /* the step loop ran past its guard of 1000 N + 1000 steps (1000 NN + 1000 for
                                the one-pair cyclic sweeps) */

typedef struct jkg_ctx jkg_ctx;

/* Size limits of the device step (JKG_ENOTSUP beyond them).  The rotation kernel
 * stages the two columns of A of a pivot in LDS (160 KiB per workgroup on gfx950):
 * N <= 10240 for real and N <= 5120 for complex problems.  Pair values are two
 * 16-bit indices.  ID_NCP=3 keeps a used-index map per wave: N <= 8192. */
#define JKG_NMAX_REAL 10240
#define JKG_NMAX_COMPLEX 5120
#define JKG_NMAX_NCP3 8192

/* The AW candidate record of the reference (src/shared/atypes.F90:18-24):
 * weight, row P < column Q (1-based), band B = Q - P, info. 24 bytes. */
typedef struct {
  double w;
  int32_t p, q, b, i;
} jkg_aw;

/* One line of the reference's per-step stderr CSV (src/D/dstep.F90:388-440):
 * STEP, OLDLEN, NEWLEN, TRIG, HYP, NORMF2D (tw); it = number of finite weights
 * (IT) of the step; build_s and transf_s = BUILDs and TRANSFs, the step's
 * device wall time up to its 2x2 pivots and from there to its end (seconds,
 * 100 MHz device clock). */
typedef struct {
  int32_t step, oldlen, newlen, trig, hyp, it;
  double tw;
  float build_s, transf_s;
} jkg_steprec;

typedef struct {
  int32_t id_ncp;         /* 0, 1 or 2: greedy walk (AW_NCP1/2 select the same pivots); 3: AW_NCP3, N <= 8192 */
  int32_t id_tru;         /* 0, 1 or 2 (pair enumeration; does not change results) */
  int32_t fused;          /* 0: FMAD(a,b,c) = a*b+c, two roundings (the reference's default build); 1: fma */
  int32_t max_steps;      /* < 0: run to convergence; else stop after this many steps */
  int32_t steps_per_sync; /* steps launched between two host polls of the device's done flag; 0: auto */
  int32_t trace_cap;      /* records available in trace */
  jkg_steprec *trace;     /* host buffer for the per-step trace, or NULL */
  /* progress hook, or NULL: called on the calling thread at every host poll of the
   * device (every steps_per_sync steps) with the number of steps executed so far
   * and whether the loop has ended; trace[0 .. min(steps, trace_cap)) already
   * holds their records, so a caller can print the CSV as the reference does
   * (line by line during the loop); for long solves that report as they go */
  void (*progress)(void *user, int32_t step, int32_t done);
  void *user;
} jkg_opts;

typedef struct {
  int32_t info;   /* >= 0: steps with transformations (the reference's INFO); < 0: error */
  int32_t tt;     /* number of trigonometric pairs (the reference's TT) */
  int32_t npos;   /* J entries equal to +1 */
  int32_t nneg;   /* J entries equal to -1 */
  int32_t n2;     /* N_2 after JSTEP_LEN */
  int32_t steps;  /* trace records written (steps executed, including the last) */
  double seconds; /* device time of the step loop (HIP events) */
  /* work done, summed over the steps: chosen pivots, finite weights, pivots
   * with a row transform (INFO&2), with a column transform (INFO&4), and
   * hyperbolic pairs whose weight scanned their two columns (ABODNDF2) */
  int64_t sum_sl, sum_it, sum_b2, sum_b4, sum_abod;
} jkg_result;

void jkg_opts_default(jkg_opts *o);
const char *jkg_strerror(int code);
int jkg_last_hip_error(void);

/* A context owns one HIP stream and the device buffers of the last problem
 * size; one context per host thread. */
int jkg_create(jkg_ctx **ctx, int device);
void jkg_destroy(jkg_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Process level: what djk.exe/zjk.exe FN N N_2 [ID_NCP [ID_TRU]] computes
 * between reading FN.Y/FN.J and writing FN.YU/FN.ZZ/FN.SY
 * (src/D/djk.F90:27-121): JSTEP_LEN, R%TRU, DSTEP_LOOP (U=I, Z=I, steps until
 * no pivot or no effective transform), SY = diag(A), ZZ = J Z^T J.
 * y is N x N (ld = n), j holds N int64 signs; yu, zz (N x N) and sy (N) are
 * written whenever the loop ran (JKG_OK): after a failed 2x2 (res->info < 0)
 * they hold U, J Z^T J and diag(A) at the failure, which the reference also
 * writes after its ERROR line (src/D/djk.F90:86-114).
 * --------------------------------------------------------------------------- */
int jkg_dsolve(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
               double *sy, const jkg_opts *opts, jkg_result *res);
int jkg_zsolve(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
               double *sy, const jkg_opts *opts, jkg_result *res);

/* jkg_dsolve_from / jkg_zsolve_from: checkpoint and resume of the same loop.
 * DSTEP_LOOP (src/D/dstep.F90:472-604; ZSTEP_LOOP, src/Z/zstep.F90) started from
 * its own state instead of (Y, I, I): a (the iterate A), u and z (the factors U
 * and Z as the steps accumulate them, before the finalisation's J Z^T J), all
 * N x N column-major, complex interleaved for the z variant, and the index step0
 * of the next step (0 with a = Y, u = z = I: jkg_dsolve's run).  The loop runs
 * until it ends or, with opts->max_steps = K >= 0, until K steps have run; then
 * (a, u, z) hold its state after the last step -- the checkpoint to resume from
 * with step0 = res->info -- and yu, zz, sy the finalised outputs of that state,
 * as jkg_dsolve writes them.  The loop ended (no pivot, no effective transform
 * or a failed 2x2) iff max_steps < 0 or res->info < step0 + max_steps.  Steps
 * keep their absolute index: the trace's STEP and res->info continue from step0,
 * res->steps counts this call's records (trace[0] is step step0), and frames
 * (jkg_set_frames) are taken at the steps s = 0 mod every.  Resuming is
 * bit-identical to running on: the state is all the loop carries from one step
 * to the next.  The plain solver's options (id_ncp, id_tru, fused) apply.
 * Argument errors: -(position). */
int jkg_dsolve_from(jkg_ctx *ctx, int n, int n2, double *a, double *u, double *z, const int64_t *j, int step0,
                    double *yu, double *zz, double *sy, const jkg_opts *opts, jkg_result *res);
int jkg_zsolve_from(jkg_ctx *ctx, int n, int n2, double *a, double *u, double *z, const int64_t *j, int step0,
                    double *yu, double *zz, double *sy, const jkg_opts *opts, jkg_result *res);

/* jkg_dsolve_noalloc: jkg_dsolve that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  y, yu, zz (N x N,
 * column-major) and sy (N) are DEVICE buffers on the context's device; j is a
 * host array.  Every working buffer lives in one fixed scratch array of the
 * library (6 MiB per GPU, held until the process ends): N <= 256 (JKG_ENOTSUP
 * beyond, outputs untouched), trace records <= 64 N + 64, ID_NCP 0-2.  The steps
 * are launched one by one on the context's stream (no graph); results, INFO and
 * trace are jkg_dsolve's bit for bit; res->seconds is host wall time.  All calls
 * on one GPU share the scratch array: calls that overlap on one GPU race. */
int jkg_dsolve_noalloc(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
                       double *sy, const jkg_opts *opts, jkg_result *res);

/* ---------------------------------------------------------------------------
 * Strategy level: the reference's DPROC plug-in procedures
 * (src/D/dtypes.F90:5-30), bound by DPROC_INIT (src/D/dstep.F90:76-134).
 * jkg_srt  = ASRT: sort dz[0..nn) by AW_CMP (W descending, then B, P, Q
 *            descending); info >= 0 on success.
 * jkg_ncp  = ANCP (AW_NCP1/AW_NCP2): greedy disjoint pivots from the sorted
 *            list; step[0..*sl) are 1-based list positions.
 * jkg_ncp3 = ANCP bound to AW_NCP3 (src/shared/atypes.F90:488-592, ID_NCP=3):
 *            the greedy walk from every start offset, the selection of the
 *            largest total weight (ties: the smallest start); N <= 8192.
 * jkg_tru  = ATRU (TRU1/TRU2, src/shared/jstep.F90:9-120); returns TT or -2.
 * jkg_jstep_len = JSTEP_LEN (src/shared/jstep.F90:124-133).
 * --------------------------------------------------------------------------- */
int jkg_srt(jkg_ctx *ctx, int nn, int nm, jkg_aw *dz, int64_t *info);
int jkg_ncp(jkg_ctx *ctx, int n, int nn, int nm, jkg_aw *dz, int n2, int32_t *sl, int32_t *step, int64_t *info);
int jkg_ncp3(jkg_ctx *ctx, int n, int nn, int nm, jkg_aw *dz, int n2, int32_t *sl, int32_t *step, int64_t *info);
/* jkg_set_select: the greedy walk's schedule for later solves on ctx (no
 * counterpart in the reference; a tuning knob): `rounds` walks per step (1..16;
 * 0: by the list length, about one per 768K pairs, 2 to 12), each but the last
 * stopping after `limit` (>= 2048, rounded up to whole super-chunks of 4096)
 * entries of its list and followed by a whole-GPU pass that keeps only the
 * entries with both indices free.  The pivots do not depend on it.  Default:
 * rounds by size, 65536 entries. */
int jkg_set_select(jkg_ctx *ctx, int rounds, int limit);
/* jkg_set_frames: pictures of |A| during later solves on ctx (the role of the
 * reference's ANIMATE frames, src/D/dstep.F90:485-495 and 549-592, written there
 * by libpvn's PVN_RVIS_*; no file format is shared): at the start of every
 * `every`-th step, before its transforms, a size x size map (row-major bytes)
 * whose pixel holds the largest |a_ij| of the rows and columns mapped to it on a
 * fixed log scale -- 0 for zero, else 1 + floor(8 (log10|a| + 16)) clamped to
 * 1..255 (8 levels per decade from 1e-16).  After each solve the first
 * min(frames recorded, max_frames) frames are in `frames` (max_frames * size^2
 * bytes, kept valid by the caller) and their number in *count.  every <= 0: off. */
int jkg_set_frames(jkg_ctx *ctx, int every, int size, int max_frames, uint8_t *frames, int32_t *count);
int jkg_tru(int id_tru, int n, const int32_t *j, int nn, int32_t *p, int32_t *q, int32_t pn[2]);
int jkg_jstep_len(int n, int n2);

/* ---------------------------------------------------------------------------
 * Step level, for callers that drive the loop themselves and for parity
 * checks of each stage against the reference.
 * jkg_dhsvd2   : DHSVD2 (src/D/dtransf.F90:614-683) of count 2x2 blocks
 *                (a: 4 per block, column-major; j: 2 per block); writes the
 *                U and Z factors (4 each), the new diagonal (2) and INFO.
 * jkg_zhsvd2   : ZHSVD2 (src/Z/ztransf.F90:925-1009), complex interleaved.
 * jkg_dweights : the weight of every pair (DMAGF2T/DMAGF2H,
 *                src/D/dstep.F90:12-72) as its sort key (see jkg_key_weight),
 *                indexed by jkg_pair_index.
 * jkg_dbuild   : DSTEP_BUILD (src/D/dstep.F90:138-282): the sorted finite
 *                candidates sorted[0..*it) and the chosen positions step[0..*sl).
 * jkg_dstep    : one DSTEP_EXEC on (a, u, z) in place; rec gets its CSV line.
 *                Returns JKG_OK; rec->newlen < 0 or rec->oldlen == 0 ends a loop.
 * --------------------------------------------------------------------------- */
int jkg_dhsvd2(jkg_ctx *ctx, int count, const double *a, const int32_t *j, double th1, int fused, double *u,
               double *z, double *d, int32_t *info);
int jkg_zhsvd2(jkg_ctx *ctx, int count, const double *a, const int32_t *j, double th1, int fused, double *u,
               double *z, double *d, int32_t *info);
int jkg_dweights(jkg_ctx *ctx, int n, const double *a, const int64_t *j, int fused, uint64_t *keys);
int jkg_zweights(jkg_ctx *ctx, int n, const double *a, const int64_t *j, int fused, uint64_t *keys);
int jkg_dbuild(jkg_ctx *ctx, int n, int n2, const double *a, const int64_t *j, int fused, int32_t *it,
               jkg_aw *sorted, int32_t *sl, int32_t *step);
int jkg_dstep(jkg_ctx *ctx, int n, int n2, double *a, const int64_t *j, double *u, double *z, int fused,
              jkg_steprec *rec);
/* jkg_dprofile: `steps` steps of the real solver from (y, j), each stage
 * launched on its own and timed with HIP events on the context's stream;
 * stage_ms[0..6] = begin, weights, sort, select, pivots, rotate, end
 * (milliseconds summed over the steps); res as for jkg_dsolve. */
int jkg_dprofile(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                 double *stage_ms, jkg_result *res);
/* ---------------------------------------------------------------------------
 * The mixed solver.  No counterpart in the reference: the fast path of this
 * library for callers that need the decomposition, not the reference's exact
 * trajectory.  Same dynamic ordering, termination and outputs as jkg_dsolve,
 * but the hyperbolic weights come from the Gram form: one FP64 matrix-core
 * product O(:,J=+1)^T O(:,J=-1) per step (O = A with a zero diagonal) and
 * O(1) work per pair, instead of a scan of two columns per pair (ABODNDF2).
 * Optionally an FP32 warm start, shortened candidate lists and a learned step
 * controller.  Results agree with jkg_dsolve's to rounding, not bit for bit.
 * --------------------------------------------------------------------------- */
#define JKG_W_PAIR 0 /* per-pair column scans, as jkg_dsolve */
#define JKG_W_GRAM 1 /* Gram form (default) */
typedef struct {
  int32_t weights;    /* JKG_W_PAIR or JKG_W_GRAM */
  int32_t fp32_warm;  /* 1: FP32 steps until the off-norm falls below switch_tol or stalls, then
                         FP64 (see jkg_warm_info); Gram-form weights only */
  double switch_tol;  /* relative off-norm that ends the FP32 phase (default 1e-5) */
  int32_t key_bits;   /* 64: full sort keys (default); 32: the warm start's keys cut to their top
                         32 bits, ties by pair index */
  int32_t controller; /* step controller: 0 none (N_2 and `threshold` every step); 1 the MLP, ctl = its
                         JKG_CTL_NPARAM parameters; 2 a phase table, ctl = 8 floats (see below) */
  const float *ctl;   /* controller parameters, or NULL */
  double threshold;   /* fixed threshold: weights below threshold * (largest weight) dropped; 0: none */
  double tol;         /* convergence test: a pair with max(|a_pq|,|a_qp|) <= tol * sqrt(|a_pp| |a_qq|) whose
                         transform would otherwise be the identity is dropped (default: DBL_EPSILON);
                         0: the reference's test only */
  float *ctl_trace;   /* per-step controller records (JKG_CTL_REC floats each), or NULL */
  int32_t ctl_trace_cap;
  int32_t ordering;   /* 0: dynamic (the reference's, default).  Cyclic sweep baselines (real; no
                         weights, sort or walk, no controller or threshold; a whole sweep without a
                         transform ends the run): 1 round robin, N/2 disjoint pairs a step; 2 and 3
                         the reference's pointwise sweeps (src/old/jk1.f90:101-146), one pair a
                         step, row-cyclic (1,2), (1,3), .., (2,3), .. or column-cyclic (1,2), (1,3),
                         (2,3), (1,4), .., with the reference's 2x2 and transforms (DHSVD2,
                         DSTEP_TRANSF) on the full matrix */
  struct jkg_warm_info *warm; /* out (fp32_warm): what the warm start did, or NULL */
  double ctl_tau_max; /* largest threshold a controller may set (0: 1e-6, the network's range) */
  float ctl_handoff;  /* log10(TW / TW_0) below which a controller hands every step back to the plain
                         full ordering, in [-16, 0) (0: -10) */
  int32_t reserved;
} jkg_mixed_opts;

/* The FP32 warm start of jkg_dsolve_mixed and jkg_zsolve_mixed: steps on FP32 copies of A, U and Z
 * (weights and 2x2 parameters in FP64) until A's relative off-norm
 * sqrt(off(A)^2 / |A|_F^2) falls below switch_tol (or has not improved by 1 %
 * in 64 steps); then U and Z are replaced by their polar factors in the
 * unitary and J-unitary groups (Newton-Schulz in FP64, X <- X (3I - X* X)/2),
 * A is recomputed as U^H Y Z in FP64 and FP64 steps finish the solve. */
typedef struct jkg_warm_info {
  int32_t steps32;        /* steps of the FP32 phase */
  int32_t warm_end;       /* 1 off-norm < switch_tol, 2 stalled, 3 converged in FP32, 4 step limit */
  int32_t ns_u, ns_z;     /* Newton-Schulz updates of U and of Z */
  int32_t fallback;       /* 1: a polar iteration could not start; the solve restarted in FP64 from Y */
  int32_t reserved;
  double off_rel;         /* relative off-norm at the switch */
  double res_u, res_z;    /* |X* X - I|_F before the last update (U, Z) */
  double seconds32;       /* device time of the FP32 phase */
  double seconds_switch;  /* device time of the hand-over */
} jkg_warm_info;

/* The step controller.  At the start of each step it sets that step's N_2 (a
 * fraction of the problem's N_2) and threshold tau (-16 for log10 tau means
 * none) from 16 features of the previous step:
 *   0 step / 16N, 1 log10(TW/TW_0)/20, 2 log10(TW/TW_prev)/2, 3 SL/N_2, 4 NL/SL,
 *   5 IT/NN, 6 last chosen position / IT, 7 trigonometric share, 8 log10(w_max/|Y|^2)/32,
 *   9 log2(N)/16, 10 N_2 fraction used, 11 (log10 tau used + 16)/12 (0: none),
 *   12 log10(TW/|Y|^2)/32, 13 smoothed feature 2, 14 SL reached N_2, 15 1.
 * MLP (16-64-64-2, ReLU): W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] as
 * FP32; the device rounds the weights to BF16 and accumulates in FP32.  Outputs
 * y clamped to [0, 1]: N_2 fraction = 0.75 + 0.25 y0, log10 tau = -16 + 10 y1.
 * Phase table t[8]: phase 0 while log10(TW/TW_0) > t[0], phase 1 while > t[1],
 * then phase 2; actions (N_2 fraction, log10 tau) in t[2..7].
 * Fallback (both): the plain full ordering on the first step and once TW has
 * fallen below 10^ctl_handoff TW_0 (default 1e-10).  A trace record holds the 16 features after the
 * step, the step's N_2 fraction and log10 tau, its device time (us) and TW. */
#define JKG_CTL_NPARAM 5378
#define JKG_CTL_REC 20
/* jkg_ctl_grad: the controller MLP's loss, sum_s w_s (y_s - t_s)^2 / (2 sum w), over ns
 * samples (x: ns x 16, t: ns x 2, w: ns or NULL) and its gradient, with the library's
 * forward and backward kernels; y (ns x 2) gets the outputs if not NULL. */
int jkg_ctl_grad(jkg_ctx *ctx, int ns, const float *x, const float *t, const float *w, const float *params,
                 float *grad, float *loss, float *y);
/* jkg_ctl_fit: trains params (in: initial, out: trained) with minibatch Adam on the
 * GPU; batches follow a permutation drawn from seed each epoch; loss[e] is the
 * full-set loss after epoch e. */
int jkg_ctl_fit(jkg_ctx *ctx, int ns, const float *x, const float *t, const float *w, int epochs, int batch,
                float lr, float weight_decay, uint64_t seed, float *params, float *loss);
void jkg_mixed_opts_default(jkg_mixed_opts *m);
int jkg_dsolve_mixed(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
                     double *sy, const jkg_opts *opts, const jkg_mixed_opts *m, jkg_result *res);
/* jkg_dweights_mixed: as jkg_dweights, with the weights of the mixed solver */
int jkg_dweights_mixed(jkg_ctx *ctx, int n, const double *a, const int64_t *j, int fused, uint64_t *keys);
/* jkg_dprofile_mixed: as jkg_dprofile, for the steps of the mixed solver */
int jkg_dprofile_mixed(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                       const jkg_mixed_opts *m, double *stage_ms, jkg_result *res);
/* The complex mixed solver: as jkg_dsolve_mixed for jkg_zsolve's problem (interleaved
 * complex y, yu, zz).  The Gram product is O(:,J=+1)^H O(:,J=-1), formed as two real
 * FP64 matrix-core products over the interleaved rows.  No step controller here. */
int jkg_zsolve_mixed(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, double *yu, double *zz,
                     double *sy, const jkg_opts *opts, const jkg_mixed_opts *m, jkg_result *res);
/* jkg_zweights_mixed: as jkg_zweights, with the weights of the complex mixed solver */
int jkg_zweights_mixed(jkg_ctx *ctx, int n, const double *a, const int64_t *j, int fused, uint64_t *keys);
/* jkg_zprofile / jkg_zprofile_mixed: as jkg_dprofile / jkg_dprofile_mixed, for a complex y */
int jkg_zprofile(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                 double *stage_ms, jkg_result *res);
int jkg_zprofile_mixed(jkg_ctx *ctx, int n, int n2, const double *y, const int64_t *j, int fused, int steps,
                       const jkg_mixed_opts *m, double *stage_ms, jkg_result *res);

/* ---------------------------------------------------------------------------
 * Test matrices with a prescribed hyperbolic spectrum (the role of the
 * reference's jkdatgen, src/var/old/jkdatgen/jkdatgen.F90:1-274 and xa2g.F90:
 * a factor G of A = G J G^T with given eigenvalues).  jkg_dgen writes an
 * N x N column-major g and j = sign(lam) such that g diag(j) g^T =
 * Q diag(lam) Q^T for an orthogonal Q, so the hyperbolic singular values of
 * (g, j) are sqrt|lam|: g = Q diag(sqrt|lam|) H, Q and H (H^T J H = J)
 * products of `rounds` rounds of random 2x2 rotations on random matchings
 * (rounds <= 0: 2 ceil(log2 N) + 2), hyperbolic angles uniform in
 * [-tmax, tmax] (tmax bounds the conditioning of H).  lam: N nonzero finite
 * values; the same seed gives the same g.  No counterpart of the reference's
 * extended-precision Bunch-Parlett factorisation: same contract, other bits.
 * --------------------------------------------------------------------------- */
int jkg_dgen(jkg_ctx *ctx, int n, const double *lam, int rounds, double tmax, uint64_t seed, double *g, int64_t *j);

/* ---------------------------------------------------------------------------
 * Batched solver: count independent real HSVDs of different sizes, each with
 * djk's semantics and arithmetic (YU, ZZ, SY and INFO of instance i are what
 * jkg_dsolve returns for it alone), advanced together by one persistent launch
 * per step (a workgroup runs a whole step of one instance; the workgroups walk
 * instance lists scheduled by decreasing cost).  ns[i]: N of instance i
 * (3..1024); n2s: N_2 per instance (NULL or 0: N/2; clamped by JSTEP_LEN);
 * y, yu, zz: the N_i x N_i matrices one after another (column-major); j: the
 * N_i int64 signs one after another; sy: the N_i values one after another;
 * info[i]: the instance's INFO.  The launch grid and the kernel variant (rows
 * staged per weight tile, walk entries per lane) are chosen by an autotuner
 * that times two steps of every candidate on this batch unless opts sets them.
 * --------------------------------------------------------------------------- */
typedef struct {
  int32_t fused;     /* FMAD policy: 0 unfused (the reference's default build), 1 fma */
  int32_t max_steps; /* < 0: to convergence; else at most this many steps per instance */
  int32_t grid;      /* workgroups of the persistent launch, 1..JKG_BATCH_GRID_MAX; 0: tuned */
  int32_t variant;   /* kernel variant 0..3; -1: tuned */
  double tol;        /* 0: djk's termination (bit-exact per instance); > 0: the mixed solver's
                        convergence test (see jkg_mixed_opts.tol), results agree to rounding */
  /* results */
  int32_t steps;         /* batched steps launched */
  int32_t tuned_grid;    /* grid used */
  int32_t tuned_variant; /* variant used */
  int32_t reserved;
  double seconds;      /* device time of the step loop */
  double tune_seconds; /* host time spent tuning */
} jkg_batch_opts;
#define JKG_BATCH_GRID_MAX 4096
void jkg_batch_opts_default(jkg_batch_opts *o);
int jkg_dsolve_batched(jkg_ctx *ctx, int count, const int32_t *ns, const int32_t *n2s, const double *y,
                       const int64_t *j, double *yu, double *zz, double *sy, int32_t *info, jkg_batch_opts *opts);

/* index of pair (p,q), 1 <= p < q <= n, in the tie-break order of AW_CMP */
uint32_t jkg_pair_index(int n, int p, int q);
/* the weight a sort key stands for (NaN for the key of a dropped pair) */
double jkg_key_weight(uint64_t key);

#ifdef __cplusplus
}
#endif
#endif
