/*
 * jk_oracle.c -- CPU restatement of venovako/JKogb's dynamic-ordering
 * J-Kogbetliantz step.  TEST INFRASTRUCTURE ONLY (the checker): it is linked
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * never by the product library.
 *
 * This file holds the precision-independent parts (pair enumeration, the AW
 * order, sorting, greedy pivot selection, J Z^T J) and instantiates the
 * floating-point core jk_core.inc twice, once per FMA policy.  Every function
 * cites the reference routine it restates (paths relative to the reference's
 * src/ directory).  The build must not contract a*b+c into fma: oracle/Makefile
 * compiles with -ffp-contract=off.
 */
#include "jk_oracle.h"
#include <math.h>
#include <float.h>
#include <limits.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------------
 * hypot: the reference calls the Fortran HYPOT intrinsic, which AMD flang
 * lowers to libm's hypot (glibc 2.35 on this image).  glibc 2.35's hypot is not
 * correctly rounded; it is Borges' algorithm ("An Improved Algorithm for
 * hypot(a,b)", ACM TOMS 47(1), 2021, arXiv:1904.09481): a naive sqrt followed
 * by one correction step, with power-of-two rescaling at the extremes and the
 * no-FMA correction formulas.  Restated here from the paper; bit-identical to
 * glibc 2.35 on 1.4e8 random inputs spanning the whole exponent range
 * (tests/test_oracle.py re-checks a sample against libm at run time).
 * ------------------------------------------------------------------------- */
static double hypot_kernel(double ax, double ay)
{
    /* ax >= ay > 0, both scaled so that no square over/underflows */
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        double d = h - ay;
        t1 = ax * (2.0 * d - ax);
        t2 = (d - 2.0 * (ax - ay)) * d;
    } else {
        double d = h - ax;
        t1 = 2.0 * d * (ax - 2.0 * ay);
        t2 = (4.0 * d - ay) * ay + d * d;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h;
}

double orc_hypot(double x, double y)
{
    if (!isfinite(x) || !isfinite(y)) {
        if (isinf(x) || isinf(y)) return INFINITY;
        return x + y;
    }
    x = fabs(x);
    y = fabs(y);
    double ax = (x < y) ? y : x;
    double ay = (x < y) ? x : y;
    if (ax > 0x1p+511) {
        if (ay <= ax * 0x1p-54) return ax + ay;
        return hypot_kernel(ax * 0x1p-600, ay * 0x1p-600) / 0x1p-600;
    }
    if (ay < 0x1p-511) {
        if (ax >= ay / 0x1p-54) return ax + ay;
        return hypot_kernel(ax / 0x1p-600, ay / 0x1p-600) * 0x1p-600;
    }
    if (ay <= ax * 0x1p-54) return ax + ay;
    return hypot_kernel(ax, ay);
}

/* JSTEP_LEN, src/shared/jstep.F90:124-133 */
int orc_jstep_len(int n, int n2)
{
    if (n2 == 0) return (n / 2 > 0) ? n / 2 : 0;
    int m = (n2 > 0) ? n2 : 0;
    return (m < n / 2) ? m : n / 2;
}

/* TRU1 (column-cyclic, id 1) and TRU2 (row-cyclic, id 2),
 * src/shared/jstep.F90:9-62 and :67-120.  Trigonometric pairs (J_p == J_q)
 * fill P/Q from the front, hyperbolic ones from the back.  Returns TT (the
 * number of trigonometric pairs) or -2 on a J entry outside {-1,+1}. */
int orc_tru(int id_tru, int n, const int32_t *J, int nn, int32_t *P, int32_t *Q, int32_t PN[2])
{
    PN[0] = PN[1] = 0;
    if (n < 0) return -1;
    if (nn < 0) return -3;
    if (n <= 1) return 0;
    for (int it = 0; it < n; ++it) {
        if (J[it] == -1) PN[1]++;
        else if (J[it] == 1) PN[0]++;
        else return -2;
    }
    int it = 1, ih = nn;
    if (id_tru == 2) {
        for (int ip = 1; ip <= n - 1; ++ip)
            for (int iq = ip + 1; iq <= n; ++iq) {
                if (it > ih) return 0; /* the reference returns with INFO still 0 */
                if (J[ip - 1] == J[iq - 1]) { P[it - 1] = ip; Q[it - 1] = iq; ++it; }
                else { P[ih - 1] = ip; Q[ih - 1] = iq; --ih; }
            }
    } else {
        for (int iq = 2; iq <= n; ++iq)
            for (int ip = 1; ip <= iq - 1; ++ip) {
                if (it > ih) return 0;
                if (J[ip - 1] == J[iq - 1]) { P[it - 1] = ip; Q[it - 1] = iq; ++it; }
                else { P[ih - 1] = ip; Q[ih - 1] = iq; --ih; }
            }
    }
    return it - 1;
}

/* AW_CMP, src/shared/atypes.F90:85-120: W descending, then band B descending,
 * then P descending, then Q descending; NaN weights last; -0 == +0. */
int orc_aw_cmp(const jk_aw *a, const jk_aw *b)
{
    if (a->w < b->w) return 1;
    if (a->w > b->w) return -1;
    if (a->w == b->w) {
        if (a->b < b->b) return 2;
        if (a->b > b->b) return -2;
        if (a->p < b->p) return 3;
        if (a->p > b->p) return -3;
        if (a->q < b->q) return 4;
        if (a->q > b->q) return -4;
        return 0;
    }
    return (b->w == b->w) ? 5 : -5;
}

static int aw_qsort_cmp(const void *x, const void *y)
{
    int c = orc_aw_cmp((const jk_aw *)x, (const jk_aw *)y);
    return (c < 0) ? -1 : (c > 0) ? 1 : 0;
}

/* AW_SRT1/AW_SRT2, src/shared/atypes.F90:218-355.  Both produce the unique
 * permutation that orders the records by AW_CMP (a total order over distinct
 * pairs), so any correct sort gives the same result. */
int64_t orc_srt(int nn, int nm, jk_aw *dz)
{
    if (nn < 0) return -1;
    if (nm < nn) return -2;
    if (nn == 0) return 0;
    qsort(dz, (size_t)nn, sizeof(jk_aw), aw_qsort_cmp);
    return 0;
}

/* AW_NCP2, src/shared/atypes.F90:430-484 (AW_NCP1, :359-426, selects the same
 * pivots on a NaN-free list): greedy walk over the sorted list, taking a pair
 * when neither of its indices is used yet, until N_2 pairs are taken.  STEP
 * holds 1-based list positions. */
static int64_t ncp_greedy(int n, int nn, const jk_aw *dz, int n2, int32_t *sl, int32_t *step)
{
    *sl = 0;
    if (n == 0 || nn == 0 || n2 == 0) return 0;
    unsigned char *r = (unsigned char *)calloc((size_t)n + 1, 1);
    int j = 1;
    for (int i = 1; i <= n2; ++i) {
        step[i - 1] = j;
        r[dz[j - 1].p] = 1;
        r[dz[j - 1].q] = 1;
        ++j;
        while (j <= nn) {
            const jk_aw *x = &dz[j - 1];
            if (x->w == x->w && !r[x->p] && !r[x->q]) break;
            ++j;
        }
        *sl = i;
        if (j > nn) break;
    }
    free(r);
    return 0;
}

/* AW_NCP3 + AW_NCPT, src/shared/atypes.F90:488-592: run the greedy walk from
 * every start offset and keep the selection of the largest total weight (the
 * positive weights summed last-to-first, then the rest first-to-last); ties go
 * to the smallest start. */
static int64_t ncp_best(int n, int nn, const jk_aw *dz, int n2, int32_t *sl, int32_t *step)
{
    *sl = 0;
    if (n == 0 || nn == 0 || n2 == 0) return 0;
    int32_t *stp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n2);
    double w = NAN;
    int k = nn + 1;
    for (int i = 1; i <= nn; ++i) {
        int jo = i - 1, mysl = 0;
        ncp_greedy(n, nn - jo, dz + jo, n2, &mysl, stp);
        if (mysl < 1) continue;
        int l = 0;
        for (int t = mysl; t >= 1; --t)
            if (dz[jo + stp[t - 1] - 1].w > 0.0) { l = t; break; }
        double wp = 0.0, wn = 0.0;
        for (int t = l; t >= 1; --t) wp = wp + dz[jo + stp[t - 1] - 1].w;
        for (int t = l + 1; t <= mysl; ++t) wn = wn + dz[jo + stp[t - 1] - 1].w;
        double myw = wp + wn;
        int start = jo + 1;
        if (!(myw <= w) || (myw == w && start < k)) {
            *sl = mysl;
            for (int t = 0; t < mysl; ++t) step[t] = stp[t] + jo;
            w = myw;
            k = start;
        }
    }
    free(stp);
    return 0;
}

int64_t orc_ncp(int id_ncp, int n, int nn, int nm, jk_aw *dz, int n2, int32_t *sl, int32_t *step)
{
    *sl = 0;
    if (n < 0) return -1;
    if (nn < 0) return -2;
    if (nm < nn) return -3;
    if (n2 < 0) return -5;
    if (id_ncp == 3) return ncp_best(n, nn, dz, n2, sl, step);
    return ncp_greedy(n, nn, dz, n2, sl, step);
}

/* JZTJ, src/D/dstep.F90:447-468: Z <- J Z^T J over the pair list. */
void orc_jztj(int n, double *Z, int ldz, const int32_t *J, int nn, const int32_t *P, const int32_t *Q)
{
    (void)n;
    for (int i = 0; i < nn; ++i) {
        int ip = P[i], iq = Q[i];
        double *zpq = &Z[(size_t)(iq - 1) * ldz + (ip - 1)];
        double *zqp = &Z[(size_t)(ip - 1) * ldz + (iq - 1)];
        double t = *zpq;
        *zpq = *zqp;
        *zqp = t;
        if (J[ip - 1] != J[iq - 1]) { *zpq = -*zpq; *zqp = -*zqp; }
    }
}

/* JZHJ, src/Z/zstep.F90:451-478: Z <- J Z^H J (the diagonal conjugated too). */
void orc_jzhj(int n, jk_z *Z, int ldz, const int32_t *J, int nn, const int32_t *P, const int32_t *Q)
{
    for (int i = 0; i < nn; ++i) {
        int ip = P[i], iq = Q[i];
        jk_z *zpq = &Z[(size_t)(iq - 1) * ldz + (ip - 1)];
        jk_z *zqp = &Z[(size_t)(ip - 1) * ldz + (iq - 1)];
        jk_z t = *zpq;
        zpq->re = zqp->re; zpq->im = -zqp->im;
        zqp->re = t.re;    zqp->im = -t.im;
        if (J[ip - 1] != J[iq - 1]) {
            zpq->re = -zpq->re; zpq->im = -zpq->im;
            zqp->re = -zqp->re; zqp->im = -zqp->im;
        }
    }
    for (int i = 0; i < n; ++i) Z[(size_t)i * ldz + i].im = -Z[(size_t)i * ldz + i].im;
}

/* ---- the floating-point core, once per FMA policy ---- */
#define PFX(x) orc_##x
#define FMAD(a, b, c) ((a) * (b) + (c))
#include "jk_core.inc"
#undef PFX
#undef FMAD

#define PFX(x) orcf_##x
#define FMAD(a, b, c) fma((a), (b), (c))
#include "jk_core.inc"
#undef PFX
#undef FMAD
