/*
 * jk_oracle.h -- CPU restatement of the J-Kogbetliantz dynamic-ordering step
 * (venovako/JKogb, src/D, src/Z, src/shared).  TEST INFRASTRUCTURE ONLY: the
 * checker that the GPU path is compared against.  Nothing in the product
 * (jkogb_amd/, host/) links or calls this library.
 *
 * Two copies of every floating-point routine are exported:
 *   orc_*   FMAD(a,b,c) = a*b+c with two roundings (the reference's default
 *           build: no -march, so nothing is contracted);
 *   orcf_*  FMAD(a,b,c) = fma(a,b,c) (the reference built with
 *           -march=x86-64-v3 -ffp-contract=fast).
 * Integer / ordering routines exist once (orc_*).
 *
 * Conventions follow the Fortran reference exactly: matrices are column-major
 * with a leading dimension, pair indices P < Q are 1-based, 2x2 blocks are
 * column-major arrays of 4 (A11, A21, A12, A22).
 */
#ifndef JK_ORACLE_H
#define JK_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The AW candidate record, src/shared/atypes.F90:18-24 (24 bytes, no padding). */
typedef struct { double w; int32_t p, q, b, i; } jk_aw;

/* COMPLEX(KIND=8) layout: real part then imaginary part. */
typedef struct { double re, im; } jk_z;

/* One line of the per-step CSV trace (src/D/dstep.F90:410-440):
 * STEP, OLDLEN, NEWLEN, NORMF2D, TRIG, HYP (the two timing columns are not
 * recorded: they are wall-clock values). */
typedef struct { int32_t step, oldlen, newlen, trig, hyp, pad; double tw; } jk_steprec;

/* ---------------- precision-independent routines ---------------- */
double  orc_hypot(double x, double y);                        /* = glibc 2.35 hypot */
int     orc_jstep_len(int n, int n2);                          /* jstep.F90:124-133 */
int     orc_tru(int id_tru, int n, const int32_t *J, int nn,
                int32_t *P, int32_t *Q, int32_t PN[2]);         /* jstep.F90:9-120 */
int     orc_aw_cmp(const jk_aw *a, const jk_aw *b);            /* atypes.F90:85-120 */
int64_t orc_srt(int nn, int nm, jk_aw *dz);                    /* atypes.F90:218-355 */
int64_t orc_ncp(int id_ncp, int n, int nn, int nm, jk_aw *dz,
                int n2, int32_t *sl, int32_t *step);            /* atypes.F90:359-592 */
void    orc_jztj(int n, double *Z, int ldz, const int32_t *J,
                 int nn, const int32_t *P, const int32_t *Q);   /* dstep.F90:447-468 */
void    orc_jzhj(int n, jk_z *Z, int ldz, const int32_t *J,
                 int nn, const int32_t *P, const int32_t *Q);   /* zstep.F90:451-478 */

/* ---------------- floating-point routines, one copy per FMA policy ---------------- */
#define JK_ORACLE_DECL(PFX)                                                              \
double  PFX##dasum2(double x1, double x2);                                               \
double  PFX##dasum4(double x1, double x2, double x3, double x4);                          \
void    PFX##dhsvd2(double A[4], const int32_t J[2], double U[4], double Z[4],           \
                    double th1, int32_t *info);                                          \
void    PFX##zhsvd2(jk_z A[4], const int32_t J[2], jk_z U[4], jk_z Z[4],                 \
                    double th1, int32_t *info);                                          \
void    PFX##dba(const double B[4], int n, double *x, double *y, int lda);               \
void    PFX##dab(const double B[4], int m, double *x, double *y);                        \
void    PFX##zba(const jk_z B[4], int n, jk_z *x, jk_z *y, int lda);                     \
void    PFX##zab(const jk_z B[4], int m, jk_z *x, jk_z *y);                              \
double  PFX##abodndf2(const double B[4], int m, const double *x, const double *y,        \
                      double aqp, double apq, int p, int q);                             \
double  PFX##abodnzf2(const jk_z B[4], int m, const jk_z *x, const jk_z *y,              \
                      jk_z aqp, jk_z apq, int p, int q);                                 \
double  PFX##dmagf2t(int n, int p, int q, const double *A, int lda, const int32_t *J);   \
double  PFX##dmagf2h(int n, int p, int q, const double *A, int lda, const int32_t *J);   \
double  PFX##zmagf2t(int n, int p, int q, const jk_z *A, int lda, const int32_t *J);     \
double  PFX##zmagf2h(int n, int p, int q, const jk_z *A, int lda, const int32_t *J);     \
int64_t PFX##dstep_build(int s, int n, const double *A, int lda, const int32_t *J,       \
                         int nn, const int32_t *P, const int32_t *Q, int id_ncp, int nm, \
                         jk_aw *dz, int tt, int n2, int32_t *sl, int32_t *step);         \
int64_t PFX##zstep_build(int s, int n, const jk_z *A, int lda, const int32_t *J,         \
                         int nn, const int32_t *P, const int32_t *Q, int id_ncp, int nm, \
                         jk_aw *dz, int tt, int n2, int32_t *sl, int32_t *step);         \
int32_t PFX##dstep_transf(int n, double *U, int ldu, double *A, int lda, double *Z,      \
                          int ldz, const int32_t *J, double *sigma, int nn, jk_aw *dz,   \
                          int sl, const int32_t *step, double *W);                       \
int32_t PFX##zstep_transf(int n, jk_z *U, int ldu, jk_z *A, int lda, jk_z *Z,            \
                          int ldz, const int32_t *J, double *sigma, int nn, jk_aw *dz,   \
                          int sl, const int32_t *step, jk_z *W);                         \
int32_t PFX##dstep_loop(int n, double *U, int ldu, double *A, int lda, double *Z,        \
                        int ldz, const int32_t *J, double *sigma, int nn,                \
                        const int32_t *P, const int32_t *Q, int id_ncp, int nm,          \
                        jk_aw *dz, int n2, int32_t *step, int tt, double *W,             \
                        jk_steprec *trace, int max_trace, int max_steps);                \
int32_t PFX##zstep_loop(int n, jk_z *U, int ldu, jk_z *A, int lda, jk_z *Z,              \
                        int ldz, const int32_t *J, double *sigma, int nn,                \
                        const int32_t *P, const int32_t *Q, int id_ncp, int nm,          \
                        jk_aw *dz, int n2, int32_t *step, int tt, jk_z *W,               \
                        jk_steprec *trace, int max_trace, int max_steps);                \
int32_t PFX##dcyc_loop(int n, double *U, int ldu, double *A, int lda, double *Z, int ldz, \
                       const int32_t *J, double *sigma, double *W, jk_steprec *trace,    \
                       int max_trace, int max_steps, int order);

JK_ORACLE_DECL(orc_)
JK_ORACLE_DECL(orcf_)

#ifdef __cplusplus
}
#endif
#endif
