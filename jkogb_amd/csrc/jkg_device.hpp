// jkg_device.hpp -- device-side scalar helpers shared by the gfx950 kernels.
//
// Every kernel of the FP64 path is compiled with -ffp-contract=off, so a*b+c
// stays two roundings unless a helper asks for fma explicitly; this is how
// the reference's default build evaluates its FMAD macro
// (src/shared/utils.F90:52-54, no -march, nothing contracted).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace jkg {

// FMAD(a,b,c) of the reference: unfused (two roundings) or fused.
template <bool FUSED>
__device__ __forceinline__ double fmad(double a, double b, double c) {
  if constexpr (FUSED) return __builtin_fma(a, b, c);
  else return a * b + c;
}
template <bool FUSED>
__device__ __forceinline__ float fmadf(float a, float b, float c) {
  if constexpr (FUSED) return __builtin_fmaf(a, b, c);
  else return a * b + c;
}

constexpr double TWOF = 1.3407807929942596e+154;   // SQRT(HUGE(0D0)), params.F90:31
constexpr double MINF = 4.9406564584124654e-324;   // smallest subnormal, params.F90:32
constexpr double TH1FIX = 0.75;                    // params.F90:34
constexpr int EH = 1024;                           // EXPONENT(HUGE(0D0))
constexpr int ET = -1021;                          // EXPONENT(TINY(0D0))

// Fortran EXPONENT(x): x = f * 2**e, 0.5 <= |f| < 1; 0 for x == 0.
__device__ __forceinline__ int f_exponent(double x) {
  return (x == 0.0) ? 0 : __builtin_amdgcn_frexp_exp(x);
}
// Fortran SCALE(x, i) = x * 2**i, one rounding.
__device__ __forceinline__ double f_scale(double x, int i) { return __builtin_ldexp(x, i); }
__device__ __forceinline__ bool sign_neg(double x) { return __builtin_signbit(x) != 0; }
__device__ __forceinline__ double f_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double f_max(double a, double b) { return (b > a) ? b : a; }
// The reference binds MAXD/MIND to C fmax/fmin (NaN-ignoring).
__device__ __forceinline__ double maxd(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ double mind(double a, double b) { return __builtin_fmin(a, b); }

// QUIET_NAN(payload), src/shared/utils.F90:144-149
__device__ __forceinline__ double quiet_nan(int payload) {
  return __longlong_as_double((long long)payload | 0x7FF8000000000000LL);
}

// hypot as the reference gets it (HYPOT -> libm hypot, glibc 2.35): Borges'
// algorithm, "An Improved Algorithm for hypot(a,b)", ACM TOMS 47(1), 2021 --
// a naive sqrt, one correction step with the no-FMA formulas, and
// power-of-two rescaling at the extremes.  All operations are IEEE-rounded
// (sqrt and / are correctly rounded on gfx950; -ffp-contract=off).
__device__ __forceinline__ double hypot_kernel(double ax, double ay) {
  double h = __builtin_sqrt(ax * ax + ay * ay);
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
  return h - (t1 + t2) / (2.0 * h);
}
__device__ __forceinline__ double hypot_ref(double x, double y) {
  if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) {
    if (__builtin_isinf(x) || __builtin_isinf(y)) return __builtin_inf();
    return x + y;
  }
  x = __builtin_fabs(x);
  y = __builtin_fabs(y);
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

// ---------------------------------------------------------------------------
// Sort keys.  The reference orders candidates by W descending, then band
// B = Q-P descending, then P descending (AW_CMP, src/shared/atypes.F90:85-120),
// with -0 == +0 and NaN weights dropped.  The pair index k enumerates the
// pairs in exactly that tie-break order (band-major, bands and rows
// descending), so an ascending stable sort of key(W) over k = 0,1,2,...
// reproduces the reference's permutation.  key(W) is the IEEE bit pattern
// mapped to an unsigned integer that grows as W decreases; NaN -> ~0.
// ---------------------------------------------------------------------------
constexpr uint64_t KEY_NAN = ~0ull;
__device__ __forceinline__ uint64_t weight_key(double w) {
  if (!(w == w)) return KEY_NAN;
  if (w == 0.0) w = 0.0;  // -0 -> +0
  uint64_t u = (uint64_t)__double_as_longlong(w);
  uint64_t ord = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);  // ascending in w
  return ~ord;  // descending in w
}
__device__ __forceinline__ double key_weight(uint64_t key) {
  uint64_t ord = ~key;
  uint64_t u = (ord & 0x8000000000000000ull) ? (ord & 0x7FFFFFFFFFFFFFFFull) : ~ord;
  return __longlong_as_double((long long)u);
}

// Pair index k <-> (p, q), 1-based p < q, N x N.  m = N - (q - p) runs over
// 1..N-1 (band N-1 first); within a band p runs from m down to 1.
//   k = m(m-1)/2 + (m - p)
__host__ __device__ __forceinline__ uint32_t pair_index(int n, int p, int q) {
  uint32_t m = (uint32_t)(n - (q - p));
  return m * (m - 1) / 2 + (m - (uint32_t)p);
}
__device__ __forceinline__ void pair_of(int n, uint32_t k, int &p, int &q) {
  // largest m with m(m-1)/2 <= k
  double fk = (double)k;
  uint32_t m = (uint32_t)((1.0 + __builtin_sqrt(1.0 + 8.0 * fk)) * 0.5);
  while ((uint64_t)m * (m - 1) / 2 > k) --m;
  while ((uint64_t)(m + 1) * m / 2 <= k) ++m;
  uint32_t r = k - m * (m - 1) / 2;
  p = (int)(m - r);
  q = p + n - (int)m;
}

// A pair as the sort carries it: 0-based p in the low, q in the high 16 bits.
__host__ __device__ __forceinline__ uint32_t pack_pq(int p0, int q0) { return (uint32_t)p0 | ((uint32_t)q0 << 16); }
__host__ __device__ __forceinline__ int pq_p(uint32_t v) { return (int)(v & 0xFFFFu); }
__host__ __device__ __forceinline__ int pq_q(uint32_t v) { return (int)(v >> 16); }
// Key of a packed pair: the sort leaves the keys in pair-index order and sorts the pairs.
__device__ __forceinline__ uint64_t pair_key(const uint64_t *keys, int n, uint32_t v) {
  return keys[pair_index(n, pq_p(v) + 1, pq_q(v) + 1)];
}

// The mixed solver's convergence test for one pair: its off-diagonal entries
// are negligible against the diagonal,
//   max(|a_pq|, |a_qp|) <= tol * sqrt(|a_pp|) * sqrt(|a_qq|)
// (the relative criterion of one-sided Jacobi methods; such entries move the
// singular values by O(tol) relatively).
__device__ __forceinline__ bool negligible(double aqp, double apq, double app, double aqq, double tol) {
  const double m = f_max(__builtin_fabs(aqp), __builtin_fabs(apq));
  return m <= tol * __builtin_sqrt(__builtin_fabs(app)) * __builtin_sqrt(__builtin_fabs(aqq));
}

// Complex FP64 value in the reference's COMPLEX(KIND=8) layout.
struct dz {
  double re, im;
};
__device__ __forceinline__ dz zmk(double r, double i) { return dz{r, i}; }
__device__ __forceinline__ dz zconj(dz a) { return dz{a.re, -a.im}; }
__device__ __forceinline__ bool zeq0(dz a) { return a.re == 0.0 && a.im == 0.0; }
__device__ __forceinline__ bool zne1(dz a) { return a.re != 1.0 || a.im != 0.0; }
__device__ __forceinline__ double absz(dz a) { return hypot_ref(a.re, a.im); }  // ABS(complex) = cabs

// Complex FP32 value of the mixed solver's warm start (A, U, Z held in FP32;
// every 2x2 parameter and weight is formed in FP64 from the promoted entries).
struct fz {
  float re, im;
};
__device__ __forceinline__ dz ldz(const double &a) { return dz{a, 0.0}; }
__device__ __forceinline__ dz ldz(const float &a) { return dz{(double)a, 0.0}; }
__device__ __forceinline__ dz ldz(const dz &a) { return a; }
__device__ __forceinline__ dz ldz(const fz &a) { return dz{(double)a.re, (double)a.im}; }
template <class E>
__device__ __forceinline__ E emk(double re, double im);
template <>
__device__ __forceinline__ dz emk<dz>(double re, double im) { return dz{re, im}; }
template <>
__device__ __forceinline__ fz emk<fz>(double re, double im) { return fz{(float)re, (float)im}; }

// Sort keys cut to their top 32 bits (the warm start's shorter keys): the low
// four radix digits are then zero in every key and the sort skips their passes;
// ties go by pair index (the sort is stable).  A cut NaN key is KEY_NAN32,
// above every finite key in either width.
constexpr uint64_t KEY_NAN32 = 0xFFFFFFFF00000000ull;
__device__ __forceinline__ bool key_nan(uint64_t k) { return k >= KEY_NAN32; }
__device__ __forceinline__ uint64_t key_cut(uint64_t k, bool cut) { return cut ? (k & KEY_NAN32) : k; }

}  // namespace jkg
