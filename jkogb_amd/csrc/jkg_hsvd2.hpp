// jkg_hsvd2.hpp -- the 2x2 hyperbolic SVD of the J-Kogbetliantz step on the
// device, FP64, one lane per 2x2 problem.  The branch structure, the
// power-of-two prescaling, the tangent formulas and the INFO bits follow the
// reference's DHSVD2/ZHSVD2 (src/D/dtransf.F90:352-683,
// src/shared/ktransf.F90:15-217, src/Z/ztransf.F90:641-1009, NDEBUG build)
// operation by operation, so that the results are bit-identical to it.
//
// 2x2 blocks are column-major arrays of 4: [0]=X11 [1]=X21 [2]=X12 [3]=X22.
// INFO bits: 1 = block was not diagonal, 2 = U != I, 4 = Z != I,
// 8 = |tanh| clamped to TH1FIX (block left partially reduced); < 0 error.
#pragma once
#include "jkg_device.hpp"

namespace jkg {

// ---- real --------------------------------------------------------------
template <bool F>
__device__ __forceinline__ void c1a(const double B[4], double &a0, double &a1) {
  double c0 = fmad<F>(B[2], a1, a0) / B[0];
  double c1 = fmad<F>(B[1], a0, a1) / B[3];
  a0 = c0;
  a1 = c1;
}
template <bool F>
__device__ __forceinline__ void c2a(const double B[4], double A[4]) {
  double C[4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    C[0 + 2 * j] = fmad<F>(B[2], A[1 + 2 * j], A[0 + 2 * j]) / B[0];
    C[1 + 2 * j] = fmad<F>(B[1], A[0 + 2 * j], A[1 + 2 * j]) / B[3];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) A[k] = C[k];
}
template <bool F>
__device__ __forceinline__ void a2c(double A[4], const double B[4]) {
  double C[4];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    C[i + 0] = fmad<F>(A[i + 2], B[1], A[i + 0]) / B[0];
    C[i + 2] = fmad<F>(A[i + 0], B[2], A[i + 2]) / B[3];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) A[k] = C[k];
}

// the shared tail of the three K-steps: downscaling of the new diagonal
__device__ __forceinline__ void kscale(double &x, int s, int &e) {
  int d = f_exponent(x) - ET;
  if (s > d) {
    x = f_scale(x, -d);
    e = s - d;
  } else {
    x = f_scale(x, -s);
    e = 0;
  }
}

template <bool F>
__device__ void khsvd2t(const double A[4], double G[2], double U[4], double Z[4], int s) {
  double x = maxd(A[2] / A[0], 0.0);
  double y = maxd(A[3] / A[0], 0.0);
  double t2 = f_min(maxd((f_scale(f_min(x, y), 1) * f_max(x, y)) / fmad<F>((x - y), (x + y), 1.0), 0.0), TWOF);
  double tu = t2 / (1.0 + __builtin_sqrt(fmad<F>(t2, t2, 1.0)));
  double cu = __builtin_sqrt(fmad<F>(tu, tu, 1.0));
  double tz = fmad<F>(y, tu, x);
  double cz = __builtin_sqrt(fmad<F>(tz, tz, 1.0));
  int e;
  y = cz / cu;
  x = y;
  kscale(x, s, e);
  G[0] = f_scale(A[0], -e) * x;
  G[1] = f_scale(A[3] / y, -s);
  U[0] = cu; U[1] = -tu; U[2] = tu; U[3] = cu;
  Z[0] = cz; Z[1] = tz; Z[2] = -tz; Z[3] = cz;
}

template <bool F>
__device__ void khsvd2u(const double A[4], double G[2], double U[4], double Z[4], int s, double th1, int &info) {
  double x = maxd(A[2] / A[0], 0.0);
  double y = maxd(A[3] / A[0], 0.0);
  double t2 = -f_min(maxd((f_scale(f_min(x, y), 1) * f_max(x, y)) / fmad<F>((y - x), (y + x), 1.0), 0.0), TWOF);
  double tu = t2 / (1.0 + __builtin_sqrt(fmad<F>(t2, t2, 1.0)));
  double tz = -fmad<F>(y, tu, x);
  t2 = __builtin_fabs(tz);
  if (t2 >= 1.0) { info = -10; return; }
  if (t2 >= th1) { tz = __builtin_copysign(TH1FIX, tz); info += 8; }
  double cu = __builtin_sqrt(fmad<F>(tu, tu, 1.0));
  double cz = __builtin_sqrt(fmad<F>(-tz, tz, 1.0));
  if (t2 < th1) {
    int e;
    y = cu / cz;
    x = y;
    kscale(x, s, e);
    G[0] = f_scale(A[0] / y, -s);
    G[1] = f_scale(A[3], -e) * x;
  }
  U[0] = cu; U[1] = -tu; U[2] = tu; U[3] = cu;
  Z[0] = cz; Z[1] = tz; Z[2] = tz; Z[3] = cz;
}

template <bool F>
__device__ void khsvd2l(const double A[4], double G[2], double U[4], double Z[4], int s, double th1, int &info) {
  double x = maxd(A[1] / A[3], 0.0);
  double y = maxd(A[0] / A[3], 0.0);
  double t2 = f_min(maxd((f_scale(f_min(x, y), 1) * f_max(x, y)) / fmad<F>((y - x), (y + x), 1.0), 0.0), TWOF);
  double tu = t2 / (1.0 + __builtin_sqrt(fmad<F>(t2, t2, 1.0)));
  double tz = fmad<F>(y, tu, -x);
  t2 = __builtin_fabs(tz);
  if (t2 >= 1.0) { info = -11; return; }
  if (t2 >= th1) { tz = __builtin_copysign(TH1FIX, tz); info += 8; }
  double cu = __builtin_sqrt(fmad<F>(tu, tu, 1.0));
  double cz = __builtin_sqrt(fmad<F>(-tz, tz, 1.0));
  if (t2 < th1) {
    int e;
    y = cu / cz;
    x = y;
    kscale(x, s, e);
    G[0] = f_scale(A[0], -e) * x;
    G[1] = f_scale(A[3] / y, -s);
  }
  U[0] = cu; U[1] = -tu; U[2] = tu; U[3] = cu;
  Z[0] = cz; Z[1] = tz; Z[2] = tz; Z[3] = cz;
}

template <bool F>
__device__ __forceinline__ void khsvd2(const double A[4], const int J[2], double G[2], double U[4], double Z[4], int s,
                                       double th1, int &info) {
  if (A[1] == 0.0) {
    if (J[0] == J[1]) {
      if (info == 1) khsvd2t<F>(A, G, U, Z, s);
    } else {
      if (info == 1) khsvd2u<F>(A, G, U, Z, s, th1, info);
    }
  } else if (A[2] == 0.0) {
    if (info == 1) khsvd2l<F>(A, G, U, Z, s, th1, info);
  }
}

__device__ __forceinline__ void swap2(double &a, double &b) {
  double t = a;
  a = b;
  b = t;
}

template <bool F>
__device__ void dhsvd2t(double A[4], const int J[2], double U[4], double Z[4], int &info) {
  double c, s, r;
  bool p, dg = true;
  if (A[1] == 0.0) c = __builtin_fabs(A[0]);
  else if (A[0] == 0.0) { c = __builtin_fabs(A[1]); dg = false; }
  else { c = hypot_ref(A[0], A[1]); dg = false; }
  if (A[2] == 0.0) s = __builtin_fabs(A[3]);
  else if (A[3] == 0.0) { s = __builtin_fabs(A[2]); dg = false; }
  else { s = hypot_ref(A[2], A[3]); dg = false; }

  if (c < s) {
    r = s;
    swap2(A[0], A[2]); swap2(A[1], A[3]);
    if (J[0] == J[1]) { swap2(Z[0], Z[2]); swap2(Z[1], Z[3]); }
    p = true;
  } else {
    r = c;
    p = false;
  }
  if (sign_neg(A[0])) { U[0] = -U[0]; U[2] = -U[2]; A[2] = -A[2]; A[0] = -A[0]; }
  if (sign_neg(A[1])) { U[1] = -U[1]; U[3] = -U[3]; A[3] = -A[3]; A[1] = -A[1]; }
  if (A[0] < A[1]) { swap2(U[0], U[1]); swap2(U[2], U[3]); swap2(A[0], A[1]); swap2(A[2], A[3]); }
  if (!dg) {
    s = A[1] / A[0];
    c = __builtin_sqrt(fmad<F>(s, s, 1.0));
    double q[4] = {c, -s, s, c};
    c2a<F>(q, U);
    c1a<F>(q, A[2], A[3]);
  }
  A[0] = r;
  A[1] = 0.0;
  if (J[0] != J[1] && p) {
    swap2(A[0], A[2]); swap2(A[1], A[3]);   // columns
    swap2(U[0], U[1]); swap2(U[2], U[3]);   // rows of U
    swap2(A[0], A[1]); swap2(A[2], A[3]);   // rows of A
    if (sign_neg(A[1])) { Z[0] = -Z[0]; Z[1] = -Z[1]; A[0] = -A[0]; A[1] = -A[1]; }
    if (sign_neg(A[0])) { U[0] = -U[0]; U[2] = -U[2]; A[0] = -A[0]; }
  } else {
    if (sign_neg(A[2])) { Z[2] = -Z[2]; Z[3] = -Z[3]; A[3] = -A[3]; A[2] = -A[2]; }
    if (sign_neg(A[3])) { U[1] = -U[1]; U[3] = -U[3]; A[3] = -A[3]; }
  }
  info = dg ? 0 : 1;
}

__device__ __forceinline__ void dhsvd2s(double A[4], const int J[2], double U[4], double Z[4], int &info) {
  int t = J[0] + J[1];
  if ((t == 2 && A[0] < A[3]) || (t == -2 && A[3] < A[0])) {
    swap2(U[0], U[1]); swap2(U[2], U[3]);
    swap2(A[0], A[3]);
    swap2(Z[0], Z[2]); swap2(Z[1], Z[3]);
  }
  if ((info & 8) == 0) { A[1] = 0.0; A[2] = 0.0; }
  if (U[0] != 1.0 || U[1] != 0.0 || U[2] != 0.0 || U[3] != 1.0) info += 2;
  if (Z[0] != 1.0 || Z[1] != 0.0 || Z[2] != 0.0 || Z[3] != 1.0) info += 4;
}

__device__ __forceinline__ void dscalea(double A[4], int &s) {
  int sc = 0x7FFFFFFF;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double aa = __builtin_fabs(A[k]);
    if (!(aa <= 1.7976931348623157e308)) { s = -((k >> 1) + 1) * 2 - ((k & 1) + 1); return; }
    int v = EH - f_exponent(f_max(aa, MINF)) - 2;
    sc = (v < sc) ? v : sc;
  }
  if (sc != 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) A[k] = f_scale(A[k], sc);
  }
  s = sc;
}

// DHSVD2: A (in) the 2x2 block, (out) the new diagonal in A[0], A[3];
// U, Z the 2x2 factors; returns INFO.
template <bool F>
__device__ int dhsvd2(double A[4], const int J[2], double U[4], double Z[4], double th1) {
  double G[2] = {0.0, 0.0}, B[4] = {0.0, 0.0, 0.0, 0.0}, C[4] = {0.0, 0.0, 0.0, 0.0};
  int s, info;
  dscalea(A, s);
  if (s <= -3) return s;
  U[0] = 1.0; U[1] = 0.0; U[2] = 0.0; U[3] = 1.0;
  Z[0] = 1.0; Z[1] = 0.0; Z[2] = 0.0; Z[3] = 1.0;
  dhsvd2t<F>(A, J, U, Z, info);
  if (info > 0) khsvd2<F>(A, J, G, B, C, s, th1, info);
  else if (info == 0 && s != 0) { A[0] = f_scale(A[0], -s); A[3] = f_scale(A[3], -s); }
  if (info < 0) return info;
  if (info > 0) {
    c2a<F>(B, U);
    a2c<F>(Z, C);
    if ((info & 8) == 0) {
      A[0] = G[0];
      A[3] = G[1];
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) A[k] = f_scale(A[k], -s);
      c2a<F>(B, A);
      a2c<F>(A, C);
    }
  }
  dhsvd2s(A, J, U, Z, info);
  return info;
}

// ---- complex -------------------------------------------------------------
template <bool F>
__device__ __forceinline__ dz zdfma(double a, dz b, dz c) { return zmk(fmad<F>(a, b.re, c.re), fmad<F>(a, b.im, c.im)); }
template <bool F>
__device__ __forceinline__ dz zzfma(dz a, dz b, dz c) {
  double f = fmad<F>(-a.im, b.im, c.re);
  double r = fmad<F>(a.re, b.re, f);
  f = fmad<F>(a.im, b.re, c.im);
  double i = fmad<F>(a.re, b.im, f);
  return zmk(r, i);
}
template <bool F>
__device__ __forceinline__ dz zzmul(dz a, dz b) {
  double f = a.im * b.im;
  double r = fmad<F>(a.re, b.re, -f);
  f = a.im * b.re;
  double i = fmad<F>(a.re, b.im, f);
  return zmk(r, i);
}
__device__ __forceinline__ dz zddiv(dz a, double b) { return zmk(a.re / b, a.im / b); }
template <bool F>
__device__ __forceinline__ dz zzdiv(dz a, dz b) {
  double f = absz(b);
  return zddiv(zzmul<F>(zconj(zddiv(b, f)), a), f);
}
__device__ __forceinline__ void zpolar(dz z, double &a, dz &e) {
  a = hypot_ref(z.re, z.im);
  double c = __builtin_copysign(mind(__builtin_fabs(z.re) / a, 1.0), z.re);
  double s = z.im / f_max(a, MINF);
  e = zmk(c, s);
}
template <bool F>
__device__ __forceinline__ void zc1a(const double B[4], dz &a0, dz &a1) {
  dz c0 = zddiv(zdfma<F>(B[2], a1, a0), B[0]);
  dz c1 = zddiv(zdfma<F>(B[1], a0, a1), B[3]);
  a0 = c0;
  a1 = c1;
}
template <bool F>
__device__ __forceinline__ void zc2a(const double B[4], dz A[4]) {
  dz C[4];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    C[0 + 2 * j] = zddiv(zdfma<F>(B[2], A[1 + 2 * j], A[0 + 2 * j]), B[0]);
    C[1 + 2 * j] = zddiv(zdfma<F>(B[1], A[0 + 2 * j], A[1 + 2 * j]), B[3]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) A[k] = C[k];
}
template <bool F>
__device__ __forceinline__ void za2c(dz A[4], const double B[4]) {
  dz C[4];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    C[i + 0] = zddiv(zdfma<F>(B[1], A[i + 2], A[i + 0]), B[0]);
    C[i + 2] = zddiv(zdfma<F>(B[2], A[i + 0], A[i + 2]), B[3]);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) A[k] = C[k];
}
__device__ __forceinline__ void zswap(dz &a, dz &b) {
  dz t = a;
  a = b;
  b = t;
}

template <bool F>
__device__ void zhsvd2t(dz A[4], const int J[2], dz U[4], dz Z[4], int &info) {
  double c, t, w;
  dz r, s;
  bool p, dg = true;
  if (zeq0(A[1])) c = absz(A[0]);
  else if (zeq0(A[0])) { c = absz(A[1]); dg = false; }
  else { c = hypot_ref(absz(A[0]), absz(A[1])); dg = false; }
  if (zeq0(A[2])) t = absz(A[3]);
  else if (zeq0(A[3])) { t = absz(A[2]); dg = false; }
  else { t = hypot_ref(absz(A[2]), absz(A[3])); dg = false; }
  if (c < t) {
    zswap(A[0], A[2]); zswap(A[1], A[3]);
    if (J[0] == J[1]) { zswap(Z[0], Z[2]); zswap(Z[1], Z[3]); }
    p = true;
    w = t;
  } else {
    p = false;
    w = c;
  }
  zpolar(A[0], c, r);
  if (zne1(r)) {
    s = zconj(r);
    U[0] = zzmul<F>(U[0], s); U[2] = zzmul<F>(U[2], s);
    A[0] = zzmul<F>(A[0], s); A[2] = zzmul<F>(A[2], s);
  }
  zpolar(A[1], t, s);
  if (zne1(s)) {
    r = zconj(s);
    U[1] = zzmul<F>(U[1], r); U[3] = zzmul<F>(U[3], r);
    A[1] = zzmul<F>(A[1], r); A[3] = zzmul<F>(A[3], r);
  }
  if (A[0].re < A[1].re) { zswap(U[0], U[1]); zswap(U[2], U[3]); zswap(A[0], A[1]); zswap(A[2], A[3]); }
  if (!dg) {
    t = A[1].re / A[0].re;
    c = __builtin_sqrt(fmad<F>(t, t, 1.0));
    double q[4] = {c, -t, t, c};
    zc2a<F>(q, U);
    zc1a<F>(q, A[2], A[3]);
  }
  A[0] = zmk(w, 0.0);
  A[1] = zmk(0.0, 0.0);
  if (J[0] != J[1] && p) {
    zswap(A[0], A[2]); zswap(A[1], A[3]);
    zswap(U[0], U[1]); zswap(U[2], U[3]);
    zswap(A[0], A[1]); zswap(A[2], A[3]);
    zpolar(A[1], t, s);
    if (zne1(s)) {
      r = zconj(s);
      Z[0] = zzmul<F>(Z[0], r); Z[1] = zzmul<F>(Z[1], r);
      A[0] = zzmul<F>(A[0], r); A[1] = zzmul<F>(A[1], r);
    }
    zpolar(A[0], c, r);
    if (zne1(r)) {
      s = zconj(r);
      U[0] = zzmul<F>(U[0], s); U[2] = zzmul<F>(U[2], s);
      A[0] = zzmul<F>(A[0], s);
    }
  } else {
    zpolar(A[2], c, r);
    if (zne1(r)) {
      s = zconj(r);
      Z[2] = zzmul<F>(Z[2], s); Z[3] = zzmul<F>(Z[3], s);
      A[2] = zzmul<F>(A[2], s); A[3] = zzmul<F>(A[3], s);
    }
    zpolar(A[3], t, s);
    if (zne1(s)) {
      r = zconj(s);
      U[1] = zzmul<F>(U[1], r); U[3] = zzmul<F>(U[3], r);
      A[3] = zzmul<F>(A[3], r);
    }
  }
  info = dg ? 0 : 1;
}

__device__ __forceinline__ void zhsvd2s(dz A[4], const int J[2], dz U[4], dz Z[4], int &info) {
  int t = J[0] + J[1];
  if ((t == 2 && A[0].re < A[3].re) || (t == -2 && A[3].re < A[0].re)) {
    zswap(U[0], U[1]); zswap(U[2], U[3]);
    zswap(A[0], A[3]);
    zswap(Z[0], Z[2]); zswap(Z[1], Z[3]);
  }
  if ((info & 8) == 0) { A[1] = zmk(0.0, 0.0); A[2] = zmk(0.0, 0.0); }
  if (zne1(U[0]) || !zeq0(U[1]) || !zeq0(U[2]) || zne1(U[3])) info += 2;
  if (zne1(Z[0]) || !zeq0(Z[1]) || !zeq0(Z[2]) || zne1(Z[3])) info += 4;
}

__device__ __forceinline__ void zscalea(dz A[4], int &s) {
  int sc = 0x7FFFFFFF;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int code = -((k >> 1) + 1) * 2 - ((k & 1) + 1);
    double aa = __builtin_fabs(A[k].re);
    if (!(aa <= 1.7976931348623157e308)) { s = code; return; }
    int v = EH - f_exponent(f_max(aa, MINF)) - 2;
    sc = (v < sc) ? v : sc;
    aa = __builtin_fabs(A[k].im);
    if (!(aa <= 1.7976931348623157e308)) { s = code; return; }
    v = EH - f_exponent(f_max(aa, MINF)) - 2;
    sc = (v < sc) ? v : sc;
  }
  if (sc != 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) A[k] = zmk(f_scale(A[k].re, sc), f_scale(A[k].im, sc));
  }
  s = sc;
}

template <bool F>
__device__ int zhsvd2(dz A[4], const int J[2], dz U[4], dz Z[4], double th1) {
  double G[2] = {0.0, 0.0}, W[4] = {0.0, 0.0, 0.0, 0.0}, B[4] = {0.0, 0.0, 0.0, 0.0}, C[4] = {0.0, 0.0, 0.0, 0.0};
  int s, info;
  zscalea(A, s);
  if (s <= -3) return s;
  U[0] = zmk(1, 0); U[1] = zmk(0, 0); U[2] = zmk(0, 0); U[3] = zmk(1, 0);
  Z[0] = zmk(1, 0); Z[1] = zmk(0, 0); Z[2] = zmk(0, 0); Z[3] = zmk(1, 0);
  zhsvd2t<F>(A, J, U, Z, info);
  if (info > 0) {
    W[0] = A[0].re; W[1] = A[1].re; W[2] = A[2].re; W[3] = A[3].re;
    khsvd2<F>(W, J, G, B, C, s, th1, info);
  } else if (info == 0) {
    if (s == 0) { W[0] = A[0].re; W[3] = A[3].re; }
    else { W[0] = f_scale(A[0].re, -s); W[3] = f_scale(A[3].re, -s); }
  }
  if (info < 0) return info;
  if (info > 0) {
    zc2a<F>(B, U);
    za2c<F>(Z, C);
    if ((info & 8) == 0) {
      W[0] = G[0];
      W[3] = G[1];
    } else {
      if (s != 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) A[k] = zmk(f_scale(A[k].re, -s), f_scale(A[k].im, -s));
      }
      zc2a<F>(B, A);
      za2c<F>(A, C);
      zhsvd2s(A, J, U, Z, info);
      return info;
    }
  }
  A[0] = zmk(W[0], 0.0);
  A[3] = zmk(W[3], 0.0);
  zhsvd2s(A, J, U, Z, info);
  return info;
}

// |x1|^2 + |x2|^2 scaled by the larger magnitude (DASUM2, src/D/dtransf.F90:12-38)
template <bool F>
__device__ __forceinline__ double dasum2(double x1, double x2) {
  double a = __builtin_fabs(x1), b = __builtin_fabs(x2);
  if (a >= b) {
    if (a == 0.0) return 0.0;
    return fmad<F>(b, (b / a), a) * a;
  }
  return fmad<F>(a, (a / b), b) * b;
}

// DASUM4, src/Z/ztransf.F90:18-75: sort the magnitudes, sum squares upwards
template <bool F>
__device__ __forceinline__ double dasum4(double x1, double x2, double x3, double x4) {
  double x[4] = {__builtin_fabs(x1), __builtin_fabs(x2), __builtin_fabs(x3), __builtin_fabs(x4)}, y[4];
  if (x[0] <= x[1]) { y[0] = x[0]; y[1] = x[1]; } else { y[0] = x[1]; y[1] = x[0]; }
  if (x[2] <= x[3]) { y[2] = x[2]; y[3] = x[3]; } else { y[2] = x[3]; y[3] = x[2]; }
  if (y[0] <= y[2]) { x[0] = y[0]; x[1] = y[2]; } else { x[0] = y[2]; x[1] = y[0]; }
  if (y[1] <= y[3]) { x[2] = y[1]; x[3] = y[3]; } else { x[2] = y[3]; x[3] = y[1]; }
  y[0] = x[0];
  if (x[1] <= x[2]) { y[1] = x[1]; y[2] = x[2]; } else { y[1] = x[2]; y[2] = x[1]; }
  y[3] = x[3];
  double s = y[0] * y[0];
  s = fmad<F>(y[1], y[1], s);
  s = fmad<F>(y[2], y[2], s);
  s = fmad<F>(y[3], y[3], s);
  return s;
}

}  // namespace jkg
