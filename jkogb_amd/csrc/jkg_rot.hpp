// jkg_rot.hpp -- the ratio-form 2x2 row/column transforms of BA/AB
// (src/D/dtransf.F90:121-242) and their descriptors (Rot, jkg_step.hpp).
#pragma once
#include "jkg_device.hpp"
#include "jkg_step.hpp"

namespace jkg {

// ratio-form 2x2 transform of the pair (x, y) (BA or AB, see Rot)
template <bool F>
__device__ __forceinline__ void rot_apply(const Rot &R, double &x, double &y) {
  double xx = fmad<F>(R.sa ? y : x, R.r1, R.sa ? x : y) * R.m1;
  double yy = fmad<F>(R.sb ? x : y, R.r2, R.sb ? y : x) * R.m2;
  x = xx;
  y = yy;
}
// descriptor of BA(B) (rows, src/D/dtransf.F90:121-183); equal to AB(B^T)
__device__ __forceinline__ Rot rot_ba(const double B[4]) {
  Rot R;
  R.sa = __builtin_fabs(B[0]) >= __builtin_fabs(B[2]);
  R.r1 = R.sa ? B[2] / B[0] : B[0] / B[2];
  R.m1 = R.sa ? B[0] : B[2];
  R.sb = __builtin_fabs(B[3]) >= __builtin_fabs(B[1]);
  R.r2 = R.sb ? B[1] / B[3] : B[3] / B[1];
  R.m2 = R.sb ? B[3] : B[1];
  R.pad0 = R.pad1 = 0;
  return R;
}
// descriptor of AB(B) (columns, src/D/dtransf.F90:187-242)
__device__ __forceinline__ Rot rot_ab(const double B[4]) {
  Rot R;
  R.sa = __builtin_fabs(B[0]) >= __builtin_fabs(B[1]);
  R.r1 = R.sa ? B[1] / B[0] : B[0] / B[1];
  R.m1 = R.sa ? B[0] : B[1];
  R.sb = __builtin_fabs(B[3]) >= __builtin_fabs(B[2]);
  R.r2 = R.sb ? B[2] / B[3] : B[3] / B[2];
  R.m2 = R.sb ? B[3] : B[2];
  R.pad0 = R.pad1 = 0;
  return R;
}

// The same transform on FP32 data (the FP32 warm start): the descriptor formed in
// FP64 and rounded once, the products in FP32.
struct RotF {
  float r1, r2, m1, m2;
  int sa, sb;
};
template <bool F>
__device__ __forceinline__ void rot_apply(const RotF &R, float &x, float &y) {
  float xx = fmadf<F>(R.sa ? y : x, R.r1, R.sa ? x : y) * R.m1;
  float yy = fmadf<F>(R.sb ? x : y, R.r2, R.sb ? y : x) * R.m2;
  x = xx;
  y = yy;
}
// the descriptor in the precision of the data it is applied to
__device__ __forceinline__ const Rot &as_rot(const Rot &R, const double *) { return R; }
__device__ __forceinline__ RotF as_rot(const Rot &R, const float *) {
  return RotF{(float)R.r1, (float)R.r2, (float)R.m1, (float)R.m2, R.sa, R.sb};
}

}  // namespace jkg
