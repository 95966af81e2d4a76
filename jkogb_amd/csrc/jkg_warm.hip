// jkg_warm.hip -- the mixed solvers' FP32 warm start (real and complex) and its
// hand-over to FP64.
//
// The warm start runs the dynamic-ordering step on FP32 copies of A, U and Z
// (every weight and 2x2 parameter still formed in FP64 from the promoted
// entries) until A's relative off-norm falls below a tolerance.  Its U and Z
// are then unitary and J-unitary only to FP32 rounding.  The hand-over makes
// them exact to FP64 rounding with Newton-Schulz iterations for the polar
// factor in the matrix group (Higham, Mackey, Mackey, Tisseur, "Computing the
// polar decomposition and the matrix sign decomposition in matrix groups",
// SIAM J. Matrix Anal. Appl. 25(4), 2004):
//   X <- X (3 I - X* X) / 2,  X* = X^H (unitary) or J X^H J (J-unitary),
// which converges quadratically from an FP32-accurate start, and then
// recomputes A = U^H Y Z in FP64 from the input Y.  The FP64 steps continue
// from there, so Y = U A Z^{-1} holds to FP64 rounding throughout and the
// result is an HSVD of Y itself, not of a perturbed matrix.  Every product is
// C = X^H Y on the FP64 matrix cores (gemm_tn_launch; complex matrices as
// interleaved rows); X B is formed as (X^H)^H B from an explicit (conjugate)
// transpose.  Real problems use the same steps with X^T for X^H.
#include "jkg_device.hpp"
#include "jkg_step.hpp"

namespace jkg {

static inline int wgrid(size_t work) {
  size_t g = (work + 255) / 256;
  if (g < 1) g = 1;
  return (int)(g > 16384 ? 16384 : g);
}

// A32 = A (rounded), U32 = Z32 = I
__global__ void k_warm_begin(int n, const dz *__restrict__ A, fz *__restrict__ A32, fz *__restrict__ U32,
                             fz *__restrict__ Z32) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const dz a = A[e];
    A32[e] = fz{(float)a.re, (float)a.im};
    const fz v = fz{(e / n == e % n) ? 1.0f : 0.0f, 0.0f};
    U32[e] = v;
    Z32[e] = v;
  }
}

// X = X32 (exact)
__global__ void k_warm_promote(size_t total, const fz *__restrict__ X32, dz *__restrict__ X) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    X[e] = ldz(X32[e]);
}

// P = X^H, 32 x 32 tiles through LDS
__global__ void __launch_bounds__(256) k_zadj(int n, const dz *__restrict__ X, dz *__restrict__ P) {
  __shared__ dz t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = bx + tx, c = by + k;
    t[k][tx] = (r < n && c < n) ? X[(size_t)c * n + r] : zmk(0.0, 0.0);
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = by + tx, c = bx + k;  // P(r, c) = conj(X(c, r))
    if (r < n && c < n) P[(size_t)c * n + r] = zconj(t[tx][k]);
  }
}

// Y = J X (rows scaled by the signs)
__global__ void k_jrows(int n, const dz *__restrict__ X, const int *__restrict__ J, dz *__restrict__ Y) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const dz x = X[e];
    Y[e] = (J[e % n] < 0) ? zmk(-x.re, -x.im) : x;
  }
}

// M = X* X's Gram form (X^H X, or X^H J X with J): B = (3 I - S M) / 2 with S = I or J
// (rows scaled), and per block sum |M - T|^2 with T = I or J.
__global__ void __launch_bounds__(256) k_ns_step(int n, const dz *__restrict__ M, const int *__restrict__ J,
                                                 dz *__restrict__ B, double *__restrict__ part) {
  __shared__ double red[256];
  const size_t total = (size_t)n * n;
  double acc = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    const dz m = M[e];
    const double sg = (J && J[r] < 0) ? -1.0 : 1.0;
    const double tgt = (r == c) ? sg : 0.0;
    const double dr = m.re - tgt, di = m.im;
    acc = __builtin_fma(dr, dr, __builtin_fma(di, di, acc));
    // S M: row r times sg; B = (3 I - S M) / 2
    B[e] = zmk(((r == c) ? 1.5 : 0.0) - 0.5 * sg * m.re, -0.5 * sg * m.im);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The real twins of the kernels above.
__global__ void k_warm_begin_real(int n, const double *__restrict__ A, float *__restrict__ A32,
                                  float *__restrict__ U32, float *__restrict__ Z32) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    A32[e] = (float)A[e];
    const float v = (e / n == e % n) ? 1.0f : 0.0f;
    U32[e] = v;
    Z32[e] = v;
  }
}

__global__ void k_warm_promote_real(size_t total, const float *__restrict__ X32, double *__restrict__ X) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    X[e] = (double)X32[e];
}

__global__ void __launch_bounds__(256) k_transpose(int n, const double *__restrict__ X, double *__restrict__ P) {
  __shared__ double t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = bx + tx, c = by + k;
    t[k][tx] = (r < n && c < n) ? X[(size_t)c * n + r] : 0.0;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = by + tx, c = bx + k;  // P(r, c) = X(c, r)
    if (r < n && c < n) P[(size_t)c * n + r] = t[tx][k];
  }
}

__global__ void k_jrows_real(int n, const double *__restrict__ X, const int *__restrict__ J, double *__restrict__ Y) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
    Y[e] = (J[e % n] < 0) ? -X[e] : X[e];
}

__global__ void __launch_bounds__(256) k_ns_step_real(int n, const double *__restrict__ M, const int *__restrict__ J,
                                                      double *__restrict__ B, double *__restrict__ part) {
  __shared__ double red[256];
  const size_t total = (size_t)n * n;
  double acc = 0.0;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    const double m = M[e];
    const double sg = (J && J[r] < 0) ? -1.0 : 1.0;
    const double dr = m - ((r == c) ? sg : 0.0);
    acc = __builtin_fma(dr, dr, acc);
    B[e] = ((r == c) ? 1.5 : 0.0) - 0.5 * sg * m;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

void warm_begin_launch(DeviceStep &d, hipStream_t s) {
  const dim3 g(wgrid((size_t)d.n * d.n));
  if (d.cplx)
    hipLaunchKernelGGL(k_warm_begin, g, dim3(256), 0, s, d.n, (const dz *)d.A, (fz *)d.A32, (fz *)d.U32,
                       (fz *)d.Z32);
  else
    hipLaunchKernelGGL(k_warm_begin_real, g, dim3(256), 0, s, d.n, (const double *)d.A, (float *)d.A32,
                       (float *)d.U32, (float *)d.Z32);
}

void warm_promote_launch(DeviceStep &d, hipStream_t s) {
  const size_t total = (size_t)d.n * d.n;
  if (d.cplx) {
    hipLaunchKernelGGL(k_warm_promote, dim3(wgrid(total)), dim3(256), 0, s, total, (const fz *)d.U32, (dz *)d.U);
    hipLaunchKernelGGL(k_warm_promote, dim3(wgrid(total)), dim3(256), 0, s, total, (const fz *)d.Z32, (dz *)d.Z);
  } else {
    hipLaunchKernelGGL(k_warm_promote_real, dim3(wgrid(total)), dim3(256), 0, s, total, (const float *)d.U32,
                       (double *)d.U);
    hipLaunchKernelGGL(k_warm_promote_real, dim3(wgrid(total)), dim3(256), 0, s, total, (const float *)d.Z32,
                       (double *)d.Z);
  }
}

// C = X^H Y (n x n, leading dimension n; complex interleaved, or real: X^T Y)
void gemm_h_launch(int n, int cplx, const void *X, const void *Y, void *C, hipStream_t s) {
  const double *x = (const double *)X, *y = (const double *)Y;
  double *c = (double *)C;
  if (!cplx) {
    gemm_tn_launch(n, x, n, nullptr, n, y, n, nullptr, n, c, 1L, (long)n, false, nullptr, s, 0);
    return;
  }
  gemm_tn_launch(2 * n, x, 2 * n, nullptr, n, y, 2 * n, nullptr, n, c, 2L, 2L * n, false, nullptr, s, 1);
  gemm_tn_launch(2 * n, x, 2 * n, nullptr, n, y, 2 * n, nullptr, n, c + 1, 2L, 2L * n, false, nullptr, s, 2);
}

// P = X^H (complex) or X^T (real)
void adj_launch(int n, int cplx, const void *X, void *P, hipStream_t s) {
  dim3 g((n + 31) / 32, (n + 31) / 32);
  if (cplx) hipLaunchKernelGGL(k_zadj, g, dim3(256), 0, s, n, (const dz *)X, (dz *)P);
  else hipLaunchKernelGGL(k_transpose, g, dim3(256), 0, s, n, (const double *)X, (double *)P);
}

void jrows_launch(int n, int cplx, const void *X, const int *J, void *Y, hipStream_t s) {
  if (cplx) hipLaunchKernelGGL(k_jrows, dim3(wgrid((size_t)n * n)), dim3(256), 0, s, n, (const dz *)X, J, (dz *)Y);
  else
    hipLaunchKernelGGL(k_jrows_real, dim3(wgrid((size_t)n * n)), dim3(256), 0, s, n, (const double *)X, J,
                       (double *)Y);
}

int ns_blocks(int n) { return wgrid((size_t)n * n) < 1024 ? wgrid((size_t)n * n) : 1024; }

void ns_step_launch(int n, int cplx, const void *M, const int *J, void *B, double *part, hipStream_t s) {
  if (cplx)
    hipLaunchKernelGGL(k_ns_step, dim3(ns_blocks(n)), dim3(256), 0, s, n, (const dz *)M, J, (dz *)B, part);
  else
    hipLaunchKernelGGL(k_ns_step_real, dim3(ns_blocks(n)), dim3(256), 0, s, n, (const double *)M, J, (double *)B,
                       part);
}

}  // namespace jkg
