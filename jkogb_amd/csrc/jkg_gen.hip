// jkg_gen.hip -- test matrices with a prescribed hyperbolic spectrum (the role
// of the reference's jkdatgen, src/var/old/jkdatgen/jkdatgen.F90:1-274,
// gendat.F90, xa2g.F90: a factor G of a symmetric A = G J G^T whose eigenvalues
// lambda are given, so the hyperbolic singular values of (G, J) are sqrt|lambda|).
//
// The reference forms A = U diag(lambda) U^T (XLAGSY) and factors it by
// Bunch-Parlett in extended precision (XSYBPC).  Here G is built directly:
//   G = Q diag(sqrt|lambda|) H,   J = sign(lambda),
// with Q orthogonal and H J-orthogonal (H^T J H = J), both products of rounds
// of 2x2 rotations on random perfect matchings of the indices: Q by Givens
// rotations of rows, H by column transforms that are hyperbolic (cosh, sinh)
// for a pair of opposite signs and trigonometric otherwise.  Then
// G J G^T = Q diag(lambda) Q^T exactly in exact arithmetic; in FP64 each round
// adds O(eps) relative to the row and column norms.  Rounds are HBM-bound
// passes: one workgroup per column pair (H) or per column (Q, every pair of
// the round applied in LDS).
#include "jkg_device.hpp"
#include "jkg_step.hpp"

namespace jkg {

// G = diag(d) (column-major, leading dimension n)
__global__ void k_gen_diag(int n, const double *__restrict__ d, double *__restrict__ G) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t c = e / n, r = e % n;
    G[e] = (r == c) ? d[r] : 0.0;
  }
}

// One round of column transforms: pair k = (a, b) with (c, s); hyperbolic if J_a != J_b:
//   [x y] <- [x y] [[c, s], [s, c]]  (c = cosh t, s = sinh t)
//   [x y] <- [x y] [[c, -s], [s, c]] (c = cos t,  s = sin t)
__global__ void __launch_bounds__(256) k_gen_cols(int n, double *__restrict__ G, const int *__restrict__ pa,
                                                  const int *__restrict__ pb, const double *__restrict__ cs,
                                                  const int *__restrict__ J) {
  const int k = blockIdx.x;
  const int a = pa[k], b = pb[k];
  const double c = cs[2 * k], s = cs[2 * k + 1];
  const bool hyp = J[a] != J[b];
  double *x = G + (size_t)a * n, *y = G + (size_t)b * n;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double xv = x[r], yv = y[r];
    x[r] = __builtin_fma(c, xv, s * yv);
    y[r] = hyp ? __builtin_fma(s, xv, c * yv) : __builtin_fma(c, yv, -(s * xv));
  }
}

// One round of row rotations, applied to every column (one workgroup per column, the
// column staged in LDS): (g_a, g_b) <- (c g_a - s g_b, s g_a + c g_b) for every pair.
__global__ void __launch_bounds__(256) k_gen_rows(int n, int npairs, double *__restrict__ G,
                                                  const int *__restrict__ pa, const int *__restrict__ pb,
                                                  const double *__restrict__ cs) {
  extern __shared__ double col[];
  double *g = G + (size_t)blockIdx.x * n;
  for (int r = threadIdx.x; r < n; r += blockDim.x) col[r] = g[r];
  __syncthreads();
  for (int k = threadIdx.x; k < npairs; k += blockDim.x) {
    const int a = pa[k], b = pb[k];
    const double c = cs[2 * k], s = cs[2 * k + 1];
    const double ga = col[a], gb = col[b];
    col[a] = __builtin_fma(c, ga, -(s * gb));
    col[b] = __builtin_fma(s, ga, c * gb);
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += blockDim.x) g[r] = col[r];
}

void gen_diag_launch(int n, const double *d, double *G, hipStream_t s) {
  size_t g = ((size_t)n * n + 255) / 256;
  hipLaunchKernelGGL(k_gen_diag, dim3((unsigned)(g > 8192 ? 8192 : g)), dim3(256), 0, s, n, d, G);
}

void gen_cols_launch(int n, int npairs, double *G, const int *pa, const int *pb, const double *cs, const int *J,
                     hipStream_t s) {
  if (npairs > 0) hipLaunchKernelGGL(k_gen_cols, dim3(npairs), dim3(256), 0, s, n, G, pa, pb, cs, J);
}

int gen_rows_prepare(int n) {
  const size_t lds = sizeof(double) * (size_t)n;
  if (lds > 160 * 1024) return -1;
  return hipFuncSetAttribute((const void *)k_gen_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
                 hipSuccess
             ? 0
             : -2;
}

void gen_rows_launch(int n, int npairs, double *G, const int *pa, const int *pb, const double *cs, hipStream_t s) {
  if (npairs > 0)
    hipLaunchKernelGGL(k_gen_rows, dim3(n), dim3(256), sizeof(double) * n, s, n, npairs, G, pa, pb, cs);
}

}  // namespace jkg

namespace jkg {

// ------------------------------------------------------------------ frames
// Per-step pictures of |A| (the role of the reference's ANIMATE frames,
// src/D/dstep.F90:485-495, 549-592, which libpvn's PVN_RVIS_* write): at the
// start of every `every`-th step (A before its transforms, as the reference
// records it) an S x S map, pixel (u, v) = the largest |a_ij| over the rows
// and columns that map to it, on a fixed logarithmic scale: 8 levels per
// decade from 1e-16 (level 1) to 1e+15.875 (255); 0 for exact zeros.  One
// workgroup per pixel row; the f-th frame of the run goes to frames[f * S * S], row-major.
template <class E>
__global__ void __launch_bounds__(256) k_frame(int n, const E *__restrict__ A, int lda, int size, int every,
                                               int max_frames, uint8_t *__restrict__ frames, const StepState *st) {
  if (!st->active) return;
  // frames at the steps s = 0 mod every; a resumed loop (step0 > 0) numbers its own from 0
  const int s = st->step, first = (st->step0 + every - 1) / every;
  if (s % every != 0 || s / every - first >= max_frames) return;
  uint8_t *f = frames + (size_t)(s / every - first) * size * size + (size_t)blockIdx.x * size;
  const int u = blockIdx.x;
  const int r0 = (int)((long)u * n / size), r1 = (int)((long)(u + 1) * n / size);
  for (int v = threadIdx.x; v < size; v += blockDim.x) {
    const int c0 = (int)((long)v * n / size), c1 = (int)((long)(v + 1) * n / size);
    double m = 0.0;
    for (int c = c0; c < c1; ++c)
      for (int r = r0; r < r1; ++r) {
        const dz a = ldz(A[(size_t)c * lda + r]);
        const double x = __builtin_sqrt(a.re * a.re + a.im * a.im);
        m = x > m ? x : m;
      }
    int lev = 0;
    if (m > 0.0) {
      lev = 1 + (int)__builtin_floor((log10(m) + 16.0) * 8.0);
      lev = lev < 1 ? 1 : (lev > 255 ? 255 : lev);
    }
    f[v] = (uint8_t)lev;
  }
}

void frame_launch(DeviceStep &d, hipStream_t s) {
  if (!d.frames) return;
  auto go = [&](auto kern, const auto *A) {
    hipLaunchKernelGGL(kern, dim3(d.frame_size), dim3(256), 0, s, d.n, A, d.lda, d.frame_size, d.frame_every,
                       d.frame_max, d.frames, d.st);
  };
  if (d.cplx && d.prec32) go(k_frame<fz>, (const fz *)d.A32);
  else if (d.cplx) go(k_frame<dz>, (const dz *)d.A);
  else if (d.prec32) go(k_frame<float>, (const float *)d.A32);
  else go(k_frame<double>, (const double *)d.A);
}

}  // namespace jkg
