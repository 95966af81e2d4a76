// jkg_gram.hip -- pair weights of the mixed solver in Gram form, and the FP64
// MFMA TN product they rest on.
//
// The reference's hyperbolic weight (DMAGF2H + ABODNDF2, src/D/dstep.F90:41-72,
// src/D/dtransf.F90:246-348) is the off-diagonal Frobenius^2 of columns p, q
// before minus after the 2x2 column transform B, rows p and q excepted, plus
// two block terms.  With x = A(:,p), y = A(:,q) over the rows r not in {p, q},
// x' = B11 x + B21 y and y' = B12 x + B22 y, that sum is a quadratic form in
//   gxx = sum x_r^2,  gyy = sum y_r^2,  gxy = sum x_r y_r:
//   W = d0 + gxx (1 - B11^2 - B12^2) + gyy (1 - B21^2 - B22^2)
//          - 2 gxy (B11 B21 + B12 B22).
// gxy for every hyperbolic pair is one product O(:,plus)^T O(:,minus), where O
// is A with its diagonal zeroed: (O^T O)_pq sums exactly the rows r != p, q,
// so no diagonal term has to be subtracted (no cancellation near convergence).
// gxx = c_p - A_qp^2 with c_p the off-diagonal column norm^2.  The product
// runs on the FP64 matrix cores; the weights are the reference's up to
// rounding (not bit for bit), which is what the mixed solver needs.
#include "jkg_device.hpp"
#include "jkg_hsvd2.hpp"
#include "jkg_step.hpp"

namespace jkg {

typedef double d4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ TN product
// C[i * csi + j * csj] = sum_{r < k} X[r, xc(i)] * Y[r, yc(j)]
//   xc(i) = xcols ? xcols[i] : i (i < m), yc(j) = ycols ? ycols[j] : j (j < nc);
// with ZD, X[r, c] and Y[r, c] read as 0 for r == c (the diagonal of O).
// A 256-lane workgroup computes a 128 x 128 tile; each wave 64 x 64 as 4 x 4
// MFMA f64 16x16x4 tiles.  Rows are staged 16 at a time through LDS (two
// buffers: the next stage's global loads are in flight during the MFMAs).
// MFMA f64 16x16x4 lane layout: A[i = l % 16][k = l / 16], B[k = l / 16][j =
// l % 16], D item t: [i = 4 t + l / 16][j = l % 16].
// Complex columns (CX > 0) are read as 2k interleaved doubles (re, im per row);
// the diagonal of O is then the pair of flat rows 2c, 2c + 1.  Over the flat
// rows, sum X Y = Re(X^H Y) (CX = 1), and with Y's pairs turned to (im, -re),
// sum X Y' = Im(X^H Y) (CX = 2): the complex product is two real ones.
constexpr int GB_M = 128, GB_N = 128, GB_K = 16, GB_LD = GB_M + 4;

// Split K: workgroup z sums the rows [z kchunk, (z + 1) kchunk) into its own
// plane C + z plane (the consumer adds the planes in a fixed order, so results
// do not depend on scheduling); kchunk is a multiple of GB_K.
template <bool ZD, int CX, class T>
__global__ void __launch_bounds__(256) k_gemm_tn(int k, int kchunk, long plane, const T *__restrict__ X, int ldx,
                                                 const int *__restrict__ xcols, int m, const T *__restrict__ Y,
                                                 int ldy, const int *__restrict__ ycols, int nc, double *__restrict__ C,
                                                 long csi, long csj, const StepState *st) {
  if (st && !st->active) return;
  const int kb = blockIdx.z * kchunk;
  const int ke = (k - kb < kchunk) ? k : kb + kchunk;
  C += blockIdx.z * plane;
  __shared__ double Xs[2][GB_K][GB_LD];
  __shared__ double Ys[2][GB_K][GB_LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  // loader: lane pair (c = tid / 2) holds rows (tid & 1) * 8 .. + 7 of column c
  const int lc = tid >> 1, lh = (tid & 1) * 8;
  const int xi = i0 + lc, yj = j0 + lc;
  const int xcol = (xi < m) ? (xcols ? xcols[xi] : xi) : -1;
  const int ycol = (yj < nc) ? (ycols ? ycols[yj] : yj) : -1;
  const T *xp = X + (size_t)(xcol < 0 ? 0 : xcol) * ldx;
  const T *yp = Y + (size_t)(ycol < 0 ? 0 : ycol) * ldy;
  double xr[8], yr[8];
  // 16-byte column starts (even leading dimension): a lane's 8 rows are four 16-byte loads
  // (rows kb + r0 + lh are a multiple of 8), issued unconditionally from the clamped column; the
  // entries to zero (no column, the diagonal of O) are cleared when the stage is written to LDS,
  // after the products, so nothing waits for the loads before then.  The last group of a chunk
  // that ends inside it is loaded row by row.
  const bool vec = sizeof(T) == 8 && (ldx & 1) == 0 && (ldy & 1) == 0;
  bool raw = false;  // xr, yr hold unmasked vector loads
  auto gload = [&](int r0) {
    const int rb = kb + r0 + lh;
    raw = vec && rb + 8 <= ke;
    if (raw) {
      typedef double d2v __attribute__((ext_vector_type(2)));
      const d2v *px = (const d2v *)(xp + rb), *py = (const d2v *)(yp + rb);
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const d2v a = px[h], b = py[h];
        xr[2 * h] = a.x;
        xr[2 * h + 1] = a.y;
        yr[2 * h] = b.x;
        yr[2 * h + 1] = b.y;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = rb + u;
        const int rd = CX ? (r >> 1) : r;
        xr[u] = (xcol >= 0 && r < ke && !(ZD && rd == xcol)) ? (double)xp[r] : 0.0;
        yr[u] = (ycol >= 0 && r < ke && !(ZD && rd == ycol)) ? (double)yp[r] : 0.0;
      }
    }
  };
  auto sstore = [&](int buf, int r0) {
    const int rb = kb + r0 + lh;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int rd = CX ? ((rb + u) >> 1) : rb + u;
      double x = xr[u], y = yr[u];
      if (raw) {
        x = (xcol < 0 || (ZD && rd == xcol)) ? 0.0 : x;
        y = (ycol < 0 || (ZD && rd == ycol)) ? 0.0 : y;
      }
      yr[u] = y;
      Xs[buf][lh + u][lc] = x;
    }
    if (CX == 2) {  // rows r0 + lh + u are even for even u: (re, im) -> (im, -re)
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        const double re = yr[u];
        yr[u] = yr[u + 1];
        yr[u + 1] = -re;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) Ys[buf][lh + u][lc] = yr[u];
  };
  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
  gload(0);
  sstore(0, 0);
  __syncthreads();
  const int nst = (ke - kb + GB_K - 1) / GB_K;
  for (int s = 0; s < nst; ++s) {
    const int buf = s & 1;
    if (s + 1 < nst) gload((s + 1) * GB_K);
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      double af[4], bf[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) af[a] = Xs[buf][kk + (lane >> 4)][wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 4; ++b) bf[b] = Ys[buf][kk + (lane >> 4)][wn * 64 + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    if (s + 1 < nst) sstore(buf ^ 1, (s + 1) * GB_K);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = i0 + wm * 64 + a * 16 + 4 * t + (lane >> 4);
        const int j = j0 + wn * 64 + b * 16 + (lane & 15);
        if (i < m && j < nc) C[i * csi + j * csj] = acc[a][b][t];
      }
}

// K split of a product with an m x nc output: enough workgroups to cover the
// chip twice (2 fit a CU), at most GRAM_KS_MAX, at least 256 rows each.
int gram_ksplit(int m, int nc, int k) {
  const int tiles = ((m + GB_M - 1) / GB_M) * ((nc + GB_N - 1) / GB_N);
  int ks = (512 + tiles - 1) / (tiles > 0 ? tiles : 1);
  const int kmax = k / 256 > 1 ? k / 256 : 1;
  if (ks > kmax) ks = kmax;
  if (ks > GRAM_KS_MAX) ks = GRAM_KS_MAX;
  return ks < 1 ? 1 : ks;
}

// Doubles of the Gram planes for an N x N problem, per part (real or imaginary):
// npl nmi <= floor(N/2) ceil(N/2) for any J split, and ks <= GRAM_KS_MAX.
size_t gram_plane_doubles(int n) { return (size_t)GRAM_KS_MAX * (size_t)(n / 2) * (size_t)(n - n / 2); }

static inline int kchunk_of(int k, int ks) { return (((k + ks - 1) / ks) + GB_K - 1) / GB_K * GB_K; }

void gemm_tn_launch(int k, const double *X, int ldx, const int *xcols, int m, const double *Y, int ldy,
                    const int *ycols, int nc, double *C, long csi, long csj, bool zero_diag, const StepState *st,
                    hipStream_t s, int cx, int ks) {  // cx: 0 real, 1 / 2 real / imaginary part of X^H Y
  if (m <= 0 || nc <= 0) return;
  dim3 grid((nc + GB_N - 1) / GB_N, (m + GB_M - 1) / GB_M, ks);
  const int kc = kchunk_of(k, ks);
  const long plane = (long)m * nc;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, k, kc, plane, X, ldx, xcols, m, Y, ldy, ycols, nc, C, csi, csj,
                       st);
  };
  if (cx == 1) go(zero_diag ? k_gemm_tn<true, 1, double> : k_gemm_tn<false, 1, double>);
  else if (cx == 2) go(zero_diag ? k_gemm_tn<true, 2, double> : k_gemm_tn<false, 2, double>);
  else go(zero_diag ? k_gemm_tn<true, 0, double> : k_gemm_tn<false, 0, double>);
}

// The warm start's Gram product on the FP32 matrix cores (twice the FP64 rate):
// the tiling of k_gemm_tn with FP32 operands in LDS and FP32 sums over each K
// chunk; the chunks' planes are added in FP64 by the weight kernel.  CX = 0: real
// columns (the diagonal of O is row c of column c); CX = 1, 2: complex columns
// read as interleaved rows, as in k_gemm_tn.
// MFMA f32 16x16x4 lane layout: A[i = l % 16][k = l / 16], B[k = l / 16][j =
// l % 16], D item t: [i = 4 (l / 16) + t][j = l % 16].
typedef float f4v __attribute__((ext_vector_type(4)));

template <int CX>
__global__ void __launch_bounds__(256) k_zgram_f32(int k, int kchunk, long plane, const float *__restrict__ X, int ldx,
                                                   const int *__restrict__ xcols, int m,
                                                   const int *__restrict__ ycols, int nc, double *__restrict__ C,
                                                   long csi, long csj, const StepState *st) {
  if (!st->active) return;
  __shared__ float Xs[2][GB_K][GB_LD];
  __shared__ float Ys[2][GB_K][GB_LD];
  const int kb = blockIdx.z * kchunk;
  const int ke = (k - kb < kchunk) ? k : kb + kchunk;
  C += blockIdx.z * plane;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int i0 = blockIdx.y * GB_M, j0 = blockIdx.x * GB_N;
  const int lc = tid >> 1, lh = (tid & 1) * 8;
  const int xi = i0 + lc, yj = j0 + lc;
  const int xcol = (xi < m) ? xcols[xi] : -1;
  const int ycol = (yj < nc) ? ycols[yj] : -1;
  constexpr int DS = CX ? 1 : 0;  // flat row r belongs to matrix row r >> DS
  const float *xp = X + (size_t)(xcol < 0 ? 0 : xcol) * ldx;
  const float *yp = X + (size_t)(ycol < 0 ? 0 : ycol) * ldx;
  float xr[8], yr[8];
  // as in k_gemm_tn: two unconditional 16-byte loads per lane when the column starts are 16-byte
  // aligned (ldx a multiple of 4) and the 8 rows lie inside the chunk, the diagonal pair and absent
  // columns zeroed when the stage is written to LDS, after the products
  const bool vec = (ldx & 3) == 0;
  bool raw = false;
  auto gload = [&](int r0) {
    const int rb = kb + r0 + lh;
    raw = vec && rb + 8 <= ke;
    if (raw) {
      const f4v *px = (const f4v *)(xp + rb), *py = (const f4v *)(yp + rb);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const f4v a = px[h], b = py[h];
        xr[4 * h] = a.x;
        xr[4 * h + 1] = a.y;
        xr[4 * h + 2] = a.z;
        xr[4 * h + 3] = a.w;
        yr[4 * h] = b.x;
        yr[4 * h + 1] = b.y;
        yr[4 * h + 2] = b.z;
        yr[4 * h + 3] = b.w;
      }
    } else {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int r = rb + u;
        xr[u] = (xcol >= 0 && r < ke && (r >> DS) != xcol) ? xp[r] : 0.0f;
        yr[u] = (ycol >= 0 && r < ke && (r >> DS) != ycol) ? yp[r] : 0.0f;
      }
    }
  };
  auto sstore = [&](int buf, int r0) {
    const int rb = kb + r0 + lh;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      float x = xr[u], y = yr[u];
      if (raw) {
        x = (xcol < 0 || ((rb + u) >> DS) == xcol) ? 0.0f : x;
        y = (ycol < 0 || ((rb + u) >> DS) == ycol) ? 0.0f : y;
      }
      yr[u] = y;
      Xs[buf][lh + u][lc] = x;
    }
    if (CX == 2) {  // rows rb + u are even for even u: (re, im) -> (im, -re)
#pragma unroll
      for (int u = 0; u < 8; u += 2) {
        const float re = yr[u];
        yr[u] = yr[u + 1];
        yr[u + 1] = -re;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) Ys[buf][lh + u][lc] = yr[u];
  };
  f4v acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f4v){0.f, 0.f, 0.f, 0.f};
  gload(0);
  sstore(0, 0);
  __syncthreads();
  const int nst = (ke - kb + GB_K - 1) / GB_K;
  for (int s = 0; s < nst; ++s) {
    const int buf = s & 1;
    if (s + 1 < nst) gload((s + 1) * GB_K);
#pragma unroll
    for (int kk = 0; kk < GB_K; kk += 4) {
      float af[4], bf[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) af[a] = Xs[buf][kk + (lane >> 4)][wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 4; ++b) bf[b] = Ys[buf][kk + (lane >> 4)][wn * 64 + b * 16 + (lane & 15)];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    if (s + 1 < nst) sstore(buf ^ 1, (s + 1) * GB_K);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int i = i0 + wm * 64 + a * 16 + 4 * (lane >> 4) + t;
        const int j = j0 + wn * 64 + b * 16 + (lane & 15);
        if (i < m && j < nc) C[i * csi + j * csj] = (double)acc[a][b][t];
      }
}

static void zgemm_tn_f32_launch(int k, const float *X, int ldx, const int *xcols, int m, const int *ycols, int nc,
                                double *C, long csi, long csj, const StepState *st, hipStream_t s, int cx, int ks) {
  if (m <= 0 || nc <= 0) return;
  dim3 grid((nc + GB_N - 1) / GB_N, (m + GB_M - 1) / GB_M, ks);
  auto kern = (cx == 1) ? k_zgram_f32<1> : ((cx == 2) ? k_zgram_f32<2> : k_zgram_f32<0>);
  hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, k, kchunk_of(k, ks), (long)m * nc, X, ldx, xcols, m, ycols, nc, C,
                     csi, csj, st);
}

// ------------------------------------------------------------------ column norms
// c[col] = sum_{r != col} A[r, col]^2 and dg[col] = A[col, col], one wave per column
// (E = float: the FP32 warm start's A, promoted).
template <class E>
__global__ void __launch_bounds__(256) k_colnorm_off(int n, const E *__restrict__ A, int lda,
                                                     double *__restrict__ c, double *__restrict__ dg,
                                                     const StepState *st) {
  if (!st->active) return;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (col >= n) return;
  const E *a = A + (size_t)col * lda;
  // a lane's rows lane, lane + 64, ... summed in that order; CB of them loaded at a time (from
  // clamped rows) before any is added
  constexpr int CB = 16;
  double s = 0.0;
  for (int r0 = lane; r0 < n; r0 += 64 * CB) {
    E v[CB];
#pragma unroll
    for (int t = 0; t < CB; ++t) v[t] = a[min(r0 + 64 * t, n - 1)];
#pragma unroll
    for (int t = 0; t < CB; ++t) {
      const int r = r0 + 64 * t;
      const double x = (double)v[t];
      if (r == col) dg[col] = x;
      else if (r < n) s = __builtin_fma(x, x, s);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) c[col] = s;
}

// ------------------------------------------------------------------ weights
// One workgroup per 32 (J = +1) x 32 (J = -1) tile of hyperbolic pairs.  The
// two off-diagonal blocks A(minus, plus) and A(plus, minus) of the tile are
// gathered through LDS so that every global read runs down a column; the
// diagonal comes from dg, staged with the columns' norms and signs.  The branch structure is DMAGF2H's; only the
// ABODNDF2 sum is replaced by the quadratic form above.
constexpr int WT = 32;

template <bool F, class E>
__global__ void __launch_bounds__(256) k_hyp_weights_gram(int n, const E *__restrict__ A, int lda,
                                                          const int *__restrict__ J, const int *__restrict__ plus,
                                                          int npl, const int *__restrict__ minus, int nmi,
                                                          const double *__restrict__ G, int ldg, int gks,
                                                          const double *__restrict__ coln,
                                                          const double *__restrict__ dg,
                                                          uint64_t *__restrict__ keys,
                                                          StepState *st, double tol, int cut) {
  if (!st->active) return;
  __shared__ double E1[WT][WT + 1];  // E1[a][b] = A(minus b, plus a)
  __shared__ double E2[WT][WT + 1];  // E2[b][a] = A(plus a, minus b)
  __shared__ int pc[WT], mc[WT];     // the tile's columns (-1: none), plus then minus
  __shared__ double cdg[2 * WT], cnr[2 * WT];  // their diagonal entries and off-diagonal norms^2
  __shared__ int cj[2 * WT];                   // and signs
  const int i0 = blockIdx.y * WT, j0 = blockIdx.x * WT, tid = threadIdx.x;
  // Every global load is issued early and unconditionally (indices clamped into their arrays,
  // what lies outside masked afterwards): the Gram entries of this lane's pairs first (they
  // need no column index), then the column data, then the two blocks.
  constexpr int PL = WT * WT / 256;  // pairs per lane
  constexpr int GPF = 2;             // Gram planes prefetched (the rest, if any, read in place)
  const size_t plane = (size_t)npl * nmi;
  double gv[PL][GPF];
#pragma unroll
  for (int i = 0; i < PL; ++i) {
    const int e = tid + 256 * i, a = e / WT, b = e % WT;
    const size_t gi = (size_t)min(i0 + a, npl - 1) * ldg + min(j0 + b, nmi - 1);
#pragma unroll
    for (int z = 0; z < GPF; ++z) gv[i][z] = (z < gks) ? G[gi + z * plane] : 0.0;
  }
  if (tid < 2 * WT) {
    const bool isp = tid < WT;
    const int g = isp ? i0 + tid : j0 + tid - WT;
    const bool in = g < (isp ? npl : nmi);
    const int c = isp ? plus[min(g, npl - 1)] : minus[min(g, nmi - 1)];
    const double vd = dg[c], vn = coln[c];
    const int vj = J[c];
    if (isp) pc[tid] = in ? c : -1;
    else mc[tid - WT] = in ? c : -1;
    cdg[tid] = vd;
    cnr[tid] = vn;
    cj[tid] = vj;
  }
  __syncthreads();
  {
    double v1[PL], v2[PL];
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      const int e = tid + 256 * i, hi = e / WT, lo = e % WT;
      const int c1 = pc[hi], r1 = mc[lo];  // column plus hi, row minus lo
      const int c2 = mc[hi], r2 = pc[lo];  // column minus hi, row plus lo
      v1[i] = (double)A[(size_t)max(c1, 0) * lda + max(r1, 0)];
      v2[i] = (double)A[(size_t)max(c2, 0) * lda + max(r2, 0)];
    }
#pragma unroll
    for (int i = 0; i < PL; ++i) {
      const int e = tid + 256 * i, hi = e / WT, lo = e % WT;
      E1[hi][lo] = (pc[hi] >= 0 && mc[lo] >= 0) ? v1[i] : 0.0;
      E2[hi][lo] = (mc[hi] >= 0 && pc[lo] >= 0) ? v2[i] : 0.0;
    }
  }
  __syncthreads();
  unsigned nfull = 0;
#pragma unroll
  for (int i = 0; i < PL; ++i) {
    const int e = tid + 256 * i, a = e / WT, b = e % WT;
    const int cp = pc[a], cm = mc[b];
    if (cp < 0 || cm < 0) continue;
    const bool pfirst = cp < cm;  // the plus column is p (the smaller index)
    const int p0 = pfirst ? cp : cm, q0 = pfirst ? cm : cp;
    const double e1 = E1[a][b], e2 = E2[b][a];  // A(cm, cp), A(cp, cm)
    const double aqp = pfirst ? e1 : e2, apq = pfirst ? e2 : e1;
    const double app = pfirst ? cdg[a] : cdg[WT + b], aqq = pfirst ? cdg[WT + b] : cdg[a];
    uint64_t key = KEY_NAN;
    const bool order = sign_neg(app) || sign_neg(aqq);
    const bool drop = tol > 0.0 && !order && negligible(aqp, apq, app, aqq, tol);
    if ((aqp != 0.0 || apq != 0.0 || order) && !drop) {
      double a2[4] = {app, aqp, apq, aqq}, u2[4], z2[4];
      int j2[2] = {pfirst ? cj[a] : cj[WT + b], pfirst ? cj[WT + b] : cj[a]};
      const int info = dhsvd2<F>(a2, j2, u2, z2, TH1FIX);
      if (info < 0) key = KEY_NAN;
      else if ((info & 1) == 0) key = weight_key(0.0);
      else if ((info & 4) == 0) key = weight_key(dasum2<F>(aqp, apq));
      else {
        ++nfull;
        const double aq = (info & 8) ? a2[1] : 0.0, ap = (info & 8) ? a2[2] : 0.0;
        double d = fmad<F>(aqp - aq, aqp + aq, 0.0);
        d = fmad<F>(apq - ap, apq + ap, d);
        const double cnp = pfirst ? cnr[a] : cnr[WT + b], cnq = pfirst ? cnr[WT + b] : cnr[a];
        const double gxx = cnp - aqp * aqp, gyy = cnq - apq * apq;
        double gxy = gv[i][0];
        for (int z = 1; z < gks; ++z)
          gxy += (z < GPF) ? gv[i][z < GPF ? z : 0] : G[(size_t)(i0 + a) * ldg + j0 + b + z * plane];
        const double b11 = z2[0], b21 = z2[1], b12 = z2[2], b22 = z2[3];
        const double w = d + gxx * (1.0 - b11 * b11 - b12 * b12) + gyy * (1.0 - b21 * b21 - b22 * b22) -
                         2.0 * gxy * (b11 * b21 + b12 * b22);
        key = (w < -1.7976931348623157e308 || !(w == w)) ? KEY_NAN : weight_key(w);
      }
    }
    const uint32_t k = pair_index(n, p0 + 1, q0 + 1);
    keys[k] = key_cut(key, cut);
  }
  // work counter: pairs whose weight needed the column sums
  for (int o = 32; o > 0; o >>= 1) nfull += __shfl_xor(nfull, o);
  __shared__ unsigned wfull[4];
  if ((tid & 63) == 0) wfull[tid >> 6] = nfull;
  __syncthreads();
  if (tid == 0 && wfull[0] + wfull[1] + wfull[2] + wfull[3])
    acc_abod_add(st, (unsigned long long)(wfull[0] + wfull[1] + wfull[2] + wfull[3]));
}

static void zgemm_tn_f32_launch(int k, const float *X, int ldx, const int *xcols, int m, const int *ycols, int nc,
                                double *C, long csi, long csj, const StepState *st, hipStream_t s, int cx, int ks);
template <class DG>
__global__ void k_warm_offnorm(int n, const double *__restrict__ c, const DG *__restrict__ dg, StepState *st,
                               double switch_tol);

void gram_hyp_weights_launch(DeviceStep &d, hipStream_t s) {
  const int ks = gram_ksplit(d.npl, d.nmi, d.n);
  dim3 g((d.nmi + WT - 1) / WT, (d.npl + WT - 1) / WT);
  auto weights = [&](auto kern, const auto *A, int cut) {
    hipLaunchKernelGGL(kern, g, dim3(256), 0, s, d.n, A, d.lda, d.J, d.plus, d.npl, d.minus, d.nmi, d.gram, d.nmi,
                       ks, d.coln, d.coln + d.n, d.sort.k0, d.st, d.neg_tol, cut);
  };
  if (d.prec32) {  // the FP32 warm start: column norms, its end test, the FP32 Gram product
    const float *A = (const float *)d.A32;
    hipLaunchKernelGGL(k_colnorm_off<float>, dim3((d.n + 3) / 4), dim3(256), 0, s, d.n, A, d.lda, d.coln,
                       d.coln + d.n, d.st);
    hipLaunchKernelGGL(k_warm_offnorm<double>, dim3(1), dim3(256), 0, s, d.n, d.coln, d.coln + d.n, d.st,
                       d.switch_tol);
    if (d.npl <= 0 || d.nmi <= 0) return;
    zgemm_tn_f32_launch(d.n, A, d.lda, d.plus, d.npl, d.minus, d.nmi, d.gram, (long)d.nmi, 1L, d.st, s, 0, ks);
    if (d.fused) weights(k_hyp_weights_gram<true, float>, A, d.cut32);
    else weights(k_hyp_weights_gram<false, float>, A, d.cut32);
    return;
  }
  if (d.npl <= 0 || d.nmi <= 0) return;
  const double *A = (const double *)d.A;
  hipLaunchKernelGGL(k_colnorm_off<double>, dim3((d.n + 3) / 4), dim3(256), 0, s, d.n, A, d.lda, d.coln,
                     d.coln + d.n, d.st);
  gemm_tn_launch(d.n, A, d.lda, d.plus, d.npl, A, d.lda, d.minus, d.nmi, d.gram, (long)d.nmi, 1L, true, d.st, s, 0,
                 ks);
  if (d.fused) weights(k_hyp_weights_gram<true, double>, A, 0);
  else weights(k_hyp_weights_gram<false, double>, A, 0);
}

// ------------------------------------------------------------------ complex
// The complex hyperbolic weight (ZMAGF2H + ABODNZF2, src/Z/zstep.F90:43-74,
// src/Z/ztransf.F90:452-637) in the same form: with x' = W11 x + W21 y,
// y' = W12 x + W22 y (W the pair's complex column transform) and
// gxy = sum conj(x_r) y_r,
//   W = d0 + gxx (1 - |W11|^2 - |W12|^2) + gyy (1 - |W21|^2 - |W22|^2)
//          - 2 Re((conj(W11) W21 + conj(W12) W22) gxy).
// G = O(:,plus)^H O(:,minus) (real and imaginary parts as two products); for a
// pair whose J = +1 column is the larger index, gxy = conj(G).

// c[col] = sum_{r != col} |A[r, col]|^2 and dg[col] = A[col, col] (complex), one wave per column.
template <class E>
__global__ void __launch_bounds__(256) k_zcolnorm_off(int n, const E *__restrict__ A, int lda,
                                                      double *__restrict__ c, dz *__restrict__ dg,
                                                      const StepState *st) {
  if (!st->active) return;
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (col >= n) return;
  const E *a = A + (size_t)col * lda;
  // as k_colnorm_off: the rows in the same order, ZCB of them loaded at a time from clamped rows
  constexpr int ZCB = 8;
  double s = 0.0;
  for (int r0 = lane; r0 < n; r0 += 64 * ZCB) {
    E v[ZCB];
#pragma unroll
    for (int t = 0; t < ZCB; ++t) v[t] = a[min(r0 + 64 * t, n - 1)];
#pragma unroll
    for (int t = 0; t < ZCB; ++t) {
      const int r = r0 + 64 * t;
      const dz x = ldz(v[t]);
      if (r == col) dg[col] = x;
      else if (r < n) s = __builtin_fma(x.im, x.im, __builtin_fma(x.re, x.re, s));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) c[col] = s;
}

// The warm start's end test, after the column norms: the relative off-norm
// sqrt(sum c / (sum c + sum |a_ii|^2)) of this step's A; below switch_tol, or
// no 1 % improvement in WARM_STALL steps, ends the FP32 phase before this step
// (the rest of the step is skipped; the loop reports done).
__device__ __forceinline__ double sq_acc(double v, double b) { return __builtin_fma(v, v, b); }
__device__ __forceinline__ double sq_acc(dz v, double b) { return __builtin_fma(v.re, v.re, __builtin_fma(v.im, v.im, b)); }
template <class DG>
__global__ void __launch_bounds__(256) k_warm_offnorm(int n, const double *__restrict__ c, const DG *__restrict__ dg,
                                                      StepState *st, double switch_tol) {
  if (!st->active) return;
  __shared__ double so[256], sd[256];
  // a lane's entries i, i + 256, ... added in that order, OB of them loaded at a time
  constexpr int OB = 8;
  double a = 0.0, b = 0.0;
  for (int i0 = threadIdx.x; i0 < n; i0 += 256 * OB) {
    double cv[OB];
    DG dv[OB];
#pragma unroll
    for (int t = 0; t < OB; ++t) {
      const int i = min(i0 + 256 * t, n - 1);
      cv[t] = c[i];
      dv[t] = dg[i];
    }
#pragma unroll
    for (int t = 0; t < OB; ++t) {
      if (i0 + 256 * t < n) {
        a += cv[t];
        b = sq_acc(dv[t], b);
      }
    }
  }
  so[threadIdx.x] = a;
  sd[threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      so[threadIdx.x] += so[threadIdx.x + h];
      sd[threadIdx.x] += sd[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double off = so[0], tot = so[0] + sd[0];
    const double rel = (tot > 0.0) ? __builtin_sqrt(off / tot) : 0.0;
    const int step = st->step;
    st->off_rel = rel;
    if (rel < 0.99 * st->off_best) {
      st->off_best = rel;
      st->off_best_step = step;
    }
    int end = 0;
    if (rel < switch_tol) end = 1;
    else if (step - st->off_best_step >= WARM_STALL) end = 2;
    if (end) {
      st->warm_end = end;
      st->done = 1;
      st->info = step;
      st->active = 0;
      st->sel_done = 1;
      st->nn_sort = 0;
    }
  }
}

constexpr int ZWT = 32;

template <bool F, class E>
__global__ void __launch_bounds__(256) k_zhyp_weights_gram(int n, const E *__restrict__ A, int lda,
                                                           const int *__restrict__ J, const int *__restrict__ plus,
                                                           int npl, const int *__restrict__ minus, int nmi,
                                                           const double *__restrict__ Gre,
                                                           const double *__restrict__ Gim, int ldg, int gks,
                                                           const double *__restrict__ coln,
                                                           const dz *__restrict__ dg, uint64_t *__restrict__ keys,
                                                           StepState *st, double tol,
                                                           int cut) {
  if (!st->active) return;
  __shared__ dz E1[ZWT][ZWT + 1];  // E1[a][b] = A(minus b, plus a)
  __shared__ dz E2[ZWT][ZWT + 1];  // E2[b][a] = A(plus a, minus b)
  __shared__ int pc[ZWT], mc[ZWT];
  const int i0 = blockIdx.y * ZWT, j0 = blockIdx.x * ZWT, tid = threadIdx.x;
  if (tid < ZWT) pc[tid] = (i0 + tid < npl) ? plus[i0 + tid] : -1;
  else if (tid < 2 * ZWT) mc[tid - ZWT] = (j0 + tid - ZWT < nmi) ? minus[j0 + tid - ZWT] : -1;
  __syncthreads();
  for (int e = tid; e < ZWT * ZWT; e += 256) {
    const int hi = e / ZWT, lo = e % ZWT;
    const int c1 = pc[hi], r1 = mc[lo];
    E1[hi][lo] = (c1 >= 0 && r1 >= 0) ? ldz(A[(size_t)c1 * lda + r1]) : zmk(0.0, 0.0);
    const int c2 = mc[hi], r2 = pc[lo];
    E2[hi][lo] = (c2 >= 0 && r2 >= 0) ? ldz(A[(size_t)c2 * lda + r2]) : zmk(0.0, 0.0);
  }
  __syncthreads();
  unsigned nfull = 0;
  for (int e = tid; e < ZWT * ZWT; e += 256) {
    const int a = e / ZWT, b = e % ZWT;
    const int cp = pc[a], cm = mc[b];
    if (cp < 0 || cm < 0) continue;
    const int p0 = cp < cm ? cp : cm, q0 = cp < cm ? cm : cp;
    const dz e1 = E1[a][b], e2 = E2[b][a];  // A(cm, cp), A(cp, cm)
    const dz aqp = cp < cm ? e1 : e2, apq = cp < cm ? e2 : e1;
    const dz app = dg[p0], aqq = dg[q0];
    uint64_t key = KEY_NAN;
    const bool diag = app.im != 0.0 || aqq.im != 0.0 || sign_neg(app.re) || sign_neg(aqq.re);
    const bool drop = tol > 0.0 && !diag && negligible(absz(aqp), absz(apq), app.re, aqq.re, tol);
    if ((!zeq0(aqp) || !zeq0(apq) || diag) && !drop) {
      dz a2[4] = {app, aqp, apq, aqq}, u2[4], z2[4];
      int j2[2] = {J[p0], J[q0]};
      const int info = zhsvd2<F>(a2, j2, u2, z2, TH1FIX);
      if (info < 0) key = KEY_NAN;
      else if ((info & 1) == 0) key = weight_key(0.0);
      else if ((info & 4) == 0) key = weight_key(dasum4<F>(aqp.re, aqp.im, apq.re, apq.im));
      else {
        ++nfull;
        const dz aq = (info & 8) ? a2[1] : zmk(0.0, 0.0), ap = (info & 8) ? a2[2] : zmk(0.0, 0.0);
        double d = fmad<F>(aqp.re - aq.re, aqp.re + aq.re, 0.0);
        d = fmad<F>(aqp.im - aq.im, aqp.im + aq.im, d);
        d = fmad<F>(apq.re - ap.re, apq.re + ap.re, d);
        d = fmad<F>(apq.im - ap.im, apq.im + ap.im, d);
        const double gxx = coln[p0] - (aqp.re * aqp.re + aqp.im * aqp.im);
        const double gyy = coln[q0] - (apq.re * apq.re + apq.im * apq.im);
        const size_t gi = (size_t)(i0 + a) * ldg + j0 + b, plane = (size_t)npl * nmi;
        double gr = Gre[gi], gi2 = Gim[gi];
        for (int z = 1; z < gks; ++z) {
          gr += Gre[gi + z * plane];
          gi2 += Gim[gi + z * plane];
        }
        const double gim = cp < cm ? gi2 : -gi2;
        const dz w11 = z2[0], w21 = z2[1], w12 = z2[2], w22 = z2[3];
        const double n11 = w11.re * w11.re + w11.im * w11.im, n12 = w12.re * w12.re + w12.im * w12.im;
        const double n21 = w21.re * w21.re + w21.im * w21.im, n22 = w22.re * w22.re + w22.im * w22.im;
        // s = conj(W11) W21 + conj(W12) W22; Re(s gxy)
        const double sr = w11.re * w21.re + w11.im * w21.im + w12.re * w22.re + w12.im * w22.im;
        const double si = w11.re * w21.im - w11.im * w21.re + w12.re * w22.im - w12.im * w22.re;
        const double w = d + gxx * (1.0 - n11 - n12) + gyy * (1.0 - n21 - n22) - 2.0 * (sr * gr - si * gim);
        key = (w < -1.7976931348623157e308 || !(w == w)) ? KEY_NAN : weight_key(w);
      }
    }
    const uint32_t k = pair_index(n, p0 + 1, q0 + 1);
    keys[k] = key_cut(key, cut);
  }
  for (int o = 32; o > 0; o >>= 1) nfull += __shfl_xor(nfull, o);
  __shared__ unsigned wfull[4];
  if ((tid & 63) == 0) wfull[tid >> 6] = nfull;
  __syncthreads();
  if (tid == 0 && wfull[0] + wfull[1] + wfull[2] + wfull[3])
    acc_abod_add(st, (unsigned long long)(wfull[0] + wfull[1] + wfull[2] + wfull[3]));
}

void zgram_hyp_weights_launch(DeviceStep &d, hipStream_t s) {
  const int ks = gram_ksplit(d.npl, d.nmi, 2 * d.n);
  double *Gre = d.gram, *Gim = d.gram + (size_t)ks * d.npl * d.nmi;
  dz *dg = (dz *)(d.coln + d.n);
  const int cut = d.prec32 ? d.cut32 : 0;
  dim3 g((d.nmi + ZWT - 1) / ZWT, (d.npl + ZWT - 1) / ZWT);
  auto weights = [&](auto kern, const auto *A) {
    hipLaunchKernelGGL(kern, g, dim3(256), 0, s, d.n, A, d.lda, d.J, d.plus, d.npl, d.minus, d.nmi, Gre, Gim, d.nmi,
                       ks, d.coln, dg, d.sort.k0, d.st, d.neg_tol, cut);
  };
  if (d.prec32) {
    const fz *A = (const fz *)d.A32;
    hipLaunchKernelGGL(k_zcolnorm_off<fz>, dim3((d.n + 3) / 4), dim3(256), 0, s, d.n, A, d.lda, d.coln, dg, d.st);
    hipLaunchKernelGGL(k_warm_offnorm<dz>, dim3(1), dim3(256), 0, s, d.n, d.coln, dg, d.st, d.switch_tol);
    if (d.npl <= 0 || d.nmi <= 0) return;
    const float *X = (const float *)d.A32;
    zgemm_tn_f32_launch(2 * d.n, X, 2 * d.lda, d.plus, d.npl, d.minus, d.nmi, Gre, (long)d.nmi, 1L, d.st, s, 1, ks);
    zgemm_tn_f32_launch(2 * d.n, X, 2 * d.lda, d.plus, d.npl, d.minus, d.nmi, Gim, (long)d.nmi, 1L, d.st, s, 2, ks);
    if (d.fused) weights(k_zhyp_weights_gram<true, fz>, A);
    else weights(k_zhyp_weights_gram<false, fz>, A);
    return;
  }
  if (d.npl <= 0 || d.nmi <= 0) return;
  const dz *A = (const dz *)d.A;
  hipLaunchKernelGGL(k_zcolnorm_off<dz>, dim3((d.n + 3) / 4), dim3(256), 0, s, d.n, A, d.lda, d.coln, dg, d.st);
  const double *X = (const double *)d.A;
  gemm_tn_launch(2 * d.n, X, 2 * d.lda, d.plus, d.npl, X, 2 * d.lda, d.minus, d.nmi, Gre, (long)d.nmi, 1L, true,
                 d.st, s, 1, ks);
  gemm_tn_launch(2 * d.n, X, 2 * d.lda, d.plus, d.npl, X, 2 * d.lda, d.minus, d.nmi, Gim, (long)d.nmi, 1L, true,
                 d.st, s, 2, ks);
  if (d.fused) weights(k_zhyp_weights_gram<true, dz>, A);
  else weights(k_zhyp_weights_gram<false, dz>, A);
}

}  // namespace jkg
