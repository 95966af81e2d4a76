// jkg_zstep.hip -- the complex (zjk) kernels of the dynamic-ordering step:
// weights (ZMAGF2T/ZMAGF2H, src/Z/zstep.F90:12-74, with DASUM4 and ABODNZF2,
// src/Z/ztransf.F90:18-75, 452-637), the batched 2x2 ZHSVD2 of the pivots,
// the complex row/column transforms BA/AB (src/Z/ztransf.F90:279-448) and the
// finalisation JZHJ (src/Z/zstep.F90:451-478).  Sort, selection, step begin
// and end are the real path's kernels (they only see keys, pairs and INFO).
// Operand orders of the complex multiply-adds are the reference's: ZZFMA is
// not symmetric in its first two arguments, so BA and AB keep their own forms.
#include "jkg_hsvd2.hpp"
#include "jkg_step.hpp"

namespace jkg {

// a complex ratio b/a as the reference forms it: real divisor -> ZDDIV
template <bool F>
__device__ __forceinline__ dz zratio(dz num, dz den) {
  return (den.im == 0.0) ? zddiv(num, den.re) : zzdiv<F>(num, den);
}

// Ratio-form descriptor of a complex 2x2 B for BA (rows) or AB (columns).
struct ZRot {
  dz r1, r2, m1, m2;
  int sa, sb, pad0, pad1;
};

// BA, src/Z/ztransf.F90:279-365 (rows: [x;y] <- B [x;y])
template <bool F>
__device__ __forceinline__ ZRot zrot_ba(const dz B[4]) {
  ZRot R;
  R.sa = absz(B[0]) >= absz(B[2]);
  R.r1 = R.sa ? zratio<F>(B[2], B[0]) : zratio<F>(B[0], B[2]);
  R.m1 = R.sa ? B[0] : B[2];
  R.sb = absz(B[3]) >= absz(B[1]);
  R.r2 = R.sb ? zratio<F>(B[1], B[3]) : zratio<F>(B[3], B[1]);
  R.m2 = R.sb ? B[3] : B[1];
  R.pad0 = R.pad1 = 0;
  return R;
}
template <bool F>
__device__ __forceinline__ void zrot_ba_apply(const ZRot &R, dz &x, dz &y) {
  dz xx, yy;
  if (R.sa) {
    xx = zzfma<F>(R.r1, y, x);
    // ztransf.F90:325: the |B22| < |B21| branch updates y with R1, as written
    yy = R.sb ? zzfma<F>(R.r2, x, y) : zzfma<F>(R.r1, y, x);
  } else {
    xx = zzfma<F>(R.r1, x, y);
    yy = R.sb ? zzfma<F>(R.r2, x, y) : zzfma<F>(R.r2, y, x);
  }
  x = zzmul<F>(xx, R.m1);
  y = zzmul<F>(yy, R.m2);
}

// AB, src/Z/ztransf.F90:369-448 (columns: [x y] <- [x y] B)
template <bool F>
__device__ __forceinline__ ZRot zrot_ab(const dz B[4]) {
  ZRot R;
  R.sa = absz(B[0]) >= absz(B[1]);
  R.r1 = R.sa ? zratio<F>(B[1], B[0]) : zratio<F>(B[0], B[1]);
  R.m1 = R.sa ? B[0] : B[1];
  R.sb = absz(B[3]) >= absz(B[2]);
  R.r2 = R.sb ? zratio<F>(B[2], B[3]) : zratio<F>(B[3], B[2]);
  R.m2 = R.sb ? B[3] : B[2];
  R.pad0 = R.pad1 = 0;
  return R;
}
template <bool F>
__device__ __forceinline__ void zrot_ab_apply(const ZRot &R, dz &x, dz &y) {
  const dz xx = R.sa ? zzfma<F>(y, R.r1, x) : zzfma<F>(x, R.r1, y);
  const dz yy = R.sb ? zzfma<F>(x, R.r2, y) : zzfma<F>(y, R.r2, x);
  x = zzmul<F>(xx, R.m1);
  y = zzmul<F>(yy, R.m2);
}

// The same transforms on FP32 data (the warm start): the descriptor is formed
// in FP64 and rounded once; the products run in FP32.
struct ZRotF {
  fz r1, r2, m1, m2;
  int sa, sb;
};
__device__ __forceinline__ fz tofz(dz a) { return fz{(float)a.re, (float)a.im}; }
__device__ __forceinline__ const ZRot &as_rot(const ZRot &R, const dz *) { return R; }
__device__ __forceinline__ ZRotF as_rot(const ZRot &R, const fz *) {
  return ZRotF{tofz(R.r1), tofz(R.r2), tofz(R.m1), tofz(R.m2), R.sa, R.sb};
}
template <bool F>
__device__ __forceinline__ fz zzfmaf(fz a, fz b, fz c) {
  float f = fmadf<F>(-a.im, b.im, c.re);
  const float r = fmadf<F>(a.re, b.re, f);
  f = fmadf<F>(a.im, b.re, c.im);
  return fz{r, fmadf<F>(a.re, b.im, f)};
}
template <bool F>
__device__ __forceinline__ fz zzmulf(fz a, fz b) {
  const float r = fmadf<F>(a.re, b.re, -(a.im * b.im));
  return fz{r, fmadf<F>(a.re, b.im, a.im * b.re)};
}
template <bool F>
__device__ __forceinline__ void zrot_ba_apply(const ZRotF &R, fz &x, fz &y) {
  fz xx, yy;
  if (R.sa) {
    xx = zzfmaf<F>(R.r1, y, x);
    yy = R.sb ? zzfmaf<F>(R.r2, x, y) : zzfmaf<F>(R.r1, y, x);
  } else {
    xx = zzfmaf<F>(R.r1, x, y);
    yy = R.sb ? zzfmaf<F>(R.r2, x, y) : zzfmaf<F>(R.r2, y, x);
  }
  x = zzmulf<F>(xx, R.m1);
  y = zzmulf<F>(yy, R.m2);
}
template <bool F>
__device__ __forceinline__ void zrot_ab_apply(const ZRotF &R, fz &x, fz &y) {
  const fz xx = R.sa ? zzfmaf<F>(y, R.r1, x) : zzfmaf<F>(x, R.r1, y);
  const fz yy = R.sb ? zzfmaf<F>(x, R.r2, y) : zzfmaf<F>(y, R.r2, x);
  x = zzmulf<F>(xx, R.m1);
  y = zzmulf<F>(yy, R.m2);
}

// ---------------------------------------------------------------- weights
// ZMAGF2T, src/Z/zstep.F90:12-39
template <bool F, class E>
__global__ void __launch_bounds__(256) k_ztrig_weights(int n, const E *__restrict__ A, int lda,
                                                        const int *__restrict__ J, uint64_t *__restrict__ keys,
                                                        const StepState *st, int nn,
                                                        double tol, int cut) {
  if (!st->active) return;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < (uint32_t)nn; k += gridDim.x * blockDim.x) {
    int p, q;
    pair_of(n, k, p, q);
    const int jp = J[p - 1], jq = J[q - 1];
    if (jp != jq) continue;
    const size_t P = p - 1, Q = q - 1;
    const dz aqp = ldz(A[P * lda + Q]), apq = ldz(A[Q * lda + P]), app = ldz(A[P * lda + P]),
             aqq = ldz(A[Q * lda + Q]);
    const bool off = !zeq0(aqp) || !zeq0(apq) || app.im != 0.0 || aqq.im != 0.0 || sign_neg(app.re) ||
                     sign_neg(aqq.re);
    const bool diag = app.im != 0.0 || aqq.im != 0.0 || sign_neg(app.re) || sign_neg(aqq.re);
    const bool order = (jp + jq == -2) ? (aqq.re < app.re) : (app.re < aqq.re);
    bool go = off || order;
    // tol > 0 (mixed solver only): the convergence test, as for the real pairs
    if (tol > 0.0 && go && !diag && !order && negligible(absz(aqp), absz(apq), app.re, aqq.re, tol)) go = false;
    keys[k] = key_cut(go ? weight_key(dasum4<F>(aqp.re, aqp.im, apq.re, apq.im)) : KEY_NAN, cut);
  }
}

// ZMAGF2H + ABODNZF2: tile of 16 (J=+1) x 16 (J=-1) columns, rows in LDS.
constexpr int ZHT = 16;
constexpr int ZHR = 32;
template <bool F>
__global__ void __launch_bounds__(256) k_zhyp_weights(int n, const dz *__restrict__ A, int lda,
                                                       const int *__restrict__ J, const int *__restrict__ plus,
                                                       int npl, const int *__restrict__ minus, int nmi,
                                                       uint64_t *__restrict__ keys,
                                                       StepState *st) {
  if (!st->active) return;
  __shared__ dz L[ZHR][2 * ZHT + 1];
  __shared__ int cols[2 * ZHT];
  __shared__ int nneed;
  const int a = threadIdx.x / ZHT, b = threadIdx.x % ZHT;
  const int ia = blockIdx.x * ZHT + a, ib = blockIdx.y * ZHT + b;
  if (threadIdx.x < 2 * ZHT) {
    const int c = threadIdx.x;
    const int g = (c < ZHT) ? blockIdx.x * ZHT + c : blockIdx.y * ZHT + (c - ZHT);
    cols[c] = (c < ZHT) ? (g < npl ? plus[g] : -1) : (g < nmi ? minus[g] : -1);
  }
  if (threadIdx.x == 0) nneed = 0;
  const bool valid = ia < npl && ib < nmi;
  const int cp = valid ? plus[ia] : 0, cm = valid ? minus[ib] : 1;
  const int p0 = cp < cm ? cp : cm, q0 = cp < cm ? cm : cp;
  const bool xplus = cp < cm;
  bool need = false;
  double d = 0.0;
  ZRot R;
  R.sa = R.sb = 1;
  R.r1 = R.r2 = R.m1 = R.m2 = zmk(0.0, 0.0);
  uint64_t key = KEY_NAN;
  if (valid) {
    const size_t P = p0, Q = q0;
    const dz app = A[P * lda + P], aqp = A[P * lda + Q], apq = A[Q * lda + P], aqq = A[Q * lda + Q];
    if (!zeq0(aqp) || !zeq0(apq) || app.im != 0.0 || aqq.im != 0.0 || sign_neg(app.re) || sign_neg(aqq.re)) {
      dz a2[4] = {app, aqp, apq, aqq}, u2[4], z2[4];
      int j2[2] = {J[p0], J[q0]};
      const int info = zhsvd2<F>(a2, j2, u2, z2, TH1FIX);
      if (info < 0) key = KEY_NAN;
      else if ((info & 1) == 0) key = weight_key(0.0);
      else if ((info & 4) == 0) key = weight_key(dasum4<F>(aqp.re, aqp.im, apq.re, apq.im));
      else {
        need = true;
        const dz aq = (info & 8) ? a2[1] : zmk(0.0, 0.0), ap = (info & 8) ? a2[2] : zmk(0.0, 0.0);
        d = fmad<F>(aqp.re - aq.re, aqp.re + aq.re, 0.0);
        d = fmad<F>(aqp.im - aq.im, aqp.im + aq.im, d);
        d = fmad<F>(apq.re - ap.re, apq.re + ap.re, d);
        d = fmad<F>(apq.im - ap.im, apq.im + ap.im, d);
        R = zrot_ab<F>(z2);
      }
    }
  }
  __syncthreads();
  if (need) atomicAdd(&nneed, 1);
  if (__syncthreads_or(need)) {
    if (threadIdx.x == 0) acc_abod_add(st, (unsigned long long)nneed);
    const int xc = xplus ? a : ZHT + b, yc = xplus ? ZHT + b : a;
    for (int r0 = 0; r0 < n; r0 += ZHR) {
      __syncthreads();
      for (int e = threadIdx.x; e < ZHR * 2 * ZHT; e += 256) {
        const int c = e / ZHR, r = e % ZHR, col = cols[c];
        L[r][c] = (col >= 0 && r0 + r < n) ? A[(size_t)col * lda + r0 + r] : zmk(0.0, 0.0);
      }
      __syncthreads();
      if (need) {
        const int rmax = (n - r0 < ZHR) ? n - r0 : ZHR;
        for (int r = 0; r < rmax; ++r) {
          const int row = r0 + r;
          if (row == p0 || row == q0) continue;
          const dz x = L[r][xc], y = L[r][yc];
          const dz xx = zzmul<F>(R.sa ? zzfma<F>(y, R.r1, x) : zzfma<F>(x, R.r1, y), R.m1);
          const dz yy = zzmul<F>(R.sb ? zzfma<F>(x, R.r2, y) : zzfma<F>(y, R.r2, x), R.m2);
          const double axi = absz(x), axx = absz(xx), ayi = absz(y), ayy = absz(yy);
          d = fmad<F>((axi - axx), (axi + axx), d);
          d = fmad<F>((ayi - ayy), (ayi + ayy), d);
        }
      }
    }
    if (need) key = (d < -1.7976931348623157e308) ? KEY_NAN : weight_key(d);
  }
  if (valid) {
    const uint32_t k = pair_index(n, p0 + 1, q0 + 1);
    keys[k] = key;
  }
}

// ---------------------------------------------------------------- pivots
template <bool F, class E>
__global__ void __launch_bounds__(256) k_zpivots(int n, const E *__restrict__ A, int lda, const int *__restrict__ J,
                                                  const int *__restrict__ piv_p, const int *__restrict__ piv_q,
                                                  const int *__restrict__ piv_pos, const uint64_t *keys,
                                                  double *piv_w, int *piv_info,
                                                  double *piv_d, ZRot *zrot, int *colpiv, StepState *st) {
  if (!st->active) return;
  const int sl = st->sl;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) st->t_transf = wall_clock64();
  if (i >= sl) return;
  const int p = piv_p[i], q = piv_q[i];
  const size_t P = p, Q = q;
  piv_w[i] = key_weight(keys[pair_index(n, p + 1, q + 1)]);  // the pivot's key, at its pair index
  dz b[4] = {ldz(A[P * lda + P]), ldz(A[P * lda + Q]), ldz(A[Q * lda + P]), ldz(A[Q * lda + Q])};
  int k[2] = {J[p], J[q]};
  dz v[4], w[4];
  const int info = zhsvd2<F>(b, k, v, w, TH1FIX);
  piv_info[i] = info;
  colpiv[p] = i;
  colpiv[q] = -(i + 2);
  if (info < 0) {
    atomicMin(&st->err, info);
    return;
  }
  piv_d[2 * i] = b[0].re;
  piv_d[2 * i + 1] = b[3].re;
  // descriptors: A rows (BA with V), U columns (AB with V^H, UH2), Z and A columns (AB with W)
  zrot[3 * i + 0] = zrot_ba<F>(v);
  const dz vh[4] = {zconj(v[0]), zconj(v[2]), zconj(v[1]), zconj(v[3])};
  zrot[3 * i + 1] = zrot_ab<F>(vh);
  zrot[3 * i + 2] = zrot_ab<F>(w);
  if (info & 2) st->nrow = 1;  // a flag: whether any pivot has a row transform
}

// ---------------------------------------------------------------- rotate
template <bool F, class E>
__global__ void __launch_bounds__(1024) k_zrotate(int n, E *__restrict__ A, int lda, E *__restrict__ U,
                                                 E *__restrict__ Z, const int *__restrict__ piv_p,
                                                 const int *__restrict__ piv_q, const int *__restrict__ piv_info,
                                                 const double *__restrict__ piv_d, const ZRot *__restrict__ zrot,
                                                 const int *__restrict__ colpiv, const StepState *st) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lz_raw[];
  E *LZ = (E *)lz_raw;
  const E *tag = nullptr;  // selects the descriptor's precision
  if (!st->active) return;
  const int sl = st->sl;
  if (sl <= 0) return;
  const int c = blockIdx.x;
  const int j = colpiv[c];
  if (j <= -2) return;
  const int err = st->err, nrow = st->nrow;
  const int info = (j >= 0) ? piv_info[j] : -1;
  const int c1 = (j >= 0) ? piv_q[j] : -1;
  const bool col_rot = (j >= 0) && err == 0 && info >= 0 && (info & 4);
  const bool diag_set = (j >= 0) && err == 0 && info >= 0 && !(info & 8);
  const int tid = threadIdx.x, bs = blockDim.x;
  if (nrow > 0 || col_rot || diag_set) {
    E *col0 = A + (size_t)c * lda;
    E *col1 = (c1 >= 0) ? A + (size_t)c1 * lda : nullptr;
    for (int r = tid; r < n; r += bs) {
      LZ[r] = col0[r];
      if (col1) LZ[n + r] = col1[r];
    }
    __syncthreads();
    if (nrow > 0) {
      for (int i = tid; i < sl; i += bs) {
        const int inf = piv_info[i];
        if (inf < 0 || !(inf & 2)) continue;
        const auto R = as_rot(zrot[3 * i + 0], tag);
        const int p = piv_p[i], q = piv_q[i];
        E x = LZ[p], y = LZ[q];
        zrot_ba_apply<F>(R, x, y);
        LZ[p] = x;
        LZ[q] = y;
        if (col1) {
          x = LZ[n + p];
          y = LZ[n + q];
          zrot_ba_apply<F>(R, x, y);
          LZ[n + p] = x;
          LZ[n + q] = y;
        }
      }
      __syncthreads();
    }
    if (col1) {
      const int p = c, q = c1;
      const auto R = as_rot(zrot[3 * (col_rot ? j : 0) + 2], tag);  // read only if col_rot
      const double d0 = diag_set ? piv_d[2 * j] : 0.0, d1 = diag_set ? piv_d[2 * j + 1] : 0.0;
      for (int r = tid; r < n; r += bs) {
        E x = LZ[r], y = LZ[n + r];
        if (col_rot) zrot_ab_apply<F>(R, x, y);
        if (diag_set) {
          if (r == p) { x = emk<E>(d0, 0.0); y = emk<E>(0.0, 0.0); }
          if (r == q) { x = emk<E>(0.0, 0.0); y = emk<E>(d1, 0.0); }
        }
        col0[r] = x;
        col1[r] = y;
      }
    } else {
      for (int r = tid; r < n; r += bs) col0[r] = LZ[r];
    }
  }
  if (j >= 0 && info >= 0) {
    if (info & 2) {
      const auto R = as_rot(zrot[3 * j + 1], tag);
      E *u0 = U + (size_t)c * lda, *u1 = U + (size_t)c1 * lda;
      for (int r = tid; r < n; r += bs) {
        E x = u0[r], y = u1[r];
        zrot_ab_apply<F>(R, x, y);
        u0[r] = x;
        u1[r] = y;
      }
    }
    if (info & 4) {
      const auto R = as_rot(zrot[3 * j + 2], tag);
      E *z0 = Z + (size_t)c * lda, *z1 = Z + (size_t)c1 * lda;
      for (int r = tid; r < n; r += bs) {
        E x = z0[r], y = z1[r];
        zrot_ab_apply<F>(R, x, y);
        z0[r] = x;
        z1[r] = y;
      }
    }
  }
}

// ---------------------------------------------------------------- setup / finalize
__global__ void k_zidentity(dz *U, dz *Z, int n, int lda) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t c = e / n, r = e % n;
    const dz v = zmk((r == c) ? 1.0 : 0.0, 0.0);
    U[c * lda + r] = v;
    Z[c * lda + r] = v;
  }
}

// SIGMA = REAL(diag(A)), A(i,i) = 0, Z <- J Z^H J (diagonal conjugated)
__global__ void k_zfinalize(dz *A, dz *Z, int n, int lda, const int *J, double *sigma) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / n), r = (int)(e % n);
    if (r == c) {
      sigma[r] = A[(size_t)c * lda + r].re;
      A[(size_t)c * lda + r] = zmk(0.0, 0.0);
      Z[(size_t)c * lda + r].im = -Z[(size_t)c * lda + r].im;
    } else if (r < c) {
      dz *zrc = &Z[(size_t)c * lda + r], *zcr = &Z[(size_t)r * lda + c];
      const dz t = *zrc;
      dz a = zconj(*zcr), b2 = zconj(t);
      if (J[r] != J[c]) {
        a = zmk(-a.re, -a.im);
        b2 = zmk(-b2.re, -b2.im);
      }
      *zrc = a;
      *zcr = b2;
    }
  }
}

static inline int zgrid(size_t work, int bs, int cap) {
  size_t g = (work + bs - 1) / bs;
  if (g < 1) g = 1;
  return (int)((g > (size_t)cap) ? cap : g);
}

size_t zrot_bytes() { return sizeof(ZRot); }

void zweights_launch(DeviceStep &d, hipStream_t s) {
  const double tol = d.wmode ? d.neg_tol : 0.0;
  const dim3 gt(zgrid(d.nn, 256, 65536));
  auto trig = [&](auto kern, const auto *A) {
    hipLaunchKernelGGL(kern, gt, dim3(256), 0, s, d.n, A, d.lda, d.J, d.sort.k0, d.st, d.nn, tol,
                       d.prec32 ? d.cut32 : 0);
  };
  if (d.prec32) {
    const fz *A = (const fz *)d.A32;
    if (d.fused) trig(k_ztrig_weights<true, fz>, A);
    else trig(k_ztrig_weights<false, fz>, A);
  } else {
    const dz *A = (const dz *)d.A;
    if (d.fused) trig(k_ztrig_weights<true, dz>, A);
    else trig(k_ztrig_weights<false, dz>, A);
  }
  if (d.wmode == 1) {
    zgram_hyp_weights_launch(d, s);
    return;
  }
  const dz *A = (const dz *)d.A;
  if (d.npl > 0 && d.nmi > 0) {
    dim3 g((d.npl + ZHT - 1) / ZHT, (d.nmi + ZHT - 1) / ZHT);
    if (d.fused)
      hipLaunchKernelGGL(k_zhyp_weights<true>, g, dim3(256), 0, s, d.n, A, d.lda, d.J, d.plus, d.npl, d.minus,
                         d.nmi, d.sort.k0, d.st);
    else
      hipLaunchKernelGGL(k_zhyp_weights<false>, g, dim3(256), 0, s, d.n, A, d.lda, d.J, d.plus, d.npl, d.minus,
                         d.nmi, d.sort.k0, d.st);
  }
}

void zpivots_launch(DeviceStep &d, hipStream_t s) {
  const int g = (d.n2max + 255) / 256;
  auto go = [&](auto kern, const auto *A) {
    hipLaunchKernelGGL(kern, dim3(g), dim3(256), 0, s, d.n, A, d.lda, d.J, d.piv_p, d.piv_q, d.piv_pos, d.sort.k0,
                       d.piv_w, d.piv_info, d.piv_d, (ZRot *)d.zrot,
                       d.colpiv, d.st);
  };
  if (d.prec32) {
    if (d.fused) go(k_zpivots<true, fz>, (const fz *)d.A32);
    else go(k_zpivots<false, fz>, (const fz *)d.A32);
  } else {
    if (d.fused) go(k_zpivots<true, dz>, (const dz *)d.A);
    else go(k_zpivots<false, dz>, (const dz *)d.A);
  }
}

void zrotate_launch(DeviceStep &d, hipStream_t s) {
  auto go = [&](auto kern, auto *A, auto *U, auto *Z, size_t esz) {
    hipLaunchKernelGGL(kern, dim3(d.n), dim3(1024), 2 * d.n * esz, s, d.n, A, d.lda, U, Z, d.piv_p, d.piv_q,
                       d.piv_info, d.piv_d, (const ZRot *)d.zrot, d.colpiv, d.st);
  };
  if (d.prec32) {
    fz *A = (fz *)d.A32, *U = (fz *)d.U32, *Z = (fz *)d.Z32;
    if (d.fused) go(k_zrotate<true, fz>, A, U, Z, sizeof(fz));
    else go(k_zrotate<false, fz>, A, U, Z, sizeof(fz));
  } else {
    dz *A = (dz *)d.A, *U = (dz *)d.U, *Z = (dz *)d.Z;
    if (d.fused) go(k_zrotate<true, dz>, A, U, Z, sizeof(dz));
    else go(k_zrotate<false, dz>, A, U, Z, sizeof(dz));
  }
}

int zstep_prepare(DeviceStep &d) {
  const size_t lds = (size_t)2 * d.n * sizeof(dz);
  if (lds > 160 * 1024) return -1;
  const void *k[4] = {(const void *)k_zrotate<true, dz>, (const void *)k_zrotate<false, dz>,
                      (const void *)k_zrotate<true, fz>, (const void *)k_zrotate<false, fz>};
  for (int i = 0; i < 4; ++i)
    if (hipFuncSetAttribute(k[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)(i < 2 ? lds : lds / 2)) !=
        hipSuccess)
      return -2;
  return 0;
}

void zinit_identity_launch(DeviceStep &d, hipStream_t s) {
  hipLaunchKernelGGL(k_zidentity, dim3(zgrid((size_t)d.n * d.n, 256, 8192)), dim3(256), 0, s, (dz *)d.U, (dz *)d.Z,
                     d.n, d.lda);
}

void zfinalize_launch(DeviceStep &d, hipStream_t s, double *sigma) {
  hipLaunchKernelGGL(k_zfinalize, dim3(zgrid((size_t)d.n * d.n, 256, 8192)), dim3(256), 0, s, (dz *)d.A, (dz *)d.Z,
                     d.n, d.lda, d.J, sigma);
}

}  // namespace jkg
