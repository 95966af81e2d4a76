// jkg_batch.hip -- many independent HSVDs of different sizes advanced by one
// persistent launch per step.
//
// Every instance follows the plain solver's arithmetic (the reference's
// DSTEP_EXEC, src/D/dstep.F90:376-443, with DSTEP_LOOP's termination), so each
// instance's YU/ZZ/SY equal what djk.exe computes for it alone -- unless the
// caller asks for the mixed solver's convergence test (tol > 0), which drops
// the pairs whose off-diagonal entries are negligible against the diagonal.  One
// workgroup runs a whole step of one instance -- weights (DMAGF2T/DMAGF2H +
// ABODNDF2), an LDS-ranked radix sort of the keys, the greedy walk (AW_NCP2),
// the 2x2 solves (DHSVD2) and the fused row/column rotations -- with
// workgroup barriers only; the workgroups of the launch loop over the
// instances the host scheduled to them (largest cost first).
#include "jkg_batch.hpp"
#include "jkg_device.hpp"
#include "jkg_hsvd2.hpp"
#include "jkg_rot.hpp"

namespace jkg {

constexpr int BT = 256;      // lanes per workgroup
constexpr int BW = BT / 64;  // waves
constexpr int BHT = 16;      // hyperbolic weight tile: 16 (J = +1) x 16 (J = -1) columns
constexpr int BMAXN = 1024;  // largest instance

struct BLds {
  double L[64][2 * BHT + 1];  // weights: a chunk of rows of the tile's 32 columns
  int cols[2 * BHT];
  int hist[256], off[256], wcnt[BW][256];
  uint32_t slot[2][2 * BW];
  double colbuf[2 * BMAXN];  // rotations: a column and its partner
  uint8_t used[BMAXN + 1];
  int flag, it, sl, err, nrow, cnt[5];
};

__device__ __forceinline__ void bsync() { __syncthreads(); }

// ------------------------------------------------------------------ (a) weights
// The trigonometric pairs: DMAGF2T (src/D/dstep.F90:12-37).
template <bool F>
__device__ void b_trig_weights(int n, const double *A, const int *J, uint64_t *keys, uint32_t *vals, double tol) {
  const int nn = n * (n - 1) / 2, lda = n;
  for (int k = threadIdx.x; k < nn; k += BT) {
    int p, q;
    pair_of(n, (uint32_t)k, p, q);
    const int jp = J[p - 1], jq = J[q - 1];
    if (jp != jq) continue;
    const size_t P = p - 1, Q = q - 1;
    const double aqp = A[P * lda + Q], apq = A[Q * lda + P], app = A[P * lda + P], aqq = A[Q * lda + Q];
    const bool order = sign_neg(app) || sign_neg(aqq) || ((jp + jq == -2) ? (aqq < app) : (app < aqq));
    bool go = (aqp != 0.0) || (apq != 0.0) || order;
    if (tol > 0.0 && go && !order && negligible(aqp, apq, app, aqq, tol)) go = false;
    keys[k] = go ? weight_key(dasum2<F>(aqp, apq)) : KEY_NAN;
    vals[k] = pack_pq(p - 1, q - 1);
  }
}

// The hyperbolic pairs of one 16 x 16 column tile: DMAGF2H + ABODNDF2
// (src/D/dstep.F90:41-72, src/D/dtransf.F90:246-348), rows in the
// reference's order, staged through LDS HR at a time.
template <bool F, int HR>
__device__ void b_hyp_tile(int n, const double *A, const int *J, const int *plus, int npl, const int *minus, int nmi,
                           int tx, int ty, uint64_t *keys, uint32_t *vals, BLds &S, double tol) {
  const int lda = n;
  const int a = threadIdx.x / BHT, b = threadIdx.x % BHT;
  const int ia = tx * BHT + a, ib = ty * BHT + b;
  if (threadIdx.x < 2 * BHT) {
    const int c = threadIdx.x;
    const int g = (c < BHT) ? tx * BHT + c : ty * BHT + (c - BHT);
    S.cols[c] = (c < BHT) ? (g < npl ? plus[g] : -1) : (g < nmi ? minus[g] : -1);
  }
  const bool valid = ia < npl && ib < nmi;
  const int cp = valid ? plus[ia] : 0, cm = valid ? minus[ib] : 1;
  const int p0 = cp < cm ? cp : cm, q0 = cp < cm ? cm : cp;
  const bool xplus = cp < cm;
  bool need = false;
  double d = 0.0, r1 = 0.0, r2 = 0.0, m1 = 0.0, m2 = 0.0;
  bool sel1 = true, sel2 = true;
  uint64_t key = KEY_NAN;
  if (valid) {
    const size_t P = p0, Q = q0;
    const double app = A[P * lda + P], aqp = A[P * lda + Q], apq = A[Q * lda + P], aqq = A[Q * lda + Q];
    const bool order = sign_neg(app) || sign_neg(aqq);
    const bool drop = tol > 0.0 && !order && negligible(aqp, apq, app, aqq, tol);
    if ((aqp != 0.0 || apq != 0.0 || order) && !drop) {
      double a2[4] = {app, aqp, apq, aqq}, u2[4], z2[4];
      int j2[2] = {J[p0], J[q0]};
      const int info = dhsvd2<F>(a2, j2, u2, z2, TH1FIX);
      if (info < 0) key = KEY_NAN;
      else if ((info & 1) == 0) key = weight_key(0.0);
      else if ((info & 4) == 0) key = weight_key(dasum2<F>(aqp, apq));
      else {
        need = true;
        const double aq = (info & 8) ? a2[1] : 0.0, ap = (info & 8) ? a2[2] : 0.0;
        d = fmad<F>(aqp - aq, aqp + aq, 0.0);
        d = fmad<F>(apq - ap, apq + ap, d);
        sel1 = __builtin_fabs(z2[0]) >= __builtin_fabs(z2[1]);
        r1 = sel1 ? z2[1] / z2[0] : z2[0] / z2[1];
        m1 = sel1 ? z2[0] : z2[1];
        sel2 = __builtin_fabs(z2[3]) >= __builtin_fabs(z2[2]);
        r2 = sel2 ? z2[2] / z2[3] : z2[3] / z2[2];
        m2 = sel2 ? z2[3] : z2[2];
      }
    }
  }
  if (__syncthreads_or(need)) {
    const int xc = xplus ? a : BHT + b, yc = xplus ? BHT + b : a;
    const int s1c = sel1 ? yc : xc, t1c = sel1 ? xc : yc;
    const bool eq = (sel1 == sel2);
    for (int r0 = 0; r0 < n; r0 += HR) {
      bsync();
      for (int e = threadIdx.x; e < HR * 2 * BHT; e += BT) {
        const int c = e / HR, r = e % HR, col = S.cols[c];
        S.L[r][c] = (col >= 0 && r0 + r < n) ? A[(size_t)col * lda + r0 + r] : 0.0;
      }
      bsync();
      if (need) {
        const int rmax = (n - r0 < HR) ? n - r0 : HR;
        for (int r = 0; r < rmax; ++r) {
          const double x = S.L[r][xc], y = S.L[r][yc], s1 = S.L[r][s1c], t1 = S.L[r][t1c];
          const double s2 = eq ? t1 : s1, t2 = eq ? s1 : t1;
          const double xx = fmad<F>(s1, r1, t1) * m1;
          const double yy = fmad<F>(s2, r2, t2) * m2;
          double dn = fmad<F>(x - xx, x + xx, d);
          dn = fmad<F>(y - yy, y + yy, dn);
          const int row = r0 + r;
          d = (row == p0 || row == q0) ? d : dn;
        }
      }
    }
    if (need) key = (d < -1.7976931348623157e308) ? KEY_NAN : weight_key(d);  // -inf removed
  }
  bsync();  // S.cols and S.L are reused by the next tile
  if (valid) {
    const uint32_t k = pair_index(n, p0 + 1, q0 + 1);
    keys[k] = key;
    vals[k] = pack_pq(p0, q0);
  }
}

// ------------------------------------------------------------------ (b) sort
// Stable LSD radix sort of (key, value) by key, 8 passes of 8 bits; a pass
// whose digit is the same for every key is skipped.  Returns the buffer (0 or
// 1) that holds the result.
__device__ int b_sort(int nk, uint64_t *k0, uint64_t *k1, uint32_t *v0, uint32_t *v1, BLds &S) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int cur = 0;
  for (int pass = 0; pass < 8; ++pass) {
    const int shift = pass * 8;
    const uint64_t *kin = cur ? k1 : k0;
    const uint32_t *vin = cur ? v1 : v0;
    uint64_t *kout = cur ? k0 : k1;
    uint32_t *vout = cur ? v0 : v1;
    S.hist[tid] = 0;
    bsync();
    for (int e = tid; e < nk; e += BT) atomicAdd(&S.hist[(int)((kin[e] >> shift) & 0xFF)], 1);
    bsync();
    if (tid == 0) S.flag = 0;
    bsync();
    if (S.hist[tid] == nk) S.flag = 1;  // identity pass
    bsync();
    if (S.flag) continue;
    // exclusive scan of the histogram (one wave per 64 bins, then the carries)
    {
      int v = S.hist[tid];
      int incl = v;
      for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
      }
      S.off[tid] = incl - v;
      if (lane == 63) S.wcnt[0][wave] = incl;
      bsync();
      int carry = 0;
      for (int w = 0; w < wave; ++w) carry += S.wcnt[0][w];
      S.off[tid] += carry;
      bsync();
    }
    for (int base = 0; base < nk; base += BT) {
      for (int i = tid; i < BW * 256; i += BT) (&S.wcnt[0][0])[i] = 0;
      bsync();
      const int idx = base + tid;
      const bool valid = idx < nk;
      const uint64_t key = valid ? kin[idx] : 0;
      const uint32_t val = valid ? vin[idx] : 0;
      const int d = (int)((key >> shift) & 0xFF);
      uint64_t match = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint64_t bal = __ballot(valid && ((d >> b) & 1));
        match &= ((d >> b) & 1) ? bal : ~bal;
      }
      const int wrank = __popcll(match & lt_mask);
      if (valid && wrank == 0) S.wcnt[wave][d] = __popcll(match);
      bsync();
      if (valid) {
        int before = 0;
        for (int w = 0; w < wave; ++w) before += S.wcnt[w][d];
        const int pos = S.off[d] + before + wrank;
        kout[pos] = key;
        vout[pos] = val;
      }
      bsync();
      {
        int add = 0;
        for (int w = 0; w < BW; ++w) add += S.wcnt[w][tid];
        S.off[tid] += add;
      }
      bsync();  // wcnt is cleared by the next round
    }
    cur ^= 1;
  }
  return cur;
}

// ------------------------------------------------------------------ (c) select
// The greedy walk over vals[0, len) (AW_NCP2, src/shared/atypes.F90:430-484):
// chunks of BT * WI entries tested against used[] once, then one acceptance
// per iteration (ballot-found wave candidates, one LDS round trip, packed
// 16-bit collision test), as the single-problem walk does.
template <int WI>
__device__ int b_walk(int n, int n2, const uint32_t *vals, int len, int *piv_p, int *piv_q, int *piv_pos, BLds &S) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int CH = BT * WI;
  for (int i = tid; i <= n; i += BT) S.used[i] = (i == n);
  bsync();
  int acc = 0, par = 0;
  for (int c = 0; c < len && acc < n2; c += CH) {
    uint32_t vv[WI], alive = 0;
#pragma unroll
    for (int i = 0; i < WI; ++i) {
      const int e = c + tid * WI + i;
      vv[i] = (e < len) ? vals[e] : 0xFFFFFFFFu;
      const uint32_t p = min(vv[i] & 0xFFFFu, (uint32_t)n), q = min(vv[i] >> 16, (uint32_t)n);
      alive |= ((S.used[p] | S.used[q]) == 0 ? 1u : 0u) << i;
    }
    for (;;) {
      const int fi = __ffs(alive) - 1;
      uint32_t fv = vv[0];
#pragma unroll
      for (int i = 1; i < WI; ++i) fv = (fi == i) ? vv[i] : fv;
      const uint64_t bal = __ballot(alive != 0);
      uint32_t widx = 0x7FFFFFFFu, wval = 0;
      if (bal != 0) {
        const int L = __ffsll((unsigned long long)bal) - 1;
        widx = (uint32_t)((wave * 64 + L) * WI + __builtin_amdgcn_readlane(fi, L));
        wval = __builtin_amdgcn_readlane(fv, L);
      }
      if (lane == 0) {
        S.slot[par][wave] = widx;
        S.slot[par][BW + wave] = wval;
      }
      bsync();
      uint32_t g = S.slot[par][BW - 1], gv = S.slot[par][2 * BW - 1];
#pragma unroll
      for (int w = BW - 2; w >= 0; --w) {
        const uint32_t si = S.slot[par][w];
        const bool ok = si != 0x7FFFFFFFu;
        g = ok ? si : g;
        gv = ok ? S.slot[par][BW + w] : gv;
      }
      par ^= 1;
      if (g == 0x7FFFFFFFu) break;
      const uint32_t pg = pq_p(gv), qg = pq_q(gv);
      if (tid == 0) {
        S.used[pg] = 1;
        S.used[qg] = 1;
      } else if (tid == 64) {
        piv_p[acc] = (int)pg;
        piv_q[acc] = (int)qg;
        piv_pos[acc] = c + (int)g + 1;
      }
      ++acc;
      if (acc >= n2) break;
      if (bal == 0) continue;
      const uint32_t P = pg * 0x10001u, Q = qg * 0x10001u;
      uint32_t kill = 0;
#pragma unroll
      for (int i = 0; i < WI; ++i) {
        const uint32_t x = vv[i] ^ P, y = vv[i] ^ Q;
        const uint32_t hz = (((x - 0x10001u) & ~x) | ((y - 0x10001u) & ~y)) & 0x80008000u;
        kill |= (hz != 0 ? 1u : 0u) << i;
      }
      alive &= ~kill;
    }
  }
  bsync();
  return acc;
}

// ------------------------------------------------------------------ the step
template <bool F, int HR, int WI>
__device__ void b_step(const BatchArgs &a, int idx, BLds &S) {
  const BInst I = a.inst[idx];
  BState *st = a.state + idx;
  const int tid = threadIdx.x;
  const int n = I.n, lda = n, nn = n * (n - 1) / 2;
  double *A = a.mat + I.a, *U = a.mat + I.u, *Z = a.mat + I.z;
  const int *J = a.ints + I.j, *plus = a.ints + I.pm, *minus = plus + I.npl;
  int *colpiv = a.ints + I.cp;
  uint64_t *k0 = a.keys + I.k, *k1 = a.keys + a.pair_total + I.k;
  uint32_t *v0 = a.vals + I.k, *v1 = a.vals + a.pair_total + I.k;
  int *piv_p = a.ints + I.pv, *piv_q = piv_p + I.n2, *piv_pos = piv_q + I.n2, *piv_info = piv_pos + I.n2;
  double *piv_w = a.dbl + I.pw, *piv_d = piv_w + I.n2;
  Rot *rot_row = a.rot + I.pr, *rot_col = rot_row + I.n2;
  const int s = st->step;
  for (int r = tid; r < n; r += BT) colpiv[r] = -1;
  // (a) weights
  b_trig_weights<F>(n, A, J, k0, v0, a.tol);
  const int ntx = (I.npl + BHT - 1) / BHT, nty = (I.nmi + BHT - 1) / BHT;
  if (I.npl > 0 && I.nmi > 0)
    for (int t = 0; t < ntx * nty; ++t)
      b_hyp_tile<F, HR>(n, A, J, plus, I.npl, minus, I.nmi, t % ntx, t / ntx, k0, v0, S, a.tol);
  bsync();
  // (b) sort
  const int cur = b_sort(nn, k0, k1, v0, v1, S);
  const uint64_t *keys = cur ? k1 : k0;
  const uint32_t *vals = cur ? v1 : v0;
  // (c) IT and the walk
  if (tid == 0) {
    int lo = 0, hi = nn;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (keys[mid] == KEY_NAN) hi = mid;
      else lo = mid + 1;
    }
    S.it = lo;
    S.err = 0;
    S.nrow = 0;
    for (int i = 0; i < 5; ++i) S.cnt[i] = 0;
  }
  bsync();
  const int sl = b_walk<WI>(n, I.n2, vals, S.it, piv_p, piv_q, piv_pos, S);
  if (sl <= 0) {
    if (tid == 0) {
      st->done = 1;
      st->info = s;
      st->steps = s + 1;
      atomicAdd(a.ndone, 1);
    }
    return;
  }
  // (d) the 2x2 of every pivot (DSTEP_TRANSF phase 1 parameters)
  for (int i = tid; i < sl; i += BT) {
    const int p = piv_p[i], q = piv_q[i];
    const size_t P = p, Q = q;
    piv_w[i] = key_weight(keys[piv_pos[i] - 1]);
    double bb[4] = {A[P * lda + P], A[P * lda + Q], A[Q * lda + P], A[Q * lda + Q]};
    int jj[2] = {J[p], J[q]};
    double v[4], w[4];
    const int info = dhsvd2<F>(bb, jj, v, w, TH1FIX);
    piv_info[i] = info;
    colpiv[p] = i;
    colpiv[q] = -(i + 2);
    if (info < 0) {
      atomicMin(&S.err, info);
      continue;
    }
    piv_d[2 * i] = bb[0];
    piv_d[2 * i + 1] = bb[3];
    rot_row[i] = rot_ba(v);
    rot_col[i] = rot_ab(w);
    if (info & 2) atomicAdd(&S.nrow, 1);
  }
  bsync();
  const int err = S.err, nrow = S.nrow;
  // (e) rotations: per column (pair) all row transforms, then the column transform (k_rotate's order)
  double *L = S.colbuf;
  for (int c = 0; c < n; ++c) {
    const int j = colpiv[c];
    if (j <= -2) continue;
    const int info = (j >= 0) ? piv_info[j] : -1;
    const int c1 = (j >= 0) ? piv_q[j] : -1;
    const bool col_rot = (j >= 0) && err == 0 && info >= 0 && (info & 4);
    const bool diag_set = (j >= 0) && err == 0 && info >= 0 && !(info & 8);
    if (nrow > 0 || col_rot || diag_set) {
      double *col0 = A + (size_t)c * lda;
      double *col1 = (c1 >= 0) ? A + (size_t)c1 * lda : nullptr;
      for (int r = tid; r < n; r += BT) {
        L[r] = col0[r];
        if (col1) L[n + r] = col1[r];
      }
      bsync();
      if (nrow > 0) {
        for (int i = tid; i < sl; i += BT) {
          const int inf = piv_info[i];
          if (inf < 0 || !(inf & 2)) continue;
          const Rot R = rot_row[i];
          const int p = piv_p[i], q = piv_q[i];
          double x = L[p], y = L[q];
          rot_apply<F>(R, x, y);
          L[p] = x;
          L[q] = y;
          if (col1) {
            x = L[n + p];
            y = L[n + q];
            rot_apply<F>(R, x, y);
            L[n + p] = x;
            L[n + q] = y;
          }
        }
        bsync();
      }
      if (col1) {
        const int p = c, q = c1;
        Rot R;
        if (col_rot) R = rot_col[j];
        const double d0 = diag_set ? piv_d[2 * j] : 0.0, d1 = diag_set ? piv_d[2 * j + 1] : 0.0;
        for (int r = tid; r < n; r += BT) {
          double x = L[r], y = L[n + r];
          if (col_rot) rot_apply<F>(R, x, y);
          if (diag_set) {
            if (r == p) {
              x = d0;
              y = 0.0;
            }
            if (r == q) {
              x = 0.0;
              y = d1;
            }
          }
          col0[r] = x;
          col1[r] = y;
        }
      } else {
        for (int r = tid; r < n; r += BT) col0[r] = L[r];
      }
    }
    if (j >= 0 && info >= 0) {
      if (info & 2) {
        const Rot R = rot_row[j];
        double *u0 = U + (size_t)c * lda, *u1 = U + (size_t)c1 * lda;
        for (int r = tid; r < n; r += BT) {
          double x = u0[r], y = u1[r];
          rot_apply<F>(R, x, y);
          u0[r] = x;
          u1[r] = y;
        }
      }
      if (info & 4) {
        const Rot R = rot_col[j];
        double *z0 = Z + (size_t)c * lda, *z1 = Z + (size_t)c1 * lda;
        for (int r = tid; r < n; r += BT) {
          double x = z0[r], y = z1[r];
          rot_apply<F>(R, x, y);
          z0[r] = x;
          z1[r] = y;
        }
      }
    }
    bsync();  // L is reused by the next column
  }
  // (f) step end: NEWLEN (DSTEP_TRANSF), SL = MIN(SL, NL), the loop test
  {
    int nl = 0;
    for (int i = tid; i < sl; i += BT) {
      const int l = piv_info[i];
      if ((l & 1) && ((l & 2) || (l & 4))) ++nl;
    }
    atomicAdd(&S.cnt[0], nl);
  }
  bsync();
  if (tid == 0) {
    const int nl = (err == 0) ? S.cnt[0] : err;
    const int sl2 = (nl < sl) ? nl : sl;
    st->steps = s + 1;
    if (sl2 <= 0) {
      st->done = 1;
      st->info = (sl2 < 0) ? sl2 : s;
      atomicAdd(a.ndone, 1);
    } else {
      st->step = s + 1;
    }
  }
}

template <bool F, int HR, int WI>
__global__ void __launch_bounds__(BT) k_batch_step(BatchArgs a) {
  __shared__ BLds S;
  const int s0 = a.sched_off[blockIdx.x], s1 = a.sched_off[blockIdx.x + 1];
  for (int s = s0; s < s1; ++s) {
    const int idx = a.sched[s];
    BState *st = a.state + idx;
    if (threadIdx.x == 0) {
      int act = !st->done;
      if (act && st->max_steps >= 0 && st->step >= st->max_steps) {
        st->done = 1;
        st->info = st->step;
        atomicAdd(a.ndone, 1);
        act = 0;
      }
      S.flag = act;
    }
    bsync();
    const int act = S.flag;
    bsync();
    if (!act) continue;
    b_step<F, HR, WI>(a, idx, S);
    bsync();
  }
}

// SY = diag(A), A's diagonal zeroed, Z := J Z^T J in place (k_finalize's work), one workgroup per instance.
__global__ void __launch_bounds__(BT) k_batch_finalize(BatchArgs a, double *sy) {
  const int idx = blockIdx.x;
  if (idx >= a.count) return;
  const BInst I = a.inst[idx];
  const int n = I.n;
  double *A = a.mat + I.a, *Z = a.mat + I.z;
  const int *J = a.ints + I.j;
  for (int e = threadIdx.x; e < n * n; e += BT) {
    const int c = e / n, r = e % n;
    if (r == c) {
      sy[I.sy + r] = A[(size_t)c * n + r];
      A[(size_t)c * n + r] = 0.0;
    } else if (r < c) {
      double *zrc = &Z[(size_t)c * n + r], *zcr = &Z[(size_t)r * n + c];
      const double t = *zrc;
      double x = *zcr, y = t;
      if (J[r] != J[c]) {
        x = -x;
        y = -y;
      }
      *zrc = x;
      *zcr = y;
    }
  }
}

// U = Z = I for every instance.
__global__ void k_batch_identity(BatchArgs a) {
  const int idx = blockIdx.y;
  if (idx >= a.count) return;
  const BInst I = a.inst[idx];
  const int n = I.n;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x) {
    const double v = (e / n == e % n) ? 1.0 : 0.0;
    a.mat[I.u + e] = v;
    a.mat[I.z + e] = v;
  }
}

// ------------------------------------------------------------------ launchers
int batch_variants() { return 4; }

void batch_step_launch(const BatchArgs &a, int grid, int variant, int fused, hipStream_t s) {
  // variants: (rows staged per LDS chunk of the weight tile, walk entries per lane)
  switch (variant * 2 + (fused ? 1 : 0)) {
    case 0: hipLaunchKernelGGL((k_batch_step<false, 64, 4>), dim3(grid), dim3(BT), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_batch_step<true, 64, 4>), dim3(grid), dim3(BT), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_batch_step<false, 32, 4>), dim3(grid), dim3(BT), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_batch_step<true, 32, 4>), dim3(grid), dim3(BT), 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_batch_step<false, 64, 8>), dim3(grid), dim3(BT), 0, s, a); break;
    case 5: hipLaunchKernelGGL((k_batch_step<true, 64, 8>), dim3(grid), dim3(BT), 0, s, a); break;
    case 6: hipLaunchKernelGGL((k_batch_step<false, 32, 8>), dim3(grid), dim3(BT), 0, s, a); break;
    default: hipLaunchKernelGGL((k_batch_step<true, 32, 8>), dim3(grid), dim3(BT), 0, s, a); break;
  }
}

void batch_identity_launch(const BatchArgs &a, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_identity, dim3(16, a.count), dim3(256), 0, s, a);
}

void batch_finalize_launch(const BatchArgs &a, double *sy, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_finalize, dim3(a.count), dim3(BT), 0, s, a, sy);
}

}  // namespace jkg
