// jkg_step.hip -- one dynamic-ordering step of the real FP64 J-Kogbetliantz
// HSVD on gfx950 (the reference's DSTEP_EXEC, src/D/dstep.F90:376-443).
//
//   begin   -- per-step control values, maps reset                    (1 WG)
//   weights -- (a) trigonometric pairs: DASUM2 of the 2x2 off-diagonals;
//              hyperbolic pairs: the 2x2 HSVD per pair, then the off-norm
//              change of columns p,q (ABODNDF2) with the rows staged in LDS
//   sort    -- (b) stable LSD radix sort of (key(W), pair index)
//   select  -- (c) exact greedy walk of the sorted list (AW_NCP2): one
//              workgroup tests super-chunks of the list against an LDS flag
//              array with all its lanes and one wave resolves the entries
//              left alive, in order, with ballots; between rounds every CU
//              filters the rest of the list against the flags (order-
//              preserving compaction), so later rounds visit only candidates
//              whose indices were free
//   pivots  -- (d) batched 2x2 HSVD of the chosen pivots, rotation descriptors
//   rotate  -- (e) one workgroup per column pair: both columns of A staged in
//              LDS, all row transforms (BA), then the column transform (AB)
//              and the new diagonal; columns of U and Z updated in place.
//              1024 lanes per workgroup: the LDS of two columns leaves two
//              workgroups per CU, and 32 waves keep more loads in flight
//   end     -- NEWLEN, NORMF2D, TRIG, HYP trace line; loop control  (1 WG)
//
// Everything follows the reference's operation order, so with identical
// inputs the results are bit-identical to it.
#include "jkg_hsvd2.hpp"
#include "jkg_step.hpp"
#include "jkg_rot.hpp"

namespace jkg {

// ------------------------------------------------------------------ helpers
// ------------------------------------------------------------------ begin
__global__ void k_step_begin(StepState *st, int nn, int n, int *colpiv) {
  __shared__ int act;
  if (threadIdx.x == 0) {
    st->t_begin = wall_clock64();
    int a = !st->done;
    if (a && st->max_steps >= 0 && st->step >= st->max_steps) {
      st->done = 1;
      st->info = st->step;
      a = 0;
    }
    st->active = a;
    st->nn_sort = a ? nn : 0;
    st->sl = 0;
    st->nl = 0;
    st->err = 0;
    st->nrow = 0;
    st->sel_done = !a;
    st->sel_round = 0;
    st->consumed = 0;
    st->list_len = 0;
    st->next_len = 0;
    act = a;
  }
  __syncthreads();
  if (!act) return;
  for (int i = threadIdx.x; i < n; i += blockDim.x) colpiv[i] = -1;
}

// ------------------------------------------------------------------ (a) weights
// DMAGF2T (src/D/dstep.F90:12-37) for the pairs with J_p == J_q.
// tol > 0 (mixed solver only): a pair whose transform would be the identity
// if its off-diagonal pair were zero and whose off-diagonal pair is negligible
// (see negligible()) is dropped like the reference drops a zero pair.
// One workgroup per 32 x 32 tile of pairs (p in [P0, P0 + 32), q in [Q0, Q0 +
// 32), p < q): the blocks A(q, p) and A(p, q) are staged in LDS with coalesced
// column loads, and the pairs are handed to the lanes along the tile's
// diagonals (q - p fixed within 32 lanes), where consecutive lanes have
// consecutive pair indices, so the keys and values leave as contiguous runs.
// E = float: the FP32 warm start (entries promoted; keys cut to their top 32 bits when cut).
constexpr int TWT = 32;
template <bool F, class E>
__global__ void __launch_bounds__(256) k_trig_weights(int n, const E *__restrict__ A, int lda,
                                                       const int *__restrict__ J, uint64_t *__restrict__ keys,
                                                       const StepState *st,
                                                       double tol, int cut) {
  const int bp = blockIdx.y, bq = blockIdx.x;
  if (bp > bq || !st->active) return;
  __shared__ double X[TWT][TWT + 1];  // X[qi][pi] = A(Q0 + qi, P0 + pi)
  __shared__ double Y[TWT][TWT + 1];  // Y[pi][qi] = A(P0 + pi, Q0 + qi)
  __shared__ double dgp[TWT], dgq[TWT];
  __shared__ int jp[TWT], jq[TWT];
  const int P0 = bp * TWT, Q0 = bq * TWT, t = threadIdx.x;
  // every load issued before any is used: addresses clamped into the matrix, entries outside
  // it cleared when the tile is written to LDS
  constexpr int TL = TWT * TWT / 256;
  E xv[TL], yv[TL];
#pragma unroll
  for (int i = 0; i < TL; ++i) {
    const int e = t + 256 * i, r = e % TWT, c = e / TWT;
    const int pc = min(P0 + c, n - 1), qr = min(Q0 + r, n - 1), qc = min(Q0 + c, n - 1), pr = min(P0 + r, n - 1);
    xv[i] = A[(size_t)pc * lda + qr];
    yv[i] = A[(size_t)qc * lda + pr];
  }
  const int dq = (t < TWT) ? P0 + t : Q0 + t - TWT;  // lanes 0..63: the diagonal of the P, then the Q block
  const int dqc = min(dq, n - 1);
  E dv = 0;
  int jv = 0;
  if (t < 2 * TWT) {
    dv = A[(size_t)dqc * lda + dqc];
    jv = J[dqc];
  }
#pragma unroll
  for (int i = 0; i < TL; ++i) {
    const int e = t + 256 * i, r = e % TWT, c = e / TWT;
    X[r][c] = (P0 + c < n && Q0 + r < n) ? (double)xv[i] : 0.0;
    Y[r][c] = (P0 + r < n && Q0 + c < n) ? (double)yv[i] : 0.0;
  }
  if (t < TWT) {
    dgp[t] = (dq < n) ? (double)dv : 0.0;
    jp[t] = (dq < n) ? jv : 0;
  } else if (t < 2 * TWT) {
    dgq[t - TWT] = (dq < n) ? (double)dv : 0.0;
    jq[t - TWT] = (dq < n) ? jv : 0;
  }
  __syncthreads();
  // slot s: diagonal dl = s / 32 - 31 (q - Q0 = p - P0 + dl), a = p - P0 = s % 32
  for (int sl = t; sl < (2 * TWT - 1) * TWT; sl += 256) {
    const int a = sl % TWT, qi = a + sl / TWT - (TWT - 1);
    if (qi < 0 || qi >= TWT) continue;
    const int P = P0 + a, Q = Q0 + qi;
    if (P >= Q || Q >= n || jp[a] != jq[qi]) continue;  // hyperbolic pairs: k_hyp_weights
    const int jj = jp[a];
    const double aqp = X[qi][a], apq = Y[a][qi], app = dgp[a], aqq = dgq[qi];
    const bool order = sign_neg(app) || sign_neg(aqq) || ((jj + jj == -2) ? (aqq < app) : (app < aqq));
    bool go = (aqp != 0.0) || (apq != 0.0) || order;
    if (tol > 0.0 && go && !order && negligible(aqp, apq, app, aqq, tol)) go = false;
    const uint32_t k = pair_index(n, P + 1, Q + 1);
    keys[k] = key_cut(go ? weight_key(dasum2<F>(aqp, apq)) : KEY_NAN, cut);
  }
}

// DMAGF2H (src/D/dstep.F90:41-72) + ABODNDF2 (src/D/dtransf.F90:246-348) for
// the pairs with J_p != J_q: one lane per pair of a 16 (J=+1) x 16 (J=-1)
// column tile; rows are streamed through LDS in chunks of 64, coalesced.
constexpr int HT = 16;
constexpr int HR = 64;
template <bool F>
__global__ void __launch_bounds__(256) k_hyp_weights(int n, const double *__restrict__ A, int lda,
                                                      const int *__restrict__ J, const int *__restrict__ plus,
                                                      int npl, const int *__restrict__ minus, int nmi,
                                                      uint64_t *__restrict__ keys,
                                                      StepState *st) {
  if (!st->active) return;
  __shared__ double L[HR][2 * HT + 1];
  __shared__ int cols[2 * HT];
  const int a = threadIdx.x / HT, b = threadIdx.x % HT;
  const int ia = blockIdx.x * HT + a, ib = blockIdx.y * HT + b;
  if (threadIdx.x < 2 * HT) {
    int c = threadIdx.x;
    int g = (c < HT) ? blockIdx.x * HT + c : blockIdx.y * HT + (c - HT);
    cols[c] = (c < HT) ? (g < npl ? plus[g] : -1) : (g < nmi ? minus[g] : -1);
  }
  const bool valid = ia < npl && ib < nmi;
  const int cp = valid ? plus[ia] : 0, cm = valid ? minus[ib] : 1;
  const int p0 = cp < cm ? cp : cm, q0 = cp < cm ? cm : cp;  // 0-based, p0 < q0
  const bool xplus = cp < cm;
  bool need = false;
  double d = 0.0, r1 = 0.0, r2 = 0.0, m1 = 0.0, m2 = 0.0;
  bool sel1 = true, sel2 = true;
  uint64_t key = KEY_NAN;
  if (valid) {
    const size_t P = p0, Q = q0;
    double app = A[P * lda + P], aqp = A[P * lda + Q], apq = A[Q * lda + P], aqq = A[Q * lda + Q];
    if (aqp != 0.0 || apq != 0.0 || sign_neg(app) || sign_neg(aqq)) {
      double a2[4] = {app, aqp, apq, aqq}, u2[4], z2[4];
      int j2[2] = {J[p0], J[q0]};
      int info = dhsvd2<F>(a2, j2, u2, z2, TH1FIX);
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
  __shared__ int nneed;
  if (threadIdx.x == 0) nneed = 0;
  __syncthreads();
  if (need) atomicAdd(&nneed, 1);
  if (__syncthreads_or(need)) {
    if (threadIdx.x == 0) acc_abod_add(st, (unsigned long long)nneed);
    const int xc = xplus ? a : HT + b, yc = xplus ? HT + b : a;
    const int s1c = sel1 ? yc : xc, t1c = sel1 ? xc : yc;
    const bool eq = (sel1 == sel2);
    for (int r0 = 0; r0 < n; r0 += HR) {
      __syncthreads();
      for (int e = threadIdx.x; e < HR * 2 * HT; e += 256) {
        const int c = e / HR, r = e % HR, col = cols[c];
        L[r][c] = (col >= 0 && r0 + r < n) ? A[(size_t)col * lda + r0 + r] : 0.0;
      }
      __syncthreads();
      if (need) {
        const int rmax = (n - r0 < HR) ? n - r0 : HR;
        for (int r = 0; r < rmax; ++r) {
          const double x = L[r][xc], y = L[r][yc], s1 = L[r][s1c], t1 = L[r][t1c];
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
  if (valid) {
    const uint32_t k = pair_index(n, p0 + 1, q0 + 1);
    keys[k] = key;
  }
}

// ------------------------------------------------------------------ (c) select
// The greedy walk (AW_NCP2, src/shared/atypes.F90:430-484): take the first
// entry of the sorted list whose two indices are both free, mark them used,
// repeat, until N_2 pivots or the end of the list.  It is a chain of
// dependent acceptances, so one workgroup walks; what it makes parallel is
// the test of the entries that are already dead.
//
// k_sel_walk: one workgroup of SW_WAVES waves walks the list in super-chunks
// of SW_CHUNK entries (SW_ITEMS consecutive entries per lane, one 16-byte
// buffer load per array, the next super-chunk's loads in flight).  All lanes
// test their entries against used[] (n + 1 flag bytes in LDS; used[n] = 1).
// used[] only grows, so an entry dead at the start of its super-chunk stays
// dead.  The entries still alive are decided by locally dominant rounds: over
// the whole workgroup when many are alive, else compacted, in list order, into
// an LDS queue that wave 0 alone resolves 64 entries at a time.
//
// Between walk rounds the rest of the list is compacted, in order, to the
// entries whose two indices are both free (k_sel_count: survivors per wave
// sub-tile; k_sel_write: each tile sums the counts of the tiles before it and
// writes its survivors with their 1-based positions in the sorted list).  The
// walk of the next round visits only those.  The pivots do not depend on the
// schedule: the filter drops only entries the walk would skip.
constexpr int SW_WAVES = 16;
constexpr int SW_ITEMS = 4;
constexpr int SW_CHUNK = SW_WAVES * 64 * SW_ITEMS;  // entries per super-chunk
static_assert(SW_ITEMS % 4 == 0 && SW_ITEMS <= 16, "16-byte loads; per-lane counts below 32");
static_assert(SW_CHUNK <= (1 << 14), "queue indices in the low 14 bits of first[]");
constexpr int SW_QCAP = SW_CHUNK;  // queue entries: every alive entry of a super-chunk
constexpr int LD_MIN = 192;        // alive entries of a super-chunk up to which wave 0 resolves them alone
constexpr int LD_MAXIT = 12;       // locally dominant rounds per super-chunk before wave 0 takes the rest
constexpr int FT_WAVES = 4;                          // waves of a filter workgroup
constexpr int FT_SUB = 2048;                         // entries per wave sub-tile
constexpr int FT_TILE = FT_WAVES * FT_SUB;           // entries per filter workgroup
constexpr int FT_G = FT_SUB / 256;                   // groups of 4 entries per lane in a sub-tile

__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Nonzero iff pair v (p | q << 16) shares an index with the pair whose
// indices, replicated to both halves, are P and Q.
__device__ __forceinline__ uint32_t pair_hits(uint32_t v, uint32_t P, uint32_t Q) {
  const uint32_t x = v ^ P, y = v ^ Q;
  return (((x - 0x10001u) & ~x) | ((y - 0x10001u) & ~y)) & 0x80008000u;
}

// First index in [lo, hi) at which the monotone predicate holds (hi if none),
// by one wave: 64 probes per round cut the interval by a factor of 64.
template <class Pred>
__device__ int wave_first_true(int lo, int hi, Pred pred) {
  const int lane = threadIdx.x & 63;
  while (lo < hi) {
    const int step = (hi - lo + 63) >> 6;
    const int p = lo + (lane + 1) * step - 1;
    const bool t = (p < hi) ? pred(p) : true;
    const uint64_t b = __ballot(t);
    const int k = b ? __ffsll((unsigned long long)b) - 1 : 64;
    const int nhi = (k < 64) ? min(lo + (k + 1) * step - 1, hi) : hi;
    lo = min(lo + k * step, hi);
    hi = nhi;
  }
  return lo;
}

// Exclusive prefix over the wave of a per-lane count in 0..7 (three ballots),
// and the wave's total.
__device__ __forceinline__ int wave_prefix_small(int cnt, int &total) {
  const uint64_t b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4);
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  total = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
  return __popcll(b0 & lt) + 2 * __popcll(b1 & lt) + 4 * __popcll(b2 & lt);
}

// Wave 0 resolves a queue of entries (packed pairs qv, tags qp) in order: the free-entry test
// against used[] gives a wave-uniform mask of alive lanes; each acceptance is the first set bit,
// a readlane of its pair and a ballot of the lanes that share an index with it (cleared from the
// mask), all in scalar registers; then the accepted lanes mark their indices in used[] and
// either record their pivots (pair, tag = 1-based list position) from index acc - acc0 on, as
// long as acc < n2, or (taken != null) set taken[tag].  Returns the new acc.
__device__ int resolve_queue(const uint32_t *qv, const uint32_t *qp, int total, uint8_t *used, int n, int acc,
                             int acc0, int n2, int *bp, int *bq, int *bpos, uint8_t *taken) {
  const int lane = threadIdx.x & 63;
  const uint32_t sent = (uint32_t)n * 0x10001u;  // the pair (n, n): dead through used[n] = 1
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int b = 0; b < total && acc < n2; b += 64) {
    const int k = b + lane;
    const uint32_t x = (k < total) ? qv[k] : sent;
    const uint32_t xp = (k < total) ? qp[k] : 0u;
    uint64_t A = __ballot((used[x & 0xFFFFu] | used[x >> 16]) == 0);
    uint64_t M = 0;
    int room = n2 - acc;
    while (A != 0 && room > 0) {
      const int L = __ffsll((unsigned long long)A) - 1;
      const uint32_t gv = __builtin_amdgcn_readlane(x, L);
      M |= 1ull << L;
      --room;
      A &= ~__ballot(pair_hits(x, (gv & 0xFFFFu) * 0x10001u, (gv >> 16) * 0x10001u) != 0);
    }
    if ((M >> lane) & 1ull) {
      if (taken) {
        taken[xp] = 1;
      } else {
        const int r = acc - acc0 + __popcll(M & lt);
        bp[r] = (int)(x & 0xFFFFu);
        bq[r] = (int)(x >> 16);
        bpos[r] = (int)xp;
      }
      used[x & 0xFFFFu] = 1;
      used[x >> 16] = 1;
    }
    acc += __popcll(M);
  }
  return acc;
}

// The queue resolved by wave 0 in locally dominant rounds over each group of 64
// entries (the workgroup's rounds below, within one wave, so without barriers: a
// wave's LDS accesses are carried out in the order issued).  An alive entry that
// comes first among the group's alive entries at both of its indices is taken by
// the sequential walk (every earlier entry at those indices is dead), and every
// entry sharing an index with a taken one dies; a round costs a few LDS round trips
// however many entries it takes.  first[] and the round counter (its tags) are the
// workgroup rounds' own.  The taken entries are recorded in list order, as many as
// N_2 leaves room for (the capped walk is a prefix of the uncapped one).
__device__ int resolve_queue_ld(const uint32_t *qv, const uint32_t *qp, int total, uint8_t *used, uint32_t *first,
                                int n, int &ld_round, int acc, int acc0, int n2, int *bp, int *bq, int *bpos) {
  const int lane = threadIdx.x & 63;
  const uint32_t sent = (uint32_t)n * 0x10001u;  // the pair (n, n): dead through used[n] = 1
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int b = 0; b < total && acc < n2; b += 64) {
    const int k = b + lane;
    const uint32_t x = (k < total) ? qv[k] : sent;
    const uint32_t xp = (k < total) ? qp[k] : 0u;
    const uint32_t xa = x & 0xFFFFu, xb = x >> 16;
    bool alive = (used[xa] | used[xb]) == 0;
    uint64_t T = 0;
    for (;;) {
      const uint64_t A = __ballot(alive);
      if (A == 0) break;
      if ((A & (A - 1)) == 0) {  // one entry left: nothing conflicts with it
        T |= A;
        if (alive) {
          used[xa] = 1;
          used[xb] = 1;
        }
        break;
      }
      if (ld_round == 0xFFFF) {  // the tags ran out: first[] anew
        for (int i = lane; i <= n; i += 64) first[i] = 0xFFFFFFFFu;
        ld_round = 0;
      }
      const uint32_t me = ((0xFFFFu - (uint32_t)ld_round) << 14) | (uint32_t)k;
      ++ld_round;
      if (alive) {
        atomicMin(&first[xa], me);
        atomicMin(&first[xb], me);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const bool win = alive && first[xa] == me && first[xb] == me;
      if (win) {
        used[xa] = 1;
        used[xb] = 1;
      }
      T |= __ballot(win);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      alive = alive && !win && (used[xa] | used[xb]) == 0;
    }
    const int room = n2 - acc;
    if ((T >> lane) & 1ull) {
      const int r = __popcll(T & lt);
      if (r < room) {
        bp[acc - acc0 + r] = (int)xa;
        bq[acc - acc0 + r] = (int)xb;
        bpos[acc - acc0 + r] = (int)xp;
      }
    }
    const int nt = __popcll(T);
    acc += (nt < room) ? nt : room;
  }
  return acc;
}

// Round r of the walk.  Round 0 walks the sorted list (positions implicit);
// round r > 0 the list the filter of round r - 1 wrote (values and 1-based
// positions in lp.*[(r - 1) & 1], st->next_len entries).  A round walks at
// most `limit` entries (0: all) from the start of its list.
template <bool HAS_POS>
__global__ void __launch_bounds__(SW_WAVES * 64) k_sel_walk(int n, const uint64_t *keys,
                                                            const uint32_t *svals, const SortStats *sst,
                                                            ListPtrs lp, uint8_t *used_g,
                                                            int *piv_p, int *piv_q, int *piv_pos, StepState *st,
                                                            int round, int limit, int nn, int n2max) {
  // LDS: the queue (values, tags; SW_QCAP each, 16-byte aligned) | n + 1 flag bytes rounded up to
  // words | this round's pivots (p, q, position; n2max each) | first[] of the locally dominant
  // rounds (n + 1 words) | taken flags of a super-chunk (SW_CHUNK bytes)
  extern __shared__ __attribute__((aligned(16))) uint32_t lds_raw[];
  __shared__ int wtot[2][SW_WAVES];
  __shared__ int sh_len;
  if (st->sel_done || st->sel_round > round) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = n >> 2;
  uint32_t *qv = lds_raw, *qp = lds_raw + SW_QCAP;
  uint32_t *lds = lds_raw + 2 * SW_QCAP;  // the flag words
  uint8_t *used = (uint8_t *)lds;
  int *bp = (int *)(lds + nw + 1), *bq = bp + n2max, *bpos = bq + n2max;
  uint32_t *first = (uint32_t *)(bpos + n2max);
  uint8_t *taken = (uint8_t *)(first + n + 1);
  for (int i = tid; i <= n; i += SW_WAVES * 64) first[i] = 0xFFFFFFFFu;
  for (int i = tid; i < SW_CHUNK / 4; i += SW_WAVES * 64) ((uint32_t *)taken)[i] = 0u;
  const uint32_t *vals = HAS_POS ? lp.val[(round - 1) & 1] : svals;
  const uint32_t *poss = HAS_POS ? lp.pos[(round - 1) & 1] : vals;
  if (!HAS_POS) {
    if (wave == 0) {
      // IT = number of finite weights (the NaN keys sort last); the largest weight is the
      // smallest key, both found by the sort's first pass
      const int it = sst->n - sst->nan_count;
      const double wmax = (it > 0) ? key_weight(sst->minkey) : 0.0;
      int len = it;
      // the mixed solver's threshold: walk only the weights >= tau * (largest weight)
      const double tau = st->tau;
      if (tau > 0.0 && it > 0 && wmax > 0.0) {
        const uint64_t kt = weight_key(tau * wmax);  // keys above kt stand for smaller weights
        len = wave_first_true(0, it, [&](int i) { return pair_key(keys, n, svals[i]) > kt; });
      }
      if (lane == 0) {
        st->it = it;
        st->wmax = wmax;
        sh_len = len;
      }
    }
    for (int i = tid; i <= nw; i += SW_WAVES * 64) lds[i] = 0u;
  } else {
    const uint32_t *ug = (const uint32_t *)used_g;  // 256-byte aligned allocation of n bytes
    for (int i = tid; i < nw; i += SW_WAVES * 64) lds[i] = ug[i];
    if (tid == 0) {
      lds[nw] = 0u;
      sh_len = st->next_len;
    }
  }
  lds_barrier();
  if (HAS_POS && tid < (n & 3)) used[4 * nw + tid] = used_g[4 * nw + tid];
  if (tid == 0) used[n] = 1;
  lds_barrier();
  const int len = sh_len;
  const int n2 = st->n2;
  const int acc0 = st->sl;
  int acc = acc0;
  int c = 0;
  const int stop = (limit > 0 && len > limit) ? limit : len;
  // Buffer loads whose range ends at `stop` rounded up to whole 16-byte groups (inside the
  // allocations): loads past it read zeros without touching memory, so every lane loads, and
  // entries at or past `stop` are masked by position.  (The range is made wave-uniform: a buffer
  // resource in vector registers would make every load a loop.)
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const int range = __builtin_amdgcn_readfirstlane(((stop + 3) >> 2) << 4);
  const __amdgpu_buffer_rsrc_t rsv = __builtin_amdgcn_make_buffer_rsrc((void *)vals, (short)0, range, 0x00020000);
  const __amdgpu_buffer_rsrc_t rsp = __builtin_amdgcn_make_buffer_rsrc((void *)poss, (short)0, range, 0x00020000);
  const int loff = tid * SW_ITEMS * 4;
  struct Chunk {
    u4 v[SW_ITEMS / 4], p[SW_ITEMS / 4];
  };
  auto load = [&](Chunk &ch, int base) {
#pragma unroll
    for (int g = 0; g < SW_ITEMS / 4; ++g) {
      ch.v[g] = __builtin_amdgcn_raw_buffer_load_b128(rsv, loff + 16 * g + base * 4, 0, 0);
      if constexpr (HAS_POS) ch.p[g] = __builtin_amdgcn_raw_buffer_load_b128(rsp, loff + 16 * g + base * 4, 0, 0);
    }
  };
  bool run = stop > 0 && acc < n2;
  int par = 0;
  int ld_round = 0;  // locally dominant rounds so far (the tag of first[])
  // Block-wide exclusive prefix of a per-lane count in 0..31 over the entries' order (lane-major
  // within a wave, then wave by wave), and the total: five ballots per wave, one word per wave in
  // LDS (double-buffered), one barrier.
  auto block_prefix = [&](int cnt, int &total) -> int {
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int pre = 0, tot = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      const uint64_t bb = __ballot((cnt >> b) & 1);
      pre += __popcll(bb & lt) << b;
      tot += __popcll(bb) << b;
    }
    if (lane == 0) wtot[par][wave] = tot;
    lds_barrier();
    int base = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SW_WAVES; ++w) {
      const int t = wtot[par][w];
      base += (w < wave) ? t : 0;
      all += t;
    }
    par ^= 1;
    total = all;
    return base + pre;
  };
  // whether any lane of the workgroup has a nonzero value (cheaper than __syncthreads_or)
  auto block_any = [&](uint32_t v) -> bool {
    const uint64_t b = __ballot(v != 0u);
    if (lane == 0) wtot[par][wave] = (b != 0ull) ? 1 : 0;
    lds_barrier();
    int t = 0;
#pragma unroll
    for (int w = 0; w < SW_WAVES; ++w) t |= wtot[par][w];
    par ^= 1;
    return t != 0;
  };
  auto chunk = [&](const Chunk &ch) {
    uint32_t vv[SW_ITEMS], pv[SW_ITEMS];
    const int e0 = c + tid * SW_ITEMS;
#pragma unroll
    for (int g = 0; g < SW_ITEMS / 4; ++g) {
      vv[4 * g] = ch.v[g].x;
      vv[4 * g + 1] = ch.v[g].y;
      vv[4 * g + 2] = ch.v[g].z;
      vv[4 * g + 3] = ch.v[g].w;
      if constexpr (HAS_POS) {
        pv[4 * g] = ch.p[g].x;
        pv[4 * g + 1] = ch.p[g].y;
        pv[4 * g + 2] = ch.p[g].z;
        pv[4 * g + 3] = ch.p[g].w;
      }
    }
    if constexpr (!HAS_POS) {
#pragma unroll
      for (int i = 0; i < SW_ITEMS; ++i) pv[i] = (uint32_t)(e0 + i + 1);
    }
    uint32_t alive = 0;
#pragma unroll
    for (int i = 0; i < SW_ITEMS; ++i)
      alive |= ((e0 + i < stop) && (used[vv[i] & 0xFFFFu] | used[vv[i] >> 16]) == 0 ? 1u : 0u) << i;
    int total;
    const int pre = block_prefix(__popc(alive), total);
    c = (c + SW_CHUNK < stop) ? c + SW_CHUNK : stop;
    if (total > 0 && total <= LD_MIN) {
      // few entries alive: the queue, resolved by wave 0 alone (rounds over the whole
      // workgroup cost more here: a barrier each)
      int o = pre;
#pragma unroll
      for (int i = 0; i < SW_ITEMS; ++i)
        if ((alive >> i) & 1u) {
          qv[o] = vv[i];
          qp[o] = pv[i];
          ++o;
        }
      lds_barrier();
      if (wave == 0) acc = resolve_queue_ld(qv, qp, total, used, first, n, ld_round, acc, acc0, n2, bp, bq, bpos);
      if (wave == 0 && lane == 0) {
        wtot[par][0] = acc;
        wtot[par][1] = ld_round;
      }
      lds_barrier();
      acc = wtot[par][0];
      ld_round = wtot[par][1];
      par ^= 1;
    } else if (total > 0) {
      // Locally dominant rounds over the whole super-chunk: an alive entry that comes first (in
      // list order) among the alive entries at both of its indices is taken by the sequential
      // walk, and every entry sharing an index with a taken one is dead.  first[x] holds the
      // least super-chunk index of an alive entry at index x, tagged with a counter that falls
      // with every round, so values of earlier rounds never win and first[] is never cleared.
      uint32_t al = alive, accm = 0;
      int it = 0;
      while (true) {
#pragma unroll
        for (int i = 0; i < SW_ITEMS; ++i)
          if (((al >> i) & 1u) && (used[vv[i] & 0xFFFFu] | used[vv[i] >> 16])) al &= ~(1u << i);
        if (!block_any(al) || it == LD_MAXIT) break;
        if (ld_round == 0xFFFF) {  // the tags ran out: first[] anew
          for (int i = tid; i <= n; i += SW_WAVES * 64) first[i] = 0xFFFFFFFFu;
          ld_round = 0;
          lds_barrier();
        }
        const uint32_t tag = (0xFFFFu - (uint32_t)ld_round) << 14;
#pragma unroll
        for (int i = 0; i < SW_ITEMS; ++i)
          if ((al >> i) & 1u) {
            const uint32_t me = tag | (uint32_t)(tid * SW_ITEMS + i);
            atomicMin(&first[vv[i] & 0xFFFFu], me);
            atomicMin(&first[vv[i] >> 16], me);
          }
        lds_barrier();
        uint32_t win = 0;
#pragma unroll
        for (int i = 0; i < SW_ITEMS; ++i)
          if ((al >> i) & 1u) {
            const uint32_t me = tag | (uint32_t)(tid * SW_ITEMS + i);
            if (first[vv[i] & 0xFFFFu] == me && first[vv[i] >> 16] == me) win |= 1u << i;
          }
#pragma unroll
        for (int i = 0; i < SW_ITEMS; ++i)
          if ((win >> i) & 1u) {
            used[vv[i] & 0xFFFFu] = 1;
            used[vv[i] >> 16] = 1;
          }
        accm |= win;
        al &= ~win;
        lds_barrier();
        ++it;
        ++ld_round;
      }
      if (it == LD_MAXIT && block_any(al)) {
        // rounds exhausted (a long chain of conflicts): the undecided entries, in list order,
        // resolved by wave 0; an entry taken by a round is never at an index of an earlier
        // undecided entry, so the flags the rounds set kill none of them wrongly
        int total2;
        int o = block_prefix(__popc(al), total2);
#pragma unroll
        for (int i = 0; i < SW_ITEMS; ++i)
          if ((al >> i) & 1u) {
            qv[o] = vv[i];
            qp[o] = (uint32_t)(tid * SW_ITEMS + i);
            ++o;
          }
        lds_barrier();
        if (wave == 0) resolve_queue(qv, qp, total2, used, n, 0, 0, 0x7FFFFFFF, nullptr, nullptr, nullptr, taken);
        lds_barrier();
#pragma unroll
        for (int i = 0; i < SW_ITEMS; ++i)
          if (((al >> i) & 1u) && taken[tid * SW_ITEMS + i]) {
            accm |= 1u << i;
            taken[tid * SW_ITEMS + i] = 0;
          }
      }
      // record the super-chunk's pivots in list order, as many as N_2 leaves room for
      int tacc;
      int r = acc + block_prefix(__popc(accm), tacc);
#pragma unroll
      for (int i = 0; i < SW_ITEMS; ++i)
        if ((accm >> i) & 1u) {
          if (r < n2) {
            bp[r - acc0] = (int)(vv[i] & 0xFFFFu);
            bq[r - acc0] = (int)(vv[i] >> 16);
            bpos[r - acc0] = (int)pv[i];
          }
          ++r;
        }
      acc = (acc + tacc < n2) ? acc + tacc : n2;
    }
    run = c < stop && acc < n2;
  };
  // two super-chunks in registers, ping-pong: the next one's loads are issued before the current
  // one is tested
  Chunk ca, cb;
  int nb = 0;
  load(ca, nb);
  nb += SW_CHUNK;
  while (run) {
    load(cb, nb);
    nb += SW_CHUNK;
    chunk(ca);
    if (!run) break;
    load(ca, nb);
    nb += SW_CHUNK;
    chunk(cb);
  }
  lds_barrier();
  {
    uint32_t *ug = (uint32_t *)used_g;
    for (int i = tid; i < nw; i += SW_WAVES * 64) ug[i] = lds[i];
    if (tid < (n & 3)) used_g[4 * nw + tid] = used[4 * nw + tid];
    for (int i = tid; i < acc - acc0; i += SW_WAVES * 64) {
      piv_p[acc0 + i] = bp[i];
      piv_q[acc0 + i] = bq[i];
      piv_pos[acc0 + i] = bpos[i];
    }
  }
  if (tid == 0) {
    st->sl = acc;
    st->consumed = c;
    st->list_len = len;
    st->sel_round = round;
    if (acc >= n2 || c >= len) st->sel_done = 1;
  }
}

// The n flag bytes of used_g (a 256-byte aligned allocation) into LDS, by words.
// Four words per lane are loaded before any is stored (from clamped addresses).
__device__ __forceinline__ void copy_used(uint8_t *used, const uint8_t *used_g, int n) {
  const int nw = n >> 2, bd = blockDim.x;
  const uint32_t *src = (const uint32_t *)used_g;
  for (int b = 0; b < nw; b += 4 * bd) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = src[min(b + k * bd + (int)threadIdx.x, nw > 0 ? nw - 1 : 0)];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (b + k * bd + (int)threadIdx.x < nw) ((uint32_t *)used)[b + k * bd + threadIdx.x] = w[k];
  }
  if ((int)threadIdx.x < (n & 3)) used[4 * nw + threadIdx.x] = used_g[4 * nw + threadIdx.x];
}

// Filter of round r: survivors of [consumed, list_len) of the round's list whose
// indices are both free.  fcount[t] = survivors of tile t, fcount[ntiles +
// FT_WAVES t + w] = survivors of its wave sub-tile w ...
__global__ void __launch_bounds__(FT_WAVES * 64) k_sel_count(int n, const uint32_t *svals, ListPtrs lp,
                                                             const uint8_t *used_g, int *fcount, int ntiles,
                                                             const StepState *st, int round) {
  extern __shared__ __attribute__((aligned(16))) uint8_t used[];
  __shared__ int wc[FT_WAVES];
  if (st->sel_done || st->sel_round != round) return;
  const int c0 = st->consumed, len = st->list_len;
  const int base = c0 + blockIdx.x * FT_TILE;
  if (base >= len) return;  // counts of tiles past the list are never read
  copy_used(used, used_g, n);
  __syncthreads();
  const uint32_t *vals = (round == 0) ? svals : lp.val[(round - 1) & 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = base + wave * FT_SUB;
  const int end = (sub + FT_SUB < len) ? sub + FT_SUB : len;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  int cnt = 0;
  // c0 is a multiple of SW_CHUNK (the walk stops at whole super-chunks or at the end of its
  // list), so every group of 4 entries is 16-byte aligned; groups reach at most 3 entries past
  // the list, inside the padded allocation.  A lane's FT_G groups are loaded before any is
  // tested (one memory latency per wave, not one per group), unconditionally: a group past the
  // end reads the last group (its entries are masked by position).
  const int glast = (end > 0) ? (end - 1) & ~3 : 0;
  u4 gv[FT_G];
#pragma unroll
  for (int g = 0; g < FT_G; ++g) gv[g] = *(const u4 *)(vals + min(sub + 4 * lane + 256 * g, glast));
#pragma unroll
  for (int g = 0; g < FT_G; ++g) {
    const int e0 = sub + 4 * lane + 256 * g;
    const uint32_t vv[4] = {gv[g].x, gv[g].y, gv[g].z, gv[g].w};
#pragma unroll
    for (int i = 0; i < 4; ++i) cnt += (e0 + i < end && (used[vv[i] & 0xFFFFu] | used[vv[i] >> 16]) == 0) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o);
  if (lane == 0) {
    wc[wave] = cnt;
    fcount[ntiles + FT_WAVES * blockIdx.x + wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < FT_WAVES; ++w) t += wc[w];
    fcount[blockIdx.x] = t;
  }
}

// ... and written in order to the next round's list (values and 1-based
// positions in the sorted list); the tile holding the list's last entry
// stores the new length in st->next_len.
__global__ void __launch_bounds__(FT_WAVES * 64) k_sel_write(int n, const uint32_t *svals, ListPtrs lp,
                                                             const uint8_t *used_g, const int *fcount, int ntiles,
                                                             StepState *st, int round) {
  extern __shared__ __attribute__((aligned(16))) uint8_t used[];
  __shared__ int red[FT_WAVES];
  if (st->sel_done || st->sel_round != round) return;
  const int c0 = st->consumed, len = st->list_len;
  const int t = blockIdx.x;
  const int base = c0 + t * FT_TILE;
  if (base >= len) return;
  copy_used(used, used_g, n);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // survivors of the tiles before this one (each tile sums them itself: no pass over the counts)
  int s = 0;
  for (int b = 0; b < t; b += 4 * (int)blockDim.x) {  // four counts per lane in flight
    int c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) c[k] = fcount[min(b + k * (int)blockDim.x + (int)threadIdx.x, t - 1)];
#pragma unroll
    for (int k = 0; k < 4; ++k) s += (b + k * (int)blockDim.x + (int)threadIdx.x < t) ? c[k] : 0;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o);
  if (lane == 0) red[wave] = s;
  __syncthreads();
  int out = 0;
  for (int w = 0; w < FT_WAVES; ++w) out += red[w];
  const int tile_pre = out;
  for (int w = 0; w < wave; ++w) out += fcount[ntiles + FT_WAVES * t + w];
  if (threadIdx.x == 0 && t == (len - 1 - c0) / FT_TILE) st->next_len = tile_pre + fcount[t];
  const uint32_t *vals = (round == 0) ? svals : lp.val[(round - 1) & 1];
  const uint32_t *poss = (round == 0) ? nullptr : lp.pos[(round - 1) & 1];
  uint32_t *oval = lp.val[round & 1], *opos = lp.pos[round & 1];
  const int sub = base + wave * FT_SUB;
  const int end = (sub + FT_SUB < len) ? sub + FT_SUB : len;
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  // a lane's groups (values and positions) loaded before any is tested, as in k_sel_count
  const int glast = (end > 0) ? (end - 1) & ~3 : 0;
  u4 gv[FT_G], gp[FT_G];
#pragma unroll
  for (int g = 0; g < FT_G; ++g) gv[g] = *(const u4 *)(vals + min(sub + 4 * lane + 256 * g, glast));
  if (poss) {
#pragma unroll
    for (int g = 0; g < FT_G; ++g) gp[g] = *(const u4 *)(poss + min(sub + 4 * lane + 256 * g, glast));
  }
#pragma unroll
  for (int g = 0; g < FT_G; ++g) {
    if (sub + 256 * g >= end) break;  // uniform over the wave
    const int e0 = sub + 4 * lane + 256 * g;
    const uint32_t vv[4] = {gv[g].x, gv[g].y, gv[g].z, gv[g].w};
    uint32_t pv[4];
    if (poss) {
      pv[0] = gp[g].x;
      pv[1] = gp[g].y;
      pv[2] = gp[g].z;
      pv[3] = gp[g].w;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) pv[i] = (uint32_t)(e0 + i + 1);
    }
    uint32_t alive = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      alive |= ((e0 + i < end) && (used[vv[i] & 0xFFFFu] | used[vv[i] >> 16]) == 0 ? 1u : 0u) << i;
    int tot;
    int o = out + wave_prefix_small(__popc(alive), tot);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if ((alive >> i) & 1u) {
        oval[o] = vv[i];
        opos[o] = pv[i];
        ++o;
      }
    out += tot;
  }
}

// ------------------------------------------------------------------ (d) pivots
template <bool F, class E>
__global__ void __launch_bounds__(256) k_pivots(int n, const E *__restrict__ A, int lda,
                                                 const int *__restrict__ J, const int *__restrict__ piv_p,
                                                 const int *__restrict__ piv_q, const int *__restrict__ piv_pos,
                                                 const uint64_t *keys, double *piv_w,
                                                 int *piv_info, double *piv_V,
                                                 double *piv_W, double *piv_d, Rot *rot_row, Rot *rot_col,
                                                 int *colpiv, StepState *st, double tol) {
  if (!st->active) return;
  const int sl = st->sl;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) st->t_transf = wall_clock64();
  if (i >= sl) return;
  const int p = piv_p[i], q = piv_q[i];
  const size_t P = p, Q = q;
  double b[4] = {(double)A[P * lda + P], (double)A[P * lda + Q], (double)A[Q * lda + P], (double)A[Q * lda + Q]};
  int k[2] = {J[p], J[q]};
  const int pos = piv_pos[i];
  if (pos > 0) {
    // the pivot's weight (for NORMF2D): its key, stored at its pair index
    piv_w[i] = key_weight(keys[pair_index(n, p + 1, q + 1)]);
  } else {
    // cyclic ordering (no list): the pair's off-diagonal mass; a negligible pair whose
    // transform would otherwise be the identity is left alone (INFO 8: no transform, no
    // new diagonal), as the dynamic ordering drops it
    piv_w[i] = dasum2<F>(b[1], b[2]);
    const bool order = sign_neg(b[0]) || sign_neg(b[3]) ||
                       (k[0] == k[1] && ((k[0] + k[1] == -2) ? (b[3] < b[0]) : (b[0] < b[3])));
    if (tol > 0.0 && !order && negligible(b[1], b[2], b[0], b[3], tol)) {
      piv_info[i] = 8;
      colpiv[p] = i;
      colpiv[q] = -(i + 2);
      return;
    }
  }
  double v[4], w[4];
  int info = dhsvd2<F>(b, k, v, w, TH1FIX);
  if (info < 0 && pos <= 0) info = 8;  // cyclic: a pair without a transform is skipped
  piv_info[i] = info;
  colpiv[p] = i;
  colpiv[q] = -(i + 2);
  if (info < 0) {
    atomicMin(&st->err, info);
    return;
  }
  if (info == 8) return;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    piv_V[4 * i + t] = v[t];
    piv_W[4 * i + t] = w[t];
  }
  piv_d[2 * i] = b[0];
  piv_d[2 * i + 1] = b[3];
  rot_row[i] = rot_ba(v);
  rot_col[i] = rot_ab(w);
  // only whether any pivot has a row transform is read (a flag, not a count: no atomics)
  if (info & 2) st->nrow = 1;
}

// ------------------------------------------------------------------ (e) rotate
// E = float: the FP32 warm start (descriptors rounded once per pivot, products in FP32).
constexpr int ROT_RPF = 4;  // rows per lane held in registers (N <= 4096 entirely)
template <bool F, class E>
__global__ void __launch_bounds__(1024) k_rotate(int n, E *__restrict__ A, int lda, E *__restrict__ U,
                                                E *__restrict__ Z, const int *__restrict__ piv_p,
                                                const int *__restrict__ piv_q, const int *__restrict__ piv_info,
                                                const double *__restrict__ piv_d, const Rot *__restrict__ rot_row,
                                                const Rot *__restrict__ rot_col, const int *__restrict__ colpiv,
                                                const StepState *st) {
  extern __shared__ __attribute__((aligned(16))) unsigned char l_raw[];
  E *L = (E *)l_raw;  // column c in L[0..n), its partner in L[n..2n)
  if (!st->active) return;
  const int sl = st->sl;
  if (sl <= 0) return;
  const int c = blockIdx.x;
  const int j = colpiv[c];
  if (j <= -2) return;  // the partner column of a pivot: done by its p column
  const int err = st->err, nrow = st->nrow;
  const int info = (j >= 0) ? piv_info[j] : -1;
  const int c1 = (j >= 0) ? piv_q[j] : -1;
  const bool col_rot = (j >= 0) && err == 0 && info >= 0 && (info & 4);
  const bool diag_set = (j >= 0) && err == 0 && info >= 0 && !(info & 8);
  const int tid = threadIdx.x, bs = blockDim.x;
  const bool a_phase = nrow > 0 || col_rot || diag_set;
  const bool do_u = j >= 0 && info >= 0 && (info & 2), do_z = j >= 0 && info >= 0 && (info & 4);
  E *col0 = A + (size_t)c * lda;
  E *col1 = (c1 >= 0) ? A + (size_t)c1 * lda : nullptr;
  E *u0 = U + (size_t)c * lda, *u1 = (c1 >= 0) ? U + (size_t)c1 * lda : nullptr;
  E *z0 = Z + (size_t)c * lda, *z1 = (c1 >= 0) ? Z + (size_t)c1 * lda : nullptr;
  // The first ROT_RPF rows per lane of A's two columns, then of the pivot's U and Z columns, are
  // loaded together up front, so the U and Z loads are in flight while A is transformed (A's
  // loads are issued first: waiting for them does not wait for U and Z); rows past ROT_RPF * bs
  // (N > 4096) go through the same steps in plain loops.
  E ra0[ROT_RPF], ra1[ROT_RPF], ru0[ROT_RPF], ru1[ROT_RPF], rz0[ROT_RPF], rz1[ROT_RPF];
#pragma unroll
  for (int k = 0; k < ROT_RPF; ++k) {
    const int r = tid + k * bs;
    const bool in = r < n;
    ra0[k] = (a_phase && in) ? col0[r] : (E)0;
    ra1[k] = (a_phase && in && col1) ? col1[r] : (E)0;
  }
#pragma unroll
  for (int k = 0; k < ROT_RPF; ++k) {
    const int r = tid + k * bs;
    const bool in = r < n;
    ru0[k] = (do_u && in) ? u0[r] : (E)0;
    ru1[k] = (do_u && in) ? u1[r] : (E)0;
    rz0[k] = (do_z && in) ? z0[r] : (E)0;
    rz1[k] = (do_z && in) ? z1[r] : (E)0;
  }
  if (a_phase) {
#pragma unroll
    for (int k = 0; k < ROT_RPF; ++k) {
      const int r = tid + k * bs;
      if (r < n) {
        L[r] = ra0[k];
        if (col1) L[n + r] = ra1[k];
      }
    }
    for (int r = tid + ROT_RPF * bs; r < n; r += bs) {
      L[r] = col0[r];
      if (col1) L[n + r] = col1[r];
    }
    __syncthreads();
    if (nrow > 0) {
      // phase 1 on A: rows p_i, q_i of every pivot with a row transform
      for (int i = tid; i < sl; i += bs) {
        const int inf = piv_info[i];
        if (inf < 0 || !(inf & 2)) continue;
        const auto R = as_rot(rot_row[i], L);
        const int p = piv_p[i], q = piv_q[i];
        E x = L[p], y = L[q];
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
      __syncthreads();
    }
    if (col1) {
      // phase 2 on A: the column transform and the new 2x2 diagonal
      const int p = c, q = c1;
      const auto R = as_rot(col_rot ? rot_col[j] : Rot{}, L);
      const E d0 = (E)(diag_set ? piv_d[2 * j] : 0.0), d1 = (E)(diag_set ? piv_d[2 * j + 1] : 0.0);
      for (int r = tid; r < n; r += bs) {
        E x = L[r], y = L[n + r];
        if (col_rot) rot_apply<F>(R, x, y);
        if (diag_set) {
          if (r == p) { x = d0; y = (E)0; }
          if (r == q) { x = (E)0; y = d1; }
        }
        col0[r] = x;
        col1[r] = y;
      }
    } else {
      for (int r = tid; r < n; r += bs) col0[r] = L[r];
    }
  }
  // phase 1 on U (columns p,q times V^T) and Z (columns p,q times W)
  if (do_u) {
    const auto R = as_rot(rot_row[j], U);
#pragma unroll
    for (int k = 0; k < ROT_RPF; ++k) {
      const int r = tid + k * bs;
      if (r < n) {
        E x = ru0[k], y = ru1[k];
        rot_apply<F>(R, x, y);
        u0[r] = x;
        u1[r] = y;
      }
    }
    for (int r = tid + ROT_RPF * bs; r < n; r += bs) {
      E x = u0[r], y = u1[r];
      rot_apply<F>(R, x, y);
      u0[r] = x;
      u1[r] = y;
    }
  }
  if (do_z) {
    const auto R = as_rot(rot_col[j], Z);
#pragma unroll
    for (int k = 0; k < ROT_RPF; ++k) {
      const int r = tid + k * bs;
      if (r < n) {
        E x = rz0[k], y = rz1[k];
        rot_apply<F>(R, x, y);
        z0[r] = x;
        z1[r] = y;
      }
    }
    for (int r = tid + ROT_RPF * bs; r < n; r += bs) {
      E x = z0[r], y = z1[r];
      rot_apply<F>(R, x, y);
      z0[r] = x;
      z1[r] = y;
    }
  }
}

// ------------------------------------------------------------------ end
// NEWLEN, NORMF2D (summed in pivot order from -0, as one thread of the
// reference does), TRIG, HYP; then SL = MIN(SL, NL) and the loop test.
// Features of the mixed solver's step controller, written after each step
// (the controller reads them at the start of the next one):
//   0 step / 16N          1 log10(TW / TW_0) / 20   2 log10(TW / TW_prev) / 2
//   3 SL / N_2max         4 NL / SL                 5 IT / NN
//   6 walk depth (last chosen position / IT)        7 trigonometric share of the transforms
//   8 log10(w_max / |Y|^2) / 32                     9 log2(N) / 16
//  10 N_2 fraction used  11 normalised log10 tau used  12 log10(TW / |Y|^2) / 32
//  13 smoothed feature 2 14 SL reached N_2           15 1
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

__device__ void controller_features(StepState *st, int s, int n, int nn, int sl, int nl, int ntr, int nhy, double tw,
                                    const int *piv_pos, float *ctl_trace, int ctl_cap) {
  const double tiny = 1e-300;
  const double twv = tw > tiny ? tw : tiny;
  if (s == 0) {
    st->tw0 = twv;
    st->twprev = twv;
    st->ema = 0.f;
  }
  float f[16];
  f[0] = fminf((float)s / (16.f * (float)n), 1.f);
  f[1] = clampf((float)log10(twv / st->tw0), -20.f, 0.f) / 20.f;
  const float rate = clampf((float)log10(twv / st->twprev), -2.f, 2.f) / 2.f;
  f[2] = rate;
  f[3] = (float)sl / (float)(st->n2max > 0 ? st->n2max : 1);
  f[4] = (float)(nl > 0 ? nl : 0) / (float)(sl > 0 ? sl : 1);
  f[5] = (float)st->it / (float)(nn > 0 ? nn : 1);
  f[6] = (sl > 0 && st->it > 0) ? fminf((float)piv_pos[sl - 1] / (float)st->it, 1.f) : 0.f;
  f[7] = (float)ntr / (float)(ntr + nhy > 0 ? ntr + nhy : 1);
  const double wm = st->wmax > tiny ? st->wmax : tiny;
  f[8] = clampf((float)log10(wm / st->ynorm2), -32.f, 0.f) / 32.f;
  f[9] = log2f((float)n) / 16.f;
  f[10] = st->n2f;
  f[11] = (st->ltau <= -16.f) ? 0.f : (st->ltau + 16.f) / 12.f;
  f[12] = clampf((float)log10(twv / st->ynorm2), -32.f, 0.f) / 32.f;
  st->ema = 0.9f * st->ema + 0.1f * rate;
  f[13] = st->ema;
  f[14] = (sl >= st->n2) ? 1.f : 0.f;
  f[15] = 1.f;
  for (int i = 0; i < 16; ++i) st->feat[i] = f[i];
  st->twprev = twv;
  if (ctl_trace && s < ctl_cap) {
    float *r = ctl_trace + (size_t)s * CTL_REC;
    for (int i = 0; i < 16; ++i) r[i] = f[i];
    r[16] = st->n2f;
    r[17] = st->ltau;
    r[18] = (float)(wall_clock64() - st->t_begin) * 0.01f;  // 100 MHz ticks -> microseconds
    r[19] = (float)tw;
  }
}

__global__ void __launch_bounds__(1024) k_step_end(const int *__restrict__ J, const int *__restrict__ piv_p,
                                                   const int *__restrict__ piv_q, const int *__restrict__ piv_info,
                                                   const double *__restrict__ piv_w, StepState *st, StepRec *trace,
                                                   int n, int nn, const int *__restrict__ piv_pos, int ctl_on,
                                                   float *ctl_trace, int ctl_cap, int sweep) {
  __shared__ int cnt[5];
  if (!st->active) return;
  const int sl = st->sl, s = st->step;
  if (threadIdx.x == 0) {
    st->acc_sl += (unsigned long long)sl;
    st->acc_it += (unsigned long long)st->it;
  }
  StepRec rec;
  rec.step = s;
  rec.oldlen = sl;
  rec.it = st->it;
  // BUILD: from the step's start to its pivots kernel; TRANSF: from there to here (100 MHz ticks)
  const unsigned long long now = wall_clock64();
  rec.build_s = (sl > 0) ? (float)(st->t_transf - st->t_begin) * 1e-8f : (float)(now - st->t_begin) * 1e-8f;
  rec.transf_s = (sl > 0) ? (float)(now - st->t_transf) * 1e-8f : 0.f;
  if (sl <= 0) {
    if (threadIdx.x == 0) {
      rec.newlen = -1;
      rec.tw = -0.0;
      rec.trig = rec.hyp = 0;
      if (s - st->step0 < st->trace_cap) trace[s - st->step0] = rec;
      st->nrec = s - st->step0 + 1;
      st->done = 1;
      st->info = s;
      if (ctl_on) controller_features(st, s, n, nn, 0, 0, 0, 0, -0.0, piv_pos, ctl_trace, ctl_cap);
    }
    return;
  }
  const int err = st->err;
  if (threadIdx.x < 5) cnt[threadIdx.x] = 0;
  // One pass over the pivots in chunks: the counts (NL, trigonometric and hyperbolic
  // transforms, row and column transforms) and TW, the weights of the pivots with a
  // transform summed in pivot order from -0 by one lane (as the reference's loop).  A
  // pivot without a transform (or with a NaN weight) adds -0, which leaves every sum
  // unchanged (x + -0 = x for all x), so the chain is one add per pivot.  All loads of
  // an entry are issued together.
  __shared__ double wbuf[1024];
  double tw = -0.0;
  int nb2 = 0, nb4 = 0, nlc = 0, nt = 0, nh = 0;
  for (int c0 = 0; c0 < sl; c0 += 1024) {
    const int m = (sl - c0 < 1024) ? sl - c0 : 1024;
    __syncthreads();  // wbuf free, cnt cleared
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const int l = piv_info[c0 + i];
      const double w = piv_w[c0 + i];
      const bool tr = (l & 2) || (l & 4);
      if (l >= 0 && (l & 2)) ++nb2;
      if (l >= 0 && (l & 4)) ++nb4;
      if (err == 0) {
        if ((l & 1) && tr) ++nlc;
        if (tr) {
          if (J[piv_p[c0 + i]] == J[piv_q[c0 + i]]) ++nt;
          else ++nh;
        }
      }
      wbuf[i] = (tr && w == w) ? w : -0.0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && err == 0) {
#pragma unroll 16
      for (int i = 0; i < m; ++i) tw += wbuf[i];
    }
  }
  atomicAdd(&cnt[0], nlc);
  atomicAdd(&cnt[1], nt);
  atomicAdd(&cnt[2], nh);
  atomicAdd(&cnt[3], nb2);
  atomicAdd(&cnt[4], nb4);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nl = (err == 0) ? cnt[0] : err;
    rec.newlen = nl;
    rec.tw = tw;
    rec.trig = (err == 0) ? cnt[1] : 0;
    rec.hyp = (err == 0) ? cnt[2] : 0;
    if (s - st->step0 < st->trace_cap) trace[s - st->step0] = rec;
    st->nrec = s - st->step0 + 1;
    st->acc_b2 += (unsigned long long)cnt[3];
    st->acc_b4 += (unsigned long long)cnt[4];
    st->nl = nl;
    if (ctl_on) controller_features(st, s, n, nn, sl, nl, rec.trig, rec.hyp, tw, piv_pos, ctl_trace, ctl_cap);
    const int sl2 = (nl < sl) ? nl : sl;
    if (sweep > 0 && nl >= 0) {
      // cyclic ordering: the run ends after a whole sweep without a transform
      st->quiet = (nl > 0) ? 0 : st->quiet + 1;
      if (st->quiet >= sweep) {
        st->done = 1;
        st->info = s;
      } else {
        st->step = s + 1;
      }
    } else if (sl2 <= 0) {
      st->done = 1;
      st->info = (sl2 < 0) ? sl2 : s;
    } else {
      st->step = s + 1;
    }
  }
}

// ------------------------------------------------------------------ setup / finalize
__global__ void k_identity(double *U, double *Z, int n, int lda) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t c = e / n, r = e % n;
    const double v = (r == c) ? 1.0 : 0.0;
    U[c * lda + r] = v;
    Z[c * lda + r] = v;
  }
}

// SIGMA = diag(A), A(i,i) = 0, Z <- J Z^T J (src/D/dstep.F90:447-468, 594-603)
__global__ void k_finalize(double *A, double *Z, int n, int lda, const int *J, double *sigma) {
  const size_t total = (size_t)n * n;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / n), r = (int)(e % n);
    if (r == c) {
      sigma[r] = A[(size_t)c * lda + r];
      A[(size_t)c * lda + r] = 0.0;
    } else if (r < c) {  // one lane per pair (r, c), r < c
      double *zrc = &Z[(size_t)c * lda + r], *zcr = &Z[(size_t)r * lda + c];
      double t = *zrc;
      double a = *zcr, b = t;
      if (J[r] != J[c]) {
        a = -a;
        b = -b;
      }
      *zrc = a;
      *zcr = b;
    }
  }
}

// ------------------------------------------------------------------ cyclic ordering
// The sweep baselines (no weights, sort or walk; positions 0 mark the pivots as
// cyclic for k_pivots).  order 1, round robin: step s takes round s mod (M - 1)
// of the schedule over M = N (+1 if N is odd: a dummy index) positions -- pair k
// pairs (r + k) and (r - k) mod (M - 1), pair 0 pairs M - 1 with r -- so every
// round is N/2 disjoint pairs and M - 1 rounds are one sweep over all pairs.
// order 2 and 3, the reference's pointwise sweeps (src/old/jk1.f90:101-146): one
// pivot a step, pair s mod NN of the row-cyclic order (1,2), (1,3), .., (1,N),
// (2,3), .. or of the column-cyclic order (1,2), (1,3), (2,3), (1,4), ..
__device__ inline long rows_before(long i, long n) { return (i - 1) * (2 * n - i) / 2; }  // pairs in rows < i
__device__ inline long cols_before(long j) { return (j - 1) * (j - 2) / 2; }              // pairs in columns < j

__global__ void k_cyc_select(int n, int order, int *piv_p, int *piv_q, int *piv_pos, StepState *st) {
  if (!st->active) return;
  if (order >= 2) {
    if (threadIdx.x == 0) {
      const long nn = (long)n * (n - 1) / 2, t = st->step % nn;
      long p, q;
      if (order == 2) {  // the row i (1-based) with rows_before(i) <= t < rows_before(i + 1)
        long i = (long)((2.0 * n - 1.0 - __builtin_sqrt((2.0 * n - 1.0) * (2.0 * n - 1.0) - 8.0 * t)) / 2.0) + 1;
        i = i < 1 ? 1 : (i > n - 1 ? n - 1 : i);
        while (i > 1 && rows_before(i, n) > t) --i;
        while (i < n - 1 && rows_before(i + 1, n) <= t) ++i;
        p = i;
        q = i + 1 + (t - rows_before(i, n));
      } else {  // the column j with cols_before(j) <= t < cols_before(j + 1)
        long j = (long)((3.0 + __builtin_sqrt(1.0 + 8.0 * t)) / 2.0);
        j = j < 2 ? 2 : (j > n ? n : j);
        while (j > 2 && cols_before(j) > t) --j;
        while (j < n && cols_before(j + 1) <= t) ++j;
        p = 1 + (t - cols_before(j));
        q = j;
      }
      piv_p[0] = (int)p - 1;
      piv_q[0] = (int)q - 1;
      piv_pos[0] = 0;
      st->sl = 1;
      st->it = 0;
      st->sel_done = 1;
    }
    return;
  }
  const int m = n + (n & 1), r = st->step % (m - 1);
  for (int k = threadIdx.x; k < m / 2; k += blockDim.x) {
    const int x = (k == 0) ? m - 1 : (r + k) % (m - 1);
    const int y = (r - k + (m - 1)) % (m - 1);
    if ((n & 1) && k == 0) continue;  // the dummy index n sits in pair 0
    const int slot = (n & 1) ? k - 1 : k;
    piv_p[slot] = x < y ? x : y;
    piv_q[slot] = x < y ? y : x;
    piv_pos[slot] = 0;
  }
  if (threadIdx.x == 0) {
    st->sl = n / 2;
    st->it = 0;
    st->sel_done = 1;
  }
}

// ------------------------------------------------------------------ launchers
static inline int grid_for(size_t work, int bs, int cap) {
  size_t g = (work + bs - 1) / bs;
  if (g < 1) g = 1;
  return (int)((g > (size_t)cap) ? cap : g);
}

void step_begin_launch(DeviceStep &d, hipStream_t s) {
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(256), 0, s, d.st, d.nn, d.n, d.colpiv);
}

void weights_launch(DeviceStep &d, hipStream_t s) {
  if (d.cplx) {
    zweights_launch(d, s);
    return;
  }
  const int nt = (d.n + TWT - 1) / TWT;
  const double tol = d.wmode ? d.neg_tol : 0.0;
  auto trig = [&](auto kern, const auto *A, int cut) {
    hipLaunchKernelGGL(kern, dim3(nt, nt), dim3(256), 0, s, d.n, A, d.lda, d.J, d.sort.k0, d.st, tol, cut);
  };
  if (d.prec32) {  // the FP32 warm start (Gram-form hyperbolic weights only)
    const float *A = (const float *)d.A32;
    if (d.fused) trig(k_trig_weights<true, float>, A, d.cut32);
    else trig(k_trig_weights<false, float>, A, d.cut32);
    gram_hyp_weights_launch(d, s);
    return;
  }
  const double *A = (const double *)d.A;
  if (d.fused) trig(k_trig_weights<true, double>, A, 0);
  else trig(k_trig_weights<false, double>, A, 0);
  if (d.wmode == 1) {
    gram_hyp_weights_launch(d, s);
    return;
  }
  if (d.npl > 0 && d.nmi > 0) {
    dim3 g((d.npl + HT - 1) / HT, (d.nmi + HT - 1) / HT);
    if (d.fused)
      hipLaunchKernelGGL(k_hyp_weights<true>, g, dim3(256), 0, s, d.n, A, d.lda, d.J, d.plus, d.npl, d.minus, d.nmi,
                         d.sort.k0, d.st);
    else
      hipLaunchKernelGGL(k_hyp_weights<false>, g, dim3(256), 0, s, d.n, A, d.lda, d.J, d.plus, d.npl, d.minus,
                         d.nmi, d.sort.k0, d.st);
  }
}

static size_t select_lds_bytes(const DeviceStep &d) {
  return (((size_t)d.n >> 2) + 1) * 4 + (size_t)3 * d.n2max * sizeof(int) + (size_t)2 * SW_QCAP * sizeof(uint32_t) +
         ((size_t)d.n + 1) * sizeof(uint32_t) + SW_CHUNK;
}

int select_prepare(DeviceStep &d) {
  const size_t lds = select_lds_bytes(d);
  if (lds > 160 * 1024 - 256) return -1;
  if (hipFuncSetAttribute((const void *)k_sel_walk<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds) != hipSuccess ||
      hipFuncSetAttribute((const void *)k_sel_walk<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)lds) != hipSuccess)
    return -2;
  return 0;
}

int select_tiles(int nn) { return (nn + FT_TILE - 1) / FT_TILE; }

size_t select_count_ints(int nn) { return (size_t)(FT_WAVES + 1) * (select_tiles(nn) + 1) + 2; }

void select_launch(DeviceStep &d, hipStream_t s) {
  const int rounds = d.sel_rounds > 0 ? d.sel_rounds : 2;
  if (d.ncp3) {
    ncp3_launch(d, s);
    return;
  }
  const int ntiles = select_tiles(d.nn);
  const size_t lds = (size_t)d.n + 1;
  const size_t ldsw = select_lds_bytes(d);
  // a round walks whole super-chunks (the filter's groups of 4 entries stay aligned)
  const int limit = ((d.select_limit + SW_CHUNK - 1) / SW_CHUNK) * SW_CHUNK;
  for (int r = 0; r < rounds; ++r) {
    const int lim = (r == rounds - 1) ? 0 : limit;
    if (r == 0)
      hipLaunchKernelGGL(k_sel_walk<false>, dim3(1), dim3(SW_WAVES * 64), ldsw, s, d.n, d.sort.k0, d.sort.v0,
                         sort_stats(d.sort), d.lists, d.used, d.piv_p, d.piv_q, d.piv_pos, d.st, r,
                         lim, d.nn, d.n2max);
    else
      hipLaunchKernelGGL(k_sel_walk<true>, dim3(1), dim3(SW_WAVES * 64), ldsw, s, d.n, d.sort.k0, d.sort.v0,
                         sort_stats(d.sort), d.lists, d.used, d.piv_p, d.piv_q, d.piv_pos, d.st, r,
                         lim, d.nn, d.n2max);
    if (r == rounds - 1) break;
    hipLaunchKernelGGL(k_sel_count, dim3(ntiles), dim3(FT_WAVES * 64), lds, s, d.n, d.sort.v0, d.lists, d.used, d.fcount, ntiles, d.st, r);
    hipLaunchKernelGGL(k_sel_write, dim3(ntiles), dim3(FT_WAVES * 64), lds, s, d.n, d.sort.v0, d.lists, d.used, d.fcount, ntiles, d.st, r);
  }
}

void pivots_launch(DeviceStep &d, hipStream_t s) {
  if (d.cplx) {
    zpivots_launch(d, s);
    return;
  }
  const int g = (d.n2max + 255) / 256;
  auto go = [&](auto kern, const auto *A) {
    hipLaunchKernelGGL(kern, dim3(g), dim3(256), 0, s, d.n, A, d.lda, d.J, d.piv_p, d.piv_q, d.piv_pos, d.sort.k0,
                       d.piv_w, d.piv_info, (double *)d.piv_V,
                       (double *)d.piv_W, d.piv_d, d.rot_row, d.rot_col, d.colpiv, d.st, d.neg_tol);
  };
  if (d.prec32) {
    if (d.fused) go(k_pivots<true, float>, (const float *)d.A32);
    else go(k_pivots<false, float>, (const float *)d.A32);
  } else {
    if (d.fused) go(k_pivots<true, double>, (const double *)d.A);
    else go(k_pivots<false, double>, (const double *)d.A);
  }
}

void rotate_launch(DeviceStep &d, hipStream_t s) {
  if (d.cplx) {
    zrotate_launch(d, s);
    return;
  }
  auto go = [&](auto kern, auto *A, auto *U, auto *Z) {
    hipLaunchKernelGGL(kern, dim3(d.n), dim3(1024), (size_t)2 * d.n * sizeof(*A), s, d.n, A, d.lda, U, Z, d.piv_p,
                       d.piv_q, d.piv_info, d.piv_d, d.rot_row, d.rot_col, d.colpiv, d.st);
  };
  if (d.prec32) {
    float *A = (float *)d.A32, *U = (float *)d.U32, *Z = (float *)d.Z32;
    if (d.fused) go(k_rotate<true, float>, A, U, Z);
    else go(k_rotate<false, float>, A, U, Z);
  } else {
    double *A = (double *)d.A, *U = (double *)d.U, *Z = (double *)d.Z;
    if (d.fused) go(k_rotate<true, double>, A, U, Z);
    else go(k_rotate<false, double>, A, U, Z);
  }
}

void step_end_launch(DeviceStep &d, hipStream_t s) {
  // cyclic orderings: the steps of one sweep (a sweep without a transform ends the run)
  const int sweep = d.cyclic == 1 ? d.n + (d.n & 1) - 1 : (d.cyclic ? (int)d.nn : 0);
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1024), 0, s, d.J, d.piv_p, d.piv_q, d.piv_info, d.piv_w, d.st,
                     d.trace, d.n, d.nn, d.piv_pos, d.ctl_mode != 0 ? 1 : 0, d.ctl_trace, d.ctl_trace_cap, sweep);
}

int step_prepare(DeviceStep &d) {
  if (select_prepare(d) != 0) return -3;
  if (d.cplx) return zstep_prepare(d);
  const size_t lds = (size_t)2 * d.n * sizeof(double);
  if (lds > 160 * 1024) return -1;  // two columns must fit the LDS of one CU
  const void *k[4] = {(const void *)k_rotate<true, double>, (const void *)k_rotate<false, double>,
                      (const void *)k_rotate<true, float>, (const void *)k_rotate<false, float>};
  for (int i = 0; i < 4; ++i)
    if (hipFuncSetAttribute(k[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)(i < 2 ? lds : lds / 2)) !=
        hipSuccess)
      return -2;
  return 0;
}

void step_launch(DeviceStep &d, hipStream_t s) {
  step_begin_launch(d, s);
  frame_launch(d, s);
  if (d.cyclic) {
    hipLaunchKernelGGL(k_cyc_select, dim3(1), dim3(256), 0, s, d.n, d.cyclic, d.piv_p, d.piv_q, d.piv_pos, d.st);
    pivots_launch(d, s);
    rotate_launch(d, s);
    step_end_launch(d, s);
    return;
  }
  if (d.ctl_mode) controller_launch(d, s);
  weights_launch(d, s);
  sort_pairs_launch(d.sort, &d.st->nn_sort, d.nn, s, d.n, (d.prec32 && d.cut32) ? 32 : 0);
  select_launch(d, s);
  pivots_launch(d, s);
  rotate_launch(d, s);
  step_end_launch(d, s);
}

void init_identity_launch(DeviceStep &d, hipStream_t s) {
  if (d.cplx) {
    zinit_identity_launch(d, s);
    return;
  }
  hipLaunchKernelGGL(k_identity, dim3(grid_for((size_t)d.n * d.n, 256, 8192)), dim3(256), 0, s, (double *)d.U,
                     (double *)d.Z, d.n, d.lda);
}

void finalize_launch(DeviceStep &d, hipStream_t s, double *sigma) {
  if (d.cplx) {
    zfinalize_launch(d, s, sigma);
    return;
  }
  hipLaunchKernelGGL(k_finalize, dim3(grid_for((size_t)d.n * d.n, 256, 8192)), dim3(256), 0, s, (double *)d.A,
                     (double *)d.Z, d.n, d.lda, d.J, sigma);
}

}  // namespace jkg
