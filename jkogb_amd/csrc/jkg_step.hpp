// jkg_step.hpp -- device buffers and launch interface of one dynamic-ordering
// step (src/D/dstep.F90:138-443 and src/Z/zstep.F90:140-447, restated for
// the GPU).  All per-step control values live in device memory (StepState),
// so a step is a fixed sequence of launches that a HIP graph can capture and
// that never needs the host between steps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "jkg_sort.hpp"

namespace jkg {

// Entries allocated past the capacity of the sort's value buffers and of the
// candidate lists: the filter reads groups of 4 entries that may reach past
// the end of a list.
constexpr int LIST_PAD = 256;

// Candidate lists of the selection rounds after the first (ping-pong): the
// packed pairs and their 1-based positions in the sorted list; passed by value.
struct ListPtrs {
  uint32_t *val[2];
  uint32_t *pos[2];
};

// One CSV line of the reference's per-step trace (src/D/dstep.F90:410-440).
struct StepRec {
  int32_t step, oldlen, newlen, trig, hyp, it;
  double tw;
  float build_s, transf_s;  // device wall time of the step's BUILD and TRANSF parts
};

// Column-rotation / row-rotation descriptor of a 2x2 factor B in the ratio
// form of BA/AB (src/D/dtransf.F90:121-242):
//   xx = FMAD(sa ? y : x, r1, sa ? x : y) * m1
//   yy = FMAD(sb ? x : y, r2, sb ? y : x) * m2
struct Rot {
  double r1, r2, m1, m2;
  int sa, sb, pad0, pad1;
};

struct StepState {
  // (the host polls the first three fields in one copy)
  int step;       // index S of the step being executed
  int active;     // this step runs (set at step begin: !done)
  int done;       // the loop has ended
  int info;       // final INFO of the loop (steps, or < 0)
  int max_steps;  // < 0: unlimited
  int n2;         // N_2 after JSTEP_LEN
  int nn_sort;    // number of keys to sort this step (NN, or 0 when inactive)
  int it;         // finite weights this step (IT)
  int sl;         // pivots chosen this step (SL)
  int nl;         // effective transforms this step (NEWLEN)
  int err;        // a pivot's 2x2 failed: min INFO (< 0) or 0
  int nrow;       // pivots with a row transform (INFO & 2)
  // selection bookkeeping
  int sel_done;   // the greedy walk has finished for this step
  int sel_round;  // round of the current list
  int list_len;   // length of the current candidate list
  int consumed;   // entries of the current list already walked
  int next_len;   // length of the list the last filter wrote (the next round's)
  int trace_cap;  // capacity of the trace buffer
  int nrec;       // trace records written by this run (last step index - step0 + 1)
  int step0;      // index of the first step of this run (a resumed loop starts past 0)
  int ntrig, nhyp;
  double tw;
  double tau;     // > 0: the walk stops at the first weight below tau * (largest weight) (mixed solver)
  // step controller of the mixed solver (features after each step, last action)
  int n2max;      // N_2 of the problem: the largest N_2 a controller may choose
  int ctl_phase;  // phase of the last step under a phase-table policy
  double wmax;    // largest weight of this step
  double tw0, twprev, ynorm2;  // TW of the first and of the previous step; ||Y||_F^2
  unsigned long long t_begin;  // wall clock (100 MHz) at the start of this step
  unsigned long long t_transf; // wall clock at the start of its TRANSF part (the 2x2 pivots)
  float ema, n2f, ltau;        // smoothed TW rate; the action of this step (N_2 / n2max, log10 tau)
  float ltau_cap;              // largest log10 tau a controller may set
  float lhandoff;              // log10(TW / TW_0) below which a controller hands back to the full ordering
  float feat[16];              // controller features after the last step
  // FP32 warm start: relative off-norm of A (this step's), the best seen and its step, why the phase ended
  double off_rel, off_best;
  int off_best_step, warm_end;
  int quiet;  // cyclic ordering: consecutive steps without a transform
  // work counters summed over the steps (bytes and flops of the run)
  unsigned long long acc_sl, acc_it, acc_b2, acc_b4, acc_abod;
  // hyperbolic weights that needed the column sums, added per workgroup into one of 64 slots
  // (blockIdx mod 64; one address for every workgroup would queue their atomics): the run's
  // count is acc_abod plus their sum
  unsigned long long acc_abod_part[64];
};

// Adds v to the work counter of hyperbolic pairs (one call per workgroup).
__device__ __forceinline__ void acc_abod_add(StepState *st, unsigned long long v) {
  const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  atomicAdd(&st->acc_abod_part[b & 63], v);
}
inline unsigned long long acc_abod_total(const StepState &s) {
  unsigned long long t = s.acc_abod;
  for (int i = 0; i < 64; ++i) t += s.acc_abod_part[i];
  return t;
}

struct DeviceStep {
  int n, lda, cplx, fused;
  int nn, npl, nmi;
  int n2max;
  // matrices (column-major, leading dimension lda): real or complex128
  void *A, *U, *Z;
  int *J;                 // int32 signs
  int *plus, *minus;      // 0-based column indices with J = +1 / -1
  SortBuffers sort;
  // selection lists (round >= 1): keys, pair indices, 1-based sorted positions
  ListPtrs lists;
  int *fcount;            // survivor counts of the filter: per tile, then per wave sub-tile
  uint8_t *used;          // N bytes: index already in a chosen pivot
  // pivots
  int *piv_p, *piv_q, *piv_pos, *piv_info;
  double *piv_w;          // weight of the pivot (for NORMF2D)
  void *piv_V, *piv_W;    // 2x2 U- and Z-factors (4 real or 4 complex each)
  double *piv_d;          // new diagonal (2 per pivot)
  Rot *rot_row, *rot_col; // descriptors (real path)
  void *zrot;             // descriptors (complex path): BA(V), AB(V^H), AB(W) per pivot
  int *colpiv;            // N: j for c = p_j, -(j+2) for c = q_j, -1 otherwise
  StepState *st;
  StepRec *trace;
  int select_limit;       // entries walked per selection round before filtering
  int sel_rounds;         // walk rounds per step (each but the last followed by a filter pass)
  // mixed solver
  int wmode;              // hyperbolic weights: 0 per pair (ABODNDF2), 1 Gram form
  double *gram;           // ks planes of npl x nmi: (O^T O)(plus, minus), O = A with a zero diagonal
  double *coln;           // 2N: off-diagonal column norms^2, then the diagonal of A
  double neg_tol;         // > 0: negligible pairs are dropped (mixed solver's convergence test)
  // step controller: 0 none, 1 MLP (ctl_master/ctl_q), 2 phase table (ctl_table)
  int ctl_mode;
  const float *ctl_master;
  const uint16_t *ctl_q;
  const float *ctl_table;
  float *ctl_trace;       // per step: 16 features, N_2 fraction, log10 tau, step time (us), TW
  int ctl_trace_cap;
  // FP32 warm start (complex mixed solver): while prec32, the step works on the
  // FP32 copies (interleaved complex float, leading dimension lda); cut32: sort
  // keys cut to 32 bits; the phase ends once the relative off-norm of A32 is
  // below switch_tol or has not improved by 1 % in WARM_STALL steps
  void *A32, *U32, *Z32;
  int prec32, cut32;
  double switch_tol;
  // ID_NCP=3 (jkg_ncp3.hip): the best-of-all-offsets selection replaces the greedy walk
  int ncp3;
  double *ncp3_wsum;      // NN: total weight of the walk from each start offset
  // cyclic ordering (the sweep baseline of the mixed solver): round-robin pivots, no weights
  int cyclic;
  // per-step frames of |A| (jkg_gen.hip): every frame_every-th step, frame_size^2 bytes each
  uint8_t *frames;
  int frame_size, frame_every, frame_max;
};

// test-matrix generator (jkg_gen.hip)
void gen_diag_launch(int n, const double *d, double *G, hipStream_t s);
void gen_cols_launch(int n, int npairs, double *G, const int *pa, const int *pb, const double *cs, const int *J,
                     hipStream_t s);
int gen_rows_prepare(int n);
void frame_launch(DeviceStep &d, hipStream_t s);
void gen_rows_launch(int n, int npairs, double *G, const int *pa, const int *pb, const double *cs, hipStream_t s);

constexpr int NCP3_NMAX = 8192;
int ncp3_prepare(int n, int n2max);
void ncp3_launch(DeviceStep &d, hipStream_t s);

constexpr int WARM_STALL = 64;

constexpr int CTL_REC = 20;  // floats per controller trace record

// sets kernel attributes (dynamic LDS) for the problem size; before capture
int step_prepare(DeviceStep &d);
void step_launch(DeviceStep &d, hipStream_t s);
void step_begin_launch(DeviceStep &d, hipStream_t s);
void finalize_launch(DeviceStep &d, hipStream_t s, double *sigma);
void init_identity_launch(DeviceStep &d, hipStream_t s);

// the pieces, exposed for the step-level C-ABI and its parity tests
// (weights/pivots/rotate/identity/finalize dispatch on d.cplx)
void weights_launch(DeviceStep &d, hipStream_t s);
size_t zrot_bytes();
void zweights_launch(DeviceStep &d, hipStream_t s);
void zpivots_launch(DeviceStep &d, hipStream_t s);
void zrotate_launch(DeviceStep &d, hipStream_t s);
int zstep_prepare(DeviceStep &d);
void zinit_identity_launch(DeviceStep &d, hipStream_t s);
void zfinalize_launch(DeviceStep &d, hipStream_t s, double *sigma);
void select_launch(DeviceStep &d, hipStream_t s);
int select_prepare(DeviceStep &d);  // the walk's LDS (flags, pivots of a round, queue) for the problem size
int select_tiles(int nn);            // workgroups of a filter pass over nn entries
size_t select_count_ints(int nn);    // ints of DeviceStep::fcount
void pivots_launch(DeviceStep &d, hipStream_t s);
void rotate_launch(DeviceStep &d, hipStream_t s);
void step_end_launch(DeviceStep &d, hipStream_t s);

// Gram-form hyperbolic weights (jkg_gram.hip) and the FP64 MFMA TN product
//   C[i * csi + j * csj] = sum_{r < k} X[r, xc(i)] * Y[r, yc(j)]
// (xc(i) = xcols ? xcols[i] : i; zero_diag reads X[c, c] and Y[c, c] as 0).
void gram_hyp_weights_launch(DeviceStep &d, hipStream_t s);
void zgram_hyp_weights_launch(DeviceStep &d, hipStream_t s);
// the step controller (jkg_mlp.hip): sets StepState::n2 and ::tau for this step
void controller_launch(DeviceStep &d, hipStream_t s);
// FP32 warm start and its hand-over (jkg_warm.hip)
void warm_begin_launch(DeviceStep &d, hipStream_t s);
void warm_promote_launch(DeviceStep &d, hipStream_t s);
void gemm_h_launch(int n, int cplx, const void *X, const void *Y, void *C, hipStream_t s);
void adj_launch(int n, int cplx, const void *X, void *P, hipStream_t s);
void jrows_launch(int n, int cplx, const void *X, const int *J, void *Y, hipStream_t s);
int ns_blocks(int n);
void ns_step_launch(int n, int cplx, const void *M, const int *J, void *B, double *part, hipStream_t s);
void gemm_tn_launch(int k, const double *X, int ldx, const int *xcols, int m, const double *Y, int ldy,
                    const int *ycols, int nc, double *C, long csi, long csj, bool zero_diag, const StepState *st,
                    hipStream_t s, int cx = 0, int ks = 1);
// split K of the Gram products: ks planes of npl x nmi doubles (per part), added by the weight kernels
constexpr int GRAM_KS_MAX = 8;
int gram_ksplit(int m, int nc, int k);
size_t gram_plane_doubles(int n);

}  // namespace jkg
