// jkg_batch.hpp -- the batched solver: many independent HSVDs, one persistent
// launch per step (jkg_batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "jkg_step.hpp"

namespace jkg {

// One instance: offsets into the pools of BatchArgs.
struct BInst {
  int n, n2, npl, nmi;
  long long a, u, z;  // A, U, Z in mat (doubles)
  long long k;        // keys / values in the pair pools
  long long j, pm, cp, pv;  // J, plus|minus lists, colpiv, pivot ints (4 * n2) in ints
  long long pw;       // pivot weights (n2) and new diagonals (2 * n2) in dbl
  long long pr;       // row and column descriptors (2 * n2) in rot
  long long sy;       // singular values in the output array
};

// Per-instance loop state.
struct BState {
  int step, done, info, max_steps, steps, pad[3];
};

struct BatchArgs {
  int count;
  double tol;            // > 0: the mixed solver's convergence test drops negligible pairs
  int *ndone;            // instances finished (counted once each)
  long long pair_total;  // pairs over all instances (the second key/value buffer starts there)
  const BInst *inst;
  BState *state;
  double *mat, *dbl;
  int *ints;
  uint64_t *keys;
  uint32_t *vals;
  Rot *rot;
  const int *sched, *sched_off;  // instances of workgroup b: sched[sched_off[b] .. sched_off[b+1])
};

int batch_variants();
void batch_step_launch(const BatchArgs &a, int grid, int variant, int fused, hipStream_t s);
void batch_identity_launch(const BatchArgs &a, hipStream_t s);
void batch_finalize_launch(const BatchArgs &a, double *sy, hipStream_t s);

}  // namespace jkg
