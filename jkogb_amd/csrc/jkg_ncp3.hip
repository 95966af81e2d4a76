// jkg_ncp3.hip -- the ID_NCP=3 selection (AW_NCP3 + AW_NCPT,
// src/shared/atypes.F90:488-592): the greedy walk of AW_NCP2 run from every
// start offset of the sorted list, keeping the selection with the largest total
// weight -- the positive weights summed from the last positive pick back to the
// first, then the others in pick order -- and among equal totals the smallest
// start (the first one met, as the reference's loop keeps it).
//
// One wave per start offset (a persistent grid strides over the offsets): the
// wave walks 64 entries at a time; within a chunk a ballot finds the first entry
// whose indices are both free, the wave takes it, and every later lane of the
// chunk drops itself if it shares an index with the taken pair (register
// compares, no LDS round trip).  The used-index bytes and the picks live in the
// wave's LDS slice.  The total of each offset goes to a per-offset array; one
// wave then applies the reference's sequential rule over the offsets
// (keep the new total unless it is <= the best so far) and walks the winner
// again to write the pivots.  O(IT^2 / 64) wave iterations per step: the
// reference's own cost is O(IT^2) on the CPU.
#include "jkg_device.hpp"
#include "jkg_step.hpp"

namespace jkg {

constexpr int NCP3_WAVES = 4;  // waves per workgroup

// LDS of one wave: the used-index bytes (rounded to 16), then n2 pick positions
__host__ __device__ __forceinline__ size_t ncp3_slice(int n, int n2) {
  return (size_t)(((n + 1 + 15) / 16) * 16) + sizeof(int) * (size_t)n2;
}

// Walks from `start` over keys/vals[0, it); returns the number of picks (at most n2),
// their positions in picks[] (wave-uniform after return).
__device__ int ncp3_walk(int n, const uint32_t *__restrict__ vals, int it, int start, int n2, uint8_t *used,
                         int *picks) {
  const int lane = threadIdx.x & 63;
  for (int i = lane; i <= n; i += 64) used[i] = 0;
  __builtin_amdgcn_wave_barrier();
  int np = 0;
  for (int base = start; base < it && np < n2; base += 64) {
    const int e = base + lane;
    int p = -1, q = -1;
    bool alive = false;
    if (e < it) {
      const uint32_t v = vals[e];
      p = pq_p(v);
      q = pq_q(v);
      alive = !used[p] && !used[q];
    }
    while (np < n2) {
      const uint64_t bal = __ballot(alive);
      if (bal == 0) break;
      const int t = __builtin_ctzll(bal);
      const int tp = __shfl(p, t), tq = __shfl(q, t);
      if (lane == 0) {
        picks[np] = base + t;
        used[tp] = 1;
        used[tq] = 1;
      }
      ++np;
      alive = alive && lane > t && p != tp && p != tq && q != tp && q != tq;
    }
    __builtin_amdgcn_wave_barrier();
  }
  __builtin_amdgcn_wave_barrier();
  return np;
}

// The reference's total of a selection: positive weights from the last positive pick
// back to the first, then the remaining picks in order (weights descend along the list).
// (keys in pair-index order, vals the sorted packed pairs, picks positions in vals)
__device__ double ncp3_total(int n, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                             const int *picks, int np) {
  auto w = [&](int t) { return key_weight(pair_key(keys, n, vals[picks[t - 1]])); };
  int l = 0;
  for (int t = np; t >= 1; --t)
    if (w(t) > 0.0) {
      l = t;
      break;
    }
  double wp = 0.0, wn = 0.0;
  for (int t = l; t >= 1; --t) wp = wp + w(t);
  for (int t = l + 1; t <= np; ++t) wn = wn + w(t);
  return wp + wn;
}

// IT (the finite keys: position of the first NaN key) and the list length of this step.
__global__ void k_ncp3_prep(const SortStats *sst, StepState *st) {
  if (!st->active || threadIdx.x != 0) return;
  const int it = sst->n - sst->nan_count;  // the NaN keys sort last
  st->it = it;
  st->list_len = it;
  st->wmax = (it > 0) ? key_weight(sst->minkey) : 0.0;  // the smallest key is the largest weight
  st->sl = 0;
}

// (the step's candidate values are always materialised: the weight kernels write them)
__global__ void __launch_bounds__(64 * NCP3_WAVES) k_ncp3_walks(int n, const uint64_t *keys, const uint32_t *vals,
                                                                double *__restrict__ wsum,
                                                                const StepState *st, int n2max) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ncp3_lds[];
  if (!st->active) return;
  const int it = st->it, n2 = (st->n2 < st->it) ? st->n2 : st->it;  // NCP gets MIN(IT, N_2)
  if (n2 <= 0) return;
  const int wave = threadIdx.x >> 6;
  unsigned char *slice = ncp3_lds + (size_t)wave * ncp3_slice(n, n2max);
  uint8_t *used = slice;
  int *picks = (int *)(slice + ((n + 1 + 15) / 16) * 16);
  const int nwaves = gridDim.x * NCP3_WAVES;
  for (int o = blockIdx.x * NCP3_WAVES + wave; o < it; o += nwaves) {
    const int np = ncp3_walk(n, vals, it, o, n2, used, picks);
    if ((threadIdx.x & 63) == 0) wsum[o] = ncp3_total(n, keys, vals, picks, np);
    __builtin_amdgcn_wave_barrier();
  }
}

// One wave: the reference's sequential choice over the offsets, then the winner's walk.
__global__ void __launch_bounds__(64) k_ncp3_pick(int n, const uint32_t *vals,
                                                  const double *__restrict__ wsum, int *piv_p, int *piv_q,
                                                  int *piv_pos, StepState *st) {
  extern __shared__ __attribute__((aligned(16))) unsigned char ncp3_lds[];
  if (!st->active) return;
  const int it = st->it, n2 = (st->n2 < st->it) ? st->n2 : st->it;
  const int lane = threadIdx.x;
  if (it <= 0 || n2 <= 0) {
    if (lane == 0) {
      st->sl = 0;
      st->sel_done = 1;
    }
    return;
  }
  // no NaN total: the largest, the first among equals (a reduction); otherwise the
  // sequential rule by one lane
  bool has_nan = false;
  double best = -HUGE_VAL;
  int k = it;
  for (int o = lane; o < it; o += 64) {
    const double w = wsum[o];
    if (!(w == w)) has_nan = true;
    else if (w > best || (w == best && o < k)) {
      best = w;
      k = o;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ob = __shfl_xor(best, off);
    const int ok = __shfl_xor(k, off);
    if (ob > best || (ob == best && ok < k)) {
      best = ob;
      k = ok;
    }
  }
  if (__any(has_nan)) {
    if (lane == 0) {
      double w = __builtin_nan("");
      k = 0;
      for (int o = 0; o < it; ++o) {
        const double v = wsum[o];
        if (!(v <= w)) {
          w = v;
          k = o;
        }
      }
    }
    k = __shfl(k, 0);
  }
  uint8_t *used = ncp3_lds;
  int *picks = (int *)(ncp3_lds + ((n + 1 + 15) / 16) * 16);
  const int np = ncp3_walk(n, vals, it, k, n2, used, picks);
  for (int t = lane; t < np; t += 64) {
    const int e = picks[t];
    const uint32_t v = vals[e];
    piv_p[t] = pq_p(v);
    piv_q[t] = pq_q(v);
    piv_pos[t] = e + 1;
  }
  if (lane == 0) {
    st->sl = np;
    st->sel_done = 1;
    st->consumed = it;
  }
}

int ncp3_prepare(int n, int n2max) {
  if (n > NCP3_NMAX) return -1;
  const size_t lds = NCP3_WAVES * ncp3_slice(n, n2max);
  if (lds > 160 * 1024) return -1;
  if (hipFuncSetAttribute((const void *)k_ncp3_walks, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
      hipSuccess)
    return -2;
  return 0;
}

// Selection of a step under ID_NCP=3 (in place of select_launch); wsum: NN doubles.
void ncp3_launch(DeviceStep &d, hipStream_t s) {
  const int n2max = d.n2max;
  hipLaunchKernelGGL(k_ncp3_prep, dim3(1), dim3(64), 0, s, sort_stats(d.sort), d.st);
  hipLaunchKernelGGL(k_ncp3_walks, dim3(1024), dim3(64 * NCP3_WAVES), NCP3_WAVES * ncp3_slice(d.n, n2max), s, d.n,
                     d.sort.k0, d.sort.v0, d.ncp3_wsum, d.st, n2max);
  hipLaunchKernelGGL(k_ncp3_pick, dim3(1), dim3(64), ncp3_slice(d.n, n2max), s, d.n, d.sort.v0, d.ncp3_wsum, d.piv_p, d.piv_q, d.piv_pos, d.st);
}

}  // namespace jkg
