// jkg_sort.hip -- device-wide stable LSD radix sort of (uint64 key, uint32 value)
// pairs, ascending, for the candidate list of the dynamic-ordering step
// (the reference's AW_SRT1/AW_SRT2, src/shared/atypes.F90:218-355, restated as
// a radix sort over keys that encode its AW_CMP order; see weight_key()).
//
// 8-bit digits, eight passes.  Each pass is three kernels:
//   hist    -- one 4096-key tile per workgroup, LDS histogram -> ghist[digit][tile]
//   scan    -- one workgroup per digit: exclusive scan of its row of tile
//              counts, plus the digit total; a digit holding every key marks the
//              pass as the identity (skipped: stability makes it a no-op)
//   scatter -- stable rank of each key inside its tile: each wave ranks its
//              own quarter of the tile against wave-private digit counters
//              (8 ballots per key give the lanes with the same digit; no
//              barrier), then the per-wave counts give each wave's offset
//              inside each digit; the tile is staged in LDS in digit order and
//              written out in that order so that each digit's run leaves as
//              consecutive addresses (base[digit] + tile offset + rank)
// Buffers ping-pong; which one holds the current data is derived on the device
// from the per-pass skip flags, so the launch sequence is fixed (graph-safe).
// The first executed pass takes the values implicitly as the key's index.
#include "jkg_sort.hpp"

namespace jkg {

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 4096 keys per workgroup
constexpr int RADIX = 256;

struct SortPassInfo {
  int cur;         // buffer index holding the input of this pass
  bool vals_iota;  // values not materialised yet: value = index
};

__device__ __forceinline__ SortPassInfo pass_info(const int *skip, int pass) {
  int moved = 0;
  for (int i = 0; i < pass; ++i) moved += skip[i] ? 0 : 1;
  return SortPassInfo{moved & 1, moved == 0 && !skip[8]};
}

__global__ void __launch_bounds__(SORT_THREADS) k_sort_hist(const uint64_t *__restrict__ k0,
                                                            const uint64_t *__restrict__ k1, const int *n_ptr,
                                                            int *__restrict__ ghist, const int *skip, int pass,
                                                            int ntiles) {
  __shared__ int h[RADIX];
  const int n = *n_ptr;
  const int tiles = (n + SORT_TILE - 1) / SORT_TILE;
  for (int i = threadIdx.x; i < RADIX; i += SORT_THREADS) h[i] = 0;
  __syncthreads();
  const int t = blockIdx.x;
  if (t < tiles) {
    SortPassInfo pi = pass_info(skip, pass);
    const uint64_t *keys = pi.cur ? k1 : k0;
    const int shift = pass * 8;
    const int base = t * SORT_TILE;
#pragma unroll 4
    for (int r = 0; r < SORT_ITEMS; ++r) {
      int idx = base + r * SORT_THREADS + threadIdx.x;
      if (idx < n) atomicAdd(&h[(int)((keys[idx] >> shift) & 0xFF)], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < RADIX; i += SORT_THREADS) ghist[i * ntiles + t] = (t < tiles) ? h[i] : 0;
}

// one workgroup per digit value
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scan(int *__restrict__ ghist, int *__restrict__ dtotal,
                                                            const int *n_ptr, int *skip, int pass, int ntiles) {
  __shared__ int part[SORT_THREADS];
  const int d = blockIdx.x;
  int *row = ghist + d * ntiles;
  const int per = (ntiles + SORT_THREADS - 1) / SORT_THREADS;
  const int lo = threadIdx.x * per;
  int s = 0;
  for (int i = 0; i < per; ++i)
    if (lo + i < ntiles) s += row[lo + i];
  part[threadIdx.x] = s;
  __syncthreads();
  // inclusive Hillis-Steele scan of the 256 partial sums
  for (int off = 1; off < SORT_THREADS; off <<= 1) {
    int v = (threadIdx.x >= off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;
  for (int i = 0; i < per; ++i)
    if (lo + i < ntiles) {
      int v = row[lo + i];
      row[lo + i] = run;
      run += v;
    }
  if (threadIdx.x == SORT_THREADS - 1) {
    int tot = part[SORT_THREADS - 1];
    dtotal[d] = tot;
    if (tot == *n_ptr) skip[pass] = 1;  // every key has this digit: identity pass
  }
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive prefix sum of v over the 256 threads (digit = thread), total in *tot.
__device__ __forceinline__ int block_excl_scan256(int v, int *wsum, int *tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int before = 0;
#pragma unroll
  for (int w = 0; w < SORT_THREADS / 64; ++w) before += (w < wave) ? wsum[w] : 0;
  if (tot) *tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  return before + inc - v;
}

// Wave w ranks the keys [1024 w, 1024 (w + 1)) of the tile, 64 at a time in lane order, against
// counters of its own (no barrier: a wave's LDS accesses are ordered); then the per-wave counts
// give each wave's offset inside each digit, the tile is staged in LDS in stable digit order and
// written out in that order.
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(uint64_t *__restrict__ k0, uint64_t *__restrict__ k1,
                                                               uint32_t *__restrict__ v0, uint32_t *__restrict__ v1,
                                                               const int *n_ptr, const int *__restrict__ ghist,
                                                               const int *__restrict__ dtotal, const int *skip,
                                                               int pass, int ntiles) {
  constexpr int NW = SORT_THREADS / 64;
  // 53 KB in all, so that three workgroups share a CU
  __shared__ int gdelta[RADIX];    // global position of a staged slot of digit d, minus the slot
  __shared__ int wcnt[NW][RADIX];  // per-wave digit counters, then the first slot of wave w's keys of digit d
  __shared__ int wsum[NW];
  __shared__ uint64_t skey[SORT_TILE];
  __shared__ uint32_t sval[SORT_TILE];
  const int n = *n_ptr;
  const int tiles = (n + SORT_TILE - 1) / SORT_TILE;
  const int t = blockIdx.x;
  if (t >= tiles || skip[pass]) return;
  const SortPassInfo pi = pass_info(skip, pass);
  const uint64_t *kin = pi.cur ? k1 : k0;
  uint64_t *kout = pi.cur ? k0 : k1;
  const uint32_t *vin = pi.cur ? v1 : v0;
  uint32_t *vout = pi.cur ? v0 : v1;
  const int shift = pass * 8;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tlen = (n - t * SORT_TILE < SORT_TILE) ? n - t * SORT_TILE : SORT_TILE;
  const int wbase = wave * (SORT_TILE / NW);

  uint64_t key[SORT_ITEMS];
  uint32_t val[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const int li = wbase + r * 64 + lane;
    const int idx = t * SORT_TILE + li;
    const bool valid = li < tlen;
    key[r] = valid ? kin[idx] : ~0ull;
    val[r] = valid ? (pi.vals_iota ? (uint32_t)idx : vin[idx]) : 0u;
  }
#pragma unroll
  for (int i = 0; i < RADIX / 64; ++i) wcnt[wave][lane + 64 * i] = 0;

  // stable rank of each key among the wave's keys of the same digit
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int rank[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const bool valid = wbase + r * 64 + lane < tlen;
    const int d = (int)((key[r] >> shift) & 0xFF);
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1);
      match &= ((d >> b) & 1) ? bal : ~bal;
    }
    const int before = valid ? wcnt[wave][d] : 0;
    const int g = __popcll(match & lt_mask);
    rank[r] = before + g;
    if (valid && g == 0) wcnt[wave][d] = before + __popcll(match);
  }
  __syncthreads();
  // digit d = thread: the waves' offsets inside the digit, the tile's digit starts, the global starts
  {
    const int d = threadIdx.x;
    int c = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      const int x = wcnt[w][d];
      wcnt[w][d] = c;
      c += x;
    }
    const int ls = block_excl_scan256(c, wsum, nullptr);  // first slot of digit d in the staged tile
#pragma unroll
    for (int w = 0; w < NW; ++w) wcnt[w][d] += ls;
    __syncthreads();
    const int gs = block_excl_scan256(dtotal[d], wsum, nullptr);
    gdelta[d] = gs + ghist[d * ntiles + t] - ls;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    if (wbase + r * 64 + lane < tlen) {
      const int d = (int)((key[r] >> shift) & 0xFF);
      const int slot = wcnt[wave][d] + rank[r];
      skey[slot] = key[r];
      sval[slot] = val[r];
    }
  }
  __syncthreads();
  // write out in staged order: neighbouring lanes hold neighbouring slots, so each digit's
  // run of this tile goes out as consecutive addresses
  for (int i = threadIdx.x; i < tlen; i += SORT_THREADS) {
    const uint64_t k = skey[i];
    const int d = (int)((k >> shift) & 0xFF);
    const int pos = gdelta[d] + i;
    kout[pos] = k;
    vout[pos] = sval[i];
  }
}

// Resets the per-pass skip flags (part of the captured launch sequence).
__global__ void k_sort_reset(int *skip, int vals_given) {
  if (threadIdx.x < 8) skip[threadIdx.x] = 0;
  if (threadIdx.x == 8) skip[8] = vals_given;
}

size_t sort_scratch_bytes(int nmax) {
  int ntiles = (nmax + SORT_TILE - 1) / SORT_TILE;
  if (ntiles < 1) ntiles = 1;
  return sizeof(int) * ((size_t)RADIX * ntiles + RADIX + 16);
}

void sort_pairs_launch(SortBuffers &b, const int *n_dev, int nmax, hipStream_t st, int vals_given) {
  int ntiles = (nmax + SORT_TILE - 1) / SORT_TILE;
  if (ntiles < 1) ntiles = 1;
  int *ghist = b.scratch;
  int *dtotal = ghist + (size_t)RADIX * ntiles;
  int *skip = dtotal + RADIX;
  hipLaunchKernelGGL(k_sort_reset, dim3(1), dim3(64), 0, st, skip, vals_given);
  for (int pass = 0; pass < 8; ++pass) {
    hipLaunchKernelGGL(k_sort_hist, dim3(ntiles), dim3(SORT_THREADS), 0, st, b.k0, b.k1, n_dev, ghist, skip, pass,
                       ntiles);
    hipLaunchKernelGGL(k_sort_scan, dim3(RADIX), dim3(SORT_THREADS), 0, st, ghist, dtotal, n_dev, skip, pass, ntiles);
    hipLaunchKernelGGL(k_sort_scatter, dim3(ntiles), dim3(SORT_THREADS), 0, st, b.k0, b.k1, b.v0, b.v1, n_dev, ghist,
                       dtotal, skip, pass, ntiles);
  }
}

}  // namespace jkg
