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
//   scatter -- stable rank of each key inside its tile (wave ballots over the 8
//              digit bits, then a cross-wave prefix in LDS), the tile staged in
//              LDS in digit order, then written out in that order so that each
//              digit's run leaves as consecutive addresses (base[digit] + tile
//              offset + rank)
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

__global__ void __launch_bounds__(SORT_THREADS) k_sort_scatter(uint64_t *__restrict__ k0, uint64_t *__restrict__ k1,
                                                               uint32_t *__restrict__ v0, uint32_t *__restrict__ v1,
                                                               const int *n_ptr, const int *__restrict__ ghist,
                                                               const int *__restrict__ dtotal, const int *skip,
                                                               int pass, int ntiles) {
  __shared__ int gbase[RADIX];      // global output position of this tile's first key of each digit
  __shared__ int lstart[RADIX];     // first slot of each digit in the tile staged in digit order
  __shared__ int run[RADIX];        // running slot of each digit while staging
  __shared__ int wcnt[4][RADIX];    // per-wave digit counts of the current round
  __shared__ int scan_tmp[RADIX];
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

  // digit starts: exclusive scan of the digit totals; this tile's count of each digit from the
  // per-tile offsets (ghist holds, per digit, the exclusive scan over the tiles)
  {
    const int d = threadIdx.x;
    const int tile_off = ghist[d * ntiles + t];
    const int tile_cnt = ((t + 1 < tiles) ? ghist[d * ntiles + t + 1] : dtotal[d]) - tile_off;
    scan_tmp[d] = dtotal[d];
    __syncthreads();
    for (int off = 1; off < RADIX; off <<= 1) {
      int v = (d >= off) ? scan_tmp[d - off] : 0;
      __syncthreads();
      scan_tmp[d] += v;
      __syncthreads();
    }
    gbase[d] = scan_tmp[d] - dtotal[d] + tile_off;
    __syncthreads();
    scan_tmp[d] = tile_cnt;
    __syncthreads();
    for (int off = 1; off < RADIX; off <<= 1) {
      int v = (d >= off) ? scan_tmp[d - off] : 0;
      __syncthreads();
      scan_tmp[d] += v;
      __syncthreads();
    }
    lstart[d] = scan_tmp[d] - tile_cnt;
    run[d] = lstart[d];
  }
  // stage the tile in LDS in stable digit order (16 rounds of 256 keys)
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < SORT_ITEMS; ++r) {
    for (int i = threadIdx.x; i < 4 * RADIX; i += SORT_THREADS) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const int li = r * SORT_THREADS + threadIdx.x;
    const int idx = t * SORT_TILE + li;
    const bool valid = li < tlen;
    uint64_t key = valid ? kin[idx] : 0;
    uint32_t val = valid ? (pi.vals_iota ? (uint32_t)idx : vin[idx]) : 0;
    int d = (int)((key >> shift) & 0xFF);
    // lanes of this wave with the same digit
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bal = __ballot(valid && ((d >> b) & 1));
      match &= ((d >> b) & 1) ? bal : ~bal;
    }
    int wrank = __popcll(match & lt_mask);
    if (valid && wrank == 0) wcnt[wave][d] = __popcll(match);
    __syncthreads();
    if (valid) {
      int before = 0;
      for (int w = 0; w < wave; ++w) before += wcnt[w][d];
      const int slot = run[d] + before + wrank;
      skey[slot] = key;
      sval[slot] = val;
    }
    __syncthreads();
    {
      const int dd = threadIdx.x;
      run[dd] += wcnt[0][dd] + wcnt[1][dd] + wcnt[2][dd] + wcnt[3][dd];
    }
    __syncthreads();
  }
  // write out in staged order: neighbouring lanes hold neighbouring slots, so each digit's
  // run of this tile goes out as consecutive addresses
  for (int i = threadIdx.x; i < tlen; i += SORT_THREADS) {
    const uint64_t key = skey[i];
    const int d = (int)((key >> shift) & 0xFF);
    const int pos = gbase[d] + (i - lstart[d]);
    kout[pos] = key;
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
