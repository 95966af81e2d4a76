// jkg_sort.hip -- device-wide stable LSD radix sort of 64-bit keys carrying
// 32-bit values, ascending, for the candidate list of the dynamic-ordering step
// (the reference's AW_SRT1/AW_SRT2, src/shared/atypes.F90:218-355, restated as
// a radix sort over keys that encode its AW_CMP order; see weight_key()).
//
// 8-bit digits; eight passes (four for keys whose low 32 bits are zero), each
// three kernels:
//   count   -- one 4096-key tile per workgroup: LDS histogram of the pass's
//              digit -> tcount[digit][tile].  The first pass reads the keys
//              (and counts the NaN keys and finds the smallest key on the way);
//              every later pass reads one byte per key, the digit the pass
//              before it wrote beside each record
//   scan    -- one workgroup per digit: exclusive scan of its row of tile
//              counts, plus the digit total
//   scatter -- stable rank of each key inside its tile: each wave ranks its
//              own quarter of the tile against wave-private digit counters
//              (8 ballots per key give the lanes with the same digit; no
//              barrier), the per-wave counts give each wave's offset inside
//              each digit; the tile is staged in LDS in digit order and written
//              out in that order, so each digit's run of the tile leaves as
//              consecutive addresses (base[digit] + tile offset + rank)
// Between passes a key travels as one record: its value above the key bits
// that later passes still sort on (the consumed low bits are dropped), in 12
// bytes while it needs more than 64 bits, then 8, then 4, and its next digit
// as one byte (the low byte of the record, written again on its own so that
// the next count reads 1 byte per key).  The last pass writes the values alone.
// At N = 4096 (8.4M keys, 23-bit values) a sort reads and writes 8 (count) +
// 8+13 + 1 + 12+13 + 1 + 12+9 + 1 + 8+9 + 1 + 8+9 + 1 + 8+9 + 1 + 8+5 + 1 + 4+4
// = 133 bytes per key, against 8 x (8 + 24) = 256 for eight passes of
// (key, value) pairs with a histogram read of the keys in each.
// For the step (pair values) the first pass leaves the NaN keys out -- the
// walk stops at the first of them -- and the later passes carry only the
// others (their count read on the device), which late in a run, when most
// pairs are negligible, is a small part of the list.  Stability keeps the
// finite keys in the order a sort of the whole list gives them.
// Every launch is the same for every step (a fixed sequence a graph replays);
// the key count is read on the device.
#include "jkg_sort.hpp"
#include "jkg_device.hpp"

namespace jkg {

constexpr int SORT_TILE = 4096;  // keys per tile
constexpr int RADIX = 256;
// count: 256 threads x 16 keys
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = SORT_TILE / SORT_THREADS;
// scatter: 512 threads x 8 keys
constexpr int SC_THREADS = 512;
// scan: counts per lane loaded together
constexpr int SCAN_B = 8;
constexpr int SC_ITEMS = SORT_TILE / SC_THREADS;

// record formats: the input keys (and values); records of up to 96 bits as three 32-bit words
// (12 bytes, one array), of up to 64 bits as words, of up to 32 bits as 32-bit words; the values
enum : int { F_ORIG = 0, F_96 = 1, F_64 = 2, F_32 = 3, F_VAL = 4 };
struct R12 {
  uint32_t w0, w1, w2;  // bits 0-31, 32-63, 64-95
};

// scratch layout (bytes): SortStats | digit totals | per-digit rows of tile counts | per tile of
// the first pass: NaN keys, smallest key
constexpr size_t SC_TOTAL = 64;
constexpr size_t SC_TILES = SC_TOTAL + RADIX * 4;

static int tiles_for(int n) {
  const int t = (n + SORT_TILE - 1) / SORT_TILE;
  return t < 1 ? 1 : t;
}

static size_t sc_tmin_off(int ntiles) { return (SC_TILES + (size_t)RADIX * ntiles * 4 + (size_t)ntiles * 4 + 7) & ~(size_t)7; }
size_t sort_scratch_bytes(int nmax) { return sc_tmin_off(tiles_for(nmax)) + (size_t)tiles_for(nmax) * 8; }
// whole tiles: the count reads 16 digits per lane, past n inside the allocation
size_t sort_digit_bytes(int nmax) { return (size_t)tiles_for(nmax) * SORT_TILE; }

__device__ __forceinline__ SortStats *sc_stats(int *s) { return (SortStats *)s; }
__device__ __forceinline__ uint32_t *sc_total(int *s) { return (uint32_t *)((char *)s + SC_TOTAL); }
__device__ __forceinline__ uint32_t *sc_tiles(int *s) { return (uint32_t *)((char *)s + SC_TILES); }
__device__ __forceinline__ int *sc_tnan(int *s, int ntiles) {
  return (int *)((char *)s + SC_TILES + (size_t)RADIX * ntiles * 4);
}
__device__ __forceinline__ unsigned long long *sc_tmin(int *s, int ntiles) {
  return (unsigned long long *)((char *)s +
                                ((SC_TILES + (size_t)RADIX * ntiles * 4 + (size_t)ntiles * 4 + 7) & ~(size_t)7));
}

// Tile histogram of the pass's digit.  FIRST: from the keys (digit at bit
// `shift`), with the NaN count and the smallest key (drop: the NaN keys are not
// counted in the histogram); otherwise from the digit bytes the previous pass wrote.
template <bool FIRST>
__global__ void __launch_bounds__(SORT_THREADS) k_sort_count(const uint64_t *__restrict__ keys,
                                                             const uint8_t *__restrict__ dig, const int *n_ptr,
                                                             int *scratch, int ntiles, int shift, int drop) {
  // HC copies of the histogram, lane l adding to copy l mod HC: the top digits (sign and
  // exponent bits) fall into few bins, and same-address LDS atomics of a wave are serialized
  constexpr int HC = 4;
  __shared__ uint32_t hc[HC][RADIX];
  __shared__ int snan;
  __shared__ unsigned long long smin;
#pragma unroll
  for (int c = 0; c < HC; ++c) hc[c][threadIdx.x] = 0u;
  uint32_t *h = hc[threadIdx.x & (HC - 1)];
  if (FIRST && threadIdx.x == 0) {
    snan = 0;
    smin = ~0ull;
  }
  __syncthreads();
  const int n = *n_ptr;
  const int tiles = (n + SORT_TILE - 1) / SORT_TILE;
  const int t = blockIdx.x;
  if (t < tiles) {
    const int base = t * SORT_TILE;
    if constexpr (FIRST) {
      uint64_t k[SORT_ITEMS];
#pragma unroll
      for (int r = 0; r < SORT_ITEMS; ++r) {
        const int idx = base + r * SORT_THREADS + threadIdx.x;
        k[r] = (idx < n) ? keys[idx] : ~0ull;
      }
      int nan = 0;
      unsigned long long mn = ~0ull;
#pragma unroll
      for (int r = 0; r < SORT_ITEMS; ++r) {
        if (base + r * SORT_THREADS + (int)threadIdx.x < n) {
          const bool isnan = key_nan(k[r]);
          if (!(drop && isnan)) atomicAdd(&h[(int)((k[r] >> shift) & 0xFF)], 1u);
          nan += isnan ? 1 : 0;
          mn = (k[r] < mn) ? k[r] : mn;
        }
      }
      for (int o = 32; o > 0; o >>= 1) {
        nan += __shfl_down(nan, o);
        const unsigned long long m2 = __shfl_down(mn, o);
        mn = (m2 < mn) ? m2 : mn;
      }
      if ((threadIdx.x & 63) == 0) {
        if (nan) atomicAdd(&snan, nan);
        atomicMin(&smin, mn);
      }
    } else {
      uint32_t w[SORT_ITEMS / 4];
      const uint32_t *dw = (const uint32_t *)(dig + base + SORT_ITEMS * threadIdx.x);
#pragma unroll
      for (int i = 0; i < SORT_ITEMS / 4; ++i) w[i] = dw[i];
      const int i0 = base + SORT_ITEMS * threadIdx.x;
#pragma unroll
      for (int i = 0; i < SORT_ITEMS; ++i)
        if (i0 + i < n) atomicAdd(&h[(w[i >> 2] >> (8 * (i & 3))) & 0xFF], 1u);
    }
  }
  __syncthreads();
  uint32_t hsum = 0u;
#pragma unroll
  for (int c = 0; c < HC; ++c) hsum += hc[c][threadIdx.x];
  sc_tiles(scratch)[threadIdx.x * ntiles + t] = hsum;  // zero for tiles past n
  if (FIRST && threadIdx.x == 0) {
    // per tile (one global atomic per workgroup on one address would queue them all)
    sc_tnan(scratch, ntiles)[t] = snan;
    sc_tmin(scratch, ntiles)[t] = smin;
  }
}

// one workgroup per digit value; first: workgroup 0 also sums the NaN counts and finds the
// smallest key over the tiles of the first pass, and sets every field of the sort's statistics
// (drop: the later passes carry the other keys)
__global__ void __launch_bounds__(SORT_THREADS) k_sort_scan(int *scratch, int ntiles, int first, int drop,
                                                            const int *n_ptr) {
  __shared__ int part[SORT_THREADS];
  __shared__ unsigned long long pmin[SORT_THREADS];
  const int d = blockIdx.x;
  if (first && d == 0) {
    int nan = 0;
    unsigned long long mn = ~0ull;
    for (int b = 0; b < ntiles; b += SCAN_B * SORT_THREADS) {  // SCAN_B tiles per lane in flight
      int tn[SCAN_B];
      unsigned long long tm[SCAN_B];
#pragma unroll
      for (int k = 0; k < SCAN_B; ++k) {
        const int i = min(b + k * SORT_THREADS + (int)threadIdx.x, ntiles - 1);
        tn[k] = sc_tnan(scratch, ntiles)[i];
        tm[k] = sc_tmin(scratch, ntiles)[i];
      }
#pragma unroll
      for (int k = 0; k < SCAN_B; ++k)
        if (b + k * SORT_THREADS + (int)threadIdx.x < ntiles) {
          nan += tn[k];
          mn = (tm[k] < mn) ? tm[k] : mn;
        }
    }
    part[threadIdx.x] = nan;
    pmin[threadIdx.x] = mn;
    __syncthreads();
    for (int o = SORT_THREADS / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        part[threadIdx.x] += part[threadIdx.x + o];
        pmin[threadIdx.x] = (pmin[threadIdx.x + o] < pmin[threadIdx.x]) ? pmin[threadIdx.x + o] : pmin[threadIdx.x];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      SortStats *ss = sc_stats(scratch);
      const int n = *n_ptr;
      ss->n = n;
      ss->nan_count = part[0];
      ss->minkey = pmin[0];
      ss->nsorted = drop ? n - part[0] : n;
    }
    __syncthreads();
  }
  uint32_t *row = sc_tiles(scratch) + (size_t)d * ntiles;
  const int per = (ntiles + SORT_THREADS - 1) / SORT_THREADS;
  const int lo = threadIdx.x * per;
  // a lane's counts SCAN_B at a time, kept in registers for the prefix pass when they fit
  int v[SCAN_B];
  auto load_counts = [&](int b) {
#pragma unroll
    for (int k = 0; k < SCAN_B; ++k) {
      const int i = lo + b + k;
      v[k] = (b + k < per && i < ntiles) ? (int)row[min(i, ntiles - 1)] : 0;
    }
  };
  int s = 0;
  for (int b = 0; b < per; b += SCAN_B) {
    load_counts(b);
#pragma unroll
    for (int k = 0; k < SCAN_B; ++k) s += v[k];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  // inclusive Hillis-Steele scan of the 256 partial sums
  for (int off = 1; off < SORT_THREADS; off <<= 1) {
    const int v = (threadIdx.x >= off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = part[threadIdx.x] - s;
  for (int b = 0; b < per; b += SCAN_B) {
    if (per > SCAN_B) load_counts(b);  // else v holds the lane's counts already
#pragma unroll
    for (int k = 0; k < SCAN_B; ++k)
      if (b + k < per && lo + b + k < ntiles) {
        row[lo + b + k] = (uint32_t)run;
        run += v[k];
      }
  }
  if (threadIdx.x == SORT_THREADS - 1) sc_total(scratch)[d] = (uint32_t)part[SORT_THREADS - 1];
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

// Exclusive prefix sum of v over the first 256 threads (digit = thread; the others pass 0 and
// get 0); every thread of the workgroup must call it.
__device__ __forceinline__ int block_excl_scan256(int v, int *wsum) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  __syncthreads();  // wsum may still be read by a previous scan
  if (lane == 63 && wave < RADIX / 64) wsum[wave] = inc;
  __syncthreads();
  int before = 0;
#pragma unroll
  for (int w = 0; w < RADIX / 64; ++w) before += (w < wave) ? wsum[w] : 0;
  return (threadIdx.x < RADIX) ? before + inc - v : 0;
}

struct PassArgs {
  const uint64_t *k0;
  const uint32_t *vin;
  uint32_t *v0;
  const uint64_t *kin;
  const uint32_t *vin_rec;
  uint64_t *kout;
  uint32_t *vout;
  uint8_t *dout;
  int pair_n;
  int drop; // IN == F_ORIG: the NaN keys are left out
  int ds;   // bit of this pass's digit in the input record
  int sh;   // bits dropped from the record (OUT != F_VAL)
  int vsh;  // bit of the value in the input record (OUT == F_VAL)
};

// One pass over the tiles: IN/OUT record formats.
// LDS of a scatter: the staged tile -- records of 8 bytes (OUT = F_96, F_64) or 4, then per slot
// the high word with the digit in its top byte (F_96) or the digit alone -- shares its bytes with
// the ranking's lane masks (used before the staging), so a pass writing 8-byte records or less
// fits three workgroups on a CU (a 12-byte one two)
template <int OUT>
struct ScatterLds {
  static constexpr size_t REC = (OUT == F_96 || OUT == F_64) ? 8 : 4;
  static constexpr size_t HI = (OUT == F_96) ? 4 : 1;
  static constexpr size_t STAGE = (size_t)SORT_TILE * (REC + HI);
  static constexpr size_t LANES = (size_t)(SC_THREADS / 64) * RADIX * 8;
  static constexpr size_t BYTES = STAGE > LANES ? STAGE : LANES;
  static constexpr int WAVES_PER_SIMD = (OUT == F_96) ? 4 : 6;  // two or three workgroups per CU
};

template <int IN, int OUT>
__global__ void __launch_bounds__(SC_THREADS, ScatterLds<OUT>::WAVES_PER_SIMD)
    k_sort_scatter(PassArgs a, const int *n_ptr, int *scratch, int ntiles) {
  constexpr int NW = SC_THREADS / 64;
  typedef ScatterLds<OUT> SL;
  __shared__ __attribute__((aligned(16))) uint8_t sbuf[SL::BYTES];
  uint64_t(*lanes)[RADIX] = (uint64_t(*)[RADIX])sbuf;  // per wave: the lanes whose key of the current round has digit d
  uint64_t *sL8 = (uint64_t *)sbuf;                     // staged records (after the ranking)
  uint32_t *sL4 = (uint32_t *)sbuf;
  uint32_t *sH4 = (uint32_t *)(sbuf + (size_t)SORT_TILE * SL::REC);
  uint8_t *sD = sbuf + (size_t)SORT_TILE * SL::REC;
  __shared__ int wcnt[NW][RADIX];       // per-wave digit counters, then the first slot of wave w's keys of digit d
  __shared__ int gdelta[RADIX];         // global position of a staged slot of digit d, minus the slot
  __shared__ int wsum[RADIX / 64];
  __shared__ int staged;                // keys of the tile written (all, or those not dropped)
  const int n = *n_ptr;
  const int tiles = (n + SORT_TILE - 1) / SORT_TILE;
  const int t = blockIdx.x;
  if (t >= tiles) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tlen = (n - t * SORT_TILE < SORT_TILE) ? n - t * SORT_TILE : SORT_TILE;
  const int wbase = wave * (SORT_TILE / NW);
  const int ds = a.ds;
  // the digit totals and this tile's offsets (thread d < 256: digit d), loaded with the records
  int dtot = 0, doff = 0;
  if (threadIdx.x < RADIX) {
    dtot = (int)sc_total(scratch)[threadIdx.x];
    doff = (int)sc_tiles(scratch)[(size_t)threadIdx.x * ntiles + t];
  }

  uint64_t L[SC_ITEMS];
  uint32_t H[SC_ITEMS];
  uint32_t vmask = 0;  // the items ranked and written
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) {
    const int li = wbase + r * 64 + lane;
    const int idx = t * SORT_TILE + li;
    const bool valid = li < tlen;
    vmask |= (valid ? 1u : 0u) << r;
    if constexpr (IN == F_ORIG) {
      L[r] = valid ? a.k0[idx] : ~0ull;
      H[r] = valid ? (a.vin ? a.vin[idx] : (uint32_t)idx) : 0u;
    } else if constexpr (IN == F_96) {
      R12 v = {~0u, ~0u, 0u};
      if (valid) v = ((const R12 *)a.kin)[idx];
      L[r] = (uint64_t)v.w0 | ((uint64_t)v.w1 << 32);
      H[r] = v.w2;
    } else if constexpr (IN == F_64) {
      L[r] = valid ? a.kin[idx] : ~0ull;
      H[r] = 0u;
    } else {
      L[r] = valid ? (uint64_t)a.vin_rec[idx] : ~0ull;
      H[r] = 0u;
    }
  }
  if constexpr (IN == F_ORIG) {
    if (a.drop) {
#pragma unroll
      for (int r = 0; r < SC_ITEMS; ++r)
        if (key_nan(L[r])) vmask &= ~(1u << r);
    }
  }
#pragma unroll
  for (int i = 0; i < RADIX / 64; ++i) {
    wcnt[wave][lane + 64 * i] = 0;
    lanes[wave][lane + 64 * i] = 0ull;
  }

  // Stable rank of each key among the wave's keys of the same digit, one round of 64 keys at a
  // time: every lane ORs its bit into the mask of its digit, reads the mask back (the lanes with
  // the same digit), and the lowest of those lanes moves the digit's counter on and clears the
  // mask.  A wave's LDS accesses are carried out in order, so no barrier is needed.
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint64_t me = 1ull << lane;
  int rank[SC_ITEMS];
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) {
    const bool valid = (vmask >> r) & 1u;
    const int d = (int)((L[r] >> ds) & 0xFF);
    if (valid) __hip_atomic_fetch_or(&lanes[wave][d], me, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    __builtin_amdgcn_wave_barrier();
    const uint64_t match = valid ? __hip_atomic_load(&lanes[wave][d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT)
                                 : 0ull;
    const int before = valid ? wcnt[wave][d] : 0;
    const int g = __popcll(match & lt_mask);
    rank[r] = before + g;
    __builtin_amdgcn_wave_barrier();
    if (valid && g == 0) {
      wcnt[wave][d] = before + __popcll(match);
      __hip_atomic_store(&lanes[wave][d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  {
    // digit d = thread (the first 256): the waves' offsets inside the digit, the tile's digit
    // starts, the global starts
    const int d = threadIdx.x;
    int c = 0;
    if (d < RADIX) {
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const int x = wcnt[w][d];
        wcnt[w][d] = c;
        c += x;
      }
    }
    const int ls = block_excl_scan256(c, wsum);  // first slot of digit d in the staged tile
    if (d == RADIX - 1) staged = ls + c;
    if (d < RADIX) {
#pragma unroll
      for (int w = 0; w < NW; ++w) wcnt[w][d] += ls;
    }
    const int gs = block_excl_scan256(dtot, wsum);
    if (d < RADIX) gdelta[d] = gs + doff - ls;
  }
  __syncthreads();
  // stage in digit order: the output record (or value) and, in the top byte of the high word,
  // the digit (an output record needs at most 88 bits, so its high word at most 24)
#pragma unroll
  for (int r = 0; r < SC_ITEMS; ++r) {
    if ((vmask >> r) & 1u) {
      const int d = (int)((L[r] >> ds) & 0xFF);
      const int slot = wcnt[wave][d] + rank[r];
      const unsigned __int128 rec = (unsigned __int128)L[r] | ((unsigned __int128)H[r] << 64);
      uint64_t oL;
      uint32_t oH = 0u;
      if constexpr (OUT == F_VAL) {
        uint32_t v = (uint32_t)(rec >> a.vsh);
        if (a.pair_n > 0) {
          int p, q;
          pair_of(a.pair_n, v, p, q);
          v = pack_pq(p - 1, q - 1);
        }
        oL = v;
      } else {
        const unsigned __int128 o = rec >> a.sh;
        oL = (uint64_t)o;
        if constexpr (OUT == F_96) oH = (uint32_t)(o >> 64);
      }
      if constexpr (OUT == F_96) {
        sL8[slot] = oL;
        sH4[slot] = oH | ((uint32_t)d << 24);
      } else if constexpr (OUT == F_64) {
        sL8[slot] = oL;
        sD[slot] = (uint8_t)d;
      } else {
        sL4[slot] = (uint32_t)oL;
        sD[slot] = (uint8_t)d;
      }
    }
  }
  __syncthreads();
  // write out in staged order: neighbouring lanes hold neighbouring slots; the low byte of an
  // output record is the next pass's digit
  const int nst = staged;
  for (int i = threadIdx.x; i < nst; i += SC_THREADS) {
    const uint32_t hd = (OUT == F_96) ? sH4[i] : (uint32_t)sD[i] << 24;
    const uint64_t l = (OUT == F_96 || OUT == F_64) ? sL8[i] : (uint64_t)sL4[i];
    const int pos = gdelta[hd >> 24] + i;
    if constexpr (OUT == F_96) {
      ((R12 *)a.kout)[pos] = R12{(uint32_t)l, (uint32_t)(l >> 32), hd & 0x00FFFFFFu};
    } else if constexpr (OUT == F_64) {
      a.kout[pos] = l;
    } else if constexpr (OUT == F_32) {
      a.vout[pos] = (uint32_t)l;
    } else {
      a.v0[pos] = (uint32_t)l;
    }
    if constexpr (OUT != F_VAL) a.dout[pos] = (uint8_t)l;
  }
}

static int fmt_for_width(int w) { return w <= 32 ? F_32 : (w <= 64 ? F_64 : F_96); }

static void launch_scatter(int in, int out, const PassArgs &a, const int *n_dev, int *scratch, int ntiles,
                           hipStream_t st) {
#define JKG_SORT_CASE(I, O)                                                                       \
  if (in == I && out == O) {                                                                      \
    hipLaunchKernelGGL((k_sort_scatter<I, O>), dim3(ntiles), dim3(SC_THREADS), 0, st, a, n_dev,   \
                       scratch, ntiles);                                                          \
    return;                                                                                       \
  }
  JKG_SORT_CASE(F_ORIG, F_96)
  JKG_SORT_CASE(F_ORIG, F_64)
  JKG_SORT_CASE(F_ORIG, F_32)
  JKG_SORT_CASE(F_ORIG, F_VAL)
  JKG_SORT_CASE(F_96, F_96)
  JKG_SORT_CASE(F_96, F_64)
  JKG_SORT_CASE(F_96, F_32)
  JKG_SORT_CASE(F_96, F_VAL)
  JKG_SORT_CASE(F_64, F_64)
  JKG_SORT_CASE(F_64, F_32)
  JKG_SORT_CASE(F_64, F_VAL)
  JKG_SORT_CASE(F_32, F_32)
  JKG_SORT_CASE(F_32, F_VAL)
#undef JKG_SORT_CASE
}

void sort_pairs_launch(SortBuffers &b, const int *n_dev, int nmax, hipStream_t st, int pair_n, int lo_bits) {
  const int ntiles = tiles_for(nmax);
  int vbits = 32;
  if (!b.vin) {
    vbits = 0;
    while (vbits < 32 && ((uint64_t)1 << vbits) < (uint64_t)(nmax > 1 ? nmax : 1)) ++vbits;
  }
  const int first = (lo_bits == 32) ? 4 : 0;
  uint64_t *kb[2] = {b.k1, b.k2};
  uint32_t *vb[2] = {b.v1, b.v2};
  uint8_t *db[2] = {b.d1, b.d2};
  const int drop = pair_n > 0 ? 1 : 0;
  const int *n_rest = &sort_stats(b)->nsorted;  // the passes after the first
  for (int pass = first; pass < 8; ++pass) {
    const int r = pass - first;  // passes run before this one
    const int lo_in = (r == 0) ? 0 : 8 * pass;
    const int in = (r == 0) ? F_ORIG : fmt_for_width(64 - lo_in + vbits);
    const int out = (pass == 7) ? F_VAL : fmt_for_width(64 - 8 * (pass + 1) + vbits);
    if (r == 0)
      hipLaunchKernelGGL(k_sort_count<true>, dim3(ntiles), dim3(SORT_THREADS), 0, st, (const uint64_t *)b.k0,
                         (const uint8_t *)nullptr, n_dev, b.scratch, ntiles, 8 * pass, drop);
    else
      hipLaunchKernelGGL(k_sort_count<false>, dim3(ntiles), dim3(SORT_THREADS), 0, st, (const uint64_t *)nullptr,
                         (const uint8_t *)db[(r - 1) & 1], n_rest, b.scratch, ntiles, 0, 0);
    hipLaunchKernelGGL(k_sort_scan, dim3(RADIX), dim3(SORT_THREADS), 0, st, b.scratch, ntiles, r == 0 ? 1 : 0,
                       drop, n_dev);
    PassArgs a;
    a.k0 = b.k0;
    a.vin = b.vin;
    a.v0 = b.v0;
    a.kin = kb[(r - 1) & 1];
    a.vin_rec = vb[(r - 1) & 1];
    a.kout = kb[r & 1];
    a.vout = vb[r & 1];
    a.dout = db[r & 1];
    a.pair_n = pair_n;
    a.drop = drop;
    a.ds = 8 * pass - lo_in;
    a.sh = 8 * (pass + 1) - lo_in;
    a.vsh = 64 - lo_in;
    launch_scatter(in, out, a, r == 0 ? n_dev : n_rest, b.scratch, ntiles, st);
  }
}

}  // namespace jkg
