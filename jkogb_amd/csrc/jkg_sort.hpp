// jkg_sort.hpp -- stable LSD radix sort of (uint64 key, uint32 value) pairs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace jkg {

struct SortBuffers {
  uint64_t *k0, *k1;   // ping-pong key buffers; the input keys are in k0
  uint32_t *v0, *v1;   // ping-pong value buffers (the input values are implicit: index)
  int *scratch;        // sort_scratch_bytes(nmax) bytes
};

struct SortResult {
  const uint64_t *keys;
  const uint32_t *vals;
  bool vals_iota;      // every pass was the identity: value = index
};

size_t sort_scratch_bytes(int nmax);
// Sorts the first *n_dev keys of b.k0 (n read on the device; nmax bounds the
// grid).  The sorted data end up in the buffer given by sort_result().
// vals_given: the input values are in b.v0 (otherwise value = index).
void sort_pairs_launch(SortBuffers &b, const int *n_dev, int nmax, hipStream_t st, int vals_given = 0);

// Location of the per-pass skip flags inside the scratch area.
inline int *sort_skip_flags(const SortBuffers &b, int nmax) {
  int ntiles = (nmax + 4095) / 4096;
  if (ntiles < 1) ntiles = 1;
  return b.scratch + (size_t)256 * ntiles + 256;
}

// skip[0..7]: pass skipped flags; skip[8]: the input values were given.
__device__ __forceinline__ SortResult sort_result_dev(const uint64_t *k0, const uint64_t *k1, const uint32_t *v0,
                                                      const uint32_t *v1, const int *skip) {
  int moved = 0;
  for (int i = 0; i < 8; ++i) moved += skip[i] ? 0 : 1;
  SortResult r;
  r.keys = (moved & 1) ? k1 : k0;
  r.vals = (moved & 1) ? v1 : v0;
  r.vals_iota = (moved == 0) && !skip[8];
  return r;
}

}  // namespace jkg
