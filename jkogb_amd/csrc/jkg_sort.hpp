// jkg_sort.hpp -- stable LSD radix sort of 64-bit keys carrying 32-bit values.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace jkg {

// The input keys stay where they are (k0, unsorted): the sort produces the
// sorted values only, and the key of sorted entry i is the input key at the
// index the value names (for the dynamic-ordering step: the pair index of the
// packed pair).  Between passes each key travels as one record holding its
// value above the key bits not sorted yet, in 12, 8 or 4 bytes (k1/k2 or
// v1/v2), and its next digit as one byte (d1/d2) for the next pass's counts.
struct SortBuffers {
  uint64_t *k0;        // input keys (read, never written)
  uint32_t *v0;        // output: the sorted values (capacity n + padding of the caller's choice)
  uint64_t *k1, *k2;   // ping-pong records of 12 or 8 bytes (2 nmax words each)
  uint32_t *v1, *v2;   // ping-pong records of 4 bytes
  uint8_t *d1, *d2;    // ping-pong next digits (sort_digit_bytes(nmax) bytes each)
  const uint32_t *vin; // input values (null: the value of key i is i)
  int *scratch;        // sort_scratch_bytes(nmax) bytes
};

// Per-sort figures the first pass finds on the way, for the selection.
struct SortStats {
  int nan_count;             // keys >= 0xFFFFFFFF00000000 (the step's NaN / dropped keys, sorted last)
  int n;                     // keys given
  unsigned long long minkey; // smallest key (~0 when there are none)
  int nsorted;               // keys the passes after the first carry: n, or n - nan_count (pair_n > 0)
};

size_t sort_scratch_bytes(int nmax);
size_t sort_digit_bytes(int nmax);
inline SortStats *sort_stats(const SortBuffers &b) { return (SortStats *)b.scratch; }

// Sorts the first *n_dev keys of b.k0 (n read on the device; nmax bounds the
// grid) by their bits [lo_bits, 64) -- lo_bits = 0 or 32 (keys whose low half
// is zero) -- and writes the sorted values to b.v0.  pair_n > 0: the values are
// pair indices k of an N = pair_n problem (the input values must be implicit)
// and leave as packed pairs pack_pq(p - 1, q - 1), and the NaN keys are dropped
// by the first pass (only the first n - nan_count values are written: the step
// never reads past them); pair_n = 0: values as given, every key sorted.
void sort_pairs_launch(SortBuffers &b, const int *n_dev, int nmax, hipStream_t st, int pair_n, int lo_bits = 0);

}  // namespace jkg
