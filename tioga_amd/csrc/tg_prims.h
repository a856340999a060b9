// tg_prims.h -- device-wide primitives (stable compaction, stable radix sort
// of key/value pairs) used by the connectivity stages.  Implemented over
// hipCUB in tg_prims.hip so the rest of the engine compiles without it.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

// out[0..count) = the indices i < n with flags[i] != 0, ascending; returns
// count (synchronises the stream to read it).
int select_flagged(const int *flags, int n, int *out, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits.
void sort_pairs(const unsigned *keys_in, unsigned *keys_out, const int *vals_in, int *vals_out,
                int n, int bits, hipStream_t s);

// out[i] = sum of in[0..i) (exclusive prefix sum).
void exclusive_sum(const int *in, int *out, int n, hipStream_t s);

}  // namespace tg
