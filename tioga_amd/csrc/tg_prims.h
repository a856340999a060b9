// tg_prims.h -- device-wide primitives (stable compaction, stable radix sort
// of key/value pairs) used by the connectivity stages.  Implemented over
// rocPRIM in tg_prims.hip so the rest of the engine compiles without it.
#pragma once

#include <hip/hip_runtime.h>

namespace tg {

struct DevBytes;
DevBytes &thread_dev_bytes();  // tg_engine.h

// Device scratch for the library calls; grows on demand and is kept.  Work
// issued on several streams at once gives each stream its own.
struct PrimTemp {
  void *p = nullptr;
  size_t n = 0;
  void *get(size_t want);
  PrimTemp() = default;
  PrimTemp(const PrimTemp &) = delete;
  PrimTemp &operator=(const PrimTemp &) = delete;
  ~PrimTemp();
};

// scratch bytes sort_pairs needs for n pairs
size_t sort_pairs_bytes(int n, int bits);

// out[0..count) = the indices i < n with flags[i] != 0, ascending; returns
// count (synchronises the stream to read it).
int select_flagged(const int *flags, int n, int *out, hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on the low `bits` key bits.
void sort_pairs(const unsigned *keys_in, unsigned *keys_out, const int *vals_in, int *vals_out,
                int n, int bits, hipStream_t s, PrimTemp *temp = nullptr);

// out[i] = sum of in[0..i) (exclusive prefix sum).
void exclusive_sum(const int *in, int *out, int n, hipStream_t s);

}  // namespace tg
