// tg_prims.hip -- hipCUB-backed device primitives (see tg_prims.h).
#include "tg_prims.h"

#include <hipcub/hipcub.hpp>

#include <stdexcept>
#include <string>

namespace tg {

namespace {

void check(hipError_t e, const char *what)
{
  if (e != hipSuccess) throw std::runtime_error(std::string("tioga_amd: ") + what + ": " + hipGetErrorString(e));
}

thread_local PrimTemp g_temp;
thread_local int *g_count = nullptr;

}  // namespace

void *PrimTemp::get(size_t want)
{
  if (want > n) {
    if (p) check(hipFree(p), "hipFree");
    p = nullptr;
    n = 0;
    check(hipMalloc(&p, want), "hipMalloc(temp)");
    n = want;
  }
  return p;
}

PrimTemp::~PrimTemp()
{
  if (p) (void)hipFree(p);
}

size_t sort_pairs_bytes(int n, int bits)
{
  size_t bytes = 0;
  check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                           (const int *)nullptr, (int *)nullptr, n, 0, bits),
        "sort(size)");
  return bytes;
}

int select_flagged(const int *flags, int n, int *out, hipStream_t s)
{
  if (n <= 0) return 0;
  if (!g_count) check(hipMalloc(&g_count, sizeof(int)), "hipMalloc(count)");
  hipcub::CountingInputIterator<int> it(0);
  size_t bytes = 0;
  check(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flags, out, g_count, n, s), "select(size)");
  check(hipcub::DeviceSelect::Flagged(g_temp.get(bytes), bytes, it, flags, out, g_count, n, s), "select");
  int c = 0;
  check(hipMemcpyAsync(&c, g_count, sizeof(int), hipMemcpyDeviceToHost, s), "count D2H");
  check(hipStreamSynchronize(s), "sync");
  return c;
}

void sort_pairs(const unsigned *keys_in, unsigned *keys_out, const int *vals_in, int *vals_out, int n,
                int bits, hipStream_t s, PrimTemp *temp)
{
  if (n <= 0) return;
  size_t bytes = 0;
  check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, n, 0, bits, s),
        "sort(size)");
  PrimTemp &t = temp ? *temp : g_temp;
  check(hipcub::DeviceRadixSort::SortPairs(t.get(bytes), bytes, keys_in, keys_out, vals_in, vals_out, n, 0, bits, s),
        "sort");
}

void exclusive_sum(const int *in, int *out, int n, hipStream_t s)
{
  if (n <= 0) return;
  size_t bytes = 0;
  check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s), "scan(size)");
  check(hipcub::DeviceScan::ExclusiveSum(g_temp.get(bytes), bytes, in, out, n, s), "scan");
}

}  // namespace tg
