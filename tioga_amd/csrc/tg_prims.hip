// tg_prims.hip -- device primitives over rocPRIM (see tg_prims.h).
#include "tg_prims.h"
#include "tg_engine.h"

#include <rocprim/rocprim.hpp>

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
    if (p) {
      check(hipFree(p), "hipFree");
      thread_dev_bytes().add(-(long long)n);
    }
    p = nullptr;
    n = 0;
    check(hipMalloc(&p, want), "hipMalloc(temp)");
    n = want;
    thread_dev_bytes().add((long long)want);
  }
  return p;
}

PrimTemp::~PrimTemp()
{
  if (p) {
    (void)hipFree(p);
    thread_dev_bytes().add(-(long long)n);
  }
}

size_t sort_pairs_bytes(int n, int bits)
{
  size_t bytes = 0;
  check(rocprim::radix_sort_pairs(nullptr, bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                  (const int *)nullptr, (int *)nullptr, (size_t)n, 0u, (unsigned)bits),
        "sort(size)");
  return bytes;
}

int select_flagged(const int *flags, int n, int *out, hipStream_t s)
{
  if (n <= 0) return 0;
  if (!g_count) check(hipMalloc(&g_count, sizeof(int)), "hipMalloc(count)");
  rocprim::counting_iterator<int> it(0);
  size_t bytes = 0;
  check(rocprim::select(nullptr, bytes, it, flags, out, g_count, (size_t)n, s), "select(size)");
  check(rocprim::select(g_temp.get(bytes), bytes, it, flags, out, g_count, (size_t)n, s), "select");
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
  check(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0u,
                                  (unsigned)bits, s),
        "sort(size)");
  PrimTemp &t = temp ? *temp : g_temp;
  check(rocprim::radix_sort_pairs(t.get(bytes), bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0u,
                                  (unsigned)bits, s),
        "sort");
}

void exclusive_sum(const int *in, int *out, int n, hipStream_t s)
{
  if (n <= 0) return;
  size_t bytes = 0;
  check(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s), "scan(size)");
  check(rocprim::exclusive_scan(g_temp.get(bytes), bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s), "scan");
}

}  // namespace tg
