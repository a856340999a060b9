// tg_adt.hip -- the ADT of ADT.C:34-135 / buildADTrecursion.c:24-116 built
// on the device, node for node identical to the reference's.
//
// The reference builds the tree recursively: a node is the element at the
// median position (nav+1)/2-1 of its segment after median.F90's partial
// quicksort along dimension level%6 reorders the segment, its box bounds the
// segment, and the left / right subtrees are the segment's two halves.  The
// quicksort's exact swap order decides every later level, so it is kept.
//
// Here all segments of one level are processed in one launch.  A segment
// longer than SMALL gets a workgroup that runs the quicksort's partition
// rounds in list form (hoare_stop in tg_common.h): flags X >= pivot and
// X <= pivot over the active range, workgroup scans turn them into the
// ascending list A and descending list B, the first K pairs with A[m] < B[m]
// are swapped in parallel, and one thread applies the loop control.
// Segments of at most SMALL elements are finished by one thread each with
// the sequential quicksort.  Node ids are the reference's preorder counter:
// the left child of node me is me+1, the right child me+nleft.
#include <stdexcept>
#include <vector>

#include "tg_engine.h"

namespace tg {

namespace {

constexpr int BT = 512;        // threads per workgroup for long segments
constexpr int PER = 8;         // consecutive positions per thread per chunk
constexpr int SMALL = 128;     // segments at most this long: one thread

struct Seg {
  int start, nav, me, parent, side, level;
};
static_assert(sizeof(Seg) <= 2 * sizeof(int4), "segment descriptors are stored in two int4 slots");

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// exclusive scan of per-thread counts a, b over the workgroup
__device__ __forceinline__ void wg_scan2(int a, int b, int &ea, int &eb, int &ta, int &tb, int *lds)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int sa = a, sb = b;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int xa = __shfl_up(sa, o), xb = __shfl_up(sb, o);
    if (lane >= o) {
      sa += xa;
      sb += xb;
    }
  }
  if (lane == 63) {
    lds[w] = sa;
    lds[nw + w] = sb;
  }
  __syncthreads();
  int oa = 0, ob = 0;
  ta = tb = 0;
  for (int q = 0; q < nw; q++) {
    int ca = lds[q], cb = lds[nw + q];
    if (q < w) {
      oa += ca;
      ob += cb;
    }
    ta += ca;
    tb += cb;
  }
  ea = oa + sa - a;
  eb = ob + sb - b;
  __syncthreads();
}

__device__ __forceinline__ int wg_sum(int v, int *lds)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane == 0) lds[w] = v;
  __syncthreads();
  int s = 0;
  for (int q = 0; q < nw; q++) s += lds[q];
  __syncthreads();
  return s;
}

struct WgState {
  int lo, hi, st;
  double pv;
};

// median_select on X[0..n) / IX[0..n) by a whole workgroup; A and B are
// n-int scratch lists.  Same permutation as the sequential function.
__device__ void wg_median(int *IX, double *X, int n, int *A, int *B, WgState &sh, int *lds, int *err)
{
  const int t = threadIdx.x;
  auto swap = [&](int a, int b) {  // 1-based positions
    double tx = X[a - 1];
    X[a - 1] = X[b - 1];
    X[b - 1] = tx;
    int ti = IX[a - 1];
    IX[a - 1] = IX[b - 1];
    IX[b - 1] = ti;
  };
  if (n < 3) {
    if (t == 0 && n == 2 && X[1] < X[0]) swap(1, 2);
    __threadfence_block();
    __syncthreads();
    return;
  }
  if (t == 0) {
    sh.lo = 1;
    sh.hi = n;
    sh.st = 0;
  }
  __syncthreads();
  // every round narrows [lo, hi], so n + 2 rounds bound the loop; the cap
  // only guards against a broken invariant (reported, never expected)
  for (int round = 0;; round++) {
    if (round > n + 2) {
      if (t == 0) atomicAdd(err, 1);
      break;
    }
    const int lo = sh.lo, hi = sh.hi;
    if (t == 0) sh.pv = median_pivot(X - 1, lo, hi);
    __syncthreads();
    const double pv = sh.pv;
    const int len = hi - lo + 1;
    int nA = 0, nB = 0;
    for (int base = 0; base < len; base += BT * PER) {
      // thread t owns PER consecutive positions from each end
      int ca = 0, cb = 0;
      unsigned fa = 0, fb = 0;
#pragma unroll
      for (int e = 0; e < PER; e++) {
        int off = base + t * PER + e;
        if (off < len) {
          if (X[lo + off - 1] >= pv) {
            fa |= 1u << e;
            ca++;
          }
          if (X[hi - off - 1] <= pv) {
            fb |= 1u << e;
            cb++;
          }
        }
      }
      int ra, rb, ta, tb;
      wg_scan2(ca, cb, ra, rb, ta, tb, lds);
#pragma unroll
      for (int e = 0; e < PER; e++) {
        int off = base + t * PER + e;
        if (fa & (1u << e)) A[nA + ra++] = lo + off;
        if (fb & (1u << e)) B[nB + rb++] = hi - off;
      }
      nA += ta;
      nB += tb;
    }
    __threadfence_block();
    __syncthreads();
    const int mab = nA < nB ? nA : nB;
    int cnt = 0;
    for (int m = t; m < mab; m += BT) cnt += (A[m] < B[m]);
    const int K = wg_sum(cnt, lds);
    for (int m = t; m < K; m += BT) swap(A[m], B[m]);
    __threadfence_block();
    __syncthreads();
    if (t == 0) {
      int i, j, l = lo, h = hi;
      hoare_stop(K, nA, nB, K < nA ? A[K] : 0, K > 0 ? A[K - 1] : 0, K < nB ? B[K] : 0, K > 0 ? B[K - 1] : 0, A[0],
                 B[0], i, j);
      sh.st = median_control(n, i, j, l, h);
      sh.lo = l;
      sh.hi = h;
    }
    __syncthreads();
    if (sh.st != 0) break;
  }
  if (t == 0 && sh.st == 2) {
    const int half = n / 2, halfp1 = half + 1;
    if (n == 2 * half) {
      if (X[halfp1 - 1] < X[half - 1]) swap(half, halfp1);
    } else if (X[sh.lo - 1] > X[sh.hi - 1]) {
      swap(sh.lo, sh.hi);
    }
  }
  __threadfence_block();
  __syncthreads();
}

// One level: a workgroup per segment of that level (2^level slots; empty
// slots have nav = 0).  Long segments are split here; short ones are queued
// for the one-thread-per-subtree kernel.
__global__ void __launch_bounds__(BT) k_adt_level(const Seg *__restrict__ in, Seg *__restrict__ out,
                                                  Seg *__restrict__ small, int *__restrict__ nsmall, int *ea,
                                                  double *X, int *A, int *B, const double *__restrict__ ebox,
                                                  int *__restrict__ adtInt, double *__restrict__ adtReal)
{
  __shared__ int lds[2 * (BT / 64)];
  __shared__ WgState sh;
  __shared__ double red[6][BT / 64];
  const int t = threadIdx.x;
  const Seg s = in[blockIdx.x];
  Seg *o = out + 2 * (size_t)blockIdx.x;
  if (s.nav <= SMALL) {
    if (t == 0) {
      if (s.nav > 0) small[atomicAdd(nsmall, 1)] = s;
      o[0] = Seg{0, 0, 0, 0, 0, 0};
      o[1] = Seg{0, 0, 0, 0, 0, 0};
    }
    return;
  }
  const int dim = s.level % 6;
  int *ix = ea + s.start;
  double *x = X + s.start;
  for (int i = t; i < s.nav; i += BT) x[i] = ebox[6 * (size_t)ix[i] + dim];
  __threadfence_block();
  __syncthreads();
  wg_median(ix, x, s.nav, A + s.start, B + s.start, sh, lds, nsmall + 1);
  // the node's box bounds the whole segment (buildADTrecursion.c:72-88)
  double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
  for (int i = t; i < s.nav; i += BT) {
    const double *b = ebox + 6 * (size_t)ix[i];
    for (int j = 0; j < 3; j++) {
      lo[j] = dmin(lo[j], b[j]);
      hi[j] = dmax(hi[j], b[j + 3]);
    }
  }
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      lo[j] = dmin(lo[j], __shfl_xor(lo[j], q));
      hi[j] = dmax(hi[j], __shfl_xor(hi[j], q));
    }
  if ((t & 63) == 0)
    for (int j = 0; j < 3; j++) {
      red[j][t >> 6] = lo[j];
      red[j + 3][t >> 6] = hi[j];
    }
  __syncthreads();
  if (t == 0) {
    double r[6];
    for (int j = 0; j < 6; j++) {
      r[j] = red[j][0];
      for (int q = 1; q < BT / 64; q++) r[j] = j < 3 ? dmin(r[j], red[j][q]) : dmax(r[j], red[j][q]);
    }
    const int nleft = (s.nav + 1) / 2, me = s.me;
    const int el = ix[nleft - 1];
    adtInt[4 * (size_t)me] = el;
    adtInt[4 * (size_t)me + 1] = adtInt[4 * (size_t)me + 2] = adtInt[4 * (size_t)me + 3] = -1;
    for (int j = 0; j < 6; j++) adtReal[6 * (size_t)me + j] = r[j];
    if (s.side > 0) adtInt[4 * (size_t)s.parent + s.side] = el;
    o[0] = Seg{s.start, nleft - 1, me + 1, me, 1, s.level + 1};
    o[1] = Seg{s.start + nleft, s.nav - nleft, me + nleft, me, 2, s.level + 1};
  }
}

// Whole subtrees of short segments, one thread each, with the sequential
// quicksort and an explicit stack (buildADTrecursion.c:24-116).
__global__ void k_adt_small(int n, const Seg *__restrict__ segs, int *ea, double *X,
                            const double *__restrict__ ebox, int *__restrict__ adtInt, double *__restrict__ adtReal)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  Seg stk[24];
  int sp = 0;
  if (segs[k].nav > 0) stk[sp++] = segs[k];
  while (sp > 0) {
    Seg s = stk[--sp];
    int *ix = ea + s.start;
    const int me = s.me;
    if (s.nav == 1) {
      adtInt[4 * (size_t)me] = ix[0];
      adtInt[4 * (size_t)me + 1] = adtInt[4 * (size_t)me + 2] = adtInt[4 * (size_t)me + 3] = -1;
      for (int j = 0; j < 6; j++) adtReal[6 * (size_t)me + j] = ebox[6 * (size_t)ix[0] + j];
      if (s.side > 0) adtInt[4 * (size_t)s.parent + s.side] = ix[0];
      continue;
    }
    double *x = X + s.start;
    const int dim = s.level % 6;
    for (int i = 0; i < s.nav; i++) x[i] = ebox[6 * (size_t)ix[i] + dim];
    median_select(ix, x, s.nav);
    const int nleft = (s.nav + 1) / 2;
    adtInt[4 * (size_t)me] = ix[nleft - 1];
    adtInt[4 * (size_t)me + 1] = adtInt[4 * (size_t)me + 2] = adtInt[4 * (size_t)me + 3] = -1;
    double r[6] = {BIGVALUE, BIGVALUE, BIGVALUE, -BIGVALUE, -BIGVALUE, -BIGVALUE};
    for (int i = 0; i < s.nav; i++) {
      const double *b = ebox + 6 * (size_t)ix[i];
      for (int j = 0; j < 3; j++) {
        r[j] = dmin(r[j], b[j]);
        r[j + 3] = dmax(r[j + 3], b[j + 3]);
      }
    }
    for (int j = 0; j < 6; j++) adtReal[6 * (size_t)me + j] = r[j];
    if (s.side > 0) adtInt[4 * (size_t)s.parent + s.side] = ix[nleft - 1];
    if (s.nav - nleft > 0) stk[sp++] = Seg{s.start + nleft, s.nav - nleft, me + nleft, me, 2, s.level + 1};
    if (nleft > 1) stk[sp++] = Seg{s.start, nleft - 1, me + 1, me, 1, s.level + 1};
  }
}

__global__ void k_iota(int n, int *p)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// slot 3 of row e: the node that holds element e (ADT.C:120-124)
__global__ void k_adt_slot3(int n, int *adtInt)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) adtInt[4 * (size_t)adtInt[4 * (size_t)i] + 3] = i;
}

}  // namespace

int build_adt_device(int n, const double *ebox, int *adtInt, double *adtReal, double ext[6], AdtScratch &w,
                     hipStream_t s)
{
  int depth = 0;
  for (int m = n; m > 0; m /= 2) depth++;
  if (n <= 0) {
    for (int i = 0; i < 3; i++) {  // ADT.C:66-86 over no elements
      double lo = BIGVALUE, hi = -BIGVALUE, d = 0.01 * (hi - lo);
      ext[2 * i] = lo - d;
      ext[2 * i + 1] = hi + d;
    }
    return 0;
  }
  w.ea.reserve(n);
  w.X.reserve(n);
  w.A.reserve(n);
  w.B.reserve(n);
  int nlev = 0;
  for (int m = n; m > SMALL; m /= 2) nlev++;
  const size_t last = (size_t)1 << nlev;
  w.segA.reserve(4 * last);
  w.segB.reserve(4 * last);
  w.small.reserve(4 * last + 2);
  w.nsmall.reserve(2);
  k_iota<<<nblocks(n, 256), 256, 0, s>>>(n, w.ea.p);
  Seg root{0, n, 0, 0, 0, 0};
  TG_HIP(hipMemcpyAsync(w.segA.p, &root, sizeof(Seg), hipMemcpyHostToDevice, s));
  TG_HIP(hipMemsetAsync(w.nsmall.p, 0, 2 * sizeof(int), s));
  Seg *in = (Seg *)w.segA.p, *out = (Seg *)w.segB.p;
  size_t count = 1;
  for (int l = 0; l < nlev; l++) {
    k_adt_level<<<(unsigned)count, BT, 0, s>>>(in, out, (Seg *)w.small.p, w.nsmall.p, w.ea.p, w.X.p, w.A.p, w.B.p,
                                              ebox, adtInt, adtReal);
    std::swap(in, out);
    count *= 2;
  }
  int cnt[2] = {0, 0};
  TG_HIP(hipMemcpyAsync(cnt, w.nsmall.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  if (cnt[1]) throw std::runtime_error("tioga_amd: ADT build: partition loop exceeded its bound");
  const int nsmall = cnt[0];
  if (nsmall > 0)
    k_adt_small<<<nblocks(nsmall, 64), 64, 0, s>>>(nsmall, (const Seg *)w.small.p, w.ea.p, w.X.p, ebox, adtInt,
                                                 adtReal);
  k_adt_small<<<nblocks((long long)count, 64), 64, 0, s>>>((int)count, in, w.ea.p, w.X.p, ebox, adtInt, adtReal);
  k_adt_slot3<<<nblocks(n, 256), 256, 0, s>>>(n, adtInt);
  double root_box[6];
  TG_HIP(hipMemcpyAsync(root_box, adtReal, sizeof root_box, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < 3; i++) {  // root bounds padded by 1 % (ADT.C:66-86)
    ext[2 * i] = root_box[i];
    ext[2 * i + 1] = root_box[i + 3];
    double d = 0.01 * (ext[2 * i + 1] - ext[2 * i]);
    ext[2 * i] -= d;
    ext[2 * i + 1] += d;
  }
  return depth;
}

}  // namespace tg
