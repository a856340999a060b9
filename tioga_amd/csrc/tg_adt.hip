// tg_adt.hip -- the ADT of ADT.C:34-135 / buildADTrecursion.c:24-116 built
// on the device, node for node identical to the reference's.
//
// The reference builds the tree recursively: a node is the element at the
// median position (nav+1)/2-1 of its segment after median.F90's partial
// quicksort along dimension level%6 reorders the segment, its box bounds the
// segment, and the left / right subtrees are the segment's two halves.  The
// quicksort's exact swap order decides every later level, so it is kept.
//
// The split sizes depend on the segment length alone (nleft-1 and nav-nleft
// with nleft = (nav+1)/2), so the host lists every segment of every level
// ahead of time, for any number of trees built together.  Per level:
//   * long segments (more than BIGMIN elements) run the quicksort's
//     partition rounds in list form across many workgroups: flags X >= pivot
//     and X <= pivot, per-chunk counts, a scan, the ascending list A and
//     descending list B, the K pairs with A[m] < B[m] swapped in parallel
//     (K found by bisection, since A[m] - B[m] increases with m), and one
//     thread applying the loop control (hoare_stop, median_control in
//     tg_common.h); once the active ranges are short, one workgroup per
//     segment finishes the rounds from the saved loop state;
//   * medium segments get one workgroup each, which runs the same rounds;
//   * segments of at most WAVEMAX elements get one wave each;
//   * segments of at most SUB elements are finished, whole subtree, by one
//     wave each in LDS after the last level (k_adt_sub).
// The node box is the reference's sequential TIOGA_MIN / TIOGA_MAX fold,
// whose result is the value at the last position of the minimal (maximal)
// class under < (so -0.0 and +0.0 tie and the later one wins); the parallel
// reductions pick that same element.  Node ids are the reference's preorder
// counter: the left child of node me is me+1, the right child me+nleft.
#include <algorithm>
#include <stdexcept>
#include <vector>

#include "tg_engine.h"

namespace tg {

namespace {

constexpr int BT = 512;        // threads per workgroup for long segments
constexpr int PER = 8;         // consecutive positions per thread per chunk
constexpr int CH = BT * PER;   // positions per chunk of a multi-workgroup round
constexpr int WAVEMAX = 2048;  // segments at most this long (and above SUB): one wave
constexpr int BIGMIN = 1 << 14;  // segments longer than this: many workgroups
constexpr int FINISH = 2048;  // active range a single workgroup finishes a long segment's rounds from

// base: offset of the segment's tree in the concatenated element arrays;
// start: the segment's first position within its tree
struct Seg {
  int base, start, nav, me, parent, side, level, tree;
};
static_assert(sizeof(Seg) == 2 * sizeof(int4), "segment descriptors are stored in two int4 slots");

struct BigState {
  double pv;
  int lo, hi, st, nA, nB, K, rounds, pad;
};
static_assert(sizeof(BigState) <= 2 * sizeof(double4), "loop states are stored in two double4 slots");

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

__device__ void wg_median_rounds(int *IX, double *X, int n, int *A, int *B, WgState &sh, int *lds, int *err,
                                 int round0);

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
  wg_median_rounds(IX, X, n, A, B, sh, lds, err, 0);
}

// the partition rounds of wg_median from the state in sh (lo, hi; st == 0),
// round0 rounds having been made already, to the end and the fix-up
__device__ void wg_median_rounds(int *IX, double *X, int n, int *A, int *B, WgState &sh, int *lds, int *err,
                                 int round0)
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
  // every round narrows [lo, hi], so n + 2 rounds bound the loop; the cap
  // only guards against a broken invariant (reported, never expected)
  for (int round = round0;; round++) {
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

// ---------------------------------------------------------------------------
// node box: the reference's sequential fold, exactly

// canonical class key: -0.0 and +0.0 compare equal under <, so they share a key
__device__ __forceinline__ unsigned long long ckey(double v) { return dkey(v + 0.0); }

struct KP {
  unsigned long long k;
  int p;
};
// min class first, then the later position; max class first, then later
__device__ __forceinline__ KP kp_min(KP a, KP b) { return (a.k < b.k || (a.k == b.k && a.p > b.p)) ? a : b; }
__device__ __forceinline__ KP kp_max(KP a, KP b) { return (a.k > b.k || (a.k == b.k && a.p > b.p)) ? a : b; }

// box of the segment's elements ix[0..nav), by one workgroup; positions are
// segment positions, the result is written to r[6] by thread 0
__device__ void wg_box(const int *ix, int nav, const double *eb, double *r)
{
  __shared__ unsigned long long rk[6][BT / 64];
  __shared__ int rp[6][BT / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  KP m[6];
  for (int j = 0; j < 3; j++) {
    m[j] = KP{~0ull, -1};
    m[j + 3] = KP{0ull, -1};
  }
  for (int i = t; i < nav; i += BT) {
    const double *b = eb + 6 * (size_t)ix[i];
    for (int j = 0; j < 3; j++) {
      m[j] = kp_min(m[j], KP{ckey(b[j]), i});
      m[j + 3] = kp_max(m[j + 3], KP{ckey(b[j + 3]), i});
    }
  }
  for (int j = 0; j < 6; j++) {
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      KP o{__shfl_xor(m[j].k, q), __shfl_xor(m[j].p, q)};
      m[j] = j < 3 ? kp_min(m[j], o) : kp_max(m[j], o);
    }
    if (lane == 0) {
      rk[j][w] = m[j].k;
      rp[j][w] = m[j].p;
    }
  }
  __syncthreads();
  if (t < 6) {
    KP a{rk[t][0], rp[t][0]};
    for (int q = 1; q < BT / 64; q++) a = t < 3 ? kp_min(a, KP{rk[t][q], rp[t][q]}) : kp_max(a, KP{rk[t][q], rp[t][q]});
    r[t] = eb[6 * (size_t)ix[a.p] + t];
  }
  __syncthreads();
}

__device__ __forceinline__ void write_node(const Seg &s, int el, const double *r, int *adtInt, double *adtReal)
{
  const size_t me = (size_t)s.base + s.me;
  adtInt[4 * me] = el;
  adtInt[4 * me + 1] = adtInt[4 * me + 2] = adtInt[4 * me + 3] = -1;
  for (int j = 0; j < 6; j++) adtReal[6 * me + j] = r[j];
  if (s.side > 0) adtInt[4 * ((size_t)s.base + s.parent) + s.side] = el;
}

// Medium segments: one workgroup each.
__global__ void __launch_bounds__(BT) k_adt_level(const Seg *__restrict__ segs, int *ea, double *X, int *A, int *B,
                                                  const double *__restrict__ ebox, int *__restrict__ adtInt,
                                                  double *__restrict__ adtReal, int *err)
{
  __shared__ int lds[2 * (BT / 64)];
  __shared__ WgState sh;
  __shared__ double r[6];
  const int t = threadIdx.x;
  const Seg s = segs[blockIdx.x];
  const size_t g = (size_t)s.base + s.start;
  const double *eb = ebox + 6 * (size_t)s.base;
  const int dim = s.level % 6;
  int *ix = ea + g;
  double *x = X + g;
  for (int i = t; i < s.nav; i += BT) x[i] = eb[6 * (size_t)ix[i] + dim];
  __threadfence_block();
  __syncthreads();
  wg_median(ix, x, s.nav, A + g, B + g, sh, lds, err);
  wg_box(ix, s.nav, eb, r);
  if (t == 0) write_node(s, ix[(s.nav + 1) / 2 - 1], r, adtInt, adtReal);
}

// ---------------------------------------------------------------------------
// Wave-level forms: one wave per segment, no workgroup barriers.  Lanes of a
// wave exchange data through memory (LDS or global), so every hand-over is
// fenced at workgroup scope around a wave barrier.

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// median_select on X[0..n) / IX[0..n) by one wave; A and B are n-int lists.
// The rounds of wg_median with ballots in place of workgroup scans: lane l of
// a 64-position chunk holds the chunk's l-th position from each end, so the
// ballot prefix keeps A ascending and B descending.  Every lane computes the
// loop control from the same values.
__device__ void wave_median(int *IX, double *X, int n, int *A, int *B, int *err)
{
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1;
  auto swap = [&](int a, int b) {  // 1-based positions
    double tx = X[a - 1];
    X[a - 1] = X[b - 1];
    X[b - 1] = tx;
    int ti = IX[a - 1];
    IX[a - 1] = IX[b - 1];
    IX[b - 1] = ti;
  };
  if (n < 3) {
    if (lane == 0 && n == 2 && X[1] < X[0]) swap(1, 2);
    wave_sync();
    return;
  }
  int lo = 1, hi = n, st = 0;
  // every round narrows [lo, hi], so n + 2 rounds bound the loop; the cap
  // only guards against a broken invariant (reported, never expected)
  for (int round = 0;; round++) {
    if (round > n + 2) {
      if (lane == 0) atomicAdd(err, 1);
      break;
    }
    const double pv = median_pivot(X - 1, lo, hi);
    const int len = hi - lo + 1;
    int nA = 0, nB = 0;
    for (int base = 0; base < len; base += 64) {
      const int off = base + lane;
      const bool fa = off < len && X[lo + off - 1] >= pv;
      const bool fb = off < len && X[hi - off - 1] <= pv;
      const unsigned long long ma = __ballot(fa), mb = __ballot(fb);
      if (fa) A[nA + __popcll(ma & lt)] = lo + off;
      if (fb) B[nB + __popcll(mb & lt)] = hi - off;
      nA += __popcll(ma);
      nB += __popcll(mb);
    }
    wave_sync();
    // A[m] - B[m] increases with m: K is the number of leading pairs with
    // A[m] < B[m]
    const int mab = nA < nB ? nA : nB;
    int K = 0;
    for (int base = 0; base < mab; base += 64) {
      const int m = base + lane;
      const int c = __popcll(__ballot(m < mab && A[m] < B[m]));
      K += c;
      if (c < 64) break;
    }
    for (int m = lane; m < K; m += 64) swap(A[m], B[m]);
    wave_sync();
    int i, j;
    hoare_stop(K, nA, nB, K < nA ? A[K] : 0, K > 0 ? A[K - 1] : 0, K < nB ? B[K] : 0, K > 0 ? B[K - 1] : 0, A[0], B[0],
               i, j);
    st = median_control(n, i, j, lo, hi);
    wave_sync();  // the lists are rewritten by the next round
    if (st != 0) break;
  }
  if (lane == 0 && st == 2) {
    const int half = n / 2, halfp1 = half + 1;
    if (n == 2 * half) {
      if (X[halfp1 - 1] < X[half - 1]) swap(half, halfp1);
    } else if (X[lo - 1] > X[hi - 1]) {
      swap(lo, hi);
    }
  }
  wave_sync();
}

// node box of the segment's elements ix[0..nav) by one wave (the fold rule
// of wg_box); every lane returns the six bounds
__device__ void wave_box(const int *ix, int nav, const double *eb, double *r)
{
  const int lane = threadIdx.x & 63;
  KP m[6];
  for (int j = 0; j < 3; j++) {
    m[j] = KP{~0ull, -1};
    m[j + 3] = KP{0ull, -1};
  }
  for (int i = lane; i < nav; i += 64) {
    const double *b = eb + 6 * (size_t)ix[i];
    for (int j = 0; j < 3; j++) {
      m[j] = kp_min(m[j], KP{ckey(b[j]), i});
      m[j + 3] = kp_max(m[j + 3], KP{ckey(b[j + 3]), i});
    }
  }
  for (int j = 0; j < 6; j++) {
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      KP o{__shfl_xor(m[j].k, q), __shfl_xor(m[j].p, q)};
      m[j] = j < 3 ? kp_min(m[j], o) : kp_max(m[j], o);
    }
    r[j] = eb[6 * (size_t)ix[m[j].p] + j];
  }
}

constexpr int WAVES = 1;  // waves per workgroup of the wave-level kernels (more workgroups fit a CU's LDS)

// Segments of SUB < nav <= WAVEMAX elements, one level: a wave each, in
// global memory.
__global__ void __launch_bounds__(64 * WAVES) k_adt_wave(int n, const Seg *__restrict__ segs, int *ea, double *X,
                                                         int *A, int *B, const double *__restrict__ ebox,
                                                         int *__restrict__ adtInt, double *__restrict__ adtReal,
                                                         int *err)
{
  const int k = blockIdx.x * WAVES + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (k >= n) return;  // whole waves leave; no workgroup barrier follows
  const Seg s = segs[k];
  const size_t g = (size_t)s.base + s.start;
  const double *eb = ebox + 6 * (size_t)s.base;
  const int dim = s.level % 6;
  int *ix = ea + g;
  double *x = X + g;
  for (int i = lane; i < s.nav; i += 64) x[i] = eb[6 * (size_t)ix[i] + dim];
  wave_sync();
  wave_median(ix, x, s.nav, A + g, B + g, err);
  double r[6];
  wave_box(ix, s.nav, eb, r);
  if (lane == 0) write_node(s, ix[(s.nav + 1) / 2 - 1], r, adtInt, adtReal);
}

// Whole subtrees of segments of at most SUB elements, one wave each, with the
// segment's element ids and keys in LDS (buildADTrecursion.c:24-116).  The
// wave takes the subtree's segments longer than LANE_MAX in preorder with
// wave_median; their children of at most LANE_MAX elements (8 or more each,
// so at most SUB / 8 = 64 of them) are then finished by one lane each with
// the sequential quicksort.  Segments inside a subtree are packed into 64
// bits relative to its root (start, nav, node, parent + 1: 10 bits each;
// side: 2; level below the root: 5).
constexpr int SUB = 512;
constexpr int LANE_MAX = 16;
constexpr int WSTK = 16;  // wave stack: pending right halves of segments > LANE_MAX, at most log2(SUB) deep
constexpr int LSTK = 8;   // lane stack: at most log2(LANE_MAX) + 1 entries
static_assert(SUB < 1024 && SUB / 8 <= 64, "subtree fields are 10 bits; at most one lane segment per lane");
static_assert(SUB >= 64 * LSTK, "lane stacks live in the A and B lists");

struct LSeg {
  int start, nav, me, parent, side, dl;
};
__device__ __forceinline__ unsigned long long lpack(const LSeg &q)
{
  return (unsigned long long)q.start | ((unsigned long long)q.nav << 10) | ((unsigned long long)q.me << 20) |
         ((unsigned long long)(q.parent + 1) << 30) | ((unsigned long long)q.side << 40) |
         ((unsigned long long)q.dl << 42);
}
__device__ __forceinline__ LSeg lunpack(unsigned long long v)
{
  return LSeg{(int)(v & 1023), (int)((v >> 10) & 1023), (int)((v >> 20) & 1023), (int)((v >> 30) & 1023) - 1,
              (int)((v >> 40) & 3), (int)((v >> 42) & 31)};
}
// the segment q of the subtree under root as a tree segment
__device__ __forceinline__ Seg lglobal(const Seg &root, const LSeg &q)
{
  Seg s = root;
  s.start = root.start + q.start;
  s.nav = q.nav;
  s.me = root.me + q.me;
  if (q.parent >= 0) {
    s.parent = root.me + q.parent;
    s.side = q.side;
  }
  s.level = root.level + q.dl;
  return s;
}

__global__ void __launch_bounds__(64 * WAVES) k_adt_sub(int n, const Seg *__restrict__ segs,
                                                        const int *__restrict__ ea, const double *__restrict__ ebox,
                                                        int *__restrict__ adtInt, double *__restrict__ adtReal,
                                                        int *err)
{
  __shared__ double sX[WAVES][SUB];
  __shared__ int sIX[WAVES][SUB];
  __shared__ unsigned long long sAB[WAVES][SUB];  // A and B lists (SUB ints each); afterwards the lane stacks
  __shared__ unsigned long long sW[WAVES][WSTK], sL[WAVES][64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k = blockIdx.x * WAVES + w;
  if (k >= n) return;  // whole waves leave; no workgroup barrier follows
  const Seg root = segs[k];
  const double *eb = ebox + 6 * (size_t)root.base;
  double *X = sX[w];
  int *IX = sIX[w], *A = (int *)sAB[w], *B = A + SUB;
  const int *src = ea + (size_t)root.base + root.start;
  for (int i = lane; i < root.nav; i += 64) IX[i] = src[i];
  int nw = 0, nl = 0;  // wave stack depth and lane list length (uniform)
  const LSeg r0{0, root.nav, 0, -1, 0, 0};
  if (root.nav > LANE_MAX) {
    if (lane == 0) sW[w][0] = lpack(r0);
    nw = 1;
  } else {
    if (lane == 0) sL[w][0] = lpack(r0);
    nl = 1;
  }
  wave_sync();
  while (nw > 0) {
    const LSeg q = lunpack(sW[w][--nw]);
    const Seg s = lglobal(root, q);
    int *ix = IX + q.start;
    double *x = X + q.start;
    const int dim = s.level % 6;
    for (int i = lane; i < q.nav; i += 64) x[i] = eb[6 * (size_t)ix[i] + dim];
    wave_sync();
    wave_median(ix, x, q.nav, A, B, err);
    double r[6];
    wave_box(ix, q.nav, eb, r);
    const int nleft = (q.nav + 1) / 2;
    if (lane == 0) write_node(s, ix[nleft - 1], r, adtInt, adtReal);
    // children: nleft - 1 and nav - nleft elements, both at least 8
    const LSeg cl{q.start, nleft - 1, q.me + 1, q.me, 1, q.dl + 1};
    const LSeg cr{q.start + nleft, q.nav - nleft, q.me + nleft, q.me, 2, q.dl + 1};
    const int dw = (cl.nav > LANE_MAX) + (cr.nav > LANE_MAX);
    if (nw + dw > WSTK || nl + 2 - dw > 64) {  // bounded by the sizes; reported, never expected
      if (lane == 0) atomicAdd(err, 1);
      return;
    }
    if (cr.nav > LANE_MAX) {
      if (lane == 0) sW[w][nw] = lpack(cr);
      nw++;
    } else {
      if (lane == 0) sL[w][nl] = lpack(cr);
      nl++;
    }
    if (cl.nav > LANE_MAX) {
      if (lane == 0) sW[w][nw] = lpack(cl);
      nw++;
    } else {
      if (lane == 0) sL[w][nl] = lpack(cl);
      nl++;
    }
    wave_sync();
  }
  // lane phase: lane l finishes list entry l with the sequential quicksort
  if (lane < nl) {
    unsigned long long *stk = sAB[w] + lane * LSTK;
    int sp = 0;
    stk[sp++] = sL[w][lane];
    while (sp > 0) {
      const LSeg q = lunpack(stk[--sp]);
      const Seg s = lglobal(root, q);
      int *ix = IX + q.start;
      if (q.nav == 1) {
        double r[6];
        for (int j = 0; j < 6; j++) r[j] = eb[6 * (size_t)ix[0] + j];
        write_node(s, ix[0], r, adtInt, adtReal);
        continue;
      }
      double *x = X + q.start;
      const int dim = s.level % 6;
      for (int i = 0; i < q.nav; i++) x[i] = eb[6 * (size_t)ix[i] + dim];
      median_select(ix, x, q.nav);
      const int nleft = (q.nav + 1) / 2;
      double r[6] = {BIGVALUE, BIGVALUE, BIGVALUE, -BIGVALUE, -BIGVALUE, -BIGVALUE};
      for (int i = 0; i < q.nav; i++) {
        const double *b = eb + 6 * (size_t)ix[i];
        for (int j = 0; j < 3; j++) {
          r[j] = dmin(r[j], b[j]);
          r[j + 3] = dmax(r[j + 3], b[j + 3]);
        }
      }
      write_node(s, ix[nleft - 1], r, adtInt, adtReal);
      if (sp + 2 > LSTK) {  // bounded by the sizes; reported, never expected
        atomicAdd(err, 1);
        break;
      }
      if (q.nav - nleft > 0) stk[sp++] = lpack(LSeg{q.start + nleft, q.nav - nleft, q.me + nleft, q.me, 2, q.dl + 1});
      if (nleft > 1) stk[sp++] = lpack(LSeg{q.start, nleft - 1, q.me + 1, q.me, 1, q.dl + 1});
    }
  }
}

// ---------------------------------------------------------------------------
// long segments: partition rounds across workgroups.  A chunk table lists
// (segment, chunk) pairs; positions are 1-based within the segment.

__global__ void __launch_bounds__(BT) k_big_load(const int2 *__restrict__ tab, const Seg *__restrict__ segs,
                                                 const int *__restrict__ ea, double *__restrict__ X,
                                                 const double *__restrict__ ebox)
{
  const int2 tc = tab[blockIdx.x];
  const Seg s = segs[tc.x];
  const size_t g = (size_t)s.base + s.start;
  const double *eb = ebox + 6 * (size_t)s.base;
  const int dim = s.level % 6;
  const int o0 = tc.y * CH, o1 = min(o0 + CH, s.nav);
  for (int o = o0 + threadIdx.x; o < o1; o += BT) X[g + o] = eb[6 * (size_t)ea[g + o] + dim];
}

__global__ void k_big_init(int nbig, const Seg *__restrict__ segs, const double *__restrict__ X,
                           BigState *__restrict__ st)
{
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbig) return;
  const Seg s = segs[b];
  BigState S{};
  S.lo = 1;
  S.hi = s.nav;
  S.st = 0;
  S.rounds = 0;
  S.pv = median_pivot(X + (size_t)s.base + s.start - 1, 1, s.nav);
  st[b] = S;
}

// R1: per chunk, the number of A and B flags
__global__ void __launch_bounds__(BT) k_big_count(const int2 *__restrict__ tab, const Seg *__restrict__ segs,
                                                  const BigState *__restrict__ st, const int *__restrict__ choff,
                                                  const double *__restrict__ X, int *__restrict__ cntA,
                                                  int *__restrict__ cntB)
{
  __shared__ int lds[2 * (BT / 64)];
  const int2 tc = tab[blockIdx.x];
  const BigState S = st[tc.x];
  const int k = choff[tc.x] + tc.y;
  if (S.st != 0) {
    if (threadIdx.x == 0) cntA[k] = cntB[k] = 0;
    return;
  }
  const Seg s = segs[tc.x];
  const double *X1 = X + (size_t)s.base + s.start - 1;
  const int len = S.hi - S.lo + 1;
  const int o0 = tc.y * CH, o1 = min(o0 + CH, len);
  int ca = 0, cb = 0;
  for (int o = o0 + threadIdx.x; o < o1; o += BT) {
    ca += X1[S.lo + o] >= S.pv;
    cb += X1[S.hi - o] <= S.pv;
  }
  int ea, eb, ta, tb;
  wg_scan2(ca, cb, ea, eb, ta, tb, lds);
  if (threadIdx.x == 0) {
    cntA[k] = ta;
    cntB[k] = tb;
  }
}

// R2: exclusive scan of the chunk counts of each segment (one workgroup each)
__global__ void __launch_bounds__(BT) k_big_scan(const int *__restrict__ choff, const int *__restrict__ nch,
                                                 BigState *__restrict__ st, int *cntA, int *cntB)
{
  __shared__ int lds[2 * (BT / 64)];
  const int b = blockIdx.x;
  if (st[b].st != 0) return;
  int *a = cntA + choff[b], *c = cntB + choff[b];
  const int n = nch[b];
  int runA = 0, runB = 0;
  for (int base = 0; base < n; base += BT) {
    const int i = base + threadIdx.x;
    const int va = i < n ? a[i] : 0, vb = i < n ? c[i] : 0;
    int ea, eb, ta, tb;
    wg_scan2(va, vb, ea, eb, ta, tb, lds);
    if (i < n) {
      a[i] = runA + ea;
      c[i] = runB + eb;
    }
    runA += ta;
    runB += tb;
  }
  if (threadIdx.x == 0) {
    st[b].nA = runA;
    st[b].nB = runB;
  }
}

// R3: the lists A (ascending) and B (descending), chunk by chunk
__global__ void __launch_bounds__(BT) k_big_fill(const int2 *__restrict__ tab, const Seg *__restrict__ segs,
                                                 const BigState *__restrict__ st, const int *__restrict__ choff,
                                                 const double *__restrict__ X, const int *__restrict__ cntA,
                                                 const int *__restrict__ cntB, int *__restrict__ A, int *__restrict__ B)
{
  __shared__ int lds[2 * (BT / 64)];
  const int2 tc = tab[blockIdx.x];
  const BigState S = st[tc.x];
  if (S.st != 0) return;
  const int k = choff[tc.x] + tc.y;
  const Seg s = segs[tc.x];
  const size_t g = (size_t)s.base + s.start;
  const double *X1 = X + g - 1;
  const int len = S.hi - S.lo + 1;
  const int o0 = tc.y * CH;
  if (o0 >= len) return;
  // thread t owns PER consecutive offsets of the chunk
  int ca = 0, cb = 0;
  unsigned fa = 0, fb = 0;
#pragma unroll
  for (int e = 0; e < PER; e++) {
    const int o = o0 + threadIdx.x * PER + e;
    if (o < len) {
      if (X1[S.lo + o] >= S.pv) {
        fa |= 1u << e;
        ca++;
      }
      if (X1[S.hi - o] <= S.pv) {
        fb |= 1u << e;
        cb++;
      }
    }
  }
  int ra, rb, ta, tb;
  wg_scan2(ca, cb, ra, rb, ta, tb, lds);
  ra += cntA[k];
  rb += cntB[k];
  int *Ag = A + g, *Bg = B + g;
#pragma unroll
  for (int e = 0; e < PER; e++) {
    const int o = o0 + threadIdx.x * PER + e;
    if (fa & (1u << e)) Ag[ra++] = S.lo + o;
    if (fb & (1u << e)) Bg[rb++] = S.hi - o;
  }
}

// first m in [0, mab) with A[m] >= B[m] (A rises, B falls), else mab
__device__ __forceinline__ int big_K(const int *Ag, const int *Bg, int mab)
{
  int lo = 0, hi = mab;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (Ag[m] < Bg[m]) lo = m + 1;
    else hi = m;
  }
  return lo;
}

// R4: swap the K pairs (A[m], B[m]), m < K
__global__ void __launch_bounds__(BT) k_big_swap(const int2 *__restrict__ tab, const Seg *__restrict__ segs,
                                                 BigState *__restrict__ st, int *ea, double *X,
                                                 const int *__restrict__ A, const int *__restrict__ B)
{
  const int2 tc = tab[blockIdx.x];
  const BigState S = st[tc.x];
  if (S.st != 0) return;
  const Seg s = segs[tc.x];
  const size_t g = (size_t)s.base + s.start;
  const int *Ag = A + g, *Bg = B + g;
  const int K = big_K(Ag, Bg, min(S.nA, S.nB));
  if (tc.y == 0 && threadIdx.x == 0) st[tc.x].K = K;
  double *X1 = X + g - 1;
  int *I1 = ea + g - 1;
  const int m1 = min((tc.y + 1) * CH, K);
  for (int m = tc.y * CH + threadIdx.x; m < m1; m += BT) {
    const int a = Ag[m], b = Bg[m];
    const double tx = X1[a];
    X1[a] = X1[b];
    X1[b] = tx;
    const int ti = I1[a];
    I1[a] = I1[b];
    I1[b] = ti;
  }
}

// R5: loop control and the next pivot, one thread per segment
__global__ void k_big_ctrl(int nbig, const Seg *__restrict__ segs, BigState *__restrict__ st, int *ea, double *X,
                           const int *__restrict__ A, const int *__restrict__ B, int count, int *__restrict__ active,
                           int *__restrict__ err)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbig) return;
  BigState S = st[b];
  if (S.st != 0) return;
  const Seg s = segs[b];
  const size_t g = (size_t)s.base + s.start;
  const int *Ag = A + g, *Bg = B + g;
  double *X1 = X + g - 1;
  int *I1 = ea + g - 1;
  const int K = S.K, nA = S.nA, nB = S.nB;
  int i, j;
  hoare_stop(K, nA, nB, K < nA ? Ag[K] : 0, K > 0 ? Ag[K - 1] : 0, K < nB ? Bg[K] : 0, K > 0 ? Bg[K - 1] : 0, Ag[0],
             Bg[0], i, j);
  int lo = S.lo, hi = S.hi;
  S.st = median_control(s.nav, i, j, lo, hi);
  S.lo = lo;
  S.hi = hi;
  auto swap = [&](int a, int c) {
    const double tx = X1[a];
    X1[a] = X1[c];
    X1[c] = tx;
    const int ti = I1[a];
    I1[a] = I1[c];
    I1[c] = ti;
  };
  if (S.st == 2) {  // median.F90's final fix-up
    const int half = s.nav / 2, halfp1 = half + 1;
    if (s.nav == 2 * half) {
      if (X1[halfp1] < X1[half]) swap(half, halfp1);
    } else if (X1[lo] > X1[hi]) {
      swap(lo, hi);
    }
  } else if (S.st == 0) {
    if (++S.rounds > s.nav + 2) {  // every round narrows [lo, hi]; never expected
      atomicAdd(err, 1);
      S.st = 3;
    } else {
      S.pv = median_pivot(X1, lo, hi);
      if (count) atomicAdd(active, 1);
    }
  }
  st[b] = S;
}

// the rounds a long segment still needs after the multi-workgroup ones, by
// one workgroup from the saved loop state (the active range is short by then)
__global__ void __launch_bounds__(BT) k_big_finish(const Seg *__restrict__ segs, const BigState *__restrict__ st,
                                                   int *ea, double *X, int *A, int *B, int *err)
{
  __shared__ int lds[2 * (BT / 64)];
  __shared__ WgState sh;
  const BigState S = st[blockIdx.x];
  if (S.st != 0) return;  // finished (uniform over the workgroup)
  const Seg s = segs[blockIdx.x];
  const size_t g = (size_t)s.base + s.start;
  if (threadIdx.x == 0) {
    sh.lo = S.lo;
    sh.hi = S.hi;
    sh.st = 0;
  }
  __syncthreads();
  wg_median_rounds(ea + g, X + g, s.nav, A + g, B + g, sh, lds, err, S.rounds);
}

// node boxes of long segments: the minimal / maximal class per bound, then
// the last position in that class
__global__ void k_big_box_init(int nbig, unsigned long long *__restrict__ bk, int *__restrict__ bp)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * nbig) return;
  bk[i] = (i % 6) < 3 ? ~0ull : 0ull;
  bp[i] = -1;
}

__global__ void __launch_bounds__(BT) k_big_box_key(const int2 *__restrict__ tab, const Seg *__restrict__ segs,
                                                    const int *__restrict__ ea, const double *__restrict__ ebox,
                                                    unsigned long long *__restrict__ bk)
{
  const int2 tc = tab[blockIdx.x];
  const Seg s = segs[tc.x];
  const size_t g = (size_t)s.base + s.start;
  const double *eb = ebox + 6 * (size_t)s.base;
  const int o0 = tc.y * CH, o1 = min(o0 + CH, s.nav);
  unsigned long long m[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
  for (int o = o0 + threadIdx.x; o < o1; o += BT) {
    const double *b = eb + 6 * (size_t)ea[g + o];
    for (int j = 0; j < 3; j++) {
      m[j] = min(m[j], ckey(b[j]));
      m[j + 3] = max(m[j + 3], ckey(b[j + 3]));
    }
  }
  for (int j = 0; j < 6; j++) {
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      const unsigned long long o = __shfl_xor(m[j], q);
      m[j] = j < 3 ? min(m[j], o) : max(m[j], o);
    }
  }
  // the waves' results folded in LDS: one atomic per key and workgroup (the
  // keys of a long segment's first levels are shared by all its workgroups)
  __shared__ unsigned long long wk[BT / 64][6];
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 6; j++) wk[threadIdx.x >> 6][j] = m[j];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    unsigned long long v = wk[0][j];
    for (int q = 1; q < BT / 64; q++) v = j < 3 ? min(v, wk[q][j]) : max(v, wk[q][j]);
    if (j < 3) atomicMin(&bk[6 * tc.x + j], v);
    else atomicMax(&bk[6 * tc.x + j], v);
  }
}

__global__ void __launch_bounds__(BT) k_big_box_pos(const int2 *__restrict__ tab, const Seg *__restrict__ segs,
                                                    const int *__restrict__ ea, const double *__restrict__ ebox,
                                                    const unsigned long long *__restrict__ bk, int *__restrict__ bp)
{
  const int2 tc = tab[blockIdx.x];
  const Seg s = segs[tc.x];
  const size_t g = (size_t)s.base + s.start;
  const double *eb = ebox + 6 * (size_t)s.base;
  unsigned long long key[6];
  for (int j = 0; j < 6; j++) key[j] = bk[6 * tc.x + j];
  const int o0 = tc.y * CH, o1 = min(o0 + CH, s.nav);
  int p[6] = {-1, -1, -1, -1, -1, -1};
  for (int o = o0 + threadIdx.x; o < o1; o += BT) {
    const double *b = eb + 6 * (size_t)ea[g + o];
    for (int j = 0; j < 6; j++)
      if (ckey(b[j]) == key[j]) p[j] = o;
  }
  for (int j = 0; j < 6; j++) {
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) p[j] = max(p[j], __shfl_xor(p[j], q));
  }
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 6; j++)
      if (p[j] >= 0) atomicMax(&bp[6 * tc.x + j], p[j]);
}

__global__ void k_big_node(int nbig, const Seg *__restrict__ segs, const int *__restrict__ ea,
                           const double *__restrict__ ebox, const int *__restrict__ bp, int *__restrict__ adtInt,
                           double *__restrict__ adtReal)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbig) return;
  const Seg s = segs[b];
  const size_t g = (size_t)s.base + s.start;
  const double *eb = ebox + 6 * (size_t)s.base;
  double r[6];
  for (int j = 0; j < 6; j++) r[j] = eb[6 * (size_t)ea[g + bp[6 * b + j]] + j];
  write_node(s, ea[g + (s.nav + 1) / 2 - 1], r, adtInt, adtReal);
}

__global__ void k_iota_trees(int n, const int *__restrict__ base, int ntree, int *__restrict__ p)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = ntree - 1;  // last tree with base <= i
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (base[m] <= i) lo = m;
    else hi = m - 1;
  }
  p[i] = i - base[lo];
}

// slot 3 of row e: the node that holds element e (ADT.C:120-124)
__global__ void k_adt_slot3(int n, const int *__restrict__ base, int ntree, int *adtInt)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = ntree - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (base[m] <= i) lo = m;
    else hi = m - 1;
  }
  const int b = base[lo];
  adtInt[4 * ((size_t)b + adtInt[4 * (size_t)i]) + 3] = i - b;
}

// the root box of every tree of a concatenated batch (tree t's elements
// start at tb[t]; an empty tree's slot is left alone)
__global__ void k_adt_roots(int ntree, const int *__restrict__ tb, const double *__restrict__ adtReal,
                            double *__restrict__ roots)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntree || tb[t + 1] == tb[t]) return;
  for (int j = 0; j < 6; j++) roots[6 * t + j] = adtReal[6 * (size_t)tb[t] + j];
}

}  // namespace

void build_adt_batch(AdtTree *trees, int ntree, AdtScratch &w, hipStream_t s)
{
  std::vector<int> base(ntree + 1, 0);
  for (int t = 0; t < ntree; t++) {
    AdtTree &T = trees[t];
    T.depth = 0;
    for (int m = T.n; m > 0; m /= 2) T.depth++;
    if (T.n < 0) throw std::runtime_error("tioga_amd: negative ADT size");
    base[t + 1] = base[t] + T.n;
    if (base[t + 1] < base[t]) throw std::runtime_error("tioga_amd: ADT batch too large");
  }
  const int N = base[ntree];
  auto empty_ext = [](double ext[6]) {  // ADT.C:66-86 over no elements
    for (int i = 0; i < 3; i++) {
      double lo = BIGVALUE, hi = -BIGVALUE, d = 0.01 * (hi - lo);
      ext[2 * i] = lo - d;
      ext[2 * i + 1] = hi + d;
    }
  };
  if (N == 0) {
    for (int t = 0; t < ntree; t++) empty_ext(trees[t].ext);
    return;
  }
  // one tree builds in its own arrays, and so do trees whose arrays already
  // lie back to back; others are concatenated
  bool contiguous = true;
  for (int t = 1; t < ntree; t++)
    contiguous = contiguous && trees[t].ebox == trees[0].ebox + 6 * (size_t)base[t] &&
                 trees[t].adtInt == trees[0].adtInt + 4 * (size_t)base[t] &&
                 trees[t].adtReal == trees[0].adtReal + 6 * (size_t)base[t];
  const bool single = (ntree == 1) || contiguous;
  const double *ebox;
  int *adtInt;
  double *adtReal;
  if (single) {
    ebox = trees[0].ebox;
    adtInt = trees[0].adtInt;
    adtReal = trees[0].adtReal;
  } else {
    w.ebox.reserve(6 * (size_t)N);
    w.adtInt.reserve(4 * (size_t)N);
    w.adtReal.reserve(6 * (size_t)N);
    for (int t = 0; t < ntree; t++)
      if (trees[t].n)
        TG_HIP(hipMemcpyAsync(w.ebox.p + 6 * (size_t)base[t], trees[t].ebox, sizeof(double) * 6 * trees[t].n,
                              hipMemcpyDeviceToDevice, s));
    ebox = w.ebox.p;
    adtInt = w.adtInt.p;
    adtReal = w.adtReal.p;
  }
  // every segment of every level, from the sizes alone
  std::vector<std::vector<Seg>> lev;
  std::vector<Seg> small, cur, nxt;
  for (int t = 0; t < ntree; t++)
    if (trees[t].n > 0) cur.push_back(Seg{base[t], 0, trees[t].n, 0, 0, 0, 0, t});
  while (!cur.empty()) {
    std::vector<Seg> here;
    nxt.clear();
    for (const Seg &q : cur) {
      if (q.nav <= SUB) {
        small.push_back(q);
        continue;
      }
      here.push_back(q);
      const int nleft = (q.nav + 1) / 2;
      if (nleft > 1) nxt.push_back(Seg{q.base, q.start, nleft - 1, q.me + 1, q.me, 1, q.level + 1, q.tree});
      if (q.nav - nleft > 0) nxt.push_back(Seg{q.base, q.start + nleft, q.nav - nleft, q.me + nleft, q.me, 2, q.level + 1, q.tree});
    }
    if (!here.empty()) lev.push_back(here);
    cur.swap(nxt);
  }
  w.ea.reserve(N);
  w.X.reserve(N);
  w.A.reserve(N);
  w.B.reserve(N);
  w.err.reserve(2);
  w.tb.reserve(ntree + 1);
  TG_HIP(hipMemcpyAsync(w.tb.p, base.data(), sizeof(int) * (ntree + 1), hipMemcpyHostToDevice, s));
  TG_HIP(hipMemsetAsync(w.err.p, 0, 2 * sizeof(int), s));
  k_iota_trees<<<nblocks(N, 256), 256, 0, s>>>(N, w.tb.p, ntree, w.ea.p);
  // upload the level lists: per level the wave segments, the workgroup
  // segments, then the long ones
  std::vector<Seg> all;
  std::vector<size_t> wav0, nwav, med0, nmed, big0, nbigv;
  for (auto &L : lev) {
    wav0.push_back(all.size());
    for (const Seg &q : L)
      if (q.nav <= WAVEMAX) all.push_back(q);
    nwav.push_back(all.size() - wav0.back());
    med0.push_back(all.size());
    for (const Seg &q : L)
      if (q.nav > WAVEMAX && q.nav <= BIGMIN) all.push_back(q);
    nmed.push_back(all.size() - med0.back());
    big0.push_back(all.size());
    for (const Seg &q : L)
      if (q.nav > BIGMIN) all.push_back(q);
    nbigv.push_back(all.size() - big0.back());
  }
  const size_t nsmall0 = all.size();
  all.insert(all.end(), small.begin(), small.end());
  w.segs.reserve(2 * std::max<size_t>(all.size(), 1));
  TG_HIP(hipMemcpyAsync(w.segs.p, all.data(), sizeof(Seg) * all.size(), hipMemcpyHostToDevice, s));
  const Seg *segs = (const Seg *)w.segs.p;
  for (size_t l = 0; l < lev.size(); l++) {
    if (nwav[l])
      k_adt_wave<<<nblocks((long long)nwav[l], WAVES), 64 * WAVES, 0, s>>>(
          (int)nwav[l], segs + wav0[l], w.ea.p, w.X.p, w.A.p, w.B.p, ebox, adtInt, adtReal, w.err.p);
    if (nmed[l])
      k_adt_level<<<(unsigned)nmed[l], BT, 0, s>>>(segs + med0[l], w.ea.p, w.X.p, w.A.p, w.B.p, ebox, adtInt, adtReal,
                                                  w.err.p);
    const int nbig = (int)nbigv[l];
    if (!nbig) continue;
    const Seg *bs = segs + big0[l];
    // chunk table of the long segments
    std::vector<int2> tab;
    std::vector<int> choff(nbig), nch(nbig);
    for (int b = 0; b < nbig; b++) {
      const Seg &q = all[big0[l] + b];
      choff[b] = (int)tab.size();
      nch[b] = (q.nav + CH - 1) / CH;
      for (int c = 0; c < nch[b]; c++) tab.push_back(int2{b, c});
    }
    const int ntab = (int)tab.size();
    w.tab.reserve(ntab);
    w.choff.reserve(2 * nbig);
    w.cntA.reserve(ntab);
    w.cntB.reserve(ntab);
    w.state.reserve(2 * (size_t)nbig);
    w.bk.reserve(6 * (size_t)nbig);
    w.bp.reserve(6 * (size_t)nbig);
    TG_HIP(hipMemcpyAsync(w.tab.p, tab.data(), sizeof(int2) * ntab, hipMemcpyHostToDevice, s));
    TG_HIP(hipMemcpyAsync(w.choff.p, choff.data(), sizeof(int) * nbig, hipMemcpyHostToDevice, s));
    TG_HIP(hipMemcpyAsync(w.choff.p + nbig, nch.data(), sizeof(int) * nbig, hipMemcpyHostToDevice, s));
    BigState *st = (BigState *)w.state.p;
    k_big_load<<<ntab, BT, 0, s>>>(w.tab.p, bs, w.ea.p, w.X.p, ebox);
    k_big_init<<<nblocks(nbig, 64), 64, 0, s>>>(nbig, bs, w.X.p, st);
    // multi-workgroup rounds while the active ranges are long (each round
    // about halves them), then one workgroup per segment for the rest
    int maxnav = 0;
    for (int b = 0; b < nbig; b++) maxnav = std::max(maxnav, all[big0[l] + b].nav);
    int rounds = 2;
    while (rounds < 16 && (maxnav >> rounds) > FINISH) rounds++;
    for (int r = 0; r < rounds; r++) {
      k_big_count<<<ntab, BT, 0, s>>>(w.tab.p, bs, st, w.choff.p, w.X.p, w.cntA.p, w.cntB.p);
      k_big_scan<<<nbig, BT, 0, s>>>(w.choff.p, w.choff.p + nbig, st, w.cntA.p, w.cntB.p);
      k_big_fill<<<ntab, BT, 0, s>>>(w.tab.p, bs, st, w.choff.p, w.X.p, w.cntA.p, w.cntB.p, w.A.p, w.B.p);
      k_big_swap<<<ntab, BT, 0, s>>>(w.tab.p, bs, st, w.ea.p, w.X.p, w.A.p, w.B.p);
      k_big_ctrl<<<nblocks(nbig, 64), 64, 0, s>>>(nbig, bs, st, w.ea.p, w.X.p, w.A.p, w.B.p, 0, w.err.p + 1,
                                                  w.err.p);
    }
    k_big_finish<<<nbig, BT, 0, s>>>(bs, st, w.ea.p, w.X.p, w.A.p, w.B.p, w.err.p);
    k_big_box_init<<<nblocks(6 * nbig, 64), 64, 0, s>>>(nbig, w.bk.p, w.bp.p);
    k_big_box_key<<<ntab, BT, 0, s>>>(w.tab.p, bs, w.ea.p, ebox, w.bk.p);
    k_big_box_pos<<<ntab, BT, 0, s>>>(w.tab.p, bs, w.ea.p, ebox, w.bk.p, w.bp.p);
    k_big_node<<<nblocks(nbig, 64), 64, 0, s>>>(nbig, bs, w.ea.p, ebox, w.bp.p, adtInt, adtReal);
  }
  if (!small.empty())
    k_adt_sub<<<nblocks((long long)small.size(), WAVES), 64 * WAVES, 0, s>>>((int)small.size(), segs + nsmall0,
                                                                               w.ea.p, ebox, adtInt, adtReal, w.err.p);
  k_adt_slot3<<<nblocks(N, 256), 256, 0, s>>>(N, w.tb.p, ntree, adtInt);
  int err = 0;
  TG_HIP(hipMemcpyAsync(&err, w.err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if (!single)
    for (int t = 0; t < ntree; t++)
      if (trees[t].n) {
        TG_HIP(hipMemcpyAsync(trees[t].adtInt, adtInt + 4 * (size_t)base[t], sizeof(int) * 4 * trees[t].n,
                              hipMemcpyDeviceToDevice, s));
        TG_HIP(hipMemcpyAsync(trees[t].adtReal, adtReal + 6 * (size_t)base[t], sizeof(double) * 6 * trees[t].n,
                              hipMemcpyDeviceToDevice, s));
      }
  // every root's box (its first adtReal entry) gathered, one copy back
  std::vector<double> roots(6 * (size_t)ntree);
  w.roots.reserve(6 * (size_t)ntree);
  k_adt_roots<<<nblocks(ntree, 64), 64, 0, s>>>(ntree, w.tb.p, adtReal, w.roots.p);
  TG_HIP(hipMemcpyAsync(roots.data(), w.roots.p, sizeof(double) * 6 * ntree, hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  if (err) throw std::runtime_error("tioga_amd: ADT build: partition loop exceeded its bound");
  for (int t = 0; t < ntree; t++) {
    double *ext = trees[t].ext;
    if (trees[t].n == 0) {
      empty_ext(ext);
      continue;
    }
    for (int i = 0; i < 3; i++) {  // root bounds padded by 1 % (ADT.C:66-86)
      ext[2 * i] = roots[6 * t + i];
      ext[2 * i + 1] = roots[6 * t + i + 3];
      double d = 0.01 * (ext[2 * i + 1] - ext[2 * i]);
      ext[2 * i] -= d;
      ext[2 * i + 1] += d;
    }
  }
}

int build_adt_device(int n, const double *ebox, int *adtInt, double *adtReal, double ext[6], AdtScratch &w,
                     hipStream_t s)
{
  AdtTree t;
  t.n = n;
  t.ebox = ebox;
  t.adtInt = adtInt;
  t.adtReal = adtReal;
  build_adt_batch(&t, 1, w, s);
  for (int i = 0; i < 6; i++) ext[i] = t.ext[i];
  return t.depth;
}

}  // namespace tg
