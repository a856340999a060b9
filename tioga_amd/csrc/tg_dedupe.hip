// tg_dedupe.hip -- coincident query points on the device, with the exact
// result of uniquenodes_octree (utils.c:620-737).
//
// The reference splits the point set recursively into the 8 octants of its
// bounding box (mid-point split, half width + TOL) while a node holds more
// than 20 points and they do not all fall in one octant; each octant's points
// are listed in reverse arrival order.  At the leaves it compares all pairs
// in list order: a pair closer than TOL in L1 with the same mesh tag sends
// the larger index to the smaller one's representative and raises the
// smaller one's resolution.  Leaves touch disjoint points, so they can be
// processed in any order; only the leaf membership and the order inside each
// leaf matter, and both follow from the split rule alone.
//
// Here each level of the octree is one launch with a workgroup per node:
// bounding box, octant counts, then a stable reversed scatter of the node's
// points into its children.  Nodes of more than BIG points are split across
// many workgroups instead (chunked box, counts, per-octant scans, scatter).
// Nodes of at most SUBD points are set aside and finished, whole subtree, by
// one wave each with the points staged in LDS (k_oct_sub).
// Several point sets (the query sets of several blocks) can be deduplicated
// together, one root each.  Leaves are queued and finished one per thread.
//
// The bounding box enters only through mid = (max+min)/2 and
// dx = (max-min)/2 + TOL, where -0.0 and +0.0 give the same octants, so the
// box may be reduced in any order.
#include <stdexcept>
#include <vector>

#include "tg_engine.h"

namespace tg {

namespace {

constexpr int BT = 256;
constexpr int LEAF = 20;  // utils.c:633: split above 20 points
constexpr int BIG = 1 << 16;  // nodes above this: many workgroups
constexpr int CHB = 16 * BT;  // points per chunk of a big node
constexpr int SUBD = 1024;  // nodes of at most this many points: whole subtree by one wave

struct Node {
  int start, len;
};
static_assert(sizeof(Node) == sizeof(int2), "octree nodes are stored as int2");

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ __forceinline__ int octant(const double *p, const double mid[3], const double dx[3])
{
  int iv[3];
  for (int j = 0; j < 3; j++) iv[j] = (int)(floor((p[j] - mid[j]) / dx[j]) + 1);
  int o = 4 * iv[0] + 2 * iv[1] + iv[2];
  return o < 0 ? 0 : (o > 7 ? 7 : o);
}

__global__ void __launch_bounds__(BT) k_oct_level(int nnode, const Node *__restrict__ in, Node *__restrict__ out,
                                                  int *__restrict__ nout, Node *__restrict__ leaves,
                                                  int *__restrict__ nleaves, Node *__restrict__ subs,
                                                  int *__restrict__ nsubs, int *ea, int *tmp,
                                                  const double *__restrict__ x)
{
  __shared__ double red[6][BT / 64];
  __shared__ int cnt[8], wcnt[BT / 64][8];
  __shared__ double box[6];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = BT / 64;
  const Node nd = in[blockIdx.x];
  int *ix = ea + nd.start;
  // bounding box of the node's points (utils.c:635-645)
  double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
  for (int i = t; i < nd.len; i += BT) {
    const double *p = x + 3 * (size_t)ix[i];
    for (int j = 0; j < 3; j++) {
      lo[j] = dmin(lo[j], p[j]);
      hi[j] = dmax(hi[j], p[j]);
    }
  }
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      lo[j] = dmin(lo[j], __shfl_xor(lo[j], q));
      hi[j] = dmax(hi[j], __shfl_xor(hi[j], q));
    }
  if (lane == 0)
    for (int j = 0; j < 3; j++) {
      red[j][w] = lo[j];
      red[j + 3][w] = hi[j];
    }
  if (t < 8) cnt[t] = 0;
  __syncthreads();
  if (t < 6) {
    double r = red[t][0];
    for (int q = 1; q < nw; q++) r = t < 3 ? dmin(r, red[t][q]) : dmax(r, red[t][q]);
    box[t] = r;
  }
  __syncthreads();
  double mid[3], dx[3];
  for (int j = 0; j < 3; j++) {
    mid[j] = (box[j + 3] + box[j]) * 0.5;
    dx[j] = (box[j + 3] - box[j]) * 0.5 + TOL;
  }
  // octant counts
  int my[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = t; i < nd.len; i += BT) my[octant(x + 3 * (size_t)ix[i], mid, dx)]++;
  for (int o = 0; o < 8; o++)
    if (my[o]) atomicAdd(&cnt[o], my[o]);
  __syncthreads();
  bool one = false;
  for (int o = 0; o < 8; o++) one = one || (cnt[o] == nd.len);
  if (one) {  // every point in one octant: the node is compared as a whole
    if (t == 0) leaves[atomicAdd(nleaves, 1)] = nd;
    return;
  }
  int cft[9];
  cft[0] = 0;
  for (int o = 0; o < 8; o++) cft[o + 1] = cft[o] + cnt[o];
  // stable scatter, each octant in reverse arrival order (utils.c:666-679)
  int run[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // points of each octant in earlier chunks
  const unsigned long long lt = (1ull << lane) - 1;
  for (int base = 0; base < nd.len; base += BT) {
    const int i = base + t;
    int o = -1;
    if (i < nd.len) o = octant(x + 3 * (size_t)ix[i], mid, dx);
    int rank_in_wave = 0;
    for (int q = 0; q < 8; q++) {
      unsigned long long m = __ballot(o == q);
      if (lane == 0) wcnt[w][q] = __popcll(m);
      if (o == q) rank_in_wave = __popcll(m & lt);
    }
    __syncthreads();
    if (o >= 0) {
      int r = run[o] + rank_in_wave;
      for (int q = 0; q < w; q++) r += wcnt[q][o];
      tmp[nd.start + cft[o] + (cnt[o] - 1 - r)] = ix[i];
    }
    for (int q = 0; q < 8; q++)
      for (int v = 0; v < nw; v++) run[q] += wcnt[v][q];
    __syncthreads();
  }
  __threadfence_block();
  __syncthreads();
  for (int i = t; i < nd.len; i += BT) ix[i] = tmp[nd.start + i];
  if (t < 8 && cnt[t] > 0) {
    Node c{nd.start + cft[t], cnt[t]};
    if (c.len > SUBD) out[atomicAdd(nout, 1)] = c;
    else if (c.len > LEAF) subs[atomicAdd(nsubs, 1)] = c;
    else leaves[atomicAdd(nleaves, 1)] = c;
  }
}

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Whole subtrees of nodes of at most SUBD points, one wave each.  The points'
// coordinates (by local id) and the node order (local ids) live in LDS; a
// node's box, octant counts and stable reversed scatter are wave ballots, as
// in k_oct_level.  Non-leaf children are pushed largest first, so the
// smallest is taken next: a pending entry is at least as large as the child
// being split below it, and the stack stays below 7 log8(SUBD / LEAF) + 8
// entries.  The subtree's leaves of two or more points are listed in LDS and
// compared at the end, one lane per leaf, as k_oct_leaf compares them.
constexpr int DW = 2;     // waves per workgroup of k_oct_sub
constexpr int DSTK = 32;  // stack entries: start | len << 16
static_assert(SUBD <= 65535, "local ids are 16 bits");

__global__ void __launch_bounds__(64 * DW) k_oct_sub(int n, const Node *__restrict__ roots, const int *__restrict__ ea,
                                                     const double *__restrict__ x, const int *__restrict__ tag,
                                                     double *rtag, int *itag, int *__restrict__ err)
{
  __shared__ double sx[DW][3][SUBD];
  __shared__ int sg[DW][SUBD];
  __shared__ unsigned short sP[DW][SUBD], sT[DW][SUBD];
  __shared__ unsigned sS[DW][DSTK], sLeaf[DW][SUBD / 2];  // leaves: start | len << 16, two or more points
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k = blockIdx.x * DW + w;
  if (k >= n) return;  // whole waves leave; no workgroup barrier follows
  const Node root = roots[k];
  const unsigned long long lt = (1ull << lane) - 1;
  unsigned short *P = sP[w], *T = sT[w];
  for (int i = lane; i < root.len; i += 64) {
    const int g = ea[root.start + i];
    sg[w][i] = g;
    for (int j = 0; j < 3; j++) sx[w][j][i] = x[3 * (size_t)g + j];
    P[i] = (unsigned short)i;
  }
  int sp = 0, nleaf = 0;
  auto leaf = [&](int st, int len) {  // uniform; at most SUBD / 2 leaves of two or more points
    if (len < 2) return;
    if (nleaf >= SUBD / 2) {  // bounded by the sizes; reported, never expected
      if (lane == 0) atomicAdd(err, 1);
      return;
    }
    if (lane == 0) sLeaf[w][nleaf] = (unsigned)st | ((unsigned)len << 16);
    nleaf++;
  };
  if (lane == 0) sS[w][0] = (unsigned)root.len << 16;
  sp = 1;
  wave_sync();
  while (sp > 0) {
    const unsigned e = sS[w][--sp];
    const int cs = (int)(e & 0xffff), cl = (int)(e >> 16);
    // bounding box (utils.c:635-645)
    double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
    for (int i = lane; i < cl; i += 64) {
      const int l = P[cs + i];
      for (int j = 0; j < 3; j++) {
        lo[j] = dmin(lo[j], sx[w][j][l]);
        hi[j] = dmax(hi[j], sx[w][j][l]);
      }
    }
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int q = 32; q > 0; q >>= 1) {
        lo[j] = dmin(lo[j], __shfl_xor(lo[j], q));
        hi[j] = dmax(hi[j], __shfl_xor(hi[j], q));
      }
    double mid[3], dx[3];
    for (int j = 0; j < 3; j++) {
      mid[j] = (hi[j] + lo[j]) * 0.5;
      dx[j] = (hi[j] - lo[j]) * 0.5 + TOL;
    }
    auto oct_of = [&](int i) {
      if (i >= cl) return -1;
      const int l = P[cs + i];
      const double p[3] = {sx[w][0][l], sx[w][1][l], sx[w][2][l]};
      return octant(p, mid, dx);
    };
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int base = 0; base < cl; base += 64) {
      const int o = oct_of(base + lane);
      for (int q = 0; q < 8; q++) cnt[q] += __popcll(__ballot(o == q));
    }
    bool one = false;
    for (int q = 0; q < 8; q++) one = one || (cnt[q] == cl);
    if (one) {  // every point in one octant: the node is compared as a whole
      leaf(cs, cl);
      continue;
    }
    int cft[9];
    cft[0] = 0;
    for (int q = 0; q < 8; q++) cft[q + 1] = cft[q] + cnt[q];
    // stable scatter, each octant in reverse arrival order (utils.c:666-679)
    int run[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int base = 0; base < cl; base += 64) {
      const int i = base + lane;
      const int o = oct_of(i);
      int rank = 0;
      for (int q = 0; q < 8; q++) {
        const unsigned long long m = __ballot(o == q);
        if (o == q) rank = run[q] + __popcll(m & lt);
        run[q] += __popcll(m);
      }
      if (o >= 0) T[cs + cft[o] + (cnt[o] - 1 - rank)] = P[cs + i];
    }
    wave_sync();
    for (int i = lane; i < cl; i += 64) P[cs + i] = T[cs + i];
    // children: leaves out, the others pushed largest first
    int ord[8], nn = 0;
    for (int q = 0; q < 8; q++) {
      if (cnt[q] == 0) continue;
      if (cnt[q] <= LEAF) {
        leaf(cs + cft[q], cnt[q]);
        continue;
      }
      int m = nn++;
      while (m > 0 && cnt[ord[m - 1]] < cnt[q]) {  // descending size, stable
        ord[m] = ord[m - 1];
        m--;
      }
      ord[m] = q;
    }
    if (sp + nn > DSTK) {  // bounded by the sizes; reported, never expected
      if (lane == 0) atomicAdd(err, 1);
      return;
    }
    if (lane == 0)
      for (int m = 0; m < nn; m++) sS[w][sp + m] = (unsigned)(cs + cft[ord[m]]) | ((unsigned)cnt[ord[m]] << 16);
    sp += nn;
    wave_sync();
  }
  wave_sync();
  // all pairs of each leaf in list order (utils.c:691-716); leaves touch
  // disjoint points
  for (int k = lane; k < nleaf; k += 64) {
    const unsigned e = sLeaf[w][k];
    const int st = (int)(e & 0xffff), len = (int)(e >> 16);
    for (int a = 0; a < len; a++) {
      const int la = P[st + a], p1 = sg[w][la];
      const double a0 = sx[w][0][la], a1 = sx[w][1][la], a2 = sx[w][2][la];
      const int t1 = tag[p1];
      for (int b = a + 1; b < len; b++) {
        const int lb = P[st + b], p2 = sg[w][lb];
        if (fabs(a0 - sx[w][0][lb]) + fabs(a1 - sx[w][1][lb]) + fabs(a2 - sx[w][2][lb]) < TOL && t1 == tag[p2]) {
          if (p1 > p2) {
            rtag[p2] = dmax(rtag[p1], rtag[p2]);
            itag[p1] = itag[p2];
          } else {
            rtag[p1] = dmax(rtag[p1], rtag[p2]);
            itag[p2] = itag[p1];
          }
        }
      }
    }
  }
}

// all pairs of a leaf in list order (utils.c:691-716)
__global__ void k_oct_leaf(int n, const Node *__restrict__ leaves, const int *__restrict__ ea,
                           const double *__restrict__ x, const int *__restrict__ tag, double *rtag, int *itag)
{
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const Node nd = leaves[k];
  const int *ix = ea + nd.start;
  for (int i = 0; i < nd.len; i++) {
    const int p1 = ix[i];
    const double a0 = x[3 * (size_t)p1], a1 = x[3 * (size_t)p1 + 1], a2 = x[3 * (size_t)p1 + 2];
    const int t1 = tag[p1];
    for (int j = i + 1; j < nd.len; j++) {
      const int p2 = ix[j];
      if (fabs(a0 - x[3 * (size_t)p2]) + fabs(a1 - x[3 * (size_t)p2 + 1]) + fabs(a2 - x[3 * (size_t)p2 + 2]) < TOL &&
          t1 == tag[p2]) {
        if (p1 > p2) {
          rtag[p2] = dmax(rtag[p1], rtag[p2]);
          itag[p1] = itag[p2];
        } else {
          rtag[p1] = dmax(rtag[p1], rtag[p2]);
          itag[p2] = itag[p1];
        }
      }
    }
  }
}

__global__ void k_iota_sets(int n, const int *__restrict__ base, int nset, int *__restrict__ ea,
                            int *__restrict__ itag)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nset - 1;  // last set with base <= i
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (base[m] <= i) lo = m;
    else hi = m - 1;
  }
  ea[i] = i;
  itag[i] = i - base[lo];
}

// ---- nodes of more than BIG points: chunk table entries (node, chunk) ----

struct BigBox {
  double mid[3], dx[3];
  int cnt[8], cft[9], one, pad[2];
};

__global__ void k_ob_init(int nbig, unsigned long long *__restrict__ bk)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 6 * nbig) bk[i] = (i % 6) < 3 ? ~0ull : 0ull;
}

__global__ void __launch_bounds__(BT) k_ob_box(const int2 *__restrict__ tab, const Node *__restrict__ nodes,
                                               const int *__restrict__ ea, const double *__restrict__ x,
                                               unsigned long long *__restrict__ bk)
{
  const int2 tc = tab[blockIdx.x];
  const Node nd = nodes[tc.x];
  const int o0 = tc.y * CHB, o1 = min(o0 + CHB, nd.len);
  unsigned long long m[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
  for (int o = o0 + threadIdx.x; o < o1; o += BT) {
    const double *p = x + 3 * (size_t)ea[nd.start + o];
    for (int j = 0; j < 3; j++) {
      const unsigned long long k = dkey(p[j] + 0.0);
      m[j] = min(m[j], k);
      m[j + 3] = max(m[j + 3], k);
    }
  }
  for (int j = 0; j < 6; j++) {
#pragma unroll
    for (int q = 32; q > 0; q >>= 1) {
      const unsigned long long o = __shfl_xor(m[j], q);
      m[j] = j < 3 ? min(m[j], o) : max(m[j], o);
    }
  }
  if ((threadIdx.x & 63) == 0)
    for (int j = 0; j < 6; j++) {
      if (j < 3) atomicMin(&bk[6 * tc.x + j], m[j]);
      else atomicMax(&bk[6 * tc.x + j], m[j]);
    }
}

__global__ void k_ob_mid(int nbig, const unsigned long long *__restrict__ bk, BigBox *__restrict__ bb)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbig) return;
  for (int j = 0; j < 3; j++) {  // utils.c:646-649
    const double lo = dunkey(bk[6 * b + j]), hi = dunkey(bk[6 * b + j + 3]);
    bb[b].mid[j] = (hi + lo) * 0.5;
    bb[b].dx[j] = (hi - lo) * 0.5 + TOL;
  }
}

__global__ void __launch_bounds__(BT) k_ob_count(const int2 *__restrict__ tab, const Node *__restrict__ nodes,
                                                 const int *__restrict__ choff, const int *__restrict__ ea,
                                                 const double *__restrict__ x, const BigBox *__restrict__ bb,
                                                 int *__restrict__ cc)
{
  __shared__ int c[8];
  const int2 tc = tab[blockIdx.x];
  const Node nd = nodes[tc.x];
  if (threadIdx.x < 8) c[threadIdx.x] = 0;
  __syncthreads();
  const int o0 = tc.y * CHB, o1 = min(o0 + CHB, nd.len);
  int my[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int o = o0 + threadIdx.x; o < o1; o += BT) my[octant(x + 3 * (size_t)ea[nd.start + o], bb[tc.x].mid, bb[tc.x].dx)]++;
  for (int q = 0; q < 8; q++)
    if (my[q]) atomicAdd(&c[q], my[q]);
  __syncthreads();
  if (threadIdx.x < 8) cc[8 * (choff[tc.x] + tc.y) + threadIdx.x] = c[threadIdx.x];
}

// per big node: per-octant exclusive scan over its chunks, octant totals,
// and whether one octant holds every point
__global__ void k_ob_scan(int nbig, const Node *__restrict__ nodes, const int *__restrict__ choff,
                          const int *__restrict__ nch, int *__restrict__ cc, BigBox *__restrict__ bb)
{
  const int b = blockIdx.x;
  const int q = threadIdx.x;  // octant
  if (q < 8) {
    int run = 0;
    int *c = cc + 8 * (size_t)choff[b];
    for (int k = 0; k < nch[b]; k++) {
      const int v = c[8 * k + q];
      c[8 * k + q] = run;
      run += v;
    }
    bb[b].cnt[q] = run;
  }
  __syncthreads();
  if (q == 0) {
    int one = 0;
    bb[b].cft[0] = 0;
    for (int o = 0; o < 8; o++) {
      bb[b].cft[o + 1] = bb[b].cft[o] + bb[b].cnt[o];
      one |= (bb[b].cnt[o] == nodes[b].len);
    }
    bb[b].one = one;
  }
}

// stable scatter, each octant in reverse arrival order (utils.c:666-679)
__global__ void __launch_bounds__(BT) k_ob_scatter(const int2 *__restrict__ tab, const Node *__restrict__ nodes,
                                                   const int *__restrict__ choff, const int *__restrict__ ea,
                                                   int *__restrict__ tmp, const double *__restrict__ x,
                                                   const BigBox *__restrict__ bb, const int *__restrict__ cc)
{
  __shared__ int wcnt[BT / 64][8];
  const int2 tc = tab[blockIdx.x];
  const Node nd = nodes[tc.x];
  const BigBox &B = bb[tc.x];
  if (B.one) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, nw = BT / 64;
  int run[8];
  for (int q = 0; q < 8; q++) run[q] = cc[8 * (choff[tc.x] + tc.y) + q];
  const unsigned long long lt = (1ull << lane) - 1;
  const int o0 = tc.y * CHB, o1 = min(o0 + CHB, nd.len);
  for (int base = o0; base < o1; base += BT) {
    const int i = base + t;
    int o = -1;
    if (i < o1) o = octant(x + 3 * (size_t)ea[nd.start + i], B.mid, B.dx);
    int rank_in_wave = 0;
    for (int q = 0; q < 8; q++) {
      unsigned long long m = __ballot(o == q);
      if (lane == 0) wcnt[w][q] = __popcll(m);
      if (o == q) rank_in_wave = __popcll(m & lt);
    }
    __syncthreads();
    if (o >= 0) {
      int r = run[o] + rank_in_wave;
      for (int q = 0; q < w; q++) r += wcnt[q][o];
      tmp[nd.start + B.cft[o] + (B.cnt[o] - 1 - r)] = ea[nd.start + i];
    }
    for (int q = 0; q < 8; q++)
      for (int v = 0; v < nw; v++) run[q] += wcnt[v][q];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BT) k_ob_copy(const int2 *__restrict__ tab, const Node *__restrict__ nodes,
                                                int *__restrict__ ea, const int *__restrict__ tmp,
                                                const BigBox *__restrict__ bb)
{
  const int2 tc = tab[blockIdx.x];
  const Node nd = nodes[tc.x];
  if (bb[tc.x].one) return;
  const int o0 = tc.y * CHB, o1 = min(o0 + CHB, nd.len);
  for (int o = o0 + threadIdx.x; o < o1; o += BT) ea[nd.start + o] = tmp[nd.start + o];
}

__global__ void k_ob_emit(int nbig, const Node *__restrict__ nodes, const BigBox *__restrict__ bb,
                          Node *__restrict__ nout, int *__restrict__ cnt, Node *__restrict__ bout,
                          Node *__restrict__ leaves, Node *__restrict__ subs)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbig) return;
  const Node nd = nodes[b];
  const BigBox &B = bb[b];
  if (B.one) {  // every point in one octant: the node is compared as a whole
    leaves[atomicAdd(&cnt[1], 1)] = nd;
    return;
  }
  for (int q = 0; q < 8; q++) {
    if (B.cnt[q] == 0) continue;
    const Node c{nd.start + B.cft[q], B.cnt[q]};
    if (c.len > BIG) bout[atomicAdd(&cnt[2], 1)] = c;
    else if (c.len > SUBD) nout[atomicAdd(&cnt[0], 1)] = c;
    else if (c.len > LEAF) subs[atomicAdd(&cnt[3], 1)] = c;
    else leaves[atomicAdd(&cnt[1], 1)] = c;
  }
}

}  // namespace

void unique_nodes_batch(DedupeSet *sets, int nset, DedupeScratch &w, hipStream_t s)
{
  std::vector<int> base(nset + 1, 0);
  for (int k = 0; k < nset; k++) {
    if (sets[k].n < 0) throw std::runtime_error("tioga_amd: negative point count");
    base[k + 1] = base[k] + sets[k].n;
    if (base[k + 1] < base[k]) throw std::runtime_error("tioga_amd: dedupe batch too large");
  }
  const int n = base[nset];
  if (n == 0) return;
  // sets that already lie back to back (every array) are worked on in place
  bool contiguous = true;
  for (int k = 1; k < nset; k++)
    contiguous = contiguous && sets[k].x == sets[0].x + 3 * (size_t)base[k] && sets[k].tag == sets[0].tag + base[k] &&
                 sets[k].rtag == sets[0].rtag + base[k] && sets[k].itag == sets[0].itag + base[k];
  const bool single = (nset == 1) || contiguous;
  const double *x;
  const int *tag;
  double *rtag;
  int *itag;
  if (single) {
    x = sets[0].x;
    tag = sets[0].tag;
    rtag = sets[0].rtag;
    itag = sets[0].itag;
  } else {
    w.x.reserve(3 * (size_t)n);
    w.tag.reserve(n);
    w.rtag.reserve(n);
    w.itag.reserve(n);
    for (int k = 0; k < nset; k++)
      if (sets[k].n) {
        TG_HIP(hipMemcpyAsync(w.x.p + 3 * (size_t)base[k], sets[k].x, sizeof(double) * 3 * sets[k].n,
                              hipMemcpyDeviceToDevice, s));
        TG_HIP(hipMemcpyAsync(w.tag.p + base[k], sets[k].tag, sizeof(int) * sets[k].n, hipMemcpyDeviceToDevice, s));
        TG_HIP(hipMemcpyAsync(w.rtag.p + base[k], sets[k].rtag, sizeof(double) * sets[k].n,
                              hipMemcpyDeviceToDevice, s));
      }
    x = w.x.p;
    tag = w.tag.p;
    rtag = w.rtag.p;
    itag = w.itag.p;
  }
  w.ea.reserve(n);
  w.tmp.reserve(n);
  w.nodesA.reserve(n / LEAF + nset + 2);
  w.nodesB.reserve(n / LEAF + nset + 2);
  w.bigA.reserve(n / BIG + nset + 2);
  w.bigB.reserve(n / BIG + nset + 2);
  w.leaves.reserve(n + 1);
  w.subs.reserve(n / LEAF + nset + 2);
  w.counts.reserve(8);
  w.sb.reserve(nset + 1);
  TG_HIP(hipMemcpyAsync(w.sb.p, base.data(), sizeof(int) * (nset + 1), hipMemcpyHostToDevice, s));
  k_iota_sets<<<nblocks(n, 256), 256, 0, s>>>(n, w.sb.p, nset, w.ea.p, itag);
  // roots
  std::vector<Node> rn, rb, rl, rs;
  for (int k = 0; k < nset; k++) {
    const Node r{base[k], sets[k].n};
    if (r.len > BIG) rb.push_back(r);
    else if (r.len > SUBD) rn.push_back(r);
    else if (r.len > LEAF) rs.push_back(r);
    else if (r.len > 0) rl.push_back(r);
  }
  if (!rs.empty()) TG_HIP(hipMemcpyAsync(w.subs.p, rs.data(), sizeof(Node) * rs.size(), hipMemcpyHostToDevice, s));
  if (!rn.empty())
    TG_HIP(hipMemcpyAsync(w.nodesA.p, rn.data(), sizeof(Node) * rn.size(), hipMemcpyHostToDevice, s));
  if (!rb.empty()) TG_HIP(hipMemcpyAsync(w.bigA.p, rb.data(), sizeof(Node) * rb.size(), hipMemcpyHostToDevice, s));
  if (!rl.empty()) TG_HIP(hipMemcpyAsync(w.leaves.p, rl.data(), sizeof(Node) * rl.size(), hipMemcpyHostToDevice, s));
  int cnt[4] = {(int)rn.size(), (int)rl.size(), (int)rb.size(), (int)rs.size()};  // normal, leaves, big, subtrees
  Node *nin = (Node *)w.nodesA.p, *nout = (Node *)w.nodesB.p;
  Node *bin = (Node *)w.bigA.p, *bout = (Node *)w.bigB.p;
  while (cnt[0] + cnt[2] > 0) {
    const int nnorm = cnt[0], nbig = cnt[2];
    const int hc[4] = {0, cnt[1], 0, cnt[3]};
    TG_HIP(hipMemcpyAsync(w.counts.p, hc, sizeof hc, hipMemcpyHostToDevice, s));
    if (nnorm)
      k_oct_level<<<nnorm, BT, 0, s>>>(nnorm, nin, nout, w.counts.p, (Node *)w.leaves.p, w.counts.p + 1,
                                       (Node *)w.subs.p, w.counts.p + 3, w.ea.p, w.tmp.p, x);
    if (nbig) {
      std::vector<Node> bn(nbig);
      TG_HIP(hipMemcpyAsync(bn.data(), bin, sizeof(Node) * nbig, hipMemcpyDeviceToHost, s));
      TG_HIP(hipStreamSynchronize(s));
      std::vector<int2> tab;
      std::vector<int> choff(nbig), nch(nbig);
      for (int b = 0; b < nbig; b++) {
        choff[b] = (int)tab.size();
        nch[b] = (bn[b].len + CHB - 1) / CHB;
        for (int c = 0; c < nch[b]; c++) tab.push_back(int2{b, c});
      }
      const int ntab = (int)tab.size();
      w.tab.reserve(ntab);
      w.choff.reserve(2 * nbig);
      w.cc.reserve(8 * (size_t)ntab);
      w.bk.reserve(6 * (size_t)nbig);
      w.bb.reserve((sizeof(BigBox) * nbig + sizeof(double4) - 1) / sizeof(double4));
      TG_HIP(hipMemcpyAsync(w.tab.p, tab.data(), sizeof(int2) * ntab, hipMemcpyHostToDevice, s));
      TG_HIP(hipMemcpyAsync(w.choff.p, choff.data(), sizeof(int) * nbig, hipMemcpyHostToDevice, s));
      TG_HIP(hipMemcpyAsync(w.choff.p + nbig, nch.data(), sizeof(int) * nbig, hipMemcpyHostToDevice, s));
      BigBox *bb = (BigBox *)w.bb.p;
      k_ob_init<<<nblocks(6 * nbig, 64), 64, 0, s>>>(nbig, w.bk.p);
      k_ob_box<<<ntab, BT, 0, s>>>(w.tab.p, bin, w.ea.p, x, w.bk.p);
      k_ob_mid<<<nblocks(nbig, 64), 64, 0, s>>>(nbig, w.bk.p, bb);
      k_ob_count<<<ntab, BT, 0, s>>>(w.tab.p, bin, w.choff.p, w.ea.p, x, bb, w.cc.p);
      k_ob_scan<<<nbig, 64, 0, s>>>(nbig, bin, w.choff.p, w.choff.p + nbig, w.cc.p, bb);
      k_ob_scatter<<<ntab, BT, 0, s>>>(w.tab.p, bin, w.choff.p, w.ea.p, w.tmp.p, x, bb, w.cc.p);
      k_ob_copy<<<ntab, BT, 0, s>>>(w.tab.p, bin, w.ea.p, w.tmp.p, bb);
      k_ob_emit<<<nblocks(nbig, 64), 64, 0, s>>>(nbig, bin, bb, nout, w.counts.p, bout, (Node *)w.leaves.p,
                                                 (Node *)w.subs.p);
    }
    TG_HIP(hipMemcpyAsync(cnt, w.counts.p, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    std::swap(nin, nout);
    std::swap(bin, bout);
  }
  if (cnt[3] > 0) {  // subtrees of at most SUBD points, their leaves included
    const int zero = 0;
    TG_HIP(hipMemcpyAsync(w.counts.p + 4, &zero, sizeof(int), hipMemcpyHostToDevice, s));
    k_oct_sub<<<nblocks(cnt[3], DW), 64 * DW, 0, s>>>(cnt[3], (const Node *)w.subs.p, w.ea.p, x, tag, rtag, itag,
                                                      w.counts.p + 4);
    int bad = 0;
    TG_HIP(hipMemcpyAsync(&bad, w.counts.p + 4, sizeof(int), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    if (bad) throw std::runtime_error("tioga_amd: octree dedupe: subtree stack exceeded its bound");
  }
  if (cnt[1] > 0)
    k_oct_leaf<<<nblocks(cnt[1], 64), 64, 0, s>>>(cnt[1], (const Node *)w.leaves.p, w.ea.p, x, tag, rtag, itag);
  if (!single)
    for (int k = 0; k < nset; k++)
      if (sets[k].n) {
        TG_HIP(hipMemcpyAsync(sets[k].rtag, rtag + base[k], sizeof(double) * sets[k].n, hipMemcpyDeviceToDevice, s));
        TG_HIP(hipMemcpyAsync(sets[k].itag, itag + base[k], sizeof(int) * sets[k].n, hipMemcpyDeviceToDevice, s));
      }
}

void unique_nodes_device(int n, const double *x, const int *tag, double *rtag, int *itag, DedupeScratch &w,
                         hipStream_t s)
{
  DedupeSet d{n, x, tag, rtag, itag};
  unique_nodes_batch(&d, 1, w, s);
}

}  // namespace tg
