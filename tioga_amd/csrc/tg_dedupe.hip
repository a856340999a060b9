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
// points into its children.  Leaves are queued and finished one per thread.
#include <stdexcept>

#include "tg_engine.h"

namespace tg {

namespace {

constexpr int BT = 256;
constexpr int LEAF = 20;  // utils.c:633: split above 20 points

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
                                                  int *__restrict__ nleaves, int *ea, int *tmp,
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
    if (c.len > LEAF) out[atomicAdd(nout, 1)] = c;
    else leaves[atomicAdd(nleaves, 1)] = c;
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

__global__ void k_iota2(int n, int *a, int *b)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = b[i] = i;
}

}  // namespace

void unique_nodes_device(int n, const double *x, const int *tag, double *rtag, int *itag, DedupeScratch &w,
                         hipStream_t s)
{
  if (n <= 0) return;
  w.ea.reserve(n);
  w.tmp.reserve(n);
  // node lists: each level's nodes partition the points still being split
  w.nodesA.reserve(n / LEAF + 2);
  w.nodesB.reserve(n / LEAF + 2);
  w.leaves.reserve(n + 1);
  w.counts.reserve(2);
  k_iota2<<<nblocks(n, 256), 256, 0, s>>>(n, w.ea.p, itag);
  int cnt[2] = {0, 0};  // [0] next-level nodes, [1] leaves
  Node root{0, n};
  if (n > LEAF) {
    TG_HIP(hipMemcpyAsync(w.nodesA.p, &root, sizeof root, hipMemcpyHostToDevice, s));
    TG_HIP(hipMemsetAsync(w.counts.p, 0, 2 * sizeof(int), s));
    Node *in = (Node *)w.nodesA.p, *out = (Node *)w.nodesB.p;
    int nin = 1;
    while (nin > 0) {
      TG_HIP(hipMemsetAsync(w.counts.p, 0, sizeof(int), s));
      k_oct_level<<<nin, BT, 0, s>>>(nin, in, out, w.counts.p, (Node *)w.leaves.p, w.counts.p + 1, w.ea.p, w.tmp.p,
                                     x);
      TG_HIP(hipMemcpyAsync(cnt, w.counts.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
      TG_HIP(hipStreamSynchronize(s));
      nin = cnt[0];
      std::swap(in, out);
    }
  } else {
    TG_HIP(hipMemcpyAsync(w.leaves.p, &root, sizeof root, hipMemcpyHostToDevice, s));
    TG_HIP(hipStreamSynchronize(s));
    cnt[1] = 1;
  }
  if (cnt[1] > 0)
    k_oct_leaf<<<nblocks(cnt[1], 64), 64, 0, s>>>(cnt[1], (const Node *)w.leaves.p, w.ea.p, x, tag, rtag, itag);
}

}  // namespace tg
