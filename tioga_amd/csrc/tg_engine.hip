// tg_engine.hip -- device kernels and orchestration of the connectivity path.
//
// Phase order follows tioga::performConnectivity (tioga.C:112-160): hole
// maps, block-pair boxes, query gathering, per-block donor search, donor
// selection and iblanking, cell iblank.  Each kernel cites the reference
// routine whose result it reproduces.  Order-dependent sequential steps of
// the reference are either reformulated as order-free reductions with the
// same result (documented at each kernel) or run on the host in tg_seq.cpp.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>

#include "tg_engine.h"
#include "tg_prims.h"
#include "tg_seq.h"

namespace tg {

DevBytes &thread_dev_bytes()
{
  static thread_local DevBytes b;
  return b;
}

void hip_check(hipError_t e, const char *what, const char *file, int line)
{
  if (e == hipSuccess) return;
  char buf[512];
  snprintf(buf, sizeof buf, "tioga_amd: HIP error %s (%s) at %s:%d", hipGetErrorString(e), what, file, line);
  fprintf(stderr, "%s\n", buf);
  throw std::runtime_error(buf);
}

namespace {

constexpr int BS = 256;
inline unsigned nb(long long n, int bs = BS) { return (unsigned)((n + bs - 1) / bs); }

__device__ __forceinline__ int type_of(const TypeInfo &ti, int c)
{
  int t = 0;
  while (t < ti.ntypes - 1 && c >= ti.cstart[t + 1]) t++;
  return t;
}

__device__ __forceinline__ void load_xyz(const double *x, int n, double p[3])
{
  p[0] = x[3 * (size_t)n];
  p[1] = x[3 * (size_t)n + 1];
  p[2] = x[3 * (size_t)n + 2];
}

// ---------------------------------------------------------------------------
// topology
// ---------------------------------------------------------------------------

// 1-based per-type connectivity -> 0-based, 8 slots per cell
__global__ void k_pack_conn(int nc, int nv, const int *__restrict__ src, int cell0, int *__restrict__ conn)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  for (int m = 0; m < 8; m++) conn[8 * (size_t)(cell0 + i) + m] = (m < nv) ? src[(size_t)nv * i + m] - BASE : -1;
}

// (node, cell) incidence pairs in cell-major order, and per-node counts
__global__ void k_incidence(int ncells, int nnodes, const int *__restrict__ conn, unsigned *__restrict__ key,
                            int *__restrict__ val, int *__restrict__ cnt)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  for (int m = 0; m < 8; m++) {
    int v = conn[8 * (size_t)c + m];
    size_t o = 8 * (size_t)c + m;
    key[o] = v >= 0 ? (unsigned)v : (unsigned)nnodes;  // unused slots sort last
    val[o] = c;
    if (v >= 0) atomicAdd(&cnt[v], 1);
  }
}

__global__ void k_flag_list(int n, const int *__restrict__ list, unsigned char *__restrict__ flag)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[list[i] - BASE] = 1;
}

// ---------------------------------------------------------------------------
// preprocess: MeshBlock::tagBoundary (MeshBlock.C:94-339)
// ---------------------------------------------------------------------------

__global__ void k_cell_res(int ncells, TypeInfo ti, const int *__restrict__ conn, const double *__restrict__ x,
                           double scale, double *__restrict__ cellRes)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int nv = ti.nv[type_of(ti, c)];
  double xv[8][3];
  for (int m = 0; m < nv; m++) load_xyz(x, conn[8 * (size_t)c + m], xv[m]);
  double vol = cell_volume(xv, nv);
  cellRes[c] = vol * scale;
}

// nodeRes = running sum of incident cell resolutions in cell order, divided by
// the incidence count (MeshBlock.C:152-185).
__global__ void k_node_res(int nnodes, const int *__restrict__ off, const int *__restrict__ n2c,
                           const double *__restrict__ cellRes, double *__restrict__ nodeRes)
{
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nnodes) return;
  double s = 0.0;
  int b = off[v], e = off[v + 1];
  for (int k = b; k < e; k++) s += cellRes[n2c[k]];
  if (e > b) s /= (e - b);
  nodeRes[v] = s;
}

// bin of each node in the 10^3 map over the block's OBB (MeshBlock.C:171-226)
__global__ void k_node_bin(int nnodes, const double *__restrict__ x, OBB o, double dx0, double dx1, double dx2,
                           unsigned *__restrict__ key, int *__restrict__ val)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  int v = nnodes - 1 - i;  // descending node id, so a stable sort keeps it
  double p[3], xd[3];
  load_xyz(x, v, p);
  double mdx[3] = {dx0, dx1, dx2};
  int idx[3];
  for (int j = 0; j < 3; j++) {
    xd[j] = o.dxc[j];
    for (int k = 0; k < 3; k++) xd[j] += (p[k] - o.xc[k]) * o.vec[j][k];
    idx[j] = (int)(xd[j] / mdx[j]);
    idx[j] = idx[j] < 0 ? 0 : (idx[j] > 9 ? 9 : idx[j]);
  }
  key[i] = (unsigned)(idx[2] * 100 + idx[1] * 10 + idx[0]);
  val[i] = v;
}

__global__ void k_scatter_bin(int n, const unsigned *__restrict__ key, const int *__restrict__ val,
                              unsigned *__restrict__ nodebin)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) nodebin[val[i]] = key[i];
}

// Cells touching an overset-boundary node: their nodes get nodeRes = BIGVALUE
// and seed the mexclude layers (MeshBlock.C:228-262).
__global__ void k_obc_cells(int ncells, TypeInfo ti, const int *__restrict__ conn,
                            const unsigned char *__restrict__ obcnode, double *__restrict__ nodeRes,
                            unsigned char *__restrict__ ex, unsigned char *__restrict__ ex1)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  int tag = 0;
  for (int m = 0; m < nv; m++) tag |= obcnode[cv[m]];
  if (!tag) return;
  for (int m = 0; m < nv; m++) {
    nodeRes[cv[m]] = BIGVALUE;
    ex[cv[m]] = 1;
    ex1[cv[m]] = 1;
  }
}

// One mexclude layer (MeshBlock.C:264-335): cells touching a marked node get
// cellRes = BIGVALUE and mark their other nodes for the next layer.
__global__ void k_mexclude(int ncells, TypeInfo ti, const int *__restrict__ conn,
                           const unsigned char *__restrict__ ex, unsigned char *__restrict__ ex1,
                           double *__restrict__ cellRes)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  for (int m = 0; m < nv; m++)
    if (ex[cv[m]] == 1) {
      cellRes[c] = BIGVALUE;
      break;
    }
  if (cellRes[c] == BIGVALUE)
    for (int m = 0; m < nv; m++)
      if (ex[cv[m]] != 1) ex1[cv[m]] = 1;
}

// ---------------------------------------------------------------------------
// hole map: holeMap.C:40-188, MeshBlock.C:666-772, utils.c:192-295
// ---------------------------------------------------------------------------

// Wall-touching cells paint their index box (floor of the offset / spacing,
// clipped to the map) with 2 (MeshBlock.C:666-772).
// cells with a wall node among their vertices
__global__ void k_wall_cells(int ncells, TypeInfo ti, const int *__restrict__ conn,
                             const unsigned char *__restrict__ wallnode, int *__restrict__ flag)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  int w = 0;
  for (int m = 0; m < nv; m++) w |= wallnode[cv[m]];
  flag[c] = w;
}

// Every walled block's wall cells in one launch: cell k of the concatenated
// lists belongs to block b with off[b] <= k < off[b+1].
struct WallSet {
  int nb;
  int off[129];
  const int *cells[128];
  const int *conn[128];
  const double *x[128];
  TypeInfo ti[128];
  int map[128];  // hole map (mesh tag - 1) of the block
};

// MeshBlock::markWallBoundary (MeshBlock.C:637-700): the map cells covered
// by the index box of each wall cell's vertices
__global__ void k_mark_wall(const WallSet *__restrict__ W, const HoleMapView *__restrict__ views,
                            const double *__restrict__ dsv, unsigned char *const *__restrict__ sams)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= W->off[W->nb]) return;
  int lo = 0, hi = W->nb - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (W->off[m] <= k) lo = m;
    else hi = m - 1;
  }
  const int c = W->cells[lo][k - W->off[lo]];
  const TypeInfo &ti = W->ti[lo];
  const int nv = ti.nv[type_of(ti, c)];
  const int *cv = W->conn[lo] + 8 * (size_t)c;
  const int mp = W->map[lo];
  const HoleMapView &h = views[mp];
  const double ds[3] = {dsv[3 * mp], dsv[3 * mp + 1], dsv[3 * mp + 2]};
  int imin[3] = {BIGINT, BIGINT, BIGINT}, imax[3] = {-BIGINT, -BIGINT, -BIGINT};
  for (int m = 0; m < nv; m++) {
    double p[3];
    load_xyz(W->x[lo], cv[m], p);
    for (int q = 0; q < 3; q++) {
      int iv = (int)floor((p[q] - h.ext[q]) / ds[q]);
      imin[q] = iv < imin[q] ? iv : imin[q];
      imax[q] = iv > imax[q] ? iv : imax[q];
    }
  }
  for (int j = 0; j < 3; j++) {
    if (imin[j] < 0) imin[j] = 0;
    if (imax[j] > h.nx[j] - 1) imax[j] = h.nx[j] - 1;
  }
  unsigned char *sam = sams[mp];
  for (int kk = imin[2]; kk <= imax[2]; kk++)
    for (int jj = imin[1]; jj <= imax[1]; jj++)
      for (int ii = imin[0]; ii <= imax[0]; ii++)
        sam[((size_t)kk * h.nx[1] + jj) * h.nx[0] + ii] = 2;
}

__global__ void k_hole_shell(int nx, int ny, int nz, unsigned char *__restrict__ sam)
{
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t n = (size_t)nx * ny * nz;
  if (i >= n) return;
  int a = (int)(i % nx), b = (int)((i / nx) % ny), c = (int)(i / ((size_t)nx * ny));
  if (a == 0 || a == nx - 1 || b == 0 || b == ny - 1 || c == 0 || c == nz - 1) sam[i] = 1;
}

// Flood fill of the outside region (utils.c:228-290).  The reference repeats
// in-place sweeps painting 0-cells that touch a 1-cell through the isym
// stencil until nothing changes; its fixed point is the set of 0-cells joined
// to the boundary shell through 0-cells, which any sweep order reaches.  Here
// each thread owns one interior line along `axis` and propagates forwards and
// backwards along it, also taking paint from the allowed cross neighbours.
__global__ void k_line_fill(int nx, int ny, int nz, int axis, int allow, unsigned char *__restrict__ sam,
                            int *__restrict__ changed)
{
  int n[3] = {nx, ny, nz};
  size_t st[3] = {1, (size_t)nx, (size_t)nx * ny};
  int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
  int m1 = n[a1] - 2, m2 = n[a2] - 2;
  if (m1 <= 0 || m2 <= 0 || n[axis] < 3) return;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)m1 * m2) return;
  int i1 = 1 + (int)(t % m1), i2 = 1 + (int)(t / m1);
  size_t base = i1 * st[a1] + i2 * st[a2];
  size_t sa = st[axis];
  bool c1 = (allow >> a1) & 1, c2 = (allow >> a2) & 1;
  int ch = 0;
  auto paintable = [&](size_t m) -> bool {
    if (c1 && (sam[m - st[a1]] == 1 || sam[m + st[a1]] == 1)) return true;
    if (c2 && (sam[m - st[a2]] == 1 || sam[m + st[a2]] == 1)) return true;
    return false;
  };
  unsigned char prev = sam[base];
  for (int i = 1; i < n[axis] - 1; i++) {
    size_t m = base + i * sa;
    unsigned char v = sam[m];
    if (v == 0 && (prev == 1 || paintable(m))) {
      v = 1;
      sam[m] = 1;
      ch = 1;
    }
    prev = v;
  }
  prev = sam[base + (size_t)(n[axis] - 1) * sa];
  for (int i = n[axis] - 2; i >= 1; i--) {
    size_t m = base + i * sa;
    unsigned char v = sam[m];
    if (v == 0 && (prev == 1 || paintable(m))) {
      v = 1;
      sam[m] = 1;
      ch = 1;
    }
    prev = v;
  }
  if (ch) *changed = 1;
}

__global__ void k_hole_final(size_t n, unsigned char *__restrict__ sam)
{
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) sam[i] = (sam[i] == 0 || sam[i] == 2) ? 1 : 0;
}

// Flood fill of one 2-D slice in LDS.  With isym 3 (2) [1] the stencil has
// no z (y) [x] neighbours, so every slice normal to that axis fills on its
// own; the interior slices are filled here, the two outer ones belong to
// the boundary shell.  The workgroup loads the slice, repeats row and
// column sweeps (the k_line_fill rule) until nothing changes, and stores it.
constexpr int SLICE_MAX = 48 * 1024;

// One sweep step of the slice fill over m <= FK cells from sl[i0] in
// direction DIR (cell k at i0 + DIR * k * step, step 1 along a row, nu along
// a column); cross neighbours at -+ xs; prev is the cell before the first.
// A cell of value 0 next to a 1 (the previous cell in the sweep or a cross
// neighbour) becomes 1.  Returns the last cell's value.
constexpr int FK = 16;
template <int DIR>
__device__ __forceinline__ unsigned char fill_chunk(unsigned char *sl, int i0, int m, int xs, unsigned char prev,
                                                    int &ch, int step = 1)
{
  unsigned char c[FK], a[FK], b[FK];
#pragma unroll
  for (int k = 0; k < FK; k++)
    if (k < m) {
      const int i = i0 + DIR * k * step;
      c[k] = sl[i];
      a[k] = sl[i - xs];
      b[k] = sl[i + xs];
    }
  unsigned set = 0;
#pragma unroll
  for (int k = 0; k < FK; k++)
    if (k < m) {
      if (c[k] == 0 && (prev == 1 || a[k] == 1 || b[k] == 1)) {
        c[k] = 1;
        set |= 1u << k;
      }
      prev = c[k];
    }
  if (set) {
    ch = 1;
#pragma unroll
    for (int k = 0; k < FK; k++)
      if ((set >> k) & 1u) sl[i0 + DIR * k * step] = 1;
  }
  return prev;
}

__global__ void __launch_bounds__(256) k_fill_slices(int axis, int nmaps, unsigned char *const *__restrict__ sams,
                                                     const int *__restrict__ nxs)
{
  extern __shared__ unsigned char sl[];  // the slice (launch-sized, at most SLICE_MAX bytes)
  __shared__ int chg;
  const int m = blockIdx.y;
  const int n[3] = {nxs[3 * m], nxs[3 * m + 1], nxs[3 * m + 2]};
  const int k = blockIdx.x + 1;  // interior slice index along axis
  if (k >= n[axis] - 1) return;
  unsigned char *sam = sams[m];
  const size_t st[3] = {1, (size_t)n[0], (size_t)n[0] * n[1]};
  const int ua = axis == 0 ? 1 : 0, va = axis == 2 ? 1 : 2;  // in-slice axes, u fastest in memory
  const int nu = n[ua], nv = n[va];
  const int t = threadIdx.x;
  for (int i = t; i < nu * nv; i += blockDim.x) {
    const int u = i % nu, v = i / nu;
    sl[i] = sam[k * st[axis] + u * st[ua] + v * st[va]];
  }
  __syncthreads();
  for (;;) {
    if (t == 0) chg = 0;
    __syncthreads();
    int ch = 0;
    // rows: along u, interior rows v = 1..nv-2, cross neighbours v +- 1.
    // Each row in register chunks of FK cells: loads first, the sequential
    // update in registers, then the stores (so the chain never waits on an
    // LDS store; cells of the neighbouring rows may be read before or after
    // their owners update them, which the repeat-until-no-change loop allows)
    for (int v = 1 + t; v < nv - 1; v += blockDim.x) {
      const int r0 = v * nu;
      unsigned char prev = sl[r0];
      for (int u0 = 1; u0 < nu - 1; u0 += FK)
        prev = fill_chunk<+1>(sl, r0 + u0, min(FK, nu - 1 - u0), nu, prev, ch);
      prev = sl[r0 + nu - 1];
      for (int u0 = nu - 2; u0 >= 1; u0 -= FK)
        prev = fill_chunk<-1>(sl, r0 + u0, min(FK, u0), nu, prev, ch);
    }
    __syncthreads();
    // columns: along v, interior columns u = 1..nu-2, cross neighbours u +- 1
    for (int u = 1 + t; u < nu - 1; u += blockDim.x) {
      unsigned char prev = sl[u];
      for (int v0 = 1; v0 < nv - 1; v0 += FK)
        prev = fill_chunk<+1>(sl, v0 * nu + u, min(FK, nv - 1 - v0), 1, prev, ch, nu);
      prev = sl[(nv - 1) * nu + u];
      for (int v0 = nv - 2; v0 >= 1; v0 -= FK)
        prev = fill_chunk<-1>(sl, v0 * nu + u, min(FK, v0), 1, prev, ch, nu);
    }
    if (ch) atomicOr(&chg, 1);
    __syncthreads();
    const int any = chg;
    __syncthreads();
    if (!any) break;
  }
  for (int i = t; i < nu * nv; i += blockDim.x) {
    const int u = i % nu, v = i / nu;
    sam[k * st[axis] + u * st[ua] + v * st[va]] = sl[i];
  }
}

// All hole maps of a connectivity at once: the same steps, one launch each,
// a two-dimensional grid of (cells or interior lines of one map, map).
struct MapSet {
  int nmap;
  unsigned char *sam[128];
  int nx[128][3];
};

// step 0: clear; step 1: boundary shell to 1; step 2: final values
__global__ void k_maps_cells(const MapSet *__restrict__ M, int step)
{
  const int m = blockIdx.y;
  const int nx = M->nx[m][0], ny = M->nx[m][1], nz = M->nx[m][2];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx * ny * nz) return;
  unsigned char *sam = M->sam[m];
  if (step == 0) {
    sam[k] = 0;
  } else if (step == 1) {
    const int a = k % nx, b = (k / nx) % ny, c = k / (nx * ny);
    if (a == 0 || a == nx - 1 || b == 0 || b == ny - 1 || c == 0 || c == nz - 1) sam[k] = 1;
  } else {
    sam[k] = (sam[k] == 0 || sam[k] == 2) ? 1 : 0;
  }
}

// k_line_fill over the interior lines along `axis` of every active map
__global__ void k_maps_line_fill(const MapSet *__restrict__ M, int axis, int allow, const int *__restrict__ active,
                                 int *__restrict__ changed)
{
  const int mp = blockIdx.y;
  if (!active[mp]) return;
  const int n[3] = {M->nx[mp][0], M->nx[mp][1], M->nx[mp][2]};
  const int a1 = (axis + 1) % 3, a2 = (axis + 2) % 3;
  const int m1 = n[a1] - 2, m2 = n[a2] - 2;
  if (m1 <= 0 || m2 <= 0 || n[axis] < 3) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m1 * m2) return;
  unsigned char *sam = M->sam[mp];
  const int st[3] = {1, n[0], n[0] * n[1]};
  const int i1 = 1 + t % m1, i2 = 1 + t / m1;
  const int base = i1 * st[a1] + i2 * st[a2];
  const int sa = st[axis];
  const bool c1 = (allow >> a1) & 1, c2 = (allow >> a2) & 1;
  int ch = 0;
  auto paintable = [&](int m) -> bool {
    if (c1 && (sam[m - st[a1]] == 1 || sam[m + st[a1]] == 1)) return true;
    if (c2 && (sam[m - st[a2]] == 1 || sam[m + st[a2]] == 1)) return true;
    return false;
  };
  unsigned char prev = sam[base];
  for (int i = 1; i < n[axis] - 1; i++) {
    const int m = base + i * sa;
    unsigned char v = sam[m];
    if (v == 0 && (prev == 1 || paintable(m))) {
      v = 1;
      sam[m] = 1;
      ch = 1;
    }
    prev = v;
  }
  prev = sam[base + (n[axis] - 1) * sa];
  for (int i = n[axis] - 2; i >= 1; i--) {
    const int m = base + i * sa;
    unsigned char v = sam[m];
    if (v == 0 && (prev == 1 || paintable(m))) {
      v = 1;
      sam[m] = 1;
      ch = 1;
    }
    prev = v;
  }
  if (ch) changed[mp] = 1;
}

// checkHoleMap (utils.c:170-185): truncation toward zero, range test on the int
__device__ __forceinline__ int check_hole(const double p[3], const HoleMapView &h)
{
  // a point a whole map span outside the map's extents has |index| >= nx
  // along that axis, whatever the rounding: out, without the divisions
  for (int i = 0; i < 3; i++) {
    const double span = h.ext[i + 3] - h.ext[i];
    if (p[i] < h.ext[i] - span || p[i] > h.ext[i + 3] + span) return 0;
  }
  int ix[3];
  for (int i = 0; i < 3; i++) {
    double dx = (h.ext[i + 3] - h.ext[i]) / h.nx[i];
    ix[i] = (int)((p[i] - h.ext[i]) / dx);
    if (ix[i] < 0 || ix[i] > h.nx[i] - 1) return 0;
  }
  return h.sam[((size_t)ix[2] * h.nx[1] + ix[1]) * h.nx[0] + ix[0]];
}

// ---------------------------------------------------------------------------
// boxes: MeshBlock::getReducedOBB (MeshBlock.C:774-838)
// ---------------------------------------------------------------------------

// Bounds, in the block's own OBB frame, of its cells whose box in the
// partner's OBB frame overlaps the partner OBB.  min/max are exact, so the
// reduction order does not matter.
__global__ void k_reduced_obb(int ncells, TypeInfo ti, const int *__restrict__ conn, const double *__restrict__ x,
                              OBB own, OBB obc, unsigned long long *__restrict__ out)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
  bool keep = false;
  if (c < ncells) {
    int nv = ti.nv[type_of(ti, c)];
    const int *cv = conn + 8 * (size_t)c;
    double bb[6] = {BIGVALUE, BIGVALUE, BIGVALUE, -BIGVALUE, -BIGVALUE, -BIGVALUE};
    double p[3], xd[3];
    for (int m = 0; m < nv; m++) {
      load_xyz(x, cv[m], p);
      to_frame(p, obc.xc, obc.vec, xd);
      for (int j = 0; j < 3; j++) bb[j] = dmin(bb[j], xd[j]);
      for (int j = 0; j < 3; j++) bb[j + 3] = dmax(bb[j + 3], xd[j]);
    }
    keep = true;
    for (int j = 0; j < 3; j++)
      if (bb[j] > obc.dxc[j]) keep = false;
    for (int j = 0; j < 3; j++)
      if (bb[j + 3] < -obc.dxc[j]) keep = false;
    if (keep)
      for (int m = 0; m < nv; m++) {
        load_xyz(x, cv[m], p);
        to_frame(p, own.xc, own.vec, xd);
        for (int j = 0; j < 3; j++) lo[j] = dmin(lo[j], xd[j]);
        for (int j = 0; j < 3; j++) hi[j] = dmax(hi[j], xd[j]);
      }
  }
  // wave-level, then block-level reduction, then one atomic per block
  __shared__ unsigned long long red[6][BS / 64];
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int j = 0; j < 3; j++) {
    unsigned long long kl = dkey(lo[j]), kh = dkey(hi[j]);
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long a = __shfl_xor(kl, o), b = __shfl_xor(kh, o);
      kl = a < kl ? a : kl;
      kh = b > kh ? b : kh;
    }
    if ((threadIdx.x & 63) == 0) {
      red[j][w] = kl;
      red[j + 3][w] = kh;
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int j = threadIdx.x;
    unsigned long long k = red[j][0];
    for (int q = 1; q < nw; q++) k = (j < 3) ? (red[j][q] < k ? red[j][q] : k) : (red[j][q] > k ? red[j][q] : k);
    if (j < 3) atomicMin(&out[j], k);
    else atomicMax(&out[j], k);
  }
}

// The same for every pair of one local block in one pass over its cells:
// each cell's vertices are loaded once and tested against every partner
// OBB; a wave with a kept cell for a pair reduces and adds its box.
struct BoxPairs {
  int np;
  int pair[128];
  OBB obc[128];
};

__global__ void __launch_bounds__(BS) k_reduced_obb_multi(int ncells, TypeInfo ti, const int *__restrict__ conn,
                                                          const double *__restrict__ x, OBB own,
                                                          const BoxPairs *__restrict__ P,
                                                          unsigned long long *__restrict__ out)
{
  // per-workgroup bounds of every pair, folded over a grid-stride loop, so
  // the global atomics come once per workgroup rather than once per wave
  __shared__ unsigned long long skl[128 * 3], skh[128 * 3];
  const int np = P->np;
  for (int q = threadIdx.x; q < 3 * np; q += blockDim.x) {
    skl[q] = dkey(BIGVALUE);
    skh[q] = dkey(-BIGVALUE);
  }
  __syncthreads();
  for (int base = blockIdx.x * blockDim.x; base < ncells; base += gridDim.x * blockDim.x) {
    const int c = base + threadIdx.x;
    const bool valid = c < ncells;
    int nv = 0;
    const int *cv = conn + 8 * (size_t)(valid ? c : 0);
    double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
    double cen[3] = {0, 0, 0}, rad = 0;
    if (valid) {
      nv = ti.nv[type_of(ti, c)];
      for (int m = 0; m < nv; m++) {
        double p[3], xd[3];
        load_xyz(x, cv[m], p);
        to_frame(p, own.xc, own.vec, xd);
        for (int j = 0; j < 3; j++) {
          lo[j] = dmin(lo[j], xd[j]);
          hi[j] = dmax(hi[j], xd[j]);
          cen[j] += p[j];
        }
      }
      for (int j = 0; j < 3; j++) cen[j] /= nv;
      for (int m = 0; m < nv; m++) {
        double p[3];
        load_xyz(x, cv[m], p);
        const double d = sqrt((p[0] - cen[0]) * (p[0] - cen[0]) + (p[1] - cen[1]) * (p[1] - cen[1]) +
                              (p[2] - cen[2]) * (p[2] - cen[2]));
        rad = dmax(rad, d);
      }
      // a margin far above the rounding of the frame transforms
      rad = rad * (1 + 1e-9) + 1e-12 * (1 + fabs(cen[0]) + fabs(cen[1]) + fabs(cen[2]));
    }
    // the wave's sphere: around the box of its cells' spheres, so it holds
    // every one of them; a partner box it misses is missed by every cell
    double wlo[3], whi[3];
    for (int j = 0; j < 3; j++) {
      double a = valid ? cen[j] - rad : BIGVALUE, b = valid ? cen[j] + rad : -BIGVALUE;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        a = dmin(a, __shfl_xor(a, o));
        b = dmax(b, __shfl_xor(b, o));
      }
      wlo[j] = a;
      whi[j] = b;
    }
    double wc[3], wr = 0;
    for (int j = 0; j < 3; j++) {
      wc[j] = 0.5 * (wlo[j] + whi[j]);
      wr += (whi[j] - wlo[j]) * (whi[j] - wlo[j]);
    }
    wr = 0.5 * sqrt(wr) * (1 + 1e-9) + 1e-12 * (1 + fabs(wc[0]) + fabs(wc[1]) + fabs(wc[2]));
    for (int k = 0; k < np; k++) {
      const OBB &obc = P->obc[k];
      {
        double cd[3];
        to_frame(wc, obc.xc, obc.vec, cd);
        bool near = true;
        for (int j = 0; j < 3; j++) near = near && (fabs(cd[j]) <= obc.dxc[j] + wr * (1 + 1e-9) + 1e-12);
        if (!__ballot(near)) continue;  // the same value in every lane
      }
      bool keep = false;
      if (valid) {
        // the cell's bounding sphere misses the partner box: the vertex test
        // below can not keep it either
        double cd[3];
        to_frame(cen, obc.xc, obc.vec, cd);
        bool near = true;
        for (int j = 0; j < 3; j++) near = near && (fabs(cd[j]) <= obc.dxc[j] + rad * (1 + 1e-9) + 1e-12);
        if (near) {
          double bb[6] = {BIGVALUE, BIGVALUE, BIGVALUE, -BIGVALUE, -BIGVALUE, -BIGVALUE};
          for (int m = 0; m < nv; m++) {
            double p[3], xd[3];
            load_xyz(x, cv[m], p);
            to_frame(p, obc.xc, obc.vec, xd);
            for (int j = 0; j < 3; j++) bb[j] = dmin(bb[j], xd[j]);
            for (int j = 0; j < 3; j++) bb[j + 3] = dmax(bb[j + 3], xd[j]);
          }
          keep = true;
          for (int j = 0; j < 3; j++)
            if (bb[j] > obc.dxc[j]) keep = false;
          for (int j = 0; j < 3; j++)
            if (bb[j + 3] < -obc.dxc[j]) keep = false;
        }
      }
      if (__ballot(keep) == 0) continue;
      for (int j = 0; j < 3; j++) {
        unsigned long long kl = keep ? dkey(lo[j]) : dkey(BIGVALUE), kh = keep ? dkey(hi[j]) : dkey(-BIGVALUE);
        for (int o = 32; o > 0; o >>= 1) {
          unsigned long long a = __shfl_xor(kl, o), b = __shfl_xor(kh, o);
          kl = a < kl ? a : kl;
          kh = b > kh ? b : kh;
        }
        if ((threadIdx.x & 63) == 0) {
          atomicMin(&skl[3 * k + j], kl);
          atomicMax(&skh[3 * k + j], kh);
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * np; q += blockDim.x) {
    const int k = q / 3, j = q % 3;
    if (skl[q] != dkey(BIGVALUE)) atomicMin(&out[6 * (size_t)P->pair[k] + j], skl[q]);
    if (skh[q] != dkey(-BIGVALUE)) atomicMax(&out[6 * (size_t)P->pair[k] + 3 + j], skh[q]);
  }
}

// ---------------------------------------------------------------------------
// query gathering: MeshBlock::getQueryPoints2 (MeshBlock.C:953-1117)
// ---------------------------------------------------------------------------

// binoff[k] = first position of bin k in the (bin, node) order, written at
// the positions where the bin changes (bins with no node keep V and are
// fixed up on the host)
__global__ void k_bin_starts(int n, const unsigned *__restrict__ sorted_bin, int *__restrict__ binoff)
{
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned k = sorted_bin[p];
  if (p == 0 || sorted_bin[p - 1] != k) binoff[k] = p;
}

// Candidate c of a pair's bin ranges (positions of the sender's bin order,
// bins in (l, k, j) order): flag when the node lies in the partner's reduced
// OBB (MeshBlock.C:1064-1081).  cand0[i] is pair i's first candidate; rng
// holds per range (first candidate, first position), per pair in bin order.
struct QueryPairs {
  const int *cand0;      // npairs + 1
  const int2 *rng;       // ranges of all pairs
  const int *rng0;       // npairs + 1: first range of each pair
  const OBB *obc;        // npairs: the partner's reduced OBB
  const int *blk;        // npairs: the sending block
  const int *invmap[128];
  const double *x[128];
};

__global__ void k_query_cand(int ncand, int npairs, const QueryPairs *__restrict__ Q, int *__restrict__ flag,
                             int *__restrict__ pos)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncand) return;
  int lo = 0, hi = npairs - 1;  // pair: last with cand0 <= c
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (Q->cand0[m] <= c) lo = m;
    else hi = m - 1;
  }
  const int i = lo;
  int a = Q->rng0[i], z = Q->rng0[i + 1] - 1;  // range: last with first candidate <= c
  while (a < z) {
    const int m = (a + z + 1) >> 1;
    if (Q->rng[m].x <= c) a = m;
    else z = m - 1;
  }
  const int p = Q->rng[a].y + (c - Q->rng[a].x);
  const int b = Q->blk[i];
  const int v = Q->invmap[b][p];
  double q[3], xd[3];
  load_xyz(Q->x[b], v, q);
  const OBB &o = Q->obc[i];
  to_frame(q, o.xc, o.vec, xd);
  flag[c] = (fabs(xd[0]) <= o.dxc[0] && fabs(xd[1]) <= o.dxc[1] && fabs(xd[2]) <= o.dxc[2]);
  pos[c] = p;
}

// positions of the selected candidates, and per pair the number selected
// (the selection is ascending, so each pair's are one run)
__global__ void k_take_pos(int n, const int *__restrict__ sel, const int *__restrict__ pos, int *__restrict__ out)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = pos[sel[j]];
}

__global__ void k_pair_runs(int npairs, int nsel, const int *__restrict__ sel, const int *__restrict__ cand0,
                            int *__restrict__ first)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > npairs) return;
  int lo = 0, hi = nsel;  // first selected index >= cand0[i]
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (sel[m] < cand0[i]) lo = m + 1;
    else hi = m;
  }
  first[i] = lo;
}

// ---------------------------------------------------------------------------
// search: MeshBlock::search (search.C:107-265)
// ---------------------------------------------------------------------------

// Cells whose box in the query-set OBB frame overlaps that OBB (search.C:123-167)
__device__ __forceinline__ void cell_filter_one(int c, int ncells, const TypeInfo &ti, const int *__restrict__ conn,
                                                const double *__restrict__ x, const OBB &obq, int *__restrict__ flag)
{
  int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
  for (int m = 0; m < nv; m++) {
    double p[3], xd[3];
    load_xyz(x, cv[m], p);
    for (int j = 0; j < 3; j++) {
      xd[j] = 0;
      for (int k = 0; k < 3; k++) xd[j] += (p[k] - obq.xc[k]) * obq.vec[j][k];
      lo[j] = dmin(lo[j], xd[j]);
      hi[j] = dmax(hi[j], xd[j]);
    }
  }
  double xc[3], dxc[3];
  for (int j = 0; j < 3; j++) {
    xc[j] = (hi[j] + lo[j]) * 0.5;
    dxc[j] = (hi[j] - lo[j]) * 0.5;
  }
  // stored reversed: the element list is LIFO (descending cell index)
  flag[ncells - 1 - c] = (fabs(xc[0]) <= (dxc[0] + obq.dxc[0]) && fabs(xc[1]) <= (dxc[1] + obq.dxc[1]) &&
                          fabs(xc[2]) <= (dxc[2] + obq.dxc[2]));
}

__global__ void k_cell_filter(int ncells, TypeInfo ti, const int *__restrict__ conn, const double *__restrict__ x,
                              OBB obq, int *__restrict__ flag)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  cell_filter_one(c, ncells, ti, conn, x, obq, flag);
}

// element list and element boxes (search.C:183-219)
__device__ __forceinline__ void elem_box_one(int e, int ncells, const int *__restrict__ sel, int sel0,
                                             const TypeInfo &ti, const int *__restrict__ conn,
                                             const double *__restrict__ x, int *__restrict__ elist,
                                             double *__restrict__ ebox)
{
  int c = ncells - 1 - (sel[e] - sel0);
  elist[e] = c;
  int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
  for (int m = 0; m < nv; m++) {
    double p[3];
    load_xyz(x, cv[m], p);
    for (int j = 0; j < 3; j++) {
      lo[j] = dmin(lo[j], p[j]);
      hi[j] = dmax(hi[j], p[j]);
    }
  }
  for (int j = 0; j < 3; j++) {
    ebox[6 * (size_t)e + j] = lo[j];
    ebox[6 * (size_t)e + 3 + j] = hi[j];
  }
}

__global__ void k_elem_box(int nelem, int ncells, const int *__restrict__ sel, int sel0, TypeInfo ti,
                           const int *__restrict__ conn, const double *__restrict__ x, int *__restrict__ elist,
                           double *__restrict__ ebox)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nelem) return;
  elem_box_one(e, ncells, sel, sel0, ti, conn, x, elist, ebox);
}

// The blocks of a batched search preparation: block a's cells are
// [coff, coff + ncells) of the concatenated flags, its elements
// [eoff, eoff + nelem) of the compacted list (element arrays back to back).
struct PrepBlk {
  TypeInfo ti;
  const int *conn;
  const double *x;
  OBB ob;
  int coff, ncells, eoff;
};

__device__ __forceinline__ int prep_find(int na, const PrepBlk *__restrict__ T, int g, bool by_elem)
{
  int lo = 0, hi = na - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((by_elem ? T[mid].eoff : T[mid].coff) <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// k_cell_filter for every block of the batch
__global__ void k_cell_filter_all(int na, int total, const PrepBlk *__restrict__ T, int *__restrict__ flag)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const PrepBlk &B = T[prep_find(na, T, g, false)];
  cell_filter_one(g - B.coff, B.ncells, B.ti, B.conn, B.x, B.ob, flag + B.coff);
}

// k_elem_box for every block of the batch (sel: the compacted flag indices)
__global__ void k_elem_box_all(int na, int total, const PrepBlk *__restrict__ T, const int *__restrict__ sel,
                               int *__restrict__ elist, double *__restrict__ ebox)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const PrepBlk &B = T[prep_find(na, T, g, true)];
  elem_box_one(g - B.eoff, B.ncells, sel + B.eoff, B.coff, B.ti, B.conn, B.x, elist + B.eoff,
               ebox + 6 * (size_t)B.eoff);
}

// start[s] = first position j of the ascending list sorted[0..n) with
// sorted[j] >= bound[s], s = 0..nseg
__global__ void k_seg_starts(int nseg, int n, const int *__restrict__ sorted, const int *__restrict__ bound,
                             int *__restrict__ start)
{
  int sg = blockIdx.x * blockDim.x + threadIdx.x;
  if (sg > nseg) return;
  const int v = bound[sg];
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (sorted[m] < v) lo = m + 1;
    else hi = m;
  }
  start[sg] = lo;
}

static_assert(sizeof(WalkNode) == 48, "walk node record must be 48 bytes");

// two grid values in one word: a | b << 16
__device__ __forceinline__ unsigned pack2(unsigned a, unsigned b) { return a | (b << 16); }

// a box (lower xyz, upper xyz; FP64 as the reference stores it) padded by TOL
// the way the walk's tests pad it, on the tree's grid, packed three words
__device__ __forceinline__ void grid_box(const double *b, const GridFrame &g, unsigned w[3])
{
  unsigned q[6];
  for (int j = 0; j < 6; j++) {
    const double v = j < 3 ? b[j] - TOL : b[j] + TOL;
    q[j] = grid_of(v, g.lo[j % 3], g.inv[j % 3]);
  }
  w[0] = pack2(q[0], q[1]);
  w[1] = pack2(q[2], q[3]);
  w[2] = pack2(q[4], q[5]);
}

// ADT arrays (adtIntegers, adtReals, element boxes) -> walk records
__device__ __forceinline__ void k_adt_node_one(int n, const TypeInfo &ti, const GridFrame &g,
                                               const int *__restrict__ adtInt, const double *__restrict__ adtReal,
                                               const double *__restrict__ ebox, const int *__restrict__ elist,
                                               WalkNode *__restrict__ nodes, int *__restrict__ rcell)
{
  int4 a = reinterpret_cast<const int4 *>(adtInt)[n];
  WalkNode N;
  const int cell = elist[a.x];
  N.cellnv = cell | (ti.nv[type_of(ti, cell)] << 28);
  rcell[n] = cell;
  N.left = a.y > -1 ? adtInt[4 * (size_t)a.y + 3] : -1;
  N.right = a.z > -1 ? adtInt[4 * (size_t)a.z + 3] : -1;
  grid_box(ebox + 6 * (size_t)a.x, g, N.e);
  if (N.left > -1) grid_box(adtReal + 6 * (size_t)N.left, g, N.l);
  else N.l[0] = N.l[1] = N.l[2] = 0;
  if (N.right > -1) grid_box(adtReal + 6 * (size_t)N.right, g, N.r);
  else N.r[0] = N.r[1] = N.r[2] = 0;
  nodes[n] = N;
}

__global__ void k_adt_nodes(int nelem, TypeInfo ti, GridFrame g, const int *__restrict__ adtInt,
                            const double *__restrict__ adtReal, const double *__restrict__ ebox,
                            const int *__restrict__ elist, WalkNode *__restrict__ nodes, int *__restrict__ rcell)
{
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nelem) return;
  k_adt_node_one(n, ti, g, adtInt, adtReal, ebox, elist, nodes, rcell);
}

// k_adt_nodes for every tree of a batch whose arrays lie back to back: tree
// t's elements are [base[t], base[t + 1]) of the concatenation, its adtInt
// entries tree-local as the reference stores them
struct AdtNodesTree {
  TypeInfo ti;
  GridFrame g;
  int base;
};
__global__ void k_adt_nodes_all(int ntree, int total, const AdtNodesTree *__restrict__ T,
                                const int *__restrict__ adtInt, const double *__restrict__ adtReal,
                                const double *__restrict__ ebox, const int *__restrict__ elist,
                                WalkNode *__restrict__ nodes, int *__restrict__ rcell)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  int lo = 0, hi = ntree - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T[mid].base <= g) lo = mid;
    else hi = mid - 1;
  }
  const size_t b = T[lo].base;
  k_adt_node_one(g - (int)b, T[lo].ti, T[lo].g, adtInt + 4 * b, adtReal + 6 * b, ebox + 6 * b, elist + b,
                 nodes + b, rcell + b);
}

// 30-bit Morton code of each query in the frame of the search structure's
// root extents, so a wave takes a compact cluster of points (coherent
// traversal paths and cell data in cache); the search order does not change
// any result, each query's walk is independent.
__device__ __forceinline__ unsigned spread10(unsigned v)
{
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Keys for one sort of every set: (set << (mb + 1)) | (skip << mb) |
// (Morton >> (30 - mb)), so each set's queries stay in the set's own range,
// the ones that walk the tree come first in Morton order, and the skipped
// ones last: a coincident copy (its representative's donor is copied later,
// search.C:251-265) or a point outside the root box, whose donor is -1
// (searchADTrecursion.C:38-45; written here).  idx = the query's index
// within its set.
__global__ void k_morton_sets(int ns, int tot, const MortonSet *__restrict__ M, int mb, unsigned *__restrict__ key,
                              int *__restrict__ idx, int *__restrict__ qsid)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= tot) return;
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (M[mid].off <= g) lo = mid;
    else hi = mid - 1;
  }
  const MortonSet &m = M[lo];
  const int i = g - m.off;
  if (ns > 1) qsid[g] = lo;  // the set of query (and position) g
  double p[3];
  load_xyz(m.xs, i, p);
  bool inroot = true;
  for (int d = 0; d < 3; d++) inroot = inroot && (p[d] >= m.ext[2 * d] - TOL);
  for (int d = 0; d < 3; d++) inroot = inroot && (p[d] <= m.ext[2 * d + 1] + TOL);
  const bool dup = m.xtag && m.xtag[i] != i;
  if (!inroot && !dup) m.donorId[i] = -1;
  unsigned c[3];
  for (int d = 0; d < 3; d++) {
    double u = (p[d] - m.ext[2 * d]) / (m.ext[2 * d + 1] - m.ext[2 * d]);
    u = u < 0 ? 0 : (u > 0.999999 ? 0.999999 : u);
    c[d] = (unsigned)(u * 1024.0);
  }
  const unsigned m30 = (spread10(c[0]) << 2) | (spread10(c[1]) << 1) | spread10(c[2]);
  const unsigned mk = mb > 0 ? (m30 >> (30 - mb)) : 0u;
  key[g] = ((unsigned)lo << (mb + 1)) | ((inroot && !dup) ? 0u : (1u << mb)) | mk;
  idx[g] = i;
}

// nin[k] = the walking queries of set k: in the sorted keys of the set's
// range they come first, so their count is the first position whose key
// has the skip bit (a binary search; no atomics).  wstart[k] = the first of
// set k's walking waves in the enumeration of all sets' walking waves
// (wstart[ns] = their total).  One workgroup.
__global__ void __launch_bounds__(256) k_walk_counts(int ns, const MortonSet *__restrict__ M, int mb,
                                                      const unsigned *__restrict__ key, int *__restrict__ nin,
                                                      int *__restrict__ wstart)
{
  for (int k = threadIdx.x; k < ns; k += blockDim.x) {
    const unsigned first_skip = ((unsigned)k << (mb + 1)) | (1u << mb);
    int lo = M[k].off, hi = M[k].off + M[k].nq;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (key[mid] < first_skip) lo = mid + 1;
      else hi = mid;
    }
    nin[k] = lo - M[k].off;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int w = 0;
    for (int k = 0; k < ns; k++) {
      wstart[k] = w;
      w += (nin[k] + 63) / 64;
    }
    wstart[ns] = w;
  }
}

// The processing order without Morton bits is each set's walking queries in
// index order (what a stable sort on (set, skip) gives), found by a
// compaction instead of a sort.  k_walk_mask: per wave of 64 queries, the
// mask of the walking ones and its population; donors of the points outside
// the root box are written as in k_morton_sets.
__global__ void __launch_bounds__(256) k_walk_mask(int ns, int tot, const MortonSet *__restrict__ M,
                                                    unsigned long long *__restrict__ mask, int *__restrict__ cnt)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  auto test = [&](const MortonSet &m) {
    const int i = g - m.off;
    double p[3];
    load_xyz(m.xs, i, p);
    bool inroot = true;
    for (int d = 0; d < 3; d++) inroot = inroot && (p[d] >= m.ext[2 * d] - TOL);
    for (int d = 0; d < 3; d++) inroot = inroot && (p[d] <= m.ext[2 * d + 1] + TOL);
    const bool dup = m.xtag && m.xtag[i] != i;
    if (!inroot && !dup) m.donorId[i] = -1;
    return inroot && !dup;
  };
  // the set of the wave's first query by wave-uniform (scalar) loads; a
  // wave across a set boundary looks each lane's set up
  const int g0 = __builtin_amdgcn_readfirstlane(g);
  int s0 = 0;
  for (int hi = ns - 1; s0 < hi;) {
    const int mid = (s0 + hi + 1) >> 1;
    if (M[mid].off <= g0) s0 = mid;
    else hi = mid - 1;
  }
  bool walk = false;
  if (s0 == ns - 1 || M[s0 + 1].off > g0 + 63) {
    if (g < tot) walk = test(M[s0]);
  } else if (g < tot) {
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (M[mid].off <= g) lo = mid;
      else hi = mid - 1;
    }
    walk = test(M[lo]);
  }
  const unsigned long long b = __ballot(walk);
  if ((threadIdx.x & 63) == 0 && g < tot) {
    mask[g >> 6] = b;
    cnt[g >> 6] = __popcll(b);
  }
}

// walking queries before index x of the concatenated sets (pre: the
// exclusive sum of the per-wave populations, pre[nw] = their total)
__device__ __forceinline__ int walkers_before(int x, int tot, const unsigned long long *__restrict__ mask,
                                              const int *__restrict__ pre)
{
  if (x >= tot) return pre[(tot + 63) >> 6];
  return pre[x >> 6] + __popcll(mask[x >> 6] & ((1ull << (x & 63)) - 1));
}

// each walking query's position: its set's offset + the walking queries of
// its set before it; idx = its index within the set, qsid = its set
__global__ void __launch_bounds__(256) k_walk_scatter(int ns, int tot, const MortonSet *__restrict__ M,
                                                       const unsigned long long *__restrict__ mask,
                                                       const int *__restrict__ pre, int *__restrict__ idx,
                                                       int *__restrict__ qsid)
{
  // each wave of the grid takes chunks of 64 queries in turn
  const int lane = threadIdx.x & 63, nw = (tot + 63) >> 6;
  const int wstep = (int)(gridDim.x * (blockDim.x >> 6));
  for (int c = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)); c < nw;
       c += wstep) {
    const unsigned long long mw = mask[c];
    if (!mw) continue;
    const int g0 = c << 6, g = g0 + lane;
    // the set of the chunk's first query by wave-uniform (scalar) loads; a
    // chunk across a set boundary looks each lane's set up
    int lo = 0;
    for (int hi = ns - 1; lo < hi;) {
      const int mid = (lo + hi + 1) >> 1;
      if (M[mid].off <= g0) lo = mid;
      else hi = mid - 1;
    }
    if (lo < ns - 1 && M[lo + 1].off <= g0 + 63) {
      int hi = ns - 1;
      lo = 0;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (M[mid].off <= g) lo = mid;
        else hi = mid - 1;
      }
    }
    if (g >= tot || !((mw >> lane) & 1ull)) continue;
    const int off = M[lo].off;
    const int t = walkers_before(g, tot, mask, pre) - walkers_before(off, tot, mask, pre);
    idx[off + t] = g - off;
    if (ns > 1) qsid[off + t] = lo;
  }
}

// nin and wstart as k_walk_counts forms them, from the compaction's counts
__global__ void __launch_bounds__(256) k_walk_counts_mask(int ns, int tot, const MortonSet *__restrict__ M,
                                                           const unsigned long long *__restrict__ mask,
                                                           const int *__restrict__ pre, int *__restrict__ nin,
                                                           int *__restrict__ wstart)
{
  for (int k = threadIdx.x; k < ns; k += blockDim.x)
    nin[k] = walkers_before(M[k].off + M[k].nq, tot, mask, pre) - walkers_before(M[k].off, tot, mask, pre);
  __syncthreads();
  if (threadIdx.x == 0) {
    int w = 0;
    for (int k = 0; k < ns; k++) {
      wstart[k] = w;
      w += (nin[k] + 63) / 64;
    }
    wstart[ns] = w;
  }
}

struct SearchView {
  int nq;
  const int *perm;  // processing order (query index of thread t), or null
  const double *xs;
  const int *xtag;
  int *donorId;
  const WalkNode *nodes;
  const int *rcell;  // the cell of each walk record
  GridFrame g;  // the tree's grid (the walk records' bounds are on it)
  double ext[6];
  const int *conn;
  const double *x;
  const double *cellRes;
  int *tieband = nullptr;  // counter of tie-band queries (resident trees), or null
  // the tree's bin grid (k_find), or boff null: bin b's list is
  // bent[boff[b] .. boff[b + 1]), 2^(16 - bshift) bins per axis
  const int *boff = nullptr;
  const uint4 *bent = nullptr;
  int bshift = 16;
};

// Several searches (one per block, each with its own tree and query set)
// run by one launch of each kernel.  Queries are numbered globally: search s
// owns [qbase[s], qbase[s+1]).  Within a search, position t (processing
// order) holds query perm[t]; the first nin[s] positions are the queries
// that search the tree (k_morton_sets), in waves of 64: search s's walking
// waves are [wstart[s], wstart[s+1]) of the enumeration the listing kernels
// take.  Candidates and the per-query outcome words carry the global position
// qbase[s] + t, and qsid maps a global position to its search.
struct SearchSet {
  const SearchView *sv;
  const int *qbase, *nin, *wstart;
  int *qsid;
  int ns, nqt;
};

// Pointers read from the table are generic; the casts tell the compiler
// that they address global memory (global rather than flat loads), and that
// the table and the tree records are not written while a search runs, so
// wave-uniform reads of them are scalar loads.
template <class T>
__device__ __forceinline__ T *as_global(T *p)
{
  return (T *)(__attribute__((address_space(1))) T *)p;
}
template <class T>
using const_ptr = const __attribute__((address_space(4))) T *;

__device__ __forceinline__ SearchView view_of(const SearchSet &T, int sid)
{
  const_ptr<SearchView> v = (const_ptr<SearchView>)T.sv + sid;
  SearchView S;
  S.nq = v->nq;
  S.perm = as_global(v->perm);
  S.xs = as_global(v->xs);
  S.xtag = as_global(v->xtag);
  S.donorId = as_global(v->donorId);
  S.nodes = as_global(v->nodes);
  S.rcell = as_global(v->rcell);
  for (int d = 0; d < 3; d++) {
    S.g.lo[d] = v->g.lo[d];
    S.g.inv[d] = v->g.inv[d];
  }
  for (int d = 0; d < 6; d++) S.ext[d] = v->ext[d];
  S.conn = as_global(v->conn);
  S.x = as_global(v->x);
  S.cellRes = as_global(v->cellRes);
  S.tieband = v->tieband;
  S.boff = as_global(v->boff);
  S.bent = as_global(v->bent);
  S.bshift = v->bshift;
  return S;
}

// the search table entry of search sid as every lane of the wave needs it:
// by wave-uniform (scalar) loads when all lanes have the same search (the
// usual case), else lane by lane
__device__ __forceinline__ SearchView view_of_lanes(const SearchSet &T, int sid)
{
  const int s0 = __builtin_amdgcn_readfirstlane(sid);
  SearchView S = view_of(T, s0);
  if (!__all(sid == s0)) S = view_of(T, sid);
  return S;
}

// the search owning walking wave wv (last s with wstart[s] <= wv)
__device__ __forceinline__ int wave_search(const SearchSet &T, int wv)
{
  int lo = 0, hi = T.ns - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (T.wstart[m] <= wv) lo = m;
    else hi = m - 1;
  }
  return lo;
}

// The grid test of a packed box (grid_box layout) for a point at grid
// coordinates q, without short-circuit branches: passes whenever the
// point's FP64 test against the padded box passes.
__device__ __forceinline__ bool in_grid(const unsigned q[3], const unsigned w[3])
{
  return (q[0] >= (w[0] & 0xffffu)) & (q[1] >= (w[0] >> 16)) & (q[2] >= (w[1] & 0xffffu)) &
         (q[0] <= (w[1] >> 16)) & (q[1] <= (w[2] & 0xffffu)) & (q[2] <= (w[2] >> 16));
}

constexpr int WAVE = 64;
constexpr int STACK = 32;
static_assert(STACK <= WAVE, "k_collect keeps its stack one entry per lane");

// Candidates of one search launch: a list of entries j = (rank | nv << 28,
// global position of the query, test outcome), appended in any order.  Per
// query, by its position (processing order, so neighbouring candidates'
// words are neighbours), the
// tests fold into best (the lowest rank of a clean hit; INT_MAX: none), then,
// for the queries without one, into last ((rank << 2) | outcome of the
// highest-ranked candidate inside its element box; -1: none), and with a
// tie-band counter into acc (candidates accepting the point, plus 2^20 per
// near-zero-weight acceptance).
// The list is cut by cell type (bucket t = 0..3: 4, 5, 6, 8 vertices), so each
// containment-test kernel reads its own type's entries with no separate pass
// that sorts them by type, and within a type into NREG regions of rcap[t]
// slots: region (t, r) holds slots tbase[t] + r * rcap[t] .. + rcap[t] - 1 and
// is appended to through its own counter used[(t * NREG + r) * USED_STRIDE] (a
// cache line each).  Workgroup b of a listing launch appends to regions b %
// NREG, so the appends of all waves do not queue on one atomic counter.  A
// region's slots past its count are unused; entries appended past rcap[t] are
// dropped, and the host sees a count above it, grows that type's regions and
// lists again before any test runs.
constexpr int NTYPE = 4, NREG = 32, USED_STRIDE = 32, NRT = NTYPE * NREG;
// workgroups of the launches over the whole list
constexpr unsigned LAST_GRID = 4096, STATS_GRID = 1024;
struct CandOut {
  int *cell, *query;
  unsigned char *status;  // test outcome per slot (0-3)
  int *used;  // per region: entries appended so far
  int rcap[NTYPE], tbase[NTYPE];  // per type: capacity of each region, first slot
  int *best, *last, *acc;  // per global position (acc: null without a tie-band counter)
};

// Candidate tests by cell type (bucket 0..3 = 4, 5, 6, 8 vertices), so each
// test kernel runs one type's code with that type's register budget; the
// outcome goes to the candidate's own slot, so the order is free.
__host__ __device__ __forceinline__ int type_bucket(int nv) { return nv == 4 ? 0 : (nv == 5 ? 1 : (nv == 6 ? 2 : 3)); }

// the regions' capacities and first slots for per-type list sizes (slots)
struct CandLayout {
  int rcap[NTYPE], tbase[NTYPE];
  size_t slots;
};
inline CandLayout cand_layout(const size_t tcap[NTYPE])
{
  CandLayout L;
  size_t at = 0;
  for (int t = 0; t < NTYPE; t++) {
    const size_t r = std::max<size_t>((tcap[t] + NREG - 1) / NREG, 1);
    L.rcap[t] = (int)std::min<size_t>(r, 0x7fffffff / NRT);
    L.tbase[t] = (int)at;
    at += (size_t)NREG * L.rcap[t];
  }
  L.slots = at;
  return L;
}

// entries held by the regions of a list (counts as the device left them)
inline long long cand_held(const int *used, const CandLayout &L)
{
  long long n = 0;
  for (int t = 0; t < NTYPE; t++)
    for (int r = 0; r < NREG; r++) n += std::min(used[(t * NREG + r) * USED_STRIDE], L.rcap[t]);
  return n;
}

// the entries region (t, r) holds, and its first slot
__device__ __forceinline__ int region_held(const CandOut &O, int t, int r)
{
  return min(O.used[(t * NREG + r) * USED_STRIDE], O.rcap[t]);
}
__device__ __forceinline__ int region_first(const CandOut &O, int t, int r) { return O.tbase[t] + r * O.rcap[t]; }

// A launch over the regions of types t0 .. t0 + nt - 1 (gridDim.x a multiple
// of nt * NREG) gives region q = b % (nt * NREG) to workgroup b and the
// others of its residue: the workgroup's first slot (its first thread's),
// the end of the region's entries, and the stride of its loop
struct RegionLoop {
  int first, end, stride;
};
__device__ __forceinline__ RegionLoop region_loop(const CandOut &O, int t0, int nt)
{
  const int nr = nt * NREG, q = (int)blockIdx.x % nr, lb = (int)blockIdx.x / nr;
  const int t = t0 + q / NREG, r = q % NREG;
  const int nb = (int)gridDim.x / nr, bs = (int)blockDim.x;
  const int f = region_first(O, t, r);
  return RegionLoop{f + lb * bs, f + region_held(O, t, r), nb * bs};
}

// A launch over every region's held entries, evenly however the types
// divide them: the entries are cut into chunks of blockDim.x, numbered
// region after region (q = t * NREG + r), and workgroup b takes chunks b,
// b + gridDim.x, ...  chunk_prefix fills pre[q] = chunks before region q
// (pre[NRT]: all) in the workgroup's LDS (blockDim.x >= 64); chunk_slot
// gives the slot of thread threadIdx.x in chunk c, or -1 past its region's
// entries.
__device__ __forceinline__ void chunk_prefix(const CandOut &O, int *pre)
{
  static_assert(NRT == 2 * WAVE, "two regions per lane of the first wave");
  if (threadIdx.x < WAVE) {
    const int l = threadIdx.x, bs = (int)blockDim.x;
    const int q0 = 2 * l, q1 = 2 * l + 1;
    const int c0 = (region_held(O, q0 / NREG, q0 % NREG) + bs - 1) / bs;
    const int c1 = (region_held(O, q1 / NREG, q1 % NREG) + bs - 1) / bs;
    int inc = c0 + c1;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
      const int v = __shfl_up(inc, o);
      if (l >= o) inc += v;
    }
    pre[q0] = inc - c0 - c1;
    pre[q1] = inc - c1;
    if (l == WAVE - 1) pre[NRT] = inc;
  }
  __syncthreads();
}
__device__ __forceinline__ int chunk_slot(const CandOut &O, const int *pre, int c)
{
  int lo = 0, hi = NRT - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (pre[m] <= c) lo = m;
    else hi = m - 1;
  }
  const int t = lo / NREG, r = lo % NREG;
  const int k = (c - pre[lo]) * (int)blockDim.x + (int)threadIdx.x;
  return k < region_held(O, t, r) ? region_first(O, t, r) + k : -1;
}

// the region (t, r) of slot j, and whether it holds a listed candidate
__device__ __forceinline__ bool slot_live(const CandOut &O, int j, int &t, int &r)
{
  t = 0;
  while (t + 1 < NTYPE && O.tbase[t + 1] <= j) t++;
  r = (j - O.tbase[t]) / O.rcap[t];
  return j - region_first(O, t, r) < O.used[(t * NREG + r) * USED_STRIDE];
}

// ADT::searchADT / searchIntersections (searchADTrecursion.C:26-101).  The
// reference walks the tree in preorder (a node's own element, then its left
// and right subtrees whose boxes hold the point), tests every element whose
// box holds the point, lets each test overwrite the result and stops at the
// first clean hit.  Box tests have no side effects, and the ADT numbers its
// nodes in preorder (buildADTrecursion.c:55-58 counts a node before its
// subtrees), so the reference tests exactly the candidates -- the elements
// whose padded box holds the point -- in increasing node id.  An element
// whose padded box holds the point lies inside every ancestor's padded
// subtree box (subtree boxes are min/max folds that include it), so the walk
// reaches every candidate.  The reference's result is therefore a function of
// the candidate set and each candidate's node id ("rank") and test outcome:
// the cell of the lowest-ranked clean hit, and when there is none, the
// outcome of the highest-ranked candidate (its cell when accepted with a
// near-zero weight in a BIGVALUE cell, else -1).  So the search runs as
//   k_collect / k_find  list (rank | nv << 28, query) pairs, in any order,
//   k_eval              every candidate's containment test, one thread each,
//                       folded into its query's best / last words (atomics),
//   k_select            the rule above, per query, from those two words,
// which keeps long candidate lists and slow Newton solves from serialising
// inside one lane.
//
// Both listing kernels test the coarse 16-bit grid boxes of the walk records
// (WalkNode), which pass whenever the reference's padded FP64 test passes,
// so they list a superset of the candidates.  k_eval makes the exact FP64
// element-box test of each listed candidate (the element box is the min/max
// of the cell's vertices, search.C:186-214) and marks the ones that fail
// OUTSIDE, which take no part in the selection.
//
// A wave gathers its lanes' candidates in an LDS buffer of WBUF pairs and
// appends the buffer to the flat list (one atomic) when it is nearly full and
// at the end.  The copy places the pairs sorted by lane, in step order within
// a lane (each pair keeps its index among its lane's pairs, and the lanes'
// counts give their offsets), so a query's candidates mostly lie together in
// the list and its containment tests run side by side (one point,
// neighbouring cells).
// A tree built for one search gets bins when it has at least one query per
// BIN_ELEMENTS_PER_QUERY elements: with fewer queries, building the bins
// costs more than they save the search (measured on a 3.9 M-element tree
// with one query per element: bins gain 0.4 ms of 37 ms; 65 trees of 1e4 to
// 7e6 elements: no difference either way).
constexpr long long BIN_ELEMENTS_PER_QUERY = 2;

constexpr int WBUF = 512;
constexpr unsigned FIND_GRID = 32768;  // workgroups of k_find (each takes work items in turn)
// entries of its own list segment each lane of k_find scans itself; the
// rest of a longer segment is scanned by the whole wave
constexpr int FIND_OWN = 24;
// workgroups of k_find_cost (each wave takes walking waves in turn)
constexpr long long FIND_COST_GRID = 2048;
struct WaveBuf {
  int *val, *qid;  // LDS, WBUF entries each: value (rank | nv << 28), position
  int lane;
  int n = 0;  // wave-uniform: pairs held
  // each lane adds at most one pair; called in wave-uniform control flow
  __device__ __forceinline__ void emit(bool hit, int v, int g, const CandOut &O)
  {
    const unsigned long long m = __ballot(hit);
    if (!m) return;
    if (hit) {
      const int k = n + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      val[k] = v;
      qid[k] = g;
    }
    n += __popcll(m);
    if (n > WBUF - WAVE) flush(O);
  }
  // Appends the held pairs to the workgroup's region of each cell type, in
  // the order held: one atomic per type present.  Called in wave-uniform
  // control flow; one wave per workgroup.
  __device__ __forceinline__ void flush(const CandOut &O)
  {
    if (n == 0) return;
    __syncthreads();  // the lanes' pair writes before other lanes read them
    const int reg = blockIdx.x & (NREG - 1);
    const unsigned long long lt = (1ull << lane) - 1;
    int cnt[NTYPE] = {0, 0, 0, 0};
    for (int j0 = 0; j0 < n; j0 += WAVE) {
      const int j = j0 + lane;
      const int t = j < n ? type_bucket((unsigned)val[j] >> 28) : -1;
#pragma unroll
      for (int u = 0; u < NTYPE; u++) cnt[u] += __popcll(__ballot(t == u));
    }
    int base[NTYPE];
#pragma unroll
    for (int u = 0; u < NTYPE; u++) {
      int b = 0;
      if (lane == 0 && cnt[u]) b = atomicAdd(O.used + (u * NREG + reg) * USED_STRIDE, cnt[u]);
      base[u] = __shfl(b, 0);
    }
    for (int j0 = 0; j0 < n; j0 += WAVE) {
      const int j = j0 + lane;
      const int t = j < n ? type_bucket((unsigned)val[j] >> 28) : -1;
#pragma unroll
      for (int u = 0; u < NTYPE; u++) {
        const unsigned long long m = __ballot(t == u);
        if (t == u) {
          const int d = base[u] + __popcll(m & lt);
          if (d < O.rcap[u]) {
            const int k = O.tbase[u] + reg * O.rcap[u] + d;
            O.cell[k] = val[j];
            O.query[k] = qid[j];
          }
        }
        base[u] += __popcll(m);
      }
    }
    __syncthreads();  // the copy before the buffer is written again
    n = 0;
  }
};

// the listed value of a walk record: its node id (the rank) with the vertex count
__device__ __forceinline__ int rank_nv(int node, int cellnv) { return node | (cellnv & (int)0xf0000000); }

// k_collect: the 64 queries of a wave (consecutive in Morton order, so
// spatially close) walk the ADT together: the wave visits a node when any
// lane's own walk would, carrying the mask of those lanes, so every lane sees
// exactly its own preorder walk while node records are wave-uniform scalar
// loads.
__global__ void __launch_bounds__(WAVE, 8) k_collect(SearchSet T, CandOut O, int skip_binned)
{
  __shared__ int bval[WBUF], bqid[WBUF];
  const int lane = threadIdx.x;
  // workgroup wv walks walking wave wv; the grid is sized for every query,
  // so the workgroups past the walking waves leave at once
  const int wv = blockIdx.x;
  if (wv >= T.wstart[T.ns]) return;
  const int sid = __builtin_amdgcn_readfirstlane(wave_search(T, wv));
  const SearchView S = view_of(T, sid);
  if (skip_binned && S.boff) return;  // listed from the tree's bins by k_find
  const const_ptr<WalkNode> nodes = (const_ptr<WalkNode>)S.nodes;
  const int qb = T.qbase[sid];
  const int t = (wv - T.wstart[sid]) * WAVE + lane;
  const int nin = T.nin[sid];
  const int i = t < nin ? S.perm[t] : -1;
  bool inroot = false;
  unsigned q[3] = {0u, 0u, 0u};  // the point on the tree's grid
  if (i >= 0) {
    double p[3];
    load_xyz(S.xs, i, p);
    inroot = true;  // the root box test, FP64 as searchADTrecursion.C:38-43 makes it
    for (int d = 0; d < 3; d++) inroot = inroot && (p[d] >= S.ext[2 * d] - TOL);
    for (int d = 0; d < 3; d++) inroot = inroot && (p[d] <= S.ext[2 * d + 1] + TOL);
    for (int d = 0; d < 3; d++) q[d] = grid_of(p[d], S.g.lo[d], S.g.inv[d]);
  }
  const int gp = qb + t;  // the query's global position (its outcome words' index)
  WaveBuf buf{bval, bqid, lane};
  // pending right subtrees: entry k (node, lane mask) held by lane k, so a
  // push is a select and a pop a readlane; the left child, when any lane
  // enters it, is visited next without the stack (preorder)
  int vnode = 0;
  unsigned vmlo = 0, vmhi = 0;
  int sp = 0;
  unsigned long long m = __ballot(inroot);
  bool more = m != 0ull;
  int node = 0;
  while (more) {
    // the whole 48-byte record in one batch of scalar loads, then the tests
    WalkNode N;
    {
      const_ptr<WalkNode> r = nodes + node;
      for (int d = 0; d < 3; d++) {
        N.e[d] = r->e[d];
        N.l[d] = r->l[d];
        N.r[d] = r->r[d];
      }
      N.cellnv = r->cellnv;
      N.left = r->left;
      N.right = r->right;
    }
    const bool act = (m >> lane) & 1ull;
    buf.emit(act & in_grid(q, N.e), rank_nv(node, N.cellnv), gp, O);
    const int r = N.right, l = N.left;
    const unsigned long long mr = (r > -1) ? __ballot(act & in_grid(q, N.r)) : 0ull;
    const unsigned long long ml = (l > -1) ? __ballot(act & in_grid(q, N.l)) : 0ull;
    if (mr) {
      if (lane == sp) {
        vnode = r;
        vmlo = (unsigned)mr;
        vmhi = (unsigned)(mr >> 32);
      }
      sp++;
    }
    if (ml) {
      node = l;
      m = ml;
    } else if (sp > 0) {
      sp--;
      node = __builtin_amdgcn_readlane(vnode, sp);
      m = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)vmhi, sp) << 32) |
          (unsigned)__builtin_amdgcn_readlane((int)vmlo, sp);
    } else {
      more = false;
    }
  }
  buf.flush(O);
}

// k_find lists the same candidate sets from the tree's bin grid instead of
// the walk.  The bins split the tree's 16-bit grid into 2^(16 - bshift)
// cells per axis; bin b lists (packed grid box, rank | nv << 28) of every
// element whose grid box overlaps it, in any order.  A point lies in exactly
// one bin, and every grid box holding the point overlaps that bin, so
// scanning the bin's list with the grid test lists the walk's candidates.
//
// The lanes of a walking wave scan their own bins' lists (independent loads,
// four at a time).  Lists differ widely in length (where many boxes overlap,
// as at the axis of a lat-long mesh, a list holds hundreds of elements), so a
// wave's scan is cut into segments of seg entries per lane, and each segment
// is a work item of its own: k_find_cost counts each walking wave's segments,
// a scan numbers the items, and k_find's workgroups take the items in turn.
__device__ __forceinline__ void find_lane(const SearchView &S, int i, unsigned q[3], int &start, int &len)
{
  start = 0;
  len = 0;
  if (i < 0) return;
  double p[3];
  load_xyz(S.xs, i, p);
  bool inroot = true;  // the root box test, FP64 as searchADTrecursion.C:38-43 makes it
  for (int d = 0; d < 3; d++) inroot = inroot && (p[d] >= S.ext[2 * d] - TOL);
  for (int d = 0; d < 3; d++) inroot = inroot && (p[d] <= S.ext[2 * d + 1] + TOL);
  for (int d = 0; d < 3; d++) q[d] = grid_of(p[d], S.g.lo[d], S.g.inv[d]);
  if (!inroot || !S.boff) return;
  const int sh = S.bshift, R = 65536 >> sh;
  const int bin = (int)(((q[2] >> sh) * (unsigned)R + (q[1] >> sh)) * (unsigned)R + (q[0] >> sh));
  start = S.boff[bin];
  len = S.boff[bin + 1] - start;
}

// nseg[w] = segments of walking wave w (at least one), 0 past the walking
// waves; and each walking query's outcome words start out empty (best,
// last and the tie-band count; no other position is read)
__global__ void __launch_bounds__(256) k_find_cost(SearchSet T, int nwave, int seg, int *__restrict__ nseg,
                                                   int *__restrict__ best, int *__restrict__ last,
                                                   int *__restrict__ acc)
{
  // nseg is cleared beforehand, so only the walking waves are visited, each
  // wave of the grid taking them in turn
  const int lane = threadIdx.x & (WAVE - 1), nwalk = min(T.wstart[T.ns], nwave);
  const int wstep = (int)(gridDim.x * (blockDim.x / WAVE));
  for (int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE)); w < nwalk;
       w += wstep) {
    const int sid = __builtin_amdgcn_readfirstlane(wave_search(T, w));
    const SearchView S = view_of(T, sid);
    const int t = (w - T.wstart[sid]) * WAVE + lane;
    if (t < T.nin[sid]) {
      const int gp = T.qbase[sid] + t;
      best[gp] = 0x7fffffff;
      last[gp] = -1;
      if (acc) acc[gp] = 0;
    }
    if (!S.boff) continue;  // a tree without bins: its waves walk (k_collect)
    const int i = t < T.nin[sid] ? S.perm[t] : -1;
    unsigned q[3] = {0u, 0u, 0u};
    int start, len;
    find_lane(S, i, q, start, len);
    for (int o = 32; o > 0; o >>= 1) len = max(len, __shfl_xor(len, o));
    if (lane == 0) nseg[w] = len > seg ? (len + seg - 1) / seg : 1;
  }
}

__global__ void __launch_bounds__(WAVE, 16) k_find(SearchSet T, CandOut O, int nwave, int seg,
                                                   const int *__restrict__ segstart)
{
  __shared__ int bval[WBUF], bqid[WBUF];
  const int lane = threadIdx.x;
  WaveBuf buf{bval, bqid, lane};
  const int items = segstart[nwave];
  for (int it = blockIdx.x; it < items; it += gridDim.x) {
    // the walking wave of item it (last w with segstart[w] <= it): most
    // waves are one item, so the search gallops out from the wave an even
    // spread would give (segstart[lo] <= it < segstart[hi], segstart[nwave]
    // = items) and bisects the bracket
    int lo, hi;
    {
      const int g = (int)min((long long)nwave - 1, (long long)it * nwave / items);
      int step = 1;
      if (segstart[g] <= it) {
        lo = g;
        hi = min(nwave, g + 1);
        while (hi < nwave && segstart[hi] <= it) {
          lo = hi;
          step <<= 1;
          hi = min(nwave, lo + step);
        }
      } else {
        hi = g;
        lo = max(0, g - 1);
        while (lo > 0 && segstart[lo] > it) {
          hi = lo;
          step <<= 1;
          lo = max(0, hi - step);
        }
      }
    }
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (segstart[mid] <= it) lo = mid;
      else hi = mid;
    }
    const int w = __builtin_amdgcn_readfirstlane(lo);
    const int k0 = (it - segstart[w]) * seg;
    const int sid = __builtin_amdgcn_readfirstlane(wave_search(T, w));
    const SearchView S = view_of(T, sid);
    const int t = (w - T.wstart[sid]) * WAVE + lane;
    const int i = t < T.nin[sid] ? S.perm[t] : -1;
    unsigned q[3] = {0u, 0u, 0u};
    int start, len;
    find_lane(S, i, q, start, len);
    const int gp = T.qbase[sid] + t;
    const int end = min(len, k0 + seg);
    // each lane scans the first FIND_OWN entries of its own list segment;
    // the rest of the longer ones, one list at a time, by the whole wave
    const int mid = min(end, k0 + FIND_OWN);
    int lmax = mid;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lmax = max(lmax, __shfl_xor(lmax, o));
    const uint4 *ent = S.bent;
    for (int k = k0; k < lmax; k += 4) {
      uint4 e[4];
#pragma unroll
      for (int u = 0; u < 4; u++) e[u] = (k + u < mid) ? ent[start + k + u] : make_uint4(0xffffu, 0u, 0u, 0u);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const unsigned wb[3] = {e[u].x, e[u].y, e[u].z};
        buf.emit((k + u < mid) && in_grid(q, wb), (int)e[u].w, gp, O);
      }
    }
    for (unsigned long long rest = __ballot(end > mid); rest; rest &= rest - 1) {
      const int L = __builtin_ctzll(rest);
      const int st = __shfl(start, L), en = __shfl(end, L), g = __shfl(gp, L), a0 = __shfl(mid, L);
      const unsigned qq[3] = {(unsigned)__shfl((int)q[0], L), (unsigned)__shfl((int)q[1], L),
                              (unsigned)__shfl((int)q[2], L)};
      for (int b = a0; b < en; b += WAVE) {
        const int k = b + lane;
        const uint4 e = k < en ? ent[st + k] : make_uint4(0xffffu, 0u, 0u, 0u);
        const unsigned wb[3] = {e.x, e.y, e.z};
        buf.emit(k < en && in_grid(qq, wb), (int)e.w, g, O);
      }
    }
  }
  buf.flush(O);
}

// Outcome of a containment test from the weights (checkContainment.C:62-85):
// 0 reject, 1 accepted with a near-zero weight in a BIGVALUE cell, 2 clean;
// OUTSIDE: the point fails the element's padded box test, so the reference
// never tests this cell (listed only by the coarse grid walk)
constexpr int OUTSIDE = 3;
template <int NV>
__device__ __forceinline__ int test_status(const double frac[8], bool big)
{
  int flag = 0;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    if ((frac[m] + TOL) * (frac[m] - 1.0 - TOL) > 0) return 0;
    if (fabs(frac[m]) < TOL && big) flag = 1;
  }
  return flag ? 1 : 2;
}

// a tested candidate's outcome: its status, and a clean hit into its query's
// best word (and the tie-band count); last is formed afterwards, only for
// the queries without a clean hit (k_last)
__device__ __forceinline__ void record_outcome(const CandOut &O, int j, int st)
{
  O.status[j] = (unsigned char)st;
  if (st == OUTSIDE || st == 0) return;
  const int g = O.query[j];
  if (st == 2) atomicMin(&O.best[g], O.cell[j] & 0x0fffffff);
  if (O.acc) atomicAdd(&O.acc[g], st == 1 ? (1 << 20) + 1 : 1);
}

// last = the highest (rank << 2 | status) among a query's candidates inside
// their element box, for the queries without a clean hit (the only ones
// k_select reads it for)
__global__ void __launch_bounds__(256) k_last(CandOut O)
{
  __shared__ int pre[NRT + 1];
  chunk_prefix(O, pre);
  for (int c = blockIdx.x; c < pre[NRT]; c += gridDim.x) {
    const int j = chunk_slot(O, pre, c);
    if (j < 0) continue;
    const int st = O.status[j];
    if (st == OUTSIDE) continue;
    const int g = O.query[j];
    if (O.best[g] != 0x7fffffff) continue;
    atomicMax(&O.last[g], ((O.cell[j] & 0x0fffffff) << 2) | st);
  }
}

// the cell of a listed candidate (rank | nv << 28): its walk record's cell
__device__ __forceinline__ int cell_of(const SearchView &S, int ranknv) { return S.rcell[ranknv & 0x0fffffff]; }

template <int NV>
__device__ __forceinline__ void cell_vertices(const SearchView &S, int c, double xv[8][3])
{
  const int *cv = S.conn + 8 * (size_t)c;
#pragma unroll
  for (int m = 0; m < NV; m++) load_xyz(S.x, cv[m], xv[m]);
}

// Long Newton solves (a few per cent of candidates run to the 500-iteration
// cap) are continued in a separate pass so they do not hold up the lanes of
// the short ones: the first pass stops every solve after NEWTON_SPLIT
// iterations and queues the unfinished ones with their state.
constexpr int NEWTON_SPLIT = 16;

struct LongQueue {
  int *j;        // candidate slot, per type region [bstart, bstart + cap)
  double4 *st;   // (u, v, w, iterations so far)
  int *n;        // per type count
  int cap;       // per type capacity
};

// containment test of every listed candidate of one cell type
template <int NV, int WAVES>
__global__ void __launch_bounds__(256, WAVES) k_eval(SearchSet T, CandOut O, LongQueue L)
{
  constexpr int B = NV == 4 ? 0 : (NV == 5 ? 1 : (NV == 6 ? 2 : 3));
  const RegionLoop R = region_loop(O, B, 1);
  for (int j = R.first + threadIdx.x; j < R.end; j += R.stride) {
    const int gp = O.query[j], sid = T.ns == 1 ? 0 : T.qsid[gp];
    const SearchView S = view_of_lanes(T, sid);
    const int c = cell_of(S, O.cell[j]);
    double p[3], xv[8][3], frac[8];
    load_xyz(S.xs, S.perm[gp - T.qbase[sid]], p);
    cell_vertices<NV>(S, c, xv);
    // the element box test of searchADTrecursion.C:66-71 on the box of
    // search.C:199-214 (min/max of the vertices; the sign of a zero bound
    // does not change x - TOL)
    bool inb = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double lo = xv[0][d], hi = xv[0][d];
#pragma unroll
      for (int mm = 1; mm < NV; mm++) {
        lo = dmin(lo, xv[mm][d]);
        hi = dmax(hi, xv[mm][d]);
      }
      inb = inb & (p[d] >= lo - TOL) & (p[d] <= hi + TOL);
    }
    if (!inb) {
      record_outcome(O, j, OUTSIDE);
      continue;
    }
    if (NV == 4) {
      nodal_weights(xv, p, frac, 4);
    } else {
      double f[8][3], u = 0.5, v = 0.5, w = 0.5;
      int it = 0, ok = 1;
      iso_coeffs(xv, p, f, NV);
      if (!newton3_steps(f, u, v, w, it, NEWTON_SPLIT, ok)) {
        const int q = atomicAdd(&L.n[B], 1);
        if (q < L.cap) {
          L.j[B * L.cap + q] = j;
          L.st[B * L.cap + q] = make_double4(u, v, w, (double)it);
          continue;
        }
        newton3_steps(f, u, v, w, it, NEWTON_MAX, ok);  // queue full: finish here
      }
      newton3_finish(it, ok, u, v, w);
      iso_weights(u, v, w, frac, NV);
    }
    record_outcome(O, j, test_status<NV>(frac, S.cellRes[c] == BIGVALUE));
  }
}

// how many listed candidates ended with each status (0 reject, 1 accepted
// with a near-zero weight, 2 clean, 3 outside the element box); a diagnostic
__global__ void __launch_bounds__(256) k_status_hist(CandOut O, int *__restrict__ hist)
{
  __shared__ int loc[4];
  if (threadIdx.x < 4) loc[threadIdx.x] = 0;
  __syncthreads();
  __shared__ int pre[NRT + 1];
  chunk_prefix(O, pre);
  for (int c = blockIdx.x; c < pre[NRT]; c += gridDim.x) {
    const int j = chunk_slot(O, pre, c);
    if (j < 0) continue;
    const int st = O.status[j];
    if (st >= 0 && st < 4) atomicAdd(&loc[st], 1);
  }
  __syncthreads();
  if (threadIdx.x < 4 && loc[threadIdx.x]) atomicAdd(&hist[threadIdx.x], loc[threadIdx.x]);
}

// candidates ranked after their query's lowest-ranked clean hit (the
// reference's walk stops at that hit and never tests them); a diagnostic
__global__ void k_after_hit(CandOut O, unsigned long long *__restrict__ out)
{
  __shared__ int pre[NRT + 1];
  chunk_prefix(O, pre);
  unsigned long long v = 0;
  for (int c = blockIdx.x; c < pre[NRT]; c += gridDim.x) {
    const int j = chunk_slot(O, pre, c);
    if (j >= 0) v += (O.cell[j] & 0x0fffffff) > O.best[O.query[j]];
  }
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}

// the queued solves, continued from their saved state to the end
template <int NV>
__global__ void __launch_bounds__(256, 4) k_eval_long(SearchSet T, CandOut O, LongQueue L)
{
  constexpr int B = NV == 5 ? 1 : (NV == 6 ? 2 : 3);
  const int n = min(L.n[B], L.cap);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int j = L.j[B * L.cap + k];
    const double4 st = L.st[B * L.cap + k];
    const int gp = O.query[j], sid = T.ns == 1 ? 0 : T.qsid[gp];
    const SearchView S = view_of_lanes(T, sid);
    const int c = cell_of(S, O.cell[j]);
    double p[3], xv[8][3], frac[8], f[8][3];
    load_xyz(S.xs, S.perm[gp - T.qbase[sid]], p);
    cell_vertices<NV>(S, c, xv);
    iso_coeffs(xv, p, f, NV);
    double u = st.x, v = st.y, w = st.z;
    int it = (int)st.w, ok = 1;
    newton3_steps(f, u, v, w, it, NEWTON_MAX, ok);
    newton3_finish(it, ok, u, v, w);
    iso_weights(u, v, w, frac, NV);
    record_outcome(O, j, test_status<NV>(frac, S.cellRes[c] == BIGVALUE));
  }
}

// the reference's result for the queries that walked (the others' donors are
// set elsewhere): the lowest-ranked clean hit, else the highest-ranked
// candidate's outcome; with a tie-band counter the queries whose donor would
// depend on the walk order are counted (more than one candidate accepts the
// point, or one accepts it with a near-zero weight in a BIGVALUE cell)
__global__ void __launch_bounds__(256) k_select(SearchSet T, CandOut O)
{
  const int lane = threadIdx.x & (WAVE - 1), nwalk = T.wstart[T.ns];
  const int step = gridDim.x * (blockDim.x / WAVE);
  for (int w = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE; w < nwalk; w += step) {
    const int sid = __builtin_amdgcn_readfirstlane(wave_search(T, w));
    const int t = (w - T.wstart[sid]) * WAVE + lane;
    if (t >= T.nin[sid]) continue;
    const SearchView S = view_of(T, sid);
    const int gp = T.qbase[sid] + t;
    const int i = S.perm[t];
    const int best = O.best[gp], last = O.last[gp];
    int res = -1;
    if (best != 0x7fffffff) res = S.rcell[best];
    else if (last >= 0 && (last & 3) == 1) res = S.rcell[last >> 2];
    if (S.tieband && O.acc) {
      const int a = O.acc[gp];
      if ((a & 0xfffff) > 1 || (a >> 20) > 0) atomicAdd(S.tieband, 1);
    }
    S.donorId[i] = res;
  }
}

// ---- bin grids of walk records (k_find's lists) ----
struct BinTree {
  const WalkNode *nodes;
  int nelem, shift;
  long long bbase;  // the tree's first bin in the batch's count / offset arrays
  int nbase;        // the tree's first element in the batch's enumeration
};

__device__ __forceinline__ int bin_tree(int ntree, const BinTree *T, int g)
{
  int lo = 0, hi = ntree - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (T[m].nbase <= g) lo = m;
    else hi = m - 1;
  }
  return lo;
}

// the bin ranges of a packed grid box (grid_box layout) at a shift
__device__ __forceinline__ void box_bins(const unsigned w[3], int sh, int lo[3], int hi[3])
{
  lo[0] = (int)((w[0] & 0xffffu) >> sh);
  lo[1] = (int)((w[0] >> 16) >> sh);
  lo[2] = (int)((w[1] & 0xffffu) >> sh);
  hi[0] = (int)((w[1] >> 16) >> sh);
  hi[1] = (int)((w[2] & 0xffffu) >> sh);
  hi[2] = (int)((w[2] >> 16) >> sh);
}

// entries of each element at its tree's shift
__global__ void k_bin_nb(int ntree, int total, const BinTree *__restrict__ T, int *__restrict__ nbe)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int a = bin_tree(ntree, T, g);
  int lo[3], hi[3];
  box_bins(T[a].nodes[g - T[a].nbase].e, T[a].shift, lo, hi);
  nbe[g] = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}

// entries of each tree: the element prefix at its ends (tot[a] = eoff[end] - eoff[start])
__global__ void k_bin_totals(int ntree, const BinTree *__restrict__ T, const int *__restrict__ eoff,
                             long long *__restrict__ tot)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ntree) return;
  tot[a] = (long long)eoff[T[a].nbase + T[a].nelem] - eoff[T[a].nbase];
}

// one (global bin, element) pair per bin of each element, at the element's prefix
__global__ void k_bin_pairs(int ntree, int total, const BinTree *__restrict__ T, const int *__restrict__ eoff,
                            unsigned *__restrict__ key, int *__restrict__ val)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int a = bin_tree(ntree, T, g);
  const int sh = T[a].shift, R = 65536 >> sh;
  int lo[3], hi[3];
  box_bins(T[a].nodes[g - T[a].nbase].e, sh, lo, hi);
  int o = eoff[g];
  for (int z = lo[2]; z <= hi[2]; z++)
    for (int y = lo[1]; y <= hi[1]; y++)
      for (int x = lo[0]; x <= hi[0]; x++) {
        key[o] = (unsigned)(T[a].bbase + (z * R + y) * R + x);
        val[o] = g;
        o++;
      }
}

// off[b] = the first sorted pair of bin b (its lower bound), off[nbin] = n
__global__ void k_bin_offsets(int nbin, int n, const unsigned *__restrict__ key, int *__restrict__ off)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbin) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (key[m] < (unsigned)b) lo = m + 1;
    else hi = m;
  }
  off[b] = lo;
}

// entry j of the sorted pairs: its element's packed grid box and rank | nv
__global__ void k_bin_entries(int ntree, int n, const BinTree *__restrict__ T, const int *__restrict__ val,
                              uint4 *__restrict__ ent)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int g = val[j];
  const int a = bin_tree(ntree, T, g);
  const int e = g - T[a].nbase;
  const WalkNode &N = T[a].nodes[e];
  ent[j] = make_uint4(N.e[0], N.e[1], N.e[2], (unsigned)rank_nv(e, N.cellnv));
}

// coincident queries take their representative's donor (search.C:251-265)
__global__ void k_copy_dup(int nq, const int *__restrict__ xtag, int *__restrict__ donorId)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int r = i;
  while (xtag[r] != r) r = xtag[r];
  if (r != i) donorId[i] = donorId[r];
}

// k_copy_dup for every set of a batched search (query g of the
// concatenation belongs to the last set k with qb[k] <= g)
__global__ void k_copy_dup_sets(const SearchView *__restrict__ sv, const int *__restrict__ qb, int ns, int nqt)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nqt) return;
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (qb[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const int i = g - qb[lo];
  const int *xtag = sv[lo].xtag;
  int *donorId = sv[lo].donorId;
  int r = i;
  while (xtag[r] != r) r = xtag[r];
  if (r != i) donorId[i] = donorId[r];
}

// ---------------------------------------------------------------------------
// donor selection: exchangeDonors.C:25-301, bookKeeping.C:74-648
// ---------------------------------------------------------------------------
//
// The reference inserts each donor record into a per-receptor linked list
// sorted by donorRes, equal keys in arrival order (linklist.c:62-104).
// Downstream only reads: whether the list is empty, the donor mesh tags, the
// max receptorRes, and the head (the least donorRes, earliest arrival), so
// these are formed here by order-free atomic reductions.  Arrival order is
// (donor block, search index), encoded as soff[donor block] + i.

struct DonorSrc {
  int nsearch, arrival0, dblk, dtag;
  const int *donorId, *isearch, *xtag;
  const double *cellRes, *res_search;
};

struct RecvTables {  // per receptor block, indexed by block id
  int *has[128];
  unsigned long long *rmax[128], *hkey[128], *tmask[128];
  int *harr[128], *hblk[128], *hrem[128];
};

struct HoleSet {
  int nmesh;
  const HoleMapView *hm;
};

// Records sent back to donors (bookKeeping.C:318-340), in receptor order.
struct RecordList {
  int n;
  int *node, *dblk, *remote;
  double *rres;
};

// Per receptor block, the head-donor arrays the records are emitted from;
// nodes of block b are numbered noff[b] .. noff[b+1]-1 in one flag array.
struct EmitTables {
  int nbk;
  const int *noff;
  const int *hblk[128], *hrem[128];
  const double *nodeRes[128];
  double scale[128];
};

// the records of every receptor block at once, in (receptor block, node)
// order: sel lists the flagged global node numbers in ascending order
__global__ void k_emit_records_all(int n, const int *__restrict__ sel, const EmitTables *__restrict__ T,
                                   RecordList L)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int g = sel[j];
  int lo = 0, hi = T->nbk - 1;  // last block with noff <= g
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (T->noff[m] <= g) lo = m;
    else hi = m - 1;
  }
  const int i = g - T->noff[lo];
  L.node[j] = i;
  L.dblk[j] = T->hblk[lo][i];
  L.remote[j] = T->hrem[lo][i];
  L.rres[j] = (T->scale[lo] > 1.0) ? -T->nodeRes[lo][i] : T->nodeRes[lo][i];
}

__global__ void k_dblk_keys(int n, const int *__restrict__ dblk, unsigned *__restrict__ key, int *__restrict__ idx)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  key[j] = (unsigned)dblk[j];
  idx[j] = j;
}

// first position of each donor block's records in the sorted keys
__global__ void k_dblk_starts(int nbk, int n, const unsigned *__restrict__ sorted, int *__restrict__ start)
{
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbk) return;
  int lo = 0, hi = n;  // first j with sorted[j] >= b
  while (lo < hi) {
    const int m = (lo + hi) >> 1;
    if (sorted[m] < (unsigned)b) lo = m + 1;
    else hi = m;
  }
  start[b] = lo;
}

struct InterpView {
  int n;
  const int *sel;  // record index for recid r
  RecordList L;
  // donor block
  TypeInfo ti;
  const int *conn;
  const double *x;
  const double *nodeRes;
  const int *iblank;  // before any reset of this phase
  const int *donorId, *isearch, *tagsearch;
  const double *xsearch;
  double scale;
  int *tmin;
  unsigned char *cancelv;
  int *inode, *nw, *cancel, *rnode, *rblk, *i2d;
  double *w;
  int *err;
};

// MeshBlock::findInterpData (bookKeeping.C:385-535) for record r, seeing the
// iblank values the reference sees at that point of its sequential walk.
// The cell's vertex count is a template constant, so every weight and vertex
// array index is static (no run-time indexed private arrays).
template <int NV>
__device__ __forceinline__ void interp_record(const InterpView &V, int r, int c, int irec, double rr2, double rr)
{
  const int meshtagrecv = V.tagsearch[irec];
  const int *cv = V.conn + 8 * (size_t)c;
  double xv[8][3], xp[3], frac[8];
  int vert[NV];
  int accept = 1;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const int v = cv[m];
    vert[m] = v;
    const int ib = (V.tmin[v] < r) ? 1 : V.iblank[v];
    if (ib <= 0 && rr2 > 0.0) {
      if (V.nodeRes[v] == BIGVALUE) accept = 0;
      if (abs(ib) == meshtagrecv) accept = 0;
    }
    load_xyz(V.x, v, xv[m]);
  }
  load_xyz(V.xsearch, irec, xp);
  nodal_weights(xv, xp, frac, NV);
  V.i2d[irec] = r;
  int bad = 0;
  const size_t n = V.n;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    V.inode[m * n + r] = vert[m];
    V.w[m * n + r] = frac[m];
    if (frac[m] < -0.2 || frac[m] > 1.2) bad = 1;
  }
  V.inode[NV * n + r] = c;
  V.w[NV * n + r] = 0.0;
  V.nw[r] = NV;
  V.cancel[r] = (accept == 0 && rr != BIGVALUE) ? 1 : 0;
  V.rnode[r] = V.isearch[3 * irec + 1];
  V.rblk[r] = V.isearch[3 * irec + 2];
  if (bad) atomicAdd(V.err, 1);
}

struct CancelTables {
  int *i2d[128];
  int *cancel[128];
};

// cancel lists (bookKeeping.C:542-570): a vertex on a donor's cancel list
// cancels its own donor's interpolation.
__global__ void k_cancel(int nnodes, const unsigned char *__restrict__ cancelv, const int *__restrict__ has,
                         const int *__restrict__ hblk, const int *__restrict__ hrem,
                         const CancelTables *__restrict__ T)
{
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nnodes || !cancelv[v] || !has[v]) return;
  int d = hblk[v];
  int iptr = T->i2d[d][hrem[v]];
  if (iptr > -1) T->cancel[d][iptr] = 1;
}

// MeshBlock::resetCoincident (bookKeeping.C:572-597)
__global__ void k_coinc1(int nq, const int *__restrict__ i2d, const int *__restrict__ xtag,
                         const int *__restrict__ cancel, int *__restrict__ ireset)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int p = i2d[i];
  if (p > -1) atomicMin(&ireset[xtag[i]], cancel[p]);
}

__global__ void k_coinc2(int nq, const int *__restrict__ i2d, const int *__restrict__ xtag,
                         const int *__restrict__ ireset, int *__restrict__ cancel)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  int p = i2d[i];
  if (p > -1 && cancel[p] == 1) cancel[p] = ireset[xtag[i]];
}

struct IblankTables {
  int *ib[128];
};

__global__ void k_set_receptors(int n, const int *__restrict__ cancel, const int *__restrict__ rnode,
                                const int *__restrict__ rblk, const IblankTables *__restrict__ T)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n && !cancel[r]) T->ib[rblk[r]][rnode[r]] = -1;
}

// ---------------------------------------------------------------------------
// The donor exchange for every block in one launch per stage.  A launch's
// threads are cut into segments, one per block, each starting at a multiple
// of BS, so a workgroup lies in one segment and finds it by a wave-uniform
// binary search over the segment table (off = first thread, n = threads in
// use).  Within a stage the blocks touch only their own arrays or write
// other blocks' entries commutatively (stores of one value, atomic min /
// max / or), and every read of another block's data comes from an earlier
// stage, so one launch gives what one launch per block gives.
// ---------------------------------------------------------------------------
template <class S>
__device__ __forceinline__ const S &seg_of(int ns, const S *__restrict__ T)
{
  const int g0 = blockIdx.x * BS;
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T[mid].off <= g0) lo = mid;
    else hi = mid - 1;
  }
  return T[lo];
}

// exchangeSearchData.C:200-255: the receivers' search arrays, every block
// pair's packet in one launch
struct QFillSeg {
  int off, n;
  const int *sel, *invmap;
  const double *xs, *nodeRes;
  int qoff, rblk, rtag;
  double *xsearch, *res_search;
  int *isearch, *tagsearch;
};
__global__ void k_fill_search_all(int ns, const QFillSeg *__restrict__ T)
{
  const QFillSeg &S = seg_of(ns, T);
  const int j = blockIdx.x * BS + threadIdx.x - S.off;
  if (j >= S.n) return;
  int nd = S.invmap[S.sel[j]];
  size_t o = S.qoff + (size_t)j;
  S.isearch[3 * o] = 0;
  S.isearch[3 * o + 1] = nd;
  S.isearch[3 * o + 2] = S.rblk;
  S.tagsearch[o] = S.rtag;
  S.xsearch[3 * o] = S.xs[3 * (size_t)nd];
  S.xsearch[3 * o + 1] = S.xs[3 * (size_t)nd + 1];
  S.xsearch[3 * o + 2] = S.xs[3 * (size_t)nd + 2];
  S.res_search[o] = S.nodeRes[nd];
}


namespace {
template <class S>
int pack_segs(std::vector<S> &v)
{
  long long o = 0;
  for (S &x : v) {
    x.off = (int)o;
    o += (long long)(x.n + BS - 1) / BS * BS;
  }
  if (o > 0x7fffffffLL) throw std::runtime_error("tioga_amd: donor exchange stage too large for one launch");
  return (int)o;
}
}  // namespace


// fills of many arrays: chunk k sets p[start, start + len) to v (width 1, 4
// or 8 bytes)
struct FillChunk {
  void *p;
  unsigned long long v;
  long long start;
  int len, width;
};
// iblank words narrowed to bytes for the write-back (piece k: n words from
// src to out[off ...]); *wide is set when a value does not fit a byte
struct PackSeg {
  int off, n;
  const int *src;
  size_t dst;
};
__global__ void k_pack_i8(int ns, const PackSeg *__restrict__ T, signed char *__restrict__ out, int *__restrict__ wide)
{
  const PackSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  if (i >= S.n) return;
  const int v = S.src[i];
  out[S.dst + i] = (signed char)v;
  if (v < -128 || v > 127) atomicOr(wide, 1);
}

__global__ void k_fill_chunks(const FillChunk *__restrict__ C)
{
  const FillChunk c = C[blockIdx.x];
  for (int i = threadIdx.x; i < c.len; i += blockDim.x) {
    const long long j = c.start + i;
    if (c.width == 1) ((unsigned char *)c.p)[j] = (unsigned char)c.v;
    else if (c.width == 4) ((unsigned *)c.p)[j] = (unsigned)c.v;
    else ((unsigned long long *)c.p)[j] = c.v;
  }
}

struct DonorSeg {
  DonorSrc D;
  int off, n;
};
__global__ void k_donor_pass1_all(int ns, const DonorSeg *__restrict__ T, int nwords, const RecvTables *__restrict__ R)
{
  const DonorSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  const DonorSrc &D = S.D;
  if (i >= S.n || D.donorId[i] < 0) return;
  int rb = D.isearch[3 * i + 2], nd = D.isearch[3 * i + 1];
  R->has[rb][nd] = 1;
  atomicMax(&R->rmax[rb][nd], dkey(D.res_search[D.xtag[i]]));
  atomicMin(&R->hkey[rb][nd], dkey(D.cellRes[D.donorId[i]]));
  int t = D.dtag - BASE;
  atomicOr(&R->tmask[rb][(size_t)nd * nwords + (t >> 6)], 1ull << (t & 63));
}
__global__ void k_donor_pass2_all(int ns, const DonorSeg *__restrict__ T, const RecvTables *__restrict__ R)
{
  const DonorSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  const DonorSrc &D = S.D;
  if (i >= S.n || D.donorId[i] < 0) return;
  int rb = D.isearch[3 * i + 2], nd = D.isearch[3 * i + 1];
  if (dkey(D.cellRes[D.donorId[i]]) == R->hkey[rb][nd]) atomicMin(&R->harr[rb][nd], D.arrival0 + i);
}
__global__ void k_donor_pass3_all(int ns, const DonorSeg *__restrict__ T, const RecvTables *__restrict__ R)
{
  const DonorSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  const DonorSrc &D = S.D;
  if (i >= S.n || D.donorId[i] < 0) return;
  int rb = D.isearch[3 * i + 2], nd = D.isearch[3 * i + 1];
  if (R->harr[rb][nd] == D.arrival0 + i) {
    R->hblk[rb][nd] = D.dblk;
    R->hrem[rb][nd] = i;
  }
}

// processDonors, per receptor block (its nodes, then its cells)
struct ProcSeg {
  int off, n, mytag;
  const double *x;
  const int *has;
  const unsigned long long *rmax, *tmask, *hkey;
  const int *hblk;
  int *iblank;
  double *nodeRes;
  unsigned char *mt, *mt1;
  int *isrec;
};
// MeshBlock::processDonors, first part (bookKeeping.C:159-238): iblank from
// the hole maps of the other meshes (those that did not donate, when the node
// has donors) and nodeRes raised to the largest receptor resolution.
__global__ void k_process1_all(int ns, const ProcSeg *__restrict__ T, HoleSet H, int nwords)
{
  const ProcSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  if (i >= S.n) return;
  int ib = 1;
  double p[3];
  load_xyz(S.x, i, p);
  bool hd = S.has[i] != 0;
  if (hd) S.nodeRes[i] = dmax(S.nodeRes[i], dunkey(S.rmax[i]));
  for (int j = 0; j < H.nmesh; j++) {
    if (j == S.mytag - BASE || !H.hm[j].exist) continue;
    if (hd && ((S.tmask[(size_t)i * nwords + (j >> 6)] >> (j & 63)) & 1ull)) continue;
    if (check_hole(p, H.hm[j])) {
      ib = 0;
      break;
    }
  }
  S.iblank[i] = ib;
  S.mt[i] = (ib == 0);
}
struct FringeSeg {
  int off, n;
  TypeInfo ti;
  const int *conn;
  const unsigned char *mt;
  unsigned char *mt1;
};
// nfringe layers around hole nodes (bookKeeping.C:240-286)
__global__ void k_fringe_all(int ns, const FringeSeg *__restrict__ T)
{
  const FringeSeg &S = seg_of(ns, T);
  const int c = blockIdx.x * BS + threadIdx.x - S.off;
  if (c >= S.n) return;
  int nv = S.ti.nv[type_of(S.ti, c)];
  const int *cv = S.conn + 8 * (size_t)c;
  int any = 0;
  for (int m = 0; m < nv; m++) any |= (S.mt[cv[m]] == 1);
  if (!any) return;
  for (int m = 0; m < nv; m++)
    if (S.mt[cv[m]] != 1) S.mt1[cv[m]] = 1;
}
// bookKeeping.C:288-340: fringe nodes become mandatory receptors; a node with
// donors is a receptor when its best donor is finer than it.
__global__ void k_process2_all(int ns, const ProcSeg *__restrict__ T, const int *__restrict__ blk_tag)
{
  const ProcSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  if (i >= S.n) return;
  if (S.mt1[i] && S.iblank[i]) S.nodeRes[i] = BIGVALUE;
  int r = 0;
  if (S.has[i] && S.iblank[i] != 0 && dunkey(S.hkey[i]) < S.nodeRes[i]) {
    S.iblank[i] = -blk_tag[S.hblk[i]];
    r = 1;
  }
  S.isrec[i] = r;
}

// findInterpData per donor block (records) and its resets (nodes)
struct InterpSeg {
  int off, n;
  InterpView V;
};
// Mandatory receptors (receptorRes == BIGVALUE) reset their donor cell's
// negative-iblank vertices to 1 and put every vertex with iblank <= 0 and
// finite nodeRes on the cancel list (bookKeeping.C:456-468).  A record
// later in the donor's record order sees such a vertex as 1: tmin[v] is the
// first record that resets v.
__global__ void k_interp_mandatory_all(int ns, const InterpSeg *__restrict__ T)
{
  const InterpSeg &S = seg_of(ns, T);
  const int r = blockIdx.x * BS + threadIdx.x - S.off;
  if (r >= S.n) return;
  const InterpView &V = S.V;
  int k = V.sel[r];
  double rr = fabs(V.L.rres[k]);
  if (!(rr == BIGVALUE && V.scale == 1.0)) return;
  int c = V.donorId[V.L.remote[k]];
  int nv = V.ti.nv[type_of(V.ti, c)];
  const int *cv = V.conn + 8 * (size_t)c;
  for (int m = 0; m < nv; m++) {
    int v = cv[m];
    if (V.iblank[v] <= 0 && V.nodeRes[v] != BIGVALUE) {
      if (V.iblank[v] < 0) atomicMin(&V.tmin[v], r);
      V.cancelv[v] = 1;
    }
  }
}
__global__ void k_interp_all(int ns, const InterpSeg *__restrict__ T)
{
  const InterpSeg &S = seg_of(ns, T);
  const int r = blockIdx.x * BS + threadIdx.x - S.off;
  if (r >= S.n) return;
  const InterpView &V = S.V;
  const int k = V.sel[r];
  const double rr2 = V.L.rres[k], rr = fabs(rr2);
  const int irec = V.L.remote[k];
  const int c = V.donorId[irec];
  switch (V.ti.nv[type_of(V.ti, c)]) {
    case 4: interp_record<4>(V, r, c, irec, rr2, rr); break;
    case 5: interp_record<5>(V, r, c, irec, rr2, rr); break;
    case 6: interp_record<6>(V, r, c, irec, rr2, rr); break;
    default: interp_record<8>(V, r, c, irec, rr2, rr); break;
  }
}

// node-wise stages of comm 3 / comm 4 (cancel lists, clearIblanks)
struct NodeSeg {
  int off, n;
  const int *tmin;
  const unsigned char *cancelv;
  const int *has, *hblk, *hrem;
  int *iblank;
};
__global__ void k_interp_reset_all(int ns, const NodeSeg *__restrict__ T)
{
  const NodeSeg &S = seg_of(ns, T);
  const int v = blockIdx.x * BS + threadIdx.x - S.off;
  if (v < S.n && S.tmin[v] != BIGINT) S.iblank[v] = 1;
}
__global__ void k_cancel_all(int ns, const NodeSeg *__restrict__ T, const CancelTables *__restrict__ C)
{
  const NodeSeg &S = seg_of(ns, T);
  const int v = blockIdx.x * BS + threadIdx.x - S.off;
  if (v >= S.n || !S.cancelv[v] || !S.has[v]) return;
  int d = S.hblk[v];
  int iptr = C->i2d[d][S.hrem[v]];
  if (iptr > -1) C->cancel[d][iptr] = 1;
}
__global__ void k_clear_neg_all(int ns, const NodeSeg *__restrict__ T)
{
  const NodeSeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  if (i < S.n && S.iblank[i] < 0) S.iblank[i] = 1;
}

// resetCoincident per block (its query points), and the receptor marks
// (its interpolation records)
struct QuerySeg {
  int off, n;
  int *i2d;
  const int *xtag;
  int *cancel;
  int *ireset;
  const int *rnode, *rblk;
};
__global__ void k_coinc1_all(int ns, const QuerySeg *__restrict__ T)
{
  const QuerySeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  if (i >= S.n) return;
  int p = S.i2d[i];
  if (p > -1) atomicMin(&S.ireset[S.xtag[i]], S.cancel[p]);
}
__global__ void k_coinc2_all(int ns, const QuerySeg *__restrict__ T)
{
  const QuerySeg &S = seg_of(ns, T);
  const int i = blockIdx.x * BS + threadIdx.x - S.off;
  if (i >= S.n) return;
  int p = S.i2d[i];
  if (p > -1 && S.cancel[p] == 1) S.cancel[p] = S.ireset[S.xtag[i]];
}
__global__ void k_set_receptors_all(int ns, const QuerySeg *__restrict__ T, const IblankTables *__restrict__ I)
{
  const QuerySeg &S = seg_of(ns, T);
  const int r = blockIdx.x * BS + threadIdx.x - S.off;
  if (r < S.n && !S.cancel[r]) I->ib[S.rblk[r]][S.rnode[r]] = -1;
}

// MeshBlock::reduce_fringes (bookKeeping.C:650-760).  iblank_reduced starts
// as the field nodes; each of the nfringe + 1 sweeps reads the previous
// sweep's values and, in every cell that is partly field or kept receptor,
// turns the 0 of a receptor node into its iblank.  A sweep only moves values
// from 0 to the node's own iblank, so the cells can run in any order.
__global__ void k_rf_init(int n, const int *__restrict__ iblank, int *__restrict__ ibr, int *__restrict__ tmp)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = iblank[i] > 0 ? iblank[i] : 0;
  ibr[i] = v;
  tmp[i] = v;
}

__global__ void k_rf_sweep(int ncells, TypeInfo ti, const int *__restrict__ conn, const int *__restrict__ ibr,
                           const int *__restrict__ iblank, int *tmp)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  int ncount = 0;
  for (int m = 0; m < nv; m++) {
    const int r = ibr[cv[m]];
    ncount += (r == 1 || r < 0);
  }
  if (ncount > 0 && ncount < nv)
    for (int m = 0; m < nv; m++) {
      const int v = cv[m];
      if (tmp[v] == 0 && iblank[v] < 0) tmp[v] = iblank[v];
    }
}

// the receptors the reduction drops: the new cancel list
__global__ void k_rf_cancel(int n, const int *__restrict__ iblank, const int *__restrict__ ibr,
                            unsigned char *__restrict__ cancelv)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cancelv[i] = (iblank[i] < 0 && ibr[i] == 0);
}

// MeshBlock::clearIblanks with iblank_reduced (bookKeeping.C:615-625)
__global__ void k_rf_clear(int n, const int *__restrict__ ibr, int *__restrict__ iblank)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (iblank[i] < 0) iblank[i] = 1;
  if (ibr[i] == 0) iblank[i] = 0;
}

// MeshBlock::getCellIblanks (highOrder.C:88-150)
__device__ __forceinline__ void cell_iblank_one(int c, const TypeInfo &ti, const int *__restrict__ conn,
                                                const int *__restrict__ iblank, int *__restrict__ ibc)
{
  int nv = ti.nv[type_of(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  int out = 1, cnt = 0;
  for (int m = 0; m < nv; m++) {
    int v = iblank[cv[m]];
    if (v == 0) {
      out = 0;
      break;
    }
    cnt += (v == -1);
  }
  if (out && cnt == nv) out = -1;
  ibc[c] = out;
}

__global__ void k_cell_iblank(int ncells, TypeInfo ti, const int *__restrict__ conn, const int *__restrict__ iblank,
                              int *__restrict__ ibc)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  cell_iblank_one(c, ti, conn, iblank, ibc);
}

// k_cell_iblank for every block
struct CellIbSeg {
  int off, n;
  TypeInfo ti;
  const int *conn, *iblank;
  int *ibc;
};
__global__ void k_cell_iblank_all(int ns, const CellIbSeg *__restrict__ T)
{
  const CellIbSeg &S = seg_of(ns, T);
  const int c = blockIdx.x * BS + threadIdx.x - S.off;
  if (c >= S.n) return;
  cell_iblank_one(c, S.ti, S.conn, S.iblank, S.ibc);
}

// ---------------------------------------------------------------------------
// dataUpdate: MeshBlock::getInterpolatedSolution (dataUpdate.C:27-103) and
// updateSolnData (dataUpdate.C:105-119)
// ---------------------------------------------------------------------------

// Records are stored structure-of-arrays (slot m of record r at m*n + r), so
// a wave's index and weight loads are coalesced; the sum over the weights
// runs in slot order m = 0..nw-1 as the reference sums it.
template <class T>
__global__ void k_fill(size_t n, T *p, T v)
{
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class T>
void fill(T *p, size_t n, T v, hipStream_t s)
{
  if (n) k_fill<T><<<nb((long long)n), BS, 0, s>>>(n, p, v);
}

// launches the donor searches S[0..ns) together (fast wave-cooperative
// path, then the queries that overflowed it); with check, the host reads
// the candidate count after the walk and walks again with a larger list
// when it did not fit, so no later kernel reads a list that overflowed
void run_search(const SearchView *S, int ns, SearchScratch &w, hipStream_t s, bool check, bool list_only = false)
{
  std::vector<int> qb(ns + 1, 0);
  for (int k = 0; k < ns; k++) {
    if (S[k].nq < 0) throw std::runtime_error("tioga_amd: negative query count");
    qb[k + 1] = qb[k] + S[k].nq;
    if (qb[k + 1] < qb[k]) throw std::runtime_error("tioga_amd: too many queries in one search launch");
  }
  const int nqt = qb[ns];
  w.last_ns = ns;
  w.last_nqt = nqt;
  if (nqt <= 0) return;
  // the processing order, walking counts and the query -> search map come
  // from the ordering step (Engine::order_sets) of this launch
  if (w.qsid.cap < (size_t)nqt || w.nin.cap < (size_t)ns || w.wstart.cap < (size_t)ns + 1)
    throw std::runtime_error("tioga_amd: search launched without its ordering step");
  w.sv.reserve(sizeof(SearchView) * ns);
  w.qb.reserve(ns + 1);
  if (!w.skip_uploads) {
    TG_HIP(hipMemcpyAsync(w.sv.p, S, sizeof(SearchView) * ns, hipMemcpyHostToDevice, s));
    TG_HIP(hipMemcpyAsync(w.qb.p, qb.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice, s));
  }
  w.counters.reserve(NRT * USED_STRIDE);
  // per query: lowest clean-hit rank, highest tested (rank, outcome), and the
  // tie-band count when a search has a counter
  bool tieband = false;
  for (int k = 0; k < ns; k++) tieband = tieband || S[k].tieband != nullptr;
  w.best.reserve(nqt);
  w.last.reserve(nqt);
  if (tieband) w.acc.reserve(nqt);
  const SearchSet T{(const SearchView *)w.sv.p, w.qb.p, w.nin.p, w.wstart.p, w.qsid.p, ns, nqt};
  long long nwave = 0;  // an upper bound of the walking waves
  for (int k = 0; k < ns; k++) nwave += (S[k].nq + WAVE - 1) / WAVE;
  // candidates from the bin grids of the trees that have one (k_find), the
  // other trees' by the walk (k_collect)
  bool bins = false, walk = false;
  for (int k = 0; k < ns; k++) {
    const bool b = w.tune.bins > 0 && S[k].boff != nullptr;
    bins = bins || b;
    walk = walk || !b;
  }
  w.last_bins = bins;
  w.last_nwave = nwave;
  if (bins) {  // the listing's work items: segments of the walking waves' bin lists
    w.nseg.reserve((size_t)nwave + 1);
    w.segstart.reserve((size_t)nwave + 1);
    fill<int>(w.nseg.p, (size_t)nwave + 1, 0, s);
    k_find_cost<<<(unsigned)std::min<long long>(nb(nwave * WAVE), FIND_COST_GRID), BS, 0, s>>>(
        T, (int)nwave, w.tune.seg, w.nseg.p, w.best.p, w.last.p, tieband ? w.acc.p : nullptr);
    exclusive_sum(w.nseg.p, w.segstart.p, (int)nwave + 1, s);
  } else {
    fill<int>(w.best.p, nqt, 0x7fffffff, s);
    fill<int>(w.last.p, nqt, -1, s);
    if (tieband) fill<int>(w.acc.p, nqt, 0, s);
  }
  if (w.tcap[0] == 0)
    for (int t = 0; t < NTYPE; t++) w.tcap[t] = std::max<size_t>((size_t)nqt, 1 << 18);
  for (;;) {
    const CandLayout Y = cand_layout(w.tcap);
    if (Y.slots > 0x7fffffff) throw std::runtime_error("tioga_amd: candidate list too large");
    const size_t slots = Y.slots;
    w.cell.reserve(slots);
    w.query.reserve(slots);
    w.status.reserve(slots);
    CandOut O{w.cell.p, w.query.p, w.status.p, w.counters.p, {}, {}, w.best.p, w.last.p, tieband ? w.acc.p : nullptr};
    for (int t = 0; t < NTYPE; t++) {
      O.rcap[t] = Y.rcap[t];
      O.tbase[t] = Y.tbase[t];
    }
    fill<int>(w.counters.p, NRT * USED_STRIDE, 0, s);
    if (bins) k_find<<<FIND_GRID, WAVE, 0, s>>>(T, O, (int)nwave, w.tune.seg, w.segstart.p);
    if (walk) k_collect<<<(unsigned)nwave, WAVE, 0, s>>>(T, O, bins ? 1 : 0);
    if (check) {
      std::vector<int> used(NRT * USED_STRIDE);
      TG_HIP(hipMemcpyAsync(used.data(), w.counters.p, sizeof(int) * used.size(), hipMemcpyDeviceToHost, s));
      TG_HIP(hipStreamSynchronize(s));
      bool again = false;
      for (int t = 0; t < NTYPE; t++) {
        int most = 0;
        for (int r = 0; r < NREG; r++) most = std::max(most, used[(t * NREG + r) * USED_STRIDE]);
        if (most > Y.rcap[t]) {  // grow this type's regions and list again (nothing was tested yet)
          w.tcap[t] = (size_t)NREG * ((size_t)most + most / 4);
          again = true;
        }
      }
      if (again) continue;
    }
    if (list_only) return;
    const int lcap = (int)std::min<size_t>(slots / 4 + 1024, 0x7fffffff / 4);
    w.lj.reserve(4 * (size_t)lcap);
    w.lst.reserve(4 * (size_t)lcap);
    w.lcount.reserve(4);
    fill<int>(w.lcount.p, 4, 0, s);
    LongQueue L{w.lj.p, w.lst.p, w.lcount.p, lcap};
    // the pyramids on the side stream: their tests (divergent Newton solves,
    // a quarter of the lanes busy) and their continued solves (a few
    // thousand threads, each hundreds of dependent iterations, bound by
    // latency and not by the device) run beside the other cell types' tests
    const unsigned eg = (unsigned)((std::max(w.tune.eval_grid, NREG) + NREG - 1) / NREG * NREG);
    if (!w.side) {
      TG_HIP(hipStreamCreateWithFlags(&w.side, hipStreamNonBlocking));
      TG_HIP(hipEventCreateWithFlags(&w.fork, hipEventDisableTiming));
      TG_HIP(hipEventCreateWithFlags(&w.join, hipEventDisableTiming));
    }
    TG_HIP(hipEventRecord(w.fork, s));
    TG_HIP(hipStreamWaitEvent(w.side, w.fork, 0));
    k_eval<5, 4><<<eg, 256, 0, w.side>>>(T, O, L);
    k_eval_long<5><<<1024, 256, 0, w.side>>>(T, O, L);
    TG_HIP(hipEventRecord(w.join, w.side));
    k_eval<6, 4><<<eg, 256, 0, s>>>(T, O, L);
    k_eval<8, 4><<<eg, 256, 0, s>>>(T, O, L);
    k_eval_long<6><<<1024, 256, 0, s>>>(T, O, L);
    k_eval_long<8><<<1024, 256, 0, s>>>(T, O, L);
    k_eval<4, 8><<<eg, 256, 0, s>>>(T, O, L);
    TG_HIP(hipStreamWaitEvent(s, w.join, 0));
    k_last<<<LAST_GRID, 256, 0, s>>>(O);
    k_select<<<(unsigned)std::min<long long>((nwave + 3) / 4, 8192), 256, 0, s>>>(T, O);
    return;
  }
}

// the listed candidates of a single search: search point index, preorder
// rank and cell of each (the element-box test is left to the caller), the
// regions' entries one after the other
__global__ void k_export_cands(SearchSet T, CandOut O, int *__restrict__ qi, int *__restrict__ rank,
                               int *__restrict__ cell)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  int t, r;
  if (j >= O.tbase[NTYPE - 1] + NREG * O.rcap[NTYPE - 1] || !slot_live(O, j, t, r)) return;
  int k = j - region_first(O, t, r);
  for (int q = 0; q < t * NREG + r; q++) k += region_held(O, q / NREG, q % NREG);
  const SearchView S = view_of(T, 0);
  const int v = O.cell[j];
  qi[k] = S.perm[O.query[j]];
  rank[k] = v & 0x0fffffff;
  cell[k] = cell_of(S, v);
}

// a search's bin grid, when one was built
inline void view_bins(SearchView &S, const BinGrid &g)
{
  S.boff = g.nbins ? g.off.p : nullptr;
  S.bent = g.nbins ? g.ent.p : nullptr;
  S.bshift = g.shift;
}

// a small device-side copy of a host struct, valid until the next call
template <class T>
struct DevStruct {
  DBuf<T> buf;
  T *put(const T &h, hipStream_t s)
  {
    buf.reserve(1);
    TG_HIP(hipMemcpyAsync(buf.p, &h, sizeof(T), hipMemcpyHostToDevice, s));
    return buf.p;
  }
};

struct Timer {
  hipEvent_t a, b;
  hipStream_t s;
  explicit Timer(hipStream_t st) : s(st)
  {
    TG_HIP(hipEventCreate(&a));
    TG_HIP(hipEventCreate(&b));
    TG_HIP(hipEventRecord(a, s));
  }
  double stop()
  {
    float ms = 0;
    TG_HIP(hipEventRecord(b, s));
    TG_HIP(hipEventSynchronize(b));
    TG_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

}  // namespace

// ===========================================================================
// Engine
// ===========================================================================

Engine::Engine()
{
  TG_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
  err_.reserve(1);
}

Engine::~Engine()
{
  drop_search_graph();
  query_set_clear();
  for (Block *b : blocks_) delete b;
  for (auto *q : qdev_) delete q;
  for (auto *q : ptmp_) delete q;
  for (auto *q : pitmp_) delete q;
  amr_release();
  if (s_) (void)hipStreamDestroy(s_);
  if (s2_) (void)hipStreamDestroy(s2_);
  if (wb_) (void)hipHostFree(wb_);
  for (hipEvent_t e : wb_ev_) (void)hipEventDestroy(e);
}

void Engine::sync() { TG_HIP(hipStreamSynchronize(s_)); }

Block *Engine::find(int btag, int *iblk)
{
  const int i = index_of(btag);
  if (i < 0) return nullptr;
  if (iblk) *iblk = i;
  return blocks_[i];
}

// Block indices, keys and per-rank indices from blocks_ (rank-major order).
void Engine::reindex_blocks()
{
  tag_to_blk_.clear();
  std::map<int, int> nper;
  for (int i = 0; i < (int)blocks_.size(); i++) {
    Block &b = *blocks_[i];
    b.lidx = nper[b.rank]++;
    tag_to_blk_[block_key(b.rank, b.tag)] = i;
  }
}

int Engine::nranks() const
{
  int r = 0;
  for (const Block *b : blocks_) r = std::max(r, b->rank + 1);
  return std::max(r, 1);
}

const std::vector<int> &Engine::partners(int r) const
{
  static const std::vector<int> none;
  return r >= 0 && r < (int)partners_.size() ? partners_[r] : none;
}

// tioga::registerGridData (tioga.C:64-86)
void Engine::register_grid(int btag, int nnodes, double *xyz, int *ibl, int nwbc, int nobc, int *wbc, int *obc,
                           int ntypes, const int *nv, const int *nc, int **vconn)
{
  if (ntypes < 1 || ntypes > MAXTYPES) throw std::runtime_error("tioga_amd: ntypes out of range");
  if (btag < 1) throw std::runtime_error("tioga_amd: mesh tags must be >= 1");
  if (cur_rank < 0) throw std::runtime_error("tioga_amd: negative rank");
  int iblk;
  Block *b = find(btag, &iblk);
  if (!b) {
    if (blocks_.size() >= 128) throw std::runtime_error("tioga_amd: at most 128 blocks");
    if (amr_mode() && cur_rank != 0) throw std::runtime_error("tioga_amd: the AMR path runs on one rank");
    b = new Block;
    b->rank = cur_rank;
    // after every block of a rank <= cur_rank: rank-major, registration order within a rank
    size_t at = 0;
    while (at < blocks_.size() && blocks_[at]->rank <= cur_rank) at++;
    blocks_.insert(blocks_.begin() + at, b);
    b->tag = btag;
    reindex_blocks();
  }
  b->tag = btag;
  b->nnodes = nnodes;
  b->ux = xyz;
  b->uiblank = ibl;
  b->nwbc = nwbc;
  b->nobc = nobc;
  b->uwbc = wbc;
  b->uobc = obc;
  b->ti.ntypes = ntypes;
  b->ti.cstart[0] = 0;
  for (int t = 0; t < ntypes; t++) {
    if (nv[t] != 4 && nv[t] != 5 && nv[t] != 6 && nv[t] != 8)
      throw std::runtime_error("tioga_amd: cells must have 4, 5, 6 or 8 vertices");
    b->ti.nv[t] = nv[t];
    b->ti.cstart[t + 1] = b->ti.cstart[t] + nc[t];
    b->uconn[t] = vconn[t];
  }
  b->ncells = b->ti.cstart[ntypes];
  if (b->ncells >= (1 << 28)) throw std::runtime_error("tioga_amd: at most 2^28 - 1 cells per block");
  b->topo_uploaded = false;
  b->unres = b->ucres = nullptr;
  connected_ = false;
  upd_.valid = false;
}

void Engine::set_gids(int btag, uint64_t *cell_gid, uint64_t *node_gid)
{
  Block *b = find(btag);
  if (!b) throw std::runtime_error("tioga_amd: gids for an unknown tag");
  b->ucell_gid = cell_gid;
  b->unode_gid = node_gid;
}

void Engine::set_resolutions_first(double *nres, double *cres)
{
  if (blocks_.empty()) throw std::runtime_error("tioga_amd: setResolutions before any registration");
  blocks_[0]->unres = nres;
  blocks_[0]->ucres = cres;
}

void Engine::set_cell_iblank_first(int *ibc)
{
  if (blocks_.empty()) throw std::runtime_error("tioga_amd: set_cell_iblank before any registration");
  blocks_[0]->uiblank_cell = ibc;
}

void Engine::set_resolutions(int btag, double *nres, double *cres)
{
  Block *b = find(btag);
  if (!b) throw std::runtime_error("tioga_amd: setResolutions for an unknown tag");
  b->unres = nres;
  b->ucres = cres;
}

void Engine::set_cell_iblank(int btag, int *ibc)
{
  Block *b = find(btag);
  if (!b) throw std::runtime_error("tioga_amd: set_cell_iblank for an unknown tag");
  b->uiblank_cell = ibc;
}

void Engine::register_solution(int btag, double *q)
{
  Block *b = find(btag);
  if (!b) throw std::runtime_error("tioga_amd: registerSolution for an unknown tag");
  b->uq = q;
}

// Node ids outside 1..nnodes would make every later kernel read out of
// bounds; reject them before anything reaches the device.
static void validate_ids(const int *ids, size_t n, int nnodes, const char *what)
{
  for (size_t i = 0; i < n; i++)
    if (ids[i] < BASE || ids[i] > nnodes)
      throw std::runtime_error(std::string("tioga_amd: ") + what + " holds node id " + std::to_string(ids[i]) +
                               " outside 1.." + std::to_string(nnodes));
}

void Engine::upload_topology(Block &b)
{
  const int N = b.ncells, V = b.nnodes;
  for (int t = 0; t < b.ti.ntypes; t++)
    validate_ids(b.uconn[t], (size_t)(b.ti.cstart[t + 1] - b.ti.cstart[t]) * b.ti.nv[t], V, "connectivity");
  b.conn.reserve(8 * (size_t)N);
  DBuf<int> stage;
  for (int t = 0; t < b.ti.ntypes; t++) {
    int nc = b.ti.cstart[t + 1] - b.ti.cstart[t];
    if (nc == 0) continue;
    size_t n = (size_t)nc * b.ti.nv[t];
    stage.reserve(n);
    TG_HIP(hipMemcpyAsync(stage.p, b.uconn[t], n * sizeof(int), hipMemcpyHostToDevice, s_));
    k_pack_conn<<<nb(nc), BS, 0, s_>>>(nc, b.ti.nv[t], stage.p, b.ti.cstart[t], b.conn.p);
  }
  // node -> cell incidence in cell order: stable sort of (node, cell) pairs
  size_t np = 8 * (size_t)N;
  DBuf<unsigned> k0, k1;
  DBuf<int> v0, v1, cnt;
  k0.reserve(np);
  k1.reserve(np);
  v0.reserve(np);
  v1.reserve(np);
  cnt.reserve(V + 1);
  fill<int>(cnt.p, V + 1, 0, s_);
  if (N) k_incidence<<<nb(N), BS, 0, s_>>>(N, V, b.conn.p, k0.p, v0.p, cnt.p);
  int bits = 1;
  while (bits < 32 && (1u << bits) <= (unsigned)V) bits++;
  sort_pairs(k0.p, k1.p, v0.p, v1.p, (int)np, bits, s_);
  b.n2c_off.reserve(V + 1);
  exclusive_sum(cnt.p, b.n2c_off.p, V + 1, s_);
  // the sorted pairs start with the valid ones (keys < V); keep them
  b.n2c.reserve(np);
  TG_HIP(hipMemcpyAsync(b.n2c.p, v1.p, np * sizeof(int), hipMemcpyDeviceToDevice, s_));
  sync();
  b.topo_uploaded = true;
}

// MeshBlock::preprocess (MeshBlock.C:74-92) + tagBoundary (MeshBlock.C:94-339)
void Engine::tag_boundary(Block &b)
{
  const int N = b.ncells, V = b.nnodes;
  b.cellRes.reserve(N);
  b.nodeRes.reserve(V);
  if (b.unres == nullptr && b.ucres == nullptr) {
    if (N) k_cell_res<<<nb(N), BS, 0, s_>>>(N, b.ti, b.conn.p, b.x.p, b.resolutionScale, b.cellRes.p);
    if (V) k_node_res<<<nb(V), BS, 0, s_>>>(V, b.n2c_off.p, b.n2c.p, b.cellRes.p, b.nodeRes.p);
  } else {
    TG_HIP(hipMemcpyAsync(b.cellRes.p, b.ucres, sizeof(double) * N, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(b.nodeRes.p, b.unres, sizeof(double) * V, hipMemcpyHostToDevice, s_));
  }
  // 10^3 bins over the block OBB, nodes in (bin, descending id) order
  for (int j = 0; j < 3; j++) b.mapdx[j] = 2 * b.obb.dxc[j] / 10;
  uscratch_.reserve(V);
  uscratch2_.reserve(V);
  iscratch_.reserve(V);
  b.invmap.reserve(V);
  b.nodebin.reserve(V);
  if (V) {
    k_node_bin<<<nb(V), BS, 0, s_>>>(V, b.x.p, b.obb, b.mapdx[0], b.mapdx[1], b.mapdx[2], uscratch_.p,
                                     iscratch_.p);
    sort_pairs(uscratch_.p, uscratch2_.p, iscratch_.p, b.invmap.p, V, 10, s_);
    k_scatter_bin<<<nb(V), BS, 0, s_>>>(V, uscratch2_.p, b.invmap.p, b.nodebin.p);
  }
  // first position of each bin in that order (binoff[1000] = V)
  {
    DBuf<int> bo;
    bo.reserve(1001);
    fill<int>(bo.p, 1001, V, s_);
    if (V) k_bin_starts<<<nb(V), BS, 0, s_>>>(V, uscratch2_.p, bo.p);
    b.binoff.assign(1001, V);
    TG_HIP(hipMemcpyAsync(b.binoff.data(), bo.p, sizeof(int) * 1001, hipMemcpyDeviceToHost, s_));
    sync();
    for (int k = 999; k >= 0; k--) b.binoff[k] = std::min(b.binoff[k], b.binoff[k + 1]);
  }
  // overset boundary and mexclude layers (the node flags depend on the
  // topology alone, so a block sharing it has them already)
  DBuf<int> lst;
  if (b.share == 0) b.obcnode.reserve(V);
  if (b.share == 0) fill<unsigned char>(b.obcnode.p, V, 0, s_);
  if (b.share == 0 && b.nobc > 0) {
    lst.reserve(b.nobc);
    TG_HIP(hipMemcpyAsync(lst.p, b.uobc, sizeof(int) * b.nobc, hipMemcpyHostToDevice, s_));
    k_flag_list<<<nb(b.nobc), BS, 0, s_>>>(b.nobc, lst.p, b.obcnode.p);
  }
  DBuf<unsigned char> ex, ex1;
  ex.reserve(V);
  ex1.reserve(V);
  fill<unsigned char>(ex.p, V, 0, s_);
  fill<unsigned char>(ex1.p, V, 0, s_);
  if (N) k_obc_cells<<<nb(N), BS, 0, s_>>>(N, b.ti, b.conn.p, b.obcnode.p, b.nodeRes.p, ex.p, ex1.p);
  for (int it = 0; it < mexclude; it++) {
    if (N) k_mexclude<<<nb(N), BS, 0, s_>>>(N, b.ti, b.conn.p, ex.p, ex1.p, b.cellRes.p);
    TG_HIP(hipMemcpyAsync(ex.p, ex1.p, V, hipMemcpyDeviceToDevice, s_));
  }
  // wall nodes, for the hole map
  if (b.share == 0) b.wallnode.reserve(V);
  if (b.share == 0) fill<unsigned char>(b.wallnode.p, V, 0, s_);
  if (b.share == 0 && b.nwbc > 0) {
    lst.reserve(b.nwbc);
    TG_HIP(hipMemcpyAsync(lst.p, b.uwbc, sizeof(int) * b.nwbc, hipMemcpyHostToDevice, s_));
    k_flag_list<<<nb(b.nwbc), BS, 0, s_>>>(b.nwbc, lst.p, b.wallnode.p);
  }
  // the cells touching a wall node, the only ones that mark the hole map
  if (b.share == 0) {
    b.nwallcells = 0;
    if (b.nwbc > 0 && N > 0) {
      iscratch_.reserve(N);
      k_wall_cells<<<nb(N), BS, 0, s_>>>(N, b.ti, b.conn.p, b.wallnode.p, iscratch_.p);
      b.wallcells.reserve(N);
      b.nwallcells = select_flagged(iscratch_.p, N, b.wallcells.p, s_);
    }
  }
  sync();
}

// tioga::profile (tioga.C:95-110).  only_tag >= 0: the blocks other than
// that one keep their preprocessed geometry (their coordinates, resolutions
// and topology must not have changed since it was computed) and only get
// iblank = 1, as MeshBlock::preprocess sets it (MeshBlock.C:80); for them the
// result equals a full preprocess, which recomputes the same values.
void Engine::profile(int only_tag)
{
  for (Block *bp : blocks_) {
    Block &b = *bp;
    const bool keep = only_tag >= 0 && b.tag != only_tag && b.topo_uploaded && b.cellRes.p != nullptr;
    if (b.share == 2 || keep) {  // another engine's (or kept) geometry: only this engine's iblank
      b.iblank.reserve(b.nnodes);
      fill<int>(b.iblank.p, b.nnodes, 1, s_);
      if (b.writeback && b.uiblank)
        for (int i = 0; i < b.nnodes; i++) b.uiblank[i] = 1;
      continue;
    }
    if (!b.topo_uploaded) upload_topology(b);
    if (b.share == 0) {
      validate_ids(b.uwbc, b.nwbc, b.nnodes, "wall boundary list");
      validate_ids(b.uobc, b.nobc, b.nnodes, "overset boundary list");
    }
    if (b.writeback && b.uiblank)
      for (int i = 0; i < b.nnodes; i++) b.uiblank[i] = 1;  // MeshBlock.C:80
    b.x.reserve(3 * (size_t)b.nnodes);
    TG_HIP(hipMemcpyAsync(b.x.p, b.ux, sizeof(double) * 3 * b.nnodes, hipMemcpyHostToDevice, s_));
    b.iblank.reserve(b.nnodes);
    fill<int>(b.iblank.p, b.nnodes, 1, s_);
    find_obb(b.ux, nullptr, b.nnodes, b.obb);
    tag_boundary(b);
  }
}

// x' = R (x0 - c) + c + t per node, in this order (the host form of the
// serving loop's moves, evaluated the same way)
__global__ void k_rigid(int n, const double *__restrict__ x0, const double *__restrict__ m, double *__restrict__ x)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d0 = x0[3 * (size_t)i] - m[9], d1 = x0[3 * (size_t)i + 1] - m[10], d2 = x0[3 * (size_t)i + 2] - m[11];
  for (int j = 0; j < 3; j++) {
    const double y = m[3 * j] * d0 + m[3 * j + 1] * d1 + m[3 * j + 2] * d2;
    x[3 * (size_t)i + j] = y + m[9 + j] + m[12 + j];
  }
}

void Engine::profile_moved(int btag, const double *x0_dev, const double rot[9], const double c[3],
                           const double t[3])
{
  Block *bm = find(btag);
  if (!bm || !bm->topo_uploaded) throw std::runtime_error("tioga_amd: moved block not adopted");
  for (Block *bp : blocks_) {
    Block &b = *bp;
    b.iblank.reserve(b.nnodes);
    fill<int>(b.iblank.p, b.nnodes, 1, s_);
    if (b.writeback && b.uiblank)
      for (int i = 0; i < b.nnodes; i++) b.uiblank[i] = 1;  // MeshBlock.C:80
  }
  Block &b = *bm;
  double m[15];
  memcpy(m, rot, 9 * sizeof(double));
  memcpy(m + 9, c, 3 * sizeof(double));
  memcpy(m + 12, t, 3 * sizeof(double));
  moved_.reserve(15);
  TG_HIP(hipMemcpyAsync(moved_.p, m, sizeof m, hipMemcpyHostToDevice, s_));
  b.x.reserve_own(3 * (size_t)b.nnodes);
  if (b.nnodes) k_rigid<<<nb(b.nnodes), BS, 0, s_>>>(b.nnodes, x0_dev, moved_.p, b.x.p);
  // the block OBB (findOBB, utils.c:25-165) from the device coordinates, with
  // the bits of the host running sums
  if (b.nnodes) {
    const ObbSet os{b.nnodes, b.x.p};
    find_obbs_device(&os, 1, &b.obb, obbw_, s_, nullptr);
  }
  tag_boundary(b);
  connected_ = false;
  upd_.valid = false;
}

void Engine::adopt_block(const Engine &src, int src_tag, int as_tag, int share, double *ux, int *uiblank,
                         bool writeback, int slot)
{
  if (src.nranks() > 1 || cur_rank != 0) throw std::runtime_error("tioga_amd: serving runs on one rank");
  const int si = src.index_of(src_tag);
  if (si < 0) throw std::runtime_error("tioga_amd: adopt of an unknown tag");
  const Block &S = *src.blocks_[si];
  if (!S.topo_uploaded || S.nodeRes.p == nullptr) throw std::runtime_error("tioga_amd: adopt before preprocess");
  if (share != 1 && share != 2) throw std::runtime_error("tioga_amd: share must be 1 or 2");
  Block *b;
  if (slot >= 0 && slot < (int)blocks_.size()) {
    b = blocks_[slot];
    tag_to_blk_.erase(block_key(0, b->tag));
  } else {
    if (blocks_.size() >= 128) throw std::runtime_error("tioga_amd: at most 128 blocks");
    b = new Block;
    blocks_.push_back(b);
    slot = (int)blocks_.size() - 1;
  }
  if (tag_to_blk_.count(block_key(0, as_tag))) throw std::runtime_error("tioga_amd: adopt under a tag in use");
  tag_to_blk_[block_key(0, as_tag)] = slot;
  b->tag = as_tag;
  b->rank = 0;
  b->lidx = slot;
  b->nnodes = S.nnodes;
  b->ncells = S.ncells;
  b->ti = S.ti;
  b->nwbc = S.nwbc;
  b->nobc = S.nobc;
  b->uwbc = S.uwbc;
  b->uobc = S.uobc;
  for (int t = 0; t < MAXTYPES; t++) b->uconn[t] = S.uconn[t];
  b->unres = S.unres;
  b->ucres = S.ucres;
  b->ucell_gid = S.ucell_gid;
  b->unode_gid = S.unode_gid;
  b->resolutionScale = S.resolutionScale;
  b->uiblank_cell = nullptr;
  b->uq = nullptr;
  b->share = share;
  b->writeback = writeback;
  b->conn.borrow(S.conn);
  b->n2c_off.borrow(S.n2c_off);
  b->n2c.borrow(S.n2c);
  b->wallnode.borrow(S.wallnode);
  b->wallcells.borrow(S.wallcells);
  b->nwallcells = S.nwallcells;
  b->obcnode.borrow(S.obcnode);
  b->topo_uploaded = true;
  if (share == 2) {
    b->ux = S.ux;
    b->uiblank = nullptr;
    b->x.borrow(S.x);
    b->cellRes.borrow(S.cellRes);
    b->nodeRes.borrow(S.nodeRes);
    b->nodebin.borrow(S.nodebin);
    b->invmap.borrow(S.invmap);
    b->binoff = S.binoff;
    b->obb = S.obb;
    memcpy(b->mapdx, S.mapdx, sizeof b->mapdx);
  } else {
    b->ux = ux;
    b->uiblank = uiblank;
    // buffers this engine owns again if they were views
    if (!b->x.own) b->x.release();
    if (!b->cellRes.own) b->cellRes.release();
    if (!b->nodeRes.own) b->nodeRes.release();
    if (!b->nodebin.own) b->nodebin.release();
    if (!b->invmap.own) b->invmap.release();
  }
  b->ninterp = 0;
  b->interp_host_valid = false;
  connected_ = false;
  upd_.valid = false;
}

__global__ void k_nonunit(int n, const int *__restrict__ ib, int *__restrict__ flag)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = ib[i] != 1;
}

__global__ void k_gather_int(int n, const int *__restrict__ idx, const int *__restrict__ src, int *__restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

void Engine::nonunit_iblank(int btag, std::vector<int> &nodes, std::vector<int> &vals)
{
  Block *b = find(btag);
  if (!b) throw std::runtime_error("tioga_amd: unknown tag");
  const int V = b->nnodes;
  nodes.clear();
  vals.clear();
  if (V == 0 || b->iblank.p == nullptr) return;
  iscratch_.reserve(V);
  iscratch2_.reserve(V);
  k_nonunit<<<nb(V), BS, 0, s_>>>(V, b->iblank.p, iscratch_.p);
  const int n = select_flagged(iscratch_.p, V, iscratch2_.p, s_);
  nodes.resize(n);
  vals.resize(n);
  if (n) {
    k_gather_int<<<nb(n), BS, 0, s_>>>(n, iscratch2_.p, b->iblank.p, iscratch_.p);
    TG_HIP(hipMemcpyAsync(nodes.data(), iscratch2_.p, sizeof(int) * n, hipMemcpyDeviceToHost, s_));
    TG_HIP(hipMemcpyAsync(vals.data(), iscratch_.p, sizeof(int) * n, hipMemcpyDeviceToHost, s_));
  }
  sync();
}

// tioga::getHoleMap (holeMap.C:40-188), single rank
void Engine::get_hole_map()
{
  int maxtag = 1;
  for (Block *b : blocks_) maxtag = std::max(maxtag, b->tag);
  std::vector<HoleMap> old;
  old.swap(hm_);
  hm_ = std::vector<HoleMap>(maxtag);
  // reuse the old maps' storage
  for (size_t i = 0; i < old.size() && i < hm_.size(); i++) std::swap(hm_[i].sam.p, old[i].sam.p), std::swap(hm_[i].sam.cap, old[i].sam.cap);
  // Per mesh tag: a wall exists if any block of the tag has wall nodes, and
  // the bounds are the min / max over the tag's blocks of their wall bounds,
  // where a block without wall nodes reports zeros (MeshBlock.C:640-647).
  // With one block per tag this is the single-rank map; with a tag on several
  // ranks it is the MPI_MAX / MPI_MIN / MPI_MAX allreduce of holeMap.C:86-112.
  std::vector<double> bmin(3 * maxtag, BIGVALUE), bmax(3 * maxtag, -BIGVALUE);
  for (HoleMap &h : hm_) h.exist = 0;
  for (Block *bp : blocks_) {
    Block &b = *bp;
    HoleMap &h = hm_[b.tag - 1];
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (b.nwbc > 0) {
      h.exist = 1;
      for (int j = 0; j < 3; j++) {
        lo[j] = BIGVALUE;
        hi[j] = -BIGVALUE;
      }
      for (int i = 0; i < b.nwbc; i++) {
        const double *p = b.ux + 3 * (size_t)(b.uwbc[i] - BASE);
        for (int j = 0; j < 3; j++) {
          lo[j] = dmin(lo[j], p[j]);
          hi[j] = dmax(hi[j], p[j]);
        }
      }
    }
    for (int j = 0; j < 3; j++) {
      bmin[3 * (b.tag - 1) + j] = std::min(bmin[3 * (b.tag - 1) + j], lo[j]);
      bmax[3 * (b.tag - 1) + j] = std::max(bmax[3 * (b.tag - 1) + j], hi[j]);
    }
  }
  std::vector<HoleMapView> views(maxtag);
  for (int i = 0; i < maxtag; i++) {
    HoleMap &h = hm_[i];
    HoleMapView &hv = views[i];
    hv.exist = h.exist;
    hv.sam = nullptr;
    if (!h.exist) continue;
    double ds[3];
    for (int j = 0; j < 3; j++) {
      h.ext[j] = bmin[3 * i + j];
      h.ext[j + 3] = bmax[3 * i + j];
      ds[j] = h.ext[j + 3] - h.ext[j];
    }
    double dsmax = dmax(dmax(ds[0], ds[1]), ds[2]);
    double dsbox = dsmax / HOLEMAPSIZE;
    for (int j = 0; j < 3; j++) {
      h.ext[j] -= (2 * dsbox);
      h.ext[j + 3] += (2 * dsbox);
      h.nx[j] = (int)floor(dmax((h.ext[j + 3] - h.ext[j]) / dsbox, 1));
    }
    size_t n = (size_t)h.nx[0] * h.nx[1] * h.nx[2];
    h.sam.reserve(n);
    if (!batched) fill<unsigned char>(h.sam.p, n, 0, s_);
    hv.sam = h.sam.p;
    memcpy(hv.nx, h.nx, sizeof hv.nx);
    memcpy(hv.ext, h.ext, sizeof hv.ext);
  }
  // the maps of every wall mesh together (batched form)
  std::vector<int> live;
  for (int i = 0; i < maxtag; i++)
    if (hm_[i].exist) live.push_back(i);
  if (live.size() > 128) throw std::runtime_error("tioga_amd: more than 128 wall meshes");
  MapSet ms{};
  MapSet *pms = nullptr;
  int allow = 7;  // isym stencil: 3 excludes z, 1 excludes x, 2 excludes y
  if (isym == 1) allow = 6;
  else if (isym == 2) allow = 5;
  else if (isym == 3) allow = 3;
  // with a 2-D stencil every slice normal to the excluded axis fills on its
  // own, in LDS when it fits
  const int saxis = isym == 3 ? 2 : (isym == 1 ? 0 : (isym == 2 ? 1 : -1));
  bool slices = saxis >= 0;
  int maxcells = 0, maxlines[3] = {0, 0, 0}, maxslices = 0;
  size_t maxslice = 1;  // bytes of the largest slice
  if (!live.empty()) {
    ms.nmap = (int)live.size();
    for (int k = 0; k < ms.nmap; k++) {
      HoleMap &h = hm_[live[k]];
      ms.sam[k] = h.sam.p;
      for (int j = 0; j < 3; j++) ms.nx[k][j] = h.nx[j];
      const long long cells = (long long)h.nx[0] * h.nx[1] * h.nx[2];
      if (cells > 0x7fffffffLL) throw std::runtime_error("tioga_amd: hole map too large");
      maxcells = std::max(maxcells, (int)cells);
      for (int a = 0; a < 3; a++) {
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        if (h.nx[a1] > 2 && h.nx[a2] > 2) maxlines[a] = std::max(maxlines[a], (h.nx[a1] - 2) * (h.nx[a2] - 2));
      }
      if (saxis >= 0) {
        const long long sb = (long long)h.nx[(saxis + 1) % 3] * h.nx[(saxis + 2) % 3];
        slices = slices && sb <= SLICE_MAX;
        maxslice = std::max<size_t>(maxslice, (size_t)std::min<long long>(sb, SLICE_MAX));
        maxslices = std::max(maxslices, h.nx[saxis] - 2);
      }
    }
    pms = (MapSet *)put_table(stw_.maps, ms);
    if (batched) k_maps_cells<<<dim3(nb(maxcells), ms.nmap), BS, 0, s_>>>(pms, 0);
  }
  {
    WallSet W{};
    std::vector<double> dsv(3 * (size_t)std::max(maxtag, 1), 1.0);
    std::vector<unsigned char *> sams(std::max(maxtag, 1), nullptr);
    for (int i = 0; i < maxtag; i++) {
      HoleMap &h = hm_[i];
      if (!h.exist) continue;
      for (int k = 0; k < 3; k++) dsv[3 * i + k] = (h.ext[k + 3] - h.ext[k]) / h.nx[k];
      sams[i] = h.sam.p;
    }
    for (Block *bp : blocks_) {
      Block &b = *bp;
      HoleMap &h = hm_[b.tag - 1];
      if (!h.exist || b.nwallcells == 0) continue;
      const int q = W.nb++;
      W.off[q + 1] = W.off[q] + b.nwallcells;
      W.cells[q] = b.wallcells.p;
      W.conn[q] = b.conn.p;
      W.x[q] = b.x.p;
      W.ti[q] = b.ti;
      W.map[q] = b.tag - 1;
    }
    if (W.nb > 0) {
      hm_view_.reserve(maxtag);
      TG_HIP(hipMemcpyAsync(hm_view_.p, views.data(), sizeof(HoleMapView) * maxtag, hipMemcpyHostToDevice, s_));
      wallds_.reserve(dsv.size());
      wallsams_.reserve(sams.size());
      TG_HIP(hipMemcpyAsync(wallds_.p, dsv.data(), sizeof(double) * dsv.size(), hipMemcpyHostToDevice, s_));
      TG_HIP(hipMemcpyAsync(wallsams_.p, sams.data(), sizeof(unsigned char *) * sams.size(),
                            hipMemcpyHostToDevice, s_));
      k_mark_wall<<<nb(W.off[W.nb]), BS, 0, s_>>>((const WallSet *)put_table(stw_.walls, W), hm_view_.p,
                                                  wallds_.p, wallsams_.p);
      sync();  // the tables above are host memory until the copies have run
    }
  }
  nmesh_ = maxtag;
  if (batched && !live.empty()) {
    const int nm = ms.nmap;
    k_maps_cells<<<dim3(nb(maxcells), nm), BS, 0, s_>>>(pms, 1);
    DBuf<int> &act = stw_.act, &chg = stw_.chg;
    act.reserve(nm);
    chg.reserve(nm);
    std::vector<int> hact(nm, 1), hchg(nm);
    if (slices && maxslices > 0)
      k_fill_slices<<<dim3(maxslices, nm), 256, maxslice, s_>>>(saxis, nm, &pms->sam[0], &pms->nx[0][0]);
    for (; !slices;) {
      TG_HIP(hipMemcpyAsync(act.p, hact.data(), sizeof(int) * nm, hipMemcpyHostToDevice, s_));
      fill<int>(chg.p, nm, 0, s_);
      for (int a = 0; a < 3; a++)
        if (((allow >> a) & 1) && maxlines[a] > 0)
          k_maps_line_fill<<<dim3(nb(maxlines[a]), nm), BS, 0, s_>>>(pms, a, allow, act.p, chg.p);
      TG_HIP(hipMemcpyAsync(hchg.data(), chg.p, sizeof(int) * nm, hipMemcpyDeviceToHost, s_));
      sync();
      bool any = false;
      for (int k = 0; k < nm; k++) {
        hact[k] = hact[k] && hchg[k];
        any = any || hact[k];
      }
      if (!any) break;
    }
    k_maps_cells<<<dim3(nb(maxcells), nm), BS, 0, s_>>>(pms, 2);
  }
  int *changed = err_.p;  // reuse the one-int device word
  for (int i = 0; i < maxtag && !batched; i++) {
    HoleMap &h = hm_[i];
    if (!h.exist) continue;
    size_t n = (size_t)h.nx[0] * h.nx[1] * h.nx[2];
    k_hole_shell<<<nb((long long)n), BS, 0, s_>>>(h.nx[0], h.nx[1], h.nx[2], h.sam.p);
    if (slices) {
      int k = 0;
      while (live[k] != i) k++;
      const size_t sb = (size_t)h.nx[(saxis + 1) % 3] * h.nx[(saxis + 2) % 3];  // <= SLICE_MAX here
      if (h.nx[saxis] > 2)
        k_fill_slices<<<dim3(h.nx[saxis] - 2, 1), 256, sb, s_>>>(saxis, 1, &pms->sam[k], &pms->nx[k][0]);
    }
    for (int round = 0; !slices; round++) {
      int hc = 0;
      TG_HIP(hipMemsetAsync(changed, 0, sizeof(int), s_));
      for (int a = 0; a < 3; a++) {
        if (!((allow >> a) & 1)) continue;
        long long lines = (long long)(h.nx[(a + 1) % 3] - 2) * (h.nx[(a + 2) % 3] - 2);
        if (lines > 0)
          k_line_fill<<<nb(lines), BS, 0, s_>>>(h.nx[0], h.nx[1], h.nx[2], a, allow, h.sam.p, changed);
      }
      TG_HIP(hipMemcpyAsync(&hc, changed, sizeof(int), hipMemcpyDeviceToHost, s_));
      sync();
      if (!hc) break;
    }
    k_hole_final<<<nb((long long)n), BS, 0, s_>>>(n, h.sam.p);
  }
  hm_view_.reserve(maxtag);
  TG_HIP(hipMemcpyAsync(hm_view_.p, views.data(), sizeof(HoleMapView) * maxtag, hipMemcpyHostToDevice, s_));
  sync();
}

// The pair that receives pair i's packets: the packet goes to the rank of
// pair i's OBB, whose map (intBoxMap) holds its pairs by receive key, the
// last one inserted under a key (exchangeBoxes.C:206-243).
int Engine::recv_pair(int i) const
{
  const int key = pairs_[i].send_tag, dst = blocks_[pairs_[i].ob]->rank;
  int ii = -1;
  for (int j = 0; j < (int)pairs_.size(); j++)
    if (pairs_[j].recv_tag == key && blocks_[pairs_[j].local]->rank == dst) ii = j;
  return ii;
}

// tioga::exchangeBoxes (exchangeBoxes.C:34-272) for every rank
void Engine::exchange_boxes()
{
  const int nb_ = (int)blocks_.size();
  int maxtag = -1;
  for (Block *b : blocks_) maxtag = std::max(maxtag, std::abs(b->tag));
  const int mxtgsqr = maxtag * maxtag;
  // each rank's pairs in its own order (every OBB of the allgather, then
  // its blocks), rank after rank; a rank's partners are the ranks of the
  // OBBs it pairs with (exchangeBoxes.C:131-176)
  const int nr = nranks();
  pairs_.clear();
  partners_.assign(nr, std::vector<int>());
  for (int r = 0; r < nr; r++) {
    std::vector<char> flag(nr, 0);
    for (int ob = 0; ob < nb_; ob++)
      for (int ib = 0; ib < nb_; ib++) {
        Block &A = *blocks_[ib], &B = *blocks_[ob];
        if (A.rank != r) continue;
        if (std::abs(B.tag) == A.tag) continue;
        if (obb_intersect(A.obb, B.obb) || obb_intersect(B.obb, A.obb)) {
          Pair p;
          p.local = ib;
          p.ob = ob;
          pairs_.push_back(p);
          flag[B.rank] = 1;
        }
      }
    for (int k = 0; k < nr; k++)
      if (flag[k]) partners_[r].push_back(k);
  }
  nsend_ = pairs_.empty() ? 0 : 1;
  const int np = (int)pairs_.size();
  kscratch_.reserve(6 * (size_t)std::max(np, 1));
  std::vector<unsigned long long> init(6 * (size_t)std::max(np, 1));
  for (int i = 0; i < np; i++)
    for (int j = 0; j < 3; j++) {
      init[6 * i + j] = dkey(BIGVALUE);
      init[6 * i + 3 + j] = dkey(-BIGVALUE);
    }
  TG_HIP(hipMemcpyAsync(kscratch_.p, init.data(), sizeof(unsigned long long) * init.size(), hipMemcpyHostToDevice,
                        s_));
  std::vector<int> npart(nb_, 0);  // pairs per local block
  for (const Pair &p : pairs_) npart[p.local]++;
  for (int i = 0; i < np; i++) {
    Pair &p = pairs_[i];
    Block &A = *blocks_[p.local], &B = *blocks_[p.ob];
    p.recv_tag = mxtgsqr * B.rank + maxtag * (A.tag - 1) + B.tag - 1;
    p.send_tag = mxtgsqr * A.rank + maxtag * (B.tag - 1) + (A.tag - 1);
    if (A.ncells && !(batched && npart[p.local] > 1))
      k_reduced_obb<<<nb(A.ncells), BS, 0, s_>>>(A.ncells, A.ti, A.conn.p, A.x.p, A.obb, B.obb,
                                                 kscratch_.p + 6 * (size_t)i);
    memcpy(p.obc.vec, B.obb.vec, sizeof(double) * 9);
  }
  if (batched) {  // one pass over each block's cells for all of its pairs
    std::vector<BoxPairs> tabs;
    std::vector<int> tblk;
    for (int ib = 0; ib < nb_; ib++) {
      Block &A = *blocks_[ib];
      BoxPairs bp{};
      for (int i = 0; i < np; i++)
        if (pairs_[i].local == ib) {
          if (bp.np == 128) throw std::runtime_error("tioga_amd: more than 128 partners of one block");
          bp.pair[bp.np] = i;
          bp.obc[bp.np] = blocks_[pairs_[i].ob]->obb;
          bp.np++;
        }
      if (bp.np < 2 || A.ncells == 0) continue;
      tabs.push_back(bp);
      tblk.push_back(ib);
    }
    if (!tabs.empty()) {
      // every block's partner table in one upload
      stw_.boxes.reserve(sizeof(BoxPairs) * tabs.size());
      TG_HIP(hipMemcpyAsync(stw_.boxes.p, tabs.data(), sizeof(BoxPairs) * tabs.size(), hipMemcpyHostToDevice, s_));
      const BoxPairs *dt = (const BoxPairs *)stw_.boxes.p;
      for (size_t k = 0; k < tabs.size(); k++) {
        Block &A = *blocks_[tblk[k]];
        k_reduced_obb_multi<<<std::min(nb(A.ncells), 1024u), BS, 0, s_>>>(A.ncells, A.ti, A.conn.p, A.x.p, A.obb,
                                                                          dt + k, kscratch_.p);
      }
    }
    sync();
  }
  std::vector<unsigned long long> keys(6 * (size_t)std::max(np, 1));
  TG_HIP(hipMemcpyAsync(keys.data(), kscratch_.p, sizeof(unsigned long long) * keys.size(), hipMemcpyDeviceToHost,
                        s_));
  sync();
  std::vector<double> rd(6 * (size_t)std::max(np, 1));
  for (int i = 0; i < np; i++) {  // MeshBlock.C:829-837
    Block &A = *blocks_[pairs_[i].local];
    double bb[6];
    for (int j = 0; j < 6; j++) bb[j] = dunkey(keys[6 * i + j]);
    for (int j = 0; j < 3; j++) {
      rd[6 * i + j] = A.obb.xc[j];
      for (int k = 0; k < 3; k++) rd[6 * i + j] += ((bb[k] + bb[k + 3]) * 0.5) * A.obb.vec[k][j];
      rd[6 * i + 3 + j] = (bb[j + 3] - bb[j]) * 0.51;
    }
  }
  pk_.pair_rd = rd;
  for (int i = 0; i < np; i++) {
    const int ii = recv_pair(i);
    if (ii < 0) continue;
    pairs_[ii].iblk_remote = pairs_[i].local;
    pairs_[ii].tag_remote = blocks_[pairs_[i].local]->tag;
    for (int j = 0; j < 3; j++) {
      pairs_[ii].obc.xc[j] = rd[6 * i + j];
      pairs_[ii].obc.dxc[j] = rd[6 * i + 3 + j];
    }
  }
}

template <class S>
const S *Engine::put_segs(DBuf<unsigned char> &b, const std::vector<S> &v)
{
  b.reserve(sizeof(S) * std::max<size_t>(v.size(), 1));
  if (!v.empty()) TG_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(S) * v.size(), hipMemcpyHostToDevice, s_));
  return (const S *)b.p;
}

// tioga::exchangeSearchData (exchangeSearchData.C:27-263), single rank
void Engine::exchange_search_data()
{
  const int np = (int)pairs_.size();
  for (Block *b : blocks_) b->nsearch = 0;
  if (nsend_ == 0) return;
  std::vector<int> cnt(np), dest(np, -1), cand0(np + 1), rng0(np + 1), lst0(np);
  std::vector<int2> rng;
  int ncand = 0;
  for (int i = 0; i < np; i++) {
    Block &A = *blocks_[pairs_[i].local];
    const OBB &obc = pairs_[i].obc;
    // bin range and bin-level test (MeshBlock.C:973-1063)
    double xv[8][3], xd[3], mdx[3], xmin[3], xmax[3];
    int imin[3], imax[3];
    obb_corners(obc.xc, obc.dxc, obc.vec, xv);
    for (int j = 0; j < 3; j++) {
      xmin[j] = BIGVALUE;
      xmax[j] = -BIGVALUE;
    }
    for (int n = 0; n < 8; n++) {
      to_frame(xv[n], A.obb.xc, A.obb.vec, xd);
      for (int j = 0; j < 3; j++) {
        xmin[j] = dmin(xmin[j], xd[j] + A.obb.dxc[j]);
        xmax[j] = dmax(xmax[j], xd[j] + A.obb.dxc[j]);
      }
    }
    for (int j = 0; j < 3; j++) {
      double delta = 0.01 * (xmax[j] - xmin[j]);
      xmin[j] -= delta;
      xmax[j] += delta;
      imin[j] = (int)dmax(xmin[j] / A.mapdx[j], 0);
      imax[j] = (int)dmin(xmax[j] / A.mapdx[j], 10 - 1);
      mdx[j] = 0.5 * A.mapdx[j];
    }
    obb_corners(obc.xc, mdx, A.obb.vec, xv);
    for (int j = 0; j < 3; j++) {
      xmin[j] = BIGVALUE;
      xmax[j] = -BIGVALUE;
    }
    for (int m = 0; m < 8; m++) {
      to_frame(xv[m], obc.xc, obc.vec, xd);
      for (int j = 0; j < 3; j++) {
        xmin[j] = dmin(xmin[j], xd[j]);
        xmax[j] = dmax(xmax[j], xd[j]);
      }
    }
    unsigned char mask[1000];
    memset(mask, 0, sizeof mask);
    for (int l = imin[2]; l <= imax[2]; l++)
      for (int k = imin[1]; k <= imax[1]; k++)
        for (int j = imin[0]; j <= imax[0]; j++) {
          double xc[3];
          xd[0] = -A.obb.dxc[0] + j * A.mapdx[0] + A.mapdx[0] * 0.5;
          xd[1] = -A.obb.dxc[1] + k * A.mapdx[1] + A.mapdx[1] * 0.5;
          xd[2] = -A.obb.dxc[2] + l * A.mapdx[2] + A.mapdx[2] * 0.5;
          for (int n = 0; n < 3; n++) {
            xc[n] = A.obb.xc[n];
            for (int ij = 0; ij < 3; ij++) xc[n] += (xd[ij] * A.obb.vec[ij][n]);
          }
          to_frame(xc, obc.xc, obc.vec, xd);
          bool out = false;
          for (int ij = 0; ij < 3 && !out; ij++) out = (xmin[ij] + xd[ij] > obc.dxc[ij]);
          if (out) continue;
          for (int ij = 0; ij < 3 && !out; ij++) out = (xmax[ij] + xd[ij] < -obc.dxc[ij]);
          if (out) continue;
          mask[l * 100 + k * 10 + j] = 1;
        }
    // candidates: the positions of the masked bins, bins in (l, k, j) order,
    // which is the order the positions already have
    rng0[i] = (int)rng.size();
    cand0[i] = ncand;
    for (int kb = 0; kb < 1000; kb++) {
      if (!mask[kb]) continue;
      const int p0 = A.binoff[kb], p1 = A.binoff[kb + 1];
      if (p1 <= p0) continue;
      rng.push_back(int2{ncand, p0});
      if (ncand > 0x7fffffff - (p1 - p0)) throw std::runtime_error("tioga_amd: too many query candidates");
      ncand += p1 - p0;
    }
  }
  rng0[np] = (int)rng.size();
  cand0[np] = ncand;
  // one flag kernel over the candidates of every pair, one compaction
  {
    QueryPairs qp{};
    DBuf<int> &dcand0 = stw_.cand0, &drng0 = stw_.rng0, &first = stw_.first;
    DBuf<int2> &drng = stw_.rng;
    dcand0.reserve(np + 1);
    drng0.reserve(np + 1);
    first.reserve(np + 1);
    drng.reserve(std::max<size_t>(rng.size(), 1));
    TG_HIP(hipMemcpyAsync(dcand0.p, cand0.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(drng0.p, rng0.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice, s_));
    if (!rng.empty())
      TG_HIP(hipMemcpyAsync(drng.p, rng.data(), sizeof(int2) * rng.size(), hipMemcpyHostToDevice, s_));
    qp.cand0 = dcand0.p;
    qp.rng = drng.p;
    qp.rng0 = drng0.p;
    for (int b = 0; b < (int)blocks_.size(); b++) {
      qp.invmap[b] = blocks_[b]->invmap.p;
      qp.x[b] = blocks_[b]->x.p;
    }
    std::vector<OBB> hobc(np);
    std::vector<int> hblk(np);
    for (int i = 0; i < np; i++) {
      hobc[i] = pairs_[i].obc;
      hblk[i] = pairs_[i].local;
    }
    DBuf<OBB> &dobc = stw_.obc;
    DBuf<int> &dblk = stw_.pblk;
    dobc.reserve(np);
    dblk.reserve(np);
    TG_HIP(hipMemcpyAsync(dobc.p, hobc.data(), sizeof(OBB) * np, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(dblk.p, hblk.data(), sizeof(int) * np, hipMemcpyHostToDevice, s_));
    qp.obc = dobc.p;
    qp.blk = dblk.p;
    QueryPairs *pq = (QueryPairs *)put_table(stw_.qpairs, qp);
    DBuf<int> &flag = stw_.qflag, &pos = stw_.qpos, &sel = stw_.qsel;
    flag.reserve(ncand);
    pos.reserve(ncand);
    sel.reserve(ncand);
    qlist_.reserve(ncand);
    int nsel = 0;
    if (ncand) {
      k_query_cand<<<nb(ncand), BS, 0, s_>>>(ncand, np, pq, flag.p, pos.p);
      nsel = select_flagged(flag.p, ncand, sel.p, s_);
    }
    std::vector<int> hf(np + 1, 0);
    if (nsel) {
      k_take_pos<<<nb(nsel), BS, 0, s_>>>(nsel, sel.p, pos.p, qlist_.p);
      k_pair_runs<<<nb(np + 1), BS, 0, s_>>>(np, nsel, sel.p, dcand0.p, first.p);
      TG_HIP(hipMemcpyAsync(hf.data(), first.p, sizeof(int) * (np + 1), hipMemcpyDeviceToHost, s_));
      sync();
    }
    for (int i = 0; i < np; i++) {
      lst0[i] = hf[i];
      cnt[i] = nsel ? hf[i + 1] - hf[i] : 0;
    }
  }
  for (int i = 0; i < np; i++) {
    const int ii = recv_pair(i);
    dest[i] = ii;
    if (ii >= 0) blocks_[pairs_[ii].local]->nsearch += cnt[i];
  }
  // the points, resolutions, mesh tags and representatives of all blocks'
  // query sets lie back to back in one allocation each, block after block,
  // so the batched dedupe works on them in place
  size_t qtot = 0;
  for (Block *bp : blocks_) qtot += bp->nsearch;
  qx_.reserve(3 * qtot);
  qres_.reserve(qtot);
  qtag_.reserve(qtot);
  qrep_.reserve(qtot);
  size_t qo = 0;
  for (Block *bp : blocks_) {
    Block &b = *bp;
    const size_t n = b.nsearch;  // views of exactly the set (a block without queries gets none)
    b.xsearch.view(qx_.p + 3 * qo, 3 * n);
    b.res_search.view(qres_.p + qo, n);
    b.tagsearch.view(qtag_.p + qo, n);
    b.xtag.view(qrep_.p + qo, n);
    qo += n;
    b.isearch.reserve(3 * std::max(n, (size_t)1));
    b.donorId.reserve(std::max(n, (size_t)1));
  }
  // every pair's query points in one launch (a segment per pair)
  std::vector<int> off(blocks_.size(), 0);
  std::vector<QFillSeg> qs;
  pk_.pair_cnt = cnt;
  pk_.pair_off.assign(np, -1);
  for (int i = 0; i < np; i++) {
    int ii = dest[i];
    if (ii >= 0) pk_.pair_off[i] = off[pairs_[ii].local];
    if (ii < 0 || cnt[i] == 0) continue;
    Block &A = *blocks_[pairs_[i].local];
    int rb = pairs_[ii].local;
    Block &R = *blocks_[rb];
    qs.push_back(QFillSeg{0, cnt[i], qlist_.p + lst0[i], A.invmap.p, A.x.p, A.nodeRes.p, off[rb],
                          pairs_[ii].iblk_remote, pairs_[ii].tag_remote, R.xsearch.p, R.res_search.p, R.isearch.p,
                          R.tagsearch.p});
    off[rb] += cnt[i];
  }
  const int totq = pack_segs(qs);
  if (totq) k_fill_search_all<<<totq / BS, BS, 0, s_>>>((int)qs.size(), put_segs(stw_.qfill, qs));
  sync();
  if (pk_.on) {  // the resolutions as sent, before the dedupe raises them
    pk_.sent_res.assign(blocks_.size(), std::vector<double>());
    for (size_t ib = 0; ib < blocks_.size(); ib++) {
      Block &b = *blocks_[ib];
      pk_.sent_res[ib].resize(b.nsearch);
      if (b.nsearch)
        TG_HIP(hipMemcpyAsync(pk_.sent_res[ib].data(), b.res_search.p, sizeof(double) * b.nsearch,
                              hipMemcpyDeviceToHost, s_));
    }
    sync();
  }
}

// MeshBlock::search (search.C:78-269), NODEGID off
void Engine::search(Block &b)
{
  const int Q = b.nsearch;
  b.nelem = 0;
  if (Q == 0) return;
  auto t0 = std::chrono::steady_clock::now();
  if (b.resident) {
    attach_resident(b);
  } else {
    OBB obq;
    const ObbSet os{Q, b.xsearch.p};
    long long st = 0;
    find_obbs_device(&os, 1, &obq, obbw_, s_, &st);
    obb_stepped_ += st;
    prep_acc_[0] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    build_search_structure(b, obq, b.nelem, b.adt_depth, b.elist, b.ebox, b.adtInt, b.adtReal, b.nodes, b.rcell,
                           b.adtExt, b.bins, b.nsearch);
  }
  bool gid_mode = true;
  for (Block *o : blocks_) gid_mode = gid_mode && (o->unode_gid != nullptr);
  if (gid_mode) {
    dedupe_gid(b);
  } else {
    auto td = std::chrono::steady_clock::now();
    unique_nodes_device(Q, b.xsearch.p, b.tagsearch.p, b.res_search.p, b.xtag.p, dedw_, s_);
    sync();
    prep_acc_[3] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - td).count();
  }
}

// MeshBlock::search (search.C:78-269) for every block at once: the query-set
// OBBs on the host (in parallel over blocks), the cell filter per block, one
// ADT build for all blocks' trees and one coincident-point dedupe for all
// query sets.  Each block's arrays come out as search() makes them.
void Engine::search_batched()
{
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](auto a, auto c) { return std::chrono::duration<double, std::milli>(c - a).count(); };
  const int nbk = (int)blocks_.size();
  std::vector<int> act;
  for (int ib = 0; ib < nbk; ib++) {
    blocks_[ib]->nelem = 0;
    if (blocks_[ib]->nsearch > 0) act.push_back(ib);
  }
  if (act.empty()) return;
  // the coincident-point dedupe of all query sets: on a stream of its own,
  // driven by a helper thread, while this thread runs the OBB batch, the
  // cell filter and the ADT build (the dedupe reads the query points and
  // writes xtag and the raised resolutions, which only the searches read;
  // those three stages never touch them)
  bool gid_mode = true;
  for (Block *o : blocks_) gid_mode = gid_mode && (o->unode_gid != nullptr);
  std::vector<DedupeSet> dsets;
  for (int ib : act) {
    Block &b = *blocks_[ib];
    dsets.push_back(DedupeSet{b.nsearch, b.xsearch.p, b.tagsearch.p, b.res_search.p, b.xtag.p});
  }
  double ded_ms = 0;
  std::exception_ptr ded_err;
  auto dedupe = [&](hipStream_t q) {
    try {
      auto a = now();
      unique_nodes_batch(dsets.data(), (int)dsets.size(), dedw_, q);
      TG_HIP(hipStreamSynchronize(q));
      ded_ms = ms(a, now());
    } catch (...) {
      ded_err = std::current_exception();
    }
  };
  std::thread ded_thread;
  if (gid_mode) {
    auto a = now();
    for (int ib : act) dedupe_gid(*blocks_[ib]);
    ded_ms = ms(a, now());
  } else if (overlap_dedupe) {
    // created on first use, so engines that never overlap hold one stream
    if (!s2_) TG_HIP(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking));
    ded_thread = std::thread(dedupe, s2_);
  } else {
    dedupe(s_);
  }
  std::vector<int> build;  // blocks whose tree is built for this query set
  for (int ib : act) {
    if (blocks_[ib]->resident) attach_resident(*blocks_[ib]);
    else build.push_back(ib);
  }
  try {
    if (!build.empty()) prepare_searches(build);
  } catch (...) {
    if (ded_thread.joinable()) ded_thread.join();
    throw;
  }
  if (ded_thread.joinable()) ded_thread.join();
  if (ded_err) std::rethrow_exception(ded_err);
  prep_acc_[3] += ded_ms;
}

// search_batched's OBB batch, cell filter and ADT build for the active
// blocks act
void Engine::prepare_searches(const std::vector<int> &act)
{
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](auto a, auto c) { return std::chrono::duration<double, std::milli>(c - a).count(); };
  const int nbk = (int)blocks_.size();
  // query-set OBBs, all sets in one device batch with the bits of the
  // sequential running sums
  auto t0 = now();
  std::vector<OBB> obq(nbk);
  {
    std::vector<ObbSet> os;
    std::vector<OBB> ob(act.size());
    for (int ib : act) os.push_back(ObbSet{blocks_[ib]->nsearch, blocks_[ib]->xsearch.p});
    long long st = 0;
    find_obbs_device(os.data(), (int)os.size(), ob.data(), obbw_, s_, &st);
    obb_stepped_ += st;
    for (size_t a = 0; a < act.size(); a++) obq[act[a]] = ob[a];
  }
  auto t1 = now();
  // cells overlapping each query OBB, as a LIFO list, and their boxes: the
  // flags of all blocks in one array (block a's cells from coff[a], each
  // block's reversed), one compaction, and each block's range of it
  std::vector<AdtTree> trees(act.size());
  const int na = (int)act.size();
  std::vector<int> coff(na + 1, 0), eoff(na + 1, 0);
  for (int a = 0; a < na; a++) {
    coff[a + 1] = coff[a] + blocks_[act[a]]->ncells;
    if (coff[a + 1] < coff[a]) throw std::runtime_error("tioga_amd: too many cells for one batched search");
  }
  const int totN = coff[na];
  iscratch_.reserve(std::max(totN, 1));
  iscratch2_.reserve(std::max(totN, 1));
  // the per-block table of the batched filter and element boxes
  std::vector<PrepBlk> pb(na);
  for (int a = 0; a < na; a++) {
    Block &b = *blocks_[act[a]];
    pb[a] = PrepBlk{b.ti, b.conn.p, b.x.p, obq[act[a]], coff[a], b.ncells, 0};
  }
  ptab_.reserve(sizeof(PrepBlk) * na);
  pseg_.reserve(2 * (size_t)(na + 1));
  TG_HIP(hipMemcpyAsync(ptab_.p, pb.data(), sizeof(PrepBlk) * na, hipMemcpyHostToDevice, s_));
  if (totN) k_cell_filter_all<<<nb(totN), BS, 0, s_>>>(na, totN, (const PrepBlk *)ptab_.p, iscratch_.p);
  const int totE = totN ? select_flagged(iscratch_.p, totN, iscratch2_.p, s_) : 0;
  if (totE) {
    int *dcoff = pseg_.p, *dstart = pseg_.p + (na + 1);
    TG_HIP(hipMemcpyAsync(dcoff, coff.data(), sizeof(int) * (na + 1), hipMemcpyHostToDevice, s_));
    k_seg_starts<<<nb(na + 1), BS, 0, s_>>>(na, totE, iscratch2_.p, dcoff, dstart);
    TG_HIP(hipMemcpyAsync(eoff.data(), dstart, sizeof(int) * (na + 1), hipMemcpyDeviceToHost, s_));
    sync();
  }
  // every tree's element boxes and ADT arrays back to back (views into one
  // allocation each, in the order of the compacted element list), so the
  // batched build works on them in place
  const size_t te = eoff[na];
  telist_.reserve(te);
  tebox_.reserve(6 * te);
  tint_.reserve(4 * te);
  treal_.reserve(6 * te);
  for (int a = 0; a < na; a++) {
    Block &b = *blocks_[act[a]];
    const int ne = eoff[a + 1] - eoff[a];
    b.nelem = ne;
    b.elist.view(telist_.p + eoff[a], ne);
    b.ebox.view(tebox_.p + 6 * (size_t)eoff[a], 6 * (size_t)ne);
    b.adtInt.view(tint_.p + 4 * (size_t)eoff[a], 4 * (size_t)ne);
    b.adtReal.view(treal_.p + 6 * (size_t)eoff[a], 6 * (size_t)ne);
    pb[a].eoff = eoff[a];
    trees[a].n = ne;
    trees[a].ebox = b.ebox.p;
    trees[a].adtInt = b.adtInt.p;
    trees[a].adtReal = b.adtReal.p;
  }
  if (te) {
    TG_HIP(hipMemcpyAsync(ptab_.p, pb.data(), sizeof(PrepBlk) * na, hipMemcpyHostToDevice, s_));
    k_elem_box_all<<<nb(te), BS, 0, s_>>>(na, (int)te, (const PrepBlk *)ptab_.p, iscratch2_.p, telist_.p, tebox_.p);
  }
  sync();
  auto t2 = now();
  build_adt_batch(trees.data(), (int)trees.size(), adtw_, s_);
  // the search records of every tree in one launch
  tnodes_.reserve(te);
  trcell_.reserve(te);
  std::vector<AdtNodesTree> nt(na);
  for (int a = 0; a < na; a++) {
    Block &b = *blocks_[act[a]];
    b.adt_depth = trees[a].depth;
    memcpy(b.adtExt, trees[a].ext, sizeof b.adtExt);
    if (b.adt_depth + 2 > STACK) throw std::runtime_error("tioga_amd: ADT deeper than the search stack");
    b.nodes.view(tnodes_.p + eoff[a], b.nelem);
    b.rcell.view(trcell_.p + eoff[a], b.nelem);
    nt[a] = AdtNodesTree{b.ti, grid_frame(trees[a].ext), eoff[a]};
  }
  if (te) {
    ntab_.reserve(sizeof(AdtNodesTree) * na);
    TG_HIP(hipMemcpyAsync(ntab_.p, nt.data(), sizeof(AdtNodesTree) * na, hipMemcpyHostToDevice, s_));
    k_adt_nodes_all<<<nb(te), BS, 0, s_>>>(na, (int)te, (const AdtNodesTree *)ntab_.p, tint_.p, treal_.p,
                                            tebox_.p, telist_.p, tnodes_.p, trcell_.p);
  }
  {  // the bin grids of every tree, one batch
    std::vector<const WalkNode *> np(na);
    std::vector<int> ne(na);
    std::vector<BinGrid *> bv(na);
    for (int a = 0; a < na; a++) {
      Block &b = *blocks_[act[a]];
      np[a] = b.nodes.p;
      ne[a] = BIN_ELEMENTS_PER_QUERY * b.nsearch >= b.nelem ? b.nelem : 0;  // as build_search_structure decides
      bv[a] = &b.bins;
    }
    build_bins(na, np.data(), ne.data(), tbins_, bv.data());
  }
  sync();
  auto t3 = now();
  prep_acc_[0] += ms(t0, t1);
  prep_acc_[1] += ms(t1, t2);
  prep_acc_[2] += ms(t2, t3);
}

// search.C:48-74: first occurrence of each node global id
void Engine::dedupe_gid(Block &b)
{
  const int Q = b.nsearch;
  std::vector<double> rs(Q);
  std::vector<int> xt(Q), is(3 * (size_t)Q);
  TG_HIP(hipMemcpyAsync(rs.data(), b.res_search.p, sizeof(double) * Q, hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(is.data(), b.isearch.p, sizeof(int) * 3 * Q, hipMemcpyDeviceToHost, s_));
  sync();
  std::unordered_map<uint64_t, int> first;
  first.reserve(2 * (size_t)Q);
  for (int i = 0; i < Q; i++) {
    uint64_t g = blocks_[is[3 * i + 2]]->unode_gid[is[3 * i + 1]];
    auto it = first.find(g);
    if (it != first.end()) {
      xt[i] = it->second;
      rs[it->second] = std::max(rs[it->second], rs[i]);
    } else {
      first[g] = i;
      xt[i] = i;
    }
  }
  for (int i = 0; i < Q; i++) rs[i] = rs[xt[i]];
  TG_HIP(hipMemcpyAsync(b.xtag.p, xt.data(), sizeof(int) * Q, hipMemcpyHostToDevice, s_));
  TG_HIP(hipMemcpyAsync(b.res_search.p, rs.data(), sizeof(double) * Q, hipMemcpyHostToDevice, s_));
  sync();
}

// The search structure over every cell of block btag: the cells overlapping
// the block's own OBB (all of them) in the reference's LIFO order, their
// ADT and walk records.
void Engine::build_resident_tree(int btag, ResidentTree &t)
{
  Block *b = find(btag);
  if (!b || !b->topo_uploaded || b->cellRes.p == nullptr) throw std::runtime_error("tioga_amd: preprocess first");
  build_search_structure(*b, b->obb, t.nelem, t.depth, t.elist, t.ebox, t.adtInt, t.adtReal, t.nodes, t.rcell,
                         t.ext, t.bins, -1);
  if (t.nelem != b->ncells) throw std::runtime_error("tioga_amd: resident tree misses cells of its block");
}

void Engine::use_resident_tree(int btag, const ResidentTree *t)
{
  Block *b = find(btag);
  if (!b) throw std::runtime_error("tioga_amd: resident tree for an unknown tag");
  b->resident = t;
  connected_ = false;
}

// a block's search arrays as views of its resident tree
void Engine::attach_resident(Block &b)
{
  const ResidentTree &t = *b.resident;
  b.nelem = t.nelem;
  b.adt_depth = t.depth;
  memcpy(b.adtExt, t.ext, sizeof b.adtExt);
  b.elist.view(t.elist.p, t.nelem);
  b.ebox.view(t.ebox.p, 6 * (size_t)t.nelem);
  b.adtInt.view(t.adtInt.p, 4 * (size_t)t.nelem);
  b.adtReal.view(t.adtReal.p, 6 * (size_t)t.nelem);
  b.nodes.view(t.nodes.p, t.nelem);
  b.rcell.view(t.rcell.p, t.nelem);
  b.bins.shift = t.bins.shift;
  b.bins.nbins = t.bins.nbins;
  if (t.bins.nbins) {
    b.bins.off.view(t.bins.off.p, (size_t)t.bins.nbins + 1);
    b.bins.ent.view(t.bins.ent.p, t.bins.ent.cap);
  }
}

// The bin grids of ntree trees (walk records nodes[a], nelem[a] of them),
// built into store's arrays; views[a] gets tree a's grid (views of store's
// arrays unless views[a] is store itself).  A tree's bins per axis are the
// power of two nearest cbrt(nelem / tune.bins), made coarser (up to 3 times)
// while its entries exceed 32 per element.  The (bin, element) pairs of all
// trees are sorted by bin (a stable radix sort, so each bin lists its
// elements in node order) and each bin's start found by binary search: no
// contended atomics.  tune.bins == 0 builds none.
void Engine::build_bins(int ntree, const WalkNode *const *nodes, const int *nelem, BinGrid &store,
                        BinGrid *const *views)
{
  for (int a = 0; a < ntree; a++) views[a]->nbins = 0;
  const int per = sw_.tune.bins;
  std::vector<BinTree> T(ntree);
  long long total = 0;
  for (int a = 0; a < ntree; a++) {
    const int n = nelem[a];
    const double r = per > 0 && n > 0 ? std::cbrt((double)n / per) : 1.0;
    const int k = std::max(0, std::min(8, (int)std::lround(std::log2(std::max(r, 1.0)))));
    T[a] = BinTree{nodes[a], n, 16 - k, 0, (int)total};
    total += n;
  }
  if (per <= 0 || total == 0) return;
  if (total >= 0x7fffffffLL) throw std::runtime_error("tioga_amd: too many tree elements for one bin batch");
  btab_.reserve(sizeof(BinTree) * ntree);
  bsums_.reserve(ntree);
  bcnt_.reserve((size_t)total + 1);
  beoff_.reserve((size_t)total + 1);
  std::vector<long long> tot(ntree);
  long long nent = 0;
  for (int round = 0;; round++) {  // coarser bins for the trees whose entries exceed the cap
    TG_HIP(hipMemcpyAsync(btab_.p, T.data(), sizeof(BinTree) * ntree, hipMemcpyHostToDevice, s_));
    fill<int>(bcnt_.p + total, 1, 0, s_);
    k_bin_nb<<<nb(total), BS, 0, s_>>>(ntree, (int)total, (const BinTree *)btab_.p, bcnt_.p);
    exclusive_sum(bcnt_.p, beoff_.p, (int)total + 1, s_);
    k_bin_totals<<<nb(ntree), BS, 0, s_>>>(ntree, (const BinTree *)btab_.p, beoff_.p, (long long *)bsums_.p);
    TG_HIP(hipMemcpyAsync(tot.data(), bsums_.p, sizeof(long long) * ntree, hipMemcpyDeviceToHost, s_));
    sync();
    bool again = false;
    nent = 0;
    for (int a = 0; a < ntree; a++) {
      nent += tot[a];
      const long long cap = 32LL * T[a].nelem + 65536;
      if (tot[a] > cap && T[a].shift < 16 && round < 3) {
        T[a].shift++;
        again = true;
      }
    }
    if (!again) break;
  }
  long long nbin = 0;
  for (int a = 0; a < ntree; a++) {
    const long long R = 65536 >> T[a].shift;
    T[a].bbase = nbin;
    nbin += R * R * R;
  }
  if (nbin + 1 > 0x7fffffffLL || nent >= 0x7fffffffLL) throw std::runtime_error("tioga_amd: bin grids too large");
  TG_HIP(hipMemcpyAsync(btab_.p, T.data(), sizeof(BinTree) * ntree, hipMemcpyHostToDevice, s_));
  const size_t ne = (size_t)std::max<long long>(nent, 1);
  bkey_.reserve(ne);
  bkey2_.reserve(ne);
  bval_.reserve(ne);
  bval2_.reserve(ne);
  store.off.reserve_own((size_t)nbin + 1);
  store.ent.reserve_own(ne);
  k_bin_pairs<<<nb(total), BS, 0, s_>>>(ntree, (int)total, (const BinTree *)btab_.p, beoff_.p, bkey_.p, bval_.p);
  int bits = 1;
  while ((1LL << bits) <= nbin) bits++;
  sort_pairs(bkey_.p, bkey2_.p, bval_.p, bval2_.p, (int)nent, bits, s_);
  k_bin_offsets<<<nb(nbin + 1), BS, 0, s_>>>((int)nbin, (int)nent, bkey2_.p, store.off.p);
  k_bin_entries<<<nb(nent), BS, 0, s_>>>(ntree, (int)nent, (const BinTree *)btab_.p, bval2_.p, store.ent.p);
  for (int a = 0; a < ntree; a++) {
    BinGrid &v = *views[a];
    if (T[a].nelem == 0) continue;
    const long long R = 65536 >> T[a].shift;
    v.shift = T[a].shift;
    v.nbins = R * R * R;
    if (&v != &store) {
      v.off.view(store.off.p + T[a].bbase, (size_t)v.nbins + 1);
      v.ent.view(store.ent.p, store.ent.cap);
    }
  }
}

// search.C:107-233: query-set OBB -> cells whose OBB-frame box overlaps it,
// as a LIFO list (descending cell index) -> element boxes -> ADT
void Engine::build_search_structure(Block &b, const OBB &obq, int &nelem, int &depth, DBuf<int> &elist,
                                    DBuf<double> &ebox, DBuf<int> &adtInt, DBuf<double> &adtReal,
                                    DBuf<WalkNode> &nodes, DBuf<int> &rcell, double ext[6], BinGrid &bins,
                                    long long nq)
{
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](auto a, auto c) { return std::chrono::duration<double, std::milli>(c - a).count(); };
  auto t1 = now();
  const int N = b.ncells;
  iscratch_.reserve(N);
  iscratch2_.reserve(N);
  if (N) k_cell_filter<<<nb(N), BS, 0, s_>>>(N, b.ti, b.conn.p, b.x.p, obq, iscratch_.p);
  int ne = select_flagged(iscratch_.p, N, iscratch2_.p, s_);
  nelem = ne;
  elist.reserve_own(ne);
  ebox.reserve_own(6 * (size_t)ne);
  adtInt.reserve_own(4 * (size_t)ne);
  adtReal.reserve_own(6 * (size_t)ne);
  if (ne) k_elem_box<<<nb(ne), BS, 0, s_>>>(ne, N, iscratch2_.p, 0, b.ti, b.conn.p, b.x.p, elist.p, ebox.p);
  sync();
  auto t2 = now();
  depth = build_adt_device(ne, ebox.p, adtInt.p, adtReal.p, ext, adtw_, s_);
  if (depth + 2 > STACK) throw std::runtime_error("tioga_amd: ADT deeper than the search stack");
  nodes.reserve_own(ne);
  rcell.reserve_own(ne);
  if (ne)
    k_adt_nodes<<<nb(ne), BS, 0, s_>>>(ne, b.ti, grid_frame(ext), adtInt.p, adtReal.p, ebox.p, elist.p, nodes.p,
                                       rcell.p);
  // bins when the tree serves many searches (nq < 0), or one search with at
  // least one query per BIN_ELEMENTS_PER_QUERY elements; else its queries walk it
  const WalkNode *np = nodes.p;
  BinGrid *bp = &bins;
  const int nbin_el = (nq < 0 || BIN_ELEMENTS_PER_QUERY * nq >= ne) ? ne : 0;
  build_bins(1, &np, &nbin_el, bins, &bp);
  sync();
  auto t3 = now();
  prep_acc_[1] += ms(t1, t2);
  prep_acc_[2] += ms(t2, t3);
}

int Engine::query_set_create(int btag, const double *xq, long nq)
{
  int ib;
  Block *b = find(btag, &ib);
  if (!b) throw std::runtime_error("tioga_amd: query set for an unknown tag");
  if (!b->topo_uploaded || b->cellRes.p == nullptr) throw std::runtime_error("tioga_amd: preprocess first");
  if (nq <= 0 || nq > 0x7fffffffL) throw std::runtime_error("tioga_amd: query count out of range");
  QuerySet *q = new QuerySet;
  q->blk = ib;
  q->nq = (int)nq;
  q->xq.reserve(3 * (size_t)nq);
  q->donor.reserve(nq);
  TG_HIP(hipMemcpyAsync(q->xq.p, xq, sizeof(double) * 3 * nq, hipMemcpyHostToDevice, s_));
  OBB obq;
  find_obb(xq, nullptr, (int)nq, obq);
  build_search_structure(*b, obq, q->nelem, q->adt_depth, q->elist, q->ebox, q->adtInt, q->adtReal, q->nodes,
                         q->rcell, q->ext, q->bins, -1);
  qsets_.push_back(q);
  return (int)qsets_.size() - 1;
}

// Morton keys of the sets M (all queries in one buffer, set k at M[k].off),
// the walk/skip partition and one sort (with no Morton bits, a compaction of
// the walking queries in index order): returns the processing order (set
// k's at the same offset; only its first nin[k] positions are read) and
// leaves each set's count of walking queries in sw_.nin; donors of points
// outside their tree's root box are written.
const int *Engine::order_sets(const std::vector<MortonSet> &M, size_t tot)
{
  const int ns = (int)M.size();
  if (tot > 0x7fffffff) throw std::runtime_error("tioga_amd: too many queries in one search launch");
  int rb = 0;
  while ((1 << rb) < ns) rb++;
  // Morton bits (sw_.tune.morton_bits, 0-30; 21 = 7 per axis, about one
  // cell of the tree per step, three 8-bit sort passes for a few sets):
  // the order only affects speed
  const int mb = std::min(std::max(sw_.tune.morton_bits, 0), 31 - rb);
  if (31 - rb < 12) throw std::runtime_error("tioga_amd: too many query sets in one search launch");
  mv1_.reserve(tot);
  sw_.nin.reserve(ns);
  mset_.reserve(sizeof(MortonSet) * ns);
  TG_HIP(hipMemcpyAsync(mset_.p, M.data(), sizeof(MortonSet) * ns, hipMemcpyHostToDevice, s_));
  sw_.wstart.reserve(ns + 1);
  sw_.qsid.reserve(tot);
  if (mb == 0) {  // no Morton order: a compaction
    const size_t nw = (tot + 63) / 64;
    wmask_.reserve(nw);
    wcnt_.reserve(nw + 1);
    wpre_.reserve(nw + 1);
    fill<int>(wcnt_.p + nw, 1, 0, s_);
    k_walk_mask<<<nb((long long)tot), BS, 0, s_>>>(ns, (int)tot, (const MortonSet *)mset_.p, wmask_.p, wcnt_.p);
    exclusive_sum(wcnt_.p, wpre_.p, (int)nw + 1, s_);
    k_walk_scatter<<<(unsigned)std::min<long long>(nb((long long)tot), 4096), BS, 0, s_>>>(
        ns, (int)tot, (const MortonSet *)mset_.p, wmask_.p, wpre_.p, mv1_.p, sw_.qsid.p);
    k_walk_counts_mask<<<1, 256, 0, s_>>>(ns, (int)tot, (const MortonSet *)mset_.p, wmask_.p, wpre_.p,
                                          sw_.nin.p, sw_.wstart.p);
    return mv1_.p;
  }
  mk0_.reserve(tot);
  mk1_.reserve(tot);
  mv0_.reserve(tot);
  k_morton_sets<<<nb((long long)tot), BS, 0, s_>>>(ns, (int)tot, (const MortonSet *)mset_.p, mb, mk0_.p, mv0_.p,
                                                   sw_.qsid.p);
  sort_pairs(mk0_.p, mk1_.p, mv0_.p, mv1_.p, (int)tot, mb + 1 + rb, s_);
  k_walk_counts<<<1, 256, 0, s_>>>(ns, (const MortonSet *)mset_.p, mb, mk1_.p, sw_.nin.p, sw_.wstart.p);
  return mv1_.p;
}

// Searches the query sets qs in one launch of each search kernel (sets of
// uneven sizes, each with its own tree, side by side); the HIP events of a
// timed pass bracket the search kernels.
void Engine::launch_query_sets(const std::vector<QuerySet *> &qs)
{
  std::vector<QuerySet *> act;
  std::vector<MortonSet> ms;
  size_t tot = 0;
  for (QuerySet *q : qs) {
    if (q->nelem == 0) {  // no cell overlaps the set: every donor -1
      fill<int>(q->donor.p, q->nq, -1, s_);
      continue;
    }
    MortonSet m{q->xq.p, nullptr, q->donor.p, {}, (int)tot, q->nq};
    memcpy(m.ext, q->ext, sizeof m.ext);
    ms.push_back(m);
    act.push_back(q);
    tot += q->nq;
  }
  if (act.empty()) {
    if (kernel_events_) {
      TG_HIP(hipEventRecord((*kernel_events_)[event_slot_++], s_));
      TG_HIP(hipEventRecord((*kernel_events_)[event_slot_++], s_));
    }
    return;
  }
  const int *perm = order_sets(ms, tot);
  std::vector<SearchView> sv;
  for (size_t k = 0; k < act.size(); k++) {
    QuerySet &q = *act[k];
    Block &b = *blocks_[q.blk];
    SearchView S;
    S.nq = q.nq;
    S.perm = perm + ms[k].off;
    S.xs = q.xq.p;
    S.xtag = nullptr;
    S.donorId = q.donor.p;
    S.nodes = q.nodes.p;
    S.rcell = q.rcell.p;
    view_bins(S, q.bins);
    S.g = grid_frame(q.ext);
    memcpy(S.ext, q.ext, sizeof S.ext);
    S.conn = b.conn.p;
    S.x = b.x.p;
    S.cellRes = b.cellRes.p;
    sv.push_back(S);
  }
  if (kernel_events_) TG_HIP(hipEventRecord((*kernel_events_)[event_slot_++], s_));
  if (timing_ && search_graph) {
    // the unchecked launch of a timed pass, recorded once into a graph (its
    // tables are the checked pass's, on the device already) and replayed
    if (!qs_exec_) {
      sw_.skip_uploads = true;
      TG_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
      run_search(sv.data(), (int)sv.size(), sw_, s_, false);
      const hipError_t e = hipStreamEndCapture(s_, &qs_graph_);
      sw_.skip_uploads = false;
      TG_HIP(e);
      TG_HIP(hipGraphInstantiate(&qs_exec_, qs_graph_, nullptr, nullptr, 0));
    }
    TG_HIP(hipGraphLaunch(qs_exec_, s_));
  } else {
    run_search(sv.data(), (int)sv.size(), sw_, s_, !timing_);
  }
  if (kernel_events_) TG_HIP(hipEventRecord((*kernel_events_)[event_slot_++], s_));
}

void Engine::drop_search_graph()
{
  if (qs_exec_) (void)hipGraphExecDestroy(qs_exec_);
  if (qs_graph_) (void)hipGraphDestroy(qs_graph_);
  qs_exec_ = nullptr;
  qs_graph_ = nullptr;
}

void Engine::set_search_bins(int bins)
{
  if (bins < 0 || bins > 4096) throw std::runtime_error("tioga_amd: bin density must be 0..4096");
  if (bins == sw_.tune.bins) return;
  sw_.tune.bins = bins;
  for (QuerySet *q : qsets_) {
    const WalkNode *np = q->nodes.p;
    BinGrid *bp = &q->bins;
    build_bins(1, &np, &q->nelem, q->bins, &bp);
  }
  sync();
}

void Engine::search_stats(long long out[10])
{
  SearchScratch &w = sw_;
  for (int k = 0; k < 10; k++) out[k] = 0;
  out[3] = w.last_nqt;
  out[9] = w.last_bins ? 1 : 0;
  if (w.last_nqt <= 0 || w.last_ns <= 0) return;
  int items = 0, h[4] = {0, 0, 0, 0};
  std::vector<int> used(NRT * USED_STRIDE);
  std::vector<int> nin(w.last_ns);
  DBuf<int> hist;
  hist.reserve(4);
  fill<int>(hist.p, 4, 0, s_);
  const CandLayout Y = cand_layout(w.tcap);
  CandOut O{w.cell.p, w.query.p, w.status.p, w.counters.p, {}, {}, w.best.p, w.last.p, nullptr};
  for (int t = 0; t < NTYPE; t++) {
    O.rcap[t] = Y.rcap[t];
    O.tbase[t] = Y.tbase[t];
  }
  k_status_hist<<<STATS_GRID, 256, 0, s_>>>(O, hist.p);
  DBuf<unsigned long long> ah;
  ah.reserve(1);
  fill<unsigned long long>(ah.p, 1, 0ull, s_);
  k_after_hit<<<STATS_GRID, 256, 0, s_>>>(O, ah.p);
  unsigned long long after = 0;
  TG_HIP(hipMemcpyAsync(&after, ah.p, sizeof after, hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(used.data(), w.counters.p, sizeof(int) * used.size(), hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(h, hist.p, sizeof h, hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(nin.data(), w.nin.p, sizeof(int) * w.last_ns, hipMemcpyDeviceToHost, s_));
  if (w.last_bins) {
    TG_HIP(hipMemcpyAsync(&items, w.segstart.p + w.last_nwave, sizeof items, hipMemcpyDeviceToHost, s_));
  }
  TG_HIP(hipStreamSynchronize(s_));
  out[0] = cand_held(used.data(), Y);
  long long walking = 0;
  for (int v : nin) walking += v;
  out[1] = w.last_bins ? items : (walking + WAVE - 1) / WAVE;
  out[2] = walking;
  for (int k = 0; k < 4; k++) out[4 + k] = h[k];
  out[8] = (long long)after;
}

void Engine::query_set_search(int id, int *donor)
{
  if (id < 0 || id >= (int)qsets_.size()) throw std::runtime_error("tioga_amd: bad query set id");
  QuerySet &q = *qsets_[id];
  launch_query_sets({&q});
  TG_HIP(hipMemcpyAsync(donor, q.donor.p, sizeof(int) * q.nq, hipMemcpyDeviceToHost, s_));
  sync();
}

void Engine::query_set_donors(int id, int *donor)
{
  if (id < 0 || id >= (int)qsets_.size()) throw std::runtime_error("tioga_amd: bad query set id");
  QuerySet &q = *qsets_[id];
  TG_HIP(hipMemcpyAsync(donor, q.donor.p, sizeof(int) * q.nq, hipMemcpyDeviceToHost, s_));
  sync();
}

int Engine::query_set_extents(int id, double ext[6]) const
{
  if (id < 0 || id >= (int)qsets_.size()) throw std::runtime_error("tioga_amd: bad query set id");
  memcpy(ext, qsets_[id]->ext, 6 * sizeof(double));
  return qsets_[id]->nelem;
}

double Engine::time_query_sets(int nrep, long long *nq)
{
  long long n = 0;
  for (QuerySet *q : qsets_) n += q->nq;
  if (nq) *nq = n;
  if (nrep <= 0 || qsets_.empty()) return 0.0;
  // events around every pass's search launch give the kernels' own duration
  std::vector<hipEvent_t> ev(2 * nrep);
  for (auto &e : ev) TG_HIP(hipEventCreate(&e));
  // establish the candidate-list capacity with one checked pass; the graph
  // of the timed launches is recorded after it (buffers may have grown)
  drop_search_graph();
  launch_query_sets(qsets_);
  sync();
  kernel_events_ = &ev;
  timing_ = true;
  Timer tm(s_);
  for (int r = 0; r < nrep; r++) launch_query_sets(qsets_);
  double pass = tm.stop() / nrep;
  timing_ = false;
  kernel_events_ = nullptr;
  double k = 0;
  for (size_t i = 0; i + 1 < ev.size(); i += 2) {
    float m = 0;
    TG_HIP(hipEventElapsedTime(&m, ev[i], ev[i + 1]));
    k += m;
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  last_search_kernel_ms_ = k / nrep;
  event_slot_ = 0;
  drop_search_graph();
  return pass;
}

void Engine::query_set_clear()
{
  for (QuerySet *q : qsets_) delete q;
  qsets_.clear();
}

void Engine::launch_search(Block &b)
{
  if (b.nsearch == 0) return;
  if (b.nelem == 0) {
    fill<int>(b.donorId.p, b.nsearch, -1, s_);
    return;
  }
  MortonSet ms{b.xsearch.p, b.xtag.p, b.donorId.p, {}, 0, b.nsearch};
  memcpy(ms.ext, b.adtExt, sizeof ms.ext);
  if (b.resident) {
    tieband_.reserve(1);
    fill<int>(tieband_.p, 1, 0, s_);
  }
  SearchView S;
  S.tieband = b.resident ? tieband_.p : nullptr;
  S.nq = b.nsearch;
  S.perm = order_sets({ms}, b.nsearch);
  S.xs = b.xsearch.p;
  S.xtag = b.xtag.p;
  S.donorId = b.donorId.p;
  S.nodes = b.nodes.p;
  S.rcell = b.rcell.p;
  view_bins(S, b.bins);
  S.g = grid_frame(b.adtExt);
  memcpy(S.ext, b.adtExt, sizeof S.ext);
  S.conn = b.conn.p;
  S.x = b.x.p;
  S.cellRes = b.cellRes.p;
  run_search(&S, 1, sw_, s_, true);
  k_copy_dup<<<nb(b.nsearch), BS, 0, s_>>>(b.nsearch, b.xtag.p, b.donorId.p);
  if (b.resident) {
    int tb = 0;
    TG_HIP(hipMemcpyAsync(&tb, tieband_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
    sync();
    last_tieband_ += tb;
  }
}

// The candidates the device search lists for block b's query set (after
// search(b) built its tree): per candidate its search point, preorder rank
// and cell, for a containment test outside the engine (the high-order
// callbacks).  Points outside the tree's root box and coincident copies list
// none.
void Engine::list_candidates(Block &b, std::vector<int> &qi, std::vector<int> &rank, std::vector<int> &cell)
{
  qi.clear();
  rank.clear();
  cell.clear();
  if (b.nsearch == 0 || b.nelem == 0) return;
  MortonSet ms{b.xsearch.p, b.xtag.p, b.donorId.p, {}, 0, b.nsearch};
  memcpy(ms.ext, b.adtExt, sizeof ms.ext);
  SearchView S;
  S.nq = b.nsearch;
  S.perm = order_sets({ms}, b.nsearch);
  S.xs = b.xsearch.p;
  S.xtag = b.xtag.p;
  S.donorId = b.donorId.p;
  S.nodes = b.nodes.p;
  S.rcell = b.rcell.p;
  view_bins(S, b.bins);
  S.g = grid_frame(b.adtExt);
  memcpy(S.ext, b.adtExt, sizeof S.ext);
  S.conn = b.conn.p;
  S.x = b.x.p;
  S.cellRes = b.cellRes.p;
  run_search(&S, 1, sw_, s_, true, true);
  const CandLayout Y = cand_layout(sw_.tcap);
  std::vector<int> counts(NRT * USED_STRIDE);
  TG_HIP(hipMemcpyAsync(counts.data(), sw_.counters.p, sizeof(int) * counts.size(), hipMemcpyDeviceToHost, s_));
  sync();
  const int used = (int)cand_held(counts.data(), Y);
  if (used <= 0) return;
  DBuf<int> dq, dr, dc;
  dq.reserve(used);
  dr.reserve(used);
  dc.reserve(used);
  const SearchSet T{(const SearchView *)sw_.sv.p, sw_.qb.p, sw_.nin.p, sw_.wstart.p, sw_.qsid.p, 1, b.nsearch};
  CandOut O{sw_.cell.p, sw_.query.p, sw_.status.p, sw_.counters.p, {}, {}, sw_.best.p, sw_.last.p, nullptr};
  for (int t = 0; t < NTYPE; t++) {
    O.rcap[t] = Y.rcap[t];
    O.tbase[t] = Y.tbase[t];
  }
  k_export_cands<<<nb((long long)Y.slots), BS, 0, s_>>>(T, O, dq.p, dr.p, dc.p);
  qi.resize(used);
  rank.resize(used);
  cell.resize(used);
  TG_HIP(hipMemcpyAsync(qi.data(), dq.p, sizeof(int) * used, hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(rank.data(), dr.p, sizeof(int) * used, hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(cell.data(), dc.p, sizeof(int) * used, hipMemcpyDeviceToHost, s_));
  sync();
}

// every block's donor search in one launch of each search kernel: each
// block's queries in Morton order within its own range of one buffer
void Engine::launch_searches()
{
  std::vector<Block *> act;
  size_t tot = 0;
  for (Block *b : blocks_) {
    if (b->nsearch == 0) continue;
    if (b->nelem == 0) {
      fill<int>(b->donorId.p, b->nsearch, -1, s_);
      continue;
    }
    act.push_back(b);
    tot += b->nsearch;
  }
  if (act.empty()) return;
  // Morton order inside each set, all sets in one sort (the order only
  // affects speed, each query's walk is independent)
  const int ns = (int)act.size();
  std::vector<MortonSet> ms(ns);
  size_t off = 0;
  for (int k = 0; k < ns; k++) {
    Block &b = *act[k];
    ms[k] = MortonSet{b.xsearch.p, b.xtag.p, b.donorId.p, {}, (int)off, b.nsearch};
    memcpy(ms[k].ext, b.adtExt, sizeof ms[k].ext);
    off += b.nsearch;
  }
  const int *perm = order_sets(ms, tot);
  bool any_resident = false;
  for (Block *b : act) any_resident = any_resident || b->resident;
  if (any_resident) {
    tieband_.reserve(1);
    fill<int>(tieband_.p, 1, 0, s_);
  }
  std::vector<SearchView> sv;
  for (int k = 0; k < ns; k++) {
    Block &b = *act[k];
    SearchView S;
    S.tieband = b.resident ? tieband_.p : nullptr;
    S.nq = b.nsearch;
    S.perm = perm + ms[k].off;
    S.xs = b.xsearch.p;
    S.xtag = b.xtag.p;
    S.donorId = b.donorId.p;
    S.nodes = b.nodes.p;
    S.rcell = b.rcell.p;
    view_bins(S, b.bins);
    S.g = grid_frame(b.adtExt);
    memcpy(S.ext, b.adtExt, sizeof S.ext);
    S.conn = b.conn.p;
    S.x = b.x.p;
    S.cellRes = b.cellRes.p;
    sv.push_back(S);
  }
  run_search(sv.data(), ns, sw_, s_, true);
  k_copy_dup_sets<<<nb(tot), BS, 0, s_>>>((const SearchView *)sw_.sv.p, sw_.qb.p, ns, (int)tot);
  last_tieband_ = 0;
  if (any_resident) {
    int tb = 0;
    TG_HIP(hipMemcpyAsync(&tb, tieband_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
    sync();
    last_tieband_ = tb;
  }
}

// tioga::exchangeDonors (exchangeDonors.C:25-301), single rank.  Every
// stage is one launch over all blocks (per-block segment tables).
void Engine::exchange_donors()
{
  const int nbk = (int)blocks_.size();
  if (nsend_ == 0 || nbk == 0) return;
  const int nwords = (nmesh_ + 63) / 64;
  DonorScratch &w = donw_;
  constexpr int FCH = 1 << 16;  // elements per fill chunk
  // fills in three batches, each its own host table (alive until the end)
  std::vector<FillChunk> fills[3];
  int fb = 0;
  auto add_fill = [&](void *p, size_t n, unsigned long long v, int width) {
    for (size_t st = 0; st < n; st += FCH)
      fills[fb].push_back(FillChunk{p, v, (long long)st, (int)std::min<size_t>(FCH, n - st), width});
  };
  auto run_fills = [&](DBuf<unsigned char> &tab) {
    const std::vector<FillChunk> &fv = fills[fb++];
    if (fv.empty()) return;
    const FillChunk *dc = put_segs(tab, fv);
    k_fill_chunks<<<(unsigned)fv.size(), BS, 0, s_>>>(dc);
  };
  RecvTables rt{};
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    const int V = b.nnodes;
    b.has_donor.reserve(V);
    b.rmax.reserve(V);
    b.head_key.reserve(V);
    b.head_arr.reserve(V);
    b.head_blk.reserve(V);
    b.head_remote.reserve(V);
    b.tagmask.reserve((size_t)V * nwords);
    add_fill(b.has_donor.p, V, 0, 4);
    add_fill(b.rmax.p, V, 0ull, 8);
    add_fill(b.head_key.p, V, ~0ull, 8);
    add_fill(b.head_arr.p, V, (unsigned)BIGINT, 4);
    add_fill(b.head_blk.p, V, 0xffffffffu, 4);
    add_fill(b.head_remote.p, V, 0xffffffffu, 4);
    add_fill(b.tagmask.p, (size_t)V * nwords, 0ull, 8);
    rt.has[ib] = b.has_donor.p;
    rt.rmax[ib] = b.rmax.p;
    rt.hkey[ib] = b.head_key.p;
    rt.tmask[ib] = b.tagmask.p;
    rt.harr[ib] = b.head_arr.p;
    rt.hblk[ib] = b.head_blk.p;
    rt.hrem[ib] = b.head_remote.p;
  }
  run_fills(w.tab[0]);
  RecvTables *prt = (RecvTables *)put_table(w.rt, rt);
  // comm1: donor records, arrival order = (donor block, search index)
  std::vector<DonorSeg> dseg;
  int arrival = 0;
  for (int ib = 0; ib < nbk; ib++) {
    Block &D = *blocks_[ib];
    const DonorSrc src{D.nsearch, arrival, ib, D.tag, D.donorId.p, D.isearch.p, D.xtag.p, D.cellRes.p,
                       D.res_search.p};
    if (D.nsearch) dseg.push_back(DonorSeg{src, 0, D.nsearch});
    arrival += D.nsearch;
  }
  if (!dseg.empty()) {
    const int tot = pack_segs(dseg);
    const DonorSeg *dt = put_segs(w.tab[1], dseg);
    const int ns = (int)dseg.size();
    k_donor_pass1_all<<<tot / BS, BS, 0, s_>>>(ns, dt, nwords, prt);
    k_donor_pass2_all<<<tot / BS, BS, 0, s_>>>(ns, dt, prt);
    k_donor_pass3_all<<<tot / BS, BS, 0, s_>>>(ns, dt, prt);
  }
  // processDonors for every block; records in (receptor block, node) order
  std::vector<int> blk_tag(nbk);
  for (int ib = 0; ib < nbk; ib++) blk_tag[ib] = blocks_[ib]->tag;
  DBuf<int> &dtag = w.dtag;
  dtag.reserve(nbk);
  TG_HIP(hipMemcpyAsync(dtag.p, blk_tag.data(), sizeof(int) * nbk, hipMemcpyHostToDevice, s_));
  HoleSet H{nmesh_, hm_view_.p};
  std::vector<int> noff(nbk + 1, 0);
  for (int ib = 0; ib < nbk; ib++) {
    noff[ib + 1] = noff[ib] + blocks_[ib]->nnodes;
    if (noff[ib + 1] < noff[ib]) throw std::runtime_error("tioga_amd: too many nodes for one donor exchange");
  }
  const int totV = noff[nbk];
  DBuf<int> &pflag = w.pflag, &psel = w.psel, &dnoff = w.dnoff;
  pflag.reserve(std::max(totV, 1));
  psel.reserve(std::max(totV, 1));
  dnoff.reserve(nbk + 1);
  TG_HIP(hipMemcpyAsync(dnoff.p, noff.data(), sizeof(int) * (nbk + 1), hipMemcpyHostToDevice, s_));
  EmitTables et{};
  et.nbk = nbk;
  et.noff = dnoff.p;
  // the hole-cut and fringe marks of every block in one buffer (block ib's
  // nodes from noff[ib])
  DBuf<unsigned char> &mt = w.mt, &mt1 = w.mt1;
  mt.reserve(std::max(totV, 1));
  mt1.reserve(std::max(totV, 1));
  std::vector<ProcSeg> pseg;
  std::vector<FringeSeg> fseg;
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    et.hblk[ib] = b.head_blk.p;
    et.hrem[ib] = b.head_remote.p;
    et.nodeRes[ib] = b.nodeRes.p;
    et.scale[ib] = b.resolutionScale;
    // exchangeDonors returns at once on a rank without partners
    // (exchangeDonors.C:36): its blocks are neither hole-cut nor fringed
    if (partners(b.rank).empty()) continue;
    if (b.nnodes)
      pseg.push_back(ProcSeg{0, b.nnodes, b.tag, b.x.p, b.has_donor.p, b.rmax.p, b.tagmask.p, b.head_key.p,
                             b.head_blk.p, b.iblank.p, b.nodeRes.p, mt.p + noff[ib], mt1.p + noff[ib],
                             pflag.p + noff[ib]});
    if (b.ncells) fseg.push_back(FringeSeg{0, b.ncells, b.ti, b.conn.p, mt.p + noff[ib], mt1.p + noff[ib]});
  }
  const int totP = pack_segs(pseg), totF = pack_segs(fseg);
  const ProcSeg *pt = put_segs(w.tab[2], pseg);
  const FringeSeg *ft = put_segs(w.tab[3], fseg);
  if (totP) k_process1_all<<<totP / BS, BS, 0, s_>>>((int)pseg.size(), pt, H, nwords);
  if (totV) TG_HIP(hipMemcpyAsync(mt1.p, mt.p, totV, hipMemcpyDeviceToDevice, s_));
  for (int it = 0; it < nfringe; it++) {
    if (totF) k_fringe_all<<<totF / BS, BS, 0, s_>>>((int)fseg.size(), ft);
    if (totV) TG_HIP(hipMemcpyAsync(mt.p, mt1.p, totV, hipMemcpyDeviceToDevice, s_));
  }
  if (totP) k_process2_all<<<totP / BS, BS, 0, s_>>>((int)pseg.size(), pt, dtag.p);
  if (pk_.on) {  // iblank and nodeRes as findInterpData sees them
    pk_.iblank_pd.assign(nbk, std::vector<int>());
    pk_.nres_pd.assign(nbk, std::vector<double>());
    for (int ib = 0; ib < nbk; ib++) {
      Block &b = *blocks_[ib];
      pk_.iblank_pd[ib].resize(b.nnodes);
      pk_.nres_pd[ib].resize(b.nnodes);
      if (!b.nnodes) continue;
      TG_HIP(hipMemcpyAsync(pk_.iblank_pd[ib].data(), b.iblank.p, sizeof(int) * b.nnodes, hipMemcpyDeviceToHost,
                            s_));
      TG_HIP(hipMemcpyAsync(pk_.nres_pd[ib].data(), b.nodeRes.p, sizeof(double) * b.nnodes,
                            hipMemcpyDeviceToHost, s_));
    }
    sync();
  }
  // the records of all receptor blocks, in (receptor block, node) order
  const int totrec = totV ? select_flagged(pflag.p, totV, psel.p, s_) : 0;
  RecordList L;
  DBuf<int> &lnode = w.lnode, &ldblk = w.ldblk, &lrem = w.lrem;
  DBuf<double> &lrres = w.lrres;
  L.n = totrec;
  L.node = lnode.reserve(totrec);
  L.dblk = ldblk.reserve(totrec);
  L.remote = lrem.reserve(totrec);
  L.rres = lrres.reserve(totrec);
  const EmitTables *pet = (const EmitTables *)put_table(w.et, et);
  if (totrec) k_emit_records_all<<<nb(totrec), BS, 0, s_>>>(totrec, psel.p, pet, L);
  if (pk_.on) {
    std::vector<int> g(totrec);
    pk_.rec_blk.resize(totrec);
    pk_.rec_node.resize(totrec);
    pk_.rec_dblk.resize(totrec);
    pk_.rec_remote.resize(totrec);
    pk_.rec_rres.resize(totrec);
    if (totrec) {
      TG_HIP(hipMemcpyAsync(g.data(), psel.p, sizeof(int) * totrec, hipMemcpyDeviceToHost, s_));
      TG_HIP(hipMemcpyAsync(pk_.rec_node.data(), L.node, sizeof(int) * totrec, hipMemcpyDeviceToHost, s_));
      TG_HIP(hipMemcpyAsync(pk_.rec_dblk.data(), L.dblk, sizeof(int) * totrec, hipMemcpyDeviceToHost, s_));
      TG_HIP(hipMemcpyAsync(pk_.rec_remote.data(), L.remote, sizeof(int) * totrec, hipMemcpyDeviceToHost, s_));
      TG_HIP(hipMemcpyAsync(pk_.rec_rres.data(), L.rres, sizeof(double) * totrec, hipMemcpyDeviceToHost, s_));
      sync();
    }
    for (int j = 0; j < totrec; j++)
      pk_.rec_blk[j] = (int)(std::upper_bound(noff.begin(), noff.end(), g[j]) - noff.begin()) - 1;
    pk_.connectivity = true;
    pk_.reduced = pk_.updated = false;
  }
  // comm2: findInterpData per donor block, records in arrival order: a
  // stable sort by donor block keeps each donor's records in list order
  DBuf<int> &dsel = w.dsel, &tmin = w.tmin, &dstart = w.dstart, &didx = w.didx;
  DBuf<unsigned> &dkey0 = w.dkey0, &dkey1 = w.dkey1;
  dsel.reserve(std::max(totrec, 1));
  didx.reserve(std::max(totrec, 1));
  dkey0.reserve(std::max(totrec, 1));
  dkey1.reserve(std::max(totrec, 1));
  dstart.reserve(nbk + 1);
  std::vector<int> ndv(nbk, 0), doff(nbk + 1, 0);
  if (totrec) {
    k_dblk_keys<<<nb(totrec), BS, 0, s_>>>(totrec, L.dblk, dkey0.p, didx.p);
    int bits = 1;
    while (bits < 31 && (1 << bits) < nbk) bits++;
    sort_pairs(dkey0.p, dkey1.p, didx.p, dsel.p, totrec, bits, s_);
    k_dblk_starts<<<nb(nbk + 1), BS, 0, s_>>>(nbk, totrec, dkey1.p, dstart.p);
    TG_HIP(hipMemcpyAsync(doff.data(), dstart.p, sizeof(int) * (nbk + 1), hipMemcpyDeviceToHost, s_));
    sync();
  }
  for (int ib = 0; ib < nbk; ib++) ndv[ib] = doff[ib + 1] - doff[ib];
  // first reset records and cancel flags of every donor block, each in one
  // buffer (block ib's nodes from noff[ib])
  tmin.reserve(std::max(totV, 1));
  w.cancel.reserve(std::max(totV, 1));
  DBuf<int> &errs = w.errs;
  errs.reserve(nbk);
  add_fill(errs.p, nbk, 0, 4);
  add_fill(tmin.p, totV, (unsigned)BIGINT, 4);
  add_fill(w.cancel.p, totV, 0, 1);
  CancelTables ct{};
  std::vector<InterpSeg> iseg;
  std::vector<NodeSeg> nseg;
  for (int ib = 0; ib < nbk; ib++) {
    Block &D = *blocks_[ib];
    const int nd = ndv[ib];
    D.ninterp = nd;
    D.interp_inode.reserve(9 * (size_t)nd);
    D.interp_w.reserve(9 * (size_t)nd);
    D.interp_nw.reserve(nd);
    D.interp_cancel.reserve(nd);
    D.interp_rnode.reserve(nd);
    D.interp_rblk.reserve(nd);
    D.interp2donor.reserve(std::max(D.nsearch, 1));
    add_fill(D.interp2donor.p, std::max(D.nsearch, 1), 0xffffffffu, 4);
    ct.i2d[ib] = D.interp2donor.p;
    ct.cancel[ib] = D.interp_cancel.p;
    if (D.nnodes)
      nseg.push_back(NodeSeg{0, D.nnodes, tmin.p + noff[ib], w.cancel.p + noff[ib], D.has_donor.p, D.head_blk.p,
                             D.head_remote.p, D.iblank.p});
    if (nd == 0) continue;
    InterpView V;
    V.n = nd;
    V.sel = dsel.p + doff[ib];
    V.L = L;
    V.ti = D.ti;
    V.conn = D.conn.p;
    V.x = D.x.p;
    V.nodeRes = D.nodeRes.p;
    V.iblank = D.iblank.p;
    V.donorId = D.donorId.p;
    V.isearch = D.isearch.p;
    V.tagsearch = D.tagsearch.p;
    V.xsearch = D.xsearch.p;
    V.scale = D.resolutionScale;
    V.tmin = tmin.p + noff[ib];
    V.cancelv = w.cancel.p + noff[ib];
    V.inode = D.interp_inode.p;
    V.nw = D.interp_nw.p;
    V.cancel = D.interp_cancel.p;
    V.rnode = D.interp_rnode.p;
    V.rblk = D.interp_rblk.p;
    V.i2d = D.interp2donor.p;
    V.w = D.interp_w.p;
    V.err = errs.p + ib;
    iseg.push_back(InterpSeg{0, nd, V});
  }
  run_fills(w.tab[4]);
  const int totI = pack_segs(iseg), totN = pack_segs(nseg);
  const InterpSeg *itab = put_segs(w.tab[5], iseg);
  const NodeSeg *ntab = put_segs(w.tab[6], nseg);
  if (totI) {
    k_interp_mandatory_all<<<totI / BS, BS, 0, s_>>>((int)iseg.size(), itab);
    k_interp_all<<<totI / BS, BS, 0, s_>>>((int)iseg.size(), itab);
  }
  if (totN) k_interp_reset_all<<<totN / BS, BS, 0, s_>>>((int)nseg.size(), ntab);
  {
    std::vector<int> nbad(nbk, 0);
    TG_HIP(hipMemcpyAsync(nbad.data(), errs.p, sizeof(int) * nbk, hipMemcpyDeviceToHost, s_));
    sync();
    for (int ib = 0; ib < nbk; ib++)
      if (nbad[ib]) {
        weight_aborts_ += nbad[ib];
        fprintf(stderr,
                "tioga_amd: %d donor weights outside [-0.2, 1.2] in block %d (the reference aborts here)\n",
                nbad[ib], blocks_[ib]->tag);
      }
  }
  // comm3: cancellations, then coincident resets (every block's reset
  // words in one buffer, block ib's from qoff[ib])
  CancelTables *pct = (CancelTables *)put_table(w.ct, ct);
  if (totN) k_cancel_all<<<totN / BS, BS, 0, s_>>>((int)nseg.size(), ntab, pct);
  std::vector<size_t> qoff(nbk + 1, 0);
  for (int ib = 0; ib < nbk; ib++) qoff[ib + 1] = qoff[ib] + blocks_[ib]->nsearch;
  DBuf<int> &ireset = w.ireset;
  ireset.reserve(std::max<size_t>(qoff[nbk], 1));
  add_fill(ireset.p, qoff[nbk], 1, 4);
  std::vector<QuerySeg> qseg, rseg;
  IblankTables it{};
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    it.ib[ib] = b.iblank.p;
    if (b.nsearch)
      qseg.push_back(QuerySeg{0, b.nsearch, b.interp2donor.p, b.xtag.p, b.interp_cancel.p, ireset.p + qoff[ib],
                              nullptr, nullptr});
    if (b.ninterp)
      rseg.push_back(QuerySeg{0, b.ninterp, nullptr, nullptr, b.interp_cancel.p, nullptr, b.interp_rnode.p,
                              b.interp_rblk.p});
  }
  run_fills(w.tab[7]);
  const int totQ = pack_segs(qseg), totR = pack_segs(rseg);
  const QuerySeg *qt = put_segs(w.tab[8], qseg);
  const QuerySeg *rtab = put_segs(w.tab[9], rseg);
  if (totQ) {
    k_coinc1_all<<<totQ / BS, BS, 0, s_>>>((int)qseg.size(), qt);
    k_coinc2_all<<<totQ / BS, BS, 0, s_>>>((int)qseg.size(), qt);
  }
  // comm4: clear negative iblanks, then mark live receptors
  if (totN) k_clear_neg_all<<<totN / BS, BS, 0, s_>>>((int)nseg.size(), ntab);
  IblankTables *pit = (IblankTables *)put_table(w.it, it);
  if (totR) k_set_receptors_all<<<totR / BS, BS, 0, s_>>>((int)rseg.size(), rtab, pit);
  sync();
}

void Engine::cell_iblanks(Block &b)
{
  b.iblank_cell.reserve(b.ncells);
  if (b.ncells) k_cell_iblank<<<nb(b.ncells), BS, 0, s_>>>(b.ncells, b.ti, b.conn.p, b.iblank.p, b.iblank_cell.p);
}

// tioga::performConnectivity (tioga.C:112-160)
void Engine::perform_connectivity()
{
  PhaseTimes t;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  auto t0 = now();
  points_ready_ = false;
  get_hole_map();
  auto t1 = now();
  exchange_boxes();
  auto t2 = now();
  exchange_search_data();
  auto t3 = now();
  for (double &a : prep_acc_) a = 0;
  obb_stepped_ = 0;
  for (Block *b : blocks_) {
    b->ninterp = 0;
    b->interp_host_valid = false;
  }
  if (batched) {
    search_batched();
  } else {
    for (Block *b : blocks_) search(*b);
  }
  auto t4 = now();
  {
    Timer tm(s_);
    if (batched) {
      launch_searches();
    } else {
      for (Block *b : blocks_) launch_search(*b);
    }
    t.search_kernel = tm.stop();
  }
  auto t5 = now();
  exchange_donors();
  {
    // every block's cell iblanks in one launch
    std::vector<CellIbSeg> cs;
    for (Block *bp : blocks_) {
      Block &b = *bp;
      b.iblank_cell.reserve(b.ncells);
      if (b.ncells) cs.push_back(CellIbSeg{0, b.ncells, b.ti, b.conn.p, b.iblank.p, b.iblank_cell.p});
    }
    const int totc = pack_segs(cs);
    if (totc) k_cell_iblank_all<<<totc / BS, BS, 0, s_>>>((int)cs.size(), put_segs(stw_.cib, cs));
  }
  sync();
  auto t6 = now();
  // write back to the user's arrays: every block's iblank and iblank_cell
  // by DMA into one page-locked staging buffer, then into the user's
  // (pageable) arrays on host threads
  {
    struct Part {
      int *dst;
      size_t off, n;
    };
    std::vector<Part> parts;
    size_t tot = 0;
    for (Block *bp : blocks_) {
      Block &b = *bp;
      b.uq = nullptr;  // solutions are registered again after connectivity (tioga.C:152-155)
      if (!b.writeback) continue;
      int *ibc = b.uiblank_cell;
      if (!ibc) {
        b.own_iblank_cell.resize(std::max(b.ncells, 1));
        ibc = b.own_iblank_cell.data();
      }
      parts.push_back(Part{b.uiblank, tot, (size_t)b.nnodes});
      tot += b.nnodes;
      parts.push_back(Part{ibc, tot, (size_t)b.ncells});
      tot += b.ncells;
    }
    if (tot > wb_cap_) {
      if (wb_) TG_HIP(hipHostFree(wb_));
      wb_ = nullptr;
      wb_cap_ = 0;
      TG_HIP(hipHostMalloc((void **)&wb_, sizeof(int) * tot, hipHostMallocDefault));
      wb_cap_ = tot;
    }
    // the DMA in pieces of at most 1 M words, an event after each, so the
    // host copies of a piece start as soon as it has landed
    struct Piece {
      int *dst;
      const int *src;
      size_t off, n;
    };
    std::vector<Piece> pieces;
    {
      size_t k = 0;
      for (Block *bp : blocks_) {
        Block &b = *bp;
        if (!b.writeback) continue;
        const int *srcs[2] = {b.iblank.p, b.iblank_cell.p};
        for (int h = 0; h < 2; h++, k++)
          for (size_t a = 0; a < parts[k].n; a += (1 << 20))
            pieces.push_back(Piece{parts[k].dst + a, srcs[h] + a, parts[k].off + a,
                                   std::min<size_t>(1 << 20, parts[k].n - a)});
      }
    }
    while (wb_ev_.size() < pieces.size()) {
      hipEvent_t e;
      TG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      wb_ev_.push_back(e);
    }
    // iblank values are 1, 0, -1 or minus a mesh tag: when every one fits a
    // byte (checked on the device), a quarter of the bytes cross the bus and
    // the host widens them; otherwise the words go as they are
    bool narrow = false;
    if (tot) {
      std::vector<PackSeg> ps;
      for (const Piece &q : pieces) ps.push_back(PackSeg{0, (int)q.n, q.src, q.off});
      const int totp = pack_segs(ps);
      wb8_.reserve(tot);
      wbflag_.reserve(1);
      fill<int>(wbflag_.p, 1, 0, s_);
      k_pack_i8<<<totp / BS, BS, 0, s_>>>((int)ps.size(), put_segs(wbtab_, ps), wb8_.p, wbflag_.p);
      int wide = 1;
      TG_HIP(hipMemcpyAsync(&wide, wbflag_.p, sizeof(int), hipMemcpyDeviceToHost, s_));
      sync();
      narrow = (wide == 0);
    }
    signed char *wb8h = (signed char *)wb_;  // the narrowed words, front of the staging buffer
    for (size_t k = 0; k < pieces.size(); k++) {
      if (narrow)
        TG_HIP(hipMemcpyAsync(wb8h + pieces[k].off, wb8_.p + pieces[k].off, pieces[k].n, hipMemcpyDeviceToHost, s_));
      else
        TG_HIP(hipMemcpyAsync(wb_ + pieces[k].off, pieces[k].src, sizeof(int) * pieces[k].n, hipMemcpyDeviceToHost,
                              s_));
      TG_HIP(hipEventRecord(wb_ev_[k], s_));
    }
    const int nt = (int)std::max<size_t>(1, std::min<size_t>(8, pieces.size()));
    std::atomic<size_t> next{0};
    std::exception_ptr cerr;
    std::mutex cmu;
    auto copier = [&] {
      for (size_t k; (k = next.fetch_add(1)) < pieces.size();) {
        const hipError_t e = hipEventSynchronize(wb_ev_[k]);
        if (e != hipSuccess) {
          std::lock_guard<std::mutex> g(cmu);
          if (!cerr) cerr = std::make_exception_ptr(std::runtime_error("tioga_amd: iblank write-back failed"));
          continue;
        }
        if (narrow) {
          const signed char *src = wb8h + pieces[k].off;
          int *dst = pieces[k].dst;
          for (size_t j = 0; j < pieces[k].n; j++) dst[j] = src[j];
        } else {
          memcpy(pieces[k].dst, wb_ + pieces[k].off, sizeof(int) * pieces[k].n);
        }
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(copier);
    copier();
    for (auto &x : th) x.join();
    sync();
    if (cerr) std::rethrow_exception(cerr);
  }
  auto t7 = now();
  t.holemap = ms(t0, t1);
  t.boxes = ms(t1, t2);
  t.searchdata = ms(t2, t3);
  t.search_prep = ms(t3, t4);
  t.donors = ms(t5, t6);
  t.total = ms(t0, t7);
  t.nsearch = 0;
  for (Block *b : blocks_) t.nsearch += b->nsearch;
  t.prep_obb = prep_acc_[0];
  t.prep_filter = prep_acc_[1];
  t.prep_adt = prep_acc_[2];
  t.prep_dedupe = prep_acc_[3];
  t.obb_stepped = obb_stepped_;
  last_times = t;
  connected_ = true;
  ho_global_ = ho_.set;
  prepare_update();
}

// tioga::reduce_fringes (tioga.C:737-885), one rank: each block's fringe
// reduction, then comm 3 (the donor record behind the head of every dropped
// receptor's donor list is cancelled, bookKeeping.C:542-570), resetCoincident,
// and comm 4 (clearIblanks with iblank_reduced, then every live record marks
// its receptor).  Cell iblanks are left as the connectivity computed them, as
// the reference leaves them.
void Engine::reduce_fringes()
{
  if (!connected_) throw std::runtime_error("tioga_amd: reduce_fringes needs performconnectivity first");
  const int nbk = (int)blocks_.size();
  std::vector<DBuf<int> *> ibr(nbk);
  std::vector<DBuf<unsigned char> *> cancelv(nbk);
  DBuf<int> tmp;
  CancelTables ct{};
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    const int V = b.nnodes, N = b.ncells;
    ibr[ib] = new DBuf<int>;
    cancelv[ib] = new DBuf<unsigned char>;
    ibr[ib]->reserve(std::max(V, 1));
    cancelv[ib]->reserve(std::max(V, 1));
    tmp.reserve(std::max(V, 1));
    ct.i2d[ib] = b.interp2donor.p;
    ct.cancel[ib] = b.interp_cancel.p;
    if (V == 0) continue;
    k_rf_init<<<nb(V), BS, 0, s_>>>(V, b.iblank.p, ibr[ib]->p, tmp.p);
    for (int it = 0; it < nfringe + 1; it++) {
      if (N) k_rf_sweep<<<nb(N), BS, 0, s_>>>(N, b.ti, b.conn.p, ibr[ib]->p, b.iblank.p, tmp.p);
      TG_HIP(hipMemcpyAsync(ibr[ib]->p, tmp.p, sizeof(int) * V, hipMemcpyDeviceToDevice, s_));
    }
    k_rf_cancel<<<nb(V), BS, 0, s_>>>(V, b.iblank.p, ibr[ib]->p, cancelv[ib]->p);
  }
  if (pk_.on) {  // the receptors each block drops, ascending (its new cancel list)
    pk_.rf_cancel.assign(nbk, std::vector<int>());
    for (int ib = 0; ib < nbk; ib++) {
      const int V = blocks_[ib]->nnodes;
      std::vector<unsigned char> m(V);
      if (V) TG_HIP(hipMemcpyAsync(m.data(), cancelv[ib]->p, V, hipMemcpyDeviceToHost, s_));
      sync();
      for (int i = 0; i < V; i++)
        if (m[i]) pk_.rf_cancel[ib].push_back(i);
    }
    pk_.reduced = true;
    pk_.updated = false;
  }
  DevStruct<CancelTables> dct;
  CancelTables *pct = dct.put(ct, s_);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (b.nnodes && b.has_donor.p)
      k_cancel<<<nb(b.nnodes), BS, 0, s_>>>(b.nnodes, cancelv[ib]->p, b.has_donor.p, b.head_blk.p,
                                            b.head_remote.p, pct);
  }
  DBuf<int> ireset;
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (b.nsearch == 0 || b.interp2donor.p == nullptr) continue;
    ireset.reserve(b.nsearch);
    fill<int>(ireset.p, b.nsearch, 1, s_);
    k_coinc1<<<nb(b.nsearch), BS, 0, s_>>>(b.nsearch, b.interp2donor.p, b.xtag.p, b.interp_cancel.p, ireset.p);
    k_coinc2<<<nb(b.nsearch), BS, 0, s_>>>(b.nsearch, b.interp2donor.p, b.xtag.p, ireset.p, b.interp_cancel.p);
  }
  IblankTables it{};
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (b.nnodes) k_rf_clear<<<nb(b.nnodes), BS, 0, s_>>>(b.nnodes, ibr[ib]->p, b.iblank.p);
    it.ib[ib] = b.iblank.p;
  }
  DevStruct<IblankTables> dit;
  IblankTables *pit = dit.put(it, s_);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (b.ninterp)
      k_set_receptors<<<nb(b.ninterp), BS, 0, s_>>>(b.ninterp, b.interp_cancel.p, b.interp_rnode.p,
                                                    b.interp_rblk.p, pit);
    b.interp_host_valid = false;
  }
  for (Block *bp : blocks_) {
    Block &b = *bp;
    if (b.writeback && b.nnodes)
      TG_HIP(hipMemcpyAsync(b.uiblank, b.iblank.p, sizeof(int) * b.nnodes, hipMemcpyDeviceToHost, s_));
  }
  sync();
  for (int ib = 0; ib < nbk; ib++) {
    delete ibr[ib];
    delete cancelv[ib];
  }
  prepare_update();
}

static void fetch_interp(Block &b, hipStream_t s)
{
  if (b.interp_host_valid) return;
  const int n = b.ninterp;
  std::vector<int> inode(9 * (size_t)n), nw(n), cancel(n), rnode(n), rblk(n);
  std::vector<double> w(9 * (size_t)n);
  if (n) {
    TG_HIP(hipMemcpyAsync(inode.data(), b.interp_inode.p, sizeof(int) * inode.size(), hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(w.data(), b.interp_w.p, sizeof(double) * w.size(), hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(nw.data(), b.interp_nw.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(cancel.data(), b.interp_cancel.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(rnode.data(), b.interp_rnode.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipMemcpyAsync(rblk.data(), b.interp_rblk.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
  }
  b.interp_host.resize(n);
  for (int r = 0; r < n; r++) {
    InterpHost &h = b.interp_host[r];
    h.cancel = cancel[r];
    h.nweights = nw[r];
    h.rinfo[0] = 0;
    h.rinfo[1] = rnode[r];
    h.rinfo[2] = rblk[r];
    for (int m = 0; m < 9; m++) {
      h.inode[m] = m <= h.nweights ? inode[(size_t)m * n + r] : 0;
      h.w[m] = m <= h.nweights ? w[(size_t)m * n + r] : 0.0;
    }
  }
  b.interp_host_valid = true;
}

void Engine::interp_records(Block &b) { fetch_interp(b, s_); }

// tioga::getDonorCount (tioga.C:533-549, dataUpdate.C:121-134)
void Engine::get_donor_count(int btag, int *dcount, int *fcount)
{
  *dcount = *fcount = 0;
  Block *b = find(btag);
  if (!b || nsend_ == 0) return;
  fetch_interp(*b, s_);
  for (auto &h : b->interp_host)
    if (!h.cancel) {
      (*dcount)++;
      *fcount += h.nweights + 1;
    }
}

// tioga::getDonorInfo (tioga.C:551-570, dataUpdate.C:136-159): per live
// record (receptor's rank, receptor node, receptor block's index on its
// rank, nweights)
void Engine::get_donor_info(int btag, int *receptors, int *indices, double *frac, int *dcount)
{
  (void)dcount;
  Block *b = find(btag);
  if (!b || nsend_ == 0) return;
  fetch_interp(*b, s_);
  size_t j = 0, k = 0;
  for (auto &h : b->interp_host) {
    if (h.cancel) continue;
    for (int m = 0; m < h.nweights + 1; m++) {
      indices[j] = h.inode[m];
      frac[j] = h.w[m];
      j++;
    }
    const Block &R = *blocks_[h.rinfo[2]];
    receptors[k++] = R.rank;
    receptors[k++] = h.rinfo[1];
    receptors[k++] = R.lidx;
    receptors[k++] = h.nweights;
  }
}

// tioga::getReceptorInfo (tioga.C:572-671, dataUpdate.C:161-178) on the
// current rank: per live record whose receptor is on this rank (node,
// receptor mesh tag, donor cell global id as two ints), in the order the
// packets arrive: donor rank, donor block, record; the cell index stands in
// for the global id without cell_gid.
void Engine::get_receptor_info(std::vector<int> &receptors)
{
  receptors.clear();
  if (nsend_ == 0) return;
  for (Block *bp : blocks_) {
    Block &b = *bp;
    fetch_interp(b, s_);
    for (auto &h : b.interp_host) {
      if (h.cancel || blocks_[h.rinfo[2]]->rank != cur_rank) continue;
      receptors.push_back(h.rinfo[1]);
      receptors.push_back(blocks_[h.rinfo[2]]->tag);
      int don = h.inode[h.nweights];
      uint64_t g = b.ucell_gid ? b.ucell_gid[don] : (uint64_t)don;
      int two[2];
      memcpy(two, &g, sizeof g);
      receptors.push_back(two[0]);
      receptors.push_back(two[1]);
    }
  }
}

template <class T>
static long copy_out(const DBuf<T> &d, size_t n, void *out, hipStream_t s)
{
  if (out && n) {
    TG_HIP(hipMemcpyAsync(out, d.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
  }
  return (long)n;
}

long Engine::block_array(int btag, const std::string &nm, void *out)
{
  Block *bp = find(btag);
  if (!bp) return -1;
  Block &b = *bp;
  const size_t Q = b.nsearch;
  if (nm == "nodeRes") return copy_out(b.nodeRes, b.nnodes, out, s_);
  if (nm == "cellRes") return copy_out(b.cellRes, b.ncells, out, s_);
  if (nm == "xsearch") return copy_out(b.xsearch, 3 * Q, out, s_);
  if (nm == "res_search") return copy_out(b.res_search, Q, out, s_);
  if (nm == "isearch") {
    // the device holds (0, node, sender block index); the reference's
    // (partner index of the sender's rank, node, sender's index on its rank)
    const long n = copy_out(b.isearch, 3 * Q, out, s_);
    if (out && nranks() > 1) {
      int *o = (int *)out;
      const std::vector<int> &pm = partners(b.rank);
      for (size_t q = 0; q < Q; q++) {
        const Block &S = *blocks_[o[3 * q + 2]];
        o[3 * q] = (int)(std::lower_bound(pm.begin(), pm.end(), S.rank) - pm.begin());
        o[3 * q + 2] = S.lidx;
      }
    }
    return n;
  }
  if (nm == "tagsearch") return copy_out(b.tagsearch, Q, out, s_);
  if (nm == "donorId") return copy_out(b.donorId, Q, out, s_);
  if (nm == "xtag") return copy_out(b.xtag, Q, out, s_);
  if (nm == "elementList") return copy_out(b.elist, Q ? b.nelem : 0, out, s_);
  if (nm == "adtIntegers") return copy_out(b.adtInt, Q ? 4 * (size_t)b.nelem : 0, out, s_);
  if (nm == "adtReals") return copy_out(b.adtReal, Q ? 6 * (size_t)b.nelem : 0, out, s_);
  if (nm == "iblank_cell") return copy_out(b.iblank_cell, b.ncells, out, s_);
  if (nm == "adtExtents") {
    if (out && Q) memcpy(out, b.adtExt, sizeof b.adtExt);
    return Q ? 6 : 0;
  }
  if (nm == "obb") {
    if (out) {
      double *o = (double *)out;
      memcpy(o, b.obb.xc, 24);
      memcpy(o + 3, b.obb.dxc, 24);
      memcpy(o + 6, b.obb.vec, 72);
    }
    return 15;
  }
  return -1;
}

long Engine::hole_map(int mtag, int nx[3], double ext[6], int *sam)
{
  if (mtag < 1 || mtag > (int)hm_.size()) return 0;
  HoleMap &h = hm_[mtag - 1];
  if (!h.exist) return 0;
  size_t n = (size_t)h.nx[0] * h.nx[1] * h.nx[2];
  if (nx) memcpy(nx, h.nx, sizeof(int) * 3);
  if (ext) memcpy(ext, h.ext, sizeof(double) * 6);
  if (sam) {
    std::vector<unsigned char> tmp(n);
    TG_HIP(hipMemcpyAsync(tmp.data(), h.sam.p, n, hipMemcpyDeviceToHost, s_));
    sync();
    for (size_t i = 0; i < n; i++) sam[i] = tmp[i];
  }
  return (long)n;
}

double Engine::time_search(int nrep, long long *nsearch)
{
  long long q = 0;
  for (Block *b : blocks_) q += b->nsearch;
  if (nsearch) *nsearch = q;
  if (!connected_ || nrep <= 0) return 0.0;
  Timer tm(s_);
  for (int r = 0; r < nrep; r++) {
    if (batched) {
      launch_searches();
    } else {
      for (Block *b : blocks_) launch_search(*b);
    }
  }
  return tm.stop() / nrep;
}


}  // namespace tg
