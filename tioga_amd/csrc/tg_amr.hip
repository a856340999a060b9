// tg_amr.hip -- AMR Cartesian background, gather direction: near-body
// receptors take their values from Cartesian patches (pdegree 0, linear
// interpolation between cell centres).
//
// Reference path, single rank:
//   * tioga::register_amr_global_data / set_amr_patch_count /
//     register_amr_local_data (tioga.C:692-713, CartGrid.C:22-70);
//   * tioga::performConnectivityAMR (tioga.C:180-216): CartGrid::preprocess
//     (CartGrid.C:77-104), MeshBlock::getUnresolvedMandatoryReceptors
//     (cartOps.C:61-160, node form), CartGrid::search (CartGrid.C:111-145),
//     the patch interpolation lists of exchangeAMRDonors.C:120-192
//     (CartBlock::insertInInterpList, CartBlock.C:196-268, with
//     cart_interp::linear_interpolation, linCartInterp.C:49-109),
//     MeshBlock::setCartIblanks (cartOps.C:32-58) and getCellIblanks;
//   * tioga::dataUpdate_AMR (tioga.C:218-334): CartBlock::getInterpolatedData
//     (CartBlock.C:44-140, get_amr_index_xyz.c:1-56) and
//     MeshBlock::updateSolnData (dataUpdate.C:105-119).
// The Cartesian-receptor direction (getCartReceptors, donor lists on the
// patches, processDonors, cancellation) is not part of this path; on one
// rank it never cancels a near-body receptor, since getCancellationData
// (CartBlock.C:505-531) only emits cancellations of the other direction.
//
// Device layout:
//   * patch table (64 B per grid) and, per refinement level, a uniform bin
//     grid whose bins list (ascending grid index) every patch of that level
//     whose box, widened by more than TOL, touches the bin;
//   * receptors of all blocks concatenated: node id, block, coordinates,
//     donor grid;
//   * interpolation records sorted by local patch (stable in (block, point)
//     order, the order the reference's lists hold), structure of arrays:
//     patch, node, block|stencil bits, stencil base index, and the six 1-D
//     basis values from which the 8 weights are formed exactly as
//     compute_weights forms them (64 B per record instead of 108 B);
//   * patch solutions packed in one buffer, nvar fields of d3nf doubles each.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>

#include "tg_engine.h"
#include "tg_prims.h"
#include "tg_seq.h"

namespace tg {

struct AmrGridDev {
  double xlo[3], dx[3];
  int dims[3], level;
};
static_assert(sizeof(AmrGridDev) == 64, "patch records are 64 bytes");

struct AmrLevelBins {
  double o[3], h[3];
  int nb[3], start;
};

struct AmrQTables {
  double *q[128];
  int nnodes[128];
};

struct AmrState {
  // CartGrid (host copy, CartGrid.h:26-76)
  int nf = 0, qstride = 0, ngrids = 0, qnodesize = 0, maxlevel = 0;
  std::vector<int> gid, level, proc, porder, lid, dims;
  std::vector<double> xlo, dx, qnode;
  bool have_grid = false;
  // CartBlocks of this rank
  int npatch = 0;
  std::vector<int> pgid;
  std::vector<int *> pibl;
  std::vector<double *> pq;
  bool connected = false;  // performConnectivityAMR done (iamrGlobal)
  // device tables
  DBuf<AmrGridDev> grids;
  DBuf<AmrLevelBins> levels;
  DBuf<int> boff, blist, plid, ppgid, pd3nf;
  // receptors (all blocks, concatenated)
  int nrcp = 0;
  std::vector<int> rcp_off;  // per block
  DBuf<int> rnode, rblk, rdonor;
  DBuf<double> rxyz;
  DBuf<unsigned char> picked;
  // records
  int nrec = 0;
  DBuf<int> rec_lp, rec_node, rec_meta, rec_base, rec_rid;
  DBuf<double> rec_phi;
  std::vector<int> rec_count;  // per local patch
  // packed patch solutions
  DBuf<double> q;
  DBuf<long long> qoff;
  DBuf<int> err;
  DBuf<AmrQTables> qt;
  // the Cartesian-receptor direction: per patch cell (local patch p's cells
  // from coff[p], x fastest) the donor count, donor mesh tags, whether a
  // donor of resolution < BIGVALUE exists, and the cancel mark; the patch
  // iblanks (d3nf per patch from ioff[p]); per block the records of the
  // interpolation list to Cartesian cells (interpListCart)
  std::vector<long long> coff, ioff;
  DBuf<long long> dcoff, dioff;
  DBuf<int> ccnt, cgood, ccancel, pibl_dev;
  DBuf<unsigned long long> ctags;
  struct CartRecs {
    int n = 0;
    DBuf<int> inode, nw, rpt, rlp;  // inode[m * n + r]; receptor: patch q index, local patch
    DBuf<double> w;
  };
  std::vector<std::unique_ptr<CartRecs>> crec;  // per block
  std::vector<double> cvals;   // the last update's values at Cartesian receptors (host)
};

namespace {

constexpr int ABS = 256;
inline unsigned anb(long long n) { return (unsigned)((n + ABS - 1) / ABS); }

__device__ __forceinline__ int type_of_cell(const TypeInfo &ti, int c)
{
  int t = 0;
  while (t + 1 < ti.ntypes && c >= ti.cstart[t + 1]) t++;
  return t;
}

// getUnresolvedMandatoryReceptors (cartOps.C:73-97), node form: every node of
// a cell whose nodes all carry nodeRes == BIGVALUE and whose cell iblank is 1
__global__ void k_amr_pick(int ncells, TypeInfo ti, const int *__restrict__ conn, const double *__restrict__ nodeRes,
                           const int *__restrict__ ibc, unsigned char *__restrict__ picked)
{
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int nv = ti.nv[type_of_cell(ti, c)];
  const int *cv = conn + 8 * (size_t)c;
  int fcount = 0;
  for (int m = 0; m < nv; m++) fcount += (nodeRes[cv[m]] == BIGVALUE);
  if (fcount == nv && ibc[c] == 1)
    for (int m = 0; m < nv; m++) picked[cv[m]] = 1;
}

__global__ void k_amr_flags(int n, const unsigned char *__restrict__ picked, int *__restrict__ flags)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = picked[i];
}

__global__ void k_amr_receptors(int n, const int *__restrict__ nodes, const double *__restrict__ x, int blk,
                                int *__restrict__ rnode, int *__restrict__ rblk, double *__restrict__ rxyz)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = nodes[i];
  rnode[i] = v;
  rblk[i] = blk;
  for (int j = 0; j < 3; j++) rxyz[3 * (size_t)i + j] = x[3 * (size_t)v + j];
}

// CartGrid::search (CartGrid.C:111-145): the first grid, in registration
// order, of the finest level whose box holds the point within +-TOL.  The
// bins of a level list every grid that can pass the test for a point in the
// bin, in ascending order, so the first grid of the list that passes is the
// first grid of the level that passes.
__global__ void k_amr_search(int n, const double *__restrict__ rxyz, int maxlevel,
                             const AmrLevelBins *__restrict__ lv, const int *__restrict__ boff,
                             const int *__restrict__ blist, const AmrGridDev *__restrict__ grids,
                             int *__restrict__ donor)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x0 = rxyz[3 * (size_t)i], x1 = rxyz[3 * (size_t)i + 1], x2 = rxyz[3 * (size_t)i + 2];
  const double xp[3] = {x0, x1, x2};
  int found = -1;
  for (int l = maxlevel - 1; l >= 0 && found < 0; l--) {
    const AmrLevelBins L = lv[l];
    if (L.nb[0] == 0) continue;
    int b[3];
    for (int a = 0; a < 3; a++) {
      double t = floor((xp[a] - L.o[a]) / L.h[a]);
      b[a] = t < 0.0 ? 0 : (t > (double)(L.nb[a] - 1) ? L.nb[a] - 1 : (int)t);
    }
    const int bin = L.start + (b[2] * L.nb[1] + b[1]) * L.nb[0] + b[0];
    for (int k = boff[bin]; k < boff[bin + 1]; k++) {
      const int j = blist[k];
      const AmrGridDev &G = grids[j];
      bool flag = true;
      for (int a = 0; a < 3; a++) flag = flag && ((xp[a] - G.xlo[a]) > -TOL);
      for (int a = 0; a < 3; a++) flag = flag && ((xp[a] - (G.xlo[a] + G.dx[a] * (G.dims[a]))) < TOL);
      if (flag) {
        found = j;
        break;
      }
    }
  }
  donor[i] = found;
}

__global__ void k_amr_keys(int n, const int *__restrict__ donor, const int *__restrict__ plid,
                           unsigned *__restrict__ key, int *__restrict__ flags)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g = donor[i];
  flags[i] = g >= 0;
  key[i] = g >= 0 ? (unsigned)plid[g] : 0u;
}

__global__ void k_amr_take(int n, const int *__restrict__ idx, const unsigned *__restrict__ all,
                           unsigned *__restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = all[idx[i]];
}

__global__ void k_amr_iota(int n, int *__restrict__ a)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ void k_amr_key_of(int n, const int *__restrict__ perm, const int *__restrict__ src,
                             unsigned *__restrict__ key)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) key[i] = (unsigned)src[perm[i]];
}

// records from list order into gather order (perm[i] = list index)
__global__ void k_amr_permute(int n, const int *__restrict__ perm, const int *__restrict__ a0,
                              const int *__restrict__ a1, const int *__restrict__ a2, const int *__restrict__ a3,
                              const int *__restrict__ a4, const double *__restrict__ ph, int *__restrict__ b0,
                              int *__restrict__ b1, int *__restrict__ b2, int *__restrict__ b3, int *__restrict__ b4,
                              double *__restrict__ phb)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = perm[i];
  b0[i] = a0[r];
  b1[i] = a1[r];
  b2[i] = a2[r];
  b3[i] = a3[r];
  b4[i] = a4[r];
  for (int m = 0; m < 6; m++) phb[(size_t)m * n + i] = ph[(size_t)m * n + r];
}

// MeshBlock::setCartIblanks (cartOps.C:49-57)
__global__ void k_amr_iblank(int n, const int *__restrict__ donor, const int *__restrict__ rnode,
                             const int *__restrict__ rblk, int blk, int *__restrict__ iblank)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (rblk[i] == blk && donor[i] != -1) iblank[rnode[i]] = -1;
}

// CartBlock::insertInInterpList (CartBlock.C:196-268) with
// linear_interpolation (linCartInterp.C:49-109) for p = 1; one thread per
// record, records in list order (rid = receptor of the record).
__global__ void k_amr_records(int nrec, const int *__restrict__ rid, const unsigned *__restrict__ lp_sorted,
                              const double *__restrict__ rxyz, const int *__restrict__ rnode,
                              const int *__restrict__ rblk, const int *__restrict__ ppgid,
                              const AmrGridDev *__restrict__ grids, int nf, int *__restrict__ rec_lp,
                              int *__restrict__ rec_node, int *__restrict__ rec_meta, int *__restrict__ rec_base,
                              double *__restrict__ rec_phi, int *__restrict__ err)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const int i = rid[r];
  const int lp = (int)lp_sorted[r];
  const AmrGridDev G = grids[ppgid[lp]];
  int ix[3], v0[3], bits = 0;
  double ph[6];
  for (int n = 0; n < 3; n++) {
    const double x = rxyz[3 * (size_t)i + n];
    ix[n] = (int)((x - G.xlo[n]) / G.dx[n]);
    double rst = (x - G.xlo[n] - ix[n] * G.dx[n]) / G.dx[n];
    if (ix[n] == G.dims[n] && fabs(rst) < TOL) {
      ix[n]--;
      rst = (x - G.xlo[n] - ix[n] * G.dx[n]) / G.dx[n];
    }
    if (!(ix[n] >= 0 && ix[n] < G.dims[n])) atomicExch(err, 1);  // CartBlock.C:246
    // create_donor_stencil (linCartInterp.C:74-95)
    const int start = (rst - 0.5 >= 0) ? ix[n] : ix[n] - 1;
    const int a0 = max(0, min(start, G.dims[n] - 1)), a1 = max(0, min(start + 1, G.dims[n] - 1));
    v0[n] = a0;
    bits |= (a1 - a0) << n;
    // compute_ref_coords (linCartInterp.C:63-72) and compute_1d_bases (:24-47), p = 1
    const double rc = (rst - 0.5 >= 0) ? ((rst - 0.5) / 1 * 2 - 1) : ((rst + 0.5 * (2 * 1 - 1)) / 1 * 2 - 1);
    ph[2 * n] = (1 - rc) / 2;
    ph[2 * n + 1] = (1 + rc) / 2;
  }
  const int sx = G.dims[0] + 2 * nf, sy = G.dims[1] + 2 * nf;
  rec_lp[r] = lp;
  rec_node[r] = rnode[i];
  rec_meta[r] = (rblk[i] << 3) | bits;
  rec_base[r] = sy * sx * (v0[2] + nf) + sx * (v0[1] + nf) + (v0[0] + nf);  // get_amr_index_xyz.c:40-42
  for (int m = 0; m < 6; m++) rec_phi[(size_t)m * nrec + r] = ph[m];
}

// CartBlock::getInterpolatedData (CartBlock.C:113-127) and updateSolnData
// (dataUpdate.C:105-119) fused: every record writes one receptor node.
__global__ void __launch_bounds__(ABS) k_amr_gather(int nrec, int nvar, int interptype, const int *__restrict__ rec_lp,
                                                    const int *__restrict__ rec_node, const int *__restrict__ rec_meta,
                                                    const int *__restrict__ rec_base, const double *__restrict__ rec_phi,
                                                    const double *__restrict__ q, const long long *__restrict__ qoff,
                                                    const int *__restrict__ pd3nf, const int *__restrict__ pstride,
                                                    const AmrQTables *__restrict__ T)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const int lp = rec_lp[r], meta = rec_meta[r], base = rec_base[r], node = rec_node[r];
  const double px0 = rec_phi[r], px1 = rec_phi[(size_t)nrec + r];
  const double py0 = rec_phi[2 * (size_t)nrec + r], py1 = rec_phi[3 * (size_t)nrec + r];
  const double pz0 = rec_phi[4 * (size_t)nrec + r], pz1 = rec_phi[5 * (size_t)nrec + r];
  const double *qp = q + qoff[lp];
  const int d3 = pd3nf[lp];
  const int sx = pstride[2 * lp], sxy = pstride[2 * lp + 1];
  const int ox = meta & 1, oy = ((meta >> 1) & 1) * sx, oz = ((meta >> 2) & 1) * sxy;
  // weights in stencil order (k, j, i), each phi_z[k]*phi_y[j]*phi_x[i]
  // (linCartInterp.C:49-61)
  double w[8];
  int off[8];
  {
    const double pz[2] = {pz0, pz1}, py[2] = {py0, py1}, px[2] = {px0, px1};
    const int dz[2] = {0, oz}, dy[2] = {0, oy}, dxo[2] = {0, ox};
    int m = 0;
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++, m++) {
          w[m] = pz[k] * py[j] * px[i];
          off[m] = base + dz[k] + dy[j] + dxo[i];
        }
  }
  double *out = T->q[meta >> 3];
  const int nn = T->nnodes[meta >> 3];
  // fields in chunks of 8: all sums of a chunk first, then the receptor's
  // values of the chunk written together (in the row layout a whole
  // 64-byte run, as 16-byte stores when the run is so aligned)
  for (int n0 = 0; n0 < nvar; n0 += 8) {
    const int kn = nvar - n0 < 8 ? nvar - n0 : 8;
    double acc[8];
#pragma unroll
    for (int kk = 0; kk < 8; kk++) {
      acc[kk] = 0;
      if (kk < kn) {
        const double *qn = qp + (size_t)d3 * (n0 + kk);
#pragma unroll
        for (int m = 0; m < 8; m++) acc[kk] += qn[off[m]] * w[m];
      }
    }
    if (interptype == 0) {
      double *o = out + (size_t)node * nvar + n0;
      if (kn == 8 && (((uintptr_t)o) & 15) == 0) {
        double2 *o2 = reinterpret_cast<double2 *>(o);
#pragma unroll
        for (int kk = 0; kk < 4; kk++) o2[kk] = make_double2(acc[2 * kk], acc[2 * kk + 1]);
      } else {
#pragma unroll
        for (int kk = 0; kk < 8; kk++)
          if (kk < kn) o[kk] = acc[kk];
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < 8; kk++)
        if (kk < kn) out[(size_t)nn * (n0 + kk) + node] = acc[kk];
    }
  }
}

}  // namespace

AmrState &Engine::amr()
{
  if (!amr_) amr_ = new AmrState;
  return *amr_;
}

namespace {

// ---- the Cartesian-receptor direction ----

// MeshBlock::getCartReceptors (getCartReceptors.C:85-154) for pdegree 0: the
// centre of patch cell (i, j, k) (get_amr_index_xyz.c:1-56 with one point),
// cells in the reference's loop order (x outermost, z innermost), kept when
// the centre lies in the block's OBB
__device__ __forceinline__ void cart_cell(const AmrGridDev &g, int t, int &i, int &j, int &k)
{
  k = t % g.dims[2];
  j = (t / g.dims[2]) % g.dims[1];
  i = t / (g.dims[2] * g.dims[1]);
}

__device__ __forceinline__ void cart_centre(const AmrGridDev &g, double qn0, int i, int j, int k, double x[3])
{
  const double cx = g.xlo[0] + i * g.dx[0], cy = g.xlo[1] + j * g.dx[1], cz = g.xlo[2] + k * g.dx[2];
  const double hx = 0.5 * g.dx[0], hy = 0.5 * g.dx[1], hz = 0.5 * g.dx[2];
  x[2] = cz + hz * (1.0 + qn0);
  x[1] = cy + hy * (1.0 + qn0);
  x[0] = cx + hx * (1.0 + qn0);
}

// the grids of one block's Cartesian search, cells numbered grid after grid
struct CartGridList {
  int n;                     // grids
  const AmrGridDev *g;       // their records
  const long long *off;      // first cell number of each, and the total (n + 1)
  const int *lid;            // their local patch ids
};

__device__ __forceinline__ int cart_grid_of(const CartGridList &L, long long t)
{
  int lo = 0, hi = L.n - 1;  // last grid whose first cell is <= t
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (L.off[m] <= t) lo = m;
    else hi = m - 1;
  }
  return lo;
}

__global__ void k_cart_flags(long long n, CartGridList L, double qn0, OBB b, int *__restrict__ flag)
{
  const long long u = (long long)blockIdx.x * ABS + threadIdx.x;
  if (u >= n) return;
  const int gi = cart_grid_of(L, u);
  const AmrGridDev g = L.g[gi];
  int i, j, k;
  cart_cell(g, (int)(u - L.off[gi]), i, j, k);
  double x[3], xd[3] = {0, 0, 0};
  cart_centre(g, qn0, i, j, k, x);
  for (int a = 0; a < 3; a++)
    for (int c = 0; c < 3; c++) xd[a] += (x[c] - b.xc[c]) * b.vec[a][c];
  flag[u] = fabs(xd[0]) <= b.dxc[0] && fabs(xd[1]) <= b.dxc[1] && fabs(xd[2]) <= b.dxc[2];
}

// search points: the centre, isearch (rank, local patch, q index of the
// cell), resolution BIGVALUE and mesh tag -1 (the reference leaves both
// unset; they matter only to coincident points, which a patch hierarchy
// does not produce)
__global__ void k_cart_points(int n, const int *__restrict__ sel, CartGridList L, double qn0, int nf,
                              double *__restrict__ xs, int *__restrict__ is, double *__restrict__ rs,
                              int *__restrict__ ts)
{
  const int q = blockIdx.x * ABS + threadIdx.x;
  if (q >= n) return;
  const int gi = cart_grid_of(L, sel[q]);
  const AmrGridDev g = L.g[gi];
  int i, j, k;
  cart_cell(g, (int)(sel[q] - L.off[gi]), i, j, k);
  double x[3];
  cart_centre(g, qn0, i, j, k, x);
  const int sx = g.dims[0] + 2 * nf, sy = g.dims[1] + 2 * nf;
  xs[3 * (size_t)q] = x[0];
  xs[3 * (size_t)q + 1] = x[1];
  xs[3 * (size_t)q + 2] = x[2];
  is[3 * (size_t)q] = 0;
  is[3 * (size_t)q + 1] = L.lid[gi];
  is[3 * (size_t)q + 2] = sy * sx * (k + nf) + sx * (j + nf) + (i + nf);
  rs[q] = BIGVALUE;
  ts[q] = -1;
}

// the patch cell (local patch p, unghosted index) of a Cartesian search
// point: amr_index_to_ijklmn (get_amr_index_xyz.c:58-110) for pdegree 0
__device__ __forceinline__ long long cart_cell_of(const int *is, const int *__restrict__ pd3,
                                                  const long long *__restrict__ coff, const int *__restrict__ pgid,
                                                  const AmrGridDev *__restrict__ G, int nf, int q)
{
  const int lp = is[3 * (size_t)q + 1], idx = is[3 * (size_t)q + 2];
  const AmrGridDev &g = G[pgid[lp]];
  const int sx = g.dims[0] + 2 * nf, sy = g.dims[1] + 2 * nf;
  const int r = idx % pd3[lp];
  const int kk = r / (sx * sy) - nf, jj = (r % (sx * sy)) / sx - nf, ii = r % sx - nf;
  return coff[lp] + ((long long)kk * g.dims[1] + jj) * g.dims[0] + ii;
}

// CartBlock::insertInDonorList (CartBlock.C:270-306): what processDonors
// reads of a cell's list -- whether it is empty, the donor mesh tags, and
// whether some donor resolution is below BIGVALUE
__global__ void k_cart_donors(int n, const int *__restrict__ donor, const int *__restrict__ is,
                              const double *__restrict__ cellRes, int tag, const int *__restrict__ pd3,
                              const long long *__restrict__ coff, const int *__restrict__ pgid,
                              const AmrGridDev *__restrict__ G, int nf, int *__restrict__ cnt,
                              unsigned long long *__restrict__ tags, int *__restrict__ good)
{
  const int q = blockIdx.x * ABS + threadIdx.x;
  if (q >= n || donor[q] == -1) return;
  const long long c = cart_cell_of(is, pd3, coff, pgid, G, nf, q);
  atomicAdd(&cnt[c], 1);
  atomicOr(&tags[c], 1ull << ((tag - 1) & 63));
  if (cellRes[donor[q]] < BIGVALUE) atomicOr(&good[c], 1);
}

struct CartHoles {
  int nmesh;
  const HoleMapView *hm;
};

__device__ __forceinline__ int cart_in_hole(const HoleMapView &h, const double x[3])
{
  double dx[3];
  int ix[3];
  for (int i = 0; i < 3; i++) dx[i] = (h.ext[i + 3] - h.ext[i]) / h.nx[i];
  for (int i = 0; i < 3; i++) {
    ix[i] = (int)((x[i] - h.ext[i]) / dx[i]);
    if (ix[i] < 0 || ix[i] > h.nx[i] - 1) return 0;
  }
  return h.sam[(size_t)ix[2] * h.nx[1] * h.nx[0] + (size_t)ix[1] * h.nx[0] + ix[0]];
}

// CartBlock::processDonors (CartBlock.C:308-504) for pdegree 0, one thread
// per patch cell (every patch's cells, numbered patch after patch from
// coff): a cell without donors inside a hole map, or with donors but inside
// the hole map of a mesh that did not donate, becomes a hole (0); a non-hole
// cell with a donor of resolution < BIGVALUE becomes a receptor (-1); the
// donors of every other cell are cancelled
__global__ void k_cart_process(long long n, int npatch, const long long *__restrict__ coff,
                               const long long *__restrict__ ioff, const int *__restrict__ pgid,
                               const AmrGridDev *__restrict__ G, double qn0, int nf, CartHoles H,
                               const int *__restrict__ cnt, const unsigned long long *__restrict__ tags,
                               const int *__restrict__ good, int *__restrict__ cancel, int *__restrict__ ibl_all)
{
  const long long c = (long long)blockIdx.x * ABS + threadIdx.x;
  if (c >= n) return;
  int lo = 0, hi = npatch - 1;  // the last patch whose first cell is <= c
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (coff[m] <= c) lo = m;
    else hi = m - 1;
  }
  const AmrGridDev g = G[pgid[lo]];
  const int t = (int)(c - coff[lo]);
  const int k = t / (g.dims[0] * g.dims[1]), j = (t / g.dims[0]) % g.dims[1], i = t % g.dims[0];
  double x[3];
  cart_centre(g, qn0, i, j, k, x);
  int *ibl = ibl_all + ioff[lo];
  const int ib = (k + nf) * (g.dims[1] + 2 * nf) * (g.dims[0] + 2 * nf) + (j + nf) * (g.dims[0] + 2 * nf) + i + nf;
  for (int h = 0; h < H.nmesh; h++) {
    if (!H.hm[h].exist) continue;
    if (cnt[c] != 0 && h < 64 && ((tags[c] >> h) & 1ull)) continue;
    if (cart_in_hole(H.hm[h], x)) {
      ibl[ib] = 0;
      break;
    }
  }
  if (ibl[ib] == 0) cancel[c] = 1;
  else if (good[c]) ibl[ib] = -1;
  else cancel[c] = 1;
}

__global__ void k_cart_cancel(int n, int *__restrict__ donor, const int *__restrict__ is, const int *__restrict__ pd3,
                              const long long *__restrict__ coff, const int *__restrict__ pgid,
                              const AmrGridDev *__restrict__ G, int nf, const int *__restrict__ cancel,
                              int *__restrict__ flag)
{
  const int q = blockIdx.x * ABS + threadIdx.x;
  if (q >= n) return;
  if (donor[q] != -1 && cancel[cart_cell_of(is, pd3, coff, pgid, G, nf, q)]) donor[q] = -1;
  flag[q] = donor[q] != -1;
}

// MeshBlock::findInterpListCart (cartOps.C:203-278): the donor cell's
// isoparametric weights at the Cartesian point, record after record
template <int NV>
__device__ void cart_record(int n, int r, int c, const int *conn, const double *x, const double xp[3], int *inode,
                            double *w, int *nw)
{
  double xv[8][3], frac[8];
  const int *cv = conn + 8 * (size_t)c;
#pragma unroll
  for (int m = 0; m < NV; m++) {
    const size_t v = cv[m];
    xv[m][0] = x[3 * v];
    xv[m][1] = x[3 * v + 1];
    xv[m][2] = x[3 * v + 2];
  }
  nodal_weights(xv, xp, frac, NV);
#pragma unroll
  for (int m = 0; m < NV; m++) {
    inode[(size_t)m * n + r] = cv[m];
    w[(size_t)m * n + r] = frac[m];
  }
  nw[r] = NV;
}

__global__ void k_cart_records(int n, const int *__restrict__ sel, TypeInfo ti, const int *__restrict__ conn,
                               const double *__restrict__ x, const double *__restrict__ xs, const int *__restrict__ is,
                               const int *__restrict__ donor, int *__restrict__ inode, double *__restrict__ w,
                               int *__restrict__ nw, int *__restrict__ rpt, int *__restrict__ rlp)
{
  const int r = blockIdx.x * ABS + threadIdx.x;
  if (r >= n) return;
  const int q = sel[r], c = donor[q];
  const int t = type_of_cell(ti, c);
  const double xp[3] = {xs[3 * (size_t)q], xs[3 * (size_t)q + 1], xs[3 * (size_t)q + 2]};
  switch (ti.nv[t]) {
    case 4: cart_record<4>(n, r, c, conn, x, xp, inode, w, nw); break;
    case 5: cart_record<5>(n, r, c, conn, x, xp, inode, w, nw); break;
    case 6: cart_record<6>(n, r, c, conn, x, xp, inode, w, nw); break;
    default: cart_record<8>(n, r, c, conn, x, xp, inode, w, nw); break;
  }
  rpt[r] = is[3 * (size_t)q + 2];
  rlp[r] = is[3 * (size_t)q + 1];
}

// the values MeshBlock::getInterpolatedSolutionAMR sends a Cartesian
// receptor (cartOps.C:366-385, 412-428): per field, vertices in order
__global__ void k_cart_values(int n, int nvar, int column, int nnodes, const double *__restrict__ q,
                              const int *__restrict__ inode, const double *__restrict__ w, const int *__restrict__ nw,
                              double *__restrict__ out)
{
  const int r = blockIdx.x * ABS + threadIdx.x;
  if (r >= n) return;
  for (int k = 0; k < nvar; k++) {
    double qq = 0;
    for (int m = 0; m < nw[r]; m++) {
      const size_t v = inode[(size_t)m * n + r];
      qq += (column ? q[(size_t)k * nnodes + v] : q[v * nvar + k]) * w[(size_t)m * n + r];
    }
    out[(size_t)r * nvar + k] = qq;
  }
}

}  // namespace

// The Cartesian-receptor direction of performConnectivityAMR on one rank
// (tioga.C:190-207): getCartReceptors and the search of each block, the donor
// half of exchangeAMRDonors (exchangeAMRDonors.C:120-287: donor lists on the
// patch cells, CartBlock::processDonors, cancellation), findInterpListCart.
// Every block's donor entries follow each other in block order (the
// reference packs interpolation and donor entries block by block and reads
// them back as all of one kind, then all of the other, exchangeAMRDonors.C:
// 118-199, which agrees only with one near-body block).
void Engine::amr_cart_receptors()
{
  AmrState &A = amr();
  const int nbk = (int)blocks_.size();
  const int G = A.ngrids;
  const double qn0 = A.qnode.empty() ? 0.0 : A.qnode[0];
  std::vector<AmrGridDev> gt(std::max(G, 1));
  for (int g = 0; g < G; g++)
    for (int n = 0; n < 3; n++) {
      gt[g].xlo[n] = A.xlo[3 * g + n];
      gt[g].dx[n] = A.dx[3 * g + n];
      gt[g].dims[n] = A.dims[3 * g + n];
      gt[g].level = A.level[g];
    }
  // 1. each block's Cartesian points, laid out as node queries are
  DBuf<int> flag, sel;
  std::vector<std::unique_ptr<DBuf<int>>> cells;  // per block: selected cell numbers
  std::vector<std::vector<AmrGridDev>> bg(nbk);
  std::vector<std::vector<long long>> boff(nbk);
  std::vector<std::vector<int>> blid(nbk);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    cells.emplace_back(new DBuf<int>);
    boff[ib].push_back(0);
    for (int g = 0; g < G; g++) {  // the grids whose box meets the block's OBB (getCartReceptors.C:89-101)
      OBB oc{};
      for (int n = 0; n < 3; n++) {
        oc.vec[n][n] = 1.0;
        oc.dxc[n] = A.dx[3 * g + n] * A.dims[3 * g + n] * 0.5;
        oc.xc[n] = A.xlo[3 * g + n] + oc.dxc[n];
      }
      if (!(obb_intersect(b.obb, oc) || obb_intersect(oc, b.obb))) continue;
      bg[ib].push_back(gt[g]);
      blid[ib].push_back(A.lid[g]);
      boff[ib].push_back(boff[ib].back() + (long long)A.dims[3 * g] * A.dims[3 * g + 1] * A.dims[3 * g + 2]);
    }
    const long long tot = boff[ib].back();
    if (tot > 0x7fffffffLL) throw std::runtime_error("tioga_amd: too many Cartesian candidate cells for one block");
    b.nsearch = 0;
    if (!tot) continue;
    DBuf<AmrGridDev> dg;
    DBuf<long long> doff;
    DBuf<int> dlid;
    const int ng = (int)bg[ib].size();
    dg.reserve(ng);
    doff.reserve(ng + 1);
    dlid.reserve(ng);
    TG_HIP(hipMemcpyAsync(dg.p, bg[ib].data(), sizeof(AmrGridDev) * ng, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(doff.p, boff[ib].data(), sizeof(long long) * (ng + 1), hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(dlid.p, blid[ib].data(), sizeof(int) * ng, hipMemcpyHostToDevice, s_));
    const CartGridList L{ng, dg.p, doff.p, dlid.p};
    flag.reserve(tot);
    cells[ib]->reserve(tot);
    k_cart_flags<<<anb(tot), ABS, 0, s_>>>(tot, L, qn0, b.obb, flag.p);
    b.nsearch = select_flagged(flag.p, (int)tot, cells[ib]->p, s_);
  }
  size_t qtot = 0;
  for (Block *b : blocks_) qtot += b->nsearch;
  qx_.reserve(3 * qtot);
  qres_.reserve(qtot);
  qtag_.reserve(qtot);
  qrep_.reserve(qtot);
  size_t qo = 0;
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    const size_t n = b.nsearch;
    b.xsearch.view(qx_.p + 3 * qo, 3 * n);
    b.res_search.view(qres_.p + qo, n);
    b.tagsearch.view(qtag_.p + qo, n);
    b.xtag.view(qrep_.p + qo, n);
    qo += n;
    b.isearch.reserve(3 * std::max(n, (size_t)1));
    b.donorId.reserve(std::max(n, (size_t)1));
    if (!n) continue;
    DBuf<AmrGridDev> dg;
    DBuf<long long> doff;
    DBuf<int> dlid;
    const int ng = (int)bg[ib].size();
    dg.reserve(ng);
    doff.reserve(ng + 1);
    dlid.reserve(ng);
    TG_HIP(hipMemcpyAsync(dg.p, bg[ib].data(), sizeof(AmrGridDev) * ng, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(doff.p, boff[ib].data(), sizeof(long long) * (ng + 1), hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(dlid.p, blid[ib].data(), sizeof(int) * ng, hipMemcpyHostToDevice, s_));
    const CartGridList L{ng, dg.p, doff.p, dlid.p};
    k_cart_points<<<anb(n), ABS, 0, s_>>>((int)n, cells[ib]->p, L, qn0, A.nf, b.xsearch.p, b.isearch.p,
                                          b.res_search.p, b.tagsearch.p);
    sync();
  }
  // 2. the searches (MeshBlock::search)
  if (batched) {
    search_batched();
    launch_searches();
  } else {
    for (Block *b : blocks_) search(*b);
    for (Block *b : blocks_) launch_search(*b);
  }
  // 3. donor lists on the patch cells, processDonors, cancellation
  A.coff.assign(A.npatch + 1, 0);
  A.ioff.assign(A.npatch + 1, 0);
  std::vector<int> pd3(std::max(A.npatch, 1));
  for (int p = 0; p < A.npatch; p++) {
    const int g = A.pgid[p];
    A.coff[p + 1] = A.coff[p] + (long long)A.dims[3 * g] * A.dims[3 * g + 1] * A.dims[3 * g + 2];
    pd3[p] = (A.dims[3 * g] + 2 * A.nf) * (A.dims[3 * g + 1] + 2 * A.nf) * (A.dims[3 * g + 2] + 2 * A.nf);
    A.ioff[p + 1] = A.ioff[p] + pd3[p];
  }
  const long long NC = A.coff[A.npatch], NI = A.ioff[A.npatch];
  if (NC > (1ll << 31) || NI > (1ll << 31)) throw std::runtime_error("tioga_amd: too many AMR patch cells");
  A.dcoff.reserve(A.npatch + 1);
  A.dioff.reserve(A.npatch + 1);
  A.ccnt.reserve(std::max(NC, 1ll));
  A.cgood.reserve(std::max(NC, 1ll));
  A.ccancel.reserve(std::max(NC, 1ll));
  A.ctags.reserve(std::max(NC, 1ll));
  A.pibl_dev.reserve(std::max(NI, 1ll));
  DBuf<int> dpd3;
  dpd3.reserve(std::max(A.npatch, 1));
  TG_HIP(hipMemcpyAsync(A.dcoff.p, A.coff.data(), sizeof(long long) * (A.npatch + 1), hipMemcpyHostToDevice, s_));
  TG_HIP(hipMemcpyAsync(A.dioff.p, A.ioff.data(), sizeof(long long) * (A.npatch + 1), hipMemcpyHostToDevice, s_));
  if (A.npatch) TG_HIP(hipMemcpyAsync(dpd3.p, pd3.data(), sizeof(int) * A.npatch, hipMemcpyHostToDevice, s_));
  if (NC) {
    TG_HIP(hipMemsetAsync(A.ccnt.p, 0, sizeof(int) * NC, s_));
    TG_HIP(hipMemsetAsync(A.cgood.p, 0, sizeof(int) * NC, s_));
    TG_HIP(hipMemsetAsync(A.ccancel.p, 0, sizeof(int) * NC, s_));
    TG_HIP(hipMemsetAsync(A.ctags.p, 0, sizeof(unsigned long long) * NC, s_));
  }
  // the patch iblanks as registered, through one host staging copy each way
  std::vector<int> ibl_host((size_t)std::max(NI, 1ll));
  for (int p = 0; p < A.npatch; p++) {
    if (!A.pibl[p]) throw std::runtime_error("tioga_amd: AMR patch iblank not registered");
    memcpy(ibl_host.data() + A.ioff[p], A.pibl[p], sizeof(int) * pd3[p]);
  }
  if (NI) TG_HIP(hipMemcpyAsync(A.pibl_dev.p, ibl_host.data(), sizeof(int) * NI, hipMemcpyHostToDevice, s_));
  for (Block *bp : blocks_) {
    Block &b = *bp;
    if (b.tag > 64 && b.nsearch) throw std::runtime_error("tioga_amd: AMR donor meshes need tags <= 64");
    if (b.nsearch)
      k_cart_donors<<<anb(b.nsearch), ABS, 0, s_>>>(b.nsearch, b.donorId.p, b.isearch.p, b.cellRes.p, b.tag, dpd3.p,
                                                    A.dcoff.p, A.ppgid.p, A.grids.p, A.nf, A.ccnt.p, A.ctags.p,
                                                    A.cgood.p);
  }
  const CartHoles H{nmesh_, hm_view_.p};
  if (nmesh_ > 64) throw std::runtime_error("tioga_amd: AMR hole maps need mesh tags <= 64");
  if (NC)
    k_cart_process<<<anb(NC), ABS, 0, s_>>>(NC, A.npatch, A.dcoff.p, A.dioff.p, A.ppgid.p, A.grids.p, qn0, A.nf, H,
                                            A.ccnt.p, A.ctags.p, A.cgood.p, A.ccancel.p, A.pibl_dev.p);
  if (NI) TG_HIP(hipMemcpyAsync(ibl_host.data(), A.pibl_dev.p, sizeof(int) * NI, hipMemcpyDeviceToHost, s_));
  sync();
  for (int p = 0; p < A.npatch; p++) memcpy(A.pibl[p], ibl_host.data() + A.ioff[p], sizeof(int) * pd3[p]);
  // 4. findInterpListCart
  A.crec.clear();
  for (int ib = 0; ib < nbk; ib++) A.crec.emplace_back(new AmrState::CartRecs);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    AmrState::CartRecs &R = *A.crec[ib];
    R.n = 0;
    if (!b.nsearch) continue;
    flag.reserve(b.nsearch);
    sel.reserve(b.nsearch);
    k_cart_cancel<<<anb(b.nsearch), ABS, 0, s_>>>(b.nsearch, b.donorId.p, b.isearch.p, dpd3.p, A.dcoff.p, A.ppgid.p,
                                                  A.grids.p, A.nf, A.ccancel.p, flag.p);
    const int n = select_flagged(flag.p, b.nsearch, sel.p, s_);
    R.n = n;
    if (!n) continue;
    R.inode.reserve(8 * (size_t)n);
    R.w.reserve(8 * (size_t)n);
    R.nw.reserve(n);
    R.rpt.reserve(n);
    R.rlp.reserve(n);
    k_cart_records<<<anb(n), ABS, 0, s_>>>(n, sel.p, b.ti, b.conn.p, b.x.p, b.xsearch.p, b.isearch.p, b.donorId.p,
                                           R.inode.p, R.w.p, R.nw.p, R.rpt.p, R.rlp.p);
  }
  sync();
}

void Engine::amr_release()
{
  delete amr_;
  amr_ = nullptr;
}

bool Engine::amr_mode() const { return amr_ && amr_->connected && amr_->npatch > 0; }

// tioga::register_amr_global_data (tioga.C:692-701), CartGrid::registerData
// (CartGrid.C:22-70): 11 ints and 6 doubles per grid
void Engine::amr_global_data(int nf, int qstride, const double *qnode, const int *idata, const double *rdata,
                             int ngrids, int qnodesize)
{
  if (ngrids < 0 || qnodesize < 0) throw std::runtime_error("tioga_amd: negative AMR grid or qnode count");
  AmrState &A = amr();
  A.nf = nf;
  A.qstride = qstride;
  A.ngrids = ngrids;
  A.qnodesize = qnodesize;
  A.qnode.assign(qnode, qnode + qnodesize);
  A.gid.resize(ngrids);
  A.level.resize(ngrids);
  A.proc.resize(ngrids);
  A.porder.resize(ngrids);
  A.lid.resize(ngrids);
  A.dims.resize(3 * (size_t)ngrids);
  A.xlo.resize(3 * (size_t)ngrids);
  A.dx.resize(3 * (size_t)ngrids);
  for (int i = 0; i < ngrids; i++) {
    const int *id = idata + 11 * (size_t)i;
    const double *rd = rdata + 6 * (size_t)i;
    A.gid[i] = id[0];
    A.level[i] = id[1];
    A.proc[i] = id[2];
    A.porder[i] = id[3];
    A.lid[i] = id[4];
    for (int n = 0; n < 3; n++) {
      A.dims[3 * i + n] = id[8 + n] - id[5 + n] + 1;
      A.xlo[3 * i + n] = rd[n];
      A.dx[3 * i + n] = rd[3 + n];
    }
  }
  A.have_grid = true;
  A.connected = false;
}

// tioga::set_amr_patch_count (tioga.C:703-708)
void Engine::amr_patch_count(int npatch)
{
  if (npatch < 0) throw std::runtime_error("tioga_amd: negative AMR patch count");
  AmrState &A = amr();
  A.npatch = npatch;
  A.pgid.assign(npatch, 0);
  A.pibl.assign(npatch, nullptr);
  A.pq.assign(npatch, nullptr);
  A.connected = false;
}

// tioga::register_amr_local_data (tioga.C:710-713)
void Engine::amr_local_data(int ipatch, int global_id, int *iblank, double *q)
{
  AmrState &A = amr();
  if (ipatch < 0 || ipatch >= A.npatch) throw std::runtime_error("tioga_amd: AMR patch index out of range");
  A.pgid[ipatch] = global_id;
  A.pibl[ipatch] = iblank;
  A.pq[ipatch] = q;
}

// tioga::performConnectivityAMR (tioga.C:180-216), gather direction
void Engine::perform_connectivity_amr()
{
  AmrState &A = amr();
  if (!A.have_grid) throw std::runtime_error("tioga_amd: AMR grid data not registered");
  const int G = A.ngrids;
  for (int p = 0; p < A.npatch; p++) {
    const int g = A.pgid[p];
    if (g < 0 || g >= G) throw std::runtime_error("tioga_amd: AMR patch global id out of range");
    if (A.porder[g] != 0)  // CartBlock.C:167-168 needs a user callback for pdegree > 0
      throw std::runtime_error("tioga_amd: AMR patches with pdegree > 0 are not available in this engine");
  }
  for (int g = 0; g < G; g++) {
    if (A.proc[g] != 0)
      throw std::runtime_error("tioga_amd: AMR grid owned by another rank; this engine runs one rank");
    if (A.lid[g] < 0 || A.lid[g] >= A.npatch)
      throw std::runtime_error("tioga_amd: AMR grid local id outside the registered patches");
    for (int n = 0; n < 3; n++)
      if (A.dims[3 * g + n] <= 0 || !(A.dx[3 * g + n] > 0))
        throw std::runtime_error("tioga_amd: AMR grid with empty extent or non-positive spacing");
  }
  // CartGrid::preprocess (CartGrid.C:77-104)
  A.maxlevel = -1;
  for (int g = 0; g < G; g++) {
    if (A.level[g] < 0) throw std::runtime_error("tioga_amd: negative AMR level");
    A.maxlevel = std::max(A.maxlevel, A.level[g]);
  }
  A.maxlevel++;
  // patch table and per-level bins
  std::vector<AmrGridDev> gt(std::max(G, 1));
  for (int g = 0; g < G; g++)
    for (int n = 0; n < 3; n++) {
      gt[g].xlo[n] = A.xlo[3 * g + n];
      gt[g].dx[n] = A.dx[3 * g + n];
      gt[g].dims[n] = A.dims[3 * g + n];
      gt[g].level = A.level[g];
    }
  std::vector<AmrLevelBins> lv(std::max(A.maxlevel, 1));
  std::vector<std::vector<int>> bins;
  for (int l = 0; l < A.maxlevel; l++) {
    AmrLevelBins &L = lv[l];
    L = AmrLevelBins{};
    double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
    double h[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, mag = 1.0;
    int cnt = 0;
    for (int g = 0; g < G; g++) {
      if (A.level[g] != l) continue;
      cnt++;
      for (int n = 0; n < 3; n++) {
        const double a = gt[g].xlo[n], b = a + gt[g].dx[n] * gt[g].dims[n];
        lo[n] = std::min(lo[n], a);
        hi[n] = std::max(hi[n], b);
        h[n] = std::min(h[n], b - a);
        mag = std::max(mag, std::max(std::fabs(a), std::fabs(b)));
      }
    }
    if (cnt == 0) continue;
    // bins of about one patch; widen the grid by a margin far above the
    // test's TOL and the rounding of the coordinates, so that every grid
    // that can pass the containment test for a point lies in its bin
    const double margin = 4 * TOL + 64 * 2.220446049250313e-16 * mag;
    for (int n = 0; n < 3; n++) {
      L.o[n] = lo[n] - 2 * margin;
      double ext = hi[n] - lo[n] + 4 * margin;
      L.h[n] = std::max(h[n], ext / 512);
      L.nb[n] = std::max(1, (int)std::ceil(ext / L.h[n]));
    }
    L.start = (int)bins.size();
    const size_t nbin = (size_t)L.nb[0] * L.nb[1] * L.nb[2];
    if (bins.size() + nbin > (size_t)1 << 28) throw std::runtime_error("tioga_amd: AMR level bin grid too large");
    bins.resize(bins.size() + nbin);
    for (int g = 0; g < G; g++) {
      if (A.level[g] != l) continue;
      int b0[3], b1[3];
      for (int n = 0; n < 3; n++) {
        const double a = gt[g].xlo[n], b = a + gt[g].dx[n] * gt[g].dims[n];
        b0[n] = std::max(0, std::min(L.nb[n] - 1, (int)std::floor((a - L.o[n] - margin) / L.h[n])));
        b1[n] = std::max(0, std::min(L.nb[n] - 1, (int)std::floor((b - L.o[n] + margin) / L.h[n])));
      }
      for (int k = b0[2]; k <= b1[2]; k++)
        for (int j = b0[1]; j <= b1[1]; j++)
          for (int i = b0[0]; i <= b1[0]; i++)
            bins[L.start + ((size_t)k * L.nb[1] + j) * L.nb[0] + i].push_back(g);
    }
  }
  std::vector<int> boff(bins.size() + 1, 0), blist;
  for (size_t b = 0; b < bins.size(); b++) {
    boff[b + 1] = boff[b] + (int)bins[b].size();
    blist.insert(blist.end(), bins[b].begin(), bins[b].end());
  }
  A.grids.reserve(gt.size());
  A.levels.reserve(lv.size());
  A.boff.reserve(boff.size());
  A.blist.reserve(std::max<size_t>(blist.size(), 1));
  A.plid.reserve(std::max(G, 1));
  TG_HIP(hipMemcpyAsync(A.grids.p, gt.data(), sizeof(AmrGridDev) * gt.size(), hipMemcpyHostToDevice, s_));
  TG_HIP(hipMemcpyAsync(A.levels.p, lv.data(), sizeof(AmrLevelBins) * lv.size(), hipMemcpyHostToDevice, s_));
  TG_HIP(hipMemcpyAsync(A.boff.p, boff.data(), sizeof(int) * boff.size(), hipMemcpyHostToDevice, s_));
  if (!blist.empty())
    TG_HIP(hipMemcpyAsync(A.blist.p, blist.data(), sizeof(int) * blist.size(), hipMemcpyHostToDevice, s_));
  if (G) TG_HIP(hipMemcpyAsync(A.plid.p, A.lid.data(), sizeof(int) * G, hipMemcpyHostToDevice, s_));
  // per local patch: registered grid, d3nf and strides
  std::vector<int> pd3nf(std::max(A.npatch, 1)), pstride(2 * std::max(A.npatch, 1));
  for (int p = 0; p < A.npatch; p++) {
    const int g = A.pgid[p];
    const long long sx = A.dims[3 * g] + 2 * A.nf, sy = A.dims[3 * g + 1] + 2 * A.nf, sz = A.dims[3 * g + 2] + 2 * A.nf;
    if (sx * sy * sz > (1ll << 30)) throw std::runtime_error("tioga_amd: AMR patch too large");
    pd3nf[p] = (int)(sx * sy * sz);
    pstride[2 * p] = (int)sx;
    pstride[2 * p + 1] = (int)(sx * sy);
  }
  A.ppgid.reserve(std::max(A.npatch, 1));
  A.pd3nf.reserve(3 * std::max(A.npatch, 1));
  if (A.npatch) {
    TG_HIP(hipMemcpyAsync(A.ppgid.p, A.pgid.data(), sizeof(int) * A.npatch, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(A.pd3nf.p, pd3nf.data(), sizeof(int) * A.npatch, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(A.pd3nf.p + A.npatch, pstride.data(), sizeof(int) * 2 * A.npatch, hipMemcpyHostToDevice,
                          s_));
  }

  // receptors of every block (getUnresolvedMandatoryReceptors), written
  // straight into the concatenated receptor arrays (sized for every node)
  const int nbk = (int)blocks_.size();
  A.rcp_off.assign(nbk + 1, 0);
  size_t maxn = 1, totn = 1;
  for (Block *b : blocks_) {
    maxn = std::max(maxn, (size_t)b->nnodes);
    totn += (size_t)b->nnodes;
  }
  if (totn > (size_t)1 << 30) throw std::runtime_error("tioga_amd: too many AMR receptor candidates");
  A.picked.reserve(maxn);
  iscratch_.reserve(maxn);
  iscratch2_.reserve(maxn);
  A.rnode.reserve(totn);
  A.rblk.reserve(totn);
  A.rdonor.reserve(totn);
  A.rxyz.reserve(3 * totn);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (!b.topo_uploaded || b.nodeRes.p == nullptr)
      throw std::runtime_error("tioga_amd: AMR connectivity needs preprocess (tioga_preprocess_grids) first");
    // the cell iblank the reference reads: the user's array, else the one
    // the last connectivity computed; with neither every cell counts as 1
    const int *ibc = b.uiblank_cell ? b.uiblank_cell : (b.own_iblank_cell.empty() ? nullptr : b.own_iblank_cell.data());
    std::vector<int> ones;
    b.iblank_cell.reserve(b.ncells);
    if (b.ncells) {
      if (!ibc) {
        ones.assign(b.ncells, 1);
        ibc = ones.data();
      }
      TG_HIP(hipMemcpyAsync(b.iblank_cell.p, ibc, sizeof(int) * b.ncells, hipMemcpyHostToDevice, s_));
    }
    int nc = 0;
    if (b.nnodes) {
      TG_HIP(hipMemsetAsync(A.picked.p, 0, b.nnodes, s_));
      if (b.ncells)
        k_amr_pick<<<anb(b.ncells), ABS, 0, s_>>>(b.ncells, b.ti, b.conn.p, b.nodeRes.p, b.iblank_cell.p, A.picked.p);
      k_amr_flags<<<anb(b.nnodes), ABS, 0, s_>>>(b.nnodes, A.picked.p, iscratch_.p);
      nc = select_flagged(iscratch_.p, b.nnodes, iscratch2_.p, s_);
    }
    const int o = A.rcp_off[ib];
    A.rcp_off[ib + 1] = o + nc;
    if (nc)
      k_amr_receptors<<<anb(nc), ABS, 0, s_>>>(nc, iscratch2_.p, b.x.p, ib, A.rnode.p + o, A.rblk.p + o,
                                                A.rxyz.p + 3 * (size_t)o);
    sync();
  }
  const int R = A.nrcp = A.rcp_off[nbk];
  // CartGrid::search
  if (R)
    k_amr_search<<<anb(R), ABS, 0, s_>>>(R, A.rxyz.p, A.maxlevel, A.levels.p, A.boff.p, A.blist.p, A.grids.p,
                                          A.rdonor.p);
  // records: receptors with a donor, stable-sorted by local patch
  A.nrec = 0;
  if (R) {
    uscratch_.reserve(R);
    uscratch2_.reserve(R);
    iscratch_.reserve(R);
    iscratch2_.reserve(R);
    k_amr_keys<<<anb(R), ABS, 0, s_>>>(R, A.rdonor.p, A.plid.p, uscratch_.p, iscratch_.p);
    const int nr = select_flagged(iscratch_.p, R, iscratch2_.p, s_);
    A.nrec = nr;
    if (nr) {
      DBuf<unsigned> keys, keys2;
      keys.reserve(nr);
      keys2.reserve(nr);
      k_amr_take<<<anb(nr), ABS, 0, s_>>>(nr, iscratch2_.p, uscratch_.p, keys.p);
      int bits = 1;
      while (bits < 32 && ((unsigned)std::max(A.npatch - 1, 0) >> bits)) bits++;
      A.rec_rid.reserve(nr);
      sort_pairs(keys.p, keys2.p, iscratch2_.p, A.rec_rid.p, nr, bits, s_);
      A.rec_lp.reserve(nr);
      A.rec_node.reserve(nr);
      A.rec_meta.reserve(nr);
      A.rec_base.reserve(nr);
      A.rec_phi.reserve(6 * (size_t)nr);
      A.err.reserve(1);
      TG_HIP(hipMemsetAsync(A.err.p, 0, sizeof(int), s_));
      k_amr_records<<<anb(nr), ABS, 0, s_>>>(nr, A.rec_rid.p, keys2.p, A.rxyz.p, A.rnode.p, A.rblk.p, A.ppgid.p,
                                             A.grids.p, A.nf, A.rec_lp.p, A.rec_node.p, A.rec_meta.p, A.rec_base.p,
                                             A.rec_phi.p, A.err.p);
      int e = 0;
      TG_HIP(hipMemcpyAsync(&e, A.err.p, sizeof(int), hipMemcpyDeviceToHost, s_));
      std::vector<unsigned> lps(nr);
      TG_HIP(hipMemcpyAsync(lps.data(), keys2.p, sizeof(unsigned) * nr, hipMemcpyDeviceToHost, s_));
      sync();
      if (e)
        throw std::runtime_error(
            "tioga_amd: an AMR receptor lies outside its donor patch's cell range (the reference asserts, "
            "CartBlock.C:246)");
      A.rec_count.assign(A.npatch, 0);
      for (unsigned v : lps) A.rec_count[v]++;
      // Gather order: by patch, then by stencil base index, so that the lanes
      // of a wave read neighbouring cells of one patch.  Every record writes
      // its own receptor, so the order leaves the results unchanged; the
      // list order stays recoverable as (patch, receptor index).
      int maxd3 = 1;
      for (int p = 0; p < A.npatch; p++) maxd3 = std::max(maxd3, pd3nf[p]);
      int bbits = 1;
      while (bbits < 32 && ((unsigned)(maxd3 - 1) >> bbits)) bbits++;
      DBuf<int> perm0, perm1;
      perm0.reserve(nr);
      perm1.reserve(nr);
      k_amr_iota<<<anb(nr), ABS, 0, s_>>>(nr, perm0.p);
      k_amr_key_of<<<anb(nr), ABS, 0, s_>>>(nr, perm0.p, A.rec_base.p, keys.p);
      sort_pairs(keys.p, keys2.p, perm0.p, perm1.p, nr, bbits, s_);
      k_amr_key_of<<<anb(nr), ABS, 0, s_>>>(nr, perm1.p, A.rec_lp.p, keys.p);
      sort_pairs(keys.p, keys2.p, perm1.p, perm0.p, nr, bits, s_);
      DBuf<int> b0, b1, b2, b3, b4;
      DBuf<double> bph;
      b0.reserve(nr);
      b1.reserve(nr);
      b2.reserve(nr);
      b3.reserve(nr);
      b4.reserve(nr);
      bph.reserve(6 * (size_t)nr);
      k_amr_permute<<<anb(nr), ABS, 0, s_>>>(nr, perm0.p, A.rec_lp.p, A.rec_node.p, A.rec_meta.p, A.rec_base.p,
                                             A.rec_rid.p, A.rec_phi.p, b0.p, b1.p, b2.p, b3.p, b4.p, bph.p);
      sync();
      auto take = [](auto &dst, auto &src) {
        std::swap(dst.p, src.p);
        std::swap(dst.cap, src.cap);
      };
      take(A.rec_lp, b0);
      take(A.rec_node, b1);
      take(A.rec_meta, b2);
      take(A.rec_base, b3);
      take(A.rec_rid, b4);
      take(A.rec_phi, bph);
    }
  }
  if (A.nrec == 0) A.rec_count.assign(A.npatch, 0);
  // the other direction: patch cells that take values from the near-body blocks
  amr_cart_receptors();
  // setCartIblanks, getCellIblanks, write back
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (R) k_amr_iblank<<<anb(R), ABS, 0, s_>>>(R, A.rdonor.p, A.rnode.p, A.rblk.p, ib, b.iblank.p);
    cell_iblanks(b);
  }
  for (Block *bp : blocks_) {
    Block &b = *bp;
    TG_HIP(hipMemcpyAsync(b.uiblank, b.iblank.p, sizeof(int) * b.nnodes, hipMemcpyDeviceToHost, s_));
    int *ibc = b.uiblank_cell;
    if (!ibc) {
      b.own_iblank_cell.resize(std::max(b.ncells, 1));
      ibc = b.own_iblank_cell.data();
    }
    TG_HIP(hipMemcpyAsync(ibc, b.iblank_cell.p, sizeof(int) * b.ncells, hipMemcpyDeviceToHost, s_));
  }
  sync();
  A.connected = true;
}

void Engine::amr_upload_q(int nvar)
{
  AmrState &A = amr();
  std::vector<long long> off(std::max(A.npatch, 1), 0);
  long long tot = 0;
  for (int p = 0; p < A.npatch; p++) {
    off[p] = tot;
    if (A.rec_count.size() == (size_t)A.npatch && A.rec_count[p] > 0) {
      const int g = A.pgid[p];
      const long long d3 = (long long)(A.dims[3 * g] + 2 * A.nf) * (A.dims[3 * g + 1] + 2 * A.nf) *
                           (A.dims[3 * g + 2] + 2 * A.nf);
      tot += d3 * nvar;
    }
  }
  A.q.reserve((size_t)std::max(tot, 1ll));
  A.qoff.reserve(off.size());
  TG_HIP(hipMemcpyAsync(A.qoff.p, off.data(), sizeof(long long) * off.size(), hipMemcpyHostToDevice, s_));
  for (int p = 0; p < A.npatch; p++) {
    if (!(A.rec_count.size() == (size_t)A.npatch && A.rec_count[p] > 0)) continue;
    if (!A.pq[p]) throw std::runtime_error("tioga_amd: AMR patch solution not registered");
    const int g = A.pgid[p];
    const long long d3 = (long long)(A.dims[3 * g] + 2 * A.nf) * (A.dims[3 * g + 1] + 2 * A.nf) *
                         (A.dims[3 * g + 2] + 2 * A.nf);
    TG_HIP(hipMemcpyAsync(A.q.p + off[p], A.pq[p], sizeof(double) * d3 * nvar, hipMemcpyHostToDevice, s_));
  }
}

void Engine::amr_tables()
{
  AmrState &A = amr();
  AmrQTables qt{};
  for (size_t ib = 0; ib < blocks_.size(); ib++) {
    qt.q[ib] = qdev_[ib]->p;
    qt.nnodes[ib] = blocks_[ib]->nnodes;
  }
  A.qt.reserve(1);
  TG_HIP(hipMemcpyAsync(A.qt.p, &qt, sizeof qt, hipMemcpyHostToDevice, s_));
  sync();
}

void Engine::amr_gather(int nvar, int interptype)
{
  AmrState &A = amr();
  if (A.nrec == 0) return;
  k_amr_gather<<<anb(A.nrec), ABS, 0, s_>>>(A.nrec, nvar, interptype, A.rec_lp.p, A.rec_node.p, A.rec_meta.p,
                                             A.rec_base.p, A.rec_phi.p, A.q.p, A.qoff.p, A.pd3nf.p,
                                             A.pd3nf.p + A.npatch, A.qt.p);
}

// tioga::dataUpdate_AMR (tioga.C:218-334) on one rank
void Engine::data_update_amr(int nvar, int interptype)
{
  for (Block *b : blocks_)
    if (b->uq == nullptr) {
      fprintf(stderr, "Solution data not set, cannot update \n");  // tioga.C:236
      return;
    }
  if (!amr_mode() || nvar <= 0) return;
  AmrState &A = amr();
  const int nbk = (int)blocks_.size();
  while ((int)qdev_.size() < nbk) qdev_.push_back(new DBuf<double>);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    size_t n = (size_t)b.nnodes * nvar;
    qdev_[ib]->reserve(n);
    TG_HIP(hipMemcpyAsync(qdev_[ib]->p, b.uq, sizeof(double) * n, hipMemcpyHostToDevice, s_));
  }
  amr_upload_q(nvar);
  amr_tables();
  // the Cartesian receptors' values, from the near-body solutions before the
  // gather writes their receptors (every value is formed before any is
  // written, as the packets of tioga.C:286-321 are)
  size_t ncv = 0;
  std::vector<size_t> cv0(A.crec.size() + 1, 0);
  for (size_t ib = 0; ib < A.crec.size(); ib++) cv0[ib + 1] = cv0[ib] + (size_t)A.crec[ib]->n * nvar;
  ncv = cv0.back();
  DBuf<double> cvals;
  cvals.reserve(std::max<size_t>(ncv, 1));
  for (size_t ib = 0; ib < A.crec.size(); ib++) {
    const AmrState::CartRecs &R = *A.crec[ib];
    if (R.n)
      k_cart_values<<<anb(R.n), ABS, 0, s_>>>(R.n, nvar, interptype == 1, blocks_[ib]->nnodes, qdev_[ib]->p,
                                              R.inode.p, R.w.p, R.nw.p, cvals.p + cv0[ib]);
  }
  if (A.nrec) amr_gather(nvar, interptype);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    TG_HIP(hipMemcpyAsync(b.uq, qdev_[ib]->p, sizeof(double) * (size_t)b.nnodes * nvar, hipMemcpyDeviceToHost, s_));
  }
  A.cvals.resize(ncv);
  if (ncv) TG_HIP(hipMemcpyAsync(A.cvals.data(), cvals.p, sizeof(double) * ncv, hipMemcpyDeviceToHost, s_));
  sync();
  // CartBlock::update (CartBlock.C:146-151): block after block, record after record
  for (size_t ib = 0; ib < A.crec.size(); ib++) {
    const AmrState::CartRecs &R = *A.crec[ib];
    if (!R.n) continue;
    std::vector<int> rpt(R.n), rlp(R.n);
    TG_HIP(hipMemcpy(rpt.data(), R.rpt.p, sizeof(int) * R.n, hipMemcpyDeviceToHost));
    TG_HIP(hipMemcpy(rlp.data(), R.rlp.p, sizeof(int) * R.n, hipMemcpyDeviceToHost));
    for (int r = 0; r < R.n; r++) {
      const int p = rlp[r], g = A.pgid[p];
      const size_t d3 = (size_t)(A.dims[3 * g] + 2 * A.nf) * (A.dims[3 * g + 1] + 2 * A.nf) * (A.dims[3 * g + 2] + 2 * A.nf);
      if (!A.pq[p]) throw std::runtime_error("tioga_amd: AMR patch solution not registered");
      for (int k = 0; k < nvar; k++) A.pq[p][(size_t)rpt[r] + d3 * k] = A.cvals[cv0[ib] + (size_t)r * nvar + k];
    }
  }
}

// Device-resident dataUpdate_AMR: q of the blocks and patches uploaded once,
// then `nrep` gathers; returns the mean device ms per update.
double Engine::time_data_update_amr(int nvar, int interptype, int nrep, long long *nrec)
{
  AmrState &A = amr();
  if (nrec) *nrec = A.nrec;
  if (!amr_mode() || nvar <= 0 || nrep <= 0) return 0.0;
  for (Block *b : blocks_)
    if (b->uq == nullptr) throw std::runtime_error("tioga_amd: solution data not registered");
  const int nbk = (int)blocks_.size();
  while ((int)qdev_.size() < nbk) qdev_.push_back(new DBuf<double>);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    size_t n = (size_t)b.nnodes * nvar;
    qdev_[ib]->reserve(n);
    TG_HIP(hipMemcpyAsync(qdev_[ib]->p, b.uq, sizeof(double) * n, hipMemcpyHostToDevice, s_));
  }
  amr_upload_q(nvar);
  amr_tables();
  amr_gather(nvar, interptype);  // warm
  sync();
  hipEvent_t e0, e1;
  TG_HIP(hipEventCreate(&e0));
  TG_HIP(hipEventCreate(&e1));
  TG_HIP(hipEventRecord(e0, s_));
  for (int i = 0; i < nrep; i++) amr_gather(nvar, interptype);
  TG_HIP(hipEventRecord(e1, s_));
  TG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  TG_HIP(hipEventElapsedTime(&ms, e0, e1));
  TG_HIP(hipEventDestroy(e0));
  TG_HIP(hipEventDestroy(e1));
  return ms / nrep;
}

// Inspection, under the reference's names where it has them:
//   per block (btag >= 1): "receptorIdCart", "donorIdCart", "rxyzCart"
//   records (btag 0), in list order: "rec_patch", "rec_info" (rank, point,
//   block), "rec_ijk" (8 stencil cells, i j k each), "rec_w" (8 weights),
//   "rec_unique_cells" (number of distinct (patch, cell) donors)
long Engine::amr_array(int btag, const std::string &name, void *out)
{
  if (!amr_ || !amr_->connected) return 0;
  AmrState &A = *amr_;
  if (name == "patch_iblank") {  // btag: the local patch index
    if (btag < 0 || btag >= A.npatch) return -1;
    const long long n = A.ioff.size() > (size_t)btag + 1 ? A.ioff[btag + 1] - A.ioff[btag] : 0;
    if (out && n) memcpy(out, A.pibl[btag], sizeof(int) * n);
    return (long)n;
  }
  if (btag >= 1 && name.compare(0, 11, "interpCart_") == 0) {  // MeshBlock::interpListCart
    int ib = -1;
    if (!find(btag, &ib)) return -1;
    if ((size_t)ib >= A.crec.size()) return 0;
    const AmrState::CartRecs &R = *A.crec[ib];
    const int n = R.n;
    std::vector<int> nw(n), inode(8 * (size_t)n), rpt(n), rlp(n);
    std::vector<double> w(8 * (size_t)n);
    if (n) {
      TG_HIP(hipMemcpy(nw.data(), R.nw.p, sizeof(int) * n, hipMemcpyDeviceToHost));
      TG_HIP(hipMemcpy(inode.data(), R.inode.p, sizeof(int) * 8 * n, hipMemcpyDeviceToHost));
      TG_HIP(hipMemcpy(w.data(), R.w.p, sizeof(double) * 8 * n, hipMemcpyDeviceToHost));
      TG_HIP(hipMemcpy(rpt.data(), R.rpt.p, sizeof(int) * n, hipMemcpyDeviceToHost));
      TG_HIP(hipMemcpy(rlp.data(), R.rlp.p, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    if (name == "interpCart_rinfo") {  // (rank, q index, local patch) as cartOps.C:258-260 stores them
      for (int r = 0; out && r < n; r++) {
        ((int *)out)[3 * r] = 0;
        ((int *)out)[3 * r + 1] = rpt[r];
        ((int *)out)[3 * r + 2] = rlp[r];
      }
      return 3l * n;
    }
    long k = 0;
    for (int r = 0; r < n; r++)
      for (int m = 0; m < nw[r]; m++, k++)
        if (out) {
          if (name == "interpCart_inode") ((int *)out)[k] = inode[(size_t)m * n + r];
          else if (name == "interpCart_w") ((double *)out)[k] = w[(size_t)m * n + r];
        }
    return name == "interpCart_inode" || name == "interpCart_w" ? k : -1;
  }
  if (btag >= 1) {
    int ib = -1;
    Block *b = find(btag, &ib);
    if (!b) return -1;
    const int o = A.rcp_off[ib], n = A.rcp_off[ib + 1] - o;
    if (name == "receptorIdCart" || name == "donorIdCart") {
      if (out && n)
        TG_HIP(hipMemcpy(out, (name == "receptorIdCart" ? A.rnode.p : A.rdonor.p) + o, sizeof(int) * n,
                         hipMemcpyDeviceToHost));
      return n;
    }
    if (name == "rxyzCart") {
      if (out && n) TG_HIP(hipMemcpy(out, A.rxyz.p + 3 * (size_t)o, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
      return 3 * (long)n;
    }
    return -1;
  }
  const int n = A.nrec;
  const long want = name == "rec_patch" ? n : name == "rec_info" ? 3l * n : name == "rec_ijk" ? 24l * n
                  : name == "rec_w" ? 8l * n : name == "rec_unique_cells" ? 1 : -1;
  if (want < 0 || !out || n == 0) return want < 0 ? -1 : (n == 0 && name != "rec_unique_cells" ? 0 : want);
  std::vector<int> glp(n), gmeta(n), gbase(n), grid(n);
  std::vector<double> gphi(6 * (size_t)n);
  TG_HIP(hipMemcpy(glp.data(), A.rec_lp.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(gmeta.data(), A.rec_meta.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(gbase.data(), A.rec_base.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(grid.data(), A.rec_rid.p, sizeof(int) * n, hipMemcpyDeviceToHost));
  TG_HIP(hipMemcpy(gphi.data(), A.rec_phi.p, sizeof(double) * gphi.size(), hipMemcpyDeviceToHost));
  // records are stored in gather order; the list order is (patch, receptor)
  std::vector<int> ord(n);
  for (int r = 0; r < n; r++) ord[r] = r;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) {
    return glp[a] != glp[b] ? glp[a] < glp[b] : grid[a] < grid[b];
  });
  std::vector<int> lp(n), meta(n), base(n), rid(n);
  std::vector<double> phi(6 * (size_t)n);
  for (int r = 0; r < n; r++) {
    const int g = ord[r];
    lp[r] = glp[g];
    meta[r] = gmeta[g];
    base[r] = gbase[g];
    rid[r] = grid[g];
    for (int m = 0; m < 6; m++) phi[(size_t)m * n + r] = gphi[(size_t)m * n + g];
  }
  if (name == "rec_patch") {
    std::copy(lp.begin(), lp.end(), (int *)out);
    return n;
  }
  if (name == "rec_info") {
    int *o = (int *)out;
    for (int r = 0; r < n; r++) {
      const int blk = meta[r] >> 3;
      o[3 * r] = 0;
      o[3 * r + 1] = rid[r] - A.rcp_off[blk];
      o[3 * r + 2] = blk;
    }
    return 3l * n;
  }
  if (name == "rec_w") {
    double *o = (double *)out;
    for (int r = 0; r < n; r++) {
      const double px[2] = {phi[r], phi[(size_t)n + r]}, py[2] = {phi[2 * (size_t)n + r], phi[3 * (size_t)n + r]},
                   pz[2] = {phi[4 * (size_t)n + r], phi[5 * (size_t)n + r]};
      int m = 0;
      for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
          for (int i = 0; i < 2; i++) o[8 * (size_t)r + m++] = pz[k] * py[j] * px[i];
    }
    return 8l * n;
  }
  // stencil cells from the base index and the per-axis step bits
  std::vector<long long> cells;
  int *o = (int *)out;
  for (int r = 0; r < n; r++) {
    const int g = A.pgid[lp[r]];
    const int sx = A.dims[3 * g] + 2 * A.nf, sy = A.dims[3 * g + 1] + 2 * A.nf;
    const int i0 = base[r] % sx - A.nf, j0 = (base[r] / sx) % sy - A.nf, k0 = base[r] / (sx * sy) - A.nf;
    const int bx = meta[r] & 1, by = (meta[r] >> 1) & 1, bz = (meta[r] >> 2) & 1;
    int m = 0;
    for (int k = 0; k < 2; k++)
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++, m++) {
          const int ci = i0 + i * bx, cj = j0 + j * by, ck = k0 + k * bz;
          if (name == "rec_ijk") {
            o[24 * (size_t)r + 3 * m] = ci;
            o[24 * (size_t)r + 3 * m + 1] = cj;
            o[24 * (size_t)r + 3 * m + 2] = ck;
          } else {
            cells.push_back(((long long)lp[r] << 32) | (long long)((ck + A.nf) * sy * sx + (cj + A.nf) * sx + ci + A.nf));
          }
        }
  }
  if (name == "rec_unique_cells") {
    std::sort(cells.begin(), cells.end());
    *o = (int)(std::unique(cells.begin(), cells.end()) - cells.begin());
    return 1;
  }
  return 24l * n;
}

}  // namespace tg
