// tg_points.hip -- tioga::performConnectivityHighOrder (tioga.C:162-178) and
// tioga::dataUpdate(..., at_points = 1) (tioga.C:336-524) for a solver that
// registers no high-order callbacks: the receptor nodes of the last
// connectivity are searched again as points, against donor cells whose cell
// iblank is 1, and a point update fills them from those donors.
//
// Stages, each for every block:
//   1. MeshBlock::getInternalNodes (highOrder.C:219-263, ihigh = 0): the
//      picked points are the nodes with iblank -1, in node order;
//   2. exchangeSearchData(1) / MeshBlock::getExtraQueryPoints
//      (highOrder.C:266-313): per OBB pair, the sender's points inside the
//      reduced OBB it received (|xd| <= dxc on every axis, xd the point in
//      that box's frame), in point order, with resolution BIGVALUE; the
//      receiver's search arrays are laid out as for node queries;
//   3. MeshBlock::search: the engine's search (batched or block by block);
//   4. MeshBlock::processPointDonors (highOrder.C:315-419, ihigh = 0): a
//      record per search point whose donor cell has cell iblank 1, with the
//      isoparametric weights of the point in that cell;
//   5. the point update: every record's values, sum over the cell's vertices
//      of q times weight in vertex order (highOrder.C:477-513), land on its
//      point (records of later donor blocks overwrite earlier ones, as the
//      packets arrive, tioga.C:442-460); then per receptor block
//      MeshBlock::clearOrphans (highOrder.C:193-206: a picked node no record
//      reached gets iblank 1) and MeshBlock::updatePointData
//      (highOrder.C:544-556: a picked node still at -1 takes its values).
#include <algorithm>
#include <stdexcept>

#include "tg_engine.h"

namespace tg {

namespace {

constexpr int PBS = 256;
inline unsigned pnb(long long n) { return (unsigned)((n + PBS - 1) / PBS); }

__global__ void k_pick(int n, const int *__restrict__ iblank, int *__restrict__ flag)
{
  const int i = blockIdx.x * PBS + threadIdx.x;
  if (i < n) flag[i] = iblank[i] == -1;
}

__global__ void k_gather_pts(int n, const int *__restrict__ node, const double *__restrict__ x,
                             double *__restrict__ px)
{
  const int i = blockIdx.x * PBS + threadIdx.x;
  if (i >= n) return;
  const size_t v = node[i];
  px[3 * (size_t)i] = x[3 * v];
  px[3 * (size_t)i + 1] = x[3 * v + 1];
  px[3 * (size_t)i + 2] = x[3 * v + 2];
}

// highOrder.C:282-297: the point in the box's frame, axis by axis, in the
// reference's summation order
__global__ void k_in_box(int n, const double *__restrict__ px, OBB b, int *__restrict__ flag)
{
  const int i = blockIdx.x * PBS + threadIdx.x;
  if (i >= n) return;
  double xd[3] = {0, 0, 0};
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) xd[j] += (px[3 * (size_t)i + k] - b.xc[k]) * b.vec[j][k];
  flag[i] = fabs(xd[0]) <= b.dxc[0] && fabs(xd[1]) <= b.dxc[1] && fabs(xd[2]) <= b.dxc[2];
}

__global__ void k_fill_points(int n, const int *__restrict__ sel, const double *__restrict__ px, int sblk, int stag,
                              double *__restrict__ xs, double *__restrict__ rs, int *__restrict__ is,
                              int *__restrict__ ts)
{
  const int j = blockIdx.x * PBS + threadIdx.x;
  if (j >= n) return;
  const int p = sel[j];
  xs[3 * (size_t)j] = px[3 * (size_t)p];
  xs[3 * (size_t)j + 1] = px[3 * (size_t)p + 1];
  xs[3 * (size_t)j + 2] = px[3 * (size_t)p + 2];
  rs[j] = BIGVALUE;
  is[3 * (size_t)j] = 0;
  is[3 * (size_t)j + 1] = p;
  is[3 * (size_t)j + 2] = sblk;
  ts[j] = stag;
}

__global__ void k_point_flags(int n, const int *__restrict__ donor, const int *__restrict__ ibc, int *__restrict__ flag)
{
  const int i = blockIdx.x * PBS + threadIdx.x;
  if (i < n) flag[i] = donor[i] > -1 && ibc[donor[i]] == 1;
}

struct PointRecs {
  int n;
  int *inode, *nw, *rpt, *rblk;  // inode[m * n + r]
  double *w;                     // w[m * n + r]
};

template <int NV>
__device__ void point_record(const PointRecs &R, int r, int c, const int *conn, const double *x, const double xp[3])
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
    R.inode[(size_t)m * R.n + r] = cv[m];
    R.w[(size_t)m * R.n + r] = frac[m];
  }
  R.nw[r] = NV;
}

__global__ void k_point_records(PointRecs R, const int *__restrict__ sel, TypeInfo ti, const int *__restrict__ conn,
                                const double *__restrict__ x, const double *__restrict__ xs,
                                const int *__restrict__ is, const int *__restrict__ donor)
{
  const int r = blockIdx.x * PBS + threadIdx.x;
  if (r >= R.n) return;
  const int i = sel[r], c = donor[i];
  int t = 0;
  while (t < ti.ntypes - 1 && c >= ti.cstart[t + 1]) t++;
  const double xp[3] = {xs[3 * (size_t)i], xs[3 * (size_t)i + 1], xs[3 * (size_t)i + 2]};
  switch (ti.nv[t]) {
    case 4: point_record<4>(R, r, c, conn, x, xp); break;
    case 5: point_record<5>(R, r, c, conn, x, xp); break;
    case 6: point_record<6>(R, r, c, conn, x, xp); break;
    default: point_record<8>(R, r, c, conn, x, xp); break;
  }
  R.rpt[r] = is[3 * (size_t)i + 1];
  R.rblk[r] = is[3 * (size_t)i + 2];
}

struct PointDest {
  double *qtmp[128];
  int *itmp[128];
};

// one donor block's records: values in the reference's order (per field,
// vertices m = 0 .. nw-1), written to their points' staging
__global__ void k_point_values(PointRecs R, const double *__restrict__ q, int nnodes, int nvar, int column,
                               PointDest D)
{
  const int r = blockIdx.x * PBS + threadIdx.x;
  if (r >= R.n) return;
  const int nw = R.nw[r], b = R.rblk[r], p = R.rpt[r];
  double *out = D.qtmp[b] + (size_t)p * nvar;
  for (int k = 0; k < nvar; k++) {
    double qq = 0;
    for (int m = 0; m < nw; m++) {
      const size_t v = R.inode[(size_t)m * R.n + r];
      qq += (column ? q[(size_t)k * nnodes + v] : q[v * nvar + k]) * R.w[(size_t)m * R.n + r];
    }
    out[k] = qq;
  }
  D.itmp[b][p] = 1;
}

__global__ void k_point_apply(int n, const int *__restrict__ node, const int *__restrict__ itmp,
                              const double *__restrict__ qtmp, int nvar, int column, int nnodes, int *__restrict__ iblank,
                              double *__restrict__ q)
{
  const int m = blockIdx.x * PBS + threadIdx.x;
  if (m >= n) return;
  const int v = node[m];
  if (itmp[m] == 0) {
    iblank[v] = 1;
    return;
  }
  if (iblank[v] != -1) return;
  for (int k = 0; k < nvar; k++) {
    const double val = qtmp[(size_t)m * nvar + k];
    if (column) q[(size_t)k * nnodes + v] = val;
    else q[(size_t)v * nvar + k] = val;
  }
}

template <class T>
void fill_dev(T *p, size_t n, T v, hipStream_t s);

template <class T>
__global__ void k_fill_dev(size_t n, T *p, T v)
{
  const size_t i = blockIdx.x * (size_t)PBS + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class T>
void fill_dev(T *p, size_t n, T v, hipStream_t s)
{
  if (n) k_fill_dev<T><<<pnb((long long)n), PBS, 0, s>>>(n, p, v);
}

}  // namespace

void Engine::perform_connectivity_points()
{
  if (!connected_) throw std::runtime_error("tioga_amd: performConnectivityHighOrder needs performconnectivity first");
  const int nbk = (int)blocks_.size();
  // 1. picked points
  DBuf<int> flag;
  for (Block *bp : blocks_) {
    Block &b = *bp;
    b.pnode.reserve(std::max(b.nnodes, 1));
    flag.reserve(std::max(b.nnodes, 1));
    b.npts = 0;
    if (b.nnodes) {
      k_pick<<<pnb(b.nnodes), PBS, 0, s_>>>(b.nnodes, b.iblank.p, flag.p);
      b.npts = select_flagged(flag.p, b.nnodes, b.pnode.p, s_);
    }
    b.pxyz.reserve(3 * (size_t)std::max(b.npts, 1));
    if (b.npts) k_gather_pts<<<pnb(b.npts), PBS, 0, s_>>>(b.npts, b.pnode.p, b.x.p, b.pxyz.p);
  }
  // 2. the points each pair sends, and where they land
  const int np = (int)pairs_.size();
  std::vector<int> cnt(np, 0), off(np + 1, 0), dest(np, -1);
  size_t cap = 0;
  for (int i = 0; i < np; i++) cap += blocks_[pairs_[i].local]->npts;
  plist_.reserve(std::max<size_t>(cap, 1));
  for (Block *b : blocks_) b->nsearch = 0;
  for (int i = 0; i < np; i++) {
    Block &A = *blocks_[pairs_[i].local];
    off[i + 1] = off[i];
    if (A.npts == 0) continue;
    flag.reserve(A.npts);
    k_in_box<<<pnb(A.npts), PBS, 0, s_>>>(A.npts, A.pxyz.p, pairs_[i].obc, flag.p);
    cnt[i] = select_flagged(flag.p, A.npts, plist_.p + off[i], s_);
    off[i + 1] = off[i] + cnt[i];
  }
  for (int i = 0; i < np; i++) {
    dest[i] = recv_pair(i);
    if (dest[i] >= 0) blocks_[pairs_[dest[i]].local]->nsearch += cnt[i];
  }
  size_t qtot = 0;
  for (Block *b : blocks_) qtot += b->nsearch;
  qx_.reserve(3 * qtot);
  qres_.reserve(qtot);
  qtag_.reserve(qtot);
  qrep_.reserve(qtot);
  size_t qo = 0;
  for (Block *bp : blocks_) {
    Block &b = *bp;
    const size_t n = b.nsearch;
    b.xsearch.view(qx_.p + 3 * qo, 3 * n);
    b.res_search.view(qres_.p + qo, n);
    b.tagsearch.view(qtag_.p + qo, n);
    b.xtag.view(qrep_.p + qo, n);
    qo += n;
    b.isearch.reserve(3 * std::max(n, (size_t)1));
    b.donorId.reserve(std::max(n, (size_t)1));
  }
  std::vector<int> at(nbk, 0);
  for (int i = 0; i < np; i++) {
    const int ii = dest[i];
    if (ii < 0 || cnt[i] == 0) continue;
    const Block &A = *blocks_[pairs_[i].local];
    const int rb = pairs_[ii].local;
    Block &R = *blocks_[rb];
    const int o = at[rb];
    k_fill_points<<<pnb(cnt[i]), PBS, 0, s_>>>(cnt[i], plist_.p + off[i], A.pxyz.p, pairs_[ii].iblk_remote,
                                               pairs_[ii].tag_remote, R.xsearch.p + 3 * (size_t)o,
                                               R.res_search.p + o, R.isearch.p + 3 * (size_t)o, R.tagsearch.p + o);
    at[rb] += cnt[i];
  }
  sync();
  // 3. the search
  if (batched) {
    search_batched();
    launch_searches();
  } else {
    for (Block *b : blocks_) search(*b);
    for (Block *b : blocks_) launch_search(*b);
  }
  // 4. point donors
  for (Block *bp : blocks_) {
    Block &D = *bp;
    D.ninterp2 = 0;
    if (D.nsearch == 0) continue;
    flag.reserve(D.nsearch);
    psel_.reserve(D.nsearch);
    k_point_flags<<<pnb(D.nsearch), PBS, 0, s_>>>(D.nsearch, D.donorId.p, D.iblank_cell.p, flag.p);
    const int n2 = select_flagged(flag.p, D.nsearch, psel_.p, s_);
    D.ninterp2 = n2;
    if (!n2) continue;
    D.p_inode.reserve(8 * (size_t)n2);
    D.p_w.reserve(8 * (size_t)n2);
    D.p_nw.reserve(n2);
    D.p_rpt.reserve(n2);
    D.p_rblk.reserve(n2);
    PointRecs R{n2, D.p_inode.p, D.p_nw.p, D.p_rpt.p, D.p_rblk.p, D.p_w.p};
    k_point_records<<<pnb(n2), PBS, 0, s_>>>(R, psel_.p, D.ti, D.conn.p, D.x.p, D.xsearch.p, D.isearch.p,
                                             D.donorId.p);
  }
  sync();
  points_ready_ = true;
}

void Engine::data_update_points(int nvar, int interptype)
{
  if (!points_ready_) throw std::runtime_error("tioga_amd: dataUpdate at points needs performConnectivityHighOrder");
  for (Block *b : blocks_)
    if (b->uq == nullptr) return;  // tioga.C:348-352
  if (nsend_ == 0 || nvar <= 0) return;
  const int nbk = (int)blocks_.size();
  const int column = interptype == 1;
  while ((int)qdev_.size() < nbk) qdev_.push_back(new DBuf<double>);
  while ((int)ptmp_.size() < nbk) {
    ptmp_.push_back(new DBuf<double>);
    pitmp_.push_back(new DBuf<int>);
  }
  PointDest D{};
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    const size_t n = (size_t)b.nnodes * nvar;
    qdev_[ib]->reserve(std::max<size_t>(n, 1));
    if (n) TG_HIP(hipMemcpyAsync(qdev_[ib]->p, b.uq, sizeof(double) * n, hipMemcpyHostToDevice, s_));
    ptmp_[ib]->reserve((size_t)std::max(b.npts, 1) * nvar);
    pitmp_[ib]->reserve(std::max(b.npts, 1));
    fill_dev<int>(pitmp_[ib]->p, b.npts, 0, s_);
    D.qtmp[ib] = ptmp_[ib]->p;
    D.itmp[ib] = pitmp_[ib]->p;
  }
  // donor blocks in order: a later block's value for a point replaces an earlier one's
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (!b.ninterp2) continue;
    PointRecs R{b.ninterp2, b.p_inode.p, b.p_nw.p, b.p_rpt.p, b.p_rblk.p, b.p_w.p};
    k_point_values<<<pnb(b.ninterp2), PBS, 0, s_>>>(R, qdev_[ib]->p, b.nnodes, nvar, column, D);
  }
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    if (b.npts)
      k_point_apply<<<pnb(b.npts), PBS, 0, s_>>>(b.npts, b.pnode.p, pitmp_[ib]->p, ptmp_[ib]->p, nvar, column,
                                                b.nnodes, b.iblank.p, qdev_[ib]->p);
  }
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    const size_t n = (size_t)b.nnodes * nvar;
    if (n) TG_HIP(hipMemcpyAsync(b.uq, qdev_[ib]->p, sizeof(double) * n, hipMemcpyDeviceToHost, s_));
    if (b.writeback && b.nnodes)
      TG_HIP(hipMemcpyAsync(b.uiblank, b.iblank.p, sizeof(int) * b.nnodes, hipMemcpyDeviceToHost, s_));
  }
  sync();
}

// The point path's arrays under the reference's member names:
// ntotalPoints, rxyz, ninterp2, and the records of interpList2 as
// interp2_inode / interp2_weights (nweights each, record after record) and
// interp2_rinfo (receptorInfo: partner index of the receptor's rank, point,
// the receptor block's index on its rank).
long Engine::points_array(int btag, const std::string &nm, void *out)
{
  Block *bp = find(btag);
  if (!bp) return -1;
  Block &b = *bp;
  auto down = [&](const void *src, size_t bytes, void *dst) {
    if (bytes) TG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s_));
  };
  if (nm == "ntotalPoints") {
    if (out) *(int *)out = b.npts;
    return 1;
  }
  if (nm == "rxyz") {
    if (out) {
      down(b.pxyz.p, sizeof(double) * 3 * (size_t)b.npts, out);
      sync();
    }
    return 3 * (long)b.npts;
  }
  if (nm == "ninterp2") {
    if (out) *(int *)out = b.ninterp2;
    return 1;
  }
  const int n = b.ninterp2;
  std::vector<int> nw(n), inode(8 * (size_t)n), rpt(n), rblk(n);
  std::vector<double> w(8 * (size_t)n);
  down(b.p_nw.p, sizeof(int) * n, nw.data());
  down(b.p_inode.p, sizeof(int) * 8 * (size_t)n, inode.data());
  down(b.p_w.p, sizeof(double) * 8 * (size_t)n, w.data());
  down(b.p_rpt.p, sizeof(int) * n, rpt.data());
  down(b.p_rblk.p, sizeof(int) * n, rblk.data());
  sync();
  long k = 0;
  if (nm == "interp2_inode" || nm == "interp2_weights") {
    for (int r = 0; r < n; r++)
      for (int m = 0; m < nw[r]; m++, k++)
        if (out) {
          if (nm == "interp2_inode") ((int *)out)[k] = inode[(size_t)m * n + r];
          else ((double *)out)[k] = w[(size_t)m * n + r];
        }
    return k;
  }
  if (nm == "interp2_rinfo") {
    if (out) {
      const std::vector<int> &pm = partners(b.rank);
      for (int r = 0; r < n; r++) {
        const Block &R = *blocks_[rblk[r]];
        ((int *)out)[3 * r] = nranks() > 1 ? (int)(std::lower_bound(pm.begin(), pm.end(), R.rank) - pm.begin()) : 0;
        ((int *)out)[3 * r + 1] = rpt[r];
        ((int *)out)[3 * r + 2] = R.lidx;
      }
    }
    return 3 * (long)n;
  }
  return -1;
}

}  // namespace tg
