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
//
// With the solver's high-order callbacks set (tioga_set_highorder_callback_)
// the same stages call them, on the host, where the reference does:
//   1. the points of each cell with cell iblank -1 (get_nodes_per_cell for
//      every such cell, then get_receptor_nodes, highOrder.C:220-245);
//   3. the device search lists each point's candidate cells; for each point
//      in order the host tests them in preorder rank with the exact element
//      box and donor_inclusion_test, up to the first that passes
//      (searchADTrecursion.C:52-101, checkContainment.C:86-93): the
//      reference's call sequence.  A coincident copy takes its
//      representative's donor and rst (the reference leaves its rst unset);
//   4. donor_frac per point donor (highOrder.C:363-383);
//   5. the values sum q[first + m * nvar + k] * w_m (highOrder.C:445-469,
//      row layout only, as in the reference); clearOrphans marks a receptor
//      cell 0 when its last orphan point lies in another mesh's hole map,
//      else 1 (highOrder.C:157-192); convert_to_modal writes the values of
//      cells still at -1 (highOrder.C:527-543).
#include <algorithm>
#include <cstring>
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
  const bool cb = ho_.set;
  // 1. picked points
  DBuf<int> flag;
  for (Block *bp : blocks_) {
    Block &b = *bp;
    b.npts = 0;
    if (cb) {  // MeshBlock::getInternalNodes with callbacks (highOrder.C:220-245)
      std::vector<int> ibc(b.ncells);
      if (b.ncells)
        TG_HIP(hipMemcpyAsync(ibc.data(), b.iblank_cell.p, sizeof(int) * b.ncells, hipMemcpyDeviceToHost, s_));
      sync();
      b.hctag.clear();
      for (int i = 0; i < b.ncells; i++)
        if (ibc[i] == -1) b.hctag.push_back(i + BASE);
      b.hppc.assign(b.hctag.size(), 0);
      for (size_t i = 0; i < b.hctag.size(); i++) {
        ho_.nodes_per_cell(&b.hctag[i], &b.hppc[i]);
        b.npts += b.hppc[i];
      }
      b.hrxyz.assign(3 * (size_t)b.npts, 0.0);
      size_t m = 0;
      for (size_t i = 0; i < b.hctag.size(); i++) {
        ho_.receptor_nodes(&b.hctag[i], &b.hppc[i], &b.hrxyz[m]);
        m += 3 * (size_t)b.hppc[i];
      }
      b.pxyz.reserve(3 * (size_t)std::max(b.npts, 1));
      if (b.npts)
        TG_HIP(hipMemcpyAsync(b.pxyz.p, b.hrxyz.data(), sizeof(double) * 3 * b.npts, hipMemcpyHostToDevice, s_));
      sync();
      continue;
    }
    b.pnode.reserve(std::max(b.nnodes, 1));
    flag.reserve(std::max(b.nnodes, 1));
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
  if (cb) {  // 3, 4 with callbacks: block by block, as tioga.C:171-176 runs them
    for (Block *b : blocks_) {
      search(*b);
      points_by_callbacks(*b);
    }
    points_ready_ = true;
    return;
  }
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

namespace {
constexpr int NFRAC = 1331;  // highOrder.C:25: the callbacks' weight and modal buffers

// utils.c:170-185 (C truncation toward zero, then the range test)
int check_hole_host(const double *x, const int nx[3], const unsigned char *sam, const double ext[6])
{
  double dx[3];
  int ix[3];
  for (int i = 0; i < 3; i++) dx[i] = (ext[i + 3] - ext[i]) / nx[i];
  for (int i = 0; i < 3; i++) {
    ix[i] = (int)((x[i] - ext[i]) / dx[i]);
    if (ix[i] < 0 || ix[i] > nx[i] - 1) return 0;
  }
  return sam[(size_t)ix[2] * nx[1] * nx[0] + (size_t)ix[1] * nx[0] + ix[0]];
}
}  // namespace

// MeshBlock::search's walk with donor_inclusion_test, then
// MeshBlock::processPointDonors with donor_frac, for block D's point queries
void Engine::points_by_callbacks(Block &D)
{
  D.h2_inode.clear();
  D.h2_woff.assign(1, 0);
  D.h2_rpt.clear();
  D.h2_rblk.clear();
  D.h2_w.clear();
  D.ninterp2 = 0;
  const int Q = D.nsearch;
  if (Q == 0) return;
  std::vector<int> qi, rank, cell;
  list_candidates(D, qi, rank, cell);
  std::vector<double> xs(3 * (size_t)Q);
  std::vector<int> xt(Q), is(3 * (size_t)Q), ibc(D.ncells);
  TG_HIP(hipMemcpyAsync(xs.data(), D.xsearch.p, sizeof(double) * xs.size(), hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(xt.data(), D.xtag.p, sizeof(int) * Q, hipMemcpyDeviceToHost, s_));
  TG_HIP(hipMemcpyAsync(is.data(), D.isearch.p, sizeof(int) * is.size(), hipMemcpyDeviceToHost, s_));
  if (D.ncells)
    TG_HIP(hipMemcpyAsync(ibc.data(), D.iblank_cell.p, sizeof(int) * D.ncells, hipMemcpyDeviceToHost, s_));
  sync();
  std::vector<int> ord(qi.size());
  for (size_t j = 0; j < ord.size(); j++) ord[j] = (int)j;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return qi[a] != qi[b] ? qi[a] < qi[b] : rank[a] < rank[b]; });
  std::vector<int> donor(Q, -1);
  std::vector<double> rst(3 * (size_t)Q, 0.0);
  for (size_t k = 0; k < ord.size();) {
    const int i = qi[ord[k]];
    bool found = false;
    for (; k < ord.size() && qi[ord[k]] == i; k++) {
      if (found) continue;
      const int c = cell[ord[k]];
      int t = 0;
      while (t < D.ti.ntypes - 1 && c >= D.ti.cstart[t + 1]) t++;
      const int nv = D.ti.nv[t];
      const int *cv = D.uconn[t] + (size_t)nv * (c - D.ti.cstart[t]);
      // the element box (search.C:199-214), tested with TOL (searchADTrecursion.C:60-66)
      bool inb = true;
      for (int d = 0; d < 3; d++) {
        double lo = BIGVALUE, hi = -BIGVALUE;
        for (int m = 0; m < nv; m++) {
          const double v = D.ux[3 * (size_t)(cv[m] - BASE) + d];
          lo = dmin(lo, v);
          hi = dmax(hi, v);
        }
        inb = inb && xs[3 * (size_t)i + d] >= lo - TOL && xs[3 * (size_t)i + d] <= hi + TOL;
      }
      if (!inb) continue;
      int icell1 = c + BASE, pass = 0;
      ho_.inclusion(&icell1, &xs[3 * (size_t)i], &pass, &rst[3 * (size_t)i]);
      if (pass) {
        donor[i] = c;
        found = true;
      }
    }
  }
  for (int i = 0; i < Q; i++)
    if (xt[i] != i) {
      donor[i] = donor[xt[i]];
      for (int j = 0; j < 3; j++) rst[3 * (size_t)i + j] = rst[3 * (size_t)xt[i] + j];
    }
  TG_HIP(hipMemcpyAsync(D.donorId.p, donor.data(), sizeof(int) * Q, hipMemcpyHostToDevice, s_));
  sync();
  std::vector<double> frac(NFRAC);
  for (int i = 0; i < Q; i++) {
    if (donor[i] < 0 || ibc[donor[i]] != 1) continue;
    int icell = donor[i] + BASE, nw = 0, inode0 = 0, ndim = NFRAC;
    ho_.frac(&icell, &xs[3 * (size_t)i], &nw, &inode0, frac.data(), &rst[3 * (size_t)i], &ndim);
    if (nw < 0 || nw > NFRAC) throw std::runtime_error("tioga_amd: donor_frac returned a weight count out of range");
    D.h2_inode.push_back(inode0);
    D.h2_w.insert(D.h2_w.end(), frac.begin(), frac.begin() + nw);
    D.h2_woff.push_back((int)D.h2_w.size());
    D.h2_rpt.push_back(is[3 * (size_t)i + 1]);
    D.h2_rblk.push_back(is[3 * (size_t)i + 2]);
  }
  D.ninterp2 = (int)D.h2_inode.size();
}

// tioga::dataUpdate(at_points = 1) with the callbacks: the reference's
// high-order data has a row layout only (highOrder.C:445-471)
void Engine::data_update_points_callbacks(int nvar, int interptype)
{
  if (interptype != 0)
    throw std::runtime_error("tioga_amd: the high-order point update has a row layout only (highOrder.C:445)");
  const int nbk = (int)blocks_.size();
  std::vector<std::vector<double>> qtmp(nbk);
  std::vector<std::vector<int>> itmp(nbk);
  for (int ib = 0; ib < nbk; ib++) {
    qtmp[ib].assign((size_t)blocks_[ib]->npts * nvar, 0.0);
    itmp[ib].assign(blocks_[ib]->npts, 0);
  }
  std::vector<double> qq(nvar);
  for (int ib = 0; ib < nbk; ib++) {
    const Block &D = *blocks_[ib];
    for (int r = 0; r < D.ninterp2; r++) {
      const size_t inode = (size_t)(D.h2_inode[r] - BASE);
      for (int k = 0; k < nvar; k++) qq[k] = 0;
      for (int m = 0; m < D.h2_woff[r + 1] - D.h2_woff[r]; m++) {
        const double w = D.h2_w[D.h2_woff[r] + m];
        for (int k = 0; k < nvar; k++) qq[k] += D.uq[inode + (size_t)m * nvar + k] * w;
      }
      const int b = D.h2_rblk[r], p = D.h2_rpt[r];
      std::copy(qq.begin(), qq.end(), qtmp[b].begin() + (size_t)p * nvar);
      itmp[b][p] = 1;
    }
  }
  // the hole maps, on the host
  std::vector<std::vector<unsigned char>> sam(hm_.size());
  for (size_t k = 0; k < hm_.size(); k++)
    if (hm_[k].exist) {
      sam[k].resize((size_t)hm_[k].nx[0] * hm_[k].nx[1] * hm_[k].nx[2]);
      TG_HIP(hipMemcpyAsync(sam[k].data(), hm_[k].sam.p, sam[k].size(), hipMemcpyDeviceToHost, s_));
    }
  sync();
  std::vector<double> qout((size_t)nvar * NFRAC);
  for (int ib = 0; ib < nbk; ib++) {
    Block &R = *blocks_[ib];
    std::vector<int> ibc(R.ncells);
    if (R.ncells)
      TG_HIP(hipMemcpyAsync(ibc.data(), R.iblank_cell.p, sizeof(int) * R.ncells, hipMemcpyDeviceToHost, s_));
    sync();
    // MeshBlock::clearOrphans (highOrder.C:157-192)
    size_t m = 0;
    for (size_t i = 0; i < R.hctag.size(); i++) {
      int reject = 0;
      for (int j = 0; j < R.hppc[i]; j++, m++) {
        if (itmp[ib][m] != 0) continue;
        reject = 2;
        for (size_t k = 0; k < hm_.size(); k++)
          if ((int)k != R.tag - BASE && hm_[k].exist &&
              check_hole_host(&R.hrxyz[3 * m], hm_[k].nx, sam[k].data(), hm_[k].ext)) {
            reject = 1;
            break;
          }
      }
      if (reject == 1) ibc[R.hctag[i] - 1] = 0;
      else if (reject == 2) ibc[R.hctag[i] - 1] = 1;
    }
    // MeshBlock::updatePointData (highOrder.C:527-543); npts is set once
    // and then carries what the callback leaves in it, as there
    int npts = NFRAC;
    m = 0;
    for (size_t i = 0; i < R.hctag.size(); i++) {
      if (ibc[R.hctag[i] - 1] == -1) {
        int index_out = 0;
        ho_.modal(&R.hctag[i], &R.hppc[i], &qtmp[ib][m], &npts, &index_out, qout.data());
        if (npts < 0 || npts > NFRAC) throw std::runtime_error("tioga_amd: convert_to_modal returned too many points");
        index_out -= BASE;
        size_t k = 0;
        for (int j = 0; j < npts; j++)
          for (int n = 0; n < nvar; n++) R.uq[(size_t)index_out + (size_t)j * nvar + n] = qout[k++];
      }
      m += (size_t)R.hppc[i] * nvar;
    }
    if (R.ncells) {
      TG_HIP(hipMemcpyAsync(R.iblank_cell.p, ibc.data(), sizeof(int) * R.ncells, hipMemcpyHostToDevice, s_));
      int *dst = R.uiblank_cell ? R.uiblank_cell : (R.own_iblank_cell.size() >= (size_t)R.ncells ? R.own_iblank_cell.data() : nullptr);
      if (dst && R.writeback) memcpy(dst, ibc.data(), sizeof(int) * R.ncells);
    }
    sync();
  }
}

void Engine::data_update_points(int nvar, int interptype)
{
  if (!points_ready_) throw std::runtime_error("tioga_amd: dataUpdate at points needs performConnectivityHighOrder");
  for (Block *b : blocks_)
    if (b->uq == nullptr) return;  // tioga.C:348-352
  if (nsend_ == 0 || nvar <= 0) return;
  if (ho_.set) {
    data_update_points_callbacks(nvar, interptype);
    return;
  }
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
  if (ho_.set) {  // callback records: one solution index each
    const int n = b.ninterp2;
    if (nm == "interp2_inode") {
      if (out) std::copy(b.h2_inode.begin(), b.h2_inode.end(), (int *)out);
      return n;
    }
    if (nm == "interp2_weights") {
      if (out) std::copy(b.h2_w.begin(), b.h2_w.end(), (double *)out);
      return (long)b.h2_w.size();
    }
    if (nm == "interp2_rinfo") {
      if (out) {
        const std::vector<int> &pm = partners(b.rank);
        for (int r = 0; r < n; r++) {
          const Block &R = *blocks_[b.h2_rblk[r]];
          ((int *)out)[3 * r] = nranks() > 1 ? (int)(std::lower_bound(pm.begin(), pm.end(), R.rank) - pm.begin()) : 0;
          ((int *)out)[3 * r + 1] = b.h2_rpt[r];
          ((int *)out)[3 * r + 2] = R.lidx;
        }
      }
      return 3 * (long)n;
    }
    return -1;
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
