// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Runs the compiled reference TIOGA (oracle/_ref/libtioga_ref.so, built from
// the sources under /root/reference/src by oracle/Makefile) on one MPI rank
// (or, for version 5 cases, on several) over a case file, and dumps every array the parity tests compare: the
// user-visible outputs (iblank, iblank_cell, donor info, updated q) and the
// internal per-block state (search arrays, donorId, xtag, ADT arrays,
// resolutions, hole maps).  Internal members are reached by compiling the
// reference headers with `private` mapped to `public`; the reference sources
// themselves are compiled unchanged.
//
// Case file (little endian):
//   int32 magic 0x41474954, version 1, nblocks, isym, nfringe, mexclude,
//         nvar, interptype, nsteps
//   per block: int32 tag, nnodes, nwbc, nobc, ntypes, has_user_res
//              f64 x[3*nnodes]; int32 wbc[nwbc]; int32 obc[nobc]
//              per type: int32 nv, nc; int32 conn[nv*nc] (1-based)
//              if has_user_res: f64 nres[nnodes]; f64 cres[ncells]
//              f64 q[nnodes*nvar]
//   per step s >= 1 (moving meshes): per block int32 moved;
//              if moved: f64 x[3*nnodes] (new coordinates, same topology)
//   version 2 adds, after the blocks: int32 has_amr; if has_amr:
//              int32 nf, qstride, ngrids, qnodesize, npatch;
//              f64 qnode[qnodesize]; int32 idata[11*ngrids];
//              f64 rdata[6*ngrids]; per patch: int32 global_id,
//              f64 q[d3nf*nvar] (d3nf = prod(dims+2*nf) of that grid).
//   With AMR patches the driver runs the gather direction of the AMR path
//   (see run_amr) instead of performConnectivity; nsteps must be 0.
//   version 3 adds, after has_amr: int32 reduce; if set, every step goes on
//   after its dataUpdate with tioga::reduce_fringes (tioga.C:737-885), the
//   full dump again and a dataUpdate on a fresh copy of q, all under the
//   name prefix "rf.".
//   version 4 adds, after reduce: int32 has_gids; if set, per block:
//              uint64 cell_gid[ncells], uint64 node_gid[nnodes], passed to
//              registerGridData (tioga.h:121-123), and every dump adds
//              "receptor_info" (tioga::getReceptorInfo, tioga.C:572-671).
//              Cases with node ids need the gid build of this driver
//              (-DTIOGA_HAS_NODEGID: the dedupe of search.C:48-74).
//   version 5 adds, after the node ids: int32 has_ranks; if set:
//              int32 rank[nblocks], the MPI rank that registers each block.
//              Run under mpiexec with that many ranks: every rank registers
//              its own blocks in file order (a mesh tag may sit on several
//              ranks: a partitioned mesh), runs the same call sequence, and
//              writes its results to out.bin.r<rank> (only the blocks it
//              owns, under their case-file index) and the packets it sends
//              to out.bin.r<rank>.pk (below).
// Result file: records {int32 namelen, name, int32 dtype(0=i32,1=f64),
//              int64 count, payload}.  Step s > 0 prefixes names with "s<s>.".
//   version 6 adds, after the ranks: int32 points; if set, every step goes on
//              after its dataUpdate with tioga::performConnectivityHighOrder
//              (tioga.C:162-178), a dump of the point path under the prefix
//              "hp." (dump_points), and tioga::dataUpdate with at_points = 1
//              on a fresh copy of q, whose q, iblank and iblank_cell are
//              dumped as hp.b<i>.q / .iblank / .iblank_cell.  points = 1: no
//              high-order callbacks (the receptor nodes searched again as
//              points); points = 2: the callbacks of tests/cxx/ho_callbacks.c
//              are set (tioga::setcallback) on a case of two single-type
//              blocks, block 0 the receptor mesh and block 1 the donor mesh,
//              and the node dataUpdate is left out (with callbacks the C ABI
//              sends dataUpdate to the point update).
//   version 7 adds, after points: int32 amr_full; if set (an AMR case with
//              one near-body block), run_amr runs tioga::performConnectivity
//              first (the hole maps) and then performConnectivityAMR's own
//              sequence up to its closing MPI_Abort (tioga.C:187-207: both
//              directions, exchangeAMRDonors included), and also dumps the
//              near-body search of the Cartesian points, interpListCart,
//              every patch's iblank and, after dataUpdate_AMR, every
//              patch's q (amr.patch<p>.*).
// Packet log (multi-rank mode): every MPI_Isend the reference makes, in call
//              order, as {int32 1, dest, tag, dtype(0=int,1=double), count,
//              payload}, and every MPI_Waitall as {int32 2}.  Each
//              parallelComm::sendRecvPackets call (parallelComm.C:96-170) is
//              two such groups: the per-destination (nints, nreals) counts,
//              then the int and double payloads.  Captured through the MPI
//              profiling interface: these two functions are defined here and
//              forward to PMPI_Isend / PMPI_Waitall.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>
#include <mpi.h>
#define private public
#include "tioga.h"
#undef private

using namespace TIOGA;

extern "C" {  // tests/cxx/ho_callbacks.c
void ho_cb_set(const double *rx, const int *rconn, int rnv, const double *dx, const int *dconn, int dnv, int dnnodes,
               int nvar);
void ho_nodes_per_cell(int *icell, int *npts);
void ho_receptor_nodes(int *icell, int *npts, double *xyz);
void ho_inclusion(int *icell, double *x, int *pass, double *rst);
void ho_frac(int *icell, double *x, int *nweights, int *inode, double *frac, double *rst, int *ndim);
void ho_modal(int *icell, int *npts_cell, double *qtmp, int *npts, int *index_out, double *qout);
}

namespace {
FILE *g_pk = nullptr;  // packet log of this rank (multi-rank mode)
}

extern "C" int MPI_Isend(const void *buf, int count, MPI_Datatype dt, int dest, int tag, MPI_Comm comm,
                         MPI_Request *req)
{
  if (g_pk) {
    const int dtype = dt == MPI_DOUBLE ? 1 : (dt == MPI_INT ? 0 : -1);
    const int32_t h[5] = {1, dest, tag, dtype, count};
    fwrite(h, 4, 5, g_pk);
    if (count > 0 && dtype >= 0) fwrite(buf, dtype ? 8 : 4, count, g_pk);
  }
  return PMPI_Isend(buf, count, dt, dest, tag, comm, req);
}

extern "C" int MPI_Waitall(int n, MPI_Request *reqs, MPI_Status *st)
{
  if (g_pk) {
    const int32_t h = 2;
    fwrite(&h, 4, 1, g_pk);
  }
  return PMPI_Waitall(n, reqs, st);
}

namespace {

struct Reader {
  FILE *f;
  template <class T> T get() {
    T v;
    if (fread(&v, sizeof(T), 1, f) != 1) { fprintf(stderr, "ref_driver: short read\n"); exit(2); }
    return v;
  }
  template <class T> void arr(std::vector<T> &v, size_t n) {
    v.resize(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) { fprintf(stderr, "ref_driver: short read\n"); exit(2); }
  }
};

struct Writer {
  FILE *f;
  std::string prefix;
  void put(const std::string &name0, int dtype, const void *p, int64_t n) {
    std::string name = prefix + name0;
    int32_t l = (int32_t)name.size();
    fwrite(&l, 4, 1, f);
    fwrite(name.data(), 1, l, f);
    fwrite(&dtype, 4, 1, f);
    fwrite(&n, 8, 1, f);
    if (n > 0 && p) fwrite(p, dtype == 0 ? 4 : 8, n, f);
  }
  void i32(const std::string &n, const int *p, int64_t c) { put(n, 0, p, p ? c : 0); }
  void f64(const std::string &n, const double *p, int64_t c) { put(n, 1, p, p ? c : 0); }
};

struct Block {
  int tag, nnodes, nwbc, nobc, ntypes, has_res, ncells, rank = 0;
  std::vector<uint64_t> cell_gid, node_gid;
  std::vector<double> x, nres, cres, q;
  std::vector<int> wbc, obc, iblank, iblank_cell;
  std::vector<int> nv, nc;
  std::vector<std::vector<int>> conn;
  std::vector<int *> vconn;
};

void dump(tioga *tg, std::vector<Block> &blocks, Writer &w, double tconn, bool gids)
{
  w.f64("time_connectivity", &tconn, 1);
  if (gids) {
    std::vector<int> ri;
    tg->getReceptorInfo(ri);
    w.i32("receptor_info", ri.data(), (int64_t)ri.size());
  }
  for (size_t b = 0; b < blocks.size(); b++) {
    Block &B = blocks[b];
    if (B.rank != tg->myid) continue;
    auto it = tg->tag_iblk_map.find(B.tag);
    MeshBlock *mb = tg->mblocks[it->second].get();
    std::string p = "b" + std::to_string(b) + ".";
    w.i32(p + "iblank", B.iblank.data(), B.nnodes);
    w.i32(p + "iblank_cell", B.iblank_cell.data(), B.ncells);
    w.f64(p + "nodeRes", mb->nodeRes, B.nnodes);
    w.f64(p + "cellRes", mb->cellRes, B.ncells);
    double obb[15];
    memcpy(obb, mb->obb->xc, 24);
    memcpy(obb + 3, mb->obb->dxc, 24);
    memcpy(obb + 6, mb->obb->vec, 72);
    w.f64(p + "obb", obb, 15);
    int ns = mb->nsearch;
    w.i32(p + "nsearch", &ns, 1);
    if (ns > 0) {
      w.f64(p + "xsearch", mb->xsearch, 3 * (int64_t)ns);
      w.f64(p + "res_search", mb->res_search, ns);
      w.i32(p + "isearch", mb->isearch, 3 * (int64_t)ns);
      w.i32(p + "tagsearch", mb->tagsearch, ns);
      w.i32(p + "donorId", mb->donorId, ns);
      w.i32(p + "xtag", mb->xtag, ns);
      if (mb->adt) {
        int ne = mb->adt->nelem;
        w.i32(p + "elementList", mb->elementList, ne);
        w.i32(p + "adtIntegers", mb->adt->adtIntegers, 4 * (int64_t)ne);
        w.f64(p + "adtReals", mb->adt->adtReals, 6 * (int64_t)ne);
        w.f64(p + "adtExtents", mb->adt->adtExtents, 6);
      }
    }
    int dcount = 0, fcount = 0;
    tg->getDonorCount(B.tag, &dcount, &fcount);
    std::vector<int> rec(4 * (size_t)dcount + 1), ind((size_t)fcount + 1);
    std::vector<double> frac((size_t)fcount + 1);
    if (dcount > 0) tg->getDonorInfo(B.tag, rec.data(), ind.data(), frac.data(), &dcount);
    w.i32(p + "dcount", &dcount, 1);
    w.i32(p + "receptors", rec.data(), 4 * (int64_t)dcount);
    w.i32(p + "indices", ind.data(), fcount);
    w.f64(p + "frac", frac.data(), fcount);
  }
  for (int m = 0; m < tg->nmesh; m++) {
    HOLEMAP &h = tg->holeMap[m];
    std::string p = "hm" + std::to_string(m + 1) + ".";
    w.i32(p + "existWall", &h.existWall, 1);
    if (!h.existWall) continue;
    w.i32(p + "nx", h.nx, 3);
    w.f64(p + "extents", h.extents, 6);
    w.i32(p + "sam", h.sam, (int64_t)h.nx[0] * h.nx[1] * h.nx[2]);
  }
}

// The point path of one rank (highOrder.C:219-419): picked points, the
// point search's arrays and the point donor records (interpList2).
void dump_points(tioga *tg, std::vector<Block> &blocks, Writer &w)
{
  for (size_t b = 0; b < blocks.size(); b++) {
    Block &B = blocks[b];
    if (B.rank != tg->myid) continue;
    MeshBlock *mb = tg->mblocks[tg->tag_iblk_map.find(B.tag)->second].get();
    std::string p = "b" + std::to_string(b) + ".";
    int np = mb->ntotalPoints, ns = mb->nsearch, n2 = mb->ninterp2;
    w.i32(p + "ntotalPoints", &np, 1);
    w.f64(p + "rxyz", mb->rxyz, 3 * (int64_t)np);
    w.i32(p + "nsearch", &ns, 1);
    if (ns > 0) {
      w.f64(p + "xsearch", mb->xsearch, 3 * (int64_t)ns);
      w.f64(p + "res_search", mb->res_search, ns);
      w.i32(p + "isearch", mb->isearch, 3 * (int64_t)ns);
      w.i32(p + "tagsearch", mb->tagsearch, ns);
      w.i32(p + "donorId", mb->donorId, ns);
      w.i32(p + "xtag", mb->xtag, ns);
    }
    std::vector<int> inode, rinfo;
    std::vector<double> wt;
    for (int i = 0; i < n2; i++) {
      INTERPLIST &l = mb->interpList2[i];
      // with callbacks a record holds one solution index (highOrder.C:365)
      inode.insert(inode.end(), l.inode, l.inode + (mb->ihigh ? 1 : l.nweights));
      wt.insert(wt.end(), l.weights, l.weights + l.nweights);
      rinfo.insert(rinfo.end(), l.receptorInfo, l.receptorInfo + 3);
    }
    w.i32(p + "ninterp2", &n2, 1);
    w.i32(p + "interp2_inode", inode.data(), (int64_t)inode.size());
    w.f64(p + "interp2_weights", wt.data(), (int64_t)wt.size());
    w.i32(p + "interp2_rinfo", rinfo.data(), (int64_t)rinfo.size());
  }
}

struct Amr {
  int nf = 0, qstride = 0, ngrids = 0, qnodesize = 0, npatch = 0;
  std::vector<double> qnode, rdata;
  std::vector<int> idata, pgid;
  std::vector<std::vector<double>> q;
  std::vector<std::vector<int>> ibl;
};

// tioga::performConnectivityAMR (tioga.C:180-216) on one rank without the
// Cartesian-receptor direction (getCartReceptors, the near-body search of
// Cartesian points and the donor half of exchangeAMRDonors) and without its
// closing MPI_Abort; then tioga::dataUpdate_AMR (tioga.C:218-334).  Every
// step is a call into the compiled reference except the loop that fills the
// patch interpolation lists, which is exchangeAMRDonors.C:120-192 written
// for a single rank (packets to self, so arrival order = packing order).
void run_amr_full(tioga *tg, std::vector<Block> &blocks, Amr &A, int nvar, int itype, Writer &w);

void run_amr(tioga *tg, std::vector<Block> &blocks, Amr &A, int nvar, int itype, Writer &w, bool full)
{
  tg->register_amr_global_data(A.nf, A.qstride, A.qnode.data(), A.idata.data(), A.rdata.data(), A.ngrids,
                               A.qnodesize);
  tg->set_amr_patch_count(A.npatch);
  for (int p = 0; p < A.npatch; p++) tg->register_amr_local_data(p, A.pgid[p], A.ibl[p].data(), A.q[p].data());
  tg->profile();
  if (full) {
    run_amr_full(tg, blocks, A, nvar, itype, w);
    return;
  }
  double t0 = MPI_Wtime();
  tg->cg->preprocess();
  for (int i = 0; i < tg->ncart; i++) tg->cb[i].preprocess(tg->cg);
  for (int ib = 0; ib < tg->nblocks; ib++) {
    MeshBlock *mb = tg->mblocks[ib].get();
    mb->getUnresolvedMandatoryReceptors();
    tg->cg->search(mb->rxyzCart, mb->donorIdCart, mb->ntotalPointsCart);
  }
  for (int i = 0; i < tg->ncart; i++) {
    tg->cb[i].clearLists();
    tg->cb[i].initializeLists();
  }
  for (int ib = 0; ib < tg->nblocks; ib++) {
    MeshBlock *mb = tg->mblocks[ib].get();
    for (int i = 0; i < mb->ntotalPointsCart; i++)
      if (mb->donorIdCart[i] != -1) {
        int gid = mb->donorIdCart[i];
        tg->cb[tg->cg->local_id[gid]].insertInInterpList(0, i, ib, &mb->rxyzCart[3 * i]);
      }
  }
  for (int ib = 0; ib < tg->nblocks; ib++) tg->mblocks[ib]->setCartIblanks();
  for (int ib = 0; ib < tg->nblocks; ib++) tg->mblocks[ib]->getCellIblanks();
  double tconn = MPI_Wtime() - t0;
  w.f64("time_connectivity", &tconn, 1);
  for (size_t b = 0; b < blocks.size(); b++) {
    Block &B = blocks[b];
    MeshBlock *mb = tg->mblocks[tg->tag_iblk_map.find(B.tag)->second].get();
    std::string p = "b" + std::to_string(b) + ".";
    w.i32(p + "iblank", B.iblank.data(), B.nnodes);
    w.i32(p + "iblank_cell", B.iblank_cell.data(), B.ncells);
    w.f64(p + "nodeRes", mb->nodeRes, B.nnodes);
    int n = mb->ntotalPointsCart;
    w.i32(p + "ncart", &n, 1);
    w.i32(p + "receptorIdCart", mb->receptorIdCart, n);
    w.i32(p + "donorIdCart", mb->donorIdCart, n);
    w.f64(p + "rxyzCart", mb->rxyzCart, 3 * (int64_t)n);
  }
  std::vector<int> rpatch, rinfo, rijk;
  std::vector<double> rw;
  for (int i = 0; i < tg->ncart; i++)
    for (INTERPLIST2 *l = tg->cb[i].interpList; l; l = l->next) {
      rpatch.push_back(i);
      rinfo.insert(rinfo.end(), l->receptorInfo, l->receptorInfo + 3);
      rijk.insert(rijk.end(), l->inode, l->inode + 3 * l->nweights);
      rw.insert(rw.end(), l->weights, l->weights + l->nweights);
    }
  w.i32("amr.rec_patch", rpatch.data(), (int64_t)rpatch.size());
  w.i32("amr.rec_info", rinfo.data(), (int64_t)rinfo.size());
  w.i32("amr.rec_ijk", rijk.data(), (int64_t)rijk.size());
  w.f64("amr.rec_w", rw.data(), (int64_t)rw.size());
  // dataUpdate_AMR needs the Cartesian communicator map that
  // exchangeAMRDonors.C:58-60 sets: on one rank, send to and receive from self
  int self = 0;
  tg->pc_cart->setMap(1, 1, &self, &self);
  // registerSolution stores into qblock, which only performConnectivity
  // allocates (tioga.C:152-155); allocate it the same way
  if (!tg->qblock) {
    tg->qblock = (double **)malloc(sizeof(double *) * tg->nblocks);
    for (int ib = 0; ib < tg->nblocks; ib++) tg->qblock[ib] = NULL;
  }
  std::vector<std::vector<double>> qs(blocks.size());
  for (size_t b = 0; b < blocks.size(); b++) {
    qs[b] = blocks[b].q;
    tg->registerSolution(blocks[b].tag, qs[b].data());
  }
  tg->dataUpdate_AMR(nvar, itype);
  for (size_t b = 0; b < blocks.size(); b++)
    w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
}

// tioga::performConnectivityAMR (tioga.C:180-216) as the reference runs
// it, after a performConnectivity that leaves the hole maps its Cartesian
// donor processing reads (CartBlock.C:308-380), without the closing
// MPI_Barrier / MPI_Abort; then tioga::dataUpdate_AMR.
void run_amr_full(tioga *tg, std::vector<Block> &blocks, Amr &A, int nvar, int itype, Writer &w)
{
  tg->performConnectivity();
  double t0 = MPI_Wtime();
  tg->cg->preprocess();
  for (int i = 0; i < tg->ncart; i++) tg->cb[i].preprocess(tg->cg);
  for (int ib = 0; ib < tg->nblocks; ib++) {
    MeshBlock *mb = tg->mblocks[ib].get();
    mb->getCartReceptors(tg->cg, tg->pc_cart);
    mb->ihigh = tg->ihigh;
    mb->search();
    mb->getUnresolvedMandatoryReceptors();
    tg->cg->search(mb->rxyzCart, mb->donorIdCart, mb->ntotalPointsCart);
  }
  tg->exchangeAMRDonors();
  for (int ib = 0; ib < tg->nblocks; ib++) tg->mblocks[ib]->getCellIblanks();
  double tconn = MPI_Wtime() - t0;
  w.f64("time_connectivity", &tconn, 1);
  for (size_t b = 0; b < blocks.size(); b++) {
    Block &B = blocks[b];
    MeshBlock *mb = tg->mblocks[tg->tag_iblk_map.find(B.tag)->second].get();
    std::string p = "b" + std::to_string(b) + ".";
    w.i32(p + "iblank", B.iblank.data(), B.nnodes);
    w.i32(p + "iblank_cell", B.iblank_cell.data(), B.ncells);
    w.f64(p + "nodeRes", mb->nodeRes, B.nnodes);
    int n = mb->ntotalPointsCart;
    w.i32(p + "ncart", &n, 1);
    w.i32(p + "receptorIdCart", mb->receptorIdCart, n);
    w.i32(p + "donorIdCart", mb->donorIdCart, n);
    w.f64(p + "rxyzCart", mb->rxyzCart, 3 * (int64_t)n);
    int ns = mb->nsearch;
    w.i32(p + "nsearch", &ns, 1);
    if (ns > 0) {
      w.f64(p + "xsearch", mb->xsearch, 3 * (int64_t)ns);
      w.i32(p + "isearch", mb->isearch, 3 * (int64_t)ns);
      w.i32(p + "donorId", mb->donorId, ns);
      w.i32(p + "xtag", mb->xtag, ns);
    }
    std::vector<int> rinfo, inode;
    std::vector<double> wt;
    for (int i = 0; i < mb->ninterpCart; i++) {
      INTERPLIST &l = mb->interpListCart[i];
      rinfo.insert(rinfo.end(), l.receptorInfo, l.receptorInfo + 3);
      inode.insert(inode.end(), l.inode, l.inode + l.nweights);
      wt.insert(wt.end(), l.weights, l.weights + l.nweights);
    }
    int nc = mb->ninterpCart;
    w.i32(p + "ninterpCart", &nc, 1);
    w.i32(p + "interpCart_rinfo", rinfo.data(), (int64_t)rinfo.size());
    w.i32(p + "interpCart_inode", inode.data(), (int64_t)inode.size());
    w.f64(p + "interpCart_w", wt.data(), (int64_t)wt.size());
  }
  std::vector<int> rpatch, rinfo, rijk;
  std::vector<double> rw;
  for (int i = 0; i < tg->ncart; i++)
    for (INTERPLIST2 *l = tg->cb[i].interpList; l; l = l->next) {
      rpatch.push_back(i);
      rinfo.insert(rinfo.end(), l->receptorInfo, l->receptorInfo + 3);
      rijk.insert(rijk.end(), l->inode, l->inode + 3 * l->nweights);
      rw.insert(rw.end(), l->weights, l->weights + l->nweights);
    }
  w.i32("amr.rec_patch", rpatch.data(), (int64_t)rpatch.size());
  w.i32("amr.rec_info", rinfo.data(), (int64_t)rinfo.size());
  w.i32("amr.rec_ijk", rijk.data(), (int64_t)rijk.size());
  w.f64("amr.rec_w", rw.data(), (int64_t)rw.size());
  for (int p = 0; p < A.npatch; p++)
    w.i32("amr.patch" + std::to_string(p) + ".iblank", A.ibl[p].data(), (int64_t)A.ibl[p].size());
  if (!tg->qblock) {
    tg->qblock = (double **)malloc(sizeof(double *) * tg->nblocks);
    for (int ib = 0; ib < tg->nblocks; ib++) tg->qblock[ib] = NULL;
  }
  std::vector<std::vector<double>> qs(blocks.size());
  for (size_t b = 0; b < blocks.size(); b++) {
    qs[b] = blocks[b].q;
    tg->registerSolution(blocks[b].tag, qs[b].data());
  }
  tg->dataUpdate_AMR(nvar, itype);
  for (size_t b = 0; b < blocks.size(); b++)
    w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
  for (int p = 0; p < A.npatch; p++)
    w.f64("amr.patch" + std::to_string(p) + ".q", A.q[p].data(), (int64_t)A.q[p].size());
}

}  // namespace

int main(int argc, char **argv)
{
  // REF_DRIVER_TIMES set: wall times of profile, performConnectivity and
  // dataUpdate to stderr (the timers build adds the reference's own per-phase
  // lines, tioga.C:715-735)
  const bool times = getenv("REF_DRIVER_TIMES") != nullptr;
  if (argc < 3) {
    fprintf(stderr, "usage: ref_driver case.bin out.bin\n");
    return 2;
  }
  MPI_Init(&argc, &argv);
  Reader r{fopen(argv[1], "rb")};
  if (!r.f) { perror(argv[1]); return 2; }
  if (r.get<int32_t>() != 0x41474954) { fprintf(stderr, "bad case\n"); return 2; }
  const int version = r.get<int32_t>();
  if (version < 1 || version > 7) { fprintf(stderr, "bad case version\n"); return 2; }
  int nblocks = r.get<int32_t>(), isym = r.get<int32_t>(), nfringe = r.get<int32_t>();
  int mexclude = r.get<int32_t>(), nvar = r.get<int32_t>(), itype = r.get<int32_t>();
  int nsteps = r.get<int32_t>();
  std::vector<Block> blocks(nblocks);
  for (auto &B : blocks) {
    B.tag = r.get<int32_t>(); B.nnodes = r.get<int32_t>(); B.nwbc = r.get<int32_t>();
    B.nobc = r.get<int32_t>(); B.ntypes = r.get<int32_t>(); B.has_res = r.get<int32_t>();
    r.arr(B.x, 3 * (size_t)B.nnodes);
    r.arr(B.wbc, B.nwbc);
    r.arr(B.obc, B.nobc);
    B.nv.resize(B.ntypes); B.nc.resize(B.ntypes); B.conn.resize(B.ntypes);
    B.ncells = 0;
    for (int t = 0; t < B.ntypes; t++) {
      B.nv[t] = r.get<int32_t>(); B.nc[t] = r.get<int32_t>();
      r.arr(B.conn[t], (size_t)B.nv[t] * B.nc[t]);
      B.ncells += B.nc[t];
    }
    for (int t = 0; t < B.ntypes; t++) B.vconn.push_back(B.conn[t].data());
    if (B.has_res) { r.arr(B.nres, B.nnodes); r.arr(B.cres, B.ncells); }
    r.arr(B.q, (size_t)B.nnodes * nvar);
    B.iblank.assign(B.nnodes, 1);
    B.iblank_cell.assign(B.ncells > 0 ? B.ncells : 1, 1);
  }
  Amr amr;
  int has_amr = version >= 2 ? r.get<int32_t>() : 0;
  if (has_amr) {
    amr.nf = r.get<int32_t>(); amr.qstride = r.get<int32_t>(); amr.ngrids = r.get<int32_t>();
    amr.qnodesize = r.get<int32_t>(); amr.npatch = r.get<int32_t>();
    r.arr(amr.qnode, amr.qnodesize);
    r.arr(amr.idata, 11 * (size_t)amr.ngrids);
    r.arr(amr.rdata, 6 * (size_t)amr.ngrids);
    amr.pgid.resize(amr.npatch);
    amr.q.resize(amr.npatch);
    amr.ibl.resize(amr.npatch);
    for (int p = 0; p < amr.npatch; p++) {
      int g = amr.pgid[p] = r.get<int32_t>();
      size_t d3nf = 1;
      for (int n = 0; n < 3; n++) d3nf *= (size_t)(amr.idata[11 * g + 8 + n] - amr.idata[11 * g + 5 + n] + 1 + 2 * amr.nf);
      r.arr(amr.q[p], d3nf * nvar);
      amr.ibl[p].assign(d3nf, 1);
    }
  }
  const int reduce = version >= 3 ? r.get<int32_t>() : 0;
  const int has_gids = version >= 4 ? r.get<int32_t>() : 0;
  if (has_gids)
    for (auto &B : blocks) {
      r.arr(B.cell_gid, B.ncells);
      r.arr(B.node_gid, B.nnodes);
    }
  const int has_ranks = version >= 5 ? r.get<int32_t>() : 0;
  if (has_ranks)
    for (auto &B : blocks) B.rank = r.get<int32_t>();
  const int points = version >= 6 ? r.get<int32_t>() : 0;
  const int amr_full = version >= 7 ? r.get<int32_t>() : 0;
  int myrank = 0, nprocs = 1;
  MPI_Comm_rank(MPI_COMM_WORLD, &myrank);
  MPI_Comm_size(MPI_COMM_WORLD, &nprocs);
  if (!has_ranks && nprocs != 1) { fprintf(stderr, "ref_driver: a case without ranks runs on one rank\n"); return 2; }
  for (auto &B : blocks)
    if (B.rank < 0 || B.rank >= nprocs) { fprintf(stderr, "ref_driver: block rank outside the run\n"); return 2; }
  std::string outp = argv[2];
  if (has_ranks) {
    outp += ".r" + std::to_string(myrank);
    g_pk = fopen((outp + ".pk").c_str(), "wb");
  }
  Writer w{fopen(outp.c_str(), "wb"), ""};
  tioga *tg = new tioga;
  tg->setCommunicator(MPI_COMM_WORLD, myrank, nprocs);
  tg->setSymmetry(isym);
  tg->setNfringe(&nfringe);
  tg->setMexclude(&mexclude);
  for (auto &B : blocks) {
    if (B.rank != myrank) continue;
    tg->registerGridData(B.tag, B.nnodes, B.x.data(), B.iblank.data(), B.nwbc, B.nobc, B.wbc.data(),
                         B.obc.data(), B.ntypes, B.nv.data(), B.nc.data(), B.vconn.data(),
                         has_gids ? B.cell_gid.data() : NULL, has_gids ? B.node_gid.data() : NULL);
    if (B.has_res) tg->setResolutions(B.tag, B.nres.data(), B.cres.data());
    tg->set_cell_iblank(B.tag, B.iblank_cell.data());
  }
  if (points == 2) {
    if (blocks.size() != 2 || blocks[0].ntypes != 1 || blocks[1].ntypes != 1 || has_ranks) {
      fprintf(stderr, "ref_driver: callback cases have two single-type blocks on one rank\n");
      return 2;
    }
    ho_cb_set(blocks[0].x.data(), blocks[0].conn[0].data(), blocks[0].nv[0], blocks[1].x.data(),
              blocks[1].conn[0].data(), blocks[1].nv[0], blocks[1].nnodes, nvar);
    tg->setcallback(ho_nodes_per_cell, ho_receptor_nodes, ho_inclusion, ho_frac, ho_modal);
  }
  if (has_amr) {
    if (has_ranks) { fprintf(stderr, "ref_driver: AMR cases run on one rank\n"); return 2; }
    run_amr(tg, blocks, amr, nvar, itype, w, amr_full != 0);
    nsteps = -1;
  }
  for (int s = 0; s <= nsteps; s++) {
    if (s > 0) {
      for (auto &B : blocks) {
        int moved = r.get<int32_t>();
        if (moved) {
          std::vector<double> nx;
          r.arr(nx, 3 * (size_t)B.nnodes);
          memcpy(B.x.data(), nx.data(), sizeof(double) * nx.size());
        }
      }
      w.prefix = "s" + std::to_string(s) + ".";
    }
    const double tp0 = MPI_Wtime();
    tg->profile();
    const double tprof = MPI_Wtime() - tp0;
    double t0 = MPI_Wtime();
    tg->performConnectivity();
    double tconn = MPI_Wtime() - t0;
    if (times) fprintf(stderr, "ref_driver step %d: profile %.6f s, performConnectivity %.6f s\n", s, tprof, tconn);
    dump(tg, blocks, w, tconn, has_gids != 0);
    // dataUpdate on a fresh copy of q so every step starts from the same field
    std::vector<std::vector<double>> qs(blocks.size());
    for (size_t b = 0; b < blocks.size() && points != 2; b++) {
      if (blocks[b].rank != myrank) continue;
      qs[b] = blocks[b].q;
      tg->registerSolution(blocks[b].tag, qs[b].data());
    }
    const double tu0 = MPI_Wtime();
    if (points != 2) tg->dataUpdate(nvar, itype);
    if (times) fprintf(stderr, "ref_driver step %d: dataUpdate %.6f s\n", s, MPI_Wtime() - tu0);
    for (size_t b = 0; b < blocks.size() && points != 2; b++)
      if (blocks[b].rank == myrank) w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
    if (points) {
      const std::string p0 = w.prefix;
      tg->performConnectivityHighOrder();
      w.prefix = p0 + "hp.";
      dump_points(tg, blocks, w);
      for (size_t b = 0; b < blocks.size(); b++) {
        if (blocks[b].rank != myrank) continue;
        qs[b] = blocks[b].q;
        tg->registerSolution(blocks[b].tag, qs[b].data());
      }
      tg->dataUpdate(nvar, itype, 1);
      for (size_t b = 0; b < blocks.size(); b++) {
        if (blocks[b].rank != myrank) continue;
        w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
        w.i32("b" + std::to_string(b) + ".iblank", blocks[b].iblank.data(), blocks[b].nnodes);
        w.i32("b" + std::to_string(b) + ".iblank_cell", blocks[b].iblank_cell.data(), blocks[b].ncells);
      }
      w.prefix = p0;
    }
    if (reduce) {
      const std::string p0 = w.prefix;
      tg->reduce_fringes();
      w.prefix = p0 + "rf.";
      dump(tg, blocks, w, 0.0, has_gids != 0);
      for (size_t b = 0; b < blocks.size(); b++) {
        if (blocks[b].rank != myrank) continue;
        qs[b] = blocks[b].q;
        tg->registerSolution(blocks[b].tag, qs[b].data());
      }
      tg->dataUpdate(nvar, itype);
      for (size_t b = 0; b < blocks.size(); b++)
        if (blocks[b].rank == myrank) w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
      w.prefix = p0;
    }
  }
  fclose(w.f);
  fclose(r.f);
  if (g_pk) fclose(g_pk);
  g_pk = nullptr;
  delete tg;
  MPI_Finalize();
  return 0;
}
