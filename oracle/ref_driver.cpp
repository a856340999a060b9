// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Runs the compiled reference TIOGA (oracle/_ref/libtioga_ref.so, built from
// the sources under /root/reference/src by oracle/Makefile) on one MPI rank
// over a case file, and dumps every array the parity tests compare: the
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
// Result file: records {int32 namelen, name, int32 dtype(0=i32,1=f64),
//              int64 count, payload}.  Step s > 0 prefixes names with "s<s>.".
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
  int tag, nnodes, nwbc, nobc, ntypes, has_res, ncells;
  std::vector<double> x, nres, cres, q;
  std::vector<int> wbc, obc, iblank, iblank_cell;
  std::vector<int> nv, nc;
  std::vector<std::vector<int>> conn;
  std::vector<int *> vconn;
};

void dump(tioga *tg, std::vector<Block> &blocks, Writer &w, double tconn)
{
  w.f64("time_connectivity", &tconn, 1);
  for (size_t b = 0; b < blocks.size(); b++) {
    Block &B = blocks[b];
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

}  // namespace

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: ref_driver case.bin out.bin\n");
    return 2;
  }
  MPI_Init(&argc, &argv);
  Reader r{fopen(argv[1], "rb")};
  if (!r.f) { perror(argv[1]); return 2; }
  if (r.get<int32_t>() != 0x41474954 || r.get<int32_t>() != 1) { fprintf(stderr, "bad case\n"); return 2; }
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
  Writer w{fopen(argv[2], "wb"), ""};
  tioga *tg = new tioga;
  tg->setCommunicator(MPI_COMM_WORLD, 0, 1);
  tg->setSymmetry(isym);
  tg->setNfringe(&nfringe);
  tg->setMexclude(&mexclude);
  for (auto &B : blocks) {
    tg->registerGridData(B.tag, B.nnodes, B.x.data(), B.iblank.data(), B.nwbc, B.nobc, B.wbc.data(),
                         B.obc.data(), B.ntypes, B.nv.data(), B.nc.data(), B.vconn.data());
    if (B.has_res) tg->setResolutions(B.tag, B.nres.data(), B.cres.data());
    tg->set_cell_iblank(B.tag, B.iblank_cell.data());
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
    tg->profile();
    double t0 = MPI_Wtime();
    tg->performConnectivity();
    double tconn = MPI_Wtime() - t0;
    dump(tg, blocks, w, tconn);
    // dataUpdate on a fresh copy of q so every step starts from the same field
    std::vector<std::vector<double>> qs(blocks.size());
    for (size_t b = 0; b < blocks.size(); b++) {
      qs[b] = blocks[b].q;
      tg->registerSolution(blocks[b].tag, qs[b].data());
    }
    tg->dataUpdate(nvar, itype);
    for (size_t b = 0; b < blocks.size(); b++)
      w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
  }
  fclose(w.f);
  fclose(r.f);
  delete tg;
  MPI_Finalize();
  return 0;
}
