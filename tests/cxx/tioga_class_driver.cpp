// tioga_class_driver.cpp -- a C++ caller of the engine's TIOGA::tioga class
// (include/tioga.h), for tests/test_cxx_class.py.
//
// It reads the case files the golden fixtures were made from (the format of
// oracle/ref_driver.cpp, versions 1-4, without AMR patches), drives the
// class through the reference's call sequence -- setCommunicator,
// setSymmetry / setNfringe / setMexclude, registerGridData (with cell and
// node global ids when the case has them), setResolutions, set_cell_iblank,
// profile, performConnectivity, registerSolution, dataUpdate,
// getDonorCount / getDonorInfo, getReceptorInfo, reduce_fringes -- and
// writes the caller-visible results under the names the reference driver
// uses: iblank, iblank_cell, dcount, receptors, indices, frac, receptor_info
// and the updated q of every step ("s<k>." and "rf." prefixes alike).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tioga.h"

namespace {

struct Reader {
  FILE *f;
  template <class T> T get()
  {
    T v;
    if (fread(&v, sizeof(T), 1, f) != 1) {
      fprintf(stderr, "tioga_class_driver: short read\n");
      exit(2);
    }
    return v;
  }
  template <class T> void arr(std::vector<T> &v, size_t n)
  {
    v.resize(n);
    if (n && fread(v.data(), sizeof(T), n, f) != n) {
      fprintf(stderr, "tioga_class_driver: short read\n");
      exit(2);
    }
  }
};

struct Writer {
  FILE *f;
  std::string prefix;
  void put(const std::string &name0, int dtype, const void *p, int64_t n)
  {
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
  std::vector<int> wbc, obc, iblank, iblank_cell, nv, nc;
  std::vector<std::vector<int>> conn;
  std::vector<int *> vconn;
  std::vector<uint64_t> cell_gid, node_gid;
};

void dump(TIOGA::tioga &tg, std::vector<Block> &blocks, Writer &w, bool gids)
{
  if (gids) {
    std::vector<int> ri;
    tg.getReceptorInfo(ri);
    w.i32("receptor_info", ri.data(), (int64_t)ri.size());
  }
  for (size_t b = 0; b < blocks.size(); b++) {
    Block &B = blocks[b];
    std::string p = "b" + std::to_string(b) + ".";
    w.i32(p + "iblank", B.iblank.data(), B.nnodes);
    w.i32(p + "iblank_cell", B.iblank_cell.data(), B.ncells);
    int dcount = 0, fcount = 0;
    tg.getDonorCount(B.tag, &dcount, &fcount);
    std::vector<int> rec(4 * (size_t)dcount + 1), ind((size_t)fcount + 1);
    std::vector<double> frac((size_t)fcount + 1);
    if (dcount > 0) tg.getDonorInfo(B.tag, rec.data(), ind.data(), frac.data(), &dcount);
    w.i32(p + "dcount", &dcount, 1);
    w.i32(p + "receptors", rec.data(), 4 * (int64_t)dcount);
    w.i32(p + "indices", ind.data(), fcount);
    w.f64(p + "frac", frac.data(), fcount);
  }
}

void update(TIOGA::tioga &tg, std::vector<Block> &blocks, Writer &w, int nvar, int itype)
{
  std::vector<std::vector<double>> qs(blocks.size());
  for (size_t b = 0; b < blocks.size(); b++) {
    qs[b] = blocks[b].q;
    tg.registerSolution(blocks[b].tag, qs[b].data());
  }
  tg.dataUpdate(nvar, itype);
  for (size_t b = 0; b < blocks.size(); b++)
    w.f64("b" + std::to_string(b) + ".q", qs[b].data(), (int64_t)qs[b].size());
}

int run(const char *in, const char *out)
{
  Reader r{fopen(in, "rb")};
  if (!r.f) {
    perror(in);
    return 2;
  }
  if (r.get<int32_t>() != 0x41474954) return 2;
  const int version = r.get<int32_t>();
  if (version < 1 || version > 4) return 2;
  int nblocks = r.get<int32_t>(), isym = r.get<int32_t>(), nfringe = r.get<int32_t>();
  int mexclude = r.get<int32_t>(), nvar = r.get<int32_t>(), itype = r.get<int32_t>();
  int nsteps = r.get<int32_t>();
  std::vector<Block> blocks(nblocks);
  for (auto &B : blocks) {
    B.tag = r.get<int32_t>();
    B.nnodes = r.get<int32_t>();
    B.nwbc = r.get<int32_t>();
    B.nobc = r.get<int32_t>();
    B.ntypes = r.get<int32_t>();
    B.has_res = r.get<int32_t>();
    r.arr(B.x, 3 * (size_t)B.nnodes);
    r.arr(B.wbc, B.nwbc);
    r.arr(B.obc, B.nobc);
    B.nv.resize(B.ntypes);
    B.nc.resize(B.ntypes);
    B.conn.resize(B.ntypes);
    B.ncells = 0;
    for (int t = 0; t < B.ntypes; t++) {
      B.nv[t] = r.get<int32_t>();
      B.nc[t] = r.get<int32_t>();
      r.arr(B.conn[t], (size_t)B.nv[t] * B.nc[t]);
      B.ncells += B.nc[t];
    }
    for (int t = 0; t < B.ntypes; t++) B.vconn.push_back(B.conn[t].data());
    if (B.has_res) {
      r.arr(B.nres, B.nnodes);
      r.arr(B.cres, B.ncells);
    }
    r.arr(B.q, (size_t)B.nnodes * nvar);
    B.iblank.assign(B.nnodes, 1);
    B.iblank_cell.assign(B.ncells > 0 ? B.ncells : 1, 1);
  }
  if (version >= 2 && r.get<int32_t>() != 0) {
    fprintf(stderr, "tioga_class_driver: AMR cases are not driven here\n");
    return 2;
  }
  const int reduce = version >= 3 ? r.get<int32_t>() : 0;
  const int has_gids = version >= 4 ? r.get<int32_t>() : 0;
  if (has_gids)
    for (auto &B : blocks) {
      r.arr(B.cell_gid, B.ncells);
      r.arr(B.node_gid, B.nnodes);
    }
  Writer w{fopen(out, "wb"), ""};
  if (!w.f) {
    perror(out);
    return 2;
  }
  TIOGA::tioga tg;
  tg.setCommunicator(0, 0, 1);
  tg.setSymmetry(isym);
  tg.setNfringe(&nfringe);
  tg.setMexclude(&mexclude);
  for (auto &B : blocks) {
    tg.registerGridData(B.tag, B.nnodes, B.x.data(), B.iblank.data(), B.nwbc, B.nobc, B.wbc.data(), B.obc.data(),
                        B.ntypes, B.nv.data(), B.nc.data(), B.vconn.data(),
                        has_gids ? B.cell_gid.data() : NULL, has_gids ? B.node_gid.data() : NULL);
    if (B.has_res) tg.setResolutions(B.tag, B.nres.data(), B.cres.data());
    tg.set_cell_iblank(B.tag, B.iblank_cell.data());
  }
  for (int s = 0; s <= nsteps; s++) {
    if (s > 0) {
      for (auto &B : blocks)
        if (r.get<int32_t>()) {
          std::vector<double> nx;
          r.arr(nx, 3 * (size_t)B.nnodes);
          memcpy(B.x.data(), nx.data(), sizeof(double) * nx.size());
        }
      w.prefix = "s" + std::to_string(s) + ".";
    }
    tg.profile();
    tg.performConnectivity();
    dump(tg, blocks, w, has_gids != 0);
    update(tg, blocks, w, nvar, itype);
    if (reduce) {
      const std::string p0 = w.prefix;
      tg.reduce_fringes();
      w.prefix = p0 + "rf.";
      dump(tg, blocks, w, has_gids != 0);
      update(tg, blocks, w, nvar, itype);
      w.prefix = p0;
    }
  }
  fclose(w.f);
  fclose(r.f);
  return 0;
}

}  // namespace

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: tioga_class_driver case.bin out.bin\n");
    return 2;
  }
  try {
    return run(argv[1], argv[2]);
  } catch (const std::exception &e) {
    fprintf(stderr, "tioga_class_driver: %s\n", e.what());
    return 1;
  }
}
