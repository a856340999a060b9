// tg_capi.cpp -- the TIOGA::tioga class and the tioga_* C ABI over the
// engine (include/tioga.h, include/tioga_amd.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

#define TIOGA_AMD_BUILDING  // the library side of tioga_comm_t: an int, never read
#include "../../include/tioga.h"
#include "../../include/tioga_amd.h"
#include "tg_engine.h"
#include "tg_seq.h"
#include "tg_serve.h"

namespace {

std::string g_err;
int g_status = 0;

void set_error(const char *where, const char *what)
{
  g_status = 1;
  g_err = std::string(where) + ": " + what;
  fprintf(stderr, "tioga_amd: %s\n", g_err.c_str());
}

template <class F>
void guarded(const char *where, F &&f)
{
  try {
    f();
  } catch (const std::exception &e) {
    set_error(where, e.what());
  }
}

}  // namespace

namespace TIOGA {

tioga::tioga() : mexclude(3), nfringe(1), eng_(new tg::Engine) {}

tioga::~tioga() { delete eng_; }

void tioga::registerGridData(int btag, int nnodes, double *xyz, int *ibl, int nwbc, int nobc, int *wbcnode,
                             int *obcnode, int ntypes, int *nv, int *nc, int **vconn, uint64_t *cell_gid,
                             uint64_t *node_gid)
{
  eng_->register_grid(btag, nnodes, xyz, ibl, nwbc, nobc, wbcnode, obcnode, ntypes, nv, nc, vconn);
  eng_->set_gids(btag, cell_gid, node_gid);
}

void tioga::registerSolution(int btag, double *q) { eng_->register_solution(btag, q); }

void tioga::profile(void)
{
  eng_->mexclude = mexclude;
  eng_->nfringe = nfringe;
  eng_->profile();
}

void tioga::performConnectivity(void) { eng_->perform_connectivity(); }

void tioga::dataUpdate(int nvar, int interptype, int at_points)
{
  if (at_points) eng_->data_update_points(nvar, interptype);
  else eng_->data_update(nvar, interptype);
}

void tioga::performConnectivityHighOrder(void) { eng_->perform_connectivity_points(); }

void tioga::setcallback(void (*f1)(int *, int *), void (*f2)(int *, int *, double *),
                        void (*f3)(int *, double *, int *, double *),
                        void (*f4)(int *, double *, int *, int *, double *, double *, int *),
                        void (*f5)(int *, int *, double *, int *, int *, double *))
{
  if (!f1 || !f2 || !f3 || !f4 || !f5) throw std::runtime_error("tioga_amd: a high-order callback is null");
  tg::HighOrderCallbacks &h = eng_->ho_;
  h.nodes_per_cell = f1;
  h.receptor_nodes = f2;
  h.inclusion = f3;
  h.frac = f4;
  h.modal = f5;
  h.set = true;
}

void tioga::getDonorCount(int btag, int *dcount, int *fcount) { eng_->get_donor_count(btag, dcount, fcount); }

void tioga::getDonorInfo(int btag, int *receptors, int *indices, double *frac, int *dcount)
{
  eng_->get_donor_info(btag, receptors, indices, frac, dcount);
}

void tioga::getReceptorInfo(std::vector<int> &receptors) { eng_->get_receptor_info(receptors); }

void tioga::setSymmetry(int syminput) { eng_->isym = syminput; }

void tioga::setResolutions(double *nres, double *cres) { eng_->set_resolutions_first(nres, cres); }

void tioga::setResolutions(int btag, double *nres, double *cres) { eng_->set_resolutions(btag, nres, cres); }

void tioga::setMexclude(int *m) { mexclude = *m; }

void tioga::setNfringe(int *n) { nfringe = *n; }

void tioga::set_cell_iblank(int *iblank_cell) { eng_->set_cell_iblank_first(iblank_cell); }

void tioga::set_cell_iblank(int btag, int *ib_cell) { eng_->set_cell_iblank(btag, ib_cell); }

void tioga::set_uniform_hex_flag(int, int) {}

void tioga::reduce_fringes(void) { eng_->reduce_fringes(); }

void tioga::register_amr_global_data(int nf, int qstride, double *qnodein, int *idata, double *rdata, int ngridsin,
                                     int qnodesize)
{
  eng_->amr_global_data(nf, qstride, qnodein, idata, rdata, ngridsin, qnodesize);
}

void tioga::set_amr_patch_count(int npatchesin) { eng_->amr_patch_count(npatchesin); }

void tioga::register_amr_local_data(int ipatch, int global_id, int *iblank, double *q)
{
  eng_->amr_local_data(ipatch, global_id, iblank, q);
}

void tioga::performConnectivityAMR(void) { eng_->perform_connectivity_amr(); }

void tioga::dataUpdate_AMR(int nvar, int interptype) { eng_->data_update_amr(nvar, interptype); }

int tioga::iamrGlobal() const { return eng_->amr_mode() ? 1 : 0; }

}  // namespace TIOGA

// ---------------------------------------------------------------------------
// C ABI (tiogaInterface.C)
// ---------------------------------------------------------------------------

namespace {

TIOGA::tioga *g_tg = nullptr;
tg::Server *g_srv = nullptr;  // serving over g_tg's blocks (tioga_amd_serve_*)
constexpr int MAXBLOCKS = 100;  // globals.h:30
struct InpData {
  std::vector<int> nv, nc;
  std::vector<int *> vconn;
};
InpData g_idata[MAXBLOCKS];

TIOGA::tioga &tg_or_throw()
{
  if (!g_tg) throw std::runtime_error("tioga_init_ has not been called");
  return *g_tg;
}

void register_va(int iblk, int *btag, int *nnodes, double *xyz, int *ibl, int *nwbc, int *nobc, int *wbcnode,
                 int *obcnode, int *ntypes, va_list ap)
{
  if (iblk < 0 || iblk >= MAXBLOCKS) throw std::runtime_error("block id out of range 1..100");
  InpData &d = g_idata[iblk];
  d.nv.resize(*ntypes);
  d.nc.resize(*ntypes);
  d.vconn.resize(*ntypes);
  for (int i = 0; i < *ntypes; i++) {
    d.nv[i] = *(va_arg(ap, int *));
    d.nc[i] = *(va_arg(ap, int *));
    d.vconn[i] = va_arg(ap, int *);
  }
  tg_or_throw().registerGridData(*btag, *nnodes, xyz, ibl, *nwbc, *nobc, wbcnode, obcnode, *ntypes, d.nv.data(),
                                 d.nc.data(), d.vconn.data());
}

int parse_itype(const char *itype)
{
  if (strstr(itype, "row")) return 0;
  if (strstr(itype, "column")) return 1;
  return -1;
}

void not_available(const char *what) { set_error(what, "not available in this engine (outside its scope)"); }

}  // namespace

extern "C" {

void tioga_init_f90_(int *) { tioga_init_(0); }

void tioga_init_(tioga_comm_t)
{
  guarded("tioga_init", [] {
    delete g_srv;
    g_srv = nullptr;
    delete g_tg;
    g_tg = nullptr;
    for (auto &d : g_idata) d = InpData();
    g_tg = new TIOGA::tioga;
  });
}

void tioga_registergrid_data_(int *btag, int *nnodes, double *xyz, int *ibl, int *nwbc, int *nobc, int *wbcnode,
                              int *obcnode, int *ntypes, ...)
{
  va_list ap;
  va_start(ap, ntypes);
  guarded("tioga_registergrid_data", [&] {
    register_va(0, btag, nnodes, xyz, ibl, nwbc, nobc, wbcnode, obcnode, ntypes, ap);
  });
  va_end(ap);
}

void tioga_registergrid_data_mb_(int *bid, int *btag, int *nnodes, double *xyz, int *ibl, int *nwbc, int *nobc,
                                 int *wbcnode, int *obcnode, int *ntypes, ...)
{
  va_list ap;
  va_start(ap, ntypes);
  guarded("tioga_registergrid_data_mb", [&] {
    register_va(*bid - 1, btag, nnodes, xyz, ibl, nwbc, nobc, wbcnode, obcnode, ntypes, ap);
  });
  va_end(ap);
}

void tioga_preprocess_grids_(void)
{
  guarded("tioga_preprocess_grids", [] { tg_or_throw().profile(); });
}

void tioga_performconnectivity_(void)
{
  guarded("tioga_performconnectivity", [] { tg_or_throw().performConnectivity(); });
}

void tioga_registersolution_(int *bid, double *q)
{
  guarded("tioga_registersolution", [&] { tg_or_throw().registerSolution(*bid, q); });
}

void tioga_dataupdate_mb_(int *nvar, char *itype)
{
  guarded("tioga_dataupdate_mb", [&] {
    int t = parse_itype(itype);
    if (t < 0) {
      printf("#tiogaInterface.C:dataupdate_:unknown data orientation\n");
      return;
    }
    // tiogaInterface.C:194-216
    if (tg_or_throw().engine()->ho_global_) {
      if (tg_or_throw().iamrGlobal())
        printf("Data udpate between high-order near-body and AMR cartesian Not implemented yet\n");
      else
        tg_or_throw().dataUpdate(*nvar, t, 1);
    } else if (tg_or_throw().iamrGlobal()) {
      tg_or_throw().dataUpdate_AMR(*nvar, t);
    } else {
      tg_or_throw().dataUpdate(*nvar, t);
    }
  });
}

int tioga_amd_dataupdate_mb_noalloc(int *nvar, char *itype, double **q_dev)
{
  int rc = -1;
  guarded("tioga_amd_dataupdate_mb_noalloc", [&] {
    const int t = tg::update_interptype(itype);
    if (t < 0) throw std::runtime_error("tioga_amd: dataUpdate: unknown data orientation (itype)");
    if (tg_or_throw().iamrGlobal())
      throw std::runtime_error("tioga_amd: the allocation-free dataUpdate serves the unstructured blocks only");
    tg_or_throw().engine()->data_update_noalloc(*nvar, t, q_dev);
    rc = 0;
  });
  return rc;
}

long tioga_amd_dataupdate_noalloc_max_bytes(void) { return (long)tg::update_scratch_bytes(); }

void tioga_dataupdate_(double *q, int *nvar, char *itype)
{
  guarded("tioga_dataupdate", [&] {
    tg_or_throw().registerSolution(0, q);
    tioga_dataupdate_mb_(nvar, itype);
  });
}

void tioga_writeoutputfiles_(int *, char *itype)
{
  if (parse_itype(itype) < 0) {
    printf("#tiogaInterface.C:dataupdate_:unknown data orientation\n");
    return;
  }
  fprintf(stderr, "tioga_amd: tioga_writeoutputfiles_ writes no Tecplot files in this engine\n");
}

void tioga_amd_register_gids(int btag, uint64_t *cell_gid, uint64_t *node_gid)
{
  guarded("tioga_amd_register_gids", [&] { tg_or_throw().engine()->set_gids(btag, cell_gid, node_gid); });
}

long tioga_amd_receptor_info(int *out)
{
  long r = -1;
  guarded("tioga_amd_receptor_info", [&] {
    std::vector<int> v;
    tg_or_throw().getReceptorInfo(v);
    if (out && !v.empty()) memcpy(out, v.data(), sizeof(int) * v.size());
    r = (long)v.size();
  });
  return r;
}

void tioga_getdonorcount_(int *btag, int *dcount, int *fcount)
{
  guarded("tioga_getdonorcount", [&] { tg_or_throw().getDonorCount(*btag, dcount, fcount); });
}

void tioga_getdonorinfo_(int *btag, int *receptors, int *indices, double *frac, int *dcount)
{
  guarded("tioga_getdonorinfo", [&] { tg_or_throw().getDonorInfo(*btag, receptors, indices, frac, dcount); });
}

void tioga_setsymmetry_(int *isym)
{
  guarded("tioga_setsymmetry", [&] { tg_or_throw().setSymmetry(*isym); });
}

void tioga_setresolutions_(double *nres, double *cres)
{
  guarded("tioga_setresolutions", [&] { tg_or_throw().setResolutions(nres, cres); });
}

void tioga_setresolutions_multi_(int *btag, double *nres, double *cres)
{
  guarded("tioga_setresolutions_multi", [&] { tg_or_throw().setResolutions(*btag, nres, cres); });
}

void tioga_setcelliblank_(int *iblank_cell)
{
  guarded("tioga_setcelliblank", [&] { tg_or_throw().set_cell_iblank(iblank_cell); });
}

void tioga_reduce_fringes_(void)
{
  guarded("tioga_reduce_fringes", [] { tg_or_throw().reduce_fringes(); });
}

void tioga_setnfringe_(int *nfringe)
{
  guarded("tioga_setnfringe", [&] { tg_or_throw().setNfringe(nfringe); });
}

void tioga_setmexclude_(int *mexclude)
{
  guarded("tioga_setmexclude", [&] { tg_or_throw().setMexclude(mexclude); });
}

void tioga_delete_(void)
{
  guarded("tioga_delete", [] {
    delete g_srv;
    g_srv = nullptr;
    delete g_tg;
    g_tg = nullptr;
    for (auto &d : g_idata) d = InpData();
  });
}

void tioga_register_amr_global_data_(int *nf, int *qstride, double *qnodein, int *idata, double *rdata,
                                     int *ngridsin, int *qnodesize)
{
  guarded("tioga_register_amr_global_data", [&] {
    tg_or_throw().register_amr_global_data(*nf, *qstride, qnodein, idata, rdata, *ngridsin, *qnodesize);
  });
}

void tioga_register_amr_patch_count_(int *npatches)
{
  guarded("tioga_register_amr_patch_count", [&] { tg_or_throw().set_amr_patch_count(*npatches); });
}

void tioga_register_amr_local_data_(int *ipatch, int *global_id, int *iblank, double *q)
{
  guarded("tioga_register_amr_local_data",
          [&] { tg_or_throw().register_amr_local_data(*ipatch, *global_id, iblank, q); });
}

void tioga_performconnectivity_amr_(void)
{
  guarded("tioga_performconnectivity_amr", [] { tg_or_throw().performConnectivityAMR(); });
}

void tioga_performconnectivity_highorder_(void)
{
  guarded("tioga_performconnectivity_highorder", [] { tg_or_throw().performConnectivityHighOrder(); });
}

void tioga_amd_dataupdate_points(int *nvar, char *itype)
{
  guarded("tioga_amd_dataupdate_points", [&] {
    const int it = parse_itype(itype);
    if (it < 0) throw std::runtime_error("unknown data orientation");
    tg_or_throw().dataUpdate(*nvar, it, 1);
  });
}

long tioga_amd_points_array(int btag, const char *name, void *out)
{
  long r = -1;
  guarded("tioga_amd_points_array", [&] { r = tg_or_throw().engine()->points_array(btag, name, out); });
  return r;
}
void tioga_set_highorder_callback_(void (*f1)(int *, int *), void (*f2)(int *, int *, double *),
                                   void (*f3)(int *, double *, int *, double *),
                                   void (*f4)(int *, double *, int *, int *, double *, double *, int *),
                                   void (*f5)(int *, int *, double *, int *, int *, double *))
{
  guarded("tioga_set_highorder_callback", [&] { tg_or_throw().setcallback(f1, f2, f3, f4, f5); });
}
void tioga_set_p4est_(void) { not_available("tioga_set_p4est"); }
void tioga_set_amr_callback_(void (*)(int *, double *, int *, double *)) { not_available("tioga_set_amr_callback"); }
void tioga_set_p4est_search_callback_(void (*)(double *, int *, int *, int *), void (*)(int *, int *))
{
  not_available("tioga_set_p4est_search_callback");
}

int tioga_amd_status(char *msg, int len)
{
  if (msg && len > 0) {
    strncpy(msg, g_err.c_str(), (size_t)len - 1);
    msg[len - 1] = 0;
  }
  return g_status;
}

void tioga_amd_clear_status(void)
{
  g_status = 0;
  g_err.clear();
}

long tioga_amd_block_array(int btag, const char *name, void *out)
{
  long r = -1;
  guarded("tioga_amd_block_array", [&] { r = tg_or_throw().engine()->block_array(btag, name, out); });
  return r;
}

long tioga_amd_hole_map(int mtag, int *nx, double *ext, int *sam)
{
  long r = 0;
  guarded("tioga_amd_hole_map", [&] { r = tg_or_throw().engine()->hole_map(mtag, nx, ext, sam); });
  return r;
}

double tioga_amd_time_search(int nrep, long long *nsearch)
{
  double r = 0;
  guarded("tioga_amd_time_search", [&] { r = tg_or_throw().engine()->time_search(nrep, nsearch); });
  return r;
}

double tioga_amd_time_dataupdate(int nvar, int interptype, int nrep, long long *nreceptors)
{
  double r = 0;
  guarded("tioga_amd_time_dataupdate",
          [&] { r = tg_or_throw().engine()->time_data_update(nvar, interptype, nrep, nreceptors); });
  return r;
}

// ---- serving (tg_serve.h) ----

int tioga_amd_serve_open2(int bg_tag, int nslots, double budget_gb, int flags)
{
  int rc = -1;
  guarded("tioga_amd_serve_open", [&] {
    delete g_srv;
    g_srv = nullptr;
    g_srv = new tg::Server(*tg_or_throw().engine(), bg_tag, nslots, budget_gb * 1e9, (flags & 1) != 0);
    rc = 0;
  });
  return rc;
}

int tioga_amd_serve_open(int bg_tag, int nslots, double budget_gb)
{
  return tioga_amd_serve_open2(bg_tag, nslots, budget_gb, 0);
}

long tioga_amd_serve_submit(int shell_tag, const double *rot, const double *center, const double *shift)
{
  long id = -1;
  guarded("tioga_amd_serve_submit", [&] {
    if (!g_srv) throw std::runtime_error("serving is not open");
    id = g_srv->submit(shell_tag, rot, center, shift);
  });
  return id;
}

}  // extern "C"

namespace {
int serve_wait_impl(long id, int *shell_iblank, int nshell, int *nbg, int *bg_nodes, int *bg_vals, int cap,
                    double *times, long long *nsearch, long long *stats)
{
  int rc = -1;
  guarded("tioga_amd_serve_wait", [&] {
    if (!g_srv) throw std::runtime_error("serving is not open");
    tg::ServeResult r;
    g_srv->wait(id, r);
    if (r.status != 0) throw std::runtime_error("request failed: " + r.err);
    if ((int)r.shell_iblank.size() != nshell) throw std::runtime_error("shell iblank size mismatch");
    memcpy(shell_iblank, r.shell_iblank.data(), sizeof(int) * nshell);
    *nbg = (int)r.bg_nodes.size();
    const int k = std::min(cap, *nbg);
    if (k > 0) {
      memcpy(bg_nodes, r.bg_nodes.data(), sizeof(int) * k);
      memcpy(bg_vals, r.bg_vals.data(), sizeof(int) * k);
    }
    times[0] = r.t_submit;
    times[1] = r.t_start;
    times[2] = r.t_done;
    times[3] = r.connectivity_ms;
    *nsearch = r.nsearch;
    if (stats) {
      stats[0] = r.nsearch;
      stats[1] = r.tieband;
      stats[2] = (long long)r.est_bytes;
      stats[3] = (long long)r.peak_bytes;
    }
    rc = 0;
  });
  return rc;
}
}  // namespace

extern "C" {

int tioga_amd_serve_wait(long id, int *shell_iblank, int nshell, int *nbg, int *bg_nodes, int *bg_vals, int cap,
                         double *times, long long *nsearch)
{
  return serve_wait_impl(id, shell_iblank, nshell, nbg, bg_nodes, bg_vals, cap, times, nsearch, nullptr);
}

int tioga_amd_serve_wait2(long id, int *shell_iblank, int nshell, int *nbg, int *bg_nodes, int *bg_vals, int cap,
                          double *times, long long *stats)
{
  long long ns = 0;
  return serve_wait_impl(id, shell_iblank, nshell, nbg, bg_nodes, bg_vals, cap, times, &ns, stats);
}

void tioga_amd_serve_close(void)
{
  guarded("tioga_amd_serve_close", [] {
    delete g_srv;
    g_srv = nullptr;
  });
}

void tioga_amd_set_rank(int rank)
{
  guarded("tioga_amd_set_rank", [&] {
    if (rank < 0) throw std::runtime_error("negative rank");
    tg_or_throw().engine()->cur_rank = rank;
  });
}

int tioga_amd_rank(void)
{
  int r = 0;
  guarded("tioga_amd_rank", [&] { r = tg_or_throw().engine()->cur_rank; });
  return r;
}

void tioga_amd_record_packets(int on)
{
  guarded("tioga_amd_record_packets", [&] { tg_or_throw().engine()->record_packets(on != 0); });
}

int tioga_amd_packet(const char *exchange, int src, int dst, int *ints, double *reals, long long *counts)
{
  int r = -1;
  guarded("tioga_amd_packet", [&] {
    std::vector<int> iv;
    std::vector<double> rv;
    long long hdr[2];
    if (!tg_or_throw().engine()->packet(exchange, src, dst, iv, rv, hdr)) return;
    counts[0] = hdr[0];
    counts[1] = hdr[1];
    counts[2] = (long long)iv.size();
    counts[3] = (long long)rv.size();
    if (ints && !iv.empty()) memcpy(ints, iv.data(), sizeof(int) * iv.size());
    if (reals && !rv.empty()) memcpy(reals, rv.data(), sizeof(double) * rv.size());
    r = 0;
  });
  return r;
}

void tioga_amd_search_graph(int on)
{
  guarded("tioga_amd_search_graph", [&] { tg_or_throw().engine()->search_graph = (on != 0); });
}

void tioga_amd_set_batching(int on)
{
  guarded("tioga_amd_set_batching", [&] { tg_or_throw().engine()->batched = (on != 0); });
}

double tioga_amd_time_dataupdate_amr(int nvar, int interptype, int nrep, long long *nrecords)
{
  double r = 0;
  guarded("tioga_amd_time_dataupdate_amr",
          [&] { r = tg_or_throw().engine()->time_data_update_amr(nvar, interptype, nrep, nrecords); });
  return r;
}

long tioga_amd_amr_array(int btag, const char *name, void *out)
{
  long r = -1;
  guarded("tioga_amd_amr_array", [&] { r = tg_or_throw().engine()->amr_array(btag, name, out); });
  return r;
}

int tioga_amd_query_set_create(int btag, const double *xq, long nq)
{
  int r = -1;
  guarded("tioga_amd_query_set_create", [&] { r = tg_or_throw().engine()->query_set_create(btag, xq, nq); });
  return r;
}

void tioga_amd_query_set_search(int id, int *donor)
{
  guarded("tioga_amd_query_set_search", [&] { tg_or_throw().engine()->query_set_search(id, donor); });
}

double tioga_amd_time_query_sets(int nrep, long long *nq)
{
  double r = 0;
  guarded("tioga_amd_time_query_sets", [&] { r = tg_or_throw().engine()->time_query_sets(nrep, nq); });
  return r;
}

void tioga_amd_query_set_clear(void)
{
  guarded("tioga_amd_query_set_clear", [] { tg_or_throw().engine()->query_set_clear(); });
}

int tioga_amd_search_stats(long long *out)
{
  int r = -1;
  guarded("tioga_amd_search_stats", [&] {
    tg_or_throw().engine()->search_stats(out);
    r = 0;
  });
  return r;
}

double tioga_amd_last_search_kernel_ms(void)
{
  double r = 0;
  guarded("tioga_amd_last_search_kernel_ms", [&] { r = tg_or_throw().engine()->last_search_kernel_ms(); });
  return r;
}

int tioga_amd_query_set_nelem(int id)
{
  int r = -1;
  guarded("tioga_amd_query_set_nelem", [&] { r = tg_or_throw().engine()->query_set_nelem(id); });
  return r;
}

void tioga_amd_query_set_donors(int id, int *donor)
{
  guarded("tioga_amd_query_set_donors", [&] { tg_or_throw().engine()->query_set_donors(id, donor); });
}

int tioga_amd_query_set_extents(int id, double *ext)
{
  int r = -1;
  guarded("tioga_amd_query_set_extents", [&] { r = tg_or_throw().engine()->query_set_extents(id, ext); });
  return r;
}

void tioga_amd_preprocess_block(int btag)
{
  guarded("tioga_amd_preprocess_block", [&] {
    TIOGA::tioga &t = tg_or_throw();
    t.engine()->mexclude = t.mexclude;
    t.engine()->nfringe = t.nfringe;
    t.engine()->profile(btag);
  });
}

int tioga_amd_device_arch(char *out, int len)
{
  int rc = -1;
  guarded("tioga_amd_device_arch", [&] {
    int dev = 0;
    hipDeviceProp_t prop;
    TG_HIP(hipGetDevice(&dev));
    TG_HIP(hipGetDeviceProperties(&prop, dev));
    std::string a(prop.gcnArchName);
    a = a.substr(0, a.find(':'));
    if (out && len > 0) {
      strncpy(out, a.c_str(), (size_t)len - 1);
      out[len - 1] = 0;
    }
    rc = 0;
  });
  return rc;
}

int tioga_amd_search_tuning(const int *set, int *get)
{
  int rc = -1;
  guarded("tioga_amd_search_tuning", [&] {
    tg::SearchTuning &t = tg_or_throw().engine()->search_tuning();
    if (set) {
      if (set[0] < 0 || set[0] > 30) throw std::runtime_error("tioga_amd: Morton bits must be 0..30");
      if (set[1] < 16 || set[1] > 65536) throw std::runtime_error("tioga_amd: listing segment must be 16..65536 entries");
      if (set[2] < 64 || set[2] > 65536) throw std::runtime_error("tioga_amd: evaluation grid must be 64..65536");
      t.morton_bits = set[0];
      t.seg = set[1];
      t.eval_grid = set[2];
    }
    if (get) {
      get[0] = t.morton_bits;
      get[1] = t.seg;
      get[2] = t.eval_grid;
    }
    rc = 0;
  });
  return rc;
}

int tioga_amd_search_bins(int bins)
{
  int r = -1;
  guarded("tioga_amd_search_bins", [&] {
    tg::Engine *e = tg_or_throw().engine();
    if (bins >= 0) e->set_search_bins(bins);
    r = e->search_tuning().bins;
  });
  return r;
}

long tioga_amd_search_capacity(long cap)
{
  long r = -1;
  guarded("tioga_amd_search_capacity", [&] {
    tg::Engine *e = tg_or_throw().engine();
    if (cap > 0) e->set_search_capacity((size_t)cap);
    r = (long)e->search_capacity();
  });
  return r;
}

void tioga_amd_phase_times(double *out)
{
  guarded("tioga_amd_phase_times", [&] {
    const tg::PhaseTimes &t = tg_or_throw().engine()->last_times;
    out[0] = t.holemap;
    out[1] = t.boxes;
    out[2] = t.searchdata;
    out[3] = t.search_prep;
    out[4] = t.search_kernel;
    out[5] = t.donors;
    out[6] = t.total;
    out[7] = (double)t.nsearch;
    out[8] = t.prep_obb;
    out[9] = t.prep_filter;
    out[10] = t.prep_adt;
    out[11] = t.prep_dedupe;
    out[12] = (double)t.obb_stepped;
  });
}

void tioga_amd_host_median(int *ix, double *x, int n, int lists)
{
  if (lists) tg::median_select_lists(ix, x, n);
  else tg::median_select(ix, x, n);
}

void tioga_amd_host_find_obb(const double *x, int n, double *out)
{
  tg::OBB o;
  tg::find_obb(x, nullptr, n, o);
  memcpy(out, o.xc, 24);
  memcpy(out + 3, o.dxc, 24);
  memcpy(out + 6, o.vec, 72);
}

int tioga_amd_host_build_adt(int nelem, const double *ebox, int *adtInt, double *adtReal, double *ext)
{
  return tg::build_adt(nelem, ebox, adtInt, adtReal, ext);
}

int tioga_amd_device_build_adts(int ntree, const int *n, const double *ebox, int *adtInt, double *adtReal,
                                double *ext)
{
  int rc = -1;
  guarded("tioga_amd_device_build_adts", [&] {
    if (ntree <= 0) throw std::runtime_error("no trees");
    hipStream_t s;
    TG_HIP(hipStreamCreate(&s));
    tg::AdtScratch w;
    std::vector<tg::AdtTree> trees(ntree);
    std::vector<tg::DBuf<double>> eb(ntree), ar(ntree);
    std::vector<tg::DBuf<int>> ai(ntree);
    size_t off = 0;
    for (int t = 0; t < ntree; t++) {
      if (n[t] < 0) throw std::runtime_error("negative tree size");
      trees[t].n = n[t];
      eb[t].reserve(6 * (size_t)n[t]);
      ai[t].reserve(4 * (size_t)n[t]);
      ar[t].reserve(6 * (size_t)n[t]);
      if (n[t])
        TG_HIP(hipMemcpyAsync(eb[t].p, ebox + 6 * off, sizeof(double) * 6 * n[t], hipMemcpyHostToDevice, s));
      trees[t].ebox = eb[t].p;
      trees[t].adtInt = ai[t].p;
      trees[t].adtReal = ar[t].p;
      off += n[t];
    }
    tg::build_adt_batch(trees.data(), ntree, w, s);
    off = 0;
    for (int t = 0; t < ntree; t++) {
      if (n[t]) {
        TG_HIP(hipMemcpyAsync(adtInt + 4 * off, ai[t].p, sizeof(int) * 4 * n[t], hipMemcpyDeviceToHost, s));
        TG_HIP(hipMemcpyAsync(adtReal + 6 * off, ar[t].p, sizeof(double) * 6 * n[t], hipMemcpyDeviceToHost, s));
      }
      for (int i = 0; i < 6; i++) ext[6 * t + i] = trees[t].ext[i];
      off += n[t];
    }
    TG_HIP(hipStreamSynchronize(s));
    TG_HIP(hipStreamDestroy(s));
    rc = 0;
  });
  return rc;
}

int tioga_amd_device_unique_nodes(int nset, const int *n, const double *x, const int *meshtag, double *rtag,
                                  int *xtag)
{
  int rc = -1;
  guarded("tioga_amd_device_unique_nodes", [&] {
    if (nset <= 0) throw std::runtime_error("no point sets");
    hipStream_t s;
    TG_HIP(hipStreamCreate(&s));
    tg::DedupeScratch w;
    std::vector<tg::DedupeSet> sets(nset);
    std::vector<tg::DBuf<double>> dx(nset), dr(nset);
    std::vector<tg::DBuf<int>> dt(nset), di(nset);
    size_t off = 0;
    for (int k = 0; k < nset; k++) {
      if (n[k] < 0) throw std::runtime_error("negative point count");
      dx[k].reserve(3 * (size_t)n[k]);
      dr[k].reserve(n[k]);
      dt[k].reserve(n[k]);
      di[k].reserve(n[k]);
      if (n[k]) {
        TG_HIP(hipMemcpyAsync(dx[k].p, x + 3 * off, sizeof(double) * 3 * n[k], hipMemcpyHostToDevice, s));
        TG_HIP(hipMemcpyAsync(dr[k].p, rtag + off, sizeof(double) * n[k], hipMemcpyHostToDevice, s));
        TG_HIP(hipMemcpyAsync(dt[k].p, meshtag + off, sizeof(int) * n[k], hipMemcpyHostToDevice, s));
      }
      sets[k] = tg::DedupeSet{n[k], dx[k].p, dt[k].p, dr[k].p, di[k].p};
      off += n[k];
    }
    tg::unique_nodes_batch(sets.data(), nset, w, s);
    off = 0;
    for (int k = 0; k < nset; k++) {
      if (n[k]) {
        TG_HIP(hipMemcpyAsync(rtag + off, dr[k].p, sizeof(double) * n[k], hipMemcpyDeviceToHost, s));
        TG_HIP(hipMemcpyAsync(xtag + off, di[k].p, sizeof(int) * n[k], hipMemcpyDeviceToHost, s));
      }
      off += n[k];
    }
    TG_HIP(hipStreamSynchronize(s));
    TG_HIP(hipStreamDestroy(s));
    rc = 0;
  });
  return rc;
}

int tioga_amd_device_find_obbs(int nset, const int *n, const double *x, double *out, long long *stepped)
{
  int rc = -1;
  guarded("tioga_amd_device_find_obbs", [&] {
    if (nset <= 0) throw std::runtime_error("no point sets");
    hipStream_t s;
    TG_HIP(hipStreamCreate(&s));
    tg::ObbScratch w;
    std::vector<tg::ObbSet> sets(nset);
    std::vector<tg::DBuf<double>> dx(nset);
    std::vector<tg::OBB> o(nset);
    size_t off = 0;
    for (int k = 0; k < nset; k++) {
      if (n[k] <= 0) throw std::runtime_error("empty point set");
      dx[k].reserve(3 * (size_t)n[k]);
      TG_HIP(hipMemcpyAsync(dx[k].p, x + 3 * off, sizeof(double) * 3 * n[k], hipMemcpyHostToDevice, s));
      sets[k] = tg::ObbSet{n[k], dx[k].p};
      off += n[k];
    }
    tg::find_obbs_device(sets.data(), nset, o.data(), w, s, stepped);
    TG_HIP(hipStreamSynchronize(s));
    TG_HIP(hipStreamDestroy(s));
    for (int k = 0; k < nset; k++) {
      memcpy(out + 15 * k, o[k].xc, 24);
      memcpy(out + 15 * k + 3, o[k].dxc, 24);
      memcpy(out + 15 * k + 6, o[k].vec, 72);
    }
    rc = 0;
  });
  return rc;
}

void tioga_amd_host_unique_nodes(const double *x, const int *meshtag, double *rtag, int *xtag, int n)
{
  tg::unique_nodes_octree(x, meshtag, rtag, xtag, n);
}

int tioga_amd_weight_aborts(void)
{
  int r = 0;
  guarded("tioga_amd_weight_aborts", [&] { r = tg_or_throw().engine()->weight_aborts(); });
  return r;
}

}  // extern "C"
