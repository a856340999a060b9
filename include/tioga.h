// tioga.h -- source-compatible TIOGA::tioga class of the MI355X engine
// (libtioga_amd.so).
//
// Public methods, argument meaning and defaults follow the reference class
// (src/tioga.h:47-261).  Private state is an engine handle, so users
// recompile against this header.  Methods whose feature is outside this
// engine (the p4est callbacks, the AMR donor callback) are not declared, so
// a build that needs them fails at compile time.
#ifndef TIOGA_AMD_TIOGA_H
#define TIOGA_AMD_TIOGA_H

#include <stdint.h>

#include <stdexcept>
#include <vector>

#define TIOGA_HAS_UINT64T

namespace tg {
class Engine;
}

namespace TIOGA {

class tioga {
 public:
  int mexclude, nfringe;

  tioga();
  ~tioga();
  tioga(const tioga &) = delete;
  tioga &operator=(const tioga &) = delete;

  // tioga.C:31-60.  One engine serves one rank on the current HIP device;
  // any communicator type is accepted and nprocs must be 1.
  template <class Comm>
  void setCommunicator(Comm, int id_proc, int nprocs)
  {
    if (nprocs != 1 || id_proc != 0)
      throw std::runtime_error("tioga_amd: one rank per engine (nprocs must be 1)");
  }

  // tioga.C:64-86; node_gid selects the global-id coincident-point dedupe
  // of search.C:48-74 (when every block has one), and cell_gid is what
  // getReceptorInfo reports for each donor cell
  void registerGridData(int btag, int nnodes, double *xyz, int *ibl, int nwbc, int nobc, int *wbcnode,
                        int *obcnode, int ntypes, int *nv, int *nc, int **vconn, uint64_t *cell_gid = NULL,
                        uint64_t *node_gid = NULL);
  void registerSolution(int btag, double *q);  // tioga.C:88-93
  void profile(void);                          // tioga.C:95-110
  void performConnectivity(void);              // tioga.C:112-160
  // tioga.C:336-524; at_points = 1 updates the point receptors of
  // performConnectivityHighOrder
  void dataUpdate(int nvar, int interptype, int at_points = 0);
  // tioga.C:162-178: without high-order callbacks the receptor nodes of the
  // last performConnectivity are searched again as points, with them the
  // callbacks' points of the receptor cells; donors are cells of cell
  // iblank 1 (highOrder.C:219-419)
  void performConnectivityHighOrder(void);
  // tioga.h:210-222: the solver's high-order callbacks (host functions),
  // called where the reference calls them: get_nodes_per_cell,
  // get_receptor_nodes, donor_inclusion_test, donor_frac, convert_to_modal
  void setcallback(void (*f1)(int *, int *), void (*f2)(int *, int *, double *),
                   void (*f3)(int *, double *, int *, double *),
                   void (*f4)(int *, double *, int *, int *, double *, double *, int *),
                   void (*f5)(int *, int *, double *, int *, int *, double *));
  void getDonorCount(int btag, int *dcount, int *fcount);        // tioga.C:533-549
  void getDonorInfo(int btag, int *receptors, int *indices, double *frac, int *dcount);  // tioga.C:551-570
  void getReceptorInfo(std::vector<int> &receptors);  // tioga.C:572-671 (single rank)

  void setSymmetry(int syminput);                              // tioga.h:166
  void setResolutions(double *nres, double *cres);             // tioga.h:168-169
  void setResolutions(int btag, double *nres, double *cres);   // tioga.h:171-177
  void setMexclude(int *mexclude_input);                       // tioga.h:179-182
  void setNfringe(int *nfringe_input);                         // tioga.h:184-187
  void set_cell_iblank(int *iblank_cell);                      // tioga.h:189-193
  void set_cell_iblank(int btag, int *ib_cell);                // tioga.h:203-209
  void set_uniform_hex_flag(int btag, int flag);               // tioga.h:195-201 (accepted; the
                                                               // general search is always used)
  void reduce_fringes(void);                                   // tioga.C:737-885

  // AMR Cartesian background, gather direction (near-body receptors take
  // their values from Cartesian patches of pdegree 0)
  void register_amr_global_data(int nf, int qstride, double *qnodein, int *idata, double *rdata, int ngridsin,
                                int qnodesize);                                  // tioga.C:692-701
  void set_amr_patch_count(int npatchesin);                                      // tioga.C:703-708
  void register_amr_local_data(int ipatch, int global_id, int *iblank, double *q);  // tioga.C:710-713
  void performConnectivityAMR(void);                           // tioga.C:180-216
  void dataUpdate_AMR(int nvar, int interptype);               // tioga.C:218-334
  // the reference's iamrGlobal: set once performConnectivityAMR ran with
  // patches; tioga_dataupdate_mb_ then routes to dataUpdate_AMR
  int iamrGlobal() const;

  tg::Engine *engine() { return eng_; }

 private:
  tg::Engine *eng_;
};

}  // namespace TIOGA

#endif
