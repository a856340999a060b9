/*
 * tioga_amd.h -- C ABI of the MI355X overset-connectivity engine
 * (libtioga_amd.so).
 *
 * The tioga_* entry points are the ones TIOGA's Fortran/C interface exports,
 * with the same names, argument meaning (every argument by pointer, 1-based
 * block ids and node lists, mesh tags >= 1) and call order
 * (init -> register -> [set*] -> preprocess -> performconnectivity ->
 * registersolution -> dataupdate -> [getdonor*] -> delete).  Each
 * declaration cites the reference function it replaces in
 * src/tiogaInterface.C.  The engine runs one rank's share of the work on the
 * current HIP device; the communicator argument is accepted and ignored, and
 * a job with more than one rank is rejected.
 *
 * The reference reports errors by printf and MPI_Abort; these entry points
 * keep its void signatures, print the error to stderr and record it, and
 * tioga_amd_status() returns it (0 = no error since the last reset).
 *
 * The tioga_amd_* entry points are extensions with no counterpart in the
 * reference: status, inspection of internal arrays, and device timing.
 */
#ifndef TIOGA_AMD_H
#define TIOGA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The communicator of tioga_init_ (tiogaInterface.C:61).  The library runs
   one rank and never reads the handle, and on x86-64 an int (MPICH's
   MPI_Comm) and a pointer (Open MPI's) both travel in the first integer
   register, so one library binds to either.  With <mpi.h> included first the
   declaration takes the caller's own MPI_Comm; otherwise int, or the type
   TIOGA_AMD_COMM_T names. */
#if defined(TIOGA_AMD_COMM_T)
typedef TIOGA_AMD_COMM_T tioga_comm_t;
#elif defined(MPI_VERSION) && !defined(TIOGA_AMD_BUILDING)
typedef MPI_Comm tioga_comm_t;
#else
typedef int tioga_comm_t;
#endif

/* tiogaInterface.C:36-59 */
void tioga_init_f90_(int *scomm);
/* tiogaInterface.C:61-82 */
void tioga_init_(tioga_comm_t tcomm);
/* tiogaInterface.C:84-106: per cell type, the variadic arguments are
   (int *nv, int *nc, int *conn) with conn 1-based, nv*nc entries. */
void tioga_registergrid_data_(int *btag, int *nnodes, double *xyz, int *ibl, int *nwbc, int *nobc,
                              int *wbcnode, int *obcnode, int *ntypes, ...);
/* tiogaInterface.C:108-130 (bid is 1-based, at most 100 blocks) */
void tioga_registergrid_data_mb_(int *bid, int *btag, int *nnodes, double *xyz, int *ibl, int *nwbc,
                                 int *nobc, int *wbcnode, int *obcnode, int *ntypes, ...);
/* tiogaInterface.C:150-153 */
void tioga_preprocess_grids_(void);
/* tiogaInterface.C:155-158 */
void tioga_performconnectivity_(void);
/* tiogaInterface.C:170-173 */
void tioga_registersolution_(int *bid, double *q);
/* tiogaInterface.C:175-216: itype contains "row" (q[node*nvar+k]) or
   "column" (q[k*nnodes+node]) */
void tioga_dataupdate_mb_(int *nvar, char *itype);
/* tiogaInterface.C:218-224 (registers q under tag 0, as the reference does) */
void tioga_dataupdate_(double *q, int *nvar, char *itype);
/* tiogaInterface.C:226-243 (Tecplot output is not produced; prints a note) */
void tioga_writeoutputfiles_(int *nvar, char *itype);
/* tiogaInterface.C:244-247 */
void tioga_getdonorcount_(int *btag, int *dcount, int *fcount);
/* tiogaInterface.C:248-251: receptors[4*dcount] = (rank, node, block, nw);
   indices/frac hold nw+1 entries per receptor, the last = donor cell, 0.0 */
void tioga_getdonorinfo_(int *btag, int *receptors, int *indices, double *frac, int *dcount);
/* tiogaInterface.C:253-256 */
void tioga_setsymmetry_(int *isym);
/* tiogaInterface.C:258-261 (first registered block) */
void tioga_setresolutions_(double *nres, double *cres);
/* tiogaInterface.C:263-266 */
void tioga_setresolutions_multi_(int *btag, double *nres, double *cres);
/* tiogaInterface.C:268-271 (first registered block) */
void tioga_setcelliblank_(int *iblank_cell);
/* tiogaInterface.C:302-305 */
void tioga_reduce_fringes_(void);
/* tiogaInterface.C:307-310 */
void tioga_setnfringe_(int *nfringe);
/* tiogaInterface.C:312-315 */
void tioga_setmexclude_(int *mexclude);
/* tiogaInterface.C:317-328 */
void tioga_delete_(void);

/* AMR Cartesian background, gather direction: near-body receptors (nodes of
   cells whose nodes all carry nodeRes == BIGVALUE and whose cell iblank is 1,
   cartOps.C:61-160) take their values from Cartesian patches of pdegree 0
   by linear interpolation between cell centres.  One rank: every grid has
   proc_id 0 and a local id of a registered patch.
   tiogaInterface.C:132-136 (11 ints: global id, level, proc, porder, local
   id, ilo[3], ihi[3]; 6 doubles: xlo[3], dx[3] per grid) */
void tioga_register_amr_global_data_(int *nf, int *qstride, double *qnodein, int *idata, double *rdata,
                                     int *ngridsin, int *qnodesize);
/* tiogaInterface.C:140-143 */
void tioga_register_amr_patch_count_(int *npatches);
/* tiogaInterface.C:145-148; q holds the patch's fields, field n at
   n*prod(dims+2*nf) (get_amr_index_xyz.c:27-45) */
void tioga_register_amr_local_data_(int *ipatch, int *global_id, int *iblank, double *q);
/* tiogaInterface.C:165-168 (tioga.C:180-216 without the Cartesian-receptor
   direction and without its closing MPI_Abort); writes the near-body iblank
   (-1 at receptors that found a patch) and iblank_cell.  After it,
   tioga_dataupdate_mb_ runs dataUpdate_AMR (tiogaInterface.C:194-204). */
void tioga_performconnectivity_amr_(void);

/* tiogaInterface.C:160-163: tioga::performConnectivityHighOrder
   (tioga.C:162-178).  Without high-order callbacks the receptor nodes of the
   last performconnectivity are searched again as points, and each found in
   a donor cell of cell iblank 1 gets a point donor record with its
   isoparametric weights (highOrder.C:219-419); tioga_dataupdate_mb_ then
   still updates the node receptors (ihighGlobal stays 0 without callbacks)
   and tioga_amd_dataupdate_points runs the point update. */
void tioga_performconnectivity_highorder_(void);

/* tiogaInterface.C:273-284: the solver's high-order callbacks, host
   functions the engine calls where the reference calls them:
   get_nodes_per_cell and get_receptor_nodes for the points of every cell of
   cell iblank -1, donor_inclusion_test on each point's candidate cells in
   the reference's walk order (the device search lists them), donor_frac
   per point donor, convert_to_modal per updated cell.  After the next
   performconnectivity, tioga_dataupdate_mb_ runs the point update
   (tiogaInterface.C:207-216), row layout only as in the reference. */
void tioga_set_highorder_callback_(void (*f1)(int *, int *), void (*f2)(int *, int *, double *),
                                   void (*f3)(int *, double *, int *, double *),
                                   void (*f4)(int *, double *, int *, int *, double *, double *, int *),
                                   void (*f5)(int *, int *, double *, int *, int *, double *));
/* Out of scope of this engine (the p4est callbacks, the AMR donor callback
   of pdegree > 0 patches); they link, print that the feature is not
   available and set the error status.  tiogaInterface.C:286-300. */
void tioga_set_p4est_(void);
void tioga_set_amr_callback_(void (*f1)(int *, double *, int *, double *));
void tioga_set_p4est_search_callback_(void (*f1)(double *xsearch, int *process_id, int *cell_id, int *npts),
                                      void (*f2)(int *pid, int *iflag));

/* ---------------- extensions (no counterpart in the reference) ---------------- */

/* Error status of the global instance: 0 when no call has failed since the
   last tioga_amd_clear_status(); the message is copied into msg (len bytes). */
int tioga_amd_status(char *msg, int len);
void tioga_amd_clear_status(void);

/* Internal per-block arrays under the reference's member names ("nodeRes",
   "cellRes", "obb", "xsearch", "res_search", "isearch", "tagsearch",
   "donorId", "xtag", "elementList", "adtIntegers", "adtReals",
   "adtExtents", "iblank_cell").  Returns the element count (-1: unknown
   name or tag) and copies into out when out != NULL. */
long tioga_amd_block_array(int btag, const char *name, void *out);

/* Hole map of mesh tag mtag as {0,1} ints (1 = inside the hole); returns
   nx*ny*nz, or 0 when the mesh has no wall. */
long tioga_amd_hole_map(int mtag, int *nx, double *ext, int *sam);

/* Device time (ms per pass, HIP events on the engine's stream) of nrep
   re-runs of the donor-search kernels over the resident search structures of
   the last performconnectivity; *nsearch = donor searches per pass. */
double tioga_amd_time_search(int nrep, long long *nsearch);

/* Device time (ms per pass) of nrep dataUpdate gather/scatter passes with q
   resident in HBM (uploaded once from the registered arrays); *nreceptors =
   interpolation records per pass. */
double tioga_amd_time_dataupdate(int nvar, int interptype, int nrep, long long *nreceptors);

/* Serving: requests that move one registered block rigidly and ask for its
   connectivity with the background block alone (the two-block
   performconnectivity on the moved coordinates, x' = R (x - c) + c + t per
   node), served by nslots concurrent slots that share the background's
   resident device data; requests start in arrival order while the device
   memory of every slot (running: the request's charge, the most a request
   on that block has held; idle: what the slot keeps) fits budget_gb; one
   request always runs.  Open after preprocess; the registered
   blocks must not change while serving.  submit returns a request id (-1 on
   error); wait blocks until that request is done and fills the moved
   block's iblank (nshell entries), the background nodes whose iblank is not
   1 (count in *nbg, up to cap entries copied), times[4] = submit, start and
   done (seconds, one steady clock) and connectivity ms, and *nsearch. */
int tioga_amd_serve_open(int bg_tag, int nslots, double budget_gb);
long tioga_amd_serve_submit(int shell_tag, const double *rot9, const double *center3, const double *shift3);
int tioga_amd_serve_wait(long id, int *shell_iblank, int nshell, int *nbg, int *bg_nodes, int *bg_vals, int cap,
                         double *times, long long *nsearch);
void tioga_amd_serve_close(void);
/* tioga_amd_serve_open with flags: bit 0 = resident-tree mode (searches
   against the background walk one tree over all its cells, built at open,
   instead of a tree per request; a request's donors may then differ from the
   per-request tree's only at queries whose donor depends on the walk order,
   which tioga_amd_serve_wait2 counts). */
int tioga_amd_serve_open2(int bg_tag, int nslots, double budget_gb, int flags);
/* tioga_amd_serve_wait with stats[4] = donor searches, tie-band queries
   (resident-tree mode), device bytes the admission charged the request, and
   device bytes its slot held at most while serving it. */
int tioga_amd_serve_wait2(long id, int *shell_iblank, int nshell, int *nbg, int *bg_nodes, int *bg_vals, int cap,
                          double *times, long long *stats);

/* Virtual ranks: one process holds the blocks of every rank of an MPI run
   (parallelComm.C:96-170 is the exchange between them).  Per-block calls
   that follow (registration, resolutions, cell iblanks, solutions, donor
   count and info, receptor info) act on the blocks of this rank, each named
   by its mesh tag as the reference's rank names its own blocks
   (tag_iblk_map, tioga.h:78); a mesh tag may sit on several ranks, a
   partitioned mesh.  Collective calls (preprocess, performconnectivity,
   dataupdate, reduce_fringes) run every rank together.  Results per rank
   equal the reference run with that many MPI ranks: every array, the rank
   and per-rank block labels in donor info and the search arrays, and the
   packets of tioga_amd_packet.  Rank 0, the default, alone is the
   single-rank reference; tioga_init_ resets it.  Each rank holds at most
   one block per mesh tag, as in the reference. */
void tioga_amd_set_rank(int rank);
int tioga_amd_rank(void);

/* The packets virtual ranks exchange, in the reference's PACKET layout
   (codetypes.h:114-120) and order (parallelComm.C:96-170): what rank src
   sends rank dst in one exchange.  tioga_amd_record_packets(1) keeps what
   the exchanges send from then on (0 stops and frees it).  exchange is one
   of "boxes" (exchangeBoxes.C:184-243), "search" (exchangeSearchData.C:
   84-112), "donors1" .. "donors4" (the four rounds of exchangeDonors.C),
   "receptor_info" (tioga.C:572-650), "update" (tioga.C:368-427) and
   "rf.donors3" (the cancellations of reduce_fringes, tioga.C:764-807); the
   exchanges after reduce_fringes are its "donors4", "update" and
   "receptor_info" asked again.  counts[0..1] = the packet's (nints, nreals)
   header, counts[2..3] = the ints and doubles written to ints / reals
   (either may be NULL to size them first; "boxes" writes the six doubles
   per pair that the reference fills of the eighteen it sends).  Returns 0,
   or -1 when src and dst do not exchange. */
void tioga_amd_record_packets(int on);

/* tioga::dataUpdate(nvar, itype, at_points = 1) (tioga.C:336-524) through
   the C ABI, which reaches it in the reference only when high-order
   callbacks are set: after tioga_performconnectivity_highorder_, every
   point receptor takes its value from its point donor; picked nodes no
   donor reached become field nodes (iblank 1, highOrder.C:193-206). */
void tioga_amd_dataupdate_points(int *nvar, char *itype);

/* The point path's arrays of block btag on the current rank, under the
   reference's member names: "ntotalPoints", "rxyz", "ninterp2",
   "interp2_inode", "interp2_weights" (nweights per record, record after
   record) and "interp2_rinfo" (3 per record).  Returns the element count
   (out may be NULL), -1 for an unknown block or name. */
long tioga_amd_points_array(int btag, const char *name, void *out);
int tioga_amd_packet(const char *exchange, int src, int dst, int *ints, double *reals, long long *counts);

/* tioga_amd_time_query_sets' timed passes replay their search launch from
   a HIP graph recorded after the checked pass (on != 0, the default) or
   launch its kernels one by one (0).  Results are the same. */
void tioga_amd_search_graph(int on);

/* Batched form (on != 0, the default): the stages of performconnectivity
   run for all blocks together (one ADT build, one coincident-point dedupe
   for every block's query set, ...); 0 runs them block by block.  The
   results are identical either way. */
void tioga_amd_set_batching(int on);

/* Device time (ms per update) of nrep dataUpdate_AMR gathers with the
   patch and block solutions resident in HBM; *nrecords = interpolation
   records per update. */
double tioga_amd_time_dataupdate_amr(int nvar, int interptype, int nrep, long long *nrecords);

/* AMR state after tioga_performconnectivity_amr_: for btag >= 1 the block's
   "receptorIdCart", "donorIdCart" (grid index or -1) and "rxyzCart"; for
   btag 0 the patch interpolation records in list order: "rec_patch",
   "rec_info" (rank, point, block), "rec_ijk" (8 stencil cells i,j,k),
   "rec_w" (8 weights) and "rec_unique_cells" (1 int: distinct donor cells).
   Returns the element count (-1: unknown name). */
long tioga_amd_amr_array(int btag, const char *name, void *out);

/* Query sets: nq points (xyz AoS, host memory) searched against block
   btag, whose search structure is built the way MeshBlock::search builds it
   (search.C:107-233: cells overlapping the query set's OBB, ADT) with no
   coincident-point dedupe.  The block must have been preprocessed.
   create returns a set id (>= 0, or -1 on error); search writes donor[nq]
   (cell index or -1); time returns the device ms of one pass over all sets
   (mean of nrep) and the total query count; clear frees every set. */
int tioga_amd_query_set_create(int btag, const double *xq, long nq);
void tioga_amd_query_set_search(int id, int *donor);
double tioga_amd_time_query_sets(int nrep, long long *nq);
/* device ms per pass of the search kernel launches alone in the last
   tioga_amd_time_query_sets (HIP events around each launch) */
double tioga_amd_last_search_kernel_ms(void);
/* counts of the last search launch: out[0] = candidate cells listed,
   out[1] = listing work items (walking waves, or their bin-list segments),
   out[2] = queries that searched a tree (inside its root box, not a
   coincident copy), out[3] = queries, out[4..7] = listed candidates whose
   containment test rejected the point, accepted it with a near-zero weight
   in a BIGVALUE cell, accepted it cleanly, or found it outside the element's
   padded box, out[8] = candidates ranked after their query's lowest-ranked
   clean hit, which the reference's walk never tests, out[9] = 1 when the
   candidates were listed from bin grids, 0 by the tree walk (10 values);
   returns 0, or -1 with the error status set */
int tioga_amd_search_stats(long long *out);
void tioga_amd_query_set_clear(void);
/* number of cells in a query set's search structure (-1: bad id) */
int tioga_amd_query_set_nelem(int id);
/* donors of query set id as the last pass over it left them (a search or
   a timed pass of tioga_amd_time_query_sets), without searching again */
void tioga_amd_query_set_donors(int id, int *donor);
/* root extents of query set id's search structure, in the order of
   adtExtents (xmin, xmax, ymin, ymax, zmin, zmax); returns its cell count
   (-1: bad id) */
int tioga_amd_query_set_extents(int id, double *ext);
/* Capacity (entries) of the donor search's candidate list: cap > 0 sets the
   capacity the next search starts from (a checked search grows it when its
   candidates do not fit and walks again); returns the current capacity. */
long tioga_amd_search_capacity(long cap);
/* Launch parameters of the donor search, none of which changes a result:
   {Morton bits of the query order (0..30), bin-list entries per query in one
   work item of the bin listing (16..65536), workgroups of each
   containment-test kernel (64..65536)}.  set (may be NULL) applies 3 values, get (may be NULL)
   receives the current ones; returns 0, or -1 with the error status set.
   tioga_amd/autotune.py chooses them by measurement. */
int tioga_amd_search_tuning(const int *set, int *get);
/* How the donor search lists each query's candidate cells, which changes no
   result: bins > 0 lists them from a bin grid built with each search tree,
   about `bins` elements per bin (a few hundred MB for a tree of millions of
   cells); 0 walks the tree instead.  bins >= 0 sets it (the registered
   query sets' grids are built again), -1 only queries; returns the current
   value, or -1 with the error status set. */
int tioga_amd_search_bins(int bins);
/* tioga_preprocess_grids_ for a solver that moved one block only: block
   btag is preprocessed in full (tioga.C:95-110), every other block keeps the
   geometry of its last preprocess (its coordinates, resolutions and
   topology must not have changed since) and gets iblank = 1, which is what
   a full preprocess gives them. */
void tioga_amd_preprocess_block(int btag);
/* The current device's architecture name (e.g. "gfx950", without feature
   suffixes) into out (len bytes); returns 0, or -1 with the error status set.
   Needs a GPU. */
int tioga_amd_device_arch(char *out, int len);

/* Host wall times (ms) of the last performconnectivity phases:
   out[0..6] = hole map, boxes, query gathering, search preparation,
   search kernels (device), donor selection, total; out[7] = searches;
   out[8..11] = inside search preparation: query-set OBB, cell filter and
   element boxes, ADT build, coincident-point dedupe; out[12] = segments of
   the OBB running sums stepped through term by term. */
void tioga_amd_phase_times(double *out);

/* Host-side sequential stages, callable without a device (they are what
   the engine runs on the host; exposed so they can be checked on their own):
   median_select of median.F90:59-210 on x[0..n) permuting ix (lists != 0
   runs the list-based partition rounds the device build uses); the
   principal-axis OBB of utils.c:25-165 (out = xc[3], dxc[3], vec[3][3]);
   the ADT of ADT.C:34-135 (returns the depth); the coincident-point dedupe
   of utils.c:620-737. */
void tioga_amd_host_median(int *ix, double *x, int n, int lists);
void tioga_amd_host_find_obb(const double *x, int n, double *out);
int tioga_amd_host_build_adt(int nelem, const double *ebox, int *adtInt, double *adtReal, double *ext);
/* The device ADT builder on ntree trees at once (tree t: n[t] boxes; all
   arrays concatenated in tree order, host memory; ext: 6 per tree); returns
   0, or -1 with the error status set.  Needs a GPU. */
int tioga_amd_device_build_adts(int ntree, const int *n, const double *ebox, int *adtInt, double *adtReal,
                                double *ext);
void tioga_amd_host_unique_nodes(const double *x, const int *meshtag, double *rtag, int *xtag, int n);
/* The device coincident-point dedupe on nset point sets at once (set k: n[k]
   points; arrays concatenated in set order, host memory; rtag raised in
   place, xtag = representative within the set); returns 0, or -1 with the
   error status set.  Needs a GPU. */
int tioga_amd_device_unique_nodes(int nset, const int *n, const double *x, const int *meshtag, double *rtag,
                                  int *xtag);
/* The device OBB of findOBB (utils.c:25-161) for nset point sets at once
   (set k: n[k] > 0 points, AoS, concatenated in set order, host memory;
   out: 15 doubles per set as tioga_amd_host_find_obb writes them); the bits
   of tioga_amd_host_find_obb.  stepped (may be NULL) receives the number of
   512-point segments whose running sums were stepped through term by term.
   Returns 0, or -1 with the error status set.  Needs a GPU. */
int tioga_amd_device_find_obbs(int nset, const int *n, const double *x, double *out, long long *stepped);

/* dataUpdate without device allocation -- no counterpart in the reference.
   For callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc): the same
   update as tioga_dataupdate_mb_ (itype "row" or "column", nvar fields),
   applied in place to the caller's device arrays q_dev[b] (the solution of
   the b-th registered block, in registration order, each
   nnodes(b) * nvar doubles).  Its staging (the per-block table and, when
   some receptor is also a donor vertex, the staged values) lives in one
   fixed device array of tioga_amd_dataupdate_noalloc_max_bytes() bytes
   (64 MiB) that each GPU holds until the process ends; an update that needs
   more is rejected with an error.  Calls that overlap on one GPU share that
   array and race.  Synchronous; returns 0, or -1 with the error status set.
   Needs a performconnectivity first. */
int tioga_amd_dataupdate_mb_noalloc(int *nvar, char *itype, double **q_dev);
long tioga_amd_dataupdate_noalloc_max_bytes(void);

/* Global ids of block btag, as the C++ registerGridData(..., cell_gid,
   node_gid) takes them (tioga.h:121-123; kept by pointer, must stay valid).
   When every block has node ids, coincident queries are deduped by node id
   (search.C:48-74, the reference's TIOGA_HAS_NODEGID build) instead of by
   position.  Call after registering the block. */
void tioga_amd_register_gids(int btag, uint64_t *cell_gid, uint64_t *node_gid);
/* tioga::getReceptorInfo (tioga.C:572-671): per live receptor (node, mesh
   tag of its block, the donor cell's 64-bit global id as two ints in memory
   order; the cell index when the block has no cell ids).  Returns the
   number of ints (-1 on error) and copies them when out != NULL. */
long tioga_amd_receptor_info(int *out);

/* Donor weights outside [-0.2, 1.2] seen (the reference aborts on one). */
int tioga_amd_weight_aborts(void);

#ifdef __cplusplus
}
#endif

#endif
