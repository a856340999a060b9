/*
 * tioga_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded restatement of TIOGA's single-rank overset
 * connectivity path (hole map, box exchange, query gathering, ADT donor
 * search, isoparametric containment, donor selection, iblanking, cell iblank,
 * dataUpdate).  It is the checker for the HIP engine in tioga_amd/: only the
 * tests, the smoke check and bench.py's cpu_baseline leg may load it, and the
 * shipped library never links or calls it.
 *
 * Every function in tioga_oracle.c cites the reference file:line it follows.
 * Build: gcc -O2 -ffp-contract=off (no fast-math), so FP64 results are the
 * ones the reference produces on x86-64 (which has no FMA contraction by
 * default either).
 */
#ifndef TIOGA_ORACLE_H
#define TIOGA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants of the reference (codetypes.h:36-55) ---- */
#define ORC_BASE 1
#define ORC_BIGVALUE 1.0e+15
#define ORC_BIGINT 2147483647
#define ORC_TOL 1.0e-10
#define ORC_HOLEMAPSIZE 192

/* ---- leaf numerics (pinned individually against the compiled reference) ---- */
void orc_solvec(double a[3][3], double b[3], int *iflag);
void orc_newton_solve(double f[8][3], double *u, double *v, double *w);
void orc_nodal_weights(const double xv[8][3], const double *xp, double frac[8], int nvert);
double orc_cell_volume(const double xv[8][3], int nvert);
void orc_median(int *ix, double *x, int n);
void orc_find_obb(const double *x, double xc[3], double dxc[3], double vec[3][3], int nnodes);
int orc_obb_intersect(const double vA[3][3], const double xA[3], const double dxA[3],
                      const double vB[3][3], const double xB[3], const double dxB[3]);
void orc_fill_hole_map(int *sam, const int nx[3], int isym);
int orc_check_hole_map(const double *x, const int nx[3], const int *sam, const double ext[6]);
void orc_unique_nodes_octree(const double *x, const int *meshtag, double *rtag, int *itag, int n);
/* ADT over nelem 6-D boxes; adtInt[4*nelem], adtReal[6*nelem], ext[6] */
void orc_build_adt(int nelem, const double *ebox, int *adtInt, double *adtReal, double ext[6]);

/* ---- whole-path driver (single rank, mirrors TIOGA::tioga) ---- */
typedef struct OrcTioga OrcTioga;

OrcTioga *orc_create(void);
void orc_destroy(OrcTioga *t);
/* tioga::registerGridData (tioga.C:64-86); pointers are kept, not copied,
   exactly as the reference keeps the user's arrays (MeshBlock.C:46-61). */
void orc_register_grid(OrcTioga *t, int btag, int nnodes, double *xyz, int *ibl,
                       int nwbc, int nobc, int *wbcnode, int *obcnode,
                       int ntypes, const int *nv, const int *nc, int **vconn);
void orc_set_resolutions(OrcTioga *t, int btag, double *nres, double *cres);
void orc_set_cell_iblank(OrcTioga *t, int btag, int *iblank_cell);
void orc_set_params(OrcTioga *t, int isym, int nfringe, int mexclude);
void orc_profile(OrcTioga *t);
void orc_perform_connectivity(OrcTioga *t);
void orc_reduce_fringes(OrcTioga *t);
void orc_register_solution(OrcTioga *t, int btag, double *q);
void orc_data_update(OrcTioga *t, int nvar, int interptype);
void orc_get_donor_count(OrcTioga *t, int btag, int *dcount, int *fcount);
void orc_get_donor_info(OrcTioga *t, int btag, int *receptors, int *indices, double *frac, int *dcount);

/* registerGridData's global ids (tioga.h:121-123; kept by pointer).  When
   every block has node ids the search dedupes coincident queries by node id
   (search.C:48-74), as the reference's TIOGA_HAS_NODEGID build does. */
void orc_set_gids(OrcTioga *t, int btag, const unsigned long long *cell_gid, const unsigned long long *node_gid);
/* uniquenode_map (search.C:48-74) */
void orc_unique_nodes_gid(const unsigned long long *gid, double *res, int *itag, int n);
/* tioga::getReceptorInfo (tioga.C:572-671): 4 ints per live receptor;
   returns the count, copies when out != NULL */
long orc_get_receptor_info(OrcTioga *t, int *out);

/* Inspection of internal per-block state (for parity of intermediate arrays).
   Returns the number of elements; copies into out when out != NULL.
   names: "nodeRes","cellRes","xsearch","res_search","isearch","tagsearch",
   "donorId","xtag","elementList","adtIntegers","adtReals","iblank_cell",
   "obb" (15 doubles: xc,dxc,vec), "receptorIdCart","donorIdCart","rxyzCart" */
long orc_block_array(OrcTioga *t, int btag, const char *name, void *out);
/* hole map of mesh tag `mtag`: returns nx0*nx1*nx2 (0 when no wall);
   fills nx[3], ext[6] and sam when non-NULL */
long orc_hole_map(OrcTioga *t, int mtag, int nx[3], double ext[6], int *sam);

/* Number of donor weights seen outside [-0.2,1.2] since load; the reference
   aborts the job on the first one (bookKeeping.C:521-529). */
int orc_weight_aborts(void);

/* Standalone donor search used as the CPU baseline: searches nq points
   against the ADT of block `btag` built over all of its cells, sequentially,
   with containment tests; returns elapsed seconds. */
double orc_search_points(OrcTioga *t, int btag, const double *xq, int nq, int *donor);
/* The same search with the structure of another point set: the ADT of block
   btag over the cells overlapping the OBB of xobb[nobb] (what
   MeshBlock::search builds for the query set xobb), searched for each of
   xq[nq] (a subset of xobb, say) without dedupe, on nthreads POSIX threads;
   *nelem = the tree's element count.  Returns the search loop's seconds. */
double orc_search_points_obb(OrcTioga *t, int btag, const double *xobb, int nobb, const double *xq, int nq,
                             int *donor, int nthreads, int *nelem);

/* ---- AMR Cartesian background, gather direction (single rank) ---- */
/* tioga::register_amr_global_data (tioga.C:692-701): 11 ints and 6 doubles
   per patch, as CartGrid::registerData (CartGrid.C:22-70) reads them */
void orc_amr_global_data(OrcTioga *t, int nf, int qstride, const double *qnode, const int *idata,
                         const double *rdata, int ngrids, int qnodesize);
void orc_amr_patch_count(OrcTioga *t, int npatch);
void orc_amr_local_data(OrcTioga *t, int ipatch, int global_id, int *iblank, double *q);
/* CartGrid::search on arbitrary points */
void orc_amr_search(OrcTioga *t, const double *x, int n, int *donor);
int orc_perform_connectivity_amr(OrcTioga *t);
void orc_data_update_amr(OrcTioga *t, int nvar, int interptype);
long orc_amr_array(OrcTioga *t, const char *name, void *out);

#ifdef __cplusplus
}
#endif
#endif
