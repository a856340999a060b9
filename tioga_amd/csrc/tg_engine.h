// tg_engine.h -- the MI355X connectivity engine behind the tioga C ABI and
// the TIOGA::tioga class.
//
// One Engine holds one rank's share of an overset system: every registered
// block's coordinates, connectivity, resolutions, search arrays, ADTs, hole
// maps and donor tables live in HBM.  The user's host arrays are kept by
// pointer exactly as the reference keeps them (MeshBlock.C:46-61): they are
// read at preprocess (coordinates, connectivity, boundary lists, user
// resolutions) and written at the end of performConnectivity (iblank,
// iblank_cell) and dataUpdate (q).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "tg_common.h"
#include "tg_prims.h"

namespace tg {

void hip_check(hipError_t e, const char *what, const char *file, int line);
#define TG_HIP(x) ::tg::hip_check((x), #x, __FILE__, __LINE__)

// Device bytes allocated through DBuf and the library primitives by the
// calling host thread (net of frees on that thread) and its high-water mark:
// a serving slot runs its engine on one thread, so these measure what a
// request holds and needs.
struct DevBytes {
  long long now = 0, peak = 0;
  void add(long long n)
  {
    now += n;
    if (now > peak) peak = now;
  }
};
DevBytes &thread_dev_bytes();

// Owning device buffer that grows on demand.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t cap = 0;
  bool own = true;  // false: a view of another engine's array (read only)
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  T *reserve(size_t n)
  {
    if (n == 0) n = 1;
    if (n > cap) {
      if (!own) throw std::runtime_error("tioga_amd: a shared device array can not grow");
      release();
      TG_HIP(hipMalloc((void **)&p, n * sizeof(T)));
      cap = n;
      thread_dev_bytes().add((long long)(n * sizeof(T)));
    }
    return p;
  }
  // reserve in memory of its own (a view is dropped first)
  T *reserve_own(size_t n)
  {
    if (!own) release();
    return reserve(n);
  }
  // a view of n elements at q inside another buffer (never grown or freed)
  void view(T *q, size_t n)
  {
    release();
    p = q;
    cap = n;
    own = false;
  }
  void borrow(const DBuf &src)
  {
    release();
    p = src.p;
    cap = src.cap;
    own = false;
  }
  void release()
  {
    if (p && own) {
      (void)hipFree(p);
      thread_dev_bytes().add(-(long long)(cap * sizeof(T)));
    }
    p = nullptr;
    cap = 0;
    own = true;
  }
  operator T *() const { return p; }
};

constexpr int MAXTYPES = 8;

// Cell types of a block: cells [cstart[t], cstart[t+1]) have nv[t] vertices.
struct TypeInfo {
  int ntypes;
  int nv[MAXTYPES];
  int cstart[MAXTYPES + 1];
};

// One ADT node as the search walk reads it (48 bytes).  The walk only
// needs to know which nodes a point may enter and which elements' boxes may
// hold it, so the bounds are kept coarse and conservative: every bound, padded
// by the search tolerance as the reference tests it (lower - TOL,
// upper + TOL), is mapped to the tree's 16-bit grid by the monotone map
// grid_of() (tg_common.h); a point whose padded test passes also passes the
// grid test, never the other way round.  The exact FP64 box test of each
// listed element is made afterwards, by the evaluation.  Bounds are packed two
// per word: (xlo | ylo << 16), (zlo | xhi << 16), (yhi | zhi << 16) for the
// node's own element box (e), the left child's subtree box (l) and the right
// child's (r).  cellnv = cell | nv << 28; left/right: child node ids in the
// ADT's numbering (-1: none).
struct __align__(16) WalkNode {
  unsigned e[3], l[3], r[3];
  int cellnv, left, right;
};

// The 16-bit grid of one tree: grid_of(x) = clamp(((float)x - lo) * inv,
// 0, 65535), truncated (tg_common.h).
struct GridFrame {
  float lo[3], inv[3];
};

// The grid of a tree whose boxes lie inside ext (adtExtents order)
inline GridFrame grid_frame(const double ext[6])
{
  GridFrame g;
  for (int d = 0; d < 3; d++) {
    const double span = ext[2 * d + 1] - ext[2 * d];
    g.lo[d] = (float)ext[2 * d];
    const double inv = span > 0.0 ? 65535.0 / span : 0.0;
    g.inv[d] = inv < 3.0e38 ? (float)inv : 3.0e38f;
  }
  return g;
}

// Hole map of one mesh tag as the kernels see it.
struct HoleMapView {
  const unsigned char *sam;  // nx*ny*nz, 1 = inside the hole
  int nx[3];
  int exist;
  double ext[6];
};

// One interpolation record of a donor block (INTERPLIST, codetypes.h:122-131).
struct InterpHost {
  int cancel, nweights, rinfo[3];
  int inode[9];
  double w[9];
};

// Bins of one tree's walk records, for listing candidates without the walk:
// the tree's 16-bit grid cut into 2^(16 - shift) bins per axis; bin b lists,
// in any order, (packed grid box, node id | nv << 28) of every element whose
// grid box overlaps it, at ent[off[b] .. off[b + 1]).  off holds indices into
// the ent array of the batch the tree was built in (its own when built alone).
struct BinGrid {
  int shift = 16;
  long long nbins = 0;  // 0: none built
  DBuf<int> off;
  DBuf<uint4> ent;
};

// A search structure kept over every cell of a block (a served background):
// searches against the block walk it instead of a tree built per query set.
struct ResidentTree {
  int nelem = 0, depth = 0;
  double ext[6] = {0, 0, 0, 0, 0, 0};
  DBuf<int> elist, adtInt;
  DBuf<double> ebox, adtReal;
  DBuf<WalkNode> nodes;
  DBuf<int> rcell;  // the cell of each walk record (node id order)
  BinGrid bins;
};

struct Block {
  // ---- registration: the user's memory, kept by pointer ----
  int tag = 0, nnodes = 0, ncells = 0;
  int rank = 0, lidx = 0;  // virtual rank, and index among that rank's blocks
  TypeInfo ti{};
  double *ux = nullptr;
  int *uiblank = nullptr;
  int nwbc = 0, nobc = 0;
  int *uwbc = nullptr, *uobc = nullptr;
  int *uconn[MAXTYPES] = {};
  double *unres = nullptr, *ucres = nullptr;
  int *uiblank_cell = nullptr;
  std::vector<int> own_iblank_cell;
  double *uq = nullptr;
  uint64_t *ucell_gid = nullptr, *unode_gid = nullptr;
  double resolutionScale = 1.0;
  bool topo_uploaded = false;
  // 0: owns everything; 1: topology (conn, incidence, wall / overset node
  // flags) is another engine's; 2: topology and geometry (coordinates,
  // resolutions, bins, OBB) are another engine's.  Shared arrays are read only.
  int share = 0;
  bool writeback = true;  // write iblank / iblank_cell back to the user's arrays

  // ---- device state ----
  DBuf<double> x;          // 3V, AoS as registered
  DBuf<int> conn;          // 8 per cell, 0-based, unused slots -1
  DBuf<int> n2c_off, n2c;  // node -> incident cells, cell order (with multiplicity)
  DBuf<unsigned char> wallnode, obcnode;
  DBuf<int> wallcells;  // cells with a wall node (hole-map marking)
  int nwallcells = 0;
  DBuf<int> iblank, iblank_cell;
  DBuf<double> cellRes, nodeRes;
  DBuf<unsigned> nodebin;  // 10^3 bin of each node in the block's OBB frame
  DBuf<int> invmap;        // nodes ordered by (bin, descending node id)
  std::vector<int> binoff; // first position of each bin in that order (1001)
  OBB obb{};
  double mapdx[3] = {0, 0, 0};

  // search arrays (MeshBlock.h: xsearch, res_search, isearch, tagsearch, donorId, xtag)
  int nsearch = 0;
  DBuf<double> xsearch, res_search;
  DBuf<int> isearch, tagsearch, donorId, xtag;
  // ADT over the cells that overlap the query set's OBB
  int nelem = 0, adt_depth = 0;
  DBuf<int> elist, adtInt;
  DBuf<double> ebox, adtReal;
  DBuf<WalkNode> nodes;
  DBuf<int> rcell;  // the cell of each walk record (node id order)
  BinGrid bins;
  double adtExt[6] = {0, 0, 0, 0, 0, 0};
  // searches against this block walk this tree (read only; another engine
  // owns it) when set; queries whose donor could depend on the walk order
  // are counted (tie band)
  const ResidentTree *resident = nullptr;

  // donor bookkeeping (receptor side: one entry per node)
  DBuf<int> has_donor, head_arr, head_blk, head_remote;
  DBuf<unsigned long long> rmax, head_key, tagmask;
  // donor side: interpolation list
  int ninterp = 0;
  DBuf<int> interp_inode, interp_nw, interp_cancel, interp_rnode, interp_rblk, interp2donor;
  DBuf<double> interp_w;
  std::vector<InterpHost> interp_host;
  DBuf<unsigned char> upd_mark;  // dataUpdate's donor-vertex marks
  bool interp_host_valid = false;
  // point receptors (performConnectivityHighOrder without callbacks,
  // tg_points.hip): the picked nodes and their coordinates, and the point
  // donor records of this block (interpList2): vertices and weights
  // (slot m of record r at m * ninterp2 + r), receptor point and block
  int npts = 0, ninterp2 = 0;
  DBuf<int> pnode, p_inode, p_nw, p_rpt, p_rblk;
  DBuf<double> pxyz, p_w;
  // with high-order callbacks (host): receptor cells (1-based), their point
  // counts and points, and the point donor records: first solution index
  // (1-based), weights at h2_woff[r] .. h2_woff[r + 1], receptor point and block
  std::vector<int> hctag, hppc, h2_inode, h2_woff, h2_rpt, h2_rblk;
  std::vector<double> hrxyz, h2_w;
};

// The solver's high-order callbacks (tiogaInterface.C:273-277, MeshBlock.h:
// 78-83), host functions.
struct HighOrderCallbacks {
  void (*nodes_per_cell)(int *, int *) = nullptr;
  void (*receptor_nodes)(int *, int *, double *) = nullptr;
  void (*inclusion)(int *, double *, int *, double *) = nullptr;
  void (*frac)(int *, double *, int *, int *, double *, double *, int *) = nullptr;
  void (*modal)(int *, int *, double *, int *, int *, double *) = nullptr;
  bool set = false;
};

// A standalone donor search: nq points against one block, with the search
// structure built the way MeshBlock::search builds it (cells overlapping the
// query set's OBB, ADT) and no coincident-point dedupe.
struct QuerySet {
  int blk = 0, nq = 0;
  DBuf<double> xq;
  DBuf<int> donor;
  int nelem = 0, adt_depth = 0;
  DBuf<int> elist, adtInt;
  DBuf<double> ebox, adtReal;
  DBuf<WalkNode> nodes;
  DBuf<int> rcell;  // the cell of each walk record (node id order)
  BinGrid bins;
  double ext[6] = {0, 0, 0, 0, 0, 0};
};

// Reusable device scratch of the ADT build (tg_adt.hip).
struct AdtScratch {
  DBuf<int> ea, A, B, err, tb, choff, cntA, cntB, adtInt;
  DBuf<double> X, ebox, adtReal, roots;
  DBuf<int4> segs;  // segment descriptors (8 ints each)
  DBuf<int2> tab;
  DBuf<double4> state;
  DBuf<unsigned long long> bk;
  DBuf<int> bp;
};

// One ADT of a batch: n element boxes (device, 6 doubles each) in; the
// reference's adtInt[4n] / adtReal[6n] (device) and the padded root extents
// ext[6] and depth (host) out.
struct AdtTree {
  int n = 0;
  const double *ebox = nullptr;
  int *adtInt = nullptr;
  double *adtReal = nullptr;
  double ext[6] = {0, 0, 0, 0, 0, 0};
  int depth = 0;
};

// Builds every tree of the batch exactly as ADT::buildADT does, level by
// level for all trees together.
void build_adt_batch(AdtTree *trees, int ntree, AdtScratch &w, hipStream_t s);

// One tree; returns its depth in levels.
int build_adt_device(int n, const double *ebox, int *adtInt, double *adtReal, double ext[6], AdtScratch &w,
                     hipStream_t s);

// Reusable device scratch of the coincident-point dedupe (tg_dedupe.hip).
struct DedupeScratch {
  DBuf<int> ea, tmp, counts, sb, choff, cc, tag, itag;
  DBuf<int2> nodesA, nodesB, bigA, bigB, leaves, tab, subs;
  DBuf<double> x, rtag;
  DBuf<unsigned long long> bk;
  DBuf<double4> bb;
};

// One point set of a dedupe batch (device arrays): n points x[3n] with mesh
// tags; rtag[n] is raised in place, itag[n] receives the representatives.
struct DedupeSet {
  int n;
  const double *x;
  const int *tag;
  double *rtag;
  int *itag;
};

// uniquenodes_octree (utils.c:620-737) on the device for every set of the
// batch (one octree root each): the reference's representatives and raised
// resolutions.
void unique_nodes_batch(DedupeSet *sets, int nset, DedupeScratch &w, hipStream_t s);
void unique_nodes_device(int n, const double *x, const int *tag, double *rtag, int *itag, DedupeScratch &w,
                         hipStream_t s);

// One point set of an OBB batch: n points, AoS xyz, in device memory.
struct ObbSet {
  int n;
  const double *x;
};

// The device table of an OBB batch (tg_obb.hip): segments [seg0, seg1) of
// the chain sums and workgroups [blk0, blk1) of the extents pass.
struct ObbSetDev {
  const double *x;
  int n, seg0, seg1, blk0, blk1;
};

// Reusable device scratch of the OBB batch.
struct ObbScratch {
  DBuf<ObbSetDev> sets;
  DBuf<double> apx, qP, qN, xc, aa, vec, lohi, part;
  DBuf<int> ep, qbad, nstep;
};

// findOBB (utils.c:25-161) for every set of the batch, with the bits of the
// sequential host computation (find_obb); *stepped (if given) receives the
// number of segments whose terms had to be added one by one.
void find_obbs_device(const ObbSet *sets, int nset, OBB *out, ObbScratch &w, hipStream_t s,
                      long long *stepped = nullptr);

// Reusable device scratch of the donor exchange (kept between calls, so a
// connectivity step allocates nothing once the sizes have been seen).
struct DonorScratch {
  DBuf<int> dtag, pflag, psel, dnoff, lnode, ldblk, lrem, dsel, tmin, dstart, didx, errs, ireset;
  DBuf<double> lrres;
  DBuf<unsigned> dkey0, dkey1;
  DBuf<unsigned char> mt, mt1, cancel, rt, et, ct, it;
  DBuf<unsigned char> tab[10];  // per-stage segment tables and fill chunks
};

// Reusable device scratch of the other connectivity phases: hole maps, box
// exchange, query gathering.
struct StepScratch {
  DBuf<unsigned char> maps, walls, boxes, qpairs, qfill, cib;
  DBuf<int> act, chg, cand0, rng0, first, pblk, qflag, qpos, qsel;
  DBuf<int2> rng;
  DBuf<OBB> obc;
};

// Parameters of the donor search, none of which changes a result: Morton
// bits of the query order, bin-list entries per lane in one work item of the
// bin listing, workgroups of the containment-test kernels, and the bin
// density (0: candidates listed by the tree walk).
struct SearchTuning {
  int morton_bits = 14;
  int seg = 256;  // bin-list entries per lane in one work item of the listing
  int eval_grid = 8192;
  // candidates listed from bin grids with about this many elements per bin
  // (built with each tree; 0: by the tree walk)
  int bins = 8;
};

struct SearchScratch {
  SearchTuning tune;
  // a launch recorded into a HIP graph reads the search tables the previous
  // launch of the same sets uploaded (identical), so it copies none itself
  bool skip_uploads = false;
  DBuf<int> cell, query, counters, lj, lcount, qb, qsid, nin, wstart;
  DBuf<unsigned char> status;    // test outcome per candidate slot
  DBuf<int> best, last, acc;     // per query outcome words of a launch
  DBuf<int> nseg, segstart;      // the bin listing's work items per walking wave
  DBuf<unsigned char> sv;  // the table of searches of one launch
  DBuf<double4> lst;
  size_t tcap[4] = {0, 0, 0, 0};  // capacity of the candidate list per cell type (0: not sized yet)
  int last_ns = 0, last_nqt = 0;  // searches and queries of the last launch
  bool last_bins = false;         // the last launch listed candidates from bin grids
  long long last_nwave = 0;       // waves the last launch was sized for
  // the continued Newton solves run on a second stream (default priority),
  // beside the tet tests; created on first use
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  SearchScratch() = default;
  SearchScratch(const SearchScratch &) = delete;
  SearchScratch &operator=(const SearchScratch &) = delete;
  ~SearchScratch()
  {
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
  }
};

// dataUpdate (tg_update.hip): the interpolation records of one donor block
// and the thread range [start, start + n) they take in a launch (start a
// multiple of the workgroup size); voff = their first row of staged values.
struct UpdSeg {
  const int *inode;
  const double *w;
  const int *nw, *cancel, *rnode, *rblk;
  int n, start, blk, voff;
};
constexpr int UPD_MAXBLK = 128;
struct UpdTable {
  double *q[UPD_MAXBLK];  // each block's solution (device)
  int nnodes[UPD_MAXBLK];
  int nseg;
  UpdSeg seg[UPD_MAXBLK];
};
// the staging bytes of the allocation-free dataUpdate (its fixed maximum)
size_t update_scratch_bytes();
// tiogaInterface.C:178-190: 0 for an itype containing "row", 1 for
// "column", -1 otherwise
int update_interptype(const char *itype);

// The query sets of a search launch, for their Morton keys in one launch:
// set k's queries are [off, off + nq) of the concatenation.  ext: the root
// box of the set's tree (adtExtents order), which frames the Morton code and
// decides which queries walk the tree at all.
struct MortonSet {
  const double *xs;
  const int *xtag;  // representatives of coincident queries, or null
  int *donorId;
  double ext[6];
  int off, nq;
};

struct AmrState;  // tg_amr.hip

struct Pair {
  int local = 0, ob = 0, iblk_remote = 0, tag_remote = 0, send_tag = 0, recv_tag = 0;
  OBB obc{};
};

struct HoleMap {
  int exist = 0, nx[3] = {0, 0, 0};
  double ext[6] = {0, 0, 0, 0, 0, 0};
  DBuf<unsigned char> sam;
};

// What the exchanges of the last performConnectivity, reduce_fringes and
// dataUpdate sent between virtual ranks, kept while packet recording is on:
// the state each packet is read from at the moment the reference builds it
// (later stages change some of it in place).
struct PacketLog {
  bool on = false;
  bool connectivity = false, reduced = false, updated = false;
  std::vector<double> pair_rd;               // per pair: the reduced OBB sent (xc, dxc)
  std::vector<int> pair_cnt, pair_off;       // per pair: query points sent, their offset at the receiver
  std::vector<std::vector<double>> sent_res; // per block: res_search as received (before the dedupe)
  std::vector<std::vector<int>> iblank_pd;   // per block: iblank after processDonors
  std::vector<std::vector<double>> nres_pd;  // per block: nodeRes after processDonors
  // the receptor records of processDonors (comm2) in send order
  std::vector<int> rec_blk, rec_node, rec_dblk, rec_remote;
  std::vector<double> rec_rres;
  std::vector<std::vector<int>> rf_cancel;   // per block: receptors reduce_fringes drops, ascending
  int nvar = 0;
  std::vector<std::vector<double>> upd_vals; // per donor block: values of its live records (dataUpdate)
};

// Per-phase device time of the last performConnectivity (ms, HIP events).
struct PhaseTimes {
  double holemap = 0, boxes = 0, searchdata = 0, search_prep = 0, search_kernel = 0, donors = 0,
         total = 0;
  long long nsearch = 0;
  // inside search_prep (host wall ms, summed over blocks)
  double prep_obb = 0, prep_filter = 0, prep_adt = 0, prep_dedupe = 0;
  long long obb_stepped = 0;  // OBB sum segments stepped term by term
};

class Engine {
 public:
  Engine();
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  // tioga.h:106-107 defaults
  int isym = 3, nfringe = 1, mexclude = 3;
  // Batched form: the stages of performConnectivity run for all blocks
  // together (one ADT build, one dedupe, ... for every block); false runs
  // them block by block.  Results are identical either way.
  bool batched = true;
  // batched search: the coincident-point dedupe on a second stream,
  // overlapped with the OBBs, cell filter and ADT build (off for the serving
  // slots, which overlap whole requests)
  bool overlap_dedupe = true;
  int block_count() const { return (int)blocks_.size(); }
  // Virtual ranks.  Every block belongs to one rank of a virtual
  // communicator; per-block calls (registration, solutions, donor info, ...)
  // name a block by its mesh tag on the current rank, as a rank of the
  // reference names its own blocks (tag_iblk_map, tioga.h:78).  A mesh tag may
  // sit on several ranks (a partitioned mesh).  Blocks are kept in rank-major
  // order, each rank's in its registration order, which is the order of the
  // reference's OBB allgather (exchangeBoxes.C:99-117); rank 0 alone is the
  // single-rank reference.
  int cur_rank = 0;
  static long long block_key(int rank, int tag) { return ((long long)rank << 32) | (unsigned)tag; }
  int index_of(int btag) const
  {
    auto it = tag_to_blk_.find(block_key(cur_rank, btag));
    return it == tag_to_blk_.end() ? -1 : it->second;
  }
  int block_nnodes(int btag) const
  {
    const int i = index_of(btag);
    return i < 0 ? -1 : blocks_[i]->nnodes;
  }
  int block_ncells(int btag) const
  {
    const int i = index_of(btag);
    return i < 0 ? -1 : blocks_[i]->ncells;
  }
  const double *block_coords(int btag) const
  {
    const int i = index_of(btag);
    return i < 0 ? nullptr : blocks_[i]->ux;
  }
  const double *block_device_coords(int btag) const
  {
    const int i = index_of(btag);
    return i < 0 ? nullptr : blocks_[i]->x.p;
  }
  const int *block_wall_nodes(int btag, int *nwbc) const
  {
    const int i = index_of(btag);
    if (i < 0) return nullptr;
    *nwbc = blocks_[i]->nwbc;
    return blocks_[i]->uwbc;
  }
  // ranks holding blocks, and rank r's communication partners after the box
  // exchange: the ranks one of its blocks shares an OBB pair with, ascending
  // (sndMap = rcvMap, exchangeBoxes.C:166-176)
  int nranks() const;
  const std::vector<int> &partners(int r) const;
  // Packets between virtual ranks in the reference's PACKET layout
  // (codetypes.h:114-120): recording on keeps what the exchanges send;
  // packet() rebuilds the packet virtual rank src sends to dst in exchange
  // `ex` ("boxes", "search", "donors1".."donors4", "receptor_info",
  // "update", "rf.donors3").  counts = the (nints, nreals) header; returns
  // false when src and dst do not exchange.
  void record_packets(bool on);
  bool packet(const std::string &ex, int src, int dst, std::vector<int> &ints, std::vector<double> &reals,
              long long counts[2]);
  PacketLog pk_;
  void interp_records(Block &b);
  void record_update(int nvar, int interptype);
  // tioga::performConnectivityHighOrder and dataUpdate(at_points = 1) for a
  // solver without high-order callbacks (tg_points.hip)
  void perform_connectivity_points();
  void data_update_points(int nvar, int interptype);
  long points_array(int btag, const std::string &name, void *out);
  bool points_ready_ = false;
  // the high-order callbacks, and the reference's ihighGlobal: a
  // performConnectivity has run with them set (tioga.C:138), which sends
  // tioga_dataupdate_mb_ to the point update (tiogaInterface.C:207-216)
  HighOrderCallbacks ho_;
  bool ho_global_ = false;
  void points_by_callbacks(Block &D);
  void data_update_points_callbacks(int nvar, int interptype);
  DBuf<int> plist_, psel_;
  std::vector<DBuf<double> *> ptmp_;
  std::vector<DBuf<int> *> pitmp_;
  // the search structure over every cell of block btag (after preprocess)
  void build_resident_tree(int btag, ResidentTree &t);
  // searches against block btag walk t (owned by another engine) from now on
  void use_resident_tree(int btag, const ResidentTree *t);
  // tioga::profile for a block moved rigidly on the device: block btag's
  // coordinates become x' = R (x0 - c) + c + t per node (x0: device
  // coordinates of the same topology, row-major R), then it is preprocessed;
  // every other block only gets iblank = 1 (profile(btag) semantics).  The
  // host coordinates the block was registered or adopted with must hold the
  // moved wall nodes (the hole map's wall bounds read them).
  void profile_moved(int btag, const double *x0_dev, const double rot[9], const double c[3], const double t[3]);
  // queries of the last search whose donor could depend on the walk order
  // of a resident tree (more than one accepting candidate, or a near-zero
  // weight in a BIGVALUE cell)
  long long last_tieband() const { return last_tieband_; }
  // timed query-set passes replay the search launch from a HIP graph (true)
  // or launch its kernels one by one (false); results are the same
  bool search_graph = true;

  void register_grid(int btag, int nnodes, double *xyz, int *ibl, int nwbc, int nobc, int *wbc, int *obc,
                     int ntypes, const int *nv, const int *nc, int **vconn);
  void set_gids(int btag, uint64_t *cell_gid, uint64_t *node_gid);
  void set_resolutions(int btag, double *nres, double *cres);
  void set_resolutions_first(double *nres, double *cres);
  void set_cell_iblank(int btag, int *ibc);
  void set_cell_iblank_first(int *ibc);
  void profile(int only_tag = -1);
  void perform_connectivity();
  void reduce_fringes();
  void register_solution(int btag, double *q);
  void data_update(int nvar, int interptype);
  void get_donor_count(int btag, int *dcount, int *fcount);
  void get_donor_info(int btag, int *receptors, int *indices, double *frac, int *dcount);
  void get_receptor_info(std::vector<int> &receptors);

  // inspection of internal arrays (same names as the reference's members)
  long block_array(int btag, const std::string &name, void *out);

  // Serving: block `as_tag` of this engine becomes a view of block
  // `src_tag` of `src` (share 1: topology; 2: topology and geometry, see
  // Block::share); ux / uiblank are this engine's host coordinate and iblank
  // arrays (ignored for share 2); writeback false keeps iblank on the
  // device.  slot >= 0 reuses that block slot (its device buffers too).
  void adopt_block(const Engine &src, int src_tag, int as_tag, int share, double *ux, int *uiblank, bool writeback,
                   int slot = -1);
  // nodes of block btag whose iblank is not 1, with their values
  void nonunit_iblank(int btag, std::vector<int> &nodes, std::vector<int> &vals);
  long hole_map(int mtag, int nx[3], double ext[6], int *sam);

  // Re-run the donor search kernels of every block over the resident ADTs
  // and search arrays `nrep` times; returns the mean device ms per pass.
  double time_search(int nrep, long long *nsearch);
  // Re-run the dataUpdate gather/scatter on device-resident q `nrep` times
  // (q uploaded once from the registered arrays); mean device ms per pass.
  double time_data_update(int nvar, int interptype, int nrep, long long *nreceptors);
  // dataUpdate on the caller's device solutions qdev[block] (registration
  // order), staging in a fixed device array: allocates no device memory
  void data_update_noalloc(int nvar, int interptype, double *const *qdev);


  // Query sets (extension): create returns an id; search runs the kernel
  // and copies donors out; time_query_sets times nrep passes over all sets.
  int query_set_create(int btag, const double *xq, long nq);
  void query_set_search(int id, int *donor);
  double time_query_sets(int nrep, long long *nq);
  void query_set_clear();
  // mean device ms per pass of the k_search launches alone in the last
  // time_query_sets (the pass also orders the queries)
  double last_search_kernel_ms() const { return last_search_kernel_ms_; }
  // counts of the last search launch: out[0] = candidates listed, out[1] =
  // queries with more candidates than the walk stages, out[2] = queries that
  // walked a tree, out[3] = queries, out[4..7] = candidates whose test
  // rejected the point, accepted it with a near-zero weight in a BIGVALUE
  // cell, accepted it cleanly, found it outside the element box, out[8] =
  // candidates ranked after their query's lowest-ranked clean hit, out[9] =
  // 1 when the candidates came from bin grids, 0 from the tree walk
  // (synchronises the stream)
  void search_stats(long long out[10]);
  int query_set_nelem(int id) const { return (id >= 0 && id < (int)qsets_.size()) ? qsets_[id]->nelem : -1; }
  // donors of set id as the last pass over it (searched or timed) left them
  void query_set_donors(int id, int *donor);
  // root extents of set id's tree (adtExtents order); returns its size
  int query_set_extents(int id, double ext[6]) const;
  // candidate-list capacity (entries) the next donor search starts from; it
  // grows whenever a checked search finds it too small
  size_t search_capacity() const { return sw_.tcap[0] + sw_.tcap[1] + sw_.tcap[2] + sw_.tcap[3]; }
  SearchTuning &search_tuning() { return sw_.tune; }
  // the bin density of the search (0: the tree walk); the query sets' bin
  // grids are built again at a new density, trees built later use it
  void set_search_bins(int bins);
  void set_search_capacity(size_t cap)
  {
    for (size_t &c : sw_.tcap) c = std::max<size_t>(cap / 4, 1);
  }

  // AMR Cartesian background, gather direction (tg_amr.hip):
  // tioga::register_amr_global_data / set_amr_patch_count /
  // register_amr_local_data (tioga.C:692-713), performConnectivityAMR
  // (tioga.C:180-216) and dataUpdate_AMR (tioga.C:218-334) on one rank.
  void amr_global_data(int nf, int qstride, const double *qnode, const int *idata, const double *rdata, int ngrids,
                       int qnodesize);
  void amr_patch_count(int npatch);
  void amr_local_data(int ipatch, int global_id, int *iblank, double *q);
  void perform_connectivity_amr();
  void data_update_amr(int nvar, int interptype);
  // true once performConnectivityAMR ran with patches (the reference's
  // iamrGlobal, which routes tioga_dataupdate_mb_ to dataUpdate_AMR)
  bool amr_mode() const;
  // device-resident dataUpdate_AMR, `nrep` gathers; mean device ms each
  double time_data_update_amr(int nvar, int interptype, int nrep, long long *nrec);
  long amr_array(int btag, const std::string &name, void *out);

  PhaseTimes last_times;
  int weight_aborts() const { return weight_aborts_; }
  hipStream_t stream() const { return s_; }

 private:
  Block *find(int btag, int *iblk = nullptr);
  void upload_topology(Block &b);
  void tag_boundary(Block &b);
  void get_hole_map();
  void exchange_boxes();
  void exchange_search_data();
  void search(Block &b);
  void search_batched();
  void prepare_searches(const std::vector<int> &act);
  void dedupe_gid(Block &b);
  ObbScratch obbw_;           // query-set OBB batches
  DonorScratch donw_;         // donor exchange
  // a per-block segment table copied to device memory that b keeps
  template <class S>
  const S *put_segs(DBuf<unsigned char> &b, const std::vector<S> &v);
  StepScratch stw_;           // hole maps, boxes, query gathering
  // a host table copied to device memory that b keeps (grown on demand)
  template <class T>
  void *put_table(DBuf<unsigned char> &b, const T &h)
  {
    b.reserve(sizeof(T));
    TG_HIP(hipMemcpyAsync(b.p, &h, sizeof(T), hipMemcpyHostToDevice, s_));
    return b.p;
  }
  DBuf<double> wallds_;       // hole-map cell sizes, per mesh tag
  DBuf<unsigned char *> wallsams_;  // hole maps, per mesh tag
  void exchange_donors();
  void cell_iblanks(Block &b);
  void launch_search(Block &b);
  void launch_searches();
  void list_candidates(Block &b, std::vector<int> &qi, std::vector<int> &rank, std::vector<int> &cell);

  void build_search_structure(Block &b, const OBB &obq, int &nelem, int &depth, DBuf<int> &elist,
                              DBuf<double> &ebox, DBuf<int> &adtInt, DBuf<double> &adtReal, DBuf<WalkNode> &nodes,
                              DBuf<int> &rcell, double ext[6], BinGrid &bins, long long nq);
  void build_bins(int ntree, const WalkNode *const *nodes, const int *nelem, BinGrid &store, BinGrid *const *views);
  BinGrid tbins_;                      // the bin grids of every batched tree
  DBuf<int> bcnt_, beoff_;             // a bin build's entries per element and their prefix
  DBuf<unsigned char> btab_;           // the tree table of a bin build
  DBuf<long long> bsums_;              // entries per tree of a bin build
  DBuf<unsigned> bkey_, bkey2_;        // a bin build's (bin, element) pairs, sorted by bin
  DBuf<int> bval_, bval2_;
  void launch_query_sets(const std::vector<QuerySet *> &qs);
  const int *order_sets(const std::vector<MortonSet> &M, size_t tot);
  SearchScratch sw_;
  bool timing_ = false;  // inside time_query_sets: candidate capacity already established
  // the timed passes replay the search launch from a HIP graph, recorded
  // once per time_query_sets call after its checked pass
  hipGraph_t qs_graph_ = nullptr;
  hipGraphExec_t qs_exec_ = nullptr;
  void drop_search_graph();
  std::vector<hipEvent_t> *kernel_events_ = nullptr;
  size_t event_slot_ = 0;
  double last_search_kernel_ms_ = 0;
  DBuf<unsigned> mk0_, mk1_;
  DBuf<unsigned long long> wmask_;  // the unsorted ordering: walking masks per wave of queries
  DBuf<int> wcnt_, wpre_;            // their populations and exclusive sum
  DBuf<unsigned char> mset_;  // the per-set Morton table of a batched search
  DBuf<double> qx_, qres_;    // every block's query points / resolutions, back to back
  DBuf<int> qtag_, qrep_;     // every block's query mesh tags / representatives
  DBuf<double> tebox_, treal_;  // every batched tree's element boxes / adtReal, back to back
  DBuf<int> tint_, telist_;     // every batched tree's adtInt / element list, back to back
  DBuf<WalkNode> tnodes_;        // every batched tree's search records, back to back
  DBuf<int> trcell_;             // every batched tree's record cells, back to back
  DBuf<unsigned char> ntab_;    // the per-tree table of the batched search records
  DBuf<unsigned char> ptab_;    // the per-block table of the batched cell filter / element boxes
  DBuf<int> pseg_;              // their cell and element offsets
  DBuf<int> mv0_, mv1_;
  std::vector<QuerySet *> qsets_;
  void sync();

  hipStream_t s_ = nullptr;
  hipStream_t s2_ = nullptr;  // work overlapped with s_ (the batched dedupe)
  int *wb_ = nullptr;         // page-locked staging of the iblank write-back
  size_t wb_cap_ = 0;
  std::vector<hipEvent_t> wb_ev_;  // one per write-back piece
  DBuf<signed char> wb8_;          // the write-back narrowed to bytes (device)
  DBuf<int> wbflag_;               // set when a value does not fit a byte
  DBuf<unsigned char> wbtab_;      // the pieces' table
  std::vector<Block *> blocks_;
  std::map<long long, int> tag_to_blk_;  // block_key(rank, tag) -> index in blocks_
  std::vector<std::vector<int>> partners_;  // per rank (exchange_boxes)
  void reindex_blocks();
  int recv_pair(int i) const;
  std::vector<HoleMap> hm_;
  DBuf<HoleMapView> hm_view_;
  int nmesh_ = 0;
  std::vector<Pair> pairs_;
  int nsend_ = 0;
  int weight_aborts_ = 0;
  bool connected_ = false;
  // scratch
  DBuf<int> iscratch_, iscratch2_;
  DBuf<unsigned> uscratch_, uscratch2_;
  DBuf<unsigned long long> kscratch_;
  DBuf<int> err_;
  AdtScratch adtw_;
  DedupeScratch dedw_;
  double prep_acc_[4] = {0, 0, 0, 0};
  long long obb_stepped_ = 0;
  std::vector<DBuf<double> *> qdev_;
  DBuf<int> tieband_;  // tie-band count of the last search over resident trees
  DBuf<double> moved_;  // the last rigid move (R, c, t) on the device
  long long last_tieband_ = 0;
  void attach_resident(Block &b);
  // dataUpdate's plan for the current interpolation records (tg_update.hip)
  struct UpdatePlan {
    UpdTable host{};
    DBuf<unsigned char> tab;      // the table on the device
    DBuf<unsigned char *> marks;  // per block: donor-vertex marks
    DBuf<int> flag;
    DBuf<double> vals;            // staged values (two-kernel form)
    long long nthreads = 0;
    int nrec = 0;
    bool valid = false, staged = false;
  } upd_;
  void prepare_update();
  DBuf<int> qlist_;  // query positions of every pair (exchange_search_data)
  // AMR Cartesian background (tg_amr.hip)
  AmrState *amr_ = nullptr;
  AmrState &amr();
  void amr_release();
  void amr_upload_q(int nvar);
  void amr_tables();
  void amr_gather(int nvar, int interptype);
  void amr_cart_receptors();
};

}  // namespace tg
