// tg_packets.cpp -- the packets virtual ranks exchange, rebuilt in the
// reference's PACKET layout (codetypes.h:114-120: nints, nreals, intData,
// realData) from what the engine's stages recorded.  One engine holds every
// rank's blocks and runs each stage for all of them at once; a multi-GPU run
// would move exactly these buffers between ranks (parallelComm.C:96-170).
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "tg_engine.h"

namespace tg {

void Engine::record_packets(bool on)
{
  pk_ = PacketLog();
  pk_.on = on;
}

// MeshBlock::getInterpolatedSolution (dataUpdate.C:27-103): the values each
// live record sends, from the solutions as they are before the update.
void Engine::record_update(int nvar, int interptype)
{
  pk_.nvar = nvar;
  pk_.upd_vals.assign(blocks_.size(), std::vector<double>());
  std::vector<double> qq(nvar);
  for (size_t ib = 0; ib < blocks_.size(); ib++) {
    Block &D = *blocks_[ib];
    interp_records(D);
    const double *q = D.uq;
    for (const InterpHost &h : D.interp_host) {
      if (h.cancel) continue;
      for (int k = 0; k < nvar; k++) qq[k] = 0;
      for (int m = 0; m < h.nweights; m++) {
        const int inode = h.inode[m];
        const double w = h.w[m];
        for (int k = 0; k < nvar; k++)
          qq[k] += (interptype == 0 ? q[(size_t)inode * nvar + k] : q[(size_t)k * D.nnodes + inode]) * w;
      }
      pk_.upd_vals[ib].insert(pk_.upd_vals[ib].end(), qq.begin(), qq.end());
    }
  }
  pk_.updated = true;
}

namespace {

template <class T>
std::vector<T> download(const DBuf<T> &d, size_t n, hipStream_t s)
{
  std::vector<T> h(n);
  if (n) {
    TG_HIP(hipMemcpyAsync(h.data(), d.p, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
  }
  return h;
}

void put_gid(std::vector<int> &ints, uint64_t g)
{
  int two[2];
  memcpy(two, &g, sizeof g);
  ints.push_back(two[0]);
  ints.push_back(two[1]);
}

}  // namespace

bool Engine::packet(const std::string &ex, int src, int dst, std::vector<int> &ints, std::vector<double> &reals,
                    long long counts[2])
{
  ints.clear();
  reals.clear();
  counts[0] = counts[1] = 0;
  if (!pk_.on || !pk_.connectivity) throw std::runtime_error("tioga_amd: no recorded connectivity");
  const std::vector<int> &pm = partners(src);
  if (!std::binary_search(pm.begin(), pm.end(), dst)) return false;
  const int np = (int)pairs_.size();
  bool gids = !blocks_.empty();
  for (const Block *b : blocks_) gids = gids && b->unode_gid != nullptr;
  const int per_node = gids ? 3 : 1;  // node id, and its 64-bit global id as two ints

  // the heads of a block's donor lists (DONORLIST, linklist.c:62-104)
  auto cancellations = [&](int ib, const std::vector<int> &clist) {
    Block &B = *blocks_[ib];
    const std::vector<int> has = download(B.has_donor, B.nnodes, s_);
    const std::vector<int> hb = download(B.head_blk, B.nnodes, s_);
    const std::vector<int> hr = download(B.head_remote, B.nnodes, s_);
    // MeshBlock::getCancellationData (bookKeeping.C:542-563)
    for (int v : clist) {
      if (!has[v]) continue;
      const Block &H = *blocks_[hb[v]];
      if (H.rank != dst) continue;
      ints.push_back(hr[v]);
      ints.push_back(H.tag);
    }
  };

  if (ex == "boxes") {  // exchangeBoxes.C:184-243
    int n = 0;
    for (int i = 0; i < np; i++) {
      const Block &A = *blocks_[pairs_[i].local], &B = *blocks_[pairs_[i].ob];
      if (A.rank != src || B.rank != dst) continue;
      ints.insert(ints.end(), {pairs_[i].send_tag, A.lidx, A.tag});
      reals.insert(reals.end(), pk_.pair_rd.begin() + 6 * (size_t)i, pk_.pair_rd.begin() + 6 * (size_t)i + 6);
      n++;
    }
    // the reference sizes the reals at 18 per pair and fills the first six
    counts[0] = 3LL * n;
    counts[1] = 18LL * n;
    return true;
  }
  if (ex == "search") {  // exchangeSearchData.C:84-112, MeshBlock.C:1083-1110
    std::vector<std::vector<int>> isr(blocks_.size());
    std::vector<std::vector<double>> xsr(blocks_.size());
    for (int i = 0; i < np; i++) {
      const Block &A = *blocks_[pairs_[i].local], &B = *blocks_[pairs_[i].ob];
      if (A.rank != src || B.rank != dst) continue;
      const int c = pk_.pair_cnt[i];
      ints.insert(ints.end(), {pairs_[i].send_tag, c * per_node, 4 * c});
      if (c == 0) continue;
      const int ii = recv_pair(i);
      if (ii < 0) throw std::runtime_error("tioga_amd: a pair's query points have no receiver");
      const int rb = pairs_[ii].local;
      Block &R = *blocks_[rb];
      if (isr[rb].empty()) {
        isr[rb] = download(R.isearch, 3 * (size_t)R.nsearch, s_);
        xsr[rb] = download(R.xsearch, 3 * (size_t)R.nsearch, s_);
      }
      for (int q = pk_.pair_off[i]; q < pk_.pair_off[i] + c; q++) {
        const int node = isr[rb][3 * (size_t)q + 1];
        ints.push_back(node);
        if (gids) put_gid(ints, A.unode_gid[node]);
        reals.insert(reals.end(), {xsr[rb][3 * (size_t)q], xsr[rb][3 * (size_t)q + 1], xsr[rb][3 * (size_t)q + 2],
                                   pk_.sent_res[rb][q]});
      }
    }
  } else if (ex == "donors1") {  // MeshBlock::getMBDonorPackets (bookKeeping.C:89-110)
    for (size_t ib = 0; ib < blocks_.size(); ib++) {
      Block &D = *blocks_[ib];
      if (D.rank != src || D.nsearch == 0) continue;
      const size_t Q = D.nsearch;
      const std::vector<int> did = download(D.donorId, Q, s_), isr = download(D.isearch, 3 * Q, s_),
                             xt = download(D.xtag, Q, s_);
      const std::vector<double> rs = download(D.res_search, Q, s_), cr = download(D.cellRes, D.ncells, s_);
      for (size_t i = 0; i < Q; i++) {
        if (did[i] < 0) continue;
        const Block &S = *blocks_[isr[3 * i + 2]];
        if (S.rank != dst) continue;
        ints.insert(ints.end(), {D.tag, isr[3 * i + 1], (int)i, S.lidx});
        reals.push_back(cr[did[i]]);
        reals.push_back(rs[xt[i]]);
      }
    }
  } else if (ex == "donors2") {  // MeshBlock::processDonors records (bookKeeping.C:309-340), exchangeDonors.C:124-145
    for (size_t j = 0; j < pk_.rec_blk.size(); j++) {
      const Block &Dn = *blocks_[pk_.rec_dblk[j]];
      if (blocks_[pk_.rec_blk[j]]->rank != src || Dn.rank != dst) continue;
      ints.push_back(pk_.rec_remote[j]);
      ints.push_back(Dn.tag);
      reals.push_back(pk_.rec_rres[j]);
    }
  } else if (ex == "donors3") {
    // cancelList as findInterpData appends to it (bookKeeping.C:454-494): per
    // mandatory record in list order, each vertex of the donor cell whose
    // iblank is <= 0 at that moment and whose nodeRes is not BIGVALUE, a
    // negative iblank reset to 1 on the way
    for (size_t ib = 0; ib < blocks_.size(); ib++) {
      Block &B = *blocks_[ib];
      if (B.rank != src) continue;
      std::vector<int> cur = pk_.iblank_pd[ib], clist;
      const std::vector<double> &nres = pk_.nres_pd[ib];
      const std::vector<int> did = download(B.donorId, B.nsearch, s_);
      for (size_t j = 0; j < pk_.rec_blk.size(); j++) {
        if (pk_.rec_dblk[j] != (int)ib) continue;
        if (!(pk_.rec_rres[j] == BIGVALUE && B.resolutionScale == 1.0)) continue;
        const int c = did[pk_.rec_remote[j]];
        int t = 0;
        while (c >= B.ti.cstart[t + 1]) t++;
        const int nv = B.ti.nv[t];
        const int *cv = B.uconn[t] + (size_t)nv * (c - B.ti.cstart[t]);
        for (int m = 0; m < nv; m++) {
          const int v = cv[m] - BASE;
          if (cur[v] <= 0 && nres[v] != BIGVALUE) {
            if (cur[v] < 0) cur[v] = 1;
            clist.push_back(v);
          }
        }
      }
      cancellations((int)ib, clist);
    }
  } else if (ex == "rf.donors3") {  // MeshBlock::reduce_fringes' cancel list (bookKeeping.C:709-730)
    if (!pk_.reduced) throw std::runtime_error("tioga_amd: no recorded reduce_fringes");
    for (size_t ib = 0; ib < blocks_.size(); ib++)
      if (blocks_[ib]->rank == src) cancellations((int)ib, pk_.rf_cancel[ib]);
  } else if (ex == "donors4" || ex == "receptor_info" || ex == "update") {
    // MeshBlock::getInterpData (bookKeeping.C:598-613), getReceptorInfo
    // (tioga.C:572-650, dataUpdate.C:161-178), tioga::dataUpdate
    // (tioga.C:368-427): the live records of every donor block
    if (ex == "update" && !pk_.updated) throw std::runtime_error("tioga_amd: no recorded dataUpdate");
    for (size_t ib = 0; ib < blocks_.size(); ib++) {
      Block &D = *blocks_[ib];
      if (D.rank != src) continue;
      interp_records(D);
      size_t live = 0;
      for (const InterpHost &h : D.interp_host) {
        if (h.cancel) continue;
        const Block &R = *blocks_[h.rinfo[2]];
        const size_t l = live++;
        if (R.rank != dst) continue;
        ints.push_back(h.rinfo[1]);
        ints.push_back(R.lidx);
        if (ex == "receptor_info") {
          const int don = h.inode[h.nweights];
          put_gid(ints, D.ucell_gid ? D.ucell_gid[don] : (uint64_t)don);
        } else if (ex == "update") {
          const double *v = pk_.upd_vals[ib].data() + l * pk_.nvar;
          reals.insert(reals.end(), v, v + pk_.nvar);
        }
      }
    }
  } else {
    throw std::runtime_error("tioga_amd: unknown exchange " + ex);
  }
  counts[0] = (long long)ints.size();
  counts[1] = (long long)reals.size();
  return true;
}

}  // namespace tg
