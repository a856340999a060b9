// tg_update.hip -- tioga::dataUpdate (tioga.C:336-524; getInterpolatedSolution
// and updateSolnData, dataUpdate.C:27-119) on device-resident solutions.
//
// Every live interpolation record r of a donor block computes, per field k,
// qq[k] = sum over m of q[inode_m][k] * w_m in the reference's order
// (m = 0 .. nw-1, dataUpdate.C:48-101) and writes it to its receptor node
// in the receptor block's q (row layout q[node * nvar + k], column layout
// q[k * nnodes + node]).  In the reference every donor block computes its
// values before any receptor is written (the values travel as packets), so
// a receptor that is also a vertex of some live record's donor cell must be
// read before it is written.  prepare_update() checks, once per
// connectivity, whether any receptor is such a vertex:
//   * none (the usual case): one fused kernel reads the donors and writes the
//     receptors; no value of a donor vertex changes during the launch;
//   * some: a gather kernel writes every record's values to a staging
//     buffer, then a scatter kernel writes them to the receptors.
// All blocks' records are handled by one launch of each kernel; a thread's
// block comes from a table of per-block segments whose thread ranges start
// at multiples of the workgroup size (a wave-uniform binary search).
//
// Two entry points share the kernels:
//   * Engine::data_update (tioga_dataupdate_mb_) copies the registered host
//     arrays to device buffers that the engine keeps, updates them and
//     copies them back; the segment table and the staging buffer are the
//     engine's too, grown on demand and kept, so repeated calls allocate
//     nothing new;
//   * Engine::data_update_noalloc (tioga_amd_dataupdate_mb_noalloc) updates
//     the caller's device arrays in place, with the segment table and the
//     staged values in `scratch`, a file-scope device array of 64 MiB; it
//     allocates no device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "tg_engine.h"

// The staging array of the allocation-free dataUpdate (see below).  In host
// code its name is only a handle for hipGetSymbolAddress; kernels get the
// device address that call returns.
static __device__ __align__(256) unsigned char scratch[size_t(64) << 20];

namespace tg {

namespace {

constexpr int UBS = 256;
inline unsigned unb(long long n) { return (unsigned)((n + UBS - 1) / UBS); }

// The staging area of the allocation-free dataUpdate: the segment table,
// then the staged values of the two-kernel form.  Each GPU the library runs
// on holds its own copy until the process ends.  Calls that overlap on one
// GPU share it and race; the table is written at the start of every call
// and the values before they are read, so nothing needs clearing.
constexpr size_t UPDATE_SCRATCH_BYTES = sizeof(::scratch);
static_assert(UPDATE_SCRATCH_BYTES == size_t(64) << 20, "dataUpdate staging is 64 MiB");

constexpr int MAXDEV = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAXDEV] = {};

// the device address of this device's `scratch` (looked up once per device,
// with that device current)
void *scratch_address()
{
  int dev = 0;
  TG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) throw std::runtime_error("tioga_amd: device index out of range for dataUpdate staging");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    TG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    if (!p) throw std::runtime_error("tioga_amd: no device address for the dataUpdate staging array");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

// the segment holding thread g (segments start at multiples of the
// workgroup size, so it is the same for a whole workgroup)
__device__ __forceinline__ int seg_of(const UpdTable *__restrict__ T, int g)
{
  int lo = 0, hi = T->nseg - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (T->seg[m].start <= g) lo = m;
    else hi = m - 1;
  }
  return lo;
}

// MODE 0: fused (gather and write the receptor); 1: gather into vals;
// 2: scatter vals to the receptors
template <int MODE>
__global__ void __launch_bounds__(UBS) k_update(const UpdTable *__restrict__ T, int nvar, int interptype,
                                                double *__restrict__ vals)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = __builtin_amdgcn_readfirstlane(seg_of(T, g));
  const UpdSeg &S = T->seg[sg];
  const int r = g - S.start;
  if (r >= S.n || S.cancel[r]) return;
  const size_t vbase = (size_t)(S.voff + r) * nvar;
  double *qr = T->q[S.rblk[r]];
  const int rnn = T->nnodes[S.rblk[r]], nd = S.rnode[r];
  if (MODE == 2) {
    for (int k = 0; k < nvar; k++) {
      const double v = vals[vbase + k];
      if (interptype == 0) qr[(size_t)nd * nvar + k] = v;
      else qr[(size_t)k * rnn + nd] = v;
    }
    return;
  }
  const double *qd = T->q[S.blk];
  const int dnn = T->nnodes[S.blk], nw = S.nw[r];
  int v[8];
  double wt[8];
#pragma unroll
  for (int m = 0; m < 8; m++)
    if (m < nw) {
      v[m] = S.inode[(size_t)m * S.n + r];
      wt[m] = S.w[(size_t)m * S.n + r];
    }
  // fields in chunks of 8, each with its own running sum over the weights in
  // order m = 0 .. nw-1 (the reference's order per field), so a donor row's
  // fields are loaded together
  for (int k0 = 0; k0 < nvar; k0 += 8) {
    const int kn = nvar - k0 < 8 ? nvar - k0 : 8;
    double acc[8];
#pragma unroll
    for (int kk = 0; kk < 8; kk++) acc[kk] = 0;
#pragma unroll
    for (int m = 0; m < 8; m++) {
      if (m >= nw) break;
#pragma unroll
      for (int kk = 0; kk < 8; kk++)
        if (kk < kn) {
          const int k = k0 + kk;
          const double qv = interptype == 0 ? qd[(size_t)v[m] * nvar + k] : qd[(size_t)k * dnn + v[m]];
          acc[kk] += qv * wt[m];
        }
    }
#pragma unroll
    for (int kk = 0; kk < 8; kk++)
      if (kk < kn) {
        const int k = k0 + kk;
        if (MODE == 1) vals[vbase + k] = acc[kk];
        else if (interptype == 0) qr[(size_t)nd * nvar + k] = acc[kk];
        else qr[(size_t)k * rnn + nd] = acc[kk];
      }
  }
}

// the vertices of every live record's donor cell (mark[b][node] = 1) ...
__global__ void k_mark_donors(const UpdTable *__restrict__ T, unsigned char *const *__restrict__ mark)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = __builtin_amdgcn_readfirstlane(seg_of(T, g));
  const UpdSeg &S = T->seg[sg];
  const int r = g - S.start;
  if (r >= S.n || S.cancel[r]) return;
  const int nw = S.nw[r];
  for (int m = 0; m < nw && m < 8; m++) mark[S.blk][S.inode[(size_t)m * S.n + r]] = 1;
}

// ... and whether a live record's receptor is one of them
__global__ void k_receptor_is_donor(const UpdTable *__restrict__ T, unsigned char *const *__restrict__ mark,
                                    int *__restrict__ flag)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int sg = __builtin_amdgcn_readfirstlane(seg_of(T, g));
  const UpdSeg &S = T->seg[sg];
  const int r = g - S.start;
  if (r >= S.n || S.cancel[r]) return;
  if (mark[S.rblk[r]][S.rnode[r]]) *flag = 1;
}

__global__ void k_zero_bytes(size_t n, unsigned char *p)
{
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_zero_int(int *p) { *p = 0; }

void launch_update(const UpdTable *dtab, long long nthreads, bool staged, int nvar, int interptype, double *vals,
                   hipStream_t s)
{
  if (nthreads <= 0) return;
  if (!staged) {
    k_update<0><<<unb(nthreads), UBS, 0, s>>>(dtab, nvar, interptype, nullptr);
  } else {
    k_update<1><<<unb(nthreads), UBS, 0, s>>>(dtab, nvar, interptype, vals);
    k_update<2><<<unb(nthreads), UBS, 0, s>>>(dtab, nvar, interptype, vals);
  }
}

int parse_itype(const char *itype)
{
  // tiogaInterface.C:178-190
  if (itype && strstr(itype, "row")) return 0;
  if (itype && strstr(itype, "column")) return 1;
  return -1;
}

}  // namespace

// The segment table of the current interpolation records (the q pointers
// are filled per call) and whether the two-kernel form is needed.
void Engine::prepare_update()
{
  upd_.valid = false;
  const int nbk = (int)blocks_.size();
  if (nbk > UPD_MAXBLK) throw std::runtime_error("tioga_amd: too many blocks for dataUpdate");
  UpdTable &T = upd_.host;
  memset(&T, 0, sizeof T);
  long long start = 0;
  int voff = 0, nseg = 0;
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    T.nnodes[ib] = b.nnodes;
    if (b.ninterp == 0) continue;
    UpdSeg &S = T.seg[nseg++];
    S.inode = b.interp_inode.p;
    S.w = b.interp_w.p;
    S.nw = b.interp_nw.p;
    S.cancel = b.interp_cancel.p;
    S.rnode = b.interp_rnode.p;
    S.rblk = b.interp_rblk.p;
    S.n = b.ninterp;
    S.blk = ib;
    S.start = (int)start;
    S.voff = voff;
    start += (b.ninterp + UBS - 1) / UBS * UBS;
    voff += b.ninterp;
    if (start > 0x7fffffffLL) throw std::runtime_error("tioga_amd: too many interpolation records");
  }
  T.nseg = nseg;
  upd_.nthreads = start;
  upd_.nrec = voff;
  upd_.staged = false;
  if (nseg > 0) {
    // receptors that are donor vertices of a live record: mark every block's
    // donor vertices, then test the receptors
    std::vector<unsigned char *> mk(nbk, nullptr);
    for (int ib = 0; ib < nbk; ib++) {
      Block &b = *blocks_[ib];
      b.upd_mark.reserve(std::max(b.nnodes, 1));
      mk[ib] = b.upd_mark.p;
      if (b.nnodes) k_zero_bytes<<<unb(b.nnodes), UBS, 0, s_>>>((size_t)b.nnodes, b.upd_mark.p);
    }
    upd_.tab.reserve(sizeof(UpdTable));
    upd_.marks.reserve(nbk);
    upd_.flag.reserve(1);
    TG_HIP(hipMemcpyAsync(upd_.tab.p, &T, sizeof T, hipMemcpyHostToDevice, s_));
    TG_HIP(hipMemcpyAsync(upd_.marks.p, mk.data(), sizeof(unsigned char *) * nbk, hipMemcpyHostToDevice, s_));
    k_zero_int<<<1, 1, 0, s_>>>(upd_.flag.p);
    const UpdTable *dt = (const UpdTable *)upd_.tab.p;
    k_mark_donors<<<unb(start), UBS, 0, s_>>>(dt, upd_.marks.p);
    k_receptor_is_donor<<<unb(start), UBS, 0, s_>>>(dt, upd_.marks.p, upd_.flag.p);
    int f = 0;
    TG_HIP(hipMemcpyAsync(&f, upd_.flag.p, sizeof(int), hipMemcpyDeviceToHost, s_));
    sync();
    upd_.staged = f != 0;
  }
  upd_.valid = true;
}

// tioga::dataUpdate with the registered host arrays (tiogaInterface.C:175-216)
void Engine::data_update(int nvar, int interptype)
{
  for (Block *b : blocks_)
    if (b->uq == nullptr) return;
  if (nsend_ == 0 || nvar <= 0) return;
  if (!upd_.valid) prepare_update();
  const int nbk = (int)blocks_.size();
  if (pk_.on) record_update(nvar, interptype);
  while ((int)qdev_.size() < nbk) qdev_.push_back(new DBuf<double>);
  UpdTable &T = upd_.host;
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    size_t n = (size_t)b.nnodes * nvar;
    qdev_[ib]->reserve(n);
    TG_HIP(hipMemcpyAsync(qdev_[ib]->p, b.uq, sizeof(double) * n, hipMemcpyHostToDevice, s_));
    T.q[ib] = qdev_[ib]->p;
  }
  upd_.tab.reserve(sizeof(UpdTable));
  TG_HIP(hipMemcpyAsync(upd_.tab.p, &T, sizeof T, hipMemcpyHostToDevice, s_));
  if (upd_.staged) upd_.vals.reserve((size_t)upd_.nrec * nvar);
  launch_update((const UpdTable *)upd_.tab.p, upd_.nthreads, upd_.staged, nvar, interptype, upd_.vals.p, s_);
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    TG_HIP(hipMemcpyAsync(b.uq, qdev_[ib]->p, sizeof(double) * (size_t)b.nnodes * nvar, hipMemcpyDeviceToHost,
                          s_));
  }
  sync();
}

// Device ms per update of nrep updates on device-resident copies of the
// registered solutions.
double Engine::time_data_update(int nvar, int interptype, int nrep, long long *nrecv)
{
  const int nbk = (int)blocks_.size();
  long long nr = 0;
  for (Block *b : blocks_) nr += b->ninterp;
  if (nrecv) *nrecv = nr;
  for (Block *b : blocks_)
    if (b->uq == nullptr) return 0.0;
  if (nsend_ == 0 || nvar <= 0 || nrep <= 0) return 0.0;
  if (!upd_.valid) prepare_update();
  while ((int)qdev_.size() < nbk) qdev_.push_back(new DBuf<double>);
  UpdTable &T = upd_.host;
  for (int ib = 0; ib < nbk; ib++) {
    Block &b = *blocks_[ib];
    size_t n = (size_t)b.nnodes * nvar;
    qdev_[ib]->reserve(n);
    TG_HIP(hipMemcpyAsync(qdev_[ib]->p, b.uq, sizeof(double) * n, hipMemcpyHostToDevice, s_));
    T.q[ib] = qdev_[ib]->p;
  }
  upd_.tab.reserve(sizeof(UpdTable));
  TG_HIP(hipMemcpyAsync(upd_.tab.p, &T, sizeof T, hipMemcpyHostToDevice, s_));
  if (upd_.staged) upd_.vals.reserve((size_t)upd_.nrec * nvar);
  sync();
  hipEvent_t a, b;
  TG_HIP(hipEventCreate(&a));
  TG_HIP(hipEventCreate(&b));
  TG_HIP(hipEventRecord(a, s_));
  for (int rep = 0; rep < nrep; rep++)
    launch_update((const UpdTable *)upd_.tab.p, upd_.nthreads, upd_.staged, nvar, interptype, upd_.vals.p, s_);
  TG_HIP(hipEventRecord(b, s_));
  TG_HIP(hipEventSynchronize(b));
  float ms = 0;
  TG_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / nrep;
}

// tioga::dataUpdate on the caller's device arrays qdev[ib] (block ib in
// registration order), with every staging buffer in `scratch`: no device
// memory is allocated.  Synchronous, like the host-array form.
void Engine::data_update_noalloc(int nvar, int interptype, double *const *qdev)
{
  if (!connected_ || !upd_.valid) throw std::runtime_error("tioga_amd: dataUpdate needs performconnectivity first");
  if (nvar <= 0) throw std::runtime_error("tioga_amd: dataUpdate needs nvar >= 1");
  if (interptype != 0 && interptype != 1) throw std::runtime_error("tioga_amd: dataUpdate: unknown interpolation type");
  const int nbk = (int)blocks_.size();
  if (nsend_ == 0 || upd_.host.nseg == 0) return;
  for (int ib = 0; ib < nbk; ib++)
    if (!qdev || !qdev[ib]) throw std::runtime_error("tioga_amd: dataUpdate needs a device solution for every block");
  const size_t tab_bytes = (sizeof(UpdTable) + 255) / 256 * 256;
  const size_t need = tab_bytes + (upd_.staged ? sizeof(double) * (size_t)upd_.nrec * nvar : 0);
  if (need > UPDATE_SCRATCH_BYTES)
    throw std::runtime_error("tioga_amd: dataUpdate without allocation needs " + std::to_string(need) +
                             " bytes of staging, more than its " + std::to_string(UPDATE_SCRATCH_BYTES));
  unsigned char *base = (unsigned char *)scratch_address();
  UpdTable T = upd_.host;
  for (int ib = 0; ib < nbk; ib++) T.q[ib] = qdev[ib];
  TG_HIP(hipMemcpyAsync(base, &T, sizeof T, hipMemcpyHostToDevice, s_));
  launch_update((const UpdTable *)base, upd_.nthreads, upd_.staged, nvar, interptype, (double *)(base + tab_bytes),
                s_);
  sync();
}

size_t update_scratch_bytes() { return UPDATE_SCRATCH_BYTES; }

int update_interptype(const char *itype) { return parse_itype(itype); }

}  // namespace tg
