// tg_obb.hip -- oriented bounding boxes of device-resident point sets with
// the bits of findOBB (utils.c:25-161).
//
// findOBB's centroid and covariance are running FP64 sums (utils.c:40-46,
// 91-99): s = fl(s + t_i) in index order, whose rounding fixes the result
// bits.  They are reproduced here without a sequential pass over the points:
//
//   Let |s| lie in the binade [2^k, 2^(k+1)) and q = 2^(k-52) be its spacing.
//   s is a multiple of q, so while every exact s + t_i stays inside the
//   binade, fl(s + t_i) = s + R(t_i), where R rounds to the nearest multiple
//   of q -- independent of s, unless t_i / q has fraction exactly 1/2 (then
//   the tie goes to the even neighbour of s + t_i, which depends on s).
//   So over a stretch of terms without such a tie, and whose partial sums of
//   R(t_i) keep s at least q inside the binade, the running sum advances by
//   the exact integer sum of R(t_i) / q in units of q.  Sufficient: with P and
//   N the sums of the positive and of the negative R(t_i) / q and A = |s| / q
//   (an integer in [2^52, 2^53)), A + N >= 2^52 + 1 and A + P <= 2^53 - 1.
//   Negative s is the mirror image (P and N swap and change sign).
//
// The points are cut into segments of OSEG.  Per segment and chain, all in
// parallel: an approximate sum (any order); a per-set prefix of those gives
// the predicted binade of s at each segment start; then P, N and a tie flag
// for the predicted binade and its two neighbours.  One wave per chain then
// walks the segments with the exact s, 64 at a time (a prefix scan of their
// increments), taking each segment in one step when its binade is one of
// the three and the test above passes, and otherwise adding the segment's
// terms one by one in index order (the reference's own sequence of
// operations).  The result is the sequential sum's bits in every
// case; only the time depends on how often a segment has to be stepped
// through (sums that keep returning to zero, ties of dyadic data).
//
// The covariance (utils.c:91-99) is the same with t_i the product of two
// centred coordinates, each rounded as the reference rounds it.  The box
// extents (utils.c:122-142) are min/max folds: TIOGA_MAX keeps the later of
// two equal values, which matters only for -0.0 versus +0.0, so each fold
// carries the index of the element it holds and ties go to the larger index
// -- an associative, commutative merge that any reduction order may use.
// The eigenvectors (kaiser.f) are computed on the host from the nine sums.
#include <stdexcept>
#include <vector>

#include "tg_engine.h"
#include "tg_seq.h"

namespace tg {

namespace {

constexpr int OSEG = 512;          // points per segment
constexpr int OW = 64;             // lanes of a segment's wave
constexpr int OPL = OSEG / OW;     // points per lane
constexpr int OWPB = 4;            // segment waves per workgroup
constexpr int EXT_PER = 8;         // points per thread of the extents pass
constexpr int EXT_BT = 256;
constexpr int EXT_CHUNK = EXT_PER * EXT_BT;
constexpr double MAGIC = 6755399441055744.0;  // 1.5 * 2^52: (y + MAGIC) - MAGIC rounds y to an integer
constexpr double LIM = 1125899906842624.0;    // 2^50: bound on |t / q| for that rounding
constexpr double LOW = 4503599627370497.0;    // 2^52 + 1
constexpr double HIGH = 9007199254740991.0;   // 2^53 - 1

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ __forceinline__ int bexp(double s)
{
  return (int)((__double_as_longlong(s) >> 52) & 0x7ff);
}

// 1 / q for a binade of biased exponent E (52 <= E <= 2046): 2^(1075 - E)
__device__ __forceinline__ double inv_quantum(int E)
{
  return __longlong_as_double((long long)(2098 - E) << 52);
}

__device__ __forceinline__ double rl(double v, int j)
{
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), j);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), j);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// a double moved across lanes by a DPP pattern (lanes the pattern does not
// write read 0.0)
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_mov(double v)
{
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffLL), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// inclusive and exclusive prefix sums over the 64 lanes: row shifts by 1,
// 2, 4, 8, then each row's total broadcast into the rows above (every
// partial sum covers a contiguous run of lanes ending at the lane itself)
__device__ __forceinline__ void wave_scan(double v, double &incl, double &excl)
{
  double x = v;
  x += dpp_mov<0x111>(x);        // row_shr:1
  x += dpp_mov<0x112>(x);        // row_shr:2
  x += dpp_mov<0x114>(x);        // row_shr:4
  x += dpp_mov<0x118>(x);        // row_shr:8
  x += dpp_mov<0x142, 0xa>(x);   // row_bcast:15 into rows 1 and 3
  x += dpp_mov<0x143, 0xc>(x);   // row_bcast:31 into rows 2 and 3
  incl = x;
  excl = dpp_mov<0x138>(x);      // wave_shr:1 (lane 0 reads 0)
}

__device__ __forceinline__ double wave_sum(double v)
{
  for (int o = OW / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, OW);
  return v;
}

// chain terms: C == 3, the coordinates (centroid); C == 6, the products of
// centred coordinates in the order aa[0], aa[4], aa[8], aa[3], aa[6], aa[7]
// (components picked by selects, so a run-time c indexes no local array)
__device__ __forceinline__ double pick(int j, const double *v) { return j == 0 ? v[0] : (j == 1 ? v[1] : v[2]); }

template <int C>
__device__ __forceinline__ double term(int c, const double *p, const double *xc)
{
  if (C == 3) return pick(c, p);
  const int pa = (c < 3) ? c : (c == 5 ? 1 : 0);
  const int pb = (c < 3) ? c : (c == 3 ? 1 : 2);
  return (pick(pa, p) - pick(pa, xc)) * (pick(pb, p) - pick(pb, xc));
}

struct SetRef {
  int set, seg;  // set index, segment index inside the set
};

__device__ __forceinline__ SetRef find_set(int g, int ns, const ObbSetDev *__restrict__ S)
{
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S[mid].seg0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return SetRef{lo, g - S[lo].seg0};
}

// approximate per-segment sums of every chain
template <int C>
__global__ void __launch_bounds__(OW *OWPB) k_obb_apx(int ns, int G, const ObbSetDev *__restrict__ S,
                                                     const double *__restrict__ xcs, double *__restrict__ apx)
{
  const int g = blockIdx.x * OWPB + (int)(threadIdx.x / OW);
  const int lane = threadIdx.x % OW;
  if (g >= G) return;
  const SetRef r = find_set(g, ns, S);
  const ObbSetDev d = S[r.set];
  double xc[3] = {0, 0, 0};
  if (C == 6)
    for (int j = 0; j < 3; j++) xc[j] = xcs[3 * r.set + j];
  double a[C];
  for (int c = 0; c < C; c++) a[c] = 0;
  const long long i0 = (long long)r.seg * OSEG;
  for (int k = 0; k < OPL; k++) {
    const long long i = i0 + lane + (long long)k * OW;
    if (i < d.n) {
      const double *p = d.x + 3 * i;
      const double pv[3] = {p[0], p[1], p[2]};
      for (int c = 0; c < C; c++) a[c] += term<C>(c, pv, xc);
    }
  }
  for (int c = 0; c < C; c++) {
    const double t = wave_sum(a[c]);
    if (lane == 0) apx[(size_t)c * G + g] = t;
  }
}

// per set and chain: exclusive prefix of the approximate sums -> predicted
// binade (biased exponent) of the running sum at each segment start
template <int C>
__global__ void __launch_bounds__(256) k_obb_pred(int ns, int G, const ObbSetDev *__restrict__ S,
                                                  const double *__restrict__ apx, int *__restrict__ ep)
{
  __shared__ double sh[256];
  const int set = blockIdx.x / C, c = blockIdx.x % C;
  const int g0 = S[set].seg0, g1 = S[set].seg1;
  const int t = threadIdx.x;
  double carry = 0;
  for (int gb = g0; gb < g1; gb += 256) {
    const int g = gb + t;
    const double v = (g < g1) ? apx[(size_t)c * G + g] : 0.0;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
      const double w = (t >= o) ? sh[t - o] : 0.0;
      __syncthreads();
      sh[t] += w;
      __syncthreads();
    }
    const double incl = sh[t];
    if (g < g1) ep[(size_t)c * G + g] = bexp(carry + (incl - v));
    carry += sh[255];
    __syncthreads();
  }
}

// per segment and chain: sums of the positive and negative quantised terms
// and the tie / range flag, for the predicted binade and its neighbours
template <int C>
__global__ void __launch_bounds__(OW *OWPB) k_obb_quant(int ns, int G, const ObbSetDev *__restrict__ S,
                                                       const double *__restrict__ xcs, const int *__restrict__ ep,
                                                       double *__restrict__ qP, double *__restrict__ qN,
                                                       int *__restrict__ qbad)
{
  const int g = blockIdx.x * OWPB + (int)(threadIdx.x / OW);
  const int lane = threadIdx.x % OW;
  if (g >= G) return;
  const SetRef r = find_set(g, ns, S);
  const ObbSetDev d = S[r.set];
  double xc[3] = {0, 0, 0};
  if (C == 6)
    for (int j = 0; j < 3; j++) xc[j] = xcs[3 * r.set + j];
  const long long i0 = (long long)r.seg * OSEG;
  double pv[OPL][3];
  bool in[OPL];
#pragma unroll
  for (int k = 0; k < OPL; k++) {
    const long long i = i0 + lane + (long long)k * OW;
    in[k] = i < d.n;
    const double *p = d.x + 3 * (in[k] ? i : 0);
    for (int j = 0; j < 3; j++) pv[k][j] = p[j];
  }
  for (int c = 0; c < C; c++) {
    const int e = ep[(size_t)c * G + g];
    int badbits = 0;
    for (int dd = 0; dd < 3; dd++) {
      const int E = e + dd - 1;
      double P = 0, N = 0;
      bool bad = false;
      if (E < 52 || E > 2046) {
        bad = true;
      } else {
        const double sc = inv_quantum(E);
#pragma unroll
        for (int k = 0; k < OPL; k++) {
          if (!in[k]) continue;
          const double y = term<C>(c, pv[k], xc) * sc;
          const double rr = (y + MAGIC) - MAGIC;
          bad = bad || !(fabs(y) < LIM) || fabs(y - rr) == 0.5;
          P += rr > 0 ? rr : 0.0;
          N += rr < 0 ? rr : 0.0;
        }
      }
      P = wave_sum(P);
      N = wave_sum(N);
      if (__ballot(bad) != 0) badbits |= 1 << dd;
      if (lane == 0) {
        qP[((size_t)c * G + g) * 3 + dd] = P;
        qN[((size_t)c * G + g) * 3 + dd] = N;
      }
    }
    if (lane == 0) qbad[(size_t)c * G + g] = badbits;
  }
}

// s advanced over one chunk of cnt <= 64 terms (lane l holds term l), with
// the sequential sum's bits: a scan as in the segment walk, each term's own
// quantisation standing for the segment sums
__device__ double chunk_walk(double s, double t, int cnt, int lane)
{
  int st = 0;
  bool scan = true;
  while (st < cnt) {
    if (!scan) {
      for (int l = st; l < cnt; l++) s = s + rl(t, l);
      break;
    }
    const int E = bexp(s);
    const bool sok = s != 0.0 && E >= 53 && E <= 2046;
    const bool neg = s < 0;
    const double sc = inv_quantum(sok ? E : 1075);
    const double A = fabs(s) * sc;
    const bool mine = lane >= st && lane < cnt;
    const double y = (neg ? -t : t) * sc;
    const double r = (y + MAGIC) - MAGIC;
    bool pass = sok && mine && fabs(y) < LIM && fabs(y - r) != 0.5;
    const double T = pass ? r : 0.0;
    double incl, ex;
    wave_scan(T, incl, ex);
    pass = pass && (A + ex + (r < 0 ? r : 0.0) >= LOW) && (A + ex + (r > 0 ? r : 0.0) <= HIGH);
    const unsigned long long fm = __ballot(mine && !pass);
    const int f = fm ? __ffsll((long long)fm) - 1 : cnt;
    if (f > st) {
      const double off = (f < cnt) ? rl(ex, f) : rl(incl, OW - 1);
      const double q = __longlong_as_double((long long)(E - 52) << 52);
      const double v = (A + off) * q;
      s = neg ? -v : v;
    } else {
      scan = false;
    }
    if (f < cnt) s = s + rl(t, f);
    st = f + 1;
  }
  return s;
}

// a segment's record: predicted binade, flags, quantised sums
struct SegRec {
  int ep, bad;
  double p0, p1, p2, n0, n1, n2;
};

__device__ __forceinline__ SegRec load_rec(size_t rec, bool have, const int *__restrict__ ep,
                                           const double *__restrict__ qP, const double *__restrict__ qN,
                                           const int *__restrict__ qbad)
{
  SegRec r;
  r.ep = ep[rec];
  r.bad = have ? qbad[rec] : 7;
  r.p0 = qP[rec * 3];
  r.p1 = qP[rec * 3 + 1];
  r.p2 = qP[rec * 3 + 2];
  r.n0 = qN[rec * 3];
  r.n1 = qN[rec * 3 + 1];
  r.n2 = qN[rec * 3 + 2];
  return r;
}

// one wave per set and chain: the exact running sum over the segments, 64
// segments at a time.  With s in binade E, lane k holds segment k's exact
// increment T_k (in units of q) for that binade; an exclusive prefix scan
// gives every segment's starting offset, so all 64 tests run at once, and s
// jumps to the start of the first segment that fails (or past the batch).
// That segment's terms are then added one by one and the scan restarts
// after it.  Up to the first failure every prefix keeps A + offset inside
// [2^52 + 1, 2^53 - 1], so all partial sums the scan forms there are
// integers below 2^53 in magnitude, hence exact.
template <int C>
__global__ void __launch_bounds__(OW) k_obb_walk(int ns, int G, const ObbSetDev *__restrict__ S,
                                                 const double *__restrict__ xcs, const int *__restrict__ ep,
                                                 const double *__restrict__ qP, const double *__restrict__ qN,
                                                 const int *__restrict__ qbad, double *__restrict__ out,
                                                 int *__restrict__ nstep)
{
  const int set = blockIdx.x / C, c = blockIdx.x % C;
  const int lane = threadIdx.x;
  const ObbSetDev d = S[set];
  double xc[3] = {0, 0, 0};
  if (C == 6)
    for (int j = 0; j < 3; j++) xc[j] = xcs[3 * set + j];
  double s = 0.0;
  int stepped = 0;
  auto rec_of = [&](int gb, bool &have) {
    const int g = gb + lane;
    have = g < d.seg1;
    return (size_t)c * G + (have ? g : d.seg0);
  };
  bool hv;
  size_t ri = rec_of(d.seg0, hv);
  SegRec cur = load_rec(ri, hv, ep, qP, qN, qbad);
  for (int gb = d.seg0; gb < d.seg1; gb += OW) {
    SegRec nxt = cur;
    if (gb + OW < d.seg1) {
      ri = rec_of(gb + OW, hv);
      nxt = load_rec(ri, hv, ep, qP, qN, qbad);
    }
    const int m = min(OW, d.seg1 - gb);
    int start = 0;
    while (start < m) {
      const int E = bexp(s);
      const bool sok = s != 0.0 && E >= 53 && E <= 2046;
      const bool neg = s < 0;
      const double sc = inv_quantum(sok ? E : 1075);
      const double A = fabs(s) * sc;
      const int dd = E - cur.ep + 1;
      const bool mine = lane >= start && lane < m;
      bool pass = sok && mine && dd >= 0 && dd <= 2 && !((cur.bad >> (dd & 3)) & 1);
      double P = dd == 0 ? cur.p0 : (dd == 1 ? cur.p1 : cur.p2);
      double N = dd == 0 ? cur.n0 : (dd == 1 ? cur.n1 : cur.n2);
      if (neg) {
        const double t = P;
        P = -N;
        N = -t;
      }
      const double T = pass ? P + N : 0.0;
      double incl, ex;
      wave_scan(T, incl, ex);
      pass = pass && (A + ex + N >= LOW) && (A + ex + P <= HIGH);
      const unsigned long long fm = __ballot(mine && !pass);
      const int f = fm ? __ffsll((long long)fm) - 1 : m;
      if (f > start) {
        // s at the start of segment f (or after the batch)
        const double off = (f < m) ? rl(ex, f) : rl(incl, OW - 1);
        const double q = __longlong_as_double((long long)(E - 52) << 52);
        const double v = (A + off) * q;
        s = neg ? -v : v;
      }
      if (f < m) {
        // segment f term by term (all its terms loaded first, so one memory
        // latency): per chunk of 64, the same scan with each term's own
        // rounding at the current binade, the first failing term (a tie, a
        // binade crossing) added by itself; a scan that makes no progress
        // hands the rest of its chunk to plain sequential adds
        stepped++;
        const long long i0 = (long long)(gb + f - d.seg0) * OSEG;
        const long long i1 = min((long long)d.n, i0 + OSEG);
        double tv[OPL];
#pragma unroll
        for (int k = 0; k < OPL; k++) {
          const long long i = i0 + (long long)k * OW + lane;
          tv[k] = 0;
          if (i < i1) {
            const double *p = d.x + 3 * i;
            const double pv[3] = {p[0], p[1], p[2]};
            tv[k] = term<C>(c, pv, xc);
          }
        }
#pragma unroll
        for (int k = 0; k < OPL; k++) {
          const long long cb = i0 + (long long)k * OW;
          if (cb < i1) s = chunk_walk(s, tv[k], (int)min((long long)OW, i1 - cb), lane);
        }
      }
      start = f + 1;
    }
    cur = nxt;
  }
  if (lane == 0) {
    if (C == 3) out[3 * set + c] = s / (double)d.n;  // utils.c:48-50
    else out[6 * set + c] = s;
    if (nstep) atomicAdd(nstep, stepped);
  }
}

// extents: per workgroup, the min and max of each frame coordinate with the
// index of the element held (ties: larger index)
struct Ext {
  double v;
  long long i;
};

__device__ __forceinline__ Ext ext_max(Ext a, Ext b)
{
  if (a.v > b.v) return a;
  if (b.v > a.v) return b;
  return a.i > b.i ? a : b;
}
__device__ __forceinline__ Ext ext_min(Ext a, Ext b)
{
  if (a.v < b.v) return a;
  if (b.v < a.v) return b;
  return a.i > b.i ? a : b;
}

__device__ __forceinline__ Ext shfl_ext(Ext e, int o)
{
  Ext r;
  r.v = __shfl_xor(e.v, o, OW);
  r.i = __shfl_xor(e.i, o, OW);
  return r;
}

__global__ void __launch_bounds__(EXT_BT) k_obb_ext(int ns, const ObbSetDev *__restrict__ S,
                                                   const double *__restrict__ xcs, const double *__restrict__ vecs,
                                                   Ext *__restrict__ part)
{
  __shared__ Ext sh[EXT_BT / OW][6];
  int lo = 0, hi = ns - 1;
  const int b = blockIdx.x;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const ObbSetDev d = S[lo];
  double xc[3], vec[3][3];
  for (int j = 0; j < 3; j++) xc[j] = xcs[3 * lo + j];
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) vec[j][k] = vecs[9 * lo + 3 * j + k];
  Ext mx[3], mn[3];
  for (int j = 0; j < 3; j++) {
    mx[j] = Ext{-BIGVALUE, -1};
    mn[j] = Ext{BIGVALUE, -1};
  }
  const long long base = (long long)(b - d.blk0) * EXT_CHUNK;
  for (int k = 0; k < EXT_PER; k++) {
    const long long i = base + threadIdx.x + (long long)k * EXT_BT;
    if (i < d.n) {
      const double *p = d.x + 3 * i;
      const double pv[3] = {p[0], p[1], p[2]};
      double e[3];
      to_frame(pv, xc, vec, e);  // utils.c:131-135
      for (int j = 0; j < 3; j++) {
        mx[j] = ext_max(mx[j], Ext{e[j], i});
        mn[j] = ext_min(mn[j], Ext{e[j], i});
      }
    }
  }
  for (int o = OW / 2; o > 0; o >>= 1)
    for (int j = 0; j < 3; j++) {
      mx[j] = ext_max(mx[j], shfl_ext(mx[j], o));
      mn[j] = ext_min(mn[j], shfl_ext(mn[j], o));
    }
  const int w = threadIdx.x / OW;
  if (threadIdx.x % OW == 0)
    for (int j = 0; j < 3; j++) {
      sh[w][j] = mn[j];
      sh[w][3 + j] = mx[j];
    }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    Ext r = sh[0][j];
    for (int q = 1; q < EXT_BT / OW; q++) r = (j < 3) ? ext_min(r, sh[q][j]) : ext_max(r, sh[q][j]);
    part[(size_t)b * 6 + j] = r;
  }
}

// per set: merge the workgroup extents -> xmin[3], xmax[3]
__global__ void __launch_bounds__(OW) k_obb_ext_merge(const ObbSetDev *__restrict__ S, const Ext *__restrict__ part,
                                                      double *__restrict__ lohi)
{
  const ObbSetDev d = S[blockIdx.x];
  const int nb = d.blk1 - d.blk0;
  Ext r[6];
  for (int j = 0; j < 3; j++) {
    r[j] = Ext{BIGVALUE, -1};
    r[3 + j] = Ext{-BIGVALUE, -1};
  }
  for (int q = threadIdx.x; q < nb; q += OW)
    for (int j = 0; j < 6; j++) {
      const Ext e = part[(size_t)(d.blk0 + q) * 6 + j];
      r[j] = (j < 3) ? ext_min(r[j], e) : ext_max(r[j], e);
    }
  for (int o = OW / 2; o > 0; o >>= 1)
    for (int j = 0; j < 6; j++) {
      const Ext e = shfl_ext(r[j], o);
      r[j] = (j < 3) ? ext_min(r[j], e) : ext_max(r[j], e);
    }
  if (threadIdx.x == 0)
    for (int j = 0; j < 6; j++) lohi[6 * blockIdx.x + j] = r[j].v;
}

template <int C>
void chain_sums(int ns, int G, ObbScratch &w, hipStream_t s, double *out)
{
  const ObbSetDev *S = w.sets.p;
  const unsigned gw = nblk(G, OWPB);
  k_obb_apx<C><<<gw, OW * OWPB, 0, s>>>(ns, G, S, w.xc.p, w.apx.p);
  k_obb_pred<C><<<ns * C, 256, 0, s>>>(ns, G, S, w.apx.p, w.ep.p);
  k_obb_quant<C><<<gw, OW * OWPB, 0, s>>>(ns, G, S, w.xc.p, w.ep.p, w.qP.p, w.qN.p, w.qbad.p);
  k_obb_walk<C><<<ns * C, OW, 0, s>>>(ns, G, S, w.xc.p, w.ep.p, w.qP.p, w.qN.p, w.qbad.p, out, w.nstep.p);
}

}  // namespace

void find_obbs_device(const ObbSet *sets, int nset, OBB *out, ObbScratch &w, hipStream_t s, long long *stepped)
{
  if (nset <= 0) return;
  // sets of fewer than 4 points take findOBB's special branch (utils.c:51-80)
  // on the host
  std::vector<int> big, small;
  for (int k = 0; k < nset; k++) {
    if (sets[k].n <= 0) throw std::runtime_error("tioga_amd: OBB of an empty point set");
    (sets[k].n < 4 ? small : big).push_back(k);
  }
  for (int k : small) {
    double x[9];
    TG_HIP(hipMemcpyAsync(x, sets[k].x, sizeof(double) * 3 * sets[k].n, hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    find_obb(x, nullptr, sets[k].n, out[k]);
  }
  if (stepped) *stepped = 0;
  if (big.empty()) return;
  const int ns = (int)big.size();
  std::vector<ObbSetDev> hs(ns);
  long long G = 0, B = 0;
  for (int a = 0; a < ns; a++) {
    const ObbSet &q = sets[big[a]];
    hs[a].x = q.x;
    hs[a].n = q.n;
    hs[a].seg0 = (int)G;
    G += (q.n + OSEG - 1) / OSEG;
    hs[a].seg1 = (int)G;
    hs[a].blk0 = (int)B;
    B += (q.n + EXT_CHUNK - 1) / EXT_CHUNK;
    hs[a].blk1 = (int)B;
  }
  if (G > (1LL << 30)) throw std::runtime_error("tioga_amd: too many points for one OBB batch");
  w.sets.reserve(ns);
  w.apx.reserve(6 * G);
  w.ep.reserve(6 * G);
  w.qP.reserve(18 * G);
  w.qN.reserve(18 * G);
  w.qbad.reserve(6 * G);
  w.xc.reserve(3 * (size_t)ns);
  w.aa.reserve(6 * (size_t)ns);
  w.vec.reserve(9 * (size_t)ns);
  w.lohi.reserve(6 * (size_t)ns);
  w.part.reserve(6 * 2 * B);  // Ext = two 8-byte words
  w.nstep.reserve(1);
  TG_HIP(hipMemcpyAsync(w.sets.p, hs.data(), sizeof(ObbSetDev) * ns, hipMemcpyHostToDevice, s));
  TG_HIP(hipMemsetAsync(w.nstep.p, 0, sizeof(int), s));
  chain_sums<3>(ns, (int)G, w, s, w.xc.p);
  chain_sums<6>(ns, (int)G, w, s, w.aa.p);
  std::vector<double> hxc(3 * (size_t)ns), haa(6 * (size_t)ns), hvec(9 * (size_t)ns), hlh(6 * (size_t)ns);
  int hstep = 0;
  TG_HIP(hipMemcpyAsync(hxc.data(), w.xc.p, sizeof(double) * hxc.size(), hipMemcpyDeviceToHost, s));
  TG_HIP(hipMemcpyAsync(haa.data(), w.aa.p, sizeof(double) * haa.size(), hipMemcpyDeviceToHost, s));
  TG_HIP(hipMemcpyAsync(&hstep, w.nstep.p, sizeof(int), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  if (stepped) *stepped = hstep;
  for (int a = 0; a < ns; a++) {
    OBB &o = out[big[a]];
    for (int j = 0; j < 3; j++) o.xc[j] = hxc[3 * a + j];
    // aa[0] aa[4] aa[8] aa[3] aa[6] aa[7] -> the symmetric matrix (utils.c:100-102)
    const double *c = &haa[6 * a];
    double m[9] = {c[0], c[3], c[4], c[3], c[1], c[5], c[4], c[5], c[2]};
    obb_axes(m, o);
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) hvec[9 * a + 3 * j + k] = o.vec[j][k];
  }
  TG_HIP(hipMemcpyAsync(w.vec.p, hvec.data(), sizeof(double) * hvec.size(), hipMemcpyHostToDevice, s));
  k_obb_ext<<<(unsigned)B, EXT_BT, 0, s>>>(ns, w.sets.p, w.xc.p, w.vec.p, (Ext *)w.part.p);
  k_obb_ext_merge<<<ns, OW, 0, s>>>(w.sets.p, (const Ext *)w.part.p, w.lohi.p);
  TG_HIP(hipMemcpyAsync(hlh.data(), w.lohi.p, sizeof(double) * hlh.size(), hipMemcpyDeviceToHost, s));
  TG_HIP(hipStreamSynchronize(s));
  for (int a = 0; a < ns; a++) obb_finish(out[big[a]], &hlh[6 * a], &hlh[6 * a + 3]);
}

}  // namespace tg
