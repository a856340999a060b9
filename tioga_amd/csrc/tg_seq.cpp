// tg_seq.cpp -- host-side sequential stages (see tg_seq.h).
#include "tg_seq.h"

#include <algorithm>
#include <math.h>
#include <string.h>

#include <utility>
#include <vector>

namespace tg {

namespace {

// Jacobi-type eigen-solver of kaiser.f:29-180 for a symmetric 3x3 matrix in
// Fortran (column-major) storage; on return the columns of a are the unit
// eigenvectors.  Same sweep order, thresholds and rotation formulas.
void jacobi3(double *a)
{
  auto A = [a](int i, int j) -> double & { return a[(j - 1) * 3 + (i - 1)]; };
  const int n = 3, npairs = 3;
  double ss = 0.0;
  for (int j = 1; j <= n; j++)
    for (int i = 1; i <= n; i++) ss = ss + A(i, j) * A(i, j);
  const double eps = 1.0e-12 * ss / n;
  int remaining = npairs;
  bool converged = false;
  for (int sweep = 0; sweep < 10 && !converged; sweep++) {
    remaining = npairs;
    for (int j = 1; j <= n - 1 && !converged; j++) {
      for (int k = j + 1; k <= n; k++) {
        double halfp = 0.0, q = 0.0;
        for (int i = 1; i <= n; i++) {
          double xj = A(i, j), xk = A(i, k);
          halfp = halfp + xj * xk;
          q = q + (xj + xk) * (xj - xk);
        }
        double p = halfp + halfp;
        double absp = fabs(p);
        if (absp < eps && q >= 0.0) {
          remaining = remaining - 1;
          if (remaining <= 0) {
            converged = true;
            break;
          }
          continue;
        }
        double absq = fabs(q), cs, sn;
        if (absp <= absq) {
          double tn = absp / absq;
          cs = 1.0 / sqrt(1.0 + tn * tn);
          sn = tn * cs;
        } else {
          double ctn = absq / absp;
          sn = 1.0 / sqrt(1.0 + ctn * ctn);
          cs = ctn * sn;
        }
        cs = sqrt((1.0 + cs) * 0.5);
        sn = sn / (cs + cs);
        if (q < 0.0) {
          double t = cs;
          cs = sn;
          sn = t;
        }
        if (p < 0.0) sn = -sn;
        for (int i = 1; i <= n; i++) {
          double t = A(i, j);
          A(i, j) = t * cs + A(i, k) * sn;
          A(i, k) = -t * sn + A(i, k) * cs;
        }
      }
    }
  }
  for (int j = 1; j <= n; j++) {
    double t = 0.0;
    for (int i = 1; i <= n; i++) t = t + A(i, j) * A(i, j);
    double e = sqrt(t);
    double s = e > 0.0 ? 1.0 / e : 0.0;
    for (int i = 1; i <= n; i++) A(i, j) = A(i, j) * s;
  }
}

}  // namespace

void obb_axes(double *aa, OBB &o)
{
  jacobi3(aa);
  for (int i = 0, m = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o.vec[i][j] = aa[m++];
}

void obb_finish(OBB &o, const double xmin[3], const double xmax[3])
{
  double xd[3];
  for (int j = 0; j < 3; j++) {
    o.dxc[j] = (xmax[j] - xmin[j]) * 0.5 * 1.01;
    xd[j] = (xmax[j] + xmin[j]) * 0.5;
  }
  for (int j = 0; j < 3; j++)
    for (int k = 0; k < 3; k++) o.xc[j] += (xd[k] * o.vec[k][j]);
}

void find_obb(const double *x, const int *idx, int n, OBB &o)
{
  auto P = [x, idx](int i) { return x + 3 * (size_t)(idx ? idx[i] : i); };
  // running sums in locals (o may alias nothing, but the compiler cannot
  // know), added in the reference's order
  double s0 = 0, s1 = 0, s2 = 0;
  if (idx) {
    for (int i = 0; i < n; i++) {
      const double *p = P(i);
      s0 += p[0];
      s1 += p[1];
      s2 += p[2];
    }
  } else {
    for (int i = 0; i < n; i++) {
      s0 += x[3 * (size_t)i];
      s1 += x[3 * (size_t)i + 1];
      s2 += x[3 * (size_t)i + 2];
    }
  }
  o.xc[0] = s0 / n;
  o.xc[1] = s1 / n;
  o.xc[2] = s2 / n;
  if (n < 4) {
    // utils.c:51-80 (the 2-point branch halves only the else-value, a
    // consequence of the unparenthesised TIOGA_MAX macro)
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) o.vec[i][j] = (i == j) ? 1 : 0;
    if (n == 1) {
      o.dxc[0] = o.dxc[1] = o.dxc[2] = 1e-3;
    } else if (n == 2) {
      for (int j = 0; j < 3; j++) {
        double d = fabs(P(1)[j] - P(0)[j]);
        o.dxc[j] = (1e-3 > d) ? 1e-3 : d * 0.5;
      }
    } else {
      // utils.c:71-77 measures every coordinate from the first point's x
      for (int i = 0; i < n; i++)
        for (int j = 0; j < 3; j++) o.dxc[j] = dmax(1e-3, fabs(P(i)[j] - P(0)[0]));
    }
    return;
  }
  double aa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  {
    const double c0 = o.xc[0], c1 = o.xc[1], c2 = o.xc[2];
    double a0 = 0, a4 = 0, a8 = 0, a3 = 0, a6 = 0, a7 = 0;
    for (int i = 0; i < n; i++) {
      const double *p = P(i);
      const double d0 = p[0] - c0, d1 = p[1] - c1, d2 = p[2] - c2;
      a0 += (d0 * d0);
      a4 += (d1 * d1);
      a8 += (d2 * d2);
      a3 += (d0 * d1);
      a6 += (d0 * d2);
      a7 += (d1 * d2);
    }
    aa[0] = a0;
    aa[4] = a4;
    aa[8] = a8;
    aa[3] = a3;
    aa[6] = a6;
    aa[7] = a7;
  }
  aa[1] = aa[3];
  aa[2] = aa[6];
  aa[5] = aa[7];
  obb_axes(aa, o);
  double xmin[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, xmax[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
  {
    const double xc[3] = {o.xc[0], o.xc[1], o.xc[2]};
    double vec[3][3];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) vec[j][k] = o.vec[j][k];
    double lo0 = xmin[0], lo1 = xmin[1], lo2 = xmin[2], hi0 = xmax[0], hi1 = xmax[1], hi2 = xmax[2];
    for (int i = 0; i < n; i++) {
      double e[3];
      to_frame(P(i), xc, vec, e);
      hi0 = dmax(hi0, e[0]);
      lo0 = dmin(lo0, e[0]);
      hi1 = dmax(hi1, e[1]);
      lo1 = dmin(lo1, e[1]);
      hi2 = dmax(hi2, e[2]);
      lo2 = dmin(lo2, e[2]);
    }
    xmin[0] = lo0;
    xmin[1] = lo1;
    xmin[2] = lo2;
    xmax[0] = hi0;
    xmax[1] = hi1;
    xmax[2] = hi2;
  }
  obb_finish(o, xmin, xmax);
}

int obb_intersect(const OBB &A, const OBB &B)
{
  const double eps = 1e-12;
  double D[3], c[3][3];
  for (int i = 0; i < 3; i++) D[i] = B.xc[i] - A.xc[i];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      c[i][j] = 0;
      for (int k = 0; k < 3; k++) c[i][j] = c[i][j] + A.vec[i][k] * B.vec[j][k];
    }
  for (int i = 0; i < 3; i++) {  // axes of A
    double r0 = A.dxc[i], r1 = 0, r = 0;
    for (int j = 0; j < 3; j++) {
      r1 += B.dxc[j] * fabs(c[i][j]);
      r += fabs(A.vec[i][j]) * D[j];
    }
    if (r > (r0 + r1 + eps)) return 0;
  }
  for (int i = 0; i < 3; i++) {  // axes of B
    double r1 = B.dxc[i], r0 = 0, r = 0;
    for (int j = 0; j < 3; j++) {
      r0 += A.dxc[j] * fabs(c[j][i]);
      r += fabs(B.vec[i][j]) * D[j];
    }
    if (r > (r0 + r1 + eps)) return 0;
  }
  for (int i = 0; i < 3; i++) {  // edge-edge cross products
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; j++) {
      int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      double r0 = A.dxc[i1] * fabs(c[i2][j]) + A.dxc[i2] * fabs(c[i1][j]);
      double r1 = B.dxc[j1] * fabs(c[i][j2]) + B.dxc[j2] * fabs(c[i][j1]);
      double d1 = 0, d2 = 0;
      for (int k = 0; k < 3; k++) {
        d2 += A.vec[i2][k] * D[k];
        d1 += A.vec[i1][k] * D[k];
      }
      double r = fabs(c[i1][j] * d2 - c[i2][j] * d1);
      if (r > (r0 + r1 + eps)) return 0;
    }
  }
  return 1;
}

void obb_corners(const double xc[3], const double dxc[3], const double vec[3][3], double xv[8][3])
{
  for (int i = 0; i < 8; i++) {
    for (int k = 0; k < 3; k++) xv[i][k] = xc[k];
    for (int k = 0; k < 3; k++) {
      int s = 2 * ((i & (1 << k)) >> k) - 1;
      for (int j = 0; j < 3; j++) xv[i][j] += s * dxc[k] * vec[k][j];
    }
  }
}

void median_select_lists(int *ix0, double *x0, int n)
{
  int *IX = ix0 - 1;
  double *X = x0 - 1;
  auto swap = [&](int a, int b) {
    std::swap(X[a], X[b]);
    std::swap(IX[a], IX[b]);
  };
  const int half = n / 2, halfp1 = half + 1;
  const bool odd = (n != 2 * half);
  if (n < 3) {
    if (n == 2 && X[2] < X[1]) swap(1, 2);
    return;
  }
  int lo = 1, hi = n;
  std::vector<int> A, B;
  for (;;) {
    double pv = median_pivot(X, lo, hi);
    A.clear();
    B.clear();
    for (int p = lo; p <= hi; p++)
      if (X[p] >= pv) A.push_back(p);
    for (int p = hi; p >= lo; p--)
      if (X[p] <= pv) B.push_back(p);
    int nA = (int)A.size(), nB = (int)B.size(), K = 0;
    while (K < nA && K < nB && A[K] < B[K]) K++;
    for (int m = 0; m < K; m++) swap(A[m], B[m]);
    int i, j;
    hoare_stop(K, nA, nB, K < nA ? A[K] : 0, K > 0 ? A[K - 1] : 0, K < nB ? B[K] : 0, K > 0 ? B[K - 1] : 0, A[0],
               B[0], i, j);
    int st = median_control(n, i, j, lo, hi);
    if (st == 1) return;
    if (st == 2) break;
  }
  if (!odd) {
    if (X[halfp1] < X[half]) swap(half, halfp1);
    return;
  }
  if (X[lo] > X[hi]) swap(lo, hi);
}

namespace {

struct AdtBuilder {
  const double *box;
  double *real;
  int *ints;
  std::vector<double> work;
  int count = -1, depth = 0;

  // buildADTrecursion.c:24-116: node = element at the median position
  // (nav+1)/2-1 along dimension level%6; node box = bounds of the subtree.
  void rec(int *ea, int side, int parent, int level, int nav)
  {
    if (nav <= 0) return;
    if (level + 1 > depth) depth = level + 1;
    if (nav == 1) {
      int me = ++count;
      ints[4 * me] = ea[0];
      ints[4 * me + 1] = ints[4 * me + 2] = ints[4 * me + 3] = -1;
      for (int j = 0; j < 6; j++) real[6 * me + j] = box[6 * ea[0] + j];
      if (side > 0) ints[4 * parent + side] = ea[0];
      return;
    }
    const int dim = level % 6;
    for (int i = 0; i < nav; i++) work[i] = box[6 * ea[i] + dim];
    median_select(ea, work.data(), nav);
    const int nleft = (nav + 1) / 2;
    int me = ++count;
    ints[4 * me] = ea[nleft - 1];
    ints[4 * me + 1] = ints[4 * me + 2] = ints[4 * me + 3] = -1;
    double *r = &real[6 * me];
    for (int j = 0; j < 3; j++) {
      r[j] = BIGVALUE;
      r[j + 3] = -BIGVALUE;
    }
    for (int i = 0; i < nav; i++) {
      const double *b = &box[6 * ea[i]];
      for (int j = 0; j < 3; j++) {
        r[j] = dmin(r[j], b[j]);
        r[j + 3] = dmax(r[j + 3], b[j + 3]);
      }
    }
    if (side > 0) ints[4 * parent + side] = ea[nleft - 1];
    if (nleft > 1) rec(ea, 1, me, level + 1, nleft - 1);
    rec(ea + nleft, 2, me, level + 1, nav - nleft);
  }
};

}  // namespace

int build_adt(int nelem, const double *ebox, int *adtInt, double *adtReal, double ext[6])
{
  for (int i = 0; i < 3; i++) {
    ext[2 * i] = BIGVALUE;
    ext[2 * i + 1] = -BIGVALUE;
  }
  for (int e = 0; e < nelem; e++)
    for (int i = 0; i < 3; i++) {
      ext[2 * i] = dmin(ext[2 * i], ebox[6 * e + i]);
      ext[2 * i + 1] = dmax(ext[2 * i + 1], ebox[6 * e + i + 3]);
    }
  for (int i = 0; i < 3; i++) {  // ADT.C:81-86: 1 % padding
    double d = 0.01 * (ext[2 * i + 1] - ext[2 * i]);
    ext[2 * i] -= d;
    ext[2 * i + 1] += d;
  }
  if (nelem <= 0) return 0;
  std::vector<int> ea(nelem);
  for (int i = 0; i < nelem; i++) ea[i] = i;
  AdtBuilder b{ebox, adtReal, adtInt};
  b.work.resize(nelem);
  b.rec(ea.data(), 0, 0, 0, nelem);
  // slot 3 of row e: the node that holds element e (ADT.C:120-124)
  for (int i = 0; i < nelem; i++) adtInt[4 * adtInt[4 * i] + 3] = i;
  return b.depth;
}

namespace {

struct Dedupe {
  const double *x;
  const int *tag;
  double *rtag;
  int *itag;

  // utils.c:660-712: split into octants above 20 points (octant members in
  // reverse arrival order), otherwise compare all pairs.
  void rec(int *ea, int nav)
  {
    bool brute = true;
    if (nav > 20) {
      brute = false;
      double lo[3] = {BIGVALUE, BIGVALUE, BIGVALUE}, hi[3] = {-BIGVALUE, -BIGVALUE, -BIGVALUE};
      for (int i = 0; i < nav; i++)
        for (int j = 0; j < 3; j++) {
          lo[j] = dmin(lo[j], x[3 * ea[i] + j]);
          hi[j] = dmax(hi[j], x[3 * ea[i] + j]);
        }
      double mid[3], dx[3];
      for (int j = 0; j < 3; j++) {
        mid[j] = (hi[j] + lo[j]) * 0.5;
        dx[j] = (hi[j] - lo[j]) * 0.5 + TOL;
      }
      auto octant = [&](int p) {
        int iv[3];
        for (int j = 0; j < 3; j++) iv[j] = (int)(floor((x[3 * p + j] - mid[j]) / dx[j]) + 1);
        return 4 * iv[0] + 2 * iv[1] + iv[2];
      };
      int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0}, start[9];
      for (int i = 0; i < nav; i++) cnt[octant(ea[i])]++;
      for (int j = 0; j < 8; j++)
        if (cnt[j] == nav) brute = true;
      if (!brute) {
        std::vector<int> tmp(nav);
        start[0] = 0;
        for (int j = 0; j < 8; j++) start[j + 1] = start[j] + cnt[j];
        for (int i = 0; i < nav; i++) {
          int o = octant(ea[i]);
          tmp[--cnt[o] + start[o]] = ea[i];
        }
        memcpy(ea, tmp.data(), sizeof(int) * nav);
        for (int j = 0; j < 8; j++)
          if (start[j + 1] > start[j]) rec(ea + start[j], start[j + 1] - start[j]);
      }
    }
    if (!brute) return;
    for (int i = 0; i < nav; i++) {
      int p1 = ea[i];
      for (int j = i + 1; j < nav; j++) {
        int p2 = ea[j];
        if (fabs(x[3 * p1] - x[3 * p2]) + fabs(x[3 * p1 + 1] - x[3 * p2 + 1]) +
                    fabs(x[3 * p1 + 2] - x[3 * p2 + 2]) <
                TOL &&
            tag[p1] == tag[p2]) {
          if (p1 > p2) {
            rtag[p2] = dmax(rtag[p1], rtag[p2]);
            itag[p1] = itag[p2];
          } else {
            rtag[p1] = dmax(rtag[p1], rtag[p2]);
            itag[p2] = itag[p1];
          }
        }
      }
    }
  }
};

}  // namespace

void unique_nodes_octree(const double *x, const int *meshtag, double *rtag, int *xtag, int n)
{
  std::vector<int> ea(n > 0 ? n : 1);
  for (int i = 0; i < n; i++) {
    ea[i] = i;
    xtag[i] = i;
  }
  Dedupe d{x, meshtag, rtag, xtag};
  d.rec(ea.data(), n);
}

}  // namespace tg
