// tg_common.h -- constants and FP64 cell numerics shared by the HIP kernels
// and the host-side sequential stages of the engine.
//
// Every routine keeps the reference's operation order, and the whole library
// is compiled with -ffp-contract=off, so device results are bit-identical to
// the reference's x86-64 results (IEEE double, no FMA contraction, correctly
// rounded division and sqrt).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#define TG_HD __host__ __device__ __forceinline__

namespace tg {

// codetypes.h:36-55
constexpr int BASE = 1;
constexpr double BIGVALUE = 1.0e15;
constexpr int BIGINT = 2147483647;
constexpr double TOL = 1.0e-10;
constexpr int HOLEMAPSIZE = 192;

TG_HD double dmin(double a, double b) { return a < b ? a : b; }  // TIOGA_MIN
TG_HD double dmax(double a, double b) { return a > b ? a : b; }  // TIOGA_MAX

// The 16-bit grid coordinate of x in a tree's grid (lo, inv per axis):
// clamp(((float)x - lo) * inv, 0, 65535), truncated.  Every step (rounding to
// float, subtracting a constant, multiplying by a constant inv >= 0, the
// clamps, truncation) is non-decreasing in x, so x >= b implies
// grid_of(x) >= grid_of(b) and x <= b implies grid_of(x) <= grid_of(b): a box
// test on grid coordinates passes whenever the FP64 test passes.
TG_HD unsigned grid_of(double x, float lo, float inv)
{
  float u = ((float)x - lo) * inv;
  u = u > 0.0f ? u : 0.0f;  // also maps NaN to 0
  u = u < 65535.0f ? u : 65535.0f;
  return (unsigned)u;
}

// Gaussian elimination on a 3x3 system with row swaps only when
// |a_ii| < 1e-8 (math.c:25-81).  Returns 0 for a singular system.
TG_HD int solve3(double a[3][3], double b[3])
{
  for (int i = 0; i < 3; i++) {
    if (fabs(a[i][i]) < 1e-8) {
      int swapped = 0;
      for (int k = i + 1; k < 3 && !swapped; k++) {
        if (a[k][i] != 0) {
          swapped = 1;
          for (int l = 0; l < 3; l++) {
            double t = a[k][l];
            a[k][l] = a[i][l];
            a[i][l] = t;
          }
          double t = b[k];
          b[k] = b[i];
          b[i] = t;
        }
      }
      if (!swapped) return 0;
    }
    for (int k = i + 1; k < 3; k++) {
      double fact = -a[k][i] / a[i][i];
      for (int j = 0; j < 3; j++) a[k][j] += fact * a[i][j];
      b[k] += fact * b[i];
    }
  }
  for (int i = 2; i >= 0; i--) {
    double s = 0;
    for (int j = i + 1; j < 3; j++) s += a[i][j] * b[j];
    b[i] = (b[i] - s) / a[i][i];
  }
  return 1;
}

// Newton iteration for the isoparametric map
//   f0 + f1 u + f2 v + f3 w + f4 uv + f5 vw + f6 wu + f7 uvw = 0
// from (1/2,1/2,1/2), full step, absolute stop sqrt(|r|^2) <= 1e-14, at most
// 500 iterations; failure yields (2,0,0) (math.c:83-146).
//
// newton3_steps runs iterations [it, itstop) from the state (u, v, w); it
// returns 1 when the iteration has ended (converged, or a singular Jacobian:
// ok = 0) and 0 when it stopped at itstop, so a solve can be continued later
// from the same state with identical arithmetic.
TG_HD int newton3_steps(const double f[8][3], double &u, double &v, double &w, int &it, int itstop, int &ok)
{
  for (; it < itstop; it++) {
    double uv = u * v, vw = v * w, wu = w * u, uvw = u * v * w;
    double rhs[3], lhs[3][3];
    for (int j = 0; j < 3; j++)
      rhs[j] = f[0][j] + f[1][j] * u + f[2][j] * v + f[3][j] * w + f[4][j] * uv + f[5][j] * vw +
               f[6][j] * wu + f[7][j] * uvw;
    double nrm = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
    if (sqrt(nrm) <= 1e-14) return 1;
    for (int j = 0; j < 3; j++) {
      lhs[j][0] = f[1][j] + f[4][j] * v + f[6][j] * w + f[7][j] * vw;
      lhs[j][1] = f[2][j] + f[5][j] * w + f[4][j] * u + f[7][j] * wu;
      lhs[j][2] = f[3][j] + f[6][j] * u + f[5][j] * v + f[7][j] * uv;
    }
    ok = solve3(lhs, rhs);
    if (!ok) return 1;
    u -= rhs[0];
    v -= rhs[1];
    w -= rhs[2];
  }
  return it >= 500;
}

constexpr int NEWTON_MAX = 500;

// the outcome of a finished solve (math.c:127-138)
TG_HD void newton3_finish(int it, int ok, double &u, double &v, double &w)
{
  if (it == NEWTON_MAX || !ok) {
    u = 2.0;
    v = w = 0.0;
  }
}

TG_HD void newton3(const double f[8][3], double &uo, double &vo, double &wo)
{
  double u = 0.5, v = 0.5, w = 0.5;
  int it = 0, ok = 1;
  newton3_steps(f, u, v, w, it, NEWTON_MAX, ok);
  newton3_finish(it, ok, u, v, w);
  uo = u;
  vo = v;
  wo = w;
}

// Coefficients of the isoparametric map of a pyramid (5), prism (6) or hex
// (8) about the point xp (math.c:196-281).
TG_HD void iso_coeffs(const double xv[8][3], const double xp[3], double f[8][3], int nvert)
{
  if (nvert == 5) {
    for (int j = 0; j < 3; j++) {
      f[0][j] = xv[0][j] - xp[j];
      f[1][j] = xv[1][j] - xv[0][j];
      f[2][j] = xv[3][j] - xv[0][j];
      f[3][j] = xv[4][j] - xv[0][j];
      f[4][j] = xv[0][j] - xv[1][j] + xv[2][j] - xv[3][j];
      f[5][j] = xv[0][j] - xv[3][j];
      f[6][j] = xv[0][j] - xv[1][j];
      f[7][j] = -xv[0][j] + xv[1][j] - xv[2][j] + xv[3][j];
    }
  } else if (nvert == 6) {
    for (int j = 0; j < 3; j++) {
      f[0][j] = xv[0][j] - xp[j];
      f[1][j] = xv[1][j] - xv[0][j];
      f[2][j] = xv[2][j] - xv[0][j];
      f[3][j] = xv[3][j] - xv[0][j];
      f[4][j] = 0;
      f[5][j] = xv[0][j] - xv[2][j] - xv[3][j] + xv[5][j];
      f[6][j] = xv[0][j] - xv[1][j] - xv[3][j] + xv[4][j];
      f[7][j] = 0.;
    }
  } else {
    for (int j = 0; j < 3; j++) {
      f[0][j] = xv[0][j] - xp[j];
      f[1][j] = xv[1][j] - xv[0][j];
      f[2][j] = xv[3][j] - xv[0][j];
      f[3][j] = xv[4][j] - xv[0][j];
      f[4][j] = xv[0][j] - xv[1][j] + xv[2][j] - xv[3][j];
      f[5][j] = xv[0][j] - xv[3][j] + xv[7][j] - xv[4][j];
      f[6][j] = xv[0][j] - xv[1][j] + xv[5][j] - xv[4][j];
      f[7][j] = -xv[0][j] + xv[1][j] - xv[2][j] + xv[3][j] + xv[4][j] - xv[5][j] + xv[6][j] - xv[7][j];
    }
  }
}

// Nodal weights from the parametric point of a pyramid, prism or hex
// (math.c:213-221, 243-253, 276-287).
TG_HD void iso_weights(double u, double v, double w, double frac[8], int nvert)
{
  if (nvert == 5) {
    double omu = 1.0 - u, omv = 1.0 - v, omw = 1.0 - w;
    frac[0] = omu * omv * omw;
    frac[1] = u * omv * omw;
    frac[2] = u * v * omw;
    frac[3] = omu * v * omw;
    frac[4] = w;
  } else if (nvert == 6) {
    double omuv = 1.0 - u - v, omw = 1.0 - w;
    frac[0] = omuv * omw;
    frac[1] = u * omw;
    frac[2] = v * omw;
    frac[3] = omuv * w;
    frac[4] = u * w;
    frac[5] = v * w;
  } else {
    double omu = 1.0 - u, omv = 1.0 - v, omw = 1.0 - w;
    frac[0] = omu * omv * omw;
    frac[1] = u * omv * omw;
    frac[2] = u * v * omw;
    frac[3] = omu * v * omw;
    frac[4] = omu * omv * w;
    frac[5] = u * omv * w;
    frac[6] = u * v * w;
    frac[7] = omu * v * w;
  }
}

// Nodal interpolation weights of point xp in a tet (direct solve), pyramid,
// prism or hex (Newton) cell (math.c:149-294).
TG_HD void nodal_weights(const double xv[8][3], const double xp[3], double frac[8], int nvert)
{
  if (nvert == 4) {
    double lhs[3][3], rhs[3];
    for (int k = 0; k < 3; k++) {
      for (int j = 0; j < 3; j++) lhs[j][k] = xv[k][j] - xv[3][j];
      rhs[k] = xp[k] - xv[3][k];
    }
    if (solve3(lhs, rhs)) {
      frac[0] = rhs[0];
      frac[1] = rhs[1];
      frac[2] = rhs[2];
      frac[3] = 1. - frac[0] - frac[1] - frac[2];
    } else {
      frac[0] = 1.0;
      frac[1] = frac[2] = frac[3] = 0;
    }
    return;
  }
  if (nvert != 5 && nvert != 6 && nvert != 8) return;
  double f[8][3], u, v, w;
  iso_coeffs(xv, xp, f, nvert);
  newton3(f, u, v, w);
  iso_weights(u, v, w, frac, nvert);
}

// a . (b x c), expanded in the order of cellVolume.f90:99-109
TG_HD double triple(const double *a, const double *b, const double *c)
{
  return a[0] * b[1] * c[2] - a[0] * b[2] * c[1] + a[1] * b[2] * c[0] - a[1] * b[0] * c[2] +
         a[2] * b[0] * c[1] - a[2] * b[1] * c[0];
}

// Cell volume by the divergence theorem over the faces of math.c:298-330
// (quads contribute the four triangles of cellVolume.f90:73-87).
TG_HD double cell_volume(const double xv[8][3], int nvert)
{
  // faces as 1-based vertex lists, 0 terminating a triangle
  constexpr int kTet[4][4] = {{1, 2, 3, 0}, {1, 4, 2, 0}, {2, 4, 3, 0}, {1, 3, 4, 0}};
  constexpr int kPyr[5][4] = {{1, 2, 3, 4}, {1, 5, 2, 0}, {2, 5, 3, 0}, {4, 3, 5, 0}, {1, 4, 5, 0}};
  constexpr int kPri[5][4] = {{1, 2, 3, 0}, {1, 4, 5, 2}, {2, 5, 6, 3}, {1, 3, 6, 4}, {4, 6, 5, 0}};
  constexpr int kHex[6][4] = {{1, 2, 3, 4}, {1, 5, 6, 2}, {2, 6, 7, 3},
                              {3, 7, 8, 4}, {1, 4, 8, 5}, {5, 8, 7, 6}};
  const int(*faces)[4];
  int nf;
  if (nvert == 4) { faces = kTet; nf = 4; }
  else if (nvert == 5) { faces = kPyr; nf = 5; }
  else if (nvert == 6) { faces = kPri; nf = 5; }
  else { faces = kHex; nf = 6; }
  double vol = 0.;
  for (int i = 0; i < nf; i++) {
    const double *p1 = xv[faces[i][0] - 1], *p2 = xv[faces[i][1] - 1], *p3 = xv[faces[i][2] - 1];
    if (faces[i][3] == 0) {
      vol = vol - 0.5 * triple(p1, p2, p3);
    } else {
      const double *p4 = xv[faces[i][3] - 1];
      vol = vol - 0.25 * triple(p1, p2, p3);
      vol = vol - 0.25 * triple(p1, p3, p4);
      vol = vol - 0.25 * triple(p1, p2, p4);
      vol = vol - 0.25 * triple(p2, p3, p4);
    }
  }
  return vol / 3.0;
}

// Partial quicksort that isolates the median of X[1..n] and permutes IX
// alongside (median.F90:59-210; 1-based like the Fortran).  Positions
// (n+1)/2 and, for even n, n/2+1 hold their sorted values on return; the
// exact permutation it leaves decides the ADT's shape, so every step is kept.
TG_HD void median_select(int *ix0, double *x0, int n)
{
  int *IX = ix0 - 1;
  double *X = x0 - 1;
  auto swap = [&](int a, int b) {
    double t = X[a];
    X[a] = X[b];
    X[b] = t;
    int it = IX[a];
    IX[a] = IX[b];
    IX[b] = it;
  };
  const int half = n / 2, halfp1 = half + 1;
  const bool odd = (n != 2 * half);
  if (n < 3) {
    if (n == 2 && X[2] < X[1]) swap(1, 2);
    return;
  }
  int lo = 1, hi = n;
  for (;;) {
    int mid = (lo + hi) / 2;
    double pv = X[mid], a = X[lo], b = X[hi];
    if (b < a) {
      double t = b;
      b = a;
      a = t;
    }
    if (pv > b) pv = b;
    else if (pv < a) pv = a;
    int i = lo, j = hi;
    for (;;) {
      while (!(X[i] >= pv)) i++;
      while (!(X[j] <= pv)) j--;
      if (i < j) {
        swap(i, j);
        i++;
        j--;
        if (i <= j) continue;
      }
      break;
    }
    if (!odd) {
      if (j == half && i == halfp1) return;
      if (j < half) lo = i;
      if (i > halfp1) hi = j;
      if (i == j) {
        if (i == half) lo = half;
        if (j == halfp1) hi = halfp1;
      }
    } else {
      if (j < halfp1) lo = i;
      if (i > halfp1) hi = j;
      if (i == j && i == halfp1) return;
    }
    if (lo < hi - 1) continue;
    break;
  }
  if (!odd) {
    if (X[halfp1] < X[half]) swap(half, halfp1);
    return;
  }
  if (X[lo] > X[hi]) swap(lo, hi);
}

// One round of median_select's partition loop stated over lists, the form
// the device runs with a whole workgroup.  A = positions in [lo,hi] with
// X >= pv (ascending), B = positions with X <= pv (descending).  The scan
// pointers of the sequential loop meet the pairs (A[m], B[m]) in order and
// swap them while A[m] < B[m]; after K swaps they stop at
//   i = A[K] if A[K] < B[K-1] else B[K-1],  j = B[K] if B[K] > A[K-1] else A[K-1]
// (i = A[0], j = B[0] when K = 0).  Given K and the four list entries this
// returns the (i, j) the sequential loop ends with.
TG_HD void hoare_stop(int K, int nA, int nB, int aK, int aK1, int bK, int bK1, int a0, int b0, int &i, int &j)
{
  if (K == 0) {
    i = a0;
    j = b0;
    return;
  }
  i = (K < nA && aK < bK1) ? aK : bK1;
  j = (K < nB && bK > aK1) ? bK : aK1;
}

// The loop-control step of median_select after a partition round ending at
// (i, j): narrows [lo, hi]; returns 1 when the median is in place (the
// function returns), 2 when the loop falls through to the final fix-up, 0
// to partition again.
TG_HD int median_control(int n, int i, int j, int &lo, int &hi)
{
  const int half = n / 2, halfp1 = half + 1;
  const bool odd = (n != 2 * half);
  if (!odd) {
    if (j == half && i == halfp1) return 1;
    if (j < half) lo = i;
    if (i > halfp1) hi = j;
    if (i == j) {
      if (i == half) lo = half;
      if (j == halfp1) hi = halfp1;
    }
  } else {
    if (j < halfp1) lo = i;
    if (i > halfp1) hi = j;
    if (i == j && i == halfp1) return 1;
  }
  return (lo < hi - 1) ? 0 : 2;
}

// median_select's pivot: X[mid] clamped to [min, max] of X[lo], X[hi]
TG_HD double median_pivot(const double *X1, int lo, int hi)
{
  int mid = (lo + hi) / 2;
  double pv = X1[mid], a = X1[lo], b = X1[hi];
  if (b < a) {
    double t = b;
    b = a;
    a = t;
  }
  if (pv > b) pv = b;
  else if (pv < a) pv = a;
  return pv;
}

// Oriented bounding box: centre, half extents, rows of vec = axes
// (codetypes.h:92-104).
struct OBB {
  double xc[3], dxc[3], vec[3][3];
};

// Coordinates of p in the frame (xc, vec), accumulated from 0 as
// utils.c:402-411 does.
TG_HD void to_frame(const double *p, const double xc[3], const double vec[3][3], double xd[3])
{
  for (int j = 0; j < 3; j++) {
    xd[j] = 0;
    for (int k = 0; k < 3; k++) xd[j] += (p[k] - xc[k]) * vec[j][k];
  }
}

// Order-preserving maps between doubles and unsigned 64-bit keys, so exact
// min/max reductions of doubles can use integer atomics.
TG_HD uint64_t dkey(double d)
{
  uint64_t u;
  memcpy(&u, &d, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
TG_HD double dunkey(uint64_t k)
{
  uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  memcpy(&d, &u, 8);
  return d;
}

}  // namespace tg
