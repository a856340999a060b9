/*
 * tioga_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Single-threaded C restatement of TIOGA's single-rank connectivity path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker; the shipped engine never calls it.
 *
 * Each function names the reference file:line range it restates.  Where the
 * reference has order-dependent behaviour (linked-list arrival order, LIFO
 * cell lists, first-hit ADT traversal, in-place mutation during
 * findInterpData) the order is reproduced as written.
 *
 * Every function is pinned against the compiled reference in oracle/_ref:
 * tests/test_oracle_golden.py compares the whole path with the golden
 * fixtures that reference wrote (tests/golden/), and
 * tests/test_oracle_closed_forms.py checks the leaf numerics against closed
 * forms.
 *
 * Build flags: -O2 -ffp-contract=off, no fast-math.
 */
#include "tioga_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>
#include <time.h>

#define BIGV ORC_BIGVALUE
#define TOLV ORC_TOL
#define MAXBLK 128

static int g_weight_aborts = 0;
int orc_weight_aborts(void) { return g_weight_aborts; }

static double dmin(double a, double b) { return a < b ? a : b; }  /* TIOGA_MIN, codetypes.h:63 */
static double dmax(double a, double b) { return a > b ? a : b; }  /* TIOGA_MAX, codetypes.h:64 */

/* ===================================================================== */
/* math.c                                                                */
/* ===================================================================== */

/* math.c:25-81 -- Gaussian elimination, pivot only when |a_ii| < 1e-8,
   swapping in the first later row with a non-zero entry in column i. */
void orc_solvec(double a[3][3], double b[3], int *iflag)
{
  const int n = 3;
  int i, j, k, l, flag;
  double fact, temp, sum;
  for (i = 0; i < n; i++) {
    if (fabs(a[i][i]) < 1e-8) {
      flag = 1;
      for (k = i + 1; k < n && flag; k++) {
        if (a[k][i] != 0) {
          flag = 0;
          for (l = 0; l < n; l++) { temp = a[k][l]; a[k][l] = a[i][l]; a[i][l] = temp; }
          temp = b[k]; b[k] = b[i]; b[i] = temp;
        }
      }
      if (flag) { *iflag = 0; return; }
    }
    for (k = i + 1; k < n; k++) {
      fact = -a[k][i] / a[i][i];
      for (j = 0; j < n; j++) a[k][j] += fact * a[i][j];
      b[k] += fact * b[i];
    }
  }
  for (i = n - 1; i >= 0; i--) {
    sum = 0;
    for (j = i + 1; j < n; j++) sum += a[i][j] * b[j];
    b[i] = (b[i] - sum) / a[i][i];
  }
  *iflag = 1;
}

/* math.c:83-146 -- Newton iteration on the trilinear map from (1/2,1/2,1/2),
   full step, absolute stop sqrt(|r|^2) <= 1e-14, at most 500 iterations;
   non-convergence or a singular Jacobian gives (2,0,0). */
void orc_newton_solve(double f[8][3], double *u1, double *v1, double *w1)
{
  int j, iter, isolflag = 1;
  const int itmax = 500;
  double u = 0.5, v = 0.5, w = 0.5;
  double uv, vw, wu, uvw, norm;
  double rhs[3], lhs[3][3];
  for (iter = 0; iter < itmax; iter++) {
    uv = u * v;
    vw = v * w;
    wu = w * u;
    uvw = u * v * w;
    for (j = 0; j < 3; j++)
      rhs[j] = f[0][j] + f[1][j] * u + f[2][j] * v + f[3][j] * w +
               f[4][j] * uv + f[5][j] * vw + f[6][j] * wu + f[7][j] * uvw;
    norm = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
    if (sqrt(norm) <= 1e-14) break;
    for (j = 0; j < 3; j++) {
      lhs[j][0] = f[1][j] + f[4][j] * v + f[6][j] * w + f[7][j] * vw;
      lhs[j][1] = f[2][j] + f[5][j] * w + f[4][j] * u + f[7][j] * wu;
      lhs[j][2] = f[3][j] + f[6][j] * u + f[5][j] * v + f[7][j] * uv;
    }
    orc_solvec(lhs, rhs, &isolflag);
    if (isolflag == 0) break;
    u -= rhs[0];
    v -= rhs[1];
    w -= rhs[2];
  }
  if (iter == itmax) { u = 2.0; v = w = 0.; }
  if (isolflag == 0) { u = 2.0; v = w = 0.; }
  *u1 = u; *v1 = v; *w1 = w;
}

/* math.c:149-294 -- nodal weights for tet (direct 3x3 solve), pyramid, prism
   and hex (Newton on the isoparametric map). */
void orc_nodal_weights(const double xv[8][3], const double *xp, double frac[8], int nvert)
{
  int j, k, isolflag;
  double lhs[3][3], rhs[3], f[8][3];
  double u, v, w, omu, omv, omw, omuv;
  switch (nvert) {
  case 4:
    for (k = 0; k < 3; k++) {
      for (j = 0; j < 3; j++) lhs[j][k] = xv[k][j] - xv[3][j];
      rhs[k] = xp[k] - xv[3][k];
    }
    orc_solvec(lhs, rhs, &isolflag);
    if (isolflag) {
      for (k = 0; k < 3; k++) frac[k] = rhs[k];
      frac[3] = 1. - frac[0] - frac[1] - frac[2];
    } else {
      frac[0] = 1.0;
      frac[1] = frac[2] = frac[3] = 0;
    }
    break;
  case 5:
    for (j = 0; j < 3; j++) {
      f[0][j] = xv[0][j] - xp[j];
      f[1][j] = xv[1][j] - xv[0][j];
      f[2][j] = xv[3][j] - xv[0][j];
      f[3][j] = xv[4][j] - xv[0][j];
      f[4][j] = xv[0][j] - xv[1][j] + xv[2][j] - xv[3][j];
      f[5][j] = xv[0][j] - xv[3][j];
      f[6][j] = xv[0][j] - xv[1][j];
      f[7][j] = -xv[0][j] + xv[1][j] - xv[2][j] + xv[3][j];
    }
    orc_newton_solve(f, &u, &v, &w);
    omu = 1.0 - u; omv = 1.0 - v; omw = 1.0 - w;
    frac[0] = omu * omv * omw;
    frac[1] = u * omv * omw;
    frac[2] = u * v * omw;
    frac[3] = omu * v * omw;
    frac[4] = w;
    break;
  case 6:
    for (j = 0; j < 3; j++) {
      f[0][j] = xv[0][j] - xp[j];
      f[1][j] = xv[1][j] - xv[0][j];
      f[2][j] = xv[2][j] - xv[0][j];
      f[3][j] = xv[3][j] - xv[0][j];
      f[4][j] = 0;
      f[5][j] = xv[0][j] - xv[2][j] - xv[3][j] + xv[5][j];
      f[6][j] = xv[0][j] - xv[1][j] - xv[3][j] + xv[4][j];
      f[7][j] = 0.;
    }
    orc_newton_solve(f, &u, &v, &w);
    omuv = 1.0 - u - v; omu = 1.0 - u; omv = 1.0 - v; omw = 1.0 - w;
    (void)omu; (void)omv;
    frac[0] = omuv * omw;
    frac[1] = u * omw;
    frac[2] = v * omw;
    frac[3] = omuv * w;
    frac[4] = u * w;
    frac[5] = v * w;
    break;
  case 8:
    for (j = 0; j < 3; j++) {
      f[0][j] = xv[0][j] - xp[j];
      f[1][j] = xv[1][j] - xv[0][j];
      f[2][j] = xv[3][j] - xv[0][j];
      f[3][j] = xv[4][j] - xv[0][j];
      f[4][j] = xv[0][j] - xv[1][j] + xv[2][j] - xv[3][j];
      f[5][j] = xv[0][j] - xv[3][j] + xv[7][j] - xv[4][j];
      f[6][j] = xv[0][j] - xv[1][j] + xv[5][j] - xv[4][j];
      f[7][j] = -xv[0][j] + xv[1][j] - xv[2][j] + xv[3][j] + xv[4][j] - xv[5][j] + xv[6][j] - xv[7][j];
    }
    orc_newton_solve(f, &u, &v, &w);
    omu = 1.0 - u; omv = 1.0 - v; omw = 1.0 - w;
    frac[0] = omu * omv * omw;
    frac[1] = u * omv * omw;
    frac[2] = u * v * omw;
    frac[3] = omu * v * omw;
    frac[4] = omu * omv * w;
    frac[5] = u * omv * w;
    frac[6] = u * v * w;
    frac[7] = omu * v * w;
    break;
  default:
    break;
  }
}

/* math.c:298-330 face tables + cellVolume.f90:20-109 (divergence theorem;
   quads split into the four triangles of cellVolume.f90:73-87). */
static const int kNumVerts[4][6] = {
    {3, 3, 3, 3, 0, 0}, {4, 3, 3, 3, 3, 0}, {3, 4, 4, 4, 3, 0}, {4, 4, 4, 4, 4, 4}};
static const int kFaceInfo[4][24] = {
    {1, 2, 3, 0, 1, 4, 2, 0, 2, 4, 3, 0, 1, 3, 4, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {1, 2, 3, 4, 1, 5, 2, 0, 2, 5, 3, 0, 4, 3, 5, 0, 1, 4, 5, 0, 0, 0, 0, 0},
    {1, 2, 3, 0, 1, 4, 5, 2, 2, 5, 6, 3, 1, 3, 6, 4, 4, 6, 5, 0, 0, 0, 0, 0},
    {1, 2, 3, 4, 1, 5, 6, 2, 2, 6, 7, 3, 3, 7, 8, 4, 1, 4, 8, 5, 5, 8, 7, 6}};

static double box_product(const double *a, const double *b, const double *c)
{
  /* cellVolume.f90:99-109, evaluated left to right */
  return a[0] * b[1] * c[2] - a[0] * b[2] * c[1] + a[1] * b[2] * c[0] - a[1] * b[0] * c[2] +
         a[2] * b[0] * c[1] - a[2] * b[1] * c[0];
}

double orc_cell_volume(const double xv[8][3], int nvert)
{
  int itype, nfaces, iface;
  switch (nvert) {
  case 4: itype = 0; nfaces = 4; break;
  case 5: itype = 1; nfaces = 5; break;
  case 6: itype = 2; nfaces = 5; break;
  default: itype = 3; nfaces = 6; break;
  }
  double vol = 0.;
  for (iface = 0; iface < nfaces; iface++) {
    const int *fc = &kFaceInfo[itype][4 * iface];
    const double *p1 = xv[fc[0] - 1], *p2 = xv[fc[1] - 1], *p3 = xv[fc[2] - 1];
    if (kNumVerts[itype][iface] == 3) {
      vol = vol - 0.5 * box_product(p1, p2, p3);
    } else {
      const double *p4 = xv[fc[3] - 1];
      vol = vol - 0.25 * box_product(p1, p2, p3);
      vol = vol - 0.25 * box_product(p1, p3, p4);
      vol = vol - 0.25 * box_product(p1, p2, p4);
      vol = vol - 0.25 * box_product(p2, p3, p4);
    }
  }
  return vol / 3.0;
}

/* ===================================================================== */
/* median.F90:59-210 (partial quicksort isolating the median, reordering  */
/* the index array alongside the values); 1-based as in the Fortran.     */
/* ===================================================================== */
void orc_median(int *ixa, double *xa, int n)
{
  int *IX = ixa - 1;
  double *X = xa - 1;
  int nby2 = n / 2, nby2p1 = nby2 + 1, odd = 1;
  int lo, hi, mid, i, j, itemp;
  double xmed, xlo, xhi, temp;
  if (n == 2 * nby2) odd = 0;
  lo = 1;
  hi = n;
  if (n < 3) {
    if (n < 2) return;
    if (X[2] < X[1]) {
      temp = X[1]; X[1] = X[2]; X[2] = temp;
      itemp = IX[1]; IX[1] = IX[2]; IX[2] = itemp;
    }
    return;
  }
L10:
  mid = (lo + hi) / 2;
  xmed = X[mid];
  xlo = X[lo];
  xhi = X[hi];
  if (xhi < xlo) { temp = xhi; xhi = xlo; xlo = temp; }
  if (xmed > xhi) xmed = xhi;
  else if (xmed < xlo) xmed = xlo;
  i = lo;
  j = hi;
L50:
  while (!(X[i] >= xmed)) i++;
  while (!(X[j] <= xmed)) j--;
  if (i < j) {
    temp = X[i]; X[i] = X[j]; X[j] = temp;
    itemp = IX[i]; IX[i] = IX[j]; IX[j] = itemp;
    i++;
    j--;
    if (i <= j) goto L50;
  }
  if (!odd) {
    if (j == nby2 && i == nby2p1) return; /* label 130: only xmed, unused by callers */
    if (j < nby2) lo = i;
    if (i > nby2p1) hi = j;
    if (i != j) goto L100;
    if (i == nby2) lo = nby2;
    if (j == nby2p1) hi = nby2p1;
  } else {
    if (j < nby2p1) lo = i;
    if (i > nby2p1) hi = j;
    if (i != j) goto L100;
    if (i == nby2p1) return;
  }
L100:
  if (lo < hi - 1) goto L10;
  if (!odd) {
    if (X[nby2p1] < X[nby2]) {
      temp = X[nby2p1]; X[nby2p1] = X[nby2]; X[nby2] = temp;
      itemp = IX[nby2p1]; IX[nby2p1] = IX[nby2]; IX[nby2] = itemp;
    }
    return;
  }
  temp = X[lo];
  itemp = IX[lo];
  if (temp > X[hi]) {
    X[lo] = X[hi]; X[hi] = temp;
    IX[lo] = IX[hi]; IX[hi] = itemp;
  }
}

/* ===================================================================== */
/* kaiser.f:29-180 (Jacobi-type JK eigen-solver), findOBB utils.c:25-165  */
/* ===================================================================== */
static void kaiser3(double *A, double eigenv[3])
{
#define AA(i, j) A[((j)-1) * 3 + ((i)-1)]
  const int n = 3;
  const double small = 1.0e-12, zero = 0.0, half = 0.5, one = 1.0;
  int i, j, k, iter = 0, ncount, nn;
  double absp, absq, cs, ctn, eps, halfp, p, q, sn, ss, tn, temp, xj, xk;
  ss = zero;
  for (j = 1; j <= n; j++)
    for (i = 1; i <= n; i++) ss = ss + AA(i, j) * AA(i, j);
  eps = small * ss / n;
  nn = n * (n - 1) / 2;
  ncount = nn;
L20:
  for (j = 1; j <= n - 1; j++) {
    for (k = j + 1; k <= n; k++) {
      halfp = zero;
      q = zero;
      for (i = 1; i <= n; i++) {
        xj = AA(i, j);
        xk = AA(i, k);
        halfp = halfp + xj * xk;
        q = q + (xj + xk) * (xj - xk);
      }
      p = halfp + halfp;
      absp = fabs(p);
      if (absp < eps && q >= zero) {
        ncount = ncount - 1;
        if (ncount <= 0) goto L160;
        continue;
      }
      absq = fabs(q);
      if (absp <= absq) {
        tn = absp / absq;
        cs = one / sqrt(one + tn * tn);
        sn = tn * cs;
      } else {
        ctn = absq / absp;
        sn = one / sqrt(one + ctn * ctn);
        cs = ctn * sn;
      }
      cs = sqrt((one + cs) * half);
      sn = sn / (cs + cs);
      if (q < zero) { temp = cs; cs = sn; sn = temp; }
      if (p < zero) sn = -sn;
      for (i = 1; i <= n; i++) {
        temp = AA(i, j);
        AA(i, j) = temp * cs + AA(i, k) * sn;
        AA(i, k) = -temp * sn + AA(i, k) * cs;
      }
    }
  }
  ncount = nn;
  iter = iter + 1;
  if (iter < 10) goto L20;
L160:
  for (j = 1; j <= n; j++) {
    temp = 0.0;
    for (i = 1; i <= n; i++) temp = temp + AA(i, j) * AA(i, j);
    eigenv[j - 1] = sqrt(temp);
  }
  for (j = 1; j <= n; j++) {
    temp = eigenv[j - 1] > zero ? one / eigenv[j - 1] : zero;
    for (i = 1; i <= n; i++) AA(i, j) = AA(i, j) * temp;
  }
#undef AA
}

void orc_find_obb(const double *x, double xc[3], double dxc[3], double vec[3][3], int nnodes)
{
  int i, j, k, m, i3;
  double aa[9], eig[3], xd[3], xmin[3], xmax[3];
  xc[0] = xc[1] = xc[2] = 0;
  for (i = 0; i < nnodes; i++) {
    i3 = 3 * i;
    xc[0] += x[i3];
    xc[1] += x[i3 + 1];
    xc[2] += x[i3 + 2];
  }
  xc[0] /= nnodes;
  xc[1] /= nnodes;
  xc[2] /= nnodes;
  if (nnodes < 4) {
    /* utils.c:51-80, including its unparenthesised TIOGA_MAX(...)*0.5 */
    vec[0][0] = vec[1][1] = vec[2][2] = 1;
    vec[0][1] = vec[1][0] = vec[1][2] = vec[2][1] = 0;
    vec[0][2] = vec[2][0] = 0;
    if (nnodes == 1) {
      dxc[0] = dxc[1] = dxc[2] = 1e-3;
    } else if (nnodes == 2) {
      for (j = 0; j < 3; j++) {
        double d = fabs(x[3 + j] - x[j]);
        dxc[j] = (1e-3 > d) ? 1e-3 : d * 0.5;
      }
    } else {
      for (i = 0; i < nnodes; i++)
        for (j = 0; j < 3; j++) dxc[j] = dmax(1e-3, fabs(x[3 * i + j] - x[0]));
    }
    return;
  }
  for (i = 0; i < 9; i++) aa[i] = 0;
  for (i = 0; i < nnodes; i++) {
    i3 = 3 * i;
    aa[0] += ((x[i3] - xc[0]) * (x[i3] - xc[0]));
    aa[4] += ((x[i3 + 1] - xc[1]) * (x[i3 + 1] - xc[1]));
    aa[8] += ((x[i3 + 2] - xc[2]) * (x[i3 + 2] - xc[2]));
    aa[3] += ((x[i3] - xc[0]) * (x[i3 + 1] - xc[1]));
    aa[6] += ((x[i3] - xc[0]) * (x[i3 + 2] - xc[2]));
    aa[7] += ((x[i3 + 1] - xc[1]) * (x[i3 + 2] - xc[2]));
  }
  aa[1] = aa[3];
  aa[2] = aa[6];
  aa[5] = aa[7];
  kaiser3(aa, eig);
  m = 0;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) vec[i][j] = aa[m++];
  for (j = 0; j < 3; j++) { xmax[j] = -BIGV; xmin[j] = BIGV; }
  for (i = 0; i < nnodes; i++) {
    i3 = 3 * i;
    for (j = 0; j < 3; j++) xd[j] = 0;
    for (j = 0; j < 3; j++)
      for (k = 0; k < 3; k++) xd[j] += (x[i3 + k] - xc[k]) * vec[j][k];
    for (j = 0; j < 3; j++) {
      xmax[j] = dmax(xmax[j], xd[j]);
      xmin[j] = dmin(xmin[j], xd[j]);
    }
  }
  for (j = 0; j < 3; j++) {
    dxc[j] = (xmax[j] - xmin[j]) * 0.5 * 1.01;
    xd[j] = (xmax[j] + xmin[j]) * 0.5;
  }
  for (j = 0; j < 3; j++)
    for (k = 0; k < 3; k++) xc[j] += (xd[k] * vec[k][j]);
}

/* utils.c:297-384 -- 15-axis separating-axis test, eps 1e-12 */
int orc_obb_intersect(const double vA[3][3], const double xA[3], const double dxA[3],
                      const double vB[3][3], const double xB[3], const double dxB[3])
{
  int i, j, k, i1, i2, j1, j2;
  double r, r0, r1, d1, d2, D[3], c[3][3];
  const double eps = 1e-12;
  for (i = 0; i < 3; i++) D[i] = xB[i] - xA[i];
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      c[i][j] = 0;
      for (k = 0; k < 3; k++) c[i][j] = c[i][j] + vA[i][k] * vB[j][k];
    }
  for (i = 0; i < 3; i++) {
    r0 = dxA[i]; r1 = 0; r = 0;
    for (j = 0; j < 3; j++) { r1 += dxB[j] * fabs(c[i][j]); r += fabs(vA[i][j]) * D[j]; }
    if (r > (r0 + r1 + eps)) return 0;
  }
  for (i = 0; i < 3; i++) {
    r1 = dxB[i]; r0 = 0; r = 0;
    for (j = 0; j < 3; j++) { r0 += dxA[j] * fabs(c[j][i]); r += fabs(vB[i][j]) * D[j]; }
    if (r > (r0 + r1 + eps)) return 0;
  }
  for (i = 0; i < 3; i++) {
    i1 = (i + 1) % 3;
    i2 = (i + 2) % 3;
    for (j = 0; j < 3; j++) {
      j1 = (j + 1) % 3;
      j2 = (j + 2) % 3;
      r0 = dxA[i1] * fabs(c[i2][j]) + dxA[i2] * fabs(c[i1][j]);
      r1 = dxB[j1] * fabs(c[i][j2]) + dxB[j2] * fabs(c[i][j1]);
      d2 = 0; d1 = 0;
      for (k = 0; k < 3; k++) { d2 += vA[i2][k] * D[k]; d1 += vA[i1][k] * D[k]; }
      r = fabs(c[i1][j] * d2 - c[i2][j] * d1);
      if (r > (r0 + r1 + eps)) return 0;
    }
  }
  return 1;
}

/* utils.c:386-400 */
static void obb_corners(const double xc[3], const double dxc[3], const double vec[3][3], double xv[8][3])
{
  int i, j, k, ik;
  for (i = 0; i < 8; i++) {
    for (k = 0; k < 3; k++) xv[i][k] = xc[k];
    for (k = 0; k < 3; k++) {
      ik = (2 * ((i & (1 << k)) >> k) - 1);
      for (j = 0; j < 3; j++) xv[i][j] += ik * dxc[k] * vec[k][j];
    }
  }
}

/* utils.c:402-411 */
static void to_obb(const double xv[3], const double xc[3], const double vec[3][3], double xd[3])
{
  int j, k;
  for (j = 0; j < 3; j++) {
    xd[j] = 0;
    for (k = 0; k < 3; k++) xd[j] += (xv[k] - xc[k]) * vec[j][k];
  }
}

/* utils.c:170-185 -- C truncation toward zero, range test on the int */
int orc_check_hole_map(const double *x, const int nx[3], const int *sam, const double ext[6])
{
  int i, ix[3];
  double dx[3];
  for (i = 0; i < 3; i++) dx[i] = (ext[i + 3] - ext[i]) / nx[i];
  for (i = 0; i < 3; i++) {
    ix[i] = (int)((x[i] - ext[i]) / dx[i]);
    if (ix[i] < 0 || ix[i] > nx[i] - 1) return 0;
  }
  return sam[ix[2] * nx[1] * nx[0] + ix[1] * nx[0] + ix[0]];
}

/* utils.c:192-295 -- boundary shell painted 1 (overwriting 2s), then
   in-place sweeps painting 0-cells next to a 1 under the isym stencil until
   nothing changes; finally sam = (sam==0 || sam==2). */
void orc_fill_hole_map(int *sam, const int ix[3], int isym)
{
  int i, j, k, m, ii, jj, kk, mm, ns2, ipaint, npaint, nneig, mk[6];
  ns2 = ix[0] * ix[1];
  for (kk = 0; kk < ix[2]; kk += (ix[2] - 1))
    for (jj = 0; jj < ix[1]; jj++)
      for (ii = 0; ii < ix[0]; ii++) sam[kk * ns2 + jj * ix[0] + ii] = 1;
  for (kk = 0; kk < ix[2]; kk++)
    for (jj = 0; jj < ix[1]; jj += (ix[1] - 1))
      for (ii = 0; ii < ix[0]; ii++) sam[kk * ns2 + jj * ix[0] + ii] = 1;
  for (kk = 0; kk < ix[2]; kk++)
    for (jj = 0; jj < ix[1]; jj++)
      for (ii = 0; ii < ix[0]; ii += (ix[0] - 1)) sam[kk * ns2 + jj * ix[0] + ii] = 1;
  npaint = ns2 * ix[2];
  while (npaint > 0) {
    npaint = 0;
    for (k = 1; k < ix[2] - 1; k++)
      for (j = 1; j < ix[1] - 1; j++)
        for (i = 1; i < ix[0] - 1; i++) {
          m = k * ns2 + j * ix[0] + i;
          if (sam[m] != 0) continue;
          if (isym == 1) {
            mk[0] = m - ns2; mk[1] = m + ns2; mk[2] = m - ix[0]; mk[3] = m + ix[0];
            nneig = 4;
          } else if (isym == 2) {
            mk[0] = m - ns2; mk[1] = m + ns2; mk[2] = m - 1; mk[3] = m + 1;
            nneig = 4;
          } else if (isym == 3) {
            mk[0] = m - ix[0]; mk[1] = m + ix[0]; mk[2] = m - 1; mk[3] = m + 1;
            nneig = 4;
          } else {
            mk[0] = m - ns2; mk[1] = m + ns2; mk[2] = m - ix[0];
            mk[3] = m + ix[0]; mk[4] = m - 1; mk[5] = m + 1;
            nneig = 6;
          }
          ipaint = 0;
          for (mm = 0; mm < nneig && ipaint == 0; mm++) ipaint = (sam[mk[mm]] == 1);
          if (ipaint) { sam[m] = 1; npaint++; }
        }
  }
  for (i = 0; i < ix[2] * ix[1] * ix[0]; i++) sam[i] = (sam[i] == 0 || sam[i] == 2);
}

/* utils.c:620-737 -- octree split above 20 points (each octant's points in
   reverse arrival order), brute-force pairs at leaves. */
static void uniq_tree(const double *coord, int *itag, double *rtag, const int *meshtag, int *ea, int nav)
{
  int i, j, ibox, p1, p2, icheck = 1;
  int npts[8], iv[3], cft[9];
  double xmax[3], xmin[3], xmid[3], dx[3], xp[3];
  if (nav > 20) {
    icheck = 0;
    xmin[0] = xmin[1] = xmin[2] = BIGV;
    xmax[0] = xmax[1] = xmax[2] = -BIGV;
    for (i = 0; i < nav; i++)
      for (j = 0; j < 3; j++) {
        xmin[j] = dmin(xmin[j], coord[3 * ea[i] + j]);
        xmax[j] = dmax(xmax[j], coord[3 * ea[i] + j]);
      }
    for (j = 0; j < 3; j++) {
      xmid[j] = (xmax[j] + xmin[j]) * 0.5;
      dx[j] = (xmax[j] - xmin[j]) * 0.5 + TOLV;
    }
    for (j = 0; j < 8; j++) npts[j] = 0;
    for (i = 0; i < nav; i++) {
      for (j = 0; j < 3; j++) {
        xp[j] = coord[3 * ea[i] + j] - xmid[j];
        iv[j] = (int)(floor(xp[j] / dx[j]) + 1);
      }
      ibox = 4 * iv[0] + 2 * iv[1] + iv[2];
      npts[ibox]++;
    }
    for (j = 0; j < 8; j++)
      if (npts[j] == nav) icheck = 1;
    if (!icheck) {
      int *tmp = (int *)malloc(sizeof(int) * nav);
      cft[0] = 0;
      for (j = 0; j < 8; j++) cft[j + 1] = cft[j] + npts[j];
      for (i = 0; i < nav; i++) {
        for (j = 0; j < 3; j++) {
          xp[j] = coord[3 * ea[i] + j] - xmid[j];
          iv[j] = (int)(floor(xp[j] / dx[j]) + 1);
        }
        ibox = 4 * iv[0] + 2 * iv[1] + iv[2];
        npts[ibox]--;
        tmp[npts[ibox] + cft[ibox]] = ea[i];
      }
      memcpy(ea, tmp, sizeof(int) * nav);
      free(tmp);
      for (j = 0; j < 8; j++)
        if (cft[j + 1] > cft[j]) uniq_tree(coord, itag, rtag, meshtag, ea + cft[j], cft[j + 1] - cft[j]);
    }
  }
  if (icheck) {
    for (i = 0; i < nav; i++) {
      p1 = ea[i];
      for (j = i + 1; j < nav; j++) {
        p2 = ea[j];
        if (fabs(coord[3 * p1] - coord[3 * p2]) + fabs(coord[3 * p1 + 1] - coord[3 * p2 + 1]) +
                    fabs(coord[3 * p1 + 2] - coord[3 * p2 + 2]) < TOLV &&
            meshtag[p1] == meshtag[p2]) {
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
}

void orc_unique_nodes_octree(const double *x, const int *meshtag, double *rtag, int *itag, int n)
{
  int i;
  int *ea = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (i = 0; i < n; i++) { ea[i] = i; itag[i] = i; }
  uniq_tree(x, itag, rtag, meshtag, ea, n);
  free(ea);
}

/* ===================================================================== */
/* ADT: ADT.C:34-135, buildADTrecursion.c:24-116                          */
/* ===================================================================== */
static void adt_rec(const double *coord, double *adtReals, double *adtWork, int *adtInt, int *ea,
                    int *cnt, int side, int parent, int level, int nav)
{
  const int ndim = 6, nd = 3;
  int i, j, ii, nleft, me;
  if (nav > 1) {
    int dimcut = level % ndim;
    for (i = 0; i < nav; i++) adtWork[i] = coord[ndim * ea[i] + dimcut];
    orc_median(ea, adtWork, nav);
    nleft = (nav + 1) / 2;
    (*cnt)++;
    me = *cnt;
    ii = me * 4;
    adtInt[ii] = ea[nleft - 1];
    adtInt[ii + 1] = adtInt[ii + 2] = adtInt[ii + 3] = -1;
    for (i = 0; i < nd; i++) {
      adtReals[ndim * me + i] = BIGV;
      adtReals[ndim * me + i + nd] = -BIGV;
    }
    for (i = 0; i < nav; i++)
      for (j = 0; j < nd; j++) {
        int a = ndim * me + j, b = ndim * ea[i] + j;
        adtReals[a] = dmin(adtReals[a], coord[b]);
        adtReals[a + nd] = dmax(adtReals[a + nd], coord[b + nd]);
      }
    if (side > 0) adtInt[4 * parent + side] = ea[nleft - 1];
    if (nleft > 1) adt_rec(coord, adtReals, adtWork, adtInt, ea, cnt, 1, me, level + 1, nleft - 1);
    adt_rec(coord, adtReals, adtWork, adtInt, ea + nleft, cnt, 2, me, level + 1, nav - nleft);
  } else if (nav == 1) {
    (*cnt)++;
    me = *cnt;
    adtInt[4 * me] = ea[0];
    adtInt[4 * me + 1] = adtInt[4 * me + 2] = adtInt[4 * me + 3] = -1;
    for (j = 0; j < ndim; j++) adtReals[ndim * me + j] = coord[ndim * ea[0] + j];
    if (side > 0) adtInt[4 * parent + side] = ea[0];
  }
}

void orc_build_adt(int nelem, const double *ebox, int *adtInt, double *adtReal, double ext[6])
{
  int i, j, cnt = -1;
  for (i = 0; i < 3; i++) { ext[2 * i] = BIGV; ext[2 * i + 1] = -BIGV; }
  for (j = 0; j < nelem; j++)
    for (i = 0; i < 3; i++) {
      ext[2 * i] = dmin(ext[2 * i], ebox[6 * j + i]);
      ext[2 * i + 1] = dmax(ext[2 * i + 1], ebox[6 * j + i + 3]);
    }
  for (i = 0; i < 3; i++) {
    double delta = 0.01 * (ext[2 * i + 1] - ext[2 * i]);
    ext[2 * i] -= delta;
    ext[2 * i + 1] += delta;
  }
  if (nelem <= 0) return;
  int *ea = (int *)malloc(sizeof(int) * nelem);
  double *work = (double *)malloc(sizeof(double) * nelem);
  for (i = 0; i < nelem; i++) ea[i] = i;
  adt_rec(ebox, adtReal, work, adtInt, ea, &cnt, 0, 0, 0, nelem);
  for (i = 0; i < nelem; i++) adtInt[4 * adtInt[4 * i] + 3] = i;
  free(ea);
  free(work);
}

/* ===================================================================== */
/* Per-block state (MeshBlock.h:39-354)                                  */
/* ===================================================================== */
typedef struct { double xc[3], dxc[3], vec[3][3]; } OrcOBB;

typedef struct DRec {
  int sender, meshtag, remoteid; /* donorData[0..2] (bookKeeping.C:135-137) */
  double donorRes, receptorRes;
  struct DRec *next;
} DRec;

typedef struct {
  int cancel, nweights, rinfo[3];
  int inode[9];
  double w[9];
} Interp;

typedef struct {
  int tag, nnodes, ncells, ntypes;
  int nv[8], nc[8];
  int *vconn[8];
  double *x;
  int *iblank;
  int nwbc, nobc;
  int *wbc, *obc;
  double *unres, *ucres;
  int *iblank_cell;
  int own_iblank_cell;
  int nfringe, mexclude;
  double resolutionScale;
  OrcOBB obb;
  double *cellRes, *nodeRes;
  int mapdims[3];
  double mapdx[3];
  int *icft, *invmap, *mapmask;
  int nsearch;
  double *xsearch, *res_search;
  int *isearch, *tagsearch, *donorId, *xtag;
  int nelem;
  double *ebox;
  int *elist, *adtInt;
  double *adtReal, adtExt[6];
  DRec **donorList;
  int donorListLen;
  Interp *interp;
  int ninterp;
  int *interp2donor;
  int *cancelList;
  int ncancel, cancelCap;
  double *q;
  /* global ids (registerGridData's cell_gid / node_gid, tioga.h:121-123) */
  const unsigned long long *cell_gid, *node_gid;
  /* receptors of the AMR Cartesian background (cartOps.C:61-160) */
  int ncart;
  int *receptorIdCart, *donorIdCart;
  double *rxyzCart;
} OrcBlock;

typedef struct { int existWall, nx[3]; int *sam; double ext[6]; } OrcHM;

typedef struct {
  int local, ob;
  int iblk_remote, tag_remote, send_tag, recv_tag;
  OrcOBB obc;
} OrcPair;

/* One interpolation record of a Cartesian patch: INTERPLIST2 as filled by
   CartBlock::insertInInterpList (CartBlock.C:196-268) for pdegree 0. */
typedef struct {
  int rinfo[3];
  int ijk[24];
  double w[8];
} AmrRec;

/* CartGrid (CartGrid.h:26-76) and the CartBlock array of one rank. */
typedef struct {
  int ngrids, nf, qstride, qnodesize, maxlevel;
  int *gid, *level, *proc, *porder, *lid, *ilo, *ihi, *dims;
  double *xlo, *dx, *qnode, xlosup[3], *dxlvl;
  int npatch;
  int *pgid, **pibl;
  double **pq;
  int *roff; /* npatch + 1 offsets of each patch's records in rec */
  AmrRec *rec;
  int nrec;
} OrcCart;

struct OrcTioga {
  int nblocks;
  OrcBlock *b[MAXBLK];
  int isym, nfringe, mexclude;
  int nmesh;
  OrcHM *hm;
  OrcPair *pairs;
  int npairs, nsend;
  OrcCart *cart;
  int amr;
};

static void cell_locate(const OrcBlock *b, int icell, int *n, int *i)
{
  int isum = 0, t;
  for (t = 0; t < b->ntypes; t++) {
    isum += b->nc[t];
    if (icell < isum) { *n = t; *i = icell - (isum - b->nc[t]); return; }
  }
  *n = b->ntypes - 1;
  *i = 0;
}

static OrcBlock *find_block(OrcTioga *t, int btag, int *iblk)
{
  int i;
  for (i = 0; i < t->nblocks; i++)
    if (t->b[i]->tag == btag) { if (iblk) *iblk = i; return t->b[i]; }
  return NULL;
}

OrcTioga *orc_create(void)
{
  OrcTioga *t = (OrcTioga *)calloc(1, sizeof(OrcTioga));
  t->isym = 3;   /* tioga.h:106 */
  t->nfringe = 1; /* tioga.h:107 */
  t->mexclude = 3;
  return t;
}

static void free_donor_lists(OrcBlock *b)
{
  int i;
  if (!b->donorList) return;
  for (i = 0; i < b->donorListLen; i++) {
    DRec *d = b->donorList[i];
    while (d) { DRec *n = d->next; free(d); d = n; }
  }
  free(b->donorList);
  b->donorList = NULL;
}

static void free_block(OrcBlock *b)
{
  free(b->cellRes); free(b->nodeRes); free(b->icft); free(b->invmap); free(b->mapmask);
  free(b->xsearch); free(b->res_search); free(b->isearch); free(b->tagsearch);
  free(b->donorId); free(b->xtag); free(b->ebox); free(b->elist); free(b->adtInt); free(b->adtReal);
  free_donor_lists(b);
  free(b->interp); free(b->interp2donor); free(b->cancelList);
  free(b->receptorIdCart); free(b->donorIdCart); free(b->rxyzCart);
  if (b->own_iblank_cell) free(b->iblank_cell);
  free(b);
}

static void free_cart(OrcCart *c)
{
  if (!c) return;
  free(c->gid); free(c->level); free(c->proc); free(c->porder); free(c->lid);
  free(c->ilo); free(c->ihi); free(c->dims); free(c->xlo); free(c->dx); free(c->qnode);
  free(c->dxlvl); free(c->pgid); free(c->pibl); free(c->pq); free(c->roff); free(c->rec);
  free(c);
}

void orc_destroy(OrcTioga *t)
{
  int i;
  if (!t) return;
  for (i = 0; i < t->nblocks; i++) free_block(t->b[i]);
  if (t->hm) {
    for (i = 0; i < t->nmesh; i++) free(t->hm[i].sam);
    free(t->hm);
  }
  free(t->pairs);
  free_cart(t->cart);
  free(t);
}

/* tioga.C:64-86 + MeshBlock.C:37-72 */
void orc_register_grid(OrcTioga *t, int btag, int nnodes, double *xyz, int *ibl, int nwbc, int nobc,
                       int *wbcnode, int *obcnode, int ntypes, const int *nv, const int *nc, int **vconn)
{
  int i;
  OrcBlock *b = find_block(t, btag, NULL);
  if (!b) {
    b = (OrcBlock *)calloc(1, sizeof(OrcBlock));
    b->nfringe = 1;
    b->mexclude = 3;
    b->resolutionScale = 1.0;
    t->b[t->nblocks++] = b;
  }
  b->tag = btag;
  b->nnodes = nnodes;
  b->x = xyz;
  b->iblank = ibl;
  b->nwbc = nwbc;
  b->nobc = nobc;
  b->wbc = wbcnode;
  b->obc = obcnode;
  b->ntypes = ntypes;
  b->ncells = 0;
  for (i = 0; i < ntypes; i++) {
    b->nv[i] = nv[i];
    b->nc[i] = nc[i];
    b->vconn[i] = vconn[i];
    b->ncells += nc[i];
  }
}

void orc_set_resolutions(OrcTioga *t, int btag, double *nres, double *cres)
{
  OrcBlock *b = find_block(t, btag, NULL);
  if (b) { b->unres = nres; b->ucres = cres; }
}

void orc_set_cell_iblank(OrcTioga *t, int btag, int *ibc)
{
  OrcBlock *b = find_block(t, btag, NULL);
  if (b) {
    if (b->own_iblank_cell) free(b->iblank_cell);
    b->own_iblank_cell = 0;
    b->iblank_cell = ibc;
  }
}

void orc_set_params(OrcTioga *t, int isym, int nfringe, int mexclude)
{
  t->isym = isym;
  t->nfringe = nfringe;
  t->mexclude = mexclude;
}

/* MeshBlock.C:94-339 */
static void tag_boundary(OrcBlock *b)
{
  int i, j, k, n, m, nvert, iex;
  int inode[8];
  double xv[8][3], vol;
  const int nn = b->nnodes;
  free(b->cellRes);
  free(b->nodeRes);
  b->cellRes = (double *)malloc(sizeof(double) * (b->ncells > 0 ? b->ncells : 1));
  b->nodeRes = (double *)malloc(sizeof(double) * (nn > 0 ? nn : 1));
  int *iflag = (int *)calloc(nn > 0 ? nn : 1, sizeof(int));
  int *iextmp = (int *)calloc(nn > 0 ? nn : 1, sizeof(int));
  int *iextmp1 = (int *)calloc(nn > 0 ? nn : 1, sizeof(int));
  if (b->unres == NULL && b->ucres == NULL) {
    for (i = 0; i < nn; i++) b->nodeRes[i] = 0.0;
    k = 0;
    for (n = 0; n < b->ntypes; n++) {
      nvert = b->nv[n];
      for (i = 0; i < b->nc[n]; i++) {
        for (m = 0; m < nvert; m++) {
          inode[m] = b->vconn[n][nvert * i + m] - ORC_BASE;
          for (j = 0; j < 3; j++) xv[m][j] = b->x[3 * inode[m] + j];
        }
        vol = orc_cell_volume((const double(*)[3])xv, nvert);
        b->cellRes[k++] = (vol * b->resolutionScale);
        for (m = 0; m < nvert; m++) {
          iflag[inode[m]]++;
          b->nodeRes[inode[m]] += (vol * b->resolutionScale);
        }
      }
    }
  } else {
    for (k = 0; k < b->ncells; k++) b->cellRes[k] = b->ucres[k];
    for (k = 0; k < nn; k++) b->nodeRes[k] = b->unres[k];
  }
  for (j = 0; j < 3; j++) {
    b->mapdims[j] = 10;
    b->mapdx[j] = 2 * b->obb.dxc[j] / b->mapdims[j];
  }
  const int nbins = b->mapdims[0] * b->mapdims[1] * b->mapdims[2];
  free(b->icft);
  free(b->invmap);
  b->icft = (int *)malloc(sizeof(int) * (nbins + 1));
  b->invmap = (int *)malloc(sizeof(int) * (nn > 0 ? nn : 1));
  for (i = 0; i < nbins + 1; i++) b->icft[i] = -1;
  b->icft[0] = 0;
  int *iptr = (int *)malloc(sizeof(int) * (nn > 0 ? nn : 1));
  for (i = 0; i < nn; i++) {
    double xd[3];
    int idx[3];
    if (iflag[i] != 0) b->nodeRes[i] /= iflag[i];
    iflag[i] = 0;
    iextmp[i] = iextmp1[i] = 0;
    for (j = 0; j < 3; j++) {
      xd[j] = b->obb.dxc[j];
      for (k = 0; k < 3; k++) xd[j] += (b->x[3 * i + k] - b->obb.xc[k]) * b->obb.vec[j][k];
      idx[j] = (int)(xd[j] / b->mapdx[j]);
    }
    int indx = idx[2] * b->mapdims[1] * b->mapdims[0] + idx[1] * b->mapdims[0] + idx[0];
    iptr[i] = b->icft[indx + 1];
    b->icft[indx + 1] = i;
  }
  int kc = 0;
  for (i = 0; i < nbins; i++) {
    int ip = b->icft[i + 1], cnt = 0;
    while (ip != -1) { b->invmap[kc++] = ip; ip = iptr[ip]; cnt++; }
    b->icft[i + 1] = b->icft[i] + cnt;
  }
  free(iptr);
  for (i = 0; i < b->nobc; i++) iflag[b->obc[i] - ORC_BASE] = 1;
  free(b->mapmask);
  b->mapmask = (int *)calloc(nbins, sizeof(int));
  for (n = 0; n < b->ntypes; n++) {
    nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      double xd[3], xmin[3], xmax[3];
      int itag = 0;
      for (j = 0; j < 3; j++) { xmin[j] = BIGV; xmax[j] = -BIGV; }
      for (m = 0; m < nvert; m++) {
        inode[m] = b->vconn[n][nvert * i + m] - ORC_BASE;
        if (iflag[inode[m]]) itag = 1;
        for (j = 0; j < 3; j++) {
          xd[j] = b->obb.dxc[j];
          for (k = 0; k < 3; k++) xd[j] += (b->x[3 * inode[m] + k] - b->obb.xc[k]) * b->obb.vec[j][k];
          xmin[j] = dmin(xd[j], xmin[j]);
          xmax[j] = dmax(xd[j], xmax[j]);
        }
      }
      for (j = 0; j < 3; j++) { xmin[j] -= TOLV; xmax[j] += TOLV; }
      for (int jj = (int)(xmin[0] / b->mapdx[0]); jj <= xmax[0] / b->mapdx[0]; jj++)
        for (int kk = (int)(xmin[1] / b->mapdx[1]); kk <= xmax[1] / b->mapdx[1]; kk++)
          for (int ll = (int)(xmin[2] / b->mapdx[2]); ll <= xmax[2] / b->mapdx[2]; ll++) {
            int i0 = jj < b->mapdims[0] - 1 ? jj : b->mapdims[0] - 1;
            int i1 = kk < b->mapdims[1] - 1 ? kk : b->mapdims[1] - 1;
            int i2 = ll < b->mapdims[2] - 1 ? ll : b->mapdims[2] - 1;
            if (i0 < 0) i0 = 0;
            if (i1 < 0) i1 = 0;
            if (i2 < 0) i2 = 0;
            b->mapmask[i2 * b->mapdims[1] * b->mapdims[0] + i1 * b->mapdims[0] + i0] = 1;
          }
      if (itag) {
        for (m = 0; m < nvert; m++) {
          b->nodeRes[inode[m]] = BIGV;
          iextmp[inode[m]] = iextmp1[inode[m]] = 1;
        }
      }
    }
  }
  for (iex = 0; iex < b->mexclude; iex++) {
    k = 0;
    for (n = 0; n < b->ntypes; n++) {
      nvert = b->nv[n];
      for (i = 0; i < b->nc[n]; i++) {
        for (m = 0; m < nvert; m++) {
          inode[m] = b->vconn[n][nvert * i + m] - ORC_BASE;
          if (iextmp[inode[m]] == 1) { b->cellRes[k] = BIGV; break; }
        }
        if (b->cellRes[k] == BIGV) {
          for (m = 0; m < nvert; m++) {
            inode[m] = b->vconn[n][nvert * i + m] - ORC_BASE;
            if (iextmp[inode[m]] != 1) iextmp1[inode[m]] = 1;
          }
        }
        k++;
      }
    }
    for (i = 0; i < nn; i++) iextmp[i] = iextmp1[i];
  }
  free(iflag);
  free(iextmp);
  free(iextmp1);
}

/* tioga.C:95-110 + MeshBlock.C:74-92 */
void orc_profile(OrcTioga *t)
{
  int ib, i;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    b->mexclude = t->mexclude;
    b->nfringe = t->nfringe;
    for (i = 0; i < b->nnodes; i++) b->iblank[i] = 1;
    orc_find_obb(b->x, b->obb.xc, b->obb.dxc, b->obb.vec, b->nnodes);
    tag_boundary(b);
  }
}

/* MeshBlock.C:637-664 */
static void wall_bounds(const OrcBlock *b, int *exist, double wbox[6])
{
  int i, j;
  if (b->nwbc <= 0) {
    *exist = 0;
    for (i = 0; i < 6; i++) wbox[i] = 0;
    return;
  }
  *exist = 1;
  wbox[0] = wbox[1] = wbox[2] = BIGV;
  wbox[3] = wbox[4] = wbox[5] = -BIGV;
  for (i = 0; i < b->nwbc; i++) {
    int i3 = 3 * (b->wbc[i] - ORC_BASE);
    for (j = 0; j < 3; j++) {
      wbox[j] = dmin(wbox[j], b->x[i3 + j]);
      wbox[j + 3] = dmax(wbox[j + 3], b->x[i3 + j]);
    }
  }
}

/* MeshBlock.C:666-772 */
static void mark_wall(const OrcBlock *b, int *sam, const int nx[3], const double ext[6])
{
  int i, j, k, m, n, nvert;
  int *cflag = (int *)calloc(b->ncells > 0 ? b->ncells : 1, sizeof(int));
  int *nflag = (int *)calloc(b->nnodes > 0 ? b->nnodes : 1, sizeof(int));
  double ds[3];
  for (i = 0; i < b->nwbc; i++) nflag[b->wbc[i] - ORC_BASE] = 1;
  m = 0;
  for (n = 0; n < b->ntypes; n++) {
    nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      for (j = 0; j < nvert; j++)
        if (nflag[b->vconn[n][nvert * i + j] - ORC_BASE] == 1) { cflag[m] = 1; break; }
      m++;
    }
  }
  for (k = 0; k < 3; k++) ds[k] = (ext[k + 3] - ext[k]) / nx[k];
  m = 0;
  for (n = 0; n < b->ntypes; n++) {
    nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      if (cflag[m] == 1) {
        int imin[3], imax[3];
        imin[0] = imin[1] = imin[2] = ORC_BIGINT;
        imax[0] = imax[1] = imax[2] = -ORC_BIGINT;
        for (j = 0; j < nvert; j++) {
          int i3 = 3 * (b->vconn[n][nvert * i + j] - ORC_BASE);
          for (k = 0; k < 3; k++) {
            int iv = (int)floor((b->x[i3 + k] - ext[k]) / ds[k]);
            if (iv < imin[k]) imin[k] = iv;
            if (iv > imax[k]) imax[k] = iv;
          }
        }
        for (j = 0; j < 3; j++) {
          if (imin[j] < 0) imin[j] = 0;
          if (imax[j] > nx[j] - 1) imax[j] = nx[j] - 1;
        }
        for (int kk = imin[2]; kk < imax[2] + 1; kk++)
          for (int jj = imin[1]; jj < imax[1] + 1; jj++)
            for (int ii = imin[0]; ii < imax[0] + 1; ii++) sam[kk * nx[1] * nx[0] + jj * nx[0] + ii] = 2;
      }
      m++;
    }
  }
  free(cflag);
  free(nflag);
}

/* holeMap.C:40-188 (single rank: the MPI reductions are identities) */
static void get_hole_map(OrcTioga *t)
{
  int i, j, ib;
  int maxtag = -ORC_BIGINT;
  for (ib = 0; ib < t->nblocks; ib++)
    if (t->b[ib]->tag > maxtag) maxtag = t->b[ib]->tag;
  if (maxtag < 1) maxtag = 1;
  if (t->hm) {
    for (i = 0; i < t->nmesh; i++) free(t->hm[i].sam);
    free(t->hm);
  }
  t->hm = (OrcHM *)calloc(maxtag, sizeof(OrcHM));
  double *bmin = (double *)malloc(sizeof(double) * 3 * maxtag);
  double *bmax = (double *)malloc(sizeof(double) * 3 * maxtag);
  for (i = 0; i < 3 * maxtag; i++) { bmin[i] = BIGV; bmax[i] = -BIGV; }
  for (ib = 0; ib < t->nblocks; ib++) {
    int ex;
    double wbox[6];
    wall_bounds(t->b[ib], &ex, wbox);
    int tg = t->b[ib]->tag;
    t->hm[tg - 1].existWall = ex;
    for (i = 0; i < 3; i++) {
      bmin[3 * (tg - 1) + i] = wbox[i];
      bmax[3 * (tg - 1) + i] = wbox[i + 3];
    }
  }
  for (i = 0; i < maxtag; i++) {
    OrcHM *h = &t->hm[i];
    if (!h->existWall) continue;
    double ds[3], dsmax, dsbox;
    for (j = 0; j < 3; j++) {
      h->ext[j] = bmin[3 * i + j];
      h->ext[j + 3] = bmax[3 * i + j];
      ds[j] = h->ext[j + 3] - h->ext[j];
    }
    dsmax = dmax(ds[0], ds[1]);
    dsmax = dmax(dsmax, ds[2]);
    dsbox = dsmax / ORC_HOLEMAPSIZE;
    for (j = 0; j < 3; j++) {
      h->ext[j] -= (2 * dsbox);
      h->ext[j + 3] += (2 * dsbox);
      h->nx[j] = (int)floor(dmax((h->ext[j + 3] - h->ext[j]) / dsbox, 1));
    }
    h->sam = (int *)calloc((size_t)h->nx[0] * h->nx[1] * h->nx[2], sizeof(int));
  }
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcHM *h = &t->hm[t->b[ib]->tag - 1];
    if (h->existWall) mark_wall(t->b[ib], h->sam, h->nx, h->ext);
  }
  t->nmesh = maxtag;
  for (i = 0; i < maxtag; i++)
    if (t->hm[i].existWall) orc_fill_hole_map(t->hm[i].sam, t->hm[i].nx, t->isym);
  free(bmin);
  free(bmax);
}

/* MeshBlock.C:774-838 -- AABB, in the block's own OBB frame, of its cells
   whose OBB-frame box (w.r.t. obc) overlaps obc; half extents x0.51. */
static void reduced_obb(const OrcBlock *b, const OrcOBB *obc, double rd[6])
{
  int i, j, m, n, nvert;
  double bbox[6], xd[3];
  for (j = 0; j < 3; j++) { rd[j] = BIGV; rd[j + 3] = -BIGV; }
  for (n = 0; n < b->ntypes; n++) {
    nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      bbox[0] = bbox[1] = bbox[2] = BIGV;
      bbox[3] = bbox[4] = bbox[5] = -BIGV;
      for (m = 0; m < nvert; m++) {
        const double *xp = &b->x[3 * (b->vconn[n][nvert * i + m] - ORC_BASE)];
        to_obb(xp, obc->xc, obc->vec, xd);
        for (j = 0; j < 3; j++) bbox[j] = dmin(bbox[j], xd[j]);
        for (j = 0; j < 3; j++) bbox[j + 3] = dmax(bbox[j + 3], xd[j]);
      }
      int fl = 0;
      for (j = 0; j < 3; j++) fl = (fl || (bbox[j] > obc->dxc[j]));
      if (fl) continue;
      fl = 0;
      for (j = 0; j < 3; j++) fl = (fl || (bbox[j + 3] < -obc->dxc[j]));
      if (fl) continue;
      for (m = 0; m < nvert; m++) {
        const double *xp = &b->x[3 * (b->vconn[n][nvert * i + m] - ORC_BASE)];
        to_obb(xp, b->obb.xc, b->obb.vec, xd);
        for (j = 0; j < 3; j++) rd[j] = dmin(rd[j], xd[j]);
        for (j = 0; j < 3; j++) rd[j + 3] = dmax(rd[j + 3], xd[j]);
      }
    }
  }
  for (j = 0; j < 6; j++) bbox[j] = rd[j];
  for (j = 0; j < 3; j++) {
    rd[j] = b->obb.xc[j];
    for (int k = 0; k < 3; k++) rd[j] += ((bbox[k] + bbox[k + 3]) * 0.5) * b->obb.vec[k][j];
    rd[j + 3] = (bbox[j + 3] - bbox[j]) * 0.51;
  }
}

/* exchangeBoxes.C:34-272, single rank (obbRecv[ob] is block ob's OBB, one
   packet to self; the send/recv keys pair (A,B) with (B,A)). */
static void exchange_boxes(OrcTioga *t)
{
  int ob, ib, i, maxtag = -1;
  for (ib = 0; ib < t->nblocks; ib++) {
    int tg = abs(t->b[ib]->tag);
    if (maxtag < tg) maxtag = tg;
  }
  free(t->pairs);
  t->pairs = (OrcPair *)calloc((size_t)t->nblocks * t->nblocks + 1, sizeof(OrcPair));
  t->npairs = 0;
  for (ob = 0; ob < t->nblocks; ob++)
    for (ib = 0; ib < t->nblocks; ib++) {
      OrcBlock *A = t->b[ib], *B = t->b[ob];
      if (abs(B->tag) == A->tag) continue;
      if (orc_obb_intersect((const double(*)[3])A->obb.vec, A->obb.xc, A->obb.dxc,
                            (const double(*)[3])B->obb.vec, B->obb.xc, B->obb.dxc) ||
          orc_obb_intersect((const double(*)[3])B->obb.vec, B->obb.xc, B->obb.dxc,
                            (const double(*)[3])A->obb.vec, A->obb.xc, A->obb.dxc)) {
        OrcPair *p = &t->pairs[t->npairs++];
        p->local = ib;
        p->ob = ob;
      }
    }
  t->nsend = t->npairs > 0 ? 1 : 0;
  /* sender side: reduced OBB of block ib's cells w.r.t. obbRecv[ob] */
  double *rd = (double *)malloc(sizeof(double) * 6 * (t->npairs + 1));
  for (i = 0; i < t->npairs; i++) {
    OrcPair *p = &t->pairs[i];
    OrcBlock *A = t->b[p->local], *B = t->b[p->ob];
    p->recv_tag = maxtag * (A->tag - 1) + B->tag - 1;
    p->send_tag = maxtag * (B->tag - 1) + (A->tag - 1);
    reduced_obb(A, &B->obb, &rd[6 * i]);
    memcpy(p->obc.vec, B->obb.vec, sizeof(double) * 9);
  }
  /* receiver side: section i lands on the pair whose recv key is i's send key */
  for (i = 0; i < t->npairs; i++) {
    int key = t->pairs[i].send_tag, ii = -1, j;
    for (j = 0; j < t->npairs; j++)
      if (t->pairs[j].recv_tag == key) ii = j; /* std::map keeps the last insert */
    if (ii < 0) continue;
    t->pairs[ii].iblk_remote = t->pairs[i].local;
    t->pairs[ii].tag_remote = t->b[t->pairs[i].local]->tag;
    for (j = 0; j < 3; j++) {
      t->pairs[ii].obc.xc[j] = rd[6 * i + j];
      t->pairs[ii].obc.dxc[j] = rd[6 * i + 3 + j];
    }
  }
  free(rd);
}

/* MeshBlock.C:953-1117 -- own nodes inside obc, visited in 10^3-bin order */
static int query_points(const OrcBlock *b, const OrcOBB *obc, int *inode)
{
  int j, k, l, m, n, ij, nints = 0;
  double xv[8][3], xd[3], xc[3], mdx[3], xmin[3], xmax[3];
  int imin[3], imax[3];
  obb_corners(obc->xc, obc->dxc, (const double(*)[3])obc->vec, xv);
  for (j = 0; j < 3; j++) { xmin[j] = BIGV; xmax[j] = -BIGV; }
  for (n = 0; n < 8; n++) {
    to_obb(xv[n], b->obb.xc, (const double(*)[3])b->obb.vec, xd);
    for (j = 0; j < 3; j++) {
      xmin[j] = dmin(xmin[j], xd[j] + b->obb.dxc[j]);
      xmax[j] = dmax(xmax[j], xd[j] + b->obb.dxc[j]);
    }
  }
  for (j = 0; j < 3; j++) {
    double delta = 0.01 * (xmax[j] - xmin[j]);
    xmin[j] -= delta;
    xmax[j] += delta;
    imin[j] = (int)dmax(xmin[j] / b->mapdx[j], 0);
    imax[j] = (int)dmin(xmax[j] / b->mapdx[j], b->mapdims[j] - 1);
    mdx[j] = 0.5 * b->mapdx[j];
  }
  obb_corners(obc->xc, mdx, (const double(*)[3])b->obb.vec, xv);
  for (j = 0; j < 3; j++) { xmin[j] = BIGV; xmax[j] = -BIGV; }
  for (m = 0; m < 8; m++) {
    to_obb(xv[m], obc->xc, (const double(*)[3])obc->vec, xd);
    for (j = 0; j < 3; j++) {
      xmin[j] = dmin(xmin[j], xd[j]);
      xmax[j] = dmax(xmax[j], xd[j]);
    }
  }
  for (l = imin[2]; l <= imax[2]; l++)
    for (k = imin[1]; k <= imax[1]; k++)
      for (j = imin[0]; j <= imax[0]; j++) {
        xd[0] = -b->obb.dxc[0] + j * b->mapdx[0] + b->mapdx[0] * 0.5;
        xd[1] = -b->obb.dxc[1] + k * b->mapdx[1] + b->mapdx[1] * 0.5;
        xd[2] = -b->obb.dxc[2] + l * b->mapdx[2] + b->mapdx[2] * 0.5;
        for (n = 0; n < 3; n++) {
          xc[n] = b->obb.xc[n];
          for (ij = 0; ij < 3; ij++) xc[n] += (xd[ij] * b->obb.vec[ij][n]);
        }
        to_obb(xc, obc->xc, (const double(*)[3])obc->vec, xd);
        int fl = 0;
        for (ij = 0; ij < 3 && !fl; ij++) fl = (fl || (xmin[ij] + xd[ij] > obc->dxc[ij]));
        if (fl) continue;
        fl = 0;
        for (ij = 0; ij < 3 && !fl; ij++) fl = (fl || (xmax[ij] + xd[ij] < -obc->dxc[ij]));
        if (fl) continue;
        int indx = l * b->mapdims[1] * b->mapdims[0] + k * b->mapdims[0] + j;
        for (m = b->icft[indx]; m < b->icft[indx + 1]; m++) {
          const double *xp = &b->x[3 * b->invmap[m]];
          to_obb(xp, obc->xc, (const double(*)[3])obc->vec, xd);
          if (fabs(xd[0]) <= obc->dxc[0] && fabs(xd[1]) <= obc->dxc[1] && fabs(xd[2]) <= obc->dxc[2])
            inode[nints++] = b->invmap[m];
        }
      }
  return nints;
}

/* exchangeSearchData.C:27-263, single rank */
static void exchange_search_data(OrcTioga *t)
{
  int ib, i, j;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    b->nsearch = 0;
    free(b->xsearch); b->xsearch = NULL;
    free(b->isearch); b->isearch = NULL;
    free(b->tagsearch); b->tagsearch = NULL;
    free(b->donorId); b->donorId = NULL;
    free(b->res_search); b->res_search = NULL;
  }
  if (t->nsend == 0) return;
  int **lists = (int **)calloc(t->npairs + 1, sizeof(int *));
  int *cnt = (int *)calloc(t->npairs + 1, sizeof(int));
  int *dest = (int *)calloc(t->npairs + 1, sizeof(int));
  for (i = 0; i < t->npairs; i++) {
    OrcBlock *A = t->b[t->pairs[i].local];
    lists[i] = (int *)malloc(sizeof(int) * (A->nnodes > 0 ? A->nnodes : 1));
    cnt[i] = query_points(A, &t->pairs[i].obc, lists[i]);
  }
  for (i = 0; i < t->npairs; i++) {
    int key = t->pairs[i].send_tag, ii = -1;
    for (j = 0; j < t->npairs; j++)
      if (t->pairs[j].recv_tag == key) ii = j;
    dest[i] = ii;
    if (ii >= 0) t->b[t->pairs[ii].local]->nsearch += cnt[i];
  }
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    if (b->nsearch < 1) continue;
    b->xsearch = (double *)malloc(sizeof(double) * 3 * b->nsearch);
    b->res_search = (double *)malloc(sizeof(double) * b->nsearch);
    b->isearch = (int *)malloc(sizeof(int) * 3 * b->nsearch);
    b->tagsearch = (int *)malloc(sizeof(int) * b->nsearch);
    b->donorId = (int *)malloc(sizeof(int) * b->nsearch);
  }
  int *off = (int *)calloc(t->nblocks, sizeof(int));
  for (i = 0; i < t->npairs; i++) {
    int ii = dest[i];
    if (ii < 0) continue;
    OrcBlock *A = t->b[t->pairs[i].local];
    int rb = t->pairs[ii].local;
    OrcBlock *R = t->b[rb];
    for (j = 0; j < cnt[i]; j++) {
      int o = off[rb]++;
      int nd = lists[i][j];
      R->isearch[3 * o] = 0;
      R->isearch[3 * o + 1] = nd;
      R->isearch[3 * o + 2] = t->pairs[ii].iblk_remote;
      R->tagsearch[o] = t->pairs[ii].tag_remote;
      R->xsearch[3 * o] = A->x[3 * nd];
      R->xsearch[3 * o + 1] = A->x[3 * nd + 1];
      R->xsearch[3 * o + 2] = A->x[3 * nd + 2];
      R->res_search[o] = A->nodeRes[nd];
    }
  }
  for (i = 0; i < t->npairs; i++) free(lists[i]);
  free(lists);
  free(cnt);
  free(dest);
  free(off);
}

/* checkContainment.C:27-96 (ihigh==0 branch) */
static void check_containment(const OrcBlock *b, int cidx[2], int adtElement, const double *xs)
{
  int n, i, m, j, nvert;
  double xv[8][3], frac[8];
  int icell = b->elist[adtElement];
  cell_locate(b, icell, &n, &i);
  nvert = b->nv[n];
  for (m = 0; m < nvert; m++) {
    int i3 = 3 * (b->vconn[n][nvert * i + m] - ORC_BASE);
    for (j = 0; j < 3; j++) xv[m][j] = b->x[i3 + j];
  }
  orc_nodal_weights((const double(*)[3])xv, xs, frac, nvert);
  cidx[0] = icell;
  cidx[1] = 0;
  for (m = 0; m < nvert; m++) {
    if ((frac[m] + TOLV) * (frac[m] - 1.0 - TOLV) > 0) { cidx[0] = -1; return; }
    if (fabs(frac[m]) < TOLV && b->cellRes[icell] == BIGV) cidx[1] = 1;
  }
}

/* searchADTrecursion.C:52-101 */
static void search_rec(const OrcBlock *b, int cidx[2], int node, const double *xs)
{
  int i, d, flag;
  const double *el = &b->ebox[6 * b->adtInt[4 * node]];
  flag = 1;
  for (i = 0; i < 3; i++) flag = (flag && (xs[i] >= el[i] - TOLV));
  for (i = 3; i < 6; i++) flag = (flag && (xs[i - 3] <= el[i] + TOLV));
  if (flag) {
    check_containment(b, cidx, b->adtInt[4 * node], xs);
    if (cidx[0] > -1 && cidx[1] == 0) return;
  }
  for (d = 1; d < 3; d++) {
    int child = b->adtInt[4 * node + d];
    if (child > -1) {
      child = b->adtInt[4 * child + 3];
      const double *bx = &b->adtReal[6 * child];
      flag = 1;
      for (i = 0; i < 3; i++) flag = (flag && (xs[i] >= bx[i] - TOLV));
      for (i = 3; i < 6; i++) flag = (flag && (xs[i - 3] <= bx[i] + TOLV));
      if (flag) {
        search_rec(b, cidx, child, xs);
        if (cidx[0] > -1 && cidx[1] == 0) return;
      }
    }
  }
}

/* searchADTrecursion.C:26-50 */
static int search_adt(const OrcBlock *b, const double *xs)
{
  int i, flag, cidx[2] = {-1, 0};
  if (b->nelem <= 0) return -1;
  flag = 1;
  for (i = 0; i < 3; i++) flag = (flag && (xs[i] >= b->adtExt[2 * i] - TOLV));
  for (i = 0; i < 3; i++) flag = (flag && (xs[i] <= b->adtExt[2 * i + 1] + TOLV));
  if (flag) search_rec(b, cidx, 0, xs);
  return cidx[0];
}

/* search.C:107-233 -- cells whose query-OBB-frame box overlaps the query
   OBB, as a LIFO list (descending cell index), their AABBs, and the ADT. */
static void build_block_adt(OrcBlock *b, const double *xq, int nq)
{
  int n, i, m, j, k, p, nvert;
  OrcOBB obq;
  double xd[3], dxc[3], xmin[3], xmax[3];
  orc_find_obb(xq, obq.xc, obq.dxc, obq.vec, nq);
  int *icell = (int *)malloc(sizeof(int) * (b->ncells > 0 ? b->ncells : 1));
  for (i = 0; i < b->ncells; i++) icell[i] = -1;
  int iptr = -1, count = 0;
  p = 0;
  for (n = 0; n < b->ntypes; n++) {
    nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      xmin[0] = xmin[1] = xmin[2] = BIGV;
      xmax[0] = xmax[1] = xmax[2] = -BIGV;
      for (m = 0; m < nvert; m++) {
        int i3 = 3 * (b->vconn[n][nvert * i + m] - ORC_BASE);
        for (j = 0; j < 3; j++) {
          xd[j] = 0;
          for (k = 0; k < 3; k++) xd[j] += (b->x[i3 + k] - obq.xc[k]) * obq.vec[j][k];
          xmin[j] = dmin(xmin[j], xd[j]);
          xmax[j] = dmax(xmax[j], xd[j]);
        }
      }
      for (j = 0; j < 3; j++) {
        xd[j] = (xmax[j] + xmin[j]) * 0.5;
        dxc[j] = (xmax[j] - xmin[j]) * 0.5;
      }
      if (fabs(xd[0]) <= (dxc[0] + obq.dxc[0]) && fabs(xd[1]) <= (dxc[1] + obq.dxc[1]) &&
          fabs(xd[2]) <= (dxc[2] + obq.dxc[2])) {
        icell[p] = iptr;
        iptr = p;
        count++;
      }
      p++;
    }
  }
  free(b->ebox);
  free(b->elist);
  free(b->adtInt);
  free(b->adtReal);
  b->nelem = count;
  b->ebox = (double *)malloc(sizeof(double) * 6 * (count > 0 ? count : 1));
  b->elist = (int *)malloc(sizeof(int) * (count > 0 ? count : 1));
  b->adtInt = (int *)malloc(sizeof(int) * 4 * (count > 0 ? count : 1));
  b->adtReal = (double *)malloc(sizeof(double) * 6 * (count > 0 ? count : 1));
  k = iptr;
  p = 0;
  while (k != -1) {
    int tn, ti;
    cell_locate(b, k, &tn, &ti);
    nvert = b->nv[tn];
    xmin[0] = xmin[1] = xmin[2] = BIGV;
    xmax[0] = xmax[1] = xmax[2] = -BIGV;
    for (m = 0; m < nvert; m++) {
      int i3 = 3 * (b->vconn[tn][nvert * ti + m] - ORC_BASE);
      for (j = 0; j < 3; j++) {
        xmin[j] = dmin(xmin[j], b->x[i3 + j]);
        xmax[j] = dmax(xmax[j], b->x[i3 + j]);
      }
    }
    for (j = 0; j < 3; j++) { b->ebox[6 * p + j] = xmin[j]; b->ebox[6 * p + 3 + j] = xmax[j]; }
    b->elist[p++] = k;
    k = icell[k];
  }
  orc_build_adt(b->nelem, b->ebox, b->adtInt, b->adtReal, b->adtExt);
  free(icell);
}

/* search.C:78-269 (NODEGID off: octree dedupe) */
/* search.C:48-74 (uniquenode_map): the first query carrying each global id
   represents all that carry it; the representative's resolution becomes
   the max over its copies and every copy takes it.  First occurrence in
   query order, found by an open-addressing table of the ids (any exact set
   lookup gives the same first occurrences). */
void orc_unique_nodes_gid(const unsigned long long *gid, double *res, int *itag, int n)
{
  int i;
  size_t cap = 16;
  while (cap < 2 * (size_t)(n > 0 ? n : 1)) cap <<= 1;
  int *slot = (int *)malloc(sizeof(int) * cap);
  for (size_t k = 0; k < cap; k++) slot[k] = -1;
  for (i = 0; i < n; i++) {
    unsigned long long g = gid[i];
    size_t h = (size_t)((g * 0x9E3779B97F4A7C15ULL) >> 17) & (cap - 1);
    while (slot[h] >= 0 && gid[slot[h]] != g) h = (h + 1) & (cap - 1);
    if (slot[h] >= 0) {
      int f = slot[h];
      itag[i] = f;
      if (res[f] < res[i]) res[f] = res[i]; /* std::max(res[f], res[i]) */
    } else {
      slot[h] = i;
      itag[i] = i;
    }
  }
  for (i = 0; i < n; i++) res[i] = res[itag[i]];
  free(slot);
}

/* the node-id mode of a single-rank run: the reference's TIOGA_HAS_NODEGID
   build, which every block's node ids select here */
static int gid_mode(const OrcTioga *t)
{
  int ib;
  if (t->nblocks == 0) return 0;
  for (ib = 0; ib < t->nblocks; ib++)
    if (!t->b[ib]->node_gid) return 0;
  return 1;
}

static void block_search(OrcTioga *t, OrcBlock *b)
{
  int i;
  if (b->nsearch == 0) return;
  build_block_adt(b, b->xsearch, b->nsearch);
  free(b->donorId);
  free(b->xtag);
  b->donorId = (int *)malloc(sizeof(int) * b->nsearch);
  b->xtag = (int *)malloc(sizeof(int) * b->nsearch);
  if (gid_mode(t)) {
    /* gid_search: the sender's node id of every query (exchangeSearchData.C:191-231) */
    unsigned long long *gs = (unsigned long long *)malloc(sizeof(unsigned long long) * b->nsearch);
    for (i = 0; i < b->nsearch; i++) gs[i] = t->b[b->isearch[3 * i + 2]]->node_gid[b->isearch[3 * i + 1]];
    orc_unique_nodes_gid(gs, b->res_search, b->xtag, b->nsearch);
    free(gs);
  } else {
    orc_unique_nodes_octree(b->xsearch, b->tagsearch, b->res_search, b->xtag, b->nsearch);
  }
  for (i = 0; i < b->nsearch; i++) {
    if (b->xtag[i] == i) b->donorId[i] = search_adt(b, &b->xsearch[3 * i]);
    else b->donorId[i] = b->donorId[b->xtag[i]];
  }
}

/* linklist.c:62-104 -- insert before the first entry whose |donorRes| is
   strictly larger, so equal keys keep arrival order. */
static void insert_in_list(DRec **head, DRec *r)
{
  DRec *cur = *head, *prev = NULL;
  while (cur != NULL) {
    if (fabs(cur->donorRes) > r->donorRes) {
      r->next = cur;
      if (prev) prev->next = r;
      else *head = r;
      return;
    }
    prev = cur;
    cur = cur->next;
  }
  r->next = NULL;
  if (prev) prev->next = r;
  else *head = r;
}

typedef struct { int k, remoteid, tag; double rres; } Rec3;

/* bookKeeping.C:143-355 */
static int process_donors(OrcBlock *b, OrcHM *hm, int nmesh, Rec3 **out)
{
  int i, j, n, m, mm, nvert, iter;
  int *iflag = (int *)malloc(sizeof(int) * (nmesh > 0 ? nmesh : 1));
  const int nn = b->nnodes;
  for (i = 0; i < nn; i++) {
    b->iblank[i] = 1;
    if (b->donorList[i] == NULL) {
      for (j = 0; j < nmesh; j++)
        if (j != (b->tag - ORC_BASE) && hm[j].existWall) {
          if (orc_check_hole_map(&b->x[3 * i], hm[j].nx, hm[j].sam, hm[j].ext)) { b->iblank[i] = 0; break; }
        }
    } else {
      DRec *d = b->donorList[i];
      for (j = 0; j < nmesh; j++) iflag[j] = 0;
      while (d != NULL) {
        iflag[d->meshtag - ORC_BASE] = 1;
        b->nodeRes[i] = dmax(b->nodeRes[i], d->receptorRes);
        d = d->next;
      }
      for (j = 0; j < nmesh; j++)
        if (j != (b->tag - ORC_BASE) && hm[j].existWall && !iflag[j]) {
          if (orc_check_hole_map(&b->x[3 * i], hm[j].nx, hm[j].sam, hm[j].ext)) { b->iblank[i] = 0; break; }
        }
    }
  }
  int *mtag = (int *)malloc(sizeof(int) * (nn > 0 ? nn : 1));
  int *mtag1 = (int *)malloc(sizeof(int) * (nn > 0 ? nn : 1));
  for (i = 0; i < nn; i++) mtag[i] = mtag1[i] = (b->iblank[i] == 0) ? 1 : 0;
  for (iter = 0; iter < b->nfringe; iter++) {
    for (n = 0; n < b->ntypes; n++) {
      nvert = b->nv[n];
      for (i = 0; i < b->nc[n]; i++) {
        const int *cv = &b->vconn[n][nvert * i];
        for (m = 0; m < nvert; m++)
          if (mtag[cv[m] - ORC_BASE] == 1)
            for (mm = 0; mm < nvert; mm++)
              if (m != mm && mtag[cv[mm] - ORC_BASE] != 1) mtag1[cv[mm] - ORC_BASE] = 1;
      }
    }
    for (i = 0; i < nn; i++) mtag[i] = mtag1[i];
  }
  for (i = 0; i < nn; i++)
    if (mtag1[i] && b->iblank[i]) b->nodeRes[i] = BIGV;
  free(mtag);
  free(mtag1);
  int nrec = 0;
  for (i = 0; i < nn; i++) {
    if (b->donorList[i] != NULL && b->iblank[i] != 0) {
      DRec *d = b->donorList[i];
      while (d != NULL) {
        if (d->donorRes < b->nodeRes[i]) { b->iblank[i] = -d->meshtag; nrec++; break; }
        d = d->next;
      }
    }
  }
  *out = (Rec3 *)malloc(sizeof(Rec3) * (nrec > 0 ? nrec : 1));
  int kk = 0;
  for (i = 0; i < nn; i++) {
    if (b->iblank[i] < 0) {
      DRec *d = b->donorList[i]; /* the head: the list is sorted ascending */
      (*out)[kk].rres = (b->resolutionScale > 1.0) ? -b->nodeRes[i] : b->nodeRes[i];
      (*out)[kk].k = d->sender;
      (*out)[kk].remoteid = d->remoteid;
      (*out)[kk].tag = d->meshtag;
      kk++;
    }
  }
  free(iflag);
  return nrec;
}

/* bookKeeping.C:357-383 */
static void init_interp_list(OrcBlock *b, int n)
{
  int i;
  free(b->interp);
  b->interp = (Interp *)calloc(n > 0 ? n : 1, sizeof(Interp));
  b->ninterp = n;
  free(b->cancelList);
  b->cancelList = NULL;
  b->ncancel = b->cancelCap = 0;
  free(b->interp2donor);
  b->interp2donor = (int *)malloc(sizeof(int) * (b->nsearch > 0 ? b->nsearch : 1));
  for (i = 0; i < b->nsearch; i++) b->interp2donor[i] = -1;
}

static void push_cancel(OrcBlock *b, int inode)
{
  if (b->ncancel == b->cancelCap) {
    b->cancelCap = b->cancelCap ? 2 * b->cancelCap : 64;
    b->cancelList = (int *)realloc(b->cancelList, sizeof(int) * b->cancelCap);
  }
  b->cancelList[b->ncancel++] = inode;
}

/* bookKeeping.C:385-535 */
static void find_interp_data(OrcBlock *b, int *recid, int irecord, double receptorRes2)
{
  int i, j, m, n, nvert, acceptFlag = 1;
  double xv[8][3], xp[3], frac[8];
  int inode[8];
  double receptorRes = fabs(receptorRes2);
  int procid = b->isearch[3 * irecord];
  int pointid = b->isearch[3 * irecord + 1];
  int blockid = b->isearch[3 * irecord + 2];
  int meshtagrecv = b->tagsearch[irecord];
  xp[0] = b->xsearch[3 * irecord];
  xp[1] = b->xsearch[3 * irecord + 1];
  xp[2] = b->xsearch[3 * irecord + 2];
  cell_locate(b, b->donorId[irecord], &n, &i);
  nvert = b->nv[n];
  for (m = 0; m < nvert; m++) {
    inode[m] = b->vconn[n][nvert * i + m] - ORC_BASE;
    if (b->iblank[inode[m]] <= 0 && receptorRes2 > 0.0) {
      if (b->nodeRes[inode[m]] == BIGV) acceptFlag = 0;
      if (abs(b->iblank[inode[m]]) == meshtagrecv) acceptFlag = 0;
    }
    for (j = 0; j < 3; j++) xv[m][j] = b->x[3 * inode[m] + j];
  }
  if (receptorRes == BIGV && b->resolutionScale == 1.0) {
    for (m = 0; m < nvert; m++) {
      if (b->iblank[inode[m]] <= 0 && b->nodeRes[inode[m]] != BIGV) {
        if (b->iblank[inode[m]] < 0) b->iblank[inode[m]] = 1;
        push_cancel(b, inode[m]);
      }
    }
  }
  orc_nodal_weights((const double(*)[3])xv, xp, frac, nvert);
  b->interp2donor[irecord] = *recid;
  Interp *it = &b->interp[*recid];
  it->cancel = 0;
  it->nweights = nvert;
  it->rinfo[0] = procid;
  it->rinfo[1] = pointid;
  it->rinfo[2] = blockid;
  for (m = 0; m < nvert; m++) {
    it->inode[m] = inode[m];
    it->w[m] = frac[m];
    if (frac[m] < -0.2 || frac[m] > 1.2) g_weight_aborts++; /* the reference calls MPI_Abort here */
  }
  it->inode[m] = b->donorId[irecord];
  it->w[m] = 0.0;
  if (acceptFlag == 0 && receptorRes != BIGV) it->cancel = 1;
  (*recid)++;
}

/* bookKeeping.C:572-597 */
static void reset_coincident(OrcBlock *b)
{
  int i, iptr;
  if (b->nsearch <= 0 || b->interp2donor == NULL) return;
  int *ireset = (int *)malloc(sizeof(int) * b->nsearch);
  for (i = 0; i < b->nsearch; i++) ireset[i] = 1;
  for (i = 0; i < b->nsearch; i++) {
    iptr = b->interp2donor[i];
    if (iptr > -1) {
      int c = b->interp[iptr].cancel;
      if (c < ireset[b->xtag[i]]) ireset[b->xtag[i]] = c;
    }
  }
  for (i = 0; i < b->nsearch; i++) {
    iptr = b->interp2donor[i];
    if (iptr > -1 && b->interp[iptr].cancel == 1) b->interp[iptr].cancel = ireset[b->xtag[i]];
  }
  free(ireset);
}

/* exchangeDonors.C:25-301, single rank (k is always 0) */
static void exchange_donors(OrcTioga *t)
{
  int ib, i;
  if (t->nsend == 0 || t->nblocks <= 0) return;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    free_donor_lists(b);
    b->donorListLen = b->nnodes;
    b->donorList = (DRec **)calloc(b->nnodes > 0 ? b->nnodes : 1, sizeof(DRec *));
  }
  /* comm1: donor blocks in order, search points in order (bookKeeping.C:89-110) */
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *D = t->b[ib];
    for (i = 0; i < D->nsearch; i++) {
      if (D->donorId[i] < 0) continue;
      OrcBlock *R = t->b[D->isearch[3 * i + 2]];
      DRec *r = (DRec *)malloc(sizeof(DRec));
      r->sender = 0;
      r->meshtag = D->tag;
      r->remoteid = i;
      r->donorRes = D->cellRes[D->donorId[i]];
      r->receptorRes = D->res_search[D->xtag[i]];
      r->next = NULL;
      insert_in_list(&R->donorList[D->isearch[3 * i + 1]], r);
    }
  }
  Rec3 **recs = (Rec3 **)calloc(t->nblocks, sizeof(Rec3 *));
  int *nrec = (int *)calloc(t->nblocks, sizeof(int));
  for (ib = 0; ib < t->nblocks; ib++) nrec[ib] = process_donors(t->b[ib], t->hm, t->nmesh, &recs[ib]);
  /* comm2 */
  int *nint = (int *)calloc((size_t)(t->nblocks > 0 ? t->nblocks : 1), sizeof(int));
  for (ib = 0; ib < t->nblocks; ib++)
    for (i = 0; i < nrec[ib]; i++) {
      int db = 0;
      find_block(t, recs[ib][i].tag, &db);
      nint[db]++;
    }
  for (ib = 0; ib < t->nblocks; ib++) init_interp_list(t->b[ib], nint[ib]);
  memset(nint, 0, sizeof(int) * t->nblocks);
  for (ib = 0; ib < t->nblocks; ib++)
    for (i = 0; i < nrec[ib]; i++) {
      int db = 0;
      find_block(t, recs[ib][i].tag, &db);
      find_interp_data(t->b[db], &nint[db], recs[ib][i].remoteid, recs[ib][i].rres);
    }
  for (ib = 0; ib < t->nblocks; ib++) t->b[ib]->ninterp = nint[ib];
  /* comm3: cancellations (bookKeeping.C:542-570) */
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ncancel; i++) {
      DRec *h = b->donorList[b->cancelList[i]];
      if (h == NULL) continue;
      int db = 0;
      OrcBlock *D = find_block(t, h->meshtag, &db);
      if (!D) continue;
      int iptr = D->interp2donor[h->remoteid];
      if (iptr > -1) D->interp[iptr].cancel = 1;
    }
  }
  for (ib = 0; ib < t->nblocks; ib++) reset_coincident(t->b[ib]);
  /* comm4: clearIblanks then setIblanks (bookKeeping.C:598-648) */
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->nnodes; i++)
      if (b->iblank[i] < 0) b->iblank[i] = 1;
  }
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ninterp; i++)
      if (!b->interp[i].cancel) t->b[b->interp[i].rinfo[2]]->iblank[b->interp[i].rinfo[1]] = -1;
  }
  for (ib = 0; ib < t->nblocks; ib++) free(recs[ib]);
  free(recs);
  free(nrec);
  free(nint);
}

/* MeshBlock::reduce_fringes (bookKeeping.C:650-760): iblank_reduced keeps the
   field nodes and, over nfringe + 1 sweeps, the receptors of cells that are
   partly field or receptor; every sweep reads the previous sweep's values
   (ibltmp) and only turns 0 into the node's negative iblank.  The receptors
   it drops (iblank < 0, iblank_reduced == 0) form the new cancel list, in
   node order.  Returns iblank_reduced (caller frees). */
static int *block_reduce_fringes(OrcBlock *b)
{
  int i, n, m, iter;
  const int nn = b->nnodes > 0 ? b->nnodes : 1;
  int *ibr = (int *)malloc(sizeof(int) * nn), *tmp = (int *)malloc(sizeof(int) * nn);
  for (i = 0; i < b->nnodes; i++) ibr[i] = b->iblank[i] > 0 ? b->iblank[i] : 0;
  for (i = 0; i < b->nnodes; i++) tmp[i] = ibr[i];
  for (iter = 0; iter < b->nfringe + 1; iter++) {
    for (n = 0; n < b->ntypes; n++) {
      const int nvert = b->nv[n];
      for (i = 0; i < b->nc[n]; i++) {
        int inode[8], ncount = 0;
        for (m = 0; m < nvert; m++) {
          inode[m] = b->vconn[n][nvert * i + m] - ORC_BASE;
          ncount += (ibr[inode[m]] == 1 || ibr[inode[m]] < 0);
        }
        if (ncount > 0 && ncount < nvert)
          for (m = 0; m < nvert; m++)
            if (tmp[inode[m]] == 0 && b->iblank[inode[m]] < 0) tmp[inode[m]] = b->iblank[inode[m]];
      }
    }
    for (i = 0; i < b->nnodes; i++) ibr[i] = tmp[i];
  }
  free(tmp);
  b->ncancel = 0;
  for (i = 0; i < b->nnodes; i++)
    if (b->iblank[i] < 0 && ibr[i] == 0) push_cancel(b, i);
  return ibr;
}

/* tioga::reduce_fringes (tioga.C:737-885) on one rank: the blocks' fringe
   reduction, then comm 3 (getCancellationData / cancelDonor,
   bookKeeping.C:542-570, with the head of each dropped receptor's donor
   list), resetCoincident, and comm 4 (clearIblanks with iblank_reduced,
   bookKeeping.C:615-625, then setIblanks for every live record).  Cell
   iblanks are not recomputed, as in the reference. */
void orc_reduce_fringes(OrcTioga *t)
{
  int ib, i;
  int **ibr = (int **)calloc(t->nblocks > 0 ? t->nblocks : 1, sizeof(int *));
  for (ib = 0; ib < t->nblocks; ib++) ibr[ib] = block_reduce_fringes(t->b[ib]);
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ncancel; i++) {
      DRec *h = b->donorList ? b->donorList[b->cancelList[i]] : NULL;
      if (h == NULL) continue;
      int db = 0;
      OrcBlock *D = find_block(t, h->meshtag, &db);
      if (!D || !D->interp2donor) continue;
      int iptr = D->interp2donor[h->remoteid];
      if (iptr > -1) D->interp[iptr].cancel = 1;
    }
  }
  for (ib = 0; ib < t->nblocks; ib++) reset_coincident(t->b[ib]);
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->nnodes; i++)
      if (b->iblank[i] < 0) b->iblank[i] = 1;
    for (i = 0; i < b->nnodes; i++)
      if (ibr[ib][i] == 0) b->iblank[i] = 0;
  }
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ninterp; i++)
      if (!b->interp[i].cancel) t->b[b->interp[i].rinfo[2]]->iblank[b->interp[i].rinfo[1]] = -1;
  }
  for (ib = 0; ib < t->nblocks; ib++) free(ibr[ib]);
  free(ibr);
}

/* highOrder.C:88-150 */
static void cell_iblanks(OrcBlock *b)
{
  int n, i, m, icell = 0;
  if (b->iblank_cell == NULL) {
    b->iblank_cell = (int *)malloc(sizeof(int) * (b->ncells > 0 ? b->ncells : 1));
    b->own_iblank_cell = 1;
  }
  for (n = 0; n < b->ntypes; n++) {
    int nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      int flag = 1, ncount = 0;
      b->iblank_cell[icell] = 1;
      for (m = 0; m < nvert && flag; m++) {
        int v = b->iblank[b->vconn[n][nvert * i + m] - ORC_BASE];
        if (v == 0) { b->iblank_cell[icell] = 0; flag = 0; }
        ncount += (v == -1);
      }
      if (flag && ncount == nvert) b->iblank_cell[icell] = -1;
      icell++;
    }
  }
}

/* tioga.C:112-160 */
void orc_perform_connectivity(OrcTioga *t)
{
  int ib;
  get_hole_map(t);
  exchange_boxes(t);
  exchange_search_data(t);
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    free(b->interp);
    b->interp = NULL;
    b->ninterp = 0;
    block_search(t, b);
  }
  exchange_donors(t);
  for (ib = 0; ib < t->nblocks; ib++) cell_iblanks(t->b[ib]);
  for (ib = 0; ib < t->nblocks; ib++) t->b[ib]->q = NULL;
}

void orc_register_solution(OrcTioga *t, int btag, double *q)
{
  OrcBlock *b = find_block(t, btag, NULL);
  if (b) b->q = q;
}

/* tioga.C:336-524 + dataUpdate.C:27-119 (at_points = 0) */
void orc_data_update(OrcTioga *t, int nvar, int interptype)
{
  int ib, i, k, m;
  for (ib = 0; ib < t->nblocks; ib++)
    if (t->b[ib]->q == NULL) return;
  if (t->nsend == 0) return;
  /* gather every live record first, then scatter (packets are built before
     any update is applied) */
  size_t total = 0;
  for (ib = 0; ib < t->nblocks; ib++)
    for (i = 0; i < t->b[ib]->ninterp; i++)
      if (!t->b[ib]->interp[i].cancel) total++;
  double *vals = (double *)malloc(sizeof(double) * (total * nvar + 1));
  int *dst = (int *)malloc(sizeof(int) * (2 * total + 1));
  double *qq = (double *)malloc(sizeof(double) * (nvar > 0 ? nvar : 1));
  size_t r = 0;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ninterp; i++) {
      Interp *it = &b->interp[i];
      if (it->cancel) continue;
      for (k = 0; k < nvar; k++) qq[k] = 0;
      for (m = 0; m < it->nweights; m++) {
        int inode = it->inode[m];
        double w = it->w[m];
        for (k = 0; k < nvar; k++) {
          if (interptype == 0) qq[k] += b->q[inode * nvar + k] * w;
          else qq[k] += b->q[(size_t)k * b->nnodes + inode] * w;
        }
      }
      dst[2 * r] = it->rinfo[1];
      dst[2 * r + 1] = it->rinfo[2];
      for (k = 0; k < nvar; k++) vals[r * nvar + k] = qq[k];
      r++;
    }
  }
  for (r = 0; r < total; r++) {
    OrcBlock *R = t->b[dst[2 * r + 1]];
    int node = dst[2 * r];
    for (k = 0; k < nvar; k++) {
      if (interptype == 0) R->q[node * nvar + k] = vals[r * nvar + k];
      else R->q[(size_t)R->nnodes * k + node] = vals[r * nvar + k];
    }
  }
  free(vals);
  free(dst);
  free(qq);
}

/* tioga.C:533-549 + dataUpdate.C:121-134 */
void orc_get_donor_count(OrcTioga *t, int btag, int *dcount, int *fcount)
{
  int i;
  *dcount = *fcount = 0;
  OrcBlock *b = find_block(t, btag, NULL);
  if (!b || t->nsend == 0) return;
  for (i = 0; i < b->ninterp; i++)
    if (!b->interp[i].cancel) { (*dcount)++; *fcount += b->interp[i].nweights + 1; }
}

/* tioga.C:551-570 + dataUpdate.C:136-159 (rank = sndMap[0] = 0) */
void orc_get_donor_info(OrcTioga *t, int btag, int *receptors, int *indices, double *frac, int *dcount)
{
  int i, m, j = 0, k = 0;
  (void)dcount;
  OrcBlock *b = find_block(t, btag, NULL);
  if (!b || t->nsend == 0) return;
  for (i = 0; i < b->ninterp; i++) {
    Interp *it = &b->interp[i];
    if (it->cancel) continue;
    for (m = 0; m < it->nweights + 1; m++) { indices[j] = it->inode[m]; frac[j] = it->w[m]; j++; }
    receptors[k++] = it->rinfo[0];
    receptors[k++] = it->rinfo[1];
    receptors[k++] = it->rinfo[2];
    receptors[k++] = it->nweights;
  }
}

/* tioga::getReceptorInfo (tioga.C:572-671) on one rank, from
   MeshBlock::getReceptorInfo (dataUpdate.C:161-178) of every block in
   registration order: per live record (receptor node, mesh tag of the
   receptor block, the donor cell's global id as two ints in memory order).
   Without cell ids the reference reads a null pointer; here the cell index
   stands in.  Returns the number of ints; copies them when out != NULL. */
long orc_get_receptor_info(OrcTioga *t, int *out)
{
  long k = 0;
  int ib, i;
  if (t->nsend == 0) return 0;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ninterp; i++) {
      Interp *it = &b->interp[i];
      if (it->cancel) continue;
      if (out) {
        int don = it->inode[it->nweights];
        unsigned long long g = b->cell_gid ? b->cell_gid[don] : (unsigned long long)don;
        out[k] = it->rinfo[1];
        out[k + 1] = t->b[it->rinfo[2]]->tag;
        memcpy(&out[k + 2], &g, sizeof g);
      }
      k += 4;
    }
  }
  return k;
}

void orc_set_gids(OrcTioga *t, int btag, const unsigned long long *cell_gid, const unsigned long long *node_gid)
{
  OrcBlock *b = find_block(t, btag, NULL);
  if (!b) return;
  b->cell_gid = cell_gid;
  b->node_gid = node_gid;
}

#define COPY_OUT(ptr, n, type)                                   \
  do {                                                           \
    if (out && (ptr)) memcpy(out, (ptr), sizeof(type) * (size_t)(n)); \
    return (n);                                                  \
  } while (0)

long orc_block_array(OrcTioga *t, int btag, const char *name, void *out)
{
  OrcBlock *b = find_block(t, btag, NULL);
  if (!b) return -1;
  if (!strcmp(name, "nodeRes")) COPY_OUT(b->nodeRes, b->nnodes, double);
  if (!strcmp(name, "cellRes")) COPY_OUT(b->cellRes, b->ncells, double);
  if (!strcmp(name, "xsearch")) COPY_OUT(b->xsearch, 3 * b->nsearch, double);
  if (!strcmp(name, "res_search")) COPY_OUT(b->res_search, b->nsearch, double);
  if (!strcmp(name, "isearch")) COPY_OUT(b->isearch, 3 * b->nsearch, int);
  if (!strcmp(name, "tagsearch")) COPY_OUT(b->tagsearch, b->nsearch, int);
  if (!strcmp(name, "donorId")) COPY_OUT(b->donorId, b->nsearch, int);
  if (!strcmp(name, "xtag")) COPY_OUT(b->xtag, b->nsearch, int);
  if (!strcmp(name, "elementList")) COPY_OUT(b->elist, b->nelem, int);
  if (!strcmp(name, "adtIntegers")) COPY_OUT(b->adtInt, 4 * b->nelem, int);
  if (!strcmp(name, "adtReals")) COPY_OUT(b->adtReal, 6 * b->nelem, double);
  if (!strcmp(name, "adtExtents")) COPY_OUT(b->adtExt, b->nsearch > 0 ? 6 : 0, double);
  if (!strcmp(name, "iblank_cell")) COPY_OUT(b->iblank_cell, b->ncells, int);
  if (!strcmp(name, "receptorIdCart")) COPY_OUT(b->receptorIdCart, b->ncart, int);
  if (!strcmp(name, "donorIdCart")) COPY_OUT(b->donorIdCart, b->ncart, int);
  if (!strcmp(name, "rxyzCart")) COPY_OUT(b->rxyzCart, 3 * b->ncart, double);
  if (!strcmp(name, "obb")) {
    if (out) {
      double *o = (double *)out;
      memcpy(o, b->obb.xc, 24);
      memcpy(o + 3, b->obb.dxc, 24);
      memcpy(o + 6, b->obb.vec, 72);
    }
    return 15;
  }
  return -1;
}

long orc_hole_map(OrcTioga *t, int mtag, int nx[3], double ext[6], int *sam)
{
  if (!t->hm || mtag < 1 || mtag > t->nmesh) return 0;
  OrcHM *h = &t->hm[mtag - 1];
  if (!h->existWall) return 0;
  long n = (long)h->nx[0] * h->nx[1] * h->nx[2];
  if (nx) memcpy(nx, h->nx, sizeof(int) * 3);
  if (ext) memcpy(ext, h->ext, sizeof(double) * 6);
  if (sam) memcpy(sam, h->sam, sizeof(int) * n);
  return n;
}

typedef struct {
  const OrcBlock *b;
  const double *xq;
  int *donor;
  int i0, i1;
} SearchChunk;

static void *search_chunk(void *arg)
{
  SearchChunk *c = (SearchChunk *)arg;
  int i;
  for (i = c->i0; i < c->i1; i++) c->donor[i] = search_adt(c->b, &c->xq[3 * i]);
  return NULL;
}

/* MeshBlock::search (search.C:107-233) for the query set xobb[nobb]: the ADT
   over the cells of block btag that overlap the OBB of xobb, then
   searchADT (searchADTrecursion.C:26-101) for each of xq[nq], without the
   coincident-point dedupe.  xq may be a subset (a sample) of xobb: its
   searches run against the tree of the whole set.  Each search only reads
   the block, so nthreads > 1 splits the queries into contiguous chunks over
   POSIX threads.  Returns the elapsed seconds of the search loop. */
double orc_search_points_obb(OrcTioga *t, int btag, const double *xobb, int nobb, const double *xq, int nq,
                             int *donor, int nthreads, int *nelem)
{
  int k;
  OrcBlock *b = find_block(t, btag, NULL);
  if (nelem) *nelem = 0;
  if (!b || nobb <= 0 || nq < 0) return 0.0;
  if (!b->cellRes) tag_boundary(b);
  build_block_adt(b, xobb, nobb);
  if (nelem) *nelem = b->nelem;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  SearchChunk ch[256];
  pthread_t th[256];
  int started[256];
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (k = 0; k < nthreads; k++) {
    ch[k].b = b;
    ch[k].xq = xq;
    ch[k].donor = donor;
    ch[k].i0 = (int)((long long)nq * k / nthreads);
    ch[k].i1 = (int)((long long)nq * (k + 1) / nthreads);
  }
  if (nthreads == 1) {
    search_chunk(&ch[0]);
  } else {
    for (k = 0; k < nthreads; k++) {
      started[k] = pthread_create(&th[k], NULL, search_chunk, &ch[k]) == 0;
      if (!started[k]) search_chunk(&ch[k]); /* no thread: run the chunk here */
    }
    for (k = 0; k < nthreads; k++)
      if (started[k]) pthread_join(th[k], NULL);
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

double orc_search_points(OrcTioga *t, int btag, const double *xq, int nq, int *donor)
{
  int i;
  OrcBlock *b = find_block(t, btag, NULL);
  if (!b || nq <= 0) return 0.0;
  if (!b->cellRes) tag_boundary(b);
  build_block_adt(b, xq, nq);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (i = 0; i < nq; i++) donor[i] = search_adt(b, &xq[3 * i]);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ======================================================================
 * AMR Cartesian background, gather direction: near-body receptors take
 * their values from Cartesian patches (pdegree 0, linear interpolation
 * between cell centres).
 * ====================================================================== */

/* CartGrid::registerData (CartGrid.C:22-70) */
void orc_amr_global_data(OrcTioga *t, int nf, int qstride, const double *qnode, const int *idata,
                         const double *rdata, int ngrids, int qnodesize)
{
  int i, n;
  OrcCart *c = t->cart;
  if (c) free_cart(c);
  c = (OrcCart *)calloc(1, sizeof(OrcCart));
  t->cart = c;
  c->ngrids = ngrids;
  c->nf = nf;
  c->qstride = qstride;
  c->qnodesize = qnodesize;
  c->gid = (int *)malloc(sizeof(int) * (ngrids + 1));
  c->level = (int *)malloc(sizeof(int) * (ngrids + 1));
  c->proc = (int *)malloc(sizeof(int) * (ngrids + 1));
  c->porder = (int *)malloc(sizeof(int) * (ngrids + 1));
  c->lid = (int *)malloc(sizeof(int) * (ngrids + 1));
  c->ilo = (int *)malloc(sizeof(int) * 3 * (ngrids + 1));
  c->ihi = (int *)malloc(sizeof(int) * 3 * (ngrids + 1));
  c->dims = (int *)malloc(sizeof(int) * 3 * (ngrids + 1));
  c->xlo = (double *)malloc(sizeof(double) * 3 * (ngrids + 1));
  c->dx = (double *)malloc(sizeof(double) * 3 * (ngrids + 1));
  c->qnode = (double *)malloc(sizeof(double) * (qnodesize + 1));
  for (i = 0; i < qnodesize; i++) c->qnode[i] = qnode[i];
  for (i = 0; i < ngrids; i++) {
    const int *id = idata + 11 * i;
    const double *rd = rdata + 6 * i;
    c->gid[i] = id[0];
    c->level[i] = id[1];
    c->proc[i] = id[2];
    c->porder[i] = id[3];
    c->lid[i] = id[4];
    for (n = 0; n < 3; n++) {
      c->ilo[3 * i + n] = id[5 + n];
      c->ihi[3 * i + n] = id[8 + n];
      c->dims[3 * i + n] = id[8 + n] - id[5 + n] + 1;
      c->xlo[3 * i + n] = rd[n];
      c->dx[3 * i + n] = rd[3 + n];
    }
  }
}

/* tioga::set_amr_patch_count (tioga.C:703-708) */
void orc_amr_patch_count(OrcTioga *t, int npatch)
{
  OrcCart *c = t->cart;
  if (!c) return;
  free(c->pgid); free(c->pibl); free(c->pq);
  c->npatch = npatch;
  c->pgid = (int *)calloc(npatch + 1, sizeof(int));
  c->pibl = (int **)calloc(npatch + 1, sizeof(int *));
  c->pq = (double **)calloc(npatch + 1, sizeof(double *));
}

/* tioga::register_amr_local_data (tioga.C:710-713), CartBlock::registerData
   (CartBlock.h:52-58) */
void orc_amr_local_data(OrcTioga *t, int ipatch, int global_id, int *iblank, double *q)
{
  OrcCart *c = t->cart;
  if (!c || ipatch < 0 || ipatch >= c->npatch) return;
  c->pgid[ipatch] = global_id;
  c->pibl[ipatch] = iblank;
  c->pq[ipatch] = q;
}

/* CartGrid::preprocess (CartGrid.C:77-104) */
static void cart_preprocess(OrcCart *c)
{
  int i, n;
  c->xlosup[0] = c->xlosup[1] = c->xlosup[2] = BIGV;
  c->maxlevel = -1;
  for (i = 0; i < c->ngrids; i++) {
    for (n = 0; n < 3; n++) c->xlosup[n] = dmin(c->xlosup[n], c->xlo[3 * i + n]);
    if (c->level[i] > c->maxlevel) c->maxlevel = c->level[i];
  }
  c->maxlevel++;
  free(c->dxlvl);
  c->dxlvl = (double *)malloc(sizeof(double) * 3 * (c->maxlevel + 1));
  for (i = 0; i < c->ngrids; i++)
    for (n = 0; n < 3; n++) c->dxlvl[3 * c->level[i] + n] = c->dx[3 * i + n];
}

/* CartGrid::search (CartGrid.C:111-145): the first patch, in registration
   order, of the finest level whose box holds the point within +-TOL. */
static int cart_search_point(const OrcCart *c, const double *x)
{
  int l, j, n;
  for (l = c->maxlevel - 1; l >= 0; l--)
    for (j = 0; j < c->ngrids; j++) {
      if (c->level[j] != l) continue;
      int flag = 1;
      for (n = 0; n < 3; n++) flag = flag && ((x[n] - c->xlo[3 * j + n]) > -TOLV);
      for (n = 0; n < 3; n++)
        flag = flag && ((x[n] - (c->xlo[3 * j + n] + c->dx[3 * j + n] * (c->dims[3 * j + n]))) < TOLV);
      if (flag) return j;
    }
  return -1;
}

void orc_amr_search(OrcTioga *t, const double *x, int n, int *donor)
{
  int i;
  if (!t->cart) return;
  cart_preprocess(t->cart);
  for (i = 0; i < n; i++) donor[i] = cart_search_point(t->cart, x + 3 * (size_t)i);
}

/* MeshBlock::getUnresolvedMandatoryReceptors (cartOps.C:61-160), node form:
   every node of a cell whose nodes all carry nodeRes == BIGVALUE and whose
   iblank_cell is 1, in node order. */
static void unresolved_mandatory(OrcBlock *b)
{
  int n, i, m, k = 0;
  char *picked = (char *)calloc(b->nnodes + 1, 1);
  for (n = 0; n < b->ntypes; n++) {
    int nvert = b->nv[n];
    for (i = 0; i < b->nc[n]; i++) {
      int fcount = 0, ibc = b->iblank_cell ? b->iblank_cell[k] : 1;
      for (m = 0; m < nvert; m++)
        if (b->nodeRes[b->vconn[n][nvert * i + m] - ORC_BASE] == BIGV) fcount++;
      if (fcount == nvert && ibc == 1)
        for (m = 0; m < nvert; m++) picked[b->vconn[n][nvert * i + m] - ORC_BASE] = 1;
      k++;
    }
  }
  free(b->receptorIdCart); free(b->donorIdCart); free(b->rxyzCart);
  b->ncart = 0;
  for (i = 0; i < b->nnodes; i++) b->ncart += picked[i];
  b->receptorIdCart = (int *)malloc(sizeof(int) * (b->ncart + 1));
  b->donorIdCart = (int *)malloc(sizeof(int) * (b->ncart + 1));
  b->rxyzCart = (double *)malloc(sizeof(double) * 3 * (b->ncart + 1));
  m = 0;
  for (i = 0; i < b->nnodes; i++)
    if (picked[i]) {
      b->receptorIdCart[m] = i;
      for (n = 0; n < 3; n++) b->rxyzCart[3 * m + n] = b->x[3 * i + n];
      m++;
    }
  free(picked);
}

/* CartBlock::insertInInterpList (CartBlock.C:196-268) with
   cart_interp::linear_interpolation (linCartInterp.C:49-109) for p = 1.
   Returns 0, or -1 where the reference's assert (CartBlock.C:246) fails. */
static int insert_interp(const OrcCart *c, int gidx, const double *x, AmrRec *r)
{
  int n, ix[3], dims[3], start[3], i, j, k, ind;
  double xlo[3], dx[3], rst[3], rc[3], px[2], py[2], pz[2];
  for (n = 0; n < 3; n++) {
    xlo[n] = c->xlo[3 * gidx + n];
    dx[n] = c->dx[3 * gidx + n];
    dims[n] = c->ihi[3 * gidx + n] - c->ilo[3 * gidx + n] + 1;
  }
  for (n = 0; n < 3; n++) {
    ix[n] = (int)((x[n] - xlo[n]) / dx[n]);
    rst[n] = (x[n] - xlo[n] - ix[n] * dx[n]) / dx[n];
    if (ix[n] == dims[n] && fabs(rst[n]) < TOLV) {
      ix[n]--;
      rst[n] = (x[n] - xlo[n] - ix[n] * dx[n]) / dx[n];
    }
    if (!(ix[n] >= 0 && ix[n] < dims[n])) return -1;
  }
  /* create_donor_stencil (linCartInterp.C:74-95) */
  for (n = 0; n < 3; n++) start[n] = (rst[n] - 0.5 >= 0) ? ix[n] : ix[n] - 1;
  ind = 0;
  for (k = 0; k < 2; k++)
    for (j = 0; j < 2; j++)
      for (i = 0; i < 2; i++) {
        int s[3] = {start[0] + i, start[1] + j, start[2] + k};
        for (n = 0; n < 3; n++) {
          int v = s[n] < dims[n] - 1 ? s[n] : dims[n] - 1;
          r->ijk[ind++] = v > 0 ? v : 0;
        }
      }
  /* compute_ref_coords (linCartInterp.C:63-72), p = 1 */
  for (n = 0; n < 3; n++)
    rc[n] = (rst[n] - 0.5 >= 0) ? ((rst[n] - 0.5) / 1 * 2 - 1) : ((rst[n] + 0.5 * (2 * 1 - 1)) / 1 * 2 - 1);
  /* compute_1d_bases + compute_weights (linCartInterp.C:24-61) */
  px[0] = (1 - rc[0]) / 2; px[1] = (1 + rc[0]) / 2;
  py[0] = (1 - rc[1]) / 2; py[1] = (1 + rc[1]) / 2;
  pz[0] = (1 - rc[2]) / 2; pz[1] = (1 + rc[2]) / 2;
  ind = 0;
  for (k = 0; k < 2; k++)
    for (j = 0; j < 2; j++)
      for (i = 0; i < 2; i++) r->w[ind++] = pz[k] * py[j] * px[i];
  return 0;
}

/* MeshBlock::setCartIblanks (cartOps.C:32-58), node form */
static void set_cart_iblanks(OrcBlock *b)
{
  int m;
  for (m = 0; m < b->ncart; m++)
    if (b->donorIdCart[m] != -1) b->iblank[b->receptorIdCart[m]] = -1;
}

/* tioga::performConnectivityAMR (tioga.C:180-216) on one rank, gather
   direction only: CartGrid/CartBlock preprocess, the unresolved mandatory
   receptors of every block, CartGrid::search, the patch interpolation lists
   in the order exchangeAMRDonors.C:120-192 inserts them (block, point),
   setCartIblanks (exchangeAMRDonors.C:271-272) and getCellIblanks.  The
   Cartesian-receptor direction (getCartReceptors, donor lists on patches,
   processDonors, cancellation) is not part of this path.
   Returns 0; -1 for a patch with pdegree > 0 (the reference needs a
   user callback there, CartBlock.C:167-168); -2 where the reference's
   stencil assert fails. */
int orc_perform_connectivity_amr(OrcTioga *t)
{
  int ib, i, p, ret = 0;
  OrcCart *c = t->cart;
  if (!c) return -3;
  cart_preprocess(c);
  for (p = 0; p < c->npatch; p++)
    if (c->porder[c->pgid[p]] != 0) return -1;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    unresolved_mandatory(b);
    for (i = 0; i < b->ncart; i++) b->donorIdCart[i] = cart_search_point(c, b->rxyzCart + 3 * (size_t)i);
  }
  /* patch lists: count, then fill in (block, point) order */
  free(c->roff); free(c->rec);
  c->roff = (int *)calloc(c->npatch + 1, sizeof(int));
  c->nrec = 0;
  for (ib = 0; ib < t->nblocks; ib++) {
    OrcBlock *b = t->b[ib];
    for (i = 0; i < b->ncart; i++)
      if (b->donorIdCart[i] != -1) { c->roff[c->lid[b->donorIdCart[i]] + 1]++; c->nrec++; }
  }
  for (p = 0; p < c->npatch; p++) c->roff[p + 1] += c->roff[p];
  c->rec = (AmrRec *)malloc(sizeof(AmrRec) * (c->nrec + 1));
  {
    int *fill = (int *)calloc(c->npatch + 1, sizeof(int));
    for (ib = 0; ib < t->nblocks; ib++) {
      OrcBlock *b = t->b[ib];
      for (i = 0; i < b->ncart; i++) {
        int g = b->donorIdCart[i];
        if (g == -1) continue;
        int lp = c->lid[g];
        AmrRec *r = &c->rec[c->roff[lp] + fill[lp]++];
        r->rinfo[0] = 0;
        r->rinfo[1] = i;
        r->rinfo[2] = ib;
        if (insert_interp(c, c->pgid[lp], b->rxyzCart + 3 * (size_t)i, r)) ret = -2;
      }
    }
    free(fill);
  }
  for (ib = 0; ib < t->nblocks; ib++) set_cart_iblanks(t->b[ib]);
  for (ib = 0; ib < t->nblocks; ib++) cell_iblanks(t->b[ib]);
  t->amr = 1;
  return ret;
}

/* tioga::dataUpdate_AMR (tioga.C:218-334) on one rank: CartBlock::
   getInterpolatedData (CartBlock.C:44-140, pdegree 0, get_amr_index_xyz.c:1-56)
   for every patch, then MeshBlock::updateSolnData (dataUpdate.C:105-119) at
   receptorIdCart of the receiving block. */
void orc_data_update_amr(OrcTioga *t, int nvar, int interptype)
{
  int p, r, m, n;
  OrcCart *c = t->cart;
  double *qq = (double *)malloc(sizeof(double) * (nvar + 1));
  if (!c || !c->rec) { free(qq); return; }
  for (p = 0; p < c->npatch; p++) {
    int g = c->pgid[p];
    int nx = c->dims[3 * g], ny = c->dims[3 * g + 1], nz = c->dims[3 * g + 2], nf = c->nf;
    int d3nf = (nx + 2 * nf) * (ny + 2 * nf) * (nz + 2 * nf);
    const double *q = c->pq[p];
    for (r = c->roff[p]; r < c->roff[p + 1]; r++) {
      const AmrRec *a = &c->rec[r];
      for (n = 0; n < nvar; n++) qq[n] = 0;
      for (m = 0; m < 8; m++) {
        int i = a->ijk[3 * m], j = a->ijk[3 * m + 1], k = a->ijk[3 * m + 2];
        int idx = (ny + 2 * nf) * (nx + 2 * nf) * (k + nf) + (nx + 2 * nf) * (j + nf) + (i + nf);
        for (n = 0; n < nvar; n++) qq[n] += q[idx + d3nf * n] * a->w[m];
      }
      OrcBlock *b = t->b[a->rinfo[2]];
      int inode = b->receptorIdCart[a->rinfo[1]];
      if (b->q == NULL) continue;
      for (n = 0; n < nvar; n++) {
        if (interptype == 0) b->q[inode * nvar + n] = qq[n];
        else b->q[b->nnodes * n + inode] = qq[n];
      }
    }
  }
  free(qq);
}

/* Patch interpolation records in list order: "rec_patch" (local patch),
   "rec_info" (3 ints: rank, point, block), "rec_ijk" (24 ints), "rec_w"
   (8 doubles). */
long orc_amr_array(OrcTioga *t, const char *name, void *out)
{
  OrcCart *c = t->cart;
  long i, n;
  int p;
  if (!c || !c->rec) return 0;
  n = c->nrec;
  if (!strcmp(name, "rec_patch")) {
    if (out)
      for (p = 0; p < c->npatch; p++)
        for (i = c->roff[p]; i < c->roff[p + 1]; i++) ((int *)out)[i] = p;
    return n;
  }
  if (!strcmp(name, "rec_info")) {
    if (out) for (i = 0; i < n; i++) memcpy((int *)out + 3 * i, c->rec[i].rinfo, 12);
    return 3 * n;
  }
  if (!strcmp(name, "rec_ijk")) {
    if (out) for (i = 0; i < n; i++) memcpy((int *)out + 24 * i, c->rec[i].ijk, 96);
    return 24 * n;
  }
  if (!strcmp(name, "rec_w")) {
    if (out) for (i = 0; i < n; i++) memcpy((double *)out + 8 * i, c->rec[i].w, 64);
    return 8 * n;
  }
  return -1;
}
