/* ho_callbacks.c -- TEST INFRASTRUCTURE ONLY.
 *
 * The five high-order callbacks of a small synthetic solver
 * (tiogaInterface.C:273-277: get_nodes_per_cell, get_receptor_nodes,
 * donor_inclusion_test, donor_frac, convert_to_modal), compiled into both
 * the reference driver (oracle/ref_driver.cpp) and the test library that
 * the engine tests hand to tioga_set_highorder_callback_, so both sides
 * call the same code.  The solver has one receptor mesh (whose receptor
 * cells give the points) and one donor mesh (whose cells are tested and
 * weighted), each of one cell type with nv vertices, 1-based connectivity.
 *
 *   nodes per cell   1 + cell % 2;
 *   receptor points  the vertex centroid, and for two points also the
 *                    midpoint of the centroid and the first vertex;
 *   inclusion        the point inside the axis-aligned box of the cell's
 *                    vertices (bounds included); rst = the point's fraction
 *                    of the box along each axis;
 *   donor weights    8 trilinear weights of rst, applied to 8 consecutive
 *                    solution rows from 1 + (cell - 1) % (nnodes - 8);
 *   modal values     one row: the mean of the cell's point values, written
 *                    at the row of the cell's first vertex.
 * Built with -ffp-contract=off on both sides, so the arithmetic is the same.
 */
#include <math.h>

static const double *g_rx, *g_dx;
static const int *g_rconn, *g_dconn;
static int g_rnv, g_dnv, g_dnnodes, g_nvar;
static long g_calls[5];

void ho_cb_set(const double *rx, const int *rconn, int rnv, const double *dx, const int *dconn, int dnv, int dnnodes,
               int nvar)
{
  g_rx = rx;
  g_rconn = rconn;
  g_rnv = rnv;
  g_dx = dx;
  g_dconn = dconn;
  g_dnv = dnv;
  g_dnnodes = dnnodes;
  g_nvar = nvar;
  for (int k = 0; k < 5; k++) g_calls[k] = 0;
}

/* calls of each callback since ho_cb_set */
long ho_cb_calls(int k) { return k >= 0 && k < 5 ? g_calls[k] : -1; }

void ho_nodes_per_cell(int *icell, int *npts)
{
  g_calls[0]++;
  *npts = 1 + (*icell % 2);
}

void ho_receptor_nodes(int *icell, int *npts, double *xyz)
{
  g_calls[1]++;
  const int *cv = g_rconn + (long)g_rnv * (*icell - 1);
  double c[3] = {0, 0, 0};
  for (int m = 0; m < g_rnv; m++)
    for (int j = 0; j < 3; j++) c[j] += g_rx[3 * (long)(cv[m] - 1) + j];
  for (int j = 0; j < 3; j++) {
    c[j] = c[j] / g_rnv;
    xyz[j] = c[j];
  }
  if (*npts == 2)
    for (int j = 0; j < 3; j++) xyz[3 + j] = 0.5 * c[j] + 0.5 * g_rx[3 * (long)(cv[0] - 1) + j];
}

void ho_inclusion(int *icell, double *x, int *pass, double *rst)
{
  g_calls[2]++;
  const int *cv = g_dconn + (long)g_dnv * (*icell - 1);
  double lo[3], hi[3];
  for (int j = 0; j < 3; j++) lo[j] = hi[j] = g_dx[3 * (long)(cv[0] - 1) + j];
  for (int m = 1; m < g_dnv; m++)
    for (int j = 0; j < 3; j++) {
      const double v = g_dx[3 * (long)(cv[m] - 1) + j];
      if (v < lo[j]) lo[j] = v;
      if (v > hi[j]) hi[j] = v;
    }
  *pass = 1;
  for (int j = 0; j < 3; j++) {
    if (x[j] < lo[j] || x[j] > hi[j]) *pass = 0;
    rst[j] = (x[j] - lo[j]) / (hi[j] - lo[j]);
  }
}

void ho_frac(int *icell, double *x, int *nweights, int *inode, double *frac, double *rst, int *ndim)
{
  (void)x;
  (void)ndim;
  g_calls[3]++;
  *nweights = 8;
  *inode = 1 + (*icell - 1) % (g_dnnodes - 8);
  for (int m = 0; m < 8; m++)
    frac[m] = ((m & 1) ? rst[0] : 1.0 - rst[0]) * ((m & 2) ? rst[1] : 1.0 - rst[1]) * ((m & 4) ? rst[2] : 1.0 - rst[2]);
}

void ho_modal(int *icell, int *npts_cell, double *qtmp, int *npts, int *index_out, double *qout)
{
  g_calls[4]++;
  const int *cv = g_rconn + (long)g_rnv * (*icell - 1);
  *npts = 1;
  *index_out = 1 + g_nvar * (cv[0] - 1);
  for (int n = 0; n < g_nvar; n++) qout[n] = *npts_cell == 2 ? 0.5 * (qtmp[n] + qtmp[g_nvar + n]) : qtmp[n];
}
