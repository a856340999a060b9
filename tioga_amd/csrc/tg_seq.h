// tg_seq.h -- the order-dependent sequential stages of the connectivity path
// that run on the host: FP64 running sums whose rounding fixes the OBB bits,
// the median-split ADT build whose quickselect order fixes the tree, and the
// octree coincident-point dedupe whose chains fix xtag.  Each keeps the
// reference's exact operation order so that the device stages downstream see
// the same inputs as the reference's.
#pragma once

#include <vector>

#include "tg_common.h"

namespace tg {

// Principal-axis OBB of n points (utils.c:25-165 with the Jacobi solver of
// kaiser.f:29-180).  x is AoS xyz; when idx is non-null the points are
// x[3*idx[i]..], in that order.
void find_obb(const double *x, const int *idx, int n, OBB &o);
// find_obb's last two steps: the eigenvectors of the covariance matrix aa[9]
// (column-major, overwritten) into o.vec (utils.c:106-118), and from the
// frame extents the half widths and the shifted centre (utils.c:148-161)
void obb_axes(double *aa, OBB &o);
void obb_finish(OBB &o, const double xmin[3], const double xmax[3]);

// 15-axis separating-axis test with eps 1e-12 (utils.c:297-384), including
// the reference's |vA[i][j]|*D[j] projection of the centre offset.
int obb_intersect(const OBB &a, const OBB &b);

// Corners of an OBB given centre, half extents and axes (utils.c:386-400).
void obb_corners(const double xc[3], const double dxc[3], const double vec[3][3], double xv[8][3]);

// 6-D alternating digital tree over nelem boxes (xmin,ymin,zmin,xmax,ymax,
// zmax) (ADT.C:34-135, buildADTrecursion.c:24-116, median.F90:59-210):
// adtInt[4*nelem], adtReal[6*nelem], ext[6] = padded root extents.
// Returns the tree depth (levels).
int build_adt(int nelem, const double *ebox, int *adtInt, double *adtReal, double ext[6]);

// median_select restated with list-based partition rounds (hoare_stop,
// median_control), the formulation the device uses; same result.
void median_select_lists(int *ix, double *x, int n);

// Coincident-point dedupe (utils.c:620-737): xtag[i] = representative of i,
// rtag (resolutions) raised to the max over each coincident group.
void unique_nodes_octree(const double *x, const int *meshtag, double *rtag, int *xtag, int n);

}  // namespace tg
