// Finite strain ellipses and the DEM viscous tensor (fabric.hip).
#pragma once

#include "stk_internal.h"

namespace stk {

enum FseMode { FSE_NO_HEALING = 0, FSE_FABRIC_DESTRUCTION = 1, FSE_ANNEALING_DESTRUCTION = 2 };

// F: 4 doubles per point (column-major 2x2), npt points, device; updated in place by the
// healing rules.  fse: 6 planes of npt (x1, x2, y1, y2, a1, a2), device.
void launch_fse(Context &c, double *F, int64_t npt, int mode, double eps, double *fse);

// ellipses between the device planes and the host layout (6 consecutive fields per point)
void fse_layout(Context &c, const double *src, double *dst, int64_t npt, bool to_points);

// getDEM on the host arrays of DEMloader: tensor n x 9, axes n x 3 (column-major).
void dem_prepare(Context &c, const double *tensor, const double *axes, int64_t n, double w);

// Fabric shape of every point: either r1, r2 per point (nodes == nullptr) or sampled at
// nshape increasing a1 nodes.  All device pointers.
struct ShapeIn {
    const double *nodes = nullptr, *r1 = nullptr, *r2 = nullptr;
    int64_t nshape = 0;
};

// fse: device ellipse planes of nel x nip points (nip 6 or 7); out: 6 x nel x 7, device.
void launch_anisotropic_tensor(Context &c, const double *fse, int nip, const ShapeIn &shape, double *out);

}  // namespace stk
