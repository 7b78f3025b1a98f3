// Stress post-processor and deformation gradient (stress.hip).
#pragma once

#include "stk_internal.h"

namespace stk {

void stress_prepare(Context &c);
void stress_set_rule(Context &c, int nip);
// out: 8 planes of nel x nip (exx, ezz, exz, eII, txx, tzz, txz, tII)
void launch_stress(Context &c, const double *U, double *out);
// F: 4 x (nel x nip), column-major 2x2 per point, advanced in place by one RK4 step
void launch_deformation_gradient(Context &c, const double *U, double dt, double *F);
// write_stats: out = (Urms, <T>, Nusselt top, Nusselt bottom); U, T device arrays
void statistics(Context &c, const double *U, const double *T, double out[4]);
// Once per mesh: natural-order element data and the Nusselt node lists of statistics.
void statistics_prepare(Context &c);
// Device address of the fixed scratch array of statistics_noalloc on `device` (the current device).
double *stats_scratch_address(int device);
// statistics on device arrays into the device array out[4], enqueued on the context stream,
// allocating nothing: its scratch is the fixed per-device array.
void statistics_noalloc(Context &c, const double *U, const double *T, double *out);

}  // namespace stk
