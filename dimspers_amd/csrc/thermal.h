// Thermal diffusion operators (thermal.hip) and the thermal step (solver.hip).
#pragma once

#include "stk_internal.h"

namespace stk {

// KT, MT: one value per stored entry of the sliced P2 node pattern (zeroed by the caller);
// F: per node (zeroed by the caller; written only when dQdT != 0).
void thermal_assemble(Context &c, const double *rho, double K, double Cp, double dQdT, double dt, double *KT,
                      double *MT, double *F);

struct ThermalStep {
    double K, Cp, dQdT, dt, rtol;
    int itmax;
};
// One step of solveDiffusion_threaded on device arrays: T (nnod) holds T0 on entry and T on
// return; fixed: 0-based nodes with their values.
void thermal_solve(Context &c, double *T, const double *rho, const ThermalStep &s, const int64_t *fixed,
                   const double *vfix, int64_t nfix, stk_report &rep);

}  // namespace stk
