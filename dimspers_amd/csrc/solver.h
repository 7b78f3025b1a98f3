// Solver entry points and the device-scalar slots they share.
#pragma once

#include "stk_internal.h"

namespace stk {

// Slots of Context::d_sc (device doubles written by reductions, read by kernels).
enum : int {
    SC_TMP = 0,      // scratch (two slots)
    SC_BB = 2,       // ||b||^2 of the inner CG
    SC_RU = 6,       // ||ru||^2 of the full residual
    SC_RC = 7,       // ||rc||^2 of the full residual
    SC_HIST = 8,     // recent inner residual norms (8 slots)
    SC_VY = 16,      // v . y of the inner right-hand side (null-vector component)
    SC_VZ = 17,      // v . z of the inner solution
    SC_PAP2 = 20,    // (p . K p, v . K p) of the inner CG (20, 21)
    SC_RZV = 22,     // (r . z, v . z) of the inner CG, ping-pong pairs (22, 23) and (24, 25)
    SC_SRD = 26,     // r . d of the current Schur iteration
    SC_SQSQ = 27,    // q . S q
    SC_SRN = 28,     // ||r||^2 after the update; then (29, 30): r . d and (r - r_last) . d of the
                     // next direction, fetched together (29 is the next iteration's r . d)
    SC_COUNT = 32
};

// Two-pass deterministic reductions into device scalar slots.  map_reduce<ND> takes a
// device function returning a double (ND = 1) or Dots<ND> (slots slot .. slot+ND-1);
// bpe = the bytes per element it reads and writes (0: not stated, 8 assumed).
struct Reducer {
    Context &c;
    explicit Reducer(Context &c);
    template <int ND = 1, typename F>
    void map_reduce(int64_t n, F f, int slot, int bpe = 0);
    void finish(int64_t npart, int slot, int nd = 1);
    double fetch(int slot);
    void fetch_many(int slot0, int n, double *out);
};

template <typename T>
void setup_solver_data(Context &c);
// Set-up of the FP32 methods after an FP64 + FP32 assembly with the multigrid; the FP64 hierarchy
// follows on demand (ensure_mg64).
void setup_fast(Context &c);
void ensure_mg64(Context &c);
void build_coarse_inverse(Context &c, bool f64);
int run_solve(Context &c, const stk_opts &o, stk_report &rep);
double full_residual(Context &c, const double *U, const double *P, double *ru, double *hc, bool project_null);
double time_kernel(Context &c, int kernel, int reps, double *bytes);

}  // namespace stk
