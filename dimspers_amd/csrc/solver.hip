// Solvers on the assembled operators.
//
//  * Schur-complement PCG (src/IterativeSolvers.jl:94-168): the reference
//    algorithm -- initial velocity solve, residual r = -G^T U with its mean
//    removed, lumped 1/eta-mass preconditioner (:189-240), Polak-Ribiere beta
//    (:163), fused solution update (:246-259) -- with the direct inner solve
//    lu(K[ifree, ifree]) (src/PardisoSolvers.jl:247-278) replaced by CG
//    preconditioned with a geometric multigrid V-cycle.
//  * Mixed precision: FP64 iterative refinement on the full saddle-point
//    residual around an FP32 Schur PCG whose inner solves are FP32 too.
//  * Learned: the mixed solver with the assembled local-Schur blocks predicted
//    by the MLP as the pressure preconditioner, falling back to the plain FP64
//    method when a refinement sweep stalls.
//
// The rigid rotation of the annulus (free slip on both walls) is a numerical
// null vector of K; the inner solves project it out of their right-hand side,
// and the caller removes the net rotation afterwards (updatevelocity2,
// src/Stokes/StokesSolver.jl:170-184), exactly as the reference does.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <type_traits>
#include <exception>


#include "device_ops.h"
#include "solver.h"
#include "fabric.h"
#include "stress.h"
#include "thermal.h"

namespace stk {

namespace {
constexpr int NT = 256;

// The FP64 or FP32 member of a pair, chosen by the value type.
template <typename T, typename A, typename B>
auto &pick(A &a64, B &a32) {
    if constexpr (std::is_same_v<T, double>) return a64;
    else return a32;
}

inline int nblk(int64_t n, int per = NT) { return (int)((n + per - 1) / per); }

// STK_TRACE=1 in the environment prints solver progress to stderr.
bool trace_on() {
    static const int on = [] {
        const char *e = std::getenv("STK_TRACE");
        return (e && *e && *e != '0') ? 1 : 0;
    }();
    return on != 0;
}
}  // namespace

template <typename T>
void DBuf<T>::zero(hipStream_t s) {
    if (n) STK_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
}
template struct DBuf<double>;
template struct DBuf<float>;
template struct DBuf<int32_t>;
template struct DBuf<uint8_t>;
template struct DBuf<int64_t>;
template struct DBuf<uint16_t>;

// ------------------------------------------------------------------ launch helpers
Reducer::Reducer(Context &c) : c(c) {}

// Copy n scalars to the host's coherent memory, then publish the sequence number (system-scope
// fence between the two, so a host that sees the number sees the values).
__global__ void k_publish(const double *__restrict__ src, int n, double *dst, unsigned *seq_ptr, unsigned seq) {
    if (threadIdx.x != 0) return;
    for (int i = 0; i < n; ++i) dst[i] = src[i];
    __threadfence_system();
    *(volatile unsigned *)seq_ptr = seq;
}

double Reducer::fetch(int slot) {
    double v;
    fetch_many(slot, 1, &v);
    return v;
}

// The host polls the sequence number instead of waiting for the stream (a round trip of ~6 instead
// of ~15 us, measured on MI355X): the values are read as soon as the kernel has written them.  After
// 2 ms of polling (the stream still had much work queued, or a kernel failed) it waits for the
// stream, which also reports a failed kernel.
void Reducer::fetch_many(int slot0, int n, double *out) {
    const unsigned seq = ++c.pub_seq;
    k_publish<<<dim3(1), dim3(64), 0, c.stream>>>(c.d_sc.p + slot0, n, c.h_pub + slot0, c.h_seq, seq);
    STK_HIP(hipGetLastError());
    volatile unsigned *hs = c.h_seq;
    const auto t0 = std::chrono::steady_clock::now();
    bool seen = false;
    for (int k = 0;; ++k) {
        if (*hs == seq) { seen = true; break; }
        if ((k & 255) == 255 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) break;
    }
    if (!seen) {
        STK_HIP(hipStreamSynchronize(c.stream));
        if (*hs != seq) throw Error(STK_E_HIP, "a published scalar did not reach the host");
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    const volatile double *hp = c.h_pub + slot0;
    for (int i = 0; i < n; ++i) out[i] = hp[i];
}

void Reducer::finish(int64_t npart, int slot, int nd) {
    constexpr int NF = 1024;
    if (nd == 1) k_reduce_final<NF, 1><<<dim3(1), dim3(NF), 0, c.stream>>>(npart, c.red64.p, c.d_sc.p + slot);
    else if (nd == 2) k_reduce_final<NF, 2><<<dim3(1), dim3(NF), 0, c.stream>>>(npart, c.red64.p, c.d_sc.p + slot);
    else throw Error(STK_E_STATE, "reduction of more than two columns");
    STK_HIP(hipGetLastError());
    c.ledger.add(8.0 * (double)(npart + 1) * nd);
}

// Element-wise kernels: bpe = bytes per element read and written (0: not stated, 8 assumed).
inline void ledger_elementwise(Context &c, int64_t n, int bpe) {
    if (bpe > 0) c.ledger.add((double)n * bpe);
    else c.ledger.add_floor(n);
}

template <int ND, typename F>
void Reducer::map_reduce(int64_t n, F f, int slot, int bpe) {
    const int nb = std::min<int64_t>(kReduceBlocks, std::max<int64_t>(1, nblk(n)));
    if ((int64_t)c.red64.n < (int64_t)nb * ND) throw Error(STK_E_STATE, "reduction buffer too small");
    k_reduce_map<NT, ND, F><<<dim3(nb), dim3(NT), 0, c.stream>>>(n, f, c.red64.p);
    STK_HIP(hipGetLastError());
    ledger_elementwise(c, n, bpe);
    finish(nb, slot, ND);
}

template <typename F>
void launch_map(Context &c, int64_t n, F f, int bpe = 0) {
    if (n == 0) return;
    k_map<NT, F><<<dim3(nblk(n)), dim3(NT), 0, c.stream>>>(n, f);
    STK_HIP(hipGetLastError());
    ledger_elementwise(c, n, bpe);
}

// Bytes of a block sparse product: structural blocks (values and a 32-bit column each), row or
// slice pointers, the gathered vector once (xvec vectors of ncols * BC), the epilogue's row vectors.
inline double spmv_ledger_bytes(const Pattern &p, int BR, int BC, int vsize, int tsize, int xvec, int rowvec) {
    const double ptrs = p.sell ? 8.0 * ((p.nrows + Pattern::kSlice - 1) / Pattern::kSlice + 1) : 4.0 * (p.nrows + 1);
    return (double)p.nnz() * (BR * BC * vsize + 4) + ptrs + (double)xvec * p.ncols * BC * tsize +
           (double)rowvec * p.nrows * BR * tsize;
}

template <typename S, typename D>
void convert_vec(Context &c, const S *src, D *dst, int64_t n) {
    launch_map(c, n, [=] __device__(int64_t i) { dst[i] = (D)src[i]; });
}

// Node-block SpMV y = A x (plus epilogue), L lanes per row node.
template <int BR, int BC, int L, typename T, typename Ep>
int64_t spmv(Context &c, const Pattern &p, const T *val, const T *x, Ep ep, double *partial) {
    const int64_t rows_per_block = NT / L;
    const int64_t nb = (p.nrows + rows_per_block - 1) / rows_per_block;
    if (partial && (int64_t)c.red64.n < nb * Ep::kDots) throw Error(STK_E_STATE, "reduction buffer too small");
    k_spmv_block<BR, BC, L, NT, T, Ep><<<dim3(nb), dim3(NT), 0, c.stream>>>(p.nrows, p.dptr.p, p.dcol.p, val, x, ep, partial, nb);
    STK_HIP(hipGetLastError());
    if (c.ledger.on) c.ledger.add(spmv_ledger_bytes(p, BR, BC, sizeof(T), sizeof(T), 1, ep.rowvec()));
    return nb;
}

// ------------------------------------------------------------------ operator wrappers
template <typename T>
struct Op22 {   // BSR 2x2 operator
    const Pattern *pat;
    const T *val;
};

// Sliced 2x2 SpMV; xl gathers the x pair of a column node (XPlain: stored x).
template <typename VT, typename T, typename Ep, typename XL>
int64_t spmv_sell_x(Context &c, const Pattern &p, const VT *val, XL xl, Ep ep, double *partial) {
    const int64_t slices_per_block = NT / Pattern::kSlice;
    const int64_t nslice = (p.nrows + Pattern::kSlice - 1) / Pattern::kSlice;
    const int64_t nb = (nslice + slices_per_block - 1) / slices_per_block;
    if (partial && (int64_t)c.red64.n < nb * Ep::kDots) throw Error(STK_E_STATE, "reduction buffer too small");
    k_spmv_sell22<NT, VT, T, Ep, XL><<<dim3(nb), dim3(NT), 0, c.stream>>>(p.nrows, p.dsptr.p, p.dscol.p, val, xl, ep,
                                                                       partial, nb);
    STK_HIP(hipGetLastError());
    if (c.ledger.on) c.ledger.add(spmv_ledger_bytes(p, 2, 2, sizeof(VT), sizeof(T), xl.xvec(), ep.rowvec()));
    return nb;
}

template <typename VT, typename T, typename Ep>
int64_t spmv_sell(Context &c, const Pattern &p, const VT *val, const T *x, Ep ep, double *partial) {
    return spmv_sell_x<VT, T, Ep>(c, p, val, XPlain<T>{x}, ep, partial);
}

// G (2x1) / G^T (1x2) product: sliced kernel when the pattern is sliced, else CSR.
template <int BR, int BC, typename T, typename Ep>
int64_t apply_rc(Context &c, const Pattern &p, const T *val, const T *x, Ep ep, double *partial = nullptr) {
    if (!p.sell) return spmv<BR, BC, 8, T, Ep>(c, p, val, x, ep, partial);
    const int64_t slices_per_block = NT / Pattern::kSlice;
    const int64_t nslice = (p.nrows + Pattern::kSlice - 1) / Pattern::kSlice;
    const int64_t nb = (nslice + slices_per_block - 1) / slices_per_block;
    if (partial && (int64_t)c.red64.n < nb * Ep::kDots) throw Error(STK_E_STATE, "reduction buffer too small");
    k_spmv_sell_rc<BR, BC, NT, T, Ep><<<dim3(nb), dim3(NT), 0, c.stream>>>(p.nrows, p.dsptr.p, p.dscol.p, val, x,
                                                                            ep, partial, nb);
    STK_HIP(hipGetLastError());
    if (c.ledger.on) c.ledger.add(spmv_ledger_bytes(p, BR, BC, sizeof(T), sizeof(T), 1, ep.rowvec()));
    return nb;
}

// Scalar SpMV on a sliced node pattern.
template <typename T, typename Ep>
int64_t spmv_sell11(Context &c, const Pattern &p, const T *val, const T *x, Ep ep, double *partial) {
    const int64_t slices_per_block = NT / Pattern::kSlice;
    const int64_t nslice = (p.nrows + Pattern::kSlice - 1) / Pattern::kSlice;
    const int64_t nb = (nslice + slices_per_block - 1) / slices_per_block;
    if (partial && (int64_t)c.red64.n < nb * Ep::kDots) throw Error(STK_E_STATE, "reduction buffer too small");
    k_spmv_sell11<NT, T, Ep><<<dim3(nb), dim3(NT), 0, c.stream>>>(p.nrows, p.dsptr.p, p.dscol.p, val, x, ep,
                                                                   partial, nb);
    STK_HIP(hipGetLastError());
    if (c.ledger.on) c.ledger.add(spmv_ledger_bytes(p, 1, 1, sizeof(T), sizeof(T), 1, ep.rowvec()));
    return nb;
}

template <int L, typename T, typename Ep>
int64_t apply22(Context &c, const Op22<T> &A, const T *x, Ep ep, double *partial = nullptr) {
    if (A.pat->sell) return spmv_sell<T, T, Ep>(c, *A.pat, A.val, x, ep, partial);
    return spmv<2, 2, L, T, Ep>(c, *A.pat, A.val, x, ep, partial);
}

// lanes per row: P2 rows carry up to 19 blocks, P1 rows 7, transfer rows 1-2 / up to 13.
template <typename T, typename Ep>
int64_t apply22_auto(Context &c, const Op22<T> &A, const T *x, Ep ep, double *partial = nullptr) {
    const double avg = (double)A.pat->nnz() / std::max<int64_t>(1, A.pat->nrows);
    if (avg > 10) return apply22<8>(c, A, x, ep, partial);
    if (avg > 3) return apply22<4>(c, A, x, ep, partial);
    return apply22<2>(c, A, x, ep, partial);
}

// y = A1 x1 - A2 x2 for two operators with the same rows in one launch (both sliced, or both
// node-block CSR with lanes per row as apply22_auto picks them); other layout pairs (small
// meshes only) take two launches.
template <typename T>
void apply22_restrict(Context &c, const Op22<T> &A1, const T *x1, const Op22<T> &A2, const T *x2, T *y) {
    const Pattern &p1 = *A1.pat, &p2 = *A2.pat;
    if (p1.nrows != p2.nrows) throw Error(STK_E_STATE, "restriction: operators with different rows");
    if (p1.sell != p2.sell) {
        EpStore<T> es{y};
        apply22_auto<T>(c, A1, x1, es);
        EpAxpyOut<T> ea{y, T(-1), T(1)};
        apply22_auto<T>(c, A2, x2, ea);
        return;
    }
    if (p1.sell) {
        const int64_t nslice = (p1.nrows + Pattern::kSlice - 1) / Pattern::kSlice;
        const int64_t nb = (nslice + NT / Pattern::kSlice - 1) / (NT / Pattern::kSlice);
        k_spmv_dual_sell22<NT, T><<<dim3((unsigned)nb), dim3(NT), 0, c.stream>>>(
            p1.nrows, p1.dsptr.p, p1.dscol.p, A1.val, x1, p2.dsptr.p, p2.dscol.p, A2.val, x2, y);
    } else {
        const double avg = (double)(p1.nnz() + p2.nnz()) / std::max<int64_t>(1, p1.nrows);
        auto launch = [&](auto lanes) {
            constexpr int L = decltype(lanes)::value;
            const int64_t nb = (p1.nrows + NT / L - 1) / (NT / L);
            k_spmv_dual_block22<L, NT, T><<<dim3((unsigned)nb), dim3(NT), 0, c.stream>>>(
                p1.nrows, p1.dptr.p, p1.dcol.p, A1.val, x1, p2.dptr.p, p2.dcol.p, A2.val, x2, y);
        };
        if (avg > 10) launch(std::integral_constant<int, 8>{});
        else launch(std::integral_constant<int, 4>{});
    }
    STK_HIP(hipGetLastError());
    if (c.ledger.on)
        c.ledger.add(spmv_ledger_bytes(p1, 2, 2, sizeof(T), sizeof(T), 1, 0) +
                     spmv_ledger_bytes(p2, 2, 2, sizeof(T), sizeof(T), 1, 1));
}

template <typename T>
void power_lmax(Context &c, const Op22<T> &A, const T *dinv, int64_t n, T *v, T *w, double *out);

// ------------------------------------------------------------------ multigrid
template <typename T>
struct MGView {
    struct Lv {
        Op22<T> A;            // operator
        const T *dinv;        // inverse diagonal
        double lmax;          // upper eigenvalue estimate of D^-1 A
        int64_t n;            // dofs
        T *b, *x, *xa, *r, *d;
        Op22<T> P, R;         // transfer to / from the next coarser level (P: this <- coarser)
        Op22<T> RK;           // R A: R K on level 0, R A of the P1 levels (pattern null when not built)
    };
    std::vector<Lv> lv;
    const T *coarse_inv = nullptr;
    int64_t coarse_n = 0;
    bool half0 = false;       // level 0 smooths with the FP16 scaled copy of K (FP32 only)
    bool gs0 = false;         // level 0 smooths with multicolour block Gauss-Seidel (FP32 only)
};

template <typename T>
MGLevelBufs<T> &level_bufs(MGLevel &L) { return pick<T>(L.w64, L.w32); }

template <typename T>
MGView<T> make_view(Context &c) {
    MGView<T> v;
    const bool f64 = sizeof(T) == 8;
    // level 0: the P2 operator K
    typename MGView<T>::Lv l0{};
    l0.A = {&c.KN, (const T *)pick<T>(c.K64, c.K32).p};
    l0.dinv = (const T *)pick<T>(c.Kdinv64, c.Kdinv32).p;
    l0.lmax = f64 ? c.Klmax64 : c.Klmax32;
    l0.n = c.nU;
    MGLevelBufs<T> &w0 = pick<T>(c.kw64, c.kw32);
    l0.b = w0.b.p; l0.x = w0.x.p; l0.xa = w0.xa.p; l0.r = w0.r.p; l0.d = w0.d.p;
    if (!c.mg.empty()) {
        l0.P = {&c.mg[0].P_pat, f64 ? (const T *)c.mg[0].P64.p : (const T *)c.mg[0].P32.p};
        l0.R = {&c.mg[0].R_pat, f64 ? (const T *)c.mg[0].R64.p : (const T *)c.mg[0].R32.p};
        const T *rk = (const T *)pick<T>(c.RK64, c.RK32).p;
        if (rk) l0.RK = {&c.GTN, rk};
    }
    v.lv.push_back(l0);
    for (size_t i = 0; i < c.mg.size(); ++i) {
        MGLevel &L = c.mg[i];
        typename MGView<T>::Lv l{};
        l.A = {&L.pat, (const T *)pick<T>(L.A64, L.A32).p};
        l.dinv = (const T *)pick<T>(L.dinv64, L.dinv32).p;
        l.lmax = f64 ? L.lmax64 : L.lmax32;
        l.n = 2 * L.nvert;
        MGLevelBufs<T> &w = level_bufs<T>(L);
        l.b = w.b.p; l.x = w.x.p; l.xa = w.xa.p; l.r = w.r.p; l.d = w.d.p;
        if (i + 1 < c.mg.size()) {
            MGLevel &C = c.mg[i + 1];
            l.P = {&C.P_pat, f64 ? (const T *)C.P64.p : (const T *)C.P32.p};
            l.R = {&C.R_pat, f64 ? (const T *)C.R64.p : (const T *)C.R32.p};
            const T *ra = (const T *)pick<T>(C.RA64, C.RA32).p;
            if (ra) l.RK = {&C.RA_pat, ra};
        }
        v.lv.push_back(l);
    }
    v.coarse_inv = (const T *)pick<T>(c.coarse_inv64, c.coarse_inv32).p;
    v.coarse_n = c.coarse_n;
    v.half0 = !f64 && c.ksm.on && v.lv.size() > 1;
    v.gs0 = !f64 && !v.half0 && c.gs.on && c.gs.filled && c.fine_smoother == STK_SMOOTH_GS && v.lv.size() > 1 &&
            v.lv[0].RK.pat != nullptr;
    if (!v.gs0 && v.lv.size() > 1 && v.lv[0].lmax == 0.0 && v.lv[0].A.val) {
        // level 0 smooths with Chebyshev, whose bound the set-up skipped: compute it now (the level-0
        // work vectors are free between V-cycles)
        double lm = 0.0;
        MGLevelBufs<T> &kw = pick<T>(c.kw64, c.kw32);
        power_lmax<T>(c, v.lv[0].A, v.lv[0].dinv, c.nU, kw.r.p, kw.d.p, &lm);
        (f64 ? c.Klmax64 : c.Klmax32) = lm;
        v.lv[0].lmax = lm;
    }
    return v;
}

// Chebyshev smoothing of A x = b with point Jacobi on [lo, hi] = [0.3, 1.1] * lmax.
// x0 is zero when `zero_start`; the result is written to *xout (one of x / xa).
// With dot_w (non-zero start only) the last step also accumulates (b . x, dot_w . x) of its
// result into block partials; *dot_nb receives their count.
template <typename T>
void chebyshev(Context &c, const typename MGView<T>::Lv &l, const T *b, T *x, T *xa, T *d, int deg,
               bool zero_start, T **xout, const T *dot_w = nullptr, int64_t *dot_nb = nullptr) {
    const double hi = 1.1 * l.lmax, lo = 0.3 * hi;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    T *cur = x, *nxt = xa;
    const T *dinv = l.dinv;
    if (zero_start && deg > 1 && l.A.pat->sell) {
        // x0 = d0 = D^-1 b / theta is never stored: the first step gathers it from b and
        // D^-1 and forms it again in the epilogue (one pass and one launch fewer)
        const T it = (T)(1.0 / theta);
        const double rn = 1.0 / (2.0 * sigma - rho);
        EpChebFirst<T> ep{b, dinv, d, nxt, it, (T)(rn * rho), (T)(2.0 * rn / delta), deg > 2};
        spmv_sell_x<T, T>(c, *l.A.pat, l.A.val, XJacobi<T>{b, dinv, it}, ep, nullptr);
        std::swap(cur, nxt);
        rho = rn;
        for (int k = 2; k < deg; ++k) {
            const double rk = 1.0 / (2.0 * sigma - rho);
            EpCheb<T> ek{b, dinv, cur, d, nxt, (T)(rk * rho), (T)(2.0 * rk / delta), true, k < deg - 1};
            apply22_auto<T>(c, l.A, cur, ek);
            std::swap(cur, nxt);
            rho = rk;
        }
        *xout = cur;
        return;
    }
    if (zero_start) {
        const T it = (T)(1.0 / theta);
        launch_map(c, l.n, [=] __device__(int64_t i) {
            const T v = it * dinv[i] * b[i];
            d[i] = v;
            cur[i] = v;
        });
    }
    // the first step of a non-zero start does not read d (c1 = 0), the last does not store it
    auto step = [&](double c1, double c2, bool first, bool last) {
        if (last && dot_w) {
            EpChebDot2<T> ep{b, dinv, cur, d, nxt, (T)c1, (T)c2, dot_w, !first};
            *dot_nb = apply22_auto<T>(c, l.A, cur, ep, c.red64.p);
        } else {
            EpCheb<T> ep{b, dinv, cur, d, nxt, (T)c1, (T)c2, !first, !last};
            apply22_auto<T>(c, l.A, cur, ep);
        }
        std::swap(cur, nxt);
    };
    // d = D^-1 (b - A x) / theta ; x += d
    if (!zero_start) step(0.0, 1.0 / theta, true, deg == 1);
    for (int k = 1; k < deg; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        step(rn * rho, 2.0 * rn / delta, false, k == deg - 1);
        rho = rn;
    }
    *xout = cur;
}

template <typename T>
T *vcycle(Context &c, MGView<T> &v, size_t l, const T *b, int deg, const T *dot_w, int64_t *dot_nb);

// Level 0 of an FP32 V-cycle with the smoother on the FP16 scaled copy K^ = S^-1 K S^-1:
// Jacobi-Chebyshev on K is Chebyshev on K^ in the variables x^ = S x, b^ = S^-1 b (same
// coefficients, same spectrum).  The residual passed to the coarse level is formed with
// the FP32 K on the unscaled iterate: smooth error components cancel to ~1e-4 of |K||x|,
// below FP16 resolution but well inside FP32.
float *vcycle0_half(Context &c, MGView<float> &v, const float *b, int deg) {
    auto &L = v.lv[0];
    Context::HalfSmoother &H = c.ksm;
    const _Float16 *Ah = reinterpret_cast<const _Float16 *>(H.A.p);
    const double hi = 1.1 * L.lmax, lo = 0.3 * hi;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta;
    float *bh = H.bh.p, *xu = H.xu.p, *d = L.d;
    const float *sinv = H.sinv.p, *s = H.s.p;
    float *cur = H.xh.p, *nxt = H.xha.p;
    auto step = [&](double c1, double c2, bool last) {
        EpChebScaled ep{bh, cur, d, nxt, (float)c1, (float)c2, last ? xu : nullptr, sinv};
        spmv_sell<_Float16, float>(c, c.KN, Ah, cur, ep, nullptr);
        std::swap(cur, nxt);
    };
    // pre-smoothing from zero: x^ = d = b^ / theta
    {
        const float it = (float)(1.0 / theta);
        float *xh = cur;
        const bool one = deg == 1;
        launch_map(c, L.n, [=] __device__(int64_t i) {
            const float bv = b[i] * sinv[i];
            const float v0 = it * bv;
            bh[i] = bv;
            d[i] = v0;
            xh[i] = v0;
            if (one) xu[i] = v0 * sinv[i];
        });
    }
    double rho = 1.0 / sigma;
    for (int k = 1; k < deg; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        step(rn * rho, 2.0 * rn / delta, k == deg - 1);
        rho = rn;
    }
    // coarse-grid correction on the FP32 residual
    EpResidual<float> er{L.r, b};
    apply22<8>(c, L.A, xu, er);
    auto &C = v.lv[1];
    EpStore<float> es{C.b};
    apply22_auto<float>(c, L.R, L.r, es);
    float *xc = vcycle<float>(c, v, 1, C.b, deg, nullptr, nullptr);
    EpAxpyScaled ea{xu, cur, s};
    apply22_auto<float>(c, L.P, xc, ea);
    // post-smoothing from the corrected iterate
    rho = 1.0 / sigma;
    step(0.0, 1.0 / theta, deg == 1);
    for (int k = 1; k < deg; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        step(rn * rho, 2.0 * rn / delta, k == deg - 1);
        rho = rn;
    }
    return xu;
}

// One multicolour block Gauss-Seidel sweep on K (FP32) with the iterate in coloured row
// order (c.gs.xq): GS_FROM_ZERO forward from x = 0 (colours in order, lower-colour blocks
// only; b in node order, copied to c.gs.bq), GS_BACKWARD (colours in reverse order, all
// blocks; the result also written in node order to z).  With dots a backward sweep
// accumulates (b . x, v . x), v the near-null vector, colour by colour into block partials;
// *dot_nb receives their count.  (Two or three sweeps each way cut the inner iterations by
// 4-12% at 256x2048 and 512x4096 and cost more than they saved.)
enum GsMode { GS_FROM_ZERO, GS_BACKWARD };

// Algorithmic bytes of one colour launch: the blocks it reads (four values and a column), the x
// of every neighbour node of its rows (x changes between launches, so each launch gathers each
// once), and per real row from zero: b read, b and x written in coloured order, D^-1, row id,
// lower-block count; backward: b and the row's own x read, x written in both orders, D^-1, row
// id (and w with the dots); slice pointers (and lengths).
double gs_launch_bytes(const Context::GSSmoother &g, int k, bool zero, bool dots, int64_t nslices) {
    const double nsl = (double)nslices;
    if (zero)
        return (double)g.blocks_low[k] * 20 + (double)g.nbr_low[k] * 8 + (double)g.rows[k] * 45 + nsl * 12;
    return (double)g.blocks[k] * 20 + (double)g.nbr[k] * 8 + (double)g.rows[k] * (dots ? 60 : 52) + (nsl + 1) * 8;
}
void gs_sweep(Context &c, GsMode mode, const float *b, float *z, bool dots, int64_t *dot_nb) {
    Context::GSSmoother &g = c.gs;
    constexpr int NTG = 256;
    const int64_t spb = NTG / kWave;
    auto nblk = [&](int k) { return (g.slice_off[k + 1] - g.slice_off[k] + spb - 1) / spb; };
    auto launch = [&](int k, bool zero, double *part, int64_t pstride) {
        const int64_t nb = nblk(k);
        if (nb == 0) return;
        const dim3 grid((unsigned)nb), block(NTG);
        const int64_t s0 = g.slice_off[k], s1 = g.slice_off[k + 1], r0 = g.row_off[k];
#define STK_GS_ARGS s0, s1, r0, g.sptr.p, g.col.p, g.val.p, g.rowid.p, g.slow.p, g.nlow.p, g.dinv.p, b, g.bq.p, \
                    g.xq.p, z, g.wq.p, part, pstride
        if (zero) k_gs_colour<NTG, true, 0><<<grid, block, 0, c.stream>>>(STK_GS_ARGS);
        else if (part) k_gs_colour<NTG, false, 2><<<grid, block, 0, c.stream>>>(STK_GS_ARGS);
        else k_gs_colour<NTG, false, 0><<<grid, block, 0, c.stream>>>(STK_GS_ARGS);
#undef STK_GS_ARGS
        STK_HIP(hipGetLastError());
        if (c.ledger.on) c.ledger.add(gs_launch_bytes(g, k, zero, part != nullptr, s1 - s0));
    };
    if (mode == GS_FROM_ZERO) {
        for (int k = 0; k < g.ncolour; ++k) launch(k, true, nullptr, 0);
        return;
    }
    int64_t ntot = 0;
    for (int k = 0; k < g.ncolour; ++k) ntot += nblk(k);
    if (dots && (int64_t)c.red64.n < 2 * ntot) throw Error(STK_E_STATE, "reduction buffer too small");
    int64_t off = 0;
    for (int k = g.ncolour - 1; k >= 0; --k) {
        launch(k, false, dots ? c.red64.p + off : nullptr, ntot);
        off += nblk(k);
    }
    if (dots) *dot_nb = ntot;
}

// Level 0 of an FP32 V-cycle with the Gauss-Seidel smoother: forward sweep from zero,
// coarse right-hand side R b - (R K) x, coarse correction x += P x_c, backward sweep (the
// pair is symmetric, as CG needs).  x stays in coloured row order throughout; R K and P
// address it through their renumbered columns / rows.  Returns x in node order (lv[0].x).
float *vcycle0_gs(Context &c, MGView<float> &v, const float *b, int deg, const float *dot_w, int64_t *dot_nb) {
    auto &L = v.lv[0];
    Context::GSSmoother &g = c.gs;
    if (dot_w && dot_w != c.rot32.p) throw Error(STK_E_STATE, "V-cycle dots are taken with the near-null vector");
    float *z = L.x;
    gs_sweep(c, GS_FROM_ZERO, b, nullptr, false, nullptr);
    auto &C = v.lv[1];
    apply22_restrict<float>(c, L.R, b, Op22<float>{&g.rkq, L.RK.val}, g.xq.p, C.b);
    float *xc = vcycle<float>(c, v, 1, C.b, deg, nullptr, nullptr);
    EpAxpyOut<float> ea{g.xq.p, 1.0f, 1.0f};
    apply22_auto<float>(c, Op22<float>{&g.pq, g.pq_val.p}, xc, ea);
    gs_sweep(c, GS_BACKWARD, nullptr, z, dot_w != nullptr, dot_nb);
    return z;
}

// x_l = V(b_l); result pointer returned (lives in lv[l].x or lv[l].xa).
// With dot_w, the level-0 result x also gets its dots (b . x, dot_w . x) accumulated
// into block partials by the last smoothing step; *dot_nb = their count (0: not done).
template <typename T>
T *vcycle(Context &c, MGView<T> &v, size_t l, const T *b, int deg, const T *dot_w, int64_t *dot_nb) {
    if (dot_nb) *dot_nb = 0;
    if constexpr (sizeof(T) == 4) {
        if (l == 0 && v.half0) return vcycle0_half(c, v, b, deg);
        if (l == 0 && v.gs0) return vcycle0_gs(c, v, b, deg, dot_w, dot_nb);
    }
    auto &L = v.lv[l];
    if (l + 1 == v.lv.size()) {
        // coarsest: dense inverse times b
        T *x = L.x;
        k_dense_matvec<T><<<dim3((unsigned)((v.coarse_n + 3) / 4)), dim3(256), 0, c.stream>>>(v.coarse_n, v.coarse_inv,
                                                                                           b, x);
        STK_HIP(hipGetLastError());
        c.ledger.add(((double)v.coarse_n * v.coarse_n + 2.0 * v.coarse_n) * sizeof(T));
        return x;
    }
    T *x;
    chebyshev<T>(c, L, b, L.x, L.xa, L.d, deg, true, &x);
    auto &C = v.lv[l + 1];
    if (L.RK.pat) {
        // b_c = R b - (R A) x in one pass: R A has the coarse level's rows, a quarter of A's
        apply22_restrict<T>(c, L.R, b, L.RK, x, C.b);
    } else {
        // r = b - A x; b_c = R r
        EpResidual<T> er{L.r, b};
        apply22_auto<T>(c, L.A, x, er);
        EpStore<T> es{C.b};
        apply22_auto<T>(c, L.R, L.r, es);
    }
    T *xc = vcycle<T>(c, v, l + 1, C.b, deg, nullptr, nullptr);
    // x += P xc
    EpAxpyOut<T> ea{x, T(1), T(1)};
    apply22_auto<T>(c, L.P, xc, ea);
    T *other = (x == L.x) ? L.xa : L.x;
    T *xo;
    chebyshev<T>(c, L, b, x, other, L.d, deg, false, &xo, dot_w, dot_nb);
    return xo;
}

// ------------------------------------------------------------------ inner CG
// Solve K z = y (BC operator) to relative tolerance rtol on the complement of the
// near-null vector v: y is overwritten by y - (v.y) v and CG runs on the projected
// operator P K P (P = I - v v^T; K p and the preconditioned residual are projected
// every iteration), so it never resolves the near-null direction whatever the
// tolerance; z comes back orthogonal to v.  With `deflate` the exact null-space
// term (v.y / lambda) v is added to z.
// Scalars live in device memory (slots of d_sc); the host reads the residual
// norms every `check_every` iterations.
template <typename T>
int64_t inner_cg(Context &c, MGView<T> &v, bool use_mg, T *y, T *z, double rtol, int itmax, int deg,
                 bool *converged, bool deflate = false, bool warm = false) {
    Reducer red(c);
    const int64_t n = c.nU;
    const bool f64 = sizeof(T) == 8;
    MGLevelBufs<T> &w = pick<T>(c.cg64, c.cg32);
    T *r = w.r.p, *p = w.d.p, *Ap = w.b.p;
    const T *rot = (const T *)pick<T>(c.rot64, c.rot32).p;
    const T *kdinv = (const T *)pick<T>(c.Kdinv64, c.Kdinv32).p;
    double *sc = c.d_sc.p;
    Op22<T> K{&c.KN, (const T *)pick<T>(c.K64, c.K32).p};

    // project the near-null component out of y; r = y; z = 0 (or the given z); |y|^2
    constexpr int B = (int)sizeof(T);
    red.map_reduce(n, [=] __device__(int64_t i) { return (double)rot[i] * (double)y[i]; }, SC_VY, 2 * B);
    red.map_reduce(n, [=] __device__(int64_t i) {
        const T v = y[i] - (T)sc[SC_VY] * rot[i];
        y[i] = v;
        r[i] = v;
        if (!warm) z[i] = T(0);
        return (double)v * (double)v;
    }, SC_BB, (warm ? 4 : 5) * B);
    if (warm) {
        // start from the given z (orthogonal to v): r = P (y - K z)
        EpResidual<T> er{r, y};
        apply22<8>(c, K, z, er);
        red.map_reduce(n, [=] __device__(int64_t i) { return (double)rot[i] * (double)r[i]; }, SC_TMP, 2 * B);
        launch_map(c, n, [=] __device__(int64_t i) { r[i] -= (T)sc[SC_TMP] * rot[i]; }, 3 * B);
    }
    const double lam = c.null_lambda;
    auto finish_null = [&]() {
        // z <- z - (v.z) v  (+ (v.y / lambda) v when deflating)
        red.map_reduce(n, [=] __device__(int64_t i) { return (double)rot[i] * (double)z[i]; }, SC_VZ, 2 * B);
        const double dl = deflate ? 1.0 / lam : 0.0;
        launch_map(c, n, [=] __device__(int64_t i) {
            z[i] += (T)(dl * sc[SC_VY] - sc[SC_VZ]) * rot[i];
        }, 3 * B);
    };
    const double bb = red.fetch(SC_BB);
    if (!std::isfinite(bb)) throw Error(STK_E_NAN, "non-finite right-hand side in the inner CG");
    if (!(bb > 0)) {
        *converged = true;
        finish_null();
        return 0;
    }
    const double tol2 = rtol * rtol * bb;
    // zz = M^-1 rin (the V-cycle's own output buffer, or w.x for Jacobi) together with
    // (rin . zz, v . zz) in slots (slot, slot + 1): from the V-cycle's last smoothing
    // step when it can return them, else from one fused pass.
    auto precond = [&](const T *rin, int slot) -> T * {
        if (use_mg) {
            int64_t nbd = 0;
            T *res = vcycle<T>(c, v, 0, rin, deg, rot, &nbd);
            if (nbd > 0) {
                red.finish(nbd, slot, 2);
            } else {
                red.map_reduce<2>(n, [=] __device__(int64_t i) {
                    return Dots<2>{{(double)rin[i] * (double)res[i], (double)rot[i] * (double)res[i]}};
                }, slot, 3 * (int)sizeof(T));
            }
            return res;
        }
        T *zo = w.x.p;
        red.map_reduce<2>(n, [=] __device__(int64_t i) {
            const T zv = kdinv[i] * rin[i];
            zo[i] = zv;
            return Dots<2>{{(double)rin[i] * (double)zv, (double)rot[i] * (double)zv}};
        }, slot, 4 * (int)sizeof(T));
        return zo;
    };
    // first direction: p = zz - (v . zz) v, r . z taken with the projected zz
    {
        T *zz = precond(r, SC_RZV);
        launch_map(c, n, [=] __device__(int64_t i) { p[i] = zz[i] - (T)sc[SC_RZV + 1] * rot[i]; }, 3 * B);
        red.map_reduce(n, [=] __device__(int64_t i) { return (double)r[i] * (double)p[i]; }, SC_RZV, 2 * B);
    }
    int par = 0;
    const int check_every = 2;
    int64_t it = 0;
    *converged = false;
    double hist[8];
    double best = bb;   // smallest |r|^2 seen at a check
    while (it < itmax) {
        const int rz_cur = SC_RZV + 2 * par, rz_new = SC_RZV + 2 * (1 - par);
        // Ap = K p with (p . Ap, v . Ap)
        EpStoreDot2<T> ed{Ap, p, rot};
        const int64_t nb = apply22<8>(c, K, p, ed, c.red64.p);
        red.finish(nb, SC_PAP2, 2);
        // z += al p, r -= al P Ap, with |r|^2
        const int slot_rr = SC_HIST + (int)(it % check_every);
        red.map_reduce(n, [=] __device__(int64_t i) {
            const T al = (T)(sc[rz_cur] / sc[SC_PAP2]);
            z[i] += al * p[i];
            const T rv = r[i] - al * (Ap[i] - (T)sc[SC_PAP2 + 1] * rot[i]);
            r[i] = rv;
            return (double)rv * (double)rv;
        }, slot_rr, 7 * B);
        ++it;
        if (it % check_every == 0 || it == itmax) {
            if (trace_on() && it == itmax) std::fprintf(stderr, "[stk] inner CG hit its cap %d\n", itmax);
            const int m = (it % check_every == 0) ? check_every : (int)(it % check_every);
            red.fetch_many(SC_HIST, m, hist);
            bool bad = false, done = false;
            for (int k = 0; k < m; ++k) {
                if (!std::isfinite(hist[k])) bad = true;
                if (hist[k] <= tol2) done = true;
            }
            if (bad) throw Error(STK_E_NAN, "non-finite residual in the inner CG");
            if (done) { *converged = true; break; }
            // A tolerance below the attainable accuracy: once the recursive residual has stagnated
            // CG drifts and the residual climbs.  Stop when it is 100x (in norm) above the smallest
            // one seen, so the returned iterate stays near the attainable floor.
            bool grew = false;
            for (int k = 0; k < m; ++k) grew |= hist[k] > 1e6 * bb || hist[k] > 1e4 * best;
            if (grew) {
                if (trace_on()) std::fprintf(stderr, "[stk] inner CG stagnated at %.3e\n", std::sqrt(best / bb));
                break;
            }
            for (int k = 0; k < m; ++k) best = std::min(best, hist[k]);
        }
        // zz = M^-1 r with (r . zz, v . zz); p = P zz + be p
        T *zz = precond(r, rz_new);
        launch_map(c, n, [=] __device__(int64_t i) {
            const T be = (T)(sc[rz_new] / sc[rz_cur]);
            p[i] = zz[i] - (T)sc[rz_new + 1] * rot[i] + be * p[i];
        }, 4 * B);
        par = 1 - par;
    }
    finish_null();
    if (trace_on()) {
        // true residual of the returned solution on the complement: |y - K z| / |y|
        T *t = Ap;
        EpResidual<T> er{t, y};
        const int64_t nb = apply22<8>(c, K, z, er, c.red64.p);
        red.finish(nb, SC_TMP);
        const double tr = red.fetch(SC_TMP);
        std::fprintf(stderr, "[stk] inner CG %lld its, recursive |r|/|b| %.3e, true %.3e\n", (long long)it,
                     std::sqrt(hist[0] / bb), std::sqrt(tr / bb));
    }
    return it;
}

// ------------------------------------------------------------------ set-up of the solver data
template <typename T>
void power_lmax(Context &c, const Op22<T> &A, const T *dinv, int64_t n, T *v, T *w, double *out) {
    Reducer red(c);
    launch_map(c, n, [=] __device__(int64_t i) {
        uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x12345ull;
        h ^= h >> 31; h *= 0xBF58476D1CE4E5B9ull; h ^= h >> 29;
        v[i] = (T)((double)(h >> 11) * 0x1.0p-53 - 0.5) * (dinv[i] != T(0) ? T(1) : T(0));
    });
    // the iteration stays on the device: the scale of each step is read from the reduction slot,
    // and the host fetches the last Rayleigh quotient only
    const double *sc = c.d_sc.p;
    for (int k = 0; k < 20; ++k) {
        EpStore<T> es{w};
        apply22_auto<T>(c, A, v, es);
        // Rayleigh quotient in the D inner product: (v, A v) / (v, D v)
        red.map_reduce<2>(n, [=] __device__(int64_t i) {
            const double vi = (double)v[i];
            return Dots<2>{{vi * (double)w[i], dinv[i] != T(0) ? vi * vi / (double)dinv[i] : 0.0}};
        }, SC_TMP);
        // v <- D^-1 A v, normalised in the D norm
        launch_map(c, n, [=] __device__(int64_t i) {
            const T s = (T)(1.0 / sqrt(fmax(sc[SC_TMP + 1], 1e-300)));
            v[i] = s * dinv[i] * w[i];
        });
    }
    double h[2];
    red.fetch_many(SC_TMP, 2, h);
    *out = h[0] / h[1];
}

template <typename T>
void diag_inverse(Context &c, const Pattern &p, const T *A, T *dinv) {
    const int32_t *ptr = p.dptr.p, *col = p.dcol.p, *scol = p.sell ? p.dscol.p : nullptr;
    const int64_t *sptr = p.sell ? p.dsptr.p : nullptr;
    launch_map(c, p.nrows, [=] __device__(int64_t node) {
        const int32_t len = ptr[node + 1] - ptr[node];
        for (int32_t k = 0; k < len; ++k) {
            const int64_t q = sptr ? sptr[node / Pattern::kSlice] + (int64_t)k * Pattern::kSlice + node % Pattern::kSlice
                                   : (int64_t)ptr[node] + k;
            if ((sptr ? scol[q] : col[q]) == node) {
                const T a0 = A[4 * q], a3 = A[4 * q + 3];
                dinv[2 * node] = a0 != T(0) ? T(1) / a0 : T(0);
                dinv[2 * node + 1] = a3 != T(0) ? T(1) / a3 : T(0);
                break;
            }
        }
    });
}

// Dense inverse (Gauss-Jordan, partial pivoting) of the coarsest operator, on the host.
// Dense row-major copy of a CSR 2x2 block operator (unsliced), zero-filled by the caller.
__global__ void k_dense_from_blocks(int64_t nrows, int64_t ld, const int32_t *__restrict__ ptr,
                                    const int32_t *__restrict__ col, const double *__restrict__ blk,
                                    double *__restrict__ A) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) {
        const int64_t cl = col[k];
        A[(2 * r) * ld + 2 * cl] = blk[4 * (int64_t)k];
        A[(2 * r) * ld + 2 * cl + 1] = blk[4 * (int64_t)k + 1];
        A[(2 * r + 1) * ld + 2 * cl] = blk[4 * (int64_t)k + 2];
        A[(2 * r + 1) * ld + 2 * cl + 1] = blk[4 * (int64_t)k + 3];
    }
}

// Identity on the diagonal of the padding rows n <= i < np.
__global__ void k_pad_identity(int64_t n, int64_t np, double *__restrict__ A) {
    const int64_t i = n + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) A[i * np + i] = 1.0;
}

// The leading n x n block of an np x np matrix, into a dense n x n array.
__global__ void k_copy_leading(int64_t n, int64_t np, const double *__restrict__ src, double *__restrict__ dst) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j < n) dst[i * n + j] = src[i * np + j];
}

// The lower triangle of a column-major symmetric matrix to a
// full row-major copy.
__global__ void k_symmetrize_lower_cm(int64_t n, const double *__restrict__ L, double *__restrict__ A) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t i = blockIdx.y;
    if (j >= n) return;
    // entry (i, j): column-major lower holds (max, min) at [max + min * n]
    const int64_t hi = i > j ? i : j, lo = i > j ? j : i;
    A[i * n + j] = L[hi + lo * n];
}

// Dense inverse of the coarsest multigrid operator, computed on the device in FP64 (from the FP64
// operator when the assembly has one): blocked Cholesky factorisation and inverse of the SPD matrix
// (spd_inverse, dense.hip), work on stream s (and, for the forward stage, on the context's dense
// stream, joined back into s), into out64 (n x n, row-major = column-major) and, when out32 is given,
// its FP32 rounding.  It reads the coarsest level's pattern and operator and writes no state of the
// context, so it may run on another stream beside the context stream's work (one call at a time: the
// dense stream and its events are the context's).
void coarse_inverse_on(Context &c, hipStream_t s, DBuf<double> &out64, DBuf<float> *out32) {
    MGLevel &L = c.mg.back();
    const int64_t n = 2 * L.nvert;
    if (L.pat.sell) throw Error(STK_E_STATE, "the coarsest multigrid level must use CSR block order");
    const int64_t nb = L.pat.nnz();
    DBuf<double> blk;
    const double *bp;
    const bool have64 = (c.asm_flags & STK_ASM_FP64) && L.A64;
    if (have64) {
        bp = L.A64.p;
    } else {
        blk.alloc(4 * nb);
        const float *a32 = L.A32.p;
        double *b = blk.p;
        k_map<NT><<<dim3(nblk(4 * nb)), dim3(NT), 0, s>>>(4 * nb, [=] __device__(int64_t i) { b[i] = (double)a32[i]; });
        STK_HIP(hipGetLastError());
        bp = blk.p;
    }
    {
        // non-finite entries (e.g. a NaN viscosity) are reported before the factorisation
        std::vector<double> hb(4 * nb);
        STK_HIP(hipMemcpyAsync(hb.data(), bp, 4 * nb * sizeof(double), hipMemcpyDeviceToHost, s));
        STK_HIP(hipStreamSynchronize(s));
        for (double v : hb)
            if (!std::isfinite(v)) throw Error(STK_E_NAN, "non-finite coarsest multigrid operator");
    }
    // padded to a multiple of 64 with the identity: the blocked inverse works on 64 x 64 tiles
    const int64_t np = (n + 63) / 64 * 64;
    DBuf<double> &dA = c.dense_A, &Y = c.dense_Y;
    dA.alloc(np * np);
    Y.alloc(np * np);
    dA.zero(s);
    k_dense_from_blocks<<<dim3((unsigned)((L.nvert + 255) / 256)), dim3(256), 0, s>>>(
        L.nvert, np, L.pat.dptr.p, L.pat.dcol.p, bp, dA.p);
    if (np > n) k_pad_identity<<<dim3((unsigned)((np - n + 255) / 256)), dim3(256), 0, s>>>(n, np, dA.p);
    STK_HIP(hipGetLastError());
    DBuf<int32_t> &info = c.dense_info;
    info.alloc(1);
    info.zero(s);
    // symmetric: the row-major matrix is its own column-major copy
    spd_inverse(s, c.dense_stream, c.dense_ev, c.dense_ev2, dA.p, Y.p, np, info.p);
    int32_t hinfo = 0;
    info.download(&hinfo, 1, s);
    STK_HIP(hipStreamSynchronize(s));
    if (hinfo != 0) throw Error(STK_E_SHAPE, "coarsest multigrid operator is not positive definite");
    out64.alloc(n * n);
    k_copy_leading<<<dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, s>>>(n, np, dA.p, out64.p);
    STK_HIP(hipGetLastError());
    if (out32) {
        out32->alloc(n * n);
        const double *src = out64.p;
        float *dst = out32->p;
        k_map<NT><<<dim3(nblk(n * n)), dim3(NT), 0, s>>>(n * n, [=] __device__(int64_t i) { dst[i] = (float)src[i]; });
        STK_HIP(hipGetLastError());
    }
    STK_HIP(hipStreamSynchronize(s));
}

// This assembly's coarsest inverse on the context stream (FP64, and its FP32 rounding unless f64).
void build_coarse_inverse(Context &c, bool f64) {
    MGLevel &L = c.mg.back();
    const int64_t n = 2 * L.nvert;
    if (c.coarse64_valid && c.coarse_inv64.n == (size_t)(n * n)) {
        // this assembly's coarsest operator is already inverted in FP64: its rounding serves the
        // FP32 preconditioner as well
        c.coarse_n = n;
        if (!f64) {
            c.coarse_inv32.alloc(n * n);
            convert_vec(c, c.coarse_inv64.p, c.coarse_inv32.p, n * n);
            STK_HIP(hipStreamSynchronize(c.stream));
        }
        return;
    }
    coarse_inverse_on(c, c.stream, c.coarse_inv64, f64 ? nullptr : &c.coarse_inv32);
    c.coarse_n = n;
    c.coarse64_valid = true;
}

template <typename T>
void alloc_bufs(MGLevelBufs<T> &w, int64_t n) {
    w.b.alloc(n); w.x.alloc(n); w.xa.alloc(n); w.r.alloc(n); w.d.alloc(n);
}

// FP16 copy of K^ = S^-1 K S^-1 (S^2 = diag K) in the sliced layout of K (STK_ASM_FP16_SMOOTHER).
// For SPD K every entry satisfies |K_ij| <= sqrt(K_ii K_jj), so |K^_ij| <= 1 and FP16 holds it
// with 2^-11 relative rounding (entries below 2^-24 flush to zero).  It streams 12 instead of
// 20 bytes per block, but its rounding (summed over the 38 terms of a row) perturbs the
// smoothed iterate by ~1e-2 relative, which costs more Krylov iterations than it saves at
// 256 x 2048, so the FP32 smoother is the default.
void build_half_smoother(Context &c) {
    Context::HalfSmoother &H = c.ksm;
    H.on = false;
    if (!(c.asm_flags & STK_ASM_FP16_SMOOTHER)) return;
    const int64_t nU = c.nU, nrows = c.KN.nrows;
    H.A.alloc(4 * c.KN.nstore());
    for (DBuf<float> *b : {&H.s, &H.sinv, &H.bh, &H.xh, &H.xha, &H.xu}) b->alloc(nU);
    float *s = H.s.p, *sinv = H.sinv.p;
    const float *dinv = c.Kdinv32.p;
    launch_map(c, nU, [=] __device__(int64_t i) {
        const float di = dinv[i];
        const float si = di > 0.0f ? sqrtf(di) : 1.0f;
        sinv[i] = si;
        s[i] = 1.0f / si;
    });
    const int64_t *sptr = c.KN.dsptr.p;
    const int32_t *scol = c.KN.dscol.p;
    const float *K = c.K32.p;
    _Float16 *A = reinterpret_cast<_Float16 *>(H.A.p);
    const int64_t npad = (int64_t)(c.KN.sptr.size() - 1) * Pattern::kSlice;
    launch_map(c, npad, [=] __device__(int64_t i) {
        const int64_t sl = i / Pattern::kSlice, lane = i % Pattern::kSlice;
        const int64_t base = sptr[sl], len = (sptr[sl + 1] - base) / Pattern::kSlice;
        for (int64_t k = 0; k < len; ++k) {
            const int64_t q = base + k * Pattern::kSlice + lane;
            const int64_t j = scol[q];
            for (int a = 0; a < 2; ++a)
                for (int bb = 0; bb < 2; ++bb) {
                    const float v = i < nrows ? K[4 * q + 2 * a + bb] * sinv[2 * i + a] * sinv[2 * j + bb] : 0.0f;
                    A[4 * q + 2 * a + bb] = (_Float16)v;
                }
        }
    });
    H.on = true;
}

// With STK_TRACE: wall time of set-up phases (synchronises the stream at each mark).
struct PhaseTimer {
    Context &c;
    const char *scope;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    PhaseTimer(Context &c, const char *scope) : c(c), scope(scope) {}
    void mark(const char *what) {
        if (!trace_on()) return;
        STK_HIP(hipStreamSynchronize(c.stream));
        const auto t1 = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[stk] %s %s: %.2f ms\n", scope, what,
                     std::chrono::duration<double, std::milli>(t1 - t0).count());
        t0 = t1;
    }
};

// After assembly: inverse diagonals, eigenvalue bounds, coarse inverse, rotation vector.
// Per-precision solver data that every method needs: inner-CG work vectors, the inverse diagonal of K,
// the Gauss-Seidel smoother's values (FP32) and the unrefined near-null vector (the nodal rigid
// rotation u_theta = r, u_r = 0, unit norm).
template <typename T>
void setup_base(Context &c, PhaseTimer &pt) {
    (void)pt;
    MGLevelBufs<T> &kw = pick<T>(c.kw64, c.kw32);
    MGLevelBufs<T> &cg = pick<T>(c.cg64, c.cg32);
    alloc_bufs(kw, c.nU);
    alloc_bufs(cg, c.nU);
    DBuf<T> &kd = pick<T>(c.Kdinv64, c.Kdinv32);
    kd.alloc(c.nU);
    const T *K = (const T *)pick<T>(c.K64, c.K32).p;
    diag_inverse<T>(c, c.KN, K, kd.p);
    if constexpr (sizeof(T) == 4) {
        // Gauss-Seidel smoother data: K's values in the coloured layout, inverse diagonal
        // blocks (from the FP64 K when there is one)
        Context::GSSmoother &g = c.gs;
        g.filled = false;
        if (g.on && c.mg_built) {
            g.val.alloc(4 * g.nstore);
            const int64_t nrc = g.row_off.back();
            g.dinv.alloc(4 * nrc);
            const float *K32 = c.K32.p;
            const double *K64 = c.K64.p;
            const int32_t *src = g.src.p;
            float *val = g.val.p;
            launch_map(c, g.nstore, [=] __device__(int64_t q) {
                const int64_t s = src[q];
                for (int j = 0; j < 4; ++j) val[4 * q + j] = s >= 0 ? K32[4 * s + j] : 0.0f;
            });
            // prolongation values in coloured row order, near-null vector in that order,
            // zeroed work vectors (padding rows stay zero)
            g.pq_val.alloc(4 * g.pq.nstore());
            {
                const float *P32 = c.mg[0].P32.p;
                const int32_t *ps = g.pq_src.p;
                float *pv = g.pq_val.p;
                launch_map(c, g.pq.nstore(), [=] __device__(int64_t q) {
                    const int64_t s = ps[q];
                    for (int j = 0; j < 4; ++j) pv[4 * q + j] = s >= 0 ? P32[4 * s + j] : 0.0f;
                });
            }
            {
                const int64_t nq = g.row_off.back();
                g.xq.alloc(2 * nq); g.bq.alloc(2 * nq); g.wq.alloc(2 * nq);
                float *xq = g.xq.p, *bq = g.bq.p;
                launch_map(c, 2 * nq, [=] __device__(int64_t i) {
                    xq[i] = 0.0f;
                    bq[i] = 0.0f;
                });
            }
            const int64_t *dp = g.diagpos.p;
            const int32_t *rid = g.rowid.p;
            float *di = g.dinv.p;
            launch_map(c, nrc, [=] __device__(int64_t q) {
                const int32_t i = rid[q];
                if (i < 0) {
                    for (int j = 0; j < 4; ++j) di[4 * q + j] = 0.0f;
                    return;
                }
                const int64_t p = dp[i];
                const double a = K64 ? K64[4 * p] : (double)K32[4 * p];
                const double bb = K64 ? K64[4 * p + 1] : (double)K32[4 * p + 1];
                const double cc = K64 ? K64[4 * p + 2] : (double)K32[4 * p + 2];
                const double d = K64 ? K64[4 * p + 3] : (double)K32[4 * p + 3];
                const double det = a * d - bb * cc;
                di[4 * q + 0] = (float)(d / det);
                di[4 * q + 1] = (float)(-bb / det);
                di[4 * q + 2] = (float)(-cc / det);
                di[4 * q + 3] = (float)(a / det);
            });
            // g.filled (the smoother in use) once the near-null vector is final, below
        }
    }
    // rigid rotation in the node frames: u_theta = r, u_r = 0 (zero on fixed dofs), unit norm
    DBuf<T> &rot = pick<T>(c.rot64, c.rot32);
    rot.alloc(c.nU);
    {
        const double *nr = c.d_noder.p;
        const uint8_t *fx = c.d_fixed.p;
        T *rp = rot.p;
        double s2 = 0;
        for (double r : c.node_r) s2 += r * r;
        // a fixed tangential dof pins the rotation: K has no near-null mode, the vector is zero
        // and every projection on it is the identity
        const double inv = c.rotation_free ? 1.0 / std::sqrt(s2) : 0.0;
        launch_map(c, c.nU, [=] __device__(int64_t i) {
            rp[i] = (i % 2 == 0 && !fx[i]) ? (T)(nr[i / 2] * inv) : T(0);
        });
    }
}

// The multigrid data of precision T: level work vectors, inverse diagonals, Chebyshev bounds, the
// coarsest inverse (and the FP16 smoother copy in FP32).
template <typename T>
void setup_mg(Context &c, PhaseTimer &pt, bool coarse = true) {
    const bool f64 = sizeof(T) == 8;
    MGLevelBufs<T> &kw = pick<T>(c.kw64, c.kw32);
    DBuf<T> &kd = pick<T>(c.Kdinv64, c.Kdinv32);
    const T *K = (const T *)pick<T>(c.K64, c.K32).p;
    if (c.mg_built) {
        for (MGLevel &L : c.mg) {
            alloc_bufs(level_bufs<T>(L), 2 * L.nvert);
            DBuf<T> &di = pick<T>(L.dinv64, L.dinv32);
            di.alloc(2 * L.nvert);
            diag_inverse<T>(c, L.pat, (const T *)pick<T>(L.A64, L.A32).p, di.p);
        }
        // R A of the levels from 1 on, from this precision's restrictions and level operators
        for (size_t k = 1; k < c.mg.size(); ++k) {
            MGLevel &C = c.mg[k], &F = c.mg[k - 1];
            const int64_t npos = C.RA_pat.nstore();
            if (npos == 0 || !C.ra_ptr) continue;
            DBuf<T> &ra = pick<T>(C.RA64, C.RA32);
            ra.alloc(4 * npos);
            k_pair_product<T><<<dim3((unsigned)((npos + 255) / 256)), dim3(256), 0, c.stream>>>(
                npos, C.ra_ptr.p, C.ra_pair.p, (const T *)pick<T>(C.R64, C.R32).p,
                (const T *)pick<T>(F.A64, F.A32).p, ra.p);
            STK_HIP(hipGetLastError());
        }
        // eigenvalue bounds (level 0 = K, then the P1 levels except the coarsest).  K's is needed by
        // Chebyshev smoothing on level 0: every FP64 V-cycle, and FP32 ones without the Gauss-Seidel
        // smoother (otherwise it is computed when a solve first asks for it, make_view)
        Op22<T> K0{&c.KN, K};
        double lm = 0.0;
        if (f64 || !c.gs.on || (c.asm_flags & STK_ASM_FP16_SMOOTHER))
            power_lmax<T>(c, K0, kd.p, c.nU, kw.r.p, kw.d.p, &lm);
        (f64 ? c.Klmax64 : c.Klmax32) = lm;
        for (size_t i = 0; i + 1 < c.mg.size(); ++i) {
            MGLevel &L = c.mg[i];
            MGLevelBufs<T> &w = level_bufs<T>(L);
            Op22<T> A{&L.pat, (const T *)pick<T>(L.A64, L.A32).p};
            power_lmax<T>(c, A, (const T *)pick<T>(L.dinv64, L.dinv32).p, 2 * L.nvert, w.r.p, w.d.p, &lm);
            (f64 ? L.lmax64 : L.lmax32) = lm;
            if (trace_on()) std::fprintf(stderr, "[stk] level %zu lmax %.4e\n", i + 1, lm);
        }
        pt.mark("diagonals and eigenvalue bounds");
        if (coarse) {
            build_coarse_inverse(c, f64);
            pt.mark("coarse inverse");
        }
        if (!f64) build_half_smoother(c);
    } else {
        Op22<T> K0{&c.KN, K};
        double lm;
        power_lmax<T>(c, K0, kd.p, c.nU, kw.r.p, kw.d.p, &lm);
        (f64 ? c.Klmax64 : c.Klmax32) = lm;
    }
}

// Near-null vector of precision T from its own multigrid (or the rounding of the FP64 one).
template <typename T>
void null_vector(Context &c, PhaseTimer &pt) {
    const bool f64 = sizeof(T) == 8;
    DBuf<T> &rot = pick<T>(c.rot64, c.rot32);
    const T *K = (const T *)pick<T>(c.K64, c.K32).p;
    if (!c.rotation_free) {
        if (f64) c.null_lambda = 0.0;
    } else if (!f64 && c.rot64) {
        convert_vec(c, c.rot64.p, rot.p, c.nU);
    } else {
        // Jacobi-Davidson steps v <- v - P (P K P)^-1 P K v, quadratically convergent
        MGView<T> view = make_view<T>(c);
        DBuf<T> y(c.nU), d(c.nU);
        Op22<T> K0{&c.KN, K};
        bool conv = false;
        double inv = 1.0, dn = 1.0;
        T *rp = rot.p;
        const T *dp = d.p;
        Reducer red(c);
        DBuf<T> keep(c.nU);
        T *kp = keep.p;
        double prev = 1.0;
        for (int step = 0; step < 4; ++step) {
            launch_map(c, c.nU, [=] __device__(int64_t i) { kp[i] = rp[i]; });
            EpStore<T> es{y.p};
            apply22<8>(c, K0, rot.p, es);
            inner_cg<T>(c, view, c.mg_built, y.p, d.p, f64 ? 1e-12 : 1e-6, 4000, 3, &conv);
            red.map_reduce(c.nU, [=] __device__(int64_t i) { return (double)dp[i] * (double)dp[i]; }, SC_TMP);
            dn = std::sqrt(red.fetch(SC_TMP));
            if (trace_on()) std::fprintf(stderr, "[stk] null vector step %d: |correction| %.3e conv %d\n", step, dn, (int)conv);
            // a step is kept only while the (quadratically convergent) correction still shrinks
            if (step > 0 && (!conv || !std::isfinite(dn) || dn > 0.1 * prev)) {
                launch_map(c, c.nU, [=] __device__(int64_t i) { rp[i] = kp[i]; });
                break;
            }
            launch_map(c, c.nU, [=] __device__(int64_t i) { rp[i] -= dp[i]; });
            red.map_reduce(c.nU, [=] __device__(int64_t i) { return (double)rp[i] * (double)rp[i]; }, SC_TMP);
            inv = 1.0 / std::sqrt(red.fetch(SC_TMP));
            launch_map(c, c.nU, [=] __device__(int64_t i) { rp[i] = (T)(inv * (double)rp[i]); });
            prev = dn;
            // Jacobi-Davidson converges quadratically: once a correction is below 1e-6 the
            // next one is at the noise floor of the inner solve, so the vector is final
            if (dn < (f64 ? 1e-6 : 1e-4)) break;
        }
        if (f64) {
            EpStoreDot<T> ed{y.p, rp};
            const int64_t nb = apply22<8>(c, K0, rp, ed, c.red64.p);
            red.finish(nb, SC_TMP);
            c.null_lambda = red.fetch(SC_TMP);
        }
        if (trace_on())
            std::fprintf(stderr, "[stk] null vector: correction solve conv %d, |v|^-1 %.6e, lambda %.6e, Klmax %.4e\n",
                         (int)conv, inv, c.null_lambda, f64 ? c.Klmax64 : c.Klmax32);
    }
    pt.mark("near-null vector");
}

// The Gauss-Seidel smoother's copy of the FP32 near-null vector in coloured row order.
void fill_gs_rot(Context &c) {
    Context::GSSmoother &g = c.gs;
    if (g.on && c.mg_built) {
        const int64_t nq = g.row_off.back();
        const int32_t *rid = g.rowid.p;
        const float *rp = c.rot32.p;
        float *wq = g.wq.p;
        launch_map(c, 2 * nq, [=] __device__(int64_t i) {
            const int32_t n = rid[i / 2];
            wq[i] = n >= 0 ? rp[2 * (int64_t)n + i % 2] : 0.0f;
        });
        g.filled = true;
    }
}

template <typename T>
void setup_solver_data(Context &c) {
    const bool f64 = sizeof(T) == 8;
    PhaseTimer pt(c, f64 ? "set-up FP64" : "set-up FP32");
    setup_base<T>(c, pt);
    setup_mg<T>(c, pt);
    null_vector<T>(c, pt);
    if constexpr (sizeof(T) == 4) fill_gs_rot(c);
    if (f64) c.have_mg64 = c.mg_built;
    else c.have_mg32 = c.mg_built;
    STK_HIP(hipStreamSynchronize(c.stream));
}

template void setup_solver_data<double>(Context &);
template void setup_solver_data<float>(Context &);

// Near-null vector for the FP32 methods: Jacobi-Davidson steps v <- z - d with (Pi K Pi) d = Pi K z,
// Pi = I - z z^T, the correction from FP64 iterative refinement around FP32 multigrid-CG solves (the
// multigrid the FP32 solves use).  d is needed to 1e-6 of its own size: the first correction of the
// nodal rotation is already ~1e-8 of it at 256 x 2048, so v ends accurate to ~1e-14.
void null_vector_mixed(Context &c, PhaseTimer &pt) {
    Reducer red(c);
    const int64_t n = c.nU;
    double *z = c.rot64.p;
    float *z32 = c.rot32.p;
    DBuf<double> &y = c.nv_y, &d = c.nv_d, &r = c.nv_r, &keep = c.nv_keep;
    DBuf<float> &r32 = c.nv_r32, &e32 = c.nv_e32;
    y.alloc(n); d.alloc(n); r.alloc(n); keep.alloc(n);
    r32.alloc(n); e32.alloc(n);
    double *yp = y.p, *dp = d.p, *rp = r.p, *kp = keep.p, *sc = c.d_sc.p;
    float *r32p = r32.p, *e32p = e32.p;
    Op22<double> K{&c.KN, c.K64.p};
    auto project = [&](double *w) {
        red.map_reduce(n, [=] __device__(int64_t i) { return z[i] * w[i]; }, SC_TMP);
        launch_map(c, n, [=] __device__(int64_t i) { w[i] -= sc[SC_TMP] * z[i]; });
    };
    auto norm = [&](const double *w) {
        red.map_reduce(n, [=] __device__(int64_t i) { return w[i] * w[i]; }, SC_TMP);
        return std::sqrt(red.fetch(SC_TMP));
    };
    double prev = 1.0;
    int sweeps = 0;
    for (int step = 0; step < 4; ++step) {
        convert_vec(c, z, z32, n);
        fill_gs_rot(c);
        MGView<float> view = make_view<float>(c);
        launch_map(c, n, [=] __device__(int64_t i) { kp[i] = z[i]; dp[i] = 0.0; });
        EpStore<double> ey{yp};
        apply22<8>(c, K, z, ey);
        project(yp);
        const double yn = norm(yp);
        if (!(yn > 0)) break;
        for (int sweep = 0; sweep < 8; ++sweep, ++sweeps) {
            // r = Pi (y - K d), d orthogonal to z
            EpStore<double> er{rp};
            apply22<8>(c, K, dp, er);
            launch_map(c, n, [=] __device__(int64_t i) { rp[i] = yp[i] - rp[i]; });
            project(rp);
            const double rn = norm(rp);
            if (!std::isfinite(rn)) throw Error(STK_E_NAN, "non-finite residual in the near-null vector set-up");
            if (rn <= 1e-6 * yn) break;
            const double scale = 1.0 / rn;
            launch_map(c, n, [=] __device__(int64_t i) { r32p[i] = (float)(rp[i] * scale); });
            bool conv;
            inner_cg<float>(c, view, true, r32p, e32p, 1e-3, 500, 2, &conv);
            launch_map(c, n, [=] __device__(int64_t i) { dp[i] += rn * (double)e32p[i]; });
            project(dp);
            // a correction this small (z has unit norm) leaves z far inside the step's 1e-6
            // criterion even if the sweep's solve were off by a factor of ten: the usual case
            // when z is the previous assembly's vector
            if (norm(dp) < 1e-7) break;
        }
        const double dn = norm(dp);
        if (trace_on()) std::fprintf(stderr, "[stk] null vector step %d: |correction| %.3e after %d sweeps\n", step, dn, sweeps);
        // a step is kept only while the (quadratically convergent) correction still shrinks
        if (step > 0 && (!std::isfinite(dn) || dn > 0.1 * prev)) {
            launch_map(c, n, [=] __device__(int64_t i) { z[i] = kp[i]; });
            break;
        }
        launch_map(c, n, [=] __device__(int64_t i) { z[i] -= dp[i]; });
        const double inv = 1.0 / norm(z);
        launch_map(c, n, [=] __device__(int64_t i) { z[i] *= inv; });
        prev = dn;
        if (dn < 1e-6) break;
    }
    convert_vec(c, z, z32, n);
    fill_gs_rot(c);
    EpStoreDot<double> ed{yp, z};
    const int64_t nb = apply22<8>(c, K, z, ed, c.red64.p);
    red.finish(nb, SC_TMP);
    c.null_lambda = red.fetch(SC_TMP);
    if (trace_on()) std::fprintf(stderr, "[stk] null vector (mixed): lambda %.6e\n", c.null_lambda);
    pt.mark("near-null vector (FP64 refinement around FP32 multigrid)");
}

// Set-up after an assembly with FP64 and FP32 operators and the multigrid: what the FP32 methods
// use (the FP32 hierarchy, the coarsest inverse, the near-null vector).  The FP64 hierarchy waits
// for the first solve that needs it (ensure_mg64).
void setup_fast(Context &c) {
    PhaseTimer pt(c, "set-up");
    // The coarsest inverse, a latency-bound chain of small launches that reads only the coarsest
    // operator, is factorised on the auxiliary stream (by a second host thread) while this thread
    // issues the per-precision set-up (diagonals, smoother values, R A, eigenvalue bounds); it is
    // joined before the near-null vector, whose multigrid solves use it.  The work and its order on
    // each stream are those of a sequential set-up, so the results are the same bit for bit.
    // The worker writes the context's coarse inverse buffers, which this thread does not touch
    // before the join (coarse64_valid was cleared by the assembly).
    const bool overlap = c.aux_stream && !c.mg.empty();
    std::exception_ptr err;
    std::thread worker;
    if (overlap) {
        STK_HIP(hipStreamSynchronize(c.stream));   // the assembled operators the factorisation reads
        worker = std::thread([&c, &err]() {
            try {
                STK_HIP(hipSetDevice(c.device));
                coarse_inverse_on(c, c.aux_stream, c.coarse_inv64, &c.coarse_inv32);
            } catch (...) {
                err = std::current_exception();
            }
        });
    }
    try {
        setup_base<double>(c, pt);
        setup_base<float>(c, pt);
        setup_mg<float>(c, pt, !overlap);
    } catch (...) {
        if (worker.joinable()) worker.join();
        throw;
    }
    if (overlap) {
        worker.join();
        if (err) std::rethrow_exception(err);
        c.coarse_n = 2 * c.mg.back().nvert;
        c.coarse64_valid = true;
        pt.mark("coarse inverse (auxiliary stream) joined");
    }
    if (c.rotation_free) {
        null_vector_mixed(c, pt);
    } else {
        c.null_lambda = 0.0;
        convert_vec(c, c.rot64.p, c.rot32.p, c.nU);
        fill_gs_rot(c);
    }
    c.have_mg32 = true;
    c.have_mg64 = false;
    STK_HIP(hipStreamSynchronize(c.stream));
}

void ensure_mg64(Context &c) {
    if (!c.mg_built || c.have_mg64 || !c.K64) return;
    PhaseTimer pt(c, "set-up FP64 multigrid");
    setup_mg<double>(c, pt);
    c.have_mg64 = true;
    STK_HIP(hipStreamSynchronize(c.stream));
}

// ------------------------------------------------------------------ Schur PCG
// Pressure-space vectors (nP) and the velocity work vectors of one Schur solve.
template <typename T>
struct SchurWork {
    T *y, *z, *du;        // nU
    T *r, *d, *q, *Sq, *rl, *pc, *dp;   // nP
};

template <typename T>
void vmean_remove(Context &c, T *a, int64_t n) {
    Reducer red(c);
    red.map_reduce(n, [=] __device__(int64_t i) { return (double)a[i]; }, SC_TMP, (int)sizeof(T));
    const double *sc = c.d_sc.p;
    launch_map(c, n, [=] __device__(int64_t i) { a[i] -= (T)(sc[SC_TMP] / (double)n); }, 2 * (int)sizeof(T));
}

template <typename T>
double vdot(Context &c, const T *a, const T *b, int64_t n) {
    Reducer red(c);
    red.map_reduce(n, [=] __device__(int64_t i) { return (double)a[i] * (double)b[i]; }, SC_TMP, 2 * (int)sizeof(T));
    return red.fetch(SC_TMP);
}

// z = (L D L^T)^-1 r with the incomplete Cholesky of M (STK_PPC_IC): forward substitution colour by
// colour (each row needs only rows of earlier colours), then the diagonal and the backward
// substitution in reverse colour order.  y is an nP work vector.
void apply_pressure_ic(Context &c, const double *r, double *y, double *z) {
    const Context::PressureIC &ic = c.pic;
    const int32_t *ord = ic.order.p;
    const int32_t *lp = ic.lptr.p, *lc = ic.lcol.p, *up = ic.uptr.p, *uc = ic.ucol.p;
    const double *lv = ic.lval.p, *uv = ic.uval.p, *di = ic.dinv.p;
    const int nc = (int)ic.colour_off.size() - 1;
    for (int k = 0; k < nc; ++k) {
        const int64_t q0 = ic.colour_off[k];
        launch_map(c, ic.colour_off[k + 1] - q0, [=] __device__(int64_t t) {
            const int64_t q = q0 + t;
            const int32_t i = ord[q];
            double v = r[i];
            for (int32_t e = lp[q]; e < lp[q + 1]; ++e) v -= lv[e] * y[lc[e]];
            y[i] = v;
        });
    }
    for (int k = nc - 1; k >= 0; --k) {
        const int64_t q0 = ic.colour_off[k];
        launch_map(c, ic.colour_off[k + 1] - q0, [=] __device__(int64_t t) {
            const int64_t q = q0 + t;
            const int32_t i = ord[q];
            double v = di[i] * y[i];
            for (int32_t e = up[q]; e < up[q + 1]; ++e) v -= uv[e] * z[uc[e]];
            z[i] = v;
        });
    }
}

// Generalised StokesPcCG: solve K du - G dp = fu, G^T du = h (h may be null = 0)
// with du, dp starting at zero.  Returns outer iterations; inner iterations accumulate.
template <typename T>
int schur_pcg(Context &c, const stk_opts &o, bool use_mg, const T *fu, const T *h, SchurWork<T> &w,
              double rtol, int itmax, int itmin, double inner_rtol, int64_t *inner_its, double *schur_res,
              bool *conv, const T *pcvec, bool learned, bool deflate = false, bool *stalled = nullptr,
              double abs_tol = 0.0) {
    const bool f64 = sizeof(T) == 8;
    MGView<T> v = make_view<T>(c);
    const int64_t nU = c.nU, nP = c.nP;
    const T *G = (const T *)pick<T>(c.G64, c.G32).p;
    const T *GT = (const T *)pick<T>(c.GT64, c.GT32).p;
    const int deg = o.cheb_degree;
    bool ic;
    // du = K^-1 fu
    launch_map(c, nU, [=] __device__(int64_t i) { w.y[i] = fu[i]; });
    *inner_its += inner_cg<T>(c, v, use_mg, w.y, w.du, inner_rtol, o.inner_itmax, deg, &ic, deflate);
    // r = h - G^T du ; mean removed
    {
        EpStore<T> es{w.r};
        apply_rc<1, 2, T>(c, c.GTN, GT, w.du, es, nullptr);
        T *r = w.r;
        launch_map(c, nP, [=] __device__(int64_t i) { r[i] = (h ? h[i] : T(0)) - r[i]; });
    }
    vmean_remove<T>(c, w.r, nP);
    launch_map(c, nP, [=] __device__(int64_t i) { w.dp[i] = T(0); });
    const double r0 = std::sqrt(vdot<T>(c, w.r, w.r, nP));
    const double tol = std::max(rtol * r0, abs_tol);
    // learned: damped Jacobi sweeps d <- d + w Dinv (r - S d) on the assembled learned Schur
    // operator, from d = Dinv r; on the sliced vertex layout each sweep is one product with
    // the update in its epilogue, alternating between dout and w.Sq so the last lands in dout
    const bool sliced_s = learned && !f64 && c.S32s && !c.mg.empty() && c.mg[0].pat.sell;
    auto apply_pc = [&](const T *rin, T *dout) {
        if (learned) {
            const T *sd = pcvec;   // inverse diagonal of S
            const int nsweep = o.pc_sweeps > 0 ? o.pc_sweeps : 3;
            if constexpr (sizeof(T) == 4) {
                if (sliced_s) {
                    T *buf[2] = {dout, w.Sq};
                    int cur = nsweep % 2;   // the last sweep writes buf[0] = dout
                    T *d0 = buf[cur];
                    launch_map(c, nP, [=] __device__(int64_t i) { d0[i] = sd[i] * rin[i]; }, 3 * (int)sizeof(T));
                    for (int s = 0; s < nsweep; ++s) {
                        EpJacobi<T> ej{rin, sd, buf[cur], buf[1 - cur], T(0.7)};
                        spmv_sell11<T>(c, c.mg[0].pat, c.S32s.p, buf[cur], ej, nullptr);
                        cur = 1 - cur;
                    }
                    return;
                }
            }
            const T *S = (const T *)pick<T>(c.S64, c.S32).p;
            T *tmp = w.Sq;
            launch_map(c, nP, [=] __device__(int64_t i) { dout[i] = sd[i] * rin[i]; }, 3 * (int)sizeof(T));
            for (int s = 0; s < nsweep; ++s) {
                EpStore<T> es{tmp};
                spmv<1, 1, 8, T>(c, c.MN, S, dout, es, nullptr);
                launch_map(c, nP, [=] __device__(int64_t i) { dout[i] += T(0.7) * sd[i] * (rin[i] - tmp[i]); },
                           5 * (int)sizeof(T));
            }
        } else if (pcvec) {
            launch_map(c, nP, [=] __device__(int64_t i) { dout[i] = pcvec[i] * rin[i]; }, 3 * (int)sizeof(T));
        } else {
            // no vector: the incomplete Cholesky of M (FP64 plain method only)
            if constexpr (sizeof(T) == 8) apply_pressure_ic(c, rin, w.Sq, dout);
            else throw Error(STK_E_ARG, "the incomplete Cholesky pressure preconditioner is FP64 only");
        }
    };
    *conv = false;
    *schur_res = 1.0;
    if (!std::isfinite(r0)) throw Error(STK_E_NAN, "non-finite initial Schur residual");
    if (!(r0 > 0)) {
        *conv = true;
        *schur_res = 0.0;
        return 0;
    }
    // The scalars of the iteration stay on the device (slots of d_sc): alpha = (r . d) /
    // (q . S q) is read by the update kernels; the host fetches once per iteration the
    // residual norm and the next (r . d) and beta numerator.
    Reducer red(c);
    double *sc = c.d_sc.p;
    constexpr int B = (int)sizeof(T);
    apply_pc(w.r, w.d);
    launch_map(c, nP, [=] __device__(int64_t i) { w.q[i] = w.d[i]; }, 2 * B);
    red.map_reduce(nP, [=] __device__(int64_t i) { return (double)w.r[i] * (double)w.d[i]; }, SC_SRD, 2 * B);
    double rd = red.fetch(SC_SRD);
    int it = 0;
    double beta_prev = 0.0;
    const bool warm = o.inner_warm != 0 && !deflate;
    if (stalled) *stalled = false;
    for (it = 1; it <= itmax; ++it) {
        vmean_remove<T>(c, w.q, nP);
        if (stalled && !(rd > 0)) {   // the learned preconditioner is not positive definite here
            *stalled = true;
            return it - 1;
        }
        // y = G q ; z = K^-1 y ; Sq = G^T z
        EpStore<T> ey{w.y};
        apply_rc<2, 1, T>(c, c.GN, G, w.q, ey, nullptr);
        // K^-1 G q_k = K^-1 G d_k + beta K^-1 G q_k-1: beta z_k-1 is a good first guess
        const bool start = warm && it > 1;
        if (start) {
            const T bp = (T)beta_prev;
            T *zz = w.z;
            launch_map(c, nU, [=] __device__(int64_t i) { zz[i] *= bp; }, 2 * B);
        }
        *inner_its += inner_cg<T>(c, v, use_mg, w.y, w.z, inner_rtol, o.inner_itmax, deg, &ic, deflate, start);
        EpStore<T> es{w.Sq};
        apply_rc<1, 2, T>(c, c.GTN, GT, w.z, es, nullptr);
        red.map_reduce(nP, [=] __device__(int64_t i) { return (double)w.q[i] * (double)w.Sq[i]; }, SC_SQSQ, 2 * B);
        // r . d of this direction: from before the loop, then from the previous iteration's fetch
        const int rds = it == 1 ? SC_SRD : SC_SRN + 1;
        launch_map(c, nP, [=] __device__(int64_t i) {
            const T al = (T)(sc[rds] / sc[SC_SQSQ]);
            w.rl[i] = w.r[i];
            w.dp[i] += al * w.q[i];
            w.r[i] -= al * w.Sq[i];
        }, 7 * B);
        launch_map(c, nU, [=] __device__(int64_t i) { w.du[i] += (T)(sc[rds] / sc[SC_SQSQ]) * w.z[i]; }, 3 * B);
        vmean_remove<T>(c, w.r, nP);
        red.map_reduce(nP, [=] __device__(int64_t i) { return (double)w.r[i] * (double)w.r[i]; }, SC_SRN, B);
        // the next direction's preconditioned residual and its dots (unused after the last
        // iteration: one preconditioner application more than needed)
        apply_pc(w.r, w.d);
        red.map_reduce<2>(nP, [=] __device__(int64_t i) {
            const double d = (double)w.d[i];
            return Dots<2>{{(double)w.r[i] * d, ((double)w.r[i] - (double)w.rl[i]) * d}};
        }, SC_SRN + 1, 3 * B);
        double h[3];
        red.fetch_many(SC_SRN, 3, h);
        const double rn = std::sqrt(h[0]);
        *schur_res = rn / r0;
        if (trace_on())
            std::fprintf(stderr, "[stk] schur it %d  |r|/|r0| %.3e  inner its %lld\n", it, rn / r0,
                         (long long)*inner_its);
        if (!std::isfinite(rn)) throw Error(STK_E_NAN, "non-finite residual in the Schur PCG");
        if (rn < tol && it >= itmin) { *conv = true; break; }
        // learned preconditioner stall: less than a tenfold reduction after 30 iterations
        if (stalled && it == 30 && rn > 0.1 * r0) {
            *stalled = true;
            break;
        }
        const double beta = h[2] / rd;
        rd = h[1];
        beta_prev = beta;
        const T be = (T)beta;
        launch_map(c, nP, [=] __device__(int64_t i) { w.q[i] = w.d[i] + be * w.q[i]; }, 3 * B);
    }
    return std::min(it, itmax);
}

// FP64 full-system residual: ru = F - K U + G P (free dofs; fixed rows are zero), rc = G^T U.
// Returns ||[ru; rc]|| / ||F_free||, optionally keeping ru and -rc.
double full_residual(Context &c, const double *U, const double *P, double *ru, double *hc, bool project_null) {
    Reducer red(c);
    const int64_t nU = c.nU, nP = c.nP;
    double *sc = c.d_sc.p;
    const uint8_t *fx = c.d_fixed.p;
    const double *F = c.F64.p;
    // ru = K U
    EpStore<double> ek{ru};
    apply22<8>(c, Op22<double>{&c.KN, c.K64.p}, U, ek);
    // t = G P ; ru = F + t - ru  (masked)
    double *t = c.work64.p;
    EpStore<double> eg{t};
    apply_rc<2, 1, double>(c, c.GN, c.G64.p, P, eg, nullptr);
    launch_map(c, nU, [=] __device__(int64_t i) { ru[i] = fx[i] ? 0.0 : F[i] + t[i] - ru[i]; }, 33);
    if (project_null) {
        // the near-null component is outside the solved space (STK_NULL_PROJECT)
        const double *v = c.rot64.p;
        red.map_reduce(nU, [=] __device__(int64_t i) { return v[i] * ru[i]; }, SC_TMP + 1, 16);
        launch_map(c, nU, [=] __device__(int64_t i) { ru[i] -= sc[SC_TMP + 1] * v[i]; }, 24);
    }
    red.map_reduce(nU, [=] __device__(int64_t i) { return ru[i] * ru[i]; }, SC_RU, 8);
    // hc = -G^T U with its mean removed: the Schur iteration works modulo constants
    // (rm0space!, src/IterativeSolvers.jl:131,153), so the continuity residual does too.
    EpStore<double> ec{hc};
    apply_rc<1, 2, double>(c, c.GTN, c.GT64.p, U, ec, nullptr);
    red.map_reduce(nP, [=] __device__(int64_t i) { return hc[i]; }, SC_TMP, 8);
    launch_map(c, nP, [=] __device__(int64_t i) { hc[i] = sc[SC_TMP] / (double)nP - hc[i]; }, 16);
    red.map_reduce(nP, [=] __device__(int64_t i) { return hc[i] * hc[i]; }, SC_RC, 8);
    double v[2];
    red.fetch_many(SC_RU, 2, v);
    if (trace_on())
        std::fprintf(stderr, "[stk] full residual: momentum %.3e continuity %.3e (relative to |f|)\n",
                     std::sqrt(v[0]) / c.fnorm, std::sqrt(v[1]) / c.fnorm);
    return std::sqrt(v[0] + v[1]) / c.fnorm;
}

template <typename T>
SchurWork<T> schur_work(Context &c) {
    const bool f64 = sizeof(T) == 8;
    SchurBufs<T> &b = pick<T>(c.sw64, c.sw32);
    b.v.alloc(3 * c.nU + 7 * c.nP);
    T *base = b.v.p;
    SchurWork<T> w;
    w.y = base; w.z = base + c.nU; w.du = base + 2 * c.nU;
    T *pp = base + 3 * c.nU;
    w.r = pp; w.d = pp + c.nP; w.q = pp + 2 * c.nP; w.Sq = pp + 3 * c.nP; w.rl = pp + 4 * c.nP;
    w.pc = pp + 5 * c.nP; w.dp = pp + 6 * c.nP;
    return w;
}

// Smallest relative Schur reduction asked of one FP32 correction sweep.
constexpr double kSweepFloor = 1e-5;

int run_solve(Context &c, const stk_opts &o, stk_report &rep) {
    std::memset(&rep, 0, sizeof(rep));
    const bool use_mg = (o.inner == STK_INNER_MG) && c.mg_built && !c.loaded;
    if (o.inner == STK_INNER_MG && !use_mg)
        throw Error(STK_E_NOMESH,
                    "multigrid not available for this system" +
                        (c.loaded ? std::string(": it is built from the element matrices of stk_assemble, and the "
                                                "operators were set by stk_set_operators")
                                  : c.mg_note.empty() ? std::string() : ": " + c.mg_note));
    const int64_t nU = c.nU, nP = c.nP;
    // warm start: keep the context's solution (the last solve's or stk_set_solution's)
    const bool warm = o.warm_start != 0 && c.U.n == (size_t)nU && c.P.n == (size_t)nP;
    if (!warm) {
        c.U.alloc(nU);
        c.P.alloc(nP);
        c.U.zero(c.stream);
        c.P.zero(c.stream);
    }
    c.work64.alloc(std::max(nU, nP));
    EventPair ev_;
    hipEvent_t e0 = ev_.a, e1 = ev_.b;
    STK_HIP(hipEventRecord(e0, c.stream));
    // count the algorithmic bytes of every kernel launched until e1 (off again on any exit)
    struct LedgerScope {
        Context::Ledger &l;
        explicit LedgerScope(Context::Ledger &l) : l(l) { l = Context::Ledger{}; l.on = true; }
        ~LedgerScope() { l.on = false; }
    } ledger_scope(c.ledger);
    // masked right-hand side
    DBuf<double> Fm(nU);
    {
        const uint8_t *fx = c.d_fixed.p;
        const double *F = c.F64.p;
        double *f = Fm.p;
        launch_map(c, nU, [=] __device__(int64_t i) { f[i] = fx[i] ? 0.0 : F[i]; });
        c.fnorm = std::sqrt(vdot<double>(c, Fm.p, Fm.p, nU));
        if (!(c.fnorm > 0)) c.fnorm = 1.0;
    }
    int status = STK_OK;
    const bool deflate = (o.null_mode == STK_NULL_DEFLATE);
    const bool project = !deflate;
    if (deflate && (o.method != STK_SOLVE_PLAIN || !(c.null_lambda > 0)))
        throw Error(STK_E_ARG, "null-space deflation needs the plain FP64 method and a near-null rotation "
                               "(no fixed tangential dof)");
    if (o.method == STK_SOLVE_PLAIN) {
        if (use_mg) ensure_mg64(c);
        SchurWork<double> w = schur_work<double>(c);
        bool conv;
        double sres;
        int64_t iits = 0;
        // the plain method runs to the Schur tolerance and checks the full residual;
        // with STK_STOP_FULL it stops as soon as the continuity residual |r| is below
        // half the full-residual target (the momentum part is the inner-solve error),
        // and continues in restarted sweeps while the full residual is not met.
        if (o.pressure_pc == STK_PPC_IC && !c.pic.valid) {
            std::vector<double> Mh(c.M64.n);
            c.M64.download(Mh.data(), Mh.size(), c.stream);
            STK_HIP(hipStreamSynchronize(c.stream));
            build_pressure_ic(c, Mh);
            if (trace_on()) std::fprintf(stderr, "[stk] pressure IC(0): %d colours, shift %.1e\n",
                                         (int)c.pic.colour_off.size() - 1, c.pic.shift);
        }
        const double *pc = o.pressure_pc == STK_PPC_IC ? nullptr : c.pc64.p;
        const double abs_tol = (o.stop == STK_STOP_FULL) ? 0.5 * o.rtol * c.fnorm : 0.0;
        DBuf<double> &ru = c.res_u, &hc = c.res_p;
        ru.alloc(nU); hc.alloc(nP);
        // from zero: the system itself; warm: the correction of (U, P), whose initial Schur
        // residual -G^T K^-1 (F + G P) is StokesPcCG_pardiso's (Rhs .+= GG*P, then U = K^-1 Rhs)
        const double *fu = Fm.p, *hh = nullptr;
        int its = 0;
        conv = false;
        sres = 1.0;
        if (warm) {
            rep.full_residual = full_residual(c, c.U.p, c.P.p, ru.p, hc.p, project);
            fu = ru.p;
            hh = hc.p;
            // the full residual target is already met: nothing to do
            if (o.stop == STK_STOP_FULL && rep.full_residual < o.rtol) {
                conv = true;
                sres = 0.0;
                fu = nullptr;
            }
        }
        if (fu) {
            its = schur_pcg<double>(c, o, use_mg, fu, hh, w, o.rtol, o.itmax, o.itmin, o.inner_rtol, &iits, &sres,
                                    &conv, pc, false, deflate, nullptr, abs_tol);
            double *U = c.U.p, *P = c.P.p;
            const double keep = warm ? 1.0 : 0.0;
            launch_map(c, nU, [=] __device__(int64_t i) { U[i] = keep * U[i] + w.du[i]; });
            launch_map(c, nP, [=] __device__(int64_t i) { P[i] = keep * P[i] + w.dp[i]; });
        }
        rep.outer_iterations = its;
        rep.inner_iterations = iits;
        rep.schur_residual = sres;
        rep.refinements = 0;
        rep.full_residual = full_residual(c, c.U.p, c.P.p, ru.p, hc.p, project);
        if (trace_on()) std::fprintf(stderr, "[stk] plain first pass: %d its, full residual %.3e\n", its, rep.full_residual);
        if (o.stop == STK_STOP_FULL) {
            // restarted sweeps of the same FP64 method on the residual system
            while (rep.full_residual >= o.rtol && rep.outer_iterations < o.itmax) {
                int k = schur_pcg<double>(c, o, use_mg, ru.p, hc.p, w, o.rtol, o.itmax - rep.outer_iterations,
                                          1, o.inner_rtol, &iits, &sres, &conv, pc, false, deflate, nullptr,
                                          abs_tol);
                double *U = c.U.p, *P = c.P.p;
                launch_map(c, nU, [=] __device__(int64_t i) { U[i] += w.du[i]; });
                launch_map(c, nP, [=] __device__(int64_t i) { P[i] += w.dp[i]; });
                rep.outer_iterations += k;
                rep.inner_iterations = iits;
                rep.refinements++;
                rep.full_residual = full_residual(c, c.U.p, c.P.p, ru.p, hc.p, project);
                if (trace_on()) std::fprintf(stderr, "[stk] plain sweep %d full residual %.3e\n", rep.refinements, rep.full_residual);
                if (k == 0) break;
            }
            conv = rep.full_residual < o.rtol;
        }
        rep.converged = conv ? 1 : 0;
    } else {
        // FP64 refinement around FP32 Schur PCG corrections
        if (!c.K32) throw Error(STK_E_STATE, "mixed solve needs the FP32 operators (STK_ASM_FP32)");
        if (!c.K64) throw Error(STK_E_STATE, "mixed solve needs the FP64 operators of its refinement (STK_ASM_FP64)");
        const bool learned = (o.method == STK_SOLVE_LEARNED);
        if (learned && !c.S32) throw Error(STK_E_STATE, "learned solve needs the learned Schur operator");
        SchurWork<float> w = schur_work<float>(c);
        DBuf<double> &ru = c.res_u, &hc = c.res_p;
        DBuf<float> &ru32 = c.res_u32, &hc32 = c.res_p32;
        ru.alloc(nU); hc.alloc(nP); ru32.alloc(nU); hc32.alloc(nP);
        int64_t iits = 0;
        double res = 1.0, prev = 1.0;
        bool use_learned = learned, last_conv = true;
        rep.full_residual = 1.0;
        for (int sweep = 0; sweep < o.refine_max; ++sweep) {
            res = full_residual(c, c.U.p, c.P.p, ru.p, hc.p, project);
            rep.full_residual = res;
            if (trace_on()) std::fprintf(stderr, "[stk] mixed sweep %d full residual %.3e\n", sweep, res);
            if (res < o.rtol) break;
            // a learned sweep whose Schur PCG missed its target and that barely reduced
            // the full residual: drop the model (a converged Schur PCG whose sweep is
            // limited by the inner-solve error is not the model's fault)
            if (sweep > 0 && use_learned && !last_conv && res > o.stall_ratio * prev) {
                use_learned = false;
                rep.fallbacks++;
            }
            prev = res;
            // scale the correction problem to O(1) for FP32
            const double scale = 1.0 / (res * c.fnorm);
            {
                double *a = ru.p, *b = hc.p;
                float *a32 = ru32.p, *b32 = hc32.p;
                launch_map(c, nU, [=] __device__(int64_t i) { a32[i] = (float)(a[i] * scale); }, 12);
                launch_map(c, nP, [=] __device__(int64_t i) { b32[i] = (float)(b[i] * scale); }, 12);
            }
            bool conv, stalled = false;
            double sres;
            const float *pc = use_learned ? c.Sdinv32.p : c.pc32.p;
            // The scaled correction problem has unit full residual; its Schur residual is
            // the continuity part of the next full residual.  Target the reduction this
            // sweep needs (mixed_schur_rtol, or less on the last sweep), but never ask FP32
            // for more than kSweepFloor relative to the sweep's initial Schur residual
            // (which may be far above 1: continuity and momentum carry different units).
            const double target = std::max(o.mixed_schur_rtol, 0.5 * o.rtol / res);
            int k = schur_pcg<float>(c, o, use_mg, ru32.p, hc32.p, w, kSweepFloor, o.itmax, 1,
                                     o.mixed_inner_rtol, &iits, &sres, &conv, pc, use_learned, false,
                                     use_learned ? &stalled : nullptr, target);
            last_conv = conv;
            if (stalled) {
                // keep the partial correction (a valid Krylov step) and continue without the model
                use_learned = false;
                rep.fallbacks++;
                if (trace_on()) std::fprintf(stderr, "[stk] learned preconditioner stalled: falling back\n");
            }
            rep.outer_iterations += k;
            rep.schur_residual = sres;
            const double inv = 1.0 / scale;
            double *U = c.U.p, *P = c.P.p;
            launch_map(c, nU, [=] __device__(int64_t i) { U[i] += inv * (double)w.du[i]; }, 20);
            launch_map(c, nP, [=] __device__(int64_t i) { P[i] += inv * (double)w.dp[i]; }, 20);
            rep.refinements++;
        }
        rep.inner_iterations = iits;
        rep.full_residual = full_residual(c, c.U.p, c.P.p, ru.p, hc.p, project);
        // fallback: finish with the plain FP64 method when the mixed sweeps did not get there
        if (rep.full_residual >= o.rtol) {
            rep.fallbacks++;
            SchurWork<double> w64 = schur_work<double>(c);
            bool conv;
            double sres;
            int64_t it64 = 0;
            if (use_mg) ensure_mg64(c);
            const bool mg64 = use_mg && c.K64 && c.mg_built && c.have_mg64;
            int k = schur_pcg<double>(c, o, mg64, ru.p, hc.p, w64, o.rtol, o.itmax, 1, o.inner_rtol, &it64, &sres,
                                      &conv, c.pc64.p, false);
            double *U = c.U.p, *P = c.P.p;
            launch_map(c, nU, [=] __device__(int64_t i) { U[i] += w64.du[i]; });
            launch_map(c, nP, [=] __device__(int64_t i) { P[i] += w64.dp[i]; });
            rep.outer_iterations += k;
            rep.inner_iterations += it64;
            rep.full_residual = full_residual(c, c.U.p, c.P.p, ru.p, hc.p, project);
        }
        rep.converged = rep.full_residual < o.rtol ? 1 : 0;
    }
    if (project) {
        // the solved space excludes the near-null direction: drop whatever component
        // of it the FP32 corrections (orthogonal to the rounded vector) left in U
        Reducer red(c);
        const double *v = c.rot64.p;
        double *U = c.U.p, *sc = c.d_sc.p;
        red.map_reduce(nU, [=] __device__(int64_t i) { return v[i] * U[i]; }, SC_TMP);
        launch_map(c, nU, [=] __device__(int64_t i) { U[i] -= sc[SC_TMP] * v[i]; });
    }
    STK_HIP(hipEventRecord(e1, c.stream));
    c.ledger.on = false;
    rep.bytes = c.ledger.bytes;
    rep.launches = c.ledger.launches;
    rep.launches_floor = c.ledger.floor_launches;
    STK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    STK_HIP(hipEventElapsedTime(&ms, e0, e1));
    rep.solve_ms = ms;
    if (!rep.converged) status = STK_E_NOTCONV;
    return status;
}

}  // namespace stk

// ------------------------------------------------------------------ post-assembly helpers
namespace stk {

template <typename S, typename D>
void convert(Context &c, const S *src, D *dst, int64_t n) {
    launch_map(c, n, [=] __device__(int64_t i) { dst[i] = (D)src[i]; });
}
template void convert<double, float>(Context &, const double *, float *, int64_t);
template void convert<float, double>(Context &, const float *, double *, int64_t);

// pc = 1 / rowsum(M) (lumpedpreconditioner, src/IterativeSolvers.jl:238-240)
template <typename T>
void lumped_inverse(Context &c, const Pattern &p, const T *M, T *pc) {
    const int32_t *ptr = p.dptr.p;
    launch_map(c, p.nrows, [=] __device__(int64_t i) {
        double s = 0;
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) s += (double)M[k];
        pc[i] = (T)(1.0 / s);
    });
}
template void lumped_inverse<double>(Context &, const Pattern &, const double *, double *);
template void lumped_inverse<float>(Context &, const Pattern &, const float *, float *);

// Inverse diagonal of a scalar CSR operator.
template <typename T>
void csr_diag_inverse(Context &c, const Pattern &p, const T *A, T *dinv) {
    const int32_t *ptr = p.dptr.p, *col = p.dcol.p;
    launch_map(c, p.nrows, [=] __device__(int64_t i) {
        T d = T(0);
        for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
            if (col[k] == i) d = A[k];
        dinv[i] = d != T(0) ? T(1) / d : T(0);
    });
}
template void csr_diag_inverse<double>(Context &, const Pattern &, const double *, double *);
template void csr_diag_inverse<float>(Context &, const Pattern &, const float *, float *);

// FP32 copy and inverse diagonal of the learned Schur operator.
void finish_learned_operator(Context &c) {
    const int64_t nnz = c.MN.nnz();
    convert<double, float>(c, c.S64.p, c.S32.p, nnz);
    csr_diag_inverse<double>(c, c.MN, c.S64.p, c.Sdinv64.p);
    convert<double, float>(c, c.Sdinv64.p, c.Sdinv32.p, c.nP);
    // the same values in the sliced layout of the vertex pattern (the multigrid's first level
    // has M's pattern), for the sliced product in the preconditioner's sweeps
    c.S32s.release();
    if (!c.mg.empty() && c.mg[0].pat.sell) {
        const Pattern &P = c.mg[0].pat;
        c.S32s.alloc(P.nstore());
        c.S32s.zero(c.stream);
        const int32_t *ptr = c.MN.dptr.p;
        const int64_t *sptr = P.dsptr.p;
        const float *S = c.S32.p;
        float *Ss = c.S32s.p;
        launch_map(c, c.nP, [=] __device__(int64_t r) {
            const int64_t base = sptr[r / Pattern::kSlice] + r % Pattern::kSlice;
            for (int32_t k = ptr[r]; k < ptr[r + 1]; ++k) Ss[base + (int64_t)(k - ptr[r]) * Pattern::kSlice] = S[k];
        });
    }
}

// y = A x for an assembled operator, host FP64 vectors.
void host_spmv(Context &c, int which, int prec, const double *xh, double *yh) {
    const Pattern *p = nullptr;
    int64_t nx = 0, ny = 0;
    switch (which) {
        case STK_MAT_K: p = &c.KN; nx = ny = c.nU; break;
        case STK_MAT_G: p = &c.GN; nx = c.nP; ny = c.nU; break;
        case STK_MAT_GT: p = &c.GTN; nx = c.nU; ny = c.nP; break;
        case STK_MAT_M: case STK_MAT_S: p = &c.MN; nx = ny = c.nP; break;
        case STK_MAT_R:
            STK_CHECK(c.mg_built, STK_E_STATE, "no multigrid hierarchy");
            p = &c.mg[0].R_pat; nx = c.nU; ny = 2 * c.mg[0].nvert;
            break;
        case STK_MAT_RK:
            STK_CHECK(c.mg_built, STK_E_STATE, "no multigrid hierarchy");
            p = &c.GTN; nx = c.nU; ny = 2 * c.nP;
            break;
        default: throw Error(STK_E_ARG, "unknown operator");
    }
    DBuf<double> x64(nx), y64(ny);
    x64.upload(xh, nx, c.stream);
    auto run = [&](auto tag, const auto *val, auto *x, auto *y) {
        using T = decltype(tag);
        EpStore<T> es{y};
        switch (which) {
            case STK_MAT_K: apply22<8>(c, Op22<T>{p, val}, x, es); break;
            case STK_MAT_R: case STK_MAT_RK: apply22_auto<T>(c, Op22<T>{p, val}, x, es); break;
            case STK_MAT_G: apply_rc<2, 1, T>(c, *p, val, x, es, nullptr); break;
            case STK_MAT_GT: apply_rc<1, 2, T>(c, *p, val, x, es, nullptr); break;
            default: spmv<1, 1, 8, T>(c, *p, val, x, es, nullptr); break;
        }
    };
    auto pick64 = [&]() -> const double * {
        switch (which) {
            case STK_MAT_K: return c.K64.p;
            case STK_MAT_G: return c.G64.p;
            case STK_MAT_GT: return c.GT64.p;
            case STK_MAT_M: return c.M64.p;
            case STK_MAT_R: return c.mg[0].R64.p;
            case STK_MAT_RK: return c.RK64.p;
            default: return c.S64.p;
        }
    };
    auto pick32 = [&]() -> const float * {
        switch (which) {
            case STK_MAT_K: return c.K32.p;
            case STK_MAT_G: return c.G32.p;
            case STK_MAT_GT: return c.GT32.p;
            case STK_MAT_M: return c.M32.p;
            case STK_MAT_R: return c.mg[0].R32.p;
            case STK_MAT_RK: return c.RK32.p;
            default: return c.S32.p;
        }
    };
    if (prec == 64) {
        const double *v = pick64();
        if (!v) throw Error(STK_E_STATE, "FP64 operator not assembled");
        run(double{}, v, (const double *)x64.p, y64.p);
    } else if (prec == 32) {
        const float *v = pick32();
        if (!v) throw Error(STK_E_STATE, "FP32 operator not assembled");
        DBuf<float> x32(nx), y32(ny);
        convert<double, float>(c, x64.p, x32.p, nx);
        run(float{}, v, (const float *)x32.p, y32.p);
        convert<float, double>(c, y32.p, y64.p, ny);
    } else {
        throw Error(STK_E_ARG, "precision must be 64 or 32");
    }
    y64.download(yh, ny, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
}

// Net-rotation removal (updatevelocity2, src/Stokes/StokesSolver.jl:170-184):
// omega = int rho u_theta / int rho r with volume_integral (src/Stress/FSE.jl:412-464),
// u_theta -= omega r at every node, Ucart = TT U.
struct Quad6 {
    double xi[6], eta[6], w[6];
};

double update_velocity(Context &c, double *Ucart_dev) {
    Quad6 q;
    {
        const double s = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
        const double g1 = (8.0 - std::sqrt(10.0) + s) / 18.0, g2 = (8.0 - std::sqrt(10.0) - s) / 18.0;
        const double w1 = (620.0 + std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0))) / 3720.0;
        const double w2 = (620.0 - std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0))) / 3720.0;
        const double x[6][2] = {{1 - 2 * g1, g1}, {g1, 1 - 2 * g1}, {g1, g1}, {1 - 2 * g2, g2}, {g2, 1 - 2 * g2}, {g2, g2}};
        for (int i = 0; i < 6; ++i) { q.xi[i] = x[i][0]; q.eta[i] = x[i][1]; q.w[i] = 0.5 * (i < 3 ? w1 : w2); }
    }
    const int64_t nel = c.nel;
    const int32_t *e2n = c.d_e2n.p;
    const double *th = c.d_elth.p, *rr = c.d_elr.p, *rho = c.d_rho.p, *U = c.U.p, *nr = c.d_noder.p;
    auto integral = [=] __device__(int64_t p, int which) {
        const double r1 = rr[p], r2 = rr[nel + p], r3 = rr[2 * nel + p];
        const double t1 = th[p], t2 = th[nel + p], t3 = th[2 * nel + p];
        const double detJ = (t2 - t1) * (r3 - r1) - (t3 - t1) * (r2 - r1);
        double V[6];
        for (int k = 0; k < 6; ++k) {
            const int64_t nd = e2n[k * nel + p];
            V[k] = rho[nd] * (which == 0 ? U[2 * nd] : nr[nd]);
        }
        double s = 0;
        for (int i = 0; i < 6; ++i) {
            const double a = q.xi[i], b = q.eta[i], t = 1.0 - a - b;
            const double N[6] = {t * (2 * t - 1), a * (2 * a - 1), b * (2 * b - 1), 4 * a * t, 4 * a * b, 4 * b * t};
            const double rq = t * r1 + a * r2 + b * r3;
            double v = 0;
            for (int k = 0; k < 6; ++k) v += N[k] * V[k];
            s += v * rq * detJ * q.w[i];
        }
        return s;
    };
    Reducer red(c);
    red.map_reduce(nel, [=] __device__(int64_t p) { return integral(p, 0); }, SC_TMP);
    red.map_reduce(nel, [=] __device__(int64_t p) { return integral(p, 1); }, SC_TMP + 1);
    double v[2];
    red.fetch_many(SC_TMP, 2, v);
    const double omega = v[0] / v[1];
    double *Um = c.U.p;
    const double *cs = c.d_nodecs.p;
    launch_map(c, c.nnod, [=] __device__(int64_t n) {
        Um[2 * n] -= omega * nr[n];
        if (Ucart_dev) {
            const double co = cs[2 * n], si = cs[2 * n + 1];
            const double a = Um[2 * n], b = Um[2 * n + 1];
            Ucart_dev[2 * n] = co * a + si * b;
            Ucart_dev[2 * n + 1] = -si * a + co * b;
        }
    });
    STK_HIP(hipStreamSynchronize(c.stream));
    return omega;
}

}  // namespace stk

// ------------------------------------------------------------------ thermal step
// solveDiffusion_threaded (src/ThermalDiffusion/AssemblerThermal.jl:1-62): assembly, then
// _prepare_matrices! (ThermalSolver.jl:54-63, alpha = 1): Rhs = M T0 + F, A = M + dt K;
// _apply_bcs! (:38-41): T[fixed] = vfix, Rhs -= A[:, fixed] vfix; the free block solved by
// Jacobi-preconditioned CG (A is SPD) to |r| <= rtol |Rhs_free| (the reference calls
// Krylov.gmres, unpreconditioned, whose default stops at about 1.5e-8 of |Rhs|).
namespace stk {

void thermal_solve(Context &c, double *T, const double *rho, const ThermalStep &s, const int64_t *fixed,
                   const double *vfix, int64_t nfix, stk_report &rep) {
    std::memset(&rep, 0, sizeof(rep));
    const int64_t nn = c.nnod, ns = c.KN.nstore();
    EventPair ev_;
    hipEvent_t e0 = ev_.a, e1 = ev_.b;
    STK_HIP(hipEventRecord(e0, c.stream));
    c.KT.alloc(ns); c.MT.alloc(ns); c.AT.alloc(ns); c.FT.alloc(nn);
    c.KT.zero(c.stream); c.MT.zero(c.stream); c.FT.zero(c.stream);
    thermal_assemble(c, rho, s.K, s.Cp, s.dQdT, s.dt, c.KT.p, c.MT.p, c.FT.p);
    c.have_thermal = true;
    double *A = c.AT.p;
    {
        const double *Kt = c.KT.p, *Mt = c.MT.p;
        const double dt = s.dt;
        launch_map(c, ns, [=] __device__(int64_t q) { A[q] = Mt[q] + dt * Kt[q]; });
    }
    ThermalWork &w = c.thw;
    DBuf<double> &b = w.b, &xf = w.xf, &t = w.t, &r = w.r, &z = w.z, &p = w.p, &Ap = w.Ap, &dinv = w.dinv;
    DBuf<uint8_t> &flag = w.flag;
    for (DBuf<double> *v : {&b, &xf, &t, &r, &z, &p, &Ap, &dinv}) v->alloc(nn);
    flag.alloc(nn);
    xf.zero(c.stream);
    flag.zero(c.stream);
    {
        EpStorePlus<double> ep{b.p, c.FT.p};
        spmv_sell11<double>(c, c.KN, c.MT.p, T, ep, nullptr);
        double *xv = xf.p;
        uint8_t *fl = flag.p;
        launch_map(c, nfix, [=] __device__(int64_t k) { xv[fixed[k]] = vfix[k]; fl[fixed[k]] = 1; });
        EpStore<double> es{t.p};
        spmv_sell11<double>(c, c.KN, A, xv, es, nullptr);
        double *bv = b.p;
        const double *tv = t.p;
        launch_map(c, nn, [=] __device__(int64_t i) { bv[i] -= tv[i]; });
    }
    // rows and columns of fixed nodes become identity (structural entries only)
    {
        const int64_t *sptr = c.KN.dsptr.p;
        const int32_t *scol = c.KN.dscol.p, *ptr = c.KN.dptr.p;
        const uint8_t *fl = flag.p;
        double *dv = dinv.p;
        launch_map(c, nn, [=] __device__(int64_t i) {
            const int64_t base = sptr[i / Pattern::kSlice] + i % Pattern::kSlice;
            const int32_t len = ptr[i + 1] - ptr[i];
            double diag = 0.0;
            for (int32_t k = 0; k < len; ++k) {
                const int64_t q = base + (int64_t)k * Pattern::kSlice;
                const int64_t j = scol[q];
                if (fl[i] || fl[j]) A[q] = (i == j && fl[i]) ? 1.0 : 0.0;
                if (j == i) diag = A[q];
            }
            dv[i] = diag != 0.0 ? 1.0 / diag : 0.0;
        });
        double *bv = b.p;
        launch_map(c, nfix, [=] __device__(int64_t k) { bv[fixed[k]] = vfix[k]; T[fixed[k]] = vfix[k]; });
    }
    // Jacobi PCG from T0
    Reducer red(c);
    double *sc = c.d_sc.p;
    const uint8_t *fl = flag.p;
    const double *bv = b.p, *dv = dinv.p;
    double *rv = r.p, *zv = z.p, *pv = p.p, *Apv = Ap.p;
    red.map_reduce(nn, [=] __device__(int64_t i) { return fl[i] ? 0.0 : bv[i] * bv[i]; }, SC_BB);
    {
        EpResidual<double> er{rv, bv};
        const int64_t nb = spmv_sell11<double>(c, c.KN, A, T, er, c.red64.p);
        red.finish(nb, SC_HIST);
    }
    red.map_reduce(nn, [=] __device__(int64_t i) {
        const double zz = dv[i] * rv[i];
        zv[i] = zz;
        pv[i] = zz;
        return rv[i] * zz;
    }, SC_RZV);
    double h[2];
    red.fetch_many(SC_BB, 1, h);
    const double bb = h[0];
    if (!std::isfinite(bb)) throw Error(STK_E_NAN, "non-finite thermal right-hand side");
    const double tol2 = s.rtol * s.rtol * bb;
    double rr = red.fetch(SC_HIST);
    int it = 0, par = 0;
    bool conv = !(rr > tol2) && std::isfinite(rr);
    while (!conv && it < s.itmax) {
        const int rz_cur = SC_RZV + par, rz_new = SC_RZV + (1 - par);
        EpStoreDot<double> ed{Apv, pv};
        const int64_t nb = spmv_sell11<double>(c, c.KN, A, pv, ed, c.red64.p);
        red.finish(nb, SC_PAP2);
        red.map_reduce(nn, [=] __device__(int64_t i) {
            const double al = sc[rz_cur] / sc[SC_PAP2];
            T[i] += al * pv[i];
            const double rn = rv[i] - al * Apv[i];
            rv[i] = rn;
            return rn * rn;
        }, SC_HIST);
        red.map_reduce(nn, [=] __device__(int64_t i) {
            const double zz = dv[i] * rv[i];
            zv[i] = zz;
            return rv[i] * zz;
        }, rz_new);
        launch_map(c, nn, [=] __device__(int64_t i) { pv[i] = zv[i] + (sc[rz_new] / sc[rz_cur]) * pv[i]; });
        ++it;
        par = 1 - par;
        rr = red.fetch(SC_HIST);
        if (!std::isfinite(rr)) throw Error(STK_E_NAN, "non-finite residual in the thermal solve");
        conv = rr <= tol2;
    }
    STK_HIP(hipEventRecord(e1, c.stream));
    STK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    STK_HIP(hipEventElapsedTime(&ms, e0, e1));
    rep.converged = conv ? 1 : 0;
    rep.outer_iterations = it;
    rep.inner_iterations = it;
    rep.schur_residual = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    rep.full_residual = rep.schur_residual;
    rep.solve_ms = ms;
}

}  // namespace stk

// ------------------------------------------------------------------ kernel timing
namespace stk {

double time_kernel(Context &c, int kernel, int reps, double *bytes) {
    STK_CHECK(reps > 0, STK_E_ARG, "reps must be positive");
    EventPair ev_;
    hipEvent_t e0 = ev_.a, e1 = ev_.b;
    // compulsory bytes: structural blocks only (four values and a 32-bit column each), the
    // slice pointers (8 bytes per 64 rows), x read once, y written once
    const int64_t nU = c.nU, nb = c.KN.nnz(), nsl = (c.KN.nrows + Pattern::kSlice - 1) / Pattern::kSlice;
    auto spmv_bytes = [&](int sv) { return (double)nb * (4 * sv + 4) + (double)(nsl + 1) * 8 + 2.0 * nU * sv; };
    std::function<void()> launch;
    DBuf<double> x64, y64, b64, d64, o64;
    DBuf<float> x32, y32, b32, d32, o32;
    switch (kernel) {
        case STK_KERNEL_KP32: {
            STK_CHECK(c.K32 && c.rot32, STK_E_STATE, "FP32 operators not assembled");
            x32.alloc(nU); y32.alloc(nU);
            float *x = x32.p;
            launch_map(c, nU, [=] __device__(int64_t i) { x[i] = 1.0f + 1e-3f * (float)(i % 17); });
            Op22<float> K{&c.KN, c.K32.p};
            float *y = y32.p;
            const float *rot = c.rot32.p;
            // as in the inner CG: Ap = K p with (p . Ap, v . Ap) block partials; p, Ap and v
            // once each besides the blocks (the partials are 16 B per 256 rows)
            launch = [&c, K, x, y, rot]() {
                EpStoreDot2<float> ed{y, x, rot};
                apply22<8>(c, K, x, ed, c.red64.p);
            };
            *bytes = spmv_bytes(4) + 1.0 * nU * 4;
            break;
        }
        case STK_KERNEL_SPMV_K32: case STK_KERNEL_CHEB_K32: {
            STK_CHECK(c.K32 && c.Kdinv32, STK_E_STATE, "FP32 operators not assembled");
            x32.alloc(nU); y32.alloc(nU); b32.alloc(nU); d32.alloc(nU); o32.alloc(nU);
            float *x = x32.p, *y = y32.p, *b = b32.p, *d = d32.p;
            launch_map(c, nU, [=] __device__(int64_t i) { x[i] = 1.0f + 1e-3f * (float)(i % 17); b[i] = 1.0f; d[i] = 0.0f; y[i] = 0.0f; });
            Op22<float> K{&c.KN, c.K32.p};
            if (kernel == STK_KERNEL_SPMV_K32) {
                launch = [&c, K, x, y]() { EpStore<float> es{y}; apply22<8>(c, K, x, es); };
                *bytes = spmv_bytes(4);
            } else {
                float *o = o32.p;
                const float *dinv = c.Kdinv32.p;
                // the first post-smoothing step of the V-cycle: d = c2 Dinv (b - K x) is
                // stored, the previous d is not read; b and Dinv read, d written besides x, x'
                launch = [&c, K, x, b, d, o, dinv]() {
                    EpCheb<float> ep{b, dinv, x, d, o, 0.0f, 0.25f, false, true};
                    apply22<8>(c, K, x, ep);
                };
                *bytes = spmv_bytes(4) + 3.0 * nU * 4;
            }
            break;
        }
        case STK_KERNEL_CHEB_KH16: {
            STK_CHECK(c.ksm.on, STK_E_STATE, "FP16 smoother copy not built");
            x32.alloc(nU); b32.alloc(nU); d32.alloc(nU); o32.alloc(nU); y32.alloc(nU);
            float *x = x32.p, *b = b32.p, *d = d32.p, *o = o32.p, *xu = y32.p;
            launch_map(c, nU, [=] __device__(int64_t i) { x[i] = 1.0f + 1e-3f * (float)(i % 17); b[i] = 1.0f; d[i] = 0.0f; });
            const _Float16 *Ah = reinterpret_cast<const _Float16 *>(c.ksm.A.p);
            const float *sinv = c.ksm.sinv.p;
            (void)xu; (void)sinv;
            launch = [&c, Ah, x, b, d, o]() {
                EpChebScaled ep{b, x, d, o, 0.5f, 0.25f, nullptr, nullptr};
                spmv_sell<_Float16, float>(c, c.KN, Ah, x, ep, nullptr);
            };
            // values 8 B + column 4 B per structural block; x^ gathered once, b^ and d read,
            // d and xo^ written
            *bytes = (double)nb * 12 + 5.0 * nU * 4;
            break;
        }
        case STK_KERNEL_VCYCLE32: {
            STK_CHECK(c.K32 && c.have_mg32, STK_E_STATE, "FP32 multigrid not set up");
            b32.alloc(nU);
            float *b = b32.p;
            launch_map(c, nU, [=] __device__(int64_t i) { b[i] = 1.0f + 1e-3f * (float)(i % 17); });
            auto view = std::make_shared<MGView<float>>(make_view<float>(c));
            launch = [&c, view, b]() { vcycle<float>(c, *view, 0, b, 2, nullptr, nullptr); };
            *bytes = 0.0;
            break;
        }
        case STK_KERNEL_GS_BWD32: {
            // the backward Gauss-Seidel sweep of the V-cycle with its dots, per colour launch
            STK_CHECK(c.gs.on && c.gs.filled, STK_E_STATE, "Gauss-Seidel smoother not set up");
            x32.alloc(nU);
            float *z = x32.p, *xq = c.gs.xq.p, *bq = c.gs.bq.p;
            const int32_t *rid = c.gs.rowid.p;
            const int64_t nq = c.gs.row_off.back();
            launch_map(c, 2 * nq, [=] __device__(int64_t i) {
                const bool real = rid[i / 2] >= 0;
                xq[i] = real ? 1.0f + 1e-3f * (float)(i % 17) : 0.0f;
                bq[i] = real ? 1.0f : 0.0f;
            });
            launch = [&c, z]() {
                int64_t nb = 0;
                gs_sweep(c, GS_BACKWARD, nullptr, z, true, &nb);
            };
            // per launch: its blocks, the neighbour x it gathers, its rows' data (gs_launch_bytes),
            // averaged over the colour launches of a sweep
            double sum = 0.0;
            for (int k = 0; k < c.gs.ncolour; ++k)
                sum += gs_launch_bytes(c.gs, k, false, true, c.gs.slice_off[k + 1] - c.gs.slice_off[k]);
            *bytes = sum / c.gs.ncolour;
            break;
        }
        case STK_KERNEL_INNER_IT32: {
            // a fixed number of inner CG iterations (tolerance unreachable), per iteration
            STK_CHECK(c.K32 && c.have_mg32, STK_E_STATE, "FP32 multigrid not set up");
            y32.alloc(nU); x32.alloc(nU);
            float *y = y32.p, *z = x32.p;
            auto view = std::make_shared<MGView<float>>(make_view<float>(c));
            launch = [&c, view, y, z, nU]() {
                launch_map(c, nU, [=] __device__(int64_t i) { y[i] = 1.0f + 1e-3f * (float)(i % 17); });
                bool conv;
                inner_cg<float>(c, *view, true, y, z, 1e-30, 20, 2, &conv);
            };
            *bytes = 0.0;
            break;
        }
        case STK_KERNEL_STRESS: case STK_KERNEL_DEFGRAD: {
            STK_CHECK(c.have_sfields, STK_E_STATE, "stress fields not set");
            x64.alloc(nU);
            double *u = x64.p;
            const double *nr = c.d_noder.p;
            // synthetic velocity: smooth in r, varying with the node index
            launch_map(c, nU, [=] __device__(int64_t i) {
                u[i] = (i % 2 ? 0.3 : 1.0) * nr[i / 2] + 1e-3 * (double)(i % 7);
            });
            const int64_t plane = c.nel * (int64_t)c.s_nip;
            const int64_t nel = c.nel;
            // compulsory: element connectivity and vertices, the velocity once, eta (+ tensor)
            const double in = (double)nel * (6 * 4 + 6 * 8) + (double)nU * 8 + (double)plane * 8;
            if (kernel == STK_KERNEL_STRESS) {
                o64.alloc(8 * plane);
                double *o = o64.p;
                launch = [&c, u, o]() { launch_stress(c, u, o); };
                *bytes = in + (c.s_aniso ? 6.0 * plane * 8 : 0.0) + 8.0 * plane * 8;
            } else {
                o64.alloc(4 * plane);
                double *F = o64.p;
                launch_map(c, 4 * plane, [=] __device__(int64_t i) { F[i] = (i % 4 == 0 || i % 4 == 3) ? 1.0 : 0.0; });
                launch = [&c, u, F]() { launch_deformation_gradient(c, u, 1e-3, F); };
                *bytes = (double)nel * (6 * 4 + 6 * 8) + (double)nU * 8 + 2.0 * 4 * plane * 8;
            }
            break;
        }
        case STK_KERNEL_G32: case STK_KERNEL_GT32: {
            STK_CHECK(c.G32 && c.GT32, STK_E_STATE, "FP32 operators not assembled");
            const bool g = kernel == STK_KERNEL_G32;
            const Pattern &pat = g ? c.GN : c.GTN;
            const int64_t nin = g ? c.nP : nU, nout = g ? nU : c.nP;
            x32.alloc(nin); y32.alloc(nout);
            float *x = x32.p, *y = y32.p;
            launch_map(c, nin, [=] __device__(int64_t i) { x[i] = 1.0f + 1e-3f * (float)(i % 17); });
            const float *val = g ? c.G32.p : c.GT32.p;
            launch = [&c, &pat, val, x, y, g]() {
                EpStore<float> es{y};
                if (g) apply_rc<2, 1, float>(c, pat, val, x, es, nullptr);
                else apply_rc<1, 2, float>(c, pat, val, x, es, nullptr);
            };
            // structural blocks (two FP32 values and a 32-bit column each), slice pointers, x and y once
            const int64_t nsl = (pat.nrows + Pattern::kSlice - 1) / Pattern::kSlice;
            *bytes = (double)pat.nnz() * (2 * 4 + 4) + (double)(nsl + 1) * 8 + (double)(nin + nout) * 4;
            break;
        }
        case STK_KERNEL_COARSE32: {
            STK_CHECK(c.coarse_inv32 && c.coarse_n > 0, STK_E_STATE, "no FP32 coarsest inverse");
            const int64_t n = c.coarse_n;
            x32.alloc(n); y32.alloc(n);
            float *x = x32.p, *y = y32.p;
            const float *A = c.coarse_inv32.p;
            launch_map(c, n, [=] __device__(int64_t i) { x[i] = 1.0f + 1e-3f * (float)(i % 17); });
            launch = [&c, A, x, y, n]() {
                k_dense_matvec<float><<<dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c.stream>>>(n, A, x, y);
                STK_HIP(hipGetLastError());
            };
            *bytes = (double)n * n * 4 + 2.0 * n * 4;
            break;
        }
        case STK_KERNEL_RESTRICT32: {
            STK_CHECK(c.mg_built && c.RK32 && c.gs.on && c.gs.filled, STK_E_STATE, "no FP32 multigrid with R K");
            const int64_t nc = 2 * c.mg[0].nvert;
            x32.alloc(nU); y32.alloc(nc);
            float *b = x32.p, *y = y32.p, *xq = c.gs.xq.p;
            launch_map(c, nU, [=] __device__(int64_t i) { b[i] = 1.0f + 1e-3f * (float)(i % 17); });
            const Op22<float> R{&c.mg[0].R_pat, c.mg[0].R32.p}, RK{&c.gs.rkq, c.RK32.p};
            launch = [&c, R, RK, b, xq, y]() { apply22_restrict<float>(c, R, b, RK, xq, y); };
            // both operators' blocks (four values and a column), their slice pointers, b and the
            // coloured iterate gathered once, y written
            const Pattern &pr = *R.pat, &pk = *RK.pat;
            auto ptrs = [](const Pattern &p) {
                return p.sell ? 8.0 * ((p.nrows + Pattern::kSlice - 1) / Pattern::kSlice + 1) : 4.0 * (p.nrows + 1);
            };
            *bytes = (double)(pr.nnz() + pk.nnz()) * 20 + ptrs(pr) + ptrs(pk) + 2.0 * nU * 4 + (double)nc * 4;
            break;
        }
        case STK_KERNEL_CGUPD32: {
            STK_CHECK(c.rot32, STK_E_STATE, "FP32 operators not assembled");
            x32.alloc(nU); y32.alloc(nU); b32.alloc(nU); d32.alloc(nU);
            float *z = x32.p, *p = y32.p, *r = b32.p, *Ap = d32.p;
            const float *rot = c.rot32.p;
            launch_map(c, nU, [=] __device__(int64_t i) {
                z[i] = 0.0f; p[i] = 1e-3f * (float)(i % 13); r[i] = 1.0f; Ap[i] = 0.5f;
            });
            Reducer red(c);
            launch = [red, z, p, r, Ap, rot, nU]() mutable {
                red.map_reduce(nU, [=] __device__(int64_t i) {
                    const float al = 1e-6f;
                    z[i] += al * p[i];
                    const float rv = r[i] - al * (Ap[i] - 1e-7f * rot[i]);
                    r[i] = rv;
                    return (double)rv * (double)rv;
                }, SC_TMP);
            };
            // z, p, r, Ap and the near-null vector read, z and r written (the block partials are
            // 8 B per 256 rows)
            *bytes = (double)nU * 7 * 4;
            break;
        }
        case STK_KERNEL_FSE: {
            // nel x 7 points of F = R(a) diag(s, 1/s) with a1/a2 = s^2 below the destruction ratio
            const int64_t npt = 7 * c.nel;
            o64.alloc(4 * npt);
            x64.alloc(6 * npt);
            double *F = o64.p, *fse = x64.p;
            launch_map(c, npt, [=] __device__(int64_t p) {
                const double a = 0.001 * (double)(p % 6283), s = 1.0 + 0.01 * (double)(p % 97);
                F[4 * p] = cos(a) * s; F[4 * p + 1] = sin(a) * s;
                F[4 * p + 2] = -sin(a) / s; F[4 * p + 3] = cos(a) / s;
            });
            launch = [&c, F, npt, fse]() { launch_fse(c, F, npt, FSE_FABRIC_DESTRUCTION, 1e3, fse); };
            // F read and written back, the ellipse written
            *bytes = (double)npt * (32 + 32 + 48);
            break;
        }
        case STK_KERNEL_ANISO: {
            STK_CHECK(c.dem.set, STK_E_STATE, "no DEM table (stk_set_dem)");
            STK_CHECK(c.d_fse && c.fse_nip > 0, STK_E_STATE, "no device ellipses (stk_fse)");
            // fabric-shape table of 16 nodes over a1 in [1, 61]
            constexpr int K = 16;
            x64.alloc(3 * K);
            double *tb = x64.p;
            launch_map(c, K, [=] __device__(int64_t k) {
                tb[k] = 1.0 + 4.0 * (double)k;
                tb[K + k] = 1.7 * (double)k / (K - 1);
                tb[2 * K + k] = 2.0 - 2.0 * (double)k / (K - 1);
            });
            o64.alloc(42 * c.nel);
            double *o = o64.p;
            const double *fse = c.d_fse.p;
            const int nip = c.fse_nip;
            launch = [&c, tb, o, fse, nip]() {
                ShapeIn s;
                s.nodes = tb; s.r1 = tb + K; s.r2 = tb + 2 * K; s.nshape = K;
                launch_anisotropic_tensor(c, fse, nip, s, o);
            };
            // per element: x1, y1, a1 of six points, the six nel x 7 planes written
            *bytes = (double)c.nel * (6 * 3 * 8 + 42 * 8);
            break;
        }
        case STK_KERNEL_COPY: {
            const int64_t n = (int64_t)1 << 26;   // 256 MiB of FP32
            x32.alloc(n); y32.alloc(n);
            float *x = x32.p, *y = y32.p;
            launch_map(c, n, [=] __device__(int64_t i) { x[i] = (float)(i % 1021); });
            // 16-byte accesses: the attainable streaming rate of the part
            launch = [&c, x, y]() {
                const float4 *xs = reinterpret_cast<const float4 *>(x);
                float4 *ys = reinterpret_cast<float4 *>(y);
                launch_map(c, n / 4, [=] __device__(int64_t i) { ys[i] = xs[i]; });
            };
            *bytes = 2.0 * n * 4;
            break;
        }
        case STK_KERNEL_SPMV_K64: case STK_KERNEL_CHEB_K64: {
            STK_CHECK(c.K64 && c.Kdinv64, STK_E_STATE, "FP64 operators not assembled");
            x64.alloc(nU); y64.alloc(nU); b64.alloc(nU); d64.alloc(nU); o64.alloc(nU);
            double *x = x64.p, *y = y64.p, *b = b64.p, *d = d64.p;
            launch_map(c, nU, [=] __device__(int64_t i) { x[i] = 1.0 + 1e-3 * (double)(i % 17); b[i] = 1.0; d[i] = 0.0; y[i] = 0.0; });
            Op22<double> K{&c.KN, c.K64.p};
            if (kernel == STK_KERNEL_SPMV_K64) {
                launch = [&c, K, x, y]() { EpStore<double> es{y}; apply22<8>(c, K, x, es); };
                *bytes = spmv_bytes(8);
            } else {
                double *o = o64.p;
                const double *dinv = c.Kdinv64.p;
                launch = [&c, K, x, b, d, o, dinv]() {
                    EpCheb<double> ep{b, dinv, x, d, o, 0.0, 0.25, false, true};
                    apply22<8>(c, K, x, ep);
                };
                *bytes = spmv_bytes(8) + 3.0 * nU * 8;
            }
            break;
        }
        default:
            throw Error(STK_E_ARG, "unknown kernel id");
    }
    launch();   // warm-up
    STK_HIP(hipEventRecord(e0, c.stream));
    for (int k = 0; k < reps; ++k) launch();
    STK_HIP(hipEventRecord(e1, c.stream));
    STK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    STK_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (kernel == STK_KERNEL_INNER_IT32) return ms / reps / 20;
    if (kernel == STK_KERNEL_GS_BWD32) return ms / reps / c.gs.ncolour;
    return ms / reps;
}

}  // namespace stk
