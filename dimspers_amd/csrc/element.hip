// Element assembly of the Stokes operators (replaces assembly_stokes_cylindric,
// src/Stokes/StokesSolver.jl:247-404, with _prepare_matrices :192-201 and _apply_bcs :186-190),
// in two steps:
//
//  1. k_element: eight lanes per element.  The first 6 or 7 lanes each integrate one point of the
//     rule (_fill_R_J! :522-541, _ip_loop! :543-639, _derivative_weights! :641-686) and leave its
//     Cartesian shape-function derivatives, the scaled viscous tensor and the weights in LDS; then
//     every lane accumulates its share of the 21 node-pair blocks of K_e (_fill_Kblk!,
//     src/src/Stokes/StokesSolver.jl:744-808), of G_e (:812-823), M_e (:825-836) and f_e (:838-845),
//     rotates them into the node frames (Q_a blk Q_b^T: TT^T K TT element by element), masks the
//     fixed dofs and stores them per element.  With the multigrid, the lanes then form the P1
//     projection of the element (E1 = Ph^T K_e Ph, the level-0 element matrix, and R K_e, the
//     element's part of R K) from the rotated K_e held in LDS.
//  2. k_gather: one thread per stored block of each operator (K, G, G^T, M, F, the P1 operator, R K)
//     sums the element blocks that land there, in element order, from a list built once per mesh.
//     Every stored position is written once, coalesced, with no atomics and no colouring: the sum
//     order is fixed, so the assembly is deterministic.
//
// The reference assembles triplets and lets sparse(I, J, V) sort and sum them (:488-520,
// src/Algebra/Sparsity.jl); the gather lists play the role of that sort, done once per mesh.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_ops.h"
#include "stk_internal.h"

namespace stk {

namespace {

constexpr int kElLanes = 8;                     // lanes per element
constexpr int kElPerBlock = 256 / kElLanes;     // elements per 256-thread block
constexpr int kIpVals = 22;                     // per point: dx[6] dz[6] d[6] omega omega/eta fx fz
// LDS values per element for an NQ-point rule: the point data (NQ x 22), then K_e (84) followed by
// the table of the P1 projection (36), then T1 (72) over K_e; an odd number, so the eight elements of
// a wave that read the same offset fall in different LDS banks.  With 6 points a 32-element block
// takes 40.2 KB (FP64): four blocks, 16 waves, per CU.
template <int NQ>
constexpr int el_lds() { return (NQ * kIpVals > 84 + 36 ? NQ * kIpVals : 84 + 36) | 1; }

// Integration rule of the assembly: ip_triangle6 (6 points, the live assembly, StokesSolver.jl:262)
// or ip_triangle7 (7 points, the threaded assembler, AssemblerStokes.jl:174); weights include the 1/2
// (src/Algebra/Quadrature.jl:80-138).  Passed by value in every launch.
struct AsmRule {
    int nq;
    double xi[7], eta[7], w[7];
};

AsmRule make_rule(int nq) {
    AsmRule q{};
    q.nq = nq;
    if (nq == 7) {
        const double g1 = (6.0 - std::sqrt(15.0)) / 21.0, g2 = (6.0 + std::sqrt(15.0)) / 21.0;
        const double w1 = (155.0 - std::sqrt(15.0)) / 1200.0, w2 = (155.0 + std::sqrt(15.0)) / 1200.0;
        const double x[7] = {1.0 / 3.0, 1.0 - 2.0 * g1, g1, g1, 1.0 - 2.0 * g2, g2, g2};
        const double y[7] = {1.0 / 3.0, g1, 1.0 - 2.0 * g1, g1, g2, 1.0 - 2.0 * g2, g2};
        const double w[7] = {0.225, w1, w1, w1, w2, w2, w2};
        for (int i = 0; i < 7; ++i) { q.xi[i] = x[i]; q.eta[i] = y[i]; q.w[i] = 0.5 * w[i]; }
    } else {
        const double s = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
        const double g1 = (8.0 - std::sqrt(10.0) + s) / 18.0, g2 = (8.0 - std::sqrt(10.0) - s) / 18.0;
        const double w1 = (620.0 + std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0))) / 3720.0;
        const double w2 = (620.0 - std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0))) / 3720.0;
        const double x[6] = {1 - 2 * g1, g1, g1, 1 - 2 * g2, g2, g2};
        const double y[6] = {g1, 1 - 2 * g1, g1, g2, 1 - 2 * g2, g2};
        for (int i = 0; i < 6; ++i) { q.xi[i] = x[i]; q.eta[i] = y[i]; q.w[i] = 0.5 * (i < 3 ? w1 : w2); }
    }
    return q;
}

template <typename T>
struct ElemArgs {
    int64_t nel;
    AsmRule rule;
    const int32_t *e2n;              // 6 x nel, natural element order
    const double *th, *r;            // 3 x nel vertex polar coordinates after fixangles!
    const double *nodecs, *rho;      // per node: (cos, sin) of the node angle; rho.rho
    const uint8_t *fixed;            // per dof
    const double *eta, *C;           // eta.ip column q at q * nel; tensor component k, column q at (k * ccol + q) * nel
    int64_t ccol;                    // columns stored per tensor component
    double g;
    int cart, bc, iso, clamp, proj;
    T *EK, *EG, *EM, *EF;            // [21][nel][4], [18][nel][2], [6][nel], [6][nel][2]
    T *E1, *ERK;                     // [36][nel], [18][nel][4] (proj only)
};

// lower node-pair block k -> (row node, column node); lblk(a, b) = a (a + 1) / 2 + b
__constant__ int8_t c_pair_row[21] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4, 5, 5, 5, 5, 5, 5};
__constant__ int8_t c_pair_col[21] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4, 0, 1, 2, 3, 4, 5};
// M_e upper entries (i <= j)
__constant__ int8_t c_m_row[6] = {0, 0, 0, 1, 1, 2};
__constant__ int8_t c_m_col[6] = {0, 1, 2, 1, 2, 2};

__device__ __forceinline__ int lblk(int a, int b) { return a * (a + 1) / 2 + b; }

template <typename T>
__device__ __forceinline__ void sincos_t(T x, T *s, T *c) {
    if constexpr (sizeof(T) == 8) sincos(x, s, c);
    else sincosf(x, s, c);
}

// 2x2 block b of K_e from its lower storage (transposed above the diagonal)
template <typename T>
__device__ __forceinline__ void kblock(const T *L, int f, int g, T v[4]) {
    if (f >= g) {
        const T *p = L + 4 * lblk(f, g);
        v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
    } else {
        const T *p = L + 4 * lblk(g, f);
        v[0] = p[0]; v[1] = p[2]; v[2] = p[1]; v[3] = p[3];
    }
}

// NQ: points of the rule (6 or 7), sizing the LDS; 4 waves per SIMD (at most 128 VGPRs) where that
// LDS allows it
template <typename T, int NQ>
__global__ __launch_bounds__(256, NQ == 6 ? 4 : 3) void k_element(ElemArgs<T> a) {
    __shared__ T lds[kElPerBlock][el_lds<NQ>()];
    __shared__ T ltab[NQ][9];                   // per point: P2 N (6), P1 NP (3)
    __shared__ T lrho[kElPerBlock][3];          // rho at the three vertices
    __shared__ T lcs[kElPerBlock][6][2];        // node frame (cos, sin)
    __shared__ uint8_t lfix[kElPerBlock][12];   // fixed flags per local dof
    __shared__ T lph[kElPerBlock][3][2];        // (cos h, sin h) of the three edges, h = half the angle step
    const int le = threadIdx.x / kElLanes, lane = threadIdx.x % kElLanes;
    // 32-bit element offsets (the host keeps 84 nel below 2^32): fewer registers per address
    const uint32_t nel = (uint32_t)a.nel;
    const uint32_t e = blockIdx.x * kElPerBlock + le;
    const bool live = e < nel;
    const int nq = a.rule.nq;
    T *L = lds[le];

    if (threadIdx.x < nq) {
        const T r = (T)a.rule.xi[threadIdx.x], s = (T)a.rule.eta[threadIdx.x], t = T(1) - r - s;
        T *N = ltab[threadIdx.x];
        N[0] = t * (T(2) * t - T(1)); N[1] = r * (T(2) * r - T(1)); N[2] = s * (T(2) * s - T(1));
        N[3] = T(4) * r * t; N[4] = T(4) * r * s; N[5] = T(4) * s * t;
        N[6] = t; N[7] = r; N[8] = s;
    }
    T th[3], rr[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        th[k] = live ? (T)a.th[k * nel + e] : T(0);
        rr[k] = live ? (T)a.r[k * nel + e] : T(1);
    }
    // every global load of the element is issued here, before the first barrier: the point's
    // viscosity and tensor columns, and (through e2n) the node data
    T eta_q = T(1), cq[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (live && lane < nq) {
        eta_q = (T)a.eta[lane * nel + e];
        if (!a.iso) {
#pragma unroll
            for (int k = 0; k < 6; ++k) cq[k] = (T)a.C[((uint32_t)k * (uint32_t)a.ccol + lane) * nel + e];
        }
    }
    if (live && lane < 6) {
        const int32_t n = a.e2n[lane * nel + e];
        if (lane < 3) lrho[le][lane] = (T)a.rho[n];
        lcs[le][lane][0] = (T)a.nodecs[2 * (int64_t)n];
        lcs[le][lane][1] = (T)a.nodecs[2 * (int64_t)n + 1];
        lfix[le][2 * lane] = a.fixed[2 * (int64_t)n];
        lfix[le][2 * lane + 1] = a.fixed[2 * (int64_t)n + 1];
    }
    if (live && a.proj && lane >= 5) {
        // Ph of edge k = lane - 5 (ea = k, eb = (k + 1) % 3): h = (theta_eb - theta_ea) / 2
        const int k = lane - 5;
        T sh, ch;
        sincos_t(T(0.5) * (th[(k + 1) % 3] - th[k]), &sh, &ch);
        lph[le][k][0] = ch;
        lph[le][k][1] = sh;
    }
    __syncthreads();

    // ---- 1. one integration point per lane
    if (live && lane < nq) {
        const int q = lane;
        const T R21 = rr[1] - rr[0], R31 = rr[2] - rr[0];
        const T Th21 = th[1] - th[0], Th31 = th[2] - th[0];
        const T detJ = Th21 * R31 - Th31 * R21;
        const T xi = (T)a.rule.xi[q], et = (T)a.rule.eta[q], tq = T(1) - xi - et;
        const T dr[6] = {-(T(4) * tq - T(1)), T(4) * xi - T(1), T(0), T(4) * (tq - xi), T(4) * et, -T(4) * et};
        const T ds[6] = {-(T(4) * tq - T(1)), T(0), T(4) * et - T(1), -T(4) * xi, T(4) * xi, T(4) * (tq - et)};
        const T thq = tq * th[0] + xi * th[1] + et * th[2];
        const T rq = tq * rr[0] + xi * rr[1] + et * rr[2];
        T sn, cs;
        sincos_t(thq, &sn, &cs);
        const T inv = T(1) / rq / detJ;
        const T jx0 = (R31 * cs - Th31 * sn * rq) * inv;
        const T jx1 = (-R21 * cs + Th21 * sn * rq) * inv;
        const T jz0 = (-R31 * sn - Th31 * cs * rq) * inv;
        const T jz1 = (R21 * sn + Th21 * cs * rq) * inv;
        const T om = rq * detJ * (T)a.rule.w[q];
        T *P = L + q * kIpVals;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            P[b] = jx0 * dr[b] + jx1 * ds[b];
            P[6 + b] = jz0 * dr[b] + jz1 * ds[b];
        }
        const T eta = eta_q;
        T d[6];   // eta11, eta33, eta55, eta13, eta15, eta35 (the StiffnessTensor order)
        if (a.iso) {
            d[0] = d[1] = T(4) / T(3); d[2] = T(1); d[3] = -T(2) / T(3); d[4] = d[5] = T(0);
#pragma unroll
            for (int k = 0; k < 6; ++k) d[k] *= om * eta;
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                T v = cq[k];
                if (a.clamp) {
                    // applybounds(eta C, 1e25, 1e18) of the top-level _fill_Kblk! (StokesSolver.jl:756)
                    v = v * eta;
                    v = v < T(1e25) ? v : T(1e25);
                    v = v > T(1e18) ? v : T(1e18);
                    d[k] = v * om;
                } else {
                    d[k] = v * (om * eta);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) P[12 + k] = d[k];
        P[18] = om;
        P[19] = om / eta;
        const T rhoq = tq * lrho[le][0] + xi * lrho[le][1] + et * lrho[le][2];
        const T cte = om * (T)a.g * rhoq;
        P[20] = cte * sn;
        P[21] = cte * cs;
    }
    __syncthreads();

    // ---- 2. accumulate: K blocks k = lane + 8 m, G pairs p = lane + 8 m, M entry / F node = lane
    T kacc[3][4], gacc[3][2], macc = T(0), facc[2] = {T(0), T(0)};
    int ka[3], kb[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const int k = lane + kElLanes * m;
        ka[m] = k < 21 ? c_pair_row[k] : 0;
        kb[m] = k < 21 ? c_pair_col[k] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) kacc[m][j] = T(0);
        gacc[m][0] = gacc[m][1] = T(0);
    }
    const int mi = lane < 6 ? c_m_row[lane] : 0, mj = lane < 6 ? c_m_col[lane] : 0;
    if (live) {
        for (int q = 0; q < nq; ++q) {
            const T *P = L + q * kIpVals;
            const T d11 = P[12], d22 = P[13], d33 = P[14], d12 = P[15], d13 = P[16], d23 = P[17];
            const T om = P[18];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                if (lane + kElLanes * m < 21) {
                    const T dxa = P[ka[m]], dza = P[6 + ka[m]], dxb = P[kb[m]], dzb = P[6 + kb[m]];
                    const T Dx0 = d11 * dxb + d13 * dzb, Dx1 = d12 * dxb + d23 * dzb, Dx2 = d13 * dxb + d33 * dzb;
                    const T Dz0 = d12 * dzb + d13 * dxb, Dz1 = d22 * dzb + d23 * dxb, Dz2 = d23 * dzb + d33 * dxb;
                    kacc[m][0] += dxa * Dx0 + dza * Dx2;
                    kacc[m][1] += dxa * Dz0 + dza * Dz2;
                    kacc[m][2] += dza * Dx1 + dxa * Dx2;
                    kacc[m][3] += dza * Dz1 + dxa * Dz2;
                }
                const int pg = lane + kElLanes * m;
                if (pg < 18) {
                    const int an = pg / 3, iv = pg % 3;
                    const T w = om * ltab[q][6 + iv];
                    gacc[m][0] += w * P[an];
                    gacc[m][1] += w * P[6 + an];
                }
            }
            if (lane < 6) {
                macc += P[19] * ltab[q][6 + mi] * ltab[q][6 + mj];
                facc[0] += P[20] * ltab[q][lane];
                facc[1] += P[21] * ltab[q][lane];
            }
        }
    }
    __syncthreads();   // the point data are read by every lane of the element before K_e replaces them

    // ---- 3. rotate to the node frames, mask the fixed dofs, store; K_e also into LDS
    if (live) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int k = lane + kElLanes * m;
            if (k < 21) {
                const int aa = ka[m], bb = kb[m];
                T v[4] = {kacc[m][0], kacc[m][1], kacc[m][2], kacc[m][3]};
                if (!a.cart) {
                    const T ca = lcs[le][aa][0], sa = lcs[le][aa][1], cb = lcs[le][bb][0], sb = lcs[le][bb][1];
                    const T t00 = ca * v[0] - sa * v[2], t01 = ca * v[1] - sa * v[3];
                    const T t10 = sa * v[0] + ca * v[2], t11 = sa * v[1] + ca * v[3];
                    v[0] = t00 * cb - t01 * sb;
                    v[1] = t00 * sb + t01 * cb;
                    v[2] = t10 * cb - t11 * sb;
                    v[3] = t10 * sb + t11 * cb;
                }
                if (a.bc) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (lfix[le][2 * aa + (j >> 1)] | lfix[le][2 * bb + (j & 1)]) v[j] = T(0);
                }
                T *o = a.EK + ((uint32_t)k * nel + e) * 4u;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    o[j] = v[j];
                    L[4 * k + j] = v[j];
                }
            }
            const int pg = lane + kElLanes * m;
            if (pg < 18) {
                const int an = pg / 3;
                T g0 = gacc[m][0], g1 = gacc[m][1];
                if (!a.cart) {
                    const T ca = lcs[le][an][0], sa = lcs[le][an][1];
                    const T r0 = ca * g0 - sa * g1, r1 = sa * g0 + ca * g1;
                    g0 = r0;
                    g1 = r1;
                }
                if (a.bc) {
                    if (lfix[le][2 * an]) g0 = T(0);
                    if (lfix[le][2 * an + 1]) g1 = T(0);
                }
                T *o = a.EG + ((uint32_t)pg * nel + e) * 2u;
                o[0] = g0;
                o[1] = g1;
            }
        }
        if (lane < 6) {
            a.EM[lane * nel + e] = macc;
            T f0 = facc[0], f1 = facc[1];
            if (!a.cart) {
                const T ca = lcs[le][lane][0], sa = lcs[le][lane][1];
                const T r0 = ca * f0 - sa * f1, r1 = sa * f0 + ca * f1;
                f0 = r0;
                f1 = r1;
            }
            T *o = a.EF + (lane * nel + e) * 2u;
            o[0] = f0;
            o[1] = f1;
        }
    }
    if (!a.proj) return;   // uniform over the grid: no barrier below is skipped by part of a block

    // ---- 4. P1 projection.  Ph (fine node f, vertex v) is the identity for f = v and, for the
    // midpoint 3 + k of edge (ea, eb) = (k, k+1 mod 3), 0.5 R(+h) towards ea and 0.5 R(-h) towards eb;
    // rows and columns of fixed dofs are zero (the multigrid acts on the constrained space).
    // T1 = K_e Ph (6 x 3 blocks); R K_e (vertex u, node f) = T1(f, u)^T with the fixed rows of f
    // masked; E1 = Ph^T T1.
    auto ph = [&](int f, int v, T P[4]) {
        P[0] = P[1] = P[2] = P[3] = T(0);
        if (f == v) { P[0] = T(1); P[3] = T(1); }
        else if (f >= 3) {
            const int k = f - 3, ea = k, eb = (k + 1) % 3;
            const T ch = T(0.5) * lph[le][k][0], sh = T(0.5) * lph[le][k][1];
            if (v == ea) { P[0] = ch; P[1] = -sh; P[2] = sh; P[3] = ch; }
            else if (v == eb) { P[0] = ch; P[1] = sh; P[2] = -sh; P[3] = ch; }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (lfix[le][2 * f + (j >> 1)] | lfix[le][2 * v + (j & 1)]) P[j] = T(0);
    };
    // the nine nonzero blocks Ph(g, v), g in {v, 3 + v, 3 + (v + 2) % 3}, tabulated once per element
    // after K_e (the point data are dead): PH + 4 (3 v + gi)
    T *PH = L + 84;
    if (live) {
        for (int idx = lane; idx < 9; idx += kElLanes) {
            const int v = idx / 3, gi = idx % 3;
            const int g = gi == 0 ? v : (gi == 1 ? 3 + v : 3 + (v + 2) % 3);
            T P[4];
            ph(g, v, P);
#pragma unroll
            for (int j = 0; j < 4; ++j) PH[4 * idx + j] = P[j];
        }
    }
    __syncthreads();

    T t1[3][4];   // T1 blocks b = lane + 8 m, held until every lane of the element has read K_e
    if (live) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int b = lane + kElLanes * m;
            T *t = t1[m];
            t[0] = t[1] = t[2] = t[3] = T(0);
            if (b < 18) {
                const int f = b / 3, v = b % 3;
                // nonzero Ph(g, v): g = v and the midpoints of the two edges at v
                const int gs[3] = {v, 3 + v, 3 + (v + 2) % 3};
#pragma unroll
                for (int gi = 0; gi < 3; ++gi) {
                    T K[4];
                    kblock(L, f, gs[gi], K);
                    const T *P = PH + 4 * (3 * v + gi);
                    t[0] += K[0] * P[0] + K[1] * P[2];
                    t[1] += K[0] * P[1] + K[1] * P[3];
                    t[2] += K[2] * P[0] + K[3] * P[2];
                    t[3] += K[2] * P[1] + K[3] * P[3];
                }
                // R K_e block (u = v, node f): entry (cu, rc) = T1(f, u)[rc][cu], zero on a fixed row dof of f
                T *rk = a.ERK + ((uint32_t)(v * 6 + f) * nel + e) * 4u;
#pragma unroll
                for (int cu = 0; cu < 2; ++cu)
#pragma unroll
                    for (int rc = 0; rc < 2; ++rc)
                        rk[cu * 2 + rc] = lfix[le][2 * f + rc] ? T(0) : t[rc * 2 + cu];
            }
        }
    }
    __syncthreads();   // K_e is read: T1 takes its place
    if (live) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int b = lane + kElLanes * m;
            if (b < 18) {
#pragma unroll
                for (int j = 0; j < 4; ++j) L[4 * b + j] = t1[m][j];
            }
        }
    }
    __syncthreads();
    if (live) {
        for (int b = lane; b < 9; b += kElLanes) {
            const int u = b / 3, v = b % 3;
            T E[4] = {T(0), T(0), T(0), T(0)};
            const int fs[3] = {u, 3 + u, 3 + (u + 2) % 3};
#pragma unroll
            for (int fi = 0; fi < 3; ++fi) {
                const T *P = PH + 4 * (3 * u + fi);
                const T *t = L + 4 * (fs[fi] * 3 + v);
                // E(u, v)[cu][cv] += sum_rc Ph(f, u)[rc][cu] T1(f, v)[rc][cv]
                E[0] += P[0] * t[0] + P[2] * t[2];
                E[1] += P[0] * t[1] + P[2] * t[3];
                E[2] += P[1] * t[0] + P[3] * t[2];
                E[3] += P[1] * t[1] + P[3] * t[3];
            }
#pragma unroll
            for (int cu = 0; cu < 2; ++cu)
#pragma unroll
                for (int cv = 0; cv < 2; ++cv)
                    a.E1[(uint32_t)((2 * u + cu) * 6 + 2 * v + cv) * nel + e] = E[cu * 2 + cv];
        }
    }
}

// out[q] = sum of the element blocks listed for stored position q (in element order).
// MODE 0: block code c at src[c * BS]; MODE 1: the same with the low bit of the code marking a
// transposed 2x2 block (K above its diagonal, code >> 1 is the block); MODE 2: the 36 x nel layout
// of the level-0 element matrices, code = (u * 3 + v) * nel + e for the 2x2 block (u, v).
template <typename T, int BS, int MODE>
__global__ __launch_bounds__(256) void k_gather(int64_t npos, const int32_t *__restrict__ ptr,
                                                const uint32_t *__restrict__ list, const T *__restrict__ src,
                                                int64_t nel, T *__restrict__ out) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= npos) return;
    T acc[BS];
#pragma unroll
    for (int j = 0; j < BS; ++j) acc[j] = T(0);
    const int32_t c0 = ptr[q], c1 = ptr[q + 1];
    for (int32_t c = c0; c < c1; ++c) {
        const uint32_t code = list[c];
        T v[BS];
        if constexpr (MODE == 2) {
            const int64_t blk = code / nel, e = code % nel;
            const int u = (int)(blk / 3), w = (int)(blk % 3);
#pragma unroll
            for (int j = 0; j < 4; ++j) v[j] = src[(int64_t)((2 * u + (j >> 1)) * 6 + 2 * w + (j & 1)) * nel + e];
        } else {
            const int64_t b = MODE == 1 ? (int64_t)(code >> 1) : (int64_t)code;
            const T *s = src + b * BS;
#pragma unroll
            for (int j = 0; j < BS; ++j) v[j] = s[j];
            if constexpr (MODE == 1) {
                if (code & 1u) { const T x = v[1]; v[1] = v[2]; v[2] = x; }
            }
        }
#pragma unroll
        for (int j = 0; j < BS; ++j) acc[j] += v[j];
    }
#pragma unroll
    for (int j = 0; j < BS; ++j) out[q * BS + j] = acc[j];
}

template <typename T, int BS, int MODE>
void gather(Context &c, const Context::GatherList &g, const T *src, T *out) {
    if (g.npos == 0) return;
    const int nb = (int)((g.npos + 255) / 256);
    k_gather<T, BS, MODE><<<dim3(nb), dim3(256), 0, c.stream>>>(g.npos, g.ptr.p, g.list.p, src, c.nel, out);
    STK_HIP(hipGetLastError());
}

}  // namespace

// Element assembly into the operators of the context's layouts (K, G, G^T sliced; M CSR; F per
// dof; with the multigrid: the level-0 element matrices E1, the P1 operator A1 and R K).
template <typename T>
void assemble_impl(Context &c, int flags, T *K, T *G, T *GT, T *M, T *F, T *A1, T *E1, T *RK) {
    const int64_t nel = c.nel;
    const bool proj = A1 != nullptr;
    if ((uint64_t)84 * (uint64_t)nel > (uint64_t)UINT32_MAX)
        throw Error(STK_E_SHAPE, "mesh too large for the element kernel's 32-bit offsets");
    DBuf<T> *bufp;
    if constexpr (sizeof(T) == 8) bufp = &c.el64;
    else bufp = &c.el32;
    DBuf<T> &buf = *bufp;
    // per element: K_e 84, G_e 36, M_e 6, f_e 12, R K_e 72
    buf.alloc((size_t)(84 + 36 + 6 + 12 + (proj ? 72 : 0)) * nel);
    ElemArgs<T> a{};
    a.nel = nel;
    a.rule = make_rule((flags & STK_ASM_7POINT) ? 7 : 6);
    a.e2n = c.d_e2n_nat.p; a.th = c.d_th_nat.p; a.r = c.d_r_nat.p;
    a.nodecs = c.d_nodecs.p; a.rho = c.d_rho.p; a.fixed = c.d_fixed.p;
    a.eta = c.d_eta.p; a.C = c.d_C.p; a.ccol = c.field_cols; a.g = c.g;
    a.cart = (flags & STK_ASM_CARTESIAN) ? 1 : 0;
    a.bc = (flags & STK_ASM_NO_BC) ? 0 : 1;
    a.iso = ((flags & STK_ASM_ISOTROPIC) || !c.have_C) ? 1 : 0;
    a.clamp = (flags & STK_ASM_CLAMP) ? 1 : 0;
    a.proj = proj ? 1 : 0;
    T *p = buf.p;
    a.EK = p; p += 84 * nel;
    a.EG = p; p += 36 * nel;
    a.EM = p; p += 6 * nel;
    a.EF = p; p += 12 * nel;
    a.ERK = proj ? p : nullptr;
    a.E1 = E1;
    const int nb = (int)((nel + kElPerBlock - 1) / kElPerBlock);
    if (a.rule.nq == 7) k_element<T, 7><<<dim3(nb), dim3(256), 0, c.stream>>>(a);
    else k_element<T, 6><<<dim3(nb), dim3(256), 0, c.stream>>>(a);
    STK_HIP(hipGetLastError());
    gather<T, 4, 1>(c, c.gK, a.EK, K);
    gather<T, 2, 0>(c, c.gG, a.EG, G);
    gather<T, 2, 0>(c, c.gGT, a.EG, GT);
    gather<T, 1, 0>(c, c.gM, a.EM, M);
    gather<T, 2, 0>(c, c.gF, a.EF, F);
    if (proj) {
        gather<T, 4, 2>(c, c.gA1, E1, A1);
        gather<T, 4, 0>(c, c.gRK, a.ERK, RK);
    }
}

template void assemble_impl<double>(Context &, int, double *, double *, double *, double *, double *, double *, double *,
                                    double *);
template void assemble_impl<float>(Context &, int, float *, float *, float *, float *, float *, float *, float *,
                                   float *);

}  // namespace stk
