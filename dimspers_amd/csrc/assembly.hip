// Element-level kernels beside the assembly proper (csrc/element.hip):
//
//  * integrate_element: one element per thread with the 6-point rule, following the reference block
//    assembly (src/Stokes/StokesSolver.jl:247-404): _fill_R_J! (:522-541), the polar -> Cartesian
//    "double Jacobian" of _derivative_weights! (:641-686), _fill_Kblk! (aniso,
//    src/src/Stokes/StokesSolver.jl:744-808; iso :688-742), _fill_Gblk! (:812-823), _fill_Mblk!
//    (:825-836), _fill_Fbblk! (:838-845).  k_schur_samples uses it to form the exact local Schur
//    blocks the learned preconditioner is trained on;
//  * k_learned_scatter: the predicted local Schur blocks added into M's pattern, colour by colour
//    (elements of one colour share no vertex, so the additions are race free and in a fixed order);
//  * k_galerkin / k_galerkin4: the coarse multigrid levels, coarse element = sum over its four
//    children of P^T E P (nested meshes of two or four triangles per cell, rotation-frame
//    interpolation blocks, masked boundary dofs), colour scatter into the coarse operator;
//  * k_fix_diag: the unit diagonal of the fixed dofs.
#include <cmath>

#include "device_ops.h"
#include "stk_internal.h"

namespace stk {

// Integration rule of the element assembly, passed by value in every launch.
struct QuadTables {
    double xi[6], eta[6], w[6];
};

namespace {

constexpr int kAsmThreads = 64;

// 6-point rule (src/Algebra/Quadrature.jl:80-107), weights include the 1/2.
QuadTables make_quad() {
    QuadTables q;
    const double s = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
    const double g1 = (8.0 - std::sqrt(10.0) + s) / 18.0;
    const double g2 = (8.0 - std::sqrt(10.0) - s) / 18.0;
    const double w1 = (620.0 + std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0))) / 3720.0;
    const double w2 = (620.0 - std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0))) / 3720.0;
    const double x[6][2] = {{1 - 2 * g1, g1}, {g1, 1 - 2 * g1}, {g1, g1},
                            {1 - 2 * g2, g2}, {g2, 1 - 2 * g2}, {g2, g2}};
    for (int i = 0; i < 6; ++i) {
        q.xi[i] = x[i][0];
        q.eta[i] = x[i][1];
        q.w[i] = 0.5 * (i < 3 ? w1 : w2);
    }
    return q;
}

// P2 shape functions and derivatives (Quadrature.jl:264-307); node order v1 v2 v3 m12 m23 m31.
template <typename T>
__device__ __forceinline__ void p2_shape(T r, T s, T N[6], T dr[6], T ds[6]) {
    const T t = T(1) - r - s;
    N[0] = t * (T(2) * t - T(1)); N[1] = r * (T(2) * r - T(1)); N[2] = s * (T(2) * s - T(1));
    N[3] = T(4) * r * t; N[4] = T(4) * r * s; N[5] = T(4) * s * t;
    dr[0] = -(T(4) * t - T(1)); dr[1] = T(4) * r - T(1); dr[2] = T(0);
    dr[3] = T(4) * (t - r); dr[4] = T(4) * s; dr[5] = -T(4) * s;
    ds[0] = -(T(4) * t - T(1)); ds[1] = T(0); ds[2] = T(4) * s - T(1);
    ds[3] = -T(4) * r; ds[4] = T(4) * r; ds[5] = T(4) * (t - s);
}

// index of lower block (a >= b) among the 21 node-pair blocks
__device__ __forceinline__ int lblk(int a, int b) { return a * (a + 1) / 2 + b; }

// Integrates one element with the 6-point rule (StokesSolver.jl:543-845): the lower
// node-pair 2x2 blocks of K_e into sK[.][tid] (LDS), M_e (upper triangle, 6) into
// sM[.][tid] (LDS), G_e (node a, pressure i, component) and f_e into registers.
// eta_at(q) gives eta.ip at integration point q, tensor_at(q, d) the six viscous
// tensor components (eta11, eta33, eta55, eta13, eta15, eta35) there.
template <typename T, typename EtaF, typename TensF>
__device__ __forceinline__ void integrate_element(const QuadTables &qt, const T th[3], const T rr[3], const T rhov[3],
                                                  T g, bool iso,
                                                  EtaF eta_at, TensF tensor_at, T (*sK)[kAsmThreads],
                                                  T (*sM)[kAsmThreads], int tid, T sG[36], T sF[12]) {
    // _fill_R_J!
    const T R21 = rr[1] - rr[0], R31 = rr[2] - rr[0];
    const T Th21 = th[1] - th[0], Th31 = th[2] - th[0];
    const T detJ = Th21 * R31 - Th31 * R21;
    for (int k = 0; k < 84; ++k) sK[k][tid] = T(0);
#pragma unroll
    for (int k = 0; k < 36; ++k) sG[k] = T(0);
    for (int k = 0; k < 6; ++k) sM[k][tid] = T(0);
#pragma unroll
    for (int k = 0; k < 12; ++k) sF[k] = T(0);

    for (int q = 0; q < 6; ++q) {
        const T xi = (T)qt.xi[q], et = (T)qt.eta[q];
        T N[6], dr[6], ds[6];
        p2_shape<T>(xi, et, N, dr, ds);
        const T NP[3] = {T(1) - xi - et, xi, et};
        const T thq = NP[0] * th[0] + NP[1] * th[1] + NP[2] * th[2];
        const T rq = NP[0] * rr[0] + NP[1] * rr[1] + NP[2] * rr[2];
        T sn, cs;
        if constexpr (sizeof(T) == 8) sincos(thq, &sn, &cs);
        else sincosf(thq, &sn, &cs);
        const T inv = T(1) / rq / detJ;
        const T jx0 = (R31 * cs - Th31 * sn * rq) * inv;
        const T jx1 = (-R21 * cs + Th21 * sn * rq) * inv;
        const T jz0 = (-R31 * sn - Th31 * cs * rq) * inv;
        const T jz1 = (R21 * sn + Th21 * cs * rq) * inv;
        const T om = rq * detJ * (T)qt.w[q];
        T dx[6], dz[6];
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            dx[b] = jx0 * dr[b] + jx1 * ds[b];
            dz[b] = jz0 * dr[b] + jz1 * ds[b];
        }
        const T eta = eta_at(q);
        T d11, d22, d33, d12, d13, d23;
        if (iso) {
            d11 = d22 = T(4) / T(3); d12 = -T(2) / T(3); d33 = T(1); d13 = d23 = T(0);
        } else {
            T d[6];
            tensor_at(q, d);
            d11 = d[0]; d22 = d[1]; d33 = d[2]; d12 = d[3]; d13 = d[4]; d23 = d[5];
        }
        const T sc = om * eta;
        d11 *= sc; d22 *= sc; d33 *= sc; d12 *= sc; d13 *= sc; d23 *= sc;
        // D B columns of node b for the x and z dofs
        T Dx[6][3], Dz[6][3];
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            Dx[b][0] = d11 * dx[b] + d13 * dz[b];
            Dx[b][1] = d12 * dx[b] + d23 * dz[b];
            Dx[b][2] = d13 * dx[b] + d33 * dz[b];
            Dz[b][0] = d12 * dz[b] + d13 * dx[b];
            Dz[b][1] = d22 * dz[b] + d23 * dx[b];
            Dz[b][2] = d23 * dz[b] + d33 * dx[b];
        }
#pragma unroll
        for (int aa = 0; aa < 6; ++aa)
#pragma unroll
            for (int b = 0; b <= aa; ++b) {
                const int k = lblk(aa, b) * 4;
                sK[k + 0][tid] += dx[aa] * Dx[b][0] + dz[aa] * Dx[b][2];
                sK[k + 1][tid] += dx[aa] * Dz[b][0] + dz[aa] * Dz[b][2];
                sK[k + 2][tid] += dz[aa] * Dx[b][1] + dx[aa] * Dx[b][2];
                sK[k + 3][tid] += dz[aa] * Dz[b][1] + dx[aa] * Dz[b][2];
            }
#pragma unroll
        for (int aa = 0; aa < 6; ++aa)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                sG[(aa * 3 + i) * 2 + 0] += om * NP[i] * dx[aa];
                sG[(aa * 3 + i) * 2 + 1] += om * NP[i] * dz[aa];
            }
        const T wm = om / eta;
        sM[0][tid] += wm * NP[0] * NP[0]; sM[1][tid] += wm * NP[0] * NP[1]; sM[2][tid] += wm * NP[0] * NP[2];
        sM[3][tid] += wm * NP[1] * NP[1]; sM[4][tid] += wm * NP[1] * NP[2]; sM[5][tid] += wm * NP[2] * NP[2];
        const T rhoq = NP[0] * rhov[0] + NP[1] * rhov[1] + NP[2] * rhov[2];
        const T cte = om * g * rhoq;
#pragma unroll
        for (int aa = 0; aa < 6; ++aa) {
            sF[2 * aa + 0] += sn * cte * N[aa];
            sF[2 * aa + 1] += cs * cte * N[aa];
        }
    }

}

// ------------------------------------------------------------------------ local Schur blocks
// Scalar (i >= j) entry of the 12x12 element matrix held as lower node-pair blocks.
__device__ __forceinline__ int kidx(int i, int j) { return lblk(i >> 1, j >> 1) * 4 + (i & 1) * 2 + (j & 1); }

// Exact local Schur block of one element, frame independent:
//   S_e = G_e^T (K_e + mu N N^T)^-1 G_e,
// N = orthonormal Cartesian rigid motions of the six nodes (two translations and the
// rotation about the node centroid), mu = trace(K_e) / 12.  K_e vanishes on these
// motions up to the curvature of the polar map and G_e^T annihilates them (rigid
// motions are divergence free), so S_e is the pressure Schur complement of the
// element on its deformation modes.  Uses sK (destroyed) and the G_e registers.
template <typename T>
__device__ void local_schur(T (*sK)[kAsmThreads], int tid, const T sG[36], const T xn[6], const T zn[6], T S[3][3]) {
    T nv[3][12];
    T xc = 0, zc = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) { xc += xn[a]; zc += zn[a]; }
    xc /= T(6); zc /= T(6);
    T nr = 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        nv[0][2 * a] = T(1); nv[0][2 * a + 1] = T(0);
        nv[1][2 * a] = T(0); nv[1][2 * a + 1] = T(1);
        nv[2][2 * a] = zn[a] - zc; nv[2][2 * a + 1] = -(xn[a] - xc);
        nr += (zn[a] - zc) * (zn[a] - zc) + (xn[a] - xc) * (xn[a] - xc);
    }
    const T s6 = T(1) / sqrt(T(6)), sr = T(1) / sqrt(nr);
#pragma unroll
    for (int i = 0; i < 12; ++i) { nv[0][i] *= s6; nv[1][i] *= s6; nv[2][i] *= sr; }
    T tr = 0;
    for (int i = 0; i < 12; ++i) tr += sK[kidx(i, i)][tid];
    const T mu = tr / T(12);
    for (int i = 0; i < 12; ++i)
        for (int j = 0; j <= i; ++j)
            sK[kidx(i, j)][tid] += mu * (nv[0][i] * nv[0][j] + nv[1][i] * nv[1][j] + nv[2][i] * nv[2][j]);
    // Cholesky in place
    for (int j = 0; j < 12; ++j) {
        T d = sK[kidx(j, j)][tid];
        for (int k = 0; k < j; ++k) { const T l = sK[kidx(j, k)][tid]; d -= l * l; }
        const T ljj = sqrt(d);
        sK[kidx(j, j)][tid] = ljj;
        const T inv = T(1) / ljj;
        for (int i = j + 1; i < 12; ++i) {
            T v = sK[kidx(i, j)][tid];
            for (int k = 0; k < j; ++k) v -= sK[kidx(i, k)][tid] * sK[kidx(j, k)][tid];
            sK[kidx(i, j)][tid] = v * inv;
        }
    }
    // X = Kr^-1 G (three columns), S = G^T X
    for (int c = 0; c < 3; ++c) {
        T x[12];
        for (int i = 0; i < 12; ++i) {
            T v = sG[((i >> 1) * 3 + c) * 2 + (i & 1)];
            for (int k = 0; k < i; ++k) v -= sK[kidx(i, k)][tid] * x[k];
            x[i] = v / sK[kidx(i, i)][tid];
        }
        for (int i = 11; i >= 0; --i) {
            T v = x[i];
            for (int k = i + 1; k < 12; ++k) v -= sK[kidx(k, i)][tid] * x[k];
            x[i] = v / sK[kidx(i, i)][tid];
        }
        for (int c2 = 0; c2 < 3; ++c2) {
            T s = 0;
            for (int i = 0; i < 12; ++i) s += sG[((i >> 1) * 3 + c2) * 2 + (i & 1)] * x[i];
            S[c2][c] = s;
        }
    }
}

// directRotation2D (src/Rheology/anisotropy.jl:503-536) of the fabric-frame tensor
// a = c = 4/3, h = -2/3, e = 1/delta with rotmatrix(cos phi, sin phi) (:287-292).
__device__ __forceinline__ void fabric_tensor(double phi, double delta, double d[6]) {
    const double x1 = cos(phi), x2 = sin(phi), y1 = -x2, y2 = x1;
    const double a = 4.0 / 3.0, c = 4.0 / 3.0, h = -2.0 / 3.0, e = 1.0 / delta;
    const double x1x1 = x1 * x1, x2x2 = x2 * x2, y1y1 = y1 * y1, y2y2 = y2 * y2, s = x1 * y2 + x2 * y1;
    d[0] = (c * x2x2 + h * x1x1 + 4.0 * e * x1x1) * x2x2 + (a * x1x1 + h * x2x2) * x1x1;
    d[1] = (c * y2y2 + h * y1y1 + 4.0 * e * y1y1) * y2y2 + (a * y1y1 + h * y2y2) * y1y1;
    d[2] = (c * x2 * y2 + h * x1 * y1) * x2 * y2 + s * s * e + (a * x1 * y1 + h * x2 * y2) * x1 * y1;
    d[3] = ((c * x2x2 + h * x1x1) * y2 + 4.0 * e * x1 * x2 * y1) * y2 + (a * x1x1 + h * x2x2) * y1y1;
    d[4] = ((c * x2x2 + h * x1x1) * y2 + 2.0 * s * e * x1) * x2 + (a * x1x1 + h * x2x2) * x1 * y1;
    d[5] = ((c * y2y2 + h * y1y1) * x2 + 2.0 * s * e * y1) * y2 + (a * y1y1 + h * y2y2) * x1 * y1;
}

__device__ __forceinline__ double splitmix_u(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

// Canonical vertex order of the learned blocks by triangle type (element id modulo the
// triangles per cell).  Two per cell: right-angle vertex, its azimuthal neighbour, its radial
// neighbour.  Four per cell ([corner k, corner k+1, centre]): the corners swapped for types 2
// and 3, so the mirror images of types 0/2 (in r) and 1/3 (in theta) get the same block.
__device__ __forceinline__ const int *canon_order(int split, int64_t e) {
    static constexpr int c2[2][3] = {{0, 1, 2}, {1, 2, 0}};
    static constexpr int c4[4][3] = {{0, 1, 2}, {0, 1, 2}, {1, 0, 2}, {1, 0, 2}};
    return split == 4 ? c4[e & 3] : c2[e & 1];
}

struct SchurSampleArgs {
    int64_t nel, nsamp, per_snap, first_snap;
    int split;                   // triangles per cell (2 or 4)
    int mode;                    // 0: the current fields of every element; 1: seeded training draws
    int ratio_kind;              // 0: delta = 10, 1: U[1, 100], 2: log-uniform [1, 1e3]
    double mean_area;
    const int32_t *iperm;        // element -> colour position
    const double *th, *r;        // 3 x nel, by position
    const double *eta;           // 6 x nel by element (ip-major)
    const double *phi, *delta;   // by element (mode 0)
    float *X;                    // 5 x nsamp features (sample-major)
    float *Tg;                   // 6 x nsamp log-Cholesky targets (mode 1, may be null)
    double *S;                   // 6 x nsamp canonical S_e / s_e (may be null)
    double *scale;               // s_e = area / eta_bar (may be null)
    QuadTables quad;
};

// Features (and, for training, exact targets) of the learned local-Schur model:
// mean log eta, cos / sin 2(phi - theta_c), log delta, and the element area relative to the
// mean area of its cell's triangles (they differ by the curvature of the polar map, which is
// the largest variation of the normalised blocks: inner, side and outer triangles of a cell
// fall in separate bands).  Four triangles per cell: the director is measured from the
// corner-corner edge (phi - theta_c, less pi/2 on the radial edges of types 1 and 3) with the
// sign of its sine flipped for the mirrored types 2 and 3, so a triangle and its rotated or
// mirrored counterpart see the same inputs.  Blocks in the canonical order of canon_order.
__global__ __launch_bounds__(kAsmThreads) void k_schur_samples(SchurSampleArgs a) {
    __shared__ double sK[84][kAsmThreads];
    __shared__ double sM[6][kAsmThreads];
    const int tid = threadIdx.x;
    const int64_t sidx = (int64_t)blockIdx.x * kAsmThreads + tid;
    if (sidx >= a.nsamp) return;
    int64_t e;
    double fscale = 1.0, phi, delta;
    if (a.mode == 0) {
        e = sidx;
        phi = a.phi[e];
        delta = a.delta[e];
    } else {
        const int64_t snap = a.first_snap + sidx / a.per_snap, k = sidx % a.per_snap;
        const uint64_t seed = 0x5EED1000ull + (uint64_t)snap;
        e = (int64_t)(splitmix_u(seed, 8 * k) * a.nel);
        if (e >= a.nel) e = a.nel - 1;
        fscale = exp((2.0 * splitmix_u(seed, 8 * k + 1) - 1.0) * 2.302585092994046);
        phi = M_PI * splitmix_u(seed, 8 * k + 2);
        const double u = splitmix_u(seed, 8 * k + 3);
        delta = a.ratio_kind == 0 ? 10.0 : a.ratio_kind == 1 ? 1.0 + 99.0 * u : pow(10.0, 3.0 * u);
    }
    const int64_t pos = a.iperm[e];
    double th[3], rr[3], rhov[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) { th[k] = a.th[k * a.nel + pos]; rr[k] = a.r[k * a.nel + pos]; }
    double tens[6];
    fabric_tensor(phi, delta, tens);
    double logeta = 0;
    for (int q = 0; q < 6; ++q) logeta += log(a.eta[(int64_t)q * a.nel + e] * fscale);
    logeta /= 6.0;
    auto eta_at = [&](int q) { return a.eta[(int64_t)q * a.nel + e] * fscale; };
    auto tensor_at = [&](int, double d[6]) { for (int k = 0; k < 6; ++k) d[k] = tens[k]; };
    double sG[36], sF[12];
    integrate_element<double>(a.quad, th, rr, rhov, 0.0, false, eta_at, tensor_at, sK, sM, tid, sG, sF);
    const double detJ = (th[1] - th[0]) * (rr[2] - rr[0]) - (th[2] - th[0]) * (rr[1] - rr[0]);
    const double area = detJ * (rr[0] + rr[1] + rr[2]) / 6.0;
    const double thc = (th[0] + th[1] + th[2]) / 3.0;
    // area relative to the mean of its cell's triangles (elements split * q .. split * q + split - 1)
    double shape;
    {
        const int64_t e0 = e - e % a.split;
        double sum = 0.0;
        for (int k = 0; k < a.split; ++k) {
            const int64_t pp = a.iperm[e0 + k];
            double pt[3], pr[3];
            for (int v = 0; v < 3; ++v) { pt[v] = a.th[v * a.nel + pp]; pr[v] = a.r[v * a.nel + pp]; }
            sum += ((pt[1] - pt[0]) * (pr[2] - pr[0]) - (pt[2] - pt[0]) * (pr[1] - pr[0])) * (pr[0] + pr[1] + pr[2]) / 6.0;
        }
        shape = area / (sum / a.split);
    }
    float *X = a.X + 5 * sidx;
    double psi = phi - thc, sgn = 1.0;
    if (a.split == 4) {
        const int t = (int)(e & 3);
        if (t & 1) psi -= 0.5 * M_PI;
        if (t >= 2) sgn = -1.0;
    }
    X[0] = (float)logeta;
    X[1] = (float)cos(2.0 * psi);
    X[2] = (float)(sgn * sin(2.0 * psi));
    X[3] = (float)log(delta);
    X[4] = (float)shape;
    const double se = area / exp(logeta);
    if (a.scale) a.scale[sidx] = se;
    if (!a.Tg && !a.S) return;
    double xn[6], zn[6];
    const int ea[3] = {0, 1, 2}, eb[3] = {1, 2, 0};
    for (int k = 0; k < 3; ++k) { xn[k] = rr[k] * sin(th[k]); zn[k] = rr[k] * cos(th[k]); }
    for (int k = 0; k < 3; ++k) {
        const double tm = 0.5 * (th[ea[k]] + th[eb[k]]), rm = 0.5 * (rr[ea[k]] + rr[eb[k]]);
        xn[3 + k] = rm * sin(tm); zn[3 + k] = rm * cos(tm);
    }
    double Sl[3][3];
    local_schur<double>(sK, tid, sG, xn, zn, Sl);
    const int *pp = canon_order(a.split, e);
    double Sc[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Sc[i][j] = Sl[pp[i]][pp[j]] / se;
    if (a.S) {
        double *S = a.S + 6 * sidx;
        S[0] = Sc[0][0]; S[1] = Sc[1][0]; S[2] = Sc[1][1]; S[3] = Sc[2][0]; S[4] = Sc[2][1]; S[5] = Sc[2][2];
    }
    if (a.Tg) {
        // log-Cholesky parameters: (log L00, L10, log L11, L20, L21, log L22)
        const double l00 = sqrt(Sc[0][0]);
        const double l10 = Sc[1][0] / l00, l20 = Sc[2][0] / l00;
        const double l11 = sqrt(Sc[1][1] - l10 * l10);
        const double l21 = (Sc[2][1] - l20 * l10) / l11;
        const double l22 = sqrt(Sc[2][2] - l20 * l20 - l21 * l21);
        float *Tg = a.Tg + 6 * sidx;
        Tg[0] = (float)log(l00); Tg[1] = (float)l10; Tg[2] = (float)log(l11);
        Tg[3] = (float)l20; Tg[4] = (float)l21; Tg[5] = (float)log(l22);
    }
}

// Predicted element blocks -> assembled S (pattern of M), colour by colour.
struct LearnedScatterArgs {
    int64_t nel, off, cnt;
    int split;                      // triangles per cell (2 or 4)
    const int32_t *perm, *e2n;      // e2n by position
    const uint8_t *sM;
    const int32_t *Mptr;
    const float *Y;                 // 6 x nel log-Cholesky parameters (denormalised), element order
    const double *scale;            // s_e
    double *S64;
};

__global__ void k_learned_scatter(LearnedScatterArgs a) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.cnt) return;
    const int64_t p = a.off + j, nel = a.nel;
    const int64_t e = a.perm[p];
    const float *y = a.Y + 6 * e;
    const double l00 = exp((double)y[0]), l10 = y[1], l11 = exp((double)y[2]);
    const double l20 = y[3], l21 = y[4], l22 = exp((double)y[5]);
    const double se = a.scale[e];
    double Sc[3][3];
    Sc[0][0] = l00 * l00;
    Sc[1][0] = Sc[0][1] = l10 * l00;
    Sc[1][1] = l10 * l10 + l11 * l11;
    Sc[2][0] = Sc[0][2] = l20 * l00;
    Sc[2][1] = Sc[1][2] = l20 * l10 + l21 * l11;
    Sc[2][2] = l20 * l20 + l21 * l21 + l22 * l22;
    const int *pp = canon_order(a.split, e);
    double Sl[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) Sl[pp[i]][pp[k]] = Sc[i][k] * se;
    int32_t n[3];
    for (int k = 0; k < 3; ++k) n[k] = a.e2n[k * nel + p];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            const int64_t bm = (int64_t)a.Mptr[n[i]] + a.sM[(i * 3 + k) * nel + p];
            a.S64[bm] += Sl[i][k];
        }
}

// ------------------------------------------------------------------------ coarse Galerkin
// Coarse P1 element (I, J, t) of level l+1 is the union of four level-l elements.
// Child table: (di, dj, type) of the fine cells (2I+di, 2J+dj).
__constant__ int8_t c_child[2][4][3] = {
    {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0}},
    {{1, 1, 0}, {1, 1, 1}, {1, 0, 1}, {0, 1, 1}},
};

template <typename T>
struct GalArgs {
    int32_t ncr_c, nct_c, ncr_f, nct_f;
    int64_t nel_c, nel_f, off, cnt;
    const int32_t *perm, *e2v;      // coarse colour order, e2v SoA by position
    const uint8_t *slot;            // 9 x nel_c by position
    const int32_t *ptr;             // coarse pattern rows
    const int64_t *sptr;            // coarse sliced layout (null: CSR order)
    const T *Ef;                    // 36 x nel_f by element id
    T *Ec;                          // 36 x nel_c by element id
    T *A;                           // coarse BSR 2x2
};

template <typename T>
__device__ __forceinline__ void tri_verts(int t, int I, int J, int vi[3], int vj[3]) {
    if (t == 0) { vi[0] = I; vj[0] = J; vi[1] = I; vj[1] = J + 1; vi[2] = I + 1; vj[2] = J; }
    else { vi[0] = I; vj[0] = J + 1; vi[1] = I + 1; vj[1] = J + 1; vi[2] = I + 1; vj[2] = J; }
}

// E += Ph^T Kf Ph for one child: Kf its 6x6 element matrix (element ef of the fine level), Ph[f][v]
// the 2x2 interpolation block from coarse vertex v to the child's fine vertex f.
template <typename T>
__device__ __forceinline__ void galerkin_child(const GalArgs<T> &a, int64_t ef, const T (&Ph)[3][3][4],
                                               T (&E)[6][6]) {
    T Kf[6][6];
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
        for (int v = 0; v < 6; ++v) Kf[u][v] = a.Ef[(int64_t)(u * 6 + v) * a.nel_f + ef];
    T T1[6][6];   // Kf Ph
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
                T s = T(0);
#pragma unroll
                for (int f = 0; f < 3; ++f) s += Kf[r][2 * f] * Ph[f][v][cc] + Kf[r][2 * f + 1] * Ph[f][v][2 + cc];
                T1[r][2 * v + cc] = s;
            }
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int cu = 0; cu < 2; ++cu)
#pragma unroll
            for (int c2 = 0; c2 < 6; ++c2) {
                T s = T(0);
#pragma unroll
                for (int f = 0; f < 3; ++f) s += Ph[f][u][cu] * T1[2 * f][c2] + Ph[f][u][2 + cu] * T1[2 * f + 1][c2];
                E[2 * u + cu][c2] += s;
            }
}

// Interpolation block: weight w, rotation by the fine-minus-coarse angle (sine sd, cosine cd), radial
// row / column zero on a fixed fine / coarse dof.
template <typename T>
__device__ __forceinline__ void galerkin_block(T w, T sd, T cd, bool ffix, bool cfix, T (&blk)[4]) {
    blk[0] = w * cd;
    blk[1] = cfix ? T(0) : -w * sd;
    blk[2] = ffix ? T(0) : w * sd;
    blk[3] = (ffix || cfix) ? T(0) : w * cd;
}

// Store the coarse element matrix (by element id) and add it into the coarse operator.
template <typename T>
__device__ __forceinline__ void galerkin_store(const GalArgs<T> &a, int64_t p, int64_t ec, const T (&E)[6][6]) {
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
        for (int v = 0; v < 6; ++v) a.Ec[(int64_t)(u * 6 + v) * a.nel_c + ec] = E[u][v];
    int32_t vid[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) vid[k] = a.e2v[k * a.nel_c + p];
    for (int u = 0; u < 3; ++u)
        for (int v = 0; v < 3; ++v) {
            const int64_t sl = a.slot[(u * 3 + v) * a.nel_c + p];
            const int64_t bm = a.sptr ? a.sptr[vid[u] / Pattern::kSlice] + sl * Pattern::kSlice + vid[u] % Pattern::kSlice
                                      : (int64_t)a.ptr[vid[u]] + sl;
            a.A[4 * bm + 0] += E[2 * u][2 * v];
            a.A[4 * bm + 1] += E[2 * u][2 * v + 1];
            a.A[4 * bm + 2] += E[2 * u + 1][2 * v];
            a.A[4 * bm + 3] += E[2 * u + 1][2 * v + 1];
        }
}

template <typename T>
__device__ __forceinline__ void sincos_g(T d, T *s, T *c) {
    if constexpr (sizeof(T) == 8) sincos(d, s, c);
    else sincosf(d, s, c);
}

// v[k] for a lane-dependent k in 0..4, by selects (a dynamic index would put v in scratch memory)
template <typename T>
__device__ __forceinline__ T pick5(const T (&v)[5], int k) {
    T r = v[0];
    r = k == 1 ? v[1] : r;
    r = k == 2 ? v[2] : r;
    r = k == 3 ? v[3] : r;
    r = k == 4 ? v[4] : r;
    return r;
}

// Four lanes per coarse element, one child each (the four lanes of an element are consecutive in a
// wavefront and share its range test); lane 0 sums the four child products in child order, as one
// thread adding the children in turn would, and stores the coarse element.
template <typename T>
__global__ __launch_bounds__(256) void k_galerkin(GalArgs<T> a) {
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t j = gid / 4;
    const int ch = (int)(gid % 4);
    if (j >= a.cnt) return;
    const int64_t p = a.off + j;
    const int64_t ec = a.perm[p];
    const int t = (int)(ec % 2);
    const int64_t qc = ec / 2;
    const int I = (int)(qc % a.ncr_c), J = (int)(qc / a.ncr_c);
    int ci[3], cj[3];
    tri_verts<T>(t, I, J, ci, cj);
    const T dthc = T(2.0 * M_PI) / T(a.nct_c);
    // A fine vertex lies half a coarse cell from a coarse one at most: the fine-minus-coarse angle is
    // dthc h with h = fj / 2 - cj in {-1, -1/2, 0, 1/2, 1}, so five sincos serve all the blocks.
    T sh[5], chh[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) sincos_g(dthc * (T(0.5) * T(k) - T(1)), &sh[k], &chh[k]);
    T E[6][6];
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
        for (int v = 0; v < 6; ++v) E[u][v] = T(0);
    {
        const int di = c_child[t][ch][0], dj = c_child[t][ch][1], tt = c_child[t][ch][2];
        const int fi0 = 2 * I + di, fj0 = 2 * J + dj;
        const int64_t ef = 2 * ((int64_t)fi0 + (int64_t)a.ncr_f * fj0) + tt;
        int fi[3], fj[3];
        tri_verts<T>(tt, fi0, fj0, fi, fj);
        // interpolation blocks Ph[fine vertex][coarse vertex]
        T Ph[3][3][4];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            // normalised position in the coarse cell
            const T rho = T(0.5) * (fi[f] - 2 * I), tau = T(0.5) * (fj[f] - 2 * J);
            T lam[3];
            if (t == 0) { lam[0] = T(1) - rho - tau; lam[1] = tau; lam[2] = rho; }
            else { lam[0] = T(1) - rho; lam[1] = rho + tau - T(1); lam[2] = T(1) - tau; }
            const bool ffix = (fi[f] == 0) || (fi[f] == a.ncr_f);
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                const bool cfix = (ci[v] == 0) || (ci[v] == a.ncr_c);
                const int k = fj[f] - 2 * cj[v] + 2;   // 2 h + 2
                galerkin_block<T>(lam[v], pick5(sh, k), pick5(chh, k), ffix, cfix, Ph[f][v]);
            }
        }
        galerkin_child<T>(a, ef, Ph, E);
    }
    const int lb = (int)(threadIdx.x % 64) & ~3;   // lane of child 0 in the wavefront
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
        for (int v = 0; v < 6; ++v) {
            const T e1 = __shfl(E[u][v], lb + 1), e2 = __shfl(E[u][v], lb + 2), e3 = __shfl(E[u][v], lb + 3);
            E[u][v] = ((E[u][v] + e1) + e2) + e3;
        }
    if (ch == 0) galerkin_store<T>(a, p, ec, E);
}

// Four triangles per cell (Grid): coarse element 4q + t = [K_t, K_t+1, centre] of cell (I, J) is the
// union of two triangles in each of two fine cells (2I + a, 2J + b): c_child4[t] lists (a, b, type).
// Corner k of a cell sits at the lattice offset (kCornerDi[k], kCornerDj[k]) (radial, azimuthal).
__constant__ int8_t c_child4[4][4][3] = {
    {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 3}},
    {{0, 1, 1}, {0, 1, 2}, {1, 1, 0}, {1, 1, 1}},
    {{1, 1, 2}, {1, 1, 3}, {1, 0, 1}, {1, 0, 2}},
    {{1, 0, 3}, {1, 0, 0}, {0, 0, 2}, {0, 0, 3}},
};
__constant__ int8_t c_cdi[4] = {0, 0, 1, 1}, c_cdj[4] = {0, 1, 1, 0};

template <typename T>
__global__ __launch_bounds__(64) void k_galerkin4(GalArgs<T> a) {
    const int64_t j = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (j >= a.cnt) return;
    const int64_t p = a.off + j;
    const int64_t ec = a.perm[p];
    const int t = (int)(ec % 4);
    const int64_t qc = ec / 4;
    const int I = (int)(qc % a.ncr_c), J = (int)(qc / a.ncr_c);
    // coarse vertices in cell-local (radial u, azimuthal v) units, their angles (coarse cell widths,
    // unwrapped) and fixed radial dofs (wall corners only)
    T cu[3], cv[3], cpos[3];
    bool cfix[3];
    for (int k = 0; k < 2; ++k) {
        const int kk = (t + k) % 4;
        cu[k] = T(c_cdi[kk]);
        cv[k] = T(c_cdj[kk]);
        cpos[k] = T(J + c_cdj[kk]);
        const int ci = I + c_cdi[kk];
        cfix[k] = (ci == 0) || (ci == a.ncr_c);
    }
    cu[2] = T(0.5); cv[2] = T(0.5); cpos[2] = T(J) + T(0.5); cfix[2] = false;
    const T det = (cu[0] - cu[2]) * (cv[1] - cv[2]) - (cu[1] - cu[2]) * (cv[0] - cv[2]);
    const T dthc = T(2.0 * M_PI) / T(a.nct_c);
    T E[6][6];
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
        for (int v = 0; v < 6; ++v) E[u][v] = T(0);
    for (int ch = 0; ch < 4; ++ch) {
        const int da = c_child4[t][ch][0], db = c_child4[t][ch][1], tt = c_child4[t][ch][2];
        const int64_t ef = 4 * ((int64_t)(2 * I + da) + (int64_t)a.ncr_f * (2 * J + db)) + tt;
        T Ph[3][3][4];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            // fine vertex f of the child: corners tt and tt+1 of its cell, then the centre
            T fu, fv, fpos;
            bool ffix = false;
            if (f < 2) {
                const int kk = (tt + f) % 4;
                fu = T(0.5) * T(da + c_cdi[kk]);
                fv = T(0.5) * T(db + c_cdj[kk]);
                fpos = T(J) + fv;
                const int fi = 2 * I + da + c_cdi[kk];
                ffix = (fi == 0) || (fi == a.ncr_f);
            } else {
                fu = T(0.5) * (T(da) + T(0.5));
                fv = T(0.5) * (T(db) + T(0.5));
                fpos = T(J) + fv;
            }
            // barycentric coordinates of the fine vertex in the coarse triangle
            const T l0 = ((fu - cu[2]) * (cv[1] - cv[2]) - (cu[1] - cu[2]) * (fv - cv[2])) / det;
            const T l1 = ((cu[0] - cu[2]) * (fv - cv[2]) - (fu - cu[2]) * (cv[0] - cv[2])) / det;
            const T lam[3] = {l0, l1, T(1) - l0 - l1};
#pragma unroll
            for (int v = 0; v < 3; ++v) {
                T sd, cd;
                sincos_g(dthc * (fpos - cpos[v]), &sd, &cd);
                galerkin_block<T>(lam[v], sd, cd, ffix, cfix[v], Ph[f][v]);
            }
        }
        galerkin_child<T>(a, ef, Ph, E);
    }
    galerkin_store<T>(a, p, ec, E);
}

// Unit diagonal on the fixed dofs of a BSR 2x2 operator (their rows/columns are zero).
template <typename T>
__global__ void k_fix_diag(int64_t nfix, const int64_t *__restrict__ dofs, const int32_t *__restrict__ ptr,
                           const int32_t *__restrict__ col, const int64_t *__restrict__ sptr,
                           const int32_t *__restrict__ scol, T *__restrict__ A) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfix) return;
    const int64_t d = dofs[i], node = d / 2, c = d % 2;
    const int32_t len = ptr[node + 1] - ptr[node];
    for (int32_t k = 0; k < len; ++k) {
        const int64_t q = sptr ? sptr[node / Pattern::kSlice] + (int64_t)k * Pattern::kSlice + node % Pattern::kSlice
                               : (int64_t)ptr[node] + k;
        if ((sptr ? scol[q] : col[q]) == node) { A[4 * q + 3 * c] = T(1); break; }
    }
}

}  // namespace

void launch_schur_samples(Context &c, const SchurSampleHost &h) {
    SchurSampleArgs a{};
    a.quad = make_quad();
    a.nel = c.nel; a.nsamp = h.nsamp; a.per_snap = h.per_snap; a.first_snap = h.first_snap;
    a.mode = h.mode; a.ratio_kind = h.ratio_kind; a.mean_area = c.mean_area; a.split = c.mesh_split;
    a.iperm = c.d_iperm.p; a.th = c.d_elth.p; a.r = c.d_elr.p; a.eta = c.d_eta.p;
    a.phi = c.d_phi.p; a.delta = c.d_delta.p;
    a.X = h.X; a.Tg = h.Tg; a.S = h.S; a.scale = h.scale;
    const int nb = (int)((h.nsamp + kAsmThreads - 1) / kAsmThreads);
    k_schur_samples<<<dim3(nb), dim3(kAsmThreads), 0, c.stream>>>(a);
    STK_HIP(hipGetLastError());
}

void launch_learned_scatter(Context &c, const float *Y, const double *scale, double *S64) {
    LearnedScatterArgs a{};
    a.nel = c.nel; a.split = c.mesh_split; a.perm = c.d_perm.p; a.e2n = c.d_e2n.p; a.sM = c.d_sM.p; a.Mptr = c.MN.dptr.p;
    a.Y = Y; a.scale = scale; a.S64 = S64;
    for (size_t k = 0; k + 1 < c.col_off.size(); ++k) {
        a.off = c.col_off[k];
        a.cnt = c.col_off[k + 1] - c.col_off[k];
        const int nb = (int)((a.cnt + 255) / 256);
        k_learned_scatter<<<dim3(nb), dim3(256), 0, c.stream>>>(a);
        STK_HIP(hipGetLastError());
    }
}

template <typename T>
void galerkin_impl(Context &c, size_t l, T *Ef, T *Ec, T *A) {
    MGLevel &F = c.mg[l], &C = c.mg[l + 1];
    GalArgs<T> a{};
    a.ncr_c = C.ncr; a.nct_c = C.nct; a.ncr_f = F.ncr; a.nct_f = F.nct;
    a.nel_c = C.nel; a.nel_f = F.nel;
    a.perm = C.perm.p; a.e2v = C.e2v.p; a.slot = C.slot.p; a.ptr = C.pat.dptr.p;
    a.sptr = C.pat.sell ? C.pat.dsptr.p : nullptr;
    a.Ef = Ef; a.Ec = Ec; a.A = A;
    for (size_t k = 0; k + 1 < C.col_off.size(); ++k) {
        a.off = C.col_off[k];
        a.cnt = C.col_off[k + 1] - C.col_off[k];
        if (C.split == 4) k_galerkin4<T><<<dim3((unsigned)((a.cnt + 63) / 64)), dim3(64), 0, c.stream>>>(a);
        else k_galerkin<T><<<dim3((unsigned)((4 * a.cnt + 255) / 256)), dim3(256), 0, c.stream>>>(a);
        STK_HIP(hipGetLastError());
    }
}

template <typename T>
void fix_diag(Context &c, const Pattern &p, int64_t nfix, const int64_t *dofs, T *A) {
    if (nfix == 0) return;
    const int nb = (int)((nfix + 255) / 256);
    k_fix_diag<T><<<dim3(nb), dim3(256), 0, c.stream>>>(nfix, dofs, p.dptr.p, p.dcol.p,
                                                        p.sell ? p.dsptr.p : nullptr, p.sell ? p.dscol.p : nullptr, A);
    STK_HIP(hipGetLastError());
}

template void galerkin_impl<double>(Context &, size_t, double *, double *, double *);
template void galerkin_impl<float>(Context &, size_t, float *, float *, float *);
template void fix_diag<double>(Context &, const Pattern &, int64_t, const int64_t *, double *);
template void fix_diag<float>(Context &, const Pattern &, int64_t, const int64_t *, float *);

}  // namespace stk
