// Thermal diffusion operators on the P2 mesh (FP64).
//
// assemble_element! (src/ThermalDiffusion/AssemblerThermal.jl:117-209) with the 7-point
// rule of diffusion_immutables(gr, 2, 3, 6, 7) (MAIN.jl:167-169): per element
//   K_e = sum_q K (dN/dx^T dN/dx + dN/dz^T dN/dz) w_q,   M_e = sum_q rho_q Cp N^T N w_q,
//   f_e = sum_q dQdT dt N w_q  (only when dQdT != 0),
// w_q = r_q detJ w_ip, rho_q the P1 interpolation of rho at the three vertices, Cartesian
// derivatives by the same double Jacobian as the Stokes assembly.  One thread per element,
// colour by colour, read-modify-write into the sliced layout of the P2 node pattern (the node
// pattern of K, one value per node pair).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_ops.h"
#include "thermal.h"

namespace stk {

namespace {

struct Rule7 {
    double xi[7], eta[7], w[7];
};

// ip_triangle7 (src/Algebra/Quadrature.jl:109-138), weights including the 1/2.
Rule7 make_rule7() {
    Rule7 q{};
    const double g1 = (6.0 - std::sqrt(15.0)) / 21.0, g2 = (6.0 + std::sqrt(15.0)) / 21.0;
    const double w1 = (155.0 - std::sqrt(15.0)) / 1200.0, w2 = (155.0 + std::sqrt(15.0)) / 1200.0;
    const double x[7] = {1.0 / 3.0, 1.0 - 2.0 * g1, g1, g1, 1.0 - 2.0 * g2, g2, g2};
    const double y[7] = {1.0 / 3.0, g1, 1.0 - 2.0 * g1, g1, g2, 1.0 - 2.0 * g2, g2};
    const double w[7] = {0.225, w1, w1, w1, w2, w2, w2};
    for (int i = 0; i < 7; ++i) {
        q.xi[i] = x[i];
        q.eta[i] = y[i];
        q.w[i] = 0.5 * w[i];
    }
    return q;
}

struct ThermalArgs {
    int64_t nel, off, cnt;
    const int32_t *perm, *e2n;          // colour-permuted SoA (6 x nel)
    const double *th, *r;               // 3 x nel, colour-permuted SoA
    const double *rho;                  // per node (the vertex values are used)
    double K, Cp, dQdT, dt;
    const uint8_t *sK;                  // node-pair slot offsets (36 x nel by position)
    const int64_t *sptr;                // sliced layout of the node pattern
    double *KT, *MT, *F;
    Rule7 rule;                         // the 7-point rule, passed by value
};

__global__ __launch_bounds__(64) void k_thermal_assemble(ThermalArgs a) {
    const int64_t j = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (j >= a.cnt) return;
    const int64_t p = a.off + j, nel = a.nel;
    int32_t n[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) n[b] = a.e2n[b * nel + p];
    double th[3], r[3], rv[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        th[v] = a.th[v * nel + p];
        r[v] = a.r[v * nel + p];
        rv[v] = a.rho[n[v]];
    }
    const double R21 = r[1] - r[0], R31 = r[2] - r[0];
    const double Th31 = th[2] - th[0], Th21 = th[1] - th[0];
    const double detJ = Th21 * R31 - R21 * Th31;
    double Ke[21], Me[21], fe[6];
#pragma unroll
    for (int k = 0; k < 21; ++k) { Ke[k] = 0.0; Me[k] = 0.0; }
#pragma unroll
    for (int b = 0; b < 6; ++b) fe[b] = 0.0;
    for (int q = 0; q < 7; ++q) {
        const double xi = a.rule.xi[q], et = a.rule.eta[q], t = 1.0 - xi - et;
        const double N[6] = {t * (2 * t - 1), xi * (2 * xi - 1), et * (2 * et - 1), 4 * xi * t, 4 * xi * et, 4 * et * t};
        const double dr[6] = {-(4 * t - 1), 4 * xi - 1, 0.0, 4 * (t - xi), 4 * et, -4 * et};
        const double ds[6] = {-(4 * t - 1), 0.0, 4 * et - 1, -4 * xi, 4 * xi, 4 * (t - et)};
        const double thq = th[0] * t + th[1] * xi + th[2] * et;
        const double rq = r[0] * t + r[1] * xi + r[2] * et;
        const double rhoq = rv[0] * t + rv[1] * xi + rv[2] * et;
        double sn, cs;
        sincos(thq, &sn, &cs);
        const double cr = cs / rq, sr = sn / rq;
        const double j00 = R31 * cr - Th31 * sn, j01 = -R21 * cr + Th21 * sn;
        const double j10 = -R31 * sr - Th31 * cs, j11 = R21 * sr + Th21 * cs;
        double dx[6], dz[6];
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            dx[b] = (j00 * dr[b] + j01 * ds[b]) / detJ;
            dz[b] = (j10 * dr[b] + j11 * ds[b]) / detJ;
        }
        const double om = rq * detJ * a.rule.w[q];
        const double kw = a.K * om, mw = om * rhoq * a.Cp;
#pragma unroll
        for (int u = 0; u < 6; ++u)
#pragma unroll
            for (int v = 0; v <= u; ++v) {
                Ke[u * (u + 1) / 2 + v] += (dx[u] * dx[v] + dz[u] * dz[v]) * kw;
                Me[u * (u + 1) / 2 + v] += N[u] * N[v] * mw;
            }
        if (a.dQdT != 0.0)
#pragma unroll
            for (int u = 0; u < 6; ++u) fe[u] += N[u] * a.dQdT * (a.dt * om);
    }
    for (int u = 0; u < 6; ++u) {
        const int64_t nu = n[u];
        const int64_t rowbase = a.sptr[nu / Pattern::kSlice] + nu % Pattern::kSlice;
        for (int v = 0; v < 6; ++v) {
            const int k = u >= v ? u * (u + 1) / 2 + v : v * (v + 1) / 2 + u;
            const int64_t q = rowbase + (int64_t)a.sK[(u * 6 + v) * nel + p] * Pattern::kSlice;
            a.KT[q] += Ke[k];
            a.MT[q] += Me[k];
        }
        if (a.dQdT != 0.0) a.F[nu] += fe[u];
    }
}

}  // namespace

void thermal_assemble(Context &c, const double *rho, double K, double Cp, double dQdT, double dt, double *KT,
                      double *MT, double *F) {
    if (!c.KN.sell) throw Error(STK_E_STATE, "the thermal operators need the sliced node pattern");
    ThermalArgs a{};
    a.rule = make_rule7();
    a.nel = c.nel;
    a.perm = c.d_perm.p; a.e2n = c.d_e2n.p; a.th = c.d_elth.p; a.r = c.d_elr.p;
    a.rho = rho;
    a.K = K; a.Cp = Cp; a.dQdT = dQdT; a.dt = dt;
    a.sK = c.d_sK.p;
    a.sptr = c.KN.dsptr.p;
    a.KT = KT; a.MT = MT; a.F = F;
    for (size_t k = 0; k + 1 < c.col_off.size(); ++k) {
        a.off = c.col_off[k];
        a.cnt = c.col_off[k + 1] - c.col_off[k];
        k_thermal_assemble<<<dim3((unsigned)((a.cnt + 63) / 64)), dim3(64), 0, c.stream>>>(a);
        STK_HIP(hipGetLastError());
    }
}

}  // namespace stk
