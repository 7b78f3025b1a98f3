// Stress post-processor and deformation gradient on the solved velocity.
//
// stress! (src/Stress/StressCalculation.jl:23-127 isotropic, :129-242 anisotropic,
// driver :244-249) evaluates, at every point of the stress rule, the strain rate from the
// Cartesian velocity, its second invariant, the deviatoric stress and its invariant.
// deformation_gradient! (:251-349) advances the 2x2 deformation gradient F at every point
// by one RK4 step of F' = (grad U)^T F.  Both map the P2 shape-function derivatives from the
// reference triangle to (theta, r) and then to Cartesian (x, z) with the "double Jacobian":
// d/dx = sin(theta) d/dr + cos(theta)/r d/dtheta, d/dz = cos(theta) d/dr - sin(theta)/r d/dtheta.
//
// One thread per element in natural element order: the eta / tensor inputs and all outputs
// are nel x nip column-major (element fastest), so every one of them is read or written
// coalesced; the 12 velocity values of an element are gathered (they are shared with the
// neighbours and mostly hit in cache).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include <map>
#include <mutex>

#include "device_ops.h"
#include "stress.h"

namespace stk {

namespace {

// Points of the stress rule, passed by value in every launch: the rule belongs to the
// context (stk_set_stress_fields), not to the process or the device.
struct StressRule {
    int nip;
    double xi[7], eta[7];
};

// Points of ip_triangle6 / ip_triangle7 (src/Algebra/Quadrature.jl:80-138), in their order.
StressRule make_rule(int nip) {
    StressRule q{};
    q.nip = nip;
    if (nip == 7) {
        const double g1 = (6.0 - std::sqrt(15.0)) / 21.0, g2 = (6.0 + std::sqrt(15.0)) / 21.0;
        const double x[7] = {1.0 / 3.0, 1.0 - 2.0 * g1, g1, g1, 1.0 - 2.0 * g2, g2, g2};
        const double y[7] = {1.0 / 3.0, g1, 1.0 - 2.0 * g1, g1, g2, 1.0 - 2.0 * g2, g2};
        for (int i = 0; i < 7; ++i) { q.xi[i] = x[i]; q.eta[i] = y[i]; }
    } else {
        const double s = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
        const double g1 = (8.0 - std::sqrt(10.0) + s) / 18.0, g2 = (8.0 - std::sqrt(10.0) - s) / 18.0;
        const double x[6] = {1.0 - 2.0 * g1, g1, g1, 1.0 - 2.0 * g2, g2, g2};
        const double y[6] = {g1, 1.0 - 2.0 * g1, g1, g2, 1.0 - 2.0 * g2, g2};
        for (int i = 0; i < 6; ++i) { q.xi[i] = x[i]; q.eta[i] = y[i]; }
    }
    return q;
}

// Cartesian derivatives of the six P2 shape functions at reference point (xi, eta) of an
// element with vertex coordinates th[3], r[3] (StressCalculation.jl:145-205).
__device__ __forceinline__ void cartesian_gradients(const double th[3], const double r[3], double xi, double et,
                                                    double dx[6], double dz[6]) {
    const double R21 = r[1] - r[0], R31 = r[2] - r[0];
    const double Th31 = th[2] - th[0], Th21 = th[1] - th[0];
    const double invdet = 1.0 / (Th21 * R31 - R21 * Th31);
    const double t = 1.0 - xi - et;
    // P2 derivatives w.r.t. (xi, eta); node order: vertices 1,2,3, midpoints 12, 23, 31
    const double dr[6] = {-(4.0 * t - 1.0), 4.0 * xi - 1.0, 0.0, 4.0 * (t - xi), 4.0 * et, -4.0 * et};
    const double ds[6] = {-(4.0 * t - 1.0), 0.0, 4.0 * et - 1.0, -4.0 * xi, 4.0 * xi, 4.0 * (t - et)};
    const double thq = th[0] * t + th[1] * xi + th[2] * et;
    const double rq = r[0] * t + r[1] * xi + r[2] * et;
    double sn, cs;
    sincos(thq, &sn, &cs);
    const double cr = cs / rq, sr = sn / rq;
    const double j00 = R31 * cr - Th31 * sn, j01 = -R21 * cr + Th21 * sn;
    const double j10 = -R31 * sr - Th31 * cs, j11 = R21 * sr + Th21 * cs;
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        dx[b] = (j00 * dr[b] + j01 * ds[b]) * invdet;
        dz[b] = (j10 * dr[b] + j11 * ds[b]) * invdet;
    }
}

struct StressArgs {
    int64_t nel;
    const int32_t *e2n;             // 6 x nel, natural element order (SoA)
    const double *th, *r;           // 3 x nel (SoA)
    const double *U;                // Cartesian velocity (x, z) per node
    const double *eta;              // nel x nip
    const double *C;                // 6 components x (nel x nip), or null (isotropic)
    double *out;                    // 8 x (nel x nip): exx, ezz, exz, eII, txx, tzz, txz, tII
    StressRule rule;
};

__global__ __launch_bounds__(256) void k_stress(StressArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nel) return;
    const int64_t nel = a.nel;
    double th[3], r[3], ux[6], uz[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) { th[j] = a.th[j * nel + e]; r[j] = a.r[j * nel + e]; }
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const int64_t n = a.e2n[b * nel + e];
        ux[b] = a.U[2 * n];
        uz[b] = a.U[2 * n + 1];
    }
    const int nip = a.rule.nip;
    const int64_t plane = nel * nip;
    for (int q = 0; q < nip; ++q) {
        double dx[6], dz[6];
        cartesian_gradients(th, r, a.rule.xi[q], a.rule.eta[q], dx, dz);
        double exx = 0, ezz = 0, duxdz = 0, duzdx = 0;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            exx += dx[b] * ux[b];
            ezz += dz[b] * uz[b];
            duxdz += dz[b] * ux[b];
            duzdx += dx[b] * uz[b];
        }
        const double exz = 0.5 * (duxdz + duzdx);
        const int64_t k = q * nel + e;
        const double eta = a.eta[k];
        double txx, tzz, txz;
        if (a.C) {
            // anisotropic law of :223-235, with the reference's naming of the tensor components
            const double h11 = a.C[0 * plane + k] * eta, h22 = a.C[1 * plane + k] * eta;
            const double h33 = a.C[2 * plane + k] * eta, h12 = a.C[3 * plane + k] * eta;
            const double h13 = a.C[4 * plane + k] * eta, h23 = a.C[5 * plane + k] * eta;
            const double u = -exx / 3.0 + 2.0 * ezz / 3.0, v = 2.0 * exx / 3.0 - ezz / 3.0;
            txx = h12 * u + h11 * v + h13 * exz;
            tzz = h22 * u + h12 * v + h23 * 2.0 * exz;
            txz = h23 * u + h13 * v + h33 * exz;
        } else {
            // isotropic law of :117-120
            txx = eta * (4.0 / 3.0 * exx - 2.0 / 3.0 * ezz);
            tzz = eta * (-2.0 / 3.0 * exx + 4.0 / 3.0 * ezz);
            txz = eta * 2.0 * exz;
        }
        a.out[0 * plane + k] = exx;
        a.out[1 * plane + k] = ezz;
        a.out[2 * plane + k] = exz;
        a.out[3 * plane + k] = sqrt(0.5 * (exx * exx + ezz * ezz) + exz * exz);
        a.out[4 * plane + k] = txx;
        a.out[5 * plane + k] = tzz;
        a.out[6 * plane + k] = txz;
        a.out[7 * plane + k] = sqrt(0.5 * (txx * txx + tzz * tzz) + txz * txz);
    }
}

struct DefGradArgs {
    int64_t nel;
    const int32_t *e2n;
    const double *th, *r;
    const double *U;                // velocity as the caller passes it (x, z per node)
    double dt;
    double *F;                      // 4 x (nel x nip): column-major 2x2 per point
    StressRule rule;
};

__global__ __launch_bounds__(256) void k_deformation_gradient(DefGradArgs a) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nel) return;
    const int64_t nel = a.nel;
    double th[3], r[3], ux[6], uz[6];
#pragma unroll
    for (int j = 0; j < 3; ++j) { th[j] = a.th[j * nel + e]; r[j] = a.r[j * nel + e]; }
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const int64_t n = a.e2n[b * nel + e];
        ux[b] = a.U[2 * n];
        uz[b] = a.U[2 * n + 1];
    }
    const int nip = a.rule.nip;
    for (int q = 0; q < nip; ++q) {
        double dx[6], dz[6];
        cartesian_gradients(th, r, a.rule.xi[q], a.rule.eta[q], dx, dz);
        double uxx = 0, uzz = 0, uxz = 0, uzx = 0;
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            uxx += dx[b] * ux[b];
            uzz += dz[b] * uz[b];
            uxz += dz[b] * ux[b];
            uzx += dx[b] * uz[b];
        }
        // L = dt (grad U)^T = dt [[dUx/dx, dUx/dz], [dUz/dx, dUz/dz]]  (:304-326)
        const double L00 = a.dt * uxx, L01 = a.dt * uxz, L10 = a.dt * uzx, L11 = a.dt * uzz;
        double *Fp = a.F + 4 * (q * nel + e);
        const double f00 = Fp[0], f10 = Fp[1], f01 = Fp[2], f11 = Fp[3];
        // classic RK4 on F' = L F (:328-338)
        auto mul = [&](double a00, double a10, double a01, double a11, double &b00, double &b10, double &b01,
                       double &b11) {
            b00 = L00 * a00 + L01 * a10;
            b10 = L10 * a00 + L11 * a10;
            b01 = L00 * a01 + L01 * a11;
            b11 = L10 * a01 + L11 * a11;
        };
        double k1[4], k2[4], k3[4], k4[4];
        mul(f00, f10, f01, f11, k1[0], k1[1], k1[2], k1[3]);
        mul(0.5 * k1[0] + f00, 0.5 * k1[1] + f10, 0.5 * k1[2] + f01, 0.5 * k1[3] + f11, k2[0], k2[1], k2[2], k2[3]);
        mul(0.5 * k2[0] + f00, 0.5 * k2[1] + f10, 0.5 * k2[2] + f01, 0.5 * k2[3] + f11, k3[0], k3[1], k3[2], k3[3]);
        mul(k3[0] + f00, k3[1] + f10, k3[2] + f01, k3[3] + f11, k4[0], k4[1], k4[2], k4[3]);
        const double f0[4] = {f00, f10, f01, f11};
#pragma unroll
        for (int i = 0; i < 4; ++i) Fp[i] = f0[i] + (k1[i] + 2.0 * (k2[i] + k3[i]) + k4[i]) * (1.0 / 6.0);
    }
}

// volume_integral_cartesian (src/Stress/FSE.jl:466-514) of two nodal fields at once: per
// element detJ (straight Cartesian triangle of the three vertices) times sum over the 6-point
// rule of w_q * sum_a N_a(q) V_a, folded into the integrated weights cw[a] = sum_q w_q N_a(q).
struct VolIntArgs {
    int64_t nel;
    const int32_t *e2n;             // 6 x nel natural order
    const double *cs, *nr;          // node cos/sin (2 per node), node radius
    const double *U;                // velocity (2 per node): S = Ux^2 + Uz^2 at the nodes
    const double *T;                // temperature per node
    double cw[6];
};

struct VolInt {
    VolIntArgs a;
    __device__ Dots<2> operator()(int64_t e) const {
        const int64_t nel = a.nel;
        double x[3], z[3];
        for (int j = 0; j < 3; ++j) {
            const int64_t n = a.e2n[j * nel + e];
            x[j] = a.nr[n] * a.cs[2 * n + 1];   // x = r sin(theta)
            z[j] = a.nr[n] * a.cs[2 * n];       // z = r cos(theta)
        }
        const double det = (x[1] - x[0]) * (z[2] - z[0]) - (z[1] - z[0]) * (x[2] - x[0]);
        double s = 0, t = 0;
        for (int b = 0; b < 6; ++b) {
            const int64_t n = a.e2n[b * nel + e];
            const double ux = a.U[2 * n], uz = a.U[2 * n + 1];
            s += a.cw[b] * (ux * ux + uz * uz);
            t += a.cw[b] * a.T[n];
        }
        return Dots<2>{{det * s, det * t}};
    }
};

struct ListSum {
    const int32_t *idx;
    const double *T;
    __device__ double operator()(int64_t i) const { return T[idx[i]]; }
};

}  // namespace

// Natural-order copies of the element connectivity and vertex coordinates (once per mesh).
void stress_prepare(Context &c) {
    if (c.d_e2n_nat) return;
    const int64_t nel = c.nel;
    std::vector<int32_t> e2n(6 * nel);
    std::vector<double> th(3 * nel), r(3 * nel);
    for (int64_t e = 0; e < nel; ++e) {
        for (int a = 0; a < 6; ++a) e2n[a * nel + e] = c.e2n[e * 6 + a];
        for (int j = 0; j < 3; ++j) {
            th[j * nel + e] = c.el_th[e * 3 + j];
            r[j * nel + e] = c.el_r[e * 3 + j];
        }
    }
    c.d_e2n_nat.upload(e2n, c.stream);
    c.d_th_nat.upload(th, c.stream);
    c.d_r_nat.upload(r, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
}

void stress_set_rule(Context &c, int nip) {
    if (nip != 6 && nip != 7) throw Error(STK_E_SHAPE, "the stress rule has 6 or 7 points");
    c.s_nip = nip;
}

void launch_stress(Context &c, const double *U, double *out) {
    StressArgs a{};
    a.nel = c.nel;
    a.e2n = c.d_e2n_nat.p; a.th = c.d_th_nat.p; a.r = c.d_r_nat.p;
    a.U = U;
    a.eta = c.d_seta.p;
    a.C = c.s_aniso ? c.d_sC.p : nullptr;
    a.out = out;
    a.rule = make_rule(c.s_nip);
    k_stress<<<dim3((unsigned)((c.nel + 255) / 256)), dim3(256), 0, c.stream>>>(a);
    STK_HIP(hipGetLastError());
}

// write_stats (src/IO/IOmanager.jl:277-322) without the file: Urms, <T>, Nusselt top/bottom.
// Once per mesh: the natural-order element data and ScratchNusselt's node lists (IOmanager.jl:338-354),
// the nodes one radius level inside each wall.
void statistics_prepare(Context &c) {
    stress_prepare(c);
    if (c.d_itop) return;
    std::vector<double> ru(c.node_r);
    std::sort(ru.begin(), ru.end());
    ru.erase(std::unique(ru.begin(), ru.end()), ru.end());
    if (ru.size() < 3) throw Error(STK_E_SHAPE, "statistics need at least three radius levels");
    std::vector<int32_t> top, bot;
    for (int64_t n = 0; n < c.nnod; ++n) {
        if (c.node_r[n] == ru[ru.size() - 2]) top.push_back((int32_t)n);
        if (c.node_r[n] == ru[1]) bot.push_back((int32_t)n);
    }
    if (top.size() < 2) throw Error(STK_E_SHAPE, "statistics need two nodes below the outer wall");
    c.nu_dr = ru[1] - ru[0];
    c.nu_dth = c.node_th[top[0]] - c.node_th[top[1]];
    c.nu_rin = ru.front();
    c.nu_rout = ru.back();
    c.d_itop.upload(top, c.stream);
    c.d_ibot.upload(bot, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
}

namespace {

// The constants that turn the four sums into write_stats' numbers.
struct StatsCombine {
    double area;   // pi (r_out^2 - r_in^2) from the extreme node radii
    double A;      // log(f) / (2 pi 2.22 (1 - f)), f = 1.22 / 2.22: radii fixed in ScratchNusselt (:341)
    double dth, dr, rin, rout, nbot;
};

StatsCombine stats_combine(const Context &c) {
    const double f = 1.22 / 2.22;
    return StatsCombine{M_PI * (c.nu_rout * c.nu_rout - c.nu_rin * c.nu_rin),
                        std::log(f) / (2.0 * M_PI * 2.22 * (1.0 - f)), c.nu_dth, c.nu_dr, c.nu_rin, c.nu_rout,
                        (double)c.d_ibot.n};
}

__host__ __device__ inline void combine_stats(const StatsCombine &k, const double h[4], double out[4]) {
    out[0] = sqrt(h[0] / k.area);
    out[1] = h[1] / k.area;
    out[2] = k.A * k.dth * (-h[2] / k.dr) * k.rout;
    out[3] = k.A * k.dth * ((h[3] - k.nbot) / k.dr) * k.rin;
}

__global__ void k_stats_combine(StatsCombine k, const double *__restrict__ h, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double hv[4] = {h[0], h[1], h[2], h[3]}, o[4];
        combine_stats(k, hv, o);
        for (int i = 0; i < 4; ++i) out[i] = o[i];
    }
}

// The four sums of write_stats into sums[0..3]: int |U|^2 and int T over the straight triangles
// (volume_integral_cartesian, src/Stress/FSE.jl:466-514: 6-point rule, P2 interpolation), and the sums
// of T over the two Nusselt node lists.  part holds 2 * kReduceBlocks partials.
void statistics_launch(Context &c, const double *U, const double *T, double *part, double *sums) {
    // integrated P2 weights of the 6-point rule (ip_triangle6, Quadrature.jl:80-107)
    VolIntArgs a{};
    {
        const double s = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
        const double g1 = (8.0 - std::sqrt(10.0) + s) / 18.0, g2 = (8.0 - std::sqrt(10.0) - s) / 18.0;
        const double q = std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0));
        const double w1 = 0.5 * (620.0 + q) / 3720.0, w2 = 0.5 * (620.0 - q) / 3720.0;
        const double xs[6] = {1.0 - 2.0 * g1, g1, g1, 1.0 - 2.0 * g2, g2, g2};
        const double ys[6] = {g1, 1.0 - 2.0 * g1, g1, g2, 1.0 - 2.0 * g2, g2};
        const double ws[6] = {w1, w1, w1, w2, w2, w2};
        for (int b = 0; b < 6; ++b) a.cw[b] = 0.0;
        for (int k = 0; k < 6; ++k) {
            const double r = xs[k], t = ys[k], u = 1.0 - r - t;
            const double N[6] = {u * (2 * u - 1), r * (2 * r - 1), t * (2 * t - 1), 4 * r * u, 4 * r * t, 4 * t * u};
            for (int b = 0; b < 6; ++b) a.cw[b] += ws[k] * N[b];
        }
    }
    a.nel = c.nel;
    a.e2n = c.d_e2n_nat.p;
    a.cs = c.d_nodecs.p;
    a.nr = c.d_noder.p;
    a.U = U;
    a.T = T;
    auto blocks = [](int64_t n) { return (int)std::min<int64_t>(kReduceBlocks, std::max<int64_t>(1, (n + 255) / 256)); };
    const int nb = blocks(c.nel);
    k_reduce_map<256, 2, VolInt><<<dim3(nb), dim3(256), 0, c.stream>>>(c.nel, VolInt{a}, part);
    k_reduce_final<256, 2><<<dim3(1), dim3(256), 0, c.stream>>>(nb, part, sums + 0);
    ListSum lt{c.d_itop.p, T}, lb{c.d_ibot.p, T};
    const int64_t ntop = (int64_t)c.d_itop.n, nbot = (int64_t)c.d_ibot.n;
    const int nbt = blocks(ntop), nbb = blocks(nbot);
    k_reduce_map<256, 1, ListSum><<<dim3(nbt), dim3(256), 0, c.stream>>>(ntop, lt, part);
    k_reduce_final<256, 1><<<dim3(1), dim3(256), 0, c.stream>>>(nbt, part, sums + 2);
    k_reduce_map<256, 1, ListSum><<<dim3(nbb), dim3(256), 0, c.stream>>>(nbot, lb, part);
    k_reduce_final<256, 1><<<dim3(1), dim3(256), 0, c.stream>>>(nbb, part, sums + 3);
    STK_HIP(hipGetLastError());
}

// Scratch of stats_noalloc: the reduction partials (at most 2 per block and kReduceBlocks blocks:
// every reduction runs a grid-stride loop over at most that many blocks, whatever the mesh size)
// and the four sums.  One copy per device, held until the process ends and shared by every call on
// that device; nothing in it has to start at zero (each launch writes what the next one reads).
constexpr int64_t kStatsScratch = 2 * kReduceBlocks + 8;
static __device__ __align__(256) double scratch[kStatsScratch];

}  // namespace

void statistics(Context &c, const double *U, const double *T, double out[4]) {
    statistics_prepare(c);
    double *part = c.red64.p, *sc = c.d_sc.p;
    if ((int64_t)c.red64.n < 2 * (int64_t)kReduceBlocks) throw Error(STK_E_STATE, "reduction buffer too small");
    statistics_launch(c, U, T, part, sc);
    double h[4];
    STK_HIP(hipMemcpyAsync(h, sc, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    STK_HIP(hipStreamSynchronize(c.stream));
    combine_stats(stats_combine(c), h, out);
}

// Device address of this device's scratch array; the current device must be the context's.
// hipGetSymbolAddress is called once per device (each device holds its own copy at its own address).
double *stats_scratch_address(int device) {
    static std::mutex mu;
    static std::map<int, double *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    double *p = nullptr;
    STK_HIP(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(STK_E_HIP, "hipGetSymbolAddress returned a null scratch address");
    addr[device] = p;
    return p;
}

void statistics_noalloc(Context &c, const double *U, const double *T, double *out) {
    if (!c.d_itop) throw Error(STK_E_STATE, "statistics data not prepared for this mesh" +
                                                (c.stats_note.empty() ? std::string() : ": " + c.stats_note));
    if (!c.stats_scratch) throw Error(STK_E_STATE, "no scratch address for this device");
    // every reduction uses at most kReduceBlocks partials of two columns, then four sums
    const int64_t need = 2 * (int64_t)kReduceBlocks + 4;
    if (need > kStatsScratch) throw Error(STK_E_SHAPE, "statistics need more scratch than the fixed array holds");
    double *part = c.stats_scratch, *sums = c.stats_scratch + 2 * kReduceBlocks;
    statistics_launch(c, U, T, part, sums);
    k_stats_combine<<<dim3(1), dim3(64), 0, c.stream>>>(stats_combine(c), sums, out);
    STK_HIP(hipGetLastError());
}

void launch_deformation_gradient(Context &c, const double *U, double dt, double *F) {
    DefGradArgs a{};
    a.nel = c.nel;
    a.e2n = c.d_e2n_nat.p; a.th = c.d_th_nat.p; a.r = c.d_r_nat.p;
    a.U = U;
    a.dt = dt;
    a.F = F;
    a.rule = make_rule(c.s_nip);
    k_deformation_gradient<<<dim3((unsigned)((c.nel + 255) / 256)), dim3(256), 0, c.stream>>>(a);
    STK_HIP(hipGetLastError());
}

}  // namespace stk
