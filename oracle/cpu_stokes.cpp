// CPU restatement of the reference Stokes path in C++/OpenMP, FP64 -- TEST INFRASTRUCTURE ONLY.
//
// It is the CPU baseline that bench.py times beside the GPU library (and nothing in the library
// calls it).  It restates, from the reference Julia sources:
//   * the element assembly of assembly_stokes_cylindric (src/Stokes/StokesSolver.jl:247-404: 6-point
//     rule :262, _fill_R_J! :522-541, _derivative_weights! :641-686, the anisotropic _fill_Kblk! of
//     src/src/Stokes/StokesSolver.jl:744-808, _fill_Gblk! / _fill_Mblk! / _fill_Fbblk! :812-845);
//   * _prepare_matrices (:192-201: the rotation TT^T K TT into (u_theta, u_r) node frames, here
//     element by element) and _apply_bcs with free slip (:186-190, src/Setup/boundary_conditions.jl);
//   * StokesPcCG (src/IterativeSolvers.jl:94-168): Schur-complement PCG, lumped 1/eta-mass
//     preconditioner (:189-240), rm0space! (:172-186), Polak-Ribiere beta, itmin 10; the inner solve
//     lu(K[ifree, ifree]) (src/PardisoSolvers.jl:247-256) replaced, per the CPU baseline plan, by
//     point-Jacobi CG on K[ifree, ifree] projected off the nodal rigid rotation (the near-null mode
//     free slip leaves), to a relative residual inner_rtol.
// Assembly, SpMV and vector operations run on all OpenMP threads.  The elements are coloured so
// that the accumulation into the global CSR matrices is race free.
//
// Usage: cpu_stokes DIR BUDGET_SECONDS [INNER_RTOL]
//   DIR holds meta.txt (nel nnod nvert nfixed g) and the raw arrays of tests/c_abi/abi_driver.c's
//   input layout (e2n.bin, el_theta.bin, el_r.bin, node_theta.bin, node_r.bin, fixed.bin, rho.bin,
//   eta.bin, C0.bin .. C5.bin).  The assembly always completes; the solve stops once BUDGET_SECONDS
//   of solve time have passed (or at convergence).  One JSON line on stdout reports the timings,
//   the iterations done and the thread count; the solution (node frames, before the net-rotation
//   removal) is written to DIR/cpu_U.bin and DIR/cpu_P.bin.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t0) { return std::chrono::duration<double>(Clock::now() - t0).count(); }

template <typename T>
std::vector<T> load(const std::string &dir, const char *name, size_t count) {
    std::vector<T> v(count);
    const std::string path = dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(T), count, f) != count) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

// 6-point rule (src/Algebra/Quadrature.jl:80-107), weights with the 1/2 of the reference triangle.
struct Rule {
    double xi[6], eta[6], w[6];
    Rule() {
        const double s = std::sqrt(38.0 - 44.0 * std::sqrt(2.0 / 5.0));
        const double g1 = (8.0 - std::sqrt(10.0) + s) / 18.0, g2 = (8.0 - std::sqrt(10.0) - s) / 18.0;
        const double q = std::sqrt(213125.0 - 53320.0 * std::sqrt(10.0));
        const double w1 = (620.0 + q) / 3720.0, w2 = (620.0 - q) / 3720.0;
        const double x[6] = {1 - 2 * g1, g1, g1, 1 - 2 * g2, g2, g2};
        const double y[6] = {g1, 1 - 2 * g1, g1, g2, 1 - 2 * g2, g2};
        for (int i = 0; i < 6; ++i) { xi[i] = x[i]; eta[i] = y[i]; w[i] = 0.5 * (i < 3 ? w1 : w2); }
    }
};

// CSR matrix over dofs.
struct Csr {
    int64_t n = 0;
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<double> val;
    // y = A x
    void apply(const double *x, double *y) const {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
            y[i] = s;
        }
    }
    int64_t find(int64_t i, int32_t j) const {
        const int32_t *b = col.data() + ptr[i], *e = col.data() + ptr[i + 1];
        const int32_t *p = std::lower_bound(b, e, j);
        return (p != e && *p == j) ? (int64_t)(p - col.data()) : -1;
    }
};

// Dof pattern from the element connectivity: row dof r lists the sorted dofs of every element node
// pair; nr_el / nc_el local row / column nodes, dofs per node dr / dc.
Csr pattern(int64_t nrows_nodes, int64_t nel, const std::vector<int32_t> &e2r, int nr_el, int dr,
            const std::vector<int32_t> &e2c, int nc_el, int dc) {
    std::vector<std::vector<int32_t>> adj(nrows_nodes);
    std::vector<int64_t> cnt(nrows_nodes + 1, 0);
    std::vector<std::vector<int32_t>> inc(nrows_nodes);
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < nr_el; ++a) inc[e2r[e * nr_el + a]].push_back((int32_t)e);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t r = 0; r < nrows_nodes; ++r) {
        std::vector<int32_t> &v = adj[r];
        for (int32_t e : inc[r])
            for (int b = 0; b < nc_el; ++b) v.push_back(e2c[(int64_t)e * nc_el + b]);
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    Csr A;
    A.n = nrows_nodes * dr;
    A.ptr.assign(A.n + 1, 0);
    for (int64_t r = 0; r < nrows_nodes; ++r)
        for (int i = 0; i < dr; ++i) A.ptr[r * dr + i + 1] = (int64_t)adj[r].size() * dc;
    for (int64_t i = 0; i < A.n; ++i) A.ptr[i + 1] += A.ptr[i];
    A.col.resize(A.ptr[A.n]);
    A.val.assign(A.ptr[A.n], 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nrows_nodes; ++r)
        for (int i = 0; i < dr; ++i) {
            int64_t k = A.ptr[r * dr + i];
            for (int32_t c : adj[r])
                for (int j = 0; j < dc; ++j) A.col[k++] = c * dc + j;
        }
    return A;
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    const int64_t n = (int64_t)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s DIR BUDGET_SECONDS [INNER_RTOL]\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const double budget = std::atof(argv[2]);
    const double inner_rtol = argc > 3 ? std::atof(argv[3]) : 1e-10;
    long long nel_l, nnod_l, nvert_l, nfixed_l;
    double g;
    {
        FILE *m = std::fopen((dir + "/meta.txt").c_str(), "r");
        if (!m || std::fscanf(m, "%lld %lld %lld %lld %lf", &nel_l, &nnod_l, &nvert_l, &nfixed_l, &g) != 5) {
            std::fprintf(stderr, "bad meta.txt\n");
            return 2;
        }
        std::fclose(m);
    }
    const int64_t nel = nel_l, nnod = nnod_l, nvert = nvert_l, nfixed = nfixed_l, nU = 2 * nnod, nP = nvert;
    // inputs in the Julia layouts: e2n 6 x nel column-major (element-major rows of 6), el_* 3 x nel,
    // eta and the tensor components nel x 6 column-major (column q at q * nel)
    const auto e2n64 = load<int64_t>(dir, "e2n.bin", 6 * (size_t)nel);
    const auto elth = load<double>(dir, "el_theta.bin", 3 * (size_t)nel);
    const auto elr = load<double>(dir, "el_r.bin", 3 * (size_t)nel);
    const auto nth = load<double>(dir, "node_theta.bin", (size_t)nnod);
    const auto nr = load<double>(dir, "node_r.bin", (size_t)nnod);
    const auto fixed1 = load<int64_t>(dir, "fixed.bin", (size_t)nfixed);
    const auto rho = load<double>(dir, "rho.bin", (size_t)nnod);
    const auto eta = load<double>(dir, "eta.bin", 6 * (size_t)nel);
    std::vector<std::vector<double>> C(6);
    for (int k = 0; k < 6; ++k) C[k] = load<double>(dir, ("C" + std::to_string(k) + ".bin").c_str(), 6 * (size_t)nel);
    std::vector<int32_t> e2n(6 * nel), e2v(3 * nel);
    for (int64_t e = 0; e < nel; ++e)
        for (int a = 0; a < 6; ++a) {
            e2n[e * 6 + a] = (int32_t)(e2n64[e * 6 + a] - 1);
            if (a < 3) e2v[e * 3 + a] = e2n[e * 6 + a];
        }
    std::vector<uint8_t> fx(nU, 0);
    for (int64_t k = 0; k < nfixed; ++k) fx[fixed1[k] - 1] = 1;
    const int threads = omp_get_max_threads();

    // ---------------------------------------------------------------- assembly
    const auto ta = Clock::now();
    Csr K = pattern(nnod, nel, e2n, 6, 2, e2n, 6, 2);
    Csr G = pattern(nnod, nel, e2n, 6, 2, e2v, 3, 1);
    Csr M = pattern(nvert, nel, e2v, 3, 1, e2v, 3, 1);
    std::vector<double> F(nU, 0.0);
    // element colouring: no two elements of a colour share a node
    std::vector<int32_t> colour(nel);
    int ncol = 0;
    {
        std::vector<uint64_t> used(nnod, 0);
        for (int64_t e = 0; e < nel; ++e) {
            uint64_t forbid = 0;
            for (int a = 0; a < 6; ++a) forbid |= used[e2n[e * 6 + a]];
            int c = 0;
            while (forbid >> c & 1ull) ++c;
            colour[e] = c;
            ncol = std::max(ncol, c + 1);
            for (int a = 0; a < 6; ++a) used[e2n[e * 6 + a]] |= 1ull << c;
        }
    }
    std::vector<std::vector<int64_t>> by_colour(ncol);
    for (int64_t e = 0; e < nel; ++e) by_colour[colour[e]].push_back(e);
    const Rule rule;
    std::vector<double> ncs(2 * nnod);
    for (int64_t n = 0; n < nnod; ++n) { ncs[2 * n] = std::cos(nth[n]); ncs[2 * n + 1] = std::sin(nth[n]); }
    for (int cidx = 0; cidx < ncol; ++cidx) {
        const std::vector<int64_t> &els = by_colour[cidx];
#pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)els.size(); ++ii) {
            const int64_t e = els[ii];
            const int32_t *n = &e2n[e * 6];
            const double th[3] = {elth[e * 3], elth[e * 3 + 1], elth[e * 3 + 2]};
            const double rr[3] = {elr[e * 3], elr[e * 3 + 1], elr[e * 3 + 2]};
            const double R21 = rr[1] - rr[0], R31 = rr[2] - rr[0], Th21 = th[1] - th[0], Th31 = th[2] - th[0];
            const double detJ = Th21 * R31 - Th31 * R21;
            double Ke[12][12] = {}, Ge[12][3] = {}, Me[3][3] = {}, fe[12] = {};
            for (int q = 0; q < 6; ++q) {
                const double xi = rule.xi[q], et = rule.eta[q], t = 1.0 - xi - et;
                const double N[6] = {t * (2 * t - 1), xi * (2 * xi - 1), et * (2 * et - 1), 4 * xi * t, 4 * xi * et, 4 * et * t};
                const double dr[6] = {-(4 * t - 1), 4 * xi - 1, 0.0, 4 * (t - xi), 4 * et, -4 * et};
                const double ds[6] = {-(4 * t - 1), 0.0, 4 * et - 1, -4 * xi, 4 * xi, 4 * (t - et)};
                const double NP[3] = {t, xi, et};
                const double thq = t * th[0] + xi * th[1] + et * th[2], rq = t * rr[0] + xi * rr[1] + et * rr[2];
                const double sn = std::sin(thq), cs = std::cos(thq);
                const double inv = 1.0 / rq / detJ;
                const double jx0 = (R31 * cs - Th31 * sn * rq) * inv, jx1 = (-R21 * cs + Th21 * sn * rq) * inv;
                const double jz0 = (-R31 * sn - Th31 * cs * rq) * inv, jz1 = (R21 * sn + Th21 * cs * rq) * inv;
                const double om = rq * detJ * rule.w[q];
                double dx[6], dz[6];
                for (int b = 0; b < 6; ++b) { dx[b] = jx0 * dr[b] + jx1 * ds[b]; dz[b] = jz0 * dr[b] + jz1 * ds[b]; }
                const double et_q = eta[(int64_t)q * nel + e], sc = om * et_q;
                // D = eta [[C11, C13, C15], [C13, C33, C35], [C15, C35, C55]] (Voigt xx, zz, xz)
                const double d11 = C[0][(int64_t)q * nel + e] * sc, d22 = C[1][(int64_t)q * nel + e] * sc;
                const double d33 = C[2][(int64_t)q * nel + e] * sc, d12 = C[3][(int64_t)q * nel + e] * sc;
                const double d13 = C[4][(int64_t)q * nel + e] * sc, d23 = C[5][(int64_t)q * nel + e] * sc;
                for (int a = 0; a < 6; ++a)
                    for (int b = 0; b < 6; ++b) {
                        const double Dx0 = d11 * dx[b] + d13 * dz[b], Dx1 = d12 * dx[b] + d23 * dz[b], Dx2 = d13 * dx[b] + d33 * dz[b];
                        const double Dz0 = d12 * dz[b] + d13 * dx[b], Dz1 = d22 * dz[b] + d23 * dx[b], Dz2 = d23 * dz[b] + d33 * dx[b];
                        Ke[2 * a][2 * b] += dx[a] * Dx0 + dz[a] * Dx2;
                        Ke[2 * a][2 * b + 1] += dx[a] * Dz0 + dz[a] * Dz2;
                        Ke[2 * a + 1][2 * b] += dz[a] * Dx1 + dx[a] * Dx2;
                        Ke[2 * a + 1][2 * b + 1] += dz[a] * Dz1 + dx[a] * Dz2;
                    }
                for (int a = 0; a < 6; ++a)
                    for (int i = 0; i < 3; ++i) { Ge[2 * a][i] += om * NP[i] * dx[a]; Ge[2 * a + 1][i] += om * NP[i] * dz[a]; }
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) Me[i][j] += om / et_q * NP[i] * NP[j];
                const double rhoq = NP[0] * rho[n[0]] + NP[1] * rho[n[1]] + NP[2] * rho[n[2]];
                const double cte = om * g * rhoq;
                for (int a = 0; a < 6; ++a) { fe[2 * a] += sn * cte * N[a]; fe[2 * a + 1] += cs * cte * N[a]; }
            }
            // node frames: block (a, b) <- Q_a blk Q_b^T with Q = [[c, -s], [s, c]] (TT^T K TT)
            double c[6], s[6];
            for (int a = 0; a < 6; ++a) { c[a] = ncs[2 * n[a]]; s[a] = ncs[2 * n[a] + 1]; }
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) {
                    const double k00 = Ke[2 * a][2 * b], k01 = Ke[2 * a][2 * b + 1], k10 = Ke[2 * a + 1][2 * b], k11 = Ke[2 * a + 1][2 * b + 1];
                    const double t00 = c[a] * k00 - s[a] * k10, t01 = c[a] * k01 - s[a] * k11;
                    const double t10 = s[a] * k00 + c[a] * k10, t11 = s[a] * k01 + c[a] * k11;
                    Ke[2 * a][2 * b] = t00 * c[b] - t01 * s[b];
                    Ke[2 * a][2 * b + 1] = t00 * s[b] + t01 * c[b];
                    Ke[2 * a + 1][2 * b] = t10 * c[b] - t11 * s[b];
                    Ke[2 * a + 1][2 * b + 1] = t10 * s[b] + t11 * c[b];
                }
            for (int a = 0; a < 6; ++a) {
                for (int i = 0; i < 3; ++i) {
                    const double g0 = Ge[2 * a][i], g1 = Ge[2 * a + 1][i];
                    Ge[2 * a][i] = c[a] * g0 - s[a] * g1;
                    Ge[2 * a + 1][i] = s[a] * g0 + c[a] * g1;
                }
                const double f0 = fe[2 * a], f1 = fe[2 * a + 1];
                fe[2 * a] = c[a] * f0 - s[a] * f1;
                fe[2 * a + 1] = s[a] * f0 + c[a] * f1;
            }
            // scatter (free-slip rows and columns of K and rows of G left out; identity later)
            for (int a = 0; a < 6; ++a)
                for (int ra = 0; ra < 2; ++ra) {
                    const int64_t row = 2 * (int64_t)n[a] + ra;
                    F[row] += fe[2 * a + ra];
                    if (fx[row]) continue;
                    for (int b = 0; b < 6; ++b)
                        for (int cb = 0; cb < 2; ++cb) {
                            const int32_t cl = 2 * n[b] + cb;
                            if (fx[cl]) continue;
                            K.val[K.find(row, cl)] += Ke[2 * a + ra][2 * b + cb];
                        }
                    for (int i = 0; i < 3; ++i) G.val[G.find(row, n[i])] += Ge[2 * a + ra][i];
                }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) M.val[M.find(n[i], n[j])] += Me[i][j];
        }
    }
    for (int64_t d = 0; d < nU; ++d)
        if (fx[d]) K.val[K.find(d, (int32_t)d)] = 1.0;
    // G^T as CSR (nP x nU)
    Csr GT;
    GT.n = nP;
    GT.ptr.assign(nP + 1, 0);
    for (int64_t k = 0; k < (int64_t)G.col.size(); ++k) GT.ptr[G.col[k] + 1]++;
    for (int64_t i = 0; i < nP; ++i) GT.ptr[i + 1] += GT.ptr[i];
    GT.col.resize(G.col.size());
    GT.val.resize(G.col.size());
    {
        std::vector<int64_t> fill(GT.ptr.begin(), GT.ptr.end() - 1);
        for (int64_t r = 0; r < G.n; ++r)
            for (int64_t k = G.ptr[r]; k < G.ptr[r + 1]; ++k) {
                const int64_t p = fill[G.col[k]]++;
                GT.col[p] = (int32_t)r;
                GT.val[p] = G.val[k];
            }
    }
    const double t_assembly = since(ta);

    // ---------------------------------------------------------------- StokesPcCG
    const auto ts = Clock::now();
    std::vector<double> pc(nP), dinv(nU), v(nU, 0.0);
    for (int64_t i = 0; i < nP; ++i) {
        double s = 0.0;
        for (int64_t k = M.ptr[i]; k < M.ptr[i + 1]; ++k) s += M.val[k];
        pc[i] = 1.0 / s;   // lumped 1/eta mass (_preconditioner(:jacobi))
    }
    for (int64_t d = 0; d < nU; ++d) dinv[d] = 1.0 / K.val[K.find(d, (int32_t)d)];
    // nodal rigid rotation u_theta = r, u_r = 0 (unit, zero on fixed dofs): the near-null mode
    {
        double s2 = 0.0;
        for (int64_t n = 0; n < nnod; ++n) s2 += fx[2 * n] ? 0.0 : nr[n] * nr[n];
        for (int64_t n = 0; n < nnod; ++n) v[2 * n] = fx[2 * n] ? 0.0 : nr[n] / std::sqrt(s2);
    }
    long long inner_total = 0;
    bool out_of_time = false;
    std::vector<double> r(nU), z(nU), p(nU), Ap(nU);
    // x = (P K P)^-1 P y by point-Jacobi CG with P = I - v v^T, y and x on free dofs.  r and p stay
    // orthogonal to v, so p.(P K p) = p.K p and r.(P z) = r.z: the projections fold into the passes.
    auto inner = [&](const std::vector<double> &y, std::vector<double> &x) {
        double yv = 0.0;
#pragma omp parallel for reduction(+ : yv) schedule(static)
        for (int64_t i = 0; i < nU; ++i) {
            r[i] = fx[i] ? 0.0 : y[i];
            x[i] = 0.0;
            yv += r[i] * v[i];
        }
        double bb = 0.0, rz = 0.0, zv = 0.0;
#pragma omp parallel for reduction(+ : bb, rz, zv) schedule(static)
        for (int64_t i = 0; i < nU; ++i) {
            r[i] -= yv * v[i];
            z[i] = dinv[i] * r[i];
            bb += r[i] * r[i];
            rz += r[i] * z[i];
            zv += z[i] * v[i];
        }
        if (!(bb > 0)) return;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nU; ++i) p[i] = z[i] - zv * v[i];
        for (int it = 0; it < 10000000; ++it) {
            double pAp = 0.0, vAp = 0.0;
#pragma omp parallel for reduction(+ : pAp, vAp) schedule(static)
            for (int64_t i = 0; i < nU; ++i) {
                double s = 0.0;
                for (int64_t k = K.ptr[i]; k < K.ptr[i + 1]; ++k) s += K.val[k] * p[K.col[k]];
                Ap[i] = s;
                pAp += p[i] * s;
                vAp += v[i] * s;
            }
            const double al = rz / pAp;
            double rr = 0.0, rz2 = 0.0;
            zv = 0.0;
#pragma omp parallel for reduction(+ : rr, rz2, zv) schedule(static)
            for (int64_t i = 0; i < nU; ++i) {
                x[i] += al * p[i];
                r[i] -= al * (Ap[i] - vAp * v[i]);
                z[i] = dinv[i] * r[i];
                rr += r[i] * r[i];
                rz2 += r[i] * z[i];
                zv += z[i] * v[i];
            }
            ++inner_total;
            if (rr <= inner_rtol * inner_rtol * bb) break;
            if (since(ts) > budget) { out_of_time = true; break; }
            const double be = rz2 / rz;
            rz = rz2;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < nU; ++i) p[i] = z[i] - zv * v[i] + be * p[i];
        }
    };
    std::vector<double> U(nU, 0.0), P(nP, 0.0), y(nU), zz(nU), Sq(nP), rs(nP), d(nP), q(nP), rl(nP);
    auto rm0 = [&](std::vector<double> &a) {
        double s = 0.0;
        for (double x : a) s += x;
        s /= (double)a.size();
        for (double &x : a) x -= s;
    };
    int outer = 0;
    bool conv = false;
    inner(F, U);                                   // U = K^-1 (F + G P), P = 0
    std::vector<double> tmp(nP);
    GT.apply(U.data(), tmp.data());
    for (int64_t i = 0; i < nP; ++i) rs[i] = -tmp[i];
    rm0(rs);
    const double r0 = std::sqrt(dot(rs, rs)), tol = 1e-8 * r0;
    for (int64_t i = 0; i < nP; ++i) d[i] = rs[i] * pc[i];
    q = d;
    while (!out_of_time && outer < 200) {
        ++outer;
        rm0(q);
        const double rd = dot(rs, d);
        G.apply(q.data(), y.data());
        inner(y, zz);
        GT.apply(zz.data(), Sq.data());
        const double al = rd / dot(q, Sq);
        rl = rs;
        for (int64_t i = 0; i < nP; ++i) { P[i] += al * q[i]; rs[i] -= al * Sq[i]; }
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nU; ++i) U[i] += al * zz[i];
        rm0(rs);
        if (std::sqrt(dot(rs, rs)) < tol && outer >= 10) { conv = true; break; }
        for (int64_t i = 0; i < nP; ++i) d[i] = rs[i] * pc[i];
        double num = 0.0;
        for (int64_t i = 0; i < nP; ++i) num += (rs[i] - rl[i]) * d[i];
        const double be = num / rd;
        for (int64_t i = 0; i < nP; ++i) q[i] = d[i] + be * q[i];
    }
    const double t_solve = since(ts);
    {
        FILE *fu = std::fopen((dir + "/cpu_U.bin").c_str(), "wb");
        FILE *fp = std::fopen((dir + "/cpu_P.bin").c_str(), "wb");
        if (!fu || !fp || std::fwrite(U.data(), 8, nU, fu) != (size_t)nU || std::fwrite(P.data(), 8, nP, fp) != (size_t)nP) {
            std::fprintf(stderr, "cannot write the solution\n");
            return 2;
        }
        std::fclose(fu);
        std::fclose(fp);
    }
    std::printf("{\"threads\": %d, \"nU\": %lld, \"nP\": %lld, \"nnz_K\": %lld, \"assembly_s\": %.4f, "
                "\"solve_s\": %.4f, \"inner_iterations\": %lld, \"outer_iterations\": %d, \"converged\": %s, "
                "\"out_of_time\": %s, \"s_per_inner_iteration\": %.6e}\n",
                threads, (long long)nU, (long long)nP, (long long)K.col.size(), t_assembly, t_solve, inner_total, outer,
                conv ? "true" : "false", out_of_time ? "true" : "false",
                inner_total > 0 ? t_solve / (double)inner_total : 0.0);
    return 0;
}
