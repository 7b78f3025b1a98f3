// Finite strain ellipses (getFSE) and the DEM viscous tensor (anisotropic_tensor).
//
// getFSE (src/Stress/FSE.jl:131-231, 297-355): at every point of the deformation gradient F
// the semi-axes and axes of the ellipse F F^T, with the healing rule of the run: none,
// fabric destruction (a1/a2 >= eps resets F to I and the ellipse to the unit circle), or
// annealing + destruction (one normalisation of all F first; axes with a1 <= 1 or a2 >= 1 count
// as destroyed).  The 2x2 symmetric eigen solve is the closed form StaticArrays uses for
// Hermitian 2x2 matrices; contraction into fused multiply-adds is off in it so the rounding
// is the one of the plain expression.
//
// anisotropic_tensor (src/Rheology/anisotropy.jl:140-285): the fabric shape (r1, r2) of each
// point picks the nearest entry of the DEM table (first local minimum of the distance along the
// sorted block values, argminsorted :114-122), whose C11, C33, C55 (clamped to the cut-off w),
// C13 are rotated into the ellipse axis (rotmatrix :287-292, directRotation2D :503-536); eta55 is
// clamped to w again and the 7th column is the mean of the first six.
//
// One thread per point (ellipses) or per element (tensor).  F is 4 doubles per point (the
// reference's layout), so a wave reads and writes one contiguous run; the device keeps the
// ellipses as six planes (field k of point p at k * npt + p), so the tensor reads only the
// three fields it uses, coalesced; the tensor planes are nel x 7 column-major (element fastest),
// written coalesced.  The DEM table (a few hundred doubles) is read through the caches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "fabric.h"

namespace stk {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ double nan_max(double x, double w) { return (x < w) ? w : x; }

// max over the points of |F11| (NaN skipped), in two passes: block partials, then one block.
__global__ void __launch_bounds__(kBlock) k_f11_partial(const double *F, int64_t npt, double *partial) {
    __shared__ double s[kBlock];
    double m = -INFINITY;
    for (int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x; p < npt; p += (int64_t)gridDim.x * kBlock) {
        const double v = fabs(F[4 * p]);
        if (v > m) m = v;
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h && s[threadIdx.x + h] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// scale = 1 / max when max > 1e30 (normalize_F!, FSE.jl:163-178), else 1
__global__ void __launch_bounds__(kBlock) k_f11_scale(const double *partial, int npart, double *scale) {
    __shared__ double s[kBlock];
    double m = -INFINITY;
    for (int i = threadIdx.x; i < npart; i += kBlock)
        if (partial[i] > m) m = partial[i];
    s[threadIdx.x] = m;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h && s[threadIdx.x + h] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) scale[0] = s[0] > 1e30 ? 1.0 / s[0] : 1.0;
}

// Ellipse of Fi Fi^T: out = (x1, x2, y1, y2, a1, a2).
__device__ void ellipse(double f11, double f21, double f12, double f22, double out[6]) {
#pragma clang fp contract(off)
    const double l11 = f11 * f11 + f12 * f12;
    const double l12 = f11 * f21 + f12 * f22;
    const double l22 = f21 * f21 + f22 * f22;
    double v1, v2, e11, e21, e12, e22;  // values ascending; e_i1, e_i2 = components of vector 1, 2
    if (l12 != 0.0) {
        const double t = (l11 + l22) / 2;
        const double d = l11 * l22 - l12 * l12;
        const double s2 = t * t - d;
        const double s = s2 < 0 ? 0.0 : sqrt(s2);
        v1 = t - s;
        v2 = t + s;
        const double u1 = v1 - l22, n1 = sqrt(u1 * u1 + l12 * l12);
        const double u2 = v2 - l22, n2 = sqrt(u2 * u2 + l12 * l12);
        e11 = u1 / n1; e21 = l12 / n1;
        e12 = u2 / n2; e22 = l12 / n2;
    } else if (l11 < l22) {
        v1 = l11; v2 = l22;
        e11 = 1.0; e21 = 0.0; e12 = 0.0; e22 = 1.0;
    } else {
        v1 = l22; v2 = l11;
        e11 = 0.0; e21 = 1.0; e12 = 1.0; e22 = 0.0;
    }
    // eigval_order (FSE.jl:189): the second value is the major one only if strictly larger
    if (v2 > v1) {
        out[0] = e12; out[1] = e11; out[2] = e22; out[3] = e21;
        out[4] = sqrt(fabs(v2)); out[5] = sqrt(fabs(v1));
    } else {
        out[0] = e11; out[1] = e12; out[2] = e21; out[3] = e22;
        out[4] = sqrt(fabs(v1)); out[5] = sqrt(fabs(v2));
    }
}

__global__ void __launch_bounds__(kBlock) k_fse(double *__restrict__ F, int64_t npt, int mode, double eps,
                                                const double *__restrict__ scale, double *__restrict__ fse) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= npt) return;
    double f[4] = {F[4 * p], F[4 * p + 1], F[4 * p + 2], F[4 * p + 3]};  // F11 F21 F12 F22
    if (mode == FSE_ANNEALING_DESTRUCTION) {
        const double sc = scale[0];
        for (int k = 0; k < 4; ++k) f[k] *= sc;
    }
    double fi[4] = {f[0], f[1], f[2], f[3]};
    if (mode != FSE_NO_HEALING) {
        // normalize_F (FSE.jl:180-187): |maximum(F)| over the entries in memory order
        double fmax = f[0];
        for (int k = 1; k < 4; ++k) fmax = (f[k] > fmax || f[k] != f[k]) ? f[k] : fmax;
        fmax = fabs(fmax);
        if (fmax > 1e30)
            for (int k = 0; k < 4; ++k) fi[k] = f[k] / fmax;
    }
    double o[6];
    ellipse(fi[0], fi[1], fi[2], fi[3], o);
    if (mode != FSE_NO_HEALING) {
        double a1 = o[4], a2 = o[5];
        if (mode == FSE_ANNEALING_DESTRUCTION && (a1 <= 1.0 || a2 >= 1.0)) {  // force_isotropy (:289-295)
            a1 = 1.0;
            a2 = 0.0;
        }
        o[4] = a1;
        o[5] = a2;
        if (!(a1 / a2 < eps)) {  // fabric destroyed: isotropic aggregate
            o[0] = 1.0; o[1] = 0.0; o[2] = 0.0; o[3] = 1.0; o[4] = 1.0; o[5] = 1.0;
            f[0] = 1.0; f[1] = 0.0; f[2] = 0.0; f[3] = 1.0;
        }
        for (int k = 0; k < 4; ++k) F[4 * p + k] = f[k];
    }
    for (int k = 0; k < 6; ++k) fse[k * npt + p] = o[k];
}

// ellipses between the device planes (field k of point p at k * npt + p) and the host layout
// (6 consecutive fields per point)
__global__ void __launch_bounds__(kBlock) k_fse_layout(const double *__restrict__ src, double *__restrict__ dst,
                                                       int64_t npt, int to_points) {
    const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (p >= npt) return;
    for (int k = 0; k < 6; ++k) {
        if (to_points) dst[6 * p + k] = src[k * npt + p];
        else dst[k * npt + p] = src[6 * p + k];
    }
}

struct AnisoArgs {
    int64_t nel;
    int nip;
    const double *fse;                // 6 planes of nel x nip points
    const double *nodes, *sr1, *sr2;  // shape table (nodes != null) or per-point r1, r2
    int64_t nshape;
    const double *a1a2, *a2a3, *tab;
    const int32_t *perm;
    int32_t nblk, sblk;
    double w;
    double *out;  // 6 x nel x 7
};

// argminsorted: first i with v[i] < v[i+1], else the last index (0-based)
__device__ __forceinline__ int argminsorted(const double *__restrict__ blk, int m, double r) {
    double prev = fabs(r - blk[0]);
    for (int i = 0; i + 1 < m; ++i) {
        const double next = fabs(r - blk[i + 1]);
        if (prev < next) return i;
        prev = next;
    }
    return m - 1;
}

// fabric shape from the table: r at a1, linear between nodes, the end values beyond
__device__ __forceinline__ void shape_from_table(const double *__restrict__ nodes, const double *__restrict__ t1,
                                                 const double *__restrict__ t2, int64_t K, double a1, double &r1,
                                                 double &r2) {
#pragma clang fp contract(off)
    if (a1 <= nodes[0]) { r1 = t1[0]; r2 = t2[0]; return; }
    if (a1 >= nodes[K - 1]) { r1 = t1[K - 1]; r2 = t2[K - 1]; return; }
    // largest k with nodes[k] <= a1 (NaN: none, k = 0 and t = NaN)
    int64_t lo = 0, hi = K - 1;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) / 2;
        if (nodes[mid] <= a1) lo = mid; else hi = mid;
    }
    const double t = (a1 - nodes[lo]) / (nodes[lo + 1] - nodes[lo]);
    r1 = t1[lo] + t * (t1[lo + 1] - t1[lo]);
    r2 = t2[lo] + t * (t2[lo + 1] - t2[lo]);
}

__device__ __forceinline__ void rotate(double x, double y, double a, double c, double e, double h, double n[6]) {
#pragma clang fp contract(off)
    const double hh = 1.0 / sqrt(y * y + x * x);
    const double x1 = x * hh, x2 = y * hh, y1 = -y * hh, y2 = x1;
    const double x1x1 = x1 * x1, x2x2 = x2 * x2, y1y1 = y1 * y1, y2y2 = y2 * y2, s = x1 * y2 + x2 * y1;
    n[0] = (c * x2x2 + h * x1x1 + 4 * e * x1x1) * x2x2 + (a * x1x1 + h * x2x2) * x1x1;
    n[1] = (c * y2y2 + h * y1y1 + 4 * e * y1y1) * y2y2 + (a * y1y1 + h * y2y2) * y1y1;
    n[2] = (c * x2 * y2 + h * x1 * y1) * x2 * y2 + s * s * e + (a * x1 * y1 + h * x2 * y2) * x1 * y1;
    n[3] = ((c * x2x2 + h * x1x1) * y2 + 4 * e * x1 * x2 * y1) * y2 + (a * x1x1 + h * x2x2) * y1y1;
    n[4] = ((c * x2x2 + h * x1x1) * y2 + 2 * s * e * x1) * x2 + (a * x1x1 + h * x2x2) * x1 * y1;
    n[5] = ((c * y2y2 + h * y1y1) * x2 + 2 * s * e * y1) * y2 + (a * y1y1 + h * y2y2) * x1 * y1;
}

__global__ void __launch_bounds__(kBlock) k_anisotropic_tensor(AnisoArgs a) {
#pragma clang fp contract(off)
    const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if (e >= a.nel) return;
    const int64_t nel = a.nel;
    const double *__restrict__ fse = a.fse;
    double *__restrict__ out = a.out;
    // the 7th point of a 7-point ellipse field is computed by the reference and then replaced
    // by the mean, so only the first six are evaluated; all loads of the element first
    const int64_t npt = nel * a.nip;
    double x[6], y[6], r1[6], r2[6];
    for (int q = 0; q < 6; ++q) {
        x[q] = fse[e + q * nel];
        y[q] = fse[2 * npt + e + q * nel];
        r1[q] = fse[4 * npt + e + q * nel];
    }
    if (a.nodes) {
        for (int q = 0; q < 6; ++q) shape_from_table(a.nodes, a.sr1, a.sr2, a.nshape, r1[q], r1[q], r2[q]);
    } else {
        for (int q = 0; q < 6; ++q) {
            r1[q] = a.sr1[e + q * nel];
            r2[q] = a.sr2[e + q * nel];
        }
    }
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int q = 0; q < 6; ++q) {
        const int i1 = argminsorted(a.a1a2, a.nblk, r1[q]);
        const int i2 = a.perm[argminsorted(a.a2a3, a.sblk, r2[q])];
        const double *__restrict__ t = a.tab + 4 * ((int64_t)a.sblk * i1 + i2);
        double n[6];
        rotate(x[q], y[q], t[0], t[1], nan_max(t[2], a.w), t[3], n);
        n[2] = nan_max(n[2], a.w);
        for (int k = 0; k < 6; ++k) {
            out[(int64_t)k * 7 * nel + q * nel + e] = n[k];
            acc[k] += n[k];
        }
    }
    for (int k = 0; k < 6; ++k) out[(int64_t)k * 7 * nel + 6 * nel + e] = acc[k] * (1.0 / 6.0);
}

}  // namespace

void launch_fse(Context &c, double *F, int64_t npt, int mode, double eps, double *fse) {
    if (npt <= 0) return;
    DBuf<double> scale;
    if (mode == FSE_ANNEALING_DESTRUCTION) {
        const int npart = (int)std::min<int64_t>(1024, (npt + kBlock - 1) / kBlock);
        DBuf<double> partial(npart);
        scale.alloc(1);
        k_f11_partial<<<dim3(npart), dim3(kBlock), 0, c.stream>>>(F, npt, partial.p);
        STK_HIP(hipGetLastError());
        k_f11_scale<<<dim3(1), dim3(kBlock), 0, c.stream>>>(partial.p, npart, scale.p);
        STK_HIP(hipGetLastError());
        k_fse<<<dim3((unsigned)((npt + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream>>>(F, npt, mode, eps,
                                                                                            scale.p, fse);
        STK_HIP(hipGetLastError());
        STK_HIP(hipStreamSynchronize(c.stream));  // partial / scale are released on return
        return;
    }
    k_fse<<<dim3((unsigned)((npt + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream>>>(F, npt, mode, eps, nullptr,
                                                                                        fse);
    STK_HIP(hipGetLastError());
}

void fse_layout(Context &c, const double *src, double *dst, int64_t npt, bool to_points) {
    if (npt <= 0) return;
    k_fse_layout<<<dim3((unsigned)((npt + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream>>>(src, dst, npt,
                                                                                               to_points ? 1 : 0);
    STK_HIP(hipGetLastError());
}

void dem_prepare(Context &c, const double *tensor, const double *axes, int64_t n, double w) {
    STK_CHECK(n >= 2, STK_E_SHAPE, "the DEM table needs at least two entries");
    std::vector<double> a1a2(n), a2a3(n);
    for (int64_t i = 0; i < n; ++i) {
        STK_CHECK(std::isfinite(axes[i]) && axes[i] > 0 && std::isfinite(axes[2 * n + i]) && axes[2 * n + i] > 0,
                  STK_E_ARG, "the DEM semi-axes ax[:,1] and ax[:,3] must be positive and finite");
        a1a2[i] = std::log10(axes[i] / 1.0);
        a2a3[i] = std::log10(1.0 / axes[2 * n + i]);
    }
    // get_stride (anisotropy.jl:95-101): the leading run of equal a1 values is one block
    int64_t sblk = 1;
    while (sblk < n && a1a2[sblk] == a1a2[sblk - 1]) ++sblk;
    STK_CHECK(sblk < n, STK_E_SHAPE, "the DEM table has a single value of ax[:,1]: no block structure");
    STK_CHECK(n % sblk == 0, STK_E_SHAPE, "the DEM table length is not a multiple of its block length");
    const int64_t nblk = n / sblk;
    // sort_axes (:103-112): the sorted a2a3 values of the first block and their stable order
    std::vector<int32_t> perm(sblk);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t i, int32_t j) { return a2a3[i] < a2a3[j]; });
    std::vector<double> a2s(sblk), a1s(nblk), tab(4 * n);
    for (int64_t k = 0; k < sblk; ++k) a2s[k] = a2a3[perm[k]];
    for (int64_t b = 0; b < nblk; ++b) a1s[b] = a1a2[b * sblk];
    // get_tensor_and_rotate! (:258-266) reads C11, C33, C55 and C13: columns 1, 3, 5, 8
    for (int64_t i = 0; i < n; ++i) {
        tab[4 * i + 0] = tensor[0 * n + i];
        tab[4 * i + 1] = tensor[2 * n + i];
        tab[4 * i + 2] = tensor[4 * n + i];
        tab[4 * i + 3] = tensor[7 * n + i];
    }
    c.dem.a1a2.upload(a1s, c.stream);
    c.dem.a2a3.upload(a2s, c.stream);
    c.dem.perm.upload(perm, c.stream);
    c.dem.tab.upload(tab, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
    c.dem.sblk = (int32_t)sblk;
    c.dem.nblk = (int32_t)nblk;
    c.dem.w = w;
    c.dem.set = true;
}

void launch_anisotropic_tensor(Context &c, const double *fse, int nip, const ShapeIn &shape, double *out) {
    AnisoArgs a{};
    a.nel = c.nel;
    a.nip = nip;
    a.fse = fse;
    a.nodes = shape.nodes;
    a.sr1 = shape.r1;
    a.sr2 = shape.r2;
    a.nshape = shape.nshape;
    a.a1a2 = c.dem.a1a2.p;
    a.a2a3 = c.dem.a2a3.p;
    a.tab = c.dem.tab.p;
    a.perm = c.dem.perm.p;
    a.nblk = c.dem.nblk;
    a.sblk = c.dem.sblk;
    a.w = c.dem.w;
    a.out = out;
    k_anisotropic_tensor<<<dim3((unsigned)((c.nel + kBlock - 1) / kBlock)), dim3(kBlock), 0, c.stream>>>(a);
    STK_HIP(hipGetLastError());
}

}  // namespace stk
