// Dense inverse of a symmetric positive definite matrix on the GPU (the coarsest multigrid level,
// n of a few thousand), FP64, in three stages over 64 x 64 tiles of the column-major matrix padded
// to a multiple of 64 (identity in the padding):
//
//  1. right-looking blocked Cholesky A = L L^T: per tile column k, the diagonal tile by four
//     wavefronts (k_potrf_tile), the tiles below it by triangular solves (k_trsm_panel), then the
//     trailing lower tiles minus L(I,k) L(J,k)^T (k_tile_gemm, SYRK form);
//  2. Y = L^-1 by blocked forward substitution: per tile row i, Y(i, j) = L(i,i)^-1 R(i, j) for j <= i
//     (k_trsm_rows), then R(m, j) -= L(m, i) Y(i, j) for the rows m > i (k_tile_gemm), on a second
//     stream, each tile row as soon as the factorisation has finished its tile column;
//  3. A^-1 = Y^T Y (k_tile_gemm, LAUUM form), written as the full symmetric matrix.
//
// Every stage is made of 64 x 64 x 64 tile products staged through LDS in chunks of 16, each of the
// 256 threads of a workgroup holding a 4 x 4 block of the result.  About n^3 flops in all.
#include <hip/hip_runtime.h>

#include <cmath>

#include "stk_internal.h"

namespace stk {

namespace {

constexpr int kT = 64;      // tile size
constexpr int kKc = 16;     // inner chunk staged in LDS

// Tile (I, J) of a column-major matrix with leading dimension ld.
__device__ __forceinline__ double *tile(double *A, int64_t ld, int I, int J) {
    return A + (int64_t)J * kT * ld + (int64_t)I * kT;
}

// Factor the lower triangle of diagonal tile k in place by four wavefronts: thread (t, q) =
// (tid % 64, tid / 64) keeps the entries c = 4 m + q of row t in registers.  Per column j the pivot
// comes through LDS from the thread that holds it, every thread forms its square root and reciprocal
// (the same values), the threads holding column j scale it and publish it in LDS, and each thread
// updates its entries of row t right of j.  Every entry of the lower triangle sees the same
// operations in the same order as in a one-row-per-lane factorisation.  The upper triangle of the
// tile is set to zero.
// A pivot that is not positive is reported in *info (1-based column, the first one only) and
// replaced by 1 so that the factorisation finishes.
constexpr int kPotrfWaves = 4;
constexpr int kPotrfPer = kT / kPotrfWaves;   // entries of a row per thread

__global__ __launch_bounds__(kT * kPotrfWaves) void k_potrf_tile(double *A, int64_t ld, int k, int *info) {
    __shared__ double col[kT];
    __shared__ double piv;
    const int t = threadIdx.x % kT, q = threadIdx.x / kT;   // q is uniform over a wavefront
    double *D = tile(A, ld, k, k);
    double a[kPotrfPer];
#pragma unroll
    for (int m = 0; m < kPotrfPer; ++m) {
        const int c = kPotrfWaves * m + q;
        a[m] = c <= t ? D[(int64_t)c * ld + t] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        const int mj = j / kPotrfWaves, qj = j % kPotrfWaves;
        if (q == qj && t == j) piv = a[mj];
        __syncthreads();
        double d = piv;
        if (!(d > 0.0)) {
            if (q == qj && t == j && *info == 0) *info = kT * k + j + 1;
            d = 1.0;
        }
        const double s = sqrt(d), rs = 1.0 / s;
        if (q == qj) {
            a[mj] = t == j ? s : (t > j ? a[mj] * rs : 0.0);
            col[t] = t > j ? a[mj] : 0.0;
        }
        __syncthreads();
        // l = L(t, j), zero on the rows t <= j; the test c > j is uniform over a wavefront.  Entries
        // above the diagonal (c > t) take updates too: they are never read and are stored as zero.
        const double l = col[t];
#pragma unroll
        for (int m = 0; m < kPotrfPer; ++m) {
            const int c = kPotrfWaves * m + q;
            if (c > j) a[m] -= l * col[c];
        }
    }
#pragma unroll
    for (int m = 0; m < kPotrfPer; ++m) {
        const int c = kPotrfWaves * m + q;
        D[(int64_t)c * ld + t] = c <= t ? a[m] : 0.0;
    }
}

// L(I, k) = A(I, k) L(k, k)^-T for the tiles I = k + 1 + blockIdx.x: lane t solves row t in
// registers, L(k, k) and the reciprocals of its diagonal are read from LDS as broadcasts; each
// forward step sums in two chains.
__global__ __launch_bounds__(kT) void k_trsm_panel(double *A, int64_t ld, int k) {
    __shared__ double L[kT][kT + 1];
    __shared__ double rdiag[kT];
    const int t = threadIdx.x;
    const int I = k + 1 + blockIdx.x;
    const double *D = tile(A, ld, k, k);
    double *P = tile(A, ld, I, k);
    double x[kT];
#pragma unroll
    for (int c = 0; c < kT; ++c) {
        L[t][c] = D[(int64_t)c * ld + t];
        x[c] = P[(int64_t)c * ld + t];
    }
    __syncthreads();
    rdiag[t] = 1.0 / L[t][t];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kT; ++j) {
        double v0 = x[j], v1 = 0.0;
#pragma unroll
        for (int m = 0; m < j; m += 2) {
            v0 -= x[m] * L[j][m];
            if (m + 1 < j) v1 -= x[m + 1] * L[j][m + 1];
        }
        x[j] = (v0 + v1) * rdiag[j];
    }
#pragma unroll
    for (int c = 0; c < kT; ++c) P[(int64_t)c * ld + t] = x[c];
}

// Y(i, j) = L(i, i)^-1 R(i, j) for j = blockIdx.x <= i, in place in Y: lane t solves column t in
// registers, L(i, i) is read from LDS as broadcasts.
__global__ __launch_bounds__(kT) void k_trsm_rows(const double *A, double *Y, int64_t ld, int i) {
    __shared__ double L[kT][kT + 1];
    const int t = threadIdx.x;
    const int j = blockIdx.x;
    const double *D = A + (int64_t)i * kT * ld + (int64_t)i * kT;
    double *B = Y + (int64_t)j * kT * ld + (int64_t)i * kT;
    __shared__ double rdiag[kT];
    double x[kT];
#pragma unroll
    for (int c = 0; c < kT; ++c) L[t][c] = D[(int64_t)c * ld + t];
    // column t of the tile: B(r, t) at B[t * ld + r]
#pragma unroll
    for (int r = 0; r < kT; ++r) x[r] = B[(int64_t)t * ld + r];
    __syncthreads();
    rdiag[t] = 1.0 / L[t][t];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kT; ++r) {
        double v0 = x[r], v1 = 0.0;
#pragma unroll
        for (int m = 0; m < r; m += 2) {
            v0 -= L[r][m] * x[m];
            if (m + 1 < r) v1 -= L[r][m + 1] * x[m + 1];
        }
        x[r] = (v0 + v1) * rdiag[r];
    }
#pragma unroll
    for (int r = 0; r < kT; ++r) B[(int64_t)t * ld + r] = x[r];
}

// The three tile-product stages.
enum : int { GEMM_SYRK = 0, GEMM_FWD = 1, GEMM_LAUUM = 2 };

struct GemmArgs {
    double *A;          // Cholesky factor (lower)
    double *Y;          // L^-1 (lower), forward and LAUUM stages
    double *Out;        // LAUUM: the full inverse
    int64_t ld;
    int k;              // SYRK: tile column; FWD: tile row i
    int nt;             // tiles per side
};

// op(A)(r, q) at pa[r * sar + q * saq], op(B)(q, c) at pb[q * sbq + c * sbc]; acc += op(A) op(B) over
// the 64 values of q, in LDS chunks of 16.
__device__ __forceinline__ void tile_mma(double acc[4][4], const double *pa, int64_t sar, int64_t saq,
                                         const double *pb, int64_t sbq, int64_t sbc, double (*As)[kT + 1],
                                         double (*Bs)[kT + 1]) {
    const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
    for (int q0 = 0; q0 < kT; q0 += kKc) {
        for (int u = 0; u < 4; ++u) {
            const int idx = tid + 256 * u;
            // consecutive threads along the unit-stride index of each operand
            int q, r;
            if (sar == 1) { r = idx % kT; q = idx / kT; }
            else { q = idx % kKc; r = idx / kKc; }
            As[q][r] = pa[(int64_t)r * sar + (int64_t)(q0 + q) * saq];
            int qb, cb;
            if (sbc == 1) { cb = idx % kT; qb = idx / kT; }
            else { qb = idx % kKc; cb = idx / kKc; }
            Bs[qb][cb] = pb[(int64_t)(q0 + qb) * sbq + (int64_t)cb * sbc];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kKc; ++q) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a[i] = As[q][4 * ty + i];
                b[i] = Bs[q][4 * tx + i];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
        }
        __syncthreads();
    }
}

// Lower-triangle tile index b -> (row, column) with row >= column, both in [0, m).
__device__ __forceinline__ void tri_index(int64_t b, int *row, int *col) {
    int r = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((int64_t)r * (r + 1) / 2 > b) --r;
    while ((int64_t)(r + 1) * (r + 2) / 2 <= b) ++r;
    *row = r;
    *col = (int)(b - (int64_t)r * (r + 1) / 2);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_tile_gemm(GemmArgs g) {
    __shared__ double As[kKc][kT + 1];
    __shared__ double Bs[kKc][kT + 1];
    const int tid = threadIdx.x, ty = tid / 16, tx = tid % 16;
    double acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    const int64_t ld = g.ld;
    double *C;
    int I, J;
    if (MODE == GEMM_SYRK) {
        // C(I, J) -= L(I, k) L(J, k)^T for k < J <= I
        int ii, jj;
        tri_index(blockIdx.x, &ii, &jj);
        I = g.k + 1 + ii;
        J = g.k + 1 + jj;
        const double *pa = g.A + (int64_t)g.k * kT * ld + (int64_t)I * kT;   // (r, q) = L(64I + r, 64k + q)
        const double *pb = g.A + (int64_t)g.k * kT * ld + (int64_t)J * kT;   // (q, c) = L(64J + c, 64k + q)
        tile_mma(acc, pa, 1, ld, pb, ld, 1, As, Bs);
        C = g.A + (int64_t)J * kT * ld + (int64_t)I * kT;
    } else if (MODE == GEMM_FWD) {
        // R(m, j) -= L(m, i) Y(i, j) for m = i + 1 + blockIdx.y, j = blockIdx.x <= i
        const int i = g.k;
        I = i + 1 + blockIdx.y;
        J = blockIdx.x;
        const double *pa = g.A + (int64_t)i * kT * ld + (int64_t)I * kT;     // (r, q) = L(64m + r, 64i + q)
        const double *pb = g.Y + (int64_t)J * kT * ld + (int64_t)i * kT;     // (q, c) = Y(64i + q, 64j + c)
        tile_mma(acc, pa, 1, ld, pb, 1, ld, As, Bs);
        C = g.Y + (int64_t)J * kT * ld + (int64_t)I * kT;
    } else {
        // X(I, J) = sum_{k >= I} Y(k, I)^T Y(k, J) for I >= J
        int ii, jj;
        tri_index(blockIdx.x, &ii, &jj);
        I = ii;
        J = jj;
        for (int k = I; k < g.nt; ++k) {
            const double *pa = g.Y + (int64_t)I * kT * ld + (int64_t)k * kT;   // (r, q) = Y(64k + q, 64I + r)
            const double *pb = g.Y + (int64_t)J * kT * ld + (int64_t)k * kT;   // (q, c) = Y(64k + q, 64J + c)
            tile_mma(acc, pa, ld, 1, pb, 1, ld, As, Bs);
        }
        double *Ci = g.Out + (int64_t)J * kT * ld + (int64_t)I * kT;
        double *Ct = g.Out + (int64_t)I * kT * ld + (int64_t)J * kT;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int rr = 4 * ty + r, cc = 4 * tx + c;
                Ci[(int64_t)cc * ld + rr] = acc[r][c];
                if (I != J) Ct[(int64_t)rr * ld + cc] = acc[r][c];   // the mirror tile
            }
        return;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) C[(int64_t)(4 * tx + c) * ld + 4 * ty + r] -= acc[r][c];
}

__global__ void k_identity_lower(double *Y, int64_t ld) {
    const int64_t j = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ld) Y[j * ld + i] = (i == j) ? 1.0 : 0.0;
}

}  // namespace

// In-place inverse of the SPD matrix A (np x np column-major, np a multiple of 64); Y is np x np
// work space.  *info (device) receives the first non-positive pivot (1-based) or stays 0.
// The forward stage runs on stream s2 beside the factorisation on s: its tile row i reads only
// L(i, i) and the column L(m > i, i), which are final once the factorisation's step i has solved its
// panel (the later trailing updates touch columns right of i), so after each panel solve s records
// `ev` and s2 waits for it.  The inverse stage waits on s for the whole forward stage (`ev2`).  Each
// tile is computed by the same kernels in the same order as on one stream.
void spd_inverse(hipStream_t s, hipStream_t s2, hipEvent_t ev, hipEvent_t ev2, double *A, double *Y, int64_t np,
                 int *info) {
    if (np % kT != 0) throw Error(STK_E_STATE, "spd_inverse needs a multiple of 64");
    const int nt = (int)(np / kT);
    GemmArgs g{A, Y, A, np, 0, nt};
    GemmArgs gf = g;
    // s2 starts after the work already on s (the matrix A is written there)
    STK_HIP(hipEventRecord(ev, s));
    STK_HIP(hipStreamWaitEvent(s2, ev, 0));
    k_identity_lower<<<dim3((unsigned)((np + 255) / 256), (unsigned)np), dim3(256), 0, s2>>>(Y, np);
    for (int k = 0; k < nt; ++k) {
        k_potrf_tile<<<dim3(1), dim3(kT * kPotrfWaves), 0, s>>>(A, np, k, info);
        const int m = nt - k - 1;
        if (m > 0) k_trsm_panel<<<dim3(m), dim3(kT), 0, s>>>(A, np, k);
        STK_HIP(hipEventRecord(ev, s));
        STK_HIP(hipStreamWaitEvent(s2, ev, 0));
        if (m > 0) {
            g.k = k;
            k_tile_gemm<GEMM_SYRK><<<dim3((unsigned)((int64_t)m * (m + 1) / 2)), dim3(256), 0, s>>>(g);
        }
        // forward tile row k on s2
        k_trsm_rows<<<dim3(k + 1), dim3(kT), 0, s2>>>(A, Y, np, k);
        if (m > 0) {
            gf.k = k;
            k_tile_gemm<GEMM_FWD><<<dim3(k + 1, m), dim3(256), 0, s2>>>(gf);
        }
    }
    STK_HIP(hipGetLastError());
    STK_HIP(hipEventRecord(ev2, s2));
    STK_HIP(hipStreamWaitEvent(s, ev2, 0));
    g.Out = A;
    k_tile_gemm<GEMM_LAUUM><<<dim3((unsigned)((int64_t)nt * (nt + 1) / 2)), dim3(256), 0, s>>>(g);
    STK_HIP(hipGetLastError());
}

}  // namespace stk
