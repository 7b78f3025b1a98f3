// Device building blocks: node-block sparse products with fused epilogues,
// deterministic two-pass reductions and elementwise maps.
//
// Sparse operators are stored as node-block CSR: row node i owns blocks
// ptr[i] .. ptr[i+1]-1, block k couples row node i with column node col[k] and
// holds BR x BC values (row-major) at val[k*BR*BC].  K is 2x2 (u_theta, u_r per
// node), G is 2x1, G^T is 1x2, the pressure matrices are 1x1.  A group of L
// lanes of one wavefront walks the blocks of one row node; the partial sums are
// combined with cross-lane shuffles inside the group.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace stk {

constexpr int kWave = 64;
constexpr int kReduceBlocks = 1024;   // partial-sum slots of a reduction

// Sum over the L-lane group (L a power of two dividing the wavefront).
template <int L, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, L);
    return v;
}

// Block-wide sum of one double per thread; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double part[NT / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    if (l == 0) part[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NT / kWave; ++i) s += part[i];
    }
    __syncthreads();
    return s;
}

// Several dot products accumulated together (epilogues with kDots > 1 return one).
template <int ND>
struct Dots {
    double v[ND];
};
__device__ __forceinline__ void add_dots(double (&s)[1], double v) { s[0] += v; }
template <int ND>
__device__ __forceinline__ void add_dots(double (&s)[ND], const Dots<ND> &v) {
#pragma unroll
    for (int d = 0; d < ND; ++d) s[d] += v.v[d];
}

// Block-wide sums of ND doubles per thread; results valid in thread 0.
template <int NT, int ND>
__device__ __forceinline__ void block_sum_n(double (&v)[ND]) {
    __shared__ double part[ND][NT / kWave];
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) v[d] += __shfl_xor(v[d], o, kWave);
        if (l == 0) part[d][w] = v[d];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0;
#pragma unroll
            for (int i = 0; i < NT / kWave; ++i) s += part[d][i];
            v[d] = s;
        }
    }
    __syncthreads();
}

// Blocks are dealt round-robin over the 8 XCDs in launch order, so consecutive row
// blocks run on different XCDs at the same time: the whole chip sweeps the rows front
// to back together, their x gathers overlap in the shared last-level cache, and rows of
// different cost (P2 vertex rows carry 19 blocks, midpoint rows 9, and vertices are
// numbered first) are spread evenly over the XCDs.  Giving each XCD one contiguous
// eighth of the rows instead measured 25% slower on K (two XCDs held all vertex rows).

// ------------------------------------------------------------------ node-block SpMV
// Epilogue interface: operator()(row_dof, Ax value) -> contribution to a dot (double);
// rowvec() (host) counts the row-length vectors it reads or writes, y included.
template <typename T>
struct EpStore {                       // y = A x
    T *y;
    static constexpr int kDots = 1;
    __device__ double operator()(int64_t i, T ax) const { y[i] = ax; return 0.0; }
    int rowvec() const { return 1; }
};
template <typename T>
struct EpStoreDot {                    // y = A x, dot += p_i (A x)_i
    T *y; const T *p;
    static constexpr int kDots = 1;
    __device__ double operator()(int64_t i, T ax) const { y[i] = ax; return (double)p[i] * (double)ax; }
    int rowvec() const { return 2; }
};
template <typename T>
struct EpResidual {                    // r = b - A x, dot += r_i^2
    T *r; const T *b;
    static constexpr int kDots = 1;
    __device__ double operator()(int64_t i, T ax) const {
        const T v = b[i] - ax; r[i] = v; return (double)v * (double)v;
    }
    int rowvec() const { return 2; }
};
template <typename T>
struct EpAxpyOut {                     // y = beta*y + alpha*A x  (alpha, beta scalars)
    T *y; T alpha, beta;
    static constexpr int kDots = 1;
    __device__ double operator()(int64_t i, T ax) const { y[i] = beta * y[i] + alpha * ax; return 0.0; }
    int rowvec() const { return 2; }
};
template <typename T>
struct EpJacobi {                      // damped Jacobi: dout = din + w Dinv (r - A din)
    static constexpr int kDots = 1;
    const T *r, *dinv, *din; T *dout; T w;
    __device__ double operator()(int64_t i, T ax) const { dout[i] = din[i] + w * dinv[i] * (r[i] - ax); return 0.0; }
    int rowvec() const { return 4; }
};
template <typename T>
struct EpStorePlus {                   // y = A x + a
    static constexpr int kDots = 1;
    T *y; const T *a;
    __device__ double operator()(int64_t i, T ax) const { y[i] = ax + a[i]; return 0.0; }
    int rowvec() const { return 2; }
};
template <typename T>
struct EpStoreDot2 {                   // y = A x, dots (p . A x, w . A x)
    static constexpr int kDots = 2;
    T *y; const T *p, *w;
    __device__ Dots<2> operator()(int64_t i, T ax) const {
        y[i] = ax;
        return Dots<2>{{(double)p[i] * (double)ax, (double)w[i] * (double)ax}};
    }
    int rowvec() const { return 3; }
};
// Chebyshev step that also returns the dots (b . xo, w . xo) of its result.
template <typename T>
struct EpChebDot2 {
    static constexpr int kDots = 2;
    const T *b, *dinv, *x; const T *d; T *xo; T c1, c2; const T *w;
    bool rd = true;   // false: c1 = 0 and d is not read (first step)
    __device__ Dots<2> operator()(int64_t i, T ax) const {
        // the last step of a smoothing: its direction is not needed afterwards
        const T dn = (rd ? c1 * d[i] : T(0)) + c2 * dinv[i] * (b[i] - ax);
        const T v = x[i] + dn;
        xo[i] = v;
        return Dots<2>{{(double)b[i] * (double)v, (double)w[i] * (double)v}};
    }
    int rowvec() const { return 5 + (rd ? 1 : 0); }
};
// Chebyshev step with point-Jacobi: d = c1 d + c2 Dinv (b - A x); xo = x + d.
// rd = false: first step (c1 = 0), d is not read; wd = false: last step, d is not stored.
template <typename T>
struct EpCheb {
    const T *b, *dinv, *x; T *d, *xo; T c1, c2;
    bool rd = true, wd = true;
    static constexpr int kDots = 1;
    __device__ double operator()(int64_t i, T ax) const {
        const T dn = (rd ? c1 * d[i] : T(0)) + c2 * dinv[i] * (b[i] - ax);
        if (wd) d[i] = dn;
        xo[i] = x[i] + dn; return 0.0;
    }
    int rowvec() const { return 4 + (rd ? 1 : 0) + (wd ? 1 : 0); }
};

// First Chebyshev step from a zero start, whose iterate x0 = d0 = s D^-1 b is formed here
// (and in the gather) rather than read: d = c1 x0 + c2 D^-1 (b - A x0); xo = x0 + d.
template <typename T>
struct EpChebFirst {
    const T *b, *dinv; T *d, *xo; T s, c1, c2;
    bool wd = true;   // false: last step, d is not stored
    static constexpr int kDots = 1;
    __device__ double operator()(int64_t i, T ax) const {
        const T di = dinv[i], bi = b[i];
        const T x0 = s * di * bi;
        const T dn = c1 * x0 + c2 * di * (bi - ax);
        if (wd) d[i] = dn;
        xo[i] = x0 + dn; return 0.0;
    }
    int rowvec() const { return 3 + (wd ? 1 : 0); }
};

// Chebyshev step in the symmetrically scaled variables of A^ = S^-1 A S^-1 (S^2 = diag A),
// where point Jacobi is the identity: d = c1 d + c2 (b^ - A^ x^); xo^ = x^ + d; on the last
// step also the unscaled xu = S^-1 xo^.
struct EpChebScaled {
    static constexpr int kDots = 1;
    const float *b, *x; float *d, *xo; float c1, c2; float *xu; const float *sinv;
    __device__ double operator()(int64_t i, float ax) const {
        const float dn = c1 * d[i] + c2 * (b[i] - ax);
        d[i] = dn;
        const float v = x[i] + dn;
        xo[i] = v;
        if (xu) xu[i] = v * sinv[i];
        return 0.0;
    }
    int rowvec() const { return 5 + (xu ? 2 : 0); }
};
// xu += A x (prolongation) and its scaled copy xh = S xu.
struct EpAxpyScaled {
    static constexpr int kDots = 1;
    float *xu, *xh; const float *s;
    __device__ double operator()(int64_t i, float ax) const {
        const float v = xu[i] + ax;
        xu[i] = v;
        xh[i] = v * s[i];
        return 0.0;
    }
    int rowvec() const { return 4; }
};

template <int BR, int BC, int L, int NT, typename T, typename Ep>
__global__ __launch_bounds__(NT) void k_spmv_block(int64_t nrows, const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ col,
                                                    const T *__restrict__ val, const T *__restrict__ x,
                                                    Ep ep, double *__restrict__ partial, int64_t nblocks) {
    constexpr int ND = Ep::kDots;
    const int64_t lb = blockIdx.x;
    const int64_t row = lb * (NT / L) + threadIdx.x / L;
    const int lane = threadIdx.x % L;
    double dsum[ND] = {};
    if (row < nrows) {
        T acc[BR];
#pragma unroll
        for (int r = 0; r < BR; ++r) acc[r] = T(0);
        const int32_t b0 = ptr[row], b1 = ptr[row + 1];
        for (int32_t k = b0 + lane; k < b1; k += L) {
            const int64_t c = col[k];
            T xv[BC];
#pragma unroll
            for (int q = 0; q < BC; ++q) xv[q] = x[c * BC + q];
#pragma unroll
            for (int r = 0; r < BR; ++r)
#pragma unroll
                for (int q = 0; q < BC; ++q) acc[r] += val[(int64_t)k * (BR * BC) + r * BC + q] * xv[q];
        }
#pragma unroll
        for (int r = 0; r < BR; ++r) acc[r] = group_sum<L>(acc[r]);
        if (lane < BR) {
            T v = acc[0];
#pragma unroll
            for (int r = 1; r < BR; ++r)
                if (lane == r) v = acc[r];
            add_dots(dsum, ep(row * BR + lane, v));
        }
    }
    if (partial) {   // partial[d * nblocks + block]
        block_sum_n<NT, ND>(dsum);
        if (threadIdx.x == 0)
#pragma unroll
            for (int d = 0; d < ND; ++d) partial[d * nblocks + lb] = dsum[d];
    }
}

// 2x2-block SpMV on the sliced-ELL layout: one lane per row node, 64 rows per slice,
// block k of the slice's rows stored contiguously (values as one 16/32-byte vector
// per lane), no cross-lane reduction.
template <typename T> struct Vec4;
template <> struct Vec4<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct Vec4<_Float16> { typedef _Float16 type __attribute__((ext_vector_type(4))); };
template <typename T> struct Vec2;
template <> struct Vec2<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct Vec2<double> { typedef double type __attribute__((ext_vector_type(2))); };

// Gather of the x pair of a column node: plainly, or formed on the fly as s * dinv * b
// (the first Chebyshev step from a zero start, whose iterate is never stored).
template <typename T>
struct XPlain {
    const T *x;
    int xvec() const { return 1; }   // gathered vectors (host, for the byte count)
    __device__ __forceinline__ typename Vec2<T>::type operator()(int64_t c) const {
        return reinterpret_cast<const typename Vec2<T>::type *>(x)[c];
    }
};
template <typename T>
struct XJacobi {
    const T *b, *dinv; T s;
    int xvec() const { return 0; }   // b and D^-1: counted once, by the epilogue that reads them
    __device__ __forceinline__ typename Vec2<T>::type operator()(int64_t c) const {
        typedef typename Vec2<T>::type V2;
        const V2 bv = reinterpret_cast<const V2 *>(b)[c], dv = reinterpret_cast<const V2 *>(dinv)[c];
        return V2{s * dv.x * bv.x, s * dv.y * bv.y};
    }
};

// Slice of the wavefront: the same for its 64 lanes, so it is made a scalar (the slice
// pointer loads become scalar loads and the block loop is uniform).  Lanes past the last
// row run the loop on the zero padding of the last slice and skip the epilogue.
template <int NT>
__device__ __forceinline__ int64_t wave_slice() {
    return (int64_t)blockIdx.x * (NT / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
}

// VT is the storage type of the values (T, or _Float16 for the scaled smoother copy); XL
// gathers the x pair of a column node.
template <int NT, typename VT, typename T, typename Ep, typename XL = XPlain<T>>
__global__ __launch_bounds__(NT) void k_spmv_sell22(int64_t nrows, const int64_t *__restrict__ sptr,
                                                     const int32_t *__restrict__ col, const VT *__restrict__ val,
                                                     XL xl, Ep ep, double *__restrict__ partial, int64_t nblocks) {
    typedef typename Vec4<VT>::type V4;
    typedef typename Vec2<T>::type V2;
    constexpr int C = 64;
    static_assert(C == kWave, "one slice per wavefront");
    constexpr int ND = Ep::kDots;
    const int64_t lb = blockIdx.x;
    const int64_t slice = wave_slice<NT>();
    const int lane = threadIdx.x % C;
    const int64_t row = slice * C + lane;
    double dsum[ND] = {};
    if (slice * C < nrows) {
        const int64_t base = sptr[slice];
        const int len = (int)((sptr[slice + 1] - base) / C);
        const V4 *__restrict__ vp = reinterpret_cast<const V4 *>(val) + base + lane;
        const int32_t *__restrict__ cp = col + base + lane;
        T a0 = T(0), a1 = T(0);
        // values and columns are streamed once: non-temporal loads keep them from
        // evicting the gathered x lines (measured: 117 -> 102 us on K at 256 x 2048)
#pragma unroll 4
        for (int k = 0; k < len; ++k) {
            const V4 v = __builtin_nontemporal_load(&vp[(int64_t)k * C]);
            const V2 xv = xl(__builtin_nontemporal_load(&cp[(int64_t)k * C]));
            a0 += (T)v.x * xv.x + (T)v.y * xv.y;
            a1 += (T)v.z * xv.x + (T)v.w * xv.y;
        }
        if (row < nrows) {
            add_dots(dsum, ep(2 * row, a0));
            add_dots(dsum, ep(2 * row + 1, a1));
        }
    }
    if (partial) {   // partial[d * nblocks + block]
        block_sum_n<NT, ND>(dsum);
        if (threadIdx.x == 0)
#pragma unroll
            for (int d = 0; d < ND; ++d) partial[d * nblocks + lb] = dsum[d];
    }
}

// y = A1 x1 - A2 x2 for two 2x2-block operators with the same rows (the coarse right-hand side
// R b - (R A) x of a multigrid level in one pass), both on the sliced layout: one lane per row
// node, the wavefront's slice a scalar as in k_spmv_sell22.
template <int NT, typename T>
__global__ __launch_bounds__(NT) void k_spmv_dual_sell22(int64_t nrows, const int64_t *__restrict__ sptr1,
                                                          const int32_t *__restrict__ col1, const T *__restrict__ v1,
                                                          const T *__restrict__ x1, const int64_t *__restrict__ sptr2,
                                                          const int32_t *__restrict__ col2, const T *__restrict__ v2,
                                                          const T *__restrict__ x2, T *__restrict__ y) {
    typedef typename Vec4<T>::type V4;
    typedef typename Vec2<T>::type V2;
    constexpr int C = kWave;
    const int64_t slice = wave_slice<NT>();
    if (slice * C >= nrows) return;
    const int lane = threadIdx.x % C;
    const int64_t row = slice * C + lane;
    T a0 = T(0), a1 = T(0);
    {
        const int64_t base = sptr1[slice];
        const int len = (int)((sptr1[slice + 1] - base) / C);
        const V4 *__restrict__ vp = reinterpret_cast<const V4 *>(v1) + base + lane;
        const int32_t *__restrict__ cp = col1 + base + lane;
#pragma unroll 4
        for (int k = 0; k < len; ++k) {
            const V4 v = __builtin_nontemporal_load(&vp[(int64_t)k * C]);
            const V2 xv = reinterpret_cast<const V2 *>(x1)[__builtin_nontemporal_load(&cp[(int64_t)k * C])];
            a0 += v.x * xv.x + v.y * xv.y;
            a1 += v.z * xv.x + v.w * xv.y;
        }
    }
    {
        const int64_t base = sptr2[slice];
        const int len = (int)((sptr2[slice + 1] - base) / C);
        const V4 *__restrict__ vp = reinterpret_cast<const V4 *>(v2) + base + lane;
        const int32_t *__restrict__ cp = col2 + base + lane;
#pragma unroll 4
        for (int k = 0; k < len; ++k) {
            const V4 v = __builtin_nontemporal_load(&vp[(int64_t)k * C]);
            const V2 xv = reinterpret_cast<const V2 *>(x2)[__builtin_nontemporal_load(&cp[(int64_t)k * C])];
            a0 -= v.x * xv.x + v.y * xv.y;
            a1 -= v.z * xv.x + v.w * xv.y;
        }
    }
    if (row < nrows) reinterpret_cast<V2 *>(y)[row] = V2{a0, a1};
}

// The same on node-block CSR (the small levels): L lanes per row node, shuffle reduction.
template <int L, int NT, typename T>
__global__ __launch_bounds__(NT) void k_spmv_dual_block22(int64_t nrows, const int32_t *__restrict__ ptr1,
                                                           const int32_t *__restrict__ col1, const T *__restrict__ v1,
                                                           const T *__restrict__ x1, const int32_t *__restrict__ ptr2,
                                                           const int32_t *__restrict__ col2, const T *__restrict__ v2,
                                                           const T *__restrict__ x2, T *__restrict__ y) {
    typedef typename Vec4<T>::type V4;
    typedef typename Vec2<T>::type V2;
    const int64_t row = (int64_t)blockIdx.x * (NT / L) + threadIdx.x / L;
    const int lane = threadIdx.x % L;
    if (row >= nrows) return;   // whole L-lane groups leave together (NT / L rows per block)
    T a0 = T(0), a1 = T(0);
    for (int32_t k = ptr1[row] + lane; k < ptr1[row + 1]; k += L) {
        const V4 v = reinterpret_cast<const V4 *>(v1)[k];
        const V2 xv = reinterpret_cast<const V2 *>(x1)[col1[k]];
        a0 += v.x * xv.x + v.y * xv.y;
        a1 += v.z * xv.x + v.w * xv.y;
    }
    for (int32_t k = ptr2[row] + lane; k < ptr2[row + 1]; k += L) {
        const V4 v = reinterpret_cast<const V4 *>(v2)[k];
        const V2 xv = reinterpret_cast<const V2 *>(x2)[col2[k]];
        a0 -= v.x * xv.x + v.y * xv.y;
        a1 -= v.z * xv.x + v.w * xv.y;
    }
    a0 = group_sum<L>(a0);
    a1 = group_sum<L>(a1);
    if (lane == 0) reinterpret_cast<V2 *>(y)[row] = V2{a0, a1};
}

// out[q] = sum over the pairs (a, b) listed for stored block q of A[a] B[b] (2x2 block products,
// in list order): a sparse product C = A B on a pattern and pair lists built once per mesh.
template <typename T>
__global__ __launch_bounds__(256) void k_pair_product(int64_t npos, const int32_t *__restrict__ ptr,
                                                      const int32_t *__restrict__ pair, const T *__restrict__ A,
                                                      const T *__restrict__ B, T *__restrict__ out) {
    typedef typename Vec4<T>::type V4;
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= npos) return;
    T c0 = T(0), c1 = T(0), c2 = T(0), c3 = T(0);
    for (int32_t k = ptr[q]; k < ptr[q + 1]; ++k) {
        const V4 a = reinterpret_cast<const V4 *>(A)[pair[2 * (int64_t)k]];
        const V4 b = reinterpret_cast<const V4 *>(B)[pair[2 * (int64_t)k + 1]];
        c0 += a.x * b.x + a.y * b.z;
        c1 += a.x * b.y + a.y * b.w;
        c2 += a.z * b.x + a.w * b.z;
        c3 += a.z * b.y + a.w * b.w;
    }
    reinterpret_cast<V4 *>(out)[q] = V4{c0, c1, c2, c3};
}

// 2x1 / 1x2 block SpMV on the sliced layout (G: velocity node rows, pressure columns;
// G^T: pressure rows, velocity node columns): one lane per row, two values per block.
template <int BR, int BC, int NT, typename T, typename Ep>
__global__ __launch_bounds__(NT) void k_spmv_sell_rc(int64_t nrows, const int64_t *__restrict__ sptr,
                                                      const int32_t *__restrict__ col, const T *__restrict__ val,
                                                      const T *__restrict__ x, Ep ep, double *__restrict__ partial,
                                                      int64_t nblocks) {
    static_assert(BR * BC == 2, "2x1 or 1x2 blocks");
    typedef typename Vec2<T>::type V2;
    constexpr int C = 64;
    constexpr int ND = Ep::kDots;
    const int64_t lb = blockIdx.x;
    const int64_t slice = wave_slice<NT>();
    const int lane = threadIdx.x % C;
    const int64_t row = slice * C + lane;
    double dsum[ND] = {};
    if (slice * C < nrows) {
        const int64_t base = sptr[slice];
        const int len = (int)((sptr[slice + 1] - base) / C);
        const V2 *__restrict__ vp = reinterpret_cast<const V2 *>(val) + base + lane;
        const int32_t *__restrict__ cp = col + base + lane;
        T a0 = T(0), a1 = T(0);
#pragma unroll 4
        for (int k = 0; k < len; ++k) {
            const V2 v = __builtin_nontemporal_load(&vp[(int64_t)k * C]);
            const int64_t cl = __builtin_nontemporal_load(&cp[(int64_t)k * C]);
            if constexpr (BR == 2) {          // 2x1: both rows times one x value
                const T xv = x[cl];
                a0 += v.x * xv;
                a1 += v.y * xv;
            } else {                          // 1x2: one row times an x pair
                a0 += v.x * x[2 * cl] + v.y * x[2 * cl + 1];
            }
        }
        if (row < nrows) {
            if constexpr (BR == 2) {
                add_dots(dsum, ep(2 * row, a0));
                add_dots(dsum, ep(2 * row + 1, a1));
            } else {
                add_dots(dsum, ep(row, a0));
            }
        }
    }
    if (partial) {
        block_sum_n<NT, ND>(dsum);
        if (threadIdx.x == 0)
#pragma unroll
            for (int d = 0; d < ND; ++d) partial[d * nblocks + lb] = dsum[d];
    }
}

// Scalar SpMV on the sliced layout of a node pattern (one value per node pair, e.g. the
// P2 thermal operators): one lane per row, fused epilogue.
template <int NT, typename T, typename Ep>
__global__ __launch_bounds__(NT) void k_spmv_sell11(int64_t nrows, const int64_t *__restrict__ sptr,
                                                     const int32_t *__restrict__ col, const T *__restrict__ val,
                                                     const T *__restrict__ x, Ep ep, double *__restrict__ partial,
                                                     int64_t nblocks) {
    constexpr int C = 64;
    constexpr int ND = Ep::kDots;
    const int64_t lb = blockIdx.x;
    const int64_t slice = wave_slice<NT>();
    const int lane = threadIdx.x % C;
    const int64_t row = slice * C + lane;
    double dsum[ND] = {};
    if (slice * C < nrows) {
        const int64_t base = sptr[slice];
        const int len = (int)((sptr[slice + 1] - base) / C);
        const T *__restrict__ vp = val + base + lane;
        const int32_t *__restrict__ cp = col + base + lane;
        T a = T(0);
#pragma unroll 4
        for (int k = 0; k < len; ++k)
            a += __builtin_nontemporal_load(&vp[(int64_t)k * C]) * x[__builtin_nontemporal_load(&cp[(int64_t)k * C])];
        if (row < nrows) add_dots(dsum, ep(row, a));
    }
    if (partial) {
        block_sum_n<NT, ND>(dsum);
        if (threadIdx.x == 0)
#pragma unroll
            for (int d = 0; d < ND; ++d) partial[d * nblocks + lb] = dsum[d];
    }
}

// One colour of a multicolour block Gauss-Seidel sweep on the coloured sliced layout:
// x_q += D_q^-1 (b_q - (K x)_q) for the coloured rows q of slices [s0, s1), one lane per
// row (q = row0 + (s - s0) * 64 + lane, block k at sptr[s] + 64 k + lane, columns in
// coloured row order).  The iterate xq, right-hand side bq, near-null vector wq and the
// inverse diagonal blocks are stored in that order, so every per-row access is coalesced;
// rowid maps q to its node n (-1: padding row).  No two nodes of a colour are coupled, so
// the in-place update is race free.
// ZERO: forward sweep from x = 0: only the row's first nlow (lower-colour) blocks meet a
// non-zero x; b is read in node order and copied to bq.
// Otherwise: full sweep; the result is also written in node order to z.  ND = 2: also
// (b . x, w . x) of the updated rows into partial[d * pstride + block].
template <int NT, bool ZERO, int ND>
__global__ __launch_bounds__(NT) void k_gs_colour(int64_t s0, int64_t s1, int64_t row0, const int64_t *__restrict__ sptr,
                                                   const int32_t *__restrict__ col, const float *__restrict__ val,
                                                   const int32_t *__restrict__ rowid, const int32_t *__restrict__ slow,
                                                   const uint8_t *__restrict__ nlow, const float *__restrict__ dinv,
                                                   const float *__restrict__ bn, float *__restrict__ bq,
                                                   float *__restrict__ xq, float *__restrict__ z,
                                                   const float *__restrict__ wq, double *__restrict__ partial,
                                                   int64_t pstride) {
    typedef typename Vec4<float>::type V4;
    typedef typename Vec2<float>::type V2;
    constexpr int C = kWave;
    const int64_t slice = s0 + wave_slice<NT>();
    const int lane = threadIdx.x % C;
    double dsum[ND > 0 ? ND : 1] = {};
    if (slice < s1) {
        const int64_t base = sptr[slice];
        const int len = ZERO ? slow[slice] : (int)((sptr[slice + 1] - base) / C);
        const int64_t q = row0 + (slice - s0) * C + lane;
        const int32_t r = rowid[q];
        const int nl = ZERO ? (int)nlow[q] : len;
        const V4 *__restrict__ vp = reinterpret_cast<const V4 *>(val) + base + lane;
        const int32_t *__restrict__ cp = col + base + lane;
        V2 *__restrict__ x2 = reinterpret_cast<V2 *>(xq);
        float a0 = 0.0f, a1 = 0.0f;
#pragma unroll 4
        for (int k = 0; k < len; ++k) {
            if (!ZERO || k < nl) {
                const V4 v = __builtin_nontemporal_load(&vp[(int64_t)k * C]);
                const V2 xv = x2[__builtin_nontemporal_load(&cp[(int64_t)k * C])];
                a0 += v.x * xv.x + v.y * xv.y;
                a1 += v.z * xv.x + v.w * xv.y;
            }
        }
        if (r >= 0) {
            V2 bv;
            if (ZERO) {
                bv = reinterpret_cast<const V2 *>(bn)[r];
                reinterpret_cast<V2 *>(bq)[q] = bv;
            } else {
                bv = reinterpret_cast<const V2 *>(bq)[q];
            }
            const V4 di = reinterpret_cast<const V4 *>(dinv)[q];
            const float r0 = bv.x - a0, r1 = bv.y - a1;
            V2 xn{di.x * r0 + di.y * r1, di.z * r0 + di.w * r1};
            if (!ZERO) {
                const V2 xo = x2[q];
                xn.x += xo.x;
                xn.y += xo.y;
                reinterpret_cast<V2 *>(z)[r] = xn;
            }
            x2[q] = xn;
            if constexpr (ND == 2) {
                const V2 wv = reinterpret_cast<const V2 *>(wq)[q];
                dsum[0] = (double)bv.x * (double)xn.x + (double)bv.y * (double)xn.y;
                dsum[1] = (double)wv.x * (double)xn.x + (double)wv.y * (double)xn.y;
            }
        }
    }
    if constexpr (ND == 2) {
        block_sum_n<NT, 2>(dsum);
        if (threadIdx.x == 0) {
            partial[blockIdx.x] = dsum[0];
            partial[pstride + blockIdx.x] = dsum[1];
        }
    }
}

// ------------------------------------------------------------------ reductions
// f(i) returns a double (ND = 1) or Dots<ND>; it may also write (fused update + dot).
template <int NT, int ND, typename F>
__global__ __launch_bounds__(NT) void k_reduce_map(int64_t n, F f, double *__restrict__ partial) {
    double s[ND] = {};
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) add_dots(s, f(i));
    block_sum_n<NT, ND>(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < ND; ++d) partial[d * gridDim.x + blockIdx.x] = s[d];
}

// Sum ND columns of npart partials (partial[d * npart + i]) in a fixed order: out[d].
// The partials were just written by another kernel, so the loop is latency-bound: each
// thread keeps four independent loads per column in flight before it adds.
template <int NT, int ND>
__global__ __launch_bounds__(NT) void k_reduce_final(int64_t npart, const double *__restrict__ partial,
                                                     double *__restrict__ out) {
    double s[ND];
    for (int d = 0; d < ND; ++d) s[d] = 0.0;
    int64_t i = threadIdx.x;
    for (; i + 3 * NT < npart; i += 4 * NT) {
        double v[ND][4];
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int u = 0; u < 4; ++u) v[d][u] = partial[d * npart + i + u * NT];
#pragma unroll
        for (int d = 0; d < ND; ++d) s[d] += (v[d][0] + v[d][1]) + (v[d][2] + v[d][3]);
    }
    for (; i < npart; i += NT)
#pragma unroll
        for (int d = 0; d < ND; ++d) s[d] += partial[d * npart + i];
    block_sum_n<NT, ND>(s);
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < ND; ++d) out[d] = s[d];
}

// Coarsest-level solve x = Ainv b: one wavefront per row, coalesced over the row, in
// 16-byte vectors when the row length allows it (the matrix is streamed once: non-temporal).
template <typename T>
__global__ __launch_bounds__(256) void k_dense_matvec(int64_t n, const T *__restrict__ Ainv,
                                                      const T *__restrict__ b, T *__restrict__ x) {
    typedef typename Vec4<T>::type V4;
    const int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= n) return;
    const T *a = Ainv + row * n;
    T s = T(0);
    if (n % 4 == 0) {
        const V4 *a4 = reinterpret_cast<const V4 *>(a);
        const V4 *b4 = reinterpret_cast<const V4 *>(b);
        for (int64_t k = lane; k < n / 4; k += kWave) {
            const V4 av = __builtin_nontemporal_load(&a4[k]);
            const V4 bv = b4[k];
            s += av.x * bv.x + av.y * bv.y + av.z * bv.z + av.w * bv.w;
        }
    } else {
        for (int64_t k = lane; k < n; k += kWave) s += a[k] * b[k];
    }
    s = group_sum<kWave>(s);
    if (lane == 0) x[row] = s;
}

template <int NT, typename F>
__global__ __launch_bounds__(NT) void k_map(int64_t n, F f) {
    const int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (i < n) f(i);
}

}  // namespace stk
