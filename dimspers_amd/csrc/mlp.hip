// Per-element MLP of the learned local-Schur preconditioner (no counterpart in
// the reference): 5 -> 64 -> 64 -> 64 -> 6 with GELU, BF16 weights and BF16
// activations as MFMA operands, FP32 accumulation (gfx950's v_mfma_f32_32x32x16_bf16),
// FP32 master weights trained with Adam.
//
// One wavefront handles a tile of 32 samples.  A layer Y[32 x N] = H[32 x K] W[K x N]
// is a sum of 32x32x16 MFMA products: lane l feeds A[i][k..k+7] and B[k..k+7][j] with
// i = j = l % 32, k = 8 (l / 32) (+ 16 per k-step), each operand one 16-byte LDS load,
// and receives C[i][j] for j = l % 32, i = 8 (r / 4) + 4 (l / 32) + r % 4 in accumulator
// register r.  (Which k a lane's eight values stand for only has to agree between A and
// B: every k enters the sum once.)
//
// Inputs: (mean log eta, cos 2(phi - theta_c), sin 2(phi - theta_c), log delta,
// area / mean area); outputs: the log-Cholesky factor (log L00, L10, log L11, L20,
// L21, log L22) of the element's local Schur block divided by area / eta.
#include <cmath>
#include <cstdio>
#include <numeric>
#include <random>

#include "device_ops.h"
#include "mlp.h"
#include "solver.h"

namespace stk {

namespace {

typedef short v8i16 __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int NTW = 64;   // threads per workgroup (one wavefront)

// packed BF16 weight layouts (uint16 offsets; K of every product a multiple of 16)
// forward, transposed [out][in_pad16]: WT1 64x16, WT2 64x64, WT3 64x64, WT4 32x64 (rows >= 6 zero)
constexpr int tW1 = 0, tW2 = 1024, tW3 = 5120, tW4 = 9216, kWT = 11264;
// backward, natural [in][out_pad16]: WN4 64x16, WN3 64x64, WN2 64x64
constexpr int nW4 = 0, nW3 = 1024, nW2 = 5120, kWN = 9216;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// acc += A[32 x K] B[K x 32]: A rows i from sA (row stride lda), B column j = jbase + i
// from row (jbase + i) of sW (stride ldb, K contiguous).  K, lda, ldb multiples of 16 and the
// LDS arrays 16-byte aligned, so each lane's eight operands are one ds_read_b128.
__device__ __forceinline__ v16f mm_tile(const uint16_t *sA, int lda, const uint16_t *sW, int ldb, int K, int jbase,
                                        v16f acc) {
    const int l = threadIdx.x & 63;
    const int i = l & 31, kq = (l >> 5) * 8;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const v8i16 a = *reinterpret_cast<const v8i16 *>(sA + i * lda + k0 + kq);
        const v8i16 b = *reinterpret_cast<const v8i16 *>(sW + (jbase + i) * ldb + k0 + kq);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf16, a), __builtin_bit_cast(v8bf16, b),
                                                      acc, 0, 0, 0);
    }
    return acc;
}

__device__ __forceinline__ int out_row(int r) { return (r >> 2) * 8 + ((threadIdx.x & 63) >> 5) * 4 + (r & 3); }

struct FwdArgs {
    int64_t n;
    const float *X;          // n x 5 raw features
    const float *norm;       // xmu[5] xsd[5] ymu[6] ysd[6]
    const uint16_t *wt;      // kWT packed
    const float *w;          // FP32 master (biases)
    float *Y;                // n x 6 outputs (denormalised when `denorm`)
    int denorm;
    // training
    float *Xn;               // n x 8 normalised input
    float *Z1, *Z2, *Z3;     // n x 64 pre-activations
    const float *T;          // n x 6 raw targets
    float *dY;               // n x 8 loss gradient w.r.t. the normalised output
    double *lossp;           // per-workgroup partial loss
    double inv_count;        // 1 / (6 n)
};

__global__ __launch_bounds__(NTW) void k_mlp_forward(FwdArgs a) {
    __shared__ __align__(16) uint16_t sW[kWT];
    __shared__ __align__(16) uint16_t sH0[32 * 16];
    __shared__ __align__(16) uint16_t sH[2][32 * 64];
    __shared__ float sb[3 * 64 + 8];
    const int l = threadIdx.x;
    for (int k = l; k < kWT; k += NTW) sW[k] = a.wt[k];
    for (int k = l; k < 64; k += NTW) {
        sb[k] = a.w[mlp::oB1 + k];
        sb[64 + k] = a.w[mlp::oB2 + k];
        sb[128 + k] = a.w[mlp::oB3 + k];
    }
    if (l < 8) sb[192 + l] = l < 6 ? a.w[mlp::oB4 + l] : 0.0f;
    __syncthreads();
    double lsum = 0.0;
    const int64_t ntiles = (a.n + 31) / 32;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t s0 = t * 32;
        // normalised input (K padded to 16; Xn keeps 8 columns)
        for (int k = l; k < 32 * 16; k += NTW) {
            const int i = k / 16, f = k % 16;
            const int64_t s = s0 + i;
            float v = 0.0f;
            if (f < 5 && s < a.n) v = (a.X[5 * s + f] - a.norm[f]) / a.norm[5 + f];
            sH0[k] = f2bf(v);
            if (a.Xn && s < a.n && f < 8) a.Xn[8 * s + f] = v;
        }
        __syncthreads();
        const uint16_t *hin = sH0;
        int ldin = 16, K = 16;
        const int toff[3] = {tW1, tW2, tW3};
        float *Zs[3] = {a.Z1, a.Z2, a.Z3};
        for (int layer = 0; layer < 3; ++layer) {
            uint16_t *hout = sH[layer & 1];
            for (int nt = 0; nt < 2; ++nt) {
                v16f acc = {0};
                acc = mm_tile(hin, ldin, sW + toff[layer], K, K, nt * 32, acc);
                const int j = nt * 32 + (l & 31);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = out_row(r);
                    const float z = acc[r] + sb[layer * 64 + j];
                    const int64_t s = s0 + i;
                    if (Zs[layer] && s < a.n) Zs[layer][64 * s + j] = z;
                    hout[i * 64 + j] = f2bf(gelu(z));
                }
            }
            __syncthreads();
            hin = hout;
            ldin = 64;
            K = 64;
        }
        // output layer (6 of 32 columns used)
        v16f acc = {0};
        acc = mm_tile(hin, 64, sW + tW4, 64, 64, 0, acc);
        const int j = l & 31;
        if (j < 6) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = out_row(r);
                const int64_t s = s0 + i;
                if (s >= a.n) continue;
                const float y = acc[r] + sb[192 + j];
                if (a.dY) {
                    const float tn = (a.T[6 * s + j] - a.norm[10 + j]) / a.norm[16 + j];
                    const float d = y - tn;
                    lsum += (double)d * (double)d;
                    a.dY[8 * s + j] = (float)(2.0 * a.inv_count) * d;
                }
                a.Y[6 * s + j] = a.denorm ? y * a.norm[16 + j] + a.norm[10 + j] : y;
            }
        }
        __syncthreads();
    }
    if (a.lossp) {
        const double s = block_sum<NTW>(lsum);
        if (l == 0) a.lossp[blockIdx.x] = s * a.inv_count;
    }
}

struct BwdArgs {
    int64_t n;
    const uint16_t *wn;      // kWN packed
    const float *Z1, *Z2, *Z3, *dY;
    float *dZ1, *dZ2, *dZ3;  // n x 64
};

__global__ __launch_bounds__(NTW) void k_mlp_backward(BwdArgs a) {
    __shared__ __align__(16) uint16_t sW[kWN];
    __shared__ __align__(16) uint16_t sD0[32 * 16];
    __shared__ __align__(16) uint16_t sD[2][32 * 64];
    const int l = threadIdx.x;
    for (int k = l; k < kWN; k += NTW) sW[k] = a.wn[k];
    __syncthreads();
    const int64_t ntiles = (a.n + 31) / 32;
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int64_t s0 = t * 32;
        for (int k = l; k < 32 * 16; k += NTW) {
            const int i = k / 16, f = k % 16;
            const int64_t s = s0 + i;
            sD0[k] = f2bf((f < 6 && s < a.n) ? a.dY[8 * s + f] : 0.0f);
        }
        __syncthreads();
        const uint16_t *din = sD0;
        int ldin = 16, K = 16;
        const int woff[3] = {nW4, nW3, nW2};
        const float *Zs[3] = {a.Z3, a.Z2, a.Z1};
        float *dZs[3] = {a.dZ3, a.dZ2, a.dZ1};
        for (int layer = 0; layer < 3; ++layer) {
            uint16_t *dout = sD[layer & 1];
            for (int nt = 0; nt < 2; ++nt) {
                v16f acc = {0};
                acc = mm_tile(din, ldin, sW + woff[layer], K, K, nt * 32, acc);
                const int j = nt * 32 + (l & 31);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int i = out_row(r);
                    const int64_t s = s0 + i;
                    float dz = 0.0f;
                    if (s < a.n) {
                        dz = acc[r] * gelu_d(Zs[layer][64 * s + j]);
                        dZs[layer][64 * s + j] = dz;
                    }
                    dout[i * 64 + j] = f2bf(dz);
                }
            }
            __syncthreads();
            din = dout;
            ldin = 64;
            K = 64;
        }
    }
}

// dW[in][out] = sum_s H[s][in] dZ[s][out] over a slice of samples, one 32x32 tile per wave.
struct WgradArgs {
    int64_t n, slice;
    const float *Xn, *Z1, *Z2, *Z3;          // H0 = Xn (n x 8), H_l = gelu(Z_l)
    const float *dZ1, *dZ2, *dZ3, *dY;      // dY: n x 8
    float *part;                            // nslices x kNParam
};

__global__ __launch_bounds__(NTW) void k_mlp_wgrad(WgradArgs a) {
    // tile table: layer, in-tile, out-tile
    const int tiles[12][3] = {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1},
                              {2, 0, 0}, {2, 0, 1}, {2, 1, 0}, {2, 1, 1}, {3, 0, 0}, {3, 1, 0}};
    const int tile = blockIdx.x % 12;
    const int64_t slice = blockIdx.x / 12;
    const int layer = tiles[tile][0], ib = tiles[tile][1] * 32, ob = tiles[tile][2] * 32;
    const int nin = layer == 0 ? 5 : 64, nout = layer == 3 ? 6 : 64;
    const float *H = layer == 0 ? a.Xn : layer == 1 ? a.Z1 : layer == 2 ? a.Z2 : a.Z3;
    const int ldh = layer == 0 ? 8 : 64;
    const float *D = layer == 0 ? a.dZ1 : layer == 1 ? a.dZ2 : layer == 2 ? a.dZ3 : a.dY;
    const int ldd = layer == 3 ? 8 : 64;
    const int l = threadIdx.x;
    const int i = l & 31, kq = (l >> 5) * 8;
    const int64_t s_begin = slice * a.slice, s_end = std::min<int64_t>(a.n, s_begin + a.slice);
    v16f acc = {0};
    for (int64_t k0 = s_begin; k0 < s_end; k0 += 16) {
        v8i16 A, B;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t s = k0 + kq + q;
            float h = 0.0f, d = 0.0f;
            if (s < s_end) {
                if (ib + i < nin) {
                    const float zv = H[ldh * s + ib + i];
                    h = layer == 0 ? zv : gelu(zv);
                }
                if (ob + i < nout) d = D[ldd * s + ob + i];
            }
            A[q] = (short)f2bf(h);
            B[q] = (short)f2bf(d);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf16, A), __builtin_bit_cast(v8bf16, B),
                                                      acc, 0, 0, 0);
    }
    const int woff = layer == 0 ? mlp::oW1 : layer == 1 ? mlp::oW2 : layer == 2 ? mlp::oW3 : mlp::oW4;
    float *P = a.part + slice * mlp::kNParam;
    const int jo = ob + (l & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int ii = ib + out_row(r);
        if (ii < nin && jo < nout) P[woff + ii * nout + jo] = acc[r];
    }
}

// Biases: db[j] = sum_s dZ[s][j] into the slice-0 partial slots (other slices hold zero
// there).  One block per bias entry (3 * 64 + 6 blocks), fixed-order block reduction.
__global__ __launch_bounds__(256) void k_mlp_bgrad(int64_t n, const float *dZ1, const float *dZ2,
                                                   const float *dZ3, const float *dY, float *part) {
    const int j = blockIdx.x;
    const int layer = j / 64, jj = j % 64;
    const float *D = layer == 0 ? dZ1 : layer == 1 ? dZ2 : layer == 2 ? dZ3 : dY;
    const int ld = layer == 3 ? 8 : 64;
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += 256) s += D[ld * k + jj];
    s = block_sum<256>(s);
    const int boff = layer == 0 ? mlp::oB1 : layer == 1 ? mlp::oB2 : layer == 2 ? mlp::oB3 : mlp::oB4;
    if (threadIdx.x == 0) part[boff + jj] = (float)s;
}

__global__ void k_sum_slices(int64_t nslice, const float *part, float *grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= mlp::kNParam) return;
    float s = 0.0f;
    for (int64_t q = 0; q < nslice; ++q) s += part[q * mlp::kNParam + k];
    grad[k] = s;
}

__global__ void k_adam(float *w, float *m, float *v, const float *g, float lr, float b1t, float b2t) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= mlp::kNParam) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / (1.0f - b1t)) / (sqrtf(vk / (1.0f - b2t)) + eps);
}

// FP32 master -> packed BF16 layouts.
__global__ void k_pack(const float *w, uint16_t *wt, uint16_t *wn) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kWT) {
        float v = 0.0f;
        if (k < tW2) {                       // WT1 [64][16]
            const int j = k / 16, i = k % 16;
            if (i < 5) v = w[mlp::oW1 + i * 64 + j];
        } else if (k < tW3) {                // WT2 [64][64]
            const int q = k - tW2, j = q / 64, i = q % 64;
            v = w[mlp::oW2 + i * 64 + j];
        } else if (k < tW4) {                // WT3
            const int q = k - tW3, j = q / 64, i = q % 64;
            v = w[mlp::oW3 + i * 64 + j];
        } else {                             // WT4 [32][64]
            const int q = k - tW4, j = q / 64, i = q % 64;
            if (j < 6) v = w[mlp::oW4 + i * 6 + j];
        }
        wt[k] = f2bf(v);
    }
    if (k < kWN) {
        float v = 0.0f;
        if (k < nW3) {                       // WN4 [64][16]
            const int i = k / 16, j = k % 16;
            if (j < 6) v = w[mlp::oW4 + i * 6 + j];
        } else if (k < nW2) {                // WN3 [64][64]
            const int q = k - nW3;
            v = w[mlp::oW3 + q];
        } else {                             // WN2
            const int q = k - nW2;
            v = w[mlp::oW2 + q];
        }
        wn[k] = f2bf(v);
    }
}

__global__ void k_gather(int64_t nb, const int32_t *idx, const float *X, const float *T, float *Xb, float *Tb) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nb) return;
    const int64_t src = idx[s];
    for (int f = 0; f < 5; ++f) Xb[5 * s + f] = X[5 * src + f];
    for (int f = 0; f < 6; ++f) Tb[6 * s + f] = T[6 * src + f];
}

}  // namespace

// ------------------------------------------------------------------ host driver
void MLP::ensure(int64_t nmax, hipStream_t s) {
    if (!w) {
        w.alloc(mlp::kNParam); m.alloc(mlp::kNParam); v.alloc(mlp::kNParam); grad.alloc(mlp::kNParam);
        wt.alloc(kWT); wn.alloc(kWN); dnorm.alloc(22);
        m.zero(s); v.zero(s);
    }
    if (cap < nmax) {
        cap = nmax;
        Xn.alloc(8 * nmax); Z1.alloc(64 * nmax); Z2.alloc(64 * nmax); Z3.alloc(64 * nmax);
        dZ1.alloc(64 * nmax); dZ2.alloc(64 * nmax); dZ3.alloc(64 * nmax); dY.alloc(8 * nmax);
        Yb.alloc(6 * nmax);
        const int64_t ns = (nmax + kSlice - 1) / kSlice;
        part.alloc(ns * mlp::kNParam);
    }
}

void MLP::pack(hipStream_t s) {
    k_pack<<<dim3((kWT + 255) / 256), dim3(256), 0, s>>>(w.p, wt.p, wn.p);
    STK_HIP(hipGetLastError());
}

void MLP::init(uint64_t seed, hipStream_t s) {
    ensure(0, s);
    std::vector<float> h(mlp::kNParam, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int nin, int nout) {
        const float a = std::sqrt(6.0f / (float)(nin + nout));
        std::uniform_real_distribution<float> U(-a, a);
        for (int k = 0; k < nin * nout; ++k) h[off + k] = U(rng);
    };
    fill(mlp::oW1, 5, 64);
    fill(mlp::oW2, 64, 64);
    fill(mlp::oW3, 64, 64);
    fill(mlp::oW4, 64, 6);
    w.upload(h, s);
    m.zero(s);
    v.zero(s);
    for (int k = 0; k < 5; ++k) { norm[k] = 0.0f; norm[5 + k] = 1.0f; }
    for (int k = 0; k < 6; ++k) { norm[10 + k] = 0.0f; norm[16 + k] = 1.0f; }
    dnorm.upload(norm, 22, s);
    pack(s);
    step = 0;
    trained = false;
    STK_HIP(hipStreamSynchronize(s));
}

void MLP::forward(const float *X, int64_t n, float *Y, bool denorm, hipStream_t s) {
    FwdArgs a{};
    a.n = n; a.X = X; a.norm = dnorm.p; a.wt = wt.p; a.w = w.p; a.Y = Y; a.denorm = denorm ? 1 : 0;
    const int64_t ntiles = (n + 31) / 32;
    const int nb = (int)std::min<int64_t>(std::max<int64_t>(ntiles, 1), 4096);
    k_mlp_forward<<<dim3(nb), dim3(NTW), 0, s>>>(a);
    STK_HIP(hipGetLastError());
}

double MLP::gradient(const float *X, const float *T, int64_t n, hipStream_t s, DBuf<double> &lossbuf) {
    ensure(n, s);
    const int64_t ntiles = (n + 31) / 32;
    const int nb = (int)std::min<int64_t>(std::max<int64_t>(ntiles, 1), 1024);
    if ((int64_t)lossbuf.n < nb) lossbuf.alloc(nb);
    FwdArgs a{};
    a.n = n; a.X = X; a.norm = dnorm.p; a.wt = wt.p; a.w = w.p; a.Y = Yb.p; a.denorm = 0;
    a.Xn = Xn.p; a.Z1 = Z1.p; a.Z2 = Z2.p; a.Z3 = Z3.p; a.T = T; a.dY = dY.p; a.lossp = lossbuf.p;
    a.inv_count = 1.0 / (6.0 * (double)n);
    // padded rows of dY must be zero for the backward and gradient kernels
    STK_HIP(hipMemsetAsync(dY.p, 0, 8 * n * sizeof(float), s));
    k_mlp_forward<<<dim3(nb), dim3(NTW), 0, s>>>(a);
    STK_HIP(hipGetLastError());
    BwdArgs b{};
    b.n = n; b.wn = wn.p; b.Z1 = Z1.p; b.Z2 = Z2.p; b.Z3 = Z3.p; b.dY = dY.p;
    b.dZ1 = dZ1.p; b.dZ2 = dZ2.p; b.dZ3 = dZ3.p;
    k_mlp_backward<<<dim3(nb), dim3(NTW), 0, s>>>(b);
    STK_HIP(hipGetLastError());
    const int64_t nslice = (n + kSlice - 1) / kSlice;
    STK_HIP(hipMemsetAsync(part.p, 0, nslice * mlp::kNParam * sizeof(float), s));
    WgradArgs g{};
    g.n = n; g.slice = kSlice; g.Xn = Xn.p; g.Z1 = Z1.p; g.Z2 = Z2.p; g.Z3 = Z3.p;
    g.dZ1 = dZ1.p; g.dZ2 = dZ2.p; g.dZ3 = dZ3.p; g.dY = dY.p; g.part = part.p;
    k_mlp_wgrad<<<dim3((int)(12 * nslice)), dim3(NTW), 0, s>>>(g);
    STK_HIP(hipGetLastError());
    k_mlp_bgrad<<<dim3(3 * 64 + 6), dim3(256), 0, s>>>(n, dZ1.p, dZ2.p, dZ3.p, dY.p, part.p);
    STK_HIP(hipGetLastError());
    k_sum_slices<<<dim3((mlp::kNParam + 255) / 256), dim3(256), 0, s>>>(nslice, part.p, grad.p);
    STK_HIP(hipGetLastError());
    std::vector<double> lp(nb);
    lossbuf.download(lp.data(), nb, s);
    STK_HIP(hipStreamSynchronize(s));
    double loss = 0;
    for (double x : lp) loss += x;
    return loss;
}

void MLP::adam(float lr, hipStream_t s) {
    ++step;
    const float b1t = std::pow(0.9f, (float)step), b2t = std::pow(0.999f, (float)step);
    k_adam<<<dim3((mlp::kNParam + 255) / 256), dim3(256), 0, s>>>(w.p, m.p, v.p, grad.p, lr, b1t, b2t);
    STK_HIP(hipGetLastError());
    pack(s);
}

// Training on seeded snapshots of exact local Schur blocks (k_schur_samples, mode 1).
void mlp_train(Context &c, const stk_train_opts &o, stk_train_report &rep) {
    STK_CHECK(c.structured, STK_E_NOMESH, "the learned model needs a structured annulus (Grid_split1 or Grid)");
    STK_CHECK(c.have_fields, STK_E_STATE, "stk_set_fields must precede training");
    STK_CHECK(o.snapshots > 0 && o.per_snapshot > 0 && o.epochs > 0 && o.batch >= 32, STK_E_ARG, "bad training options");
    if (!c.mlp) c.mlp = new MLP();
    MLP &net = *c.mlp;
    EventPair ev_;
    hipEvent_t e0 = ev_.a, e1 = ev_.b;
    STK_HIP(hipEventRecord(e0, c.stream));
    net.init(o.seed, c.stream);
    const int64_t ns = (int64_t)o.snapshots * o.per_snapshot;
    DBuf<float> X(5 * ns), T(6 * ns);
    SchurSampleHost h;
    h.nsamp = ns; h.per_snap = o.per_snapshot; h.first_snap = 0; h.mode = 1; h.ratio_kind = o.ratio_kind;
    h.X = X.p; h.Tg = T.p;
    launch_schur_samples(c, h);
    // normalisation of features and targets from the training set
    std::vector<float> hx(5 * ns), ht(6 * ns);
    X.download(hx.data(), 5 * ns, c.stream);
    T.download(ht.data(), 6 * ns, c.stream);
    STK_HIP(hipStreamSynchronize(c.stream));
    auto stats = [&](const std::vector<float> &a, int d, int f, float *mu, float *sd) {
        double s1 = 0, s2 = 0;
        for (int64_t k = 0; k < ns; ++k) { const double v = a[d * k + f]; s1 += v; s2 += v * v; }
        const double m = s1 / ns, var = std::max(s2 / ns - m * m, 1e-12);
        *mu = (float)m;
        *sd = (float)std::sqrt(var);
    };
    for (int f = 0; f < 5; ++f) stats(hx, 5, f, &net.norm[f], &net.norm[5 + f]);
    for (int f = 0; f < 6; ++f) stats(ht, 6, f, &net.norm[10 + f], &net.norm[16 + f]);
    for (int64_t k = 0; k < ns; ++k)
        for (int f = 0; f < 6; ++f)
            if (!std::isfinite(ht[6 * k + f])) throw Error(STK_E_NAN, "non-finite local Schur target");
    net.dnorm.upload(net.norm, 22, c.stream);
    const int64_t B = o.batch;
    net.ensure(B, c.stream);
    DBuf<float> Xb(5 * B), Tb(6 * B);
    DBuf<int32_t> idx(ns);
    DBuf<double> lossbuf;
    std::vector<int32_t> perm(ns);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(o.seed ^ 0xA5A5A5A5ull);
    const int64_t steps_per_epoch = ns / B;
    const int64_t total = steps_per_epoch * o.epochs;
    double first = -1, last = 0;
    int64_t step = 0;
    for (int ep = 0; ep < o.epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        idx.upload(perm, c.stream);
        for (int64_t k = 0; k < steps_per_epoch; ++k, ++step) {
            k_gather<<<dim3((int)((B + 255) / 256)), dim3(256), 0, c.stream>>>(B, idx.p + k * B, X.p, T.p, Xb.p, Tb.p);
            STK_HIP(hipGetLastError());
            const double loss = net.gradient(Xb.p, Tb.p, B, c.stream, lossbuf);
            if (first < 0) first = loss;
            last = loss;
            const double frac = (double)step / (double)std::max<int64_t>(1, total - 1);
            const double lr = o.lr_final + 0.5 * (o.lr - o.lr_final) * (1.0 + std::cos(M_PI * frac));
            net.adam((float)lr, c.stream);
        }
    }
    // held-out loss on a fresh set of snapshots
    const int64_t nv = std::min<int64_t>(ns, 65536);
    DBuf<float> Xv(5 * nv), Tv(6 * nv);
    SchurSampleHost hv = h;
    hv.nsamp = nv; hv.first_snap = o.snapshots; hv.X = Xv.p; hv.Tg = Tv.p;
    launch_schur_samples(c, hv);
    const double vloss = net.gradient(Xv.p, Tv.p, nv, c.stream, lossbuf);
    STK_HIP(hipEventRecord(e1, c.stream));
    STK_HIP(hipEventSynchronize(e1));
    float ms = 0;
    STK_HIP(hipEventElapsedTime(&ms, e0, e1));
    net.trained = true;
    rep.initial_loss = first;
    rep.final_loss = last;
    rep.validation_loss = vloss;
    rep.samples = ns;
    rep.steps = (int32_t)total;
    rep.ms = ms;
}

// Inference on every element and assembly of the learned Schur operator S (pattern of M).
void learned_setup(Context &c) {
    STK_CHECK(c.mlp && c.mlp->trained, STK_E_STATE, "no trained model (stk_mlp_train or stk_mlp_params)");
    // the features and the canonical vertex order of the blocks follow the triangle types of a cell
    STK_CHECK(c.structured, STK_E_NOMESH, "the learned model needs a structured annulus (Grid_split1 or Grid)");
    STK_CHECK(c.have_fabric && c.have_fields, STK_E_STATE, "stk_set_fabric and stk_set_fields must precede");
    STK_CHECK(c.assembled && c.K64, STK_E_STATE, "assemble first");
    DBuf<float> X(5 * c.nel), Y(6 * c.nel);
    DBuf<double> sc(c.nel);
    SchurSampleHost h;
    h.nsamp = c.nel; h.mode = 0; h.X = X.p; h.scale = sc.p;
    launch_schur_samples(c, h);
    c.mlp->forward(X.p, c.nel, Y.p, true, c.stream);
    c.S64.alloc(c.MN.nnz());
    c.S64.zero(c.stream);
    launch_learned_scatter(c, Y.p, sc.p, c.S64.p);
    c.S32.alloc(c.MN.nnz());
    c.Sdinv64.alloc(c.nP);
    c.Sdinv32.alloc(c.nP);
    finish_learned_operator(c);
    STK_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace stk
