// The learned local-Schur model: parameter layout and host-side state.
#pragma once

#include "stk_internal.h"

namespace stk {

namespace mlp {
// FP32 master parameters, weights stored [in][out] row-major.
constexpr int kIn = 5, kH = 64, kOut = 6;
constexpr int oW1 = 0;
constexpr int oB1 = oW1 + kIn * kH;
constexpr int oW2 = oB1 + kH;
constexpr int oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH;
constexpr int oB3 = oW3 + kH * kH;
constexpr int oW4 = oB3 + kH;
constexpr int oB4 = oW4 + kH * kOut;
constexpr int kNParam = oB4 + kOut;   // 9094
}  // namespace mlp

struct MLP {
    static constexpr int64_t kSlice = 256;   // samples per weight-gradient slice
    DBuf<float> w, m, v, grad;               // master weights, Adam moments, gradient
    DBuf<uint16_t> wt, wn;                   // packed BF16 (forward / backward layouts)
    DBuf<float> dnorm;                       // feature and target normalisation on the device
    float norm[22] = {0};                    // xmu[5] xsd[5] ymu[6] ysd[6]
    // training work space
    int64_t cap = 0;
    DBuf<float> Xn, Z1, Z2, Z3, dZ1, dZ2, dZ3, dY, Yb, part;
    int64_t step = 0;
    bool trained = false;

    void ensure(int64_t nmax, hipStream_t s);
    void init(uint64_t seed, hipStream_t s);
    void pack(hipStream_t s);
    void forward(const float *X, int64_t n, float *Y, bool denorm, hipStream_t s);
    double gradient(const float *X, const float *T, int64_t n, hipStream_t s, DBuf<double> &lossbuf);
    void adam(float lr, hipStream_t s);
};

void mlp_train(Context &c, const stk_train_opts &o, stk_train_report &rep);
void learned_setup(Context &c);
void finish_learned_operator(Context &c);

}  // namespace stk
