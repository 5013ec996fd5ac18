// surrogate.h -- the 1D-convolutional charge surrogate U -> Q (BF16 MFMA,
// FP32 accumulate) shared by its C ABI and the surrogate-seeded SCF.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "negf_internal.h"

namespace negf {

constexpr int SG_KW = 9;      // kernel width
constexpr int SG_HALO = 4;    // (SG_KW - 1) / 2
constexpr int SG_C = 64;      // hidden channels
constexpr int SG_CP = 32;     // padded channel count of the input and output layers
constexpr int SG_LAYERS = 6;  // conv layers: 3 -> 64, 4 x 64 -> 64, 64 -> 1
constexpr int SG_IN = 3;      // real input channels: U / 0.5 eV, edge distance, per-sample feature
constexpr int SG_TILE = 256;  // sites per workgroup of the convolution kernel

__host__ __device__ constexpr int sg_cin(int l) { return l == 0 ? SG_CP : SG_C; }
__host__ __device__ constexpr int sg_cout(int l) { return l == SG_LAYERS - 1 ? SG_CP : SG_C; }

struct Surrogate {
    int N = 0;
    double mean_Q = 0.0, scale_Q = 1.0;
    std::vector<int> woff, boff;  // offsets of each layer in the flat parameter vector
    int nparam = 0;
    DevBuf d_w, d_m, d_v, d_g, d_wb, d_wt, d_part, d_partb, d_loss, d_zero;
    DevBuf d_act[SG_LAYERS + 1];  // x0 .. x5 (bf16 [B][N][C]); q (float [B][N]) in d_act[SG_LAYERS]
    DevBuf d_dz[2];               // ping-pong gradient / tangent buffers (bf16 [B][N][64])
    DevBuf d_bU, d_bQ, d_bF, d_idx;  // training minibatch gather buffers
    int capB = 0;
    long long adam_t = 0;
    hipStream_t st = nullptr;

    ~Surrogate();
    int init(int N_, unsigned long long seed);
    int pack(hipStream_t s);
    int ensure_batch(int B);
    // activations of B samples; U double [B][N] (eV), feat double [B] or null
    int forward(int B, const double* dU, const double* dfeat, hipStream_t s);
    // Q = mean + scale * q of the last forward into double [B][N]
    int output(int B, double* dQ, hipStream_t s);
    // dQ/dU . V for the last forward (V, out double [B][N])
    int jvp(int B, const double* dV, double* dout, hipStream_t s);
    // normalised-MSE loss and its parameter gradient (into d_g) for the last forward
    int backward(int B, const double* dQt, double* loss, hipStream_t s);
    int adam(float lr, hipStream_t s);
};

}  // namespace negf

struct negf_surrogate {
    negf::Surrogate s;
};
