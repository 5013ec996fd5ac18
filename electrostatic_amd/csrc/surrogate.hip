// surrogate.hip -- 1D-convolutional surrogate U -> Q of a tube's NEGF
// charge, with hand-written BF16 MFMA (FP32 accumulate) inference, tangent
// and training kernels for gfx950.  It has no counterpart in the reference:
// it seeds and preconditions the batched Broyden SCF, while the exact NEGF
// computes every residual that decides convergence.
//
// Network (sites n = 0..N-1, zero padding at the chain ends):
//   x0[n] = (U_n / 0.5 eV, exp(-min(n, N-1-n) / 16), f)   3 channels, padded to 32
//   5 x   x = tanh(conv9(x) + b)                          64 channels
//   q[n]  = conv9(x) + b                                  1 channel, padded to 32
//   Q_n   = mean_Q + scale_Q q_n
// f is a per-sample scalar feature (e.g. the drain bias), 0 when unused.
//
// Layouts: activations bf16 [b][n][c] (channels fastest); weights bf16
// [co][tap][ci] with FP32 master copies for Adam; the transposed, flipped
// copy Wt[ci][tap][co] = W[co][8 - tap][ci] turns the data gradient into a
// forward convolution.
//
// Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16:
// M = 16 output channels per MFMA (two per wavefront), N = 16 sites per MFMA
// tile (16 tiles = 256 sites per workgroup), K = 9 taps x input channels, with
// the site tile and a 4-site halo on each side staged in LDS.  Operand layout: lane l
// holds A[i][k] with i = l % 16 and B[k][j] with j = l % 16 for the same
// eight k of its 16-lane group, and receives D[i][j] for i = 4 (l / 16) + v,
// v = 0..3, j = l % 16.  (Which eight k a lane group holds does not matter:
// A and B use the same assignment, so the sum over k is unchanged.)
#include "surrogate.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "negf_material.h"

namespace negf {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

enum SgEpi { EPI_TANH = 0, EPI_LINEAR_OUT = 1, EPI_DTANH = 2 };

__device__ __forceinline__ f4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f4& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Y = epilogue(conv9(X, W) [+ bias]) for one sample (blockIdx.y) and
// SG_TILE sites, as an implicit GEMM on v_mfma_f32_16x16x32_bf16
// (A = weights [co][ci], B = staged inputs [ci][site], D [co][site]).
// X bf16 [B][N][CIN], W bf16 [COUT][9][CIN]; COUT/32 wavefronts, each owning
// 32 output channels whose 9 x CIN/32 x 2 weight fragments stay in its
// registers across the tile's SG_TILE/16 site tiles, so that every input
// fragment read from LDS feeds two MFMAs (one per 16-channel block) and the
// weights are read once per workgroup.
//   EPI_TANH:       Y bf16 [B][N][COUT] = tanh(acc + b)
//   EPI_LINEAR_OUT: Yf float [B][N] = acc(co = 0) (+ b[0] if bias)
//   EPI_DTANH:      Y bf16 [B][N][COUT] = acc * (1 - xp^2), xp bf16 [B][N][COUT]
// tanh for a result rounded to BF16 (relative error below 1e-5, a
// thousandth of a BF16 ulp): tanh|x| = (1 - e) / (1 + e) with e = exp(-2|x|)
// from v_exp_f32 (2^y) and v_rcp_f32, and the odd Taylor series
// x - x^3/3 + 2x^5/15 below |x| = 1/16, where 1 - e cancels (truncation
// 17/315 x^7, below 4e-9 relative there).
__device__ __forceinline__ float sg_tanh(float x)
{
    const float a = fabsf(x);
    const float e = __builtin_amdgcn_exp2f(-2.8853900817779268f * a);  // -2 / ln 2
    const float big = (1.0f - e) * __builtin_amdgcn_rcpf(1.0f + e);
    const float x2 = a * a;
    const float small = a * (1.0f - x2 * (1.0f / 3.0f - x2 * (2.0f / 15.0f)));
    return copysignf(a < 0.0625f ? small : big, x);
}

template <int EPI>
__device__ __forceinline__ void sg_epilogue(const f4& acc, int cb, size_t site, int COUT, const float* bias,
                                            __bf16* Y, float* Yf, const __bf16* xp)
{
    if (EPI == EPI_TANH) {
        __attribute__((aligned(8))) __bf16 o[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = (__bf16)sg_tanh(acc[v] + bias[cb + v]);
        *reinterpret_cast<uint2*>(Y + site * COUT + cb) = *reinterpret_cast<const uint2*>(o);
    } else if (EPI == EPI_LINEAR_OUT) {
        if (cb == 0) Yf[site] = acc[0] + (bias ? bias[0] : 0.0f);
    } else {
        const uint2 raw = *reinterpret_cast<const uint2*>(xp + site * COUT + cb);
        const __bf16* xq = reinterpret_cast<const __bf16*>(&raw);
        __attribute__((aligned(8))) __bf16 o[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float y = (float)xq[v];
            o[v] = (__bf16)(acc[v] * (1.0f - y * y));
        }
        *reinterpret_cast<uint2*>(Y + site * COUT + cb) = *reinterpret_cast<const uint2*>(o);
    }
}

template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(COUT / 32 * 64) void sg_conv_kernel(int N, const __bf16* __restrict__ X,
                                                                 const __bf16* __restrict__ W,
                                                                 const float* __restrict__ bias,
                                                                 __bf16* __restrict__ Y, float* __restrict__ Yf,
                                                                 const __bf16* __restrict__ xp)
{
    constexpr int NT = COUT / 32 * 64;  // threads
    constexpr int ROWS = SG_TILE + 2 * SG_HALO;
    // LDS rows: 64 channels (eight 16-byte chunks) are stored with chunk c of
    // row r at chunk c ^ (r & 7), so that the 16 rows of an operand fragment
    // read 16 distinct bank groups; narrower rows are padded by 16 bytes
    constexpr bool SW = CIN == 64;
    constexpr int RS = SW ? CIN : CIN + 8;  // row stride (elements)
    constexpr int V8 = CIN / 8;
    constexpr int KC = CIN / 32;  // 32-channel k blocks
    __shared__ __attribute__((aligned(16))) __bf16 xs[ROWS * RS];
    const int b = blockIdx.y;
    const int n0 = blockIdx.x * SG_TILE;
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    // stage rows n0 - 4 .. n0 + SG_TILE + 3 of the sample, eight loads in flight per thread
    const __bf16* xb = X + (size_t)b * N * CIN;
    constexpr int TOT = ROWS * V8, PER = (TOT + NT - 1) / NT, CH = 8;
#pragma unroll
    for (int q0 = 0; q0 < PER; q0 += CH) {
        bf16x8 v[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int idx = threadIdx.x + (q0 + q) * NT;
            const int r = idx / V8, c8 = idx - r * V8;
            const int n = min(max(n0 - SG_HALO + r, 0), N - 1);  // clamped; rows outside the sample are zeroed below
            v[q] = *reinterpret_cast<const bf16x8*>(xb + (size_t)n * CIN + c8 * 8);
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int idx = threadIdx.x + (q0 + q) * NT;
            if (q0 + q < PER && idx < TOT) {
                const int r = idx / V8, c8 = idx - r * V8;
                const int n = n0 - SG_HALO + r;
                if (n < 0 || n >= N) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) v[q][e] = (__bf16)0.0f;
                }
                *reinterpret_cast<bf16x8*>(xs + r * RS + (SW ? (c8 ^ (r & 7)) : c8) * 8) = v[q];
            }
        }
    }
    // this wavefront's weight fragments (after the staging, whose buffers are dead now)
    bf16x8 a[2][SG_KW][KC];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const __bf16* wrow = W + (size_t)(w * 32 + h * 16 + i) * SG_KW * CIN + 8 * g;
#pragma unroll
        for (int tap = 0; tap < SG_KW; ++tap)
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) a[h][tap][kc] = *reinterpret_cast<const bf16x8*>(wrow + tap * CIN + kc * 32);
    }
    __syncthreads();
    // two site tiles at a time: four independent accumulators per lane
#pragma unroll 1
    for (int t0 = 0; t0 < SG_TILE / 16; t0 += 2) {
        f4 acc[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int u = 0; u < 2; ++u) acc[h][u] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < SG_KW; ++tap) {
#pragma unroll
            for (int kc = 0; kc < KC; ++kc) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    // site n0 + 16 (t0 + u) + i + tap - 4 is staged row 16 (t0 + u) + i + tap
                    const int row = 16 * (t0 + u) + i + tap, c8 = kc * 4 + g;
                    const bf16x8 bb =
                        *reinterpret_cast<const bf16x8*>(xs + row * RS + (SW ? (c8 ^ (row & 7)) : c8) * 8);
                    acc[0][u] = mfma_bf16(a[0][tap][kc], bb, acc[0][u]);
                    acc[1][u] = mfma_bf16(a[1][tap][kc], bb, acc[1][u]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int n = n0 + 16 * (t0 + u) + i;
            if (n >= N) continue;
            const size_t site = (size_t)b * N + n;
#pragma unroll
            for (int h = 0; h < 2; ++h) sg_epilogue<EPI>(acc[h][u], w * 32 + h * 16 + 4 * g, site, COUT, bias, Y, Yf, xp);
        }
    }
}

// Partial weight gradients of one tap over one chunk of (sample, site) rows:
//   P[k][co][tap][ci] = sum_{s in chunk k} dZ[s][co] X[s + tap - 4][ci]
// (X rows outside the sample read as zero), as an MFMA GEMM with M = co,
// N = ci, K = sites.  Both tiles are staged transposed ([c][site]) so that
// the eight consecutive sites of an operand fragment are contiguous.
// grid (9, chunks); COUT/16 wavefronts.
template <int CIN, int COUT>
__global__ __launch_bounds__(COUT / 16 * 64) void sg_wgrad_kernel(int N, long long total, int sites_per_chunk,
                                                                  const __bf16* __restrict__ dZ,
                                                                  const __bf16* __restrict__ X,
                                                                  float* __restrict__ P)
{
    constexpr int NT = CIN / 16;  // 16-column tiles per wave
    constexpr int KS = 32;        // sites per MFMA step
    constexpr int LS = KS + 8;
    __shared__ __attribute__((aligned(16))) __bf16 dzT[COUT * LS];
    __shared__ __attribute__((aligned(16))) __bf16 xT[CIN * LS];
    const int tap = blockIdx.x;
    const int chunk = blockIdx.y;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = lane & 15, g = lane >> 4;
    const long long s_begin = (long long)chunk * sites_per_chunk;
    const long long s_end = std::min<long long>(total, s_begin + sites_per_chunk);
    constexpr int NTHR = COUT / 16 * 64;
    f4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
    for (long long s0 = s_begin; s0 < s_end; s0 += KS) {
        __syncthreads();
        for (int q = threadIdx.x; q < KS * COUT; q += NTHR) {
            const int s = q / COUT, c = q - s * COUT;
            const long long gs = s0 + s;
            dzT[c * LS + s] = gs < s_end ? dZ[gs * COUT + c] : (__bf16)0.0f;
        }
        for (int q = threadIdx.x; q < KS * CIN; q += NTHR) {
            const int s = q / CIN, c = q - s * CIN;
            const long long gs = s0 + s;
            __bf16 v = (__bf16)0.0f;
            if (gs < s_end) {
                const long long bb = gs / N;
                const long long n = gs - bb * N + tap - SG_HALO;
                if (n >= 0 && n < N) v = X[(bb * N + n) * CIN + c];
            }
            xT[c * LS + s] = v;
        }
        __syncthreads();
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&dzT[(w * 16 + i) * LS + 8 * g]);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 bb = *reinterpret_cast<const bf16x8*>(&xT[(t * 16 + i) * LS + 8 * g]);
            acc[t] = mfma_bf16(a, bb, acc[t]);
        }
    }
    float* Pk = P + (size_t)chunk * COUT * SG_KW * CIN;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int co = w * 16 + 4 * g + v;
            const int ci = t * 16 + i;
            Pk[((size_t)co * SG_KW + tap) * CIN + ci] = acc[t][v];
        }
}

// Partial bias gradients: Pb[k][co] = sum over chunk k's rows of dZ[s][co].
__global__ void sg_bgrad_kernel(long long total, int sites_per_chunk, int cout, const __bf16* __restrict__ dZ,
                                float* __restrict__ Pb)
{
    const int chunk = blockIdx.x;
    const int co = threadIdx.x;
    if (co >= cout) return;
    const long long s0 = (long long)chunk * sites_per_chunk;
    const long long s1 = std::min<long long>(total, s0 + sites_per_chunk);
    float a = 0.f;
    for (long long s = s0; s < s1; ++s) a += (float)dZ[s * cout + co];
    Pb[(size_t)chunk * cout + co] = a;
}

// out[j] = sum_k P[k][j] in chunk order (deterministic).
__global__ void sg_reduce_kernel(int nchunks, int len, const float* __restrict__ P, float* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float a = 0.f;
    for (int k = 0; k < nchunks; ++k) a += P[(size_t)k * len + j];
    out[j] = a;
}

__global__ void sg_input_kernel(int N, long long total, const double* __restrict__ U,
                                const double* __restrict__ feat, __bf16* __restrict__ X0)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const long long b = q / N;
    const int n = (int)(q - b * N);
    __attribute__((aligned(16))) __bf16 x[SG_CP];
#pragma unroll
    for (int c = 0; c < SG_CP; ++c) x[c] = (__bf16)0.0f;
    x[0] = (__bf16)(float)(U[q] * 2.0);
    const int d = min(n, N - 1 - n);
    x[1] = (__bf16)expf(-(float)d * 0.0625f);
    x[2] = (__bf16)(float)(feat ? feat[b] : 0.0);
    bf16x8* dst = reinterpret_cast<bf16x8*>(X0 + q * SG_CP);
#pragma unroll
    for (int c8 = 0; c8 < SG_CP / 8; ++c8) dst[c8] = reinterpret_cast<const bf16x8*>(x)[c8];
}

// tangent of the input: channel 0 = 2 V (per eV), other channels 0
__global__ void sg_tangent_input_kernel(long long total, const double* __restrict__ V, __bf16* __restrict__ T0)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    __attribute__((aligned(16))) __bf16 x[SG_CP];
#pragma unroll
    for (int c = 0; c < SG_CP; ++c) x[c] = (__bf16)0.0f;
    x[0] = (__bf16)(float)(V[q] * 2.0);
    bf16x8* dst = reinterpret_cast<bf16x8*>(T0 + q * SG_CP);
#pragma unroll
    for (int c8 = 0; c8 < SG_CP / 8; ++c8) dst[c8] = reinterpret_cast<const bf16x8*>(x)[c8];
}

__global__ void sg_output_kernel(long long n, const float* __restrict__ q, double mean, double scale,
                                 double* __restrict__ Q)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Q[i] = mean + scale * (double)q[i];
}

// Mean squared error of the normalised output; dq = 2 (q - t) / n in
// channel 0 of the padded gradient buffer; one partial sum per block.
__global__ __launch_bounds__(256) void sg_loss_kernel(long long n, const float* __restrict__ q,
                                                      const double* __restrict__ Qt, double mean, double inv_scale,
                                                      __bf16* __restrict__ dq, float* __restrict__ part)
{
    __shared__ float red[256];
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float t = (float)((Qt[i] - mean) * inv_scale);
        const float d = q[i] - t;
        l = d * d;
        __attribute__((aligned(16))) __bf16 x[SG_CP];
#pragma unroll
        for (int c = 0; c < SG_CP; ++c) x[c] = (__bf16)0.0f;
        x[0] = (__bf16)(2.0f * d / (float)n);
        bf16x8* dst = reinterpret_cast<bf16x8*>(dq + i * SG_CP);
#pragma unroll
        for (int c8 = 0; c8 < SG_CP / 8; ++c8) dst[c8] = reinterpret_cast<const bf16x8*>(x)[c8];
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void sg_adam_kernel(int n, float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                               float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const float gj = gr[j];
    const float mj = b1 * m[j] + (1.f - b1) * gj;
    const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    w[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

// FP32 masters W [COUT][9][CIN] -> bf16 Wb (same layout, padding rows and
// columns forced to zero) and Wt [CIN][9][COUT] = W[co][8 - tap][ci].
__global__ void sg_pack_kernel(int cin, int cout, int cin_real, int cout_real, const float* __restrict__ W,
                               __bf16* __restrict__ Wb, __bf16* __restrict__ Wt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cin * cout * SG_KW) return;
    const int co = j / (SG_KW * cin);
    const int r = j - co * SG_KW * cin;
    const int tap = r / cin, ci = r - tap * cin;
    const float x = (co < cout_real && ci < cin_real) ? W[j] : 0.f;
    Wb[j] = (__bf16)x;
    Wt[((size_t)ci * SG_KW + (SG_KW - 1 - tap)) * cout + co] = (__bf16)x;
}

__global__ void sg_gather_kernel(int N, int B, const int* __restrict__ idx, const double* __restrict__ U,
                                 const double* __restrict__ Q, const double* __restrict__ F, double* __restrict__ bU,
                                 double* __restrict__ bQ, double* __restrict__ bF)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)N * B) return;
    const int b = (int)(q / N), n = (int)(q - (long long)b * N);
    const size_t src = (size_t)idx[b] * N + n;
    bU[q] = U[src];
    bQ[q] = Q[src];
    if (n == 0) bF[b] = F ? F[idx[b]] : 0.0;
}

static int real_cin(int l) { return l == 0 ? SG_IN : SG_C; }
static int real_cout(int l) { return l == SG_LAYERS - 1 ? 1 : SG_C; }

Surrogate::~Surrogate()
{
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
}

int Surrogate::init(int N_, unsigned long long seed)
{
    N = N_;
    int off = 0;
    for (int l = 0; l < SG_LAYERS; ++l) {
        woff.push_back(off);
        off += sg_cout(l) * SG_KW * sg_cin(l);
        boff.push_back(off);
        off += sg_cout(l);
    }
    nparam = off;
    // uniform initialisation with variance 1 / fan-in (real channels only)
    std::vector<float> w(nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < SG_LAYERS; ++l) {
        const float lim = std::sqrt(3.0f / (float)(real_cin(l) * SG_KW));
        std::uniform_real_distribution<float> dist(-lim, lim);
        for (int co = 0; co < real_cout(l); ++co)
            for (int tap = 0; tap < SG_KW; ++tap)
                for (int ci = 0; ci < real_cin(l); ++ci)
                    w[woff[l] + ((size_t)co * SG_KW + tap) * sg_cin(l) + ci] = dist(rng);
    }
    int rc;
    NEGF_HIP_TRY(hipStreamCreate(&st));
    for (DevBuf* b : {&d_w, &d_m, &d_v, &d_g})
        if ((rc = b->ensure(sizeof(float) * nparam))) return rc;
    if ((rc = d_wb.ensure(2 * (size_t)nparam))) return rc;
    if ((rc = d_wt.ensure(2 * (size_t)nparam))) return rc;
    std::vector<float> z(nparam, 0.f);
    NEGF_HIP_TRY(hipMemcpy(d_w.p, w.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
    NEGF_HIP_TRY(hipMemcpy(d_m.p, z.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
    NEGF_HIP_TRY(hipMemcpy(d_v.p, z.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
    NEGF_HIP_TRY(hipMemcpy(d_g.p, z.data(), sizeof(float) * nparam, hipMemcpyHostToDevice));
    if ((rc = pack(st))) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(st));
    return NEGF_OK;
}

int Surrogate::pack(hipStream_t s)
{
    for (int l = 0; l < SG_LAYERS; ++l) {
        const int cin = sg_cin(l), cout = sg_cout(l);
        const int n = cin * cout * SG_KW;
        hipLaunchKernelGGL(sg_pack_kernel, dim3((n + 255) / 256), dim3(256), 0, s, cin, cout, real_cin(l),
                           real_cout(l), d_w.as<float>() + woff[l], d_wb.as<__bf16>() + woff[l],
                           d_wt.as<__bf16>() + woff[l]);
    }
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

int Surrogate::ensure_batch(int B)
{
    if (B <= capB) return NEGF_OK;
    int rc;
    const size_t sites = (size_t)B * N;
    for (int l = 0; l < SG_LAYERS; ++l)
        if ((rc = d_act[l].ensure(2 * sites * sg_cin(l)))) return rc;
    if ((rc = d_act[SG_LAYERS].ensure(sizeof(float) * sites))) return rc;
    for (DevBuf* b : {&d_dz[0], &d_dz[1]})
        if ((rc = b->ensure(2 * sites * SG_C))) return rc;
    capB = B;
    return NEGF_OK;
}

template <int CIN, int COUT, int EPI>
static void conv(int N, int B, const __bf16* X, const __bf16* W, const float* bias, __bf16* Y, float* Yf,
                 const __bf16* xp, hipStream_t s)
{
    dim3 g((N + SG_TILE - 1) / SG_TILE, B);
    hipLaunchKernelGGL((sg_conv_kernel<CIN, COUT, EPI>), g, dim3(COUT / 32 * 64), 0, s, N, X, W, bias, Y, Yf, xp);
}

int Surrogate::forward(int B, const double* dU, const double* dfeat, hipStream_t s)
{
    int rc;
    if ((rc = ensure_batch(B))) return rc;
    const long long sites = (long long)B * N;
    hipLaunchKernelGGL(sg_input_kernel, dim3((unsigned)((sites + 255) / 256)), dim3(256), 0, s, N, sites, dU, dfeat,
                       d_act[0].as<__bf16>());
    const __bf16* wb = d_wb.as<__bf16>();
    const float* w = d_w.as<float>();
    conv<SG_CP, SG_C, EPI_TANH>(N, B, d_act[0].as<__bf16>(), wb + woff[0], w + boff[0], d_act[1].as<__bf16>(),
                                nullptr, nullptr, s);
    for (int l = 1; l < SG_LAYERS - 1; ++l)
        conv<SG_C, SG_C, EPI_TANH>(N, B, d_act[l].as<__bf16>(), wb + woff[l], w + boff[l],
                                   d_act[l + 1].as<__bf16>(), nullptr, nullptr, s);
    conv<SG_C, SG_CP, EPI_LINEAR_OUT>(N, B, d_act[SG_LAYERS - 1].as<__bf16>(), wb + woff[SG_LAYERS - 1],
                                      w + boff[SG_LAYERS - 1], nullptr, d_act[SG_LAYERS].as<float>(), nullptr, s);
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

int Surrogate::output(int B, double* dQ, hipStream_t s)
{
    const long long n = (long long)B * N;
    hipLaunchKernelGGL(sg_output_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                       d_act[SG_LAYERS].as<float>(), mean_Q, scale_Q, dQ);
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

// forward-mode tangent: t_{l+1} = conv(W_l, t_l) (1 - x_{l+1}^2), dq = conv(W_5, t_5)
int Surrogate::jvp(int B, const double* dV, double* dout, hipStream_t s)
{
    if (B > capB) {
        set_error("surrogate tangent needs a forward of at least %d samples first", B);
        return NEGF_EARG;
    }
    const long long sites = (long long)B * N;
    __bf16* t0 = d_dz[0].as<__bf16>();
    __bf16* t1 = d_dz[1].as<__bf16>();
    hipLaunchKernelGGL(sg_tangent_input_kernel, dim3((unsigned)((sites + 255) / 256)), dim3(256), 0, s, sites, dV,
                       t0);
    const __bf16* wb = d_wb.as<__bf16>();
    conv<SG_CP, SG_C, EPI_DTANH>(N, B, t0, wb + woff[0], nullptr, t1, nullptr, d_act[1].as<__bf16>(), s);
    for (int l = 1; l < SG_LAYERS - 1; ++l) {
        conv<SG_C, SG_C, EPI_DTANH>(N, B, t1, wb + woff[l], nullptr, t0, nullptr, d_act[l + 1].as<__bf16>(), s);
        std::swap(t0, t1);
    }
    float* q = d_act[SG_LAYERS].as<float>();  // the forward output is consumed; the tangent reuses it
    conv<SG_C, SG_CP, EPI_LINEAR_OUT>(N, B, t1, wb + woff[SG_LAYERS - 1], nullptr, nullptr, q, nullptr, s);
    hipLaunchKernelGGL(sg_output_kernel, dim3((unsigned)((sites + 255) / 256)), dim3(256), 0, s, sites, q, 0.0,
                       scale_Q, dout);
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

int Surrogate::backward(int B, const double* dQt, double* loss, hipStream_t s)
{
    int rc;
    const long long sites = (long long)B * N;
    const int lblocks = (int)((sites + 255) / 256);
    if ((rc = d_loss.ensure(sizeof(float) * lblocks))) return rc;
    hipLaunchKernelGGL(sg_loss_kernel, dim3(lblocks), dim3(256), 0, s, sites, d_act[SG_LAYERS].as<float>(), dQt,
                       mean_Q, 1.0 / scale_Q, d_dz[0].as<__bf16>(), d_loss.as<float>());
    const int spc = 2048;  // rows per gradient chunk
    const int nchunks = (int)((sites + spc - 1) / spc);
    if ((rc = d_part.ensure(sizeof(float) * (size_t)nchunks * SG_C * SG_KW * SG_C))) return rc;
    if ((rc = d_partb.ensure(sizeof(float) * (size_t)nchunks * SG_C))) return rc;
    int cur = 0;
    for (int l = SG_LAYERS - 1; l >= 0; --l) {
        const int cin = sg_cin(l), cout = sg_cout(l);
        const __bf16* dz = d_dz[cur].as<__bf16>();
        const __bf16* x = d_act[l].as<__bf16>();
        if (cin == SG_C && cout == SG_C)
            hipLaunchKernelGGL((sg_wgrad_kernel<SG_C, SG_C>), dim3(SG_KW, nchunks), dim3(256), 0, s, N, sites, spc,
                               dz, x, d_part.as<float>());
        else if (cin == SG_C)
            hipLaunchKernelGGL((sg_wgrad_kernel<SG_C, SG_CP>), dim3(SG_KW, nchunks), dim3(128), 0, s, N, sites, spc,
                               dz, x, d_part.as<float>());
        else
            hipLaunchKernelGGL((sg_wgrad_kernel<SG_CP, SG_C>), dim3(SG_KW, nchunks), dim3(256), 0, s, N, sites, spc,
                               dz, x, d_part.as<float>());
        const int lw = cout * SG_KW * cin;
        hipLaunchKernelGGL(sg_reduce_kernel, dim3((lw + 255) / 256), dim3(256), 0, s, nchunks, lw,
                           d_part.as<float>(), d_g.as<float>() + woff[l]);
        hipLaunchKernelGGL(sg_bgrad_kernel, dim3(nchunks), dim3(64), 0, s, sites, spc, cout, dz, d_partb.as<float>());
        hipLaunchKernelGGL(sg_reduce_kernel, dim3(1), dim3(64), 0, s, nchunks, cout, d_partb.as<float>(),
                           d_g.as<float>() + boff[l]);
        if (l == 0) break;
        // data gradient into layer l's input, through the tanh that made it
        const __bf16* wt = d_wt.as<__bf16>() + woff[l];
        __bf16* dzn = d_dz[1 - cur].as<__bf16>();
        if (cout == SG_C)
            conv<SG_C, SG_C, EPI_DTANH>(N, B, dz, wt, nullptr, dzn, nullptr, x, s);
        else
            conv<SG_CP, SG_C, EPI_DTANH>(N, B, dz, wt, nullptr, dzn, nullptr, x, s);
        cur = 1 - cur;
    }
    NEGF_HIP_TRY(hipGetLastError());
    if (loss) {
        std::vector<float> h(lblocks);
        NEGF_HIP_TRY(hipMemcpyAsync(h.data(), d_loss.p, sizeof(float) * lblocks, hipMemcpyDeviceToHost, s));
        NEGF_HIP_TRY(hipStreamSynchronize(s));
        double a = 0.0;
        for (float v : h) a += v;
        *loss = a / (double)sites;
    }
    return NEGF_OK;
}

int Surrogate::adam(float lr, hipStream_t s)
{
    ++adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)adam_t), c2 = 1.f - std::pow(b2, (float)adam_t);
    hipLaunchKernelGGL(sg_adam_kernel, dim3((nparam + 255) / 256), dim3(256), 0, s, nparam, d_w.as<float>(),
                       d_g.as<float>(), d_m.as<float>(), d_v.as<float>(), lr, b1, b2, 1e-8f, c1, c2);
    NEGF_HIP_TRY(hipGetLastError());
    return pack(s);
}

}  // namespace negf

using namespace negf;

extern "C" {

int negf_surrogate_create(int N, unsigned long long seed, negf_surrogate** out)
{
    clear_error();
    if (N < 16 || N > NEGF_MAX_SITES || !out) {
        set_error("invalid surrogate arguments (N = %d)", N);
        return NEGF_EARG;
    }
    negf_surrogate* h = new negf_surrogate;
    const int rc = h->s.init(N, seed);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_surrogate_destroy(negf_surrogate* s)
{
    if (!s) return NEGF_EARG;
    delete s;
    return NEGF_OK;
}

int negf_surrogate_num_params(const negf_surrogate* s) { return s ? s->s.nparam : -1; }

int negf_surrogate_num_sites(const negf_surrogate* s) { return s ? s->s.N : -1; }

int negf_surrogate_get_params(negf_surrogate* s, float* w, double* norm)
{
    clear_error();
    if (!s) {
        set_error("null surrogate");
        return NEGF_EARG;
    }
    NEGF_HIP_TRY(hipStreamSynchronize(s->s.st));
    if (w) NEGF_HIP_TRY(hipMemcpy(w, s->s.d_w.p, sizeof(float) * s->s.nparam, hipMemcpyDeviceToHost));
    if (norm) {
        norm[0] = s->s.mean_Q;
        norm[1] = s->s.scale_Q;
    }
    return NEGF_OK;
}

int negf_surrogate_set_params(negf_surrogate* s, const float* w, const double* norm)
{
    clear_error();
    if (!s) {
        set_error("null surrogate");
        return NEGF_EARG;
    }
    if (norm && !(norm[1] > 0.0)) {
        set_error("surrogate output scale must be positive");
        return NEGF_EARG;
    }
    NEGF_HIP_TRY(hipStreamSynchronize(s->s.st));
    if (w) NEGF_HIP_TRY(hipMemcpy(s->s.d_w.p, w, sizeof(float) * s->s.nparam, hipMemcpyHostToDevice));
    if (norm) {
        s->s.mean_Q = norm[0];
        s->s.scale_Q = norm[1];
    }
    const int rc = s->s.pack(s->s.st);
    if (rc) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(s->s.st));
    return NEGF_OK;
}

int negf_surrogate_forward(negf_surrogate* s, int B, const double* d_U, const double* d_feat, double* d_Q,
                           void* stream)
{
    clear_error();
    if (!s || B <= 0 || !d_U || !d_Q) {
        set_error("invalid surrogate forward arguments");
        return NEGF_EARG;
    }
    hipStream_t st = stream ? (hipStream_t)stream : s->s.st;
    int rc = s->s.forward(B, d_U, d_feat, st);
    if (!rc) rc = s->s.output(B, d_Q, st);
    return rc;
}

int negf_surrogate_jvp(negf_surrogate* s, int B, const double* d_U, const double* d_feat, const double* d_V,
                       double* d_out, void* stream)
{
    clear_error();
    if (!s || B <= 0 || !d_U || !d_V || !d_out) {
        set_error("invalid surrogate tangent arguments");
        return NEGF_EARG;
    }
    hipStream_t st = stream ? (hipStream_t)stream : s->s.st;
    int rc = s->s.forward(B, d_U, d_feat, st);
    if (!rc) rc = s->s.jvp(B, d_V, d_out, st);
    return rc;
}

int negf_surrogate_gradient(negf_surrogate* s, int B, const double* d_U, const double* d_feat, const double* d_Q,
                            float* grad, double* loss)
{
    clear_error();
    if (!s || B <= 0 || !d_U || !d_Q) {
        set_error("invalid surrogate gradient arguments");
        return NEGF_EARG;
    }
    int rc = s->s.forward(B, d_U, d_feat, s->s.st);
    if (!rc) rc = s->s.backward(B, d_Q, loss, s->s.st);
    if (rc) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(s->s.st));
    if (grad) NEGF_HIP_TRY(hipMemcpy(grad, s->s.d_g.p, sizeof(float) * s->s.nparam, hipMemcpyDeviceToHost));
    return NEGF_OK;
}

int negf_surrogate_fit_output_scale(negf_surrogate* s, int nsamples, const double* d_Q)
{
    clear_error();
    if (!s || nsamples <= 0 || !d_Q) {
        set_error("invalid surrogate scale arguments");
        return NEGF_EARG;
    }
    std::vector<double> h((size_t)nsamples * s->s.N);
    NEGF_HIP_TRY(hipMemcpy(h.data(), d_Q, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    double a = 0.0;
    for (double v : h) a += v;
    const double mean = a / (double)h.size();
    double q = 0.0;
    for (double v : h) q += (v - mean) * (v - mean);
    s->s.mean_Q = mean;
    s->s.scale_Q = std::sqrt(q / (double)h.size());
    if (!(s->s.scale_Q > 0.0)) s->s.scale_Q = 1.0;
    return NEGF_OK;
}

// Adam on minibatches of nsamples device pairs (U, Q) (double [S][N]) with
// optional per-sample features (double [S]); seeded shuffle per epoch;
// learning rate linear from lr to lr_end.
int negf_surrogate_train(negf_surrogate* s, int nsamples, const double* d_U, const double* d_feat, const double* d_Q,
                         int epochs, int batch, double lr, double lr_end, unsigned long long seed, double* last_loss)
{
    clear_error();
    if (!s || nsamples <= 0 || !d_U || !d_Q || epochs <= 0 || batch <= 0 || batch > nsamples) {
        set_error("invalid surrogate training arguments");
        return NEGF_EARG;
    }
    Surrogate& m = s->s;
    const int N = m.N;
    int rc;
    if ((rc = m.d_bU.ensure(sizeof(double) * (size_t)batch * N))) return rc;
    if ((rc = m.d_bQ.ensure(sizeof(double) * (size_t)batch * N))) return rc;
    if ((rc = m.d_bF.ensure(sizeof(double) * (size_t)batch))) return rc;
    const int steps_per_epoch = nsamples / batch;
    if ((rc = m.d_idx.ensure(sizeof(int) * (size_t)steps_per_epoch * batch))) return rc;
    std::vector<int> order(nsamples);
    for (int i = 0; i < nsamples; ++i) order[i] = i;
    std::mt19937_64 rng(seed);
    const long long total_steps = (long long)steps_per_epoch * epochs;
    long long step = 0;
    double l = 0.0;
    const long long bs = (long long)batch * N;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        NEGF_HIP_TRY(hipMemcpyAsync(m.d_idx.p, order.data(), sizeof(int) * (size_t)steps_per_epoch * batch,
                                    hipMemcpyHostToDevice, m.st));
        for (int k = 0; k < steps_per_epoch; ++k) {
            hipLaunchKernelGGL(sg_gather_kernel, dim3((unsigned)((bs + 255) / 256)), dim3(256), 0, m.st, N, batch,
                               m.d_idx.as<int>() + (size_t)k * batch, d_U, d_Q, d_feat, m.d_bU.as<double>(),
                               m.d_bQ.as<double>(), m.d_bF.as<double>());
            NEGF_HIP_TRY(hipGetLastError());
            if ((rc = m.forward(batch, m.d_bU.as<double>(), d_feat ? m.d_bF.as<double>() : nullptr, m.st))) return rc;
            const bool want = (k == steps_per_epoch - 1) && last_loss;
            if ((rc = m.backward(batch, m.d_bQ.as<double>(), want ? &l : nullptr, m.st))) return rc;
            const double frac = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 1.0;
            if ((rc = m.adam((float)(lr + (lr_end - lr) * frac), m.st))) return rc;
            ++step;
        }
        // the shuffled order is read by the gathers just queued
        NEGF_HIP_TRY(hipStreamSynchronize(m.st));
        if (last_loss) *last_loss = l;
    }
    NEGF_HIP_TRY(hipStreamSynchronize(m.st));
    return NEGF_OK;
}

}  // extern "C"
