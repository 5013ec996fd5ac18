// negf_scf.hip -- Broyden modified-second update on the device and the
// self-consistent loop of one tube with a dense electrostatic stand-in.
//
// Broyden (Execute_Broyden_Modified_Second_Algorithm_Parallel
// Broyden_Second_Parallel.cpp:213-486, serial order :9-211): one workgroup of
// 1024 threads runs a whole update.  Every reduction has a fixed order
// (lane-strided partial sums, a fixed butterfly per wavefront, waves summed
// in index order), so results are identical from run to run.  The history
// dot products of one update are spread over the 16 wavefronts, one history
// column per wavefront at a time, so an update costs two block barriers per
// phase rather than two per history column.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "negf_internal.h"
#include "negf_material.h"
#include "surrogate.h"

namespace negf {

static constexpr int kBroydenThreads = 1024;
static constexpr int kWaves = kBroydenThreads / 64;

struct BroydenDev {
    double* n_in;
    double* n_prev;
    double* F;
    double* dF;
    double* Norm;
    double* sum_vec;
    double* V;          // [max_step + 1][S]: dF / |dF|^2 per update
    double* W;          // [max_step + 1][S]
    double* intermed;   // [max_step + 1] history dots V_it . dF past the LDS copy
    double* intermed2;  // [max_step + 1] history dots V_it . F past the LDS copy
    double* out;        // [0] max norm, [1] l2 norm, [2] |dF|^2
    double* PFp;        // previous preconditioned residual (preconditioned update only)
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}

// The history's dot products in one pass over V: a_it = V_it . dF and
// b_it = V_it . F for it in [1, m_hi], each summed lane-strided and then
// across the wavefront; wave w takes it = 1 + w, 1 + w + kWaves, ...  Entries
// below kHistLds go to LDS, the rest to intermed (a) and intermed2 (b).
static constexpr int kHistLds = 512;

__device__ void history_dots2(const BroydenDev& b, int S, int m_hi, int wave, int lane, double* ha, double* hb)
{
    for (int it = 1 + wave; it <= m_hi; it += kWaves) {
        const double* __restrict__ v = b.V + (size_t)it * S;
        double a = 0.0, c = 0.0;
#pragma unroll 4
        for (int s = lane; s < S; s += 64) {
            const double vs = v[s];
            a = fma(vs, b.dF[s], a);
            c = fma(vs, b.F[s], c);
        }
        a = wave_sum(a);
        c = wave_sum(c);
        if (lane == 0) {
            if (it < kHistLds) {
                ha[it] = a;
                hb[it] = c;
            } else {
                b.intermed[it] = a;
                b.intermed2[it] = c;
            }
        }
    }
}

// One Broyden update of one system by one workgroup of kBroydenThreads.
// PC: the mixing term fraction * F is replaced by a preconditioned residual
// PF = P F (P linear and fixed between history resets), and fraction * dF by
// PF - PF_previous; everything else is the reference update.
//
// The reference (Broyden_Parallel_General.cpp:106-170) forms
// sum_vec = sum_{it<=m} W_it (V_it . F) after appending V_m, W_m, which needs
// V_it . dF first (for W_m) and a second pass over V for V_it . F.  Both dot
// products of the old rows come from one pass here, and the new row's
// V_m . F is summed while V_m is written; the sums over it keep the
// reference's order (it = 1..m, the new row last).
template <bool PC>
__device__ void broyden_update(int S, int m, double fraction, int norm_type, const double* __restrict__ n_out,
                               const double* __restrict__ PF, const BroydenDev& b)
{
    __shared__ double red[3][kWaves];
    __shared__ double ha[kHistLds], hb[kHistLds];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;

    // norms and F, dF (:31-88)
    double nsum = 0.0, nmax = 0.0, den = 0.0;
    for (int s = tid; s < S; s += kBroydenThreads) {
        const double Fc = b.n_in[s] - n_out[s];
        double nrm;
        if (norm_type == 0) {
            nrm = fabs(Fc);
            nsum = fma(Fc, Fc, nsum);
        } else {
            nrm = fabs(Fc / (b.n_in[s] + n_out[s]));
            nsum = fma(nrm, nrm, nsum);
        }
        b.Norm[s] = nrm;
        nmax = fmax(nmax, nrm);
        const double d = Fc - b.F[s];
        b.dF[s] = d;
        b.F[s] = Fc;
        den = fma(d, d, den);
    }
    nsum = wave_sum(nsum);
    nmax = wave_max(nmax);
    den = wave_sum(den);
    if (lane == 0) {
        red[0][wave] = nsum;
        red[1][wave] = nmax;
        red[2][wave] = den;
    }
    __syncthreads();
    double norm_sum = 0.0, bnorm = 0.0, denom = 0.0;
    for (int w = 0; w < kWaves; ++w) {
        norm_sum += red[0][w];
        bnorm = fmax(bnorm, red[1][w]);
        denom += red[2][w];
    }
    if (tid == 0) {
        b.out[0] = bnorm;
        b.out[1] = sqrt(norm_sum);
        b.out[2] = denom;
    }

    double bm = 0.0;   // V_m . F
    if (m > 0) {
        history_dots2(b, S, m - 1, wave, lane, ha, hb);
        __syncthreads();
        // W_m = -fraction dF + (n_in - n_prev) - sum_{it<m} W_it (V_it . dF)
        // (:106-135); the same W_it loads give the old rows' part of sum_vec
        double vf = 0.0;
        for (int s = tid; s < S; s += kBroydenThreads) {
            double sa = 0.0, sb = 0.0;
            const int mL = min(m - 1, kHistLds - 1);
            int it = 1;
#pragma unroll 4
            for (; it <= mL; ++it) {
                const double w = b.W[(size_t)it * S + s];
                sa = fma(w, ha[it], sa);
                sb = fma(w, hb[it], sb);
            }
            for (; it <= m - 1; ++it) {
                const double w = b.W[(size_t)it * S + s];
                sa = fma(w, b.intermed[it], sa);
                sb = fma(w, b.intermed2[it], sb);
            }
            const double d = b.dF[s];
            const double pd = PC ? PF[s] - b.PFp[s] : fraction * d;
            const double vm = d / denom;
            b.V[(size_t)m * S + s] = vm;
            b.W[(size_t)m * S + s] = -pd + (b.n_in[s] - b.n_prev[s]) - sa;
            b.sum_vec[s] = sb;
            vf = fma(vm, b.F[s], vf);
        }
        vf = wave_sum(vf);
        if (lane == 0) red[0][wave] = vf;
        __syncthreads();
        for (int w = 0; w < kWaves; ++w) bm += red[0][w];
    }
    // sum_vec (:137-170) and the new input (:172-186), scalar = 1
    for (int s = tid; s < S; s += kBroydenThreads) {
        const double prev = b.n_in[s];
        const double sv = m > 0 ? fma(b.W[(size_t)m * S + s], bm, b.sum_vec[s]) : 0.0;
        if (m > 0) b.sum_vec[s] = sv;
        const double pf = PC ? PF[s] : fraction * b.F[s];
        if (PC) b.PFp[s] = PF[s];
        b.n_prev[s] = prev;
        b.n_in[s] = prev - pf - sv;
    }
}

__global__ __launch_bounds__(kBroydenThreads) void broyden_step_kernel(int S, int m, double fraction, int norm_type,
                                                                         const double* __restrict__ n_out,
                                                                         BroydenDev b)
{
    broyden_update<false>(S, m, fraction, norm_type, n_out, nullptr, b);
}

// U_i = U_ext_i - V_gate - sum_j K_ij n_j: one wavefront per row, lanes
// strided over j.
__global__ __launch_bounds__(256) void coulomb_potential_kernel(int N, const double* __restrict__ K,
                                                                const double* __restrict__ n,
                                                                const double* __restrict__ U_ext, double V_gate,
                                                                double* __restrict__ U)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= N) return;
    const double* __restrict__ k = K + (size_t)row * N;
    double acc = 0.0;
#pragma unroll 4
    for (int j = lane; j < N; j += 64) acc = fma(k[j], n[j], acc);
    acc = wave_sum(acc);
    if (lane == 0) U[row] = (U_ext ? U_ext[row] : 0.0) - V_gate - acc;
}

// ---------------------------------------------------------------- Broyden
class Broyden {
   public:
    ~Broyden()
    {
        for (DevBuf* d : {&d_state, &d_hist, &d_out, &d_nout}) d->release();
        h_out.release();
    }
    int init(int S_, double fraction_, int max_step_, int norm_type_)
    {
        S = S_;
        fraction = fraction_;
        max_step = max_step_;
        norm_type = norm_type_;
        int rc;
        if ((rc = d_state.ensure(sizeof(double) * 6 * (size_t)S))) return rc;
        if ((rc = d_hist.ensure(sizeof(double) * (2 * (size_t)(max_step + 1) * S + 2 * (max_step + 1))))) return rc;
        if ((rc = d_out.ensure(sizeof(double) * 4))) return rc;
        if ((rc = d_nout.ensure(sizeof(double) * (size_t)S))) return rc;
        if ((rc = h_out.ensure(sizeof(double) * 4))) return rc;
        return reset(0, 0.0, nullptr);
    }
    BroydenDev dev() const
    {
        BroydenDev b;
        double* st = d_state.as<double>();
        b.n_in = st;
        b.n_prev = st + S;
        b.F = st + 2 * (size_t)S;
        b.dF = st + 3 * (size_t)S;
        b.Norm = st + 4 * (size_t)S;
        b.sum_vec = st + 5 * (size_t)S;
        double* h = d_hist.as<double>();
        b.V = h;
        b.W = h + (size_t)(max_step + 1) * S;
        b.intermed = h + 2 * (size_t)(max_step + 1) * S;
        b.intermed2 = b.intermed + (max_step + 1);
        b.out = d_out.as<double>();
        b.PFp = nullptr;
        return b;
    }
    // Reset_Broyden_Parallel (Broyden_Parallel_General.cpp:250-355).
    int reset(int keep_n_in, double init_value, hipStream_t st)
    {
        BroydenDev b = dev();
        int rc;
        if (!keep_n_in && (rc = fill_device(b.n_in, S, init_value, st))) return rc;
        if ((rc = fill_device(b.n_prev, 5LL * S, 0.0, st))) return rc;
        if ((rc = fill_device(d_hist.as<double>(), 2LL * (max_step + 1) * S + 2 * (max_step + 1), 0.0, st))) return rc;
        step = 1;
        return NEGF_OK;
    }
    // One update from the device output d_n_out; reads back the norms.
    int update(const double* d_n_out, double* max_norm, double* l2, hipStream_t st)
    {
        const int m = step - 1;
        if (m > max_step) {
            set_error("Broyden_Step %d exceeds Broyden_threshold_maxstep %d", step, max_step);
            return NEGF_ENONCONV;
        }
        hipLaunchKernelGGL(broyden_step_kernel, dim3(1), dim3(kBroydenThreads), 0, st, S, m, fraction, norm_type,
                           d_n_out, dev());
        NEGF_HIP_TRY(hipGetLastError());
        NEGF_HIP_TRY(hipMemcpyAsync(h_out.p, d_out.p, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        const double* o = h_out.as<double>();
        if (max_norm) *max_norm = o[0];
        if (l2) *l2 = o[1];
        step += 1;
        if (!std::isfinite(o[0]) || !std::isfinite(o[1])) {
            set_error("Broyden: non-finite norm");
            return NEGF_ENAN;
        }
        return NEGF_OK;
    }

    int S = 0, max_step = 0, norm_type = 0, step = 1;
    double fraction = 0.1;
    DevBuf d_state, d_hist, d_out, d_nout;
    HostBuf h_out;
};

// ---------------------------------------------------------------- SCF
struct Scf {
    ~Scf()
    {
        if (st) (void)hipStreamSynchronize(st);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (DevBuf* d : {&d_K, &d_U, &d_Q, &d_Uext, &d_V, &d_step}) d->release();
        if (h_back) (void)hipHostFree(h_back);
        if (st) (void)hipStreamDestroy(st);
    }
    Tube* tube = nullptr;
    negf_scf_params p{};
    Broyden br;
    std::vector<double> K;
    DevBuf d_K, d_U, d_Q, d_Uext;
    bool has_ext = false;
    hipStream_t st = nullptr;
    // One iteration captured in a graph and replayed: gate voltage and
    // Broyden step live on the device; the Broyden kernel writes the norms
    // and the plans' status words straight into mapped, coherent host memory
    // (ScfBack), so the replay holds no copy.
    bool use_graph = true;
    hipGraphExec_t gexec = nullptr;
    long long g_paths = -1;
    bool g_ext = false;
    DevBuf d_V, d_step;
    void* h_back = nullptr;    // host address
    void* d_back = nullptr;    // the same memory as the device sees it
};

struct ScfBack {
    double norm, l2;
    int status[2];
};

// Dense stand-in for the Poisson solve + gather (not in the reference):
// K_ij [V per e] = ke/eps_r (1/sqrt(d^2 + R^2) - 1/sqrt(d^2 + R_img^2)),
// d = axial distance of rings i, j in nm, ke = q/(4 pi eps0) * 1e9 V nm.
static void coulomb_kernel(const std::vector<double>& ptd, const negf_scf_params& p, std::vector<double>& K)
{
    const double q = 1.602176634e-19, eps0 = 8.8541878128e-12, pi = 3.14159265358979323846;
    const double ke = q / (4.0 * pi * eps0) * 1.0e9;
    const size_t N = ptd.size();
    K.resize(N * N);
    for (size_t i = 0; i < N; ++i)
        for (size_t j = 0; j < N; ++j) {
            const double d = ptd[i] - ptd[j];
            const double d2 = d * d;
            K[i * N + j] = ke / p.eps_r *
                           (1.0 / std::sqrt(d2 + p.R_ring_nm * p.R_ring_nm) -
                            1.0 / std::sqrt(d2 + p.R_image_nm * p.R_image_nm));
        }
}

// ---------------------------------------------------------------- batched SCF
// Many equal tubes coupled to their own charge through one shared dense
// kernel K (N x N): every iteration is a GEMM U_t = U_ext,t - V_t - K n_t for
// all tubes, one batched NEGF solve, and one Broyden workgroup per tube.

// Partial products Yp[z][t][i] = sum_{j in chunk z} A[i][j] X[t][j]
// (A row-major M x K, rows lda apart; X rows ldx apart; T rows) on
// v_mfma_f64_16x16x4f64, chunk z = [z kchunk, (z + 1) kchunk); Yp[z] is
// T x M.  A workgroup of 4 wavefronts computes a 32 (i) x 32 (t) tile, one
// 16 x 16 quadrant per wavefront, over 32-wide j blocks staged in LDS; the
// next block's global loads are issued before the present block's MFMAs.
// Lane l supplies A[i0 + l % 16][j + l / 16] and X[t0 + l % 16][j + l / 16];
// result register v of lane l is the entry (i = l / 16 + 4 v, t = l % 16).
// The caller sums the chunks in order.
typedef __attribute__((ext_vector_type(4))) double d4;
__global__ __launch_bounds__(256) void mfma_dgemm_kernel(int M, int K, int T, const double* __restrict__ A,
                                                         size_t lda, const double* __restrict__ X, size_t ldx,
                                                         int kchunk, double* __restrict__ Yp)
{
    __shared__ double As[32][33];
    __shared__ double Xs[32][33];
    const int i0 = blockIdx.x * 32, t0 = blockIdx.y * 32;
    const int jb = blockIdx.z * kchunk, je = min(K, jb + kchunk);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int qi = w >> 1, qt = w & 1;
    double ra[4], rx[4];
    auto fetch = [&](int j0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            const int r = e >> 5, c = e & 31;
            const int i = i0 + r, t = t0 + r, j = j0 + c;
            ra[q] = (i < M && j < je) ? A[(size_t)i * lda + j] : 0.0;
            rx[q] = (t < T && j < je) ? X[(size_t)t * ldx + j] : 0.0;
        }
    };
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    fetch(jb);
    for (int j0 = jb; j0 < je; j0 += 32) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = tid + 256 * q;
            As[e >> 5][e & 31] = ra[q];
            Xs[e >> 5][e & 31] = rx[q];
        }
        __syncthreads();
        if (j0 + 32 < je) fetch(j0 + 32);
#pragma unroll
        for (int k0 = 0; k0 < 32; k0 += 4) {
            const double a = As[qi * 16 + (lane & 15)][k0 + (lane >> 4)];
            const double x = Xs[qt * 16 + (lane & 15)][k0 + (lane >> 4)];
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, x, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    double* Y = Yp + (size_t)blockIdx.z * T * M;
    const int t = t0 + qt * 16 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = i0 + qi * 16 + (lane >> 4) + 4 * v;
        if (i < M && t < T) Y[(size_t)t * M + i] = acc[v];
    }
}

// K-split of a GEMM with M output rows and inner length K: enough
// workgroups to cover the chip (about 1024), chunks a multiple of 32 wide
static int gemm_splits(int M, int K, int T)
{
    const int tiles = ((M + 31) / 32) * ((T + 31) / 32);
    int z = std::max(1, std::min(8, 1024 / std::max(tiles, 1)));
    while (z > 1 && (K + 32 * z - 1) / (32 * z) < 4) --z;  // keep >= 4 blocks per chunk
    return z;
}

static int gemm_chunk(int K, int z) { return ((K + z - 1) / z + 31) / 32 * 32; }

// Yp[z] (T x M) = A (M x K, rows lda apart) X^T over chunk z of the inner index
static void launch_dgemm(int M, int K, int T, int z, const double* A, size_t lda, const double* X, size_t ldx,
                         double* Yp, hipStream_t st)
{
    dim3 g((M + 31) / 32, (T + 31) / 32, z);
    hipLaunchKernelGGL(mfma_dgemm_kernel, g, dim3(256), 0, st, M, K, T, A, lda, X, ldx, gemm_chunk(K, z), Yp);
}

// Y = sum of the z partial products, in chunk order
__device__ __forceinline__ double sum_parts(const double* __restrict__ P, int z, long long NT, long long q)
{
    double a = P[q];
    for (int k = 1; k < z; ++k) a += P[k * NT + q];
    return a;
}

// U_t = U_ext,t - V_t - KN_t, scattered into each tube's potential and
// contact potentials; also kept in Uall for the caller.
__global__ void batch_potential_kernel(int N, int T, const double* __restrict__ Uext, const double* __restrict__ V,
                                       const double* __restrict__ KN, int nz, double* const* __restrict__ dU,
                                       double* const* __restrict__ dUc, double* __restrict__ Uall)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long NT = (long long)N * T;
    if (q >= NT) return;
    const int t = (int)(q / N), i = (int)(q - (long long)t * N);
    const double u = (Uext ? Uext[q] : 0.0) - V[t] - sum_parts(KN, nz, NT, q);
    Uall[q] = u;
    dU[t][i] = u;
    if (i == 0) dUc[t][0] = u;
    if (i == N - 1) dUc[t][1] = u;
}

// One Broyden update per active tube (workgroup t); step counters on the
// device so that the launch is identical every iteration.
template <bool PC>
__global__ __launch_bounds__(kBroydenThreads) void broyden_batch_kernel(int S, int max_step, double fraction,
                                                                          int norm_type, const double* __restrict__ n_out,
                                                                          const double* __restrict__ PF,
                                                                          BroydenDev base, size_t state_stride,
                                                                          size_t hist_stride, int* __restrict__ steps,
                                                                          const int* __restrict__ active)
{
    const int t = blockIdx.x;
    if (!active[t]) return;
    const int m = steps[t] - 1;
    BroydenDev b;
    b.n_in = base.n_in + t * state_stride;
    b.n_prev = base.n_prev + t * state_stride;
    b.F = base.F + t * state_stride;
    b.dF = base.dF + t * state_stride;
    b.Norm = base.Norm + t * state_stride;
    b.sum_vec = base.sum_vec + t * state_stride;
    b.V = base.V + t * hist_stride;
    b.W = base.W + t * hist_stride;
    b.intermed = base.intermed + t * hist_stride;
    b.intermed2 = base.intermed2 + t * hist_stride;
    b.out = base.out + (size_t)t * 4;
    b.PFp = PC ? base.PFp + t * state_stride : nullptr;
    if (m > max_step) {
        if (threadIdx.x == 0) b.out[3] = 1.0;  // step limit reached
        return;
    }
    broyden_update<PC>(S, m, fraction, norm_type, n_out + (size_t)t * S, PC ? PF + (size_t)t * S : nullptr, b);
    if (threadIdx.x == 0) {
        steps[t] = m + 2;
        b.out[3] = 0.0;
    }
}

// After the Broyden updates (one workgroup per tube): the tubes' Broyden
// outputs (max norm, l2 norm, |dF|^2, step-limit flag) and the plans' status
// words go to the host's read-back record; an active tube whose max norm
// reached max_norm keeps its U and Q (Uo, Qo) and leaves the active set, as
// does one at its step limit.  The host applies the same tests to the same
// values, so its bookkeeping follows without copies.
__global__ __launch_bounds__(256) void batch_track_kernel(int N, const double* __restrict__ out, double max_norm,
                                                          int* __restrict__ active, const double* __restrict__ U,
                                                          const double* __restrict__ Q, double* __restrict__ Uo,
                                                          double* __restrict__ Qo, double* __restrict__ back,
                                                          int* __restrict__ back_status,
                                                          const int* __restrict__ st_eq,
                                                          const int* __restrict__ st_neq)
{
    __shared__ int keep;
    const int t = blockIdx.x;
    if (threadIdx.x == 0) {
        const double nrm = out[4 * t], lim = out[4 * t + 3];
        for (int k = 0; k < 4; ++k) back[4 * t + k] = out[4 * t + k];
        int kp = 0;
        if (active[t]) {
            if (lim != 0.0)
                active[t] = 0;
            else if (nrm <= max_norm)  // false for NaN: the host stops on it
                kp = 1;
        }
        keep = kp;
        if (t == 0) {
            back_status[0] = *st_eq;
            back_status[1] = st_neq ? *st_neq : 0;
        }
        __threadfence_system();
    }
    __syncthreads();
    if (!keep) return;
    const size_t o = (size_t)t * N;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        Uo[o + i] = U[o + i];
        Qo[o + i] = Q[o + i];
    }
    if (threadIdx.x == 0) active[t] = 0;
}

// R[t][s] = n_in,t[s] - Q[t][s] (the reference's F = n_in - n_out)
__global__ void batch_residual_kernel(int N, int T, const double* __restrict__ n_in, size_t stride,
                                      const double* __restrict__ Q, double* __restrict__ R)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)N * T) return;
    const int t = (int)(q / N), i = (int)(q - (long long)t * N);
    R[q] = n_in[(size_t)t * stride + i] - Q[q];
}

// Mode-space factor of the preconditioner's correction:
// Y[t][k] = (sum of the z partial products) * c_t lam_k / (1 + c_t lam_k)
__global__ void batch_modescale_kernel(int R, int T, const double* __restrict__ lam, const double* __restrict__ c,
                                       const double* __restrict__ parts, int nz, double* __restrict__ Y)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long RT = (long long)R * T;
    if (q >= RT) return;
    const int t = (int)(q / R), k = (int)(q - (long long)t * R);
    const double cl = c[t] * lam[k];
    Y[q] = sum_parts(parts, nz, RT, q) * (cl / (1.0 + cl));
}

// X = Rv - (sum of the z partial products)
__global__ void batch_minus_parts_kernel(long long NT, const double* __restrict__ Rv,
                                         const double* __restrict__ parts, int nz, double* __restrict__ X)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < NT) X[q] = Rv[q] - sum_parts(parts, nz, NT, q);
}

// c_t = max(0, mean_s J[t][s]) for the tubes in use: the site-averaged
// response dQ/dU . 1 (capacitance per site, e per eV); one workgroup per tube
__global__ __launch_bounds__(256) void batch_rowmean_kernel(int N, const double* __restrict__ J,
                                                            double* __restrict__ c)
{
    __shared__ double red[4];
    const int t = blockIdx.x;
    double a = 0.0;
    for (int s = threadIdx.x; s < N; s += 256) a += J[(size_t)t * N + s];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = (red[0] + red[1] + red[2] + red[3]) / N;
        c[t] = m > 0.0 ? m : 0.0;
    }
}

// n_in,t -= P R_t (a preconditioned step without history, for the seed)
// One seed step n_in -= beta P (n_in - Q_surrogate)
__global__ void batch_seed_step_kernel(int N, int T, double* __restrict__ n_in, size_t stride,
                                       const double* __restrict__ PR, double beta)
{
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long long)N * T) return;
    const int t = (int)(q / N), i = (int)(q - (long long)t * N);
    n_in[(size_t)t * stride + i] -= beta * PR[q];
}

class BatchScf {
   public:
    ~BatchScf()
    {
        drop_graph();
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        if (h_back) (void)hipHostFree(h_back);
    }
    int init(negf_batch* b, negf_tube* const* tubes, int T_, const negf_scf_params& prm)
    {
        batch = b;
        T = T_;
        p = prm;
        N = tubes[0]->t.N;
        this->tubes.assign(tubes, tubes + T_);
        for (int t = 0; t < T; ++t) {
            if (tubes[t]->t.N != N) {
                set_error("batched SCF needs tubes of equal length (shared electrostatic kernel)");
                return NEGF_EARG;
            }
            hU.push_back(tubes[t]->t.d_U.as<double>());
            hUc.push_back(tubes[t]->t.d_Uc.as<double>());
        }
        coulomb_kernel(tubes[0]->t.ptd, p, K);
        M = p.Broyden_threshold_maxstep;
        state_stride = 7 * (size_t)N;
        hist_stride = 2 * (size_t)(M + 1) * N + 2 * (M + 1);
        int rc;
        NEGF_HIP_TRY(hipStreamCreate(&st));
        if ((rc = d_K.ensure(sizeof(double) * K.size()))) return rc;
        if ((rc = d_Uext.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_V.ensure(sizeof(double) * T))) return rc;
        nz = gemm_splits(N, N, T);
        if ((rc = d_KN.ensure(sizeof(double) * N * T * nz))) return rc;
        if ((rc = d_U.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_Q.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_Uo.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_Qo.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_ptrU.ensure(sizeof(double*) * T))) return rc;
        if ((rc = d_ptrUc.ensure(sizeof(double*) * T))) return rc;
        if ((rc = d_state.ensure(sizeof(double) * state_stride * T))) return rc;
        if ((rc = d_hist.ensure(sizeof(double) * hist_stride * T))) return rc;
        if ((rc = d_out.ensure(sizeof(double) * 4 * T))) return rc;
        if ((rc = d_steps.ensure(sizeof(int) * T))) return rc;
        if ((rc = d_active.ensure(sizeof(int) * T))) return rc;
        {
            // read-back record: 4 doubles per tube, then 2 status ints
            hipError_t eh = hipHostMalloc(&h_back, sizeof(double) * 4 * T + 2 * sizeof(int),
                                          hipHostMallocMapped | hipHostMallocCoherent);
            if (eh == hipSuccess) eh = hipHostGetDevicePointer(&d_back, h_back, 0);
            if (eh != hipSuccess) {
                set_error("mapped host buffer for the batched SCF read-back: %s", hipGetErrorString(eh));
                return NEGF_EALLOC;
            }
        }
        NEGF_HIP_TRY(hipMemcpy(d_K.p, K.data(), sizeof(double) * K.size(), hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_ptrU.p, hU.data(), sizeof(double*) * T, hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_ptrUc.p, hUc.data(), sizeof(double*) * T, hipMemcpyHostToDevice));
        if ((rc = fill_device(d_Uext.as<double>(), (long long)N * T, 0.0, st))) return rc;
        if ((rc = fill_device(d_state.as<double>(), (long long)(state_stride * T), 0.0, st))) return rc;
        for (int t = 0; t < T; ++t)
            if ((rc = fill_device(d_state.as<double>() + t * state_stride, N, p.NS_initial_deposit_value, st)))
                return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        return NEGF_OK;
    }
    BroydenDev base() const
    {
        BroydenDev b;
        double* s0 = d_state.as<double>();
        b.n_in = s0;
        b.n_prev = s0 + N;
        b.F = s0 + 2 * (size_t)N;
        b.dF = s0 + 3 * (size_t)N;
        b.Norm = s0 + 4 * (size_t)N;
        b.sum_vec = s0 + 5 * (size_t)N;
        double* h = d_hist.as<double>();
        b.V = h;
        b.W = h + (size_t)(M + 1) * N;
        b.intermed = h + 2 * (size_t)(M + 1) * N;
        b.intermed2 = b.intermed + (M + 1);
        b.out = d_out.as<double>();
        b.PFp = s0 + 6 * (size_t)N;
        return b;
    }
    int enqueue_iteration()
    {
        int rc = enqueue_potential();
        if (rc) return rc;
        rc = batch_enqueue(batch, d_Q.as<double>(), st, d_active.as<int>());
        if (rc) return rc;
        hipLaunchKernelGGL(broyden_batch_kernel<false>, dim3(T), dim3(kBroydenThreads), 0, st, N, M,
                           p.Broyden_fraction, p.Broyden_norm_type, d_Q.as<double>(), nullptr, base(), state_stride,
                           hist_stride, d_steps.as<int>(), d_active.as<int>());
        NEGF_HIP_TRY(hipGetLastError());
        return enqueue_track();
    }
    int enqueue_track()
    {
        const int* st_eq = nullptr;
        const int* st_neq = nullptr;
        batch_status_words(batch, &st_eq, &st_neq);
        double* back = static_cast<double*>(d_back);
        hipLaunchKernelGGL(batch_track_kernel, dim3(T), dim3(256), 0, st, N, d_out.as<const double>(),
                           p.Broyden_max_norm, d_active.as<int>(), d_U.as<const double>(), d_Q.as<const double>(),
                           d_Uo.as<double>(), d_Qo.as<double>(), back, reinterpret_cast<int*>(back + 4 * T), st_eq,
                           st_neq);
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        if (gexec_pc) (void)hipGraphExecDestroy(gexec_pc);
        gexec_pc = nullptr;
    }
    // Capture fn() on st into *ge unless a graph exists that was captured
    // over the batch's present plans (same build generation).
    template <class F>
    int capture(hipGraphExec_t* ge, long long* gen, F fn)
    {
        const long long g_now = batch_generation(batch);
        if (*ge && *gen == g_now) return NEGF_OK;
        if (*ge) (void)hipGraphExecDestroy(*ge);
        *ge = nullptr;
        NEGF_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int erc = fn();
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (erc) {
            if (g) (void)hipGraphDestroy(g);
            return erc;
        }
        if (e != hipSuccess) {
            set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
            return NEGF_EDEVICE;
        }
        const hipError_t e2 = hipGraphInstantiate(ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            *ge = nullptr;
            set_error("hipGraphInstantiate: %s", hipGetErrorString(e2));
            return NEGF_EDEVICE;
        }
        *gen = g_now;
        return NEGF_OK;
    }

    // Host part of iteration it: when a tube's adaptive correction falls due
    // (flag_adaptive_integration_limits), its integrand is sampled at this
    // iteration's potential, so the potential is written first, as in
    // Solve_NEGF (NEGF_Common.cpp:1314-1337); then the batch's plans are
    // brought up to date.
    int prepare_iteration(int it)
    {
        int rc;
        if (adaptive_due(it)) {
            if ((rc = enqueue_potential())) return rc;
            NEGF_HIP_TRY(hipStreamSynchronize(st));
        }
        bool rebuilt = false;
        return batch_prepare(batch, it, &rebuilt);
    }
    bool adaptive_due(int it) const
    {
        for (const negf_tube* t : tubes)
            if (t->t.p.flag_adaptive_integration_limits && it % t->t.p.integrand_correction_interval == 0) return true;
        return false;
    }

    // After the iteration's synchronisation, from the read-back record that
    // the track kernel wrote: fold nonzero plan status words into the batch's
    // sticky root flag, stop on a non-finite norm, and mirror the device's
    // retirement of converged tubes (which already kept their U and Q).
    struct LoopState {
        std::vector<int> active, keep, done_it;
        std::vector<double> fin;
        int remaining = 0;
        bool overflow = false;
    };
    int after_iteration(int it, LoopState& ls)
    {
        int rc;
        const volatile double* o = static_cast<const volatile double*>(h_back);
        const volatile int* bs = reinterpret_cast<const volatile int*>(o + 4 * T);
        if ((bs[0] | bs[1]) && (rc = batch_poll_status(batch))) return rc;
        for (int t = 0; t < T; ++t) {
            ls.keep[t] = 0;
            if (!ls.active[t]) continue;
            if (o[4 * t + 3] != 0.0) {
                ls.overflow = true;
                ls.active[t] = 0;
                --ls.remaining;
                continue;
            }
            const double nrm = o[4 * t];
            if (!std::isfinite(nrm)) {
                set_error("batched SCF: non-finite Broyden norm for tube %d at iteration %d", t, it + 1);
                return NEGF_ENAN;
            }
            if (nrm <= p.Broyden_max_norm) {  // the track kernel kept U, Q and cleared the flag
                ls.keep[t] = 1;
                ls.active[t] = 0;
                ls.done_it[t] = it + 1;
                ls.fin[t] = nrm;
                --ls.remaining;
            }
        }
        return NEGF_OK;
    }
    int finish_run(int it, const LoopState& ls, double* U_out, double* Q_out, int* iters, double* norms,
                   int* total_iters)
    {
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        if (U_out) NEGF_HIP_TRY(hipMemcpy(U_out, d_Uo.p, sizeof(double) * N * T, hipMemcpyDeviceToHost));
        if (Q_out) NEGF_HIP_TRY(hipMemcpy(Q_out, d_Qo.p, sizeof(double) * N * T, hipMemcpyDeviceToHost));
        if (iters) std::copy(ls.done_it.begin(), ls.done_it.end(), iters);
        if (norms) std::copy(ls.fin.begin(), ls.fin.end(), norms);
        if (total_iters) *total_iters = it;
        if (ls.overflow) {
            set_error("Broyden_Step has exceeded the Broyden_Threshold_MaxStep for some tube");
            return NEGF_ENONCONV;
        }
        return NEGF_OK;
    }
    int start_run(const double* V, LoopState& ls)
    {
        NEGF_HIP_TRY(hipMemcpyAsync(d_V.p, V, sizeof(double) * T, hipMemcpyHostToDevice, st));
        std::vector<int> steps(T, 1);
        ls.active.assign(T, 1);
        ls.keep.assign(T, 0);
        ls.done_it.assign(T, -1);
        ls.fin.assign(T, 0.0);
        ls.remaining = T;
        ls.overflow = false;
        NEGF_HIP_TRY(hipMemcpyAsync(d_steps.p, steps.data(), sizeof(int) * T, hipMemcpyHostToDevice, st));
        NEGF_HIP_TRY(hipMemcpyAsync(d_active.p, ls.active.data(), sizeof(int) * T, hipMemcpyHostToDevice, st));
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        return NEGF_OK;
    }

    // Run every tube to Broyden_Norm <= Broyden_max_norm (or its step
    // limit); per-tube iteration counts and final norms.
    int run(const double* V, double* U_out, double* Q_out, int* iters, double* norms, int* total_iters)
    {
        int rc;
        LoopState ls;
        if ((rc = start_run(V, ls))) return rc;
        int it = 0;
        while (ls.remaining > 0) {
            if (it > M + 1) {
                ls.overflow = true;
                break;
            }
            if ((rc = prepare_iteration(it))) return rc;
            if (use_graph) {
                if ((rc = capture(&gexec, &gen_graph, [&] { return enqueue_iteration(); }))) return rc;
                NEGF_HIP_TRY(hipGraphLaunch(gexec, st));
            } else if ((rc = enqueue_iteration())) {
                return rc;
            }
            NEGF_HIP_TRY(hipStreamSynchronize(st));
            if ((rc = after_iteration(it, ls))) return rc;
            ++it;
        }
        return finish_run(it, ls, U_out, Q_out, iters, norms, total_iters);
    }

    // ---- surrogate-seeded, preconditioned iteration
    // P_t = Z diag(1 / (1 + c_t lam)) Z^T = I - Z diag(c_t lam / (1 + c_t lam)) Z^T,
    // applied with the modes of lam_k >= kPcModeCut * max lam only: a dropped
    // mode's factor differs from 1 by c_t lam_k / (1 + c_t lam_k) < c_t
    // kPcModeCut max lam, and P only sets the mixing, not the fixed point.
    static constexpr double kPcModeCut = 1e-2;
    int set_preconditioner(const double* Z, const double* lam)
    {
        int rc;
        double lmax = 0.0;
        for (int k = 0; k < N; ++k) lmax = std::max(lmax, lam[k]);
        std::vector<int> keep;
        for (int k = 0; k < N; ++k)
            if (lam[k] > 0.0 && lam[k] >= kPcModeCut * lmax) keep.push_back(k);
        R_pc = std::max<int>(1, (int)keep.size());
        const size_t NR = (size_t)N * R_pc;
        if ((rc = d_Z.ensure(sizeof(double) * NR))) return rc;
        if ((rc = d_Zt.ensure(sizeof(double) * NR))) return rc;
        if ((rc = d_lam.ensure(sizeof(double) * R_pc))) return rc;
        if ((rc = d_c.ensure(sizeof(double) * T))) return rc;
        if ((rc = d_R.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_Rh.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_PF.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_ones.ensure(sizeof(double) * N * T))) return rc;
        if ((rc = d_J.ensure(sizeof(double) * N * T))) return rc;
        z_pc1 = gemm_splits(R_pc, N, T);
        z_pc2 = gemm_splits(N, R_pc, T);
        const size_t kn = std::max({(size_t)N * T * nz, (size_t)R_pc * T * z_pc1, (size_t)N * T * z_pc2});
        if ((rc = d_KN.ensure(sizeof(double) * kn))) return rc;
        // Zr: N x R (row i: the kept modes' entries at site i); Zrt: R x N
        std::vector<double> zr(NR, 0.0), zrt(NR, 0.0), lr(R_pc, 0.0);
        for (int q = 0; q < (int)keep.size(); ++q) {
            const int k = keep[q];
            lr[q] = lam[k];
            for (int i = 0; i < N; ++i) {
                zr[(size_t)i * R_pc + q] = Z[(size_t)i * N + k];
                zrt[(size_t)q * N + i] = Z[(size_t)i * N + k];
            }
        }
        NEGF_HIP_TRY(hipMemcpy(d_Z.p, zr.data(), sizeof(double) * NR, hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_Zt.p, zrt.data(), sizeof(double) * NR, hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_lam.p, lr.data(), sizeof(double) * R_pc, hipMemcpyHostToDevice));
        if ((rc = fill_device(d_ones.as<double>(), (long long)N * T, 1.0, st))) return rc;
        if ((rc = fill_device(d_c.as<double>(), T, 0.0, st))) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        has_pc = true;
        drop_graph();
        return NEGF_OK;
    }
    // U_t = U_ext,t - V_t - K n_in,t into the tubes and d_U
    int enqueue_potential()
    {
        launch_dgemm(N, N, T, nz, d_K.as<double>(), (size_t)N, base().n_in, state_stride, d_KN.as<double>(), st);
        const long long NT = (long long)N * T;
        hipLaunchKernelGGL(batch_potential_kernel, dim3((unsigned)((NT + 255) / 256)), dim3(256), 0, st, N, T,
                           d_Uext.as<double>(), d_V.as<double>(), d_KN.as<double>(), nz,
                           d_ptrU.as<double*>(), d_ptrUc.as<double*>(), d_U.as<double>());
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    // d_PF = P (n_in - Q) from the charge Q (T x N): R = n_in - Q,
    // Y = Zr^T R, Y_k *= c_t lam_k / (1 + c_t lam_k), PF = R - Zr Y
    int enqueue_precondition(const double* Q)
    {
        const long long NT = (long long)N * T, RT = (long long)R_pc * T;
        const dim3 g((unsigned)((NT + 255) / 256));
        hipLaunchKernelGGL(batch_residual_kernel, g, dim3(256), 0, st, N, T, base().n_in, state_stride, Q,
                           d_R.as<double>());
        launch_dgemm(R_pc, N, T, z_pc1, d_Zt.as<double>(), (size_t)N, d_R.as<double>(), (size_t)N,
                     d_KN.as<double>(), st);
        hipLaunchKernelGGL(batch_modescale_kernel, dim3((unsigned)((RT + 255) / 256)), dim3(256), 0, st, R_pc, T,
                           d_lam.as<double>(), d_c.as<double>(), d_KN.as<double>(), z_pc1, d_Rh.as<double>());
        launch_dgemm(N, R_pc, T, z_pc2, d_Z.as<double>(), (size_t)R_pc, d_Rh.as<double>(), (size_t)R_pc,
                     d_KN.as<double>(), st);
        hipLaunchKernelGGL(batch_minus_parts_kernel, g, dim3(256), 0, st, NT, d_R.as<double>(), d_KN.as<double>(),
                           z_pc2, d_PF.as<double>());
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    // c_t from the surrogate's tangent dQ/dU . 1 at the present d_U
    int enqueue_capacitance()
    {
        int rc = sg->forward(T, d_U.as<double>(), has_feat ? d_feat.as<double>() : nullptr, st);
        if (!rc) rc = sg->jvp(T, d_ones.as<double>(), d_J.as<double>(), st);
        if (rc) return rc;
        hipLaunchKernelGGL(batch_rowmean_kernel, dim3(T), dim3(256), 0, st, N, d_J.as<double>(), d_c.as<double>());
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    int enqueue_iteration_pc()
    {
        int rc = enqueue_potential();
        if (!rc) rc = batch_enqueue(batch, d_Q.as<double>(), st, d_active.as<int>());
        if (!rc) rc = enqueue_precondition(d_Q.as<double>());
        if (rc) return rc;
        hipLaunchKernelGGL(broyden_batch_kernel<true>, dim3(T), dim3(kBroydenThreads), 0, st, N, M,
                           p.Broyden_fraction, p.Broyden_norm_type, d_Q.as<double>(), d_PF.as<double>(), base(),
                           state_stride, hist_stride, d_steps.as<int>(), d_active.as<int>());
        NEGF_HIP_TRY(hipGetLastError());
        return enqueue_track();
    }
    // Seed steps are half steps: with full steps the scalar capacitance
    // overshoots on the high-gate tubes, whose seed then oscillates (128
    // tubes x 2048 sites: 3 full steps leave a tube unconverged after 300
    // exact iterations; 6 half steps need 23 against 29 after 6 full ones).
    static constexpr double kSeedStep = 0.5;
    // Seeded, preconditioned run: seed_iters preconditioned steps on the
    // surrogate charge (a fixed-point iteration of the surrogate SCF), then
    // the Broyden loop on the exact NEGF charge with the mixing term P F.  The
    // capacitances come from the surrogate at the start, half way through
    // the seed, at the seed, and (with a fresh history) at iteration
    // refresh_iter of the exact loop if refresh_iter > 0.
    int run_pc(const double* V, int seed_iters, int refresh_iter, double* U_out, double* Q_out, int* iters,
               double* norms, int* total_iters, double* cap)
    {
        int rc;
        LoopState ls;
        if ((rc = start_run(V, ls))) return rc;
        if ((rc = enqueue_potential())) return rc;
        if ((rc = enqueue_capacitance())) return rc;
        const long long NT = (long long)N * T;
        for (int k = 0; k < seed_iters; ++k) {
            // the capacitances follow the seed once, half way
            if (k > 0 && k == seed_iters / 2 && (rc = enqueue_capacitance())) return rc;
            if ((rc = sg->forward(T, d_U.as<double>(), has_feat ? d_feat.as<double>() : nullptr, st))) return rc;
            if ((rc = sg->output(T, d_Q.as<double>(), st))) return rc;
            if ((rc = enqueue_precondition(d_Q.as<double>()))) return rc;
            hipLaunchKernelGGL(batch_seed_step_kernel, dim3((unsigned)((NT + 255) / 256)), dim3(256), 0, st, N, T,
                               base().n_in, state_stride, d_PF.as<double>(), kSeedStep);
            NEGF_HIP_TRY(hipGetLastError());
            if ((rc = enqueue_potential())) return rc;
        }
        if (seed_iters > 0 && (rc = enqueue_capacitance())) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        int it = 0;
        while (ls.remaining > 0) {
            if (it > M + 1) {
                ls.overflow = true;
                break;
            }
            if (it > 0 && it == refresh_iter) {
                // new capacitances at the present potential; fresh history, n_in kept
                if ((rc = enqueue_capacitance())) return rc;
                std::vector<int> steps(T, 1);
                NEGF_HIP_TRY(hipMemcpyAsync(d_steps.p, steps.data(), sizeof(int) * T, hipMemcpyHostToDevice, st));
                NEGF_HIP_TRY(hipStreamSynchronize(st));
            }
            if ((rc = prepare_iteration(it))) return rc;
            if (use_graph) {
                if ((rc = capture(&gexec_pc, &gen_graph_pc, [&] { return enqueue_iteration_pc(); }))) return rc;
                NEGF_HIP_TRY(hipGraphLaunch(gexec_pc, st));
            } else if ((rc = enqueue_iteration_pc())) {
                return rc;
            }
            NEGF_HIP_TRY(hipStreamSynchronize(st));
            if ((rc = after_iteration(it, ls))) return rc;
            ++it;
        }
        rc = finish_run(it, ls, U_out, Q_out, iters, norms, total_iters);
        if (cap) NEGF_HIP_TRY(hipMemcpy(cap, d_c.p, sizeof(double) * T, hipMemcpyDeviceToHost));
        return rc;
    }

    Surrogate* sg = nullptr;
    bool has_pc = false, has_feat = false;
    DevBuf d_Z, d_Zt, d_lam, d_c, d_R, d_Rh, d_PF, d_ones, d_J, d_feat;
    hipGraphExec_t gexec_pc = nullptr;
    long long gen_graph = -1, gen_graph_pc = -1;  // batch generation each graph was captured over

    negf_batch* batch = nullptr;
    std::vector<negf_tube*> tubes;
    int T = 0, N = 0, M = 0;
    int nz = 1;  // K-split of the GEMMs (partials in d_KN)
    int R_pc = 0, z_pc1 = 1, z_pc2 = 1;   // preconditioner modes and its GEMMs' K-splits
    negf_scf_params p{};
    std::vector<double> K;
    std::vector<double*> hU, hUc;
    size_t state_stride = 0, hist_stride = 0;
    DevBuf d_K, d_Uext, d_V, d_KN, d_U, d_Q, d_Uo, d_Qo, d_ptrU, d_ptrUc, d_state, d_hist, d_out, d_steps,
        d_active;
    void* h_back = nullptr;   // mapped read-back record (host address)
    void* d_back = nullptr;   // the same memory as the device sees it
    bool has_ext = false, use_graph = true;
    hipStream_t st = nullptr;
    hipGraphExec_t gexec = nullptr;
};

// The potential of one SCF iteration written where the loop and the tube
// read it: U_i = U_ext_i - V - sum_j K_ij n_j (the sum as in
// coulomb_potential_kernel) into U and tU, the contact potentials U_0 and
// U_{N-1} into tUc.  V is read from the device so that one captured
// iteration serves every bias point.
__global__ __launch_bounds__(256) void scf_potential_kernel(int N, const double* __restrict__ K,
                                                            const double* __restrict__ n,
                                                            const double* __restrict__ U_ext,
                                                            const double* __restrict__ V, double* __restrict__ U,
                                                            double* __restrict__ tU, double* __restrict__ tUc)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= N) return;
    const double* __restrict__ k = K + (size_t)row * N;
    double acc = 0.0;
#pragma unroll 4
    for (int j = lane; j < N; j += 64) acc = fma(k[j], n[j], acc);
    acc = wave_sum(acc);
    if (lane == 0) {
        const double u = (U_ext ? U_ext[row] : 0.0) - V[0] - acc;
        U[row] = u;
        tU[row] = u;
        if (row == 0) tUc[0] = u;
        if (row == N - 1) tUc[1] = u;
    }
}

static void scf_enqueue_potential(Scf& s, hipStream_t q)
{
    Tube& t = *s.tube;
    hipLaunchKernelGGL(scf_potential_kernel, dim3((t.N + 3) / 4), dim3(256), 0, q, t.N, s.d_K.as<double>(),
                       s.br.dev().n_in, s.has_ext ? s.d_Uext.as<double>() : nullptr, s.d_V.as<double>(),
                       s.d_U.as<double>(), t.d_U.as<double>(), t.d_Uc.as<double>());
}

// With part: the tube's equilibrium charge from the plan's partial rows
// (rows r = 0..R-1 of site s at part[r S + s], summed in order from 0.0 as
// the plan's own row sum of at most 16 rows does) into eq, its charge
// Q = (eq + noneq) - rho0 (as combine_charge_kernel) into n_out; without,
// n_out holds Q already.  Then one Broyden update with the step counter on
// the device (the host checks the step limit before each launch), then the
// norms and the plans' status words to the host's read-back record.  Each
// thread forms the Q entries that it reads first in the update.
__global__ __launch_bounds__(kBroydenThreads) void broyden_scf_kernel(int S, double fraction, int norm_type,
                                                                        const double* __restrict__ part, int R,
                                                                        double* __restrict__ eq,
                                                                        const double* __restrict__ neq,
                                                                        const double* __restrict__ rho0,
                                                                        double* n_out, BroydenDev b,
                                                                        int* __restrict__ step,
                                                                        const int* __restrict__ st_eq,
                                                                        const int* __restrict__ st_neq,
                                                                        ScfBack* __restrict__ back)
{
    if (part) {
        for (int s = threadIdx.x; s < S; s += kBroydenThreads) {
            double e = 0.0;
#pragma unroll 4
            for (int r = 0; r < R; ++r) e += part[(size_t)r * S + s];
            eq[s] = e;
            n_out[s] = (e + (neq ? neq[s] : 0.0)) - rho0[s];
        }
    }
    const int m = *step - 1;
    broyden_update<false>(S, m, fraction, norm_type, n_out, nullptr, b);
    if (threadIdx.x == 0) {   // b.out[0..1] were written by this thread
        *step = m + 2;
        back->norm = b.out[0];
        back->l2 = b.out[1];
        back->status[0] = *st_eq;
        back->status[1] = st_neq ? *st_neq : 0;
        __threadfence_system();
    }
}

// Potential, solve, Broyden update and read-back: nothing on the host.
static int scf_enqueue_iteration(Scf& s, hipStream_t q)
{
    Tube& t = *s.tube;
    scf_enqueue_potential(s, q);
    NEGF_HIP_TRY(hipGetLastError());
    // the row sum and the charge go into the Broyden kernel when the plan's
    // own row sum is the in-order one (at most 16 rows)
    const double* part = nullptr;
    int R = 0;
    const bool fuse = t.eq_rows(&part, &R) && R <= 16;
    if (!fuse) part = nullptr;
    int rc = t.enqueue_solve(s.d_Q.as<double>(), q, !fuse);
    if (rc) return rc;
    hipLaunchKernelGGL(broyden_scf_kernel, dim3(1), dim3(kBroydenThreads), 0, q, s.br.S, s.br.fraction,
                       s.br.norm_type, part, R, t.d_eq.as<double>(),
                       t.neq_status() ? t.d_neq.as<const double>() : nullptr, t.d_rho0.as<const double>(),
                       s.d_Q.as<double>(), s.br.dev(), s.d_step.as<int>(), t.eq_status(), t.neq_status(),
                       static_cast<ScfBack*>(s.d_back));
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

// Capture one iteration unless the graph was captured over the tube's
// present plans and external-potential setting.
static int scf_ensure_graph(Scf& s)
{
    if (s.gexec && s.g_paths == s.tube->paths_version && s.g_ext == s.has_ext) return NEGF_OK;
    if (s.gexec) (void)hipGraphExecDestroy(s.gexec);
    s.gexec = nullptr;
    NEGF_HIP_TRY(hipStreamBeginCapture(s.st, hipStreamCaptureModeThreadLocal));
    const int erc = scf_enqueue_iteration(s, s.st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s.st, &g);
    if (erc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        s.tube->neq_zero = false;   // a zero fill may have been captured and dropped
        if (erc) return erc;
        set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
        return NEGF_EDEVICE;
    }
    const hipError_t e2 = hipGraphInstantiate(&s.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) {
        s.gexec = nullptr;
        s.tube->neq_zero = false;
        set_error("hipGraphInstantiate: %s", hipGetErrorString(e2));
        return NEGF_EDEVICE;
    }
    s.g_paths = s.tube->paths_version;
    s.g_ext = s.has_ext;
    return NEGF_OK;
}

// One iteration replayed from the graph: the host part of the solve first
// (with an adaptive correction due, the integrand is sampled at this
// iteration's potential, enqueued and finished beforehand), then one launch
// and one synchronisation.
static int scf_graph_iteration(Scf& s, int it, double* norm, double* l2)
{
    Tube& t = *s.tube;
    int rc;
    if (t.p.flag_adaptive_integration_limits && it % t.p.integrand_correction_interval == 0) {
        scf_enqueue_potential(s, s.st);
        NEGF_HIP_TRY(hipGetLastError());
        NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    }
    if ((rc = t.prepare(it))) return rc;
    if ((rc = scf_ensure_graph(s))) return rc;
    volatile ScfBack* back = static_cast<volatile ScfBack*>(s.h_back);
    back->norm = std::nan("");
    back->l2 = std::nan("");
    back->status[0] = 0;
    back->status[1] = 0;
    NEGF_HIP_TRY(hipGraphLaunch(s.gexec, s.st));
    NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    if ((back->status[0] | back->status[1]) && (rc = t.check_status())) return rc;
    const double o[2] = {back->norm, back->l2};
    *norm = o[0];
    *l2 = o[1];
    s.br.step += 1;
    if (!std::isfinite(o[0]) || !std::isfinite(o[1])) {
        set_error("Broyden: non-finite norm");
        return NEGF_ENAN;
    }
    return NEGF_OK;
}

}  // namespace negf

using namespace negf;

struct negf_broyden {
    Broyden b;
    hipStream_t st = nullptr;
};
struct negf_scf {
    Scf s;
};

extern "C" {

int negf_broyden_create(int S, double fraction, int max_step, int norm_type, negf_broyden** out)
{
    clear_error();
    if (!out || S <= 0 || max_step < 0 || (norm_type != 0 && norm_type != 1)) {
        set_error("invalid Broyden arguments");
        return NEGF_EARG;
    }
    negf_broyden* h = new negf_broyden;
    hipError_t e = hipStreamCreate(&h->st);
    if (e != hipSuccess) {
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        delete h;
        return NEGF_EDEVICE;
    }
    int rc = h->b.init(S, fraction, max_step, norm_type);
    if (rc == NEGF_OK) rc = hipStreamSynchronize(h->st) == hipSuccess ? NEGF_OK : NEGF_EDEVICE;
    if (rc) {
        (void)hipStreamDestroy(h->st);
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_broyden_destroy(negf_broyden* b)
{
    if (!b) return NEGF_EARG;
    (void)hipStreamSynchronize(b->st);
    (void)hipStreamDestroy(b->st);
    delete b;
    return NEGF_OK;
}

int negf_broyden_set_nin(negf_broyden* b, const double* n_in)
{
    clear_error();
    if (!b || !n_in) {
        set_error("null argument");
        return NEGF_EARG;
    }
    NEGF_HIP_TRY(hipStreamSynchronize(b->st));
    NEGF_HIP_TRY(hipMemcpy(b->b.dev().n_in, n_in, sizeof(double) * b->b.S, hipMemcpyHostToDevice));
    return NEGF_OK;
}

int negf_broyden_get_nin(const negf_broyden* b, double* n_in)
{
    clear_error();
    if (!b || !n_in) {
        set_error("null argument");
        return NEGF_EARG;
    }
    NEGF_HIP_TRY(hipStreamSynchronize(b->st));
    NEGF_HIP_TRY(hipMemcpy(n_in, b->b.dev().n_in, sizeof(double) * b->b.S, hipMemcpyDeviceToHost));
    return NEGF_OK;
}

int negf_broyden_step(negf_broyden* b, const double* n_out, double* max_norm, double* l2, double* norm_per_site)
{
    clear_error();
    if (!b || !n_out) {
        set_error("null argument");
        return NEGF_EARG;
    }
    NEGF_HIP_TRY(hipMemcpyAsync(b->b.d_nout.p, n_out, sizeof(double) * b->b.S, hipMemcpyHostToDevice, b->st));
    int rc = b->b.update(b->b.d_nout.as<double>(), max_norm, l2, b->st);
    if (rc) return rc;
    if (norm_per_site)
        NEGF_HIP_TRY(hipMemcpy(norm_per_site, b->b.dev().Norm, sizeof(double) * b->b.S, hipMemcpyDeviceToHost));
    return NEGF_OK;
}

int negf_broyden_step_device(negf_broyden* b, const double* d_n_out, double* max_norm, double* l2, void* stream)
{
    clear_error();
    if (!b || !d_n_out) {
        set_error("null argument");
        return NEGF_EARG;
    }
    hipStream_t st = stream ? (hipStream_t)stream : b->st;
    return b->b.update(d_n_out, max_norm, l2, st);
}

const double* negf_broyden_device_nin(const negf_broyden* b) { return b ? b->b.dev().n_in : nullptr; }

int negf_broyden_reset(negf_broyden* b, int keep_n_in, double init_value)
{
    clear_error();
    if (!b) {
        set_error("null argument");
        return NEGF_EARG;
    }
    int rc = b->b.reset(keep_n_in, init_value, b->st);
    if (rc) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(b->st));
    return NEGF_OK;
}

int negf_broyden_get_step(const negf_broyden* b) { return b ? b->b.step : -1; }

int negf_scf_create(negf_tube* t, const negf_scf_params* p, negf_scf** out)
{
    clear_error();
    if (!t || !p || !out || p->eps_r <= 0 || p->Broyden_max_norm <= 0 || p->Broyden_threshold_maxstep < 0) {
        set_error("invalid SCF arguments");
        return NEGF_EARG;
    }
    negf_scf* h = new negf_scf;
    Scf& s = h->s;
    s.tube = &t->t;
    s.p = *p;
    const int N = s.tube->N;
    int rc = NEGF_OK;
    hipError_t e = hipStreamCreate(&s.st);
    if (e != hipSuccess) {
        set_error("hipStreamCreate: %s", hipGetErrorString(e));
        delete h;
        return NEGF_EDEVICE;
    }
    coulomb_kernel(s.tube->ptd, s.p, s.K);
    if (!rc) rc = s.d_K.ensure(sizeof(double) * (size_t)N * N);
    if (!rc) rc = s.d_U.ensure(sizeof(double) * N);
    if (!rc) rc = s.d_Q.ensure(sizeof(double) * N);
    if (!rc) rc = s.br.init(N, p->Broyden_fraction, p->Broyden_threshold_maxstep, p->Broyden_norm_type);
    if (!rc) rc = s.d_V.ensure(sizeof(double));
    if (!rc) rc = s.d_step.ensure(sizeof(int));
    if (!rc) {
        hipError_t eh = hipHostMalloc(&s.h_back, sizeof(ScfBack), hipHostMallocMapped | hipHostMallocCoherent);
        if (eh == hipSuccess) eh = hipHostGetDevicePointer(&s.d_back, s.h_back, 0);
        if (eh != hipSuccess) {
            set_error("mapped host buffer for the SCF read-back: %s", hipGetErrorString(eh));
            rc = NEGF_EALLOC;
        }
    }
    if (!rc) {
        // Set_Broyden_Parallel: n_in starts at the initial deposit value
        rc = s.br.reset(0, p->NS_initial_deposit_value, s.st);
    }
    if (!rc && hipMemcpy(s.d_K.p, s.K.data(), sizeof(double) * s.K.size(), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("hipMemcpy of the electrostatic kernel failed");
        rc = NEGF_EDEVICE;
    }
    if (!rc && hipStreamSynchronize(s.st) != hipSuccess) rc = NEGF_EDEVICE;
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_scf_destroy(negf_scf* s)
{
    if (!s) return NEGF_EARG;
    delete s;
    return NEGF_OK;
}

int negf_scf_kernel(const negf_scf* s, double* K)
{
    clear_error();
    if (!s || !K) {
        set_error("null argument");
        return NEGF_EARG;
    }
    std::copy(s->s.K.begin(), s->s.K.end(), K);
    return NEGF_OK;
}

int negf_scf_use_graph(negf_scf* h, int on)
{
    clear_error();
    if (!h) {
        set_error("null argument");
        return NEGF_EARG;
    }
    h->s.use_graph = on != 0;
    return NEGF_OK;
}

int negf_scf_set_external_potential(negf_scf* h, const double* U_ext)
{
    clear_error();
    if (!h) {
        set_error("null argument");
        return NEGF_EARG;
    }
    Scf& s = h->s;
    if (!U_ext) {
        s.has_ext = false;
        return NEGF_OK;
    }
    int rc = s.d_Uext.ensure(sizeof(double) * s.tube->N);
    if (rc) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    NEGF_HIP_TRY(hipMemcpy(s.d_Uext.p, U_ext, sizeof(double) * s.tube->N, hipMemcpyHostToDevice));
    s.has_ext = true;
    return NEGF_OK;
}

// c_TransportSolver::Solve (Transport.cpp:359-539): iterate electrostatics ->
// NEGF -> Broyden until Broyden_Norm <= Broyden_max_norm, then
// Reset_ForNextBiasStep.
int negf_scf_bias_point(negf_scf* h, double V_gate, double* U_out, double* Q_out, int* iters, double* final_norm)
{
    clear_error();
    if (!h || !U_out || !Q_out) {
        set_error("null argument");
        return NEGF_EARG;
    }
    Scf& s = h->s;
    Tube& t = *s.tube;
    const int N = t.N;
    double norm = 1.0, l2 = 0.0;
    int it = 0, rc = NEGF_OK;
    if (s.use_graph) {
        const int step = s.br.step;
        NEGF_HIP_TRY(hipStreamSynchronize(s.st));
        NEGF_HIP_TRY(hipMemcpy(s.d_V.p, &V_gate, sizeof(double), hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(s.d_step.p, &step, sizeof(int), hipMemcpyHostToDevice));
    }
    do {
        if (s.br.step > s.p.Broyden_threshold_maxstep) {
            set_error("Broyden_Step has exceeded the Broyden_Threshold_MaxStep");
            rc = NEGF_ENONCONV;
            break;
        }
        if (s.use_graph) {
            if ((rc = scf_graph_iteration(s, it, &norm, &l2))) break;
            ++it;
            continue;
        }
        hipLaunchKernelGGL(coulomb_potential_kernel, dim3((N + 3) / 4), dim3(256), 0, s.st, N, s.d_K.as<double>(),
                           s.br.dev().n_in, s.has_ext ? s.d_Uext.as<double>() : nullptr, V_gate,
                           s.d_U.as<double>());
        NEGF_HIP_TRY(hipGetLastError());
        if ((rc = t.set_potential_device(s.d_U.as<double>(), s.st))) break;
        if ((rc = t.solve(it, s.d_Q.as<double>(), s.st))) break;
        if ((rc = s.br.update(s.d_Q.as<double>(), &norm, &l2, s.st))) break;
        if ((rc = t.check_status())) break;
        ++it;
    } while (norm > s.p.Broyden_max_norm);
    NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    NEGF_HIP_TRY(hipMemcpy(U_out, s.d_U.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    NEGF_HIP_TRY(hipMemcpy(Q_out, s.d_Q.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (iters) *iters = it;
    if (final_norm) *final_norm = norm;
    if (rc) return rc;
    // Reset_ForNextBiasStep
    if ((rc = s.br.reset(s.p.reset_with_previous_charge_distribution, s.p.NS_initial_deposit_value, s.st))) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    return NEGF_OK;
}

struct negf_batch_scf {
    BatchScf s;
};

int negf_batch_scf_create(negf_batch* b, negf_tube* const* tubes, int ntubes, const negf_scf_params* p,
                          negf_batch_scf** out)
{
    clear_error();
    if (!b || !tubes || ntubes <= 0 || !p || !out || p->eps_r <= 0 || p->Broyden_max_norm <= 0 ||
        p->Broyden_threshold_maxstep < 0) {
        set_error("invalid batched SCF arguments");
        return NEGF_EARG;
    }
    negf_batch_scf* h = new negf_batch_scf;
    int rc = h->s.init(b, tubes, ntubes, *p);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_batch_scf_destroy(negf_batch_scf* s)
{
    if (!s) return NEGF_EARG;
    delete s;
    return NEGF_OK;
}

int negf_batch_scf_set_external_potential(negf_batch_scf* s, const double* U_ext)
{
    clear_error();
    if (!s) {
        set_error("null argument");
        return NEGF_EARG;
    }
    // The captured iteration reads d_Uext by address, so setting or clearing
    // U_ext only rewrites that buffer (zeros when cleared): no graph goes stale.
    BatchScf& b = s->s;
    const long long NT = (long long)b.N * b.T;
    NEGF_HIP_TRY(hipStreamSynchronize(b.st));
    if (!U_ext) {
        int rc = fill_device(b.d_Uext.as<double>(), NT, 0.0, b.st);
        if (rc) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(b.st));
        b.has_ext = false;
        return NEGF_OK;
    }
    for (long long i = 0; i < NT; ++i)
        if (!std::isfinite(U_ext[i])) {
            set_error("U_ext[%lld] is not finite", i);
            return NEGF_EARG;
        }
    NEGF_HIP_TRY(hipMemcpy(b.d_Uext.p, U_ext, sizeof(double) * NT, hipMemcpyHostToDevice));
    b.has_ext = true;
    return NEGF_OK;
}

int negf_batch_scf_set_input_charge(negf_batch_scf* s, const double* n_in)
{
    clear_error();
    if (!s || !n_in) {
        set_error("null argument");
        return NEGF_EARG;
    }
    BatchScf& b = s->s;
    for (int t = 0; t < b.T; ++t)
        NEGF_HIP_TRY(hipMemcpy(b.d_state.as<double>() + t * b.state_stride, n_in + (size_t)t * b.N,
                               sizeof(double) * b.N, hipMemcpyHostToDevice));
    return NEGF_OK;
}

int negf_batch_scf_get_input_charge(const negf_batch_scf* s, double* n_in)
{
    clear_error();
    if (!s || !n_in) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const BatchScf& b = s->s;
    NEGF_HIP_TRY(hipStreamSynchronize(b.st));
    for (int t = 0; t < b.T; ++t)
        NEGF_HIP_TRY(hipMemcpy(n_in + (size_t)t * b.N, b.d_state.as<double>() + t * b.state_stride,
                               sizeof(double) * b.N, hipMemcpyDeviceToHost));
    return NEGF_OK;
}

int negf_batch_scf_use_graph(negf_batch_scf* s, int on)
{
    if (!s) return NEGF_EARG;
    s->s.use_graph = on != 0;
    if (!on) s->s.drop_graph();
    return NEGF_OK;
}

int negf_batch_scf_get_kernel(const negf_batch_scf* s, double* K)
{
    clear_error();
    if (!s || !K) {
        set_error("null argument");
        return NEGF_EARG;
    }
    std::copy(s->s.K.begin(), s->s.K.end(), K);
    return NEGF_OK;
}

int negf_batch_scf_set_preconditioner(negf_batch_scf* s, const double* Z, const double* lam)
{
    clear_error();
    if (!s || !Z || !lam) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return s->s.set_preconditioner(Z, lam);
}

int negf_batch_scf_set_surrogate(negf_batch_scf* s, negf_surrogate* sg, const double* feat)
{
    clear_error();
    if (!s || !sg) {
        set_error("null argument");
        return NEGF_EARG;
    }
    BatchScf& b = s->s;
    if (sg->s.N != b.N) {
        set_error("surrogate has %d sites, the tubes %d", sg->s.N, b.N);
        return NEGF_EARG;
    }
    b.sg = &sg->s;
    b.has_feat = feat != nullptr;
    if (feat) {
        int rc = b.d_feat.ensure(sizeof(double) * b.T);
        if (rc) return rc;
        NEGF_HIP_TRY(hipMemcpy(b.d_feat.p, feat, sizeof(double) * b.T, hipMemcpyHostToDevice));
    }
    b.drop_graph();
    return NEGF_OK;
}

int negf_batch_scf_run_surrogate(negf_batch_scf* s, const double* V_gate, int seed_iters, int refresh_iter,
                                 double* U_out, double* Q_out, int* iters, double* final_norms,
                                 int* total_iterations, double* capacitance)
{
    clear_error();
    if (!s || !V_gate || seed_iters < 0) {
        set_error("invalid arguments");
        return NEGF_EARG;
    }
    BatchScf& b = s->s;
    if (!b.sg || !b.has_pc) {
        set_error("set a surrogate and a preconditioner basis first");
        return NEGF_EARG;
    }
    int rc = fill_device(b.d_hist.as<double>(), (long long)(b.hist_stride * b.T), 0.0, b.st);
    if (rc) return rc;
    for (int t = 0; t < b.T; ++t)
        if ((rc = fill_device(b.d_state.as<double>() + t * b.state_stride + b.N, 6LL * b.N, 0.0, b.st))) return rc;
    return b.run_pc(V_gate, seed_iters, refresh_iter, U_out, Q_out, iters, final_norms, total_iterations,
                    capacitance);
}

int negf_batch_scf_run(negf_batch_scf* s, const double* V_gate, double* U_out, double* Q_out, int* iters,
                       double* final_norms, int* total_iterations)
{
    clear_error();
    if (!s || !V_gate) {
        set_error("null argument");
        return NEGF_EARG;
    }
    // each run starts a fresh Broyden history (Reset_Broyden_Parallel), keeping n_in
    BatchScf& b = s->s;
    int rc = fill_device(b.d_hist.as<double>(), (long long)(b.hist_stride * b.T), 0.0, b.st);
    if (rc) return rc;
    for (int t = 0; t < b.T; ++t)
        if ((rc = fill_device(b.d_state.as<double>() + t * b.state_stride + b.N, 6LL * b.N, 0.0, b.st))) return rc;
    return b.run(V_gate, U_out, Q_out, iters, final_norms, total_iterations);
}

}  // extern "C"
