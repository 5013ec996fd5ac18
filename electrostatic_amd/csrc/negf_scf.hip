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
    double* intermed;   // [max_step + 1]
    double* out;        // [0] max norm, [1] l2 norm, [2] |dF|^2
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

// intermed[it] = sum_s Vrow(it)[s] * x[s] for it in [1, m_hi]; wave w takes
// it = 1 + w, 1 + w + kWaves, ...
__device__ void history_dots(const BroydenDev& b, int S, int m_hi, const double* __restrict__ x, int wave,
                             int lane)
{
    for (int it = 1 + wave; it <= m_hi; it += kWaves) {
        const double* __restrict__ v = b.V + (size_t)it * S;
        double acc = 0.0;
        for (int s = lane; s < S; s += 64) acc = fma(v[s], x[s], acc);
        acc = wave_sum(acc);
        if (lane == 0) b.intermed[it] = acc;
    }
}

__global__ __launch_bounds__(kBroydenThreads) void broyden_step_kernel(int S, int m, double fraction, int norm_type,
                                                                         const double* __restrict__ n_out,
                                                                         BroydenDev b)
{
    __shared__ double red[3][kWaves];
    __shared__ double hist[512];
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

    if (m > 0) {
        // sum_vec = W^T (V dF) over the previous updates (:106-135)
        history_dots(b, S, m - 1, b.dF, wave, lane);
        __syncthreads();
        for (int it = tid + 1; it <= m - 1 && it < 512; it += kBroydenThreads) hist[it] = b.intermed[it];
        __syncthreads();
        for (int s = tid; s < S; s += kBroydenThreads) {
            double sv = 0.0;
            for (int it = 1; it <= m - 1; ++it) sv = fma(b.W[(size_t)it * S + s], it < 512 ? hist[it] : b.intermed[it], sv);
            const double d = b.dF[s];
            b.V[(size_t)m * S + s] = d / denom;
            b.W[(size_t)m * S + s] = -fraction * d + (b.n_in[s] - b.n_prev[s]) - sv;
        }
        __syncthreads();
        // sum_vec = W^T (V F) including this update (:137-170)
        history_dots(b, S, m, b.F, wave, lane);
        __syncthreads();
        for (int it = tid + 1; it <= m && it < 512; it += kBroydenThreads) hist[it] = b.intermed[it];
        __syncthreads();
        for (int s = tid; s < S; s += kBroydenThreads) {
            double sv = 0.0;
            for (int it = 1; it <= m; ++it) sv = fma(b.W[(size_t)it * S + s], it < 512 ? hist[it] : b.intermed[it], sv);
            b.sum_vec[s] = sv;
        }
    }
    // new input (:172-186), scalar = 1
    for (int s = tid; s < S; s += kBroydenThreads) {
        const double prev = b.n_in[s];
        const double sv = m > 0 ? b.sum_vec[s] : 0.0;
        b.n_prev[s] = prev;
        b.n_in[s] = prev - fraction * b.F[s] - sv;
    }
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
        if ((rc = d_hist.ensure(sizeof(double) * (2 * (size_t)(max_step + 1) * S + (max_step + 1))))) return rc;
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
        b.out = d_out.as<double>();
        return b;
    }
    // Reset_Broyden_Parallel (Broyden_Parallel_General.cpp:250-355).
    int reset(int keep_n_in, double init_value, hipStream_t st)
    {
        BroydenDev b = dev();
        int rc;
        if (!keep_n_in && (rc = fill_device(b.n_in, S, init_value, st))) return rc;
        if ((rc = fill_device(b.n_prev, 5LL * S, 0.0, st))) return rc;
        if ((rc = fill_device(d_hist.as<double>(), 2LL * (max_step + 1) * S + (max_step + 1), 0.0, st))) return rc;
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
        for (DevBuf* d : {&d_K, &d_U, &d_Q, &d_Uext}) d->release();
        if (st) (void)hipStreamDestroy(st);
    }
    Tube* tube = nullptr;
    negf_scf_params p{};
    Broyden br;
    std::vector<double> K;
    DevBuf d_K, d_U, d_Q, d_Uext;
    bool has_ext = false;
    hipStream_t st = nullptr;
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
    do {
        if (s.br.step > s.p.Broyden_threshold_maxstep) {
            set_error("Broyden_Step has exceeded the Broyden_Threshold_MaxStep");
            rc = NEGF_ENONCONV;
            break;
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

}  // extern "C"
