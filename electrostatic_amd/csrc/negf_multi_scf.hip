// negf_multi_scf.hip -- self-consistent loop over a batch of tubes of any
// lengths with one Broyden system over the concatenated sites of all tubes,
// as c_TransportSolver runs several nanostructures: every tube's Solve_NEGF
// writes its block of n_curr_out at its site offset
// (Transport.cpp:417-431, Broyden_Parallel_General.cpp:42-66) and one Broyden
// update runs over all of them (Transport.cpp:435, :596-622).
//
// One iteration, replayed from a captured hipGraph:
//   1. electrostatic stand-in, tube by tube: U_i = U_ext,i - V_t - sum_j
//      K(x_i - x_j) n_j with the dense 1/r-with-image kernel of negf_scf,
//      applied as a direct sum from a table of its values (the ring spacing
//      alternates, so K depends only on i - j and the parity of i);
//   2. the batched NEGF solve of all tubes (one RGF launch sequence);
//   3. the Broyden update over all S sites in three launches: per-block
//      partial sums (norms, |dF|^2, dF.F and the history dot products
//      V_j.dF, V_j.F), one workgroup that adds the partials in block order
//      (fixed order: bit-reproducible), and the per-site history update.
// The Broyden step counter lives on the device, so the launches are the same
// every iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "negf_internal.h"
#include "negf_material.h"

namespace negf {

static constexpr int kMsThreads = 256;   // threads per workgroup
static constexpr int kMsBlocks = 1024;   // workgroups of the per-site Broyden launches

__device__ __forceinline__ double ms_wave_sum(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ __forceinline__ double ms_wave_max(double v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}
// Sum (or max) over a workgroup of kMsThreads in a fixed order; every thread
// gets the result.
template <bool MAX>
__device__ __forceinline__ double block_reduce(double v, double* sh)
{
    v = MAX ? ms_wave_max(v) : ms_wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int k = 1; k < kMsThreads / 64; ++k) r = MAX ? fmax(r, sh[k]) : r + sh[k];
    return r;
}

struct BigBroyden {
    double *n_in, *n_prev, *F, *dF;
    double *V, *W;  // [max_step + 1][S], rows 1.. used
    double* part;   // [blocks][4 + 2 max_step] partial sums of one update
    double* coef;   // [0] max norm [1] l2 [2] |dF|^2 [3] m; [4 ..] a_j, then b_j
    int* step;      // Broyden_Step (the reference's counter, starting at 1)
    long long S;
    int max_step;
};

// Pass 1 (Execute_Broyden_Modified_Second_Algorithm_Parallel
// Broyden_Second_Parallel.cpp:31-170 in the CPU order of :9-211): F = n_in -
// n_out, the site norms, dF = F - F_previous, and per workgroup the partial
// sums of |F|^2, max|F|, |dF|^2, dF.F and, for the history rows j < m,
// V_j.dF and V_j.F.
__global__ __launch_bounds__(kMsThreads) void big_broyden_pass1(BigBroyden b, int norm_type,
                                                               const double* __restrict__ n_out)
{
    __shared__ double sh[kMsThreads / 64];
    const int m = *b.step - 1;
    const long long S = b.S;
    const long long stride = (long long)gridDim.x * kMsThreads;
    const long long s0 = (long long)blockIdx.x * kMsThreads + threadIdx.x;
    double nsum = 0.0, nmax = 0.0, den = 0.0, dff = 0.0;
    for (long long s = s0; s < S; s += stride) {
        const double Fc = b.n_in[s] - n_out[s];
        double nrm;
        if (norm_type == 0) {
            nrm = fabs(Fc);
            nsum = fma(Fc, Fc, nsum);
        } else {
            nrm = fabs(Fc / (b.n_in[s] + n_out[s]));
            nsum = fma(nrm, nrm, nsum);
        }
        nmax = fmax(nmax, nrm);
        const double d = Fc - b.F[s];
        b.dF[s] = d;
        b.F[s] = Fc;
        den = fma(d, d, den);
        dff = fma(d, Fc, dff);
    }
    double* out = b.part + (size_t)blockIdx.x * (4 + 2 * b.max_step);
    const double r0 = block_reduce<false>(nsum, sh);
    const double r1 = block_reduce<true>(nmax, sh);
    const double r2 = block_reduce<false>(den, sh);
    const double r3 = block_reduce<false>(dff, sh);
    if (threadIdx.x == 0) {
        out[0] = r0;
        out[1] = r1;
        out[2] = r2;
        out[3] = r3;
    }
    for (int j = 1; j < m; ++j) {
        const double* __restrict__ v = b.V + (size_t)j * S;
        double a = 0.0, c = 0.0;
        for (long long s = s0; s < S; s += stride) {
            a = fma(v[s], b.dF[s], a);
            c = fma(v[s], b.F[s], c);
        }
        const double ra = block_reduce<false>(a, sh);
        const double rc = block_reduce<false>(c, sh);
        if (threadIdx.x == 0) {
            out[4 + 2 * (j - 1)] = ra;
            out[5 + 2 * (j - 1)] = rc;
        }
    }
}

// Pass 2 (one workgroup): the partials of pass 1 added in a fixed order
// (thread-strided over the blocks, then the fixed workgroup reduction), the
// coefficients of the history update, and the step counter advanced.
__global__ __launch_bounds__(kMsThreads) void big_broyden_pass2(BigBroyden b, int nblocks)
{
    __shared__ double sh[kMsThreads / 64];
    __shared__ double tot[4];
    const int m = *b.step - 1;
    const int stride = 4 + 2 * b.max_step;
    const int nq = 4 + 2 * (m > 1 ? m - 1 : 0);
    for (int q = 0; q < nq; ++q) {
        double acc = 0.0;
        for (int k = threadIdx.x; k < nblocks; k += kMsThreads) {
            const double x = b.part[(size_t)k * stride + q];
            acc = (q == 1) ? fmax(acc, x) : acc + x;
        }
        acc = (q == 1) ? block_reduce<true>(acc, sh) : block_reduce<false>(acc, sh);
        if (threadIdx.x == 0) {
            if (q < 4) {
                tot[q] = acc;
            } else {
                const int j = 1 + (q - 4) / 2;
                if ((q - 4) % 2 == 0)
                    b.coef[4 + j] = acc;                   // a_j = V_j . dF
                else
                    b.coef[4 + b.max_step + 1 + j] = acc;  // b_j = V_j . F (j < m)
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double denom = tot[2];
        b.coef[0] = tot[1];
        b.coef[1] = sqrt(tot[0]);
        b.coef[2] = denom;
        b.coef[3] = (double)m;
        if (m > 0) b.coef[4 + b.max_step + 1 + m] = tot[3] / denom;  // V_m . F with V_m = dF / |dF|^2
        *b.step = m + 2;
    }
}

// Pass 3 (:106-186): W_m = -fraction dF + (n_in - n_prev) - sum_{j<m} W_j a_j,
// V_m = dF / |dF|^2, then n_in -= fraction F + sum_{j<=m} W_j b_j.
__global__ __launch_bounds__(kMsThreads) void big_broyden_pass3(BigBroyden b, double fraction)
{
    const int m = (int)b.coef[3];
    const double denom = b.coef[2];
    const double* __restrict__ a = b.coef + 4;
    const double* __restrict__ bb = b.coef + 4 + b.max_step + 1;
    const long long S = b.S;
    const long long stride = (long long)gridDim.x * kMsThreads;
    for (long long s = (long long)blockIdx.x * kMsThreads + threadIdx.x; s < S; s += stride) {
        const double prev = b.n_in[s];
        double sv = 0.0;
        if (m > 0) {
            double sa = 0.0;
            for (int j = 1; j <= m - 1; ++j) sa = fma(b.W[(size_t)j * S + s], a[j], sa);
            const double d = b.dF[s];
            const double wm = -fraction * d + (prev - b.n_prev[s]) - sa;
            b.V[(size_t)m * S + s] = d / denom;
            b.W[(size_t)m * S + s] = wm;
            for (int j = 1; j <= m - 1; ++j) sv = fma(b.W[(size_t)j * S + s], bb[j], sv);
            sv = fma(wm, bb[m], sv);
        }
        b.n_prev[s] = prev;
        b.n_in[s] = prev - fraction * b.F[s] - sv;
    }
}

__global__ void big_broyden_first_step(int* step) { *step = 1; }

// The iteration's norms (coef[0..1]) and the batch plans' status words into
// the host's mapped, coherent read-back record: no copy in the iteration.
__global__ void multi_readback_kernel(const double* __restrict__ coef, const int* __restrict__ st_eq,
                                      const int* __restrict__ st_neq, double* __restrict__ back)
{
    if (threadIdx.x == 0) {
        back[0] = coef[0];
        back[1] = coef[1];
        int* bs = reinterpret_cast<int*>(back + 2);
        bs[0] = *st_eq;
        bs[1] = st_neq ? *st_neq : 0;
        __threadfence_system();
    }
}

// Electrostatic stand-in: U_i = U_ext,i - V_t - sum_j K(i, j) n_j within each
// tube, K(i, j) = kt[i % 2][i - j + D].  The (rows x columns) plane of each
// tube is cut into tiles of MS_ROWS x MS_TILE; a workgroup sums one tile's
// columns for its rows into a partial row (partial[c][i] for column tile c),
// and a second launch adds the partials of a row in column-tile order (a
// fixed order) and forms U.  Within a tile, thread (q, p) (p = thread % 2)
// owns the MS_RPT rows row0 + p + 2 (MS_RPT q + r) of parity p, whose kernel
// values for 8 consecutive columns lie in a window of 2 MS_RPT + 6
// consecutive entries of table p: per group of 8 columns a thread loads that
// window and the 8 charges once from LDS and does 8 MS_RPT FMAs.  Columns
// are staged MS_COLS at a time (charges and the two table stretches they
// need, padded by one double every 16 against LDS bank conflicts).
#define MS_RPT 8
#define MS_ROWS (kMsThreads * MS_RPT)
#define MS_TILE 2048
#define MS_COLS 256
#define MS_SPAN (MS_ROWS + MS_COLS - 1)
#define MS_PAD(e) ((e) + ((e) >> 4))
struct MsTile {
    int tube, row0, col0;
    long long part;  // partial row of this tile's column block: partial + part + i
};
struct MsRows {
    int tube, row0, ntiles;
    long long part;  // the tube's partial rows: partial + part + c N + i
};
__global__ __launch_bounds__(kMsThreads) void multi_potential_tile_kernel(
    const MsTile* __restrict__ tiles, const long long* __restrict__ off, const double* __restrict__ kt, int D,
    const double* __restrict__ n, double* __restrict__ partial)
{
    __shared__ double ns[MS_COLS];
    __shared__ double ks[2][MS_PAD(MS_SPAN) + 1];
    const MsTile tl = tiles[blockIdx.x];
    const long long o = off[tl.tube];
    const int N = (int)(off[tl.tube + 1] - o);
    const int jend = min(N, tl.col0 + MS_TILE);
    const int p = threadIdx.x & 1, q = threadIdx.x >> 1;
    const int ib = tl.row0 + p + 2 * MS_RPT * q;  // first row of this thread (row0 even)
    double acc[MS_RPT];
#pragma unroll
    for (int r = 0; r < MS_RPT; ++r) acc[r] = 0.0;
    for (int j0 = tl.col0; j0 < jend; j0 += MS_COLS) {
        // table entries d = dlo + e, e in [0, MS_SPAN), for both parities
        const int dlo = tl.row0 - j0 - (MS_COLS - 1);
        __syncthreads();
        ns[threadIdx.x] = (j0 + (int)threadIdx.x < jend) ? n[o + j0 + threadIdx.x] : 0.0;
        for (int e = threadIdx.x; e < 2 * MS_SPAN; e += kMsThreads) {
            const int pp = e >= MS_SPAN, ee = e - pp * MS_SPAN;
            const int d = dlo + ee;
            ks[pp][MS_PAD(ee)] = (d >= -D && d <= D) ? kt[(size_t)pp * (2 * D + 1) + d + D] : 0.0;
        }
        __syncthreads();
        const double* __restrict__ kp = ks[p];
        // entry of d = ib - j0 - jg - 7 + k is e = ib - row0 - jg + MS_COLS - 8 + k
        const int e0 = ib - tl.row0 + MS_COLS - 8;
        for (int jg = 0; jg < MS_COLS; jg += 8) {
            double w[2 * MS_RPT + 6], x[8];
#pragma unroll
            for (int k = 0; k < 2 * MS_RPT + 6; ++k) w[k] = kp[MS_PAD(e0 - jg + k)];
#pragma unroll
            for (int u = 0; u < 8; ++u) x[u] = ns[jg + u];
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int r = 0; r < MS_RPT; ++r) acc[r] = fma(w[2 * r - u + 7], x[u], acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < MS_RPT; ++r) {
        const int i = ib + 2 * r;
        if (i < N) partial[tl.part + i] = acc[r];
    }
}

// U from the partial rows (column tiles in order), into the concatenated U,
// the tube's own potential and its contact potentials U_0, U_{N-1}.
__global__ __launch_bounds__(kMsThreads) void multi_potential_finish(
    const MsRows* __restrict__ rows, const long long* __restrict__ off, const double* __restrict__ partial,
    const double* __restrict__ Uext, const double* __restrict__ V, double* const* __restrict__ dU,
    double* const* __restrict__ dUc, double* __restrict__ Uall)
{
    const MsRows rw = rows[blockIdx.x];
    const long long o = off[rw.tube];
    const int N = (int)(off[rw.tube + 1] - o);
    const int i = rw.row0 + threadIdx.x;
    if (i >= N) return;
    double acc = 0.0;
    for (int c = 0; c < rw.ntiles; ++c) acc += partial[rw.part + (long long)c * N + i];
    const double u = Uext[o + i] - V[rw.tube] - acc;
    Uall[o + i] = u;
    dU[rw.tube][i] = u;
    if (i == 0) dUc[rw.tube][0] = u;
    if (i == N - 1) dUc[rw.tube][1] = u;
}

class MultiScf {
   public:
    ~MultiScf()
    {
        drop_graph();
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        if (h_back) (void)hipHostFree(h_back);
    }
    int init(negf_batch* b, negf_tube* const* tb, int T_, const negf_scf_params& prm)
    {
        batch = b;
        T = T_;
        p = prm;
        M = p.Broyden_threshold_maxstep;
        if (M < 1) {
            set_error("Broyden_threshold_maxstep must be positive");
            return NEGF_EARG;
        }
        tubes.assign(tb, tb + T_);
        off.assign(T + 1, 0);
        NEGF_HIP_TRY(hipStreamCreate(&st));
        int rc;
        std::vector<long long> boff(T + 1);
        if ((rc = negf_batch_offsets(batch, boff.data()))) return rc;
        int Nmax = 0;
        const double acc = tubes[0]->t.p.acc;
        std::vector<double*> hU(T), hUc(T);
        std::vector<MsTile> tiles;
        std::vector<MsRows> rows;
        long long pbase = 0;
        for (int t = 0; t < T; ++t) {
            const Tube& tu = tubes[t]->t;
            off[t + 1] = off[t] + tu.N;
            if (boff[t + 1] != off[t + 1]) {
                set_error("the tubes must be those of the batch, in the batch's order");
                return NEGF_EARG;
            }
            if (tu.p.acc != acc) {
                set_error("all tubes of the electrostatic stand-in need the same bond length acc");
                return NEGF_EARG;
            }
            Nmax = std::max(Nmax, tu.N);
            hU[t] = tu.d_U.as<double>();
            hUc[t] = tu.d_Uc.as<double>();
            const int C = (tu.N + MS_TILE - 1) / MS_TILE;
            for (int r = 0; r < tu.N; r += MS_ROWS)
                for (int c = 0; c < C; ++c)
                    tiles.push_back(MsTile{t, r, c * MS_TILE, pbase + (long long)c * tu.N});
            for (int r = 0; r < tu.N; r += kMsThreads) rows.push_back(MsRows{t, r, C, pbase});
            pbase += (long long)C * tu.N;
        }
        S = off[T];
        ntiles = (int)tiles.size();
        nrows = (int)rows.size();
        // Kernel table: ring r sits at the axial position (acc / 4) (3 r + 4 -
        // (r % 2)) plus the tube's centring (ring_positions of the tube), so
        // x_i - x_j = (acc / 4) (3 d - (i % 2) + (j % 2)) with d = i - j.
        D = Nmax - 1;
        std::vector<double> kt(2 * (size_t)(2 * D + 1));
        const double q = 1.602176634e-19, eps0 = 8.8541878128e-12, pi = 3.14159265358979323846;
        const double ke = q / (4.0 * pi * eps0) * 1.0e9;
        for (int pi_ = 0; pi_ < 2; ++pi_)
            for (int d = -D; d <= D; ++d) {
                const int pj = ((pi_ - d) % 2 + 2) % 2;
                const double x = 0.25 * acc * (3.0 * d - pi_ + pj) / 1.0e-9;
                const double x2 = x * x;
                kt[(size_t)pi_ * (2 * D + 1) + d + D] =
                    ke / p.eps_r *
                    (1.0 / std::sqrt(x2 + p.R_ring_nm * p.R_ring_nm) - 1.0 / std::sqrt(x2 + p.R_image_nm * p.R_image_nm));
            }
        const size_t hist = (size_t)(M + 1) * S;
        if ((rc = d_kt.ensure(sizeof(double) * kt.size()))) return rc;
        if ((rc = d_off.ensure(sizeof(long long) * (T + 1)))) return rc;
        if ((rc = d_tiles.ensure(sizeof(MsTile) * tiles.size()))) return rc;
        if ((rc = d_rows.ensure(sizeof(MsRows) * rows.size()))) return rc;
        if ((rc = d_partial.ensure(sizeof(double) * pbase))) return rc;
        if ((rc = d_ptrU.ensure(sizeof(double*) * T))) return rc;
        if ((rc = d_ptrUc.ensure(sizeof(double*) * T))) return rc;
        if ((rc = d_Uext.ensure(sizeof(double) * S))) return rc;
        if ((rc = d_V.ensure(sizeof(double) * T))) return rc;
        if ((rc = d_U.ensure(sizeof(double) * S))) return rc;
        if ((rc = d_Q.ensure(sizeof(double) * S))) return rc;
        if ((rc = d_state.ensure(sizeof(double) * 4 * S))) return rc;
        if ((rc = d_hist.ensure(sizeof(double) * 2 * hist))) return rc;
        if ((rc = d_part.ensure(sizeof(double) * kMsBlocks * (4 + 2 * (size_t)M)))) return rc;
        if ((rc = d_coef.ensure(sizeof(double) * (6 + 2 * (size_t)M)))) return rc;
        if ((rc = d_step.ensure(sizeof(int)))) return rc;
        {
            // read-back record: 2 doubles (norms), then 2 status ints
            hipError_t eh = hipHostMalloc(&h_back, 2 * sizeof(double) + 2 * sizeof(int),
                                          hipHostMallocMapped | hipHostMallocCoherent);
            if (eh == hipSuccess) eh = hipHostGetDevicePointer(&d_back, h_back, 0);
            if (eh != hipSuccess) {
                set_error("mapped host buffer for the multi-tube SCF read-back: %s", hipGetErrorString(eh));
                return NEGF_EALLOC;
            }
        }
        NEGF_HIP_TRY(hipMemcpy(d_kt.p, kt.data(), sizeof(double) * kt.size(), hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_off.p, off.data(), sizeof(long long) * (T + 1), hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_tiles.p, tiles.data(), sizeof(MsTile) * tiles.size(), hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_rows.p, rows.data(), sizeof(MsRows) * rows.size(), hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_ptrU.p, hU.data(), sizeof(double*) * T, hipMemcpyHostToDevice));
        NEGF_HIP_TRY(hipMemcpy(d_ptrUc.p, hUc.data(), sizeof(double*) * T, hipMemcpyHostToDevice));
        if ((rc = fill_device(d_Uext.as<double>(), S, 0.0, st))) return rc;
        if ((rc = fill_device(d_state.as<double>(), S, p.NS_initial_deposit_value, st))) return rc;
        if ((rc = reset(1))) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        return NEGF_OK;
    }
    BigBroyden bro() const
    {
        BigBroyden b;
        double* s0 = d_state.as<double>();
        b.n_in = s0;
        b.n_prev = s0 + S;
        b.F = s0 + 2 * S;
        b.dF = s0 + 3 * S;
        b.V = d_hist.as<double>();
        b.W = b.V + (size_t)(M + 1) * S;
        b.part = d_part.as<double>();
        b.coef = d_coef.as<double>();
        b.step = d_step.as<int>();
        b.S = S;
        b.max_step = M;
        return b;
    }
    // Reset_Broyden_Parallel (Broyden_Parallel_General.cpp:250-355): history
    // and the previous iterates cleared, the step counter back to 1; n_in
    // kept or set to the initial value.
    int reset(int keep_n_in)
    {
        int rc;
        BigBroyden b = bro();
        if (!keep_n_in && (rc = fill_device(b.n_in, S, p.NS_initial_deposit_value, st))) return rc;
        if ((rc = fill_device(b.n_prev, 3 * S, 0.0, st))) return rc;
        if ((rc = fill_device(d_hist.as<double>(), 2 * (long long)(M + 1) * S, 0.0, st))) return rc;
        hipLaunchKernelGGL(big_broyden_first_step, dim3(1), dim3(1), 0, st, d_step.as<int>());
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }

    // U = U_ext - V - K n_in for every tube, into the tubes
    int enqueue_potential()
    {
        hipLaunchKernelGGL(multi_potential_tile_kernel, dim3(ntiles), dim3(kMsThreads), 0, st, d_tiles.as<MsTile>(),
                           d_off.as<long long>(), d_kt.as<double>(), D, bro().n_in, d_partial.as<double>());
        hipLaunchKernelGGL(multi_potential_finish, dim3(nrows), dim3(kMsThreads), 0, st, d_rows.as<MsRows>(),
                           d_off.as<long long>(), d_partial.as<double>(), d_Uext.as<double>(), d_V.as<double>(),
                           d_ptrU.as<double*>(), d_ptrUc.as<double*>(), d_U.as<double>());
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    int enqueue_iteration()
    {
        int rc = enqueue_potential();
        if (rc) return rc;
        rc = batch_enqueue(batch, d_Q.as<double>(), st, nullptr);
        if (rc) return rc;
        const int nb = (int)std::min<long long>(kMsBlocks, (S + kMsThreads - 1) / kMsThreads);
        hipLaunchKernelGGL(big_broyden_pass1, dim3(nb), dim3(kMsThreads), 0, st, bro(), p.Broyden_norm_type,
                           d_Q.as<double>());
        hipLaunchKernelGGL(big_broyden_pass2, dim3(1), dim3(kMsThreads), 0, st, bro(), nb);
        hipLaunchKernelGGL(big_broyden_pass3, dim3(nb), dim3(kMsThreads), 0, st, bro(), p.Broyden_fraction);
        NEGF_HIP_TRY(hipGetLastError());
        // Broyden_Norm, the l2 norm and the plans' status words to the host
        const int* st_eq = nullptr;
        const int* st_neq = nullptr;
        batch_status_words(batch, &st_eq, &st_neq);
        hipLaunchKernelGGL(multi_readback_kernel, dim3(1), dim3(64), 0, st, d_coef.as<const double>(), st_eq, st_neq,
                           static_cast<double*>(d_back));
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
    }
    // Capture one iteration unless a graph exists over the batch's present plans.
    int ensure_graph()
    {
        const long long g_now = batch_generation(batch);
        if (gexec && gen == g_now) return NEGF_OK;
        drop_graph();
        NEGF_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int erc = enqueue_iteration();
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
        const hipError_t e2 = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            gexec = nullptr;
            set_error("hipGraphInstantiate: %s", hipGetErrorString(e2));
            return NEGF_EDEVICE;
        }
        gen = g_now;
        return NEGF_OK;
    }
    bool adaptive_due(int it) const
    {
        for (const negf_tube* t : tubes)
            if (t->t.p.flag_adaptive_integration_limits && it % t->t.p.integrand_correction_interval == 0) return true;
        return false;
    }
    // One bias point (c_TransportSolver::Solve Transport.cpp:359-539): iterate
    // until Broyden_Norm <= Broyden_max_norm, or max_iters iterations when
    // max_iters > 0; then Reset_ForNextBiasStep (Transport.cpp:713-717).
    int bias_point(const double* V, int max_iters, double* U_out, double* Q_out, int* iters, double* final_norm,
                   double* ms_iter)
    {
        int rc;
        NEGF_HIP_TRY(hipMemcpyAsync(d_V.p, V, sizeof(double) * T, hipMemcpyHostToDevice, st));
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        NEGF_HIP_TRY(hipEventCreate(&e0));
        NEGF_HIP_TRY(hipEventCreate(&e1));
        double norm = 1.0, t_ms = 0.0;
        int it = 0;
        rc = NEGF_OK;
        while (true) {
            if (it > M) {
                set_error("Broyden_Step has exceeded the Broyden_threshold_maxstep");
                rc = NEGF_ENONCONV;
                break;
            }
            if (adaptive_due(it)) {
                // the adaptive correction samples this iteration's potential
                if ((rc = enqueue_potential())) break;
                NEGF_HIP_TRY(hipStreamSynchronize(st));
            }
            bool rebuilt = false;
            if ((rc = batch_prepare(batch, it, &rebuilt))) break;
            if (use_graph && (rc = ensure_graph())) break;  // (re)capture outside the timed span
            NEGF_HIP_TRY(hipEventRecord(e0, st));
            if (use_graph) {
                NEGF_HIP_TRY(hipGraphLaunch(gexec, st));
            } else if ((rc = enqueue_iteration())) {
                break;
            }
            NEGF_HIP_TRY(hipEventRecord(e1, st));
            NEGF_HIP_TRY(hipStreamSynchronize(st));
            float ms = 0.f;
            NEGF_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            t_ms += ms;
            const volatile double* back = static_cast<const volatile double*>(h_back);
            const volatile int* bs = reinterpret_cast<const volatile int*>(back + 2);
            if ((bs[0] | bs[1]) && (rc = batch_poll_status(batch))) break;
            norm = back[0];
            ++it;
            if (!std::isfinite(norm)) {
                set_error("multi-tube SCF: non-finite Broyden norm at iteration %d", it);
                rc = NEGF_ENAN;
                break;
            }
            if (norm <= p.Broyden_max_norm || (max_iters > 0 && it >= max_iters)) break;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (iters) *iters = it;
        if (final_norm) *final_norm = norm;
        if (ms_iter) *ms_iter = it > 0 ? t_ms / it : 0.0;
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        if (U_out) NEGF_HIP_TRY(hipMemcpy(U_out, d_U.p, sizeof(double) * S, hipMemcpyDeviceToHost));
        if (Q_out) NEGF_HIP_TRY(hipMemcpy(Q_out, d_Q.p, sizeof(double) * S, hipMemcpyDeviceToHost));
        if (rc) return rc;
        if ((rc = reset(p.reset_with_previous_charge_distribution))) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(st));
        return NEGF_OK;
    }

    negf_batch* batch = nullptr;
    std::vector<negf_tube*> tubes;
    std::vector<long long> off;
    int T = 0, M = 0, D = 0, ntiles = 0, nrows = 0;
    long long S = 0;
    negf_scf_params p{};
    DevBuf d_kt, d_off, d_tiles, d_rows, d_partial, d_ptrU, d_ptrUc, d_Uext, d_V, d_U, d_Q, d_state, d_hist, d_part, d_coef, d_step;
    void* h_back = nullptr;   // mapped read-back record (host address)
    void* d_back = nullptr;   // the same memory as the device sees it
    bool use_graph = true;
    long long gen = -1;
    hipStream_t st = nullptr;
    hipGraphExec_t gexec = nullptr;
};

}  // namespace negf

using namespace negf;

struct negf_multi_scf {
    MultiScf s;
};

extern "C" {

int negf_multi_scf_create(negf_batch* b, negf_tube* const* tubes, int ntubes, const negf_scf_params* p,
                          negf_multi_scf** out)
{
    clear_error();
    if (!b || !tubes || ntubes <= 0 || !p || !out) {
        set_error("invalid multi-tube SCF arguments");
        return NEGF_EARG;
    }
    for (int t = 0; t < ntubes; ++t)
        if (!tubes[t]) {
            set_error("tube %d is NULL", t);
            return NEGF_EARG;
        }
    negf_multi_scf* h = new negf_multi_scf;
    const int rc = h->s.init(b, tubes, ntubes, *p);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_multi_scf_destroy(negf_multi_scf* s)
{
    if (!s) return NEGF_EARG;
    delete s;
    return NEGF_OK;
}

int negf_multi_scf_num_sites(const negf_multi_scf* s, long long* S)
{
    clear_error();
    if (!s || !S) {
        set_error("null argument");
        return NEGF_EARG;
    }
    *S = s->s.S;
    return NEGF_OK;
}

int negf_multi_scf_set_external_potential(negf_multi_scf* h, const double* U_ext)
{
    clear_error();
    if (!h || !U_ext) {
        set_error("null argument");
        return NEGF_EARG;
    }
    MultiScf& s = h->s;
    NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    NEGF_HIP_TRY(hipMemcpy(s.d_Uext.p, U_ext, sizeof(double) * s.S, hipMemcpyHostToDevice));
    return NEGF_OK;
}

int negf_multi_scf_use_graph(negf_multi_scf* h, int on)
{
    clear_error();
    if (!h) {
        set_error("null argument");
        return NEGF_EARG;
    }
    h->s.use_graph = on != 0;
    if (!h->s.use_graph) h->s.drop_graph();
    return NEGF_OK;
}

int negf_multi_scf_bias_point(negf_multi_scf* h, const double* V_gate, int max_iters, double* U_out,
                              double* Q_out, int* iters, double* final_norm, double* ms_per_iter)
{
    clear_error();
    if (!h || !V_gate) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return h->s.bias_point(V_gate, max_iters, U_out, Q_out, iters, final_norm, ms_per_iter);
}

int negf_multi_scf_get_input_charge(const negf_multi_scf* h, double* n_in)
{
    clear_error();
    if (!h || !n_in) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const MultiScf& s = h->s;
    NEGF_HIP_TRY(hipStreamSynchronize(s.st));
    NEGF_HIP_TRY(hipMemcpy(n_in, s.d_state.p, sizeof(double) * s.S, hipMemcpyDeviceToHost));
    return NEGF_OK;
}

}  // extern "C"
