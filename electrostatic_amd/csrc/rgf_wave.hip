// rgf_wave.hip -- segmented, division-free recursive-Green's-function kernels
// for gfx950 (MI355X), energies across lanes.
//
// Problem (reference NEGF_Common.cpp:3667-3754, :3377-3497): for each energy
// E of a contour or real-axis path, the retarded Green's function of the
// tridiagonal matrix A = E - H - Sigma_L - Sigma_R (diagonal alpha_n = E - U_n,
// links -t_n with t_n = beta for n even and gamma for n odd) is needed on the
// diagonal (equilibrium charge) and in its first/last columns
// (non-equilibrium charge).  The reference runs a host-serial RGF sweep per
// energy with two complex divisions per site.
//
// Formulation (same matrix, same quantities):
//   u_{n+1} = a_n u_n - c_n u_{n-1},  a_n = alpha_n / t_n, c_n = t_{n-1}/t_n
//   u_0 = 1, u_{-1} = Sigma_L / t_{-1}           (left lead as boundary state)
//   v_{n-1} = b_n v_n - d_n v_{n+1},  b_n = alpha_n / t_{n-1}, d_n = 1/c_n
//   v_{N-1} = 1, v_N = Sigma_R / t_{N-1}        (right lead as boundary state)
//   D = t_{N-1} u_N - Sigma_R u_{N-1}
//   G_nn = u_n v_n / D,  G_{n,0} = v_n / D,  G_{n,N-1} = u_n / D.
// u is the left-connected (forward-stable) continuant, v the right-connected
// (backward-stable) one; their ratios are the RGF Y_n and X_n.  No per-site
// division remains.
//
// Mapping: a lane is one chain (one energy of one mode of one tube); the 64
// lanes of a wavefront are 64 chains of the same tube, so the site index and
// the potential U_n are wave-uniform (scalar loads, SGPR operands).  Each
// chain is cut into P segments of L sites; one wavefront handles one
// (64-chain group, segment) item.  Three kernels:
//   transfer_kernel  2x2 transfer matrix of each segment (both basis columns),
//                    renormalised by exact powers of two;
//   scan_kernel      Hillis-Steele prefix and suffix products over the P
//                    segments of a chain (lanes = segments here) give each
//                    segment its exact left state u and right state v, and D;
//   sweep_kernel     backward v over the segment with register checkpoints
//                    every K sites, then per K-site sub-block: v recomputed
//                    into registers, u forward, G and the charge contribution
//                    on the fly.  The 64 lanes' contributions to each site are
//                    summed through an LDS transpose and written as one
//                    partial row per chain group; a final kernel sums the
//                    groups of a tube in a fixed order.
#include "negf_device.h"

namespace negf {

struct WTube {
    int N;         // sites
    int L;         // segment length (multiple of K)
    int P;         // segments per chain (power of two <= 64)
    int nsub;      // sub-blocks of a full segment
    double t_odd;  // gamma (odd links; lead tau)
    double Uc0, Uc1;
    const double* U;  // device potential, N values
    int chain0, nchain;   // chains [chain0, chain0 + nchain)
    int item0;            // first (group, segment) item of this tube
    int group0;           // first global chain group of this tube
    long long part_off;   // this tube's partial rows (groups x N)
    int sample[3];        // non-equilibrium integrand sample sites (-1: none)
    int pad;
    const double* Ucp;    // device contact potentials (2 values), or null: Uc0/Uc1
};

struct WLaunch {
    ChainDesc* chains;
    const WTube* tubes;
    const int* item_tube;     // tube of each item
    int nitems;
    double* Tm;               // [item][9][64]: transfer matrix (8) + exponent
    double* St;               // [item][10][64]: u state (4) + e, v state (4) + e
    double* Dd;               // [chain][3]: D (re, im) + exponent
    double* part;             // partial charge rows [group][N]
    double2* diagG;           // optional [chain][N]
    double* chain_out;        // per-chain outputs
    double* dos_part;         // [item][64] per-segment DOS traces
    const int* sblk_tube;     // scan blocks: tube
    const int* sblk_chain;    // scan blocks: first local chain
    int* status;
    int kind;                 // ChainKind of the launch
};

__device__ __forceinline__ int exp_of(double m) { return m == 0.0 ? 0 : __builtin_amdgcn_frexp_exp(m); }
__device__ __forceinline__ double amax4(double2 a, double2 b)
{
    return fmax(fmax(fabs(a.x), fabs(a.y)), fmax(fabs(b.x), fabs(b.y)));
}
__device__ __forceinline__ void scale2(double2& a, int k)
{
    a.x = __builtin_amdgcn_ldexp(a.x, k);
    a.y = __builtin_amdgcn_ldexp(a.y, k);
}
__device__ __forceinline__ void renorm(double2& a, double2& b, int& e)
{
    int k = exp_of(amax4(a, b));
    scale2(a, -k);
    scale2(b, -k);
    e += k;
}
__device__ __forceinline__ void renorm_mat(double2& m00, double2& m01, double2& m10, double2& m11, int& e)
{
    int k = exp_of(fmax(amax4(m00, m01), amax4(m10, m11)));
    scale2(m00, -k);
    scale2(m01, -k);
    scale2(m10, -k);
    scale2(m11, -k);
    e += k;
}

// Per-lane constants at a site of parity p = n & 1.
struct Coefs {
    double erit[2], it[2], eiit[2];     // Re E / t_n, 1 / t_n, Im E / t_n
    double eritm[2], itm[2], eiitm[2];  // same with t_{n-1}
    double c[2], d[2];                  // t_{n-1}/t_n, t_n/t_{n-1}
};

__device__ __forceinline__ Coefs make_coefs(double2 E, double te, double to)
{
    Coefs k;
    double t[2] = {te, to};
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double tn = t[p], tm = t[1 - p];
        k.it[p] = 1.0 / tn;
        k.erit[p] = E.x * k.it[p];
        k.eiit[p] = E.y * k.it[p];
        k.itm[p] = 1.0 / tm;
        k.eritm[p] = E.x * k.itm[p];
        k.eiitm[p] = E.y * k.itm[p];
        k.c[p] = tm * k.it[p];
        k.d[p] = tn * k.itm[p];
    }
    return k;
}

// forward step at site n: (x, y) = (u_n, u_{n-1}) -> (u_{n+1}, u_n)
template <int P>
__device__ __forceinline__ void fwd_step(const Coefs& k, double U, double2& x, double2& y)
{
    double ar = fma(-U, k.it[P], k.erit[P]);
    double ai = k.eiit[P];
    double2 nx;
    nx.x = fma(ar, x.x, fma(-ai, x.y, -k.c[P] * y.x));
    nx.y = fma(ar, x.y, fma(ai, x.x, -k.c[P] * y.y));
    y = x;
    x = nx;
}

// backward step at site n: (x, y) = (v_n, v_{n+1}) -> (v_{n-1}, v_n)
template <int P>
__device__ __forceinline__ void bwd_step(const Coefs& k, double U, double2& x, double2& y)
{
    double br = fma(-U, k.itm[P], k.eritm[P]);
    double bi = k.eiitm[P];
    double2 nx;
    nx.x = fma(br, x.x, fma(-bi, x.y, -k.d[P] * y.x));
    nx.y = fma(br, x.y, fma(bi, x.x, -k.d[P] * y.y));
    y = x;
    x = nx;
}

// forward step on both basis columns of the transfer matrix
template <int P>
__device__ __forceinline__ void mat_step(const Coefs& k, double U, double2& m00, double2& m01, double2& m10,
                                         double2& m11)
{
    double ar = fma(-U, k.it[P], k.erit[P]);
    double ai = k.eiit[P];
    double cc = k.c[P];
    double2 n0, n1;
    n0.x = fma(ar, m00.x, fma(-ai, m00.y, -cc * m10.x));
    n0.y = fma(ar, m00.y, fma(ai, m00.x, -cc * m10.y));
    n1.x = fma(ar, m01.x, fma(-ai, m01.y, -cc * m11.x));
    n1.y = fma(ar, m01.y, fma(ai, m01.x, -cc * m11.y));
    m10 = m00;
    m11 = m01;
    m00 = n0;
    m01 = n1;
}

__device__ __forceinline__ void mat_mul(const double2 a00, const double2 a01, const double2 a10, const double2 a11,
                                        const double2 b00, const double2 b01, const double2 b10, const double2 b11,
                                        double2& r00, double2& r01, double2& r10, double2& r11)
{
    r00 = cadd(cmul(a00, b00), cmul(a01, b10));
    r01 = cadd(cmul(a00, b01), cmul(a01, b11));
    r10 = cadd(cmul(a10, b00), cmul(a11, b10));
    r11 = cadd(cmul(a10, b01), cmul(a11, b11));
}

// Decode item -> (tube, group-in-tube, segment) and this lane's chain.
struct ItemCtx {
    int t, g, s, chain, valid;
};
__device__ __forceinline__ ItemCtx item_ctx(const WLaunch& lp, int item, int lane, WTube& tube)
{
    ItemCtx c;
    c.t = lp.item_tube[item];
    tube = lp.tubes[c.t];
    int local = item - tube.item0;
    c.g = local / tube.P;
    c.s = local - c.g * tube.P;
    int lc = c.g * 64 + lane;
    c.valid = lc < tube.nchain;
    c.chain = tube.chain0 + (c.valid ? lc : tube.nchain - 1);
    return c;
}

// ------------------------------------------------------------ transfer
template <int K>
__global__ __launch_bounds__(256) void transfer_kernel(WLaunch lp)
{
    const int lane = threadIdx.x & 63;
    const int item = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= lp.nitems) return;
    WTube tube;
    ItemCtx ic = item_ctx(lp, item, lane, tube);
    const ChainDesc ch = lp.chains[ic.chain];
    const Coefs k = make_coefs(make_double2(ch.Er, ch.Ei), ch.t_even, tube.t_odd);
    const int lo = min(tube.N, ic.s * tube.L);
    const int hi = min(tube.N, lo + tube.L);
    const double* __restrict__ U = tube.U;
    double2 m00 = make_double2(1.0, 0.0), m01 = make_double2(0.0, 0.0);
    double2 m10 = make_double2(0.0, 0.0), m11 = make_double2(1.0, 0.0);
    int eT = 0;
    for (int p0 = lo; p0 < hi; p0 += K) {
        const int len = min(K, hi - p0);
        if (len == K) {
#pragma unroll
            for (int i = 0; i < K; i += 2) {
                mat_step<0>(k, U[p0 + i], m00, m01, m10, m11);
                mat_step<1>(k, U[p0 + i + 1], m00, m01, m10, m11);
            }
        } else {
            for (int i = 0; i < len; ++i) {
                if (i & 1)
                    mat_step<1>(k, U[p0 + i], m00, m01, m10, m11);
                else
                    mat_step<0>(k, U[p0 + i], m00, m01, m10, m11);
            }
        }
        renorm_mat(m00, m01, m10, m11, eT);
    }
    double* T = lp.Tm + (long long)item * 9 * 64 + lane;
    T[0 * 64] = m00.x;
    T[1 * 64] = m00.y;
    T[2 * 64] = m01.x;
    T[3 * 64] = m01.y;
    T[4 * 64] = m10.x;
    T[5 * 64] = m10.y;
    T[6 * 64] = m11.x;
    T[7 * 64] = m11.y;
    T[8 * 64] = (double)eT;
}

// ------------------------------------------------------------ scan
// One thread per (chain, segment); a block holds 256/P chains of one tube.
__global__ __launch_bounds__(256) void scan_kernel(WLaunch lp)
{
    const int t = lp.sblk_tube[blockIdx.x];
    const WTube tube = lp.tubes[t];
    const int P = tube.P;
    const int s = threadIdx.x & (P - 1);
    const int lc = lp.sblk_chain[blockIdx.x] + threadIdx.x / P;  // local chain
    if (lc >= ((tube.nchain + 63) & ~63)) return;  // whole chain beyond the last group
    const bool valid = lc < tube.nchain;
    const int lcc = valid ? lc : tube.nchain - 1;   // descriptor of padding lanes
    const int g = lc >> 6, lane = lc & 63;          // slot of this (possibly padding) lane
    const int item = tube.item0 + g * P + s;
    const int chain = tube.chain0 + lcc;
    const ChainDesc ch = lp.chains[chain];
    const double te = ch.t_even, to = tube.t_odd;
    const int N = tube.N;
    const double tau = ((N - 1) & 1) ? to : te;  // t_{N-1}
    const double2 SL = make_double2(ch.SLr, ch.SLi);
    const double2 SR = make_double2(ch.SRr, ch.SRi);
    const int lo = min(N, s * tube.L);

    const double* T = lp.Tm + (long long)item * 9 * 64 + lane;
    double2 t00 = make_double2(T[0], T[64]), t01 = make_double2(T[128], T[192]);
    double2 t10 = make_double2(T[256], T[320]), t11 = make_double2(T[384], T[448]);
    const int tE = (int)T[512];

    // prefix: Pre_s = T_s ... T_0
    double2 m00 = t00, m01 = t01, m10 = t10, m11 = t11;
    int eT = tE;
    for (int d = 1; d < P; d <<= 1) {
        double2 o00 = make_double2(__shfl_up(m00.x, d, P), __shfl_up(m00.y, d, P));
        double2 o01 = make_double2(__shfl_up(m01.x, d, P), __shfl_up(m01.y, d, P));
        double2 o10 = make_double2(__shfl_up(m10.x, d, P), __shfl_up(m10.y, d, P));
        double2 o11 = make_double2(__shfl_up(m11.x, d, P), __shfl_up(m11.y, d, P));
        int oe = __shfl_up(eT, d, P);
        if (s >= d) {
            double2 r00, r01, r10, r11;
            mat_mul(m00, m01, m10, m11, o00, o01, o10, o11, r00, r01, r10, r11);
            m00 = r00;
            m01 = r01;
            m10 = r10;
            m11 = r11;
            eT += oe;
        }
    }
    const double2 sigL = cscale(1.0 / to, SL);  // u_{-1}; t_{-1} is an odd link
    double2 wx = cadd(m00, cmul(m01, sigL));
    double2 wy = cadd(m10, cmul(m11, sigL));
    int we = eT;
    renorm(wx, wy, we);
    double2 ux = make_double2(__shfl_up(wx.x, 1, P), __shfl_up(wx.y, 1, P));
    double2 uy = make_double2(__shfl_up(wy.x, 1, P), __shfl_up(wy.y, 1, P));
    int ue = __shfl_up(we, 1, P);
    if (s == 0) {
        ux = make_double2(1.0, 0.0);
        uy = sigL;
        ue = 0;
    }
    const double2 uN = make_double2(__shfl(wx.x, P - 1, P), __shfl(wx.y, P - 1, P));
    const double2 uN1 = make_double2(__shfl(wy.x, P - 1, P), __shfl(wy.y, P - 1, P));
    const int eN = __shfl(we, P - 1, P);
    const double2 Dm = csub(cscale(tau, uN), cmul(SR, uN1));  // D = Dm 2^eN

    // suffix: Suf_s = T_{P-1} ... T_s
    m00 = t00;
    m01 = t01;
    m10 = t10;
    m11 = t11;
    eT = tE;
    for (int d = 1; d < P; d <<= 1) {
        double2 o00 = make_double2(__shfl_down(m00.x, d, P), __shfl_down(m00.y, d, P));
        double2 o01 = make_double2(__shfl_down(m01.x, d, P), __shfl_down(m01.y, d, P));
        double2 o10 = make_double2(__shfl_down(m10.x, d, P), __shfl_down(m10.y, d, P));
        double2 o11 = make_double2(__shfl_down(m11.x, d, P), __shfl_down(m11.y, d, P));
        int oe = __shfl_down(eT, d, P);
        if (s + d < P) {
            double2 r00, r01, r10, r11;
            mat_mul(o00, o01, o10, o11, m00, m01, m10, m11, r00, r01, r10, r11);
            m00 = r00;
            m01 = r01;
            m10 = r10;
            m11 = r11;
            eT += oe;
        }
    }
    const double2 sigR = cscale(1.0 / tau, SR);
    double2 zx, zy;
    int ze;
    {
        // Suf^{-1} = adj(M) 2^eT / det(T product), det = t_{lo-1}/t_{N-1}
        double id = (lo < N) ? (tau / to) : 1.0;
        zx = cscale(id, csub(cmul(m11, sigR), m01));
        zy = cscale(id, csub(m00, cmul(m10, sigR)));
        ze = eT;
        renorm(zx, zy, ze);
    }
    double2 ax = make_double2(__shfl_down(zx.x, 1, P), __shfl_down(zx.y, 1, P));
    double2 ay = make_double2(__shfl_down(zy.x, 1, P), __shfl_down(zy.y, 1, P));
    int ae = __shfl_down(ze, 1, P);
    const int hi = min(N, lo + tube.L);
    if (s == P - 1 || hi >= N) {
        ax = sigR;  // (v_N, v_{N-1})
        ay = make_double2(1.0, 0.0);
        ae = 0;
    }
    // states: u = (u_lo, u_{lo-1}), v = (v_{hi-1}, v_hi).  Padding lanes of the
    // last group carry the last chain's values so the sweep stays finite.
    double* S = lp.St + (long long)item * 10 * 64 + lane;
    S[0 * 64] = ux.x;
    S[1 * 64] = ux.y;
    S[2 * 64] = uy.x;
    S[3 * 64] = uy.y;
    S[4 * 64] = (double)ue;
    S[5 * 64] = ay.x;
    S[6 * 64] = ay.y;
    S[7 * 64] = ax.x;
    S[8 * 64] = ax.y;
    S[9 * 64] = (double)ae;
    if (s == 0 && valid) {
        double* Dp = lp.Dd + (long long)chain * 3;
        Dp[0] = Dm.x;
        Dp[1] = Dm.y;
        Dp[2] = (double)eN;
        // transmission Gamma_L Gamma_R |G_{0,N-1}|^2, G_{0,N-1} = 1/D
        if (lp.kind == KIND_DOS && lp.chain_out && ch.out_slot >= 0) {
            double iD2 = 1.0 / cabs2(Dm);
            double gN = __builtin_amdgcn_ldexp(1.0, -eN);
            double GamL = -2.0 * SL.y, GamR = -2.0 * SR.y;
            lp.chain_out[(long long)ch.out_slot * 2 + 1] = GamL * GamR * iD2 * gN * gN;
        }
    }
}

// ------------------------------------------------------------ sweep
template <int NS>
struct Ckpt {
    double2 x[NS], y[NS];
    int e[NS];
    __device__ __forceinline__ void put(int j, const double2& a, const double2& b, int ex)
    {
#pragma unroll
        for (int i = 0; i < NS; ++i)
            if (i == j) {
                x[i] = a;
                y[i] = b;
                e[i] = ex;
            }
    }
    __device__ __forceinline__ void get(int j, double2& a, double2& b, int& ex) const
    {
#pragma unroll
        for (int i = 0; i < NS; ++i)
            if (i == j) {
                a = x[i];
                b = y[i];
                ex = e[i];
            }
    }
};

// Sum the per-lane contributions red[i][lane] (i < K) over the 64 lanes of
// the wave; red is this wave's LDS tile [K][65] (rows padded against bank
// conflicts); every lane returns the sum for site (lane % K).  LDS operations of one wavefront execute in order,
// so no barrier is needed between the writes and the reads.
template <int K>
__device__ __forceinline__ double lane_sum_transpose(const double* red, int lane)
{
    __builtin_amdgcn_wave_barrier();
    constexpr int R = 64 / K;  // lanes per site
    constexpr int C = 64 / R;  // columns summed by each lane
    const int i = lane % K;
    const int r = lane / K;
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) acc += red[i * 65 + r * C + c];
#pragma unroll
    for (int d = K; d < 64; d <<= 1) acc += __shfl_xor(acc, d, 64);
    __builtin_amdgcn_wave_barrier();
    return acc;
}

// Pass-3 sites of one sub-block [p0, p0+len): contributions q[i], optional
// G output and integrand samples, forward u step.
struct SubCtx {
    double2 C, Gs;          // KIND_EQ: contribution Re(C u v), G = Gs u v
    double kL, kR, aL, aR;  // KIND_NONEQ / DOS weights of |v|^2 and |u|^2
};

template <int K, int KIND, bool WRITE_G, bool FULL, bool SAMPLE>
__device__ __forceinline__ void site_loop(const Coefs& k, const double (&Us)[K], const double2 (&vr)[K], double2& ux,
                                          double2& uy, const SubCtx& c, int p0, int len, bool valid,
                                          double2* Grow, double* red, int lane, double& dos_acc,
                                          const WTube& tube, const ChainDesc& ch, double* chain_out)
{
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (!FULL && KIND != KIND_DOS) red[i * 65 + lane] = 0.0;
        if (FULL || i < len) {
            if (KIND == KIND_EQ) {
                double2 g = cmul(ux, vr[i]);
                red[i * 65 + lane] = fma(c.C.x, g.x, -c.C.y * g.y);
                if (WRITE_G && valid) Grow[p0 + i] = cmul(c.Gs, g);
            } else if (KIND == KIND_NONEQ) {
                double av = cabs2(vr[i]);
                double au = cabs2(ux);
                red[i * 65 + lane] = fma(c.kL, av, c.kR * au);
                if (SAMPLE && valid) {
#pragma unroll
                    for (int w = 0; w < 3; ++w)
                        if (p0 + i == tube.sample[w])
                            chain_out[(long long)ch.out_slot * 3 + w] = ch.iL * c.aL * av + ch.iR * c.aR * au;
                }
            } else {
                dos_acc += c.aL * cabs2(vr[i]) + c.aR * cabs2(ux);
            }
            if (i & 1)
                fwd_step<1>(k, Us[i], ux, uy);
            else
                fwd_step<0>(k, Us[i], ux, uy);
        }
    }
}

template <int K, int NS, int KIND, bool WRITE_G>
__global__ __launch_bounds__(256, 2) void sweep_kernel(WLaunch lp)
{
    __shared__ double red_all[4][K * 65];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int item = blockIdx.x * 4 + wv;
    if (item >= lp.nitems) return;
    double* red = red_all[wv];
    WTube tube;
    ItemCtx ic = item_ctx(lp, item, lane, tube);
    const ChainDesc ch = lp.chains[ic.chain];
    const Coefs k = make_coefs(make_double2(ch.Er, ch.Ei), ch.t_even, tube.t_odd);
    const int N = tube.N;
    const int lo = min(N, ic.s * tube.L);
    const int hi = min(N, lo + tube.L);
    const int nsub = (hi - lo + K - 1) / K;  // wave-uniform
    const double* __restrict__ U = tube.U;
    const double* S = lp.St + (long long)item * 10 * 64 + lane;
    double2 ux = make_double2(S[0], S[64]), uy = make_double2(S[128], S[192]);
    int ue = (int)S[256];
    double2 vx0 = make_double2(S[320], S[384]), vy0 = make_double2(S[448], S[512]);
    int ve0 = (int)S[576];
    const double* Dp = lp.Dd + (long long)ic.chain * 3;
    const double2 iD = crecip(make_double2(Dp[0], Dp[1]));
    const int eN = (int)Dp[2];
    const double iD2 = cabs2(iD);
    const double vmask = ic.valid ? 1.0 : 0.0;
    const double2 coef = make_double2(ch.cr * vmask, ch.ci * vmask);
    const double GamL = -2.0 * ch.SLi, GamR = -2.0 * ch.SRi;  // Gamma = i(S - S^+)
    double* prow = lp.part + tube.part_off + (long long)ic.g * N;

    // ---------------- pass 2: v checkpoints ----------------
    Ckpt<NS> ck;
    {
        double2 x = vx0, y = vy0;
        int e = ve0;
        ck.put(nsub - 1, x, y, e);
        for (int j = nsub - 1; j >= 1; --j) {
            const int p0 = lo + j * K;
            const int len = min(K, hi - p0);
            if (len == K) {
#pragma unroll
                for (int i = K - 1; i >= 0; i -= 2) {
                    bwd_step<1>(k, U[p0 + i], x, y);
                    bwd_step<0>(k, U[p0 + i - 1], x, y);
                }
            } else {
                for (int i = len - 1; i >= 0; --i) {
                    if (i & 1)
                        bwd_step<1>(k, U[p0 + i], x, y);
                    else
                        bwd_step<0>(k, U[p0 + i], x, y);
                }
            }
            renorm(x, y, e);
            ck.put(j - 1, x, y, e);
        }
    }

    // ---------------- pass 3 ----------------
    double dos_acc = 0.0;
    double2* Grow = WRITE_G ? (lp.diagG + (long long)ic.chain * N) : nullptr;
    for (int j = 0; j < nsub; ++j) {
        const int p0 = lo + j * K;
        const int len = min(K, hi - p0);
        double2 cx, cy;
        int ce;
        ck.get(j, cx, cy, ce);
        double2 vr[K];
        double Us[K];
#pragma unroll
        for (int i = 0; i < K; ++i) Us[i] = (i < len) ? U[p0 + i] : 0.0;
        {
            double2 x = cx, y = cy;  // (v_{q-1}, v_q)
            if (len == K) {
                vr[K - 1] = x;
#pragma unroll
                for (int i = K - 1; i >= 1; --i) {
                    if (i & 1)
                        bwd_step<1>(k, Us[i], x, y);
                    else
                        bwd_step<0>(k, Us[i], x, y);
                    vr[i - 1] = x;
                }
            } else {
#pragma unroll
                for (int i = K - 1; i >= 0; --i) {
                    if (i == len - 1) vr[i] = x;
                    if (i < len - 1) {
                        if ((i + 1) & 1)
                            bwd_step<1>(k, Us[i + 1], x, y);
                        else
                            bwd_step<0>(k, Us[i + 1], x, y);
                        vr[i] = x;
                    }
                }
            }
        }
        renorm(ux, uy, ue);
        const double sc_uv = __builtin_amdgcn_ldexp(1.0, ue + ce - eN);
        const double2 C = cscale(sc_uv, cmul(coef, iD));
        const double sv = __builtin_amdgcn_ldexp(1.0, ce - eN);
        const double su = __builtin_amdgcn_ldexp(1.0, ue - eN);
        const double aL = GamL * sv * sv * iD2;
        const double aR = GamR * su * su * iD2;
        const double kL = ch.cL * vmask * aL;
        const double kR = ch.cR * vmask * aR;
        SubCtx sc;
        sc.C = C;
        sc.Gs = cscale(sc_uv, iD);
        sc.kL = kL;
        sc.kR = kR;
        sc.aL = aL;
        sc.aR = aR;
        bool samp = false;
        if (KIND == KIND_NONEQ && lp.chain_out) {
#pragma unroll
            for (int w = 0; w < 3; ++w) samp |= (tube.sample[w] >= p0 && tube.sample[w] < p0 + len);
        }
        if (samp)
            site_loop<K, KIND, WRITE_G, false, true>(k, Us, vr, ux, uy, sc, p0, len, ic.valid, Grow, red, lane,
                                                     dos_acc, tube, ch, lp.chain_out);
        else if (len == K)
            site_loop<K, KIND, WRITE_G, true, false>(k, Us, vr, ux, uy, sc, p0, len, ic.valid, Grow, red, lane,
                                                     dos_acc, tube, ch, lp.chain_out);
        else
            site_loop<K, KIND, WRITE_G, false, false>(k, Us, vr, ux, uy, sc, p0, len, ic.valid, Grow, red, lane,
                                                      dos_acc, tube, ch, lp.chain_out);
        if (KIND != KIND_DOS) {
            double r = lane_sum_transpose<K>(red, lane);
            if (lane < len) prow[p0 + lane] = r;
        }
    }
    if (KIND == KIND_DOS) lp.dos_part[(long long)item * 64 + lane] = dos_acc;
}

// Lead self-energies of every chain, Sigma = tau g tau^+ with tau = t_odd
// (get_Sigma_at_contacts NEGF_Common.cpp:4165-4177, CNT.cpp:348-351); one
// thread per chain.
__global__ __launch_bounds__(256) void sigma_kernel(ChainDesc* chains, const int* chain_tube, int nchains,
                                                    const WTube* tubes, int use_decimation, int dec_layers,
                                                    int dec_max_iter, int dec_rel_error, int* status)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchains) return;
    const WTube tube = tubes[chain_tube[c]];
    ChainDesc* ch = chains + c;
    const double2 E = make_double2(ch->Er, ch->Ei);
    const double te = ch->t_even, to = tube.t_odd;
    const double Uc0 = tube.Ucp ? tube.Ucp[0] : tube.Uc0;
    const double Uc1 = tube.Ucp ? tube.Ucp[1] : tube.Uc1;
    double2 gL, gR;
    int okL = 1, okR = 1;
    if (use_decimation) {
        gL = decimation_gf(make_double2(E.x - Uc0, E.y), te, to, dec_layers, dec_max_iter, dec_rel_error);
        gR = decimation_gf(make_double2(E.x - Uc1, E.y), te, to, dec_layers, dec_max_iter, dec_rel_error);
    } else {
        gL = surface_gf(E, Uc0, te, to, &okL);
        gR = surface_gf(E, Uc1, te, to, &okR);
    }
    if (!okL || !okR) atomicOr(status, 1);
    ch->SLr = to * to * gL.x;
    ch->SLi = to * to * gL.y;
    ch->SRr = to * to * gR.x;
    ch->SRi = to * to * gR.y;
}

// Sum partial rows in a fixed order.  Row r of group t starts at
// src + src_off[t] + r*N_t; output row c (rows [c*chunk, (c+1)*chunk)) goes to
// dst + dst_off[t] + c*N_t.
__global__ __launch_bounds__(256) void reduce_rows_kernel(const double* __restrict__ src,
                                                          const long long* __restrict__ src_off,
                                                          const int* __restrict__ nrows,
                                                          const int* __restrict__ Ns, int chunk,
                                                          double* __restrict__ dst,
                                                          const long long* __restrict__ dst_off)
{
    const int t = blockIdx.z;
    const int N = Ns[t];
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int r_begin = c * chunk;
    const int r_end = min(nrows[t], r_begin + chunk);
    if (n >= N || r_begin >= r_end) return;
    const double* p = src + src_off[t] + n;
    double acc = 0.0;
    for (int r = r_begin; r < r_end; ++r) acc += p[(long long)r * N];
    dst[dst_off[t] + (long long)c * N + n] = acc;
}

// Per-chain DOS trace: sum the per-segment partials in a fixed order.
__global__ void dos_finish_kernel(WLaunch lp, int ntubes)
{
    const int chain_g = blockIdx.x * blockDim.x + threadIdx.x;
    for (int t = 0; t < ntubes; ++t) {
        const WTube tube = lp.tubes[t];
        const int lc = chain_g - tube.chain0;
        if (lc < 0 || lc >= tube.nchain) continue;
        const int g = lc >> 6, lane = lc & 63;
        double acc = 0.0;
        for (int s = 0; s < tube.P; ++s) acc += lp.dos_part[(long long)(tube.item0 + g * tube.P + s) * 64 + lane];
        const ChainDesc ch = lp.chains[chain_g];
        if (ch.out_slot >= 0) lp.chain_out[(long long)ch.out_slot * 2] = acc;
    }
}

}  // namespace negf
