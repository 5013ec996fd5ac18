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
// division remains.  Where |Im E| allows, the full-segment kernels run the
// same recurrences divided by Im E site by site (four FMAs per complex step
// instead of six; see "scaled recurrence" below).
//
// Mapping: a lane is one chain (one energy of one mode of one tube); the 64
// lanes of a wavefront are 64 chains of the same tube, so the site index and
// the potential U_n are wave-uniform (scalar loads, SGPR operands).  Each
// chain is cut into P segments of L sites; one wavefront handles one
// (64-chain group, segment) item.  Three kernels:
//   transfer_kernel  2x2 transfer matrix of each segment (both basis columns),
//                    renormalised by exact powers of two;
//   scan_kernel      walks the P segment matrices of a chain (u forward, v
//                    backward, one matrix-vector product per segment) and
//                    gives each segment its left state u and right state v,
//                    and D;
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
    int ext;              // index of the tube in its caller (the active mask)
    const double* Ucp;    // device contact potentials (2 values), or null: Uc0/Uc1
    int fused;            // 1: transfer matrices and walks in one launch (transfer_scan_kernel);
                          // 2: and the equilibrium sweep too (rgf_eq_fused_kernel, two rows per group)
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
    double* site_out;         // KIND_DOS, optional [chain][N]: the site's spectral weight
    double* chain_out;        // per-chain outputs
    double* dos_part;         // [item][64] per-segment DOS traces
    const int* sblk_tube;     // chain groups: tube
    const int* sblk_chain;    // chain groups: group index within the tube
    int* status;
    int kind;                 // ChainKind of the launch
    const int* item_list;     // sweep launches: items of this launch (or null: all)
    int nlist;
    const int* tube_active;   // optional [ext]: 0 skips the tube's work (its outputs stay as they were)
    int dec_use, dec_layers, dec_max_iter, dec_rel_error;  // lead surface GF by decimation (else closed form)
    int single;               // one tube: its descriptor is tube1 (a kernel argument, no table load)
    WTube tube1;
};

// Tube of a chain group or item: from the kernel argument when the launch has
// one tube (saves a dependent load at the start of every wavefront).
__device__ __forceinline__ int group_tube(const WLaunch& lp, int gi, int& g)
{
    if (lp.single) {
        g = gi;
        return 0;
    }
    g = lp.sblk_chain[gi];
    return lp.sblk_tube[gi];
}
__device__ __forceinline__ WTube tube_of(const WLaunch& lp, int t) { return lp.single ? lp.tube1 : lp.tubes[t]; }

__device__ __forceinline__ bool tube_skipped(const WLaunch& lp, const WTube& tube)
{
    return lp.tube_active && !lp.tube_active[__builtin_amdgcn_readfirstlane(tube.ext)];
}

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

// ---- scaled recurrence
// Dividing the three-term relation t_{n-1} y_{n-1} - alpha_n y_n + t_n y_{n+1}
// = 0 site by site by Im E: with y_n = s_n w_n and s_{n+1} = s_n ai_n,
// ai_n = Im E / t_n (forward) and the mirrored scale for the backward
// direction, both directions become
//   w_{n+1}  = (a'_n + i) w_n  - tau_{p(n-1)} w_{n-1}
//   w'_{n-1} = (a'_n + i) w'_n - tau_{p(n)}   w'_{n+1}
// with a'_n = (Re E - U_n) / Im E and tau_p = (t_p / Im E)^2: the imaginary
// part of the diagonal coefficient is exactly 1, so a complex step is four
// FMAs (the unscaled step needs six).  In a segment [lo, hi) with the local
// scales s_lo = 1 (forward) and s'_{hi-1} = 1 (backward), the product
// s_n s'_n is the same at every site: C = ai_lo ... ai_{hi-2}, so
// u_n v_n = C w_n w'_n.  The scales shrink w by |Im E| / t per site against
// u, so the scaled form is used only when t / |Im E| <= 2^24 (at most 2^384
// per 16-site renormalisation block).
struct SCoefs {
    double ie, er;  // 1 / Im E, Re E / Im E
    double tau[2];  // (t_p / Im E)^2
    double ai[2];   // Im E / t_p (index with static parities or selects: a
                    // run-time index makes the compiler move the whole
                    // struct to LDS, which slowed the transfer kernel by 40%)
};

__device__ __forceinline__ SCoefs make_scoefs(double2 E, double te, double to)
{
    SCoefs k;
    k.ie = 1.0 / E.y;
    k.er = E.x * k.ie;
    const double a = te * k.ie, b = to * k.ie;
    k.tau[0] = a * a;
    k.tau[1] = b * b;
    k.ai[0] = E.y / te;
    k.ai[1] = E.y / to;
    return k;
}

__device__ __forceinline__ bool scaled_ok(double2 E, double te, double to)
{
    return fabs(E.y) * 16777216.0 >= fmax(fabs(te), fabs(to));
}

// forward scaled step at site n (parity P): (x, y) = (w_n, w_{n-1}) -> (w_{n+1}, w_n)
template <int P>
__device__ __forceinline__ void fwd_s(const SCoefs& k, double U, double2& x, double2& y)
{
    const double ar = fma(-U, k.ie, k.er);
    const double tc = k.tau[1 - P];
    double2 nx;
    nx.x = fma(ar, x.x, fma(-tc, y.x, -x.y));
    nx.y = fma(ar, x.y, fma(-tc, y.y, x.x));
    y = x;
    x = nx;
}

// backward scaled step at site n (parity P): (x, y) = (w'_n, w'_{n+1}) -> (w'_{n-1}, w'_n)
template <int P>
__device__ __forceinline__ void bwd_s(const SCoefs& k, double U, double2& x, double2& y)
{
    const double ar = fma(-U, k.ie, k.er);
    const double tc = k.tau[P];
    double2 nx;
    nx.x = fma(ar, x.x, fma(-tc, y.x, -x.y));
    nx.y = fma(ar, x.y, fma(-tc, y.y, x.x));
    y = x;
    x = nx;
}

// scaled forward step on both basis columns
template <int P>
__device__ __forceinline__ void mat_s(const SCoefs& k, double U, double2& m00, double2& m01, double2& m10,
                                      double2& m11)
{
    const double ar = fma(-U, k.ie, k.er);
    const double tc = k.tau[1 - P];
    double2 n0, n1;
    n0.x = fma(ar, m00.x, fma(-tc, m10.x, -m00.y));
    n0.y = fma(ar, m00.y, fma(-tc, m10.y, m00.x));
    n1.x = fma(ar, m01.x, fma(-tc, m11.x, -m01.y));
    n1.y = fma(ar, m01.y, fma(-tc, m11.y, m01.x));
    m10 = m00;
    m11 = m01;
    m00 = n0;
    m01 = n1;
}

// Double-double product (h, l) *= (bh, bl) (Dekker; the FMA gives the exact
// product error).  Contraction is off so the error terms stay exact.
__device__ __forceinline__ void dd_mul(double& h, double& l, double bh, double bl)
{
#pragma clang fp contract(off)
    const double p = h * bh;
    double e = fma(h, bh, -p);
    e = fma(h, bl, e);
    e = fma(l, bh, e);
    const double s = p + e;
    l = e - (s - p);
    h = s;
}

// Product of the scale factors ai over cnt >= 0 consecutive sites whose
// first site has parity p0, as m 2^e with |m| in [0.5, 1): the pair product
// ai_0 ai_1 is kept exactly as a double-double and raised to cnt/2 by binary
// powering, so the result is within about one ulp of the exact product of
// the rounded factors (a plain power would amplify the rounding of the pair
// product cnt/2 times).  cnt is wave-uniform.
__device__ __forceinline__ void scale_prod(const SCoefs& k, int p0, int cnt, double& m, int& e)
{
    const int pairs = cnt >> 1;
    double ph = k.ai[0] * k.ai[1];
    double pl = fma(k.ai[0], k.ai[1], -ph);
    const int k0 = exp_of(ph);
    ph = __builtin_amdgcn_ldexp(ph, -k0);
    pl = __builtin_amdgcn_ldexp(pl, -k0);
    double rh = 1.0, rl = 0.0;
    for (int n = pairs; n; n >>= 1) {
        if (n & 1) dd_mul(rh, rl, ph, pl);
        if (n > 1) dd_mul(ph, pl, ph, pl);
    }
    if (cnt & 1) dd_mul(rh, rl, p0 ? k.ai[1] : k.ai[0], 0.0);  // the last site has parity p0
    const double r = rh + rl;
    const int k1 = exp_of(r);
    m = __builtin_amdgcn_ldexp(r, -k1);
    e = k0 * pairs + k1;
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
    c.t = lp.single ? 0 : __builtin_amdgcn_readfirstlane(lp.item_tube[item]);
    tube = tube_of(lp, c.t);
    // the tube is the same for the whole wavefront: keep its integers scalar
    tube.N = __builtin_amdgcn_readfirstlane(tube.N);
    tube.L = __builtin_amdgcn_readfirstlane(tube.L);
    tube.P = __builtin_amdgcn_readfirstlane(tube.P);
    tube.item0 = __builtin_amdgcn_readfirstlane(tube.item0);
    tube.chain0 = __builtin_amdgcn_readfirstlane(tube.chain0);
    tube.nchain = __builtin_amdgcn_readfirstlane(tube.nchain);
    int local = item - tube.item0;
    c.g = local / tube.P;
    c.s = local - c.g * tube.P;
    int lc = c.g * 64 + lane;
    c.valid = lc < tube.nchain;
    c.chain = tube.chain0 + (c.valid ? lc : tube.nchain - 1);
    return c;
}

// ------------------------------------------------------------ transfer
// The wave's segment of U is staged once in LDS with coalesced loads; the
// per-site reads are then wave-uniform LDS broadcasts instead of a vector
// load (and its wait) per site.
// All loads are issued before the first LDS write (clamped addresses, so they
// are unconditional): a loop that writes each value before loading the next
// waits a full memory round trip per iteration.
__device__ __forceinline__ void stage_segment(double* us, const double* __restrict__ U, int lo, int hi, int lane)
{
    constexpr int R = NEGF_SEG_MAX / 64;
    const int n = hi - lo;  // wave-uniform
    if (n > 0) {
        double v[R];
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = U[lo + min(lane + 64 * q, n - 1)];
#pragma unroll
        for (int q = 0; q < R; ++q)
            if (lane + 64 * q < n) us[lane + 64 * q] = v[q];
    }
    __builtin_amdgcn_wave_barrier();
}

// Lead self-energies of one chain, Sigma = tau g tau^+ with tau = t_odd
// (get_Sigma_at_contacts NEGF_Common.cpp:4165-4177, CNT.cpp:348-351), written
// into its ChainDesc for the scan and the sweeps.  DEC: the surface GF by
// decimation (CNT.cpp:358-361) instead of the closed form.
template <bool DEC>
__device__ __forceinline__ void lead_sigma(const WLaunch& lp, ChainDesc* out, const ChainDesc& ch, const WTube& tube,
                                           double2* SL = nullptr, double2* SR = nullptr)
{
    const double2 E = make_double2(ch.Er, ch.Ei);
    const double te = ch.t_even, to = tube.t_odd;
    const double Uc0 = tube.Ucp ? tube.Ucp[0] : tube.Uc0;
    const double Uc1 = tube.Ucp ? tube.Ucp[1] : tube.Uc1;
    double2 gL, gR;
    int okL = 1, okR = 1;
    if (DEC) {
        gL = decimation_gf(make_double2(E.x - Uc0, E.y), te, to, lp.dec_layers, lp.dec_max_iter, lp.dec_rel_error);
        gR = decimation_gf(make_double2(E.x - Uc1, E.y), te, to, lp.dec_layers, lp.dec_max_iter, lp.dec_rel_error);
    } else {
        gL = surface_gf(E, Uc0, te, to, &okL);
        gR = surface_gf(E, Uc1, te, to, &okR);
    }
    const double t2 = to * to;
    if (SL) {
        *SL = make_double2(t2 * gL.x, t2 * gL.y);
        *SR = make_double2(t2 * gR.x, t2 * gR.y);
    }
    if (!out) return;
    if (!okL || !okR) atomicOr(lp.status, 1);
    out->SLr = t2 * gL.x;
    out->SLi = t2 * gL.y;
    out->SRr = t2 * gR.x;
    out->SRi = t2 * gR.y;
}

// Lead self-energies by decimation, one thread per chain, before the transfer
// launch (the decimation's layer arrays live in scratch, which the transfer
// kernel does not carry).
__global__ __launch_bounds__(256) void sigma_dec_kernel(WLaunch lp, const int* chain_tube, int nchains)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchains) return;
    const WTube tube = lp.tubes[chain_tube[c]];
    lead_sigma<true>(lp, lp.chains + c, lp.chains[c], tube);
}

// Step overloads: plain (Coefs) or scaled (SCoefs) recurrence.
template <int P>
__device__ __forceinline__ void mstep(const Coefs& k, double U, double2& a, double2& b, double2& c, double2& d)
{
    mat_step<P>(k, U, a, b, c, d);
}
template <int P>
__device__ __forceinline__ void mstep(const SCoefs& k, double U, double2& a, double2& b, double2& c, double2& d)
{
    mat_s<P>(k, U, a, b, c, d);
}
template <int P>
__device__ __forceinline__ void fstep(const Coefs& k, double U, double2& x, double2& y)
{
    fwd_step<P>(k, U, x, y);
}
template <int P>
__device__ __forceinline__ void fstep(const SCoefs& k, double U, double2& x, double2& y)
{
    fwd_s<P>(k, U, x, y);
}
template <int P>
__device__ __forceinline__ void bstep(const Coefs& k, double U, double2& x, double2& y)
{
    bwd_step<P>(k, U, x, y);
}
template <int P>
__device__ __forceinline__ void bstep(const SCoefs& k, double U, double2& x, double2& y)
{
    bwd_s<P>(k, U, x, y);
}

// Both basis columns through n sites of the staged segment (the first at an
// even site), renormalised every K sites.
template <int K, class KC, class UP>
__device__ __forceinline__ void transfer_loop(const KC& k, UP us, int n, double2& m00, double2& m01,
                                              double2& m10, double2& m11, int& eT)
{
    for (int q0 = 0; q0 < n; q0 += K) {
        const int len = min(K, n - q0);
        if (len == K) {
#pragma unroll
            for (int i = 0; i < K; i += 2) {
                mstep<0>(k, us[q0 + i], m00, m01, m10, m11);
                mstep<1>(k, us[q0 + i + 1], m00, m01, m10, m11);
            }
        } else {
            for (int i = 0; i < len; ++i) {
                if (i & 1)
                    mstep<1>(k, us[q0 + i], m00, m01, m10, m11);
                else
                    mstep<0>(k, us[q0 + i], m00, m01, m10, m11);
            }
        }
        renorm_mat(m00, m01, m10, m11, eT);
    }
}

// Two independent full segments of L = K NS sites stepped together (twice the
// independent FMA chains of one lane), each renormalised every K sites.
struct Mat2x {
    double2 m00, m01, m10, m11;
    int e;
};
template <int K, int NS, class KC>
__device__ __forceinline__ void transfer_loop2(const KC& k, const double* ua, const double* ub, Mat2x& a, Mat2x& b)
{
#pragma unroll 1
    for (int q0 = 0; q0 < K * NS; q0 += K) {
#pragma unroll
        for (int i = 0; i < K; i += 2) {
            mstep<0>(k, ua[q0 + i], a.m00, a.m01, a.m10, a.m11);
            mstep<0>(k, ub[q0 + i], b.m00, b.m01, b.m10, b.m11);
            mstep<1>(k, ua[q0 + i + 1], a.m00, a.m01, a.m10, a.m11);
            mstep<1>(k, ub[q0 + i + 1], b.m00, b.m01, b.m10, b.m11);
        }
        renorm_mat(a.m00, a.m01, a.m10, a.m11, a.e);
        renorm_mat(b.m00, b.m01, b.m10, b.m11, b.e);
    }
}

// (256, 4): up to 4 workgroups of 4 waves per CU, enough registers for the
// lead self-energies without spills (4, 6 or 8 waves per SIMD time the same).
template <int K>
__global__ __launch_bounds__(256, 4) void transfer_kernel(WLaunch lp)
{
    __shared__ double us_all[4][NEGF_SEG_MAX];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    if (item >= lp.nitems) return;
    WTube tube;
    ItemCtx ic = item_ctx(lp, item, lane, tube);
    if (tube.fused || tube_skipped(lp, tube)) return;  // whole wavefront
    const ChainDesc ch = lp.chains[ic.chain];
    // the first segment's wavefront also computes its chains' lead self-energies
    // in closed form (only the scan and the sweeps, later launches, read them)
    if (!lp.dec_use && ic.s == 0 && ic.valid) lead_sigma<false>(lp, lp.chains + ic.chain, ch, tube);
    const double2 E = make_double2(ch.Er, ch.Ei);
    const int lo = min(tube.N, ic.s * tube.L);
    const int hi = min(tube.N, lo + tube.L);
    double* us = us_all[wv];
    stage_segment(us, tube.U, lo, hi, lane);
    double2 m00 = make_double2(1.0, 0.0), m01 = make_double2(0.0, 0.0);
    double2 m10 = make_double2(0.0, 0.0), m11 = make_double2(1.0, 0.0);
    int eT = 0;
    if (hi > lo && __all(scaled_ok(E, ch.t_even, tube.t_odd))) {
        // scaled basis: column 1 starts at w_{lo-1} = u_{lo-1} ai_{lo-1};
        // the rows return to u with s_L = C ai_{hi-1} and s_{L-1} = C
        const SCoefs ks = make_scoefs(E, ch.t_even, tube.t_odd);
        const double a0 = ((lo - 1) & 1) ? ks.ai[1] : ks.ai[0];
        m11 = make_double2(a0, 0.0);
        // U is wave-uniform and read-only here: through the constant address
        // space it arrives by scalar loads into SGPRs instead of LDS reads
        typedef const __attribute__((address_space(4))) double cdbl;
        const cdbl* ug = (const cdbl*)(uintptr_t)(tube.U + lo);
        transfer_loop<K>(ks, ug, hi - lo, m00, m01, m10, m11, eT);
        double mC;
        int eC;
        scale_prod(ks, lo & 1, hi - lo - 1, mC, eC);
        const double r0 = mC * (((hi - 1) & 1) ? ks.ai[1] : ks.ai[0]);
        m00 = cscale(r0, m00);
        m01 = cscale(r0, m01);
        m10 = cscale(mC, m10);
        m11 = cscale(mC, m11);
        eT += eC;
        renorm_mat(m00, m01, m10, m11, eT);
    } else {
        const Coefs k = make_coefs(E, ch.t_even, tube.t_odd);
        transfer_loop<K>(k, us, hi - lo, m00, m01, m10, m11, eT);
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
// Segment states.  With T_s the transfer matrix of segment s (sites
// [lo, hi)) of the recurrence u_{n+1} = a_n u_n - c_n u_{n-1} on column
// vectors (u_n, u_{n-1}):
//   u: U_0 = (u_0, u_{-1}) = (1, Sigma_L/t_{-1});  U_hi = T_s U_lo
//   v: the same recurrence read backwards, V_lo = T_s^{-1} V_hi with
//      T_s^{-1} = adj(T_s) / det(T_s), det(T_s) = t_{lo-1}/t_{hi-1}
//      (product of the per-site determinants c_n);
//      V_N = (v_N, v_{N-1}) = (Sigma_R/t_{N-1}, 1)
//   D = t_{N-1} u_N - Sigma_R u_{N-1}.
// One workgroup of SCAN_WAVES wavefronts per (64-chain group, walk), lane =
// chain, so every Tm read and St write is a coalesced row.  The u walk applies
// T_0, T_1, ... to U_0, the v walk T_{P-1}^{-1}, T_{P-2}^{-1}, ... to V_N; each
// writes the entry state of every segment it passes.  A walk is cut into
// SCAN_WAVES contiguous ranges of segments, one per wave, taken in turn (the
// state is handed on through LDS).  Every wave requests the matrices of its
// range one turn ahead of its own (the first wave when the kernel starts), so
// the loads of the next range arrive while the current one is walked and the
// first range's loads do not queue behind everybody else's.  The two walks of
// a chain group are separate workgroups, so neither waits for the other.
// The walks use matrix-vector products only: near a band edge a segment
// matrix is close to a defective matrix whose entries are of the order of the
// segment length while the state is close to its eigenvector, and a product
// of two such matrices rounds at |T_a| |T_b| eps in directions that the later
// walk does not damp (about 1e-9 relative in G_nn 1e-5 eV from the outer band
// edges of a 4096-site chain), where a matrix-vector product from the
// normalised state rounds at |T| eps.
// The state is renormalised every SCAN_RENORM steps only (off the dependency
// chain of the other steps): the matrices' mantissas have real and imaginary
// parts below 1 in magnitude, so a state grows by less than a factor 4 per
// step, and the stored entry states are exact (mantissa, exponent) pairs
// whether or not they were just renormalised.
#define SCAN_WAVES 4
#define SCAN_CACHE 8
#define SCAN_RENORM 4

struct Mat2 {
    double2 a, b, c, d;  // [[a, b], [c, d]]
    double e;            // exponent: the matrix is (a, b, c, d) 2^e (kept as
                         // loaded: converting it at load time would make the
                         // loading wave wait for each load in turn)
};

__device__ __forceinline__ Mat2 load_T(const WLaunch& lp, long long item, int lane)
{
    const double* T = lp.Tm + item * 9 * 64 + lane;
    Mat2 m;
    m.a = make_double2(T[0], T[64]);
    m.b = make_double2(T[128], T[192]);
    m.c = make_double2(T[256], T[320]);
    m.d = make_double2(T[384], T[448]);
    m.e = T[512];
    return m;
}

// (x, y) <- M (x, y)
__device__ __forceinline__ void mat2_apply(const Mat2& m, double2& x, double2& y, int& ex)
{
    const double2 nx = cadd(cmul(m.a, x), cmul(m.b, y));
    const double2 ny = cadd(cmul(m.c, x), cmul(m.d, y));
    x = nx;
    y = ny;
    ex += (int)m.e;
    renorm(x, y, ex);
}

// (x, y) <- M (x, y) without renormalisation
__device__ __forceinline__ void mat2_apply_raw(const Mat2& m, double2& x, double2& y, int& ex)
{
    const double2 nx = cadd(cmul(m.a, x), cmul(m.b, y));
    const double2 ny = cadd(cmul(m.c, x), cmul(m.d, y));
    x = nx;
    y = ny;
    ex += (int)m.e;
}

// (x, y) <- M^{-1} (x, y) without renormalisation
__device__ __forceinline__ void mat2_apply_inv_raw(const Mat2& m, double idet, double2& x, double2& y, int& ex)
{
    const double2 nx = cscale(idet, csub(cmul(m.d, x), cmul(m.b, y)));
    const double2 ny = cscale(idet, csub(cmul(m.a, y), cmul(m.c, x)));
    x = nx;
    y = ny;
    ex += (int)m.e;
}

// (x, y) <- M^{-1} (x, y) = adj(M) (x, y) / det, idet = 1/det of the true matrix
__device__ __forceinline__ void mat2_apply_inv(const Mat2& m, double idet, double2& x, double2& y, int& ex)
{
    const double2 nx = cscale(idet, csub(cmul(m.d, x), cmul(m.b, y)));
    const double2 ny = cscale(idet, csub(cmul(m.a, y), cmul(m.c, x)));
    x = nx;
    y = ny;
    ex += (int)m.e;
    renorm(x, y, ex);
}

// Loads of up to SCAN_CACHE transfer matrices, walk steps [i0, i0 + SCAN_CACHE)
// of the walk (clamped to the last segment, so the loads are unconditional
// and the compiler can count them exactly).
template <class SEG>
__device__ __forceinline__ void scan_load(const WLaunch& lp, long long item0, SEG seg, int i0, int P, int lane,
                                          Mat2 (&buf)[SCAN_CACHE])
{
#pragma unroll
    for (int q = 0; q < SCAN_CACHE; ++q) buf[q] = load_T(lp, item0 + seg(min(i0 + q, P - 1)), lane);
}

__global__ __launch_bounds__(64 * SCAN_WAVES) void scan_kernel(WLaunch lp, int ngroups)
{
    __shared__ double hand[5][64];  // state handed from one range to the next
    const int lane = threadIdx.x & 63;
    const int r = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // range, counted from the walk's start
    const int gi = blockIdx.x >> 1;
    const int dir = blockIdx.x & 1;  // 0: u walk, 1: v walk
    if (gi >= ngroups) return;       // uniform over the block
    int g;
    const int t = group_tube(lp, gi, g);
    const WTube tube = tube_of(lp, t);
    if (tube.fused || tube_skipped(lp, tube)) return;  // uniform over the block
    const int P = tube.P, L = tube.L, N = tube.N;
    const int lc = g * 64 + lane;
    const bool valid = lc < tube.nchain;
    const int chain = tube.chain0 + (valid ? lc : tube.nchain - 1);
    const ChainDesc ch = lp.chains[chain];
    const double te = ch.t_even, to = tube.t_odd;
    const double tau = ((N - 1) & 1) ? to : te;  // t_{N-1}
    const double2 SL = make_double2(ch.SLr, ch.SLi);
    const double2 SR = make_double2(ch.SRr, ch.SRi);
    const long long item0 = tube.item0 + (long long)g * P;
    const int per = (P + SCAN_WAVES - 1) / SCAN_WAVES;
    const int r0 = min(P, r * per), cnt = min(P, r0 + per) - r0;  // walk steps [r0, r0 + cnt)
    // segment of walk step i
    auto seg = [dir, P](int i) { return dir == 0 ? i : P - 1 - i; };
    Mat2 buf[SCAN_CACHE];
    if (r == 0) scan_load(lp, item0, seg, r0, P, lane, buf);
    const double idet_odd = te / to;  // 1/det T_s when hi - 1 is even
    for (int k = 0; k < SCAN_WAVES; ++k) {
        if (r == k + 1 && cnt > 0) scan_load(lp, item0, seg, r0, P, lane, buf);
        if (r == k && cnt > 0) {
            double2 x, y;
            int ex = 0;
            if (k == 0) {
                if (dir == 0) {
                    x = make_double2(1.0, 0.0);  // u_0
                    y = cscale(1.0 / to, SL);    // u_{-1}; t_{-1} is an odd link
                } else {
                    x = cscale(1.0 / tau, SR);   // v_N
                    y = make_double2(1.0, 0.0);  // v_{N-1}
                }
            } else {
                x = make_double2(hand[0][lane], hand[1][lane]);
                y = make_double2(hand[2][lane], hand[3][lane]);
                ex = (int)hand[4][lane];
            }
            for (int c = 0; c < cnt; c += SCAN_CACHE) {
                if (c > 0) scan_load(lp, item0, seg, r0 + c, P, lane, buf);
#pragma unroll
                for (int q = 0; q < SCAN_CACHE; ++q) {
                    if (c + q < cnt) {
                        const int s = seg(r0 + c + q);
                        double* S = lp.St + (item0 + s) * 10 * 64 + lane;
                        if (dir == 0) {
                            S[0 * 64] = x.x;  // (u_lo, u_{lo-1})
                            S[1 * 64] = x.y;
                            S[2 * 64] = y.x;
                            S[3 * 64] = y.y;
                            S[4 * 64] = (double)ex;
                            mat2_apply_raw(buf[q], x, y, ex);
                        } else {
                            S[5 * 64] = y.x;  // (v_{hi-1}, v_hi)
                            S[6 * 64] = y.y;
                            S[7 * 64] = x.x;
                            S[8 * 64] = x.y;
                            S[9 * 64] = (double)ex;
                            const int lo = min(N, s * L);
                            const int hi = min(N, lo + L);
                            // det T_s = t_{lo-1} / t_{hi-1}; t_{lo-1} is an odd link (lo even)
                            if (lo < hi) mat2_apply_inv_raw(buf[q], ((hi - 1) & 1) ? 1.0 : idet_odd, x, y, ex);
                        }
                        if (q % SCAN_RENORM == SCAN_RENORM - 1) renorm(x, y, ex);
                    }
                }
            }
            hand[0][lane] = x.x;
            hand[1][lane] = x.y;
            hand[2][lane] = y.x;
            hand[3][lane] = y.y;
            hand[4][lane] = (double)ex;
        }
        // the hand-off goes through LDS only: wait for the LDS writes, then
        // meet (a full __syncthreads would also wait for the walk's stores)
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
    }
    if (dir == 0 && r == 0 && valid) {
        // the u walk ended at the end of the chain: (u_N, u_{N-1})
        double2 uN = make_double2(hand[0][lane], hand[1][lane]);
        double2 uNm1 = make_double2(hand[2][lane], hand[3][lane]);
        int eN = (int)hand[4][lane];
        renorm(uN, uNm1, eN);
        const double2 Dm = csub(cscale(tau, uN), cmul(SR, uNm1));  // D = Dm 2^eN
        double* Dp = lp.Dd + (long long)chain * 3;
        Dp[0] = Dm.x;
        Dp[1] = Dm.y;
        Dp[2] = (double)eN;
        // transmission Gamma_L Gamma_R |G_{0,N-1}|^2, G_{0,N-1} = 1/D
        if (lp.kind == KIND_DOS && lp.chain_out && ch.out_slot >= 0) {
            const double iD2 = 1.0 / cabs2(Dm);
            const double gN = __builtin_amdgcn_ldexp(1.0, -eN);
            const double GamL = -2.0 * SL.y, GamR = -2.0 * SR.y;
            lp.chain_out[(long long)ch.out_slot * 2 + 1] = GamL * GamR * iD2 * gN * gN;
        }
    }
}

// ------------------------------------------------------------ fused transfer + scan
// Chain groups of tubes with FUSED_MIN_SEGS..FUSED_SEGS segments take one
// launch for both steps, and their segment matrices never leave the CU (the
// separate launches write them to memory and read them back twice, once per
// walk).  One workgroup of FUSED_WAVES wavefronts per (64-chain group, half of
// 32 chains).  Wave w computes the transfer matrices of segments w (lanes
// 0-31) and w + FUSED_WAVES (lanes 32-63) of the half's chains into LDS, in
// the arithmetic of transfer_kernel; then the first 32 lanes of wave 0 walk
// them forward (u) and those of wave 1 backward (v), in the arithmetic of
// scan_kernel, writing the same segment states and D.  The states are
// renormalised every SCAN_RENORM steps of the walk (by exact powers of two,
// so the stored (mantissa, exponent) pairs describe the same numbers for any
// schedule).
#define FUSED_WAVES 16
#define FUSED_SEGS (2 * FUSED_WAVES)
#define FUSED_MIN_SEGS 17


// One walk over the segment matrices in LDS (V: the v walk), by the 32 lanes
// of one wavefront, in chunks of CH steps whose matrices are read from LDS
// together, the state renormalised once per chunk (by exact powers of two, so
// the chunk length changes no value); the arithmetic is scan_kernel's.  The
// states go to the segment-state array in memory, as scan_kernel writes them.
#define FUSED_CHUNK 4
template <bool V, int CH = FUSED_CHUNK>
__device__ __forceinline__ void fused_walk(const WLaunch& lp, const double* Tl, const double (*sig)[32],
                                           const WTube& tube, const ChainDesc& ch, int g, int hf, int cl,
                                           bool valid, int chain)
{
    const int P = tube.P, L = tube.L, N = tube.N;
    const double te = ch.t_even, to = tube.t_odd;
    const double tau = ((N - 1) & 1) ? to : te;  // t_{N-1}
    const double2 SL = make_double2(sig[0][cl], sig[1][cl]);
    const double2 SR = make_double2(sig[2][cl], sig[3][cl]);
    double2 x, y;
    if (!V) {
        x = make_double2(1.0, 0.0);  // u_0
        y = cscale(1.0 / to, SL);    // u_{-1}; t_{-1} is an odd link
    } else {
        x = cscale(1.0 / tau, SR);   // v_N
        y = make_double2(1.0, 0.0);  // v_{N-1}
    }
    int ex = 0;
    const double idet_odd = te / to;  // 1/det T_s when hi - 1 is even
    const long long item0 = tube.item0 + (long long)g * P;
    double* Sb = lp.St + item0 * 10 * 64 + (V ? 5 : 0) * 64 + hf * 32 + cl;
    for (int i0 = 0; i0 < P; i0 += CH) {
        Mat2 m[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int i = min(i0 + q, P - 1);
            const double* T = Tl + (V ? P - 1 - i : i) * 9 * 32 + cl;
            m[q].a = make_double2(T[0 * 32], T[1 * 32]);
            m[q].b = make_double2(T[2 * 32], T[3 * 32]);
            m[q].c = make_double2(T[4 * 32], T[5 * 32]);
            m[q].d = make_double2(T[6 * 32], T[7 * 32]);
            m[q].e = T[8 * 32];
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            if (i0 + q < P) {  // wave-uniform
                const int sg = V ? P - 1 - (i0 + q) : i0 + q;
                double* S = Sb + (long long)sg * 10 * 64;
                if (!V) {
                    S[0 * 64] = x.x;  // (u_lo, u_{lo-1})
                    S[1 * 64] = x.y;
                    S[2 * 64] = y.x;
                    S[3 * 64] = y.y;
                    S[4 * 64] = (double)ex;
                    mat2_apply_raw(m[q], x, y, ex);
                } else {
                    S[0 * 64] = y.x;  // (v_{hi-1}, v_hi)
                    S[1 * 64] = y.y;
                    S[2 * 64] = x.x;
                    S[3 * 64] = x.y;
                    S[4 * 64] = (double)ex;
                    const int hi = min(N, sg * L + L);
                    mat2_apply_inv_raw(m[q], ((hi - 1) & 1) ? 1.0 : idet_odd, x, y, ex);
                }
                if (q == CH - 1) renorm(x, y, ex);
            }
        }
    }
    if (!V && valid) {
        // the u walk ended at the end of the chain: (u_N, u_{N-1})
        renorm(x, y, ex);
        const double2 Dm = csub(cscale(tau, x), cmul(SR, y));  // D = Dm 2^ex
        double* Dp = lp.Dd + (long long)chain * 3;
        Dp[0] = Dm.x;
        Dp[1] = Dm.y;
        Dp[2] = (double)ex;
        // transmission Gamma_L Gamma_R |G_{0,N-1}|^2, G_{0,N-1} = 1/D
        if (lp.kind == KIND_DOS && lp.chain_out && ch.out_slot >= 0) {
            const double iD2 = 1.0 / cabs2(Dm);
            const double gN = __builtin_amdgcn_ldexp(1.0, -ex);
            const double GamL = -2.0 * SL.y, GamR = -2.0 * SR.y;
            lp.chain_out[(long long)ch.out_slot * 2 + 1] = GamL * GamR * iD2 * gN * gN;
        }
    }
}

// The walk of fused_walk with the states written to LDS (Sl[(segment * 8 +
// c) * CW + chain], c 0-3: u_lo, u_{lo-1}; 4-7: v_{hi-1}, v_hi; the v
// exponent to Se[segment * CW + chain]), each step split over the two halves
// of the wavefront (lane h * 32 + c: chain c): half 0 forms the new x, half 1 the
// new y, each from its two matrix entries, and one v_permlane32_swap per
// dword gives every lane both (with both operands the same register N, the
// swap's first result is the lower half's N in every lane and its second the
// upper half's).  u walk: N = P x + Q y with (P, Q) = (a, b) | (c, d); v walk:
// N = (P x + Q y) / det with (P, Q) = (d, -b) | (-c, a).  A half stores the
// value it formed itself, so the state rows 0-1 | 2-3 (u) and 6-7 | 4-5 (v)
// come from halves 0 | 1.  With vprog the v walk publishes, after each chunk,
// how many segments (from the right end) have their states stored (release:
// a wavefront that reads the count with acquire sees those states).
__device__ __forceinline__ double swap_lo(double v, double& hi)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto l = __builtin_amdgcn_permlane32_swap((unsigned)b, (unsigned)b, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap((unsigned)(b >> 32), (unsigned)(b >> 32), false, false);
    hi = __longlong_as_double((long long)(((unsigned long long)h[1] << 32) | l[1]));
    return __longlong_as_double((long long)(((unsigned long long)h[0] << 32) | l[0]));
}

template <bool V, int CH, int CW = 32>
__device__ __forceinline__ void fused_walk_split(const WLaunch& lp, const double* Tl, const double (*sig)[CW],
                                                 const WTube& tube, const ChainDesc& ch, int cl, int h, bool valid,
                                                 int chain, double* Sl, int* Se, int* vprog = nullptr)
{
    const int P = tube.P, L = tube.L, N = tube.N;
    const double te = ch.t_even, to = tube.t_odd;
    const double tau = ((N - 1) & 1) ? to : te;  // t_{N-1}
    const double2 SL = make_double2(sig[0][cl], sig[1][cl]);
    const double2 SR = make_double2(sig[2][cl], sig[3][cl]);
    double2 x, y;
    if (!V) {
        x = make_double2(1.0, 0.0);  // u_0
        y = cscale(1.0 / to, SL);    // u_{-1}; t_{-1} is an odd link
    } else {
        x = cscale(1.0 / tau, SR);   // v_N
        y = make_double2(1.0, 0.0);  // v_{N-1}
    }
    double2 own = h ? y : x;  // the component this half forms
    int ex = 0;
    const double idet_odd = te / to;  // 1/det T_s when hi - 1 is even
    // this half's matrix entries (doubles of the 9-double slot): u: a b | c d; v: d b | c a
    const int oP = V ? (h ? 4 : 6) : (h ? 4 : 0);
    const int oQ = V ? (h ? 0 : 2) : (h ? 6 : 2);
    const double sP = (V && h) ? -1.0 : 1.0, sQ = (V && !h) ? -1.0 : 1.0;
    const int srow = V ? 6 - 2 * h : 2 * h;  // state rows this half stores
    for (int i0 = 0; i0 < P; i0 += CH) {
        double2 pm[CH], qm[CH];
        double em[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int i = min(i0 + q, P - 1);
            const double* T = Tl + (V ? P - 1 - i : i) * 9 * CW + cl;
            pm[q] = make_double2(sP * T[oP * CW], sP * T[(oP + 1) * CW]);  // signs: exact
            qm[q] = make_double2(sQ * T[oQ * CW], sQ * T[(oQ + 1) * CW]);
            em[q] = T[8 * CW];
        }
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            if (i0 + q < P) {  // wave-uniform
                const int sg = V ? P - 1 - (i0 + q) : i0 + q;
                double* S = Sl + (sg * 8 + srow) * CW + cl;
                S[0] = own.x;
                S[CW] = own.y;
                if (V && h == 0) Se[sg * CW + cl] = ex;
                double2 n = cadd(cmul(pm[q], x), cmul(qm[q], y));
                if (V) {
                    const int hi = min(N, sg * L + L);
                    n = cscale(((hi - 1) & 1) ? 1.0 : idet_odd, n);
                }
                own = n;
                double hx, hy;
                x.x = swap_lo(n.x, hx);
                x.y = swap_lo(n.y, hy);
                y = make_double2(hx, hy);
                ex += (int)em[q];
                if (q == CH - 1) {
                    renorm(x, y, ex);
                    own = h ? y : x;
                }
            }
        }
        if (V && vprog)
            __hip_atomic_store(vprog, min(P, i0 + CH), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (!V && valid && h == 0) {
        // the u walk ended at the end of the chain: (u_N, u_{N-1})
        renorm(x, y, ex);
        const double2 Dm = csub(cscale(tau, x), cmul(SR, y));  // D = Dm 2^ex
        double* Dp = lp.Dd + (long long)chain * 3;
        Dp[0] = Dm.x;
        Dp[1] = Dm.y;
        Dp[2] = (double)ex;
    }
}

__global__ __launch_bounds__(64 * FUSED_WAVES, 1) void transfer_scan_kernel(WLaunch lp, int ngroups)
{
    __shared__ double Tl[FUSED_SEGS * 9 * 32];                // [segment][component][chain]
    __shared__ double us_all[FUSED_WAVES][2][NEGF_SEG_MAX];   // potentials of each wave's two segments
    __shared__ double sig[4][32];                             // SL, SR of the half's chains
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gi = blockIdx.x >> 1;
    const int hf = blockIdx.x & 1;
    if (gi >= ngroups) return;  // uniform over the block
    int g;
    const int t = group_tube(lp, gi, g);
    WTube tube = tube_of(lp, t);
    tube.N = __builtin_amdgcn_readfirstlane(tube.N);
    tube.L = __builtin_amdgcn_readfirstlane(tube.L);
    tube.P = __builtin_amdgcn_readfirstlane(tube.P);
    tube.nchain = __builtin_amdgcn_readfirstlane(tube.nchain);
    // (a half without chains still runs, on copies of the group's last chain:
    // the sweeps read the states of all 64 lanes and mask the contributions)
    if (tube.fused != 1 || tube_skipped(lp, tube)) return;  // uniform over the block
    const int P = tube.P, L = tube.L, N = tube.N;
    const int h = lane >> 5, cl = lane & 31;
    const int lc = g * 64 + hf * 32 + cl;
    const bool valid = lc < tube.nchain;
    const int chain = tube.chain0 + (valid ? lc : tube.nchain - 1);
    const ChainDesc ch = lp.chains[chain];
    const double2 E = make_double2(ch.Er, ch.Ei);
    const double te = ch.t_even, to = tube.t_odd;

    // potentials of the wave's two segments (zero past the end of the tube);
    // every load is issued before the first LDS write
    {
        constexpr int R = 2 * NEGF_SEG_MAX / 64;
        double uv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = lane + 64 * q;
            const int hh = i >= L ? 1 : 0;
            const int site = (w + FUSED_WAVES * hh) * L + (i - hh * L);
            uv[q] = tube.U[min(site, N - 1)];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = lane + 64 * q;
            const int hh = i >= L ? 1 : 0;
            const int site = (w + FUSED_WAVES * hh) * L + (i - hh * L);
            if (i < 2 * L) us_all[w][hh][i - hh * L] = site < N ? uv[q] : 0.0;
        }
    }
    // lead self-energies: the lanes of segment 0 (closed form, also stored for
    // the sweeps), or as the decimation launch left them in the chain table
    if (w == 0 && h == 0) {
        double2 SL = make_double2(ch.SLr, ch.SLi), SR = make_double2(ch.SRr, ch.SRi);
        if (!lp.dec_use) lead_sigma<false>(lp, valid ? lp.chains + chain : nullptr, ch, tube, &SL, &SR);
        sig[0][cl] = SL.x;
        sig[1][cl] = SL.y;
        sig[2][cl] = SR.x;
        sig[3][cl] = SR.y;
    }
    __builtin_amdgcn_wave_barrier();

    const int s = w + FUSED_WAVES * h;  // this lane's segment
    const int lo = min(N, s * L);
    const int n = s < P ? min(N, lo + L) - lo : 0;
    const double* us = us_all[w][h];
    double2 m00 = make_double2(1.0, 0.0), m01 = make_double2(0.0, 0.0);
    double2 m10 = make_double2(0.0, 0.0), m11 = make_double2(1.0, 0.0);
    int eT = 0;
    if (__all(scaled_ok(E, te, to))) {
        const SCoefs ks = make_scoefs(E, te, to);
        if (n > 0) {
            const int hi = lo + n;
            m11 = make_double2(((lo - 1) & 1) ? ks.ai[1] : ks.ai[0], 0.0);
            transfer_loop<NEGF_SUB_K>(ks, us, n, m00, m01, m10, m11, eT);
            double mC;
            int eC;
            scale_prod(ks, lo & 1, n - 1, mC, eC);
            const double r0 = mC * (((hi - 1) & 1) ? ks.ai[1] : ks.ai[0]);
            m00 = cscale(r0, m00);
            m01 = cscale(r0, m01);
            m10 = cscale(mC, m10);
            m11 = cscale(mC, m11);
            eT += eC;
            renorm_mat(m00, m01, m10, m11, eT);
        }
    } else {
        const Coefs k = make_coefs(E, te, to);
        if (n > 0) transfer_loop<NEGF_SUB_K>(k, us, n, m00, m01, m10, m11, eT);
    }
    if (s < P) {
        double* T = Tl + s * 9 * 32 + cl;
        T[0 * 32] = m00.x;
        T[1 * 32] = m00.y;
        T[2 * 32] = m01.x;
        T[3 * 32] = m01.y;
        T[4 * 32] = m10.x;
        T[5 * 32] = m10.y;
        T[6 * 32] = m11.x;
        T[7 * 32] = m11.y;
        T[8 * 32] = (double)eT;
    }
    __syncthreads();
    if (w >= 2 || lane >= 32) return;
    if (w == 0)
        fused_walk<false>(lp, Tl, sig, tube, ch, g, hf, cl, valid, chain);
    else
        fused_walk<true>(lp, Tl, sig, tube, ch, g, hf, cl, valid, chain);
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
// H = 2: the two halves of the wavefront (lanes 0-31, 32-63) are summed
// separately; lane i < K holds the first half's sum of site i and lane 32 + i
// the second half's.
template <int K, int H = 1>
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
    for (int d = K; d < 64 / H; d <<= 1) acc += __shfl_xor(acc, d, 64);
    __builtin_amdgcn_wave_barrier();
    return acc;
}

// Pass-3 sites of one sub-block [p0, p0+len): contributions q[i], optional
// G output and integrand samples, forward u step.
struct SubCtx {
    double2 C, Gs;          // KIND_EQ: contribution Re(C u v), G = Gs u v
    double kL, kR, aL, aR;  // KIND_NONEQ / DOS weights of |v|^2 and |u|^2
};

// Row outputs of a sweep wavefront ([chain][N] arrays: diag G^R and the
// per-site DOS).  A lane's values are staged in LDS and written as whole
// 128-byte pieces of one chain's row (8 complex or 16 real sites, one piece
// per 8 or 16 lanes of a store), instead of one 16- or 8-byte store per lane
// and site into 64 different rows, whose partly written lines leave the L2
// one by one (2.7x the G^R bytes in HBM writes).
struct RowOut {
    char* base = nullptr;  // row of the wavefront's first chain
    int N = 0;             // row length (sites)
    int nvalid = 0;        // lanes 0 .. nvalid-1 have a row
    double* st = nullptr;  // the wavefront's LDS staging, 1024 doubles
};

// write sites [s0, s0 + n) of the staged values st[site % S][lane]
template <class T>
__device__ __forceinline__ void row_flush(const RowOut& o, int s0, int n, int lane)
{
    constexpr int S = 128 / (int)sizeof(T);  // sites per 128-byte piece (lanes per chain)
    constexpr int CPI = 64 / S;              // chains per store
    __builtin_amdgcn_wave_barrier();
    const T* st = reinterpret_cast<const T*>(o.st);
    T* base = reinterpret_cast<T*>(o.base);
#pragma unroll 2
    for (int r = 0; r < 64 / CPI; ++r) {  // two at a time: the sweep's registers are nearly all in use
        const int cc = CPI * r + lane / S, sl = lane % S;
        const T v = st[sl * 64 + cc];
        if (sl < n && cc < o.nvalid) base[(long long)cc * o.N + s0 + sl] = v;
    }
    __builtin_amdgcn_wave_barrier();
}

template <int K, int KIND, bool WRITE_G, bool FULL, bool SAMPLE>
__device__ __forceinline__ void site_loop(const Coefs& k, const double (&Us)[K], const double2 (&vr)[K], double2& ux,
                                          double2& uy, const SubCtx& c, int p0, int len, bool valid,
                                          const RowOut& go, const RowOut& so, double* red, int lane, double& dos_acc,
                                          const WTube& tube, const ChainDesc& ch, double* chain_out)
{
    static_assert(K <= 16, "one DOS row piece per sub-block");
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (!FULL && KIND != KIND_DOS) red[i * 65 + lane] = 0.0;
        if (FULL || i < len) {
            if (KIND == KIND_EQ) {
                double2 g = cmul(ux, vr[i]);
                red[i * 65 + lane] = fma(c.C.x, g.x, -c.C.y * g.y);
                if (WRITE_G) {
                    reinterpret_cast<double2*>(go.st)[(i & 7) * 64 + lane] = cmul(c.Gs, g);
                    if ((i & 7) == 7 || (!FULL && i == len - 1)) row_flush<double2>(go, p0 + (i & ~7), (i & 7) + 1, lane);
                }
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
                const double a = c.aL * cabs2(vr[i]) + c.aR * cabs2(ux);
                dos_acc += a;
                if (so.base) {
                    so.st[i * 64 + lane] = a;
                    if (i == K - 1 || (!FULL && i == len - 1)) row_flush<double>(so, p0, i + 1, lane);
                }
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
    __shared__ double us_all[4][NEGF_SEG_MAX];
    constexpr bool ROWS = WRITE_G || KIND == KIND_DOS;  // row outputs staged in LDS
    __shared__ double rowst_all[ROWS ? 4 : 1][ROWS ? 1024 : 1];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int idx = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    if (idx >= lp.nlist) return;
    const int item = lp.item_list ? __builtin_amdgcn_readfirstlane(lp.item_list[idx]) : idx;
    double* red = red_all[wv];
    WTube tube;
    ItemCtx ic = item_ctx(lp, item, lane, tube);
    if (tube_skipped(lp, tube)) return;  // whole wavefront
    const ChainDesc ch = lp.chains[ic.chain];
    const Coefs k = make_coefs(make_double2(ch.Er, ch.Ei), ch.t_even, tube.t_odd);
    const int N = tube.N;
    const int lo = min(N, ic.s * tube.L);
    const int hi = min(N, lo + tube.L);
    const int nsub = (hi - lo + K - 1) / K;  // wave-uniform
    double* us = us_all[wv];
    stage_segment(us, tube.U, lo, hi, lane);
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
            const int q0 = j * K;
            const int len = min(K, hi - lo - q0);
            if (len == K) {
#pragma unroll
                for (int i = K - 1; i >= 0; i -= 2) {
                    bwd_step<1>(k, us[q0 + i], x, y);
                    bwd_step<0>(k, us[q0 + i - 1], x, y);
                }
            } else {
                for (int i = len - 1; i >= 0; --i) {
                    if (i & 1)
                        bwd_step<1>(k, us[q0 + i], x, y);
                    else
                        bwd_step<0>(k, us[q0 + i], x, y);
                }
            }
            renorm(x, y, e);
            ck.put(j - 1, x, y, e);
        }
    }

    // ---------------- pass 3 ----------------
    double2 iW = make_double2(0.0, 0.0);  // KIND_EQ: 1 / mantissa of the segment Wronskian
    int eW = 0;                           // and its exponent
    double dos_acc = 0.0;
    RowOut go, so;
    {
        const long long row0 = (long long)(tube.chain0 + ic.g * 64) * N;
        const int nvalid = min(64, tube.nchain - ic.g * 64);
        if (WRITE_G) go = RowOut{reinterpret_cast<char*>(lp.diagG + row0), N, nvalid, rowst_all[ROWS ? wv : 0]};
        if (KIND == KIND_DOS && lp.site_out)
            so = RowOut{reinterpret_cast<char*>(lp.site_out + row0), N, nvalid, rowst_all[ROWS ? wv : 0]};
    }
    for (int j = 0; j < nsub; ++j) {
        const int p0 = lo + j * K;
        const int len = min(K, hi - p0);
        double2 cx, cy;
        int ce;
        ck.get(j, cx, cy, ce);
        double2 vr[K];
        double2 v1;  // v_{p0+1}
        double Us[K];
#pragma unroll
        for (int i = 0; i < K; ++i) Us[i] = (i < len) ? us[p0 - lo + i] : 0.0;
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
            v1 = y;  // the last step (or the checkpoint itself when len == 1) left v_{p0+1} here
        }
        renorm(ux, uy, ue);
        if (KIND == KIND_EQ && j == 0) {
            // G_nn = u_n v_n / W with the Wronskian W = t_lo (u_{lo+1} v_lo - u_lo v_{lo+1})
            // of this sweep's own u and v (see eq_full_body): W = Wm 2^(ue + ce)
            double2 u1 = ux, u0 = uy;
            fwd_step<0>(k, Us[0], u1, u0);
            iW = crecip(cscale(ch.t_even, csub(cmul(u1, vr[0]), cmul(u0, v1))));
            eW = ue + ce;
        }
        const double sc_uv = __builtin_amdgcn_ldexp(1.0, ue + ce - eN);
        const double sc_w = __builtin_amdgcn_ldexp(1.0, ue + ce - eW);
        const double2 C = KIND == KIND_EQ ? cscale(sc_w, cmul(coef, iW)) : cscale(sc_uv, cmul(coef, iD));
        const double sv = __builtin_amdgcn_ldexp(1.0, ce - eN);
        const double su = __builtin_amdgcn_ldexp(1.0, ue - eN);
        const double aL = GamL * sv * sv * iD2;
        const double aR = GamR * su * su * iD2;
        const double kL = ch.cL * vmask * aL;
        const double kR = ch.cR * vmask * aR;
        SubCtx sc;
        sc.C = C;
        sc.Gs = KIND == KIND_EQ ? cscale(sc_w, iW) : cscale(sc_uv, iD);
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
            site_loop<K, KIND, WRITE_G, false, true>(k, Us, vr, ux, uy, sc, p0, len, ic.valid, go, so, red, lane,
                                                     dos_acc, tube, ch, lp.chain_out);
        else if (len == K)
            site_loop<K, KIND, WRITE_G, true, false>(k, Us, vr, ux, uy, sc, p0, len, ic.valid, go, so, red, lane,
                                                     dos_acc, tube, ch, lp.chain_out);
        else
            site_loop<K, KIND, WRITE_G, false, false>(k, Us, vr, ux, uy, sc, p0, len, ic.valid, go, so, red, lane,
                                                      dos_acc, tube, ch, lp.chain_out);
        if (KIND != KIND_DOS) {
            double r = lane_sum_transpose<K>(red, lane);
            if (lane < len) prow[p0 + lane] = r;
        }
    }
    if (KIND == KIND_DOS) lp.dos_part[(long long)item * 64 + lane] = dos_acc;
}

// Passes 2 and 3 of the equilibrium sweep over a full segment of K NS sites
// staged in us: (x, y) is the right state (exponent e), (ux, uy) the left
// state.  G_nn = u_n v_n / W with W the Wronskian
//   W = t_lo (u_{lo+1} v_lo - u_lo v_{lo+1})
// of the two solutions this sweep itself runs through the segment (u forward
// from the left state, v backward from the right state), taken at the
// segment's first link once v_lo and v_{lo+1} are known.  The exact
// recurrences keep W constant along the chain, so this is the same G as
// u_n v_n / D with D = t_{N-1} u_N - Sigma_R u_{N-1}; evaluated per segment
// it cancels the normalisation error that the segment products of the scan
// accumulate in u and v (which grows like the chain length times the
// segment length near band edges, where the transfer matrices are nearly
// defective), leaving only the error of the direct recurrences inside one
// segment.  wp and wq give W's mantissa as wp u1 v0 - wq u0 v1 in the
// recurrence's own variables (t_lo, t_lo for the plain recurrence; Im E and
// t_lo^2 / Im E for the scaled one, where the segment constant C cancels).
// Pass 2 of a full-segment equilibrium sweep: the v checkpoints (x, y, e) at
// the right end of every sub-block, from the segment's right-end state.
template <int K, int NS, class KC>
__device__ __forceinline__ void eq_checkpoints(const KC& k, const double* us, double2 x, double2 y, int e,
                                               double2 (&cx)[NS], double2 (&cy)[NS], int (&ce)[NS])
{
    cx[NS - 1] = x;
    cy[NS - 1] = y;
    ce[NS - 1] = e;
#pragma unroll
    for (int j = NS - 1; j >= 1; --j) {
#pragma unroll
        for (int i = K - 1; i >= 1; i -= 2) {
            bstep<1>(k, us[j * K + i], x, y);
            bstep<0>(k, us[j * K + i - 1], x, y);
        }
        renorm(x, y, e);
        cx[j - 1] = x;
        cy[j - 1] = y;
        ce[j - 1] = e;
    }
}

// Pass 3: per sub-block, v recomputed from its checkpoint, then the forward
// u steps with the contributions, summed over the lanes of each half (H) into
// the partial row prow.
template <int K, int NS, class KC, int H = 1>
__device__ __forceinline__ void eq_sweep_pass3(const KC& k, const double* us, double2 ux, double2 uy,
                                               const double2 (&cx)[NS], const double2 (&cy)[NS],
                                               const int (&ce)[NS], double2 coef, double wp, double wq,
                                               double* red, int lane, double* prow)
{
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double2 vr[K];
        {
            double2 xx = cx[j], yy = cy[j];
            vr[K - 1] = xx;
#pragma unroll
            for (int i = K - 1; i >= 1; --i) {
                if (i & 1)
                    bstep<1>(k, us[j * K + i], xx, yy);
                else
                    bstep<0>(k, us[j * K + i], xx, yy);
                vr[i - 1] = xx;
            }
        }
        if (j == 0) {
            // ut = (coef / W) u, with W from (u_lo, u_{lo+1}) and (v_lo, v_{lo+1});
            // the exponents of u and v cancel between u v and W
            double2 u1 = ux, u0 = uy;
            fstep<0>(k, us[0], u1, u0);  // u1 = u_{lo+1}, u0 = u_lo
            const double2 Wm = csub(cscale(wp, cmul(u1, vr[0])), cscale(wq, cmul(u0, vr[1])));
            const double2 C0 = cmul(coef, crecip(Wm));
            ux = cmul(C0, ux);
            uy = cmul(C0, uy);
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            red[i * 65 + lane] = fma(ux.x, vr[i].x, -ux.y * vr[i].y);
            if (i & 1)
                fstep<1>(k, us[j * K + i], ux, uy);
            else
                fstep<0>(k, us[j * K + i], ux, uy);
        }
        const double r = lane_sum_transpose<K, H>(red, lane);
        if ((lane & (64 / H - 1)) < K && prow) prow[j * K + (lane & (K - 1))] = r;
        if (j + 1 < NS) {
            const int d = ce[j + 1] - ce[j];
            scale2(ux, d);
            scale2(uy, d);
        }
    }
}

template <int K, int NS, class KC, int H = 1>
__device__ __forceinline__ void eq_full_body(const KC& k, const double* us, double2 ux, double2 uy, double2 x,
                                             double2 y, int e, double2 coef, double wp, double wq, double* red,
                                             int lane, double* prow)
{
    double2 cx[NS], cy[NS];
    int ce[NS];
    eq_checkpoints<K, NS>(k, us, x, y, e, cx, cy, ce);
    eq_sweep_pass3<K, NS, KC, H>(k, us, ux, uy, cx, cy, ce, coef, wp, wq, red, lane, prow);
}

// Equilibrium sweep of a full segment of NS sub-blocks (the common case),
// with the chain coefficient folded into the left state: with
//   ut = (coef / W) u
// (a solution of the same linear recurrence) the contribution of site n is
// Re(ut_n v_n), two instructions.  Between sub-blocks ut only changes by the
// exact power of two of the next v checkpoint, so u needs no renormalisation:
// |ut_n v_n| = |coef G_nn| stays bounded.  Checkpoint indices are static.
// Waves whose chains all have t / |Im E| <= 2^24 run the scaled recurrence:
// the entry states become w_lo = u_lo, w_{lo-1} = ai_{lo-1} u_{lo-1},
// w'_{hi-1} = v_{hi-1}, w'_hi = ai_{hi-1} v_hi (u_n v_n = C w_n w'_n with the
// segment constant C, which also divides W and so drops out).
template <int K, int NS>
__global__ __launch_bounds__(256, NS <= 4 ? 3 : 2) void sweep_eq_full_kernel(WLaunch lp)
{
    __shared__ double red_all[4][K * 65];
    __shared__ double us_all[4][K * NS];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int idx = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    if (idx >= lp.nlist) return;
    const int item = __builtin_amdgcn_readfirstlane(lp.item_list[idx]);
    double* red = red_all[wv];
    WTube tube;
    ItemCtx ic = item_ctx(lp, item, lane, tube);
    if (tube_skipped(lp, tube)) return;  // whole wavefront
    const ChainDesc ch = lp.chains[ic.chain];
    const double2 E = make_double2(ch.Er, ch.Ei);
    const int lo = ic.s * tube.L;  // the segment is [lo, lo + K NS), lo even
    double* us = us_all[wv];
    stage_segment(us, tube.U, lo, lo + K * NS, lane);
    const double* S = lp.St + (long long)item * 10 * 64 + lane;
    double2 ux = make_double2(S[0], S[64]), uy = make_double2(S[128], S[192]);
    double2 x = make_double2(S[320], S[384]), y = make_double2(S[448], S[512]);
    const int e = (int)S[576];
    const double vmask = ic.valid ? 1.0 : 0.0;
    const double2 coef = make_double2(ch.cr * vmask, ch.ci * vmask);
    double* prow = lp.part + tube.part_off + (long long)ic.g * tube.N + lo;
    const double te = ch.t_even;  // t_lo: lo is even

    if (__all(scaled_ok(E, ch.t_even, tube.t_odd))) {
        const SCoefs ks = make_scoefs(E, ch.t_even, tube.t_odd);
        uy = cscale(ks.ai[1], uy);  // site lo - 1 is odd
        y = cscale(ks.ai[1], y);    // site hi - 1 is odd
        // W / C = t_lo (ai_lo w_{lo+1} w'_lo - w_lo w'_{lo+1} / ai_lo)
        eq_full_body<K, NS>(ks, us, ux, uy, x, y, e, coef, E.y, te * te * ks.ie, red, lane, prow);
    } else {
        const Coefs k = make_coefs(E, ch.t_even, tube.t_odd);
        eq_full_body<K, NS>(k, us, ux, uy, x, y, e, coef, te, te, red, lane, prow);
    }
}

// ------------------------------------------------------------ fused equilibrium charge
// Equilibrium charge of tubes whose 17..32 segments are all full (N a
// multiple of L = K NS), without diag G^R: transfer matrices, walks and the
// sweep in one launch, the segment states kept in LDS.  One workgroup of
// EQF_WAVES wavefronts per (64-chain group, half of 32 chains); in phase 1
// wavefront w takes segments w + EQF_WAVES q, q = 0..3 (lanes 0-31: q = 0, 2;
// lanes 32-63: q = 1, 3), in phase 3 work items from a counter.
//   phase 1  transfer matrices of its segments into Tl (transfer_kernel's
//            arithmetic), the lead self-energies by the lanes of segment 0;
//   phase 2  the u walk (wavefront 0) and the v walk (wavefront 1) over Tl,
//            writing the states into LDS (scan_kernel's arithmetic); the v
//            walk publishes its progress, and wavefronts 2.. run the
//            checkpoint pass of one sweep item each as soon as its segments'
//            v states are stored;
//   phase 3  the sweep of each segment (sweep_eq_full_kernel's arithmetic),
//            two segments per item (one per half of the wavefront), each
//            half's 32 contributions summed into its own partial row: the tube
//            has two rows per chain group.
// LDS: Tl (72 KiB) is dead after phase 2 and holds the sweep's transpose
// buffers in phase 3; the states (68 KiB), the staged potentials of a pass
// (16 KiB) and the lead self-energies (1 KiB) complete the 157 KiB.  Two
// wavefronts per SIMD (the sweep's registers), one workgroup per CU.
#define EQF_WAVES 8
#define EQF_WAIT_MAX (1 << 22)  // bound on a wavefront's polls of the v walk's progress
#define EQF_TL_BYTES (FUSED_SEGS * 9 * 32 * 8)
#define EQF_ST_BYTES (FUSED_SEGS * 8 * 32 * 8)
#define EQF_SE_BYTES (FUSED_SEGS * 32 * 4)
#define EQF_US_BYTES (EQF_WAVES * 2 * NEGF_SEG_MAX * 8)
#define EQF_LDS_BYTES (EQF_TL_BYTES + EQF_ST_BYTES + EQF_SE_BYTES + EQF_US_BYTES + 4 * 32 * 8)

template <int K, int NS>
__global__ __launch_bounds__(64 * EQF_WAVES, 1) void rgf_eq_fused_kernel(WLaunch lp, int ngroups)
{
    static_assert(EQF_WAVES * K * 65 * 8 <= EQF_TL_BYTES, "transpose buffers must fit in the Tl region");
    constexpr int L = K * NS;
    __shared__ __align__(16) unsigned char smem[EQF_LDS_BYTES];
    double* Tl = reinterpret_cast<double*>(smem);
    double* Sl = reinterpret_cast<double*>(smem + EQF_TL_BYTES);
    int* Se = reinterpret_cast<int*>(smem + EQF_TL_BYTES + EQF_ST_BYTES);
    double* us_all = reinterpret_cast<double*>(smem + EQF_TL_BYTES + EQF_ST_BYTES + EQF_SE_BYTES);
    double(*sig)[32] =
        reinterpret_cast<double(*)[32]>(smem + EQF_TL_BYTES + EQF_ST_BYTES + EQF_SE_BYTES + EQF_US_BYTES);
    __shared__ int next_item;  // phase 3 work counter
    __shared__ int v_done;     // segments (from the right end) whose v states the v walk has stored
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int gi = blockIdx.x >> 1;
    const int hf = blockIdx.x & 1;
    if (gi >= ngroups) return;  // uniform over the block
    int g;
    const int t = group_tube(lp, gi, g);
    WTube tube = tube_of(lp, t);
    tube.N = __builtin_amdgcn_readfirstlane(tube.N);
    tube.P = __builtin_amdgcn_readfirstlane(tube.P);
    tube.nchain = __builtin_amdgcn_readfirstlane(tube.nchain);
    if (tube.fused != 2 || tube_skipped(lp, tube)) return;  // uniform over the block
    const int P = tube.P, N = tube.N;
    const int h = lane >> 5, cl = lane & 31;
    const int lc = g * 64 + hf * 32 + cl;
    const bool valid = lc < tube.nchain;
    const int chain = tube.chain0 + (valid ? lc : tube.nchain - 1);
    const ChainDesc ch = lp.chains[chain];
    const double2 E = make_double2(ch.Er, ch.Ei);
    const double te = ch.t_even, to = tube.t_odd;
    const bool scaled = __all(scaled_ok(E, te, to));
    double* us_w = us_all + w * 2 * L;  // [half][L] of this wavefront's pass
    const int nwork = (P + 1) / 2;      // sweep items
    const int npre = min(EQF_WAVES - 2, nwork);
    // segment of item j in half hh (negative: none, when P is odd)
    auto seg_of = [P](int j, int hh) { return P - 1 - 2 * j - hh; };
    // the potentials of item j's two segments into us_w (zeros for a missing one)
    auto stage_item = [&](int j) {
        constexpr int R = 2 * L / 64;
        double uv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = lane + 64 * q, hh = i / L;
            const int sg = seg_of(j, hh);
            uv[q] = tube.U[max(sg, 0) * L + (i - hh * L)];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = lane + 64 * q, hh = i / L;
            us_w[i] = seg_of(j, hh) >= 0 ? uv[q] : 0.0;
        }
    };

    // ---- phase 1: transfer matrices.  Each lane steps its two segments
    // (pass 0 and pass 1) together; their potentials are staged in the
    // segments' own Tl slots (a wavefront reads them there before it writes
    // the matrices over them, and no other wavefront touches those slots).
    {
        const int sa = w + EQF_WAVES * h, sb = w + EQF_WAVES * (2 + h);  // this lane's segments
        constexpr int R = 4 * L / 64;
        double uv[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = lane + 64 * q, k4 = i / L;  // k4 = 2 pass + half
            const int site = (w + EQF_WAVES * k4) * L + (i - k4 * L);
            uv[q] = tube.U[min(site, N - 1)];
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = lane + 64 * q, k4 = i / L;
            const int sg = w + EQF_WAVES * k4;
            if (sg < P) Tl[sg * 9 * 32 + (i - k4 * L)] = uv[q];
        }
        if (threadIdx.x == 0) {
            next_item = npre;
            v_done = 0;
        }
        if (w == 0 && h == 0) {
            double2 SL = make_double2(ch.SLr, ch.SLi), SR = make_double2(ch.SRr, ch.SRi);
            if (!lp.dec_use) lead_sigma<false>(lp, valid ? lp.chains + chain : nullptr, ch, tube, &SL, &SR);
            sig[0][cl] = SL.x;
            sig[1][cl] = SL.y;
            sig[2][cl] = SR.x;
            sig[3][cl] = SR.y;
        }
        __builtin_amdgcn_wave_barrier();
        // a segment past the end (P < 32) runs on segment 0's potentials; its result is not stored
        const double* ua = Tl + (sa < P ? sa : 0) * 9 * 32;
        const double* ub = Tl + (sb < P ? sb : 0) * 9 * 32;
        Mat2x A, B;
        A.m00 = B.m00 = make_double2(1.0, 0.0);
        A.m01 = B.m01 = make_double2(0.0, 0.0);
        A.m10 = B.m10 = make_double2(0.0, 0.0);
        A.m11 = B.m11 = make_double2(1.0, 0.0);
        A.e = B.e = 0;
        if (scaled) {
            const SCoefs ks = make_scoefs(E, te, to);
            A.m11 = B.m11 = make_double2(ks.ai[1], 0.0);  // site lo - 1 is odd
            transfer_loop2<K, NS>(ks, ua, ub, A, B);
            double mC;
            int eC;
            scale_prod(ks, 0, L - 1, mC, eC);
            const double r0 = mC * ks.ai[1];  // site hi - 1 is odd
            Mat2x* M[2] = {&A, &B};
#pragma unroll
            for (int z = 0; z < 2; ++z) {
                M[z]->m00 = cscale(r0, M[z]->m00);
                M[z]->m01 = cscale(r0, M[z]->m01);
                M[z]->m10 = cscale(mC, M[z]->m10);
                M[z]->m11 = cscale(mC, M[z]->m11);
                M[z]->e += eC;
                renorm_mat(M[z]->m00, M[z]->m01, M[z]->m10, M[z]->m11, M[z]->e);
            }
        } else {
            const Coefs k = make_coefs(E, te, to);
            transfer_loop2<K, NS>(k, ua, ub, A, B);
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read its potentials
        const Mat2x* M[2] = {&A, &B};
        const int sg2[2] = {sa, sb};
#pragma unroll
        for (int z = 0; z < 2; ++z) {
            if (sg2[z] < P) {
                double* T = Tl + sg2[z] * 9 * 32 + cl;
                T[0 * 32] = M[z]->m00.x;
                T[1 * 32] = M[z]->m00.y;
                T[2 * 32] = M[z]->m01.x;
                T[3 * 32] = M[z]->m01.y;
                T[4 * 32] = M[z]->m10.x;
                T[5 * 32] = M[z]->m10.y;
                T[6 * 32] = M[z]->m11.x;
                T[7 * 32] = M[z]->m11.y;
                T[8 * 32] = (double)M[z]->e;
            }
        }
    }
    __syncthreads();

    // ---- phase 2: the walks (wavefronts 0 and 1), states into LDS.  The
    // other wavefronts run pass 2 (the v checkpoints) of items 0..npre-1
    // meanwhile, each as soon as the v walk has stored the states of its two
    // segments; the checkpoints stay in their registers for phase 3.
    double2 cx[NS], cy[NS];
    int ce[NS];
    int pre = -1;  // this wavefront's item with its checkpoints already taken
    if (w < 2) {
        if (w == 0)
            fused_walk_split<false, 8>(lp, Tl, sig, tube, ch, cl, h, valid, chain, Sl, Se);
        else
            fused_walk_split<true, 8>(lp, Tl, sig, tube, ch, cl, h, valid, chain, Sl, Se, &v_done);
    } else if (w - 2 < npre) {
        pre = w - 2;
        const int need = min(P, 2 * pre + 2);  // the v walk stores segment P-2-2 pre at its step 2 pre + 1
        for (int it = 0;; ++it) {
            const int d = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&v_done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
            if (d >= need) break;
            if (it >= EQF_WAIT_MAX) {  // never expected: flag it and go on rather than wait without end
                if (lane == 0) atomicOr(lp.status, 2);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        stage_item(pre);
        __builtin_amdgcn_wave_barrier();
        const int sr = max(seg_of(pre, h), 0);
        const double* S = Sl + sr * 8 * 32 + cl;
        double2 x = make_double2(S[128], S[160]), y = make_double2(S[192], S[224]);
        const int e = Se[sr * 32 + cl];
        const double* us = us_w + h * L;
        // opaque copies: otherwise the compiler keeps these coefficients live
        // into phase 3 (which computes the same ones), past the register budget
        double2 Eo = E;
        double teo = te, too = to;
        asm volatile("" : "+v"(Eo.x), "+v"(Eo.y), "+v"(teo), "+v"(too));
        if (scaled) {
            const SCoefs ks = make_scoefs(Eo, teo, too);
            y = cscale(ks.ai[1], y);  // site hi - 1 is odd
            eq_checkpoints<K, NS>(ks, us, x, y, e, cx, cy, ce);
        } else {
            const Coefs k = make_coefs(Eo, teo, too);
            eq_checkpoints<K, NS>(k, us, x, y, e, cx, cy, ce);
        }
    }
    __syncthreads();

    // ---- phase 3: the sweeps.  Work item j is segment P-1-2j in the first
    // half of a wavefront and segment P-2-2j in the second (items from the
    // right end, whose v states the v walk stores first); after their
    // pre-taken items, wavefronts take items from a counter, so the older
    // wavefronts of a SIMD, which the hardware issues first, take more of
    // them and none is left alone long.
    double* red = Tl + w * K * 65;
    const double vmask = valid ? 1.0 : 0.0;
    // pass 3 of item j (its v checkpoints in cx, cy, ce)
    auto finish_item = [&](int j) {
        const int sgm = seg_of(j, h);
        const bool has = sgm >= 0;
        const int sr = has ? sgm : 0;  // an idle half reads a real segment's states (masked below)
        const double* S = Sl + sr * 8 * 32 + cl;
        double2 ux = make_double2(S[0], S[32]), uy = make_double2(S[64], S[96]);
        const double msk = has ? vmask : 0.0;
        const double2 coef = make_double2(ch.cr * msk, ch.ci * msk);
        double* prow = has ? lp.part + tube.part_off + (long long)(2 * g + hf) * N + (long long)sr * L : nullptr;
        const double* us = us_w + h * L;
        if (scaled) {
            const SCoefs ks = make_scoefs(E, te, to);
            uy = cscale(ks.ai[1], uy);  // site lo - 1 is odd
            eq_sweep_pass3<K, NS, SCoefs, 2>(ks, us, ux, uy, cx, cy, ce, coef, E.y, te * te * ks.ie, red, lane,
                                             prow);
        } else {
            const Coefs k = make_coefs(E, te, to);
            eq_sweep_pass3<K, NS, Coefs, 2>(k, us, ux, uy, cx, cy, ce, coef, te, te, red, lane, prow);
        }
        __builtin_amdgcn_wave_barrier();  // us_w and red are reused by the next item
    };
    if (pre >= 0) finish_item(pre);
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(&next_item, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= nwork) break;  // wave-uniform
        stage_item(item);
        __builtin_amdgcn_wave_barrier();
        const int sgm = seg_of(item, h);
        const bool has = sgm >= 0;
        const int sr = has ? sgm : 0;  // an idle half reads a real segment's states (masked below)
        const double* S = Sl + sr * 8 * 32 + cl;
        double2 ux = make_double2(S[0], S[32]), uy = make_double2(S[64], S[96]);
        double2 x = make_double2(S[128], S[160]), y = make_double2(S[192], S[224]);
        const int e = Se[sr * 32 + cl];
        const double msk = has ? vmask : 0.0;
        const double2 coef = make_double2(ch.cr * msk, ch.ci * msk);
        double* prow = has ? lp.part + tube.part_off + (long long)(2 * g + hf) * N + (long long)sr * L : nullptr;
        const double* us = us_w + h * L;
        if (scaled) {
            const SCoefs ks = make_scoefs(E, te, to);
            uy = cscale(ks.ai[1], uy);  // site lo - 1 is odd
            y = cscale(ks.ai[1], y);    // site hi - 1 is odd
            eq_full_body<K, NS, SCoefs, 2>(ks, us, ux, uy, x, y, e, coef, E.y, te * te * ks.ie, red, lane, prow);
        } else {
            const Coefs k = make_coefs(E, te, to);
            eq_full_body<K, NS, Coefs, 2>(k, us, ux, uy, x, y, e, coef, te, te, red, lane, prow);
        }
        __builtin_amdgcn_wave_barrier();  // us_w and red are reused by the next item
    }
}

// Scaled form of the non-equilibrium full-segment sweep (every chain of the
// wave has t / |Im E| <= 2^24).  With u_n = s_n w_n and v_n = s'_n w'_n
// (s_lo = 1, s'_{hi-1} = 1, s_{n+1} = s_n ai_n, s'_{n+1} = s'_n / ai_n) the
// contribution of site n is
//   Q_n |w_n|^2 + P_n |w'_n|^2,
//   Q_n = sgnR rR^2 2^(2(e_w - e_D)) s_n^2,  P_n = sgnL rL^2 2^(2(e_v - e_D)) s'_n^2,
// where e_w is the exponent of the renormalised forward state and e_v that of
// the sub-block's checkpoint.  Q and P are carried site by site (times ai_n^2
// and 1 / ai_n^2) and by exact powers of two where w is renormalised or a new
// checkpoint starts; P starts from s'_lo = C.  Range: with t / |Im E| <= 2^24
// a state grows by at most about 2^27 per site, so |w|^2 within a 16-site
// sub-block stays below 2^864 and Q, P (about contribution / |w|^2) stay
// normal for every contribution above about 2^-150; smaller ones (far below
// the charge's rounding) may flush to zero, as they do in the plain sweep.
template <int K, int NS>
__device__ __forceinline__ void neq_full_scaled(const SCoefs& k, const double* us, double2 ux, double2 uy, int ue,
                                                double2 x, double2 y, int e, int eN, double sgnL, double sgnR,
                                                double rL, double rR, double* red, int lane, double* prow)
{
    uy = cscale(k.ai[1], uy);  // site lo - 1 is odd
    y = cscale(k.ai[1], y);    // site hi - 1 is odd
    double mC;
    int eC;
    scale_prod(k, 0, K * NS - 1, mC, eC);
    double2 cx[NS], cy[NS];
    int ce[NS];
    cx[NS - 1] = x;
    cy[NS - 1] = y;
    ce[NS - 1] = e;
#pragma unroll
    for (int j = NS - 1; j >= 1; --j) {
#pragma unroll
        for (int i = K - 1; i >= 1; i -= 2) {
            bwd_s<1>(k, us[j * K + i], x, y);
            bwd_s<0>(k, us[j * K + i - 1], x, y);
        }
        renorm(x, y, e);
        cx[j - 1] = x;
        cy[j - 1] = y;
        ce[j - 1] = e;
    }
    int ew = ue;
    renorm(ux, uy, ew);
    double Q = sgnR * rR * rR * __builtin_amdgcn_ldexp(1.0, 2 * (ew - eN));
    double P = sgnL * rL * rL * mC * mC * __builtin_amdgcn_ldexp(1.0, 2 * (ce[0] - eN + eC));
    const double a2[2] = {k.ai[0] * k.ai[0], k.ai[1] * k.ai[1]};
    const double ia2[2] = {1.0 / a2[0], 1.0 / a2[1]};
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double2 vr[K];
        {
            double2 xx = cx[j], yy = cy[j];
            vr[K - 1] = xx;
#pragma unroll
            for (int i = K - 1; i >= 1; --i) {
                if (i & 1)
                    bwd_s<1>(k, us[j * K + i], xx, yy);
                else
                    bwd_s<0>(k, us[j * K + i], xx, yy);
                vr[i - 1] = xx;
            }
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            red[i * 65 + lane] = fma(P, cabs2(vr[i]), Q * cabs2(ux));
            if (i & 1)
                fwd_s<1>(k, us[j * K + i], ux, uy);
            else
                fwd_s<0>(k, us[j * K + i], ux, uy);
            Q *= a2[i & 1];
            P *= ia2[i & 1];
        }
        const double r = lane_sum_transpose<K>(red, lane);
        if (lane < K) prow[j * K + lane] = r;
        if (j + 1 < NS) {
            P = __builtin_amdgcn_ldexp(P, 2 * (ce[j + 1] - ce[j]));
            int kk = exp_of(amax4(ux, uy));
            scale2(ux, -kk);
            scale2(uy, -kk);
            Q = __builtin_amdgcn_ldexp(Q, 2 * kk);
        }
    }
}

// Non-equilibrium sweep of a full segment (no integrand samples): the
// contribution of site n is kL Gamma_L |G_n0|^2 + kR Gamma_R |G_n,N-1|^2 with
// G_n0 = v_n / D and G_n,N-1 = u_n / D.  The weights are folded into the
// states: ut = sqrt|kR Gamma_R| 2^(e_u - e_D) u / |D| (a solution of the same
// recurrence, never renormalised: |ut_n| is |G_n,N-1| up to the constant) and,
// per sub-block, vt = sqrt|kL Gamma_L| 2^(e_v - e_D) v / |D|; the signs of the
// two weights multiply the squared moduli.
template <int K, int NS>
__global__ __launch_bounds__(256, NS <= 4 ? 3 : 2) void sweep_neq_full_kernel(WLaunch lp)
{
    __shared__ double red_all[4][K * 65];
    __shared__ double us_all[4][K * NS];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int idx = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    if (idx >= lp.nlist) return;
    const int item = __builtin_amdgcn_readfirstlane(lp.item_list[idx]);
    double* red = red_all[wv];
    WTube tube;
    ItemCtx ic = item_ctx(lp, item, lane, tube);
    if (tube_skipped(lp, tube)) return;  // whole wavefront
    const ChainDesc ch = lp.chains[ic.chain];
    const int N = tube.N;
    const int lo = ic.s * tube.L;
    double* us = us_all[wv];
    stage_segment(us, tube.U, lo, lo + K * NS, lane);
    const double* S = lp.St + (long long)item * 10 * 64 + lane;
    double2 ux = make_double2(S[0], S[64]), uy = make_double2(S[128], S[192]);
    const int ue = (int)S[256];
    double2 x = make_double2(S[320], S[384]), y = make_double2(S[448], S[512]);
    int e = (int)S[576];
    const double* Dp = lp.Dd + (long long)ic.chain * 3;
    const double iD2 = 1.0 / cabs2(make_double2(Dp[0], Dp[1]));
    const int eN = (int)Dp[2];
    const double vmask = ic.valid ? 1.0 : 0.0;
    const double wL = ch.cL * vmask * (-2.0 * ch.SLi) * iD2;  // kL Gamma_L / |Dm|^2
    const double wR = ch.cR * vmask * (-2.0 * ch.SRi) * iD2;
    const double sgnL = wL < 0.0 ? -1.0 : 1.0, sgnR = wR < 0.0 ? -1.0 : 1.0;
    const double rL = sqrt(fabs(wL)), rR = sqrt(fabs(wR));
    double* prow = lp.part + tube.part_off + (long long)ic.g * N + lo;
    const double2 E = make_double2(ch.Er, ch.Ei);
    if (__all(scaled_ok(E, ch.t_even, tube.t_odd))) {
        neq_full_scaled<K, NS>(make_scoefs(E, ch.t_even, tube.t_odd), us, ux, uy, ue, x, y, e, eN, sgnL, sgnR,
                               rL, rR, red, lane, prow);
        return;
    }
    const Coefs k = make_coefs(E, ch.t_even, tube.t_odd);

    double2 cx[NS], cy[NS];
    int ce[NS];
    cx[NS - 1] = x;
    cy[NS - 1] = y;
    ce[NS - 1] = e;
#pragma unroll
    for (int j = NS - 1; j >= 1; --j) {
#pragma unroll
        for (int i = K - 1; i >= 1; i -= 2) {
            bwd_step<1>(k, us[j * K + i], x, y);
            bwd_step<0>(k, us[j * K + i - 1], x, y);
        }
        renorm(x, y, e);
        cx[j - 1] = x;
        cy[j - 1] = y;
        ce[j - 1] = e;
    }
    {
        const double su = rR * __builtin_amdgcn_ldexp(1.0, ue - eN);
        ux = cscale(su, ux);
        uy = cscale(su, uy);
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        double2 vr[K];
        {
            const double sv = rL * __builtin_amdgcn_ldexp(1.0, ce[j] - eN);
            double2 xx = cscale(sv, cx[j]), yy = cscale(sv, cy[j]);
            vr[K - 1] = xx;
#pragma unroll
            for (int i = K - 1; i >= 1; --i) {
                if (i & 1)
                    bwd_step<1>(k, us[j * K + i], xx, yy);
                else
                    bwd_step<0>(k, us[j * K + i], xx, yy);
                vr[i - 1] = xx;
            }
        }
#pragma unroll
        for (int i = 0; i < K; ++i) {
            red[i * 65 + lane] = fma(sgnL, cabs2(vr[i]), sgnR * cabs2(ux));
            if (i & 1)
                fwd_step<1>(k, us[j * K + i], ux, uy);
            else
                fwd_step<0>(k, us[j * K + i], ux, uy);
        }
        const double r = lane_sum_transpose<K>(red, lane);
        if (lane < K) prow[j * K + lane] = r;
    }
}

// One-pass fixed-order sum of the group rows of every tube: a block covers
// 16 sites x 16 row slices (slice k sums rows k, k+16, ...), the slices are
// then added in slice order.
__global__ __launch_bounds__(256) void reduce_rows1_kernel(const double* __restrict__ src,
                                                           const long long* __restrict__ src_off,
                                                           const int* __restrict__ nrows,
                                                           const int* __restrict__ Ns, double* __restrict__ dst,
                                                           const long long* __restrict__ dst_off,
                                                           const WTube* __restrict__ tubes,
                                                           const int* __restrict__ tube_active)
{
    __shared__ double red[16][17];
    const int t = blockIdx.y;
    if (tube_active && !tube_active[tubes[t].ext]) return;  // uniform over the block
    const int N = Ns[t];
    const int sl = threadIdx.x & 15;   // site within the block
    const int k = threadIdx.x >> 4;    // row slice
    const int n = blockIdx.x * 16 + sl;
    const int R = nrows[t];
    double acc = 0.0;
    if (n < N) {
        const double* p = src + src_off[t] + n;
        // unrolled so that the loads are in flight together; the additions
        // keep their order
#pragma unroll 8
        for (int r = k; r < R; r += 16) acc += p[(long long)r * N];
    }
    red[k][sl] = acc;
    __syncthreads();
    if (k == 0 && n < N) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) s += red[q][sl];
        dst[dst_off[t] + n] = s;
    }
}

// The same sums when no tube has more than 16 rows (each slice of
// reduce_rows1_kernel then holds at most one row): one thread per site adds
// its rows in row order from 0.0, which gives the same bits, and a block
// covers 256 sites, so a batch of short row lists is not launched as 16-site
// blocks with idle slices.
__global__ __launch_bounds__(256) void reduce_rows_few_kernel(const double* __restrict__ src,
                                                              const long long* __restrict__ src_off,
                                                              const int* __restrict__ nrows,
                                                              const int* __restrict__ Ns, double* __restrict__ dst,
                                                              const long long* __restrict__ dst_off,
                                                              const WTube* __restrict__ tubes,
                                                              const int* __restrict__ tube_active)
{
    const int t = blockIdx.y;
    if (tube_active && !tube_active[tubes[t].ext]) return;  // uniform over the block
    const int N = Ns[t];
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const int R = nrows[t];
    const double* p = src + src_off[t] + n;
    double v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = r < R ? p[(long long)r * N] : 0.0;  // the loads in flight together
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r)
        if (r < R) acc += v[r];
    dst[dst_off[t] + n] = acc;
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
