// rgf_seg.hip -- segmented, division-free recursive-Green's-function kernels
// for gfx950 (MI355X).
//
// Problem (reference NEGF_Common.cpp:3667-3754, :3377-3497): for each energy
// E of a contour or real-axis path, the retarded Green's function of the
// tridiagonal matrix A = E - H - Sigma_L - Sigma_R (diagonal alpha_n = E - U_n,
// links -t_n with t_n = beta for n even and gamma for n odd) is needed on the
// diagonal (equilibrium charge) and in its first/last columns
// (non-equilibrium charge).  The reference runs a host-serial RGF sweep per
// energy with two complex divisions per site.
//
// Formulation used here (same matrix, same quantities):
//   u_{n+1} = a_n u_n - c_n u_{n-1},  a_n = alpha_n / t_n, c_n = t_{n-1}/t_n
//   u_0 = 1, u_{-1} = Sigma_L / t_{-1}           (left lead as boundary state)
//   v_{n-1} = b_n v_n - d_n v_{n+1},  b_n = alpha_n / t_{n-1}, d_n = 1/c_n
//   v_{N-1} = 1, v_N = Sigma_R / t_{N-1}        (right lead as boundary state)
//   D = t_{N-1} u_N - Sigma_R u_{N-1}
//   G_nn = u_n v_n / D,  G_{n,0} = v_n / D,  G_{n,N-1} = u_n / D.
// u is the left-connected (forward-stable) continuant, v the right-connected
// (backward-stable) one; their ratios are the RGF Y_n and X_n.  There are no
// per-site divisions.
//
// Parallel structure: each chain (one energy of one mode of one tube) is cut
// into 16 segments, one per lane of a 16-lane DPP row; a wavefront carries 4
// chains and a block 8 wavefronts (32 chains).  Per lane:
//   pass 1  2x2 transfer matrix of the segment (forward recurrence on two
//           basis columns), renormalised by exact powers of two;
//   scan    Hillis-Steele prefix and suffix products across the 16 lanes give
//           each segment its exact left state u and right state v;
//   pass 2  backward v over the segment, storing the v state at every K-site
//           sub-block boundary in registers;
//   pass 3  per sub-block: recompute v backward into registers, then run u
//           forward and form G on the fly.
// Charge contributions are accumulated into a per-block LDS row with
// ds_add_f64 and written out as one partial row per block; a second kernel
// sums the partial rows of each tube in a fixed order.
#include "negf_device.h"

namespace negf {

struct St {  // value = (x, y) * 2^e
    double2 x, y;
    int e;
};

__device__ __forceinline__ int exp_of(double m)
{
    // exponent k with m = f * 2^k, f in [0.5, 1); 0 for m == 0
    return m == 0.0 ? 0 : __builtin_amdgcn_frexp_exp(m);
}

__device__ __forceinline__ double amax4(double2 a, double2 b)
{
    return fmax(fmax(fabs(a.x), fabs(a.y)), fmax(fabs(b.x), fabs(b.y)));
}

__device__ __forceinline__ void renorm(double2& a, double2& b, int& e)
{
    int k = exp_of(amax4(a, b));
    a.x = __builtin_amdgcn_ldexp(a.x, -k);
    a.y = __builtin_amdgcn_ldexp(a.y, -k);
    b.x = __builtin_amdgcn_ldexp(b.x, -k);
    b.y = __builtin_amdgcn_ldexp(b.y, -k);
    e += k;
}

__device__ __forceinline__ void renorm_mat(double2& m00, double2& m01, double2& m10, double2& m11, int& e)
{
    int k = exp_of(fmax(amax4(m00, m01), amax4(m10, m11)));
    double2* p[4] = {&m00, &m01, &m10, &m11};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        p[i]->x = __builtin_amdgcn_ldexp(p[i]->x, -k);
        p[i]->y = __builtin_amdgcn_ldexp(p[i]->y, -k);
    }
    e += k;
}

// Lane constants for the recurrences at a site of parity p (p = n & 1).
struct Coefs {
    double erit[2], it[2];    // Re(E)/t_n, 1/t_n
    double eiit[2];           // Im(E)/t_n
    double eritm[2], itm[2];  // Re(E)/t_{n-1}, 1/t_{n-1}
    double eiitm[2];
    double c[2];              // t_{n-1}/t_n
    double d[2];              // t_n/t_{n-1}
};

// forward step: (x, y) = (u_n, u_{n-1}) -> (u_{n+1}, u_n)
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

// backward step: (x, y) = (v_n, v_{n+1}) -> (v_{n-1}, v_n)
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

// pass-1 step on both basis columns of the transfer matrix
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

__device__ __forceinline__ double2 shfl_up2(double2 v, int d)
{
    return make_double2(__shfl_up(v.x, d, NEGF_SEGMENTS), __shfl_up(v.y, d, NEGF_SEGMENTS));
}
__device__ __forceinline__ double2 shfl_down2(double2 v, int d)
{
    return make_double2(__shfl_down(v.x, d, NEGF_SEGMENTS), __shfl_down(v.y, d, NEGF_SEGMENTS));
}
__device__ __forceinline__ double2 shfl_idx2(double2 v, int src)
{
    return make_double2(__shfl(v.x, src, NEGF_SEGMENTS), __shfl(v.y, src, NEGF_SEGMENTS));
}

// 2x2 complex product R = A * B
__device__ __forceinline__ void mat_mul(const double2 a00, const double2 a01, const double2 a10, const double2 a11,
                                        const double2 b00, const double2 b01, const double2 b10, const double2 b11,
                                        double2& r00, double2& r01, double2& r10, double2& r11)
{
    r00 = cadd(cmul(a00, b00), cmul(a01, b10));
    r01 = cadd(cmul(a00, b01), cmul(a01, b11));
    r10 = cadd(cmul(a10, b00), cmul(a11, b10));
    r11 = cadd(cmul(a10, b01), cmul(a11, b11));
}

// Register-resident checkpoint store with a wave-uniform index.
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

struct SiteCtx {
    double2 C, Gs;     // KIND_EQ: contribution Re(C u v); G = Gs u v
    double kL, kR;     // KIND_NONEQ: contribution kL |v|^2 + kR |u|^2
    double aL, aR;     // A_nn pieces: Gamma_L |G_n0|^2 = aL |v|^2, Gamma_R |G_nN|^2 = aR |u|^2
    double iL, iR;     // integrand-sample coefficients
    int nbase, len;
    bool valid, lds;
};

// Pass-3 sites of one sub-block: G from (u, v), contribution, forward u step.
// U points at the sub-block's first site in the transposed table.
template <int K, int KIND, bool WRITE_G, bool FULL, bool SAMPLE>
__device__ __forceinline__ void sub_block(const Coefs& k, const double* U, const double2 (&vr)[K], double2& ux,
                                          double2& uy, const SiteCtx& c, double* Qs, double* part_row,
                                          double2* Grow, const TubeDesc& tube, double* chain_out, int out_slot,
                                          double& dos_acc)
{
#pragma unroll
    for (int i = 0; i < K; ++i) {
        if (FULL || i < c.len) {
            const int n = c.nbase + i;
            const double Un = U[(long long)i * NEGF_SEGMENTS];
            if (KIND == KIND_EQ) {
                double2 g = cmul(ux, vr[i]);
                double q = fma(c.C.x, g.x, -c.C.y * g.y);
                if (c.valid) {
                    if (c.lds)
                        __hip_atomic_fetch_add(&Qs[n], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    else
                        unsafeAtomicAdd(&part_row[n], q);
                    if (WRITE_G) Grow[n] = cmul(c.Gs, g);
                }
            } else if (KIND == KIND_NONEQ) {
                double av = cabs2(vr[i]);
                double au = cabs2(ux);
                double q = fma(c.kL, av, c.kR * au);
                if (c.valid) {
                    if (c.lds)
                        __hip_atomic_fetch_add(&Qs[n], q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    else
                        unsafeAtomicAdd(&part_row[n], q);
                    if (SAMPLE) {
#pragma unroll
                        for (int w = 0; w < 3; ++w)
                            if (n == tube.sample[w]) {
                                double v = c.iL * c.aL * av + c.iR * c.aR * au;
                                chain_out[(long long)out_slot * 3 + w] = v;
                            }
                    }
                }
            } else {
                double av = cabs2(vr[i]);
                double au = cabs2(ux);
                dos_acc += c.aL * av + c.aR * au;
            }
            if (i & 1)
                fwd_step<1>(k, Un, ux, uy);
            else
                fwd_step<0>(k, Un, ux, uy);
        }
    }
}

template <int K, int NS, int KIND, bool WRITE_G>
__global__ __launch_bounds__(NEGF_SEGMENTS* NEGF_CHAINS_PER_WAVE* 8) void rgf_seg_kernel(LaunchParams lp)
{
    extern __shared__ double Qs[];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int row = lane >> 4;
    const int s = lane & 15;
    const int blk = blockIdx.x;
    const int nchain_blk = lp.block_nchain[blk];
    const int chain_local = wave * NEGF_CHAINS_PER_WAVE + row;
    const bool valid = chain_local < nchain_blk;
    const TubeDesc tube = lp.tubes[lp.block_tube[blk]];
    const int N = tube.N;
    const int Lf = tube.seg_len;
    double* part_row = lp.part + lp.block_part[blk];

    if (KIND != KIND_DOS) {
        if (tube.lds_ok) {
            for (int n = tid; n < N; n += blockDim.x) Qs[n] = 0.0;
        }
        __syncthreads();
    }

    const int cidx = lp.block_chain0[blk] + (valid ? chain_local : 0);
    const ChainDesc ch = lp.chains[cidx];
    const double te = ch.t_even, to = tube.t_odd;
    const double2 E = make_double2(ch.Er, ch.Ei);

    const double2 SL = make_double2(ch.SLr, ch.SLi);
    const double2 SR = make_double2(ch.SRr, ch.SRi);

    Coefs k;
    {
        double t[2] = {te, to};  // t_n by parity of n
        for (int p = 0; p < 2; ++p) {
            double tn = t[p], tm = t[1 - p];
            k.it[p] = 1.0 / tn;
            k.erit[p] = E.x / tn;
            k.eiit[p] = E.y / tn;
            k.itm[p] = 1.0 / tm;
            k.eritm[p] = E.x / tm;
            k.eiitm[p] = E.y / tm;
            k.c[p] = tm / tn;
            k.d[p] = tn / tm;
        }
    }
    const double tau = ((N - 1) & 1) ? to : te;  // t_{N-1}
    const int lo = min(N, s * Lf);
    const int hi = min(N, lo + Lf);
    const int nsub = (hi - lo + K - 1) / K;
    const double* Ut = lp.Ut + tube.u_off;
    // site n = s*Lf + j lives at Ut[j*16 + s]
#define UAT(j) Ut[(long long)(j) * NEGF_SEGMENTS + s]

    // ---------------- pass 1: segment transfer matrix ----------------
    double2 m00 = make_double2(1.0, 0.0), m01 = make_double2(0.0, 0.0);
    double2 m10 = make_double2(0.0, 0.0), m11 = make_double2(1.0, 0.0);
    int eT = 0;
    for (int j = 0; j < nsub; ++j) {
        const int p0 = j * K;
        const int len = min(K, hi - lo - p0);
        if (len == K) {
#pragma unroll
            for (int i = 0; i < K; i += 2) {
                mat_step<0>(k, UAT(p0 + i), m00, m01, m10, m11);
                mat_step<1>(k, UAT(p0 + i + 1), m00, m01, m10, m11);
            }
        } else {
            for (int i = 0; i < len; ++i) {
                if (i & 1)
                    mat_step<1>(k, UAT(p0 + i), m00, m01, m10, m11);
                else
                    mat_step<0>(k, UAT(p0 + i), m00, m01, m10, m11);
            }
        }
        renorm_mat(m00, m01, m10, m11, eT);
    }
    // keep T_s for the suffix scan
    const double2 t00 = m00, t01 = m01, t10 = m10, t11 = m11;
    const int tE = eT;

    // ---------------- prefix scan: Pre_s = T_s ... T_0 ----------------
#pragma unroll
    for (int d = 1; d < NEGF_SEGMENTS; d <<= 1) {
        double2 o00 = shfl_up2(m00, d), o01 = shfl_up2(m01, d);
        double2 o10 = shfl_up2(m10, d), o11 = shfl_up2(m11, d);
        int oe = __shfl_up(eT, d, NEGF_SEGMENTS);
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
    // w_s = Pre_s (1, Sigma_L / t_{-1}); t_{-1} is an odd link -> t_odd
    const double2 sigL = cscale(1.0 / to, SL);
    double2 wx = cadd(m00, cmul(m01, sigL));
    double2 wy = cadd(m10, cmul(m11, sigL));
    int we = eT;
    renorm(wx, wy, we);
    // left state of this segment: (u_lo, u_{lo-1})
    St Ul;
    Ul.x = shfl_up2(wx, 1);
    Ul.y = shfl_up2(wy, 1);
    Ul.e = __shfl_up(we, 1, NEGF_SEGMENTS);
    if (s == 0) {
        Ul.x = make_double2(1.0, 0.0);
        Ul.y = sigL;
        Ul.e = 0;
    }
    // (u_N, u_{N-1}) from lane 15
    const double2 uN = shfl_idx2(wx, 15), uN1 = shfl_idx2(wy, 15);
    const int eN = __shfl(we, 15, NEGF_SEGMENTS);
    const double2 Dm = csub(cscale(tau, uN), cmul(SR, uN1));  // D = Dm * 2^eN
    const double2 iD = crecip(Dm);
    const double iD2 = cabs2(iD);

    // ---------------- suffix scan: Suf_s = T_15 ... T_s ----------------
    m00 = t00;
    m01 = t01;
    m10 = t10;
    m11 = t11;
    eT = tE;
#pragma unroll
    for (int d = 1; d < NEGF_SEGMENTS; d <<= 1) {
        double2 o00 = shfl_down2(m00, d), o01 = shfl_down2(m01, d);
        double2 o10 = shfl_down2(m10, d), o11 = shfl_down2(m11, d);
        int oe = __shfl_down(eT, d, NEGF_SEGMENTS);
        if (s + d < NEGF_SEGMENTS) {
            double2 r00, r01, r10, r11;
            mat_mul(o00, o01, o10, o11, m00, m01, m10, m11, r00, r01, r10, r11);
            m00 = r00;
            m01 = r01;
            m10 = r10;
            m11 = r11;
            eT += oe;
        }
    }
    // z_s = Suf_s^{-1} (v_N, v_{N-1}) with (v_N, v_{N-1}) = (Sigma_R/tau, 1);
    // Suf^{-1} = adj(M) 2^eT / det(T-product), det = t_{lo-1}/t_{N-1}
    const double2 sigR = cscale(1.0 / tau, SR);
    double2 zx, zy;
    int ze;
    {
        double detT = (lo < N) ? (to / tau) : 1.0;
        double id = 1.0 / detT;
        // adj = [[m11, -m01], [-m10, m00]]
        zx = cscale(id, csub(cmul(m11, sigR), m01));
        zy = cscale(id, csub(m00, cmul(m10, sigR)));
        ze = eT;
        renorm(zx, zy, ze);
    }
    // right state of this segment in backward form: (v_{hi-1}, v_hi)
    St Vr;
    {
        double2 ax = shfl_down2(zx, 1), ay = shfl_down2(zy, 1);
        int ae = __shfl_down(ze, 1, NEGF_SEGMENTS);
        if (s == NEGF_SEGMENTS - 1 || hi >= N) {
            ax = sigR;
            ay = make_double2(1.0, 0.0);
            ae = 0;
        }
        // (ax, ay) = (v_hi, v_{hi-1})
        Vr.x = ay;
        Vr.y = ax;
        Vr.e = ae;
    }

    // ---------------- pass 2: v checkpoints ----------------
    Ckpt<NS> ck;
    {
        double2 x = Vr.x, y = Vr.y;
        int e = Vr.e;
        ck.put(nsub - 1, x, y, e);
        for (int j = nsub - 1; j >= 1; --j) {
            const int p0 = j * K;  // sub-block j covers [lo+p0, lo+p0+len)
            const int len = min(K, hi - lo - p0);
            if (len == K) {
#pragma unroll
                for (int i = K - 1; i >= 0; i -= 2) {
                    bwd_step<1>(k, UAT(p0 + i), x, y);
                    bwd_step<0>(k, UAT(p0 + i - 1), x, y);
                }
            } else {
                for (int i = len - 1; i >= 0; --i) {
                    if (i & 1)
                        bwd_step<1>(k, UAT(p0 + i), x, y);
                    else
                        bwd_step<0>(k, UAT(p0 + i), x, y);
                }
            }
            // state = (v_{lo+p0-1}, v_{lo+p0}): end state of sub-block j-1
            renorm(x, y, e);
            ck.put(j - 1, x, y, e);
        }
    }

    // ---------------- pass 3: G on the fly ----------------
    double2 ux = Ul.x, uy = Ul.y;
    int ue = Ul.e;
    const double2 coef = make_double2(ch.cr, ch.ci);
    const double2 GamLR = make_double2(-2.0 * SL.y, -2.0 * SR.y);  // Gamma = i(S - S^+)
    double dos_acc = 0.0;
    const bool lds = tube.lds_ok;
    double2* Grow = WRITE_G ? (lp.diagG + (long long)cidx * N) : nullptr;
    for (int j = 0; j < nsub; ++j) {
        const int p0 = j * K;
        const int len = min(K, hi - lo - p0);
        double2 cx, cy;
        int ce;
        ck.get(j, cx, cy, ce);
        // v for the sub-block: vr[i] = v_{lo+p0+i}
        double2 vr[K];
        {
            double2 x = cx, y = cy;  // (v_{q-1}, v_q)
            if (len == K) {
                vr[K - 1] = x;
#pragma unroll
                for (int i = K - 1; i >= 1; i -= 2) {
                    bwd_step<1>(k, UAT(p0 + i), x, y);
                    vr[i - 1] = x;
                    if (i - 1 >= 1) {
                        bwd_step<0>(k, UAT(p0 + i - 1), x, y);
                        vr[i - 2] = x;
                    }
                }
            } else {
#pragma unroll
                for (int i = K - 1; i >= 0; --i) {
                    if (i == len - 1) vr[i] = x;
                    if (i < len - 1) {
                        if ((i + 1) & 1)
                            bwd_step<1>(k, UAT(p0 + i + 1), x, y);
                        else
                            bwd_step<0>(k, UAT(p0 + i + 1), x, y);
                        vr[i] = x;
                    }
                }
            }
        }
        renorm(ux, uy, ue);
        // scale factors
        const double sc_uv = __builtin_amdgcn_ldexp(1.0, ue + ce - eN);
        double2 C = cscale(sc_uv, cmul(coef, iD));
        double2 Gs = cscale(sc_uv, iD);
        const double sv = __builtin_amdgcn_ldexp(1.0, ce - eN);
        const double su = __builtin_amdgcn_ldexp(1.0, ue - eN);
        const double kL = ch.cL * GamLR.x * sv * sv * iD2;
        const double kR = ch.cR * GamLR.y * su * su * iD2;
        const int nbase = lo + p0;
        const bool full = (len == K);
        bool samp = false;
        if (KIND == KIND_NONEQ) {
            for (int q = 0; q < 3; ++q)
                samp |= (tube.sample[q] >= nbase && tube.sample[q] < nbase + len);
        }
        SiteCtx sc{C, Gs, kL, kR, GamLR.x * sv * sv * iD2, GamLR.y * su * su * iD2, ch.iL, ch.iR,
                   nbase, len, valid, lds};
        if (full && !__any(samp)) {
            sub_block<K, KIND, WRITE_G, true, false>(k, Ut + (long long)p0 * NEGF_SEGMENTS + s, vr, ux, uy, sc,
                                                     Qs, part_row, Grow, tube, lp.chain_out, ch.out_slot,
                                                     dos_acc);
        } else {
            sub_block<K, KIND, WRITE_G, false, true>(k, Ut + (long long)p0 * NEGF_SEGMENTS + s, vr, ux, uy, sc,
                                                     Qs, part_row, Grow, tube, lp.chain_out, ch.out_slot,
                                                     dos_acc);
        }
    }
#undef UAT

    if (KIND == KIND_DOS) {
        // row reduction of the per-lane DOS trace (fixed order)
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1) dos_acc += __shfl_xor(dos_acc, d, NEGF_SEGMENTS);
        if (valid && s == 0) {
            // T = Gamma_L Gamma_R |G_{0,N-1}|^2, G_{0,N-1} = u_0 / D = 1/D
            double gN = __builtin_amdgcn_ldexp(1.0, -eN);
            double T = GamLR.x * GamLR.y * iD2 * gN * gN;
            lp.chain_out[(long long)ch.out_slot * 2 + 0] = dos_acc;
            lp.chain_out[(long long)ch.out_slot * 2 + 1] = T;
        }
        return;
    }
    if (lds) {
        __syncthreads();
        for (int n = tid; n < N; n += blockDim.x) part_row[n] = Qs[n];
    }
}

// Lead self-energies of every chain, Sigma = tau g tau^+ with tau = t_odd
// (get_Sigma_at_contacts NEGF_Common.cpp:4165-4177, CNT.cpp:348-351); one
// thread per chain of a block.
__global__ __launch_bounds__(64) void sigma_kernel(LaunchParams lp)
{
    const int blk = blockIdx.x;
    const int c = threadIdx.x;
    if (c >= lp.block_nchain[blk]) return;
    const TubeDesc tube = lp.tubes[lp.block_tube[blk]];
    ChainDesc* ch = lp.chains + lp.block_chain0[blk] + c;
    const double2 E = make_double2(ch->Er, ch->Ei);
    const double te = ch->t_even, to = tube.t_odd;
    double2 gL, gR;
    int okL = 1, okR = 1;
    if (lp.use_decimation) {
        gL = decimation_gf(make_double2(E.x - tube.Uc0, E.y), te, to, lp.dec_layers, lp.dec_max_iter,
                           lp.dec_rel_error);
        gR = decimation_gf(make_double2(E.x - tube.Uc1, E.y), te, to, lp.dec_layers, lp.dec_max_iter,
                           lp.dec_rel_error);
    } else {
        gL = surface_gf(E, tube.Uc0, te, to, &okL);
        gR = surface_gf(E, tube.Uc1, te, to, &okR);
    }
    if (!okL || !okR) atomicOr(lp.status, 1);
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

// Transposed potential table: Ut[j*16 + s] = U[s*Lf + j]
__global__ void transpose_u_kernel(const double* __restrict__ U, int N, int Lf, double* __restrict__ Ut)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = Lf * NEGF_SEGMENTS;
    if (idx >= total) return;
    const int j = idx / NEGF_SEGMENTS, s = idx % NEGF_SEGMENTS;
    const int n = s * Lf + j;
    Ut[idx] = (n < N) ? U[n] : 0.0;
}

__global__ void fill_kernel(double* p, long long n, double v)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace negf
