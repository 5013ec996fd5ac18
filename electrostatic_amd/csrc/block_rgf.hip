// block_rgf.hip -- dense-block NEGF path on FP64 MFMA: Sancho-Rubio lead
// surface Green's functions and the block recursive Green's function sweep
// (diag G^R) for chains of 32 x 32 blocks.
//
// The reference declares this path but leaves it stubbed (2-D MatrixBlock
// algebra in Matrix_Block_Util.H:123-135, 431-438, 684-692; decimation
// DecimationTechnique NEGF_Common.cpp:4095-4153 only for 1-D blocks); the
// mathematics follows the scalar sweep of Compute_RhoEq
// (NEGF_Common.cpp:3667-3754) with blocks in place of scalars.
//
// Material: every block has the same real symmetric on-site matrix H0 plus a
// per-orbital real potential v_b, and every block couples to the next by the
// same real matrix T = H_{b,b+1}; the leads continue the chain with the
// contact potential added to H0.  All left-connected and full block Green's
// functions are then complex symmetric, which the backward step uses:
//   G_bb = g_b + X G_{b+1,b+1} X^T,   X = g_b T.
//
// Layout.  Products run on v_mfma_f64_16x16x4f64 (lane l: A[i][k] with
// i = l % 16, k = l / 16; B[k][j] with j = l % 16; result register v holds
// D[l / 16 + 4 v][l % 16], as a GPU probe against numpy established for this
// FP64 instruction).  Complex matrices are two real planes in LDS (row stride
// 33 doubles).  The block RGF sweep gives each energy one wavefront, which
// runs all four 16 x 16 quadrants of every product and keeps the matrix of the
// pivoted Gauss-Jordan inverse in its registers; the Sancho-Rubio kernel gives
// each (energy, lead) a workgroup of four wavefronts, one quadrant each, with
// the inverse spread over the workgroup.  The forward sweep stores each g_b
// (16 KiB) to a device scratch array that the backward sweep reads back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "negf_internal.h"

namespace negf {

typedef __attribute__((ext_vector_type(4))) double d4;

constexpr int BM = 32;               // orbitals per block
constexpr int BLD = 33;              // LDS row stride (doubles)
constexpr int BPL = BM * BLD;        // doubles per LDS plane
constexpr int BTHREADS = 256;        // 4 wavefronts (Sancho-Rubio: one quadrant each)
constexpr int BGL = BM * BM;         // doubles per plane in global memory (row-major, unpadded)

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

struct Quad {
    int lane, qi, qj;
    __device__ Quad()
    {
        lane = threadIdx.x & 63;
        const int w = threadIdx.x >> 6;
        qi = w >> 1;
        qj = w & 1;
    }
    __device__ int row(int v) const { return qi * 16 + (lane >> 4) + 4 * v; }
    __device__ int col() const { return qj * 16 + (lane & 15); }
};

// (cre, cim) += A B over this wavefront's quadrant; AC/BC: complex operand;
// TA: A(i,k) is read at A[k][i] (A stored transposed); TB: B(k,j) at B[j][k];
// LA, LB: row strides (BLD for LDS planes, BM for a real matrix read from
// global memory, which stays in the L1/L2 caches).
template <bool AC, bool BC, bool TA, bool TB, int LA = BLD, int LB = BLD>
__device__ __forceinline__ void mm(d4& cre, d4& cim, const double* Are, const double* Aim, const double* Bre,
                                   const double* Bim, const Quad& q)
{
    const int i = q.qi * 16 + (q.lane & 15), j = q.qj * 16 + (q.lane & 15), kk = q.lane >> 4;
#pragma unroll 2
    for (int k0 = 0; k0 < BM; k0 += 4) {
        const int k = k0 + kk;
        const int ia = TA ? k * LA + i : i * LA + k;
        const int ib = TB ? j * LB + k : k * LB + j;
        const double ar = Are[ia], br = Bre[ib];
        if (AC && BC) {
            const double ai = Aim[ia], bi = Bim[ib];
            cre = mfma64(ar, br, cre);
            cre = mfma64(-ai, bi, cre);
            cim = mfma64(ar, bi, cim);
            cim = mfma64(ai, br, cim);
        } else if (AC) {
            const double ai = Aim[ia];
            cre = mfma64(ar, br, cre);
            cim = mfma64(ai, br, cim);
        } else {
            const double bi = Bim[ib];
            cre = mfma64(ar, br, cre);
            cim = mfma64(ar, bi, cim);
        }
    }
}

__device__ __forceinline__ void store_quad(double* Mre, double* Mim, const d4& cre, const d4& cim, const Quad& q)
{
    const int j = q.col();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        Mre[q.row(v) * BLD + j] = cre[v];
        Mim[q.row(v) * BLD + j] = cim[v];
    }
}

__device__ __forceinline__ double mag2(double re, double im) { return re * re + im * im; }

// max that keeps a NaN (so that a NaN never reads as converged)
__device__ __forceinline__ double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }

// Pivot candidate key: |a|^2 (non-negative, so its bits order as an
// unsigned integer) with the low 5 mantissa bits replaced by 31 - row, so
// that the largest key is the largest magnitude and, among equal ones, the
// smallest row.  0 marks "no candidate".
__device__ __forceinline__ unsigned long long pivot_key(double m, int i)
{
    return ((unsigned long long)__double_as_longlong(m) & ~31ull) | (unsigned long long)(31 - i);
}

// 64-bit value of another lane by a DPP move (CTRL: a compile-time DPP
// control word; all lanes active).
template <int CTRL>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v & 0xffffffffull), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

// max over each aligned group of 8 lanes: quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], then row_half_mirror (lane i <- lane 7 - i within 8 lanes)
__device__ __forceinline__ unsigned long long max8(unsigned long long v)
{
    unsigned long long o = dpp64<0xB1>(v);
    v = o > v ? o : v;
    o = dpp64<0x4E>(v);
    v = o > v ? o : v;
    o = dpp64<0x141>(v);
    return o > v ? o : v;
}

// Inverse of the complex matrix (Mre, Mim), in place, by Gauss-Jordan with
// partial pivoting.  The matrix lives in registers during the elimination:
// thread t owns column j = t / 8 and rows t % 8 + 8 r (r = 0..3), so the 8
// owners of a column are 8 lanes of one wavefront.  Rows are never swapped:
// step k takes as pivot the largest |a_ik| among the rows not yet used (ties
// to the smallest row), and the result is read back through the pivot order,
// A^-1[i][j] = M[perm[i]][q(j)] with perm[q(j)] = j.  Per step the 8 owners
// of column k publish it (and the pivot row) through LDS behind one barrier;
// every thread takes the pivot row's entry of its column from the owning
// lane of its own wavefront by a shuffle.  The owner of the pivot entry
// publishes 1 / pivot with the pivot row, so no thread divides after the
// barrier.  (Wre, Wim) provide the column buffers (2 x 32 entries each) and
// the reciprocal pivots (Wre[64..67]).
__device__ void gj_inverse(double* Mre, double* Mim, double* Wre, double* Wim, int* perm)
{
    const int t = threadIdx.x, j = t >> 3, c = t & 7, lane = t & 63;
    double ar[4], ai[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        ar[r] = Mre[(c + 8 * r) * BLD + j];
        ai[r] = Mim[(c + 8 * r) * BLD + j];
    }
    unsigned used = 0u;  // bit i: row i has been a pivot (the same in every thread)
    int q = 0;           // the step whose pivot row is j
    for (int k = 0; k < BM; ++k) {
        double* cre = Wre + 32 * (k & 1);
        double* cim = Wim + 32 * (k & 1);
        // pivot candidates of the owners of column k (every lane runs the
        // shuffles; xor offsets 1, 2, 4 stay inside each 8-lane column group)
        unsigned long long kb = 0ull;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = c + 8 * r;
            const unsigned long long kk = ((used >> i) & 1u) ? 0ull : pivot_key(mag2(ar[r], ai[r]), i);
            kb = kk > kb ? kk : kb;
        }
        kb = max8(kb);
        double* pinv = Wre + 64 + 2 * (k & 1);
        if (j == k) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                cre[c + 8 * r] = ar[r];
                cim[c + 8 * r] = ai[r];
            }
            const int pk = kb == 0ull ? __builtin_ctz(~used) : 31 - (int)(kb & 31ull);
            if (c == 0) perm[k] = pk;
            if (c == (pk & 7)) {
                const int rk = pk >> 3;
                const double a = rk == 0 ? ar[0] : rk == 1 ? ar[1] : rk == 2 ? ar[2] : ar[3];
                const double b = rk == 0 ? ai[0] : rk == 1 ? ai[1] : rk == 2 ? ai[2] : ai[3];
                const double inv = 1.0 / mag2(a, b);
                pinv[0] = a * inv;
                pinv[1] = -b * inv;
            }
        }
        __syncthreads();
        const int p = __builtin_amdgcn_readfirstlane(perm[k]);
        used |= 1u << p;
        q = (p == j) ? k : q;
        const double pr = pinv[0], pi = pinv[1];  // 1 / pivot
        // the pivot row's entry of column j: lane (column group) + p % 8, register p / 8
        const int rp = p >> 3;
        const double sr = rp == 0 ? ar[0] : rp == 1 ? ar[1] : rp == 2 ? ar[2] : ar[3];
        const double si = rp == 0 ? ai[0] : rp == 1 ? ai[1] : rp == 2 ? ai[2] : ai[3];
        const int src = (lane & ~7) | (p & 7);
        const double rr = __shfl(sr, src, 64), ri = __shfl(si, src, 64);
        const bool jk = (j == k);
        const double nr = jk ? pr : rr * pr - ri * pi;  // new pivot row, column j
        const double ni = jk ? pi : rr * pi + ri * pr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = c + 8 * r;
            const double fr = cre[i], fi = cim[i];
            const double a0 = jk ? 0.0 : ar[r], b0 = jk ? 0.0 : ai[r];
            const double ur = a0 - (fr * nr - fi * ni);
            const double ui = b0 - (fr * ni + fi * nr);
            ar[r] = (i == p) ? nr : ur;
            ai[r] = (i == p) ? ni : ui;
        }
    }
    // read back through the pivot order
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Mre[(c + 8 * r) * BLD + j] = ar[r];
        Mim[(c + 8 * r) * BLD + j] = ai[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int src = perm[c + 8 * r] * BLD + q;
        ar[r] = Mre[src];
        ai[r] = Mim[src];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        Mre[(c + 8 * r) * BLD + j] = ar[r];
        Mim[(c + 8 * r) * BLD + j] = ai[r];
    }
    __syncthreads();
}

__device__ __forceinline__ void load_real(double* dst, const double* __restrict__ src)
{
    for (int e = threadIdx.x; e < BM * BM; e += BTHREADS) dst[(e >> 5) * BLD + (e & 31)] = src[e];
}

__device__ __forceinline__ void load_cplx(double* dre, double* dim, const double* __restrict__ src)
{
    for (int e = threadIdx.x; e < BM * BM; e += BTHREADS) {
        dre[(e >> 5) * BLD + (e & 31)] = src[e];
        dim[(e >> 5) * BLD + (e & 31)] = src[BGL + e];
    }
}

__device__ __forceinline__ void store_cplx(double* __restrict__ dst, const double* sre, const double* sim)
{
    for (int e = threadIdx.x; e < BM * BM; e += BTHREADS) {
        dst[e] = sre[(e >> 5) * BLD + (e & 31)];
        dst[BGL + e] = sim[(e >> 5) * BLD + (e & 31)];
    }
}

// ---------------------------------------------------------------- one
// wavefront per energy.  The block RGF sweep gives each energy a wavefront
// of its own: the whole 32 x 32 product runs on that wavefront's MFMA
// (four 16 x 16 quadrants, index qi * 2 + qj; result register v of lane l
// holds row qi * 16 + l / 16 + 4 v, column qj * 16 + l % 16), and the
// inverse keeps the matrix in its registers.  Lanes of one wavefront exchange
// data through LDS in program order, so no workgroup barrier is needed.
constexpr int WPB = 4;  // energies (wavefronts) per workgroup
static_assert(64 * WPB == BTHREADS, "the workgroup-wide loads stride by BTHREADS");

struct Quad4 {
    d4 re[4], im[4];
};

__device__ __forceinline__ void zero(Quad4& c)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        c.re[q] = d4{0, 0, 0, 0};
        c.im[q] = d4{0, 0, 0, 0};
    }
}

// LDS written by lanes of this wavefront is visible to its later reads.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// c += A B over all four quadrants (same operand conventions as mm); SYM:
// the product is known to be symmetric and quadrant (1, 0) is skipped.
template <bool AC, bool BC, bool TA, bool TB, int LA = BLD, int LB = BLD, bool SYM = false>
__device__ __forceinline__ void wmm(Quad4& c, const double* Are, const double* Aim, const double* Bre,
                                    const double* Bim, int lane)
{
    const int il = lane & 15, kk = lane >> 4;
#pragma unroll 2
    for (int k0 = 0; k0 < BM; k0 += 4) {
        const int k = k0 + kk;
        double ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = h * 16 + il;
            const int ia = TA ? k * LA + i : i * LA + k;
            const int ib = TB ? i * LB + k : k * LB + i;
            ar[h] = Are[ia];
            ai[h] = AC ? Aim[ia] : 0.0;
            br[h] = Bre[ib];
            bi[h] = BC ? Bim[ib] : 0.0;
        }
#pragma unroll
        for (int qi = 0; qi < 2; ++qi)
#pragma unroll
            for (int qj = 0; qj < 2; ++qj) {
                const int q = qi * 2 + qj;
                if (SYM && q == 2) continue;
                if (AC && BC) {
                    c.re[q] = mfma64(ar[qi], br[qj], c.re[q]);
                    c.re[q] = mfma64(-ai[qi], bi[qj], c.re[q]);
                    c.im[q] = mfma64(ar[qi], bi[qj], c.im[q]);
                    c.im[q] = mfma64(ai[qi], br[qj], c.im[q]);
                } else if (AC) {
                    c.re[q] = mfma64(ar[qi], br[qj], c.re[q]);
                    c.im[q] = mfma64(ai[qi], br[qj], c.im[q]);
                } else {
                    c.re[q] = mfma64(ar[qi], br[qj], c.re[q]);
                    c.im[q] = mfma64(ar[qi], bi[qj], c.im[q]);
                }
            }
    }
}

__device__ __forceinline__ int wrow(int lane, int qi, int v) { return qi * 16 + (lane >> 4) + 4 * v; }
__device__ __forceinline__ int wcol(int lane, int qj) { return qj * 16 + (lane & 15); }

__device__ __forceinline__ void store_quad4(double* Mre, double* Mim, const Quad4& c, int lane)
{
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = wrow(lane, q >> 1, v) * BLD + wcol(lane, q & 1);
            Mre[o] = c.re[q][v];
            Mim[o] = c.im[q][v];
        }
}

// store of a symmetric matrix of which quadrants 0, 1 and 3 are held:
// quadrant 1 is written to both of its mirror positions.
__device__ __forceinline__ void store_quad4_sym(double* Mre, double* Mim, const Quad4& c, int lane)
{
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q == 2) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = wrow(lane, q >> 1, v), j = wcol(lane, q & 1);
            Mre[i * BLD + j] = c.re[q][v];
            Mim[i * BLD + j] = c.im[q][v];
            if (q == 1) {
                Mre[j * BLD + i] = c.re[q][v];
                Mim[j * BLD + i] = c.im[q][v];
            }
        }
    }
}

// 32-bit pivot key: |a|^2 rounded to float (non-negative, so its bits order
// as an unsigned integer) with the low 5 mantissa bits replaced by 31 - row:
// the largest key is a largest magnitude (to float precision, enough for a
// pivot choice) and, among equal ones, the smallest row.  0: no candidate.
__device__ __forceinline__ unsigned pivot_key32(double m, int i)
{
    return (__float_as_uint((float)m) & ~31u) | (unsigned)(31 - i);
}

template <int CTRL>
__device__ __forceinline__ unsigned max_dpp(unsigned v)
{
    const unsigned o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
    return o > v ? o : v;
}

// max over the 64 lanes, read from lane 63: quad_perm [1,0,3,2], quad_perm
// [2,3,0,1], row_half_mirror and row_mirror give every lane its row's
// maximum; row_bcast:15 and row_bcast:31 carry the rows into lane 63.
__device__ __forceinline__ unsigned wave_max32(unsigned v)
{
    v = max_dpp<0xB1>(v);
    v = max_dpp<0x4E>(v);
    v = max_dpp<0x141>(v);
    v = max_dpp<0x140>(v);
    v = max_dpp<0x142>(v);
    v = max_dpp<0x143>(v);
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ double dpp_swap1(double x)
{
    return __longlong_as_double((long long)dpp64<0xB1>((unsigned long long)__double_as_longlong(x)));
}

__device__ __forceinline__ double readlane_f64(double x, int src)
{
    const long long b = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)(b & 0xffffffffll), src);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)((unsigned long long)b >> 32), src);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// 1 / d for d > 0: hardware reciprocal estimate refined by two Newton-Raphson
// steps y <- y + y (1 - d y), each of which squares the relative error.
__device__ __forceinline__ double recip_nr(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    y = fma(y, fma(-d, y, 1.0), y);
    return fma(y, fma(-d, y, 1.0), y);
}

// Inverse of the complex matrix (Mre, Mim) (LDS planes, row stride BLD) by
// one wavefront, in place: Gauss-Jordan with partial pivoting.  Lane l holds
// row i = l / 2, columns h * 16 .. h * 16 + 15 (h = l % 2) in registers.
// Rows are never swapped: step k takes as pivot the largest |a_ik| among the
// rows not yet used (|a_ik|^2 compared in float precision, ties to the
// smallest row), found by a 64-lane maximum.
// A pivot row is not divided by its pivot: each row keeps a scale s_i (1 until
// it is the pivot, then 1 / pivot) and holds M / s_i, which leaves every
// update in the form a_ij -= (a_ik / pivot) a_pj with the stored values; the
// scales are applied once at the end.  The pivot element and the pivot row's
// entries in column k + 1 are read from their lanes (v_readlane), so the next
// pivot search waits on no memory; every lane fetches the pivot row's other
// entries of its columns from the pivot row's lane of the same half by lane
// shuffles (ds_bpermute), so no LDS store, wait and reload sits in a step.
// perm (32 ints of LDS) records the pivot order.  The result is written back
// through the pivot order: A^-1[q(i)][perm[j]] = s_i M[i][j], with q(i) the
// step whose pivot was row i.
__device__ __forceinline__ void gj_inverse_wave(double* Mre, double* Mim, int* perm, int lane)
{
    const int i = lane >> 1, h = lane & 1;
    double ar[16], ai[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        ar[r] = Mre[i * BLD + h * 16 + r];
        ai[r] = Mim[i * BLD + h * 16 + r];
    }
    unsigned used = 0u;  // bit x: row x has been a pivot (wave-uniform)
    int qrow = 0;
    double sr = 1.0, si = 0.0;
    for (int kh = 0; kh < 2; ++kh) {
#pragma unroll
        for (int kr = 0; kr < 16; ++kr) {
            const int k = kh * 16 + kr;
            // pivot: column k is register kr of the lanes with h == kh
            unsigned key = 0u;
            if (h == kh && !((used >> i) & 1u)) key = pivot_key32(mag2(ar[kr], ai[kr]), i);
            key = wave_max32(key);
            const int p = key == 0u ? __builtin_ctz(~used) : 31 - (int)(key & 31u);
            used |= 1u << p;
            if (i == p) {
                if (h == 0) perm[k] = p;
                qrow = k;
            }
            // 1 / pivot from the lane holding (p, k)
            const double pr = readlane_f64(ar[kr], 2 * p + kh), pim = readlane_f64(ai[kr], 2 * p + kh);
            const double iv = recip_nr(pr * pr + pim * pim);
            const double ir = pr * iv, ii = -pim * iv;
            // multiplier of this row: a_ik / pivot (a_ik from the lane of this row holding column k)
            const double orr = dpp_swap1(ar[kr]), oi = dpp_swap1(ai[kr]);
            const double fr = (h == kh) ? ar[kr] : orr, fi = (h == kh) ? ai[kr] : oi;
            double mr = fr * ir - fi * ii, mi = fr * ii + fi * ir;
            if (i == p) {
                mr = 0.0;
                mi = 0.0;
                sr = ir;
                si = ii;
            }
            // register c (column k + 1 in the lanes of its half) first, with the
            // pivot row's entries read from its two lanes: the next pivot search
            // then waits on no shuffle
            const int c = (kr + 1) & 15;
            {
                const double x0r = readlane_f64(ar[c], 2 * p), x0i = readlane_f64(ai[c], 2 * p);
                const double x1r = readlane_f64(ar[c], 2 * p + 1), x1i = readlane_f64(ai[c], 2 * p + 1);
                const double xr = h ? x1r : x0r, xi = h ? x1i : x0i;
                ar[c] = fma(-mr, xr, fma(mi, xi, ar[c]));
                ai[c] = fma(-mr, xi, fma(-mi, xr, ai[c]));
            }
            // the other entries of the pivot row, from the lane of its same half
            // (the pivot row itself is not updated: m = 0)
            const int src = 2 * p + h;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                if (r == c) continue;
                const double2 x = make_double2(__shfl(ar[r], src, 64), __shfl(ai[r], src, 64));
                ar[r] = fma(-mr, x.x, fma(mi, x.y, ar[r]));
                ai[r] = fma(-mr, x.y, fma(-mi, x.x, ai[r]));
            }
            if (h == kh) {
                ar[kr] = (i == p) ? 1.0 : -mr;
                ai[kr] = (i == p) ? 0.0 : -mi;
            }
        }
    }
    wave_sync();
    int pj[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) pj[r] = perm[h * 16 + r];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int o = qrow * BLD + pj[r];
        Mre[o] = sr * ar[r] - si * ai[r];
        Mim[o] = sr * ai[r] + si * ar[r];
    }
    wave_sync();
}

// Sancho-Rubio decimation for one energy (blockIdx.x) and one lead
// (blockIdx.y: 0 = left lead, surface = its last block, alpha = T^T;
// 1 = right lead, surface = its first block, alpha = T), on-site H0 + Uc.
// Iterates until max |alpha|, |beta| < tol (at most max_iter times); writes
// the self-energy Sigma_L = T^T g T or Sigma_R = T g T^T (global complex
// planes) and the iteration count (negative if not converged).
__global__ __launch_bounds__(BTHREADS) void sancho_rubio_kernel(int nE, const double2* __restrict__ E,
                                                                const double* __restrict__ H0g,
                                                                const double* __restrict__ Tg, double Uc0, double Uc1,
                                                                double tol, int max_iter, double* __restrict__ Sigma,
                                                                int* __restrict__ iters)
{
    __shared__ double Tl[BPL], eps_s[2][BPL], eps[2][BPL], a[2][BPL], b[2][BPL], G[2][BPL], t1[2][BPL];
    __shared__ int perm[BM];
    __shared__ double wmax[4];
    const int e = blockIdx.x, side = blockIdx.y;
    if (e >= nE) return;
    const double2 z = E[e];
    const double Er = z.x - (side == 0 ? Uc0 : Uc1), Ei = z.y;
    const int tid = threadIdx.x;
    const Quad q;
    load_real(Tl, Tg);
    for (int e2 = tid; e2 < BM * BM; e2 += BTHREADS) {
        const int i = e2 >> 5, j = e2 & 31;
        const double h = H0g[e2];
        eps_s[0][i * BLD + j] = h;
        eps_s[1][i * BLD + j] = 0.0;
        eps[0][i * BLD + j] = h;
        eps[1][i * BLD + j] = 0.0;
        // alpha = H_{0,1} of the semi-infinite direction: T^T (left lead), T (right lead)
        const double t = Tg[e2], tt = Tg[j * BM + i];
        a[0][i * BLD + j] = side == 0 ? tt : t;
        a[1][i * BLD + j] = 0.0;
        b[0][i * BLD + j] = side == 0 ? t : tt;
        b[1][i * BLD + j] = 0.0;
    }
    __syncthreads();
    int it = 0;
    bool conv = false;
    while (it < max_iter) {
        ++it;
        for (int e2 = tid; e2 < BM * BM; e2 += BTHREADS) {
            const int i = e2 >> 5, j = e2 & 31;
            G[0][i * BLD + j] = (i == j ? Er : 0.0) - eps[0][i * BLD + j];
            G[1][i * BLD + j] = (i == j ? Ei : 0.0) - eps[1][i * BLD + j];
        }
        __syncthreads();
        gj_inverse(G[0], G[1], t1[0], t1[1], perm);
        // ga = G a, gb = G b
        d4 r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0}, r2 = {0, 0, 0, 0}, i2 = {0, 0, 0, 0};
        mm<true, true, false, false>(r1, i1, G[0], G[1], a[0], a[1], q);
        mm<true, true, false, false>(r2, i2, G[0], G[1], b[0], b[1], q);
        __syncthreads();
        store_quad(G[0], G[1], r1, i1, q);    // G <- ga
        store_quad(t1[0], t1[1], r2, i2, q);  // t1 <- gb
        __syncthreads();
        d4 agb_r = {0, 0, 0, 0}, agb_i = {0, 0, 0, 0}, bga_r = {0, 0, 0, 0}, bga_i = {0, 0, 0, 0};
        d4 an_r = {0, 0, 0, 0}, an_i = {0, 0, 0, 0}, bn_r = {0, 0, 0, 0}, bn_i = {0, 0, 0, 0};
        mm<true, true, false, false>(agb_r, agb_i, a[0], a[1], t1[0], t1[1], q);
        mm<true, true, false, false>(bga_r, bga_i, b[0], b[1], G[0], G[1], q);
        mm<true, true, false, false>(an_r, an_i, a[0], a[1], G[0], G[1], q);
        mm<true, true, false, false>(bn_r, bn_i, b[0], b[1], t1[0], t1[1], q);
        __syncthreads();
        double m = 0.0;
        const int j = q.col();
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int o = q.row(v) * BLD + j;
            eps_s[0][o] += agb_r[v];
            eps_s[1][o] += agb_i[v];
            eps[0][o] += agb_r[v] + bga_r[v];
            eps[1][o] += agb_i[v] + bga_i[v];
            a[0][o] = an_r[v];
            a[1][o] = an_i[v];
            b[0][o] = bn_r[v];
            b[1][o] = bn_i[v];
            m = nan_max(m, nan_max(mag2(an_r[v], an_i[v]), mag2(bn_r[v], bn_i[v])));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = nan_max(m, __shfl_xor(m, off));
        if (q.lane == 0) wmax[tid >> 6] = m;
        __syncthreads();
        const double mm4 = nan_max(nan_max(wmax[0], wmax[1]), nan_max(wmax[2], wmax[3]));
        __syncthreads();
        if (mm4 < tol * tol) {
            conv = true;
            break;
        }
    }
    // surface Green's function g = (E - eps_s)^-1 into G
    for (int e2 = tid; e2 < BM * BM; e2 += BTHREADS) {
        const int i = e2 >> 5, j = e2 & 31;
        G[0][i * BLD + j] = (i == j ? Er : 0.0) - eps_s[0][i * BLD + j];
        G[1][i * BLD + j] = (i == j ? Ei : 0.0) - eps_s[1][i * BLD + j];
    }
    __syncthreads();
    gj_inverse(G[0], G[1], t1[0], t1[1], perm);
    d4 r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
    if (side == 0) {
        // Sigma_L = T^T (g T)
        mm<true, false, false, false>(r1, i1, G[0], G[1], Tl, nullptr, q);
        __syncthreads();
        store_quad(t1[0], t1[1], r1, i1, q);
        __syncthreads();
        r1 = d4{0, 0, 0, 0};
        i1 = d4{0, 0, 0, 0};
        mm<false, true, true, false>(r1, i1, Tl, nullptr, t1[0], t1[1], q);
    } else {
        // Sigma_R = T (g T^T)
        mm<true, false, false, true>(r1, i1, G[0], G[1], Tl, nullptr, q);
        __syncthreads();
        store_quad(t1[0], t1[1], r1, i1, q);
        __syncthreads();
        r1 = d4{0, 0, 0, 0};
        i1 = d4{0, 0, 0, 0};
        mm<false, true, false, false>(r1, i1, Tl, nullptr, t1[0], t1[1], q);
    }
    double* out = Sigma + ((size_t)e * 2 + side) * 2 * BGL;
    const int j = q.col();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        out[q.row(v) * BM + j] = r1[v];
        out[BGL + q.row(v) * BM + j] = i1[v];
    }
    if (tid == 0) iters[e * 2 + side] = conv ? it : -it;
}

// Block RGF, one wavefront per energy: energy e0 + el with el = blockIdx.x *
// WPB + wavefront (el < n).  Forward sweep of left-connected g_b (stored to
// gstore[el][b]), backward sweep of G_bb; writes part[e][b*32 + o] =
// Im(coef_e G_e,bb,oo) and, if diag, G_e,bb,oo.  T is shared in LDS by the
// workgroup's wavefronts; H0 and the stored g_b are read through the caches.
// KEEP_X: the forward product g_b T (computed for S) is stored beside g_b and
// read back as the backward's X instead of being computed again.
template <bool KEEP_X>
__global__ __launch_bounds__(64 * WPB, 1) void block_rgf_kernel(int e0, int n, int nb, const double2* __restrict__ E,
                                                             const double2* __restrict__ coef,
                                                             const double* __restrict__ H0g,
                                                             const double* __restrict__ Tg,
                                                             const double* __restrict__ v,
                                                             const double* __restrict__ Sigma,
                                                             double* __restrict__ gstore, double* __restrict__ part,
                                                             double2* __restrict__ diag)
{
    __shared__ double Ts[BPL];
    __shared__ double pl[WPB][4][BPL];  // per wavefront: two complex planes
    __shared__ int perm_s[WPB][BM];
    load_real(Ts, Tg);
    __syncthreads();  // the only workgroup barrier
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int el = blockIdx.x * WPB + w;
    if (el >= n) return;  // whole wavefront
    const int e = e0 + el;
    const double Er = E[e].x, Ei = E[e].y;
    const double cr = coef[e].x, ci = coef[e].y;
    const double* SL = Sigma + (size_t)e * 4 * BGL;
    const double* SR = SL + 2 * BGL;
    const size_t nstore = nb > 1 ? nb - 1 : 1;
    double* gs = gstore + (size_t)el * nstore * (KEEP_X ? 2 : 1) * 2 * BGL;  // g_0 .. g_{nb-2}
    double* xs = gs + nstore * 2 * BGL;                                      // (KEEP_X) X_0 .. X_{nb-2}
    double* Cre = pl[w][0];
    double* Cim = pl[w][1];
    double* Pre = pl[w][2];
    double* Pim = pl[w][3];
    // ---- forward sweep
    for (int b = 0; b < nb; ++b) {
        Quad4 S;
        zero(S);
        if (b > 0) {
            Quad4 P;
            zero(P);
            wmm<true, false, false, false>(P, Cre, Cim, Ts, nullptr, lane);  // g_{b-1} T
            store_quad4(Pre, Pim, P, lane);
            if (KEEP_X) {
                double* dst = xs + (size_t)(b - 1) * 2 * BGL;
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int vv = 0; vv < 4; ++vv) {
                        const int o = wrow(lane, q >> 1, vv) * BM + wcol(lane, q & 1);
                        dst[o] = P.re[q][vv];
                        dst[BGL + o] = P.im[q][vv];
                    }
            }
            wave_sync();
            wmm<false, true, true, false, BLD, BLD, true>(S, Ts, nullptr, Pre, Pim, lane);  // T^T g_{b-1} T
        }
        // A = E - H0 - diag(v_b) - S (- Sigma_L on the first block, - Sigma_R on the last);
        // A is complex symmetric: quadrant (1, 0) is the mirror of quadrant (0, 1)
        const double* vb = v + (size_t)b * BM;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q == 2) continue;
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int i = wrow(lane, q >> 1, vv), j = wcol(lane, q & 1);
                double ar = -H0g[i * BM + j] - S.re[q][vv];
                double ai = -S.im[q][vv];
                if (i == j) {
                    ar += Er - vb[i];
                    ai += Ei;
                }
                if (b == 0) {
                    ar -= SL[i * BM + j];
                    ai -= SL[BGL + i * BM + j];
                }
                if (b == nb - 1) {
                    ar -= SR[i * BM + j];
                    ai -= SR[BGL + i * BM + j];
                }
                Cre[i * BLD + j] = ar;
                Cim[i * BLD + j] = ai;
                if (q == 1) {
                    Cre[j * BLD + i] = ar;
                    Cim[j * BLD + i] = ai;
                }
            }
        }
        wave_sync();
        gj_inverse_wave(Cre, Cim, perm_s[w], lane);
        if (b < nb - 1) {
            double* dst = gs + (size_t)b * 2 * BGL;
            for (int x = lane; x < BGL; x += 64) {
                dst[x] = Cre[(x >> 5) * BLD + (x & 31)];
                dst[BGL + x] = Cim[(x >> 5) * BLD + (x & 31)];
            }
        }
    }
    // ---- backward sweep: C holds G_{nb-1} = g_{nb-1}
    auto emit = [&](int b) {
        if (lane < BM) {
            const double gr = Cre[lane * BLD + lane], gi = Cim[lane * BLD + lane];
            part[(size_t)e * nb * BM + (size_t)b * BM + lane] = cr * gi + ci * gr;
            if (diag) diag[(size_t)e * nb * BM + (size_t)b * BM + lane] = make_double2(gr, gi);
        }
    };
    emit(nb - 1);
    for (int b = nb - 2; b >= 0; --b) {
        const double* gb = gs + (size_t)b * 2 * BGL;
        Quad4 X;
        if (KEEP_X) {
            const double* src = xs + (size_t)b * 2 * BGL;  // X = g_b T from the forward sweep
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    const int o = wrow(lane, q >> 1, vv) * BM + wcol(lane, q & 1);
                    X.re[q][vv] = src[o];
                    X.im[q][vv] = src[BGL + o];
                }
        } else {
            zero(X);
            wmm<true, false, false, false, BM, BLD>(X, gb, gb + BGL, Ts, nullptr, lane);  // X = g_b T
        }
        store_quad4(Pre, Pim, X, lane);
        wave_sync();
        Quad4 Y;
        zero(Y);
        wmm<true, true, false, false>(Y, Pre, Pim, Cre, Cim, lane);  // Y = X G_{b+1}
        wave_sync();
        store_quad4(Cre, Cim, Y, lane);
        wave_sync();
        Quad4 G;  // G_b is complex symmetric: quadrants 0, 1, 3
        zero(G);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q == 2) continue;
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int o = wrow(lane, q >> 1, vv) * BM + wcol(lane, q & 1);
                G.re[q][vv] = gb[o];
                G.im[q][vv] = gb[BGL + o];
            }
        }
        wmm<true, true, false, true, BLD, BLD, true>(G, Cre, Cim, Pre, Pim, lane);  // G_b = g_b + Y X^T
        wave_sync();
        store_quad4_sym(Cre, Cim, G, lane);
        wave_sync();
        emit(b);
    }
}

// out[j] = sum_e part[e][j] in energy order.
__global__ void block_reduce_kernel(int nE, int len, const double* __restrict__ part, double* __restrict__ out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    double a = 0.0;
    for (int e = 0; e < nE; ++e) a += part[(size_t)e * len + j];
    out[j] = a;
}

struct BlockChain {
    int nb = 0;
    int m = BM;  // orbitals per block of the caller; the device blocks are BM wide (padded)
    std::vector<double> H0, T;
    double Uc[2] = {0.0, 0.0};
    DevBuf d_H0, d_T, d_v, d_E, d_coef, d_Sigma, d_iters, d_g, d_part, d_out, d_diag;
    size_t scratch_limit = (size_t)80 << 30;  // bytes of g_b (and g_b T) storage per launch
    float ms_sr = 0.f, ms_rgf = 0.f;
    hipStream_t st = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    ~BlockChain()
    {
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
        for (auto& x : ev)
            if (x) (void)hipEventDestroy(x);
    }
};

}  // namespace negf

using namespace negf;

struct negf_block_chain {
    BlockChain c;
};

extern "C" {

// Blocks of m < BM orbitals run as BM-wide blocks: the extra orbitals have
// the on-site energy kPadLevel and no couplings, so they are decoupled from
// the caller's orbitals (G is block diagonal in the two sets), their Green's
// function 1 / (E - kPadLevel - Sigma = 0) stays finite on every path, and
// their outputs are dropped.
static constexpr double kPadLevel = 1.0e4;  // eV

int negf_block_chain_create(int m, int nblocks, const double* H0, const double* T, negf_block_chain** out)
{
    clear_error();
    if (m < 1 || m > BM || nblocks < 1 || !H0 || !T || !out) {
        set_error("block chain needs 1 <= m <= %d, nblocks >= 1 and H0, T", BM);
        return NEGF_EARG;
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            if (H0[i * m + j] != H0[j * m + i]) {
                set_error("H0 must be symmetric");
                return NEGF_EARG;
            }
    negf_block_chain* h = new negf_block_chain;
    BlockChain& c = h->c;
    c.nb = nblocks;
    c.m = m;
    c.H0.assign(BM * BM, 0.0);
    c.T.assign(BM * BM, 0.0);
    for (int i = 0; i < BM; ++i)
        for (int j = 0; j < BM; ++j) {
            c.H0[i * BM + j] = (i < m && j < m) ? H0[i * m + j] : (i == j ? kPadLevel : 0.0);
            c.T[i * BM + j] = (i < m && j < m) ? T[i * m + j] : 0.0;
        }
    H0 = c.H0.data();
    T = c.T.data();
    int rc = 0;
    auto fail = [&](int r) {
        delete h;
        return r;
    };
    if (hipStreamCreate(&c.st) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(NEGF_EDEVICE);
    }
    for (auto& x : c.ev)
        if (hipEventCreate(&x) != hipSuccess) {
            set_error("hipEventCreate failed");
            return fail(NEGF_EDEVICE);
        }
    if ((rc = c.d_H0.ensure(sizeof(double) * BM * BM))) return fail(rc);
    if ((rc = c.d_T.ensure(sizeof(double) * BM * BM))) return fail(rc);
    if ((rc = c.d_v.ensure(sizeof(double) * (size_t)nblocks * BM))) return fail(rc);
    if (hipMemcpy(c.d_H0.p, H0, sizeof(double) * BM * BM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.d_T.p, T, sizeof(double) * BM * BM, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c.d_v.p, 0, sizeof(double) * (size_t)nblocks * BM) != hipSuccess) {
        set_error("block chain upload failed");
        return fail(NEGF_EDEVICE);
    }
    *out = h;
    return NEGF_OK;
}

int negf_block_chain_dims(const negf_block_chain* h, int* m, int* nblocks)
{
    if (!h || !m || !nblocks) return NEGF_EARG;
    *m = h->c.m;
    *nblocks = h->c.nb;
    return NEGF_OK;
}

int negf_block_chain_destroy(negf_block_chain* h)
{
    if (!h) return NEGF_EARG;
    delete h;
    return NEGF_OK;
}

int negf_block_chain_set_potential(negf_block_chain* h, const double* v, const double* Uc)
{
    clear_error();
    if (!h || !v || !Uc) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const int m = h->c.m, nb = h->c.nb;
    if (m == BM) {
        NEGF_HIP_TRY(hipMemcpy(h->c.d_v.p, v, sizeof(double) * (size_t)nb * BM, hipMemcpyHostToDevice));
    } else {
        std::vector<double> pad((size_t)nb * BM, 0.0);
        for (int b = 0; b < nb; ++b) std::copy(v + (size_t)b * m, v + (size_t)(b + 1) * m, pad.begin() + (size_t)b * BM);
        NEGF_HIP_TRY(hipMemcpy(h->c.d_v.p, pad.data(), sizeof(double) * pad.size(), hipMemcpyHostToDevice));
    }
    h->c.Uc[0] = Uc[0];
    h->c.Uc[1] = Uc[1];
    return NEGF_OK;
}

int negf_block_rgf(negf_block_chain* h, int nE, const negf_complex* E, const negf_complex* coef, double sr_tol,
                   int sr_max_iter, double* acc, negf_complex* diagG, int* sr_iters)
{
    clear_error();
    if (!h || nE < 1 || !E || !coef || !(sr_tol > 0) || sr_max_iter < 1) {
        set_error("invalid block RGF arguments");
        return NEGF_EARG;
    }
    BlockChain& c = h->c;
    const int nb = c.nb;
    const size_t len = (size_t)nb * BM;
    int rc;
    if ((rc = c.d_E.ensure(sizeof(double2) * nE))) return rc;
    if ((rc = c.d_coef.ensure(sizeof(double2) * nE))) return rc;
    if ((rc = c.d_Sigma.ensure(sizeof(double) * (size_t)nE * 4 * BGL))) return rc;
    if ((rc = c.d_iters.ensure(sizeof(int) * 2 * nE))) return rc;
    if ((rc = c.d_part.ensure(sizeof(double) * (size_t)nE * len))) return rc;
    if ((rc = c.d_out.ensure(sizeof(double) * len))) return rc;
    if (diagG && (rc = c.d_diag.ensure(sizeof(double2) * (size_t)nE * len))) return rc;
    // g_b for every block and energy; also g_b T when all energies fit in one sweep that way
    const size_t per_g = (size_t)std::max(nb - 1, 1) * 2 * BGL * sizeof(double);
    const bool keep_x = 2 * per_g * (size_t)nE <= c.scratch_limit;
    const size_t per_energy = keep_x ? 2 * per_g : per_g;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)nE, c.scratch_limit / per_energy));
    if ((rc = c.d_g.ensure(per_energy * chunk))) return rc;
    NEGF_HIP_TRY(hipMemcpyAsync(c.d_E.p, E, sizeof(double2) * nE, hipMemcpyHostToDevice, c.st));
    NEGF_HIP_TRY(hipMemcpyAsync(c.d_coef.p, coef, sizeof(double2) * nE, hipMemcpyHostToDevice, c.st));
    NEGF_HIP_TRY(hipMemsetAsync(c.d_iters.p, 0, sizeof(int) * 2 * nE, c.st));
    NEGF_HIP_TRY(hipEventRecord(c.ev[0], c.st));
    hipLaunchKernelGGL(sancho_rubio_kernel, dim3(nE, 2), dim3(BTHREADS), 0, c.st, nE, c.d_E.as<double2>(),
                       c.d_H0.as<double>(), c.d_T.as<double>(), c.Uc[0], c.Uc[1], sr_tol, sr_max_iter,
                       c.d_Sigma.as<double>(), c.d_iters.as<int>());
    NEGF_HIP_TRY(hipGetLastError());
    NEGF_HIP_TRY(hipEventRecord(c.ev[1], c.st));
    for (int e0 = 0; e0 < nE; e0 += chunk) {
        const int n = std::min(chunk, nE - e0);
        if (keep_x)
            hipLaunchKernelGGL(block_rgf_kernel<true>, dim3((n + WPB - 1) / WPB), dim3(64 * WPB), 0, c.st, e0, n,
                               nb, c.d_E.as<double2>(), c.d_coef.as<double2>(), c.d_H0.as<double>(),
                               c.d_T.as<double>(), c.d_v.as<double>(), c.d_Sigma.as<double>(), c.d_g.as<double>(),
                               c.d_part.as<double>(), diagG ? c.d_diag.as<double2>() : nullptr);
        else
            hipLaunchKernelGGL(block_rgf_kernel<false>, dim3((n + WPB - 1) / WPB), dim3(64 * WPB), 0, c.st, e0, n,
                               nb, c.d_E.as<double2>(), c.d_coef.as<double2>(), c.d_H0.as<double>(),
                               c.d_T.as<double>(), c.d_v.as<double>(), c.d_Sigma.as<double>(), c.d_g.as<double>(),
                               c.d_part.as<double>(), diagG ? c.d_diag.as<double2>() : nullptr);
        NEGF_HIP_TRY(hipGetLastError());
    }
    NEGF_HIP_TRY(hipEventRecord(c.ev[2], c.st));
    hipLaunchKernelGGL(block_reduce_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, c.st, nE, (int)len,
                       c.d_part.as<double>(), c.d_out.as<double>());
    NEGF_HIP_TRY(hipGetLastError());
    std::vector<int> it(2 * nE);
    NEGF_HIP_TRY(hipMemcpyAsync(it.data(), c.d_iters.p, sizeof(int) * 2 * nE, hipMemcpyDeviceToHost, c.st));
    const int m = c.m;
    if (m == BM) {
        if (acc) NEGF_HIP_TRY(hipMemcpyAsync(acc, c.d_out.p, sizeof(double) * len, hipMemcpyDeviceToHost, c.st));
        if (diagG)
            NEGF_HIP_TRY(hipMemcpyAsync(diagG, c.d_diag.p, sizeof(double2) * (size_t)nE * len,
                                        hipMemcpyDeviceToHost, c.st));
    } else {
        // the caller's m orbitals of every BM-wide block (rows of BM, first m)
        if (acc)
            NEGF_HIP_TRY(hipMemcpy2DAsync(acc, sizeof(double) * m, c.d_out.p, sizeof(double) * BM,
                                          sizeof(double) * m, nb, hipMemcpyDeviceToHost, c.st));
        if (diagG)
            NEGF_HIP_TRY(hipMemcpy2DAsync(diagG, sizeof(double2) * m, c.d_diag.p, sizeof(double2) * BM,
                                          sizeof(double2) * m, (size_t)nE * nb, hipMemcpyDeviceToHost, c.st));
    }
    NEGF_HIP_TRY(hipStreamSynchronize(c.st));
    NEGF_HIP_TRY(hipEventElapsedTime(&c.ms_sr, c.ev[0], c.ev[1]));
    NEGF_HIP_TRY(hipEventElapsedTime(&c.ms_rgf, c.ev[1], c.ev[2]));
    if (sr_iters) std::copy(it.begin(), it.end(), sr_iters);
    for (int k = 0; k < 2 * nE; ++k)
        if (it[k] <= 0) {
            set_error("Sancho-Rubio did not reach the tolerance for energy %d within %d iterations", k / 2,
                      sr_max_iter);
            return NEGF_ENONCONV;
        }
    return NEGF_OK;
}

int negf_block_self_energies(negf_block_chain* h, int nE, const negf_complex* E, double sr_tol, int sr_max_iter,
                              double* sigma, int* sr_iters)
{
    clear_error();
    if (!h || nE < 1 || !E || !sigma || !(sr_tol > 0) || sr_max_iter < 1) {
        set_error("invalid self-energy arguments");
        return NEGF_EARG;
    }
    BlockChain& c = h->c;
    int rc;
    if ((rc = c.d_E.ensure(sizeof(double2) * nE))) return rc;
    if ((rc = c.d_Sigma.ensure(sizeof(double) * (size_t)nE * 4 * BGL))) return rc;
    if ((rc = c.d_iters.ensure(sizeof(int) * 2 * nE))) return rc;
    NEGF_HIP_TRY(hipMemcpyAsync(c.d_E.p, E, sizeof(double2) * nE, hipMemcpyHostToDevice, c.st));
    NEGF_HIP_TRY(hipMemsetAsync(c.d_iters.p, 0, sizeof(int) * 2 * nE, c.st));
    hipLaunchKernelGGL(sancho_rubio_kernel, dim3(nE, 2), dim3(BTHREADS), 0, c.st, nE, c.d_E.as<double2>(),
                       c.d_H0.as<double>(), c.d_T.as<double>(), c.Uc[0], c.Uc[1], sr_tol, sr_max_iter,
                       c.d_Sigma.as<double>(), c.d_iters.as<int>());
    NEGF_HIP_TRY(hipGetLastError());
    std::vector<int> it(2 * nE);
    NEGF_HIP_TRY(hipMemcpyAsync(it.data(), c.d_iters.p, sizeof(int) * 2 * nE, hipMemcpyDeviceToHost, c.st));
    const int m = c.m;
    if (m == BM) {
        NEGF_HIP_TRY(hipMemcpyAsync(sigma, c.d_Sigma.p, sizeof(double) * (size_t)nE * 4 * BGL,
                                    hipMemcpyDeviceToHost, c.st));
    }
    std::vector<double> full(m == BM ? 0 : (size_t)nE * 4 * BGL);
    if (m != BM)
        NEGF_HIP_TRY(hipMemcpyAsync(full.data(), c.d_Sigma.p, sizeof(double) * full.size(), hipMemcpyDeviceToHost,
                                    c.st));
    NEGF_HIP_TRY(hipStreamSynchronize(c.st));
    // the m x m leading submatrix of every [energy][lead][plane] matrix
    for (size_t q = 0; m != BM && q < (size_t)nE * 4; ++q)
        for (int i = 0; i < m; ++i)
            std::copy(full.data() + q * BGL + (size_t)i * BM, full.data() + q * BGL + (size_t)i * BM + m,
                      sigma + q * m * m + (size_t)i * m);
    if (sr_iters) std::copy(it.begin(), it.end(), sr_iters);
    for (int k = 0; k < 2 * nE; ++k)
        if (it[k] <= 0) {
            set_error("Sancho-Rubio did not reach the tolerance for energy %d within %d iterations", k / 2,
                      sr_max_iter);
            return NEGF_ENONCONV;
        }
    return NEGF_OK;
}

int negf_block_chain_set_scratch_limit(negf_block_chain* h, size_t bytes)
{
    clear_error();
    if (!h || bytes == 0) {
        set_error("invalid scratch limit");
        return NEGF_EARG;
    }
    h->c.scratch_limit = bytes;
    return NEGF_OK;
}

int negf_block_last_time(const negf_block_chain* h, float* ms_sancho_rubio, float* ms_rgf)
{
    if (!h) return NEGF_EARG;
    if (ms_sancho_rubio) *ms_sancho_rubio = h->c.ms_sr;
    if (ms_rgf) *ms_rgf = h->c.ms_rgf;
    return NEGF_OK;
}

}  // extern "C"

namespace negf {
// op 0: C = A B (complex), op 1: C = A^-1 (complex), op 2: C = A B with
// B real (B im plane ignored); ops 3, 4, 5: the same by one wavefront (the
// block RGF sweep's forms).  A, B, C: [re plane, im plane] of 32 x 32.
__global__ __launch_bounds__(BTHREADS) void block_probe_kernel(int op, const double* __restrict__ A,
                                                               const double* __restrict__ Bm, double* __restrict__ Cm)
{
    __shared__ double a[2][BPL], b[2][BPL];
    __shared__ int perm[BM];
    const Quad q;
    load_cplx(a[0], a[1], A);
    load_cplx(b[0], b[1], Bm);
    __syncthreads();
    if (op >= 3) {
        if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            if (op == 4) {
                gj_inverse_wave(a[0], a[1], perm, lane);
            } else {
                Quad4 c;
                zero(c);
                if (op == 3)
                    wmm<true, true, false, false>(c, a[0], a[1], b[0], b[1], lane);
                else
                    wmm<true, false, false, false>(c, a[0], a[1], b[0], nullptr, lane);
                wave_sync();
                store_quad4(a[0], a[1], c, lane);
            }
        }
        __syncthreads();
        store_cplx(Cm, a[0], a[1]);
        return;
    }
    if (op == 1) {
        gj_inverse(a[0], a[1], b[0], b[1], perm);
        store_cplx(Cm, a[0], a[1]);
        return;
    }
    d4 r = {0, 0, 0, 0}, i = {0, 0, 0, 0};
    if (op == 0)
        mm<true, true, false, false>(r, i, a[0], a[1], b[0], b[1], q);
    else
        mm<true, false, false, false>(r, i, a[0], a[1], b[0], nullptr, q);
    __syncthreads();
    store_quad(a[0], a[1], r, i, q);
    __syncthreads();
    store_cplx(Cm, a[0], a[1]);
}
}  // namespace negf

extern "C" int negf_block_probe(int op, const double* A, const double* B, double* C)
{
    negf::clear_error();
    if (op < 0 || op > 5 || !A || !B || !C) {
        negf::set_error("invalid probe arguments");
        return NEGF_EARG;
    }
    negf::DevBuf d;
    if (d.ensure(sizeof(double) * 3 * 2 * BGL)) return NEGF_EALLOC;
    double* p = d.as<double>();
    NEGF_HIP_TRY(hipMemcpy(p, A, sizeof(double) * 2 * BGL, hipMemcpyHostToDevice));
    NEGF_HIP_TRY(hipMemcpy(p + 2 * BGL, B, sizeof(double) * 2 * BGL, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(negf::block_probe_kernel, dim3(1), dim3(BTHREADS), 0, 0, op, p, p + 2 * BGL, p + 4 * BGL);
    NEGF_HIP_TRY(hipGetLastError());
    NEGF_HIP_TRY(hipDeviceSynchronize());
    NEGF_HIP_TRY(hipMemcpy(C, p + 4 * BGL, sizeof(double) * 2 * BGL, hipMemcpyDeviceToHost));
    return NEGF_OK;
}
