// negf_device.h -- device-side complex helpers and shared launch descriptors
// for the gfx950 NEGF kernels.  FP64 throughout.
#pragma once
#include <hip/hip_runtime.h>

#include "negf_types.h"

namespace negf {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    double inv = 1.0 / (b.x * b.x + b.y * b.y);
    return make_double2((a.x * b.x + a.y * b.y) * inv, (a.y * b.x - a.x * b.y) * inv);
}
__device__ __forceinline__ double2 crecip(double2 b)
{
    double inv = 1.0 / (b.x * b.x + b.y * b.y);
    return make_double2(b.x * inv, -b.y * inv);
}
// Principal square root.
__device__ __forceinline__ double2 csqrt(double2 z)
{
    double r = hypot(z.x, z.y);
    if (r == 0.0) return make_double2(0.0, 0.0);
    if (z.x >= 0.0) {
        double t = sqrt(0.5 * (r + z.x));
        return make_double2(t, z.y / (2.0 * t));
    }
    double t = sqrt(0.5 * (r - z.x));
    return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// Closed-form surface Green's function of the semi-infinite two-site-periodic
// lead (reference CNT.cpp:354-393): z = E-U, F = z^2+g^2-b^2,
// S = sqrt(F^2-4 z^2 g^2), root (F+S)/(2 g^2 z) if Im<0 else (F-S)/(...).
// Sets *ok = 0 when neither root has Im<0 (and returns 0).
__device__ __forceinline__ double2 surface_gf(double2 E, double U, double beta, double gamma, int* ok)
{
    double2 z = make_double2(E.x - U, E.y);
    double2 z2 = cmul(z, z);
    double g2 = gamma * gamma;
    double2 F = make_double2(z2.x + g2 - beta * beta, z2.y);
    double2 disc = csub(cmul(F, F), cscale(4.0 * g2, z2));
    double2 S = csqrt(disc);
    double2 D = cscale(2.0 * g2, z);
    double2 v1 = cdiv(cadd(F, S), D);
    double2 v2 = cdiv(csub(F, S), D);
    if (v1.y < 0.0) { *ok = 1; return v1; }
    if (v2.y < 0.0) { *ok = 1; return v2; }
    // neither root: g = 0, the zero-initialised block the reference keeps
    // (CNT.cpp:376-381); flagged in the launch status word
    *ok = 0;
    return make_double2(0.0, 0.0);
}

// 1-D decimation (reference NEGF_Common.cpp:1436-1508 condensed Hamiltonian
// and :4095-4153 iteration).  Runs max_iter+2 updates when rel_error == 0,
// as the integer-typed tolerance of the reference does.
__device__ __forceinline__ double2 decimation_gf(double2 EmU, double beta, double gamma, int P,
                                                 int max_iter, int rel_error)
{
    const double hb[2] = {-beta, -gamma};
    double2 Hs[NEGF_MAX_DECIMATION_LAYERS];
    Hs[P - 1] = crecip(EmU);
    double2 Ht_kP = Hs[P - 1];
    for (int k = P - 2; k >= 1; --k) {
        int j = k % 2;
        Hs[k] = crecip(csub(EmU, cscale(hb[j] * hb[j], Hs[k + 1])));
        Ht_kP = cscale(hb[j], cmul(Hs[k], Ht_kP));
    }
    double2 Ck = Hs[1];
    for (int k = 2; k < P; ++k) {
        int j = (k - 1) % 2;
        Ck = cadd(Hs[k], cscale(hb[j] * hb[j], cmul(cmul(Hs[k], Ck), Hs[k])));
    }
    int id = (P - 1) % 2;
    double2 Xi_s = cscale(hb[0] * hb[0], Hs[1]);
    double2 Xi = cadd(Xi_s, cscale(hb[id] * hb[id], Ck));
    double2 Pi = cscale(hb[0] * hb[id], Ht_kP);
    double2 mu, nu, gam, zeta;
    {
        double2 mu0 = csub(EmU, Xi_s);
        double2 nu0 = csub(EmU, Xi);
        double2 nu0i = crecip(nu0);
        double2 gd = cconj(Pi);
        double2 t1 = cmul(Pi, nu0i);
        double2 t2 = cmul(t1, gd);
        double2 t3 = cmul(gd, nu0i);
        mu = csub(mu0, t2);
        nu = csub(csub(nu0, t2), cmul(t3, Pi));
        gam = cmul(t1, Pi);
        zeta = cmul(t3, gd);
    }
    double err = 1.0;
    int i = 0;
    while (err > (double)rel_error) {
        double2 mu_old = mu;
        double2 nui = crecip(nu);
        double2 t1 = cmul(cmul(zeta, nui), gam);
        double2 t2 = cmul(cmul(gam, nui), zeta);
        mu = csub(mu, t1);
        nu = csub(csub(nu, t1), t2);
        gam = cmul(cmul(gam, nui), gam);
        zeta = cmul(cmul(zeta, nui), zeta);
        double2 e = cdiv(csub(mu, mu_old), cadd(mu, mu_old));
        err = sqrt(cabs2(e));
        if (i > max_iter) break;
        ++i;
    }
    return crecip(mu);
}

}  // namespace negf
