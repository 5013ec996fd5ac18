// quadrature.cpp -- host-side Gauss-Legendre nodes and integration paths.
//
// Same algorithm as the reference (Newton iteration on the Legendre
// recurrence to |dz| <= 3e-11, Quadrature::Gauss_Legendre
// Source/Utils/CodeUtils/CodeUtil.cpp:640-684) so that the energy grids, and
// hence the integrated charge, agree with it node for node; line and
// circular-arc maps as in c_IntegrationPath::Define_GaussLegendrePoints
// (Source/Solver/Transport/NEGF/IntegrationPath.cpp:41-147).
#include <cmath>
#include <vector>

#include "negf_internal.h"

namespace negf {

static const double kPi = 3.14159265358979323846;

void gauss_legendre(int n, double* x, double* w)
{
    const int half = (n + 1) / 2;
    for (int i = 1; i <= half; ++i) {
        double z = std::cos(kPi * (i - 0.25) / (n + 0.5));
        double dz, dp;
        do {
            // P_n(z) and P_{n-1}(z) by the three-term recurrence
            double p0 = 1.0, p_1 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p_2 = p_1;
                p_1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p_1 - (j - 1.0) * p_2) / j;
            }
            dp = n * (z * p0 - p_1) / (z * z - 1.0);
            double znew = z - p0 / dp;
            dz = znew - z;
            z = znew;
        } while (std::fabs(dz) > 3.0e-11);
        x[i - 1] = -z;
        x[n - i] = z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * dp * dp);
        w[n - i] = w[i - 1];
    }
}

void gl_path(double ar, double ai, double br, double bi, int n, int type, double* E, double* w, double* mul)
{
    std::vector<double> x(n), ww(n);
    gauss_legendre(n, x.data(), ww.data());
    if (type == 0) {
        const double hr = 0.5 * (br - ar), hi = 0.5 * (bi - ai);
        const double cr = 0.5 * (br + ar), ci = 0.5 * (bi + ai);
        for (int i = 0; i < n; ++i) {
            E[2 * i] = hr * x[i] + cr;
            E[2 * i + 1] = hi * x[i] + ci;
            mul[2 * i] = hr;
            mul[2 * i + 1] = hi;
            w[i] = ww[i];
        }
        return;
    }
    // arc through a (left end, below) and b (top), centred on the line Im = Im(b)
    const double cre = (br * br - ar * ar - ai * ai) / (2.0 * (br - ar));
    const double cim = bi;
    const double R = cre - br;
    double th1 = std::asin(std::fabs(ai) / R);
    if (cre > ar) th1 = kPi - th1;
    const double th2 = kPi;
    for (int i = 0; i < n; ++i) {
        const double th = 0.5 * (th2 - th1) * x[i] + 0.5 * (th2 + th1);
        const double c = std::cos(th), s = std::sin(th);
        E[2 * i] = cre + R * c;
        E[2 * i + 1] = cim + R * s;
        // (th2-th1)/2 * R * e^{i th} * i
        const double f = 0.5 * (th2 - th1) * R;
        mul[2 * i] = -f * s;
        mul[2 * i + 1] = f * c;
        w[i] = ww[i];
    }
}

}  // namespace negf
