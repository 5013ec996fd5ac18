/*
 * negf_writers.cpp -- TEST INFRASTRUCTURE ONLY.  The reference's per-step
 * output files written with the same iostream manipulators it uses, so the
 * product's writers can be compared byte for byte.
 *
 *   _U.dat     Write_PotentialAtSites NEGF_Common.cpp:3043-3078
 *   _Qout.dat  Write_InducedCharge :3092-3100 -> Write_Table1D :4223-4254
 *   _norm.dat  Write_ChargeNorm :3102-3110 -> Write_Table1D
 */
#include <fstream>
#include <iomanip>
#include <string>

#include "negf_oracle.h"

static void table(const std::string &name, const std::string &header, int N,
                  const double *x, const double *y)
{
    std::ofstream f(name.c_str());
    f << header << "\n";
    for (int e = 0; e < N; ++e)
        f << std::setprecision(15) << std::setw(35) << x[e] << std::setw(35)
          << y[e] << "\n";
}

extern "C" void orc_write_step_files(const char *prefix, int N,
                                     const double *ptd, const double *Q,
                                     const double *U, const double *norm)
{
    std::string pre(prefix);
    if (U) {
        std::ofstream f((pre + "_U.dat").c_str());
        for (int l = 0; l < N; ++l)
            f << l << std::setw(35) << ptd[l] << std::setw(35) << U[l] << "\n";
    }
    if (Q)
        table(pre + "_Qout.dat",
              "'axial location / (nm)', 'Induced charge per site / (e)'", N,
              ptd, Q);
    if (norm)
        table(pre + "_norm.dat", "'axial location / (nm)', 'norm", N, ptd,
              norm);
}
