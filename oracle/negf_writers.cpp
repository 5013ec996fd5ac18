/*
 * negf_writers.cpp -- TEST INFRASTRUCTURE ONLY.  The reference's per-step
 * output files written with the same iostream manipulators it uses, so the
 * product's writers can be compared byte for byte.
 *
 *   _U.dat     Write_PotentialAtSites NEGF_Common.cpp:3043-3078
 *   _Qout.dat  Write_InducedCharge :3092-3100 -> Write_Table1D :4223-4254
 *   _norm.dat  Write_ChargeNorm :3102-3110 -> Write_Table1D
 *   _Qin.dat   Write_InputInducedCharge :3081-3092 -> Write_Table1D
 *   _chargeComp.dat  Write_ChargeComponents :2999-3023
 *   I.dat      Define_FileHeaderForCurrent :73-82, Write_Current :4818-4840
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

extern "C" void orc_write_qin(const char *prefix, int N, const double *ptd,
                              const double *Qin)
{
    table(std::string(prefix) + "_Qin.dat",
          "'axial location / (nm)', 'Induced charge per site / (e)'", N, ptd,
          Qin);
}

extern "C" void orc_write_charge_components(const char *filename, int N,
                                            const double *ptd,
                                            const double *eq,
                                            const double *neq,
                                            const double *r0,
                                            const double *Q)
{
    std::ofstream f(filename);
    f << "axis, RhoEq, RhoNonEq, Rho0, Rho_Induced"
      << "\n";
    for (int n = 0; n < N; ++n)
        f << ptd[n] << std::setw(20) << eq[n] << std::setw(20) << neq[n]
          << std::setw(20) << r0[n] << std::setw(20) << Q[n] << "\n";
}

extern "C" void orc_write_current_header(const char *filename)
{
    std::ofstream f(filename, std::ios::app);
    f << "'step', 'Vds' , 'Vgs', ";
    for (int k = 0; k < 2; ++k) f << ", 'I at contact_" << k + 1 << "',";
    f << "'Avg_intg_pts', 'Total_Iter', 'Broyden_Fraction', \
                      'Broyden_Scalar', 'Conductance / (2q^2/h)'"
      << "\n";
}

extern "C" void orc_write_current_line(const char *filename, int step,
                                       double Vds, double Vgs,
                                       const double *I, double avg_pts,
                                       int iters, double conductance)
{
    std::ofstream f(filename, std::ios_base::app);
    f << std::setw(10) << step << std::setw(15) << Vds << std::setw(15) << Vgs;
    for (int k = 0; k < 2; ++k) f << std::setw(20) << I[k];
    f << std::setw(10) << avg_pts << std::setw(10) << iters << std::setw(20)
      << conductance << "\n";
}
