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
 *   transport_char.dat  Write_Eql_Characteristics :1931-1970
 *   Ept_<e>.dat  Compute_DensityOfStates :2778-2796 -> Write_Table1D
 */
#include <cmath>
#include <complex>
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

/* Write_Eql_Characteristics NEGF_Common.cpp:1931-1970 (the energies arrive
 * real, E_total_vec :2806-2815); FermiFunction :1790-1795 on complex E - mu. */
extern "C" void orc_write_transport_char(const char *filename, int nE,
                                         const double *E, const double *mu,
                                         const double *kT, const double *dos,
                                         const double *tran,
                                         const double *cond)
{
    std::ofstream outfile(filename);
    outfile << " E_r, Fermi_Contact1, Fermi_Contact2, DOS_r, Transmission_r, "
               "Conductance_r \n";
    auto fermi = [](std::complex<double> z, double kt) {
        return (1. / (std::exp(z / kt) + 1.)).real();
    };
    for (int i = 0; i < nE; ++i)
        outfile << std::setw(20) << E[i] << std::setw(20)
                << fermi(std::complex<double>(E[i], 0.) - mu[0], kT[0])
                << std::setw(20)
                << fermi(std::complex<double>(E[i], 0.) - mu[1], kT[1])
                << std::setw(35) << dos[i] << std::setw(35) << tran[i]
                << std::setw(35) << cond[i] << "\n";
}

/* Ept_<e>.dat, NEGF_Common.cpp:2778-2796: Write_Table1D of the site
 * positions and the LDOS at energy e, header "PTD LDOS_r at E=" +
 * std::to_string(E.real()). */
extern "C" void orc_write_ldos(const char *folder, int nE, const double *E,
                               int N, const double *ptd, const double *ldos)
{
    for (int e = 0; e < nE; ++e)
        table(std::string(folder) + "/Ept_" + std::to_string(e) + ".dat",
              "PTD LDOS_r at E=" + std::to_string(E[e]), N, ptd,
              ldos + (size_t)e * N);
}
