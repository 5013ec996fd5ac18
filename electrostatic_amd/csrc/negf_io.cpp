// negf_io.cpp -- per-step output files in the reference's column formats.
//
//   <prefix>_Qout.dat  Write_InducedCharge -> Write_Table1D
//                      (NEGF_Common.cpp:3092-3100, :4223-4254): a header line,
//                      then "%35.15g%35.15g" (setprecision(15), setw(35)).
//   <prefix>_norm.dat  Write_ChargeNorm (:3102-3110), same table format.
//   <prefix>_U.dat     Write_PotentialAtSites (:3043-3078): no header,
//                      "l, setw(35) PTD, setw(35) U" at the stream's default
//                      precision (6 significant digits).
//   <prefix>_Qin.dat   Write_InputInducedCharge (:3081-3092), table format.
//   <prefix>_chargeComp.dat  Write_ChargeComponents (:2999-3023).
//   I.dat              Define_FileHeaderForCurrent (:73-82), Write_Current
//                      (:4818-4840), appended one line per step.
//   restart            Read_Table1D (:236-286): header line, N rows of
//                      "position value", line count checked.
//   transport_char.dat Write_Eql_Characteristics (:1931-1970): header, then
//                      "%20g%20g%20g%35g%35g%35g" (setw, default precision).
//   Ept_<e>.dat        Compute_DensityOfStates (:2778-2796) -> Write_Table1D,
//                      header "PTD LDOS_r at E=" + std::to_string(E) ("%f").
#include <cmath>
#include <cstdio>
#include <string>

#include "../../include/negf_amd.h"

namespace negf {
void set_error(const char* fmt, ...);
void clear_error();
}  // namespace negf

static int write_table(const std::string& name, const char* header, int N, const double* x, const double* y)
{
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) {
        negf::set_error("cannot open %s for writing", name.c_str());
        return NEGF_EARG;
    }
    std::fprintf(f, "%s\n", header);
    for (int i = 0; i < N; ++i) std::fprintf(f, "%35.15g%35.15g\n", x[i], y[i]);
    std::fclose(f);
    return NEGF_OK;
}

extern "C" int negf_write_step_files(const char* prefix, int N, const double* ptd_nm, const double* Q,
                                     const double* U, const double* norm)
{
    negf::clear_error();
    if (!prefix || N <= 0 || !ptd_nm) {
        negf::set_error("invalid writer arguments");
        return NEGF_EARG;
    }
    const std::string pre(prefix);
    int rc;
    if (U) {
        const std::string name = pre + "_U.dat";
        FILE* f = std::fopen(name.c_str(), "w");
        if (!f) {
            negf::set_error("cannot open %s for writing", name.c_str());
            return NEGF_EARG;
        }
        for (int l = 0; l < N; ++l) std::fprintf(f, "%d%35.6g%35.6g\n", l, ptd_nm[l], U[l]);
        std::fclose(f);
    }
    if (Q && (rc = write_table(pre + "_Qout.dat", "'axial location / (nm)', 'Induced charge per site / (e)'", N,
                               ptd_nm, Q)))
        return rc;
    if (norm && (rc = write_table(pre + "_norm.dat", "'axial location / (nm)', 'norm", N, ptd_nm, norm))) return rc;
    return NEGF_OK;
}

extern "C" int negf_write_qin(const char* prefix, int N, const double* ptd_nm, const double* Qin)
{
    negf::clear_error();
    if (!prefix || N <= 0 || !ptd_nm || !Qin) {
        negf::set_error("invalid writer arguments");
        return NEGF_EARG;
    }
    return write_table(std::string(prefix) + "_Qin.dat", "'axial location / (nm)', 'Induced charge per site / (e)'",
                       N, ptd_nm, Qin);
}

extern "C" int negf_write_charge_components(const char* filename, int N, const double* ptd_nm, const double* eq_im,
                                            const double* noneq_re, const double* rho0_im, const double* Q)
{
    negf::clear_error();
    if (!filename || N <= 0 || !ptd_nm || !eq_im || !noneq_re || !rho0_im || !Q) {
        negf::set_error("invalid writer arguments");
        return NEGF_EARG;
    }
    FILE* f = std::fopen(filename, "w");
    if (!f) {
        negf::set_error("cannot open %s for writing", filename);
        return NEGF_EARG;
    }
    std::fprintf(f, "axis, RhoEq, RhoNonEq, Rho0, Rho_Induced\n");
    for (int n = 0; n < N; ++n)
        std::fprintf(f, "%g%20g%20g%20g%20g\n", ptd_nm[n], eq_im[n], noneq_re[n], rho0_im[n], Q[n]);
    std::fclose(f);
    return NEGF_OK;
}

extern "C" int negf_write_current_header(const char* filename)
{
    negf::clear_error();
    FILE* f = filename ? std::fopen(filename, "a") : nullptr;
    if (!f) {
        negf::set_error("cannot open current file for appending");
        return NEGF_EARG;
    }
    // the reference's header literal is line-spliced, so the indentation of
    // its second source line stays inside the string (22 spaces)
    std::fprintf(f, "'step', 'Vds' , 'Vgs', , 'I at contact_1',, 'I at contact_2',");
    std::fprintf(f, "'Avg_intg_pts', 'Total_Iter', 'Broyden_Fraction', %22s'Broyden_Scalar', "
                    "'Conductance / (2q^2/h)'\n", "");
    std::fclose(f);
    return NEGF_OK;
}

extern "C" int negf_write_current_line(const char* filename, int step, double Vds, double Vgs, const double* I,
                                       double avg_intg_pts, int total_iterations, double conductance)
{
    negf::clear_error();
    FILE* f = (filename && I) ? std::fopen(filename, "a") : nullptr;
    if (!f) {
        negf::set_error("cannot open current file for appending");
        return NEGF_EARG;
    }
    std::fprintf(f, "%10d%15g%15g%20g%20g%10g%10d%20g\n", step, Vds, Vgs, I[0], I[1], avg_intg_pts,
                 total_iterations, conductance);
    std::fclose(f);
    return NEGF_OK;
}

extern "C" int negf_read_charge_distribution(const char* filename, int N, double* Q)
{
    negf::clear_error();
    if (!filename || N <= 0 || !Q) {
        negf::set_error("invalid reader arguments");
        return NEGF_EARG;
    }
    FILE* f = std::fopen(filename, "r");
    if (!f) {
        negf::set_error("Failed to read file %s", filename);
        return NEGF_EARG;
    }
    // count lines (the reference requires header + N rows)
    int lines = 0, c, last = '\n';
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n') ++lines;
        last = c;
    }
    if (last != '\n') ++lines;
    if (lines - 1 != N) {
        std::fclose(f);
        negf::set_error("Assert size, %d, is not equal to the filesize-1, %d !", N, lines - 1);
        return NEGF_EARG;
    }
    std::rewind(f);
    while ((c = std::fgetc(f)) != EOF && c != '\n') {
    }
    for (int l = 0; l < N; ++l) {
        double pos, val;
        if (std::fscanf(f, "%lf %lf", &pos, &val) != 2) {
            std::fclose(f);
            negf::set_error("malformed row %d in %s", l, filename);
            return NEGF_EARG;
        }
        Q[l] = val;
    }
    std::fclose(f);
    return NEGF_OK;
}

// Compute_Conductance (NEGF_Common.cpp:1873-1928): G(E) = -T(E) dF/dE with
// second-order one-sided differences at the ends, central inside, and the
// reference's step dE = E_0 - E_1 (negative on an ascending grid); returns
// sum_i G_i dE in units of 2q^2/h.
extern "C" int negf_conductance(int nE, const double* E, const double* T, double mu0, double kT0, double* cond,
                                double* total)
{
    negf::clear_error();
    if (nE < 3 || !E || !T || !cond || !total || kT0 <= 0) {
        negf::set_error("invalid conductance arguments");
        return NEGF_EARG;
    }
    auto F = [&](double e) { return 1.0 / (std::exp((e - mu0) / kT0) + 1.0); };
    const double dE = E[0] - E[1];
    const double inv = 1.0 / (2 * dE);
    double sum = 0.0;
    const int n = nE - 1;
    for (int i = 0; i < nE; ++i) {
        double d;
        if (i == 0)
            d = (-3 * F(E[0]) + 4 * F(E[1]) - F(E[2])) * inv;
        else if (i == n)
            d = (3 * F(E[n]) - 4 * F(E[n - 1]) + F(E[n - 2])) * inv;
        else
            d = (F(E[i + 1]) - F(E[i - 1])) * inv;
        cond[i] = -1.0 * T[i] * d;
        sum += cond[i] * dE;
    }
    *total = sum;
    return NEGF_OK;
}

extern "C" int negf_write_transport_char(const char* filename, int nE, const double* E, const double mu[2],
                                         const double kT[2], const double* dos, const double* trans,
                                         const double* cond)
{
    negf::clear_error();
    if (!filename || nE <= 0 || !E || !mu || !kT || !dos || !trans || !cond) {
        negf::set_error("invalid writer arguments");
        return NEGF_EARG;
    }
    FILE* f = std::fopen(filename, "w");
    if (!f) {
        negf::set_error("cannot open %s for writing", filename);
        return NEGF_EARG;
    }
    std::fprintf(f, " E_r, Fermi_Contact1, Fermi_Contact2, DOS_r, Transmission_r, Conductance_r \n");
    for (int i = 0; i < nE; ++i) {
        // FermiFunction (:1790-1795) of the real energy: 1 / (exp((E - mu)/kT) + 1)
        const double f1 = 1.0 / (std::exp((E[i] - mu[0]) / kT[0]) + 1.0);
        const double f2 = 1.0 / (std::exp((E[i] - mu[1]) / kT[1]) + 1.0);
        std::fprintf(f, "%20g%20g%20g%35g%35g%35g\n", E[i], f1, f2, dos[i], trans[i], cond[i]);
    }
    std::fclose(f);
    return NEGF_OK;
}

extern "C" int negf_write_ldos(const char* folder, int nE, const double* E, int N, const double* ptd_nm,
                               const double* ldos)
{
    negf::clear_error();
    if (!folder || nE <= 0 || N <= 0 || !E || !ptd_nm || !ldos) {
        negf::set_error("invalid writer arguments");
        return NEGF_EARG;
    }
    for (int e = 0; e < nE; ++e) {
        char head[128];
        std::snprintf(head, sizeof(head), "PTD LDOS_r at E=%f", E[e]);
        const std::string name = std::string(folder) + "/Ept_" + std::to_string(e) + ".dat";
        const int rc = write_table(name, head, N, ptd_nm, ldos + (size_t)e * N);
        if (rc) return rc;
    }
    return NEGF_OK;
}
