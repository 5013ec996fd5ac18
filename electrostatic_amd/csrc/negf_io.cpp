// negf_io.cpp -- per-step output files in the reference's column formats.
//
//   <prefix>_Qout.dat  Write_InducedCharge -> Write_Table1D
//                      (NEGF_Common.cpp:3092-3100, :4223-4254): a header line,
//                      then "%35.15g%35.15g" (setprecision(15), setw(35)).
//   <prefix>_norm.dat  Write_ChargeNorm (:3102-3110), same table format.
//   <prefix>_U.dat     Write_PotentialAtSites (:3043-3078): no header,
//                      "l, setw(35) PTD, setw(35) U" at the stream's default
//                      precision (6 significant digits).
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
