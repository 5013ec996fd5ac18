// scf_caller.cpp -- a compiled C++ caller of include/negf_amd.h (test
// infrastructure).  It drives one (17,0) tube through a metal-voltage sweep
// in the order of the reference's transport driver
// (c_TransportSolver::Solve, Transport.cpp:359-539): electrostatics, then
// per tube Solve_NEGF (set the potential, solve into the device charge
// vector), then the Broyden update on the device, until Broyden_Norm <=
// Broyden_max_norm; the step files are written at the end of each bias
// point and Broyden is reset for the next one (Transport.cpp:713-717).  The
// electrostatic solve is the library's dense 1/r-with-image stand-in for the
// MLMG solve and gather: U = U_ext - V_gate - K n_in.
//
// usage: scf_caller <U_ext.bin> <out.bin> <step-file dir>
//   U_ext.bin: N doubles; out.bin receives, per bias point, N doubles of U,
//   then the iteration count and the final norm (as doubles).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "negf_amd.h"

static void check(int rc, const char* what)
{
    if (rc != NEGF_OK) {
        std::fprintf(stderr, "%s failed: %d %s\n", what, rc, negf_last_error());
        std::exit(2);
    }
}

static void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s failed: %s\n", what, hipGetErrorString(e));
        std::exit(3);
    }
}

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s U_ext.bin out.bin step_dir\n", argv[0]);
        return 1;
    }
    // Initialize_NEGF: the tube of the N = 400 small SCF workload
    negf_tube_params p;
    negf_tube_params_default(&p);
    p.num_unitcells = 100;
    p.num_modes = 1;
    p.E_valence_min = -10.0;
    p.E_pole_max = 3.0;
    p.E_zPlus_imag = 1e-5;
    p.eq_integration_pts[0] = 30;
    p.eq_integration_pts[1] = 30;
    p.eq_integration_pts[2] = 64;
    p.num_noneq_paths = 1;
    p.noneq_integration_pts[0] = 200;
    const double mu[2] = {-1.0, -1.0}, T[2] = {298.0, 298.0};
    negf_tube* tube = nullptr;
    check(negf_tube_create(&p, &tube), "negf_tube_create");
    check(negf_tube_set_contacts(tube, mu, T), "negf_tube_set_contacts");
    const int N = negf_tube_num_sites(tube);
    std::vector<double> ptd(N);
    check(negf_tube_positions(tube, ptd.data()), "negf_tube_positions");

    std::vector<double> U_ext(N);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(U_ext.data(), sizeof(double), N, f) != (size_t)N) {
        std::fprintf(stderr, "cannot read %d doubles from %s\n", N, argv[1]);
        return 1;
    }
    std::fclose(f);

    // electrostatic stand-in: the library's dense kernel K
    negf_scf_params sp;
    sp.eps_r = 3.9;
    sp.R_ring_nm = 0.665;
    sp.R_image_nm = 10.0;
    sp.Broyden_fraction = 0.1;
    sp.Broyden_max_norm = 1e-6;
    sp.Broyden_norm_type = 0;
    sp.Broyden_threshold_maxstep = 200;
    sp.reset_with_previous_charge_distribution = 1;
    sp.NS_initial_deposit_value = 0.0;
    negf_scf* scf = nullptr;
    check(negf_scf_create(tube, &sp, &scf), "negf_scf_create");
    std::vector<double> K((size_t)N * N);
    check(negf_scf_kernel(scf, K.data()), "negf_scf_kernel");
    check(negf_scf_destroy(scf), "negf_scf_destroy");

    // Set_Broyden_Parallel (Broyden_Parallel_General.cpp:68-248)
    negf_broyden* br = nullptr;
    check(negf_broyden_create(N, sp.Broyden_fraction, sp.Broyden_threshold_maxstep, sp.Broyden_norm_type, &br),
          "negf_broyden_create");

    hipStream_t st;
    hip_check(hipStreamCreate(&st), "hipStreamCreate");
    double *d_U = nullptr, *d_Q = nullptr;
    hip_check(hipMalloc((void**)&d_U, sizeof(double) * N), "hipMalloc");
    hip_check(hipMalloc((void**)&d_Q, sizeof(double) * N), "hipMalloc");
    std::vector<double> U(N), Q(N), n_in(N), norm_site(N, 0.0);

    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 1;
    for (int b = 0; b <= 10; ++b) {
        const double V = 0.1 * b;
        double max_norm = 1.0, l2 = 0.0;
        int it = 0;
        do {
            if (negf_broyden_get_step(br) > sp.Broyden_threshold_maxstep) {
                std::fprintf(stderr, "Broyden_Step exceeded Broyden_threshold_maxstep\n");
                return 4;
            }
            // electrostatics and gather: U = U_ext - V - K n_in
            check(negf_broyden_get_nin(br, n_in.data()), "negf_broyden_get_nin");
            for (int i = 0; i < N; ++i) {
                double s = 0.0;
                for (int j = 0; j < N; ++j) s += K[(size_t)i * N + j] * n_in[j];
                U[i] = U_ext[i] - V - s;
            }
            hip_check(hipMemcpyAsync(d_U, U.data(), sizeof(double) * N, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
            // Solve_NEGF (NEGF_Common.cpp:1309-1340)
            check(negf_tube_set_potential_device(tube, d_U, st), "negf_tube_set_potential_device");
            int npts = 0;
            check(negf_solve_device(tube, it, d_Q, &npts, st), "negf_solve_device");
            int rs = negf_tube_status(tube);
            if (rs != NEGF_OK && rs != NEGF_EROOT) check(rs, "negf_tube_status");
            // Perform_SelfConsistencyAlgorithm (Transport.cpp:596-622)
            check(negf_broyden_step_device(br, d_Q, &max_norm, &l2, st), "negf_broyden_step_device");
            ++it;
        } while (max_norm > sp.Broyden_max_norm);
        // Write_Data (NEGF_Common.cpp:1408-1418) at the end of the step
        hip_check(hipMemcpy(Q.data(), d_Q, sizeof(double) * N, hipMemcpyDeviceToHost), "hipMemcpy");
        char prefix[512];
        std::snprintf(prefix, sizeof(prefix), "%s/step%04d", argv[3], b + 1);
        check(negf_write_step_files(prefix, N, ptd.data(), Q.data(), U.data(), norm_site.data()),
              "negf_write_step_files");
        std::fwrite(U.data(), sizeof(double), N, out);
        const double tail[2] = {(double)it, max_norm};
        std::fwrite(tail, sizeof(double), 2, out);
        // Reset_ForNextBiasStep (Transport.cpp:713-717)
        check(negf_broyden_reset(br, sp.reset_with_previous_charge_distribution, sp.NS_initial_deposit_value),
              "negf_broyden_reset");
    }
    std::fclose(out);
    hip_check(hipFree(d_U), "hipFree");
    hip_check(hipFree(d_Q), "hipFree");
    hip_check(hipStreamDestroy(st), "hipStreamDestroy");
    negf_broyden_destroy(br);
    negf_tube_destroy(tube);
    std::printf("scf_caller: 11 bias points done\n");
    return 0;
}
