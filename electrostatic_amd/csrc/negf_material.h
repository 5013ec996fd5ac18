// negf_material.h -- host-side state of one NEGF tube (Level-2 material).
#pragma once
#include <complex>
#include <vector>

#include "negf_internal.h"

namespace negf {

// Energy limits of Define_EnergyLimits (NEGF_Common.cpp:1687-1787).
struct Limits {
    double kT_contact[2] = {0, 0}, kT_min = 0, kT_max = 0, mu_min = 0, mu_max = 0;
    bool noneq = false;
    std::complex<double> E_zPlus, E_left, E_right, E_rightmost, E_zeta, E_eta;
    int num_poles_formula = 0;
    std::vector<std::complex<double>> poles;   // enclosed poles (equilibrium only)
};

// Concatenated Gauss-Legendre paths.
struct Path {
    std::vector<std::complex<double>> E, mul;
    std::vector<double> w;
    std::vector<int> sizes;
};

void compute_limits(const negf_tube_params& p, const double mu[2], const double T[2], Limits& L);
// Gauss-Legendre path from a to b (type 0 line, 1 arc) appended to out.
void add_path(Path& out, std::complex<double> a, std::complex<double> b, int n, int type);
// FermiFunction (NEGF_Common.cpp:1788-1794) at complex E - mu.
std::complex<double> fermi(std::complex<double> E_minus_mu, double kT);
int fill_device(double* p, long long n, double v, hipStream_t st);

class Tube {
   public:
    ~Tube();
    int init(const negf_tube_params& p);
    int set_potential_host(const double* U, const double Uc[2]);
    // Copy a device potential; the contact potentials become U[0], U[N-1].
    int set_potential_device(const double* dU, hipStream_t st);
    void set_contacts(const double mu[2], const double T[2]);
    // Host part of a solve: limits and paths (plans rebuilt if needed).
    int prepare(int iter);
    // Enqueue one solve; Q (device, N) is ready on stream st when it returns.
    int solve(int iter, double* dQ, hipStream_t st);
    // The launches of a solve whose prepare() has run, all on stream st and
    // nothing on the host (capturable).  combine = false leaves out the last
    // two steps, the equilibrium plan's row sum into d_eq (eq_rows) and
    // Q = (eq + noneq) - rho0 (d_eq, d_neq, d_rho0), for a caller that fuses
    // both into its next kernel; dQ is then unused.
    int enqueue_solve(double* dQ, hipStream_t st, bool combine = true);
    bool eq_rows(const double** part, int* nrows) const { return eq_plan.single_tube_rows(part, nrows); }
    // Device status words of the plans enqueue_solve() launches (nullptr for
    // a plan it does not launch); nonzero after a launch: check_status().
    const int* eq_status() const { return eq_plan.status_ptr(); }
    const int* neq_status() const { return L.noneq && n_neq > 0 ? neq_plan.status_ptr() : nullptr; }
    // Synchronise; record rootless surface GFs of the last solves.
    int check_status();
    // check_status, then NEGF_EROOT if any were recorded since the last call.
    int take_root_flag();
    int total_integration_pts() const;
    int dos(int nE, const double* E_real, double* dos_out, double* trans_out);
    int dos_complex(int nE, const std::complex<double>* E, double* dos_out, double* trans_out, double* ldos);
    int dos_path(int which, std::vector<std::complex<double>>& E) const;
    // Terminal currents (A) over the current non-equilibrium path.
    int current(double I[2]);

    negf_tube_params p{};
    int N = 0;
    std::vector<int> mode_id, degen;
    std::vector<double> beta, ptd;
    int atoms_per_unitcell = 0;
    double spin_degen = 2.0;
    double mu[2] = {0, 0}, T[2] = {298, 298};
    Limits L;
    bool ec_updated = false;
    int root_flag = 0;
    // non-equilibrium path layout (may be adapted)
    int np = 1;
    std::vector<int> pts;
    std::vector<double> intercuts, density;
    double E_at_max = 0.0;
    Path neq_path;
    std::vector<double> integrand;   // [3][n_neq_E] after an adaptive update
    std::vector<ChainDesc> eq_host, neq_host;   // chain lists of the current paths
    long long paths_version = 0;                // bumped at every path rebuild
    int n_eq = 0, n_eq_pts = 0, n_neq = 0, n_neq_E = 0;

    hipStream_t stream = nullptr;
    // work enqueued on stream by solve() that a batch solve must still wait for
    bool stream_pending = false;
    std::vector<double> U_host;
    double Uc_host[2] = {0, 0};
    DevBuf d_U, d_Uc, d_zero, d_rho0, d_eq, d_neq, d_Q;
    // d_neq holds zeros (filled, and no non-equilibrium launch since)
    bool neq_zero = false;
    DevBuf eq_chains, neq_chains, d_integ, dos_chains, dos_pairs;
    HostBuf h_stage;
    RgfPlan eq_plan, neq_plan;

    // RGF tube descriptor of this tube; seg_len < 0: the tube's own setting
    RgfTube rgf_tube(const double* dU, const double* dUc, int nchain, int seg_len = -1) const;

    DecParams dec() const;

   private:
    int upload_chains(DevBuf& buf, const std::vector<ChainDesc>& ch);
    void eq_chains_of(const Path& path, double cm, bool with_fermi, std::vector<ChainDesc>& out) const;
    int compute_rho0();
    void noneq_path(Path& out) const;
    int build_noneq_plan();
    void adapt_intercuts();
    int find_integrand_max();
    int update_paths(bool correct);
};

}  // namespace negf

// Opaque handle of the C ABI.
struct negf_tube {
    negf::Tube t;
};
