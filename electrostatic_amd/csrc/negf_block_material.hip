// negf_block_material.hip -- equilibrium charge and self-consistent loop of
// a dense-block chain (blocks of 32 orbitals, negf_block_rgf) with the
// tube material's energy paths:
//   RhoEq  Compute_RhoEq (NEGF_Common.cpp:3667-3754) over the contour of
//          Define_IntegrationPaths (:1796-1800) with the Fermi factor of
//          Define_EnergyLimits (:1687-1787);
//   poles  Compute_GR_atPoles (:3833-3952), enclosed in equilibrium;
//   Rho0   Compute_Rho0 (:3957-4092) at mu = 0, U = 0;
//   Q      RhoEq - Rho0 per orbital (the charge combine of Solve_NEGF).
// The reference's 2-D MatrixBlock algebra that this path would need is
// stubbed (Matrix_Block_Util.H:123-135, 431-438, 684-692), so the block
// material has no reference counterpart beyond these formulas.  Only
// equilibrium contacts are supported: the non-equilibrium charge needs the
// first and last block columns of G, which the block sweep does not form.
//
// The self-consistent loop mirrors the tube's (c_TransportSolver::Solve,
// Transport.cpp:359-539) with a caller-supplied linear electrostatic
// stand-in U_i = U_ext_i - V_gate - sum_j K_ij n_in_j over the orbitals and
// Broyden modified-second on the orbital charge; the lead potentials are
// those of one given orbital of each end block (the tube takes the end
// sites' potential, Nanostructure.cpp:494-504) or the end blocks' means.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "negf_material.h"

namespace negf {

using cplx = std::complex<double>;
static constexpr double kPiB = 3.14159265358979323846;

struct BlockMaterial {
    negf_block_chain* c = nullptr;  // not owned
    negf_tube_params p{};
    int m = 0, nb = 0, S = 0;
    double spin = 2.0, sr_tol = 1e-12;
    int sr_max_iter = 200;
    double mu[2] = {0, 0}, T[2] = {298, 298};
    Limits L;
    std::vector<double> v, rho0, n_prev;
    double Uc[2] = {0, 0};
    int n_pts = 0;
    bool have_prev = false;

    int set_potential(const double* vv, const double uc[2])
    {
        std::copy(vv, vv + S, v.begin());
        Uc[0] = uc[0];
        Uc[1] = uc[1];
        return negf_block_chain_set_potential(c, vv, uc);
    }

    // Energies and block coefficients (the sweep accumulates Im(coef G)).
    void eq_energies(std::vector<cplx>& E, std::vector<cplx>& coef, int* n_path) const
    {
        Path eq;
        add_path(eq, L.E_right, L.E_zeta, p.eq_integration_pts[0], 0);
        add_path(eq, L.E_zeta, L.E_eta, p.eq_integration_pts[1], 0);
        add_path(eq, L.E_eta, L.E_left, p.eq_integration_pts[2], 1);
        const double cm = -spin / kPiB;
        for (size_t e = 0; e < eq.E.size(); ++e) {
            E.push_back(eq.E[e]);
            coef.push_back(cm * eq.w[e] * eq.mul[e] * fermi(eq.E[e] - L.mu_min, L.kT_min));
        }
        *n_path = (int)eq.E.size();
        // Im(-2i kT_min s G) at each enclosed pole
        for (const cplx& Ep : L.poles) {
            E.push_back(Ep);
            coef.push_back(cplx(0.0, -2.0 * L.kT_min * spin));
        }
    }

    int run(const std::vector<cplx>& E, const std::vector<cplx>& coef, double* acc)
    {
        return negf_block_rgf(c, (int)E.size(), reinterpret_cast<const negf_complex*>(E.data()),
                              reinterpret_cast<const negf_complex*>(coef.data()), sr_tol, sr_max_iter, acc, nullptr,
                              nullptr);
    }

    int compute_rho0()
    {
        Limits L0;
        const double mu0[2] = {0.0, 0.0};
        compute_limits(p, mu0, T, L0);
        Path path;
        add_path(path, L0.E_zPlus, L0.E_zeta, 30, 0);
        add_path(path, L0.E_zeta, L0.E_eta, 30, 0);
        add_path(path, L0.E_eta, L0.E_left, 30, 1);
        std::vector<cplx> E(path.E), coef(path.E.size());
        const double cm = -spin / kPiB;
        for (size_t e = 0; e < E.size(); ++e) coef[e] = cm * path.w[e] * path.mul[e];
        const std::vector<double> zero(S, 0.0);
        const double uc0[2] = {0.0, 0.0};
        const std::vector<double> v_keep(v);
        const double uc_keep[2] = {Uc[0], Uc[1]};
        int rc;
        if ((rc = negf_block_chain_set_potential(c, zero.data(), uc0))) return rc;
        rc = run(E, coef, rho0.data());
        const int rc2 = negf_block_chain_set_potential(c, v_keep.data(), uc_keep);
        return rc ? rc : rc2;
    }

    int limits()
    {
        compute_limits(p, mu, T, L);
        if (L.noneq) {
            set_error("block material: contacts out of equilibrium (|mu_L - mu_R| = %g eV) need the first and "
                      "last block columns of G, which the block sweep does not form",
                      std::fabs(mu[0] - mu[1]));
            return NEGF_ENOTSUP;
        }
        return NEGF_OK;
    }

    int solve(double* Q)
    {
        std::vector<cplx> E, coef;
        eq_energies(E, coef, &n_pts);
        int rc;
        if ((rc = run(E, coef, Q))) return rc;
        for (int i = 0; i < S; ++i) Q[i] -= rho0[i];
        return NEGF_OK;
    }
};

}  // namespace negf

using namespace negf;

struct negf_block_material {
    BlockMaterial b;
};

extern "C" {

int negf_block_material_create(negf_block_chain* c, const negf_tube_params* p, double spin_degen, double sr_tol,
                               int sr_max_iter, negf_block_material** out)
{
    clear_error();
    if (!c || !p || !out || !(spin_degen > 0) || !(sr_tol > 0) || sr_max_iter < 1) {
        set_error("invalid block material arguments");
        return NEGF_EARG;
    }
    negf_block_material* h = new negf_block_material;
    BlockMaterial& b = h->b;
    b.c = c;
    b.p = *p;
    b.spin = spin_degen;
    b.sr_tol = sr_tol;
    b.sr_max_iter = sr_max_iter;
    int rc = negf_block_chain_dims(c, &b.m, &b.nb);
    if (!rc) {
        b.S = b.m * b.nb;
        b.v.assign(b.S, 0.0);
        b.rho0.assign(b.S, 0.0);
        b.mu[0] = p->contact_mu[0];
        b.mu[1] = p->contact_mu[1];
        b.T[0] = p->contact_T[0];
        b.T[1] = p->contact_T[1];
        rc = b.limits();
    }
    if (!rc) {
        const double uc0[2] = {0.0, 0.0};
        rc = b.set_potential(b.v.data(), uc0);
    }
    if (!rc) rc = b.compute_rho0();
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_block_material_destroy(negf_block_material* h)
{
    if (!h) return NEGF_EARG;
    delete h;
    return NEGF_OK;
}

int negf_block_material_set_potential(negf_block_material* h, const double* v, const double Uc[2])
{
    clear_error();
    if (!h || !v || !Uc) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return h->b.set_potential(v, Uc);
}

int negf_block_material_set_contacts(negf_block_material* h, const double mu[2], const double T[2])
{
    clear_error();
    if (!h || !mu || !T) {
        set_error("null argument");
        return NEGF_EARG;
    }
    BlockMaterial& b = h->b;
    const double mu_keep[2] = {b.mu[0], b.mu[1]}, T_keep[2] = {b.T[0], b.T[1]};
    b.mu[0] = mu[0];
    b.mu[1] = mu[1];
    b.T[0] = T[0];
    b.T[1] = T[1];
    const int rc = b.limits();
    if (rc) {  // keep the previous (equilibrium) contacts
        std::copy(mu_keep, mu_keep + 2, b.mu);
        std::copy(T_keep, T_keep + 2, b.T);
        b.limits();
    }
    return rc;
}

int negf_block_material_rho0(const negf_block_material* h, double* rho0)
{
    clear_error();
    if (!h || !rho0) {
        set_error("null argument");
        return NEGF_EARG;
    }
    std::copy(h->b.rho0.begin(), h->b.rho0.end(), rho0);
    return NEGF_OK;
}

int negf_block_solve(negf_block_material* h, double* Q, int* n_intg_pts)
{
    clear_error();
    if (!h || !Q) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const int rc = h->b.solve(Q);
    if (n_intg_pts) *n_intg_pts = h->b.n_pts;
    return rc;
}

int negf_block_scf_bias_point(negf_block_material* h, const double* K, const double* U_ext, double V_gate,
                              const int* contact_orbital, const negf_scf_params* sp, int max_iters, double* U_out,
                              double* Q_out, int* iters, double* final_norm)
{
    clear_error();
    if (!h || !K || !sp || !U_out || !Q_out || max_iters < 1) {
        set_error("invalid block SCF arguments");
        return NEGF_EARG;
    }
    BlockMaterial& b = h->b;
    const int S = b.S, m = b.m, nb = b.nb;
    if (contact_orbital && (contact_orbital[0] < 0 || contact_orbital[0] >= m || contact_orbital[1] < 0 ||
                            contact_orbital[1] >= m)) {
        set_error("contact orbitals must be in [0, %d)", m);
        return NEGF_EARG;
    }
    negf_broyden* br = nullptr;
    int rc = negf_broyden_create(S, sp->Broyden_fraction, sp->Broyden_threshold_maxstep, sp->Broyden_norm_type, &br);
    if (rc) return rc;
    std::vector<double> n_in(S, sp->NS_initial_deposit_value), U(S), Q(S);
    if (sp->reset_with_previous_charge_distribution && b.have_prev) n_in = b.n_prev;
    rc = negf_broyden_set_nin(br, n_in.data());
    double norm = 0.0;
    int it = 0;
    bool done = false;
    while (!rc && it < max_iters) {
        // U = U_ext - V_gate - K n_in; the leads take the mean potential of the end blocks
        for (int i = 0; i < S; ++i) {
            double s = 0.0;
            const double* Ki = K + (size_t)i * S;
            for (int j = 0; j < S; ++j) s += Ki[j] * n_in[j];
            U[i] = (U_ext ? U_ext[i] : 0.0) - V_gate - s;
        }
        double uc[2] = {0.0, 0.0};
        if (contact_orbital) {
            uc[0] = U[contact_orbital[0]];
            uc[1] = U[(size_t)(nb - 1) * m + contact_orbital[1]];
        } else {
            for (int o = 0; o < m; ++o) {
                uc[0] += U[o] / m;
                uc[1] += U[(size_t)(nb - 1) * m + o] / m;
            }
        }
        if ((rc = b.set_potential(U.data(), uc))) break;
        if ((rc = b.solve(Q.data()))) break;
        ++it;
        double l2 = 0.0;
        if ((rc = negf_broyden_step(br, Q.data(), &norm, &l2, nullptr))) break;
        if ((rc = negf_broyden_get_nin(br, n_in.data()))) break;
        if (!std::isfinite(norm)) {
            set_error("block SCF: non-finite Broyden norm");
            rc = NEGF_ENAN;
            break;
        }
        if (norm <= sp->Broyden_max_norm) {
            done = true;
            break;
        }
    }
    negf_broyden_destroy(br);
    std::copy(U.begin(), U.end(), U_out);
    std::copy(Q.begin(), Q.end(), Q_out);
    if (iters) *iters = it;
    if (final_norm) *final_norm = norm;
    if (rc) return rc;
    b.n_prev = n_in;
    b.have_prev = true;
    if (!done) {
        set_error("block SCF: Broyden norm %g above %g after %d iterations", norm, sp->Broyden_max_norm, it);
        return NEGF_ENONCONV;
    }
    return NEGF_OK;
}

}  // extern "C"
