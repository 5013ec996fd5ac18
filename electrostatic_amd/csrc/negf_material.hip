// negf_material.hip -- Level-2 NEGF material: one zigzag CNT in mode space
// (c_NEGF_Common<T> + c_CNT, NEGF_Common.H:478-605, CNT.H:13-103).
//
// A tube owns its mode set, energy limits and integration paths, two RGF
// plans (equilibrium contour + poles, non-equilibrium real axis) whose chains
// are all (mode, energy) pairs, and the device arrays of the charge
// components.  A solve enqueues both plans and one combine kernel:
//   Q_n = Im RhoEq_n + Re RhoNonEq_n - Im Rho0_n      (NEGF_Common.cpp:2885-2893)
// The per-energy weights of the reference's accumulation kernels are folded
// into per-chain coefficients on the host:
//   RhoEq   (:3742-3754)  Im(cm G w mul f degen) = Re(-i cm w mul f degen * G)
//   poles   (:3937-3948)  Im(-2i kT s G degen)   = Re(-2 kT s degen * G)
//   RhoNonEq(:3377-3497)  Re(cm sum_k A_k f_k w mul degen)
//                         = sum_k (cm w degen Re(f_k mul)) Gamma_k |G_nk|^2
// (A_k = Gamma_k |G_nk|^2 is real; mul is real on a real-axis line).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "negf_internal.h"
#include "negf_material.h"

namespace negf {

using cplx = std::complex<double>;

static constexpr double kPi = 3.14159265358979323846;
static constexpr double kKbEvPerK = 8.617333262e-5;   // PhysConst::kb_eVperK

// ---------------------------------------------------------------- kernels
__global__ void combine_charge_kernel(int N, const double* __restrict__ eq, const double* __restrict__ neq,
                                      const double* __restrict__ rho0, double* __restrict__ Q)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const double nq = neq ? neq[n] : 0.0;
    Q[n] = (eq[n] + nq) - rho0[n];
}

__global__ void fill_value_kernel(double* __restrict__ p, long long n, double v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// U_contact = U at the end sites (Obtain_PotentialAtSites Nanostructure.cpp:494-504).
__global__ void end_sites_kernel(int N, const double* __restrict__ U, double* __restrict__ Uc)
{
    if (threadIdx.x == 0) {
        Uc[0] = U[0];
        Uc[1] = U[N - 1];
    }
}

int fill_device(double* p, long long n, double v, hipStream_t st)
{
    if (n <= 0) return NEGF_OK;
    hipLaunchKernelGGL(fill_value_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

// ---------------------------------------------------------------- CNT modes
// Get_Bandgap_Of_Mode (CNT.cpp:155-176): smallest of the six |3p - k| c.
static double mode_gap(int m, int n, double acc, double gamma, int p)
{
    const double R = acc * std::sqrt(3.0 * (m * m + n * n + m * n)) / (2.0 * kPi);
    const double c = acc * gamma / R;
    const int ks[3] = {2 * m + n, m + 2 * n, m - n};
    double g = 1e300;
    for (int k : ks) g = std::min(g, c * std::fabs(3.0 * p - k));
    return g;
}

// In-place Lomuto quicksort by key, last element as pivot, elements <= pivot
// moved left.  The reference sorts the mode gaps this way
// (CNT.cpp:178-225); the same partition order keeps its choice of
// representative among degenerate modes.
static void lomuto_sort(std::vector<double>& key, std::vector<int>& id, int lo, int hi)
{
    while (lo < hi) {
        const double pivot = key[hi];
        int j = lo;
        for (int i = lo; i <= hi; ++i)
            if (!(key[i] > pivot)) {
                std::swap(key[i], key[j]);
                std::swap(id[i], id[j]);
                ++j;
            }
        const int pos = j - 1;
        lomuto_sort(key, id, lo, pos - 1);
        lo = pos + 1;
    }
}

// Define_SortedModeVector (CNT.cpp:227-286): modes sorted by gap, modes whose
// gap is within 1e-6 of the first of a run counted as one degenerate mode.
static void select_modes(int m, int n, double acc, double gamma, std::vector<int>& mode_id,
                         std::vector<double>& gap, std::vector<int>& degen)
{
    const int atoms_per_ring = (n == 0) ? m : m - 1;
    std::vector<double> g(atoms_per_ring);
    std::vector<int> id(atoms_per_ring);
    for (int p = 1; p <= atoms_per_ring; ++p) {
        g[p - 1] = mode_gap(m, n, acc, gamma, p);
        id[p - 1] = p;
    }
    lomuto_sort(g, id, 0, atoms_per_ring - 1);
    mode_id.assign(1, id[0]);
    gap.assign(1, g[0]);
    degen.assign(1, 1);
    for (int p = 1; p < atoms_per_ring; ++p) {
        if (g[p] - gap.back() > 1e-6) {
            mode_id.push_back(id[p]);
            gap.push_back(g[p]);
            degen.push_back(1);
        } else {
            degen.back() += 1;
        }
    }
}

// get_beta (CNT.cpp:288-292): 2 gamma cos(pi J / m).
static double mode_beta(int J, int m, double gamma) { return 2.0 * gamma * std::cos(kPi * J / m); }

// Axial ring positions in nm (get_AtomPosition_ZigZag_CNT CNT.cpp:93-132,
// before the tube offset): rings at acc, 1.5 acc, 2.5 acc, 3 acc within each
// 3 acc unit cell, centred.
static void ring_positions(int num_unitcells, double acc, std::vector<double>& ptd)
{
    const double centre = -(num_unitcells / 2.0) * 3.0 * acc - acc / 2.0;
    static const double frac[4] = {1.0, 1.5, 2.5, 3.0};
    ptd.resize((size_t)num_unitcells * 4);
    for (size_t r = 0; r < ptd.size(); ++r) {
        const double y = frac[r % 4] * acc + 3.0 * acc * (double)(r / 4) + centre;
        ptd[r] = y / 1.0e-9;
    }
}

// ---------------------------------------------------------------- limits
// Define_EnergyLimits (NEGF_Common.cpp:1687-1787).
void compute_limits(const negf_tube_params& p, const double mu[2], const double T[2], Limits& L)
{
    for (int c = 0; c < 2; ++c) L.kT_contact[c] = kKbEvPerK * T[c];
    L.mu_min = std::min(mu[0], mu[1]);
    L.mu_max = std::max(mu[0], mu[1]);
    L.kT_min = std::min(L.kT_contact[0], L.kT_contact[1]);
    L.kT_max = std::max(L.kT_contact[0], L.kT_contact[1]);
    L.noneq = std::fabs(L.mu_min - L.mu_max) > 1e-8 || std::fabs(L.kT_min - L.kT_max) > 0.01;
    const double zi = p.E_zPlus_imag;
    L.E_zPlus = cplx(0.0, zi);
    L.E_left = cplx(p.E_valence_min, zi);
    L.E_rightmost = cplx(L.mu_max + p.Fermi_tail_factors[1] * L.kT_max, zi);
    L.num_poles_formula = (int)((p.E_pole_max - kPi * L.kT_max) / (2.0 * kPi * L.kT_max) + 1);
    const double zeta_im = 2 * L.num_poles_formula * kPi * L.kT_max;
    L.poles.clear();
    if (L.noneq) {
        L.E_right = cplx(L.mu_min - p.Fermi_tail_factors[0] * L.kT_max, zi);
        L.E_zeta = cplx(L.E_right.real(), zeta_im);
        L.E_eta = cplx(L.E_right.real() - 14 * L.kT_max, zeta_im);
    } else {
        L.E_right = L.E_rightmost;
        for (int q = 0; q < L.num_poles_formula; ++q)
            L.poles.push_back(cplx(L.mu_min, kPi * L.kT_max * (2 * q + 1)));
        L.E_zeta = cplx(L.E_right.real(), zeta_im);
        const double tmax = std::max(p.Fermi_tail_factors[0], p.Fermi_tail_factors[1]);
        L.E_eta = cplx(L.mu_min - tmax * L.kT_max, zeta_im);
    }
}

// FermiFunction (NEGF_Common.cpp:1788-1794): 1 / (exp((E - mu)/kT) + 1).
cplx fermi(cplx E_minus_mu, double kT) { return 1.0 / (std::exp(E_minus_mu / kT) + 1.0); }

void add_path(Path& out, cplx a, cplx b, int n, int type)
{
    if (n <= 0) return;
    const size_t k0 = out.E.size();
    out.E.resize(k0 + n);
    out.w.resize(k0 + n);
    out.mul.resize(k0 + n);
    gl_path(a.real(), a.imag(), b.real(), b.imag(), n, type, reinterpret_cast<double*>(out.E.data() + k0),
            out.w.data() + k0, reinterpret_cast<double*>(out.mul.data() + k0));
    out.sizes.push_back(n);
}

// ---------------------------------------------------------------- Tube
Tube::~Tube()
{
    if (stream) (void)hipStreamSynchronize(stream);
    DevBuf* bufs[] = {&d_U, &d_Uc, &d_zero, &d_rho0, &d_eq, &d_neq, &d_Q, &eq_chains, &neq_chains, &d_integ,
                      &dos_chains, &dos_pairs};
    for (DevBuf* b : bufs) b->release();
    for (HostBuf* b : {&h_stage}) b->release();
    if (stream) (void)hipStreamDestroy(stream);
}

int Tube::init(const negf_tube_params& prm)
{
    p = prm;
    if (p.type_id[0] <= 0 || p.type_id[1] != 0) {
        set_error("only zigzag (m, 0) tubes are supported (type_id = %d %d)", p.type_id[0], p.type_id[1]);
        return NEGF_ENOTSUP;
    }
    if (p.num_unitcells <= 0 || p.acc <= 0 || p.gamma <= 0) {
        set_error("invalid tube geometry (num_unitcells=%d, acc=%g, gamma=%g)", p.num_unitcells, p.acc, p.gamma);
        return NEGF_EARG;
    }
    if (p.num_noneq_paths < 1 || p.num_noneq_paths > 8) {
        set_error("num_noneq_paths must be in [1, 8]");
        return NEGF_EARG;
    }
    for (int i = 0; i < 3; ++i)
        if (p.eq_integration_pts[i] <= 0) {
            set_error("eq_integration_pts must be positive");
            return NEGF_EARG;
        }
    for (int i = 0; i < p.num_noneq_paths; ++i)
        if (p.noneq_integration_pts[i] < 0) {
            set_error("noneq_integration_pts must be non-negative");
            return NEGF_EARG;
        }
    if (p.flag_adaptive_integration_limits && p.integrand_correction_interval <= 0) {
        set_error("integrand_correction_interval must be positive");
        return NEGF_EARG;
    }
    N = 4 * p.num_unitcells;   // rings_per_unitcell * num_unitcells (CNT.H:74)
    int P, Lseg;
    if (!choose_segments(N, p.segment_length, &P, &Lseg)) {
        set_error("tube of %d sites exceeds the supported maximum %d", N, max_sites());
        return NEGF_ENOTSUP;
    }
    std::vector<int> ids, dg;
    std::vector<double> gaps;
    select_modes(p.type_id[0], p.type_id[1], p.acc, p.gamma, ids, gaps, dg);
    const int nm = std::max(1, std::min<int>(p.num_modes, (int)ids.size()));
    mode_id.assign(ids.begin(), ids.begin() + nm);
    degen.assign(dg.begin(), dg.begin() + nm);
    beta.resize(nm);
    for (int m = 0; m < nm; ++m) beta[m] = mode_beta(mode_id[m], p.type_id[0], p.gamma);
    ring_positions(p.num_unitcells, p.acc, ptd);
    atoms_per_unitcell = 4 * p.type_id[0];
    spin_degen = 2.0;
    mu[0] = p.contact_mu[0];
    mu[1] = p.contact_mu[1];
    T[0] = p.contact_T[0];
    T[1] = p.contact_T[1];
    np = p.num_noneq_paths;
    pts.assign(p.noneq_integration_pts, p.noneq_integration_pts + np);
    intercuts.assign(np > 1 ? np - 1 : 0, 0.0);
    for (int c = 0; c + 1 < np; ++c) intercuts[c] = p.noneq_percent_intercuts[c];
    density.clear();
    // noneq_integration_pts_density is read only with adaptive limits
    // (NEGF_Common.cpp:527-539); when given it is used from the first correction
    if (p.flag_adaptive_integration_limits) {
        bool given = false;
        for (int c = 0; c < np; ++c) given |= p.noneq_integration_pts_density[c] != 0;
        if (given)
            for (int c = 0; c < np; ++c) density.push_back((double)p.noneq_integration_pts_density[c]);
    }

    NEGF_HIP_TRY(hipStreamCreate(&stream));
    int rc;
    if ((rc = d_U.ensure(sizeof(double) * N))) return rc;
    if ((rc = d_Uc.ensure(sizeof(double) * 2))) return rc;
    if ((rc = d_zero.ensure(sizeof(double) * std::max(N, 2)))) return rc;
    if ((rc = d_rho0.ensure(sizeof(double) * N))) return rc;
    if ((rc = d_eq.ensure(sizeof(double) * N))) return rc;
    if ((rc = d_neq.ensure(sizeof(double) * N))) return rc;
    if ((rc = d_Q.ensure(sizeof(double) * N))) return rc;
    if ((rc = fill_device(d_U.as<double>(), N, 0.0, stream))) return rc;
    if ((rc = fill_device(d_Uc.as<double>(), 2, 0.0, stream))) return rc;
    if ((rc = fill_device(d_zero.as<double>(), std::max(N, 2), 0.0, stream))) return rc;
    if ((rc = fill_device(d_neq.as<double>(), N, 0.0, stream))) return rc;
    neq_zero = true;
    U_host.assign(N, 0.0);
    Uc_host[0] = Uc_host[1] = 0.0;

    // Initialize_NEGF (NEGF_Common.cpp:1256-1306): limits, paths, Rho0.
    compute_limits(p, mu, T, L);
    if ((rc = compute_rho0())) return rc;
    if ((rc = update_paths(false))) return rc;
    ec_updated = false;
    return NEGF_OK;
}

RgfTube Tube::rgf_tube(const double* dU, const double* dUc, int nchain, int seg_len) const
{
    RgfTube tb;
    tb.N = N;
    tb.dU = dU;
    tb.dUc = dUc;
    tb.t_odd = p.gamma;
    tb.chain0 = 0;
    tb.nchain = nchain;
    tb.out_off = 0;
    tb.L = seg_len >= 0 ? seg_len : p.segment_length > 0 ? p.segment_length : lone_tube_segment_length(N);
    // else-if order of the reference (NEGF_Common.cpp:3469-3495)
    tb.sample[0] = N / 2;
    tb.sample[1] = N / 4;
    tb.sample[2] = N * 3 / 4;
    if (tb.sample[1] == tb.sample[0]) tb.sample[1] = -1;
    if (tb.sample[2] == tb.sample[0] || tb.sample[2] == tb.sample[1]) tb.sample[2] = -1;
    return tb;
}

DecParams Tube::dec() const
{
    DecParams d;
    d.use = p.use_decimation;
    d.layers = p.decimation_layers;
    d.max_iter = p.decimation_max_iter;
    d.rel_error = p.decimation_rel_error;
    return d;
}

int Tube::upload_chains(DevBuf& buf, const std::vector<ChainDesc>& ch)
{
    int rc = buf.ensure(sizeof(ChainDesc) * std::max<size_t>(1, ch.size()));
    if (rc) return rc;
    NEGF_HIP_TRY(hipStreamSynchronize(stream));
    NEGF_HIP_TRY(hipMemcpy(buf.p, ch.data(), sizeof(ChainDesc) * ch.size(), hipMemcpyHostToDevice));
    return NEGF_OK;
}

// Equilibrium-type chains over a path list: coef = -i cm w mul [f] degen.
void Tube::eq_chains_of(const Path& path, double cm, bool with_fermi, std::vector<ChainDesc>& out) const
{
    const int nm = (int)beta.size();
    const size_t nE = path.E.size();
    for (int m = 0; m < nm; ++m)
        for (size_t e = 0; e < nE; ++e) {
            const cplx E = path.E[e];
            cplx c = cm * path.w[e] * path.mul[e];
            if (with_fermi) c *= fermi(E - L.mu_min, L.kT_min);
            c *= (double)degen[m];
            const cplx coef = cplx(0.0, -1.0) * c;
            ChainDesc d;
            std::memset(&d, 0, sizeof(d));
            d.Er = E.real();
            d.Ei = E.imag();
            d.cr = coef.real();
            d.ci = coef.imag();
            d.t_even = beta[m];
            d.degen = degen[m];
            d.out_slot = -1;
            out.push_back(d);
        }
}

// Compute_Rho0 (NEGF_Common.cpp:3957-4092) over Define_IntegrationPaths
// (:1796-1800): paths E_zPlus -> E_zeta -> E_eta (lines, 30 points each) and
// the arc E_eta -> E_contour_left (30 points), limits at mu = 0, U = 0.
int Tube::compute_rho0()
{
    Limits L0;
    const double mu0[2] = {0.0, 0.0};
    compute_limits(p, mu0, T, L0);
    Path path;
    add_path(path, L0.E_zPlus, L0.E_zeta, 30, 0);
    add_path(path, L0.E_zeta, L0.E_eta, 30, 0);
    add_path(path, L0.E_eta, L0.E_left, 30, 1);
    std::vector<ChainDesc> ch;
    const double cm = -spin_degen / kPi;
    eq_chains_of(path, cm, false, ch);
    int rc;
    DevBuf buf;
    if ((rc = upload_chains(buf, ch))) return rc;
    RgfPlan plan;
    std::vector<RgfTube> tubes{rgf_tube(d_zero.as<double>(), d_zero.as<double>(), (int)ch.size())};
    if ((rc = plan.build(KIND_EQ, false, tubes, buf.as<ChainDesc>(), (int)ch.size(), dec()))) return rc;
    if ((rc = plan.enqueue(d_rho0.as<double>(), nullptr, nullptr, stream))) return rc;
    int status = 0;
    if ((rc = plan.check_status(stream, &status))) return rc;
    root_flag |= status & 1;
    return NEGF_OK;
}

// Generate_NonEq_Paths (NEGF_Common.cpp:2018-2059).
void Tube::noneq_path(Path& out) const
{
    out = Path();
    const cplx len = L.E_rightmost - L.E_right;
    std::vector<cplx> lo(np), hi(np);
    lo[0] = L.E_right;
    hi[np - 1] = L.E_rightmost;
    for (int q = 0; q + 1 < np; ++q) hi[q] = L.E_right + (intercuts[q] / 100.0) * len;
    for (int q = 1; q < np; ++q) lo[q] = hi[q - 1];
    for (int q = 0; q < np; ++q) {
        if (pts[q] > 0)
            add_path(out, lo[q], hi[q], pts[q], 0);
        else
            out.sizes.push_back(0);
    }
}

int Tube::build_noneq_plan()
{
    noneq_path(neq_path);
    std::vector<ChainDesc> ch;
    const int nm = (int)beta.size();
    const size_t nE = neq_path.E.size();
    n_neq_E = (int)nE;
    if (nE == 0) {
        n_neq = 0;
        neq_host.clear();
        return NEGF_OK;
    }
    const double cm = -spin_degen / (2.0 * kPi);
    for (int m = 0; m < nm; ++m)
        for (size_t e = 0; e < nE; ++e) {
            const cplx E = neq_path.E[e];
            const cplx fL = fermi(E - mu[0], L.kT_contact[0]);
            const cplx fR = fermi(E - mu[1], L.kT_contact[1]);
            const double w = neq_path.w[e];
            const cplx mul = neq_path.mul[e];
            ChainDesc d;
            std::memset(&d, 0, sizeof(d));
            d.Er = E.real();
            d.Ei = E.imag();
            d.cL = cm * w * degen[m] * (fL * mul).real();
            d.cR = cm * w * degen[m] * (fR * mul).real();
            d.iL = cm * degen[m] * fL.real();
            d.iR = cm * degen[m] * fR.real();
            d.t_even = beta[m];
            d.degen = degen[m];
            d.out_slot = (int)(m * nE + e);
            ch.push_back(d);
        }
    int rc;
    if ((rc = upload_chains(neq_chains, ch))) return rc;
    n_neq = (int)ch.size();
    neq_host = ch;
    if ((rc = d_integ.ensure(sizeof(double) * 3 * n_neq))) return rc;
    std::vector<RgfTube> tubes{rgf_tube(d_U.as<double>(), d_Uc.as<double>(), n_neq)};
    return neq_plan.build(KIND_NONEQ, false, tubes, neq_chains.as<ChainDesc>(), n_neq, dec());
}

// Find_NonEq_Percent_Intercuts_Adaptively (NEGF_Common.cpp:2062-2185); the
// point densities per kT are fixed at the first call (or given as input).
// The reference keeps them in an int vector (noneq_integration_pts_density,
// NEGF_Common.H:404), so a density computed here is truncated toward zero
// when it is stored, as there.
void Tube::adapt_intercuts()
{
    const double kT0 = L.kT_contact[0];
    const double E_star = (E_at_max - mu[0]) / kT0;
    const double E_min = (L.E_right.real() - mu[0]) / kT0;
    const double E_max = (L.E_rightmost.real() - mu[0]) / kT0;
    const double span = E_max - E_min;
    auto even_round = [](double x) { return (int)(std::round(x / 2) * 2); };
    auto as_int = [](double x) { return (double)(int)x; };
    if (np == 3) {
        if (density.empty())
            density = {as_int(pts[0] / ((intercuts[0] / 100.) * span)),
                       as_int(pts[1] / (((intercuts[1] - intercuts[0]) / 100.) * span)),
                       as_int(pts[2] / (((100. - intercuts[1]) / 100.) * span))};
        const double E_lo = std::max(E_star - p.kT_window_around_singularity[0], E_min);
        const double E_hi = std::min(E_star + p.kT_window_around_singularity[1], E_max);
        intercuts[0] = 100 * (E_lo - E_min) / span;
        intercuts[1] = 100 * (E_hi - E_min) / span;
        pts[0] = even_round((E_lo - E_min) * density[0]);
        pts[1] = even_round((E_hi - E_lo) * density[1]);
        pts[2] = even_round((E_max - E_hi) * density[2]);
    } else if (np == 2) {
        if (density.empty())
            density = {as_int(pts[0] / ((intercuts[0] / 100.) * span)),
                       as_int(pts[1] / (((100 - intercuts[0]) / 100.) * span))};
        const double E_hi = E_star + p.kT_window_around_singularity[1];
        intercuts[0] = 100 * (E_hi - E_min) / span;
        pts[0] = even_round((E_hi - E_min) * density[0]);
        pts[1] = even_round((E_max - E_hi) * density[1]);
    }
}

// E of the largest |integrand| at site N/2 (first maximum in path order,
// NEGF_Common.cpp:3527-3545); E_at_max keeps its value if all are zero.
int Tube::find_integrand_max()
{
    if (n_neq == 0) return NEGF_OK;
    int rc;
    if ((rc = fill_device(d_integ.as<double>(), 3LL * n_neq, 0.0, stream))) return rc;
    neq_zero = false;
    if ((rc = neq_plan.enqueue(d_neq.as<double>(), nullptr, d_integ.as<double>(), stream))) return rc;
    std::vector<double> h(3 * (size_t)n_neq);
    NEGF_HIP_TRY(hipMemcpyAsync(h.data(), d_integ.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    NEGF_HIP_TRY(hipStreamSynchronize(stream));
    integrand.assign(3 * (size_t)n_neq_E, 0.0);
    const int nm = (int)beta.size();
    for (int w = 0; w < 3; ++w)
        for (int e = 0; e < n_neq_E; ++e) {
            double s = 0.0;
            for (int m = 0; m < nm; ++m) s += h[((size_t)m * n_neq_E + e) * 3 + w];
            integrand[(size_t)w * n_neq_E + e] = s;
        }
    double maxv = 0.0;
    for (int e = 0; e < n_neq_E; ++e)
        if (maxv < std::fabs(integrand[e])) {
            maxv = std::fabs(integrand[e]);
            E_at_max = neq_path.E[e].real();
        }
    return NEGF_OK;
}

// Update_IntegrationPaths (NEGF_Common.cpp:2188-2212).
int Tube::update_paths(bool correct)
{
    Path eq;
    add_path(eq, L.E_right, L.E_zeta, p.eq_integration_pts[0], 0);
    add_path(eq, L.E_zeta, L.E_eta, p.eq_integration_pts[1], 0);
    add_path(eq, L.E_eta, L.E_left, p.eq_integration_pts[2], 1);
    n_eq_pts = (int)eq.E.size();
    std::vector<ChainDesc> ch;
    eq_chains_of(eq, -spin_degen / kPi, true, ch);
    // poles (Compute_GR_atPoles :3833-3952, enclosed only in equilibrium)
    const int nm = (int)beta.size();
    for (int m = 0; m < nm; ++m)
        for (const cplx& E : L.poles) {
            // Im(-2i kT_min s degen G) = Re(-2 kT_min s degen * G)
            ChainDesc d;
            std::memset(&d, 0, sizeof(d));
            d.Er = E.real();
            d.Ei = E.imag();
            d.cr = -2.0 * L.kT_min * spin_degen * degen[m];
            d.ci = 0.0;
            d.t_even = beta[m];
            d.degen = degen[m];
            d.out_slot = -1;
            ch.push_back(d);
        }
    int rc;
    if ((rc = upload_chains(eq_chains, ch))) return rc;
    n_eq = (int)ch.size();
    eq_host = ch;
    ++paths_version;
    std::vector<RgfTube> tubes{rgf_tube(d_U.as<double>(), d_Uc.as<double>(), n_eq)};
    if ((rc = eq_plan.build(KIND_EQ, false, tubes, eq_chains.as<ChainDesc>(), n_eq, dec()))) return rc;
    if (L.noneq) {
        if ((rc = build_noneq_plan())) return rc;
        if (correct) {
            if ((rc = find_integrand_max())) return rc;
            adapt_intercuts();
            if ((rc = build_noneq_plan())) return rc;
        }
    } else {
        n_neq = 0;
        n_neq_E = 0;
        neq_host.clear();
    }
    return NEGF_OK;
}

int Tube::set_potential_host(const double* U, const double Uc[2])
{
    std::copy(U, U + N, U_host.begin());
    Uc_host[0] = Uc[0];
    Uc_host[1] = Uc[1];
    NEGF_HIP_TRY(hipStreamSynchronize(stream));
    NEGF_HIP_TRY(hipMemcpy(d_U.p, U, sizeof(double) * N, hipMemcpyHostToDevice));
    NEGF_HIP_TRY(hipMemcpy(d_Uc.p, Uc, sizeof(double) * 2, hipMemcpyHostToDevice));
    return NEGF_OK;
}

int Tube::set_potential_device(const double* dU, hipStream_t st)
{
    NEGF_HIP_TRY(hipMemcpyAsync(d_U.p, dU, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(end_sites_kernel, dim3(1), dim3(64), 0, st, N, d_U.as<double>(), d_Uc.as<double>());
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

void Tube::set_contacts(const double m_[2], const double T_[2])
{
    mu[0] = m_[0];
    mu[1] = m_[1];
    T[0] = T_[0];
    T[1] = T_[1];
    ec_updated = true;
}

// Solve_NEGF (NEGF_Common.cpp:1309-1340) + Compute_InducedCharge (:2850-2900).
// The host part of Solve_NEGF (:1314-1337): limits and paths when the
// contacts changed or an adaptive correction falls due at this iteration.
int Tube::prepare(int iter)
{
    const bool correct = p.flag_adaptive_integration_limits && (iter % p.integrand_correction_interval == 0);
    if (ec_updated) compute_limits(p, mu, T, L);
    if (ec_updated || correct) {
        int rc = update_paths(correct);
        if (rc) return rc;
        ec_updated = false;
    }
    return NEGF_OK;
}

int Tube::solve(int iter, double* dQ, hipStream_t st)
{
    int rc;
    if (st != stream) {
        // order the tube's own stream after the caller's work (potential upload)
        hipEvent_t ev;
        NEGF_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        NEGF_HIP_TRY(hipEventRecord(ev, st));
        NEGF_HIP_TRY(hipStreamWaitEvent(stream, ev, 0));
        NEGF_HIP_TRY(hipEventDestroy(ev));
    }
    if ((rc = prepare(iter))) return rc;
    if ((rc = enqueue_solve(dQ, stream))) return rc;
    stream_pending = true;
    if (st != stream) {
        hipEvent_t ev;
        NEGF_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        NEGF_HIP_TRY(hipEventRecord(ev, stream));
        NEGF_HIP_TRY(hipStreamWaitEvent(st, ev, 0));
        NEGF_HIP_TRY(hipEventDestroy(ev));
    }
    return NEGF_OK;
}

int Tube::enqueue_solve(double* dQ, hipStream_t st, bool combine)
{
    int rc;
    if ((rc = eq_plan.enqueue(d_eq.as<double>(), nullptr, nullptr, st, nullptr, nullptr, combine))) return rc;
    const bool have_neq = L.noneq && n_neq > 0;
    if (have_neq) {
        neq_zero = false;
        if ((rc = neq_plan.enqueue(d_neq.as<double>(), nullptr, nullptr, st))) return rc;
    } else if (!neq_zero) {
        // zeros once after the last non-equilibrium solve (the charge
        // components read them)
        if ((rc = fill_device(d_neq.as<double>(), N, 0.0, st))) return rc;
        neq_zero = true;
    }
    if (!combine) return NEGF_OK;
    hipLaunchKernelGGL(combine_charge_kernel, dim3((N + 255) / 256), dim3(256), 0, st, N, d_eq.as<double>(),
                       d_neq.as<double>(), d_rho0.as<double>(), dQ);
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

int Tube::check_status()
{
    int s1 = 0, s2 = 0, rc;
    if ((rc = eq_plan.check_status(stream, &s1))) return rc;
    if (n_neq > 0 && (rc = neq_plan.check_status(stream, &s2))) return rc;
    root_flag |= (s1 | s2) & 1;
    return NEGF_OK;
}

int Tube::take_root_flag()
{
    int rc = check_status();
    if (rc) return rc;
    const int f = root_flag;
    root_flag = 0;
    if (f) {
        set_error("surface Green's function: no root with Im<0 at some energy (g = 0 used)");
        return NEGF_EROOT;
    }
    return NEGF_OK;
}

int Tube::total_integration_pts() const
{
    // get_Total_Integration_Pts (NEGF_Common.cpp:1826-1841)
    if (L.noneq) {
        int s = 0;
        for (int q : pts) s += q;
        return s;
    }
    return n_eq_pts;
}

// Compute_DensityOfStates (NEGF_Common.cpp:2640-2777) on real energies + E_zPlus:
// DOS = spin sum_m degen_m trace_m / (2 pi) / atoms_per_unitcell,
// T = sum_m degen_m Gamma_L Gamma_R |G_0,N-1|^2.
int Tube::dos(int nE, const double* E_real, double* dos_out, double* trans_out)
{
    std::vector<cplx> E(nE);
    for (int e = 0; e < nE; ++e) E[e] = cplx(E_real[e], p.E_zPlus_imag);
    return dos_complex(nE, E.data(), dos_out, trans_out, nullptr);
}

// Compute_DensityOfStates (NEGF_Common.cpp:2390-2847) at the energies E:
// per energy the trace of the spectral function and the transmission
// (DOS = spin_degen sum_n LDOS_n / atoms per unit cell, T = Gamma_L Gamma_R
// |G_0,N-1|^2, summed over the modes with their degeneracies), and
// optionally LDOS_n = sum_modes degen A_nn / (2 pi) per site ([nE][N],
// the values the reference writes to Ept_<e>.dat, :2778-2796).
int Tube::dos_complex(int nE, const cplx* E, double* dos_out, double* trans_out, double* ldos)
{
    const int nm = (int)beta.size();
    std::vector<ChainDesc> ch;
    for (int m = 0; m < nm; ++m)
        for (int e = 0; e < nE; ++e) {
            ChainDesc d;
            std::memset(&d, 0, sizeof(d));
            d.Er = E[e].real();
            d.Ei = E[e].imag();
            d.t_even = beta[m];
            d.degen = degen[m];
            d.out_slot = m * nE + e;
            ch.push_back(d);
        }
    int rc;
    if ((rc = upload_chains(dos_chains, ch))) return rc;
    if ((rc = dos_pairs.ensure(sizeof(double) * 2 * ch.size()))) return rc;
    DevBuf d_site;
    if (ldos && (rc = d_site.ensure(sizeof(double) * ch.size() * (size_t)N))) return rc;
    RgfPlan plan;
    std::vector<RgfTube> tubes{rgf_tube(d_U.as<double>(), d_Uc.as<double>(), (int)ch.size())};
    if ((rc = plan.build(KIND_DOS, false, tubes, dos_chains.as<ChainDesc>(), (int)ch.size(), dec()))) return rc;
    if ((rc = plan.enqueue(nullptr, nullptr, dos_pairs.as<double>(), stream, nullptr,
                           ldos ? d_site.as<double>() : nullptr)))
        return rc;
    std::vector<double> h(2 * ch.size());
    NEGF_HIP_TRY(hipMemcpyAsync(h.data(), dos_pairs.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    std::vector<double> hs;
    if (ldos) {
        hs.resize(ch.size() * (size_t)N);
        NEGF_HIP_TRY(hipMemcpyAsync(hs.data(), d_site.p, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, stream));
    }
    int status = 0;
    if ((rc = plan.check_status(stream, &status))) return rc;
    root_flag |= status & 1;
    for (int e = 0; e < nE; ++e) {
        double tr = 0.0, tt = 0.0;
        for (int m = 0; m < nm; ++m) {
            tr += degen[m] * h[2 * ((size_t)m * nE + e)] / (2.0 * kPi);
            tt += degen[m] * h[2 * ((size_t)m * nE + e) + 1];
        }
        dos_out[e] = spin_degen * tr / atoms_per_unitcell;
        trans_out[e] = tt;
        if (ldos)
            for (int n = 0; n < N; ++n) {
                double v = 0.0;
                for (int m = 0; m < nm; ++m) v += degen[m] * hs[((size_t)m * nE + e) * N + n] / (2.0 * kPi);
                ldos[(size_t)e * N + n] = v;
            }
    }
    return NEGF_OK;
}

// Energies of the DOS paths (NEGF_Common.cpp:1807-1815, 2210-2249):
// which = 0, the flat-band path of Initialize_NEGF: a Gauss-Legendre line
// from flatband_dos_integration_limits[0] + i eta to [1] + i eta with
// flatband_dos_integration_pts points; which = 1, the per-step path of
// Update_IntegrationPaths: the non-equilibrium path when the contacts differ,
// else a line from Re(E_eta) (imaginary part 0, as the reference passes the
// real part alone) to E_rightmost with flatband_dos_integration_pts points.
int Tube::dos_path(int which, std::vector<cplx>& E) const
{
    const int n = p.flatband_dos_integration_pts;
    E.clear();
    if (which == 1 && L.noneq) {
        E = neq_path.E;
        return NEGF_OK;
    }
    if (n <= 0) {
        set_error("flatband_dos_integration_pts must be positive");
        return NEGF_EARG;
    }
    std::vector<double> Ev(2 * n), w(n), mul(2 * n);
    if (which == 0)
        gl_path(p.flatband_dos_integration_limits[0], p.E_zPlus_imag, p.flatband_dos_integration_limits[1],
                p.E_zPlus_imag, n, 0, Ev.data(), w.data(), mul.data());
    else
        gl_path(L.E_eta.real(), 0.0, L.E_rightmost.real(), L.E_rightmost.imag(), n, 0, Ev.data(), w.data(),
                mul.data());
    for (int e = 0; e < n; ++e) E.push_back(cplx(Ev[2 * e], Ev[2 * e + 1]));
    return NEGF_OK;
}

// Compute_Current (NEGF_Common.cpp:4321-4760) on the non-equilibrium path.
// At a contact site the reference's IF_k = Gamma_k f_k A_kk - Gamma_k Gn_kk
// reduces to Gamma_k A^j_kk (f_k - f_j), j the other contact, and
// Gamma_L A^R_00 = Gamma_R A^L_{N-1,N-1} = T(E) = Gamma_L Gamma_R |G_{0,N-1}|^2:
//   I_0 = s q/h sum_E g w T(E) Re((f_L - f_R) mul),  I_1 = -I_0 (same sum with
// f_R - f_L).  T(E) per chain comes from the DOS launch over the same chains.
int Tube::current(double I[2])
{
    I[0] = I[1] = 0.0;
    if (!L.noneq || n_neq == 0) return NEGF_OK;
    int rc;
    if ((rc = dos_pairs.ensure(sizeof(double) * 2 * (size_t)n_neq))) return rc;
    RgfPlan plan;
    std::vector<RgfTube> tubes{rgf_tube(d_U.as<double>(), d_Uc.as<double>(), n_neq)};
    if ((rc = plan.build(KIND_DOS, false, tubes, neq_chains.as<ChainDesc>(), n_neq, dec()))) return rc;
    if ((rc = plan.enqueue(nullptr, nullptr, dos_pairs.as<double>(), stream))) return rc;
    std::vector<double> h(2 * (size_t)n_neq);
    NEGF_HIP_TRY(hipMemcpyAsync(h.data(), dos_pairs.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, stream));
    int status = 0;
    if ((rc = plan.check_status(stream, &status))) return rc;
    root_flag |= status & 1;
    const double q_e = 1.602176634e-19, h_eVperHz = 4.135667696e-15;   // PhysConst
    const double cm = spin_degen * q_e / h_eVperHz;
    const int nm = (int)beta.size();
    double s0 = 0.0, s1 = 0.0;
    for (int m = 0; m < nm; ++m)
        for (int e = 0; e < n_neq_E; ++e) {
            const cplx E = neq_path.E[e];
            const cplx fL = fermi(E - mu[0], L.kT_contact[0]);
            const cplx fR = fermi(E - mu[1], L.kT_contact[1]);
            const double Tm = h[2 * ((size_t)m * n_neq_E + e) + 1];
            const double base = cm * degen[m] * neq_path.w[e] * Tm;
            s0 += base * ((fL - fR) * neq_path.mul[e]).real();
            s1 += base * ((fR - fL) * neq_path.mul[e]).real();
        }
    I[0] = s0;
    I[1] = s1;
    return NEGF_OK;
}

}  // namespace negf

// ======================================================================== C ABI
using namespace negf;


extern "C" {

void negf_tube_params_default(negf_tube_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->type_id[0] = 17;
    p->type_id[1] = 0;
    p->acc = 1.42e-10;              // CNT.H:17
    p->gamma = 2.5;                 // CNT.H:18
    p->num_modes = 1;
    p->num_unitcells = 100;
    p->E_valence_min = -14.0;       // NEGF_Common.H:389
    p->E_pole_max = 10.0;           // NEGF_Common.H:390
    p->E_zPlus_imag = 1e-8;         // NEGF_Common.H:388
    p->Fermi_tail_factors[0] = 14.0;
    p->Fermi_tail_factors[1] = 14.0;
    p->eq_integration_pts[0] = 30;  // required input in the reference; the
    p->eq_integration_pts[1] = 30;  // value of its example inputs
    p->eq_integration_pts[2] = 30;
    p->num_noneq_paths = 1;
    p->noneq_integration_pts[0] = 100;
    p->contact_mu[0] = p->contact_mu[1] = 0.0;
    p->contact_T[0] = p->contact_T[1] = 298.0;
    p->use_decimation = 0;
    p->decimation_layers = 10;
    p->decimation_max_iter = 30;
    p->decimation_rel_error = 0;    // int 1.e-5 truncates to 0 (NEGF_Common.H:475)
    p->flag_adaptive_integration_limits = 0;
    p->integrand_correction_interval = 500;
    p->kT_window_around_singularity[0] = 1.0;
    p->kT_window_around_singularity[1] = 1.0;
    p->flatband_dos_integration_pts = 100;
    p->flatband_dos_integration_limits[0] = -1.0;  // NEGF_Common.H:417
    p->flatband_dos_integration_limits[1] = 1.0;
}

int negf_tube_create(const negf_tube_params* p, negf_tube** out)
{
    clear_error();
    if (!p || !out) {
        set_error("null argument");
        return NEGF_EARG;
    }
    negf_tube_params q = *p;
    if (q.num_noneq_paths > 1) {
        // uniform intercuts when none are given (Read_NonEqPathParams :478-491)
        bool given = false;
        for (int c = 0; c + 1 < q.num_noneq_paths; ++c) given |= q.noneq_percent_intercuts[c] != 0.0;
        if (!given)
            for (int c = 0; c + 1 < q.num_noneq_paths; ++c)
                q.noneq_percent_intercuts[c] = (c + 1) * (100.0 / q.num_noneq_paths);
    }
    negf_tube* h = new negf_tube;
    int rc = h->t.init(q);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_tube_destroy(negf_tube* t)
{
    if (!t) return NEGF_EARG;
    delete t;
    return NEGF_OK;
}

int negf_tube_num_sites(const negf_tube* t) { return t ? t->t.N : -1; }

int negf_tube_positions(const negf_tube* t, double* ptd_nm)
{
    clear_error();
    if (!t || !ptd_nm) {
        set_error("null argument");
        return NEGF_EARG;
    }
    std::copy(t->t.ptd.begin(), t->t.ptd.end(), ptd_nm);
    return NEGF_OK;
}

int negf_tube_modes(const negf_tube* t, int* mode_id, double* beta, int* degen)
{
    clear_error();
    if (!t) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const int nm = (int)t->t.beta.size();
    for (int m = 0; m < nm; ++m) {
        if (mode_id) mode_id[m] = t->t.mode_id[m];
        if (beta) beta[m] = t->t.beta[m];
        if (degen) degen[m] = t->t.degen[m];
    }
    return nm;
}

int negf_tube_set_potential(negf_tube* t, const double* U, const double Uc[2])
{
    clear_error();
    if (!t || !U || !Uc) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return t->t.set_potential_host(U, Uc);
}

int negf_tube_set_potential_device(negf_tube* t, const double* d_U, void* stream)
{
    clear_error();
    if (!t || !d_U) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return t->t.set_potential_device(d_U, (hipStream_t)stream);
}

int negf_tube_set_contacts(negf_tube* t, const double mu[2], const double T[2])
{
    clear_error();
    if (!t || !mu || !T || T[0] <= 0 || T[1] <= 0) {
        set_error("invalid contact arguments");
        return NEGF_EARG;
    }
    t->t.set_contacts(mu, T);
    return NEGF_OK;
}

int negf_compute_rho0(negf_tube* t, double* rho0_im)
{
    clear_error();
    if (!t || !rho0_im) {
        set_error("null argument");
        return NEGF_EARG;
    }
    NEGF_HIP_TRY(hipMemcpyAsync(rho0_im, t->t.d_rho0.p, sizeof(double) * t->t.N, hipMemcpyDeviceToHost,
                                t->t.stream));
    NEGF_HIP_TRY(hipStreamSynchronize(t->t.stream));
    return NEGF_OK;
}

int negf_solve_device(negf_tube* t, int iter, double* d_Q, int* n_intg_pts, void* stream)
{
    clear_error();
    if (!t || !d_Q) {
        set_error("null argument");
        return NEGF_EARG;
    }
    int rc = t->t.solve(iter, d_Q, (hipStream_t)stream);
    if (rc) return rc;
    if (n_intg_pts) *n_intg_pts = t->t.total_integration_pts();
    return NEGF_OK;
}

int negf_solve(negf_tube* t, int iter, double* Q, int* n_intg_pts)
{
    clear_error();
    if (!t || !Q) {
        set_error("null argument");
        return NEGF_EARG;
    }
    Tube& tb = t->t;
    int rc = tb.solve(iter, tb.d_Q.as<double>(), tb.stream);
    if (rc) return rc;
    NEGF_HIP_TRY(hipMemcpyAsync(Q, tb.d_Q.p, sizeof(double) * tb.N, hipMemcpyDeviceToHost, tb.stream));
    if ((rc = tb.check_status())) return rc;
    if (n_intg_pts) *n_intg_pts = tb.total_integration_pts();
    return NEGF_OK;
}

int negf_tube_status(negf_tube* t)
{
    clear_error();
    if (!t) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return t->t.take_root_flag();
}

int negf_get_charge_components(const negf_tube* t, double* eq_im, double* noneq_re, double* rho0_im)
{
    clear_error();
    if (!t) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const Tube& tb = t->t;
    NEGF_HIP_TRY(hipStreamSynchronize(tb.stream));
    const size_t b = sizeof(double) * tb.N;
    if (eq_im) NEGF_HIP_TRY(hipMemcpy(eq_im, tb.d_eq.p, b, hipMemcpyDeviceToHost));
    if (noneq_re) NEGF_HIP_TRY(hipMemcpy(noneq_re, tb.d_neq.p, b, hipMemcpyDeviceToHost));
    if (rho0_im) NEGF_HIP_TRY(hipMemcpy(rho0_im, tb.d_rho0.p, b, hipMemcpyDeviceToHost));
    return NEGF_OK;
}

int negf_tube_limits(const negf_tube* t, double* out)
{
    clear_error();
    if (!t || !out) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const Limits& L = t->t.L;
    out[0] = L.mu_min;
    out[1] = L.mu_max;
    out[2] = L.kT_min;
    out[3] = L.kT_max;
    out[4] = L.noneq ? 1.0 : 0.0;
    out[5] = (double)L.poles.size();
    out[6] = L.E_right.real();
    out[7] = L.E_zeta.imag();
    out[8] = L.E_eta.real();
    out[9] = L.E_rightmost.real();
    return NEGF_OK;
}

int negf_tube_noneq_layout(const negf_tube* t, int* npaths, int* pts, double* intercuts)
{
    clear_error();
    if (!t || !npaths) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const Tube& tb = t->t;
    *npaths = tb.np;
    for (int q = 0; q < tb.np; ++q) {
        if (pts) pts[q] = tb.pts[q];
        if (intercuts && q + 1 < tb.np) intercuts[q] = tb.intercuts[q];
    }
    return NEGF_OK;
}

int negf_tube_noneq_integrand(const negf_tube* t, int max_pts, double* E_real, double* integrand)
{
    clear_error();
    if (!t) {
        set_error("null argument");
        return NEGF_EARG;
    }
    const Tube& tb = t->t;
    const int n = (int)tb.integrand.size() / 3;
    if (max_pts < n) return n;
    for (int e = 0; e < n; ++e) {
        if (E_real) E_real[e] = tb.neq_path.E[e].real();
        if (integrand)
            for (int w = 0; w < 3; ++w) integrand[(size_t)w * n + e] = tb.integrand[(size_t)w * n + e];
    }
    return n;
}

int negf_tube_current(negf_tube* t, double* I)
{
    clear_error();
    if (!t || !I) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return t->t.current(I);
}

int negf_tube_dos(negf_tube* t, int nE, const double* E_real, double* dos, double* trans)
{
    clear_error();
    if (!t || nE <= 0 || !E_real || !dos || !trans) {
        set_error("invalid DOS arguments");
        return NEGF_EARG;
    }
    return t->t.dos(nE, E_real, dos, trans);
}

int negf_tube_dos_complex(negf_tube* t, int nE, const negf_complex* E, double* dos, double* trans, double* ldos)
{
    clear_error();
    if (!t || nE <= 0 || !E || !dos || !trans) {
        set_error("invalid DOS arguments");
        return NEGF_EARG;
    }
    std::vector<cplx> Ec(nE);
    for (int e = 0; e < nE; ++e) Ec[e] = cplx(E[e].re, E[e].im);
    return t->t.dos_complex(nE, Ec.data(), dos, trans, ldos);
}

int negf_tube_dos_path(negf_tube* t, int which, int max_pts, negf_complex* E, int* n)
{
    clear_error();
    if (!t || !n || (which != 0 && which != 1)) {
        set_error("invalid DOS path arguments");
        return NEGF_EARG;
    }
    std::vector<cplx> Ev;
    int rc = t->t.dos_path(which, Ev);
    if (rc) return rc;
    *n = (int)Ev.size();
    if (E) {
        if ((int)Ev.size() > max_pts) {
            set_error("DOS path has %d points, buffer holds %d", (int)Ev.size(), max_pts);
            return NEGF_EARG;
        }
        for (size_t e = 0; e < Ev.size(); ++e) {
            E[e].re = Ev[e].real();
            E[e].im = Ev[e].imag();
        }
    }
    return NEGF_OK;
}

}  // extern "C"
