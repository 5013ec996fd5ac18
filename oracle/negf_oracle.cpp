/*
 * negf_oracle.cpp -- CPU oracle for the NEGF charge path.
 *
 * TEST INFRASTRUCTURE ONLY (see negf_oracle.h).  A restatement of the
 * reference ELEQTRONeX algorithm, loop for loop, against which the tests
 * compare the HIP product, and the timed CPU baseline.  All citations are paths in the
 * reference repository.
 */
#include "negf_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

/* Constants: Source/Utils/SelectWarpXUtils/WarpXConst.H:18-63 */
const double kPi = 3.14159265358979323846;
const double kKb_eVperK = 8.617333262e-5;
const double kQe = 1.602176634e-19;
const double kEp0 = 8.8541878128e-12;
const double kHeVperHz = 4.135667696e-15; /* PhysConst::h_eVperHz */

struct C {
    double re, im;
};
inline C mk(double r, double i) { return C{r, i}; }
inline C operator+(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
inline C operator-(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
inline C operator-(C a) { return C{-a.re, -a.im}; }
inline C operator*(C a, C b)
{
    return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline C operator*(double s, C a) { return C{s * a.re, s * a.im}; }
inline C operator*(C a, double s) { return C{s * a.re, s * a.im}; }
inline C operator/(C a, C b)
{
    double d = b.re * b.re + b.im * b.im;
    return C{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
inline C operator/(C a, double s) { return C{a.re / s, a.im / s}; }
inline C conj(C a) { return C{a.re, -a.im}; }
inline double abs2(C a) { return a.re * a.re + a.im * a.im; }
inline C csqrt(C a)
{
    std::complex<double> r = std::sqrt(std::complex<double>(a.re, a.im));
    return C{r.real(), r.imag()};
}
inline C cexp(C a)
{
    std::complex<double> r = std::exp(std::complex<double>(a.re, a.im));
    return C{r.real(), r.imag()};
}
inline C toC(orc_c a) { return C{a.re, a.im}; }
inline orc_c toO(C a) { return orc_c{a.re, a.im}; }

/* One mode of one tube: the scalar chain the reference's 1-D MatrixBlock
 * specialisation reduces to (Matrix_Block_Util.H 1-D ops are element-wise). */
struct Chain {
    int N;
    C Hb[2], Hc[2]; /* CNT.cpp:322-335: [-beta, -gamma] */
    double beta, gamma;
};

Chain make_chain(int N, double beta, double gamma)
{
    Chain ch;
    ch.N = N;
    ch.beta = beta;
    ch.gamma = gamma;
    ch.Hb[0] = mk(-beta, 0.);
    ch.Hc[0] = mk(-beta, 0.);
    ch.Hb[1] = mk(-gamma, 0.);
    ch.Hc[1] = mk(-gamma, 0.);
    return ch;
}

/* RGF sweeps: NEGF_Common.cpp:4036-4052 (identical blocks in
 * Compute_RhoEq 3667-3714, poles 3866-3911, RhoNonEq 3288-3335). */
void rgf(int N, const C *alpha, const C Hb[2], const C Hc[2], C *Xt, C *X,
         C *Yt, C *Y)
{
    Y[0] = mk(0., 0.);
    Yt[0] = mk(0., 0.);
    for (int n = 1; n < N; ++n) {
        int p = (n - 1) % 2;
        Yt[n] = Hc[p] / (alpha[n - 1] - Y[n - 1]);
        Y[n] = Hb[p] * Yt[n];
    }
    X[N - 1] = mk(0., 0.);
    Xt[N - 1] = mk(0., 0.);
    for (int n = N - 2; n > -1; n--) {
        int p = n % 2;
        Xt[n] = Hb[p] / (alpha[n + 1] - X[n + 1]);
        X[n] = Hc[p] * Xt[n];
    }
}

/* CNT.cpp:354-393 closed form; Im<0 root rule. Returns 0 if a root with
 * Im<0 exists, 1 otherwise (then the smaller-modulus root is returned). */
int surface_gf(C E, double U, double beta, double gamma, C *g)
{
    C EmU = E - mk(U, 0.);
    C EmU_sq = EmU * EmU;
    double gamma_sq = gamma * gamma;
    C Factor = EmU_sq + mk(gamma_sq - beta * beta, 0.);
    C Sqrt = csqrt(Factor * Factor - 4. * EmU_sq * gamma_sq);
    C Denom = 2. * gamma_sq * EmU;
    C val1 = (Factor + Sqrt) / Denom;
    C val2 = (Factor - Sqrt) / Denom;
    if (val1.im < 0.) {
        *g = val1;
        return 0;
    }
    if (val2.im < 0.) {
        *g = val2;
        return 0;
    }
    /* neither root: the reference leaves gr at its zero initial value
     * (MatrixBlock of GpuComplex, CNT.cpp:376-381); flagged for tests */
    *g = mk(0., 0.);
    return 1;
}

/* Compute_CondensedHamiltonian NEGF_Common.cpp:1436-1508 and
 * DecimationTechnique :4095-4153, 1-D (diagonal block) specialisation. */
C decimation_gf(C EmU, double beta, double gamma, int P, int max_iter,
                int rel_error)
{
    C Hb[2] = {mk(-beta, 0.), mk(-gamma, 0.)};
    std::vector<C> H_tilde(std::max(P, 2), mk(0., 0.));
    C one = mk(1., 0.);
    H_tilde[P - 1] = one / EmU;
    C H_tilde_kP = H_tilde[P - 1];
    for (int k = P - 2; k >= 1; --k) {
        int j = k % 2;
        C temp1 = EmU - Hb[j] * H_tilde[k + 1] * conj(Hb[j]);
        H_tilde[k] = one / temp1;
        H_tilde_kP = H_tilde[k] * Hb[j] * H_tilde_kP;
    }
    C C_tilde_kk = H_tilde[1];
    for (int k = 2; k < P; ++k) {
        int j = (k - 1) % 2;
        C_tilde_kk = H_tilde[k] +
                     H_tilde[k] * conj(Hb[j]) * C_tilde_kk * Hb[j] * H_tilde[k];
    }
    int id = (P - 1) % 2;
    C Xi_s = Hb[0] * H_tilde[1] * conj(Hb[0]);
    C Xi = Xi_s + Hb[id] * C_tilde_kk * conj(Hb[id]);
    C Pi = Hb[0] * H_tilde_kP * Hb[id];

    C mu, nu, gam, zeta;
    {
        C mu0 = EmU - Xi_s;
        C nu0 = EmU - Xi;
        C gamma0 = Pi;
        C nu0_inverse = one / nu0;
        C gamma0_dagger = conj(gamma0);
        C temp1 = gamma0 * nu0_inverse;
        C temp2 = temp1 * gamma0_dagger;
        C temp3 = gamma0_dagger * nu0_inverse;
        mu = mu0 - temp2;
        nu = nu0 - temp2 - temp3 * gamma0;
        gam = temp1 * gamma0;
        zeta = temp3 * gamma0_dagger;
    }
    double error = 1;
    int i = 0;
    while (error > rel_error) {
        C mu_old = mu;
        C nu_inverse = one / nu;
        C temp1 = zeta * nu_inverse * gam;
        C temp2 = gam * nu_inverse * zeta;
        mu = mu - temp1;
        nu = nu - temp1 - temp2;
        gam = gam * nu_inverse * gam;
        zeta = zeta * nu_inverse * zeta;
        C sum = mu + mu_old;
        C error_mat = (mu - mu_old) * (one / sum);
        C fro = csqrt(error_mat * error_mat); /* FrobeniusNorm, 1-D */
        error = std::sqrt(abs2(fro));
        if (i > max_iter) break;
        i++;
    }
    return one / mu;
}

/* get_Sigma_at_contacts NEGF_Common.cpp:4165-4177: Sigma = tau g tau^+,
 * tau = gamma (CNT.cpp:348-351). */
int sigma_contact(const orc_tube *t, C E, double Uc, double beta, C *sigma)
{
    C g;
    int st = 0;
    if (t && t->use_decimation) {
        g = decimation_gf(E - mk(Uc, 0.), beta, t->gamma, t->decimation_layers,
                          t->decimation_max_iter, t->decimation_rel_error);
    } else {
        st = surface_gf(E, Uc, beta, t ? t->gamma : 0., &g);
    }
    double tau = t->gamma;
    *sigma = mk(tau, 0.) * g * mk(tau, 0.);
    return st;
}

C fermi(C z, double kT)
{
    /* FermiFunction NEGF_Common.cpp:1790-1795 */
    C one = mk(1., 0.);
    return one / (cexp(z / kT) + one);
}

struct Path {
    std::vector<C> E, mul;
    std::vector<double> w;
};

Path gl_path(C a, C b, int n, int type)
{
    Path p;
    p.E.resize(n);
    p.mul.resize(n);
    p.w.resize(n);
    orc_gl_path(toO(a), toO(b), n, type, reinterpret_cast<orc_c *>(p.E.data()),
                p.w.data(), reinterpret_cast<orc_c *>(p.mul.data()));
    return p;
}

/* alpha_n = E + minusHa_n (NEGF_Common.cpp:3638-3642) with minusHa = -U
 * (Add_PotentialToHamiltonian :1656-1666); contacts: alpha -= Sigma_c
 * (:3644-3656). */
int fill_alpha(const orc_tube *t, const Chain &ch, const double *U,
               const double Uc[2], C E, C *alpha, C sigma[2])
{
    int st = 0;
    for (int n = 0; n < ch.N; ++n) alpha[n] = E + mk(-U[n], 0.);
    for (int c = 0; c < 2; ++c) st |= sigma_contact(t, E, Uc[c], ch.beta, &sigma[c]);
    alpha[0] = alpha[0] - sigma[0];
    alpha[ch.N - 1] = alpha[ch.N - 1] - sigma[1];
    return st;
}

struct Work {
    std::vector<C> alpha, X, Xt, Y, Yt;
    explicit Work(int N) : alpha(N), X(N), Xt(N), Y(N), Yt(N) {}
};

/* Equilibrium-type accumulation over one path list: out_mode[n] +=
 * DiagMult(((((cm*G)*w)*mul)*f)) in the reference's operand order
 * (Compute_RhoEq kernel NEGF_Common.cpp:3742-3754, Compute_Rho0 :4073-4084).
 * f_mode: 0 -> no Fermi factor (Rho0), 1 -> f(E-mu_min; kT_min). */
int accumulate_paths(const orc_tube *t, const std::vector<Path> &paths,
                     const double *U, const double Uc[2], double cm,
                     int f_mode, double mu_min, double kT_min,
                     std::vector<std::vector<C>> &acc)
{
    int st = 0;
    int N = t->N;
    Work wk(N);
    for (int m = 0; m < t->nmodes; ++m) {
        Chain ch = make_chain(N, t->beta[m], t->gamma);
        for (const Path &path : paths) {
            for (size_t e = 0; e < path.E.size(); ++e) {
                C E = path.E[e];
                C sigma[2];
                st |= fill_alpha(t, ch, U, Uc, E, wk.alpha.data(), sigma);
                rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
                    wk.Yt.data(), wk.Y.data());
                C nF = f_mode ? fermi(E - mk(mu_min, 0.), kT_min) : mk(1., 0.);
                C weight = mk(path.w[e], 0.);
                C mul = path.mul[e];
                for (int n = 0; n < N; ++n) {
                    C G = mk(1., 0.) / (wk.alpha[n] - wk.X[n] - wk.Y[n]);
                    C r = cm * G * weight * mul;
                    if (f_mode) r = r * nF;
                    acc[m][n] = acc[m][n] + r * double(t->degen[m]);
                }
            }
        }
    }
    return st;
}

void sum_modes(const orc_tube *t, const std::vector<std::vector<C>> &acc,
               orc_c *out)
{
    /* MatrixBlock::DiagSum Matrix_Block_Util.H:787-795 */
    for (int n = 0; n < t->N; ++n) {
        C s = mk(0., 0.);
        for (int m = 0; m < t->nmodes; ++m) s = s + acc[m][n];
        out[n] = toO(s);
    }
}

std::vector<Path> noneq_paths(const orc_tube *t, const orc_limits &L)
{
    /* Generate_NonEq_Paths NEGF_Common.cpp:2018-2059 */
    int np = t->num_noneq_paths;
    std::vector<C> pmin(np), pmax(np);
    C Ecr = toC(L.E_contour_right), Erm = toC(L.E_rightmost);
    pmin[0] = Ecr;
    pmax[np - 1] = Erm;
    C len = Erm - Ecr;
    for (int p = 0; p < np - 1; ++p)
        pmax[p] = Ecr + (t->noneq_intercuts[p] / 100.) * len;
    for (int p = 1; p < np; ++p) pmin[p] = pmax[p - 1];
    std::vector<Path> paths;
    for (int p = 0; p < np; ++p)
        paths.push_back(gl_path(pmin[p], pmax[p], t->noneq_pts[p], 0));
    return paths;
}

}  // namespace

extern "C" {

/* Quadrature::Gauss_Legendre Source/Utils/CodeUtils/CodeUtil.cpp:640-684 */
void orc_gauss_legendre(int n, double *x, double *w)
{
    double eps = 3.0e-11;
    int m = (n + 1) / 2;
    double z = 0.;
    for (int i = 1; i <= m; ++i) {
        z = std::cos(kPi * (i - 0.25) / (n + 0.5));
        double z1 = z;
        double pp = 0.;
        do {
            double p1 = 1.0;
            double p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (std::fabs(z - z1) > eps);
        x[i - 1] = -z;
        x[n - i] = z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* c_IntegrationPath::Define_GaussLegendrePoints IntegrationPath.cpp:41-147 */
void orc_gl_path(orc_c a_, orc_c b_, int n, int type, orc_c *E, double *w,
                 orc_c *mul)
{
    C a = toC(a_), b = toC(b_);
    std::vector<double> x(n), ww(n);
    orc_gauss_legendre(n, x.data(), ww.data());
    if (type == 0) {
        for (int i = 0; i < n; ++i) {
            E[i] = toO((b - a) * 0.5 * x[i] + (b + a) * 0.5);
            mul[i] = toO((b - a) * 0.5);
            w[i] = ww[i];
        }
    } else {
        double Ecr = (b.re * b.re - a.re * a.re - a.im * a.im) /
                     (2 * (b.re - a.re));
        double Eci = b.im;
        double R = Ecr - b.re;
        double theta1 = std::asin(std::fabs(a.im) / R);
        if (Ecr > a.re) theta1 = kPi - theta1;
        double theta2 = kPi;
        for (int i = 0; i < n; ++i) {
            double th = (theta2 - theta1) * 0.5 * x[i] + (theta2 + theta1) * 0.5;
            E[i] = orc_c{Ecr + R * std::cos(th), Eci + R * std::sin(th)};
            C eith = cexp(mk(0., th));
            mul[i] = toO((theta2 - theta1) * 0.5 * R * eith * mk(0., 1.));
            w[i] = ww[i];
        }
    }
}

/* c_CNT::Get_Bandgap_Of_Mode CNT.cpp:155-176 and Define_SortedModeVector
 * CNT.cpp:227-286 (same partition/quickSort so degenerate modes keep the
 * reference's representative). */
static double cnt_gap(int m, int n, double acc, double gamma, int p)
{
    double R_cnt = acc * std::sqrt(3. * (m * m + n * n + m * n)) / (2. * kPi);
    double c = acc * gamma / R_cnt;
    double E[6] = {c * std::fabs(3. * p - (2 * m + n)),
                   c * std::fabs(3. * p - (m + 2 * n)),
                   c * std::fabs(3. * p - (m - n)),
                   c * std::fabs(-3. * p + (2 * m + n)),
                   c * std::fabs(-3. * p + (m + 2 * n)),
                   c * std::fabs(3. * p - (m - n))};
    double Emin = E[0];
    for (int i = 1; i < 6; ++i)
        if (Emin > E[i]) Emin = E[i];
    return Emin;
}

static int qs_partition(std::vector<double> &v, std::vector<int> &id, int low,
                        int high, double pivot)
{
    int i = low, j = low;
    while (i <= high) {
        if (v[i] > pivot) {
            i++;
        } else {
            std::swap(v[i], v[j]);
            std::swap(id[i], id[j]);
            i++;
            j++;
        }
    }
    return j - 1;
}

static void qs(std::vector<double> &v, std::vector<int> &id, int low, int high)
{
    if (low < high) {
        double pivot = v[high];
        int pos = qs_partition(v, id, low, high, pivot);
        qs(v, id, low, pos - 1);
        qs(v, id, pos + 1, high);
    }
}

int orc_cnt_sorted_modes(int m, int n, double acc, double gamma, int *mode,
                         double *gap, int *degen)
{
    int atoms_per_ring = (n == 0) ? m : m - 1; /* CNT.cpp:30-40 */
    std::vector<double> Eg(atoms_per_ring);
    std::vector<int> idx(atoms_per_ring);
    for (int p = 1; p < atoms_per_ring + 1; ++p) {
        Eg[p - 1] = cnt_gap(m, n, acc, gamma, p);
        idx[p - 1] = p;
    }
    qs(Eg, idx, 0, atoms_per_ring - 1);
    int count = 0;
    mode[0] = idx[0];
    gap[0] = Eg[0];
    degen[0] = 1;
    count = 1;
    for (int p = 1; p < atoms_per_ring; ++p) {
        if ((Eg[p] - gap[count - 1]) > 1e-6) {
            mode[count] = idx[p];
            gap[count] = Eg[p];
            degen[count] = 1;
            count++;
        } else {
            degen[count - 1] += 1;
        }
    }
    return count;
}

/* c_CNT::get_beta CNT.cpp:288-292 */
double orc_cnt_beta(int J, int m, double gamma)
{
    double arg_im = -kPi * J / m;
    return 2. * gamma * std::cos(-1 * arg_im);
}

/* Ring axial positions (nm): get_AtomPosition_ZigZag_CNT CNT.cpp:93-132,
 * PTD per Generate_AtomLocations NEGF_Common.cpp:1074-1080 (before offset). */
void orc_cnt_ring_positions(int m, int num_unitcells, double acc,
                            double *ptd_nm)
{
    (void)m;
    int rings_per_unitcell = 4;
    double center = -(num_unitcells / 2.) * 3. * acc - acc / 2.;
    int Nr = num_unitcells * rings_per_unitcell;
    for (int r = 0; r < Nr; ++r) {
        int uc = r / rings_per_unitcell;
        double off = 3 * acc * uc;
        double y = 0.;
        int k = r % rings_per_unitcell;
        if (k == 0) y = acc + off;
        if (k == 1) y = 1.5 * acc + off;
        if (k == 2) y = 2.5 * acc + off;
        if (k == 3) y = 3 * acc + off;
        y += center;
        ptd_nm[r] = y / 1.e-9;
    }
}

int orc_surface_gf(orc_c E, double U, double beta, double gamma, orc_c *g)
{
    C r;
    int st = surface_gf(toC(E), U, beta, gamma, &r);
    *g = toO(r);
    return st;
}

orc_c orc_decimation_gf(orc_c EmU, double beta, double gamma, int layers,
                        int max_iter, int rel_error)
{
    return toO(decimation_gf(toC(EmU), beta, gamma, layers, max_iter, rel_error));
}

orc_c orc_fermi(orc_c z, double kT) { return toO(fermi(toC(z), kT)); }

/* Define_EnergyLimits NEGF_Common.cpp:1687-1787 */
void orc_energy_limits(const orc_tube *t, const double mu[2], orc_limits *L)
{
    for (int c = 0; c < 2; ++c) L->kT_contact[c] = kKb_eVperK * t->T[c];
    L->mu_min = mu[0];
    L->mu_max = mu[0];
    L->kT_min = L->kT_contact[0];
    L->kT_max = L->kT_contact[0];
    L->flag_noneq = 0;
    for (int c = 1; c < 2; ++c) {
        if (L->mu_min > mu[c]) L->mu_min = mu[c];
        if (L->mu_max < mu[c]) L->mu_max = mu[c];
        if (L->kT_min > L->kT_contact[c]) L->kT_min = L->kT_contact[c];
        if (L->kT_max < L->kT_contact[c]) L->kT_max = L->kT_contact[c];
    }
    if (std::fabs(L->mu_min - L->mu_max) > 1e-8) L->flag_noneq = 1;
    if (std::fabs(L->kT_min - L->kT_max) > 0.01) L->flag_noneq = 1;

    C E_zPlus = mk(0., t->E_zPlus_imag);
    L->E_zPlus = toO(E_zPlus);
    C Eleft = mk(t->E_valence_min, 0.) + E_zPlus;
    C Erm = mk(L->mu_max + t->tail_upper * L->kT_max, 0.) + E_zPlus;
    L->E_contour_left = toO(Eleft);
    L->E_rightmost = toO(Erm);
    int num_poles = int((t->E_pole_max - kPi * L->kT_max) /
                            (2. * kPi * L->kT_max) +
                        1);
    L->num_poles_formula = num_poles;
    if (L->flag_noneq) {
        C Ecr = mk(L->mu_min - t->tail_lower * L->kT_max, 0.) + E_zPlus;
        L->E_contour_right = toO(Ecr);
        L->num_poles = 0;
        L->E_zeta = orc_c{Ecr.re, 2 * num_poles * kPi * L->kT_max};
        L->E_eta = orc_c{Ecr.re - 14 * L->kT_max, L->E_zeta.im};
    } else {
        L->E_contour_right = toO(Erm);
        L->num_poles = std::min(num_poles, ORC_MAX_POLES);
        for (int p = 0; p < L->num_poles; ++p)
            L->poles[p] = orc_c{L->mu_min, kPi * L->kT_max * (2 * p + 1)};
        L->E_zeta = orc_c{Erm.re, 2 * num_poles * kPi * L->kT_max};
        double tmax = std::max(t->tail_lower, t->tail_upper);
        L->E_eta = orc_c{L->mu_min - tmax * L->kT_max, L->E_zeta.im};
    }
}

void orc_rgf(int N, const orc_c *alpha, const orc_c Hb[2], const orc_c Hc[2],
             orc_c *Xt, orc_c *X, orc_c *Yt, orc_c *Y)
{
    C hb[2] = {toC(Hb[0]), toC(Hb[1])}, hc[2] = {toC(Hc[0]), toC(Hc[1])};
    rgf(N, reinterpret_cast<const C *>(alpha), hb, hc,
        reinterpret_cast<C *>(Xt), reinterpret_cast<C *>(X),
        reinterpret_cast<C *>(Yt), reinterpret_cast<C *>(Y));
}

/* Compute_Rho0 NEGF_Common.cpp:3957-4092 with the path of
 * Define_IntegrationPaths :1798-1816 (30/30/30, built at init with mu=0,
 * U=0, U_contact=0: Initialize_NEGF :1275-1306). */
int orc_rho0(const orc_tube *t, orc_c *Rho0)
{
    double mu0[2] = {0., 0.};
    orc_limits L;
    orc_energy_limits(t, mu0, &L);
    std::vector<Path> paths;
    paths.push_back(gl_path(toC(L.E_zPlus), toC(L.E_zeta), 30, 0));
    paths.push_back(gl_path(toC(L.E_zeta), toC(L.E_eta), 30, 0));
    paths.push_back(gl_path(toC(L.E_eta), toC(L.E_contour_left), 30, 1));
    std::vector<double> U0(t->N, 0.);
    double Uc0[2] = {0., 0.};
    std::vector<std::vector<C>> acc(t->nmodes, std::vector<C>(t->N, mk(0., 0.)));
    double cm = -1. * t->spin_degen / kPi;
    int st = accumulate_paths(t, paths, U0.data(), Uc0, cm, 0, 0., 1., acc);
    sum_modes(t, acc, Rho0);
    return st;
}

/* Compute_RhoEq NEGF_Common.cpp:3581-3781 over Update_IntegrationPaths
 * :2188-2203, plus Compute_GR_atPoles :3784-3954 when in equilibrium. */
int orc_rho_eq(const orc_tube *t, const double *U, const double Uc[2],
               orc_c *RhoEq, int *npts)
{
    orc_limits L;
    orc_energy_limits(t, t->mu, &L);
    std::vector<Path> paths;
    paths.push_back(gl_path(toC(L.E_contour_right), toC(L.E_zeta), t->eq_pts[0], 0));
    paths.push_back(gl_path(toC(L.E_zeta), toC(L.E_eta), t->eq_pts[1], 0));
    paths.push_back(gl_path(toC(L.E_eta), toC(L.E_contour_left), t->eq_pts[2], 1));
    std::vector<std::vector<C>> acc(t->nmodes, std::vector<C>(t->N, mk(0., 0.)));
    double cm = -1. * t->spin_degen / kPi;
    int st = accumulate_paths(t, paths, U, Uc, cm, 1, L.mu_min, L.kT_min, acc);
    if (npts) *npts = t->eq_pts[0] + t->eq_pts[1] + t->eq_pts[2];
    if (!L.flag_noneq) {
        /* GR_atPoles kernel :3937-3948, added at :3773-3776 */
        int N = t->N;
        Work wk(N);
        std::vector<std::vector<C>> gp(t->nmodes, std::vector<C>(N, mk(0., 0.)));
        C pole_const = mk(0., -2 * L.kT_min * t->spin_degen);
        for (int m = 0; m < t->nmodes; ++m) {
            Chain ch = make_chain(N, t->beta[m], t->gamma);
            for (int p = 0; p < L.num_poles; ++p) {
                C E = toC(L.poles[p]);
                C sigma[2];
                st |= fill_alpha(t, ch, U, Uc, E, wk.alpha.data(), sigma);
                rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
                    wk.Yt.data(), wk.Y.data());
                for (int n = 0; n < N; ++n) {
                    C G = mk(1., 0.) / (wk.alpha[n] - wk.X[n] - wk.Y[n]);
                    gp[m][n] = gp[m][n] + (pole_const * G) * double(t->degen[m]);
                }
            }
            for (int n = 0; n < N; ++n) acc[m][n] = acc[m][n] + gp[m][n];
        }
    }
    sum_modes(t, acc, RhoEq);
    return st;
}

/* Compute_RhoNonEq NEGF_Common.cpp:3115-3578 with the O(N^2) contact walks
 * of the device kernel :3377-3497.  integrand (optional): 3*npts values
 * (Hsize/2, Hsize/4, 3*Hsize/4 samples, :3457-3495). */
int orc_rho_noneq(const orc_tube *t, const double *U, const double Uc[2],
                  orc_c *RhoNonEq, double *integrand, int *npts)
{
    orc_limits L;
    orc_energy_limits(t, t->mu, &L);
    int N = t->N;
    std::vector<std::vector<C>> acc(t->nmodes, std::vector<C>(N, mk(0., 0.)));
    if (!L.flag_noneq) {
        sum_modes(t, acc, RhoNonEq);
        if (npts) *npts = 0;
        return 0;
    }
    std::vector<Path> paths = noneq_paths(t, L);
    int total = 0;
    for (auto &p : paths) total += int(p.E.size());
    if (npts) *npts = total;
    if (integrand)
        for (int i = 0; i < 3 * total; ++i) integrand[i] = 0.;
    double cm = -1 * t->spin_degen / (2 * kPi);
    int gcid[2] = {0, N - 1};
    int st = 0;
    Work wk(N);
    for (int m = 0; m < t->nmodes; ++m) {
        Chain ch = make_chain(N, t->beta[m], t->gamma);
        int e_prev = 0;
        for (const Path &path : paths) {
            for (size_t e = 0; e < path.E.size(); ++e) {
                C E = path.E[e];
                C sigma[2];
                st |= fill_alpha(t, ch, U, Uc, E, wk.alpha.data(), sigma);
                C fc[2];
                for (int c = 0; c < 2; ++c)
                    fc[c] = fermi(E - mk(t->mu[c], 0.), L.kT_contact[c]);
                rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
                    wk.Yt.data(), wk.Y.data());
                C Gkk[2], Gam[2];
                for (int k = 0; k < 2; ++k) {
                    int kg = gcid[k];
                    Gkk[k] = mk(1., 0.) / (wk.alpha[kg] - wk.X[kg] - wk.Y[kg]);
                    Gam[k] = mk(0., 1.) * (sigma[k] - conj(sigma[k]));
                }
                C weight = mk(path.w[e], 0.);
                C mul = path.mul[e];
                int e_glo = e_prev + int(e);
                for (int n = 0; n < N; ++n) {
                    C AnF = mk(0., 0.);
                    for (int k = 0; k < 2; ++k) {
                        int kg = gcid[k];
                        C temp = Gkk[k];
                        for (int q = kg; q < n; ++q) temp = (-1. * wk.Xt[q]) * temp;
                        for (int q = kg; q > n; q--) temp = (-1. * wk.Yt[q]) * temp;
                        C A_nn = temp * Gam[k] * conj(temp);
                        AnF = AnF + A_nn * fc[k];
                    }
                    C r = cm * AnF * weight * mul;
                    acc[m][n] = acc[m][n] + r * double(t->degen[m]);
                    if (integrand) {
                        double v = ((cm * AnF) * double(t->degen[m])).re;
                        if (n == int(N / 2))
                            integrand[e_glo] += v;
                        else if (n == int(N / 4))
                            integrand[total + e_glo] += v;
                        else if (n == int(N * 3 / 4))
                            integrand[2 * total + e_glo] += v;
                    }
                }
            }
            e_prev += int(path.E.size());
        }
    }
    sum_modes(t, acc, RhoNonEq);
    return st;
}

/* Compute_InducedCharge NEGF_Common.cpp:2850-2900:
 * Q_n = Im RhoEq + Re RhoNonEq - Im Rho0.  comps (optional, 3N):
 * Im RhoEq, Re RhoNonEq, Im Rho0. */
int orc_solve_negf(const orc_tube *t, const double *U, const double Uc[2],
                   const orc_c *Rho0, double *Q, double *comps, int *npts)
{
    int N = t->N;
    std::vector<orc_c> eq(N), neq(N);
    int n_eq = 0, n_neq = 0;
    int st = orc_rho_eq(t, U, Uc, eq.data(), &n_eq);
    orc_limits L;
    orc_energy_limits(t, t->mu, &L);
    if (L.flag_noneq)
        st |= orc_rho_noneq(t, U, Uc, neq.data(), nullptr, &n_neq);
    else
        for (int n = 0; n < N; ++n) neq[n] = orc_c{0., 0.};
    for (int n = 0; n < N; ++n) {
        Q[n] = (eq[n].im + neq[n].re) - Rho0[n].im;
        if (comps) {
            comps[n] = eq[n].im;
            comps[N + n] = neq[n].re;
            comps[2 * N + n] = Rho0[n].im;
        }
    }
    /* get_Total_Integration_Pts :1826-1841 */
    if (npts) *npts = L.flag_noneq ? n_neq : n_eq;
    return st;
}

/* Find_NonEq_Percent_Intercuts_Adaptively NEGF_Common.cpp:2062-2185 for
 * num_noneq_paths == 3 or 2, driven by the Hsize/2 integrand maximum
 * (:3525-3555).  The point densities per kT are computed at the first call
 * unless given (flag_noneq_integration_pts_density) and kept; E_at_max keeps
 * its previous value when every integrand sample is zero.  Updates
 * t->noneq_intercuts, t->noneq_pts, t->noneq_density and t->E_at_max. */
int orc_adaptive_intercuts(orc_tube *t, const double *U, const double Uc[2])
{
    orc_limits L;
    orc_energy_limits(t, t->mu, &L);
    if (!L.flag_noneq) return 0;
    std::vector<Path> paths = noneq_paths(t, L);
    int total = 0;
    for (auto &p : paths) total += int(p.E.size());
    std::vector<orc_c> neq(t->N);
    std::vector<double> integ(3 * std::max(total, 1));
    int np = 0;
    int st = orc_rho_noneq(t, U, Uc, neq.data(), integ.data(), &np);
    double maxv = 0.;
    int e_glo = 0;
    for (auto &p : paths)
        for (size_t e = 0; e < p.E.size(); ++e, ++e_glo)
            if (maxv < std::fabs(integ[e_glo])) {
                maxv = std::fabs(integ[e_glo]);
                t->E_at_max = p.E[e].re;
            }
    double kT0 = L.kT_contact[0];
    double E_star = (t->E_at_max - t->mu[0]) / kT0;
    double E_min = (L.E_contour_right.re - t->mu[0]) / kT0;
    double E_max = (L.E_rightmost.re - t->mu[0]) / kT0;
    /* noneq_integration_pts_density is an amrex::Vector<int>
     * (NEGF_Common.H:404): the computed densities are truncated on assignment */
    double *d = t->noneq_density;
    auto as_int = [](double x) { return (double)(int)x; };
    if (t->num_noneq_paths == 3) {
        if (!t->has_density) {
            d[0] = as_int(t->noneq_pts[0] / ((t->noneq_intercuts[0] / 100.) * (E_max - E_min)));
            d[1] = as_int(t->noneq_pts[1] /
                          (((t->noneq_intercuts[1] - t->noneq_intercuts[0]) / 100.) *
                           (E_max - E_min)));
            d[2] = as_int(t->noneq_pts[2] /
                          (((100. - t->noneq_intercuts[1]) / 100.) * (E_max - E_min)));
            t->has_density = 1;
        }
        double E_lower = std::max(E_star - t->kT_window[0], E_min);
        double E_upper = std::min(E_star + t->kT_window[1], E_max);
        t->noneq_intercuts[0] = 100 * (E_lower - E_min) / (E_max - E_min);
        t->noneq_intercuts[1] = 100 * (E_upper - E_min) / (E_max - E_min);
        t->noneq_pts[0] = int(std::round(((E_lower - E_min) * d[0]) / 2) * 2);
        t->noneq_pts[1] = int(std::round(((E_upper - E_lower) * d[1]) / 2) * 2);
        t->noneq_pts[2] = int(std::round(((E_max - E_upper) * d[2]) / 2) * 2);
    } else if (t->num_noneq_paths == 2) {
        if (!t->has_density) {
            d[0] = as_int(t->noneq_pts[0] / ((t->noneq_intercuts[0] / 100.) * (E_max - E_min)));
            d[1] = as_int(t->noneq_pts[1] /
                          (((100 - t->noneq_intercuts[0]) / 100.) * (E_max - E_min)));
            t->has_density = 1;
        }
        double E_upper = E_star + t->kT_window[1];
        t->noneq_intercuts[0] = 100 * (E_upper - E_min) / (E_max - E_min);
        t->noneq_pts[0] = int(std::round(((E_upper - E_min) * d[0]) / 2) * 2);
        t->noneq_pts[1] = int(std::round(((E_max - E_upper) * d[1]) / 2) * 2);
    }
    return st;
}

/* Generic checker for a batch of energies on one scalar chain: acc_re[n] +=
 * Re(coef_k * G_nn(E_k)); diagG (optional) [nE][N].  Same RGF, Sigma and
 * alpha construction as the reference functions above. */
int orc_accumulate_diagG(int N, const double *U, double t_even, double t_odd,
                         const double Uc[2], double lead_beta,
                         double lead_gamma, int nE, const orc_c *E,
                         const orc_c *coef, double *acc_re, orc_c *diagG)
{
    orc_tube t;
    std::memset(&t, 0, sizeof(t));
    t.gamma = lead_gamma;
    Chain ch = make_chain(N, t_even, t_odd);
    Work wk(N);
    int st = 0;
    for (int k = 0; k < nE; ++k) {
        C Ek = toC(E[k]);
        for (int n = 0; n < N; ++n) wk.alpha[n] = Ek + mk(-U[n], 0.);
        C s[2];
        for (int c = 0; c < 2; ++c) {
            C g;
            st |= surface_gf(Ek, Uc[c], lead_beta, lead_gamma, &g);
            s[c] = mk(lead_gamma, 0.) * g * mk(lead_gamma, 0.);
        }
        wk.alpha[0] = wk.alpha[0] - s[0];
        wk.alpha[N - 1] = wk.alpha[N - 1] - s[1];
        rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
            wk.Yt.data(), wk.Y.data());
        C ck = toC(coef[k]);
        for (int n = 0; n < N; ++n) {
            C G = mk(1., 0.) / (wk.alpha[n] - wk.X[n] - wk.Y[n]);
            if (acc_re) acc_re[n] += (ck * G).re;
            if (diagG) diagG[size_t(k) * N + n] = toO(G);
        }
    }
    return st;
}

/* The same accumulation with the energy list split into nthreads
 * contiguous chunks, one OpenMP thread each (the reference's loop order
 * inside a chunk); chunk sums are added in chunk order, so the result is
 * reproducible for a given nthreads.  Used to time the CPU baseline on all
 * host cores. */
int orc_accumulate_diagG_par(int nthreads, int N, const double *U,
                             double t_even, double t_odd, const double Uc[2],
                             double lead_beta, double lead_gamma, int nE,
                             const orc_c *E, const orc_c *coef, double *acc_re)
{
    if (nthreads < 1) nthreads = 1;
    std::vector<std::vector<double>> part(nthreads, std::vector<double>(N, 0.));
    std::vector<int> st(nthreads, 0);
#pragma omp parallel for num_threads(nthreads) schedule(static, 1)
    for (int c = 0; c < nthreads; ++c) {
        const int k0 = int((long long)nE * c / nthreads);
        const int k1 = int((long long)nE * (c + 1) / nthreads);
        if (k1 > k0)
            st[c] = orc_accumulate_diagG(N, U, t_even, t_odd, Uc, lead_beta,
                                         lead_gamma, k1 - k0, E + k0, coef + k0,
                                         part[c].data(), nullptr);
    }
    int s = 0;
    for (int c = 0; c < nthreads; ++c) {
        s |= st[c];
        for (int n = 0; n < N; ++n) acc_re[n] += part[c][n];
    }
    return s;
}

/* Generic checker for the non-equilibrium integrand: acc_re[n] +=
 * coefL_k * Gamma_L |G_n0|^2 + coefR_k * Gamma_R |G_n,N-1|^2 with the
 * reference's O(N^2) walk order (NEGF_Common.cpp:3411-3428). */
int orc_accumulate_noneq(int N, const double *U, double t_even, double t_odd,
                         const double Uc[2], double lead_beta,
                         double lead_gamma, int nE, const orc_c *E,
                         const double *coefL, const double *coefR,
                         double *acc_re)
{
    Chain ch = make_chain(N, t_even, t_odd);
    Work wk(N);
    int st = 0;
    int gcid[2] = {0, N - 1};
    for (int k = 0; k < nE; ++k) {
        C Ek = toC(E[k]);
        for (int n = 0; n < N; ++n) wk.alpha[n] = Ek + mk(-U[n], 0.);
        C s[2];
        for (int c = 0; c < 2; ++c) {
            C g;
            st |= surface_gf(Ek, Uc[c], lead_beta, lead_gamma, &g);
            s[c] = mk(lead_gamma, 0.) * g * mk(lead_gamma, 0.);
        }
        wk.alpha[0] = wk.alpha[0] - s[0];
        wk.alpha[N - 1] = wk.alpha[N - 1] - s[1];
        rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
            wk.Yt.data(), wk.Y.data());
        double cf[2] = {coefL[k], coefR[k]};
        for (int n = 0; n < N; ++n) {
            double a = 0.;
            for (int c = 0; c < 2; ++c) {
                int kg = gcid[c];
                C temp = mk(1., 0.) / (wk.alpha[kg] - wk.X[kg] - wk.Y[kg]);
                for (int q = kg; q < n; ++q) temp = (-1. * wk.Xt[q]) * temp;
                for (int q = kg; q > n; q--) temp = (-1. * wk.Yt[q]) * temp;
                C Gam = mk(0., 1.) * (s[c] - conj(s[c]));
                a += cf[c] * (temp * Gam * conj(temp)).re;
            }
            acc_re[n] += a;
        }
    }
    return st;
}

/* Compute_DensityOfStates NEGF_Common.cpp:2390-2800: DOS(E) =
 * spin * sum_n DiagDotSum(degen, A_nn)/(2 pi) / atoms_per_unitcell and
 * T(E) = DiagDotSum(degen, Gamma_0 A_tk[1]) (Caroli).  ldos optional [nE][N]
 * holds DiagDotSum(degen, A_nn)/(2 pi). */
int orc_dos_transmission(const orc_tube *t, const double *U,
                         const double Uc[2], int nE, const orc_c *E,
                         int atoms_per_unitcell, double *dos, double *trans,
                         double *ldos)
{
    int N = t->N;
    Work wk(N);
    int st = 0;
    int gcid[2] = {0, N - 1};
    for (int e = 0; e < nE; ++e) {
        double tr_dos = 0., tr_T = 0.;
        std::vector<C> ld(N, mk(0., 0.));
        C Tsum = mk(0., 0.);
        for (int m = 0; m < t->nmodes; ++m) {
            Chain ch = make_chain(N, t->beta[m], t->gamma);
            C sigma[2];
            C Ee = toC(E[e]);
            st |= fill_alpha(t, ch, U, Uc, Ee, wk.alpha.data(), sigma);
            rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
                wk.Yt.data(), wk.Y.data());
            C Gkk[2], Gam[2];
            for (int k = 0; k < 2; ++k) {
                int kg = gcid[k];
                Gkk[k] = mk(1., 0.) / (wk.alpha[kg] - wk.X[kg] - wk.Y[kg]);
                Gam[k] = mk(0., 1.) * (sigma[k] - conj(sigma[k]));
            }
            for (int n = 0; n < N; ++n) {
                C A = mk(0., 0.);
                C Atk1 = mk(0., 0.);
                for (int k = 0; k < 2; ++k) {
                    int kg = gcid[k];
                    C temp = Gkk[k];
                    for (int q = kg; q < n; ++q) temp = (-1. * wk.Xt[q]) * temp;
                    for (int q = kg; q > n; q--) temp = (-1. * wk.Yt[q]) * temp;
                    C A_nn = temp * Gam[k] * conj(temp);
                    C A_kn = Gkk[k] * Gam[k] * conj(temp);
                    for (int q = kg + 1; q < N; ++q) A_kn = (-1. * wk.Xt[q - 1]) * A_kn;
                    for (int q = kg - 1; q >= 0; q--) A_kn = (-1. * wk.Yt[q + 1]) * A_kn;
                    int ct = (k == 0) ? N - 1 : 0; /* CNT.cpp:343-344 */
                    if (k == 1 && n == ct) Atk1 = A_kn;
                    A = A + A_nn;
                }
                C val = (mk(double(t->degen[m]), 0.) * A) / (2. * kPi);
                ld[n] = ld[n] + val;
                C T12 = Gam[0] * Atk1;
                Tsum = Tsum + mk(double(t->degen[m]), 0.) * T12;
            }
        }
        for (int n = 0; n < N; ++n) {
            tr_dos += ld[n].re;
            if (ldos) ldos[size_t(e) * N + n] = ld[n].re;
        }
        tr_T = Tsum.re;
        dos[e] = t->spin_degen * tr_dos / atoms_per_unitcell;
        trans[e] = tr_T;
    }
    return st;
}

/* Compute_Current NEGF_Common.cpp:4321-4760 (non-equilibrium path, contact
 * sites only): IF_k = Gamma_k f_k A_kk - Gamma_k Gn_kk with A = sum_j A^j,
 * Gn = sum_j A^j f_j; I_k += s q/h Re(IF_k g w mul).  Zero in equilibrium
 * (no non-equilibrium path). */
int orc_current(const orc_tube *t, const double *U, const double Uc[2],
                double I[2])
{
    I[0] = I[1] = 0.;
    orc_limits L;
    orc_energy_limits(t, t->mu, &L);
    if (!L.flag_noneq) return 0;
    int N = t->N;
    std::vector<Path> paths = noneq_paths(t, L);
    double cm = t->spin_degen * kQe / kHeVperHz;
    int gcid[2] = {0, N - 1};
    Work wk(N);
    int st = 0;
    for (int m = 0; m < t->nmodes; ++m) {
        Chain ch = make_chain(N, t->beta[m], t->gamma);
        for (const Path &path : paths)
            for (size_t e = 0; e < path.E.size(); ++e) {
                C E = path.E[e];
                C sigma[2];
                st |= fill_alpha(t, ch, U, Uc, E, wk.alpha.data(), sigma);
                C fc[2];
                for (int c = 0; c < 2; ++c)
                    fc[c] = fermi(E - mk(t->mu[c], 0.), L.kT_contact[c]);
                rgf(N, wk.alpha.data(), ch.Hb, ch.Hc, wk.Xt.data(), wk.X.data(),
                    wk.Yt.data(), wk.Y.data());
                C Gkk[2], Gam[2];
                for (int k = 0; k < 2; ++k) {
                    int kg = gcid[k];
                    Gkk[k] = mk(1., 0.) / (wk.alpha[kg] - wk.X[kg] - wk.Y[kg]);
                    Gam[k] = mk(0., 1.) * (sigma[k] - conj(sigma[k]));
                }
                for (int kk = 0; kk < 2; ++kk) {
                    int n = gcid[kk];
                    C A = mk(0., 0.), Gn = mk(0., 0.);
                    for (int k = 0; k < 2; ++k) {
                        int kg = gcid[k];
                        C temp = Gkk[k];
                        for (int q = kg; q < n; ++q) temp = (-1. * wk.Xt[q]) * temp;
                        for (int q = kg; q > n; q--) temp = (-1. * wk.Yt[q]) * temp;
                        C A_nn = temp * Gam[k] * conj(temp);
                        A = A + A_nn;
                        Gn = Gn + A_nn * fc[k];
                    }
                    C IF = Gam[kk] * fc[kk] * A - Gam[kk] * Gn;
                    C cur = (mk(cm, 0.) * IF * double(t->degen[m])) *
                            mk(path.w[e], 0.) * path.mul[e];
                    I[kk] += cur.re;
                }
            }
    }
    return st;
}

/* ---------------- Broyden modified-second (CPU order) ---------------- */
struct orc_broyden {
    int S, max_step, norm_type, step;
    double fraction, scalar, norm_sum_prev;
    std::vector<double> n_in, n_prev_in, F, dF, Norm, sum_vec, intermed;
    std::vector<double> VmatTran; /* [site][iter] (Broyden_Parallel_General.cpp:150-167) */
    std::vector<double> Wmat;     /* [iter][site] */
};

orc_broyden *orc_broyden_create(int S, double fraction, int max_step,
                                int norm_type)
{
    /* Set_Broyden_Parallel Broyden_Parallel_General.cpp:68-248 */
    orc_broyden *b = new orc_broyden;
    b->S = S;
    b->max_step = max_step;
    b->norm_type = norm_type;
    b->step = 1;
    b->fraction = fraction;
    b->scalar = 1.;
    b->norm_sum_prev = 1.e10;
    b->n_in.assign(S, 0.);
    b->n_prev_in.assign(S, 0.);
    b->F.assign(S, 0.);
    b->dF.assign(S, 0.);
    b->Norm.assign(S, 0.);
    b->sum_vec.assign(S, 0.);
    b->intermed.assign(max_step + 1, 0.);
    b->VmatTran.assign(size_t(S) * (max_step + 1), 0.);
    b->Wmat.assign(size_t(max_step + 1) * S, 0.);
    return b;
}

void orc_broyden_free(orc_broyden *b) { delete b; }

void orc_broyden_set_nin(orc_broyden *b, const double *n_in)
{
    for (int s = 0; s < b->S; ++s) b->n_in[s] = n_in[s];
}

void orc_broyden_get_nin(const orc_broyden *b, double *n_in)
{
    for (int s = 0; s < b->S; ++s) n_in[s] = b->n_in[s];
}

int orc_broyden_get_step(const orc_broyden *b) { return b->step; }

/* Execute_Broyden_Modified_Second_Algorithm_Parallel_SkipGPU
 * Broyden_Second_Parallel.cpp:9-211.  Returns 1 if m exceeds the history
 * capacity (the reference aborts at Transport.cpp:389-393). */
int orc_broyden_step(orc_broyden *b, const double *n_out, double *max_norm,
                     double *l2, double *norm_per_site)
{
    int S = b->S;
    int MS = b->max_step + 1;
    auto V = [&](int site, int iter) -> double & {
        return b->VmatTran[size_t(site) * MS + iter];
    };
    auto W = [&](int iter, int site) -> double & {
        return b->Wmat[size_t(iter) * S + site];
    };
    std::fill(b->Norm.begin(), b->Norm.end(), 0.);
    std::fill(b->sum_vec.begin(), b->sum_vec.end(), 0.);
    std::fill(b->intermed.begin(), b->intermed.end(), 0.);
    double norm_sum = 0.;
    if (b->norm_type == 0) {
        for (int s = 0; s < S; ++s) {
            double Fc = b->n_in[s] - n_out[s];
            b->Norm[s] = std::fabs(Fc);
            norm_sum += std::pow(Fc, 2);
        }
    } else {
        for (int s = 0; s < S; ++s) {
            double Fc = b->n_in[s] - n_out[s];
            b->Norm[s] = std::fabs(Fc / (b->n_in[s] + n_out[s]));
            norm_sum += std::pow(b->Norm[s], 2);
        }
    }
    double bnorm = b->Norm[0];
    for (int s = 1; s < S; ++s)
        if (bnorm < b->Norm[s]) bnorm = b->Norm[s];
    norm_sum = std::sqrt(norm_sum);
    double denom = 0.;
    for (int s = 0; s < S; ++s) {
        double Fc = b->n_in[s] - n_out[s];
        b->dF[s] = Fc - b->F[s];
        b->F[s] = Fc;
        denom += std::pow(b->dF[s], 2.);
    }
    b->norm_sum_prev = norm_sum;
    int m = b->step - 1;
    if (m > b->max_step) return 1;
    if (m > 0) {
        for (int it = 1; it <= m - 1; ++it) {
            double sum = 0.;
            for (int s = 0; s < S; ++s) sum += V(s, it) * b->dF[s];
            b->intermed[it] = sum;
        }
        for (int s = 0; s < S; ++s) {
            double sum = 0.;
            for (int it = 1; it <= m - 1; ++it) sum += W(it, s) * b->intermed[it];
            b->sum_vec[s] = sum;
        }
        for (int s = 0; s < S; ++s) {
            double delta_n = b->n_in[s] - b->n_prev_in[s];
            V(s, m) = b->dF[s] / denom;
            W(m, s) = -b->fraction * b->dF[s] + delta_n - b->sum_vec[s];
        }
        std::fill(b->intermed.begin(), b->intermed.end(), 0.);
        for (int it = 1; it <= m; ++it) {
            double sum = 0.;
            for (int s = 0; s < S; ++s) sum += V(s, it) * b->F[s];
            b->intermed[it] = sum;
        }
        std::fill(b->sum_vec.begin(), b->sum_vec.end(), 0.);
        for (int s = 0; s < S; ++s) {
            double sum = 0.;
            for (int it = 1; it <= m; ++it) sum += W(it, s) * b->intermed[it];
            b->sum_vec[s] = sum;
        }
    }
    for (int s = 0; s < S; ++s) {
        b->n_prev_in[s] = b->n_in[s];
        b->n_in[s] = b->n_prev_in[s] - b->scalar * b->fraction * b->F[s] -
                     b->scalar * b->sum_vec[s];
    }
    b->step += 1;
    if (max_norm) *max_norm = bnorm;
    if (l2) *l2 = norm_sum;
    if (norm_per_site)
        for (int s = 0; s < S; ++s) norm_per_site[s] = b->Norm[s];
    return 0;
}

/* Reset_Broyden_Parallel Broyden_Parallel_General.cpp:250-355 */
void orc_broyden_reset(orc_broyden *b, int keep_n_in, double init_value)
{
    if (!keep_n_in) std::fill(b->n_in.begin(), b->n_in.end(), init_value);
    std::fill(b->n_prev_in.begin(), b->n_prev_in.end(), 0.);
    std::fill(b->F.begin(), b->F.end(), 0.);
    std::fill(b->dF.begin(), b->dF.end(), 0.);
    std::fill(b->Norm.begin(), b->Norm.end(), 0.);
    std::fill(b->sum_vec.begin(), b->sum_vec.end(), 0.);
    std::fill(b->intermed.begin(), b->intermed.end(), 0.);
    std::fill(b->VmatTran.begin(), b->VmatTran.end(), 0.);
    std::fill(b->Wmat.begin(), b->Wmat.end(), 0.);
    b->step = 1;
    b->scalar = 1.;
    b->norm_sum_prev = 1.e10;
}

/* Dense electrostatic stand-in for MLMG + gather (not in the reference):
 * K_ij [V per e] = ke/eps_r * (1/sqrt(d^2+R^2) - 1/sqrt(d^2+R_img^2)). */
void orc_coulomb_kernel(int N, const double *ptd_nm, const orc_coulomb *c,
                        double *K)
{
    double ke = kQe / (4. * kPi * kEp0) * 1.e9; /* V nm per e */
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double d = ptd_nm[i] - ptd_nm[j];
            double d2 = d * d;
            K[size_t(i) * N + j] =
                ke / c->eps_r *
                (1. / std::sqrt(d2 + c->R_ring_nm * c->R_ring_nm) -
                 1. / std::sqrt(d2 + c->R_image_nm * c->R_image_nm));
        }
}

/* One bias point of c_TransportSolver::Solve Transport.cpp:359-539 with the
 * dense stand-in for Parts 1-2 (U = U_ext - V_gate - K n_in; U_contact = U at the
 * end sites as in Nanostructure.cpp:494-504), Part 3 NEGF, Part 4 Broyden,
 * loop while Broyden_Norm > Broyden_max_norm, then Reset_ForNextBiasStep. */
int orc_scf_bias_point(const orc_tube *t, const orc_c *Rho0, const double *K,
                       double V_gate, const orc_scf_params *sp,
                       orc_broyden *b, double *U_out, double *Q_out,
                       int *iters, double *final_norm)
{
    int N = t->N;
    std::vector<double> U(N), Q(N), nin(N);
    double norm = 1.;
    int it = 0;
    int st = 0;
    do {
        if (orc_broyden_get_step(b) > sp->max_step) {
            st = 2;
            break;
        }
        orc_broyden_get_nin(b, nin.data());
        for (int i = 0; i < N; ++i) {
            double s = 0.;
            for (int j = 0; j < N; ++j) s += K[size_t(i) * N + j] * nin[j];
            U[i] = (sp->U_ext ? sp->U_ext[i] : 0.) - V_gate - s;
        }
        double Uc[2] = {U[0], U[N - 1]};
        st |= orc_solve_negf(t, U.data(), Uc, Rho0, Q.data(), nullptr, nullptr);
        double l2;
        if (orc_broyden_step(b, Q.data(), &norm, &l2, nullptr)) {
            st = 2;
            break;
        }
        it++;
    } while (norm > sp->max_norm);
    for (int i = 0; i < N; ++i) {
        U_out[i] = U[i];
        Q_out[i] = Q[i];
    }
    if (iters) *iters = it;
    if (final_norm) *final_norm = norm;
    orc_broyden_reset(b, sp->reset_with_previous, sp->initial_value);
    return st;
}

}  /* extern "C" */
