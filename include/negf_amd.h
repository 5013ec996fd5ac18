/*
 * negf_amd.h -- C ABI of the MI355X-native NEGF charge path.
 *
 * Drop-in boundary for the reference ELEQTRONeX NEGF material path
 * (AMReX-Microelectronics/Electrostatic).  Plain pointers and sizes only;
 * every entry point returns an int status (NEGF_OK on success) and
 * negf_last_error() describes the last failure of the calling thread.
 * Device pointers (d_*) are hipMalloc'd memory owned by the caller;
 * stream arguments are hipStream_t passed as void* (NULL = default stream).
 *
 * The reference has no C ABI of its own: its boundary is the C++ class
 * c_Nanostructure_Base (Source/Solver/Transport/Nanostructure.H:15-45) over
 * c_NEGF_Common<T> (Source/Solver/Transport/NEGF/NEGF_Common.H:478-605).  Each
 * declaration below names the reference member it replaces.
 */
#ifndef NEGF_AMD_H_
#define NEGF_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } negf_complex;

/* Status codes.  Where the reference aborts (amrex::Abort) these are returned. */
#define NEGF_OK 0
#define NEGF_EARG 1      /* invalid argument                                   */
#define NEGF_EALLOC 2    /* device allocation failed                           */
#define NEGF_EROOT 3     /* diagnostic only: some energy had no surface-GF root */
                         /* with Im<0; g = 0 was used there, as the reference  */
                         /* does (CNT.cpp:376-381).  Returned by the status    */
                         /* calls, never by the compute calls.                 */
#define NEGF_ENONCONV 4  /* Broyden_Step > Broyden_threshold_maxstep            */
                         /* (Transport.cpp:389-393)                             */
#define NEGF_ENAN 5      /* non-finite value                                   */
#define NEGF_EDEVICE 6   /* HIP runtime error                                  */
#define NEGF_ENOTSUP 7   /* configuration outside what this build supports     */

/* Message of the last failing call on this thread ("" if none). */
const char *negf_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int negf_version(void);
/* Largest chain length (sites) the segmented RGF kernels accept. */
int negf_max_sites(void);

/* ------------------------------------------------------------------------
 * Level 1: batched RGF over an energy list for one scalar tridiagonal chain.
 * Replaces the per-energy loop of c_NEGF_Common::Compute_RhoEq / Compute_Rho0
 * / Compute_GR_atPoles (NEGF_Common.cpp:3630-3759, 4002-4089, 3833-3952):
 * host RGF sweep (:3667-3714) + per-site kernel (:3742-3754).
 * ------------------------------------------------------------------------ */
typedef struct {
    int N;                    /* sites                                        */
    double t_even, t_odd;     /* |hopping| on links (n,n+1), n even / n odd:  */
                              /* Hb = Hc = -t (CNT.cpp:322-335)               */
    double Uc[2];             /* contact potentials (U_contact)               */
    int use_decimation;       /* 0: closed-form lead GF, 1: decimation        */
    int decimation_layers, decimation_max_iter, decimation_rel_error;
    int segment_length;       /* RGF segment length in sites, a multiple of  */
                              /* 16 up to 128 (0: default 128); performance   */
                              /* only, results do not depend on it           */
} negf_chain_params;

/* acc[n] = sum_k Re(coef_k * G_nn(E_k)) for n < N (overwritten).
 * d_diagG (optional, [nE][N]) receives G_nn(E_k).  Leads: Sigma = t_odd^2 g
 * with g the closed-form surface GF of the (t_even, t_odd) chain. */
int negf_rgf_diag_accumulate(const negf_chain_params *p, const double *d_U,
                             int nE, const negf_complex *d_E,
                             const negf_complex *d_coef, double *d_acc,
                             negf_complex *d_diagG, void *stream);

/* negf_rgf_diag_accumulate without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results.  The workspace (chain
 * descriptors, segment matrices and states, partial rows, tables) lives in
 * one fixed 64 MiB device array of the library, which every GPU the library
 * runs on holds until the process ends; a problem whose workspace exceeds it
 * (about 5e7 energy-sites at the default segment length, e.g. 8192 energies
 * x 4096 sites needs 45 MB) fails with NEGF_EARG before anything is
 * launched.  All calls on one GPU share that array: calls that overlap in
 * time on one GPU (from other host threads or streams) race, and the caller
 * must serialise them.  Synchronises the stream before returning. */
int negf_rgf_diag_accumulate_noalloc(const negf_chain_params *p,
                                     const double *d_U, int nE,
                                     const negf_complex *d_E,
                                     const negf_complex *d_coef,
                                     double *d_acc, negf_complex *d_diagG,
                                     void *stream);

/* Non-equilibrium form (Compute_RhoNonEq kernel NEGF_Common.cpp:3377-3497):
 * acc[n] = sum_k coefL_k Gamma_L |G_n0|^2 + coefR_k Gamma_R |G_n,N-1|^2.
 * d_integrand (optional, [nE][3]) receives the same integrand with
 * coefficients iL_k, iR_k at sites N/2, N/4, 3N/4 (:3457-3495); pass NULL
 * d_iL/d_iR to reuse coefL/coefR. */
int negf_rgf_noneq_accumulate(const negf_chain_params *p, const double *d_U,
                              int nE, const negf_complex *d_E,
                              const double *d_coefL, const double *d_coefR,
                              const double *d_iL, const double *d_iR,
                              double *d_acc, double *d_integrand,
                              void *stream);

/* Per-energy trace of the spectral function and transmission
 * (Compute_DensityOfStates NEGF_Common.cpp:2640-2777): for each E_k,
 * dos[k] = sum_n (Gamma_L |G_n0|^2 + Gamma_R |G_n,N-1|^2),
 * trans[k] = Gamma_L Gamma_R |G_0,N-1|^2 (device arrays of nE). */
int negf_rgf_dos(const negf_chain_params *p, const double *d_U, int nE,
                 const negf_complex *d_E, double *d_dos, double *d_trans,
                 void *stream);

/* Reusable plan for the diagonal form: chain descriptors and block tables are
 * built once; each execute enqueues the lead self-energies, the potential
 * transpose, the RGF kernel and the fixed-order reduction on the stream
 * without host synchronisation.  The production form of
 * negf_rgf_diag_accumulate for loops that call it repeatedly. */
typedef struct negf_rgf_plan negf_rgf_plan;
int negf_rgf_plan_create(const negf_chain_params *p, const double *d_U,
                         int nE, const negf_complex *d_E,
                         const negf_complex *d_coef, int want_diagG,
                         negf_rgf_plan **out);
int negf_rgf_plan_execute(negf_rgf_plan *plan, double *d_acc,
                          negf_complex *d_diagG, void *stream);
/* Synchronise the stream; NEGF_EROOT (diagnostic) if any energy had no
 * surface-GF root with Im<0 since the last check. */
int negf_rgf_plan_status(negf_rgf_plan *plan, void *stream);
/* hipEvent timing of the RGF kernel alone (on the execute stream). */
int negf_rgf_plan_enable_timing(negf_rgf_plan *plan, int on);
int negf_rgf_plan_kernel_time(negf_rgf_plan *plan, double *total_ms,
                              int *launches);
int negf_rgf_plan_destroy(negf_rgf_plan *plan);

/* Gauss-Legendre path (c_IntegrationPath::Define_GaussLegendrePoints
 * IntegrationPath.cpp:41-147 over Quadrature::Gauss_Legendre
 * CodeUtil.cpp:640-684): type 0 line, 1 circular arc.  Host arrays of n. */
int negf_gl_path(negf_complex a, negf_complex b, int n, int type,
                 negf_complex *E, double *w, negf_complex *mul);

/* ------------------------------------------------------------------------
 * Level 2: NEGF material (one zigzag CNT in mode space).  Mirrors
 * c_NEGF_Common<T> + c_CNT (CNT.H:13-103) with the reference's input keys.
 * ------------------------------------------------------------------------ */
typedef struct negf_tube negf_tube;

typedef struct {
    /* CNT_default.* / <name>.* (CNT.cpp:13-21) */
    int type_id[2];                 /* (m, n); zigzag n = 0               */
    double acc;                     /* C-C bond length (m)                 */
    double gamma;                   /* hopping (eV)                        */
    int num_modes;                  /* BLOCK_SIZE: lowest distinct modes   */
    /* NS_default.* / <name>.* (NEGF_Common.cpp:288-696) */
    int num_unitcells;
    double E_valence_min, E_pole_max, E_zPlus_imag;
    double Fermi_tail_factors[2];
    int eq_integration_pts[3];
    int num_noneq_paths;
    int noneq_integration_pts[8];
    double noneq_percent_intercuts[8];
    double contact_mu[2];           /* contact_mu (eV)                     */
    double contact_T[2];            /* contact_T (K)                       */
    int use_decimation, decimation_layers, decimation_max_iter;
    int decimation_rel_error;       /* int, as in NEGF_Common.H:475        */
    int flag_adaptive_integration_limits;
    int integrand_correction_interval;
    double kT_window_around_singularity[2];
    int noneq_integration_pts_density[8]; /* points per kT of each noneq    */
                                    /* path (NEGF_Common.cpp:537-539), read */
                                    /* with adaptive limits; all 0: not    */
                                    /* given, set from the first correction */
    int flatband_dos_integration_pts;
    double flatband_dos_integration_limits[2];  /* eV (NEGF_Common.H:417)  */
    int segment_length;             /* RGF segment length (0: chosen from  */
                                    /* the tube length); no reference      */
                                    /* counterpart (performance only)      */
} negf_tube_params;

/* Fill p with the reference defaults (NEGF_Common.H:384-476, CNT.H:17-18). */
void negf_tube_params_default(negf_tube_params *p);

/* Initialize_NEGF_Params + Initialize_NEGF (NEGF_Common.cpp:1256-1306):
 * modes, Hamiltonian, contacts, energy limits, paths, Rho0. */
int negf_tube_create(const negf_tube_params *p, negf_tube **out);
int negf_tube_destroy(negf_tube *t);
/* Hsize_glo (get_Hsize CNT.H:74). */
int negf_tube_num_sites(const negf_tube *t);
/* h_PTD_glo_vec: axial ring positions in nm (NEGF_Common.cpp:1074-1080). */
int negf_tube_positions(const negf_tube *t, double *ptd_nm);
/* Selected modes: beta and degeneracy (Define_SortedModeVector CNT.cpp:227-286). */
int negf_tube_modes(const negf_tube *t, int *mode_id, double *beta, int *degen);
/* Potential at sites and contacts: U_n = -<phi>, U_contact = U at end sites
 * (Obtain_PotentialAtSites Nanostructure.cpp:494-504). Host arrays. */
int negf_tube_set_potential(negf_tube *t, const double *U, const double Uc[2]);
/* Device form used by a device-resident SCF loop: d_U (N doubles) is copied
 * on the stream and the contact potentials become U[0], U[N-1]. */
int negf_tube_set_potential_device(negf_tube *t, const double *d_U, void *stream);
/* set_Contact_Electrochemical_Potential (NEGF_Common.cpp:356-368) and
 * contact temperatures; limits and paths are rebuilt at the next solve. */
int negf_tube_set_contacts(negf_tube *t, const double mu[2], const double T[2]);
/* Compute_Rho0 (NEGF_Common.cpp:3957-4092); Im Rho0 per site to host. */
int negf_compute_rho0(negf_tube *t, double *rho0_im);
/* Solve_NEGF (NEGF_Common.cpp:1309-1340) -> Compute_InducedCharge
 * (:2850-2900): Q_n = Im RhoEq + Re RhoNonEq - Im Rho0 into Q (host, N).
 * iter is the SCF iteration (adaptive limits every
 * integrand_correction_interval).  *n_intg_pts = get_Total_Integration_Pts. */
int negf_solve(negf_tube *t, int iter, double *Q, int *n_intg_pts);
/* Same, with Q left in device memory d_Q (N doubles). */
int negf_solve_device(negf_tube *t, int iter, double *d_Q, int *n_intg_pts,
                      void *stream);
/* Synchronise the tube's work; NEGF_EROOT (diagnostic) if any energy of the
 * solves since the last check had no surface-GF root with Im<0. */
int negf_tube_status(negf_tube *t);
/* Charge components of the last solve (_chargeComp.dat, :3000-3023):
 * Im RhoEq, Re RhoNonEq, Im Rho0 (host arrays of N). */
int negf_get_charge_components(const negf_tube *t, double *eq_im,
                               double *noneq_re, double *rho0_im);
/* Energy limits / path summary after the last update:
 * out[0..9] = mu_min, mu_max, kT_min, kT_max, flag_noneq, num_poles,
 * E_contour_right.re, E_zeta.im, E_eta.re, E_rightmost.re. */
int negf_tube_limits(const negf_tube *t, double *out);
/* Non-equilibrium integration layout after the last update: npaths, and
 * per path points and percent intercuts. */
int negf_tube_noneq_layout(const negf_tube *t, int *npaths, int *pts,
                           double *intercuts);
/* Non-equilibrium integrand of the last adaptive update at sites N/2, N/4,
 * 3N/4 (NonEq_Integrand, _Source, _Drain; NEGF_Common.cpp:3457-3495,
 * _integrand.dat): E_real[n], integrand[3][n].  Returns n (points), or the
 * required size if max_pts is smaller. */
int negf_tube_noneq_integrand(const negf_tube *t, int max_pts, double *E_real,
                              double *integrand);
/* Compute_Current (NEGF_Common.cpp:4321-4760) at the present potential over
 * the non-equilibrium path: I[k] in A at contact k (I[1] = -I[0]); zero in
 * equilibrium. */
int negf_tube_current(negf_tube *t, double *I);
/* Compute_DensityOfStates over nE real energies (E_zPlus added):
 * DOS = spin * sum_modes degen * trace / (2 pi) / atoms_per_unitcell and
 * T = sum_modes degen * Gamma_L Gamma_R |G_0,N-1|^2 (host arrays). */
int negf_tube_dos(negf_tube *t, int nE, const double *E_real, double *dos,
                  double *trans);
/* The same at complex energies E (host array of nE), with the per-site
 * LDOS_n = sum_modes degen A_nn / (2 pi) (:2733-2742) into ldos ([nE][N],
 * optional): the values of the reference's Ept_<e>.dat files. */
int negf_tube_dos_complex(negf_tube *t, int nE, const negf_complex *E,
                          double *dos, double *trans, double *ldos);
/* Energies of the reference's DOS paths: which = 0, the flat-band path of
 * Initialize_NEGF (NEGF_Common.cpp:1807-1815): flatband_dos_integration_pts
 * Gauss-Legendre points from flatband_dos_integration_limits[0] + i eta to
 * [1] + i eta; which = 1, the per-step path of Update_IntegrationPaths
 * (:2210-2249): the non-equilibrium path when the contacts differ, else
 * flatband_dos_integration_pts points from Re(E_eta) (imaginary part 0) to
 * E_rightmost.  *n receives the count; E (optional, max_pts) the energies. */
int negf_tube_dos_path(negf_tube *t, int which, int max_pts, negf_complex *E,
                       int *n);

/* ------------------------------------------------------------------------
 * Batched solve of many tubes (no reference counterpart: the reference loops
 * over tubes, each with its own per-energy host loop, Transport.cpp:417-431).
 * All tubes' equilibrium chains form one RGF launch sequence and all biased
 * tubes' non-equilibrium chains a second one; Q of tube i lands at
 * d_Q[offsets[i] .. offsets[i+1]).  Tubes keep their own parameters,
 * contacts and potentials (negf_tube_set_* as usual) and must share the lead
 * (decimation) settings.
 * ------------------------------------------------------------------------ */
typedef struct negf_batch negf_batch;
int negf_batch_create(negf_tube *const *tubes, int ntubes, negf_batch **out);
int negf_batch_destroy(negf_batch *b);
/* offsets: ntubes + 1 values (cumulative site counts). */
int negf_batch_offsets(const negf_batch *b, long long *offsets);
/* Solve_NEGF for every tube at SCF iteration iter; d_Q is ready on stream. */
int negf_batch_solve(negf_batch *b, int iter, double *d_Q, void *stream);
/* Replay the launch sequence from a captured hipGraph (re-captured when a
 * tube's paths change or d_Q moves). */
int negf_batch_use_graph(negf_batch *b, int on);
/* RGF segment length for the whole batch (0: default). */
int negf_batch_set_segment_length(negf_batch *b, int segment_length);
/* Time the solve for segment lengths 32..128 and keep the fastest. */
int negf_batch_autotune(negf_batch *b, double *d_Q, int *segment_length,
                        double *ms);
/* negf_batch_autotune with a persistent cache (no reference counterpart):
 * cache_path is a text file of "key segment_length ms" lines, the key naming
 * the device (architecture, compute units) and a hash of every tube's sites
 * and chain counts.  A hit applies the stored length without timing
 * (*from_cache = 1); a miss tunes and appends the result. */
int negf_batch_autotune_cached(negf_batch *b, double *d_Q, const char *cache_path,
                               int *segment_length, double *ms, int *from_cache);
/* hipEvent time of the last solve's launch sequence. */
int negf_batch_enable_timing(negf_batch *b, int on);
int negf_batch_last_time(negf_batch *b, double *ms);
/* Energy-site count of one batched solve (chains x sites). */
int negf_batch_energy_sites(const negf_batch *b, double *energy_sites);
/* Synchronise; NEGF_EROOT (diagnostic) as for negf_tube_status. */
int negf_batch_status(negf_batch *b);

/* ------------------------------------------------------------------------
 * Broyden modified-second on device (Execute_Broyden_Modified_Second_
 * Algorithm_Parallel Broyden_Second_Parallel.cpp:213-486; CPU ordering
 * :9-211; Set_/Reset_Broyden_Parallel Broyden_Parallel_General.cpp:68-355).
 * ------------------------------------------------------------------------ */
typedef struct negf_broyden negf_broyden;
/* norm_type: 0 absolute, 1 relative (Broyden_norm_type). */
int negf_broyden_create(int S, double fraction, int max_step, int norm_type,
                        negf_broyden **out);
int negf_broyden_destroy(negf_broyden *b);
int negf_broyden_set_nin(negf_broyden *b, const double *n_in);
int negf_broyden_get_nin(const negf_broyden *b, double *n_in);
/* One update from n_out (host array of S).  Returns NEGF_ENONCONV when the
 * history is full.  max_norm = Broyden_Norm, l2 = Broyden_NormSum_Curr;
 * norm_per_site (optional) = h_Norm. */
int negf_broyden_step(negf_broyden *b, const double *n_out, double *max_norm,
                      double *l2, double *norm_per_site);
/* Device form: n_out and the updated n_in stay on the device. */
int negf_broyden_step_device(negf_broyden *b, const double *d_n_out,
                             double *max_norm, double *l2, void *stream);
const double *negf_broyden_device_nin(const negf_broyden *b);
int negf_broyden_reset(negf_broyden *b, int keep_n_in, double init_value);
int negf_broyden_get_step(const negf_broyden *b);

/* ------------------------------------------------------------------------
 * SCF driver (c_TransportSolver::Solve Transport.cpp:359-539) over one tube
 * with a dense 1/r-with-image electrostatic stand-in for the MLMG solve and
 * gather: U_i = U_ext_i - V_gate - sum_j K_ij n_in_j.
 * ------------------------------------------------------------------------ */
typedef struct {
    double eps_r, R_ring_nm, R_image_nm;   /* Coulomb stand-in            */
    double Broyden_fraction, Broyden_max_norm;
    int Broyden_norm_type, Broyden_threshold_maxstep;
    int reset_with_previous_charge_distribution;
    double NS_initial_deposit_value;
} negf_scf_params;

typedef struct negf_scf negf_scf;
int negf_scf_create(negf_tube *t, const negf_scf_params *p, negf_scf **out);
int negf_scf_destroy(negf_scf *s);
/* One bias point: iterate until Broyden_Norm <= Broyden_max_norm, then
 * Reset_ForNextBiasStep.  U_out, Q_out: host arrays of N. */
int negf_scf_bias_point(negf_scf *s, double V_gate, double *U_out,
                        double *Q_out, int *iters, double *final_norm);
/* Fixed external potential U_ext (host array of N, eV) added to every
 * iterate: U = U_ext - V_gate - K n_in.  NULL removes it.  Stands in for the
 * fixed part of the electrostatic solution (work functions, fixed charges). */
int negf_scf_set_external_potential(negf_scf *s, const double *U_ext);
/* Dense kernel K (N x N, row-major, V per e) used by the stand-in. */
int negf_scf_kernel(const negf_scf *s, double *K);
/* No reference counterpart (performance only): on (the default), each
 * iteration of negf_scf_bias_point replays one captured hipGraph (potential,
 * solve, Broyden update, read-back), recaptured when the tube's paths
 * change; off, the same launches are enqueued one by one.  Results are the
 * same bit for bit. */
int negf_scf_use_graph(negf_scf *s, int on);

/* Batched SCF of equal tubes (no reference counterpart: one SCF loop per
 * tube there).  The tubes of batch b (same order) share one dense
 * electrostatic kernel K (N x N, as negf_scf); each tube t has its own
 * external potential U_ext[t], metal voltage V_gate[t] and Broyden history.
 * An iteration is one FP64 GEMM for all tubes, one batched NEGF solve and one
 * Broyden workgroup per tube, replayed from a captured hipGraph; a tube stops
 * iterating once its Broyden_Norm <= Broyden_max_norm.  Arrays are host,
 * tube-major ([T][N]). */
typedef struct negf_batch_scf negf_batch_scf;
int negf_batch_scf_create(negf_batch *b, negf_tube *const *tubes, int ntubes,
                          const negf_scf_params *p, negf_batch_scf **out);
int negf_batch_scf_destroy(negf_batch_scf *s);
int negf_batch_scf_set_external_potential(negf_batch_scf *s, const double *U_ext);
/* Broyden input charge (initial guess) of every tube, and read-back. */
int negf_batch_scf_set_input_charge(negf_batch_scf *s, const double *n_in);
int negf_batch_scf_get_input_charge(const negf_batch_scf *s, double *n_in);
int negf_batch_scf_use_graph(negf_batch_scf *s, int on);
/* Run all tubes to convergence from the present input charge (fresh Broyden
 * history); U_out, Q_out: [T][N] at each tube's converged iteration; iters,
 * final_norms: [T]; total_iterations: batched iterations run. */
int negf_batch_scf_run(negf_batch_scf *s, const double *V_gate, double *U_out,
                       double *Q_out, int *iters, double *final_norms,
                       int *total_iterations);

/* Self-consistent loop over the tubes of batch b (any lengths, same order)
 * with ONE Broyden system over their concatenated sites, as the reference's
 * transport driver runs several nanostructures: each tube's Solve_NEGF fills
 * its block of n_curr_out at its site offset (Transport.cpp:417-431,
 * Broyden_Parallel_General.cpp:42-66) and one Broyden update covers all of
 * them (Transport.cpp:435); Broyden_Norm is the maximum over all sites.  The
 * electrostatic stand-in is negf_scf's dense 1/r-with-image kernel within
 * each tube (applied as a direct sum, not stored), with a per-tube metal
 * voltage V_gate[t] and the external potential U_ext.  One iteration
 * (stand-in, batched NEGF solve, three Broyden launches) is replayed from a
 * captured hipGraph.  Arrays are host, concatenated in batch order
 * (negf_batch_offsets); S = total sites. */
typedef struct negf_multi_scf negf_multi_scf;
int negf_multi_scf_create(negf_batch *b, negf_tube *const *tubes, int ntubes,
                          const negf_scf_params *p, negf_multi_scf **out);
int negf_multi_scf_destroy(negf_multi_scf *s);
int negf_multi_scf_num_sites(const negf_multi_scf *s, long long *S);
int negf_multi_scf_set_external_potential(negf_multi_scf *s, const double *U_ext);
int negf_multi_scf_use_graph(negf_multi_scf *s, int on);
/* One bias point: iterate until Broyden_Norm <= Broyden_max_norm (or for
 * max_iters iterations when max_iters > 0), then Reset_ForNextBiasStep
 * (Transport.cpp:713-717).  U_out, Q_out (optional, [S]): the last
 * iteration's potential and NEGF charge; ms_per_iter: mean device time of
 * one iteration. */
int negf_multi_scf_bias_point(negf_multi_scf *s, const double *V_gate,
                              int max_iters, double *U_out, double *Q_out,
                              int *iters, double *final_norm,
                              double *ms_per_iter);
/* Broyden input charge n_in ([S]). */
int negf_multi_scf_get_input_charge(const negf_multi_scf *s, double *n_in);

/* ------------------------------------------------------------------------
 * Dense-block chains on FP64 MFMA.  The reference declares this path but
 * leaves it stubbed: the 2-D MatrixBlock multiply/inverse/dagger are empty
 * (Matrix_Block_Util.H:123-135, 431-438, 684-692) and DecimationTechnique
 * (NEGF_Common.cpp:4095-4153) works only for 1-D blocks.  A chain of
 * nblocks blocks of 1 <= m <= 32 orbitals has on-site H0 + diag(v_b) (H0 real
 * symmetric m x m, row-major; v_b real, per orbital) and the same real
 * coupling T = H_{b,b+1} between consecutive blocks; the leads continue the
 * chain with on-site H0 + Uc[0] (left) and H0 + Uc[1] (right).  Blocks of
 * m < 32 run as 32-wide blocks whose extra orbitals are decoupled (on-site
 * 1e4 eV, no couplings) and left out of every output, at the cost of m = 32.
 * ------------------------------------------------------------------------ */
typedef struct negf_block_chain negf_block_chain;
int negf_block_chain_create(int m, int nblocks, const double *H0, const double *T,
                            negf_block_chain **out);
int negf_block_chain_destroy(negf_block_chain *c);
/* v: [nblocks][m] orbital potentials (eV); Uc[2]: lead potentials. */
int negf_block_chain_set_potential(negf_block_chain *c, const double *v, const double *Uc);
/* For every energy E[e]: lead surface Green's functions by Sancho-Rubio
 * decimation (until max|alpha|, |beta| < sr_tol, at most sr_max_iter
 * iterations), Sigma_L = T^T g_L T, Sigma_R = T g_R T^T, then the block RGF
 * sweep for diag G^R.  acc[b*m + o] = sum_e Im(coef[e] G_e,bb,oo) (the block
 * form of the charge accumulation of NEGF_Common.cpp:3742-3754);
 * diagG [nE][nblocks*m] (may be NULL); sr_iters [nE][2] (may be NULL).
 * NEGF_ENONCONV if some decimation did not reach sr_tol. */
int negf_block_rgf(negf_block_chain *c, int nE, const negf_complex *E,
                   const negf_complex *coef, double sr_tol, int sr_max_iter,
                   double *acc, negf_complex *diagG, int *sr_iters);
/* The lead self-energies alone: sigma [nE][2 leads][re, im planes][m*m]
 * (row-major), sr_iters [nE][2] (may be NULL). */
int negf_block_self_energies(negf_block_chain *c, int nE, const negf_complex *E,
                             double sr_tol, int sr_max_iter, double *sigma, int *sr_iters);
/* Self-check of the device building blocks on one 32 x 32 complex matrix
 * pair (re plane then im plane, row-major): op 0: C = A B, op 1: C = A^-1
 * (pivoted Gauss-Jordan), op 2: C = A Re(B), each by a workgroup of four
 * wavefronts; ops 3, 4, 5: the same three by one wavefront (the forms the
 * block RGF sweep uses).  No reference counterpart. */
int negf_block_probe(int op, const double *A, const double *B, double *C);
/* Device time (ms) of the decimation and of the RGF sweep of the last call. */
int negf_block_last_time(const negf_block_chain *c, float *ms_sancho_rubio, float *ms_rgf);
/* Upper bound (bytes) on the device scratch that holds the stored g_b
 * (16 KiB per block and energy) and, when all energies fit in one sweep that
 * way, the stored products g_b T (another 16 KiB); negf_block_rgf runs the
 * energies in as many sweeps as needed to stay below it (default 80 GiB; at
 * least one energy per sweep).  Results do not depend on it.  No reference
 * counterpart. */
int negf_block_chain_set_scratch_limit(negf_block_chain *c, size_t bytes);
/* m (orbitals per block) and nblocks of a chain. */
int negf_block_chain_dims(const negf_block_chain *c, int *m, int *nblocks);

/* Block material: the tube material's equilibrium contour, enclosed poles,
 * Rho0 and charge combine over a block chain (Compute_RhoEq
 * NEGF_Common.cpp:3667-3754 on the paths of :1796-1800, Compute_GR_atPoles
 * :3833-3952, Compute_Rho0 :3957-4092, Q = RhoEq - Rho0 as in Solve_NEGF).
 * The contour settings come from p (contact_mu, contact_T, E_valence_min,
 * E_pole_max, E_zPlus_imag, Fermi_tail_factors, eq_integration_pts; the
 * tube-geometry fields are ignored); spin_degen multiplies every orbital.
 * Equilibrium contacts only: NEGF_ENOTSUP when |mu_L - mu_R| opens the
 * non-equilibrium window (the block sweep forms only diagonal blocks of G).
 * The chain c must outlive the material. */
typedef struct negf_block_material negf_block_material;
int negf_block_material_create(negf_block_chain *c, const negf_tube_params *p,
                               double spin_degen, double sr_tol, int sr_max_iter,
                               negf_block_material **out);
int negf_block_material_destroy(negf_block_material *m);
/* v: [nblocks][m] orbital potentials; Uc: lead potentials. */
int negf_block_material_set_potential(negf_block_material *m, const double *v,
                                      const double Uc[2]);
int negf_block_material_set_contacts(negf_block_material *m, const double mu[2],
                                     const double T[2]);
/* Rho0 per orbital (computed at creation, U = 0, mu = 0). */
int negf_block_material_rho0(const negf_block_material *m, double *rho0);
/* Q[b*m + o] = RhoEq - Rho0 at the current potential (Solve_NEGF,
 * Nanostructure.H:115-120); n_intg_pts: the contour's points
 * (get_Total_Integration_Pts, NEGF_Common.cpp:1826-1841; poles not counted). */
int negf_block_solve(negf_block_material *m, double *Q, int *n_intg_pts);
/* One bias point of the self-consistent loop over the orbitals with the
 * linear stand-in U_i = U_ext_i - V_gate - sum_j K_ij n_in_j (K: S x S
 * row-major, S = nblocks*m; U_ext may be NULL), lead potentials U of
 * orbital contact_orbital[0] of the first block and contact_orbital[1] of
 * the last (NULL: the end blocks' mean U), Broyden modified-second on n
 * (Broyden_* fields of sp; reset_with_previous_charge_distribution starts
 * from the last bias point's charge).  U_out, Q_out: S values;
 * NEGF_ENONCONV after max_iters.  No reference counterpart beyond the
 * loop of c_TransportSolver::Solve (Transport.cpp:359-539). */
int negf_block_scf_bias_point(negf_block_material *m, const double *K, const double *U_ext,
                              double V_gate, const int *contact_orbital,
                              const negf_scf_params *sp, int max_iters,
                              double *U_out, double *Q_out, int *iters, double *final_norm);

/* ------------------------------------------------------------------------
 * Charge surrogate (no reference counterpart; a new component next to the
 * NEGF charge path).  A 1D-convolutional network maps a tube's potential
 * U (eV, N sites) and one per-sample feature f (e.g. the drain bias) to an
 * estimate of the NEGF induced charge Q: 6 layers (3 -> 64, 4 x 64 -> 64,
 * 64 -> 1), kernel width 9, tanh, BF16 weights and activations with FP32
 * accumulation on MFMA.  It is trained on the device from (U, f, Q) pairs
 * that the NEGF solver produced, and used only to seed and precondition
 * the batched SCF.  Array arguments named d_* are device pointers,
 * sample-major ([B][N] doubles, f: [B] doubles or NULL); stream NULL means
 * the surrogate's own stream.
 * ------------------------------------------------------------------------ */
typedef struct negf_surrogate negf_surrogate;
int negf_surrogate_create(int N, unsigned long long seed, negf_surrogate **out);
int negf_surrogate_destroy(negf_surrogate *s);
int negf_surrogate_num_params(const negf_surrogate *s);
int negf_surrogate_num_sites(const negf_surrogate *s);
/* Flat FP32 parameters (per layer W[cout][9][cin] then b[cout]; cin 32 for
 * the first layer, cout 32 for the last, padding ignored) and the output
 * normalisation norm = {mean_Q, scale_Q}. */
int negf_surrogate_get_params(negf_surrogate *s, float *w, double *norm);
int negf_surrogate_set_params(negf_surrogate *s, const float *w, const double *norm);
/* Q estimate of B samples. */
int negf_surrogate_forward(negf_surrogate *s, int B, const double *d_U,
                           const double *d_feat, double *d_Q, void *stream);
/* Tangent dQ/dU . V at U (B samples). */
int negf_surrogate_jvp(negf_surrogate *s, int B, const double *d_U,
                       const double *d_feat, const double *d_V,
                       double *d_out, void *stream);
/* Normalised mean-squared loss against d_Q and its parameter gradient
 * (host float[num_params]). */
int negf_surrogate_gradient(negf_surrogate *s, int B, const double *d_U,
                            const double *d_feat, const double *d_Q,
                            float *grad, double *loss);
/* mean_Q and scale_Q from the charges of nsamples samples. */
int negf_surrogate_fit_output_scale(negf_surrogate *s, int nsamples, const double *d_Q);
/* Adam over shuffled minibatches; the learning rate goes linearly from lr
 * to lr_end; last_loss: the loss of the last minibatch. */
int negf_surrogate_train(negf_surrogate *s, int nsamples, const double *d_U,
                         const double *d_feat, const double *d_Q, int epochs,
                         int batch, double lr, double lr_end,
                         unsigned long long seed, double *last_loss);

/* ------------------------------------------------------------------------
 * Surrogate-seeded, preconditioned batched SCF (no reference counterpart).
 * The Broyden mixing term fraction * F of the reference update
 * (Broyden_Second_Parallel.cpp:172-186, :106-135) becomes P_t F with the
 * Kerker-type preconditioner P_t = Z diag(1 / (1 + c_t lam)) Z^T, where
 * K = Z diag(lam) Z^T is the shared electrostatic kernel and c_t the
 * site-averaged capacitance dQ/dU . 1 that the charge surrogate predicts
 * for tube t.  Every iterate is still judged by the exact NEGF residual
 * and the same Broyden_max_norm.
 * ------------------------------------------------------------------------ */

/* The shared kernel K (N x N, row-major) of a batched SCF. */
int negf_batch_scf_get_kernel(const negf_batch_scf *s, double *K);
/* Eigenbasis of K: Z (N x N row-major, column k = eigenvector k), lam[N].
 * The preconditioner I - Z_r diag(c lam / (1 + c lam)) Z_r^T keeps the modes
 * with lam >= 1e-2 max lam (the rest would change it by less than c lam). */
int negf_batch_scf_set_preconditioner(negf_batch_scf *s, const double *Z, const double *lam);
/* The surrogate (same N) and its per-tube feature feat[T] (host, or NULL). */
int negf_batch_scf_set_surrogate(negf_batch_scf *s, negf_surrogate *sg, const double *feat);
/* Run all tubes to Broyden_Norm <= Broyden_max_norm from the present input
 * charge: seed_iters preconditioned fixed-point half steps on the surrogate
 * charge first (the seed; n_in -= P_t (n_in - Q_surrogate) / 2), then
 * Broyden with P_t F on the exact NEGF charge.
 * c_t comes from the surrogate at the start, half way through the seed, at
 * the seed, and at iteration refresh_iter of the exact loop (fresh history,
 * n_in kept; <= 0: never).  Outputs as negf_batch_scf_run; capacitance [T]
 * (may be NULL): the c_t in use at the end. */
int negf_batch_scf_run_surrogate(negf_batch_scf *s, const double *V_gate, int seed_iters,
                                 int refresh_iter, double *U_out, double *Q_out, int *iters,
                                 double *final_norms, int *total_iterations,
                                 double *capacitance);

/* ------------------------------------------------------------------------
 * Output writers (NEGF_Common.cpp:1343-1418, 3044-3112, 4223-4254):
 * <prefix>_Qout.dat, <prefix>_U.dat, <prefix>_norm.dat in the reference's
 * column formats.
 * ------------------------------------------------------------------------ */
int negf_write_step_files(const char *prefix, int N, const double *ptd_nm,
                          const double *Q, const double *U,
                          const double *norm);

/* <prefix>_Qin.dat (Write_InputInducedCharge NEGF_Common.cpp:3081-3092). */
int negf_write_qin(const char *prefix, int N, const double *ptd_nm,
                   const double *Qin);
/* _chargeComp.dat (Write_ChargeComponents NEGF_Common.cpp:2999-3023):
 * Im RhoEq, Re RhoNonEq, Im Rho0 and Q per site. */
int negf_write_charge_components(const char *filename, int N,
                                 const double *ptd_nm, const double *eq_im,
                                 const double *noneq_re, const double *rho0_im,
                                 const double *Q);
/* I.dat header (Define_FileHeaderForCurrent NEGF_Common.cpp:73-82) and one
 * appended row per step (Write_Current :4818-4840). */
int negf_write_current_header(const char *filename);
int negf_write_current_line(const char *filename, int step, double Vds,
                            double Vgs, const double *I, double avg_intg_pts,
                            int total_iterations, double conductance);
/* Restart: read a charge distribution written as <prefix>_Qout.dat
 * (Read_Table1D NEGF_Common.cpp:236-286; header + N rows required). */
int negf_read_charge_distribution(const char *filename, int N, double *Q);
/* transport_char.dat (Write_Eql_Characteristics NEGF_Common.cpp:1931-1970):
 * header line, then per energy "E f_1 f_2 DOS T G" at setw 20/20/20/35/35/35
 * and the stream's default precision, f_k = 1/(exp((E - mu_k)/kT_k) + 1). */
int negf_write_transport_char(const char *filename, int nE, const double *E_real,
                              const double mu[2], const double kT[2],
                              const double *dos, const double *trans,
                              const double *cond);
/* <folder>/Ept_<e>.dat for e < nE (NEGF_Common.cpp:2778-2796 via
 * Write_Table1D :4223-4254): header "PTD LDOS_r at E=<E_e, %f>", then N rows
 * "position LDOS" at setprecision(15) setw(35).  ldos: [nE][N]. */
int negf_write_ldos(const char *folder, int nE, const double *E_real, int N,
                    const double *ptd_nm, const double *ldos);
/* Compute_Conductance (NEGF_Common.cpp:1873-1928): per-energy conductance
 * -T dF/dE (units of 2q^2/h, finite differences on the given grid) and its
 * sum times the reference's step E_0 - E_1. */
int negf_conductance(int nE, const double *E, const double *T, double mu0,
                     double kT0, double *cond, double *total);

#ifdef __cplusplus
}
#endif
#endif
