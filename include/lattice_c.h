/*
 * lattice_c.h -- C ABI of the MI355X (gfx950) lattice evolution library.
 *
 * This is the drop-in boundary for the reference's Fortran evolution path
 * (jonathanbraden/lattice-dev).  Every entry point is a plain C function
 * taking plain pointers and sizes; status is an int (0 = ok, < 0 = error,
 * message via lat_last_error).  An opaque lat_ctx owns all device memory,
 * FFT twiddle tables and one HIP stream.  Host pointers are borrowed for the
 * duration of a call only.  A context is not thread-safe; use one per thread.
 *
 * Layout at this boundary is the reference's: Fortran column-major, x
 * fastest.  A real field f(nx,ny,nz) is f[i + nx*(j + ny*k)].  A half-complex
 * transform is FFTW's r2c layout Fk(nx/2+1, ny, nz) of interleaved doubles.
 * Field sets passed to lat_set_state/lat_get_state are the reference's AoS
 * field-fastest arrays fld(nfld, [0:]nx[+1], ...) with an optional ghost
 * layer of width `ghost` (0 or 1) on every side.
 *
 * Each declaration names the reference interface it replaces (file:line,
 * relative to the reference repository root).
 */
#ifndef LATTICE_C_H
#define LATTICE_C_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lat_ctx lat_ctx;

enum lat_precision { LAT_F64 = 0, LAT_MIXED = 1 };
enum lat_laplacian { LAT_LAP_SPECTRAL = 0, LAT_LAP_STENCIL = 1 };
enum lat_timevar { LAT_CONFORMAL = 0, LAT_COSMIC = 1 };
enum lat_model { LAT_MODEL_PHI4 = 0, LAT_MODEL_TWOFIELD = 1 };

/* Energy/Friedmann diagnostics slots of lat_energies(). */
enum lat_energy_slot {
    LAT_E_A = 0,      /* scale factor a                                   */
    LAT_E_HUBBLE,     /* H                                                */
    LAT_E_RHO,        /* rho = KE + PE + GE                               */
    LAT_E_KE,
    LAT_E_PE,
    LAT_E_GE,
    LAT_E_GRAV,       /* -3 H^2                                           */
    LAT_E_RESID,      /* (rho - 3H^2) / rho                               */
    LAT_E_PMOM,       /* -p^2/12/a^4 (conformal) or -3 p_y^2/16/a^2       */
    LAT_E_MEAN1,      /* <phi_1>                                          */
    LAT_E_MEAN2,      /* <phi_2>                                          */
    LAT_E_GEGRAD,     /* FD gradient energy, GRAD2 form (grad_en,
                         Discrete/Entropy/homlat.f90:88-115); stencil
                         contexts only, NaN for the spectral Laplacian    */
    LAT_NE
};

/* Runtime configuration (replaces the reference's compile-time params.f90,
 * macros.h and the HAM= model selection of Pseudospec/mpi_dev/Makefile:29-35). */
typedef struct lat_config {
    int nx, ny, nz;      /* lattice (params.f90:6); each must be 2^a * 3^b, 8..1024 */
    int nfld;            /* fields: 1 (phi4) or 2 (twofield)                     */
    double len;          /* box side L: dx = L/nx, dk = 2 pi / L (params.f90:11-13) */
    int precision;       /* LAT_F64 or LAT_MIXED                                  */
    int laplacian;       /* LAT_LAP_SPECTRAL (SPECTRAL) or LAT_LAP_STENCIL (DISCRETE) */
    double stencil[5];   /* c0,c1,c2,c3,cc (hamiltonian_conformal.f90:41)          */
    int time_var;        /* LAT_CONFORMAL (hamiltonian_conformal.f90) or LAT_COSMIC */
    int model;           /* LAT_MODEL_*                                           */
    double model_params[4]; /* phi4: lambda; twofield: m^2, g^2, m_chi^2 (the
                               chi mass of Discrete/Entropy/model.f90:35-55,
                               modparam.f90:9 m2psi; 0 = the Pseudospec form) */
    int symp6_tree;      /* symp6 w1 constant: 0 openmp_dev, 1 mpi_dev            */
    int gw;              /* 1: evolve the metric perturbations u_ij with the fields
                            (grav-waves/data-gw.f90:14 hij, pi_hij)                  */
    double gw_mpl;       /* reduced Planck mass in program units (0 means 1); the
                            u_ij source coefficient is 2 / gw_mpl^2                */
    int reserved[4];
} lat_config;

/* ---- lifetime ---- */
int  lat_create(const lat_config *cfg, lat_ctx **out);
void lat_destroy(lat_ctx *ctx);

/* ---- batches of realisations (ensembles, coupling scans) ----
 * No reference counterpart: the reference runs one realisation per process.
 * A batch holds nmembers independent realisations of one lattice size, each
 * with its own fields, FRW background and model parameters (model_params:
 * nmembers x 4, nullable = cfg->model_params for all).  lat_symp / lat_split /
 * lat_symp_o2step on a batch advance every member with one launch per sweep
 * (grid y = member).  Per-member calls (state, IC, energies, spectra) go
 * through lat_member views, which borrow the batch's memory and must not
 * outlive it; lat_destroy on a view frees the view and the scratch that its
 * own calls allocated on demand (e.g. the k-space arrays of the output
 * spectra of a mixed-form stencil batch, which holds none).  Members of a
 * batch share dt; the metric perturbations are not available in batches. */
int  lat_create_batch(const lat_config *cfg, int nmembers, const double *model_params, lat_ctx **out);
int  lat_num_members(const lat_ctx *ctx);
int  lat_member(lat_ctx *batch, int member, lat_ctx **view);

/* Measured HBM bandwidth (no reference counterpart): a plain 16-byte-per-lane
 * device copy of `bytes` (read + write counted) repeated `reps` times on the
 * context's stream; *gbps = 2 * bytes * reps / time. */
int  lat_copy_bandwidth(lat_ctx *ctx, size_t bytes, int reps, double *gbps);

/* Autotuner (no reference counterpart): for lattices of up to 256 per axis,
 * times two launch shapes of each sweep of the spectral Laplacian (x R2C and
 * x C2R rows per workgroup, y and z columns per tile) on the context's arrays
 * and keeps the faster; fields are not modified.  Ensembles tune every size
 * class at creation.  lat_tuning reports the choices (0 = default shape):
 * out[4] = {x R2C, y, z, x C2R}. */
/* Form of the spectral Laplacian in the field kicks (no reference
 * counterpart; the reference has the 3-D form only): 0 = five HBM sweeps
 * (x R2C, y, z * -k^2, y, x C2R + kick), 1 = the same operator as the sum of
 * one-axis second derivatives in three sweeps (x: drift + D_xx; y: + D_yy;
 * z: + D_zz, kick, reductions).  1 is for the mixed form at cubic 128, 256,
 * 512 or 1024 (error otherwise); default 0. */
int  lat_set_laplacian_form(lat_ctx *ctx, int form);
int  lat_laplacian_form(lat_ctx *ctx);
/* Row order of the k-space arrays inside the field force pipeline (no
 * reference counterpart; transforms handed to the caller are always in
 * FFTW order): 1 = blocked (the rows of 64 consecutive kz planes of one ky
 * adjacent, which shortens the z pass's stride and lengthens the y pass's;
 * cubic 256 and 1024), 0 = FFTW order (default).  The transforms agree
 * bitwise; the reductions are grouped differently, so the results agree to
 * rounding. */
int  lat_set_kspace_order(lat_ctx *ctx, int blocked);
int  lat_autotune(lat_ctx *ctx);
int  lat_tuning(lat_ctx *ctx, int *out);
int  lat_set_tuning(lat_ctx *ctx, const int *in);   /* in[4] of 0/1, as lat_tuning reports */
/* Placement tuner (no reference counterpart): how fast the sweeps stream
 * depends on where the field, momentum and k-space arrays land in HBM (the
 * x R2C of a 1024^3 mixed state ran 3.07-3.55 ms per launch over four
 * allocations in one process, each keeping its figure; the distance between
 * the field and momentum arrays sets it).  Allocates candidates - 1 further
 * sets of these arrays (fewer if memory runs out), each with a spacer of its
 * own size between its field and momentum arrays, times the sweeps of one
 * force evaluation on each and on the context's own (on copies of the
 * state), keeps the fastest set holding the state unchanged and frees the
 * others.  candidates in 1..8; ms_out (nullable) receives `candidates`
 * times, in ms per force evaluation, one per set (0: the context's own
 * arrays; -1: not allocated).  Returns the index of the set kept (>= 0) or
 * an error (< 0).  Single contexts only (not batches or member views);
 * device pointers from lat_device_state taken before are invalid
 * afterwards. */
int  lat_place_tune(lat_ctx *ctx, int candidates, float *ms_out);

/* Ensembles of uneven lattice sizes (no reference counterpart): nclass cubic
 * size classes, class k with n[k]^3 sites, count[k] members and box side
 * len[k] (nullable: base->len); model_params holds 4 values per member in
 * class-major order (nullable).  All classes share one HIP stream.
 * lat_ensemble_symp advances every member of class k by nsteps steps of
 * dt[k]: the first call with given (order, nsteps, dt) is captured into a
 * hipGraph, later calls replay it.  Two launch modes: every FFT sweep of a
 * force evaluation as ONE launch covering all classes in lockstep (up to 4
 * classes of spectral, cubic lattices of 16, 32, 48, 64, 96, 128, 192 or 256
 * sites per side), or one launch per sweep per class.  At creation the
 * autotuner times both and keeps the faster, and each class's arrays are
 * put on the fastest of eight HBM placements (as lat_place_tune does for a
 * single context).  lat_ensemble_single_launch(ens,
 * 1 or 0) forces a mode, (ens, -1) only queries; it returns 1 when the single
 * launch is in use (0 if not, or if the classes do not allow it).  Members
 * are addressed class-major through lat_ensemble_member views. */
typedef struct lat_ensemble lat_ensemble;
int  lat_ensemble_create(const lat_config *base, int nclass, const int *n, const int *count,
                         const double *len, const double *model_params, lat_ensemble **out);
void lat_ensemble_destroy(lat_ensemble *ens);
const char *lat_ensemble_last_error(const lat_ensemble *ens);
int  lat_ensemble_size(const lat_ensemble *ens);
int  lat_ensemble_member(lat_ensemble *ens, int index, lat_ctx **view);
int  lat_ensemble_use_graph(lat_ensemble *ens, int enable);   /* default 1 */
int  lat_ensemble_single_launch(lat_ensemble *ens, int enable); /* -1: query */
/* launch shapes of size class k, as lat_set_tuning / lat_tuning */
int  lat_ensemble_set_tuning(lat_ensemble *ens, int k, const int *in);
int  lat_ensemble_tuning(const lat_ensemble *ens, int k, int *out);
int  lat_ensemble_symp(lat_ensemble *ens, int order, const double *dt, int nsteps);
int  lat_ensemble_synchronize(lat_ensemble *ens);
const char *lat_last_error(const lat_ctx *ctx);
int  lat_synchronize(lat_ctx *ctx);

/* ---- spectral operators (module fftw3) ---- */
/* Replaces laplacian_3d(n1,n2,n3,f,Fk,dk,planf,planb),
 * Pseudospec/openmp_dev/fftw_mod.f90:177-207 (generic laplacian_spectral :13-15).
 * f (host, n1*n2*n3 doubles) is overwritten in place with its Laplacian.  The
 * reference's Fk scratch and FFTW plans are owned by the context. */
int  lat_laplacian_3d(lat_ctx *ctx, int n1, int n2, int n3, double *f, double dk);

/* Replaces gradient_squared_3d_spectral,
 * Pseudospec/mpi_dev/fftw_mod_mpi.f90:279-327 (correct-sign form). */
int  lat_gradient_squared_3d(lat_ctx *ctx, const double *f, double *grad2, double dk);

/* Replaces fftw_execute_dft_r2c(planf, f, Fk) as called at
 * Pseudospec/openmp_dev/fftw_mod.f90:191 and analysis_spec.f90:77, and
 * p3dfft_ftran_r2c (Discrete/Entropy/p3dfft/build/wrap.f:1-29): unnormalised
 * forward R2C into FFTW layout. */
int  lat_fft_r2c_3d(lat_ctx *ctx, const double *f, double *Fk);
/* Replaces fftw_execute_dft_c2r(planb, Fk, f) (fftw_mod.f90:203) and
 * p3dfft_btran_c2r: unnormalised backward C2R.  Fk is not modified. */
int  lat_fft_c2r_3d(lat_ctx *ctx, const double *Fk, double *f);

/* Finite-difference Laplacian STENCIL(c,LAPLACIAN)/(cc*dx^2) with periodic
 * wrap: Pseudospec/openmp_dev/macros.h:26-33 with wrap_fields
 * (hamiltonian_conformal.f90:309-330). */
int  lat_stencil_laplacian(lat_ctx *ctx, const double *f, double *lap);

/* ---- spectra (module analysis) ---- */
/* Replaces init_spectrum_3d(nx,ny,nz), analysis_spec.f90:48-57: returns ns. */
int  lat_num_bins(int nx, int ny, int nz);
/* Replaces spectrum_3d(S, f, Fk, plan), analysis_spec.f90:59-102.  S has ns
 * entries; Fk (nullable) receives the R2C of f in FFTW layout, as the
 * reference leaves it in its Fk argument. */
int  lat_spectrum_3d(lat_ctx *ctx, double *S, const double *f, double *Fk);
/* Replaces crossspec_3d(Fk1, Fk2, S_real, S_imag), analysis_spec.f90:145-178.
 * Fk1, Fk2 in FFTW layout. */
int  lat_crossspec_3d(lat_ctx *ctx, const double *Fk1, const double *Fk2, double *S_re, double *S_im);
/* Device-resident output spectra of make_output (evolve_spectral.f90:523-551):
 * out[0..ns) S_phi1, [ns..2ns) S_pi1, [2ns..3ns) Re S(pi1,phi1), [3ns..4ns) Im. */
int  lat_output_spectra(lat_ctx *ctx, int field, double *out);
/* Discrete-tree spectra on a cubic lattice (nx = ny = nz = n), with
 * ns_d = lat_num_bins_discrete(n) = int(sqrt(3) * (n/2) + 2) bins
 * (Discrete/Entropy/params.f90:45):
 *   kind 0 replaces spectrum(f, S), Discrete/Entropy/analysis.f90:89-114:
 *          two-bin weights, S / W / n^6;
 *   kind 1 replaces spectrum_bin(f, S), analysis.f90:22-52:
 *          integer count per floor(|k|/dk) bin, S / W. */
int  lat_num_bins_discrete(int n);
int  lat_spectrum_discrete(lat_ctx *ctx, const double *f, double *S, int kind);
/* The `nofft` variants on a given transform Fk (FFTW layout (n/2+1, n, n)),
 * summed over the stored half space without the Hermitian mirror:
 *   kind 0 replaces spectrum_nofft(famp, S), analysis.f90:118-147: two-bin
 *          weights, S / W / n^6;
 *   kind 1 replaces spectrum_bin_nofft(famp, S), analysis.f90:54-84: counts,
 *          x-Nyquist plane skipped (:69).  The reference bins the module
 *          array Fk there, not its argument famp; pass the array it holds. */
int  lat_spectrum_discrete_k(lat_ctx *ctx, const double *Fk, double *S, int kind);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  lat_spectrum_discrete_k's spectrum (kind 0 or 1, same
 * arithmetic, same order of additions) of a DEVICE-resident transform Fk_dev
 * (FFTW layout (n/2+1, n, n), rows unpadded) into the DEVICE array S_dev
 * (ns_d doubles), enqueued on the context's stream with no host
 * synchronisation and no allocation: the block histograms and channels live
 * in a library-static device array of 32 MiB per GPU, enough for every
 * supported lattice up to 1024^3 (larger needs are rejected with an error).
 * Every call on a GPU shares that array: calls that overlap on one GPU
 * (several contexts' streams, several host threads) race. */
int  lat_spectrum_discrete_k_noalloc(lat_ctx *ctx, const double *Fk_dev, double *S_dev, int kind);
/* Replaces crossspec_nofft(famp1, famp2, S1, S2), analysis.f90:149-186:
 * Re and Im of F1 conj(F2), two-bin weights, x-Nyquist plane skipped, S / W. */
int  lat_crossspec_discrete_k(lat_ctx *ctx, const double *Fk1, const double *Fk2, double *S1, double *S2);

/* ---- gravitational waves (module GW_data / GW_Evolution, grav-waves/) ----
 * State: hij(2,3,nx,ny,nz) and pi_hij(2,3,nx,ny,nz), the allocation shape of
 * grav-waves/data-gw.f90:28,33 (Fortran order, component pair fastest):
 * hij(1,b) is the diagonal u_bb, hij(2,1..3) the off-diagonals u_12, u_13, u_23.
 * Evolution (lat_symp and the split calls, when lat_config.gw = 1):
 *   drift  u += h pi_u / yscl^2       (the field drift coefficient; cosmic form
 *                                       of gw_kinetic_cosmic, evolve-gw.f90:12-19)
 *   kick   pi_u += hA (lap u + (2/M_p^2) S_ij),  S_ij = sum_l d_i phi_l d_j phi_l
 *          by central differences, hA the field kick's Laplacian coefficient
 *          (the source the stub gw_source_cosmic, evolve-gw.f90:21-36, leaves empty).
 * The metric perturbations do not act back on the background. */
int  lat_set_gw(lat_ctx *ctx, const double *hij, const double *pi_hij);
int  lat_get_gw(lat_ctx *ctx, double *hij, double *pi_hij);
/* S_ij of the resident fields, 6 SoA arrays of nvol: xx, xy, xz, yy, yz, zz. */
int  lat_gw_source(lat_ctx *ctx, double *Sij);
/* GW spectra of the resident u_ij: TT projection Lambda_ij,lm(k) = P_il P_jm
 * - P_ij P_lm / 2, P_ij = delta_ij - k_i k_j / k^2 (gw-projectors.f90:13-26,
 * a stub), then shell binning with the spectrum_3d weights:
 * out[0..ns) = sum_ij |Lambda pi_u^(k)|^2, out[ns..2ns) = sum_ij |Lambda u^(k)|^2. */
int  lat_gw_spectrum(lat_ctx *ctx, double *out);

/* Replaces make_output(time) with dump_rho (openmp_dev/evolve_spectral.f90:
 * 305-415,523-551; mpi_dev/evolve_spectral.f90:311-313,455-479): appends one
 * '(30(ES22.15,2X))' line to <dir>/energy_spec.out (time, a, rho, KE, PE, GE,
 * -3H^2, (rho-3H^2)/rho, -ysclp^2/12/a^4, <phi_1>, then <phi_2> for tree 0 =
 * openmp_dev or <pi_1> for tree 1 = mpi_dev) and the ns-line spectrum block
 * ((b-1) dk, S_phi1, S_pi1, Re and Im of the cross spectrum, 2*nfld-2 unset
 * columns written as 0) plus an empty record to <dir>/spectrum.out. */
int  lat_make_output(lat_ctx *ctx, double time, int tree, const char *dir);

/* Replaces make_output() of the Discrete tree (Discrete/Entropy/iofile.f90:
 * 317-335) on a cubic finite-difference context: get_variables (:432-582),
 * the PSD block of output_reg (:388-430; outputfile.f90:200-216,526-735) and
 * output_log_file (:617-647).  Appends one '(32g25.16e2)' line to
 * <dir>/LOG.out: t, a, H, <rho>, <P>, <ln rho> - ln<rho>, <P/rho>, KE, PE,
 * GE (GRAD2 form), -3H^2, rho - 3H^2, <phi_l> (l = 1..nfld), <pi_l>/a^3 --
 * with rho = KE + GE + PE and P = KE - PE - GE/3 per site (T = sum pi^2,
 * G = STENCIL(c,GRAD2), V = modelv = 2V; homlat.f90:32-44) -- and one block
 * of ns_d lines to <dir>/PSD: t, a, H, (k-1) dk and the spectra of
 * ln rho - <ln rho> and d ln rho/dt (iofile.f90:416-423) and the Re, Im of
 * their cross spectrum (half space without the x-Nyquist plane, two-bin
 * weights, S / W), then two empty records.  A new file starts with head()
 * (outputfile.f90:1216-1250) for one process: the revision line with L, dt,
 * mpl and seed, rank and decomposition lines, the model line and the
 * variable list.  vals (nullable) receives the LOG.out values (12 + 2 nfld),
 * psd (nullable) the four PSD columns (4 x ns_d, column-major as written);
 * dir = NULL computes without writing. */
int  lat_discrete_output(lat_ctx *ctx, double time, double dt, double mpl, int seed, const char *dir, double *vals,
                         double *psd);

/* Replace the checkpoint write/read of Discrete/Entropy/driver.f90:68-118
 * (make_check_pt / read_check_pt, scafld.f90:224-284) for one process:
 * <dir>/fld_<i>.raw and fldp_<i>.raw (field i = 1..nfld, native FP64, x
 * fastest, no ghosts) and <dir>/checkpt.dat ("yscl ysclp", "time",
 * "outframenum", list-directed).  lat_restore sets the state from them. */
int  lat_checkpoint(lat_ctx *ctx, const char *dir, double time, int outframenum);
int  lat_restore(lat_ctx *ctx, const char *dir, double *time, int *outframenum);

/* ---- learned corrector (no reference counterpart) ----
 * A 4-layer 3x3x3 periodic convolutional network (8 -> 32 -> 32 -> 32 -> 4
 * channels, SiLU, BF16 weights and activations, FP32 accumulation on MFMA)
 * that maps the state after one coarse leapfrog step of coarse_factor * dt to
 * the residual of coarse_factor Yoshida-4 steps of dt.  It is trained on the
 * GPU (hand-written forward, backward and Adam kernels) on pairs that the
 * plain integrator generates, and the plain integrator is the fallback when
 * the Friedmann-constraint change of a corrected step exceeds a threshold.
 * Contexts: single, two-field, FP64, cubic with n a multiple of 16.
 * Weights of layer l: cout x 27 cin with k = tap * cin + ci, tap =
 * (dz+1)*9 + (dy+1)*3 + (dx+1) for the neighbour at (x+dx, y+dy, z+dz);
 * input channels (phi, chi, pi_phi, pi_chi, <phi>, <pi_phi>, log10(g^2/m^2) - 5,
 * log10(s_chi) / 10), output channels (phi, chi, pi_phi, pi_chi).  The
 * network sees and corrects fluctuations only, at unit scale (the last four
 * input channels are uniform: what the normalisation removes and the
 * dynamics depend on): input (v - <v>) / s_v, target
 * (r - <r>) / (ratio s_v) per channel, with <.> the spatial mean and s_v the
 * spatial standard deviation of the coarse state at hand; a corrected step
 * adds ratio s_v (y - <y>), leaving the homogeneous mode at the coarse step's
 * value.  ratio[4] (residual fluctuation RMS / s_v) is fixed at the first
 * training pair, or by set_norm. */
typedef struct lat_corrector lat_corrector;
int  lat_corrector_create(unsigned long long seed, lat_corrector **out);
void lat_corrector_destroy(lat_corrector *corr);
const char *lat_corrector_last_error(const lat_corrector *corr);
int  lat_corrector_set_layer(lat_corrector *corr, int layer, const float *w, const float *b);
int  lat_corrector_get_layer(lat_corrector *corr, int layer, float *w, float *b);
int  lat_corrector_set_norm(lat_corrector *corr, const double *ratio);
int  lat_corrector_get_norm(const lat_corrector *corr, double *ratio);
/* normalised network output [site][4] for the current state of ctx */
int  lat_corrector_infer(lat_corrector *corr, lat_ctx *ctx, float *out);
/* loss (MSE) and gradients (per layer: weights then biases) against target [site][4] */
int  lat_corrector_gradients(lat_corrector *corr, lat_ctx *ctx, const float *target, double *loss, float *grads);
int  lat_corrector_train(lat_corrector *corr, lat_ctx *ctx, double dt, int coarse_factor, int npairs, double lr,
                         double *losses);
/* stats[3] = fallbacks, largest accepted constraint change, force evaluations */
int  lat_corrector_step(lat_corrector *corr, lat_ctx *ctx, double dt, int coarse_factor, int nsteps,
                        double threshold, double *stats);
/* lat_corrector_step with the plain integrator as judge of the error: every
 * judge_every-th accepted corrected step is also taken by the plain
 * integrator from the same start; the errors of the corrected step and of
 * the leapfrog step alone against it (per channel, RMS difference over the
 * plain step's spatial standard deviation, the largest over the four
 * channels) are measured and the plain step is kept.  Once a judged
 * corrected step's error exceeds err_tol or the leapfrog step's, the
 * corrector is switched off for the rest of the call.  stats[7] = fallbacks,
 * force evaluations, judged steps, largest judged error (corrected), step
 * index at which it was switched off (-1: never), largest accepted
 * constraint change, largest judged error (leapfrog alone). */
int  lat_corrector_step_judged(lat_corrector *corr, lat_ctx *ctx, double dt, int coarse_factor, int nsteps,
                               double threshold, int judge_every, double err_tol, double *stats);

/* ---- state (module Hamiltonian globals fld, fldp, yscl, ysclp,
 *      hamiltonian_conformal.f90:22-24; hamiltonian_cosmic.f90:27,31) ---- */
int  lat_set_state(lat_ctx *ctx, const double *fld, const double *fldp, int ghost,
                   double yscl, double ysclp);
int  lat_get_state(lat_ctx *ctx, double *fld, double *fldp, int ghost,
                   double *yscl, double *ysclp);
int  lat_get_scalars(lat_ctx *ctx, double *yscl, double *ysclp);
int  lat_set_scalars(lat_ctx *ctx, double yscl, double ysclp);

/* ---- evolution (module Hamiltonian / integrator) ---- */
/* Replaces Hamiltonian_Split(dt, hamind): hamiltonian_conformal.f90:66-81
 * (1 drift, 2 grav, 3 kick) / hamiltonian_cosmic.f90:74-89 (1 grav, 2 drift). */
int  lat_split(lat_ctx *ctx, double dt, int hamind);
/* Replaces symp_o2step(dt, w1, w2): hamiltonian_conformal.f90:50-64. */
int  lat_symp_o2step(lat_ctx *ctx, double dt, double w1, double w2);
/* Replaces symp2/symp4/symp6/symp8(dt, nsteps): Pseudospec/mpi_dev/integrator.f90:11-110
 * and openmp_dev/evolve_spectral.f90:129-227.  State stays on the device. */
int  lat_symp(lat_ctx *ctx, int order, double dt, int nsteps);
/* Host-only (no device work): the linear sequence of split operators that
 * lat_symp(order, dt, nsteps) executes, as (type, h) pairs with type 1 = drift
 * (fields_kin), 2 = grav (grav_kinetic), 3 = kick (fields_grad_pot), mapped
 * through Hamiltonian_Split's index order for time_var.  Returns the number of
 * operators (writes at most max_ops), or < 0 for a bad order. */
int  lat_schedule(int time_var, int order, double dt, int nsteps, int symp6_tree, int *types, double *h,
                  int max_ops);
/* Replaces calc_metric(): hamiltonian_conformal.f90:263-304 / hamiltonian_cosmic.f90:351-403. */
int  lat_calc_metric(lat_ctx *ctx);
/* Replaces dump_rho (openmp_dev/evolve_spectral.f90:305-415) / scalar_rho
 * (mpi_dev/hamiltonian_cosmic.f90:268-349) and the get_scale_factor,
 * get_hubble, grav_energy functions: out[LAT_NE]. */
int  lat_energies(lat_ctx *ctx, double *out);

/* ---- initial conditions ---- */
/* Replaces sample(gamma, m2eff) + the assignment in init_fields
 * (Pseudospec/openmp_dev/evolve_spectral.f90:262-303,421-506), on the device:
 * field `field` of fld (momentum = 0) or fldp (momentum = 1) becomes
 * mean + sample.  The compiler RNG of the reference is replaced by Philox-4x32-10
 * keyed by (seed, stream); mode (i,j,k) uses counter i + (nx/2+1)(j + ny k). */
int  lat_sample_field(lat_ctx *ctx, int field, int momentum, double mean, double gamma, double m2eff,
                      double mpl, double kcut_fac, unsigned long long seed, unsigned long long stream);

/* ---- device pointers for in-process callers (state stays resident) ---- */
/* Returns device addresses of the SoA state arrays (fld, fldp: nfld*nvol
 * elements of double for LAT_F64, float for LAT_MIXED fluctuations). */
int  lat_device_state(lat_ctx *ctx, void **fld, void **fldp);

/* ---- measurement (no counterpart in the reference, which only prints
 *      wall-clock times at evolve_spectral.f90:79,99) ---- */
enum lat_stage {
    LAT_ST_X_R2C = 0, /* x-direction R2C (drift fused)            */
    LAT_ST_Y_FWD,     /* y forward column FFT                      */
    LAT_ST_Z_LAP,     /* z forward * -k^2 * z inverse              */
    LAT_ST_Y_INV,     /* y inverse column FFT                      */
    LAT_ST_X_C2R,     /* x-direction C2R + kick + reductions       */
    LAT_ST_SCALAR,    /* background scalar update                  */
    LAT_ST_DRIFT,     /* standalone drift sweep                    */
    LAT_ST_STENCIL,   /* FD stencil + kick                         */
    LAT_ST_REDUCE,    /* state reductions                          */
    LAT_ST_GW_SRC,    /* u_ij source S_ij = sum d_i phi d_j phi     */
    LAT_ST_SEP_X,     /* separable Laplacian: drift + D_xx          */
    LAT_ST_SEP_Y,     /* separable Laplacian: lap += D_yy           */
    LAT_ST_SEP_Z,     /* separable Laplacian: D_zz + kick + reductions */
    LAT_NSTAGE
};
/* enable = 1 starts per-launch HIP-event timing on the context stream (and
 * clears the accumulators); 0 stops it. */
int  lat_profile(lat_ctx *ctx, int enable);
/* Accumulated kernel milliseconds and launch counts per stage. */
int  lat_profile_read(lat_ctx *ctx, double *ms, long long *launches);

#ifdef __cplusplus
}
#endif
#endif
