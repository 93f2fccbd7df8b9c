/*
 * latoracle.h -- CPU FP64 oracle for the lattice evolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C with OpenMP,
 * the arithmetic of the reference Fortran code's evolution path so that the
 * HIP implementation can be checked against it.  Only tests/, the smoke()
 * entry of __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  The product library (liblattice_amd.so) never links or calls it.
 *
 * Layout conventions (same as the reference, Fortran column-major):
 *   real field  f[i + nx*(j + ny*k)], i fastest;
 *   half-complex Fk[i + nxh*(j + ny*k)], nxh = nx/2+1 (FFTW r2c layout);
 *   field sets are SoA: fld[l*nvol + site].
 *
 * Pinning status per function is listed in the header of latoracle.c.
 */
#ifndef LATORACLE_H
#define LATORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT (FFTW conventions: forward e^{-2 pi i}, all unnormalised) ---- */
int  orc_fft_c2c_1d(int n, double *re_im /* 2n interleaved */, int sign);
int  orc_r2c_3d(int nx, int ny, int nz, const double *f, double *Fk /* 2*nxh*ny*nz */);
int  orc_c2r_3d(int nx, int ny, int nz, const double *Fk, double *f);

/* ---- Spectral operators ---- */
int  orc_laplacian_3d(int nx, int ny, int nz, double *f, double dk);
int  orc_gradient_squared_3d(int nx, int ny, int nz, const double *f, double *grad2, double dk);
int  orc_derivative_3d(int nx, int ny, int nz, const double *f, double *df, double dk, int axis);

/* ---- Finite-difference stencil: lap = STENCIL(c)/(cc*dx^2), periodic ---- */
int  orc_stencil_laplacian(int nx, int ny, int nz, const double *f, double *lap,
                           const double c[5], double dx);
/* sum over sites of STENCIL(c, GRAD2) (Discrete homlat grad_en form) */
double orc_stencil_grad2_sum(int nx, int ny, int nz, const double *f, const double c[5]);
/* Periodic ghost fill of an AoS ghosted array fld(nfld, 0:nx+1, 0:ny+1, 0:nz+1). */
int  orc_wrap_fields(int nfld, int nx, int ny, int nz, double *fld_ghost);

/* ---- Spectra ---- */
int  orc_num_bins(int nx, int ny, int nz);
int  orc_spectrum_3d(int nx, int ny, int nz, const double *Fk, double *S);
int  orc_crossspec_3d(int nx, int ny, int nz, const double *Fk1, const double *Fk2,
                      double *S_re, double *S_im);
/* Discrete-tree variants: kind 0 = `spectrum` (weighted, /n^6),
 * kind 1 = `spectrum_bin` (integer counts, bin floor(p)+1). */
int  orc_spectrum_discrete(int n, const double *Fk, double *S, int kind);
/* `nofft` variants on a half-complex array: kind 0 = spectrum_nofft,
 * kind 1 = spectrum_bin_nofft; and crossspec_nofft. */
int  orc_spectrum_discrete_k(int n, const double *Fk, double *S, int kind);
int  orc_crossspec_discrete_k(int n, const double *Fk1, const double *Fk2, double *S1, double *S2);

/* ---- Evolution state ---- */
enum { ORC_CONFORMAL = 0, ORC_COSMIC = 1 };
enum { ORC_LAP_SPECTRAL = 0, ORC_LAP_STENCIL = 1 };
enum { ORC_MODEL_PHI4 = 0, ORC_MODEL_TWOFIELD = 1 };

typedef struct orc_sim {
    int nx, ny, nz, nfld;
    double len;            /* box side; dx = len/nx, dk = 2 pi / len */
    int time_var;          /* ORC_CONFORMAL / ORC_COSMIC */
    int lap_kind;          /* ORC_LAP_SPECTRAL / ORC_LAP_STENCIL */
    double c[5];           /* stencil c0,c1,c2,c3,cc */
    int model;             /* ORC_MODEL_* */
    double mp[4];          /* phi4: mp[0]=lambda; twofield: mp[0]=m2, mp[1]=g2 */
    double *fld;           /* nfld*nvol */
    double *fldp;          /* nfld*nvol */
    double *scratch;       /* nvol, Laplacian work array */
    double yscl, ysclp;
    /* Metric perturbations u_ij (6 components xx,xy,xz,yy,yz,zz, SoA) and
     * their momenta.  gw = 0: absent.  Drift u += h pi_u / yscl^2 (the field
     * drift coefficient); kick pi_u += hA (lap u + gw_coup S_ij) with hA the
     * field kick's Laplacian coefficient and S_ij = sum_l d_i phi_l d_j phi_l
     * by second-order central differences.  No back-reaction on yscl. */
    int gw;
    double gw_coup;        /* 2 / M_p^2 */
    double *gwu, *gwp;     /* 6*nvol each */
} orc_sim;

int  orc_split(orc_sim *s, double h, int hamind);
int  orc_symp_o2step(orc_sim *s, double dt, double w1, double w2);
/* order 2,4,6,8; symp6_tree: 0 = openmp_dev constants, 1 = mpi_dev constants */
int  orc_symp(orc_sim *s, int order, double dt, int nsteps, int symp6_tree);
int  orc_calc_metric(orc_sim *s);

/* Energy/Friedmann diagnostics, out[ORC_NE] */
enum {
    ORC_E_A = 0, ORC_E_HUB, ORC_E_RHO, ORC_E_KE, ORC_E_PE, ORC_E_GE, ORC_E_GRAV,
    ORC_E_RESID, ORC_E_PMOM, ORC_E_MEAN1, ORC_E_MEAN2, ORC_E_GEGRAD, ORC_NE
};
int  orc_energies(orc_sim *s, double *out);

/* Discrete-tree diagnostics of make_output (Discrete/Entropy/iofile.f90:317-335)
 * on a stencil simulation.  log[12 + 2 nfld] = output_log_file's line
 * (:643): t, a, H, scalar_en, prsave, lnrhoave, eosave, KE, PE, GE (GRAD2),
 * grav_en, hamiltonian, <phi_l>, <pi_l>/a^3; lnrho[nvol], dlnrho[nvol] = the
 * two fields output_reg bins (:416-423); psd[4 ns_d] = the PSD columns of
 * those fields (get_fourier_stats, outputfile.f90:526-735, and
 * crossspec_nofft of (dlnrho, lnrho)). */
int  orc_discrete_output(orc_sim *s, double time, double *log, double *lnrho, double *dlnrho, double *psd);

/* Model plug-in */
double orc_potential(const orc_sim *s, const double *phi_site);
void   orc_modeldv(const orc_sim *s, const double *phi_site, double *dv);

/* ---- Gravitational waves (textbook; no reference arithmetic exists) ---- */
int  orc_gw_source(int nx, int ny, int nz, int nfld, const double *fld, double dk,
                   double *Sij /* 6*nvol: xx,xy,xz,yy,yz,zz */);
int  orc_tt_project(int nx, int ny, int nz, double *Uk /* 6 half-complex arrays */);
int  orc_gw_spectrum(int nx, int ny, int nz, const double *Uk, const double *dUk,
                     double dk, double *S);
/* S_ij = sum_l d_i phi_l d_j phi_l with d_a f = (f(x+e_a) - f(x-e_a)) / (2 dx), periodic */
int  orc_gw_source_fd(int nx, int ny, int nz, int nfld, const double *fld, double dx,
                      double *Sij /* 6*nvol: xx,xy,xz,yy,yz,zz */);

/* ---- Initial conditions (evolve_spectral.f90:421-506 sampler, Philox RNG) ---- */
void orc_philox_normal_pairs(unsigned long long seed, unsigned long long stream,
                             unsigned long long count, double *u1, double *u2);
int  orc_sample(int nx, int ny, int nz, double len, double gamma, double m2eff,
                double mpl, double kcut_fac, unsigned long long seed,
                unsigned long long stream, double *out);
/* the sampler's pieces: FFTW RODFT10 (DST-II) in place, and the real-space
 * convolution kernel before the noise multiply */
int  orc_dst2(int n, double *x);
int  orc_sample_kernel(int nx, int ny, int nz, double len, double gamma, double m2eff,
                       double mpl, double kcut_fac, double *out);

#ifdef __cplusplus
}
#endif
#endif
