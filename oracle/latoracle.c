/*
 * latoracle.c -- CPU FP64 restatement of the reference evolution hot path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP library, never the product.
 *
 * Function -> reference source it restates -> what pins it (tests/):
 *   orc_fft_c2c_1d, orc_r2c_3d, orc_c2r_3d
 *       FFTW r2c/c2r as called at Pseudospec/openmp_dev/fftw_mod.f90:191,203
 *       (FFTW 3.3.2 is not vendored; its published DFT definition is restated)
 *       -> numpy.fft goldens (tests/golden/fft_*.npz), round trip, Parseval.
 *   orc_laplacian_3d   Pseudospec/openmp_dev/fftw_mod.f90:177-207
 *       -> plane-wave closed forms incl. Nyquist, brute-force DFT on 4^3.
 *   orc_gradient_squared_3d   Pseudospec/mpi_dev/fftw_mod_mpi.f90:279-327
 *       -> plane-wave closed form |grad cos(k.x)|^2.
 *   orc_stencil_laplacian / orc_stencil_grad2_sum
 *       Pseudospec/openmp_dev/macros.h:26-33, hamiltonian_conformal.f90:176-196,
 *       Discrete/Entropy/macros.h:19-39, homlat.f90:90-147
 *       -> stencil eigenvalue on plane waves; GE_grad == GE_lap identity.
 *   orc_wrap_fields    Pseudospec/openmp_dev/hamiltonian_conformal.f90:309-330
 *       -> ghost == periodic neighbour on random data.
 *   orc_num_bins, orc_spectrum_3d, orc_crossspec_3d
 *       Pseudospec/openmp_dev/analysis_spec.f90:48-57,59-102,145-178
 *       -> full-spectrum (numpy fftn) brute force; sum rule on W; plane wave.
 *   orc_spectrum_discrete   Discrete/Entropy/analysis.f90:22-52,89-114
 *   orc_spectrum_discrete_k / orc_crossspec_discrete_k  analysis.f90:54-84,118-186
 *       -> same brute force with integer counts / n^6 normalisation / half
 *          space without mirror.
 *   orc_split / orc_symp_o2step / orc_symp
 *       Pseudospec/openmp_dev/hamiltonian_conformal.f90:50-207,
 *       Pseudospec/mpi_dev/hamiltonian_cosmic.f90:58-213,
 *       Pseudospec/mpi_dev/integrator.f90:11-110,
 *       Pseudospec/openmp_dev/evolve_spectral.f90:129-227
 *       -> homogeneous limit vs scipy solve_ivp of the Hamiltonian ODEs;
 *          convergence order 2/4/6/8; time reversibility; Friedmann
 *          constraint conservation; free-field mode rotation.
 *   orc_calc_metric    hamiltonian_conformal.f90:263-304, hamiltonian_cosmic.f90:351-403
 *       -> residual == 0 after the call; homogeneous ODE initial data.
 *   orc_energies       Pseudospec/openmp_dev/evolve_spectral.f90:305-415,
 *       Pseudospec/mpi_dev/hamiltonian_cosmic.f90:268-349, Discrete/Entropy/homlat.f90:90-213
 *       -> Parseval GE == sum k^2|F|^2/N^6; KE/PE by direct numpy sums.
 *   orc_gw_source / orc_gw_source_fd / orc_tt_project / orc_gw_spectrum and
 *   the u_ij pieces of the drift and kick (gw = 1)
 *       no reference arithmetic (grav-waves/evolve-gw.f90:21-36 and
 *       gw-projectors.f90:13-26 are empty stubs; the drift follows the
 *       evolve-gw.f90:12-19 cosmic form h += pi_h dt / a^3): PARITY UNPINNED
 *       against the reference; pinned by projector invariants (transverse,
 *       traceless, idempotent), plane-wave source closed forms (spectral and
 *       central-difference), the free-wave cos(k t) solution in a static
 *       background, and the vanishing TT part of a one-dimensional source.
 *   orc_sample, orc_sample_kernel, orc_dst2
 *       Pseudospec/openmp_dev/evolve_spectral.f90:421-506
 *       with a Philox-4x32-10 stream in place of the compiler's random_number
 *       -> DST-II vs the direct O(n^2) sum and scipy; the kernel's lattice
 *          transform vs the continuum transform of its radial profile (norm,
 *          gamma, cutoff, interpolation); seed-averaged mode power vs
 *          2 |G_k|^2 inside and |G_k|^2 on the kx = 0 / Nyquist planes (the
 *          Rayleigh amplitude and the C2R's Hermitian handling).
 *   orc_discrete_output
 *       Discrete/Entropy/iofile.f90:388-582,617-647 (get_variables, output_reg,
 *       output_log_file), outputfile.f90:526-735 (the PSD column binning)
 *       -> per-site densities and pressures against numpy direct sums; the
 *          two spectra against numpy rfftn with the same weights (half space,
 *          no x-Nyquist plane).
 */
#include "latoracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define TWOPI_L 6.283185307179586476925286766559005768L

/* ------------------------------------------------------------------ */
/* 1-D complex FFT: iterative Stockham (radices 4,2,3,5, generic DFT)  */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    int nfac;
    int fac[64];
    cplx *w;      /* w[m] = exp(-2 pi i m / n), m = 0..n-1 */
    cplx *stw;    /* per stage f: stw[stoff[f] + k*(r-1) + q-1] = w[(k*q*n/(ns*r)) % n] */
    int stoff[64];
} fft_plan;

static void plan_init(fft_plan *p, int n)
{
    p->n = n;
    p->nfac = 0;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nfac++] = 2; m /= 2; }
    while (m % 3 == 0) { p->fac[p->nfac++] = 3; m /= 3; }
    while (m % 5 == 0) { p->fac[p->nfac++] = 5; m /= 5; }
    for (int f = 7; m > 1; f += 2)
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int k = 0; k < n; ++k) {
        long double ang = -TWOPI_L * (long double)k / (long double)n;
        p->w[k] = (double)cosl(ang) + I * (double)sinl(ang);
    }
    /* the stage twiddles, copied from w[] at the indices the butterflies use */
    size_t tot = 0;
    int ns = 1;
    for (int f = 0; f < p->nfac; ++f) {
        p->stoff[f] = (int)tot;
        tot += (size_t)ns * (size_t)(p->fac[f] - 1);
        ns *= p->fac[f];
    }
    p->stw = (cplx *)malloc(sizeof(cplx) * (tot > 0 ? tot : 1));
    ns = 1;
    for (int f = 0; f < p->nfac; ++f) {
        const int r = p->fac[f];
        const int tstride = n / (ns * r);
        for (int k = 0; k < ns; ++k)
            for (int q = 1; q < r; ++q)
                p->stw[p->stoff[f] + k * (r - 1) + q - 1] = p->w[(int)(((long long)k * q * tstride) % n)];
        ns *= r;
    }
}

static void plan_free(fft_plan *p) { free(p->w); free(p->stw); }

/* Out-of-place Stockham: x (length n, stride 1) -> result in x; y is scratch.
 * sign = -1 forward (FFTW_FORWARD), +1 backward. */
static void fft_exec(const fft_plan *p, cplx *x, cplx *y, int sign)
{
    const int n = p->n;
    cplx *in = x, *out = y;
    int ns = 1; /* length of already-combined sub-transforms */
    for (int f = 0; f < p->nfac; ++f) {
        const int r = p->fac[f];
        const int nr = n / r;
        const cplx *tws = p->stw + p->stoff[f];
        for (int j = 0; j < nr; ++j) {
            const int k = j % ns;
            cplx v[64];
            cplx t[64];
            /* gather with twiddles w^(k*q*n/(ns*r)) */
            const cplx *tk = tws + (size_t)k * (r - 1);
            v[0] = in[j] * p->w[0];
            for (int q = 1; q < r; ++q) {
                cplx tw = tk[q - 1];
                if (sign > 0) tw = conj(tw);
                v[q] = in[j + q * nr] * tw;
            }
            /* radix-r DFT */
            if (r == 2) {
                t[0] = v[0] + v[1];
                t[1] = v[0] - v[1];
            } else if (r == 4) {
                cplx a0 = v[0] + v[2], a1 = v[0] - v[2];
                cplx a2 = v[1] + v[3], a3 = v[1] - v[3];
                cplx rot = (sign < 0) ? (-I * a3) : (I * a3);
                t[0] = a0 + a2;
                t[2] = a0 - a2;
                t[1] = a1 + rot;
                t[3] = a1 - rot;
            } else {
                for (int u = 0; u < r; ++u) {
                    cplx s = 0;
                    for (int q = 0; q < r; ++q) {
                        int widx = (int)(((long long)u * q % r) * (n / r));
                        cplx tw = p->w[widx];
                        if (sign > 0) tw = conj(tw);
                        s += v[q] * tw;
                    }
                    t[u] = s;
                }
            }
            const int base = (j / ns) * ns * r + k;
            for (int q = 0; q < r; ++q) out[base + q * ns] = t[q];
        }
        cplx *tmp = in; in = out; out = tmp;
        ns *= r;
    }
    if (in != x) memcpy(x, in, sizeof(cplx) * (size_t)n);
}

int orc_fft_c2c_1d(int n, double *re_im, int sign)
{
    if (n <= 0) return -1;
    fft_plan p;
    plan_init(&p, n);
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *y = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int i = 0; i < n; ++i) x[i] = re_im[2 * i] + I * re_im[2 * i + 1];
    fft_exec(&p, x, y, sign);
    for (int i = 0; i < n; ++i) { re_im[2 * i] = creal(x[i]); re_im[2 * i + 1] = cimag(x[i]); }
    free(x); free(y); plan_free(&p);
    return 0;
}

/* complex FFT along y (stride nxh) and z (stride nxh*ny) of a half-complex
 * array.  Columns go through in groups of CB adjacent x-indices, gathered
 * row by row (CB consecutive complex values per row) into a contiguous
 * buffer, transformed one by one, and scattered back. */
#define CB 8
static void fft_cols(int nxh, int ncol, int nother, size_t s_col, size_t s_other, cplx *F, const fft_plan *p,
                     int sign)
{
    const int ngrp = (nxh + CB - 1) / CB;
#pragma omp parallel
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)ncol * CB);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)ncol);
#pragma omp for collapse(2) schedule(static)
        for (int o = 0; o < nother; ++o)
            for (int g = 0; g < ngrp; ++g) {
                const int i0 = g * CB, cnt = (nxh - i0) < CB ? (nxh - i0) : CB;
                cplx *col = F + i0 + s_other * o;
                for (int j = 0; j < ncol; ++j)
                    for (int c = 0; c < cnt; ++c) a[(size_t)c * ncol + j] = col[s_col * j + c];
                for (int c = 0; c < cnt; ++c) fft_exec(p, a + (size_t)c * ncol, b, sign);
                for (int j = 0; j < ncol; ++j)
                    for (int c = 0; c < cnt; ++c) col[s_col * j + c] = a[(size_t)c * ncol + j];
            }
        free(a); free(b);
    }
}
#undef CB

static void fft_yz(int nxh, int ny, int nz, cplx *F, int sign)
{
    fft_plan py, pz;
    plan_init(&py, ny);
    plan_init(&pz, nz);
    const size_t sy = (size_t)nxh, sz = (size_t)nxh * ny;
    if (sign < 0) {
        fft_cols(nxh, ny, nz, sy, sz, F, &py, sign);
        fft_cols(nxh, nz, ny, sz, sy, F, &pz, sign);
    } else {
        fft_cols(nxh, nz, ny, sz, sy, F, &pz, sign);
        fft_cols(nxh, ny, nz, sy, sz, F, &py, sign);
    }
    plan_free(&py);
    plan_free(&pz);
}

/* R2C: full complex FFT of each real x-row (imag 0), keep kx = 0..nx/2; then
 * complex FFTs along y and z.  FFTW/numpy rfftn convention, unnormalised. */
static void r2c_3d(int nx, int ny, int nz, const double *f, cplx *F)
{
    const int nxh = nx / 2 + 1;
    fft_plan px;
    plan_init(&px, nx);
#pragma omp parallel
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for schedule(static)
        for (long row = 0; row < (long)ny * nz; ++row) {
            const double *src = f + (size_t)nx * row;
            for (int i = 0; i < nx; ++i) a[i] = src[i];
            fft_exec(&px, a, b, -1);
            cplx *dst = F + (size_t)nxh * row;
            for (int i = 0; i < nxh; ++i) dst[i] = a[i];
        }
        free(a); free(b);
    }
    plan_free(&px);
    fft_yz(nxh, ny, nz, F, -1);
}

/* C2R: inverse complex FFTs along z and y, then per x-row the Hermitian
 * completion X[nx-k] = conj(X[k]) with Im X[0] = Im X[nx/2] = 0 (the
 * numpy irfftn / FFTW c2r row convention), inverse FFT, real part.
 * Unnormalised.  F is copied, not modified. */
static void c2r_3d(int nx, int ny, int nz, const cplx *Fin, double *f)
{
    const int nxh = nx / 2 + 1;
    size_t nk = (size_t)nxh * ny * nz;
    cplx *F = (cplx *)malloc(sizeof(cplx) * nk);
    memcpy(F, Fin, sizeof(cplx) * nk);
    fft_yz(nxh, ny, nz, F, +1);
    fft_plan px;
    plan_init(&px, nx);
#pragma omp parallel
    {
        cplx *a = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
        cplx *b = (cplx *)malloc(sizeof(cplx) * (size_t)nx);
#pragma omp for schedule(static)
        for (long row = 0; row < (long)ny * nz; ++row) {
            const cplx *src = F + (size_t)nxh * row;
            for (int i = 0; i < nxh; ++i) a[i] = src[i];
            a[0] = creal(a[0]);
            if (nx % 2 == 0) a[nx / 2] = creal(a[nx / 2]);
            for (int i = nxh; i < nx; ++i) a[i] = conj(src[nx - i]);
            fft_exec(&px, a, b, +1);
            double *dst = f + (size_t)nx * row;
            for (int i = 0; i < nx; ++i) dst[i] = creal(a[i]);
        }
        free(a); free(b);
    }
    plan_free(&px);
    free(F);
}

int orc_r2c_3d(int nx, int ny, int nz, const double *f, double *Fk)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
    r2c_3d(nx, ny, nz, f, (cplx *)Fk);
    return 0;
}

int orc_c2r_3d(int nx, int ny, int nz, const double *Fk, double *f)
{
    if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
    c2r_3d(nx, ny, nz, (const cplx *)Fk, f);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Spectral operators                                                  */
/* ------------------------------------------------------------------ */

/* wavenumber magnitude index as in fftw_mod.f90:193-194 (1-based j):
 * jj = j-1 if j <= n/2+1 else n+1-j.  0-based: m = j0 <= n/2 ? j0 : n-j0. */
static inline int wmag(int j0, int n) { return (j0 <= n / 2) ? j0 : n - j0; }
/* signed wavenumber (fftw_mod_mpi.f90:298,310 correct form) */
static inline int wsig(int j0, int n) { return (j0 <= n / 2) ? j0 : j0 - n; }

/* fftw_mod.f90:177-207: f <- F^-1[-dk^2 (ii^2+jj^2+kk^2) F[f]] / (n1 n2 n3),
 * Nyquist planes multiplied (not zeroed), normalisation as three divisions. */
int orc_laplacian_3d(int nx, int ny, int nz, double *f, double dk)
{
    const int nxh = nx / 2 + 1;
    size_t nk = (size_t)nxh * ny * nz;
    cplx *F = (cplx *)malloc(sizeof(cplx) * nk);
    r2c_3d(nx, ny, nz, f, F);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            int kk = wmag(k, nz), jj = wmag(j, ny);
            for (int i = 0; i < nxh; ++i) {
                double rad2 = (double)(i * i) + (double)(jj * jj) + (double)(kk * kk);
                F[i + (size_t)nxh * (j + (size_t)ny * k)] *= -rad2 * (dk * dk);
            }
        }
    c2r_3d(nx, ny, nz, F, f);
    size_t nv = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (size_t s = 0; s < nv; ++s) f[s] = f[s] / (double)nx / (double)ny / (double)nz;
    free(F);
    return 0;
}

/* d f / d x_axis spectrally, normalised by 1/(nx ny nz).  Building block of
 * gradient_squared_3d_spectral (fftw_mod_mpi.f90:279-327) and of the GW
 * source.  Signed wavenumbers as fftw_mod_mpi.f90:298,310: index j0 <= n/2
 * maps to +j0 (Nyquist kept with + sign), larger j0 to j0 - n; x uses i0. */
static void spectral_derivative(int nx, int ny, int nz, const cplx *F, double *df, double dk, int axis)
{
    const int nxh = nx / 2 + 1;
    size_t nk = (size_t)nxh * ny * nz;
    cplx *G = (cplx *)malloc(sizeof(cplx) * nk);
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nxh; ++i) {
                int w;
                if (axis == 0) w = i;
                else if (axis == 1) w = wsig(j, ny);
                else w = wsig(k, nz);
                size_t id = i + (size_t)nxh * (j + (size_t)ny * k);
                G[id] = I * ((double)w * dk) * F[id];
            }
    c2r_3d(nx, ny, nz, G, df);
    size_t nv = (size_t)nx * ny * nz;
    double norm = (double)nx * (double)ny * (double)nz;
#pragma omp parallel for schedule(static)
    for (size_t s = 0; s < nv; ++s) df[s] /= norm;
    free(G);
}

int orc_derivative_3d(int nx, int ny, int nz, const double *f, double *df, double dk, int axis)
{
    if (axis < 0 || axis > 2) return -1;
    const int nxh = nx / 2 + 1;
    cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * ny * nz);
    r2c_3d(nx, ny, nz, f, F);
    spectral_derivative(nx, ny, nz, F, df, dk, axis);
    free(F);
    return 0;
}

/* fftw_mod_mpi.f90:279-327: |grad f|^2 = sum_j (F^-1[i k_j F])^2 / N^6 */
int orc_gradient_squared_3d(int nx, int ny, int nz, const double *f, double *grad2, double dk)
{
    const int nxh = nx / 2 + 1;
    size_t nv = (size_t)nx * ny * nz;
    cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * ny * nz);
    double *d = (double *)malloc(sizeof(double) * nv);
    r2c_3d(nx, ny, nz, f, F);
    memset(grad2, 0, sizeof(double) * nv);
    for (int ax = 2; ax >= 0; --ax) {
        spectral_derivative(nx, ny, nz, F, d, dk, ax);
        for (size_t s = 0; s < nv; ++s) grad2[s] += d[s] * d[s];
    }
    free(F); free(d);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Finite-difference stencil (macros.h STENCIL with RANK0..3)          */
/* ------------------------------------------------------------------ */

static inline size_t pidx(int i, int j, int k, int nx, int ny, int nz)
{
    i = (i + nx) % nx; j = (j + ny) % ny; k = (k + nz) % nz;
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
}

static inline double stencil_site(const double *f, int i, int j, int k, int nx, int ny, int nz,
                                  const double c[5])
{
#define F(x, y, z) f[pidx(i + (x), j + (y), k + (z), nx, ny, nz)]
    double r0 = F(0, 0, 0);
    double r1 = F(-1, 0, 0) + F(1, 0, 0) + F(0, -1, 0) + F(0, 1, 0) + F(0, 0, -1) + F(0, 0, 1);
    double r2 = 0.0, r3 = 0.0;
    if (c[2] != 0.0)
        r2 = F(-1, -1, 0) + F(1, -1, 0) + F(-1, 1, 0) + F(1, 1, 0) + F(-1, 0, -1) + F(1, 0, -1) +
             F(-1, 0, 1) + F(1, 0, 1) + F(0, -1, -1) + F(0, 1, -1) + F(0, -1, 1) + F(0, 1, 1);
    if (c[3] != 0.0)
        r3 = F(-1, -1, -1) + F(1, -1, -1) + F(-1, 1, -1) + F(1, 1, -1) + F(-1, -1, 1) +
             F(1, -1, 1) + F(-1, 1, 1) + F(1, 1, 1);
#undef F
    return c[0] * r0 + c[1] * r1 + c[2] * r2 + c[3] * r3;
}

int orc_stencil_laplacian(int nx, int ny, int nz, const double *f, double *lap, const double c[5],
                          double dx)
{
    const double elap = 1.0 / (dx * dx) / c[4];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                lap[pidx(i, j, k, nx, ny, nz)] = stencil_site(f, i, j, k, nx, ny, nz, c) * elap;
    return 0;
}

/* homlat.f90:90-115 GRAD2 stencil: sum of c_r * (f(nbr)-f(0))^2 over ranks */
double orc_stencil_grad2_sum(int nx, int ny, int nz, const double *f, const double c[5])
{
    double total = 0.0;
    for (int k = 0; k < nz; ++k) {
        double plane = 0.0;
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                double f0 = f[pidx(i, j, k, nx, ny, nz)];
                double acc[4] = {0, 0, 0, 0};
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dxx = -1; dxx <= 1; ++dxx) {
                            int rank = abs(dxx) + abs(dy) + abs(dz);
                            if (rank == 0) continue;
                            double d = f[pidx(i + dxx, j + dy, k + dz, nx, ny, nz)] - f0;
                            acc[rank] += d * d;
                        }
                plane += c[1] * acc[1] + c[2] * acc[2] + c[3] * acc[3];
            }
        total += plane;
    }
    return total;
}

/* hamiltonian_conformal.f90:309-330: x faces, then y, then z (edges and
 * corners follow from the sequential copies). */
int orc_wrap_fields(int nfld, int nx, int ny, int nz, double *g)
{
    const int sx = nx + 2, sy = ny + 2, sz = nz + 2;
#define G(l, i, j, k) g[(l) + (size_t)nfld * ((i) + (size_t)sx * ((j) + (size_t)sy * (k)))]
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int l = 0; l < nfld; ++l) {
                G(l, 0, j, k) = G(l, nx, j, k);
                G(l, nx + 1, j, k) = G(l, 1, j, k);
            }
    for (int k = 0; k < sz; ++k)
        for (int i = 0; i < sx; ++i)
            for (int l = 0; l < nfld; ++l) {
                G(l, i, 0, k) = G(l, i, ny, k);
                G(l, i, ny + 1, k) = G(l, i, 1, k);
            }
    for (int j = 0; j < sy; ++j)
        for (int i = 0; i < sx; ++i)
            for (int l = 0; l < nfld; ++l) {
                G(l, i, j, 0) = G(l, i, j, nz);
                G(l, i, j, nz + 1) = G(l, i, j, 1);
            }
#undef G
    return 0;
}

/* ------------------------------------------------------------------ */
/* Spectra (analysis_spec.f90)                                          */
/* ------------------------------------------------------------------ */

/* analysis_spec.f90:48-57 */
int orc_num_bins(int nx, int ny, int nz)
{
    int ns = (nx / 2 + 1) * (nx / 2 + 1) + (ny / 2 + 1) * (ny / 2 + 1) + (nz / 2 + 1) * (nz / 2 + 1);
    return (int)floor(sqrt((double)ns + 1.0));
}

/* analysis_spec.f90:59-102 (binning part; the caller supplies Fk = R2C(f)).
 * Loop over all nx x-indices, i > nx/2 re-reads Fk(nx-i) at the same (j,k).
 * Per-plane partial bins are summed in plane order (deterministic). */
static void bin_modes(int nx, int ny, int nz, const cplx *F1, const cplx *F2, double *S1, double *S2,
                      double *W, int ns)
{
    const int nxh = nx / 2 + 1;
    double *part = (double *)calloc((size_t)nz * 3 * ns, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double *ps1 = part + (size_t)k * 3 * ns, *ps2 = ps1 + ns, *pw = ps2 + ns;
        int kk = wmag(k, nz);
        for (int j = 0; j < ny; ++j) {
            int jj = wmag(j, ny);
            for (int i0 = 0; i0 < nx; ++i0) {
                int ii = wmag(i0, nx);
                double p = sqrt((double)(ii * ii + jj * jj + kk * kk));
                int l = (int)floor(p);
                double d1 = (double)l - p, d2 = (double)(l + 1) - p;
                double c1 = (1.0 - d1 * d1) * (1.0 - d1 * d1);
                double c2 = (1.0 - d2 * d2) * (1.0 - d2 * d2);
                size_t id = ii + (size_t)nxh * (j + (size_t)ny * k);
                cplx a = F1[id] * conj(F2 ? F2[id] : F1[id]);
                ps1[l] += c1 * creal(a);
                if (S2) ps2[l] += c1 * cimag(a);
                pw[l] += c1;
                /* On non-cubic lattices the top shell's second bin is l+1 = ns,
                 * past the end of S(ns) (the reference writes out of bounds
                 * there); that weight is dropped. */
                if (l + 1 < ns) {
                    ps1[l + 1] += c2 * creal(a);
                    if (S2) ps2[l + 1] += c2 * cimag(a);
                    pw[l + 1] += c2;
                }
            }
        }
    }
    memset(S1, 0, sizeof(double) * ns);
    if (S2) memset(S2, 0, sizeof(double) * ns);
    memset(W, 0, sizeof(double) * ns);
    for (int k = 0; k < nz; ++k) {
        double *ps1 = part + (size_t)k * 3 * ns, *ps2 = ps1 + ns, *pw = ps2 + ns;
        for (int b = 0; b < ns; ++b) {
            S1[b] += ps1[b];
            if (S2) S2[b] += ps2[b];
            W[b] += pw[b];
        }
    }
    free(part);
}

int orc_spectrum_3d(int nx, int ny, int nz, const double *Fk, double *S)
{
    int ns = orc_num_bins(nx, ny, nz);
    double *W = (double *)malloc(sizeof(double) * ns);
    bin_modes(nx, ny, nz, (const cplx *)Fk, NULL, S, NULL, W, ns);
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0) S[b] = S[b] / W[b];
    free(W);
    return ns;
}

/* analysis_spec.f90:145-178 */
int orc_crossspec_3d(int nx, int ny, int nz, const double *Fk1, const double *Fk2, double *S_re,
                     double *S_im)
{
    int ns = orc_num_bins(nx, ny, nz);
    double *W = (double *)malloc(sizeof(double) * ns);
    bin_modes(nx, ny, nz, (const cplx *)Fk1, (const cplx *)Fk2, S_re, S_im, W, ns);
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0) { S_re[b] /= W[b]; S_im[b] /= W[b]; }
    free(W);
    return ns;
}

/* Discrete/Entropy/analysis.f90: ns = sqrt3*(n/2)+2 (params.f90:45).
 * kind 0: `spectrum` (:89-114) full-x weighted, S/W/n^6.
 * kind 1: `spectrum_bin` (:22-52) integer counts, bin floor(p)+1, full x. */
int orc_spectrum_discrete(int n, const double *Fk, double *S, int kind)
{
    const double sqrt3 = 1.7320508075688772935274463415059;
    int ns = (int)(sqrt3 * (double)(n / 2) + 2.0);
    const int nxh = n / 2 + 1;
    const cplx *F = (const cplx *)Fk;
    double *W = (double *)calloc((size_t)ns, sizeof(double));
    memset(S, 0, sizeof(double) * ns);
    for (int k = 0; k < n; ++k) {
        int kk = wmag(k, n);
        for (int j = 0; j < n; ++j) {
            int jj = wmag(j, n);
            for (int i0 = 0; i0 < n; ++i0) {
                int ii = wmag(i0, n);
                size_t id = ii + (size_t)nxh * (j + (size_t)n * k);
                double a2 = creal(F[id] * conj(F[id]));
                double p = sqrt((double)(ii * ii + jj * jj + kk * kk));
                if (kind == 1) {
                    int l = (int)floor(p); /* S(l+1) with 1-based l+1 */
                    S[l] += a2;
                    W[l] += 1.0;
                } else {
                    int l = (int)floor(p);
                    double d1 = (double)l - p, d2 = (double)(l + 1) - p;
                    double c1 = (1.0 - d1 * d1) * (1.0 - d1 * d1);
                    double c2 = (1.0 - d2 * d2) * (1.0 - d2 * d2);
                    S[l] += c1 * a2; W[l] += c1;
                    if (l + 1 < ns) { S[l + 1] += c2 * a2; W[l + 1] += c2; }
                }
            }
        }
    }
    double n6 = pow((double)n, 6);
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0) S[b] = (kind == 1) ? S[b] / W[b] : S[b] / W[b] / n6;
    free(W);
    return ns;
}

/* Discrete/Entropy/analysis.f90 `nofft` variants on a given half-complex
 * array: the loop runs over the stored half space only (i = 1..nn, no mirror,
 * weight 1 per stored mode).
 *   kind 0: spectrum_nofft (:118-147): all stored x-indices, two-bin weights,
 *           S/W/n^6;
 *   kind 1: spectrum_bin_nofft (:54-84): x-index < nn (the x-Nyquist plane
 *           skipped, :69), integer count per floor(p) bin, S/count. */
int orc_spectrum_discrete_k(int n, const double *Fk, double *S, int kind)
{
    const double sqrt3 = 1.7320508075688772935274463415059;
    int ns = (int)(sqrt3 * (double)(n / 2) + 2.0);
    const int nxh = n / 2 + 1;
    const cplx *F = (const cplx *)Fk;
    double *W = (double *)calloc((size_t)ns, sizeof(double));
    memset(S, 0, sizeof(double) * ns);
    for (int k = 0; k < n; ++k) {
        int kk = wmag(k, n);
        for (int j = 0; j < n; ++j) {
            int jj = wmag(j, n);
            for (int ii = 0; ii < nxh; ++ii) {
                if (kind == 1 && ii == n / 2) continue;
                size_t id = ii + (size_t)nxh * (j + (size_t)n * k);
                double a2 = creal(F[id] * conj(F[id]));
                double p = sqrt((double)(ii * ii + jj * jj + kk * kk));
                int l = (int)floor(p);
                if (kind == 1) {
                    S[l] += a2;
                    W[l] += 1.0;
                } else {
                    double d1 = (double)l - p, d2 = (double)(l + 1) - p;
                    double c1 = (1.0 - d1 * d1) * (1.0 - d1 * d1);
                    double c2 = (1.0 - d2 * d2) * (1.0 - d2 * d2);
                    S[l] += c1 * a2; W[l] += c1;
                    if (l + 1 < ns) { S[l + 1] += c2 * a2; W[l + 1] += c2; }
                }
            }
        }
    }
    double n6 = pow((double)n, 6);
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0) S[b] = (kind == 1) ? S[b] / W[b] : S[b] / W[b] / n6;
    free(W);
    return ns;
}

/* crossspec_nofft (Discrete/Entropy/analysis.f90:149-186): stored half space
 * with the x-Nyquist plane skipped, two-bin weights on F1 conj(F2), S/W. */
int orc_crossspec_discrete_k(int n, const double *Fk1, const double *Fk2, double *S1, double *S2)
{
    const double sqrt3 = 1.7320508075688772935274463415059;
    int ns = (int)(sqrt3 * (double)(n / 2) + 2.0);
    const int nxh = n / 2 + 1;
    const cplx *F1 = (const cplx *)Fk1, *F2 = (const cplx *)Fk2;
    double *W = (double *)calloc((size_t)ns, sizeof(double));
    memset(S1, 0, sizeof(double) * ns);
    memset(S2, 0, sizeof(double) * ns);
    for (int k = 0; k < n; ++k) {
        int kk = wmag(k, n);
        for (int j = 0; j < n; ++j) {
            int jj = wmag(j, n);
            for (int ii = 0; ii < n / 2; ++ii) {
                size_t id = ii + (size_t)nxh * (j + (size_t)n * k);
                cplx a = F1[id] * conj(F2[id]);
                double p = sqrt((double)(ii * ii + jj * jj + kk * kk));
                int l = (int)floor(p);
                double d1 = (double)l - p, d2 = (double)(l + 1) - p;
                double c1 = (1.0 - d1 * d1) * (1.0 - d1 * d1);
                double c2 = (1.0 - d2 * d2) * (1.0 - d2 * d2);
                S1[l] += c1 * creal(a); S2[l] += c1 * cimag(a); W[l] += c1;
                if (l + 1 < ns) { S1[l + 1] += c2 * creal(a); S2[l + 1] += c2 * cimag(a); W[l + 1] += c2; }
            }
        }
    }
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0) { S1[b] /= W[b]; S2[b] /= W[b]; }
    free(W);
    return ns;
}

/* ------------------------------------------------------------------ */
/* Model plug-in                                                       */
/* ------------------------------------------------------------------ */

/* hamiltonian_conformal.f90:215-240 (V = phi^4/4, lambda scaled) and
 * hamiltonian_cosmic.f90:215-232 (V = m2 phi^2/2 + g2 phi^2 chi^2 / 2), with
 * the chi mass term m2psi psi^2 of Discrete/Entropy/model.f90:35-55
 * (modelv = 2V = (m2phi + g2 psi^2) phi^2 + m2psi psi^2) as mp[2]. */
double orc_potential(const orc_sim *s, const double *f)
{
    if (s->model == ORC_MODEL_PHI4) return 0.25 * s->mp[0] * (f[0] * f[0]) * (f[0] * f[0]);
    double p2 = f[0] * f[0], c2 = f[1] * f[1];
    return 0.5 * s->mp[0] * p2 + 0.5 * s->mp[1] * p2 * c2 + 0.5 * s->mp[2] * c2;
}

void orc_modeldv(const orc_sim *s, const double *f, double *dv)
{
    if (s->model == ORC_MODEL_PHI4) { dv[0] = s->mp[0] * f[0] * f[0] * f[0]; return; }
    dv[0] = s->mp[0] * f[0] + s->mp[1] * f[0] * (f[1] * f[1]);
    dv[1] = s->mp[1] * (f[0] * f[0]) * f[1] + s->mp[2] * f[1];
}

/* ------------------------------------------------------------------ */
/* Split pieces                                                        */
/* ------------------------------------------------------------------ */

static inline size_t nvol_of(const orc_sim *s) { return (size_t)s->nx * s->ny * s->nz; }
static inline double dx_of(const orc_sim *s) { return s->len / (double)s->nx; }
static inline double dk_of(const orc_sim *s)
{
    return 6.2831853071795864769252867665590 / s->len;
}
static inline double scale_factor(const orc_sim *s)
{
    return s->time_var == ORC_CONFORMAL ? s->yscl : pow(s->yscl, 2.0 / 3.0);
}

/* Laplacian of field l into s->scratch (spectral: fftw_mod.f90:177-207;
 * stencil: STENCIL(c)/(cc dx^2), the a-dependence applied by the caller). */
static void field_laplacian(orc_sim *s, int l)
{
    size_t nv = nvol_of(s);
    const double *f = s->fld + (size_t)l * nv;
    if (s->lap_kind == ORC_LAP_SPECTRAL) {
        memcpy(s->scratch, f, sizeof(double) * nv);
        orc_laplacian_3d(s->nx, s->ny, s->nz, s->scratch, dk_of(s));
    } else {
        orc_stencil_laplacian(s->nx, s->ny, s->nz, f, s->scratch, s->c, dx_of(s));
    }
}

/* ordered sum over z-planes of per-plane partials */
static double plane_sum(const double *part, int nz)
{
    double t = 0.0;
    for (int k = 0; k < nz; ++k) t += part[k];
    return t;
}

/* Hamiltonian_fields_kin: hamiltonian_conformal.f90:92-118 /
 * hamiltonian_cosmic.f90:100-127 (identical arithmetic; yscl is a or y). */
static void fields_kin(orc_sim *s, double h)
{
    size_t nv = nvol_of(s);
    const size_t plane = (size_t)s->nx * s->ny;
    double *part = (double *)calloc((size_t)s->nz, sizeof(double));
    const double y2 = s->yscl * s->yscl;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s->nz; ++k) {
        double acc = 0.0;
        for (int l = 0; l < s->nfld; ++l) {
            double *f = s->fld + (size_t)l * nv + plane * k;
            const double *p = s->fldp + (size_t)l * nv + plane * k;
            for (size_t q = 0; q < plane; ++q) {
                f[q] = f[q] + h * p[q] / y2;
                acc += p[q] * p[q];
            }
        }
        part[k] = acc;
    }
    double KE2 = plane_sum(part, s->nz) / (double)nv;
    free(part);
    s->ysclp = s->ysclp + h * KE2 / (s->yscl * s->yscl * s->yscl);
    if (s->gw) {
        /* u_ij drift with the field drift coefficient; no back-reaction */
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < 6 * nv; ++q) s->gwu[q] = s->gwu[q] + h * s->gwp[q] / y2;
    }
}

/* Hamiltonian_grav_kinetic: conformal a -= h p_a/6 (hamiltonian_conformal.f90:86-90);
 * cosmic y -= h p_y 3/8 (hamiltonian_cosmic.f90:94-98). */
static void grav_kinetic(orc_sim *s, double h)
{
    if (s->time_var == ORC_CONFORMAL) s->yscl = s->yscl - h * s->ysclp / 6.0;
    else s->yscl = s->yscl - h * s->ysclp * (3.0 / 8.0);
}

/* GW kick pi_u += hA (lap u_c + gw_coup S_c), hA = h a^2 (conformal) or
 * h y^2 / a^2 (cosmic): the coefficient of the Laplacian in the field kick
 * (hamiltonian_conformal.f90:183-186, hamiltonian_cosmic.f90:153-157).
 * No reference arithmetic exists (grav-waves/evolve-gw.f90:21-36 is a stub). */
static void gw_kick(orc_sim *s, double h)
{
    size_t nv = nvol_of(s);
    const double a = scale_factor(s), y = s->yscl;
    const double hA = s->time_var == ORC_CONFORMAL ? h * (y * y) : h * (y * y) / (a * a);
    double *S = (double *)malloc(sizeof(double) * 6 * nv);
    orc_gw_source_fd(s->nx, s->ny, s->nz, s->nfld, s->fld, dx_of(s), S);
    for (int c = 0; c < 6; ++c) {
        const double *u = s->gwu + (size_t)c * nv;
        if (s->lap_kind == ORC_LAP_SPECTRAL) {
            memcpy(s->scratch, u, sizeof(double) * nv);
            orc_laplacian_3d(s->nx, s->ny, s->nz, s->scratch, dk_of(s));
        } else {
            orc_stencil_laplacian(s->nx, s->ny, s->nz, u, s->scratch, s->c, dx_of(s));
        }
        double *p = s->gwp + (size_t)c * nv;
        const double *src = S + (size_t)c * nv;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < nv; ++q) p[q] = p[q] + hA * (s->scratch[q] + s->gw_coup * src[q]);
    }
    free(S);
}

/* Hamiltonian_fields_grad_pot: hamiltonian_conformal.f90:120-207,
 * hamiltonian_cosmic.f90:129-213 (spectral and DISCRETE branches). */
static void fields_grad_pot(orc_sim *s, double h)
{
    size_t nv = nvol_of(s);
    const size_t plane = (size_t)s->nx * s->ny;
    const int nz = s->nz;
    const double a = scale_factor(s);
    double *ge_part = (double *)calloc((size_t)nz, sizeof(double));
    double *pe_part = (double *)calloc((size_t)nz, sizeof(double));
    double GE2 = 0.0;
    /* the kick of field l reads only fld (unchanged by the kick) */
    for (int l = 0; l < s->nfld; ++l) {
        field_laplacian(s, l);
        double *lap = s->scratch;
        if (s->time_var == ORC_COSMIC) {
            const double a2 = a * a;
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < nv; ++q) lap[q] = lap[q] / a2; /* hamiltonian_cosmic.f90:153 */
        }
        const double y = s->yscl;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k) {
            double acc = 0.0;
            for (size_t q = plane * k; q < plane * (k + 1); ++q) {
                double fs[2], dv[2];
                for (int m = 0; m < s->nfld; ++m) fs[m] = s->fld[(size_t)m * nv + q];
                orc_modeldv(s, fs, dv);
                double *p = s->fldp + (size_t)l * nv + q;
                if (s->time_var == ORC_CONFORMAL)
                    *p = *p + h * ((y * y) * lap[q] - (y * y) * (y * y) * dv[l]);
                else
                    *p = *p + h * (y * y) * (lap[q] - dv[l]);
                acc -= fs[l] * lap[q];
            }
            ge_part[k] = acc;
        }
        GE2 += plane_sum(ge_part, nz);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double acc = 0.0;
        for (size_t q = plane * k; q < plane * (k + 1); ++q) {
            double fs[2];
            for (int m = 0; m < s->nfld; ++m) fs[m] = s->fld[(size_t)m * nv + q];
            acc += orc_potential(s, fs);
        }
        pe_part[k] = acc;
    }
    double PE = plane_sum(pe_part, nz);
    free(ge_part); free(pe_part);
    if (s->gw) gw_kick(s, h);
    const double y = s->yscl;
    if (s->time_var == ORC_CONFORMAL) {
        GE2 = GE2 / (y * y);
        PE = PE / (double)nv;
        GE2 = GE2 / (double)nv;
        s->ysclp = s->ysclp - (y * y * y) * (GE2 + 4.0 * PE) * h;
    } else {
        PE = PE / (double)nv;
        GE2 = GE2 / (double)nv;
        s->ysclp = s->ysclp - y * ((1.0 / 3.0) * GE2 + 2.0 * PE) * h;
    }
}

/* Hamiltonian_Split: conformal 1=drift 2=grav 3=kick (hamiltonian_conformal.f90:66-81);
 * cosmic 1=grav 2=drift 3=kick (hamiltonian_cosmic.f90:74-89). */
int orc_split(orc_sim *s, double h, int hamind)
{
    int conf = (s->time_var == ORC_CONFORMAL);
    switch (hamind) {
    case 1: if (conf) fields_kin(s, h); else grav_kinetic(s, h); return 0;
    case 2: if (conf) grav_kinetic(s, h); else fields_kin(s, h); return 0;
    case 3: fields_grad_pot(s, h); return 0;
    default: return -1;
    }
}

/* symp_o2step: hamiltonian_conformal.f90:50-64 with n_Hamiltonian_terms = 3 */
int orc_symp_o2step(orc_sim *s, double dt, double w1, double w2)
{
    orc_split(s, w1 * dt / 2.0, 2);
    orc_split(s, w1 * dt, 3);
    orc_split(s, w1 * dt / 2.0, 2);
    orc_split(s, (w1 + w2) * dt / 2.0, 1);
    return 0;
}

/* mpi_dev/integrator.f90:11-110; openmp_dev/evolve_spectral.f90:129-227 */
int orc_symp(orc_sim *s, int order, double dt, int nsteps, int symp6_tree)
{
    if (nsteps <= 0) return 0;
    if (order == 2) {
        orc_split(s, dt / 2.0, 1);
        for (int j = 1; j <= nsteps - 1; ++j) orc_symp_o2step(s, dt, 1.0, 1.0);
        orc_symp_o2step(s, dt, 1.0, 0.0);
        return 0;
    }
    if (order == 4) {
        const double w1 = 1.0 / (2.0 - pow(2.0, 1.0 / 3.0));
        const double w0 = 1.0 - 2.0 * w1;
        orc_split(s, w1 * dt / 2.0, 1);
        for (int i = 1; i <= nsteps; ++i) {
            orc_symp_o2step(s, dt, w1, w0);
            orc_symp_o2step(s, dt, w0, w1);
            orc_symp_o2step(s, dt, w1, (i == nsteps) ? 0.0 : w1);
        }
        return 0;
    }
    if (order == 6) {
        const double w1 = symp6_tree ? -1.17767998417887 : -1.177679984177887;
        const double w2 = 0.235573213359357, w3 = 0.784513610477560;
        const double w0 = 1.0 - 2.0 * (w1 + w2 + w3);
        orc_split(s, w3 * dt / 2.0, 1);
        for (int j = 1; j <= nsteps; ++j) {
            orc_symp_o2step(s, dt, w3, w2);
            orc_symp_o2step(s, dt, w2, w1);
            orc_symp_o2step(s, dt, w1, w0);
            orc_symp_o2step(s, dt, w0, w1);
            orc_symp_o2step(s, dt, w1, w2);
            orc_symp_o2step(s, dt, w2, w3);
            orc_symp_o2step(s, dt, w3, (j == nsteps) ? 0.0 : w3);
        }
        return 0;
    }
    if (order == 8) {
        const double w1 = 0.74167036435061295344822780, w2 = -0.40910082580003159399730010,
                     w3 = 0.19075471029623837995387626, w4 = -0.57386247111608226665638733,
                     w5 = 0.29906418130365592384446354, w6 = 0.33462491824529818378495798,
                     w7 = 0.31529309239676659663205666;
        const double w0 = 1.0 - 2.0 * (w1 + w2 + w3 + w4 + w5 + w6 + w7);
        const double seq[16] = {w7, w6, w5, w4, w3, w2, w1, w0, w1, w2, w3, w4, w5, w6, w7, 0};
        orc_split(s, w7 * dt / 2.0, 1);
        for (int j = 1; j <= nsteps; ++j) {
            for (int q = 0; q < 14; ++q) orc_symp_o2step(s, dt, seq[q], seq[q + 1]);
            orc_symp_o2step(s, dt, w7, (j == nsteps) ? 0.0 : w7);
        }
        return 0;
    }
    return -1;
}

/* ------------------------------------------------------------------ */
/* Energies and calc_metric                                            */
/* ------------------------------------------------------------------ */

/* sums: sum_l sum phi_l lap_l (raw Laplacian, no a factor), sum pi^2, sum V */
static void raw_sums(orc_sim *s, double *sphilap, double *spi2, double *sV, double *mean)
{
    size_t nv = nvol_of(s);
    const size_t plane = (size_t)s->nx * s->ny;
    const int nz = s->nz;
    double *part = (double *)calloc((size_t)nz, sizeof(double));
    double tot = 0.0;
    for (int l = 0; l < s->nfld; ++l) {
        field_laplacian(s, l);
        const double *f = s->fld + (size_t)l * nv;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k) {
            double acc = 0.0;
            for (size_t q = plane * k; q < plane * (k + 1); ++q) acc -= f[q] * s->scratch[q];
            part[k] = acc;
        }
        tot += plane_sum(part, nz);
    }
    *sphilap = tot;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double acc = 0.0;
        for (int l = 0; l < s->nfld; ++l)
            for (size_t q = plane * k; q < plane * (k + 1); ++q) {
                double p = s->fldp[(size_t)l * nv + q];
                acc += p * p;
            }
        part[k] = acc;
    }
    *spi2 = plane_sum(part, nz);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double acc = 0.0;
        for (size_t q = plane * k; q < plane * (k + 1); ++q) {
            double fs[2];
            for (int m = 0; m < s->nfld; ++m) fs[m] = s->fld[(size_t)m * nv + q];
            acc += orc_potential(s, fs);
        }
        part[k] = acc;
    }
    *sV = plane_sum(part, nz);
    for (int l = 0; l < 2; ++l) {
        mean[l] = 0.0;
        if (l >= s->nfld) continue;
        const double *f = s->fld + (size_t)l * nv;
#pragma omp parallel for schedule(static)
        for (int k = 0; k < nz; ++k) {
            double acc = 0.0;
            for (size_t q = plane * k; q < plane * (k + 1); ++q) acc += f[q];
            part[k] = acc;
        }
        mean[l] = plane_sum(part, nz) / (double)nv;
    }
    free(part);
}

/* calc_metric: conformal p_a = -sqrt(12 a^4 rho) (hamiltonian_conformal.f90:263-304);
 * cosmic p_y = -a sqrt(16 rho/3) (hamiltonian_cosmic.f90:351-403). */
int orc_calc_metric(orc_sim *s)
{
    double sphilap, spi2, sV, mean[2];
    raw_sums(s, &sphilap, &spi2, &sV, mean);
    double nv = (double)nvol_of(s);
    double a = scale_factor(s);
    double knorm = (s->time_var == ORC_CONFORMAL) ? 1.0 / pow(s->yscl, 6) : 1.0 / pow(s->yscl, 4);
    double KE = 0.5 * spi2 * knorm / nv;
    double PE = sV / nv;
    double GE = 0.5 * sphilap / (a * a) / nv;
    double rho = KE + PE + GE;
    if (s->time_var == ORC_CONFORMAL)
        s->ysclp = -sqrt(pow(s->yscl, 4) * 12.0 * rho);
    else
        s->ysclp = -a * sqrt(rho * 16.0 / 3.0);
    return 0;
}

/* dump_rho (evolve_spectral.f90:305-415) / scalar_rho (hamiltonian_cosmic.f90:268-349) */
int orc_energies(orc_sim *s, double *out)
{
    double sphilap, spi2, sV, mean[2];
    raw_sums(s, &sphilap, &spi2, &sV, mean);
    double nv = (double)nvol_of(s);
    double a = scale_factor(s);
    double knorm, hub;
    if (s->time_var == ORC_CONFORMAL) {
        knorm = 1.0 / pow(s->yscl, 6);
        hub = -s->ysclp / 6.0 / (s->yscl * s->yscl);
    } else {
        knorm = 1.0 / pow(s->yscl, 4);
        hub = -s->ysclp / s->yscl / 4.0;
    }
    double KE = 0.5 * spi2 * knorm / nv;
    double PE = sV / nv;
    double GE = 0.5 * sphilap / nv / (a * a);
    double rho = KE + PE + GE;
    double grav = -3.0 * hub * hub;
    out[ORC_E_A] = a;
    out[ORC_E_HUB] = hub;
    out[ORC_E_RHO] = rho;
    out[ORC_E_KE] = KE;
    out[ORC_E_PE] = PE;
    out[ORC_E_GE] = GE;
    out[ORC_E_GRAV] = grav;
    out[ORC_E_RESID] = (rho + grav) / rho;
    out[ORC_E_PMOM] = (s->time_var == ORC_CONFORMAL) ? -s->ysclp * s->ysclp / 12.0 / pow(a, 4)
                                                     : -s->ysclp * s->ysclp * 3.0 / 16.0 / (a * a);
    out[ORC_E_MEAN1] = mean[0];
    out[ORC_E_MEAN2] = mean[1];
    out[ORC_E_GEGRAD] = 0.0;
    if (s->lap_kind == ORC_LAP_STENCIL) {
        double dx = dx_of(s), g = 0.0;
        for (int l = 0; l < s->nfld; ++l)
            g += orc_stencil_grad2_sum(s->nx, s->ny, s->nz, s->fld + (size_t)l * nvol_of(s), s->c);
        out[ORC_E_GEGRAD] = 0.25 / ((a * dx) * (a * dx)) / s->c[4] * g / nv; /* homlat.f90:101-108 */
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Discrete-tree diagnostics (make_output of Discrete/Entropy)          */
/* ------------------------------------------------------------------ */

/* STENCIL(c, O) at one site for the operators of Discrete/Entropy/macros.h:20-24,
 * each summed over the 3^3 - 1 neighbours in RANK1..3 order (a rank's sum in
 * the macro's neighbour order, then c0*RANK0 + c1*RANK1 + c2*RANK2 + c3*RANK3):
 *   lap[l]  = LAPLACE(x,y,z) = f_l(nbr)                      (rank 0 included)
 *   grad2   = GRAD2 = sum_l (f_l(nbr) - f_l(0))^2            (rank 0 term 0)
 *   piphi[l]= PI_PHI = (p_l(nbr) - p_l(0)) (f_l(nbr) - f_l(0)) (rank 0 term 0) */
static const int RK1[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
static const int RK2[12][3] = {{-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}, {-1, 0, -1}, {1, 0, -1},
                               {-1, 0, 1},  {1, 0, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};
static const int RK3[8][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {-1, 1, 1},  {1, 1, 1}};

static void site_ops(const orc_sim *s, int i, int j, int k, double *lap, double *grad2, double *piphi)
{
    const size_t nv = nvol_of(s);
    const int nx = s->nx, ny = s->ny, nz = s->nz, nf = s->nfld;
    const size_t o0 = pidx(i, j, k, nx, ny, nz);
    double rl[4][2] = {{0}}, rg[4] = {0}, rp[4][2] = {{0}};
    for (int l = 0; l < nf; ++l) rl[0][l] = s->fld[(size_t)l * nv + o0];
    for (int r = 1; r <= 3; ++r) {
        const int cnt = r == 1 ? 6 : (r == 2 ? 12 : 8);
        for (int q = 0; q < cnt; ++q) {
            const int *d = r == 1 ? RK1[q] : (r == 2 ? RK2[q] : RK3[q]);
            const size_t o = pidx(i + d[0], j + d[1], k + d[2], nx, ny, nz);
            double g = 0.0;
            for (int l = 0; l < nf; ++l) {
                const double df = s->fld[(size_t)l * nv + o] - s->fld[(size_t)l * nv + o0];
                rl[r][l] += s->fld[(size_t)l * nv + o];
                g += df * df;
                rp[r][l] += (s->fldp[(size_t)l * nv + o] - s->fldp[(size_t)l * nv + o0]) * df;
            }
            rg[r] += g;
        }
    }
    const double *c = s->c;
    for (int l = 0; l < nf; ++l) {
        lap[l] = c[0] * rl[0][l] + c[1] * rl[1][l] + c[2] * rl[2][l] + c[3] * rl[3][l];
        piphi[l] = c[0] * 0.0 + c[1] * rp[1][l] + c[2] * rp[2][l] + c[3] * rp[3][l];
    }
    *grad2 = c[0] * 0.0 + c[1] * rg[1] + c[2] * rg[2] + c[3] * rg[3];
}

/* modelv = 2V (Discrete/Entropy/model.f90:35-40) */
static double modelv(const orc_sim *s, const double *f) { return 2.0 * orc_potential(s, f); }

int orc_discrete_output(orc_sim *s, double time, double *log_out, double *lnrho, double *dlnrho, double *psd)
{
    if (s->lap_kind != ORC_LAP_STENCIL || s->nx != s->ny || s->nx != s->nz) return -1;
    const size_t nv = nvol_of(s);
    const int n = s->nx, nf = s->nfld;
    const double a = scale_factor(s), dx = dx_of(s), cc = s->c[4];
    const double hub = s->time_var == ORC_CONFORMAL ? -s->ysclp / s->yscl / 6.0 / s->yscl
                                                    : -s->ysclp / s->yscl / 4.0; /* scafld.f90:67-75 */
    /* homlat.f90:32-44 get_energy_coeffs */
    const double ekin = 0.5 / pow(a, 6), egrad = 0.25 / ((a * dx) * (a * dx)) / cc, epot = 0.5;
    const double lap_coeff = 1.0 / ((a * dx) * (a * dx)) / cc, piphi_coeff = 0.5 / ((a * dx) * (a * dx)) / cc;
    double *rho = (double *)malloc(sizeof(double) * nv), *prs = (double *)malloc(sizeof(double) * nv);
    double lnr = 0.0, eos = 0.0, KE = 0.0, GE = 0.0, PE = 0.0, phiave[2] = {0, 0}, phidot[2] = {0, 0};
    /* get_variables (iofile.f90:478-582): one loop over the sites */
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const size_t q = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                double f[2] = {0, 0}, T = 0.0, lap[2], g, pp[2];
                for (int l = 0; l < nf; ++l) {
                    f[l] = s->fld[(size_t)l * nv + q];
                    T += s->fldp[(size_t)l * nv + q] * s->fldp[(size_t)l * nv + q];
                    phiave[l] += f[l];
                    phidot[l] += s->fldp[(size_t)l * nv + q];
                }
                site_ops(s, i, j, k, lap, &g, pp);
                const double V = modelv(s, f);
                rho[q] = ekin * T + egrad * g + epot * V;
                prs[q] = ekin * T - (1.0 / 3.0) * egrad * g - epot * V;
                GE += g;
                PE += V;
                KE += T;
                lnr += log(rho[q]);
                eos += prs[q] / rho[q];
            }
    const double n3 = (double)n * (double)n * (double)n;
    GE = egrad * GE / n3;
    PE = epot * PE / n3;
    KE = ekin * KE / n3;
    eos /= n3;
    lnr /= n3;
    for (int l = 0; l < 2; ++l) { phiave[l] /= n3; phidot[l] /= n3; }
    const double rhoave = GE + PE + KE;                  /* :578 */
    const double prsave = KE - PE - (1.0 / 3.0) * GE;     /* :579 */
    const double lnrhoave = lnr - log(rhoave);            /* :580 */
    /* output_log_file (:643) */
    const double grav = -3.0 * hub * hub;
    int m = 0;
    log_out[m++] = time;
    log_out[m++] = a;
    log_out[m++] = hub;
    log_out[m++] = KE + GE + PE;    /* scalar_en, homlat.f90:194-196 */
    log_out[m++] = prsave;
    log_out[m++] = lnrhoave;
    log_out[m++] = eos;
    log_out[m++] = KE;
    log_out[m++] = PE;
    log_out[m++] = GE;
    log_out[m++] = grav;
    log_out[m++] = GE + KE + PE + grav; /* hamiltonian, homlat.f90:198-200 */
    for (int l = 0; l < nf; ++l) log_out[m++] = phiave[l];
    for (int l = 0; l < nf; ++l) log_out[m++] = phidot[l] / (a * a * a);
    /* output_reg (:416-423): the two fields */
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const size_t q = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                double lap[2], g, pp[2];
                site_ops(s, i, j, k, lap, &g, pp);
                double num = 0.0;
                for (int l = 0; l < nf; ++l) num += s->fldp[(size_t)l * nv + q] * (lap_coeff * lap[l]);
                for (int l = 0; l < nf; ++l) num += piphi_coeff * pp[l];
                lnrho[q] = log(rho[q]) - log(rhoave) - lnrhoave;
                dlnrho[q] = -3.0 * hub * (prs[q] / rho[q] - prsave / rhoave) + num / rho[q] / (a * a * a);
            }
    free(rho);
    free(prs);
    /* PSD columns: two-bin weighted spectra over the stored half space without
     * the x-Nyquist plane (outputfile.f90:569-596,656-663, S/W, no n^6), and
     * crossspec_nofft(F(dlnrho), F(lnrho)) (analysis.f90:149-186) */
    const int ns = (int)(1.7320508075688772935274463415059 * (double)(n / 2) + 2.0);
    const int nxh = n / 2 + 1;
    cplx *F1 = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * n * n);
    cplx *F2 = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * n * n);
    r2c_3d(n, n, n, lnrho, F1);
    r2c_3d(n, n, n, dlnrho, F2);
    double *W = (double *)calloc((size_t)ns, sizeof(double));
    memset(psd, 0, sizeof(double) * 4 * (size_t)ns);
    for (int k = 0; k < n; ++k) {
        const int kk = wmag(k, n);
        for (int j = 0; j < n; ++j) {
            const int jj = wmag(j, n);
            for (int ii = 0; ii < n / 2; ++ii) {
                const size_t id = ii + (size_t)nxh * (j + (size_t)n * k);
                const double p = sqrt((double)(ii * ii + jj * jj + kk * kk));
                const int l = (int)floor(p);
                const double d1 = (double)l - p, d2 = (double)(l + 1) - p;
                const double c1 = (1.0 - d1 * d1) * (1.0 - d1 * d1), c2 = (1.0 - d2 * d2) * (1.0 - d2 * d2);
                const cplx x = F2[id] * conj(F1[id]);
                const double v[4] = {creal(F1[id] * conj(F1[id])), creal(F2[id] * conj(F2[id])), creal(x), cimag(x)};
                for (int ch = 0; ch < 4; ++ch) psd[(size_t)ch * ns + l] += c1 * v[ch];
                W[l] += c1;
                if (l + 1 < ns) {
                    for (int ch = 0; ch < 4; ++ch) psd[(size_t)ch * ns + l + 1] += c2 * v[ch];
                    W[l + 1] += c2;
                }
            }
        }
    }
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0)
            for (int ch = 0; ch < 4; ++ch) psd[(size_t)ch * ns + b] /= W[b];
    free(W);
    free(F1);
    free(F2);
    return 12 + 2 * nf;
}

/* ------------------------------------------------------------------ */
/* Gravitational waves (textbook; see header: parity unpinned)         */
/* ------------------------------------------------------------------ */

/* S_ij = sum_fields d_i phi d_j phi, spectral derivatives; component order
 * xx, xy, xz, yy, yz, zz. */
int orc_gw_source(int nx, int ny, int nz, int nfld, const double *fld, double dk, double *Sij)
{
    size_t nv = (size_t)nx * ny * nz;
    const int nxh = nx / 2 + 1;
    double *d = (double *)malloc(sizeof(double) * nv * 3);
    cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * ny * nz);
    memset(Sij, 0, sizeof(double) * nv * 6);
    static const int ci[6] = {0, 0, 0, 1, 1, 2}, cj[6] = {0, 1, 2, 1, 2, 2};
    for (int l = 0; l < nfld; ++l) {
        r2c_3d(nx, ny, nz, fld + (size_t)l * nv, F);
        for (int ax = 0; ax < 3; ++ax) spectral_derivative(nx, ny, nz, F, d + ax * nv, dk, ax);
        for (int c = 0; c < 6; ++c)
            for (size_t q = 0; q < nv; ++q) Sij[c * nv + q] += d[ci[c] * nv + q] * d[cj[c] * nv + q];
    }
    free(d); free(F);
    return 0;
}

int orc_gw_source_fd(int nx, int ny, int nz, int nfld, const double *fld, double dx, double *Sij)
{
    size_t nv = (size_t)nx * ny * nz;
    const double r = 0.5 / dx;
    memset(Sij, 0, sizeof(double) * nv * 6);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        const int km = (k + nz - 1) % nz, kp = (k + 1) % nz;
        for (int j = 0; j < ny; ++j) {
            const int jm = (j + ny - 1) % ny, jp = (j + 1) % ny;
            for (int i = 0; i < nx; ++i) {
                const int im = (i + nx - 1) % nx, ip = (i + 1) % nx;
                const size_t q = i + (size_t)nx * (j + (size_t)ny * k);
                for (int l = 0; l < nfld; ++l) {
                    const double *f = fld + (size_t)l * nv;
                    double g[3];
                    g[0] = (f[ip + (size_t)nx * (j + (size_t)ny * k)] - f[im + (size_t)nx * (j + (size_t)ny * k)]) * r;
                    g[1] = (f[i + (size_t)nx * (jp + (size_t)ny * k)] - f[i + (size_t)nx * (jm + (size_t)ny * k)]) * r;
                    g[2] = (f[i + (size_t)nx * (j + (size_t)ny * kp)] - f[i + (size_t)nx * (j + (size_t)ny * km)]) * r;
                    Sij[0 * nv + q] += g[0] * g[0];
                    Sij[1 * nv + q] += g[0] * g[1];
                    Sij[2 * nv + q] += g[0] * g[2];
                    Sij[3 * nv + q] += g[1] * g[1];
                    Sij[4 * nv + q] += g[1] * g[2];
                    Sij[5 * nv + q] += g[2] * g[2];
                }
            }
        }
    }
    return 0;
}

/* In-place TT projection of 6 half-complex symmetric-tensor arrays:
 * U^TT_ij = Lambda_ij,lm U_lm, Lambda = P_il P_jm - 1/2 P_ij P_lm,
 * P_ij = delta_ij - khat_i khat_j (signed wavenumbers); k = 0 mode zeroed. */
int orc_tt_project(int nx, int ny, int nz, double *Uk_)
{
    const int nxh = nx / 2 + 1;
    size_t nk = (size_t)nxh * ny * nz;
    cplx *U = (cplx *)Uk_;
    static const int cidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nxh; ++i) {
                size_t id = i + (size_t)nxh * (j + (size_t)ny * k);
                double kv[3] = {(double)i, (double)wsig(j, ny), (double)wsig(k, nz)};
                double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                if (k2 == 0.0) { for (int c = 0; c < 6; ++c) U[c * nk + id] = 0; continue; }
                double P[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.0 : 0.0) - kv[a] * kv[b] / k2;
                cplx u[3][3], r[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) u[a][b] = U[cidx[a][b] * nk + id];
                cplx tr = 0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) tr += P[a][b] * u[a][b];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        cplx acc = 0;
                        for (int l = 0; l < 3; ++l)
                            for (int m = 0; m < 3; ++m) acc += P[a][l] * P[b][m] * u[l][m];
                        r[a][b] = acc - 0.5 * P[a][b] * tr;
                    }
                U[0 * nk + id] = r[0][0]; U[1 * nk + id] = r[0][1]; U[2 * nk + id] = r[0][2];
                U[3 * nk + id] = r[1][1]; U[4 * nk + id] = r[1][2]; U[5 * nk + id] = r[2][2];
            }
    return 0;
}

/* GW energy spectrum per shell: sum over TT components of |dU_ij|^2 (off-
 * diagonal counted twice) with the analysis_spec.f90:86-101 two-bin weights,
 * full-x mirror loop, S/W.  Uk may be NULL (only dUk used). */
int orc_gw_spectrum(int nx, int ny, int nz, const double *Uk_, const double *dUk_, double dk, double *S)
{
    (void)Uk_; (void)dk;
    const int nxh = nx / 2 + 1;
    size_t nk = (size_t)nxh * ny * nz;
    const cplx *D = (const cplx *)dUk_;
    int ns = orc_num_bins(nx, ny, nz);
    double *W = (double *)calloc((size_t)ns, sizeof(double));
    memset(S, 0, sizeof(double) * ns);
    static const double mult[6] = {1, 2, 2, 1, 2, 1};
    for (int k = 0; k < nz; ++k) {
        int kk = wmag(k, nz);
        for (int j = 0; j < ny; ++j) {
            int jj = wmag(j, ny);
            for (int i0 = 0; i0 < nx; ++i0) {
                int ii = wmag(i0, nx);
                size_t id = ii + (size_t)nxh * (j + (size_t)ny * k);
                double a2 = 0.0;
                for (int c = 0; c < 6; ++c) a2 += mult[c] * creal(D[c * nk + id] * conj(D[c * nk + id]));
                double p = sqrt((double)(ii * ii + jj * jj + kk * kk));
                int l = (int)floor(p);
                double d1 = (double)l - p, d2 = (double)(l + 1) - p;
                double c1 = (1.0 - d1 * d1) * (1.0 - d1 * d1);
                double c2 = (1.0 - d2 * d2) * (1.0 - d2 * d2);
                S[l] += c1 * a2; W[l] += c1;
                if (l + 1 < ns) { S[l + 1] += c2 * a2; W[l + 1] += c2; }
            }
        }
    }
    for (int b = 0; b < ns; ++b)
        if (W[b] != 0.0) S[b] /= W[b];
    free(W);
    return ns;
}

/* ------------------------------------------------------------------ */
/* Initial conditions: Philox-4x32-10 (Salmon et al., SC'11)           */
/* ------------------------------------------------------------------ */

static void philox4x32_10(uint32_t ctr[4], const uint32_t key_in[2], uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0; k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniforms for counter `idx` of `stream`: u1 in (0,1], u2 in [0,1), 53-bit. */
static void philox_uniform2(uint64_t seed, uint64_t stream, uint64_t idx, double *u1, double *u2)
{
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox4x32_10(ctr, key, o);
    const double inv53 = 1.0 / 9007199254740992.0;
    uint64_t a = ((uint64_t)(o[0] >> 5) << 26) | (o[1] >> 6);
    uint64_t b = ((uint64_t)(o[2] >> 5) << 26) | (o[3] >> 6);
    *u1 = 1.0 - (double)a * inv53;
    *u2 = (double)b * inv53;
}

void orc_philox_normal_pairs(unsigned long long seed, unsigned long long stream, unsigned long long count,
                             double *u1, double *u2)
{
    for (unsigned long long q = 0; q < count; ++q) philox_uniform2(seed, stream, q, &u1[q], &u2[q]);
}

/* FFTW RODFT10 (DST-II): Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1)/n), via a
 * length-4n complex FFT of the odd extension: Y_k = -Im Z_{k+1}. */
static void dst2(int n, double *x)
{
    int m = 4 * n;
    fft_plan p;
    plan_init(&p, m);
    cplx *z = (cplx *)calloc((size_t)m, sizeof(cplx));
    cplx *t = (cplx *)malloc(sizeof(cplx) * (size_t)m);
    for (int j = 0; j < n; ++j) { z[2 * j + 1] = x[j]; z[m - 2 * j - 1] = -x[j]; }
    fft_exec(&p, z, t, -1);
    for (int k = 0; k < n; ++k) x[k] = -cimag(z[k + 1]);
    free(z); free(t); plan_free(&p);
}

int orc_dst2(int n, double *x)
{
    if (n <= 0) return -1;
    dst2(n, x);
    return 0;
}

/* evolve_spectral.f90:436-491: the real-space convolution kernel of sample()
 * before the noise multiply -- the oversampled radial profile
 * ker(kk) = kk (kk^2 + m2eff)^gamma exp(-(kk/kcut)^2) at kk = (k - 1/2) dkos,
 * DST-II, times norm / k, laid out on the lattice by linear interpolation at
 * radius |(i,j,k) - nn| os (centre: (4 ker(1) - ker(2)) / 3). */
int orc_sample_kernel(int nx, int ny, int nz, double len, double gamma, double m2eff, double mpl,
                      double kcut_fac, double *out)
{
    const int os = 16;
    const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    const int nos = nmax * os * os;
    const int nnx = nx / 2 + 1, nny = ny / 2 + 1, nnz = nz / 2 + 1;
    int nn = nnx < nny ? nnx : nny; nn = nn < nnz ? nn : nnz;
    const double twopi = 6.2831853071795864769252867665590;
    const double dx = len / (double)nx, dk = twopi / len;
    const double dxos = dx / os, dkos = dk / (2.0 * os);
    const double kcut = kcut_fac * (double)nn * dk / 2.0;
    const double nvol = (double)nx * (double)ny * (double)nz;
    const double norm = 0.5 / (nvol * sqrt(twopi * dk * dk * dk) * mpl) * (dkos / dxos);
    double *ker = (double *)malloc(sizeof(double) * (size_t)nos);
    for (int k = 1; k <= nos; ++k) {
        double kk = ((double)k - 0.5) * dkos;
        ker[k - 1] = kk * pow(kk * kk + m2eff, gamma) * exp(-(kk / kcut) * (kk / kcut));
    }
    dst2(nos, ker);
    for (int k = 1; k <= nos; ++k) ker[k - 1] = norm * ker[k - 1] / (double)k;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double di = i - nn, dj = j - nn, dkk = k - nn;
                double kk = sqrt(di * di + dj * dj + dkk * dkk) * os;
                int l = (int)floor(kk);
                double v;
                if (l > 0) v = ker[l - 1] + (kk - l) * (ker[l] - ker[l - 1]);
                else v = (4.0 * ker[0] - ker[1]) / 3.0;
                out[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1))] = v;
            }
    free(ker);
    return 0;
}

/* evolve_spectral.f90:421-506 with the RNG replaced: the per-mode Rayleigh
 * amplitude sqrt(-2 ln u1) and phase 2 pi u2 come from Philox counter
 * idx = i + nxh*(j + ny*k) of `stream` under `seed`. */
int orc_sample(int nx, int ny, int nz, double len, double gamma, double m2eff, double mpl,
               double kcut_fac, unsigned long long seed, unsigned long long stream, double *out)
{
    const double twopi = 6.2831853071795864769252867665590;
    orc_sample_kernel(nx, ny, nz, len, gamma, m2eff, mpl, kcut_fac, out);
    const int nxh = nx / 2 + 1;
    cplx *F = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * ny * nz);
    r2c_3d(nx, ny, nz, out, F);
    for (size_t id = 0; id < (size_t)nxh * ny * nz; ++id) {
        double u1, u2;
        philox_uniform2(seed, stream, id, &u1, &u2);
        double amp = sqrt(-2.0 * log(u1));
        double ph = twopi * u2;
        F[id] *= amp * (cos(ph) + I * sin(ph));
    }
    c2r_3d(nx, ny, nz, F, out);
    free(F);
    return 0;
}
