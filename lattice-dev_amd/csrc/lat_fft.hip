// lat_fft.hip -- 3-D R2C/C2R pipeline of the spectral Laplacian on gfx950.
//
// The Laplacian of laplacian_3d (Pseudospec/openmp_dev/fftw_mod.f90:177-207)
// runs as five HBM sweeps per field:
//   1. x: R2C of each contiguous real row (N/2-point complex FFT of the row
//      packed as z[n] = f[2n] + i f[2n+1], then the real-FFT split), with the
//      pending drifts phi += c*pi fused in front (the drift sweep disappears);
//   2. y: forward complex FFT of tiles of TC adjacent kx columns;
//   3. z: forward FFT, multiply by -dk^2 (i^2+jj^2+kk^2)/(nx ny nz), inverse FFT;
//   4. y: inverse;
//   5. x: C2R of every field's rows, then the kick of all fields at each site
//      (V' needs all fields) with the GE/PE/KE reductions fused in.
// Every 1-D transform is LDS-resident (fft_core.h).  The k-space array is
// FFTW's half-complex layout with the row pitch padded to P >= nx/2+1.
#include "sweeps.h"

namespace lat {

int grid_for(int tiles, int cap)
{
    int g = tiles < cap ? tiles : cap;
    if (g >= 8) g &= ~7;
    return g < 1 ? 1 : g;
}

// ---------------------------------------------------------------- kernels
// One lattice size per launch: block coordinates from blockIdx, LDS static.
template <typename T, int N, int L, bool DRIFT, int MR = 4, bool TWL = false, int KB = 0>
__global__ __launch_bounds__(TPB) void k_x_r2c(T *__restrict__ f, const T *__restrict__ pi, cpx<T> *__restrict__ Fk,
                                               const cpx<T> *__restrict__ tw, int rows, int P,
                                               const DevScalars *__restrict__ dsc, int field, MemberStride ms)
{
    __shared__ __attribute__((aligned(16))) cpx<T> buf[r2c_buf<T, N, L>()];
    __shared__ cpx<T> twl[tw_lds<T, N, TWL>()];
    x_r2c_body<T, N, L, DRIFT, MR, TWL, KB>(f, pi, Fk, tw, rows, P, dsc, field, ms, hw_blk(), buf, twl);
}

template <typename T, int N, int TC, int MODE, int MR = 16, bool PF = true, int MINW = 1, int NT = TPB,
          bool TWL = false, bool FUSE = false, bool PP = false, int BLK = 0>
__global__ __launch_bounds__(NT, MINW) void k_col(cpx<T> *__restrict__ Fk, const cpx<T> *__restrict__ tw, int nxh,
                                             int n_other, size_t s_axis, size_t s_other, size_t s_arr,
                                             int n_arr, int ny, double kfac, size_t ms_k, size_t s_blk)
{
    __shared__ cpx<T> twl[tw_lds<T, N, TWL>()];
    __shared__ __attribute__((aligned(16))) cpx<T> buf[col_buf<T, N, TC>()];
    __shared__ __attribute__((aligned(16))) cpx<T> buf2[PP ? col_buf<T, N, TC>() : 1];
    col_body<T, N, TC, MODE, MR, PF, MINW, NT, TWL, FUSE, PP, BLK>(Fk, tw, nxh, n_other, s_axis, s_other, s_arr,
                                                                   n_arr, ny, kfac, ms_k, hw_blk(), buf, buf2, twl,
                                                                   s_blk);
}

template <typename T, int N, int L, int EPI, int MR = 4, bool PRE = true, bool TWL = false, int MINW = 1,
          int KB = 0>
__global__ __launch_bounds__(TPB, MINW) void k_x_c2r(const cpx<T> *__restrict__ Fk, const cpx<T> *__restrict__ tw,
                                               int rows, int P, size_t nk, size_t nvol, int nf, T *__restrict__ out,
                                               const T *__restrict__ fld, T *__restrict__ fldp,
                                               const DevScalars *__restrict__ dsc, ModelDev md,
                                               double *__restrict__ partials, MemberStride ms)
{
    __shared__ __attribute__((aligned(16))) cpx<T> buf[c2r_buf<T, N, L>()];
    __shared__ cpx<T> twl[tw_lds<T, N, TWL>()];
    x_c2r_body<T, N, L, EPI, MR, PRE, TWL, MINW, KB>(Fk, tw, rows, P, nk, nvol, nf, out, fld, fldp, dsc, md, partials,
                                                 ms, hw_blk(), buf, twl);
}

template <typename T, int N, int L, int MR, int KB = 0>
static void r2c_go(Ctx &c, const cpx<T> *tw, T *f, const T *pi, cpx<T> *Fk, int drift, int field)
{
    const int rows = c.ny * c.nz;
    const int g = grid_for((rows + L - 1) / L, R2C_GRID);
    constexpr bool TW = r2c_twl<T>();
    if (drift)
        hipLaunchKernelGGL((k_x_r2c<T, N, L, true, MR, TW, KB>), dim3(g, c.B), dim3(TPB), 0, c.stream, f, pi, Fk,
                           tw, rows, c.P, c.dsc, field, c.ms);
    else
        hipLaunchKernelGGL((k_x_r2c<T, N, L, false, MR, TW, KB>), dim3(g, c.B), dim3(TPB), 0, c.stream, f, pi, Fk,
                           tw, rows, c.P, c.dsc, field, c.ms);
}

template <typename T, int N, int KB>
static void r2c_shape(Ctx &c, const cpx<T> *tw, T *f, const T *pi, cpx<T> *Fk, int drift, int field)
{
    constexpr int L = r2c_lines<T, N>();
    constexpr int MR = r2c_maxradix<T>();
    if constexpr (tunable<N>() && L > 1) {
        if (c.tune.r2c == 1) return r2c_go<T, N, L / 2, MR, KB>(c, tw, f, pi, Fk, drift, field);
    }
    r2c_go<T, N, L, MR, KB>(c, tw, f, pi, Fk, drift, field);
}

template <typename T, int N>
static void r2c_size(Ctx &c, const cpx<T> *tw, T *f, const T *pi, cpx<T> *Fk, int drift, int field, int layout)
{
    if constexpr (kblk_size<N>()) {
        if (layout) return r2c_shape<T, N, KBLK>(c, tw, f, pi, Fk, drift, field);
    }
    r2c_shape<T, N, 0>(c, tw, f, pi, Fk, drift, field);
}

template <typename T>
int launch_x_r2c(Ctx &c, const cpx<T> *tw, T *f, const T *pi, cpx<T> *Fk, int drift, int field, int layout)
{
    switch (c.nx) {
#define CASE(NN)                                                  \
    case NN:                                                      \
        r2c_size<T, NN>(c, tw, f, pi, Fk, drift, field, layout);  \
        break;
        LAT_FOR_SIZES(CASE)
#undef CASE
    default:
        c.last_error = "unsupported nx";
        return -2;
    }
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

template <typename T, int N, int TC, int MODE, int BLK>
static void col_go(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int n_other, size_t s_axis, size_t s_other, int n_arrays,
                   double kfac, size_t s_blk)
{
    using C = ColCfg<T, N, MODE == 2 ? 2 : 0>;
    constexpr int NT = col_nt<T, N, MODE, TC>();
    const int g = grid_for(((c.nxh + TC - 1) / TC) * n_other * n_arrays, C::GRID);
    hipLaunchKernelGGL((k_col<T, N, TC, MODE, C::MR, C::PF, 1, NT, true, C::FUSE, C::PP, BLK>), dim3(g, c.B),
                       dim3(NT), 0, c.stream, Fk, tw, c.nxh, n_other, s_axis, s_other, c.nk, n_arrays, c.ny, kfac,
                       c.ms.k, s_blk);
}

template <typename T, int N, int MODE, int BLK>
static void col_shape(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int n_other, size_t s_axis, size_t s_other,
                      int n_arrays, double kfac, int v, size_t s_blk)
{
    constexpr int TC0 = col_tc<T, N>(MODE, 0), TC1 = col_tc<T, N>(MODE, 1);
    if constexpr (tunable<N>() && TC1 != TC0) {
        if (v == 1) return col_go<T, N, TC1, MODE, BLK>(c, tw, Fk, n_other, s_axis, s_other, n_arrays, kfac, s_blk);
    }
    col_go<T, N, TC0, MODE, BLK>(c, tw, Fk, n_other, s_axis, s_other, n_arrays, kfac, s_blk);
}

// blocked order (layout != 0, cubic lattices of kblk_size): the y pass walks
// kz tiles at the 2-level offsets and ky lines KBLK rows apart, the z pass kz
// lines at the 2-level offsets and ky tiles KBLK rows apart
template <typename T, int N, int MODE>
static void col_mode(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int axis, int n_other, size_t s_axis, size_t s_other,
                     int n_arrays, double kfac, int v, int layout)
{
    if constexpr (kblk_size<N>()) {
        if (layout) {
            const size_t P = (size_t)c.P, s_blk = (size_t)c.ny * KBLK * P;
            if (axis == 1)
                return col_shape<T, N, MODE, 1>(c, tw, Fk, n_other, KBLK * P, P, n_arrays, kfac, v, s_blk);
            return col_shape<T, N, MODE, 2>(c, tw, Fk, n_other, P, KBLK * P, n_arrays, kfac, v, s_blk);
        }
    }
    col_shape<T, N, MODE, 0>(c, tw, Fk, n_other, s_axis, s_other, n_arrays, kfac, v, 0);
}

template <typename T>
int launch_col(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int axis, int mode, int n_arrays, int layout)
{
    const int n = axis == 1 ? c.ny : c.nz;
    const int n_other = axis == 1 ? c.nz : c.ny;
    const size_t s_axis = axis == 1 ? (size_t)c.P : (size_t)c.P * c.ny;
    const size_t s_other = axis == 1 ? (size_t)c.P * c.ny : (size_t)c.P;
    const double kfac = c.dk * c.dk / (double)c.nvol;
    const int v = axis == 1 ? c.tune.coly : c.tune.colz;
    switch (n) {
#define CASE(NN)                                                                                  \
    case NN:                                                                                      \
        if (mode == 0) col_mode<T, NN, 0>(c, tw, Fk, axis, n_other, s_axis, s_other, n_arrays, kfac, v, layout); \
        else if (mode == 1) col_mode<T, NN, 1>(c, tw, Fk, axis, n_other, s_axis, s_other, n_arrays, kfac, v, layout); \
        else col_mode<T, NN, 2>(c, tw, Fk, axis, n_other, s_axis, s_other, n_arrays, kfac, v, layout);          \
        break;
        LAT_FOR_SIZES(CASE)
#undef CASE
    default:
        c.last_error = "unsupported ny/nz";
        return -2;
    }
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

template <typename T, int N, int L, int KB = 0>
static int c2r_go(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int nf, int epi, T *out, const T *fld, T *fldp,
                  const ModelDev &md)
{
    const int rows = c.ny * c.nz;
    const int g = grid_for((rows + L - 1) / L, c2r_grid_n<T, N>());
    if constexpr (KB == 0) {
        if (epi == EPI_STORE) {
            hipLaunchKernelGGL((k_x_c2r<T, N, L, EPI_STORE, 4, true, true>), dim3(g, c.B), dim3(TPB), 0, c.stream, Fk,
                               tw, rows, c.P, c.nk, c.nvol, nf, out, fld, fldp, c.dsc, md, c.partials, c.ms);
            return g;
        }
        if (epi == EPI_GWKICK) {
            hipLaunchKernelGGL((k_x_c2r<T, N, L, EPI_GWKICK, 4, true, true>), dim3(g, c.B), dim3(TPB), 0, c.stream,
                               Fk, tw, rows, c.P, c.nk, c.nvol, nf, out, fld, fldp, c.dsc, md, c.partials, c.ms);
            return g;
        }
    } else {
        // the blocked order serves the field kicks and reductions only
        if (epi != EPI_KICK && epi != EPI_REDUCE) {
            c.last_error = "blocked k-space order: kick and reduce epilogues only";
            return -1;
        }
    }
    if (epi == EPI_KICK)
        hipLaunchKernelGGL((k_x_c2r<T, N, L, EPI_KICK, 4, true, true, 1, KB>), dim3(g, c.B), dim3(TPB), 0, c.stream,
                           Fk, tw, rows, c.P, c.nk, c.nvol, nf, out, fld, fldp, c.dsc, md, c.partials, c.ms);
    else
        hipLaunchKernelGGL((k_x_c2r<T, N, L, EPI_REDUCE, 4, true, true, 1, KB>), dim3(g, c.B), dim3(TPB), 0,
                           c.stream, Fk, tw, rows, c.P, c.nk, c.nvol, nf, out, fld, fldp, c.dsc, md, c.partials,
                           c.ms);
    return g;
}

template <typename T, int N, int KB>
static int c2r_shape(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int nf, int epi, T *out, const T *fld, T *fldp,
                     const ModelDev &md)
{
    constexpr int L = c2r_lines<N>();
    if constexpr (tunable<N>() && L > 1) {
        if (c.tune.c2r == 1) return c2r_go<T, N, L / 2, KB>(c, tw, Fk, nf, epi, out, fld, fldp, md);
    }
    return c2r_go<T, N, L, KB>(c, tw, Fk, nf, epi, out, fld, fldp, md);
}

template <typename T, int N>
static int c2r_size(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int nf, int epi, T *out, const T *fld, T *fldp,
                    const ModelDev &md, int layout)
{
    if constexpr (kblk_size<N>()) {
        if (layout) return c2r_shape<T, N, KBLK>(c, tw, Fk, nf, epi, out, fld, fldp, md);
    }
    return c2r_shape<T, N, 0>(c, tw, Fk, nf, epi, out, fld, fldp, md);
}

template <typename T>
int launch_x_c2r(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int nf, int epi, T *out, const T *fld, T *fldp, int layout)
{
    ModelDev md{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1], c.gw_coup};
    if (nf < 1 || nf > 2) {
        c.last_error = "nf must be 1 or 2";
        return -1;
    }
    int g;
    switch (c.nx) {
#define CASE(NN)                                                       \
    case NN:                                                           \
        g = c2r_size<T, NN>(c, tw, Fk, nf, epi, out, fld, fldp, md, layout); \
        break;
        LAT_FOR_SIZES(CASE)
#undef CASE
    default:
        c.last_error = "unsupported nx";
        return -2;
    }
    LAT_HIP(&c, hipGetLastError());
    return g;
}

template int launch_x_r2c<double>(Ctx &, const cpx<double> *, double *, const double *, cpx<double> *, int, int,
                                  int);
template int launch_x_r2c<float>(Ctx &, const cpx<float> *, float *, const float *, cpx<float> *, int, int, int);
template int launch_col<double>(Ctx &, const cpx<double> *, cpx<double> *, int, int, int, int);
template int launch_col<float>(Ctx &, const cpx<float> *, cpx<float> *, int, int, int, int);
template int launch_x_c2r<double>(Ctx &, const cpx<double> *, cpx<double> *, int, int, double *, const double *,
                                  double *, int);
template int launch_x_c2r<float>(Ctx &, const cpx<float> *, cpx<float> *, int, int, float *, const float *, float *,
                                 int);

} // namespace lat
