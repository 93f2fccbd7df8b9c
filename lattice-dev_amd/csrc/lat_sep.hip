// lat_sep.hip -- the spectral Laplacian as a sum of one-axis terms.
//
// laplacian_3d (Pseudospec/openmp_dev/fftw_mod.f90:177-207) multiplies the
// 3-D transform by -dk^2 (i^2 + jj^2 + kk^2), a sum of one-axis factors, so
// the same Laplacian is D_xx f + D_yy f + D_zz f with
// D_aa f = F_a^-1[-dk^2 k_a^2 F_a f] / n_a the 1-D spectral second derivative
// along axis a (Nyquist multiplied, as in the 3-D form).  The kick then runs
// as three sweeps instead of five:
//   x: pending drifts, then D_xx of every row (real FFT, factor, inverse),
//      written to lap;
//   y: the real field read as complex pairs (x = 2m, 2m+1 as re, im): -ky^2
//      is real and even, so one complex transform, factor and inverse gives
//      D_yy of both real lines; lap += D_yy f;
//   z: likewise D_zz, and the kick of every site with lap + D_zz f, the
//      GE / PE / KE / mean reductions fused.
// The y and z sweeps reuse the out-of-place LDS passes and the fused
// last-forward / first-inverse pass of the 3-D column kernel.
#include "sweeps.h"

namespace lat {

// ---------------------------------------------------------------- x sweep
template <typename T, int N, int L, bool DRIFT, int MR>
__global__ __launch_bounds__(TPB) void k_sep_x(T *__restrict__ f, const T *__restrict__ pi, T *__restrict__ lap,
                                               const cpx<T> *__restrict__ tw, int rows,
                                               const DevScalars *__restrict__ dsc, int field, MemberStride ms,
                                               double kf)
{
    constexpr int M = N / 2;
    constexpr int LS = M + 2; // even pitch: rows start 16-byte aligned
    constexpr bool TWL = true;
    __shared__ __attribute__((aligned(16))) cpx<T> buf[L * LS];
    __shared__ cpx<T> twl[N];
    LAT_TWIDDLES_TO_LDS(N, TPB, twl)
    f += blockIdx.y * ms.f;
    if (pi) pi += blockIdx.y * ms.f;
    lap += blockIdx.y * ms.f;
    dsc += blockIdx.y;
    int nd = 0;
    T dc0 = 0, dc1 = 0, nu = 0;
    if constexpr (DRIFT) {
        nd = dsc->nd;
        dc0 = T(dsc->dc[0]);
        dc1 = T(dsc->dc[1]);
        if (dsc->mixed && field >= 0) nu = T(dsc->nu[field]);
    }
    const T kfT = T(kf);
    int t0, t1, ts;
    tile_range((rows + L - 1) / L, hw_blk(), t0, t1, ts);
    constexpr int VW = 16 / sizeof(T);
    using VT = Vec16<T>;
    for (int t = t0; t < t1; t += ts) {
        const int r0 = t * L;
        // 1. rows in (pending drifts applied and written back)
        for (int idx = threadIdx.x; idx < L * N / VW; idx += TPB) {
            const int line = idx / (N / VW), i0 = (idx - line * (N / VW)) * VW;
            const int row = r0 + line;
            VT v{};
            if (row < rows) {
                VT *src = reinterpret_cast<VT *>(f + (size_t)row * N + i0);
                v = *src;
                if constexpr (DRIFT) {
                    if (nd > 0) {
                        const VT p = *reinterpret_cast<const VT *>(pi + (size_t)row * N + i0);
#pragma unroll
                        for (int e = 0; e < VW; ++e) {
                            T x = v.e[e] + dc0 * p.e[e];
                            if (nd > 1) x = x + dc1 * p.e[e];
                            if constexpr (sizeof(T) == 4) x = x - nu;
                            v.e[e] = x;
                        }
                        *src = v;
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < VW; e += 2) buf[line * LS + (i0 + e) / 2] = cpx<T>{v.e[e], v.e[e + 1]};
        }
        lds_sync();
        // 2. forward: z[n] = f[2n] + i f[2n+1], M-point FFT
        lds_fft_range<T, M, -1, L, TPB, false, 2, 0, num_passes(M, MR), MR>(buf, LS, 1, twp);
        // 3. per pair (k, M-k): real-FFT split X[k], X[M-k]; factor
        //    -dk^2 k^2 / N; Hermitian pre-split of the inverse (as the x C2R
        //    of the 3-D pipeline: Im X[0] = Im X[M] = 0)
        for (int idx = threadIdx.x; idx < L * (M / 2 + 1); idx += TPB) {
            const int line = idx / (M / 2 + 1), k = idx - line * (M / 2 + 1);
            cpx<T> *z = buf + line * LS;
            const int km = (M - k) % M;
            const cpx<T> a = z[k], b = z[km];
            cpx<T> xk, xm; // X[k], X[M-k] (k = 0: X[0], X[M])
            {
                const cpx<T> w = twp[k];
                const cpx<T> s = {T(0.5) * (a.x + b.x), T(0.5) * (a.y - b.y)};
                const cpx<T> d = {T(0.5) * (a.x - b.x), T(0.5) * (a.y + b.y)};
                const cpx<T> wd = cmul(w, d);
                xk = {s.x + wd.y, s.y - wd.x};
            }
            {
                const cpx<T> w = twp[M - k];
                const cpx<T> s = {T(0.5) * (b.x + a.x), T(0.5) * (b.y - a.y)};
                const cpx<T> d = {T(0.5) * (b.x - a.x), T(0.5) * (b.y + a.y)};
                const cpx<T> wd = cmul(w, d);
                xm = {s.x + wd.y, s.y - wd.x};
            }
            const T fk = -T(k * k) * kfT, fm = -T((M - k) * (M - k)) * kfT;
            xk = {xk.x * fk, xk.y * fk};
            xm = {xm.x * fm, xm.y * fm};
            if (k == 0) {
                xk.y = 0;
                xm.y = 0;
            }
            {
                cpx<T> w = twp[k];
                w.y = -w.y;
                const cpx<T> s = {xk.x + xm.x, xk.y - xm.y};
                const cpx<T> d = {xk.x - xm.x, xk.y + xm.y};
                const cpx<T> wd = cmul(w, d);
                z[k] = {s.x - wd.y, s.y + wd.x};
            }
            if (k != 0 && k != M - k) {
                cpx<T> w = twp[M - k];
                w.y = -w.y;
                const cpx<T> s = {xm.x + xk.x, xm.y - xk.y};
                const cpx<T> d = {xm.x - xk.x, xm.y + xk.y};
                const cpx<T> wd = cmul(w, d);
                z[M - k] = {s.x - wd.y, s.y + wd.x};
            }
        }
        lds_sync();
        // 4. inverse M-point FFT: rows of D_xx f (the factor carries 1/N)
        lds_fft_range<T, M, +1, L, TPB, false, 2, 0, num_passes(M, MR), MR>(buf, LS, 1, twp);
        for (int idx = threadIdx.x; idx < L * N / VW; idx += TPB) {
            const int line = idx / (N / VW), i0 = (idx - line * (N / VW)) * VW;
            const int row = r0 + line;
            if (row < rows)
                *reinterpret_cast<VT *>(lap + (size_t)row * N + i0) =
                    *reinterpret_cast<const VT *>(reinterpret_cast<const T *>(buf + line * LS) + i0);
        }
        lds_sync();
    }
}

// ---------------------------------------------------------------- y / z sweeps
// Block sum of NRED FP64 channels for NT threads, one partial per block.
template <int NT>
__device__ __forceinline__ void block_reduce_nt(double (&v)[NRED], double *out)
{
    __shared__ double red[NRED][NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NRED; ++c) {
        double x = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[c][wid] = x;
    }
    lds_sync();
    if (threadIdx.x < NRED) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) s += red[threadIdx.x][w];
        out[threadIdx.x] = s;
    }
}

constexpr int SEP_ACC = 0;  // lap += D_aa f
constexpr int SEP_KICK = 1; // kick with lap + D_aa f, reductions

// Columns of the real fields read as complex pairs: ncol = nx/2 columns,
// tiles of TC adjacent columns along axis a (stride s_axis complex), the
// other transverse axis at stride s_other, fields at s_arr.
template <typename T, int N, int TC, int EPI, int MR, int NT>
__global__ __launch_bounds__(NT) void k_sep_col(T *__restrict__ fld, T *__restrict__ lap, T *__restrict__ fldp,
                                               const cpx<T> *__restrict__ tw, int ncol, int n_other,
                                               size_t s_axis, size_t s_other, size_t s_arr, int n_arr, double kf,
                                               const DevScalars *__restrict__ dsc, ModelDev md,
                                               double *__restrict__ partials, MemberStride ms)
{
    constexpr bool TWL = true;
    __shared__ cpx<T> twl[N];
    __shared__ cpx<T> buf[TC * N], buf2[TC * N];
    LAT_TWIDDLES_TO_LDS(N, NT, twl)
    cpx<T> *F = reinterpret_cast<cpx<T> *>(fld + blockIdx.y * ms.f);
    cpx<T> *Lc = reinterpret_cast<cpx<T> *>(lap + blockIdx.y * ms.f);
    cpx<T> *Pc = reinterpret_cast<cpx<T> *>(fldp + blockIdx.y * ms.f);
    dsc += blockIdx.y;
    const Blk bk = hw_blk();
    constexpr int NP = num_passes(N, MR);
    static_assert(NP >= 2, "fused passes need two passes");
    constexpr int R0 = radix_at(N, 0, MR);
    constexpr int NR0 = N / R0;
    constexpr int TOT0 = TC * NR0;
    constexpr int B0 = cdiv(TOT0, NT);
    const int nct = ncol / TC; // ncol is a multiple of TC (host check)
    const T kfT = T(kf);
    // kick coefficients (mixed form: fluctuations about FP64 means)
    T hA = 0, hB = 0;
    float pb32[2] = {0, 0};
    double pb64[2] = {0, 0};
    T mu[2] = {0, 0};
    if constexpr (EPI == SEP_KICK) {
        md.p0 = dsc->mp[0];
        md.p1 = dsc->mp[1];
        md.p2 = dsc->mp[2];
        hA = T(dsc->hA);
        hB = T(dsc->hB);
        if constexpr (sizeof(T) == 4) {
            for (int fi = 0; fi < 2; ++fi) {
                pb64[fi] = dsc->mean_phi[fi];
                pb32[fi] = float(pb64[fi]);
                mu[fi] = T(dsc->mu[fi]);
            }
        }
    }
    double acc[NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    int t0, t1, ts;
    tile_range(nct * n_other * n_arr, bk, t0, t1, ts);
    cpx<T> v[B0][R0];
    auto tile_off = [&](int t, int &ct, int &arr) -> size_t {
        ct = t % nct;
        const int rest = t / nct;
        const int o = rest % n_other;
        arr = rest / n_other;
        return (size_t)arr * s_arr + (size_t)o * s_other + (size_t)ct * TC;
    };
    auto load = [&](int t) {
        int ct, arr;
        const cpx<T> *base = F + tile_off(t, ct, arr);
#pragma unroll
        for (int b = 0; b < B0; ++b) {
            const int g = threadIdx.x + b * NT;
            const int line = g % TC, j = g / TC;
#pragma unroll
            for (int q = 0; q < R0; ++q) {
                cpx<T> a = {0, 0};
                if (TOT0 % NT == 0 || g < TOT0) a = base[(size_t)(j + q * NR0) * s_axis + line];
                v[b][q] = a;
            }
        }
    };
    auto mul = [&](int, int n) -> T {
        const int kk = (n <= N / 2) ? n : N - n;
        return -T(kk * kk) * kfT;
    };
    // the next tile's first-pass loads stay in flight behind the LDS passes
    // (prefetch) in the accumulate sweep; the kick's batched epilogue needs
    // those registers
    constexpr bool PF = EPI == SEP_ACC;
    if (PF && t0 < t1) load(t0);
    for (int t = t0; t < t1; t += ts) {
        int ct, arr;
        const size_t off = tile_off(t, ct, arr);
        if (!PF) load(t);
#pragma unroll
        for (int b = 0; b < B0; ++b) {
            const int g = threadIdx.x + b * NT;
            if (TOT0 % NT == 0 || g < TOT0) {
                const int line = g % TC, j = g / TC;
                dft_small<R0, -1>(v[b]);
#pragma unroll
                for (int q = 0; q < R0; ++q) buf[(j * R0 + q) * TC + line] = v[b][q];
            }
        }
        if (PF && t + ts < t1) load(t + ts);
        lds_sync();
        lds_fft_range_pp<T, N, -1, TC, NT, 1, NP - 1, MR, false, 0>(buf, buf2, twp);
        constexpr int FB = (NP - 2) & 1;
        fused_last_first_pp<T, N, TC, NT, MR>(FB ? buf2 : buf, FB ? buf : buf2, twp, mul);
        lds_fft_range_pp<T, N, +1, TC, NT, 1, NP - 1, MR, true, FB ^ 1>(buf, buf2, twp);
        static_assert(((FB ^ 1) ^ ((NP - 2) & 1)) == 1, "inverse result in buf2");
        // last inverse pass into LDS too; the epilogue then issues a whole
        // batch of its global loads before using any of them
        stockham_pass_pp<T, N, sched_radix(N, NP - 1, MR, true), sched_ns(N, NP - 1, MR, true), +1, TC, NT>(
            buf2, buf, twp);
        constexpr int PER = TC * N / NT; // elements per thread
        static_assert((TC * N) % NT == 0, "whole elements per thread");
        auto elem = [&](int i) -> size_t {
            const int idx = threadIdx.x + i * NT;
            return (size_t)(idx / TC) * s_axis + (idx % TC);
        };
        if constexpr (EPI == SEP_ACC) {
            cpx<T> *__restrict__ lb = Lc + off;
            cpx<T> lv[PER];
#pragma unroll
            for (int i = 0; i < PER; ++i) lv[i] = lb[elem(i)];
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const cpx<T> a = buf[threadIdx.x + i * NT];
                lb[elem(i)] = cpx<T>{lv[i].x + a.x, lv[i].y + a.y};
            }
        } else {
            const int other = (n_arr == 2) ? 1 - arr : arr;
            const size_t rel = off - (size_t)arr * s_arr; // position within one field
            const cpx<T> *__restrict__ fo = F + (size_t)other * s_arr + rel;
            const cpx<T> *__restrict__ fs = F + off;
            const cpx<T> *__restrict__ lb = Lc + off;
            cpx<T> *__restrict__ pb = Pc + off;
            constexpr int G = PER < 8 ? PER : 8; // elements per batch of loads
            static_assert(PER % G == 0, "whole batches");
#pragma unroll 1
            for (int g0 = 0; g0 < PER; g0 += G) {
                cpx<T> lv[G], ph[G], po[G], pp[G];
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    const size_t e = elem(g0 + i);
                    lv[i] = lb[e];
                    ph[i] = fs[e];
                    po[i] = fo[e];
                    pp[i] = pb[e];
                }
#pragma unroll
                for (int i = 0; i < G; ++i) {
                    const cpx<T> a = buf[threadIdx.x + (g0 + i) * NT];
                    const T lapv[2] = {lv[i].x + a.x, lv[i].y + a.y};
                    const T phs[2] = {ph[i].x, ph[i].y}, pho[2] = {po[i].x, po[i].y};
                    T pv[2] = {pp[i].x, pp[i].y};
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        T phi[2] = {0, 0}, dv[2] = {0, 0};
                        phi[arr] = phs[s];
                        if (n_arr == 2) phi[other] = pho[s];
                        if constexpr (sizeof(T) == 4) {
                            model_ddv(md, pb32, phi, dv);
                            if (arr == 0) acc[RC_PE] += model_dvv(md, pb64, phi);
                        } else {
                            model_dv(md, phi, dv);
                            if (arr == 0) acc[RC_PE] += model_v(md, phi);
                        }
                        T q = pv[s] + (hA * lapv[s] - hB * dv[arr]);
                        if constexpr (sizeof(T) == 4) q = q - mu[arr];
                        pv[s] = q;
                        acc[RC_GE] -= double(phs[s]) * double(lapv[s]);
                        acc[RC_KE] += double(q) * double(q);
                        acc[RC_S1 + arr] += double(q);
                        acc[RC_M1 + arr] += double(phs[s]);
                    }
                    pb[elem(g0 + i)] = cpx<T>{pv[0], pv[1]};
                }
            }
        }
        lds_sync();
    }
    if constexpr (EPI == SEP_KICK)
        block_reduce_nt<NT>(acc, partials + (size_t)blockIdx.y * MAX_GRID * NRED + (size_t)bk.x * NRED);
}

// ---------------------------------------------------------------- launch
// Shapes: the x sweep as the x R2C (rows per block, radix), the y and z
// sweeps as the fused z pass of the 3-D pipeline (8 columns per tile,
// radix <= 16, 512 threads) at nx = 1024; fewer threads where the tile is
// smaller.
template <typename T, int N>
static int sep_go(Ctx &c, int drift, int &npart)
{
    constexpr int L = r2c_lines<T, N>();
    constexpr int MR = r2c_maxradix<T>();
    constexpr int TC = clampi(8192 / N, 8, 64) < N / 2 ? clampi(8192 / N, 8, 64) : N / 2;
    constexpr int CMR = 16;
    constexpr int NT = 512;
    static_assert((N / 2) % TC == 0, "whole tiles");
    const int rows = c.ny * c.nz;
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
    T *fld = (T *)c.fld, *fldp = (T *)c.fldp, *lap = (T *)c.lap;
    const double dk2 = c.dk * c.dk;
    for (int l = 0; l < c.nfld; ++l) {
        const int pi = prof_begin(c, LAT_ST_SEP_X);
        const int g = grid_for((rows + L - 1) / L, R2C_GRID);
        T *f = fld + (size_t)l * c.nvol;
        if (drift)
            hipLaunchKernelGGL((k_sep_x<T, N, L, true, MR>), dim3(g, c.B), dim3(TPB), 0, c.stream, f,
                               fldp + (size_t)l * c.nvol, lap + (size_t)l * c.nvol, twx, rows, c.dsc, l, c.ms,
                               dk2 / c.nx);
        else
            hipLaunchKernelGGL((k_sep_x<T, N, L, false, MR>), dim3(g, c.B), dim3(TPB), 0, c.stream, f,
                               fldp + (size_t)l * c.nvol, lap + (size_t)l * c.nvol, twx, rows, c.dsc, l, c.ms,
                               dk2 / c.nx);
        prof_end(c, pi);
    }
    ModelDev md{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1], c.gw_coup};
    const int ncol = c.nx / 2;
    const size_t plane = (size_t)ncol * c.ny, s_arr = c.nvol / 2;
    {
        const int pi = prof_begin(c, LAT_ST_SEP_Y);
        const int g = grid_for((ncol / TC) * c.nz * c.nfld, 8192);
        hipLaunchKernelGGL((k_sep_col<T, N, TC, SEP_ACC, CMR, NT>), dim3(g, c.B), dim3(NT), 0, c.stream, fld, lap,
                           fldp, twy, ncol, c.nz, (size_t)ncol, plane, s_arr, c.nfld, dk2 / c.ny, c.dsc, md,
                           c.partials, c.ms);
        prof_end(c, pi);
    }
    const int pi = prof_begin(c, LAT_ST_SEP_Z);
    const int g = grid_for((ncol / TC) * c.ny * c.nfld, MAX_GRID);
    hipLaunchKernelGGL((k_sep_col<T, N, TC, SEP_KICK, CMR, NT>), dim3(g, c.B), dim3(NT), 0, c.stream, fld, lap, fldp,
                       twz, ncol, c.ny, plane, (size_t)ncol, s_arr, c.nfld, dk2 / c.nz, c.dsc, md, c.partials,
                       c.ms);
    prof_end(c, pi);
    npart = g;
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

bool sep_supported(const Ctx &c)
{
    return !c.view && c.nx == c.ny && c.ny == c.nz && (c.nx == 1024 || c.nx == 512 || c.nx == 256 || c.nx == 128);
}

template <typename T>
int launch_sep_kick(Ctx &c, int drift, int *npart)
{
    if (!sep_supported(c) || c.nfld < 1 || c.nfld > 2) {
        c.last_error = "separable Laplacian: unsupported lattice";
        return -2;
    }
    switch (c.nx) {
    case 128: return sep_go<T, 128>(c, drift, *npart);
    case 256: return sep_go<T, 256>(c, drift, *npart);
    case 512: return sep_go<T, 512>(c, drift, *npart);
    default: return sep_go<T, 1024>(c, drift, *npart);
    }
}

template int launch_sep_kick<float>(Ctx &, int, int *);

} // namespace lat
