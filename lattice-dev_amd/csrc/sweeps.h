// sweeps.h -- device bodies of the three FFT sweep kernels (x R2C, y/z
// column passes, x C2R + epilogue) and their launch shapes.
//
// Each body takes its block coordinates (Blk: block index and grid size along
// the tile loop, member index) and its LDS (transform buffers and twiddle
// table as separate restrict pointers) as arguments, so the same code
// runs as a kernel of its own (lat_fft.hip: one lattice size per launch,
// coordinates from blockIdx, static LDS) or inside the multi-size kernels of
// lat_multi.hip (one launch for every size class of an ensemble, coordinates
// from a class table, dynamic LDS).
#pragma once
#include "lat_internal.h"
#include "model.h"

namespace lat {

// block coordinates of one sweep: x = block index within the tile loop,
// gx = number of such blocks, y = member of the batch
struct Blk {
    int x, gx, y;
};
__device__ __forceinline__ Blk hw_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y}; }
// member index of a body: non-negative, below the 65536 grid limit
__device__ __forceinline__ void assume_blk(const Blk &bk) { __builtin_assume(bk.y >= 0 && bk.y < 65536); }

// ---------------------------------------------------------------- tiling
// Grid-stride loop over T tiles with an XCD-aware assignment: hardware hands
// consecutive workgroup ids to the 8 XCDs round-robin, so block b runs on XCD
// b%8.  Each XCD takes one contiguous eighth of the tiles, so tiles that share
// cache lines are processed by the same XCD (same L2) at the same time.
__device__ __forceinline__ void tile_range(int T, const Blk &bk, int &start, int &end, int &step)
{
    const int G = bk.gx, b = bk.x;
    // block coordinates are non-negative (as blockIdx / gridDim are): lets
    // the compiler keep unsigned index arithmetic downstream
    __builtin_assume(G > 0);
    __builtin_assume(b >= 0 && b < G);
    __builtin_assume(T >= 0);
    if ((G & 7) == 0 && T >= G) {
        const int xcd = b & 7, slot = b >> 3, G8 = G >> 3;
        const int C = (T + 7) >> 3;
        start = xcd * C + slot;
        end = min((xcd + 1) * C, T);
        step = G8;
    } else {
        start = b;
        end = T;
        step = G;
    }
}

__device__ __forceinline__ int wmag(int j0, int n) { return (j0 <= n / 2) ? j0 : n - j0; }

// ---------------------------------------------------------------- blocked k-space order
// The force pipeline of cubic lattices may keep its k-space arrays in a
// blocked row order: the KB rows (ky, kz) with kz in one block of KB planes
// and the same ky are adjacent,
//   row(ky, kz) = ((kz / KB) * N + ky) * KB + kz % KB      (FFTW order: kz * N + ky),
// which shortens the z pass's stride from N rows to one row inside a block
// (the y pass's stride grows from 1 to KB rows).  The x sweeps walk the
// real-space rows y + N z in order and map each to its k-space row.  KB = 64
// matches the first radix-16 pass of a 1024-point column (positions j + 64 q:
// block q, row j), which keeps the z pass's addressing in compile-time parts.
constexpr int KBLK = 64;
// lattice sizes with the blocked order compiled in (the force pipeline of
// cubic lattices of these sizes uses it when switched on)
template <int N>
constexpr bool kblk_size() { return N == 256 || N == 1024; }
template <int N, int KB>
__device__ __forceinline__ int kspace_row(int row)
{
    if constexpr (KB == 0) {
        return row;
    } else {
        static_assert((KB & (KB - 1)) == 0 && N % KB == 0, "power-of-two block dividing N");
        const unsigned r = (unsigned)row;
        const unsigned ky = r % N, kz = r / N;
        return (int)(((kz / KB) * N + ky) * KB + kz % KB);
    }
}

// 16-byte vector of reals (4 floats or 2 doubles) for vector loads/stores.
template <typename T>
struct alignas(16) Vec16 {
    T e[16 / sizeof(T)];
};

// ---------------------------------------------------------------- block reduction
template <int NCH>
__device__ __forceinline__ void block_reduce_store(double (&v)[NCH], double *out)
{
    __shared__ double red[NCH][TPB / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        double x = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[c][wid] = x;
    }
    lds_sync();
    if (threadIdx.x < NCH) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) s += red[threadIdx.x][w];
        out[threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------- x R2C
// With DRIFT: phi += dc0*pi (+ dc1*pi) first, written back (mixed form: on
// the fluctuations, then minus the re-centering shift nu[field]).
// Twiddle tables: TWL = true copies the N-entry table into LDS once per
// workgroup, so the FFT passes read twiddles from LDS instead of through the
// vector memory path.
// TWLDS: the body's LDS slot for the table (NTW entries when TWL).
#define LAT_TWIDDLES_TO_LDS(NTW, NTHREADS, TWLDS)                                  \
    const cpx<T> *twp = tw;                                                       \
    if constexpr (TWL) {                                                          \
        cpx<T> *twl_ = (TWLDS);                                                   \
        for (int i_ = threadIdx.x; i_ < (NTW); i_ += (NTHREADS)) twl_[i_] = tw[i_]; \
        lds_sync();                                                          \
        twp = twl_;                                                               \
    }

// LDS of each body in complex elements: the transform buffer(s) and the
// twiddle table (when TWL), passed to the bodies as separate restrict
// pointers so that buffer stores and table loads are known not to alias.
template <typename T, int N, int L>
constexpr int r2c_buf() { return L * (N / 2 + 1); }
template <typename T, int N, int TC>
constexpr int col_buf() { return TC * N; }
template <typename T, int N, int L>
constexpr int c2r_buf() { return 2 * L * (N / 2 + 2); }
template <typename T, int N, bool TWL>
constexpr int tw_lds() { return TWL ? N : 1; }
template <typename T, int N, int L, bool TWL>
constexpr int r2c_lds() { return r2c_buf<T, N, L>() + (TWL ? N : 0); }
template <typename T, int N, int TC, bool PP, bool TWL>
constexpr int col_lds() { return col_buf<T, N, TC>() * (PP ? 2 : 1) + (TWL ? N : 0); }
template <typename T, int N, int L, bool TWL>
constexpr int c2r_lds() { return c2r_buf<T, N, L>() + (TWL ? N : 0); }

template <typename T, int N, int L, bool DRIFT, int MR = 4, bool TWL = false, int KB = 0>
__device__ __forceinline__ void x_r2c_body(T *__restrict__ f, const T *__restrict__ pi, cpx<T> *__restrict__ Fk,
                                           const cpx<T> *__restrict__ tw, int rows, int P,
                                           const DevScalars *__restrict__ dsc, int field, MemberStride ms,
                                           const Blk bk, cpx<T> *__restrict__ buf, cpx<T> *__restrict__ twl)
{
    constexpr int M = N / 2;
    constexpr int LS = M + 1;
    LAT_TWIDDLES_TO_LDS(N, TPB, twl)
    assume_blk(bk);
    f += bk.y * ms.f;
    if (pi) pi += bk.y * ms.f;
    Fk += bk.y * ms.k;
    dsc += bk.y;
    int nd = 0;
    T dc0 = 0, dc1 = 0, nu = 0;
    if constexpr (DRIFT) {
        nd = dsc->nd;
        dc0 = T(dsc->dc[0]);
        dc1 = T(dsc->dc[1]);
        if (dsc->mixed && field >= 0) nu = T(dsc->nu[field]);
    }
    int t0, t1, ts;
    tile_range((rows + L - 1) / L, bk, t0, t1, ts);
    constexpr int VW = 16 / sizeof(T); // reals per 16-byte lane access
    using VT = Vec16<T>;
    for (int t = t0; t < t1; t += ts) {
        const int r0 = t * L;
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
        lds_fft_range<T, M, -1, L, TPB, false, 2, 0, num_passes(M, MR), MR>(buf, LS, 1, twp);
        // real-FFT split: X[k] = (Z[k] + Z*[M-k])/2 - i W^k (Z[k] - Z*[M-k])/2
        for (int idx = threadIdx.x; idx < L * (M / 2 + 1); idx += TPB) {
            const int line = idx / (M / 2 + 1), k = idx - line * (M / 2 + 1);
            cpx<T> *z = buf + line * LS;
            const int km = (M - k) % M;
            const cpx<T> a = z[k], b = z[km];
            {
                const cpx<T> w = twp[k];
                const cpx<T> s = {T(0.5) * (a.x + b.x), T(0.5) * (a.y - b.y)};
                const cpx<T> d = {T(0.5) * (a.x - b.x), T(0.5) * (a.y + b.y)};
                const cpx<T> wd = cmul(w, d);
                z[k] = {s.x + wd.y, s.y - wd.x};
            }
            if (k != M - k) {
                const int k2 = M - k;
                const cpx<T> w = twp[k2];
                const cpx<T> s = {T(0.5) * (b.x + a.x), T(0.5) * (b.y - a.y)};
                const cpx<T> d = {T(0.5) * (b.x - a.x), T(0.5) * (b.y + a.y)};
                const cpx<T> wd = cmul(w, d);
                z[k2] = {s.x + wd.y, s.y - wd.x};
            }
        }
        lds_sync();
        if constexpr (sizeof(T) == 4) {
            // pairs of complex per lane (16 bytes); the odd last element k = M alone
            for (int idx = threadIdx.x; idx < L * (M / 2 + 1); idx += TPB) {
                const int line = idx / (M / 2 + 1), kp = idx - line * (M / 2 + 1);
                const int row = r0 + line;
                if (row >= rows) continue;
                cpx<T> *dst = Fk + (size_t)kspace_row<N, KB>(row) * P + 2 * kp;
                const cpx<T> *s = buf + line * LS + 2 * kp;
                if (kp < M / 2) *reinterpret_cast<VT *>(dst) = VT{{s[0].x, s[0].y, s[1].x, s[1].y}};
                else *dst = s[0];
            }
        } else {
            for (int idx = threadIdx.x; idx < L * (M + 1); idx += TPB) {
                const int line = idx / (M + 1), k = idx - line * (M + 1);
                const int row = r0 + line;
                if (row < rows) Fk[(size_t)kspace_row<N, KB>(row) * P + k] = buf[line * LS + k];
            }
        }
        lds_sync();
    }
}

// ---------------------------------------------------------------- y / z columns
// MODE 0 forward, 1 inverse, 2 forward * (-k^2 kfac) inverse (z axis only).
// FUSE (MODE 2): the last forward pass, the -k^2 factor and the first pass of
// the inverse (radices in reverse order) run in registers in one step, which
// saves one LDS write + read of the tile and a barrier pair.
// PP (with FUSE): the LDS passes run out of place between two tile buffers.
// BLK (blocked k-space order, KBLK planes per block; s_blk = N * KBLK * P):
//   0: FFTW order, line offsets n * s_axis, tile offsets o * s_other;
//   1: y pass on the blocked order: the tile's kz = o at (o / KBLK) s_blk +
//      (o % KBLK) s_other (s_other = P), lines along ky with s_axis = KBLK P;
//   2: z pass on the blocked order: kz = n at (n / KBLK) s_blk + (n % KBLK)
//      s_axis (s_axis = P), tiles along ky with s_other = KBLK P.
template <typename T, int N, int TC, int MODE, int MR = 16, bool PF = true, int MINW = 1, int NT = TPB,
          bool TWL = false, bool FUSE = false, bool PP = false, int BLK = 0>
__device__ __forceinline__ void col_body(cpx<T> *__restrict__ Fk, const cpx<T> *__restrict__ tw, int nxh, int n_other,
                                         size_t s_axis, size_t s_other, size_t s_arr, int n_arr, int ny, double kfac,
                                         size_t ms_k, const Blk bk, cpx<T> *__restrict__ buf,
                                         cpx<T> *__restrict__ buf2, cpx<T> *__restrict__ twl, size_t s_blk = 0)
{
    auto line_off = [&](int n) -> size_t {
        if constexpr (BLK == 2) return (size_t)((unsigned)n / KBLK) * s_blk + (size_t)((unsigned)n % KBLK) * s_axis;
        else return (size_t)n * s_axis;
    };
    auto tile_off = [&](int o) -> size_t {
        if constexpr (BLK == 1) return (size_t)((unsigned)o / KBLK) * s_blk + (size_t)((unsigned)o % KBLK) * s_other;
        else return (size_t)o * s_other;
    };
    LAT_TWIDDLES_TO_LDS(N, NT, twl)
    assume_blk(bk);
    Fk += bk.y * ms_k;
    // Software pipeline per tile of TC adjacent columns:
    //   pass 0 straight from HBM into registers (radix R0, no twiddles),
    //   results to LDS, then the next tile's pass-0 loads are issued into the
    //   same (now dead) registers so they are in flight during the remaining
    //   LDS passes; the last pass writes straight to HBM.
    constexpr int NP = num_passes(N, MR);
    constexpr int SIGN0 = (MODE == 1) ? +1 : -1;
    constexpr int R0 = radix_at(N, 0, MR);
    constexpr int NR0 = N / R0;
    constexpr int TOT0 = TC * NR0;
    constexpr int B0 = cdiv(TOT0, NT);
    const int nct = (nxh + TC - 1) / TC;
    int t0, t1, ts;
    tile_range(nct * n_other * n_arr, bk, t0, t1, ts);
    cpx<T> v[B0][R0];
    auto tile_base = [&](int t, int &ncol, int &ct, int &o) -> cpx<T> * {
        ct = t % nct;
        const int rest = t / nct;
        o = rest % n_other;
        const int arr = rest / n_other;
        ncol = min(TC, nxh - ct * TC);
        return Fk + (size_t)arr * s_arr + tile_off(o) + (size_t)ct * TC;
    };
    auto load = [&](int t) {
        int ncol, ct, o;
        const cpx<T> *base = tile_base(t, ncol, ct, o);
#pragma unroll
        for (int b = 0; b < B0; ++b) {
            const int g = threadIdx.x + b * NT;
            const int line = g % TC, j = g / TC;
#pragma unroll
            for (int q = 0; q < R0; ++q) {
                cpx<T> a = {0, 0};
                if ((TOT0 % NT == 0 || g < TOT0) && line < ncol) a = base[line_off(j + q * NR0) + line];
                v[b][q] = a;
            }
        }
    };
    if (PF && t0 < t1) load(t0);
    for (int t = t0; t < t1; t += ts) {
        int ncol, ct, o;
        cpx<T> *base = tile_base(t, ncol, ct, o);
        if (!PF) load(t);
#pragma unroll
        for (int b = 0; b < B0; ++b) {
            const int g = threadIdx.x + b * NT;
            if (TOT0 % NT == 0 || g < TOT0) {
                const int line = g % TC, j = g / TC;
                dft_small<R0, SIGN0>(v[b]);
#pragma unroll
                for (int q = 0; q < R0; ++q) buf[(j * R0 + q) * TC + line] = v[b][q];
            }
        }
        if (PF && t + ts < t1) load(t + ts);
        lds_sync();
        auto store = [&](int line, int pos, cpx<T> a) {
            if (line < ncol) base[line_off(pos) + line] = a;
        };
        if constexpr (MODE == 2 && FUSE && PP && NP >= 2) {
            // pass 0 left the tile in buf; forward passes 1 .. NP-2, fused,
            // inverse passes 1 .. NP-2 alternate buffers; the data ends in buf2
            lds_fft_range_pp<T, N, -1, TC, NT, 1, NP - 1, MR, false, 0>(buf, buf2, twp);
            constexpr int FB = (NP - 2) & 1; // buffer holding the forward result
            const int jj = wmag(o, ny);
            const float kf = float(kfac);
            auto mul = [&](int c, int n) -> T {
                const int ii = ct * TC + c;
                const int kk = wmag(n, N);
                if constexpr (sizeof(T) == 4) {
                    return -float(ii * ii + jj * jj + kk * kk) * kf;
                } else {
                    const double rad2 = double(ii * ii) + double(jj * jj) + double(kk * kk);
                    return T(-rad2 * kfac);
                }
            };
            fused_last_first_pp<T, N, TC, NT, MR>(FB ? buf2 : buf, FB ? buf : buf2, twp, mul);
            lds_fft_range_pp<T, N, +1, TC, NT, 1, NP - 1, MR, true, FB ^ 1>(buf, buf2, twp);
            static_assert(((FB ^ 1) ^ ((NP - 2) & 1)) == 1, "inverse result in buf2");
            stockham_pass_out<T, N, NP - 1, +1, TC, NT, true, 1, MR, true>(buf2, 1, TC, twp, store);
        } else if constexpr (MODE == 2 && FUSE && NP >= 2) {
            lds_fft_range<T, N, -1, TC, NT, true, 1, 1, NP - 1, MR>(buf, 1, TC, twp);
            const int jj = wmag(o, ny);
            const float kf = float(kfac);
            auto mul = [&](int c, int n) -> T {
                const int ii = ct * TC + c;
                const int kk = wmag(n, N);
                if constexpr (sizeof(T) == 4) {
                    // |k|^2 in lattice units is an integer below 2^24 (N <= 1024):
                    // exact in FP32, one rounding in the product
                    return -float(ii * ii + jj * jj + kk * kk) * kf;
                } else {
                    const double rad2 = double(ii * ii) + double(jj * jj) + double(kk * kk);
                    return T(-rad2 * kfac);
                }
            };
            fused_last_first<T, N, TC, NT, MR>(buf, twp, mul);
            lds_fft_range<T, N, +1, TC, NT, true, 1, 1, NP - 1, MR, true>(buf, 1, TC, twp);
            stockham_pass_out<T, N, NP - 1, +1, TC, NT, true, 1, MR, true>(buf, 1, TC, twp, store);
        } else if constexpr (MODE == 2) {
            lds_fft_range<T, N, -1, TC, NT, true, 1, 1, NP, MR>(buf, 1, TC, twp);
            const int jj = wmag(o, ny);
            for (int idx = threadIdx.x; idx < TC * N; idx += NT) {
                const int c = idx % TC, n = idx / TC;
                const int ii = ct * TC + c;
                const int kk = wmag(n, N);
                const double rad2 = double(ii * ii) + double(jj * jj) + double(kk * kk);
                const T m = T(-rad2 * kfac);
                cpx<T> a = buf[idx];
                buf[idx] = {a.x * m, a.y * m};
            }
            lds_sync();
            lds_fft_range<T, N, +1, TC, NT, true, 1, 0, NP - 1, MR>(buf, 1, TC, twp);
            stockham_pass_out<T, N, NP - 1, +1, TC, NT, true, 1, MR>(buf, 1, TC, twp, store);
        } else if constexpr (NP == 1) {
            // single-pass transform: pass 0 was the whole FFT
            for (int idx = threadIdx.x; idx < TC * N; idx += NT) store(idx % TC, idx / TC, buf[idx]);
        } else {
            lds_fft_range<T, N, SIGN0, TC, NT, true, 1, 1, NP - 1, MR>(buf, 1, TC, twp);
            stockham_pass_out<T, N, NP - 1, SIGN0, TC, NT, true, 1, MR>(buf, 1, TC, twp, store);
        }
        lds_sync();
    }
}

// ---------------------------------------------------------------- x C2R + epilogue
template <typename T, int N, int L, int EPI, int MR = 4, bool PRE = true, bool TWL = false, int MINW = 1,
          int KB = 0>
__device__ __forceinline__ void x_c2r_body(const cpx<T> *__restrict__ Fk, const cpx<T> *__restrict__ tw, int rows,
                                           int P, size_t nk, size_t nvol, int nf, T *__restrict__ out,
                                           const T *__restrict__ fld, T *__restrict__ fldp,
                                           const DevScalars *__restrict__ dsc, ModelDev md,
                                           double *__restrict__ partials, MemberStride ms, const Blk bk,
                                           cpx<T> *__restrict__ buf, cpx<T> *__restrict__ twl)
{
    constexpr int M = N / 2;
    constexpr int LS = M + 2; // even pitch: every row starts 16-byte aligned (buf 16-byte aligned)
    constexpr int MAXF = 2;
    LAT_TWIDDLES_TO_LDS(N, TPB, twl)
    assume_blk(bk);
    Fk += bk.y * ms.k;
    if (out) out += bk.y * ms.f;
    if (fld) fld += bk.y * ms.f;
    if (fldp) fldp += bk.y * ms.f;
    dsc += bk.y;
    partials += (size_t)bk.y * MAX_GRID * NRED;
    md.p0 = dsc->mp[0];
    md.p1 = dsc->mp[1];
    md.p2 = dsc->mp[2];
    double acc[NRED] = {0, 0, 0, 0, 0, 0, 0, 0}; // RedChan order
    T hA = 0, hB = 0;
    if constexpr (EPI == EPI_KICK || EPI == EPI_GWKICK) {
        hA = T(dsc->hA);
        hB = T(dsc->hB);
    }
    float pb32[MAXF] = {0, 0};
    double pb64[MAXF] = {0, 0};
    T mu[MAXF] = {0, 0};
    if constexpr (sizeof(T) == 4 && (EPI == EPI_KICK || EPI == EPI_REDUCE)) {
        for (int fi = 0; fi < MAXF; ++fi) {
            pb64[fi] = dsc->mean_phi[fi];
            pb32[fi] = float(pb64[fi]);
            if constexpr (EPI == EPI_KICK) mu[fi] = T(dsc->mu[fi]);
        }
    }
    constexpr int VW = 16 / sizeof(T); // reals per 16-byte lane access
    using VT = Vec16<T>;
    // k-space units of 16 bytes per row: complex pairs in FP32 (the odd last
    // element k = M is a unit of its own), single complex in FP64
    constexpr int U = sizeof(T) == 4 ? M / 2 + 1 : M + 1;
    constexpr int FIT = (MAXF * L * U + TPB - 1) / TPB; // k-space units per thread
    constexpr int NV = L * N / VW;                      // real 16-byte slots per field
    constexpr int IT = (NV + TPB - 1) / TPB;            // slots per thread
    int t0, t1, ts;
    tile_range((rows + L - 1) / L, bk, t0, t1, ts);
    for (int t = t0; t < t1; t += ts) {
        const int r0 = t * L;
        // 1. issue the k-space loads of this tile
        VT fr[FIT];
#pragma unroll
        for (int q = 0; q < FIT; ++q) {
            const int idx = threadIdx.x + q * TPB;
            const int fl = idx / U, u = idx - fl * U;
            const int fi = fl / L, line = fl - fi * L;
            const int row = r0 + line;
            fr[q] = VT{};
            if (idx < nf * L * U && row < rows) {
                const cpx<T> *src = Fk + (size_t)fi * nk + (size_t)kspace_row<N, KB>(row) * P;
                if (sizeof(T) == 8 || u < M / 2) {
                    fr[q] = *reinterpret_cast<const VT *>(src + (sizeof(T) == 4 ? 2 * u : u));
                } else {
                    const cpx<T> v = src[M];
                    fr[q].e[0] = v.x;
                    fr[q].e[1] = v.y;
                }
            }
        }
        // 2. issue the epilogue's state loads now (PRE), so that their latency
        //    overlaps the transform below
        VT phv[IT][MAXF], pv[IT][MAXF];
        auto load_state = [&](int it) {
            const int idx = threadIdx.x + it * TPB;
            const int line = idx / (N / VW), i0 = (idx - line * (N / VW)) * VW;
            const int row = r0 + line;
            const bool ok = idx < NV && row < rows;
            const size_t site = (size_t)row * N + i0;
#pragma unroll
            for (int fi = 0; fi < MAXF; ++fi) {
                phv[it][fi] = VT{};
                pv[it][fi] = VT{};
                if (ok && fi < nf) {
                    phv[it][fi] = *reinterpret_cast<const VT *>(fld + (size_t)fi * nvol + site);
                    pv[it][fi] = *reinterpret_cast<const VT *>(fldp + (size_t)fi * nvol + site);
                }
            }
        };
        if constexpr (EPI != EPI_STORE && PRE) {
#pragma unroll
            for (int it = 0; it < IT; ++it) load_state(it);
        }
        // 3. k-space units into LDS
#pragma unroll
        for (int q = 0; q < FIT; ++q) {
            const int idx = threadIdx.x + q * TPB;
            const int fl = idx / U, u = idx - fl * U;
            if (idx < nf * L * U) {
                cpx<T> *d = buf + fl * LS;
                if constexpr (sizeof(T) == 8) {
                    d[u] = cpx<T>{fr[q].e[0], fr[q].e[1]};
                } else {
                    d[2 * u] = cpx<T>{fr[q].e[0], fr[q].e[1]};
                    if (u < M / 2) d[2 * u + 1] = cpx<T>{fr[q].e[2], fr[q].e[3]};
                }
            }
        }
        lds_sync();
        // Hermitian pre-split: Z[k] = (X[k] + X*[M-k]) + i e^{+2 pi i k/N} (X[k] - X*[M-k]),
        // with Im X[0] = Im X[M] = 0 (imaginary parts of the x-DC and x-Nyquist
        // coefficients are discarded, as numpy irfftn / FFTW c2r rows do).
        for (int idx = threadIdx.x; idx < nf * L * (M / 2 + 1); idx += TPB) {
            const int fl = idx / (M / 2 + 1), k = idx - fl * (M / 2 + 1);
            cpx<T> *z = buf + fl * LS;
            cpx<T> a = z[k], b = z[M - k];
            if (k == 0) { a.y = 0; b.y = 0; }
            {
                cpx<T> w = twp[k];
                w.y = -w.y; // e^{+2 pi i k/N}
                const cpx<T> s = {a.x + b.x, a.y - b.y};
                const cpx<T> d = {a.x - b.x, a.y + b.y};
                const cpx<T> wd = cmul(w, d);
                z[k] = {s.x - wd.y, s.y + wd.x};
            }
            if (k != 0 && k != M - k) {
                const int k2 = M - k;
                cpx<T> w = twp[k2];
                w.y = -w.y;
                const cpx<T> s = {b.x + a.x, b.y - a.y};
                const cpx<T> d = {b.x - a.x, b.y + a.y};
                const cpx<T> wd = cmul(w, d);
                z[k2] = {s.x - wd.y, s.y + wd.x};
            }
        }
        lds_sync();
        if (nf == 2) lds_fft_range<T, M, +1, 2 * L, TPB, false, 2, 0, num_passes(M, MR), MR>(buf, LS, 1, twp);
        else lds_fft_range<T, M, +1, L, TPB, false, 2, 0, num_passes(M, MR), MR>(buf, LS, 1, twp);
        // 4. epilogue: VW consecutive sites per thread, every global access a
        //    16-byte vector per lane (4 sites in FP32, 2 in FP64).  Field loops
        //    are unrolled over MAXF with an nf guard so that every register
        //    array is indexed by compile-time constants.
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const int idx = threadIdx.x + it * TPB;
            const int line = idx / (N / VW), i0 = (idx - line * (N / VW)) * VW;
            const int row = r0 + line;
            if (idx >= NV || row >= rows) continue;
            const size_t site = (size_t)row * N + i0;
            VT lap[MAXF];
#pragma unroll
            for (int fi = 0; fi < MAXF; ++fi)
                if (fi < nf)
                    lap[fi] =
                        *reinterpret_cast<const VT *>(reinterpret_cast<const T *>(buf + (fi * L + line) * LS) + i0);
            if constexpr (EPI == EPI_STORE) {
#pragma unroll
                for (int fi = 0; fi < MAXF; ++fi)
                    if (fi < nf) *reinterpret_cast<VT *>(out + (size_t)fi * nvol + site) = lap[fi];
            } else if constexpr (EPI == EPI_GWKICK) {
                // tensor components: phv holds the source S, pv the momenta pi_u
                if constexpr (!PRE) load_state(it);
                const T gwc = T(md.gwc);
#pragma unroll
                for (int fi = 0; fi < MAXF; ++fi) {
                    if (fi >= nf) break;
#pragma unroll
                    for (int v = 0; v < VW; ++v)
                        pv[it][fi].e[v] = pv[it][fi].e[v] + hA * (lap[fi].e[v] + gwc * phv[it][fi].e[v]);
                    *reinterpret_cast<VT *>(fldp + (size_t)fi * nvol + site) = pv[it][fi];
                }
            } else {
                if constexpr (!PRE) load_state(it);
#pragma unroll
                for (int v = 0; v < VW; ++v) {
                    T phi[MAXF] = {0, 0}, dv[MAXF] = {0, 0};
#pragma unroll
                    for (int fi = 0; fi < MAXF; ++fi)
                        if (fi < nf) phi[fi] = phv[it][fi].e[v];
                    if constexpr (sizeof(T) == 4) {
                        // mixed form: phi, p hold fluctuations; V' and V as differences
                        if constexpr (EPI == EPI_KICK) model_ddv(md, pb32, phi, dv);
                        acc[RC_PE] += model_dvv(md, pb64, phi);
                    } else {
                        if constexpr (EPI == EPI_KICK) model_dv(md, phi, dv);
                        acc[RC_PE] += model_v(md, phi);
                    }
                    acc[RC_M1] += double(phi[0]);
                    acc[RC_M2] += double(phi[1]);
#pragma unroll
                    for (int fi = 0; fi < MAXF; ++fi) {
                        if (fi >= nf) break;
                        T p = pv[it][fi].e[v];
                        if constexpr (EPI == EPI_KICK) {
                            p = p + (hA * lap[fi].e[v] - hB * dv[fi]);
                            if constexpr (sizeof(T) == 4) p = p - mu[fi];
                            pv[it][fi].e[v] = p;
                        }
                        acc[RC_GE] -= double(phi[fi]) * double(lap[fi].e[v]);
                        acc[RC_KE] += double(p) * double(p);
                        acc[RC_S1 + fi] += double(p);
                    }
                }
                if constexpr (EPI == EPI_KICK) {
#pragma unroll
                    for (int fi = 0; fi < MAXF; ++fi)
                        if (fi < nf) *reinterpret_cast<VT *>(fldp + (size_t)fi * nvol + site) = pv[it][fi];
                }
            }
        }
        lds_sync();
    }
    if constexpr (EPI == EPI_KICK || EPI == EPI_REDUCE)
        block_reduce_store<NRED>(acc, partials + (size_t)bk.x * NRED);
}

// ---------------------------------------------------------------- launch helpers
// Launch shapes, chosen from sweeps of tile width, maximum radix, prefetch
// and block size at 1024^3 on MI355X (see the measurement notes in the
// repository documentation) and scaled by the line length N for other sizes.
constexpr int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// rows per workgroup: x R2C 2048/(N/2) complex elements (FP32 at N = 1024:
// 2 rows, one block per tile), x C2R 1024/(N/2)
template <typename T, int N>
constexpr int r2c_lines() { return (sizeof(T) == 4 && N == 1024) ? 2 : clampi(2048 / (N / 2), 1, 512); }
template <int N>
constexpr int c2r_lines() { return clampi(1024 / (N / 2), 1, 256); }
template <typename T>
constexpr int r2c_maxradix() { return sizeof(T) == 8 ? 4 : 8; }

// Grid sizes (1024^3 sweep of 1024 ... all tiles): the x R2C runs fastest
// with one block per tile; the x C2R + kick with 8192 blocks in FP32 and
// 1024 in FP64 (it writes one reduction partial per block, at most
// MAX_GRID); the y passes with 65536; the fused z pass with 8192 / 1024.
constexpr int R2C_GRID = 1 << 30;
template <typename T>
constexpr int c2r_grid() { return sizeof(T) == 8 ? 1024 : 8192; }
static_assert(c2r_grid<float>() <= MAX_GRID && c2r_grid<double>() <= MAX_GRID, "one partial per block");
// Lattices up to 256 per side: at most 2048 blocks, so the scalar kernel
// that sums one reduction partial per block has a quarter of them to read
// (at 256^3 that kernel was 10 % of the mixed-form step with 8192).
template <typename T, int N>
constexpr int c2r_grid_n() { return (N <= 256 && c2r_grid<T>() > 2048) ? 2048 : c2r_grid<T>(); }

// y/z column passes: TC adjacent kx columns per tile.  y passes: 8 columns
// at N = 1024 (128 B rows in FP64), radix <= 16, 512 threads.  Fused z pass:
// last forward / first inverse pass in registers (FUSE), LDS passes out of
// place between two tile buffers (PP), the next tile's first-pass loads in
// flight during the LDS passes (PF), 512 threads; FP32 8 columns, radix
// <= 16; FP64 4 columns, radix <= 8.
template <typename T, int N, int MODE>
struct ColCfg {
    static constexpr bool Z = MODE == 2;
    static constexpr bool F64 = sizeof(T) == 8;
    // radix-32 passes (1024 = 32 x 32: two passes, one LDS round trip per
    // transform) where the 1024^3 sweeps measured a gain: every y pass and
    // the FP64 z pass, 16 FP32 / 8 FP64 columns per tile
    static constexpr bool R32 = N == 1024 && !(Z && !F64);
    static constexpr int TC = R32 ? (F64 ? 8 : 16)
                                  : (Z && F64) ? clampi(4096 / N, 4, 64) : clampi(8192 / N, 8, 64);
    static constexpr int MR = R32 ? 32 : (Z && F64) ? 8 : 16;
    static constexpr bool PF = Z && !R32;
    static constexpr int NT = (R32 && F64) ? 256 : 512;
    static constexpr bool FUSE = Z;
    static constexpr bool PP = Z && !R32;
    static constexpr int GRID = Z ? ((F64 && !R32) ? 1024 : 8192) : 65536;
};

#define LAT_FOR_SIZES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(48) X(96) X(192)

// Sizes up to 256 carry a second launch shape per sweep that the autotuner
// (lat_autotune) may pick: half the rows per workgroup for the x sweeps; for
// the y/z sweeps the largest power-of-two column count up to half the
// default (whole 8-byte-multiple row segments: 192^3 takes 16, not 21), run
// by exactly the threads of its first pass, rounded to whole waves (the
// default shape keeps its 512 threads).  Measured on MI355X: 128^3 y pass
// 0.418 -> 0.357 ms and fused z pass 0.546 -> 0.402 ms (32 columns, 256
// threads), 192^3 y pass 0.462 -> 0.398 ms (16 columns, 192 threads).
template <int N>
constexpr bool tunable() { return N <= 256; }
constexpr int pow2_floor(int v) { return v >= 2 ? 2 * pow2_floor(v / 2) : 1; }
template <typename T, int N>
constexpr int col_tc(int mode, int v)
{
    return v ? clampi(pow2_floor((mode == 2 ? ColCfg<T, N, 2>::TC : ColCfg<T, N, 0>::TC) / 2), 4, 64)
             : (mode == 2 ? ColCfg<T, N, 2>::TC : ColCfg<T, N, 0>::TC);
}
template <typename T, int N, int MODE, int TC>
constexpr int col_nt()
{
    using C = ColCfg<T, N, MODE == 2 ? 2 : 0>;
    if constexpr (!tunable<N>() || TC == col_tc<T, N>(MODE == 2 ? 2 : 0, 0)) {
        return C::NT;
    } else {
        constexpr int t = TC * (N / radix_at(N, 0, C::MR));
        return clampi((t + 63) / 64 * 64, 64, 512);
    }
}

// Twiddles in LDS (TWL) in every sweep: the 1024^3 sweeps measured a gain
// (or no loss) in each.
template <typename T>
constexpr bool r2c_twl() { return true; }

} // namespace lat
