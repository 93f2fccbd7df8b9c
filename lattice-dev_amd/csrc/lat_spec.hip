// lat_spec.hip -- shell-binned power and cross spectra on gfx950.
//
// Restates the binning loop of spectrum_3d / crossspec_3d
// (Pseudospec/openmp_dev/analysis_spec.f90:84-101,157-177): for every mode,
// p = |k|/dk, l = floor(p), the two-bin weights c = (1-(l-p)^2)^2 and
// (1-(l+1-p)^2)^2 go to bins l and l+1 (0-based).  The reference loops over
// all nx x-indices and re-reads Fk(nx-i) for i > nx/2; on the half-complex
// array that is a multiplicity of 2 for 0 < kx < nx/2.
//
// Determinism: each wave walks one (ky,kz) row at a time with the 64 lanes
// on consecutive kx (4 waves per workgroup share one LDS histogram).  Along a
// row p is non-decreasing in kx, so equal bins form contiguous lane
// segments; a segmented DPP scan sums each segment and only segment tails
// update the histogram, one wave after the other in wave order (no atomics).
// Workgroup histograms are reduced in block order by k_bin_reduce (groups of
// 64 blocks, then the groups).  Results are bitwise reproducible run to run.
#include <mutex>
#include <string>

#include "lat_internal.h"

namespace lat {

// TT projection of a symmetric tensor at wavevector kv (signed integers):
// sum_ab |(Lambda U)_ab|^2 with Lambda_ab,lm = P_al P_bm - P_ab P_lm / 2,
// P = 1 - k k^T / k^2 (gw-projectors.f90:13-26 is a stub); 0 at k = 0.
// U holds the components xx, xy, xz, yy, yz, zz.
__device__ __forceinline__ double tt_norm2(const double kv[3], const double ur[6], const double ui[6])
{
    const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    if (k2 == 0.0) return 0.0;
    double P[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) P[a][b] = (a == b ? 1.0 : 0.0) - kv[a] * kv[b] / k2;
    const int ix[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double mr[3][3], mi[3][3]; // M = P U
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int l = 0; l < 3; ++l) {
                sr += P[a][l] * ur[ix[l][b]];
                si += P[a][l] * ui[ix[l][b]];
            }
            mr[a][b] = sr;
            mi[a][b] = si;
        }
    double tr = 0.0, ti = 0.0; // P_lm U_lm = trace(P U)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        tr += mr[a][a];
        ti += mi[a][a];
    }
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            double rr = -0.5 * P[a][b] * tr, ri = -0.5 * P[a][b] * ti; // (M P)_ab - P_ab tr / 2
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                rr += mr[a][m] * P[m][b];
                ri += mi[a][m] * P[m][b];
            }
            acc += rr * rr + ri * ri;
        }
    return acc;
}

// channel sets: mode 0 -> |F1|^2 ; mode 1 -> Re, Im of F1 conj(F2) ;
// mode 2 -> |F2|^2, |F1|^2, Re, Im of F1 conj(F2) ; mode 3 -> sum_ab |Lambda U|^2
// of the six arrays F1 + c*nk ; mode 4 -> no value channel (W only).  With
// WCH the weight W is binned too, as the last channel; it depends on the
// geometry and the flags only, so launch_bin bins it once (mode 4) and runs
// the value modes without it.
// flags (Discrete/Entropy/analysis.f90 variants):
//   BIN_COUNTS  integer-count bins floor(p) (spectrum_bin, :22-52) instead of
//               the two-bin weights;
//   BIN_HALF    weight 1 per stored half-space mode, no Hermitian mirror (the
//               `nofft` routines, :54-186);
//   BIN_NO_NYQ  skip the x-Nyquist plane (spectrum_bin_nofft :69,
//               crossspec_nofft :166).
constexpr int BIN_ROWCHUNKS = 9; // 64-mode chunks per k-space row (nx <= 1024)
constexpr int BIN_WAVES = 4;      // waves per workgroup, sharing one histogram

// Cross-lane moves for the segmented scan with DPP (GFX9 data-parallel
// primitives: row_shr:n within 16-lane rows, row_bcast:15 / row_bcast:31
// across rows) instead of LDS-routed shuffles.  A lane whose source is out
// of range, or whose row is masked off, receives the `old` value: key -1 (no
// match) and 0.0.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_key(int v)
{
    return __builtin_amdgcn_update_dpp(-1, v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_val(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// one step of the segmented inclusive scan: add the moved partial sums where
// the moved key equals the lane's key (keys are sorted along the wave, so
// equal keys are one contiguous segment)
template <int CTRL, int RM, int NC>
__device__ __forceinline__ void seg_scan_step(int key, double (&v1)[NC], double (&v2)[NC])
{
    const int ok = dpp_key<CTRL, RM>(key);
    double o1[NC], o2[NC];
#pragma unroll
    for (int ch = 0; ch < NC; ++ch) {
        o1[ch] = dpp_val<CTRL, RM>(v1[ch]);
        o2[ch] = dpp_val<CTRL, RM>(v2[ch]);
    }
    if (ok == key) {
#pragma unroll
        for (int ch = 0; ch < NC; ++ch) {
            v1[ch] += o1[ch];
            v2[ch] += o2[ch];
        }
    }
}
template <int NC>
__device__ __forceinline__ void seg_scan(int key, double (&v1)[NC], double (&v2)[NC])
{
    seg_scan_step<0x111, 0xF>(key, v1, v2); // row_shr:1
    seg_scan_step<0x112, 0xF>(key, v1, v2); // row_shr:2
    seg_scan_step<0x114, 0xF>(key, v1, v2); // row_shr:4
    seg_scan_step<0x118, 0xF>(key, v1, v2); // row_shr:8
    seg_scan_step<0x142, 0xA>(key, v1, v2); // row_bcast:15 into rows 1, 3
    seg_scan_step<0x143, 0xC>(key, v1, v2); // row_bcast:31 into rows 2, 3
}

template <typename T, int NCH, int MODE, bool WCH>
__global__ __launch_bounds__(64 * BIN_WAVES) void k_bin(const cpx<T> *__restrict__ F1, const cpx<T> *__restrict__ F2,
                                                        int nx, int ny, int nz, int nxh, int P, int ns,
                                                        double *__restrict__ part, int flags, size_t nk)
{
    constexpr int NC = NCH + (WCH ? 1 : 0); // binned channels
    extern __shared__ double hist[];       // NC * ns
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = threadIdx.x; q < NC * ns; q += 64 * BIN_WAVES) hist[q] = 0.0;
    __syncthreads();
    const int rows = ny * nz;
    const bool counts = (flags & BIN_COUNTS) != 0;
    // one 64-lane chunk of a row: value(s) a, b of mode i = i0 + lane
    // one 64-lane chunk of a row: segment sums of bins l (v1) and l + 1 (v2)
    // at the segment tails (tail = true)
    auto chunk = [&](bool has_row, int row, int j, int k, int jj, int kk, int i0, cpx<T> a, cpx<T> b, int &key,
                     bool &tail, double (&v1)[NC], double (&v2)[NC]) {
        const int i = i0 + lane;
        const bool valid = has_row && i < nxh && !((flags & BIN_NO_NYQ) && 2 * i == nx);
        key = 0x7fffffff;
        if (valid) {
            const double p = sqrt(double(i * i + jj * jj + kk * kk));
            const int l = (int)floor(p);
            const double d1 = double(l) - p, d2 = double(l + 1) - p;
            const double c1 = counts ? 1.0 : (1.0 - d1 * d1) * (1.0 - d1 * d1);
            const double c2 = counts ? 0.0 : (1.0 - d2 * d2) * (1.0 - d2 * d2);
            const double mult = ((flags & BIN_HALF) || i == 0 || (2 * i == nx)) ? 1.0 : 2.0;
            double val[NCH > 0 ? NCH : 1];
            if constexpr (MODE == 4) {
                val[0] = 0.0;
            } else if constexpr (MODE == 3) {
                const size_t id = (size_t)row * P + i;
                const double kv[3] = {double(i), double(j <= ny / 2 ? j : j - ny), double(k <= nz / 2 ? k : k - nz)};
                double ur[6], ui[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    const cpx<T> u = F1[(size_t)q * nk + id];
                    ur[q] = double(u.x);
                    ui[q] = double(u.y);
                }
                val[0] = tt_norm2(kv, ur, ui);
            } else if constexpr (MODE == 0) {
                val[0] = double(a.x) * a.x + double(a.y) * a.y;
            } else {
                const double re = double(a.x) * b.x + double(a.y) * b.y;
                const double im = double(a.y) * b.x - double(a.x) * b.y;
                if constexpr (MODE == 1) {
                    val[0] = re;
                    val[1] = im;
                } else {
                    val[0] = double(b.x) * b.x + double(b.y) * b.y;
                    val[1] = double(a.x) * a.x + double(a.y) * a.y;
                    val[2] = re;
                    val[3] = im;
                }
            }
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                v1[ch] = mult * c1 * val[ch];
                v2[ch] = mult * c2 * val[ch];
            }
            if constexpr (WCH) {
                v1[NCH] = mult * c1;
                v2[NCH] = mult * c2;
            }
            key = l;
        } else {
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) v1[ch] = v2[ch] = 0.0;
        }
        // segmented inclusive scan over lanes with equal key
        seg_scan<NC>(key, v1, v2);
        const int nkey = __shfl_down(key, 1, 64);
        tail = valid && (lane == 63 || nkey != key);
    };
    // Histogram updates, one wave at a time in wave order (a barrier after
    // each), so the order of the additions into every bin is fixed.  Within
    // a wave the tails have distinct keys; bin l + 1 of one segment may be
    // bin l of the next, so the l + 1 additions follow the l additions in
    // program order.
    auto update = [&](int key, bool tail, const double (&v1)[NC], const double (&v2)[NC]) {
        for (int w = 0; w < BIN_WAVES; ++w) {
            if (wave == w && tail) {
#pragma unroll
                for (int ch = 0; ch < NC; ++ch) hist[ch * ns + key] += v1[ch];
                if (key + 1 < ns) {
#pragma unroll
                    for (int ch = 0; ch < NC; ++ch) hist[ch * ns + key + 1] += v2[ch];
                }
            }
            __syncthreads();
        }
    };
    // each wave walks its own row; the block's BIN_WAVES rows advance together
    for (int base = blockIdx.x * BIN_WAVES; base < rows; base += gridDim.x * BIN_WAVES) {
        const int row = base + wave;
        const bool has_row = row < rows;
        const int rr = has_row ? row : 0;
        const int j = rr % ny, k = rr / ny;
        const int jj = (j <= ny / 2) ? j : ny - j;
        const int kk = (k <= nz / 2) ? k : nz - k;
        int key;
        bool tail;
        double v1[NC], v2[NC];
        if constexpr (MODE == 3 || MODE == 4) {
            for (int i0 = 0; i0 < nxh; i0 += 64) {
                chunk(has_row, rr, j, k, jj, kk, i0, cpx<T>{0, 0}, cpx<T>{0, 0}, key, tail, v1, v2);
                update(key, tail, v1, v2);
            }
        } else {
            // the whole row's loads are issued first (up to BIN_ROWCHUNKS
            // chunks of 64 modes), so their latency is paid once per row
            cpx<T> a[BIN_ROWCHUNKS], b[BIN_ROWCHUNKS];
            const cpx<T> *r1 = F1 + (size_t)rr * P;
            const cpx<T> *r2 = MODE != 0 ? F2 + (size_t)rr * P : nullptr;
#pragma unroll
            for (int c = 0; c < BIN_ROWCHUNKS; ++c) {
                const int i = c * 64 + lane;
                a[c] = b[c] = cpx<T>{0, 0};
                if (has_row && i < nxh) {
                    a[c] = r1[i];
                    if constexpr (MODE != 0) b[c] = r2[i];
                }
            }
#pragma unroll
            for (int c = 0; c < BIN_ROWCHUNKS; ++c)
                if (c * 64 < nxh) {
                    chunk(has_row, rr, j, k, jj, kk, c * 64, a[c], b[c], key, tail, v1, v2);
                    update(key, tail, v1, v2);
                }
        }
    }
    __syncthreads();
    double *dst = part + (size_t)blockIdx.x * NC * ns;
    for (int q = threadIdx.x; q < NC * ns; q += 64 * BIN_WAVES) dst[q] = hist[q];
}

// out[g][q] = sum over blocks b of group g (in order) of part[b][q]; with
// group = nblk this is the final, single-group sum
__global__ __launch_bounds__(256) void k_bin_reduce(const double *__restrict__ part, int nblk, int group, int nq,
                                                    double *__restrict__ out)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    const int g = blockIdx.y;
    if (q >= nq) return;
    const int b0 = g * group, b1 = min(nblk, b0 + group);
    double s = 0.0;
    for (int b = b0; b < b1; ++b) s += part[(size_t)b * nq + q];
    out[(size_t)g * nq + q] = s;
}

// The two fixed-order reduction stages over the block histograms of nq
// doubles each: groups of BIN_GROUP blocks, then the groups, into out.
static void bin_reduce(Ctx &c, int g, int nq, double *out)
{
    const int ngrp = (g + BIN_GROUP - 1) / BIN_GROUP;
    double *grp = c.spec_part + (size_t)BIN_BLOCKS * nq;
    hipLaunchKernelGGL(k_bin_reduce, dim3((nq + 255) / 256, ngrp), dim3(256), 0, c.stream, c.spec_part, g, BIN_GROUP,
                       nq, grp);
    hipLaunchKernelGGL(k_bin_reduce, dim3((nq + 255) / 256, 1), dim3(256), 0, c.stream, grp, ngrp, ngrp, nq, out);
}

template <typename T>
int launch_bin(Ctx &c, const cpx<T> *F1, const cpx<T> *F2, int mode, int nbins, int flags)
{
    const int ns = nbins > 0 ? nbins : c.ns;
    if (ns > c.spec_cap) {
        c.last_error = "spectrum bin count exceeds the allocated histogram";
        return -1;
    }
    const int nch = (mode == 0 || mode == 3) ? 1 : (mode == 1 ? 2 : 4);
    const int rows = c.ny * c.nz;
    const int gr = (rows + BIN_WAVES - 1) / BIN_WAVES; // row groups
    const int g = gr < BIN_BLOCKS ? gr : BIN_BLOCKS;
    if (mode != 3 && c.nxh > BIN_ROWCHUNKS * 64) {
        c.last_error = "spectrum rows longer than the binning kernel's row buffer";
        return -1;
    }
    // weight channel: from the cache slot of (ns, flags), binned on a miss
    Ctx &owner = (c.view && c.parent) ? *c.parent : c;
    const long long key = ((long long)flags << 32) | (unsigned)ns;
    int slot = -1;
    for (int q = 0; q < Ctx::WCACHE; ++q)
        if (owner.wkey[q] == key) slot = q;
    const size_t cap = (size_t)c.spec_cap;
    if (slot < 0) {
        slot = owner.wnext;
        owner.wnext = (owner.wnext + 1) % Ctx::WCACHE;
        owner.wkey[slot] = -1;
        hipLaunchKernelGGL((k_bin<T, 0, 4, true>), dim3(g), dim3(64 * BIN_WAVES), sizeof(double) * ns, c.stream, F1, F2,
                           c.nx, c.ny, c.nz, c.nxh, c.P, ns, c.spec_part, flags, c.nk);
        LAT_HIP(&c, hipGetLastError());
        bin_reduce(c, g, ns, c.spec_out + (5 + slot) * cap);
        LAT_HIP(&c, hipGetLastError());
        owner.wkey[slot] = key;
    }
    const size_t lds = sizeof(double) * nch * ns;
    if (mode == 0)
        hipLaunchKernelGGL((k_bin<T, 1, 0, false>), dim3(g), dim3(64 * BIN_WAVES), lds, c.stream, F1, F2, c.nx, c.ny, c.nz,
                           c.nxh, c.P, ns, c.spec_part, flags, c.nk);
    else if (mode == 1)
        hipLaunchKernelGGL((k_bin<T, 2, 1, false>), dim3(g), dim3(64 * BIN_WAVES), lds, c.stream, F1, F2, c.nx, c.ny, c.nz,
                           c.nxh, c.P, ns, c.spec_part, flags, c.nk);
    else if (mode == 2)
        hipLaunchKernelGGL((k_bin<T, 4, 2, false>), dim3(g), dim3(64 * BIN_WAVES), lds, c.stream, F1, F2, c.nx, c.ny, c.nz,
                           c.nxh, c.P, ns, c.spec_part, flags, c.nk);
    else
        hipLaunchKernelGGL((k_bin<T, 1, 3, false>), dim3(g), dim3(64 * BIN_WAVES), lds, c.stream, F1, F2, c.nx, c.ny, c.nz,
                           c.nxh, c.P, ns, c.spec_part, flags, c.nk);
    LAT_HIP(&c, hipGetLastError());
    bin_reduce(c, g, nch * ns, c.spec_out);
    LAT_HIP(&c, hipGetLastError());
    // W after the nch value channels, as the callers read it
    LAT_HIP(&c, hipMemcpyAsync(c.spec_out + (size_t)nch * ns, c.spec_out + (5 + slot) * cap, sizeof(double) * ns,
                               hipMemcpyDeviceToDevice, c.stream));
    return nch;
}

// ---------------------------------------------------------------- allocation-free spectra
// The Discrete-tree nofft spectra of a device-resident transform with every
// piece of scratch (block histograms, group sums, weight and value channels)
// in one file-scope device array.  Host code never passes the array's name
// to a kernel: its device address is looked up once per device at context
// creation (noalloc_setup) and kept in noalloc_ptr.  Every call on a device
// rewrites the whole of the part it uses (each block stores its full
// histogram, each reduction stage its full output), so nothing has to start
// at zero; calls that overlap on one device race on the array.
constexpr size_t NOALLOC_DOUBLES = (size_t)4 << 20; // 32 MiB
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];
constexpr int NOALLOC_MAXDEV = 64;
static double *noalloc_ptr[NOALLOC_MAXDEV];
static std::mutex noalloc_mu;

int noalloc_setup(Ctx &c)
{
    if (c.device < 0 || c.device >= NOALLOC_MAXDEV) {
        c.last_error = "device index beyond the allocation-free scratch table";
        return -2;
    }
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[c.device]) return 0;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(&c, e, "hipGetSymbolAddress(scratch)");
    noalloc_ptr[c.device] = (double *)ptr;
    return 0;
}

// S[b] = raw[b] / W[b] / div where W[b] != 0, else raw[b]: finish_spectra's
// arithmetic on the device
__global__ __launch_bounds__(256) void k_spec_finish(const double *__restrict__ raw, const double *__restrict__ W,
                                                     int ns, double div, double *__restrict__ S)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= ns) return;
    double s = raw[b];
    if (W[b] != 0.0) s = s / W[b] / div;
    S[b] = s;
}

int spectrum_discrete_k_noalloc(Ctx &c, const double *Fk_dev, double *S_dev, int kind)
{
    if (c.device < 0 || c.device >= NOALLOC_MAXDEV || !noalloc_ptr[c.device]) {
        c.last_error = "allocation-free scratch not set up on this device";
        return -2;
    }
    const int ns = lat_num_bins_discrete(c.nx);
    const int rows = c.ny * c.nz;
    const int gr = (rows + BIN_WAVES - 1) / BIN_WAVES;
    const int g = gr < BIN_BLOCKS ? gr : BIN_BLOCKS;
    const int ngrp = (g + BIN_GROUP - 1) / BIN_GROUP;
    const size_t need = ((size_t)g + ngrp + 2) * ns;
    if (need > NOALLOC_DOUBLES) {
        c.last_error = "lattice too large for the allocation-free scratch (" + std::to_string(need * 8) +
                       " bytes needed, " + std::to_string(NOALLOC_DOUBLES * 8) + " held)";
        return -2;
    }
    if (c.nxh > BIN_ROWCHUNKS * 64) {
        c.last_error = "spectrum rows longer than the binning kernel's row buffer";
        return -2;
    }
    double *part = noalloc_ptr[c.device];
    double *grp = part + (size_t)g * ns;
    double *W = grp + (size_t)ngrp * ns;
    double *raw = W + ns;
    const int flags = kind == 0 ? BIN_HALF : (BIN_HALF | BIN_NO_NYQ | BIN_COUNTS);
    const cpx<double> *F = (const cpx<double> *)Fk_dev;
    const int P = c.nxh; // FFTW layout: rows of nx/2 + 1, unpadded
    auto reduce = [&](double *out) {
        hipLaunchKernelGGL(k_bin_reduce, dim3((ns + 255) / 256, ngrp), dim3(256), 0, c.stream, part, g, BIN_GROUP,
                           ns, grp);
        hipLaunchKernelGGL(k_bin_reduce, dim3((ns + 255) / 256, 1), dim3(256), 0, c.stream, grp, ngrp, ngrp, ns,
                           out);
    };
    // the weight channel, then the value channel, with the kernels, grid and
    // reduction order of launch_bin (same additions in the same order)
    hipLaunchKernelGGL((k_bin<double, 0, 4, true>), dim3(g), dim3(64 * BIN_WAVES), sizeof(double) * ns, c.stream, F, F,
                       c.nx, c.ny, c.nz, c.nxh, P, ns, part, flags, c.nk);
    reduce(W);
    hipLaunchKernelGGL((k_bin<double, 1, 0, false>), dim3(g), dim3(64 * BIN_WAVES), sizeof(double) * ns, c.stream, F,
                       F, c.nx, c.ny, c.nz, c.nxh, P, ns, part, flags, c.nk);
    reduce(raw);
    const double n3 = (double)c.nvol;
    hipLaunchKernelGGL(k_spec_finish, dim3((ns + 255) / 256), dim3(256), 0, c.stream, raw, W, ns,
                       kind == 0 ? n3 * n3 : 1.0, S_dev);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

template int launch_bin<double>(Ctx &, const cpx<double> *, const cpx<double> *, int, int, int);
template int launch_bin<float>(Ctx &, const cpx<float> *, const cpx<float> *, int, int, int);

} // namespace lat
