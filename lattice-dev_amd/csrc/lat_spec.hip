// lat_spec.hip -- shell-binned power and cross spectra on gfx950.
//
// Restates the binning loop of spectrum_3d / crossspec_3d
// (Pseudospec/openmp_dev/analysis_spec.f90:84-101,157-177): for every mode,
// p = |k|/dk, l = floor(p), the two-bin weights c = (1-(l-p)^2)^2 and
// (1-(l+1-p)^2)^2 go to bins l and l+1 (0-based).  The reference loops over
// all nx x-indices and re-reads Fk(nx-i) for i > nx/2; on the half-complex
// array that is a multiplicity of 2 for 0 < kx < nx/2.
//
// Determinism: one wave per workgroup walks one (ky,kz) row at a time with
// the 64 lanes on consecutive kx.  Along a row p is non-decreasing in kx, so
// equal bins form contiguous lane segments; a segmented shuffle scan sums each
// segment and only segment tails update the wave-private LDS histogram, in
// program order (no atomics).  Workgroup histograms are reduced in block
// order by k_bin_reduce.  Results are bitwise reproducible run to run.
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
// of the six arrays F1 + c*nk.  W is always the last channel.
// flags (Discrete/Entropy/analysis.f90 variants):
//   BIN_COUNTS  integer-count bins floor(p) (spectrum_bin, :22-52) instead of
//               the two-bin weights;
//   BIN_HALF    weight 1 per stored half-space mode, no Hermitian mirror (the
//               `nofft` routines, :54-186);
//   BIN_NO_NYQ  skip the x-Nyquist plane (spectrum_bin_nofft :69,
//               crossspec_nofft :166).
template <typename T, int NCH, int MODE>
__global__ __launch_bounds__(64) void k_bin(const cpx<T> *__restrict__ F1, const cpx<T> *__restrict__ F2, int nx,
                                            int ny, int nz, int nxh, int P, int ns, double *__restrict__ part,
                                            int flags, size_t nk)
{
    extern __shared__ double hist[]; // (NCH+1) * ns
    const int lane = threadIdx.x;
    for (int q = lane; q < (NCH + 1) * ns; q += 64) hist[q] = 0.0;
    __syncthreads();
    const int rows = ny * nz;
    for (int row = blockIdx.x; row < rows; row += gridDim.x) {
        const int j = row % ny, k = row / ny;
        const int jj = (j <= ny / 2) ? j : ny - j;
        const int kk = (k <= nz / 2) ? k : nz - k;
        for (int i0 = 0; i0 < nxh; i0 += 64) {
            const int i = i0 + lane;
            const bool counts = (flags & BIN_COUNTS) != 0;
            const bool valid = i < nxh && !((flags & BIN_NO_NYQ) && 2 * i == nx);
            double v1[NCH + 1], v2[NCH + 1];
            int key = 0x7fffffff;
            if (valid) {
                const double p = sqrt(double(i * i + jj * jj + kk * kk));
                const int l = (int)floor(p);
                const double d1 = double(l) - p, d2 = double(l + 1) - p;
                const double c1 = counts ? 1.0 : (1.0 - d1 * d1) * (1.0 - d1 * d1);
                const double c2 = counts ? 0.0 : (1.0 - d2 * d2) * (1.0 - d2 * d2);
                const double mult = ((flags & BIN_HALF) || i == 0 || (2 * i == nx)) ? 1.0 : 2.0;
                const size_t id = (size_t)row * P + i;
                double val[NCH];
                if constexpr (MODE == 3) {
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
                    const cpx<T> a = F1[id];
                    val[0] = double(a.x) * a.x + double(a.y) * a.y;
                } else {
                    const cpx<T> a = F1[id];
                    const cpx<T> b = F2[id];
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
                v1[NCH] = mult * c1;
                v2[NCH] = mult * c2;
                key = l;
            } else {
#pragma unroll
                for (int ch = 0; ch <= NCH; ++ch) v1[ch] = v2[ch] = 0.0;
            }
            // segmented inclusive scan over lanes with equal key
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int okey = __shfl_up(key, off, 64);
#pragma unroll
                for (int ch = 0; ch <= NCH; ++ch) {
                    const double o1 = __shfl_up(v1[ch], off, 64);
                    const double o2 = __shfl_up(v2[ch], off, 64);
                    if (lane >= off && okey == key) {
                        v1[ch] += o1;
                        v2[ch] += o2;
                    }
                }
            }
            const int nkey = __shfl_down(key, 1, 64);
            const bool tail = valid && (lane == 63 || nkey != key);
            // bin l first, then bin l+1 (which may be another segment's bin l):
            // the workgroup is one wave, the barriers order the two phases.
            if (tail) {
#pragma unroll
                for (int ch = 0; ch <= NCH; ++ch) hist[ch * ns + key] += v1[ch];
            }
            __syncthreads();
            if (tail && key + 1 < ns) {
#pragma unroll
                for (int ch = 0; ch <= NCH; ++ch) hist[ch * ns + key + 1] += v2[ch];
            }
            __syncthreads();
        }
    }
    __syncthreads();
    double *dst = part + (size_t)blockIdx.x * (NCH + 1) * ns;
    for (int q = lane; q < (NCH + 1) * ns; q += 64) dst[q] = hist[q];
}

// out[q] = sum over blocks b (in order) of part[b][q]
__global__ __launch_bounds__(256) void k_bin_reduce(const double *__restrict__ part, int nblk, int nq,
                                                    double *__restrict__ out)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(size_t)b * nq + q];
    out[q] = s;
}

constexpr int BIN_BLOCKS = 1024;

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
    const int g = rows < BIN_BLOCKS ? rows : BIN_BLOCKS;
    const size_t lds = sizeof(double) * (nch + 1) * ns;
    if (mode == 0)
        hipLaunchKernelGGL((k_bin<T, 1, 0>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh, c.P,
                           ns, c.spec_part, flags, c.nk);
    else if (mode == 1)
        hipLaunchKernelGGL((k_bin<T, 2, 1>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh, c.P,
                           ns, c.spec_part, flags, c.nk);
    else if (mode == 2)
        hipLaunchKernelGGL((k_bin<T, 4, 2>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh, c.P,
                           ns, c.spec_part, flags, c.nk);
    else
        hipLaunchKernelGGL((k_bin<T, 1, 3>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh, c.P,
                           ns, c.spec_part, flags, c.nk);
    LAT_HIP(&c, hipGetLastError());
    const int nq = (nch + 1) * ns;
    hipLaunchKernelGGL(k_bin_reduce, dim3((nq + 255) / 256), dim3(256), 0, c.stream, c.spec_part, g, nq, c.spec_out);
    LAT_HIP(&c, hipGetLastError());
    return nch;
}

template int launch_bin<double>(Ctx &, const cpx<double> *, const cpx<double> *, int, int, int);
template int launch_bin<float>(Ctx &, const cpx<float> *, const cpx<float> *, int, int, int);

} // namespace lat
