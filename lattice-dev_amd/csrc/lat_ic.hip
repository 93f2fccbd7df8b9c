// lat_ic.hip -- Gaussian-random-field initial conditions on the device.
//
// Restates `sample(gamma, m2eff)` of Pseudospec/openmp_dev/evolve_spectral.f90:421-506:
//   1. oversampled radial profile ker(k) = kk (kk^2+m2eff)^gamma exp(-(kk/kcut)^2),
//      kk = (k-1/2) dkos, k = 1..nos, nos = max(n)*16^2           (:430-453)
//   2. DST-II (FFTW RODFT10) of the profile, then ker(k) *= norm/k (:459-463)
//   3. 3-D kernel by linear interpolation at radius |x - x_nn| * os   (:465-490)
//   4. R2C, per-mode Rayleigh amplitude sqrt(-2 ln a) and phase e^{2 pi i p} (:494-500)
//   5. C2R (unnormalised; the 1/nvol is inside norm)                  (:503)
// The compiler intrinsic random_number of the reference is replaced by a
// Philox-4x32-10 stream (Salmon et al., SC'11): counter = FFTW-layout mode
// index i + (nx/2+1)(j + ny k), 64-bit stream id, 64-bit seed.  Only the first
// K = floor(r_max*os)+2 outputs of the DST are used by step 3, so the DST is
// evaluated as that many direct sums with exact integer argument reduction.
#include <algorithm>
#include <cmath>

#include "lat_internal.h"

namespace lat {

int launch_stencil_io(Ctx &c, const double *f, double *out);

__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1,
                                       uint32_t out[4])
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        const uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

__global__ void k_ker_src(double *__restrict__ X, int nos, double dkos, double m2eff, double gamma, double kcut)
{
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nos; j += gridDim.x * blockDim.x) {
        const double kk = ((double)(j + 1) - 0.5) * dkos;
        X[j] = kk * pow(kk * kk + m2eff, gamma) * exp(-(kk / kcut) * (kk / kcut));
    }
}

// ker[k] = norm/(k+1) * 2 sum_j X_j sin(pi (2j+1)(k+1) / (2 nos)); one workgroup per k.
__global__ __launch_bounds__(256) void k_dst(const double *__restrict__ X, double *__restrict__ ker, int nos,
                                             double norm)
{
    const int k = blockIdx.x;
    const long long m4 = 4LL * nos;
    const long long kp = k + 1;
    double s = 0.0;
    for (int j = threadIdx.x; j < nos; j += 256) {
        const long long m = ((2LL * j + 1) * kp) % m4;
        s += X[j] * sinpi((double)m / (double)(2LL * nos));
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) ker[k] = norm * (2.0 * red[0]) / (double)(k + 1);
}

__global__ void k_fill_kernel3d(double *__restrict__ out, const double *__restrict__ ker, int nx, int ny, int nz,
                                int nn, int os)
{
    const size_t nvol = (size_t)nx * ny * nz;
    for (size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x; s < nvol; s += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(s % nx) + 1, j = (int)((s / nx) % ny) + 1, k = (int)(s / ((size_t)nx * ny)) + 1;
        const double di = i - nn, dj = j - nn, dk = k - nn;
        const double kk = sqrt(di * di + dj * dj + dk * dk) * os;
        const int l = (int)floor(kk);
        out[s] = (l > 0) ? ker[l - 1] + (kk - l) * (ker[l] - ker[l - 1]) : (4.0 * ker[0] - ker[1]) / 3.0;
    }
}

__global__ void k_randomize(cpx<double> *__restrict__ F, int nxh, int P, int ny, int nz, uint64_t seed,
                            uint64_t stream)
{
    const size_t n = (size_t)nxh * ny * nz;
    const double inv53 = 1.0 / 9007199254740992.0;
    const double twopi = 6.2831853071795864769252867665590;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        uint32_t o[4];
        philox((uint32_t)id, (uint32_t)(id >> 32), (uint32_t)stream, (uint32_t)(stream >> 32), (uint32_t)seed,
               (uint32_t)(seed >> 32), o);
        const uint64_t a = ((uint64_t)(o[0] >> 5) << 26) | (o[1] >> 6);
        const uint64_t b = ((uint64_t)(o[2] >> 5) << 26) | (o[3] >> 6);
        const double u1 = 1.0 - (double)a * inv53, u2 = (double)b * inv53;
        const double amp = sqrt(-2.0 * log(u1));
        const double ph = twopi * u2;
        const double cr = amp * cos(ph), ci = amp * sin(ph);
        const size_t i = id % nxh, row = id / nxh;
        cpx<double> &f = F[row * P + i];
        const cpx<double> v = f;
        f = {v.x * cr - v.y * ci, v.x * ci + v.y * cr};
    }
}

template <typename T>
__global__ void k_store_field(T *__restrict__ dst, const double *__restrict__ src, size_t n, double add)
{
    for (size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x; s < n; s += (size_t)gridDim.x * blockDim.x)
        dst[s] = (T)(add + src[s]);
}

// Returns 0 or an error; result (unnormalised C2R of the randomised kernel) in c.io64[0..nvol).
int sample_to_io(Ctx &c, double gamma, double m2eff, double mpl, double kcut_fac, uint64_t seed, uint64_t stream)
{
    const int os = 16;
    const int nmax = std::max(c.nx, std::max(c.ny, c.nz));
    const int nos = nmax * os * os;
    const int nn = std::min(c.nx / 2 + 1, std::min(c.ny / 2 + 1, c.nz / 2 + 1));
    const double twopi = 6.2831853071795864769252867665590;
    const double dx = c.cfg.len / c.nx, dk = twopi / c.cfg.len;
    const double dxos = dx / os, dkos = dk / (2.0 * os);
    const double kcut = kcut_fac * (double)nn * dk / 2.0;
    const double nvol = (double)c.nx * (double)c.ny * (double)c.nz;
    const double norm = 0.5 / (nvol * std::sqrt(twopi * dk * dk * dk) * mpl) * (dkos / dxos);
    const double dmax_x = c.nx - nn, dmax_y = c.ny - nn, dmax_z = c.nz - nn;
    const double rmax = std::sqrt(std::max(dmax_x * dmax_x, (double)(nn - 1) * (nn - 1)) +
                                  std::max(dmax_y * dmax_y, (double)(nn - 1) * (nn - 1)) +
                                  std::max(dmax_z * dmax_z, (double)(nn - 1) * (nn - 1)));
    int K = (int)std::floor(rmax * os) + 2;
    if (K > nos) K = nos;
    double *X = nullptr, *ker = nullptr;
    LAT_HIP(&c, hipMallocAsync((void **)&X, sizeof(double) * nos, c.stream));
    LAT_HIP(&c, hipMallocAsync((void **)&ker, sizeof(double) * (K + 1), c.stream));
    hipLaunchKernelGGL(k_ker_src, dim3((nos + 255) / 256), dim3(256), 0, c.stream, X, nos, dkos, m2eff, gamma, kcut);
    hipLaunchKernelGGL(k_dst, dim3(K), dim3(256), 0, c.stream, X, ker, nos, norm);
    hipLaunchKernelGGL(k_fill_kernel3d, dim3(8192), dim3(256), 0, c.stream, c.io64, ker, c.nx, c.ny, c.nz, nn, os);
    LAT_HIP(&c, hipGetLastError());
    cpx<double> *F = (cpx<double> *)c.Fk64;
    int rc;
    if ((rc = launch_x_r2c<double>(c, c.twx64, c.io64, nullptr, F, 0))) return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 0, 1))) return rc;
    if ((rc = launch_col<double>(c, c.twz64, F, 2, 0, 1))) return rc;
    hipLaunchKernelGGL(k_randomize, dim3(8192), dim3(256), 0, c.stream, F, c.nxh, c.P, c.ny, c.nz, seed, stream);
    LAT_HIP(&c, hipGetLastError());
    if ((rc = launch_col<double>(c, c.twz64, F, 2, 1, 1))) return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 1, 1))) return rc;
    rc = launch_x_c2r<double>(c, c.twx64, F, 1, EPI_STORE, c.io64, nullptr, nullptr);
    if (rc < 0) return rc;
    LAT_HIP(&c, hipFreeAsync(X, c.stream));
    LAT_HIP(&c, hipFreeAsync(ker, c.stream));
    return 0;
}

template <typename T>
int store_io_field(Ctx &c, T *dst, double add)
{
    hipLaunchKernelGGL((k_store_field<T>), dim3(8192), dim3(256), 0, c.stream, dst, c.io64, c.nvol, add);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

template int store_io_field<double>(Ctx &, double *, double);
template int store_io_field<float>(Ctx &, float *, double);

} // namespace lat
