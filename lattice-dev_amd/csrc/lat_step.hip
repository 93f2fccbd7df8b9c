// lat_step.hip -- drift, finite-difference kick, reductions and the device
// update of the FRW background scalars.
//
// Every split sub-step of the reference ends in a lattice-wide FP64 reduction
// feeding a scalar update that the next sub-step reads
// (hamiltonian_conformal.f90:116-117, 203-205; hamiltonian_cosmic.f90:125-126,
// 209-211).  Here the sweep kernels write per-workgroup FP64 partials and one
// single-workgroup kernel (k_scalar) reduces them in a fixed order and applies
// the scalar part of the split operators between two kicks, so a whole
// Yoshida step runs without a host round trip.
#include "lat_internal.h"
#include "model.h"

namespace lat {

template <int NCH>
__device__ __forceinline__ void block_reduce_store2(double (&v)[NCH], double *out)
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
    __syncthreads();
    if (threadIdx.x < NCH) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < TPB / 64; ++w) s += red[threadIdx.x][w];
        out[threadIdx.x] = s;
    }
}

// Per-site potential terms of the two forms: full V, V' (FP64 form) or the
// difference forms about the homogeneous parts (mixed form, T = float).
template <typename T>
__device__ __forceinline__ void site_dv(const ModelDev &m, const float *pb32, const T *f, T *dv)
{
    if constexpr (sizeof(T) == 4) model_ddv(m, pb32, f, dv);
    else model_dv(m, f, dv);
}
template <typename T>
__device__ __forceinline__ double site_v(const ModelDev &m, const double *pb64, const T *f)
{
    if constexpr (sizeof(T) == 4) return model_dvv(m, pb64, f);
    else return model_v(m, f);
}

// ---------------------------------------------------------------- drift
// phi_l += dc[q] * pi_l for the pending drifts q < nd (hamiltonian_conformal.f90:98,
// hamiltonian_cosmic.f90:107); the p-update was done by k_scalar.  Mixed
// form: on the fluctuations, minus the re-centering shift nu[l].
// recenter = 0: tensor components u_ij, never re-centred.
template <typename T>
__global__ __launch_bounds__(TPB) void k_drift(T *__restrict__ f, const T *__restrict__ p, size_t nvol, int nf,
                                               const DevScalars *__restrict__ dsc, int recenter, size_t ms)
{
    f += blockIdx.y * ms;
    p += blockIdx.y * ms;
    dsc += blockIdx.y;
    const int nd = dsc->nd;
    T dc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) dc[q] = T(dsc->dc[q]);
    const bool rc = recenter && dsc->mixed;
    const size_t n = nvol * nf;
    for (size_t i = blockIdx.x * (size_t)TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
        T x = f[i];
        const T pi = p[i];
        for (int q = 0; q < nd; ++q) x = x + dc[q] * pi;
        if (rc) x = x - T(dsc->nu[i / nvol]);
        f[i] = x;
    }
}

// The same drift, 16 bytes per lane and one lane per vector: grid (vectors of
// one field / TPB, members, fields), so the field index (for the re-centring
// shift) is blockIdx.z and no lane divides.  Needs nvol a multiple of the
// vector width and 16-byte aligned field and member offsets.
template <typename T>
__global__ __launch_bounds__(TPB) void k_drift_v(T *__restrict__ f, const T *__restrict__ p, size_t nvol,
                                                 const DevScalars *__restrict__ dsc, int recenter, size_t ms)
{
    constexpr int V = 16 / sizeof(T);
    using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
    if (i >= nvol / V) return;
    const size_t off = blockIdx.y * ms + blockIdx.z * nvol;
    dsc += blockIdx.y;
    const int nd = dsc->nd;
    const T sh = (recenter && dsc->mixed) ? T(dsc->nu[blockIdx.z]) : T(0);
    Vec x = reinterpret_cast<const Vec *>(f + off)[i];
    const Vec pv = reinterpret_cast<const Vec *>(p + off)[i];
    T *xs = reinterpret_cast<T *>(&x);
    const T *ps = reinterpret_cast<const T *>(&pv);
    for (int q = 0; q < nd; ++q) {
        const T d = T(dsc->dc[q]);
#pragma unroll
        for (int e = 0; e < V; ++e) xs[e] = xs[e] + d * ps[e];
    }
    if (recenter && dsc->mixed) {
#pragma unroll
        for (int e = 0; e < V; ++e) xs[e] = xs[e] - sh;
    }
    reinterpret_cast<Vec *>(f + off)[i] = x;
}

template <typename T>
int launch_drift(Ctx &c)
{
    const int pi = prof_begin(c, LAT_ST_DRIFT);
    for (int part = 0; part < (c.gw ? 2 : 1); ++part) {
        const int nf = part == 0 ? c.nfld : 6;
        const size_t n = (size_t)nf * c.nvol;
        T *f = (T *)(part == 0 ? c.fld : c.gwu);
        const T *p = (const T *)(part == 0 ? c.fldp : c.gwp);
        const size_t ms = part == 0 ? c.ms.f : (size_t)6 * c.nvol;
        constexpr size_t V = 16 / sizeof(T);
        const bool vec = c.nvol % V == 0 && ms % V == 0 && ((uintptr_t)f | (uintptr_t)p) % 16 == 0 &&
                         c.nvol / V / TPB < 0x7fffffffu;
        if (vec) {
            const unsigned gx = (unsigned)((c.nvol / V + TPB - 1) / TPB);
            hipLaunchKernelGGL((k_drift_v<T>), dim3(gx, c.B, nf), dim3(TPB), 0, c.stream, f, p, c.nvol, c.dsc,
                               part == 0 ? 1 : 0, ms);
        } else {
            int g = (int)((n + TPB - 1) / TPB);
            if (g > 8192) g = 8192;
            hipLaunchKernelGGL((k_drift<T>), dim3(g, c.B), dim3(TPB), 0, c.stream, f, p, c.nvol, nf, c.dsc,
                               part == 0 ? 1 : 0, ms);
        }
    }
    prof_end(c, pi);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- u_ij source
// S_ij = sum_l d_i phi_l d_j phi_l with second-order central differences
// d_a f = (f(x + e_a) - f(x - e_a)) / (2 dx) on the periodic lattice; no
// reference arithmetic exists (grav-waves/evolve-gw.f90:21-36 is a stub).
// Component order xx, xy, xz, yy, yz, zz.  In the mixed form the fields hold
// fluctuations about homogeneous means, whose gradients vanish.
__device__ __forceinline__ void gw_pair(int q, int &a, int &b)
{
    a = q < 3 ? 0 : (q < 5 ? 1 : 2);
    b = q < 3 ? q : (q < 5 ? q - 2 : 2);
}

template <typename T>
__device__ __forceinline__ void site_grad(const T *__restrict__ f, int nx, int ny, int nz, int i, int j, int k,
                                          T r, T g[3])
{
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
    const size_t row = (size_t)nx * ((size_t)j + (size_t)ny * k);
    g[0] = (f[row + ip] - f[row + im]) * r;
    g[1] = (f[(size_t)nx * ((size_t)jp + (size_t)ny * k) + i] - f[(size_t)nx * ((size_t)jm + (size_t)ny * k) + i]) * r;
    g[2] = (f[(size_t)nx * ((size_t)j + (size_t)ny * kp) + i] - f[(size_t)nx * ((size_t)j + (size_t)ny * km) + i]) * r;
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_gw_source(const T *__restrict__ fld, int nf, int nx, int ny, int nz, T r,
                                                   int c0, int nc, T *__restrict__ out)
{
    const size_t nvol = (size_t)nx * ny * nz;
    for (size_t s = blockIdx.x * (size_t)TPB + threadIdx.x; s < nvol; s += (size_t)gridDim.x * TPB) {
        const int i = (int)(s % nx);
        const int j = (int)((s / nx) % ny);
        const int k = (int)(s / ((size_t)nx * ny));
        T g[2][3] = {{0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int l = 0; l < 2; ++l)
            if (l < nf) site_grad(fld + (size_t)l * nvol, nx, ny, nz, i, j, k, r, g[l]);
        for (int q = 0; q < nc; ++q) {
            int a, b;
            gw_pair(c0 + q, a, b);
            out[(size_t)q * nvol + s] = g[0][a] * g[0][b] + g[1][a] * g[1][b];
        }
    }
}

template <typename T>
int launch_gw_source(Ctx &c, int c0, int nc, T *dst)
{
    if (c0 < 0 || nc < 1 || c0 + nc > 6) {
        c.last_error = "bad tensor component range";
        return -1;
    }
    size_t blocks = (c.nvol + TPB - 1) / TPB;
    int g = grid_for(blocks > (size_t)MAX_GRID ? MAX_GRID : (int)blocks);
    const int pi = prof_begin(c, LAT_ST_GW_SRC);
    hipLaunchKernelGGL((k_gw_source<T>), dim3(g), dim3(TPB), 0, c.stream, (const T *)c.fld, c.nfld, c.nx, c.ny, c.nz,
                       T(0.5 / c.dx), c0, nc, dst);
    prof_end(c, pi);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- state reductions
// partials: PE = sum V, KE = sum pi^2, M1/M2 = sum phi_l (no Laplacian).
template <typename T>
__global__ __launch_bounds__(TPB) void k_reduce_state(const T *__restrict__ f, const T *__restrict__ p, size_t nvol,
                                                      int nf, ModelDev md, const DevScalars *__restrict__ dsc,
                                                      double *__restrict__ partials, size_t ms)
{
    f += blockIdx.y * ms;
    p += blockIdx.y * ms;
    dsc += blockIdx.y;
    partials += (size_t)blockIdx.y * MAX_GRID * NRED;
    md.p0 = dsc->mp[0];
    md.p1 = dsc->mp[1];
    md.p2 = dsc->mp[2];
    double pb64[2] = {0, 0};
    if constexpr (sizeof(T) == 4) {
        pb64[0] = dsc->mean_phi[0];
        pb64[1] = dsc->mean_phi[1];
    }
    double acc[NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)TPB + threadIdx.x; i < nvol; i += (size_t)gridDim.x * TPB) {
        T phi[2] = {0, 0};
        for (int l = 0; l < nf; ++l) {
            phi[l] = f[(size_t)l * nvol + i];
            const double pp = p[(size_t)l * nvol + i];
            acc[RC_KE] += pp * pp;
            acc[RC_S1 + l] += pp;
        }
        acc[RC_PE] += site_v(md, pb64, phi);
        acc[RC_M1] += double(phi[0]);
        acc[RC_M2] += double(phi[1]);
    }
    block_reduce_store2<NRED>(acc, partials + (size_t)blockIdx.x * NRED);
}

template <typename T>
int launch_reduce_state(Ctx &c)
{
    ModelDev md{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1]};
    const size_t blocks = (c.nvol + TPB - 1) / TPB;
    int g = grid_for(blocks > (size_t)MAX_GRID ? MAX_GRID : (int)blocks);
    const int pi = prof_begin(c, LAT_ST_REDUCE);
    hipLaunchKernelGGL((k_reduce_state<T>), dim3(g, c.B), dim3(TPB), 0, c.stream, (const T *)c.fld,
                       (const T *)c.fldp, c.nvol, c.nfld, md, c.dsc, c.partials, c.ms.f);
    prof_end(c, pi);
    LAT_HIP(&c, hipGetLastError());
    return g;
}

// ---------------------------------------------------------------- FD stencil
// lap_raw = STENCIL(c)/(cc dx^2) with periodic indexing (macros.h:26-33 and
// wrap_fields, hamiltonian_conformal.f90:309-330); the a-dependence of the
// reference's b-coefficients (hamiltonian_conformal.f90:177-179,
// hamiltonian_cosmic.f90:185-187) lives in the kick coefficients hA.
// EPI_KICK: pi += hA lap - hB V', reduce GE, PE, KE'.  EPI_REDUCE: reduce only.
// EPI_STORE: write lap_raw of field 0 into out.
struct StencilC {
    double c0, c1, c2, c3, elap;
};

// STENCIL(c, LAPLACIAN) of f at site (xs[1], ys[1], zs[1]); xs/ys/zs hold the
// periodic neighbour indices (-1, 0, +1) along each axis.  Not yet scaled by
// 1/(cc dx^2).
template <typename T>
__device__ __forceinline__ T stencil_sum(const T *__restrict__ f, const int xs[3], const int ys[3], const int zs[3],
                                         int nx, int ny, T c0, T c1, T c2, T c3, bool has2, bool has3, T *center)
{
#define AT(a, b, cc_) f[(size_t)xs[a] + (size_t)nx * ((size_t)ys[b] + (size_t)ny * zs[cc_])]
    const T f0 = AT(1, 1, 1);
    const T r1 = AT(0, 1, 1) + AT(2, 1, 1) + AT(1, 0, 1) + AT(1, 2, 1) + AT(1, 1, 0) + AT(1, 1, 2);
    T v = c0 * f0 + c1 * r1;
    if (has2) {
        const T r2 = AT(0, 0, 1) + AT(2, 0, 1) + AT(0, 2, 1) + AT(2, 2, 1) + AT(0, 1, 0) + AT(2, 1, 0) +
                     AT(0, 1, 2) + AT(2, 1, 2) + AT(1, 0, 0) + AT(1, 2, 0) + AT(1, 0, 2) + AT(1, 2, 2);
        v = v + c2 * r2;
    }
    if (has3) {
        const T r3 = AT(0, 0, 0) + AT(2, 0, 0) + AT(0, 2, 0) + AT(2, 2, 0) + AT(0, 0, 2) + AT(2, 0, 2) +
                     AT(0, 2, 2) + AT(2, 2, 2);
        v = v + c3 * r3;
    }
#undef AT
    if (center) *center = f0;
    return v;
}

// STENCIL(c, GRAD2) at one site (Discrete/Entropy/macros.h:21-38, summed in
// grad_en, homlat.f90:88-115): sum over the ranks r of c_r times the squared
// differences to the rank-r neighbours, each summed over the fields
// (GRAD2 = sum over fld(:) of (f(nbr) - f(0))^2), in RANK1..3's neighbour
// order; the rank-0 term vanishes.  FP64 differences (in the mixed form the
// homogeneous parts cancel in every difference).
template <typename T>
__device__ __forceinline__ double grad2_site(const T *__restrict__ fld, size_t nvol, int nf, const int xs[3],
                                             const int ys[3], const int zs[3], int nx, int ny, const StencilC &sc)
{
    auto d2 = [&](int a, int b, int cc_) {
        double s = 0.0;
        for (int l = 0; l < nf; ++l) {
            const T *f = fld + (size_t)l * nvol;
            const double d = double(f[(size_t)xs[a] + (size_t)nx * ((size_t)ys[b] + (size_t)ny * zs[cc_])]) -
                             double(f[(size_t)xs[1] + (size_t)nx * ((size_t)ys[1] + (size_t)ny * zs[1])]);
            s += d * d;
        }
        return s;
    };
    const double r1 = d2(0, 1, 1) + d2(2, 1, 1) + d2(1, 0, 1) + d2(1, 2, 1) + d2(1, 1, 0) + d2(1, 1, 2);
    double v = sc.c1 * r1;
    if (sc.c2 != 0.0) {
        const double r2 = d2(0, 0, 1) + d2(2, 0, 1) + d2(0, 2, 1) + d2(2, 2, 1) + d2(0, 1, 0) + d2(2, 1, 0) +
                          d2(0, 1, 2) + d2(2, 1, 2) + d2(1, 0, 0) + d2(1, 2, 0) + d2(1, 0, 2) + d2(1, 2, 2);
        v += sc.c2 * r2;
    }
    if (sc.c3 != 0.0) {
        const double r3 = d2(0, 0, 0) + d2(2, 0, 0) + d2(0, 2, 0) + d2(2, 2, 0) + d2(0, 0, 2) + d2(2, 0, 2) +
                          d2(0, 2, 2) + d2(2, 2, 2);
        v += sc.c3 * r3;
    }
    return v;
}

template <typename T, int EPI>
__global__ __launch_bounds__(TPB) void k_stencil(const T *__restrict__ fld, T *__restrict__ fldp, T *__restrict__ out,
                                                 int nx, int ny, int nz, int nf, StencilC sc, ModelDev md,
                                                 const DevScalars *__restrict__ dsc, double *__restrict__ partials,
                                                 size_t ms)
{
    fld += blockIdx.y * ms;
    fldp += blockIdx.y * ms;
    dsc += blockIdx.y;
    partials += (size_t)blockIdx.y * MAX_GRID * NRED;
    md.p0 = dsc->mp[0];
    md.p1 = dsc->mp[1];
    md.p2 = dsc->mp[2];
    const size_t nvol = (size_t)nx * ny * nz;
    const T c0 = T(sc.c0), c1 = T(sc.c1), c2 = T(sc.c2), c3 = T(sc.c3), el = T(sc.elap);
    const bool has2 = sc.c2 != 0.0, has3 = sc.c3 != 0.0;
    T hA = 0, hB = 0;
    if constexpr (EPI == EPI_KICK) {
        hA = T(dsc->hA);
        hB = T(dsc->hB);
    }
    float pb32[2] = {0, 0};
    double pb64[2] = {0, 0};
    T mu[2] = {0, 0};
    if constexpr (sizeof(T) == 4) {
        for (int l = 0; l < 2; ++l) {
            pb64[l] = dsc->mean_phi[l];
            pb32[l] = float(pb64[l]);
            if constexpr (EPI == EPI_KICK) mu[l] = T(dsc->mu[l]);
        }
    }
    double acc[NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t s = blockIdx.x * (size_t)TPB + threadIdx.x; s < nvol; s += (size_t)gridDim.x * TPB) {
        const int i = (int)(s % nx);
        const int j = (int)((s / nx) % ny);
        const int k = (int)(s / ((size_t)nx * ny));
        const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
        const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
        const int km = k == 0 ? nz - 1 : k - 1, kp = k == nz - 1 ? 0 : k + 1;
        const int xs[3] = {im, i, ip}, ys[3] = {jm, j, jp}, zs[3] = {km, k, kp};
        T lap[2] = {0, 0}, phi[2] = {0, 0};
        for (int l = 0; l < nf; ++l) {
            T f0;
            const T v = stencil_sum(fld + (size_t)l * nvol, xs, ys, zs, nx, ny, c0, c1, c2, c3, has2, has3, &f0);
            lap[l] = v * el;
            phi[l] = f0;
        }
        if constexpr (EPI == EPI_STORE) {
            for (int l = 0; l < nf; ++l) out[(size_t)l * nvol + s] = lap[l];
        } else {
            T dv[2] = {0, 0};
            if constexpr (EPI == EPI_KICK) site_dv(md, pb32, phi, dv);
            acc[RC_PE] += site_v(md, pb64, phi);
            for (int l = 0; l < nf; ++l) {
                T p = fldp[(size_t)l * nvol + s];
                if constexpr (EPI == EPI_KICK) {
                    p = p + (hA * lap[l] - hB * dv[l]);
                    if constexpr (sizeof(T) == 4) p = p - mu[l];
                    fldp[(size_t)l * nvol + s] = p;
                }
                acc[RC_GE] -= double(phi[l]) * double(lap[l]);
                acc[RC_KE] += double(p) * double(p);
                acc[RC_S1 + l] += double(p);
            }
            acc[RC_M1] += double(phi[0]);
            acc[RC_M2] += double(phi[1]);
            if constexpr (EPI == EPI_REDUCE) acc[RC_X] += grad2_site(fld, nvol, nf, xs, ys, zs, nx, ny, sc);
        }
    }
    if constexpr (EPI != EPI_STORE) block_reduce_store2<NRED>(acc, partials + (size_t)blockIdx.x * NRED);
}

// 7-point stencil (c2 = c3 = 0) + kick / reduce, marching along z: a
// TXV x TYV block (64 x 4) owns an (x, y) tile and a chunk of KCV (32)
// planes; each thread keeps its column's values at k-1, k, k+1 in registers,
// so per site it loads the new k+1 value and the four in-plane neighbours
// (which the tile's other lanes load too) instead of seven scattered values.
// The neighbour sum is formed in stencil_sum's order, so the Laplacian is the
// same as there.
constexpr int ST7_TX = 64, ST7_TY = 4, ST7_KC = 32;

template <typename T, int EPI, int TXV = ST7_TX, int TYV = ST7_TY, int KCV = ST7_KC>
__global__ __launch_bounds__(TPB) void k_stencil7(const T *__restrict__ fld, T *__restrict__ fldp, int nx, int ny,
                                                  int nz, int nf, StencilC sc, ModelDev md,
                                                  const DevScalars *__restrict__ dsc, double *__restrict__ partials,
                                                  size_t ms)
{
    static_assert(TXV * TYV == TPB, "one site column per thread");
    fld += blockIdx.y * ms;
    fldp += blockIdx.y * ms;
    dsc += blockIdx.y;
    partials += (size_t)blockIdx.y * MAX_GRID * NRED;
    md.p0 = dsc->mp[0];
    md.p1 = dsc->mp[1];
    md.p2 = dsc->mp[2];
    const size_t nvol = (size_t)nx * ny * nz, plane = (size_t)nx * ny;
    const T c0 = T(sc.c0), c1 = T(sc.c1), el = T(sc.elap);
    T hA = 0, hB = 0;
    if constexpr (EPI == EPI_KICK) {
        hA = T(dsc->hA);
        hB = T(dsc->hB);
    }
    float pb32[2] = {0, 0};
    double pb64[2] = {0, 0};
    T mu[2] = {0, 0};
    if constexpr (sizeof(T) == 4) {
        for (int l = 0; l < 2; ++l) {
            pb64[l] = dsc->mean_phi[l];
            pb32[l] = float(pb64[l]);
            if constexpr (EPI == EPI_KICK) mu[l] = T(dsc->mu[l]);
        }
    }
    double acc[NRED] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int ntx = (nx + TXV - 1) / TXV, nty = (ny + TYV - 1) / TYV, nkc = (nz + KCV - 1) / KCV;
    const int items = ntx * nty * nkc;
    const int tx = threadIdx.x % TXV, ty = threadIdx.x / TXV;
    for (int it = blockIdx.x; it < items; it += gridDim.x) {
        const int xt = it % ntx, rest = it / ntx, yt = rest % nty, kc = rest / nty;
        const int i = xt * TXV + tx, j = yt * TYV + ty;
        if (i >= nx || j >= ny) continue;
        const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
        const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
        const int k0 = kc * KCV, k1 = min(k0 + KCV, nz);
        const size_t col = (size_t)i + (size_t)nx * j;
        const size_t oxm = (size_t)im + (size_t)nx * j, oxp = (size_t)ip + (size_t)nx * j;
        const size_t oym = (size_t)i + (size_t)nx * jm, oyp = (size_t)i + (size_t)nx * jp;
        // Software pipeline: the loads of plane k+1 (its four in-plane
        // neighbours, the column value at k+2, pi at k+1) are issued before
        // plane k's arithmetic, so a column waits on memory once per plane
        // with the next plane's loads in flight.
        T prev[2] = {0, 0}, cur[2] = {0, 0}, nxt[2] = {0, 0}, nb[2][4] = {}, pp[2] = {0, 0};
        auto load_plane = [&](int k, T (&n)[2][4], T (&x)[2], T (&q)[2]) {
            const size_t pk = plane * (size_t)k, pn = plane * (size_t)(k == nz - 1 ? 0 : k + 1);
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                if (l >= nf) break;
                const T *f = fld + (size_t)l * nvol;
                n[l][0] = f[oxm + pk];
                n[l][1] = f[oxp + pk];
                n[l][2] = f[oym + pk];
                n[l][3] = f[oyp + pk];
                x[l] = f[col + pn];
                q[l] = fldp[(size_t)l * nvol + col + pk];
            }
        };
        for (int l = 0; l < nf; ++l) {
            const T *f = fld + (size_t)l * nvol;
            prev[l] = f[col + plane * (size_t)(k0 == 0 ? nz - 1 : k0 - 1)];
            cur[l] = f[col + plane * (size_t)k0];
        }
        load_plane(k0, nb, nxt, pp);
        for (int k = k0; k < k1; ++k) {
            const size_t pk = plane * (size_t)k;
            T nb2[2][4] = {}, nxt2[2] = {0, 0}, pp2[2] = {0, 0};
            if (k + 1 < k1) load_plane(k + 1, nb2, nxt2, pp2);
            T lap[2] = {0, 0}, phi[2] = {0, 0};
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                if (l >= nf) break;
                const T r1 = nb[l][0] + nb[l][1] + nb[l][2] + nb[l][3] + prev[l] + nxt[l];
                lap[l] = (c0 * cur[l] + c1 * r1) * el;
                phi[l] = cur[l];
            }
            const size_t s = col + pk;
            T dv[2] = {0, 0};
            if constexpr (EPI == EPI_KICK) site_dv(md, pb32, phi, dv);
            acc[RC_PE] += site_v(md, pb64, phi);
            for (int l = 0; l < nf; ++l) {
                T p = pp[l];
                if constexpr (EPI == EPI_KICK) {
                    p = p + (hA * lap[l] - hB * dv[l]);
                    if constexpr (sizeof(T) == 4) p = p - mu[l];
                    fldp[(size_t)l * nvol + s] = p;
                }
                acc[RC_GE] -= double(phi[l]) * double(lap[l]);
                acc[RC_KE] += double(p) * double(p);
                acc[RC_S1 + l] += double(p);
            }
            acc[RC_M1] += double(phi[0]);
            acc[RC_M2] += double(phi[1]);
            if constexpr (EPI == EPI_REDUCE) {
                // STENCIL(c, GRAD2) with c2 = c3 = 0, neighbours in RANK1 order
                // (x-, x+, y-, y+, z-, z+), each summed over the fields
                auto d2 = [&](T v0, T v1) {
                    const double d0 = double(v0) - double(cur[0]);
                    double s2 = d0 * d0;
                    if (nf > 1) {
                        const double d1 = double(v1) - double(cur[1]);
                        s2 += d1 * d1;
                    }
                    return s2;
                };
                const double r1 = d2(nb[0][0], nb[1][0]) + d2(nb[0][1], nb[1][1]) + d2(nb[0][2], nb[1][2]) +
                                  d2(nb[0][3], nb[1][3]) + d2(prev[0], prev[1]) + d2(nxt[0], nxt[1]);
                acc[RC_X] += sc.c1 * r1;
            }
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                prev[l] = cur[l];
                cur[l] = nxt[l];
                nxt[l] = nxt2[l];
                pp[l] = pp2[l];
#pragma unroll
                for (int e = 0; e < 4; ++e) nb[l][e] = nb2[l][e];
            }
        }
    }
    block_reduce_store2<NRED>(acc, partials + (size_t)blockIdx.x * NRED);
}

template <typename T>
int launch_stencil(Ctx &c, int epi)
{
    const double *cs = c.cfg.stencil;
    StencilC sc{cs[0], cs[1], cs[2], cs[3], 1.0 / (c.dx * c.dx) / cs[4]};
    ModelDev md{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1]};
    size_t blocks = (c.nvol + TPB - 1) / TPB;
    int g = grid_for(blocks > (size_t)MAX_GRID ? MAX_GRID : (int)blocks);
    T *fld = (T *)c.fld, *fldp = (T *)c.fldp;
    if (cs[2] == 0.0 && cs[3] == 0.0 && (epi == EPI_KICK || epi == EPI_REDUCE)) {
        const int items = ((c.nx + ST7_TX - 1) / ST7_TX) * ((c.ny + ST7_TY - 1) / ST7_TY) *
                          ((c.nz + ST7_KC - 1) / ST7_KC);
        g = grid_for(items > MAX_GRID ? MAX_GRID : items);
        if (epi == EPI_KICK)
            hipLaunchKernelGGL((k_stencil7<T, EPI_KICK>), dim3(g, c.B), dim3(TPB), 0, c.stream, fld, fldp, c.nx,
                               c.ny, c.nz, c.nfld, sc, md, c.dsc, c.partials, c.ms.f);
        else
            hipLaunchKernelGGL((k_stencil7<T, EPI_REDUCE>), dim3(g, c.B), dim3(TPB), 0, c.stream, fld, fldp, c.nx,
                               c.ny, c.nz, c.nfld, sc, md, c.dsc, c.partials, c.ms.f);
        LAT_HIP(&c, hipGetLastError());
        return g;
    }
    if (epi == EPI_KICK)
        hipLaunchKernelGGL((k_stencil<T, EPI_KICK>), dim3(g, c.B), dim3(TPB), 0, c.stream, fld, fldp, (T *)nullptr,
                           c.nx, c.ny, c.nz, c.nfld, sc, md, c.dsc, c.partials, c.ms.f);
    else if (epi == EPI_REDUCE)
        hipLaunchKernelGGL((k_stencil<T, EPI_REDUCE>), dim3(g, c.B), dim3(TPB), 0, c.stream, fld, fldp,
                           (T *)nullptr, c.nx, c.ny, c.nz, c.nfld, sc, md, c.dsc, c.partials, c.ms.f);
    else
        hipLaunchKernelGGL((k_stencil<T, EPI_STORE>), dim3(g), dim3(TPB), 0, c.stream, fld, fldp, (T *)c.io64, c.nx,
                           c.ny, c.nz, 1, sc, md, c.dsc, c.partials, (size_t)0);
    LAT_HIP(&c, hipGetLastError());
    return g;
}

// FD path of the u_ij kick: pi_u += hA (STENCIL(c) u / (cc dx^2) + gwc S_ij),
// S_ij from central differences of the fields, all six components per site.
template <typename T>
__global__ __launch_bounds__(TPB) void k_gw_stencil_kick(const T *__restrict__ u, T *__restrict__ pu,
                                                         const T *__restrict__ fld, int nf, int nx, int ny, int nz,
                                                         StencilC sc, T r, T gwc,
                                                         const DevScalars *__restrict__ dsc)
{
    const size_t nvol = (size_t)nx * ny * nz;
    const T c0 = T(sc.c0), c1 = T(sc.c1), c2 = T(sc.c2), c3 = T(sc.c3), el = T(sc.elap);
    const bool has2 = sc.c2 != 0.0, has3 = sc.c3 != 0.0;
    const T hA = T(dsc->hA);
    for (size_t s = blockIdx.x * (size_t)TPB + threadIdx.x; s < nvol; s += (size_t)gridDim.x * TPB) {
        const int i = (int)(s % nx);
        const int j = (int)((s / nx) % ny);
        const int k = (int)(s / ((size_t)nx * ny));
        const int xs[3] = {i == 0 ? nx - 1 : i - 1, i, i == nx - 1 ? 0 : i + 1};
        const int ys[3] = {j == 0 ? ny - 1 : j - 1, j, j == ny - 1 ? 0 : j + 1};
        const int zs[3] = {k == 0 ? nz - 1 : k - 1, k, k == nz - 1 ? 0 : k + 1};
        T g[2][3] = {{0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int l = 0; l < 2; ++l)
            if (l < nf) site_grad(fld + (size_t)l * nvol, nx, ny, nz, i, j, k, r, g[l]);
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            int a, b;
            gw_pair(q, a, b);
            const T src = g[0][a] * g[0][b] + g[1][a] * g[1][b];
            const T lap = stencil_sum(u + (size_t)q * nvol, xs, ys, zs, nx, ny, c0, c1, c2, c3, has2, has3,
                                      (T *)nullptr) * el;
            T *p = pu + (size_t)q * nvol + s;
            *p = *p + hA * (lap + gwc * src);
        }
    }
}

template <typename T>
int launch_gw_stencil_kick(Ctx &c)
{
    const double *cs = c.cfg.stencil;
    StencilC sc{cs[0], cs[1], cs[2], cs[3], 1.0 / (c.dx * c.dx) / cs[4]};
    size_t blocks = (c.nvol + TPB - 1) / TPB;
    int g = grid_for(blocks > (size_t)MAX_GRID ? MAX_GRID : (int)blocks);
    hipLaunchKernelGGL((k_gw_stencil_kick<T>), dim3(g), dim3(TPB), 0, c.stream, (const T *)c.gwu, (T *)c.gwp,
                       (const T *)c.fld, c.nfld, c.nx, c.ny, c.nz, sc, T(0.5 / c.dx), T(c.gw_coup), c.dsc);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

// FP64 stencil of one host-interface field (io64 -> io64 + nvol).
__global__ __launch_bounds__(TPB) void k_stencil_io(const double *__restrict__ f, double *__restrict__ out, int nx,
                                                    int ny, int nz, StencilC sc)
{
    const size_t nvol = (size_t)nx * ny * nz;
    for (size_t s = blockIdx.x * (size_t)TPB + threadIdx.x; s < nvol; s += (size_t)gridDim.x * TPB) {
        const int i = (int)(s % nx);
        const int j = (int)((s / nx) % ny);
        const int k = (int)(s / ((size_t)nx * ny));
        const int xs[3] = {i == 0 ? nx - 1 : i - 1, i, i == nx - 1 ? 0 : i + 1};
        const int ys[3] = {j == 0 ? ny - 1 : j - 1, j, j == ny - 1 ? 0 : j + 1};
        const int zs[3] = {k == 0 ? nz - 1 : k - 1, k, k == nz - 1 ? 0 : k + 1};
        double r[4] = {0, 0, 0, 0};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int cc = 0; cc < 3; ++cc) {
                    const int rank = (a != 1) + (b != 1) + (cc != 1);
                    r[rank] += f[(size_t)xs[a] + (size_t)nx * ((size_t)ys[b] + (size_t)ny * zs[cc])];
                }
        out[s] = (sc.c0 * r[0] + sc.c1 * r[1] + sc.c2 * r[2] + sc.c3 * r[3]) * sc.elap;
    }
}

int launch_stencil_io(Ctx &c, const double *f, double *out)
{
    const double *cs = c.cfg.stencil;
    StencilC sc{cs[0], cs[1], cs[2], cs[3], 1.0 / (c.dx * c.dx) / cs[4]};
    size_t blocks = (c.nvol + TPB - 1) / TPB;
    int g = blocks > 8192 ? 8192 : (int)blocks;
    hipLaunchKernelGGL(k_stencil_io, dim3(g), dim3(TPB), 0, c.stream, f, out, c.nx, c.ny, c.nz, sc);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------- scalars
struct ScalarEnv {
    int time_var;
    double nvol;
    ModelDev md;
};

// One workgroup: ordered FP64 reduction of `npart` partials (Neumaier
// compensated per channel), then the split operators' scalar parts in order.
constexpr int SCALAR_TPB = 1024; // 8 reduction channels x 128 lanes
__global__ __launch_bounds__(SCALAR_TPB) void k_scalar(DevScalars *__restrict__ sc,
                                                       const double *__restrict__ partials, int npart, ScalarOps ops,
                                                       ScalarEnv env)
{
    // one workgroup per member of a batch
    sc += blockIdx.x;
    partials += (size_t)blockIdx.x * MAX_GRID * NRED;
    env.md.p0 = sc->mp[0];
    env.md.p1 = sc->mp[1];
    env.md.p2 = sc->mp[2];
    constexpr int LANES = SCALAR_TPB / NRED;
    __shared__ double sh[NRED][LANES];
    const int ch = threadIdx.x / LANES, lane = threadIdx.x % LANES;
    if (npart > 0) {
        double s = 0.0, comp = 0.0;
        auto add = [&](double x) {
            const double t = s + x;
            comp += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
            s = t;
        };
        // the loads of 8 partials are issued before their (ordered) sums, so
        // the chain waits on memory once per 8 terms, not per term
        int g = lane;
        for (; g + 7 * LANES < npart; g += 8 * LANES) {
            double xs[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) xs[u] = partials[(size_t)(g + u * LANES) * NRED + ch];
#pragma unroll
            for (int u = 0; u < 8; ++u) add(xs[u]);
        }
        for (; g < npart; g += LANES) add(partials[(size_t)g * NRED + ch]);
        sh[ch][lane] = s + comp;
        __syncthreads();
        if (lane == 0) {
            double tot = 0.0, c2 = 0.0;
            for (int q = 0; q < LANES; ++q) {
                const double x = sh[ch][q];
                const double t = tot + x;
                c2 += (fabs(tot) >= fabs(x)) ? (tot - t) + x : (x - t) + tot;
                tot = t;
            }
            sc->red[ch] = tot + c2;
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    sc->nd = 0;
    const bool mixed = sc->mixed != 0;
    const int nf = env.md.nfld;
    if (mixed && npart > 0) {
        sc->sdphi[0] = sc->red[RC_M1];
        sc->sdphi[1] = sc->red[RC_M2];
    }
    double y = sc->yscl, p = sc->ysclp;
    const bool conf = env.time_var == LAT_CONFORMAL;
    // Mixed form: fold the mean of the pending drifted fluctuations into the
    // FP64 part (the drift sweep subtracts nu).  Must precede any use of
    // mean_phi by a kick, which sees the shifted split.
    bool nu_done = false;
    auto apply_nu = [&]() {
        if (!mixed || nu_done) return;
        nu_done = true;
        for (int l = 0; l < 2; ++l) {
            double nu = 0.0;
            if (sc->nd > 0 && l < nf) {
                nu = (double)(float)(sc->sdphi[l] / env.nvol);
                sc->mean_phi[l] += nu;
                sc->sdphi[l] -= nu * env.nvol;
            }
            sc->nu[l] = nu;
        }
    };
    for (int q = 0; q < ops.n; ++q) {
        const double h = ops.h[q];
        switch (ops.type[q]) {
        case OP_FINISH_KICK: {
            // hamiltonian_conformal.f90:172,203-205; hamiltonian_cosmic.f90:153,209-211
            const double kh = sc->kick_h;
            double PE = sc->red[RC_PE] / env.nvol;
            if (mixed) PE += model_v64(env.md, sc->mean_phi);
            if (conf) {
                const double GE2 = sc->red[RC_GE] / (y * y) / env.nvol;
                p = p - (y * y * y) * (GE2 + 4.0 * PE) * kh;
            } else {
                const double a = pow(y, 2.0 / 3.0);
                const double GE2 = sc->red[RC_GE] / (a * a) / env.nvol;
                p = p - y * ((1.0 / 3.0) * GE2 + 2.0 * PE) * kh;
            }
            break;
        }
        case OP_DRIFT: {
            // hamiltonian_conformal.f90:98-99,116-117 / hamiltonian_cosmic.f90:107-108,125-126
            const double cdr = h / (y * y);
            if (sc->nd < 4) sc->dc[sc->nd++] = cdr;
            double K2 = sc->red[RC_KE] / env.nvol;
            if (mixed) {
                // <pi^2> = sum_l (pb_l^2 + 2 pb_l <dpi_l>) + <dpi^2>
                double cross = 0.0, hom = 0.0;
                for (int l = 0; l < nf; ++l) {
                    hom += sc->mean_pi[l] * sc->mean_pi[l];
                    cross += 2.0 * sc->mean_pi[l] * sc->red[RC_S1 + l];
                }
                K2 = hom + (cross + sc->red[RC_KE]) / env.nvol;
                for (int l = 0; l < nf; ++l) {
                    sc->mean_phi[l] += cdr * sc->mean_pi[l];
                    sc->sdphi[l] += cdr * sc->red[RC_S1 + l];
                }
            }
            p = p + h * K2 / (y * y * y);
            break;
        }
        case OP_GRAV:
            // hamiltonian_conformal.f90:89 / hamiltonian_cosmic.f90:97
            y = conf ? y - h * p / 6.0 : y - h * p * (3.0 / 8.0);
            break;
        case OP_PREP_KICK: {
            sc->kick_h = h;
            if (conf) {
                sc->hA = h * (y * y);
                sc->hB = h * (y * y) * (y * y);
            } else {
                const double a = pow(y, 2.0 / 3.0);
                sc->hA = h * (y * y) / (a * a);
                sc->hB = h * (y * y);
            }
            if (mixed) {
                // homogeneous part of the kick in FP64; the kick sweep applies
                // hA lap - hB dV' - mu to the fluctuations
                apply_nu();
                double dvm[2];
                model_dv64(env.md, sc->mean_phi, dvm);
                for (int l = 0; l < nf; ++l) {
                    const double mu = (double)(float)(sc->red[RC_S1 + l] / env.nvol);
                    sc->mu[l] = mu;
                    sc->mean_pi[l] += -sc->hB * dvm[l] + mu;
                }
            }
            break;
        }
        case OP_CLEAR_DRIFT:
            sc->nd = 0;
            break;
        default:
            break;
        }
    }
    apply_nu();
    sc->yscl = y;
    sc->ysclp = p;
}

int launch_scalar(Ctx &c, const ScalarOps &ops, int npart)
{
    ScalarEnv env{c.cfg.time_var, (double)c.nvol,
                  ModelDev{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1]}};
    const int pi = prof_begin(c, LAT_ST_SCALAR);
    hipLaunchKernelGGL(k_scalar, dim3(c.B), dim3(SCALAR_TPB), 0, c.stream, c.dsc, c.partials, npart, ops, env);
    prof_end(c, pi);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

template int launch_drift<double>(Ctx &);
template int launch_drift<float>(Ctx &);
template int launch_reduce_state<double>(Ctx &);
template int launch_reduce_state<float>(Ctx &);
template int launch_stencil<double>(Ctx &, int);
template int launch_stencil<float>(Ctx &, int);
template int launch_gw_source<double>(Ctx &, int, int, double *);
template int launch_gw_source<float>(Ctx &, int, int, float *);
template int launch_gw_stencil_kick<double>(Ctx &);
template int launch_gw_stencil_kick<float>(Ctx &);

} // namespace lat
