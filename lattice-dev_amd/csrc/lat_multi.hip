// lat_multi.hip -- one launch per sweep for every size class of an ensemble.
//
// An ensemble of uneven lattice sizes (e.g. 32 x 64^3 + 48 x 128^3 +
// 16 x 192^3) advances its classes in lockstep.  Each sweep of the spectral
// force evaluation (x R2C per field, y forward, fused z, y inverse, x C2R +
// kick) is ONE launch whose 1-D grid is the concatenation of the classes'
// grids: class k owns blocks [off_k, off_k + B_k gx_k) (padded to a multiple
// of 8), block b of that range works as block (b - off_k) % gx_k of member
// (b - off_k) / gx_k with exactly the tile loop, radix schedule and
// reduction-partial slot of the class's own launch, so every member's
// arithmetic is bitwise that of the one-class launch.  The FFT bodies are
// specialised on the line length, so the kernel switches on the class's
// size; LDS is dynamic, sized for the largest class of the launch.
#include "sweeps.h"

namespace lat {

#define LAT_FOR_MULTI_SIZES(X) X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

bool multi_supported(int n)
{
    switch (n) {
#define CASE(NN) case NN:
        LAT_FOR_MULTI_SIZES(CASE)
#undef CASE
        return true;
    default:
        return false;
    }
}

// block ranges of the classes in the concatenated grid
struct MultiMap {
    int ncls;
    int off[MULTI_MAXC + 1]; // first block of class k; off[ncls] = total
    int gx[MULTI_MAXC];      // blocks per member
    int nb[MULTI_MAXC];      // members
};

// Class and block coordinates of this block; false for padding blocks.  The
// class index is uniform per block; the selects keep every kernel-argument
// access at a compile-time offset.
__device__ __forceinline__ bool multi_block(const MultiMap &m, int &k, Blk &bk)
{
    const int b = blockIdx.x;
    k = 0;
    int off = m.off[0], gx = m.gx[0], nb = m.nb[0];
#pragma unroll
    for (int j = 1; j < MULTI_MAXC; ++j)
        if (j < m.ncls && b >= m.off[j]) {
            k = j;
            off = m.off[j];
            gx = m.gx[j];
            nb = m.nb[j];
        }
    const int rel = b - off;
    const int member = rel / gx;
    if (member >= nb) return false;
    bk = Blk{rel - member * gx, gx, member};
    return true;
}

template <typename C, int MAXC>
__device__ __forceinline__ C pick(const C (&c)[MAXC], int k)
{
    C r = c[0];
#pragma unroll
    for (int j = 1; j < MAXC; ++j)
        if (j == k) r = c[j];
    return r;
}

// ---------------------------------------------------------------- shapes
// The shape of class (N, tuning variant V) is the one the one-class
// launchers in lat_fft.hip use.
template <typename T, int N, int V>
constexpr int r2c_L()
{
    constexpr int L0 = r2c_lines<T, N>();
    return (V == 1 && tunable<N>() && L0 > 1) ? L0 / 2 : L0;
}
template <int N, int V>
constexpr int c2r_L()
{
    constexpr int L0 = c2r_lines<N>();
    return (V == 1 && tunable<N>() && L0 > 1) ? L0 / 2 : L0;
}
template <typename T, int N, int MODE, int V>
constexpr int col_TC()
{
    constexpr int TC0 = col_tc<T, N>(MODE, 0), TC1 = col_tc<T, N>(MODE, 1);
    return (V == 1 && tunable<N>() && TC1 != TC0) ? TC1 : TC0;
}

// ---------------------------------------------------------------- x R2C
template <typename T>
struct R2CCls {
    T *f;
    const T *pi;
    cpx<T> *Fk;
    const cpx<T> *tw;
    const DevScalars *dsc;
    MemberStride ms;
    int rows, P, sel; // sel = 2 N + V
};
template <typename T>
struct R2CArgs {
    MultiMap m;
    R2CCls<T> c[MULTI_MAXC];
    int field;
};

template <typename T, int N, int V, bool DRIFT>
__device__ __forceinline__ void r2c_case(const R2CCls<T> &c, int field, const Blk &bk, cpx<T> *lds)
{
    constexpr int L = r2c_L<T, N, V>();
    x_r2c_body<T, N, L, DRIFT, r2c_maxradix<T>(), r2c_twl<T>()>(c.f, c.pi, c.Fk, c.tw, c.rows, c.P, c.dsc, field,
                                                               c.ms, bk, lds, lds + r2c_buf<T, N, L>());
}

template <typename T, bool DRIFT>
__global__ __launch_bounds__(TPB) void k_x_r2c_multi(R2CArgs<T> a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lat_dyn_lds[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(lat_dyn_lds);
    int k;
    Blk bk;
    if (!multi_block(a.m, k, bk)) return;
    const R2CCls<T> c = pick(a.c, k);
    switch (c.sel) {
#define CASE(NN)                                                   \
    case 2 * NN: r2c_case<T, NN, 0, DRIFT>(c, a.field, bk, lds); break; \
    case 2 * NN + 1: r2c_case<T, NN, 1, DRIFT>(c, a.field, bk, lds); break;
        LAT_FOR_MULTI_SIZES(CASE)
#undef CASE
    default:
        break;
    }
}

// ---------------------------------------------------------------- y / z columns
template <typename T>
struct ColCls {
    cpx<T> *Fk;
    const cpx<T> *tw;
    size_t s_axis, s_other, s_arr, ms_k;
    double kfac;
    int nxh, n_other, n_arr, ny, sel;
};
template <typename T>
struct ColArgs {
    MultiMap m;
    ColCls<T> c[MULTI_MAXC];
};

constexpr int COL_MULTI_NT = 512;

template <typename T, int N, int V, int MODE>
__device__ __forceinline__ void col_case(const ColCls<T> &c, const Blk &bk, cpx<T> *lds)
{
    // one block size for every class of the launch: the default 512 threads
    // (the bodies loop or idle lanes where a shape has more or fewer
    // first-pass butterflies)
    using C = ColCfg<T, N, MODE == 2 ? 2 : 0>;
    constexpr int TC = col_TC<T, N, MODE, V>();
    constexpr int B = col_buf<T, N, TC>();
    col_body<T, N, TC, MODE, C::MR, C::PF, 1, COL_MULTI_NT, true, C::FUSE, C::PP>(
        c.Fk, c.tw, c.nxh, c.n_other, c.s_axis, c.s_other, c.s_arr, c.n_arr, c.ny, c.kfac, c.ms_k, bk, lds, lds + B,
        lds + (C::PP ? 2 * B : B));
}

template <typename T, int MODE>
__global__ __launch_bounds__(COL_MULTI_NT) void k_col_multi(ColArgs<T> a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lat_dyn_lds[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(lat_dyn_lds);
    int k;
    Blk bk;
    if (!multi_block(a.m, k, bk)) return;
    const ColCls<T> c = pick(a.c, k);
    switch (c.sel) {
#define CASE(NN)                                            \
    case 2 * NN: col_case<T, NN, 0, MODE>(c, bk, lds); break; \
    case 2 * NN + 1: col_case<T, NN, 1, MODE>(c, bk, lds); break;
        LAT_FOR_MULTI_SIZES(CASE)
#undef CASE
    default:
        break;
    }
}

// ---------------------------------------------------------------- x C2R + kick
template <typename T>
struct C2RCls {
    const cpx<T> *Fk;
    const cpx<T> *tw;
    const T *fld;
    T *fldp;
    const DevScalars *dsc;
    double *partials;
    ModelDev md;
    MemberStride ms;
    size_t nk, nvol;
    int rows, P, sel;
};
template <typename T>
struct C2RArgs {
    MultiMap m;
    C2RCls<T> c[MULTI_MAXC];
    int nf;
};

template <typename T, int N, int V, int EPI>
__device__ __forceinline__ void c2r_case(const C2RCls<T> &c, int nf, const Blk &bk, cpx<T> *lds)
{
    constexpr int L = c2r_L<N, V>();
    x_c2r_body<T, N, L, EPI, 4, true, true, 1>(c.Fk, c.tw, c.rows, c.P, c.nk, c.nvol, nf, nullptr, c.fld, c.fldp,
                                               c.dsc, c.md, c.partials, c.ms, bk, lds, lds + c2r_buf<T, N, L>());
}

template <typename T, int EPI>
__global__ __launch_bounds__(TPB) void k_x_c2r_multi(C2RArgs<T> a)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lat_dyn_lds[];
    cpx<T> *lds = reinterpret_cast<cpx<T> *>(lat_dyn_lds);
    int k;
    Blk bk;
    if (!multi_block(a.m, k, bk)) return;
    const C2RCls<T> c = pick(a.c, k);
    switch (c.sel) {
#define CASE(NN)                                               \
    case 2 * NN: c2r_case<T, NN, 0, EPI>(c, a.nf, bk, lds); break; \
    case 2 * NN + 1: c2r_case<T, NN, 1, EPI>(c, a.nf, bk, lds); break;
        LAT_FOR_MULTI_SIZES(CASE)
#undef CASE
    default:
        break;
    }
}

// ---------------------------------------------------------------- host side
// Per-class launch geometry: blocks per member and LDS bytes of the shape.
struct Geo {
    int gx;
    size_t lds;
};

template <typename T, int N, int V>
static Geo r2c_geo(const Ctx &c)
{
    constexpr int L = r2c_L<T, N, V>();
    return {grid_for((c.ny * c.nz + L - 1) / L, R2C_GRID), sizeof(cpx<T>) * r2c_lds<T, N, L, r2c_twl<T>()>()};
}
template <typename T, int N, int V, int MODE>
static Geo col_geo(const Ctx &c, int n_other, int n_arr)
{
    using C = ColCfg<T, N, MODE == 2 ? 2 : 0>;
    constexpr int TC = col_TC<T, N, MODE, V>();
    return {grid_for(((c.nxh + TC - 1) / TC) * n_other * n_arr, C::GRID),
            sizeof(cpx<T>) * col_lds<T, N, TC, C::PP, true>()};
}
template <typename T, int N, int V>
static Geo c2r_geo(const Ctx &c)
{
    constexpr int L = c2r_L<N, V>();
    return {grid_for((c.ny * c.nz + L - 1) / L, c2r_grid_n<T, N>()), sizeof(cpx<T>) * c2r_lds<T, N, L, true>()};
}

// Fills the block map; returns the total number of blocks.
static int map_classes(MultiMap &m, const std::vector<Geo> &geo, Ctx *const *cs, int ncls)
{
    m.ncls = ncls;
    int total = 0;
    for (int k = 0; k < ncls; ++k) {
        m.off[k] = total;
        m.gx[k] = geo[k].gx;
        m.nb[k] = cs[k]->B;
        total += (geo[k].gx * cs[k]->B + 7) & ~7;
    }
    for (int k = ncls; k <= MULTI_MAXC; ++k) m.off[k] = total;
    for (int k = ncls; k < MULTI_MAXC; ++k) {
        m.gx[k] = 1;
        m.nb[k] = 0;
    }
    return total;
}

static size_t max_lds(const std::vector<Geo> &geo)
{
    size_t b = 0;
    for (const Geo &g : geo) b = g.lds > b ? g.lds : b;
    return b;
}

#define LAT_MULTI_SIZE_SWITCH(NVAL, VVAL, CALL)                                   \
    switch (NVAL) {                                                               \
    case 16: CALL(16, VVAL); break;                                               \
    case 32: CALL(32, VVAL); break;                                               \
    case 48: CALL(48, VVAL); break;                                               \
    case 64: CALL(64, VVAL); break;                                               \
    case 96: CALL(96, VVAL); break;                                               \
    case 128: CALL(128, VVAL); break;                                             \
    case 192: CALL(192, VVAL); break;                                             \
    case 256: CALL(256, VVAL); break;                                             \
    default: return -2;                                                           \
    }

static int check_classes(Ctx *const *cs, int ncls)
{
    if (ncls < 1 || ncls > MULTI_MAXC) return -1;
    for (int k = 0; k < ncls; ++k)
        if (!multi_supported(cs[k]->nx) || cs[k]->ny != cs[k]->nx || cs[k]->nz != cs[k]->nx) return -2;
    return 0;
}

// Dynamic LDS above the default limit has to be allowed per kernel.
template <typename K>
static int allow_lds(Ctx &c, K kernel, size_t bytes)
{
    if (bytes <= 65536) return 0;
    LAT_HIP(&c, hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return 0;
}

template <typename T>
int launch_x_r2c_multi(Ctx *const *cs, int ncls, int field, int drift)
{
    int rc = check_classes(cs, ncls);
    if (rc) return rc;
    R2CArgs<T> a{};
    std::vector<Geo> geo(ncls);
    for (int k = 0; k < ncls; ++k) {
        Ctx &c = *cs[k];
        const int v = c.tune.r2c == 1 ? 1 : 0;
#define GEO(NN, VV) geo[k] = (VV) ? r2c_geo<T, NN, 1>(c) : r2c_geo<T, NN, 0>(c)
        LAT_MULTI_SIZE_SWITCH(c.nx, v, GEO)
#undef GEO
        const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
        a.c[k] = R2CCls<T>{(T *)c.fld + (size_t)field * c.nvol, (const T *)c.fldp + (size_t)field * c.nvol,
                           (cpx<T> *)c.Fk + (size_t)field * c.nk, twx, c.dsc, c.ms, c.ny * c.nz, c.P,
                           2 * c.nx + v};
    }
    a.field = field;
    const int blocks = map_classes(a.m, geo, cs, ncls);
    const size_t lds = max_lds(geo);
    Ctx &c0 = *cs[0];
    if (drift) {
        if ((rc = allow_lds(c0, k_x_r2c_multi<T, true>, lds))) return rc;
        hipLaunchKernelGGL((k_x_r2c_multi<T, true>), dim3(blocks), dim3(TPB), lds, c0.stream, a);
    } else {
        if ((rc = allow_lds(c0, k_x_r2c_multi<T, false>, lds))) return rc;
        hipLaunchKernelGGL((k_x_r2c_multi<T, false>), dim3(blocks), dim3(TPB), lds, c0.stream, a);
    }
    LAT_HIP(&c0, hipGetLastError());
    return 0;
}

template <typename T, int MODE>
static int col_multi_mode(Ctx *const *cs, int ncls, int axis, int n_arrays)
{
    ColArgs<T> a{};
    std::vector<Geo> geo(ncls);
    for (int k = 0; k < ncls; ++k) {
        Ctx &c = *cs[k];
        const int n_other = axis == 1 ? c.nz : c.ny;
        const int v = (axis == 1 ? c.tune.coly : c.tune.colz) == 1 ? 1 : 0;
#define GEO(NN, VV) geo[k] = (VV) ? col_geo<T, NN, 1, MODE>(c, n_other, n_arrays) : col_geo<T, NN, 0, MODE>(c, n_other, n_arrays)
        LAT_MULTI_SIZE_SWITCH(c.nx, v, GEO)
#undef GEO
        const void *tw = axis == 1 ? (sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32)
                                   : (sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
        a.c[k] = ColCls<T>{(cpx<T> *)c.Fk,
                           (const cpx<T> *)tw,
                           axis == 1 ? (size_t)c.P : (size_t)c.P * c.ny,
                           axis == 1 ? (size_t)c.P * c.ny : (size_t)c.P,
                           c.nk,
                           c.ms.k,
                           c.dk * c.dk / (double)c.nvol,
                           c.nxh,
                           n_other,
                           n_arrays,
                           c.ny,
                           2 * c.nx + v};
    }
    const int blocks = map_classes(a.m, geo, cs, ncls);
    const size_t lds = max_lds(geo);
    Ctx &c0 = *cs[0];
    int rc = allow_lds(c0, k_col_multi<T, MODE>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((k_col_multi<T, MODE>), dim3(blocks), dim3(COL_MULTI_NT), lds, c0.stream, a);
    LAT_HIP(&c0, hipGetLastError());
    return 0;
}

template <typename T>
int launch_col_multi(Ctx *const *cs, int ncls, int axis, int mode, int n_arrays)
{
    int rc = check_classes(cs, ncls);
    if (rc) return rc;
    if (mode == 0) return col_multi_mode<T, 0>(cs, ncls, axis, n_arrays);
    if (mode == 1) return col_multi_mode<T, 1>(cs, ncls, axis, n_arrays);
    if (axis != 2) return -1;
    return col_multi_mode<T, 2>(cs, ncls, axis, n_arrays);
}

template <typename T>
int launch_x_c2r_multi(Ctx *const *cs, int ncls, int nf, int epi, int *npart)
{
    int rc = check_classes(cs, ncls);
    if (rc) return rc;
    if (nf < 1 || nf > 2 || (epi != EPI_KICK && epi != EPI_REDUCE)) return -1;
    C2RArgs<T> a{};
    std::vector<Geo> geo(ncls);
    for (int k = 0; k < ncls; ++k) {
        Ctx &c = *cs[k];
        const int v = c.tune.c2r == 1 ? 1 : 0;
#define GEO(NN, VV) geo[k] = (VV) ? c2r_geo<T, NN, 1>(c) : c2r_geo<T, NN, 0>(c)
        LAT_MULTI_SIZE_SWITCH(c.nx, v, GEO)
#undef GEO
        const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
        ModelDev md{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1], c.gw_coup};
        a.c[k] = C2RCls<T>{(const cpx<T> *)c.Fk, twx, (const T *)c.fld, (T *)c.fldp, c.dsc, c.partials, md, c.ms,
                           c.nk, c.nvol, c.ny * c.nz, c.P, 2 * c.nx + v};
        npart[k] = geo[k].gx;
    }
    a.nf = nf;
    const int blocks = map_classes(a.m, geo, cs, ncls);
    const size_t lds = max_lds(geo);
    Ctx &c0 = *cs[0];
    if (epi == EPI_KICK) {
        if ((rc = allow_lds(c0, k_x_c2r_multi<T, EPI_KICK>, lds))) return rc;
        hipLaunchKernelGGL((k_x_c2r_multi<T, EPI_KICK>), dim3(blocks), dim3(TPB), lds, c0.stream, a);
    } else {
        if ((rc = allow_lds(c0, k_x_c2r_multi<T, EPI_REDUCE>, lds))) return rc;
        hipLaunchKernelGGL((k_x_c2r_multi<T, EPI_REDUCE>), dim3(blocks), dim3(TPB), lds, c0.stream, a);
    }
    LAT_HIP(&c0, hipGetLastError());
    return 0;
}

template int launch_x_r2c_multi<double>(Ctx *const *, int, int, int);
template int launch_x_r2c_multi<float>(Ctx *const *, int, int, int);
template int launch_col_multi<double>(Ctx *const *, int, int, int, int);
template int launch_col_multi<float>(Ctx *const *, int, int, int, int);
template int launch_x_c2r_multi<double>(Ctx *const *, int, int, int, int *);
template int launch_x_c2r_multi<float>(Ctx *const *, int, int, int, int *);

} // namespace lat
