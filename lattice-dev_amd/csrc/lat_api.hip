// lat_api.hip -- C ABI (include/lattice_c.h): context, state transfer, the
// drop-in operators, and the device-resident symplectic executor.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lat_internal.h"

namespace lat {

int hip_fail(Ctx *c, hipError_t e, const char *where)
{
    if (c) {
        c->last_error = std::string(where) + ": " + hipGetErrorString(e);
    }
    return -3;
}

size_t elem_size(const Ctx &c) { return c.prec == LAT_F64 ? sizeof(double) : sizeof(float); }

int launch_stencil_io(Ctx &c, const double *f, double *out);

static bool size_ok(int n)
{
    static const int ok[] = {8, 16, 32, 64, 128, 256, 512, 1024, 48, 96, 192};
    for (int v : ok)
        if (v == n) return true;
    return false;
}

template <typename T>
static cpx<T> *make_twiddles(Ctx &c, int n, int *rc)
{
    std::vector<cpx<T>> h(n);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < n; ++m) {
        long double a = -twopi * (long double)m / (long double)n;
        h[m].x = (T)cosl(a);
        h[m].y = (T)sinl(a);
    }
    cpx<T> *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(cpx<T>) * n);
    if (e != hipSuccess) { *rc = hip_fail(&c, e, "hipMalloc twiddles"); return nullptr; }
    e = hipMemcpy(d, h.data(), sizeof(cpx<T>) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) { *rc = hip_fail(&c, e, "hipMemcpy twiddles"); return nullptr; }
    return d;
}

static int ensure(Ctx &c, void **p, size_t bytes, const char *what)
{
    if (*p) return 0;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return hip_fail(&c, e, what);
    if (c.view) c.owned.push_back(*p);
    return 0;
}

static int sync(Ctx &c)
{
    LAT_HIP(&c, hipStreamSynchronize(c.stream));
    return 0;
}

static int pull_scalars(Ctx &c)
{
    LAT_HIP(&c, hipMemcpyAsync(&c.host_sc, c.dsc, sizeof(DevScalars), hipMemcpyDeviceToHost, c.stream));
    return sync(c);
}

static int push_scalars(Ctx &c)
{
    LAT_HIP(&c, hipMemcpyAsync(c.dsc, &c.host_sc, sizeof(DevScalars), hipMemcpyHostToDevice, c.stream));
    return sync(c);
}

// Validity of red[RC_KE] (sum pi^2 of the current state), per member: a view
// reads and writes its batch's flag for that member.
static bool get_ke_valid(const Ctx &c)
{
    if (c.view && c.parent) return c.parent->mvalid[c.member] != 0;
    if (c.B > 1) {
        for (char v : c.mvalid)
            if (!v) return false;
        return true;
    }
    return c.ke_valid;
}

static void set_ke_valid(Ctx &c, bool v)
{
    c.ke_valid = v;
    if (c.view && c.parent) c.parent->mvalid[c.member] = v;
    if (c.B > 1) std::fill(c.mvalid.begin(), c.mvalid.end(), (char)v);
}

// ---------------------------------------------------------------- profiling
int prof_begin(Ctx &c, int stage)
{
    if (!c.prof.on) return -1;
    const int i = (int)c.prof.used.size() * 2;
    while ((int)c.prof.pool.size() < i + 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        c.prof.pool.push_back(e);
    }
    if (hipEventRecord(c.prof.pool[i], c.stream) != hipSuccess) return -1;
    c.prof.used.push_back({stage, i});
    return (int)c.prof.used.size() - 1;
}

void prof_end(Ctx &c, int idx)
{
    if (idx < 0) return;
    (void)hipEventRecord(c.prof.pool[c.prof.used[idx].second + 1], c.stream);
}

// Folds recorded event pairs into the per-stage accumulators.
static int prof_collect(Ctx &c)
{
    if (c.prof.used.empty()) return 0;
    int rc = sync(c);
    if (rc) return rc;
    for (const auto &u : c.prof.used) {
        float ms = 0.f;
        LAT_HIP(&c, hipEventElapsedTime(&ms, c.prof.pool[u.second], c.prof.pool[u.second + 1]));
        c.prof.ms[u.first] += ms;
        c.prof.cnt[u.first] += 1;
    }
    c.prof.used.clear();
    return 0;
}

struct StageTimer {
    Ctx &c;
    int idx;
    StageTimer(Ctx &cc, int stage) : c(cc), idx(prof_begin(cc, stage)) {}
    ~StageTimer() { prof_end(c, idx); }
};

// ---------------------------------------------------------------- Laplacian pipeline
// The blocked k-space order (sweeps.h) in the field force pipeline: cubic
// 256 and 1024 when switched on (lat_set_kspace_order).
static int kblk_active(const Ctx &c)
{
    return c.kblk && c.nx == c.ny && c.nx == c.nz && (c.nx == 256 || c.nx == 1024) ? 1 : 0;
}

// Forward R2C (x, y, z) of the nfld state fields into Fk; drift fused if asked.
// layout 1: the blocked order (zmode 2 only: the transforms stay internal).
template <typename T>
static int forward_state(Ctx &c, bool drift, int zmode, int layout = 0)
{
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
    T *fld = (T *)c.fld;
    const T *fldp = (const T *)c.fldp;
    cpx<T> *Fk = (cpx<T> *)c.Fk;
    int rc;
    for (int l = 0; l < c.nfld; ++l) {
        StageTimer st(c, LAT_ST_X_R2C);
        rc = launch_x_r2c<T>(c, twx, fld + (size_t)l * c.nvol, fldp + (size_t)l * c.nvol, Fk + (size_t)l * c.nk,
                             drift ? 1 : 0, l, layout);
        if (rc) return rc;
    }
    {
        StageTimer st(c, LAT_ST_Y_FWD);
        if ((rc = launch_col<T>(c, twy, Fk, 1, 0, c.nfld, layout))) return rc;
    }
    StageTimer st(c, LAT_ST_Z_LAP);
    if ((rc = launch_col<T>(c, twz, Fk, 2, zmode, c.nfld, layout))) return rc;
    return 0;
}

// Laplacian of the state into the C2R epilogue: EPI_KICK or EPI_REDUCE.
// Returns the number of reduction partials (>0) or an error (<0).
// Kick of the metric perturbations after the field kick of the same step (the
// fields hold the drifted phi the source needs).  Spectral: three pairs of
// components, each S_ij pair -> x R2C (pending drift of u fused) -> y -> z lap
// -> y -> x C2R with pi_u += hA (lap + gw_coup S).  FD: one stencil sweep.
template <typename T>
static int gw_kick(Ctx &c, bool drift)
{
    if (c.cfg.laplacian == LAT_LAP_STENCIL) {
        StageTimer st(c, LAT_ST_STENCIL);
        return launch_gw_stencil_kick<T>(c); // drift of u applied by launch_drift
    }
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
    T *u = (T *)c.gwu, *p = (T *)c.gwp, *src = (T *)c.gwsrc;
    cpx<T> *Fk = (cpx<T> *)c.Fk;
    int rc;
    for (int c0 = 0; c0 < 6; c0 += 2) {
        if ((rc = launch_gw_source<T>(c, c0, 2, src))) return rc;
        for (int q = 0; q < 2; ++q) {
            StageTimer st(c, LAT_ST_X_R2C);
            const size_t off = (size_t)(c0 + q) * c.nvol;
            if ((rc = launch_x_r2c<T>(c, twx, u + off, p + off, Fk + (size_t)q * c.nk, drift ? 1 : 0, -1))) return rc;
        }
        {
            StageTimer st(c, LAT_ST_Y_FWD);
            if ((rc = launch_col<T>(c, twy, Fk, 1, 0, 2))) return rc;
        }
        {
            StageTimer st(c, LAT_ST_Z_LAP);
            if ((rc = launch_col<T>(c, twz, Fk, 2, 2, 2))) return rc;
        }
        {
            StageTimer st(c, LAT_ST_Y_INV);
            if ((rc = launch_col<T>(c, twy, Fk, 1, 1, 2))) return rc;
        }
        StageTimer st(c, LAT_ST_X_C2R);
        rc = launch_x_c2r<T>(c, twx, Fk, 2, EPI_GWKICK, nullptr, src, p + (size_t)c0 * c.nvol);
        if (rc < 0) return rc;
    }
    return 0;
}

template <typename T>
static int state_force(Ctx &c, bool drift, int epi)
{
    const bool gw = c.gw && epi == EPI_KICK;
    if constexpr (sizeof(T) == 4) {
        if (c.sep && epi == EPI_KICK && !c.gw && c.cfg.laplacian == LAT_LAP_SPECTRAL && sep_supported(c)) {
            int rc = ensure(c, &c.lap, sizeof(float) * c.ms.f * c.B, "hipMalloc lap");
            if (rc) return rc;
            int g = 0;
            rc = launch_sep_kick<float>(c, drift ? 1 : 0, &g);
            return rc ? rc : g;
        }
    }
    if (c.cfg.laplacian == LAT_LAP_STENCIL) {
        if (drift) {
            int rc = launch_drift<T>(c);
            if (rc) return rc;
        }
        int g;
        {
            StageTimer st(c, LAT_ST_STENCIL);
            g = launch_stencil<T>(c, epi);
        }
        if (g < 0) return g;
        if (gw) {
            int rc = gw_kick<T>(c, drift);
            if (rc) return rc;
        }
        return g;
    }
    const int layout = kblk_active(c);
    int rc = forward_state<T>(c, drift, 2, layout);
    if (rc) return rc;
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    {
        StageTimer st(c, LAT_ST_Y_INV);
        if ((rc = launch_col<T>(c, twy, (cpx<T> *)c.Fk, 1, 1, c.nfld, layout))) return rc;
    }
    int g;
    {
        StageTimer st(c, LAT_ST_X_C2R);
        g = launch_x_c2r<T>(c, twx, (cpx<T> *)c.Fk, c.nfld, epi, nullptr, (const T *)c.fld, (T *)c.fldp, layout);
    }
    if (g < 0) return g;
    if (gw && (rc = gw_kick<T>(c, drift))) return rc;
    return g;
}

static int force(Ctx &c, bool drift, int epi)
{
    return c.prec == LAT_F64 ? state_force<double>(c, drift, epi) : state_force<float>(c, drift, epi);
}

// ---------------------------------------------------------------- op executor
struct Op {
    int type; // OP_DRIFT, OP_GRAV, or 99 = kick
    double h;
};
constexpr int OP_KICK = 99;

static void map_split(const Ctx &c, std::vector<Op> &ops, double h, int hamind)
{
    // hamiltonian_conformal.f90:70-80 / hamiltonian_cosmic.f90:78-88
    const bool conf = c.cfg.time_var == LAT_CONFORMAL;
    if (hamind == 3) ops.push_back({OP_KICK, h});
    else if ((hamind == 1) == conf) ops.push_back({OP_DRIFT, h});
    else ops.push_back({OP_GRAV, h});
}

static void o2step(const Ctx &c, std::vector<Op> &ops, double dt, double w1, double w2)
{
    // symp_o2step, hamiltonian_conformal.f90:50-64 (n_Hamiltonian_terms = 3)
    map_split(c, ops, w1 * dt / 2.0, 2);
    map_split(c, ops, w1 * dt, 3);
    map_split(c, ops, w1 * dt / 2.0, 2);
    map_split(c, ops, (w1 + w2) * dt / 2.0, 1);
}

static int build_symp(const Ctx &c, std::vector<Op> &ops, int order, double dt, int nsteps)
{
    if (nsteps <= 0) return 0;
    if (order == 2) { // mpi_dev/integrator.f90:11-22
        map_split(c, ops, dt / 2.0, 1);
        for (int j = 1; j <= nsteps - 1; ++j) o2step(c, ops, dt, 1.0, 1.0);
        o2step(c, ops, dt, 1.0, 0.0);
    } else if (order == 4) { // integrator.f90:24-44
        const double w1 = 1.0 / (2.0 - std::pow(2.0, 1.0 / 3.0));
        const double w0 = 1.0 - 2.0 * w1;
        map_split(c, ops, w1 * dt / 2.0, 1);
        for (int i = 1; i <= nsteps; ++i) {
            o2step(c, ops, dt, w1, w0);
            o2step(c, ops, dt, w0, w1);
            o2step(c, ops, dt, w1, i == nsteps ? 0.0 : w1);
        }
    } else if (order == 6) { // integrator.f90:46-71 / evolve_spectral.f90:169-194
        const double w1 = c.cfg.symp6_tree ? -1.17767998417887 : -1.177679984177887;
        const double w2 = 0.235573213359357, w3 = 0.784513610477560;
        const double w0 = 1.0 - 2.0 * (w1 + w2 + w3);
        map_split(c, ops, w3 * dt / 2.0, 1);
        for (int j = 1; j <= nsteps; ++j) {
            o2step(c, ops, dt, w3, w2);
            o2step(c, ops, dt, w2, w1);
            o2step(c, ops, dt, w1, w0);
            o2step(c, ops, dt, w0, w1);
            o2step(c, ops, dt, w1, w2);
            o2step(c, ops, dt, w2, w3);
            o2step(c, ops, dt, w3, j == nsteps ? 0.0 : w3);
        }
    } else if (order == 8) { // integrator.f90:73-110
        const double w1 = 0.74167036435061295344822780, w2 = -0.40910082580003159399730010,
                     w3 = 0.19075471029623837995387626, w4 = -0.57386247111608226665638733,
                     w5 = 0.29906418130365592384446354, w6 = 0.33462491824529818378495798,
                     w7 = 0.31529309239676659663205666;
        const double w0 = 1.0 - 2.0 * (w1 + w2 + w3 + w4 + w5 + w6 + w7);
        const double seq[16] = {w7, w6, w5, w4, w3, w2, w1, w0, w1, w2, w3, w4, w5, w6, w7, 0.0};
        map_split(c, ops, w7 * dt / 2.0, 1);
        for (int j = 1; j <= nsteps; ++j) {
            for (int q = 0; q < 14; ++q) o2step(c, ops, dt, seq[q], seq[q + 1]);
            o2step(c, ops, dt, w7, j == nsteps ? 0.0 : w7);
        }
    } else {
        return -1;
    }
    return 0;
}

// Runs a linear sequence of split operators on the device.  Between two
// kicks, the scalar kernel applies (in order) the pending kick's p-update, the
// drifts' and grav steps' scalar parts, and computes the next kick's
// coefficients; the drifts' field updates are fused into the next x-R2C sweep
// (spectral) or run as one drift sweep (stencil, end of sequence).
static int run_ops(Ctx &c, const std::vector<Op> &ops)
{
    int npart = 0;
    if (!get_ke_valid(c)) {
        npart = c.prec == LAT_F64 ? launch_reduce_state<double>(c) : launch_reduce_state<float>(c);
        if (npart < 0) return npart;
    }
    ScalarOps so{};
    int ndrift = 0;
    bool kick_pending = false;
    auto flush_drifts = [&]() -> int {
        // more drifts than the fused sweep applies: settle them now
        int rc = launch_scalar(c, so, npart);
        if (rc) return rc;
        npart = 0;
        so = ScalarOps{};
        kick_pending = false;
        rc = c.prec == LAT_F64 ? launch_drift<double>(c) : launch_drift<float>(c);
        ndrift = 0;
        return rc;
    };
    auto begin_seg = [&]() {
        so = ScalarOps{};
        if (kick_pending) {
            so.type[so.n] = OP_FINISH_KICK;
            so.h[so.n++] = 0.0;
        }
    };
    begin_seg();
    for (const Op &op : ops) {
        if (op.type == OP_KICK) {
            so.type[so.n] = OP_PREP_KICK;
            so.h[so.n++] = op.h;
            int rc = launch_scalar(c, so, npart);
            if (rc) return rc;
            int g = force(c, ndrift > 0, EPI_KICK);
            if (g < 0) return g;
            npart = g;
            ndrift = 0;
            kick_pending = true;
            begin_seg();
        } else {
            if (op.type == OP_DRIFT && ndrift == 2) {
                int rc = flush_drifts();
                if (rc) return rc;
                begin_seg();
            }
            if (so.n >= MAX_OPS - 1) {
                c.last_error = "op segment overflow";
                return -1;
            }
            so.type[so.n] = op.type;
            so.h[so.n++] = op.h;
            if (op.type == OP_DRIFT) ++ndrift;
        }
    }
    if (so.n > 0 || npart > 0) {
        int rc = launch_scalar(c, so, npart);
        if (rc) return rc;
    }
    if (ndrift > 0) {
        int rc = c.prec == LAT_F64 ? launch_drift<double>(c) : launch_drift<float>(c);
        if (rc) return rc;
    }
    set_ke_valid(c, true);
    return 0;
}

// Reduce GE (with the Laplacian), PE, KE and field means of the current state.
static int state_sums(Ctx &c)
{
    int g = force(c, false, EPI_REDUCE);
    if (g < 0) return g;
    ScalarOps none{};
    int rc = launch_scalar(c, none, g);
    if (rc) return rc;
    set_ke_valid(c, true);
    return pull_scalars(c);
}

static double scale_factor(const Ctx &c, double y)
{
    return c.cfg.time_var == LAT_CONFORMAL ? y : std::pow(y, 2.0 / 3.0);
}

static double host_v(const Ctx &c, const double *f)
{
    const double *mp = c.cfg.model_params;
    if (c.cfg.model == LAT_MODEL_PHI4) return 0.25 * mp[0] * (f[0] * f[0]) * (f[0] * f[0]);
    return 0.5 * mp[0] * (f[0] * f[0]) + 0.5 * mp[1] * (f[0] * f[0]) * (f[1] * f[1]) + 0.5 * mp[2] * (f[1] * f[1]);
}

// Lattice sums of the physical fields after state_sums(): sum pi^2, sum V,
// sum -phi lap_raw, and the field means.  The mixed form reassembles them
// from the FP64 homogeneous parts and the FP64-accumulated fluctuation sums.
struct PhysSums {
    double pi2, v, ge, mean[2], mean_pi[2];
};

static PhysSums phys_sums(const Ctx &c)
{
    const DevScalars &s = c.host_sc;
    const double nv = (double)c.nvol;
    PhysSums r{};
    r.ge = s.red[RC_GE];
    if (!s.mixed) {
        r.pi2 = s.red[RC_KE];
        r.v = s.red[RC_PE];
        for (int l = 0; l < 2; ++l) {
            r.mean[l] = l < c.nfld ? s.red[RC_M1 + l] / nv : 0.0;
            r.mean_pi[l] = l < c.nfld ? s.red[RC_S1 + l] / nv : 0.0;
        }
        return r;
    }
    r.pi2 = s.red[RC_KE];
    for (int l = 0; l < c.nfld; ++l) {
        r.pi2 += nv * s.mean_pi[l] * s.mean_pi[l] + 2.0 * s.mean_pi[l] * s.red[RC_S1 + l];
        r.mean[l] = s.mean_phi[l] + s.red[RC_M1 + l] / nv;
        r.mean_pi[l] = s.mean_pi[l] + s.red[RC_S1 + l] / nv;
    }
    r.v = nv * host_v(c, s.mean_phi) + s.red[RC_PE];
    return r;
}

// ---------------------------------------------------------------- host transforms (FP64)
int ensure_io(Ctx &c)
{
    int rc = ensure(c, (void **)&c.io64, sizeof(double) * c.nvol * 2, "hipMalloc io64");
    if (rc) return rc;
    if (c.prec == LAT_F64) {
        c.Fk64 = c.Fk;
        return 0;
    }
    return ensure(c, &c.Fk64, sizeof(cpx<double>) * c.nk * 2, "hipMalloc Fk64");
}

static int host_r2c(Ctx &c, const double *f_host)
{
    int rc = ensure_io(c);
    if (rc) return rc;
    LAT_HIP(&c, hipMemcpyAsync(c.io64, f_host, sizeof(double) * c.nvol, hipMemcpyHostToDevice, c.stream));
    cpx<double> *F = (cpx<double> *)c.Fk64;
    if ((rc = launch_x_r2c<double>(c, c.twx64, c.io64, nullptr, F, 0))) return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 0, 1))) return rc;
    return launch_col<double>(c, c.twz64, F, 2, 0, 1);
}

static int fk_to_host(Ctx &c, const cpx<double> *F, double *Fk_host)
{
    LAT_HIP(&c, hipMemcpy2DAsync(Fk_host, sizeof(cpx<double>) * c.nxh, F, sizeof(cpx<double>) * c.P,
                                 sizeof(cpx<double>) * c.nxh, (size_t)c.ny * c.nz, hipMemcpyDeviceToHost, c.stream));
    return 0;
}

static int fk_from_host(Ctx &c, cpx<double> *F, const double *Fk_host)
{
    LAT_HIP(&c, hipMemcpy2DAsync(F, sizeof(cpx<double>) * c.P, Fk_host, sizeof(cpx<double>) * c.nxh,
                                 sizeof(cpx<double>) * c.nxh, (size_t)c.ny * c.nz, hipMemcpyHostToDevice, c.stream));
    return 0;
}

// derivative multiply: G = i k_axis dk * F / nvol (signed wavenumbers)
__global__ void k_deriv_mult(const cpx<double> *__restrict__ F, cpx<double> *__restrict__ G, int nxh, int P, int ny,
                             int nz, int axis, double fac)
{
    const size_t n = (size_t)P * ny * nz;
    for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < n; id += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(id % P);
        const int j = (int)((id / P) % ny);
        const int k = (int)(id / ((size_t)P * ny));
        if (i >= nxh) continue;
        int w;
        if (axis == 0) w = i;
        else if (axis == 1) w = (j <= ny / 2) ? j : j - ny;
        else w = (k <= nz / 2) ? k : k - nz;
        const double m = (double)w * fac;
        const cpx<double> a = F[id];
        G[id] = {-m * a.y, m * a.x};
    }
}

__global__ void k_sq_acc(const double *__restrict__ d, double *__restrict__ acc, size_t n, int first)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc[i] = first ? d[i] * d[i] : acc[i] + d[i] * d[i];
}

// ---------------------------------------------------------------- mixed-form helpers
template <typename T>
static int upload_fields(Ctx &c, void *dst, const double *soa, size_t n)
{
    // the copy runs on the null stream, which does not order against the
    // context's non-blocking stream: drain it first
    int rc = sync(c);
    if (rc) return rc;
    if (sizeof(T) == 8) {
        LAT_HIP(&c, hipMemcpy(dst, soa, sizeof(double) * n, hipMemcpyHostToDevice));
        return 0;
    }
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)soa[i];
    LAT_HIP(&c, hipMemcpy(dst, tmp.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
static int download_fields(Ctx &c, double *soa, const void *src, size_t n)
{
    int rc = sync(c);
    if (rc) return rc;
    if (sizeof(T) == 8) {
        LAT_HIP(&c, hipMemcpy(soa, src, sizeof(double) * n, hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<float> tmp(n);
    LAT_HIP(&c, hipMemcpy(tmp.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) soa[i] = tmp[i];
    return 0;
}

} // namespace lat

using namespace lat;

#define CTX_OR_FAIL(ctx)            \
    if (!(ctx)) return -1;          \
    Ctx &c = (ctx)->c;              \
    LAT_HIP(&c, hipSetDevice(c.device))

// per-member operations take a single context or a member view, not a batch
#define ONE_MEMBER(c)                                                          \
    if ((c).B != 1) {                                                          \
        (c).last_error = "batch context: use lat_member for per-member calls"; \
        return -2;                                                             \
    }

extern "C" {

} // extern "C"

// B realisations of one configuration; mparams (B x 4, nullable) overrides
// lat_config.model_params per member.
static int create_ctx(const lat_config *cfg, int B, const double *mparams, lat_ctx **out)
{
    if (!cfg || !out || B < 1) return -1;
    *out = nullptr;
    if (B > 1 && cfg->gw) return -2; // u_ij are evolved in single contexts only
    if (!size_ok(cfg->nx) || !size_ok(cfg->ny) || !size_ok(cfg->nz)) return -2;
    if (cfg->nfld < 1 || cfg->nfld > 2) return -2;
    if (cfg->model == LAT_MODEL_TWOFIELD && cfg->nfld != 2) return -2;
    if (cfg->model == LAT_MODEL_PHI4 && cfg->nfld != 1) return -2;
    if (cfg->precision != LAT_F64 && cfg->precision != LAT_MIXED) return -2;
    if (!(cfg->len > 0.0)) return -2;
    if (cfg->laplacian == LAT_LAP_STENCIL && cfg->stencil[4] == 0.0) return -2;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return -4;
    lat_ctx *h = new lat_ctx();
    Ctx &c = h->c;
    c.cfg = *cfg;
    (void)hipGetDevice(&c.device);
    (void)hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, c.device);
    c.nx = cfg->nx;
    c.ny = cfg->ny;
    c.nz = cfg->nz;
    c.nxh = c.nx / 2 + 1;
    c.P = (c.nxh + 7) / 8 * 8;
    c.nvol = (size_t)c.nx * c.ny * c.nz;
    c.nk = (size_t)c.P * c.ny * c.nz;
    c.dx = cfg->len / c.nx;
    c.dk = 6.2831853071795864769252867665590 / cfg->len;
    c.prec = cfg->precision;
    c.nfld = cfg->nfld;
    c.ns = lat_num_bins(c.nx, c.ny, c.nz);
    {
        int nmax = c.nx > c.ny ? c.nx : c.ny;
        nmax = nmax > c.nz ? nmax : c.nz;
        const int nd = lat_num_bins_discrete(nmax);
        c.spec_cap = c.ns > nd ? c.ns : nd;
    }
    c.B = B;
    c.mvalid.assign(B, 0);
    c.ms.f = (size_t)c.nfld * c.nvol;
    c.ms.k = (size_t)(c.nfld > 2 ? c.nfld : 2) * c.nk;
    int rc = 0;
    hipError_t e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete h; return -3; }
    const size_t es = elem_size(c);
    rc |= ensure(c, &c.fld, es * B * c.ms.f, "hipMalloc fld");
    rc |= ensure(c, &c.fldp, es * B * c.ms.f, "hipMalloc fldp");
    if (cfg->laplacian == LAT_LAP_SPECTRAL || c.prec == LAT_F64)
        rc |= ensure(c, &c.Fk, 2 * es * B * c.ms.k, "hipMalloc Fk");
    c.gw = cfg->gw != 0;
    c.gw_coup = 2.0 / ((cfg->gw_mpl > 0.0 ? cfg->gw_mpl : 1.0) * (cfg->gw_mpl > 0.0 ? cfg->gw_mpl : 1.0));
    if (c.gw) {
        rc |= ensure(c, &c.gwu, es * 6 * c.nvol, "hipMalloc gwu");
        rc |= ensure(c, &c.gwp, es * 6 * c.nvol, "hipMalloc gwp");
        if (cfg->laplacian == LAT_LAP_SPECTRAL) rc |= ensure(c, &c.gwsrc, es * 2 * c.nvol, "hipMalloc gwsrc");
    }
    rc |= ensure(c, (void **)&c.dsc, sizeof(DevScalars) * B, "hipMalloc dsc");
    rc |= ensure(c, (void **)&c.partials, sizeof(double) * MAX_GRID * NRED * B, "hipMalloc partials");
    if (rc) { lat_destroy(h); return -3; }
    c.twx64 = make_twiddles<double>(c, c.nx, &rc);
    c.twy64 = make_twiddles<double>(c, c.ny, &rc);
    c.twz64 = make_twiddles<double>(c, c.nz, &rc);
    c.twx32 = make_twiddles<float>(c, c.nx, &rc);
    c.twy32 = make_twiddles<float>(c, c.ny, &rc);
    c.twz32 = make_twiddles<float>(c, c.nz, &rc);
    if (rc) { lat_destroy(h); return -3; }
    memset(&c.host_sc, 0, sizeof(DevScalars));
    c.host_sc.yscl = 1.0;
    c.host_sc.mixed = c.prec == LAT_MIXED ? 1 : 0;
    c.host_sc.mp[0] = cfg->model_params[0];
    c.host_sc.mp[1] = cfg->model_params[1];
    c.host_sc.mp[2] = cfg->model_params[2];
    {
        std::vector<DevScalars> all(B, c.host_sc);
        for (int m = 0; m < B && mparams; ++m) {
            all[m].mp[0] = mparams[4 * m];
            all[m].mp[1] = mparams[4 * m + 1];
            all[m].mp[2] = mparams[4 * m + 2];
        }
        c.host_sc = all[0];
        if (hipMemcpy(c.dsc, all.data(), sizeof(DevScalars) * B, hipMemcpyHostToDevice) != hipSuccess) {
            lat_destroy(h);
            return -3;
        }
    }
    (void)hipMemset(c.fld, 0, es * B * c.ms.f);
    (void)hipMemset(c.fldp, 0, es * B * c.ms.f);
    (void)hipMemset(c.partials, 0, sizeof(double) * MAX_GRID * NRED * B);
    if (c.gw) {
        (void)hipMemset(c.gwu, 0, es * 6 * c.nvol);
        (void)hipMemset(c.gwp, 0, es * 6 * c.nvol);
    }
    set_ke_valid(c, false);
    if ((rc = noalloc_setup(c))) { lat_destroy(h); return rc; }
    *out = h;
    return 0;
}

extern "C" {

int lat_create(const lat_config *cfg, lat_ctx **out) { return create_ctx(cfg, 1, nullptr, out); }

int lat_create_batch(const lat_config *cfg, int nmembers, const double *model_params, lat_ctx **out)
{
    return create_ctx(cfg, nmembers, model_params, out);
}

int lat_num_members(const lat_ctx *h) { return h ? h->c.B : -1; }

int lat_member(lat_ctx *h, int m, lat_ctx **out)
{
    CTX_OR_FAIL(h);
    if (!out || m < 0 || m >= c.B || c.view) return -1;
    // the scratch buffers that single-member operations allocate on demand
    // are allocated here, once, and borrowed by every view
    int rc = ensure_io(c);
    if (!rc) rc = ensure_spec(c);
    if (rc) return rc;
    const size_t es = elem_size(c);
    lat_ctx *v = new lat_ctx();
    Ctx &w = v->c;
    w.cfg = c.cfg;
    w.cfg.model_params[0] = 0.0; // filled from the member's scalars below
    w.B = 1;
    w.ms = c.ms;
    w.view = true;
    w.parent = &c;
    w.nx = c.nx; w.ny = c.ny; w.nz = c.nz; w.nxh = c.nxh; w.P = c.P;
    w.nvol = c.nvol; w.nk = c.nk; w.dx = c.dx; w.dk = c.dk; w.prec = c.prec; w.nfld = c.nfld;
    w.stream = c.stream;
    w.device = c.device;
    w.cus = c.cus;
    w.kblk = c.kblk;
    w.fld = (char *)c.fld + es * c.ms.f * m;
    w.fldp = (char *)c.fldp + es * c.ms.f * m;
    w.Fk = c.Fk ? (void *)((char *)c.Fk + 2 * es * c.ms.k * m) : nullptr;
    w.Fk2 = c.Fk2;
    w.io64 = c.io64;
    w.Fk64 = c.prec == LAT_F64 ? w.Fk : c.Fk64;
    w.gw = false;
    w.gw_coup = c.gw_coup;
    w.dsc = c.dsc + m;
    w.partials = c.partials + (size_t)MAX_GRID * NRED * m;
    w.spec_part = c.spec_part;
    w.spec_out = c.spec_out;
    w.ns = c.ns;
    w.spec_cap = c.spec_cap;
    w.twx64 = c.twx64; w.twy64 = c.twy64; w.twz64 = c.twz64;
    w.twx32 = c.twx32; w.twy32 = c.twy32; w.twz32 = c.twz32;
    w.tune = c.tune;
    if ((rc = pull_scalars(w))) { delete v; return rc; }
    w.cfg.model_params[0] = w.host_sc.mp[0];
    w.cfg.model_params[1] = w.host_sc.mp[1];
    w.cfg.model_params[2] = w.host_sc.mp[2];
    w.member = m;
    *out = v;
    return 0;
}

void lat_destroy(lat_ctx *h)
{
    if (!h) return;
    Ctx &c = h->c;
    if (c.view) { // borrows everything from its batch but its on-demand scratch
        if (!c.owned.empty()) {
            (void)hipSetDevice(c.device);
            (void)hipStreamSynchronize(c.stream);
            for (void *p : c.owned) (void)hipFree(p);
        }
        delete h;
        return;
    }
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    void *ptrs[] = {c.fld, c.fldp, c.Fk, c.Fk2, c.io64, (c.Fk64 != c.Fk) ? c.Fk64 : nullptr, c.dsc, c.partials,
                    c.spec_part, c.spec_out, c.twx64, c.twy64, c.twz64, c.twx32, c.twy32, c.twz32,
                    c.gwu, c.gwp, c.gwsrc, c.gwk, c.lap, c.disc_part};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : c.prof.pool) (void)hipEventDestroy(e);
    if (c.stream && c.owns_stream) (void)hipStreamDestroy(c.stream);
    delete h;
}

const char *lat_last_error(const lat_ctx *h) { return h ? h->c.last_error.c_str() : "null context"; }

int lat_synchronize(lat_ctx *h)
{
    CTX_OR_FAIL(h);
    return sync(c);
}

int lat_num_bins(int nx, int ny, int nz)
{
    // analysis_spec.f90:55-56
    long ns = (long)(nx / 2 + 1) * (nx / 2 + 1) + (long)(ny / 2 + 1) * (ny / 2 + 1) + (long)(nz / 2 + 1) * (nz / 2 + 1);
    return (int)std::floor(std::sqrt((double)ns + 1.0));
}

int lat_fft_r2c_3d(lat_ctx *h, const double *f, double *Fk)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!f || !Fk) return -1;
    int rc = host_r2c(c, f);
    if (rc) return rc;
    if ((rc = fk_to_host(c, (const cpx<double> *)c.Fk64, Fk))) return rc;
    return sync(c);
}

int lat_fft_c2r_3d(lat_ctx *h, const double *Fk, double *f)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!f || !Fk) return -1;
    int rc = ensure_io(c);
    if (rc) return rc;
    cpx<double> *F = (cpx<double> *)c.Fk64;
    if ((rc = fk_from_host(c, F, Fk))) return rc;
    if ((rc = launch_col<double>(c, c.twz64, F, 2, 1, 1))) return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 1, 1))) return rc;
    rc = launch_x_c2r<double>(c, c.twx64, F, 1, EPI_STORE, c.io64, nullptr, nullptr);
    if (rc < 0) return rc;
    LAT_HIP(&c, hipMemcpyAsync(f, c.io64, sizeof(double) * c.nvol, hipMemcpyDeviceToHost, c.stream));
    return sync(c);
}

int lat_laplacian_3d(lat_ctx *h, int n1, int n2, int n3, double *f, double dk)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!f) return -1;
    if (n1 != c.nx || n2 != c.ny || n3 != c.nz) {
        c.last_error = "lattice size differs from the context";
        return -2;
    }
    int rc = ensure_io(c);
    if (rc) return rc;
    cpx<double> *F = (cpx<double> *)c.Fk64;
    LAT_HIP(&c, hipMemcpyAsync(c.io64, f, sizeof(double) * c.nvol, hipMemcpyHostToDevice, c.stream));
    if ((rc = launch_x_r2c<double>(c, c.twx64, c.io64, nullptr, F, 0))) return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 0, 1))) return rc;
    const double saved = c.dk;
    c.dk = dk;
    rc = launch_col<double>(c, c.twz64, F, 2, 2, 1);
    c.dk = saved;
    if (rc) return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 1, 1))) return rc;
    rc = launch_x_c2r<double>(c, c.twx64, F, 1, EPI_STORE, c.io64, nullptr, nullptr);
    if (rc < 0) return rc;
    LAT_HIP(&c, hipMemcpyAsync(f, c.io64, sizeof(double) * c.nvol, hipMemcpyDeviceToHost, c.stream));
    return sync(c);
}

int lat_gradient_squared_3d(lat_ctx *h, const double *f, double *grad2, double dk)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!f || !grad2) return -1;
    int rc = host_r2c(c, f);
    if (rc) return rc;
    // second FP64 half-complex buffer: upper half of the Fk64 allocation
    cpx<double> *F = (cpx<double> *)c.Fk64;
    cpx<double> *G = F + c.nk;
    double *acc = c.io64 + c.nvol;
    for (int ax = 2; ax >= 0; --ax) {
        hipLaunchKernelGGL(k_deriv_mult, dim3(4096), dim3(256), 0, c.stream, F, G, c.nxh, c.P, c.ny, c.nz, ax,
                           dk / (double)c.nvol);
        LAT_HIP(&c, hipGetLastError());
        if ((rc = launch_col<double>(c, c.twz64, G, 2, 1, 1))) return rc;
        if ((rc = launch_col<double>(c, c.twy64, G, 1, 1, 1))) return rc;
        rc = launch_x_c2r<double>(c, c.twx64, G, 1, EPI_STORE, c.io64, nullptr, nullptr);
        if (rc < 0) return rc;
        hipLaunchKernelGGL(k_sq_acc, dim3(4096), dim3(256), 0, c.stream, c.io64, acc, c.nvol, ax == 2 ? 1 : 0);
        LAT_HIP(&c, hipGetLastError());
    }
    LAT_HIP(&c, hipMemcpyAsync(grad2, acc, sizeof(double) * c.nvol, hipMemcpyDeviceToHost, c.stream));
    return sync(c);
}

int lat_stencil_laplacian(lat_ctx *h, const double *f, double *lap)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!f || !lap) return -1;
    if (c.cfg.stencil[4] == 0.0) {
        c.last_error = "stencil coefficients not set (cc = 0)";
        return -2;
    }
    int rc = ensure_io(c);
    if (rc) return rc;
    LAT_HIP(&c, hipMemcpyAsync(c.io64, f, sizeof(double) * c.nvol, hipMemcpyHostToDevice, c.stream));
    if ((rc = launch_stencil_io(c, c.io64, c.io64 + c.nvol))) return rc;
    LAT_HIP(&c, hipMemcpyAsync(lap, c.io64 + c.nvol, sizeof(double) * c.nvol, hipMemcpyDeviceToHost, c.stream));
    return sync(c);
}

} // extern "C"

int lat::ensure_disc(Ctx &c)
{
    return ensure(c, (void **)&c.disc_part, sizeof(double) * MAX_GRID * 16, "hipMalloc disc_part");
}

int lat::ensure_spec(Ctx &c)
{
    const size_t cap = (size_t)c.spec_cap;
    int rc = ensure(c, (void **)&c.spec_part, sizeof(double) * (BIN_BLOCKS + BIN_BLOCKS / BIN_GROUP) * 5 * cap,
                    "hipMalloc spec_part");
    rc |= ensure(c, (void **)&c.spec_out, sizeof(double) * (5 + Ctx::WCACHE) * cap, "hipMalloc spec_out");
    return rc;
}

extern "C" {

// S(b) / W(b) / div where W(b) != 0 (analysis_spec.f90:101)
static void finish_spectra(const std::vector<double> &raw, int ns, int nch, double *const *outs, double div = 1.0)
{
    const double *W = raw.data() + (size_t)nch * ns;
    for (int ch = 0; ch < nch; ++ch)
        for (int b = 0; b < ns; ++b) {
            double s = raw[(size_t)ch * ns + b];
            if (W[b] != 0.0) s = s / W[b] / div;
            outs[ch][b] = s;
        }
}

int lat_spectrum_3d(lat_ctx *h, double *S, const double *f, double *Fk)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!S || !f) return -1;
    int rc = ensure_spec(c);
    if (rc) return rc;
    if ((rc = host_r2c(c, f))) return rc;
    const cpx<double> *F = (const cpx<double> *)c.Fk64;
    if ((rc = launch_bin<double>(c, F, F, 0)) < 0) return rc;
    std::vector<double> raw(2 * c.ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 2 * c.ns, hipMemcpyDeviceToHost, c.stream));
    if (Fk && (rc = fk_to_host(c, F, Fk))) return rc;
    if ((rc = sync(c))) return rc;
    double *outs[1] = {S};
    finish_spectra(raw, c.ns, 1, outs);
    return 0;
}

int lat_num_bins_discrete(int n)
{
    // Discrete/Entropy/params.f90:45, integer conversion truncates
    return (int)(1.7320508075688772935274463415059 * (double)(n / 2) + 2.0);
}

int lat_spectrum_discrete(lat_ctx *h, const double *f, double *S, int kind)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!S || !f || (kind != 0 && kind != 1)) return -1;
    if (c.nx != c.ny || c.nx != c.nz) {
        c.last_error = "Discrete-tree spectra need a cubic lattice";
        return -2;
    }
    int rc = ensure_spec(c);
    if (rc) return rc;
    if ((rc = host_r2c(c, f))) return rc;
    const int ns = lat_num_bins_discrete(c.nx);
    const cpx<double> *F = (const cpx<double> *)c.Fk64;
    if ((rc = launch_bin<double>(c, F, F, 0, ns, kind == 1 ? BIN_COUNTS : 0)) < 0) return rc;
    std::vector<double> raw(2 * (size_t)ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, c.stream));
    if ((rc = sync(c))) return rc;
    double *outs[1] = {S};
    // kind 0: S/W/n^6 (analysis.f90:113); kind 1: S/W (analysis.f90:50)
    const double n3 = (double)c.nvol;
    finish_spectra(raw, ns, 1, outs, kind == 0 ? n3 * n3 : 1.0);
    return 0;
}

int lat_spectrum_discrete_k(lat_ctx *h, const double *Fk, double *S, int kind)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!S || !Fk || (kind != 0 && kind != 1)) return -1;
    if (c.nx != c.ny || c.nx != c.nz) {
        c.last_error = "Discrete-tree spectra need a cubic lattice";
        return -2;
    }
    int rc = ensure_spec(c);
    if (rc) return rc;
    if ((rc = ensure_io(c))) return rc;
    cpx<double> *F = (cpx<double> *)c.Fk64;
    if ((rc = fk_from_host(c, F, Fk))) return rc;
    const int ns = lat_num_bins_discrete(c.nx);
    // spectrum_nofft: half space, weights; spectrum_bin_nofft: half space
    // without the x-Nyquist plane, counts
    const int flags = kind == 0 ? BIN_HALF : (BIN_HALF | BIN_NO_NYQ | BIN_COUNTS);
    if ((rc = launch_bin<double>(c, F, F, 0, ns, flags)) < 0) return rc;
    std::vector<double> raw(2 * (size_t)ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, c.stream));
    if ((rc = sync(c))) return rc;
    double *outs[1] = {S};
    const double n3 = (double)c.nvol;
    finish_spectra(raw, ns, 1, outs, kind == 0 ? n3 * n3 : 1.0); // analysis.f90:146 / :82
    return 0;
}

int lat_spectrum_discrete_k_noalloc(lat_ctx *h, const double *Fk_dev, double *S_dev, int kind)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!S_dev || !Fk_dev || (kind != 0 && kind != 1)) return -1;
    if (c.nx != c.ny || c.nx != c.nz) {
        c.last_error = "Discrete-tree spectra need a cubic lattice";
        return -2;
    }
    return spectrum_discrete_k_noalloc(c, Fk_dev, S_dev, kind);
}

int lat_crossspec_discrete_k(lat_ctx *h, const double *Fk1, const double *Fk2, double *S1, double *S2)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!Fk1 || !Fk2 || !S1 || !S2) return -1;
    if (c.nx != c.ny || c.nx != c.nz) {
        c.last_error = "Discrete-tree spectra need a cubic lattice";
        return -2;
    }
    int rc = ensure_spec(c);
    if (rc) return rc;
    if ((rc = ensure_io(c))) return rc;
    cpx<double> *F1 = (cpx<double> *)c.Fk64;
    cpx<double> *F2 = F1 + c.nk;
    if ((rc = fk_from_host(c, F1, Fk1))) return rc;
    if ((rc = fk_from_host(c, F2, Fk2))) return rc;
    const int ns = lat_num_bins_discrete(c.nx);
    if ((rc = launch_bin<double>(c, F1, F2, 1, ns, BIN_HALF | BIN_NO_NYQ)) < 0) return rc;
    std::vector<double> raw(3 * (size_t)ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 3 * ns, hipMemcpyDeviceToHost, c.stream));
    if ((rc = sync(c))) return rc;
    double *outs[2] = {S1, S2};
    finish_spectra(raw, ns, 2, outs); // analysis.f90:183-184, no n^6
    return 0;
}

int lat_crossspec_3d(lat_ctx *h, const double *Fk1, const double *Fk2, double *S_re, double *S_im)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!Fk1 || !Fk2 || !S_re || !S_im) return -1;
    int rc = ensure_spec(c);
    if (rc) return rc;
    if ((rc = ensure_io(c))) return rc;
    cpx<double> *F1 = (cpx<double> *)c.Fk64;
    cpx<double> *F2 = F1 + c.nk;
    if ((rc = fk_from_host(c, F1, Fk1))) return rc;
    if ((rc = fk_from_host(c, F2, Fk2))) return rc;
    if ((rc = launch_bin<double>(c, F1, F2, 1)) < 0) return rc;
    std::vector<double> raw(3 * c.ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 3 * c.ns, hipMemcpyDeviceToHost, c.stream));
    if ((rc = sync(c))) return rc;
    double *outs[2] = {S_re, S_im};
    finish_spectra(raw, c.ns, 2, outs);
    return 0;
}

} // extern "C"

template <typename T>
static int output_spectra_t(Ctx &c, int field, double *out)
{
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
    int rc = ensure(c, &c.Fk, 2 * sizeof(cpx<T>) * c.nk, "hipMalloc Fk");
    if (rc) return rc;
    cpx<T> *Fphi = (cpx<T> *)c.Fk;
    cpx<T> *Fpi = Fphi + c.nk;
    T *fld = (T *)c.fld + (size_t)field * c.nvol;
    T *fldp = (T *)c.fldp + (size_t)field * c.nvol;
    if ((rc = launch_x_r2c<T>(c, twx, fld, nullptr, Fphi, 0))) return rc;
    if ((rc = launch_x_r2c<T>(c, twx, fldp, nullptr, Fpi, 0))) return rc;
    if ((rc = launch_col<T>(c, twy, Fphi, 1, 0, 2))) return rc;
    if ((rc = launch_col<T>(c, twz, Fphi, 2, 0, 2))) return rc;
    // crossspec_3d(Fk = pi^, Fk2 = phi^) as called at evolve_spectral.f90:534-538
    if ((rc = launch_bin<T>(c, Fpi, Fphi, 2)) < 0) return rc;
    std::vector<double> raw(5 * c.ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 5 * c.ns, hipMemcpyDeviceToHost, c.stream));
    cpx<T> f0[2];
    if (sizeof(T) == 4) {
        LAT_HIP(&c, hipMemcpyAsync(&f0[0], Fphi, sizeof(cpx<T>), hipMemcpyDeviceToHost, c.stream));
        LAT_HIP(&c, hipMemcpyAsync(&f0[1], Fpi, sizeof(cpx<T>), hipMemcpyDeviceToHost, c.stream));
        if ((rc = pull_scalars(c))) return rc;
    }
    if ((rc = sync(c))) return rc;
    if (sizeof(T) == 4) {
        // Mixed form: the FP32 transform holds the fluctuations only.  Bin 0 is
        // the k = 0 mode alone (weight 1): the FP64 mean times the volume plus
        // the fluctuation sum, which is the transform's own k = 0 coefficient
        // (its rounding is that of every other mode of the same transform).
        const double nv = (double)c.nvol;
        const double fphi = nv * c.host_sc.mean_phi[field] + (double)f0[0].x;
        const double fpi = nv * c.host_sc.mean_pi[field] + (double)f0[1].x;
        raw[0] = fphi * fphi;
        raw[(size_t)c.ns] = fpi * fpi;
        raw[2 * (size_t)c.ns] = fpi * fphi;
        raw[3 * (size_t)c.ns] = 0.0;
    }
    double *outs[4] = {out, out + c.ns, out + 2 * c.ns, out + 3 * c.ns};
    finish_spectra(raw, c.ns, 4, outs);
    return 0;
}

extern "C" {

int lat_output_spectra(lat_ctx *h, int field, double *out)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!out || field < 0 || field >= c.nfld) return -1;
    int rc = ensure_spec(c);
    if (rc) return rc;
    return c.prec == LAT_F64 ? output_spectra_t<double>(c, field, out) : output_spectra_t<float>(c, field, out);
}

} // extern "C"

// internal component (xx, xy, xz, yy, yz, zz) of the boundary pair hij(a+1, b+1):
// a = 0 diagonal u_bb, a = 1 off-diagonal u_12, u_13, u_23
static int gw_comp(int a, int b)
{
    static const int diag[3] = {0, 3, 5}, off[3] = {1, 2, 4};
    return a == 0 ? diag[b] : off[b];
}

template <typename T>
static int gw_spectrum_t(Ctx &c, double *out)
{
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
    int rc = ensure(c, &c.gwk, 6 * sizeof(cpx<T>) * c.nk, "hipMalloc gwk");
    if (rc) return rc;
    cpx<T> *K = (cpx<T> *)c.gwk;
    const int ns = c.ns;
    std::vector<double> raw(2 * (size_t)ns);
    for (int ch = 0; ch < 2; ++ch) {
        T *arr = (T *)(ch == 0 ? c.gwp : c.gwu);
        for (int q = 0; q < 6; ++q)
            if ((rc = launch_x_r2c<T>(c, twx, arr + (size_t)q * c.nvol, nullptr, K + (size_t)q * c.nk, 0, -1)))
                return rc;
        if ((rc = launch_col<T>(c, twy, K, 1, 0, 6))) return rc;
        if ((rc = launch_col<T>(c, twz, K, 2, 0, 6))) return rc;
        if ((rc = launch_bin<T>(c, K, nullptr, 3)) < 0) return rc;
        LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * 2 * ns, hipMemcpyDeviceToHost, c.stream));
        if ((rc = sync(c))) return rc;
        double *outs[1] = {out + (size_t)ch * ns};
        finish_spectra(raw, ns, 1, outs);
    }
    return 0;
}

extern "C" {

int lat_set_gw(lat_ctx *h, const double *hij, const double *pi_hij)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!hij || !pi_hij) return -1;
    if (!c.gw) {
        c.last_error = "context created without gw";
        return -2;
    }
    std::vector<double> u(6 * c.nvol), p(6 * c.nvol);
    for (size_t s = 0; s < c.nvol; ++s)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 2; ++a) {
                const size_t q = (size_t)gw_comp(a, b) * c.nvol + s;
                u[q] = hij[a + 2 * (b + 3 * s)];
                p[q] = pi_hij[a + 2 * (b + 3 * s)];
            }
    int rc = c.prec == LAT_F64 ? upload_fields<double>(c, c.gwu, u.data(), u.size())
                               : upload_fields<float>(c, c.gwu, u.data(), u.size());
    if (rc) return rc;
    return c.prec == LAT_F64 ? upload_fields<double>(c, c.gwp, p.data(), p.size())
                             : upload_fields<float>(c, c.gwp, p.data(), p.size());
}

int lat_get_gw(lat_ctx *h, double *hij, double *pi_hij)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!hij || !pi_hij) return -1;
    if (!c.gw) {
        c.last_error = "context created without gw";
        return -2;
    }
    std::vector<double> u(6 * c.nvol), p(6 * c.nvol);
    int rc = c.prec == LAT_F64 ? download_fields<double>(c, u.data(), c.gwu, u.size())
                               : download_fields<float>(c, u.data(), c.gwu, u.size());
    if (rc) return rc;
    rc = c.prec == LAT_F64 ? download_fields<double>(c, p.data(), c.gwp, p.size())
                           : download_fields<float>(c, p.data(), c.gwp, p.size());
    if (rc) return rc;
    for (size_t s = 0; s < c.nvol; ++s)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 2; ++a) {
                const size_t q = (size_t)gw_comp(a, b) * c.nvol + s;
                hij[a + 2 * (b + 3 * s)] = u[q];
                pi_hij[a + 2 * (b + 3 * s)] = p[q];
            }
    return 0;
}

int lat_gw_source(lat_ctx *h, double *Sij)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!Sij) return -1;
    const size_t es = elem_size(c);
    int rc = ensure(c, &c.gwsrc, es * 2 * c.nvol, "hipMalloc gwsrc");
    if (rc) return rc;
    for (int c0 = 0; c0 < 6; c0 += 2) {
        rc = c.prec == LAT_F64 ? launch_gw_source<double>(c, c0, 2, (double *)c.gwsrc)
                               : launch_gw_source<float>(c, c0, 2, (float *)c.gwsrc);
        if (rc) return rc;
        rc = c.prec == LAT_F64 ? download_fields<double>(c, Sij + (size_t)c0 * c.nvol, c.gwsrc, 2 * c.nvol)
                               : download_fields<float>(c, Sij + (size_t)c0 * c.nvol, c.gwsrc, 2 * c.nvol);
        if (rc) return rc;
    }
    return 0;
}

int lat_gw_spectrum(lat_ctx *h, double *out)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!out) return -1;
    if (!c.gw) {
        c.last_error = "context created without gw";
        return -2;
    }
    int rc = ensure_spec(c);
    if (rc) return rc;
    return c.prec == LAT_F64 ? gw_spectrum_t<double>(c, out) : gw_spectrum_t<float>(c, out);
}

int lat_set_state(lat_ctx *h, const double *fld, const double *fldp, int ghost, double yscl, double ysclp)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!fld || !fldp || ghost < 0 || ghost > 1) return -1;
    const int g = ghost, nf = c.nfld;
    const size_t sx = c.nx + 2 * g, sy = c.ny + 2 * g;
    std::vector<double> soa((size_t)nf * c.nvol), soap((size_t)nf * c.nvol);
    // fld carries the ghost layer (hamiltonian_conformal.f90:22 SIRANGE), fldp never does (:23 IRANGE)
    for (int k = 0; k < c.nz; ++k)
        for (int j = 0; j < c.ny; ++j)
            for (int i = 0; i < c.nx; ++i) {
                const size_t s = (size_t)i + (size_t)c.nx * ((size_t)j + (size_t)c.ny * k);
                const size_t gs = (size_t)(i + g) + sx * ((size_t)(j + g) + sy * (size_t)(k + g));
                for (int l = 0; l < nf; ++l) {
                    soa[(size_t)l * c.nvol + s] = fld[(size_t)l + (size_t)nf * gs];
                    soap[(size_t)l * c.nvol + s] = fldp[(size_t)l + (size_t)nf * s];
                }
            }
    double mphi[2] = {0, 0}, mpi[2] = {0, 0};
    if (c.prec == LAT_MIXED) {
        // split into FP64 homogeneous parts and fluctuations (stored in FP32)
        for (int l = 0; l < nf; ++l) {
            double a = 0.0, b = 0.0;
            for (size_t s = 0; s < c.nvol; ++s) {
                a += soa[(size_t)l * c.nvol + s];
                b += soap[(size_t)l * c.nvol + s];
            }
            mphi[l] = a / (double)c.nvol;
            mpi[l] = b / (double)c.nvol;
            for (size_t s = 0; s < c.nvol; ++s) {
                soa[(size_t)l * c.nvol + s] -= mphi[l];
                soap[(size_t)l * c.nvol + s] -= mpi[l];
            }
        }
    }
    int rc = c.prec == LAT_F64 ? upload_fields<double>(c, c.fld, soa.data(), soa.size())
                               : upload_fields<float>(c, c.fld, soa.data(), soa.size());
    if (rc) return rc;
    rc = c.prec == LAT_F64 ? upload_fields<double>(c, c.fldp, soap.data(), soap.size())
                           : upload_fields<float>(c, c.fldp, soap.data(), soap.size());
    if (rc) return rc;
    if ((rc = pull_scalars(c))) return rc;
    c.host_sc.yscl = yscl;
    c.host_sc.ysclp = ysclp;
    c.host_sc.nd = 0;
    for (int l = 0; l < 2; ++l) {
        c.host_sc.mean_phi[l] = mphi[l];
        c.host_sc.mean_pi[l] = mpi[l];
        c.host_sc.nu[l] = c.host_sc.mu[l] = c.host_sc.sdphi[l] = 0.0;
    }
    set_ke_valid(c, false);
    return push_scalars(c);
}

int lat_get_state(lat_ctx *h, double *fld, double *fldp, int ghost, double *yscl, double *ysclp)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (ghost < 0 || ghost > 1) return -1;
    int rc = sync(c);
    if (rc) return rc;
    const int g = ghost, nf = c.nfld;
    const size_t sx = c.nx + 2 * g, sy = c.ny + 2 * g, sz = c.nz + 2 * g;
    std::vector<double> soa((size_t)nf * c.nvol), soap((size_t)nf * c.nvol);
    if ((rc = pull_scalars(c))) return rc;
    const bool mixed = c.host_sc.mixed != 0;
    if (fld) {
        rc = c.prec == LAT_F64 ? download_fields<double>(c, soa.data(), c.fld, soa.size())
                               : download_fields<float>(c, soa.data(), c.fld, soa.size());
        if (rc) return rc;
        if (mixed)
            for (int l = 0; l < nf; ++l)
                for (size_t s = 0; s < c.nvol; ++s) soa[(size_t)l * c.nvol + s] += c.host_sc.mean_phi[l];
        for (size_t kk = 0; kk < sz; ++kk)
            for (size_t jj = 0; jj < sy; ++jj)
                for (size_t ii = 0; ii < sx; ++ii) {
                    // periodic image for ghost cells (wrap_fields, hamiltonian_conformal.f90:309-330)
                    const int i = (int)((ii - g + c.nx) % c.nx), j = (int)((jj - g + c.ny) % c.ny),
                              k = (int)((kk - g + c.nz) % c.nz);
                    const size_t s = (size_t)i + (size_t)c.nx * ((size_t)j + (size_t)c.ny * k);
                    const size_t gs = ii + sx * (jj + sy * kk);
                    for (int l = 0; l < nf; ++l) fld[(size_t)l + (size_t)nf * gs] = soa[(size_t)l * c.nvol + s];
                }
    }
    if (fldp) {
        rc = c.prec == LAT_F64 ? download_fields<double>(c, soap.data(), c.fldp, soap.size())
                               : download_fields<float>(c, soap.data(), c.fldp, soap.size());
        if (rc) return rc;
        if (mixed)
            for (int l = 0; l < nf; ++l)
                for (size_t s = 0; s < c.nvol; ++s) soap[(size_t)l * c.nvol + s] += c.host_sc.mean_pi[l];
        for (size_t s = 0; s < c.nvol; ++s)
            for (int l = 0; l < nf; ++l) fldp[(size_t)l + (size_t)nf * s] = soap[(size_t)l * c.nvol + s];
    }
    if ((rc = pull_scalars(c))) return rc;
    if (yscl) *yscl = c.host_sc.yscl;
    if (ysclp) *ysclp = c.host_sc.ysclp;
    return 0;
}

int lat_get_scalars(lat_ctx *h, double *yscl, double *ysclp)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    int rc = pull_scalars(c);
    if (rc) return rc;
    if (yscl) *yscl = c.host_sc.yscl;
    if (ysclp) *ysclp = c.host_sc.ysclp;
    return 0;
}

int lat_set_scalars(lat_ctx *h, double yscl, double ysclp)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    int rc = pull_scalars(c);
    if (rc) return rc;
    c.host_sc.yscl = yscl;
    c.host_sc.ysclp = ysclp;
    return push_scalars(c);
}

int lat_split(lat_ctx *h, double dt, int hamind)
{
    CTX_OR_FAIL(h);
    if (hamind < 1 || hamind > 3) {
        c.last_error = "Undefined Hamiltonian term"; // hamiltonian_conformal.f90:78
        return -1;
    }
    std::vector<Op> ops;
    map_split(c, ops, dt, hamind);
    return run_ops(c, ops);
}

int lat_symp_o2step(lat_ctx *h, double dt, double w1, double w2)
{
    CTX_OR_FAIL(h);
    std::vector<Op> ops;
    o2step(c, ops, dt, w1, w2);
    return run_ops(c, ops);
}

int lat_schedule(int time_var, int order, double dt, int nsteps, int symp6_tree, int *types, double *hh,
                 int max_ops)
{
    Ctx c{};
    c.cfg.time_var = time_var;
    c.cfg.symp6_tree = symp6_tree;
    std::vector<Op> ops;
    if (build_symp(c, ops, order, dt, nsteps)) return -1;
    const bool conf = time_var == LAT_CONFORMAL;
    for (size_t q = 0; q < ops.size() && (int)q < max_ops; ++q) {
        if (types) types[q] = ops[q].type == OP_KICK ? 3 : (ops[q].type == OP_DRIFT ? 1 : 2);
        if (hh) hh[q] = ops[q].h;
    }
    (void)conf;
    return (int)ops.size();
}

int lat_symp(lat_ctx *h, int order, double dt, int nsteps)
{
    CTX_OR_FAIL(h);
    std::vector<Op> ops;
    if (build_symp(c, ops, order, dt, nsteps)) {
        c.last_error = "symp order must be 2, 4, 6 or 8";
        return -1;
    }
    return run_ops(c, ops);
}

int lat_calc_metric(lat_ctx *h)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    int rc = state_sums(c);
    if (rc) return rc;
    const double nv = (double)c.nvol, y = c.host_sc.yscl;
    const double a = scale_factor(c, y);
    const bool conf = c.cfg.time_var == LAT_CONFORMAL;
    const double knorm = conf ? 1.0 / std::pow(y, 6) : 1.0 / std::pow(y, 4);
    const PhysSums ps = phys_sums(c);
    const double KE = 0.5 * ps.pi2 * knorm / nv;
    const double PE = ps.v / nv;
    const double GE = 0.5 * ps.ge / (a * a) / nv;
    const double rho = KE + PE + GE;
    c.host_sc.ysclp = conf ? -std::sqrt(std::pow(y, 4) * 12.0 * rho) : -a * std::sqrt(rho * 16.0 / 3.0);
    return push_scalars(c);
}

int lat_energies(lat_ctx *h, double *out)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!out) return -1;
    int rc = state_sums(c);
    if (rc) return rc;
    const DevScalars &s = c.host_sc;
    const double nv = (double)c.nvol, y = s.yscl;
    const double a = scale_factor(c, y);
    const bool conf = c.cfg.time_var == LAT_CONFORMAL;
    const double knorm = conf ? 1.0 / std::pow(y, 6) : 1.0 / std::pow(y, 4);
    const double hub = conf ? -s.ysclp / 6.0 / (y * y) : -s.ysclp / y / 4.0;
    const PhysSums ps = phys_sums(c);
    const double KE = 0.5 * ps.pi2 * knorm / nv;
    const double PE = ps.v / nv;
    const double GE = 0.5 * ps.ge / nv / (a * a);
    const double rho = KE + PE + GE;
    const double grav = -3.0 * hub * hub;
    out[LAT_E_A] = a;
    out[LAT_E_HUBBLE] = hub;
    out[LAT_E_RHO] = rho;
    out[LAT_E_KE] = KE;
    out[LAT_E_PE] = PE;
    out[LAT_E_GE] = GE;
    out[LAT_E_GRAV] = grav;
    out[LAT_E_RESID] = (rho + grav) / rho;
    out[LAT_E_PMOM] = conf ? -s.ysclp * s.ysclp / 12.0 / std::pow(a, 4) : -s.ysclp * s.ysclp * 3.0 / 16.0 / (a * a);
    out[LAT_E_MEAN1] = ps.mean[0];
    out[LAT_E_MEAN2] = c.nfld > 1 ? ps.mean[1] : 0.0;
    // grad_en (Discrete/Entropy/homlat.f90:88-115): enorm = 1/(4 (a dx)^2 cc)
    // times the lattice sum of STENCIL(c, GRAD2), over n^3; by summation by
    // parts it equals GE (grad_en_wlap, :120-147).  The spectral path has no
    // stencil form of the gradient energy: NaN.
    out[LAT_E_GEGRAD] = c.cfg.laplacian == LAT_LAP_STENCIL
                            ? 0.25 / ((a * c.dx) * (a * c.dx)) / c.cfg.stencil[4] * s.red[RC_X] / nv
                            : std::nan("");
    return 0;
}

} // extern "C"

// Fortran ES22.15 edit descriptor as written by '(30(ES22.15,2X))': one
// digit before the point, 15 after, exponent E+dd (3-digit exponents drop the
// E), right-justified in 22 columns.
static std::string fortran_es22(double x)
{
    char buf[64];
    if (std::isnan(x)) {
        snprintf(buf, sizeof buf, "%22s", "NaN");
        return buf;
    }
    if (std::isinf(x)) {
        snprintf(buf, sizeof buf, "%22s", x > 0 ? "Infinity" : "-Infinity");
        return buf;
    }
    snprintf(buf, sizeof buf, "%.15E", x);
    std::string t(buf);
    const size_t epos = t.find('E');
    std::string mant = t.substr(0, epos), ex = t.substr(epos + 1); // ex = "+dd" or "+ddd"
    std::string field = ex.size() > 3 ? mant + ex : mant + "E" + ex;
    if (field.size() < 22) field = std::string(22 - field.size(), ' ') + field;
    return field;
}

static int append_line(Ctx &c, const std::string &path, const std::vector<double> &vals)
{
    FILE *f = fopen(path.c_str(), "a");
    if (!f) {
        c.last_error = "cannot open " + path;
        return -5;
    }
    std::string line;
    for (size_t q = 0; q < vals.size(); ++q) {
        if (q) line += "  ";
        line += fortran_es22(vals[q]);
    }
    line += "\n";
    fputs(line.c_str(), f);
    fclose(f);
    return 0;
}

extern "C" {

int lat_discrete_output(lat_ctx *h, double time, double dt, double mpl, int seed, const char *dir, double *vals,
                        double *psd)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (c.cfg.laplacian != LAT_LAP_STENCIL) {
        c.last_error = "Discrete-tree output needs a finite-difference (stencil) context";
        return -2;
    }
    if (c.nx != c.ny || c.nx != c.nz) {
        c.last_error = "Discrete-tree output needs a cubic lattice";
        return -2;
    }
    int rc = pull_scalars(c);
    if (rc) return rc;
    return discrete_output(c, time, dt, mpl, seed, dir, vals, psd);
}

int lat_make_output(lat_ctx *h, double time, int tree, const char *dir)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!dir || (tree != 0 && tree != 1)) return -1;
    double e[LAT_NE];
    int rc = lat_energies(h, e);
    if (rc) return rc;
    const PhysSums ps = phys_sums(c);
    const double a = e[LAT_E_A], ysp = c.host_sc.ysclp;
    // dump_rho: openmp_dev/evolve_spectral.f90:403-405 ends with <phi_2> (an
    // out-of-bounds read for one field; 0 here), mpi_dev/evolve_spectral.f90:311-313 with <pi_1>
    std::vector<double> row = {time, a, e[LAT_E_RHO], e[LAT_E_KE], e[LAT_E_PE], e[LAT_E_GE], e[LAT_E_GRAV],
                               (e[LAT_E_RHO] + e[LAT_E_GRAV]) / e[LAT_E_RHO],
                               -(ysp * ysp) / 12.0 / std::pow(a, 4), e[LAT_E_MEAN1],
                               tree == 0 ? e[LAT_E_MEAN2] : ps.mean_pi[0]};
    const std::string d(dir);
    if ((rc = append_line(c, d + "/energy_spec.out", row))) return rc;
    // make_output spectra (evolve_spectral.f90:523-551): S_phi1, S_pi1, Re/Im cross;
    // spec(ns, 2*nfld+2) has 2*nfld-2 further columns the reference never sets (0 here)
    std::vector<double> sp(4 * (size_t)c.ns);
    if ((rc = lat_output_spectra(h, 0, sp.data()))) return rc;
    const int ncol = 2 * c.nfld + 2;
    for (int b = 0; b < c.ns; ++b) {
        std::vector<double> vals(1 + ncol, 0.0);
        vals[0] = b * c.dk;
        for (int q = 0; q < 4; ++q) vals[1 + q] = sp[(size_t)q * c.ns + b];
        if ((rc = append_line(c, d + "/spectrum.out", vals))) return rc;
    }
    FILE *f = fopen((d + "/spectrum.out").c_str(), "a");
    if (!f) return -5;
    fputs("\n", f); // write(97,*): an empty record
    fclose(f);
    return 0;
}

// Checkpoint files of Discrete/Entropy/driver.f90:68-118 and scafld.f90:224-284:
// fld_<i>.raw / fldp_<i>.raw hold field i (1-based) as native FP64, x fastest,
// no ghosts; checkpt.dat holds "yscl ysclp", "time", "outframenum" as
// list-directed records.
int lat_checkpoint(lat_ctx *h, const char *dir, double time, int outframenum)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!dir) return -1;
    const size_t nv = c.nvol, nf = c.nfld;
    std::vector<double> aos(nf * nv), aosp(nf * nv), raw(nv);
    double y = 0, yp = 0;
    int rc = lat_get_state(h, aos.data(), aosp.data(), 0, &y, &yp);
    if (rc) return rc;
    const std::string d(dir);
    for (size_t l = 0; l < nf; ++l)
        for (int which = 0; which < 2; ++which) {
            const std::vector<double> &src = which ? aosp : aos;
            for (size_t q = 0; q < nv; ++q) raw[q] = src[l + nf * q];
            const std::string path = d + (which ? "/fldp_" : "/fld_") + std::to_string(l + 1) + ".raw";
            FILE *f = fopen(path.c_str(), "wb");
            if (!f) {
                c.last_error = "cannot write " + path;
                return -5;
            }
            const size_t w = fwrite(raw.data(), sizeof(double), nv, f);
            fclose(f);
            if (w != nv) {
                c.last_error = "short write " + path;
                return -5;
            }
        }
    FILE *f = fopen((d + "/checkpt.dat").c_str(), "w");
    if (!f) {
        c.last_error = "cannot write checkpt.dat";
        return -5;
    }
    fprintf(f, " %.17g %.17g\n %.17g\n %d\n", y, yp, time, outframenum);
    fclose(f);
    return 0;
}

int lat_restore(lat_ctx *h, const char *dir, double *time, int *outframenum)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (!dir) return -1;
    const size_t nv = c.nvol, nf = c.nfld;
    std::vector<double> aos(nf * nv), aosp(nf * nv), raw(nv);
    const std::string d(dir);
    for (size_t l = 0; l < nf; ++l)
        for (int which = 0; which < 2; ++which) {
            const std::string path = d + (which ? "/fldp_" : "/fld_") + std::to_string(l + 1) + ".raw";
            FILE *f = fopen(path.c_str(), "rb");
            if (!f) {
                c.last_error = "cannot read " + path;
                return -5;
            }
            const size_t r = fread(raw.data(), sizeof(double), nv, f);
            fclose(f);
            if (r != nv) {
                c.last_error = "short read " + path;
                return -5;
            }
            std::vector<double> &dst = which ? aosp : aos;
            for (size_t q = 0; q < nv; ++q) dst[l + nf * q] = raw[q];
        }
    FILE *f = fopen((d + "/checkpt.dat").c_str(), "r");
    if (!f) {
        c.last_error = "cannot read checkpt.dat";
        return -5;
    }
    double y = 0, yp = 0, t = 0;
    int frame = 0;
    const int got = fscanf(f, "%lf %lf %lf %d", &y, &yp, &t, &frame);
    fclose(f);
    if (got != 4) {
        c.last_error = "malformed checkpt.dat";
        return -5;
    }
    if (time) *time = t;
    if (outframenum) *outframenum = frame;
    return lat_set_state(h, aos.data(), aosp.data(), 0, y, yp);
}

int lat_device_state(lat_ctx *h, void **fld, void **fldp)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (fld) *fld = c.fld;
    if (fldp) *fldp = c.fldp;
    set_ke_valid(c, false); // the caller may write the state
    return 0;
}

int lat_sample_field(lat_ctx *h, int field, int momentum, double mean, double gamma, double m2eff, double mpl,
                     double kcut_fac, unsigned long long seed, unsigned long long stream)
{
    CTX_OR_FAIL(h);
    ONE_MEMBER(c);
    if (field < 0 || field >= c.nfld || momentum < 0 || momentum > 1 || !(mpl != 0.0)) return -1;
    int rc = ensure_io(c);
    if (rc) return rc;
    if ((rc = sample_to_io(c, gamma, m2eff, mpl, kcut_fac, seed, stream))) return rc;
    void *base = momentum ? c.fldp : c.fld;
    if (c.prec == LAT_F64) rc = store_io_field<double>(c, (double *)base + (size_t)field * c.nvol, mean);
    else rc = store_io_field<float>(c, (float *)base + (size_t)field * c.nvol, 0.0);
    if (rc) return rc;
    set_ke_valid(c, false);
    if (c.prec == LAT_MIXED) {
        // the homogeneous value goes to the FP64 part, the sample is the fluctuation
        if ((rc = pull_scalars(c))) return rc;
        (momentum ? c.host_sc.mean_pi : c.host_sc.mean_phi)[field] = mean;
        return push_scalars(c);
    }
    return sync(c);
}

int lat_profile(lat_ctx *h, int enable)
{
    CTX_OR_FAIL(h);
    int rc = prof_collect(c);
    if (rc) return rc;
    if (enable) {
        for (int s = 0; s < LAT_NSTAGE; ++s) {
            c.prof.ms[s] = 0.0;
            c.prof.cnt[s] = 0;
        }
    }
    c.prof.on = enable != 0;
    return 0;
}

int lat_profile_read(lat_ctx *h, double *ms, long long *launches)
{
    CTX_OR_FAIL(h);
    int rc = prof_collect(c);
    if (rc) return rc;
    for (int s = 0; s < LAT_NSTAGE; ++s) {
        if (ms) ms[s] = c.prof.ms[s];
        if (launches) launches[s] = c.prof.cnt[s];
    }
    return 0;
}

} // extern "C"

// ---------------------------------------------------------------- copy bandwidth
// Plain device copy with 16-byte accesses per lane: the measured HBM
// bandwidth quoted beside the vendor peak, and the calibration kernel for
// the HBM byte counters (read + write of a known size).  One element per
// lane and one block per 256 elements: on MI355X this shape copies at
// ~6.2 TB/s (read + write), against ~4.7 TB/s for a grid-stride loop over
// 8192 blocks.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" int lat_copy_bandwidth(lat_ctx *h, size_t bytes, int reps, double *gbps)
{
    CTX_OR_FAIL(h);
    if (!gbps || reps < 1 || bytes < 16) return -1;
    const size_t n = bytes / 16;
    if ((n + 255) / 256 > 0x7fffffffu) {
        c.last_error = "copy size too large for one launch";
        return -1;
    }
    const dim3 grid((unsigned)((n + 255) / 256));
    void *a = nullptr, *b = nullptr;
    LAT_HIP(&c, hipMalloc(&a, n * 16));
    if (hipMalloc(&b, n * 16) != hipSuccess) {
        (void)hipFree(a);
        c.last_error = "hipMalloc copy buffer";
        return -3;
    }
    hipEvent_t e0, e1;
    int rc = 0;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -3;
    if (!rc) {
        (void)hipMemsetAsync(a, 1, n * 16, c.stream);
        hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, c.stream, (const uint4 *)a, (uint4 *)b, n);
        (void)hipEventRecord(e0, c.stream);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_copy16, grid, dim3(256), 0, c.stream, (const uint4 *)a, (uint4 *)b, n);
        (void)hipEventRecord(e1, c.stream);
        if (hipEventSynchronize(e1) != hipSuccess) rc = -3;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        *gbps = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    if (rc) c.last_error = "copy bandwidth timing failed";
    return rc;
}

// ---------------------------------------------------------------- autotuner
// Times the two launch shapes of each sweep on the context's own arrays (the
// x R2C without drift and the y/z passes write only the k-space scratch; the
// x C2R runs its reduce-only epilogue, which writes only the partials) and
// keeps the faster.  Fields and momenta are not modified.
template <typename T>
static int autotune_t(Ctx &c, int reps)
{
    if (c.cfg.laplacian != LAT_LAP_SPECTRAL) return 0;
    const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
    const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
    const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
    cpx<T> *Fk = (cpx<T> *)c.Fk;
    hipEvent_t e0, e1;
    LAT_HIP(&c, hipEventCreate(&e0));
    LAT_HIP(&c, hipEventCreate(&e1));
    int *knobs[4] = {&c.tune.r2c, &c.tune.coly, &c.tune.colz, &c.tune.c2r};
    int rc = 0;
    for (int kind = 0; kind < 4 && !rc; ++kind) {
        float best = 0.f;
        int arg = 0;
        for (int v = 0; v < 2 && !rc; ++v) {
            *knobs[kind] = v;
            auto run = [&]() -> int {
                switch (kind) {
                case 0: return launch_x_r2c<T>(c, twx, (T *)c.fld, nullptr, Fk, 0, 0);
                case 1: return launch_col<T>(c, twy, Fk, 1, 0, c.nfld);
                case 2: return launch_col<T>(c, twz, Fk, 2, 2, c.nfld);
                default: {
                    const int g = launch_x_c2r<T>(c, twx, Fk, c.nfld, EPI_REDUCE, nullptr, (const T *)c.fld,
                                                  (T *)c.fldp);
                    return g < 0 ? g : 0;
                }
                }
            };
            if ((rc = run())) break;
            if (hipEventRecord(e0, c.stream) != hipSuccess) { rc = -3; break; }
            for (int r = 0; r < reps && !rc; ++r) rc = run();
            if (rc || hipEventRecord(e1, c.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
                rc = rc ? rc : -3;
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (v == 0 || ms < best) {
                best = ms;
                arg = v;
            }
        }
        *knobs[kind] = arg;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_ke_valid(c, false);
    return rc;
}

static int autotune(Ctx &c, int reps)
{
    return c.prec == LAT_F64 ? autotune_t<double>(c, reps) : autotune_t<float>(c, reps);
}

// ---------------------------------------------------------------- placement tuner
// How fast the sweeps stream depends on where the field, momentum and k-space
// arrays land in HBM: the x R2C of a 1024^3 mixed state, timed on four such
// allocations in one process, ran 3.07, 3.21, 3.49 and 3.55 ms per launch,
// each allocation keeping its own figure on every later call.  The tuner
// allocates further sets of these arrays (with the metric perturbations,
// their momenta and source scratch where the context has them), times one
// force evaluation's sweeps on each (one pending drift of size zero, so that
// the x R2C reads the momenta as it does in a step; kick coefficients zero),
// keeps the fastest set holding the state unchanged and frees the others.
struct Placement {
    void *fld = nullptr, *fldp = nullptr, *Fk = nullptr;
    void *gwu = nullptr, *gwp = nullptr, *gwsrc = nullptr; // contexts with the metric perturbations
};

template <typename T>
static int place_time(Ctx &c, int reps, float *ms)
{
    auto run = [&]() -> int {
        const int g = state_force<T>(c, true, EPI_KICK);
        return g < 0 ? g : 0;
    };
    hipEvent_t e0, e1;
    LAT_HIP(&c, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -3;
    }
    int rc = run();
    if (!rc && hipEventRecord(e0, c.stream) != hipSuccess) rc = -3;
    for (int r = 0; r < reps && !rc; ++r) rc = run();
    if (!rc && (hipEventRecord(e1, c.stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) rc = -3;
    if (!rc) (void)hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

static int place_tune(Ctx &c, int cand, int reps)
{
    const size_t es = elem_size(c), B = (size_t)c.B;
    const size_t bf = es * B * c.ms.f, bk = c.Fk ? 2 * es * B * c.ms.k : 0;
    const size_t bg = c.gw ? es * 6 * c.nvol : 0, bs = c.gwsrc ? es * 2 * c.nvol : 0;
    std::vector<Placement> pl(1);
    pl[0] = {c.fld, c.fldp, c.Fk, c.gwu, c.gwp, c.gwsrc};
    // Set k puts a spacer of its own size between its field and momentum
    // arrays (and between the metric perturbations and their momenta), so
    // that even where consecutive allocations are consecutive in memory each
    // set has the two at another distance (a spacer of 66 MB or 1026 MB made
    // the 1024^3 mixed x R2C 3.20 ms per launch, none or 130 MB 3.49 ms, in
    // each of several processes).
    static const size_t spacer_mb[7] = {66, 1026, 34, 258, 18, 514, 2050};
    for (int k = 1; k < cand; ++k) {
        const size_t sp = spacer_mb[(k - 1) % 7] << 20;
        // a further set only while it leaves 8 GiB of device memory free
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr < 2 * bf + bk + 2 * bg + bs + sp + (size_t(8) << 30)) {
            (void)hipGetLastError();
            break;
        }
        Placement q;
        void *spacer = nullptr, *spacer2 = nullptr;
        bool ok = hipMalloc(&q.fld, bf) == hipSuccess && hipMalloc(&spacer, sp) == hipSuccess &&
                  hipMalloc(&q.fldp, bf) == hipSuccess && (!bk || hipMalloc(&q.Fk, bk) == hipSuccess);
        if (ok && bg)
            ok = hipMalloc(&q.gwu, bg) == hipSuccess && hipMalloc(&spacer2, sp) == hipSuccess &&
                 hipMalloc(&q.gwp, bg) == hipSuccess && (!bs || hipMalloc(&q.gwsrc, bs) == hipSuccess);
        for (void *p : {spacer, spacer2})
            if (p) (void)hipFree(p);
        if (!ok) {
            (void)hipGetLastError(); // out of memory: tune over the sets allocated so far
            for (void *p : {q.fld, q.fldp, q.Fk, q.gwu, q.gwp, q.gwsrc})
                if (p) (void)hipFree(p);
            break;
        }
        pl.push_back(q);
    }
    const int K = (int)pl.size();
    int rc = 0;
    DevScalars *tdsc = nullptr;
    DevScalars *const dsc0 = c.dsc;
    void *const Fk64_0 = c.Fk64;
    const bool prof0 = c.prof.on;
    std::vector<float> t(K, 0.f);
    auto use = [&](const Placement &q) {
        c.fld = q.fld;
        c.fldp = q.fldp;
        c.Fk = q.Fk;
        if (Fk64_0 && Fk64_0 == pl[0].Fk) c.Fk64 = q.Fk; // the FP64 context's Fk64 is its Fk
        c.gwu = q.gwu;
        c.gwp = q.gwp;
        c.gwsrc = q.gwsrc;
    };
    auto copy_state = [&](const Placement &to, const Placement &from) -> int {
        if (hipMemcpyAsync(to.fld, from.fld, bf, hipMemcpyDeviceToDevice, c.stream) != hipSuccess ||
            hipMemcpyAsync(to.fldp, from.fldp, bf, hipMemcpyDeviceToDevice, c.stream) != hipSuccess)
            return -3;
        if (bg && (hipMemcpyAsync(to.gwu, from.gwu, bg, hipMemcpyDeviceToDevice, c.stream) != hipSuccess ||
                   hipMemcpyAsync(to.gwp, from.gwp, bg, hipMemcpyDeviceToDevice, c.stream) != hipSuccess))
            return -3;
        return 0;
    };
    if (K > 1) {
        // timing scalars: the context's own (per member) with one zero drift
        // pending, no re-centering and a zero kick
        std::vector<DevScalars> ts(B);
        if (hipMalloc((void **)&tdsc, sizeof(DevScalars) * B) != hipSuccess ||
            hipMemcpyAsync(ts.data(), dsc0, sizeof(DevScalars) * B, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess) {
            rc = -3;
        } else {
            for (DevScalars &m : ts) {
                m.nd = 1;
                for (double &d : m.dc) d = 0.0;
                for (int l = 0; l < 2; ++l) m.nu[l] = m.mu[l] = 0.0;
                m.hA = m.hB = 0.0;
            }
            if (hipMemcpy(tdsc, ts.data(), sizeof(DevScalars) * B, hipMemcpyHostToDevice) != hipSuccess) rc = -3;
        }
        c.prof.on = false;
        c.dsc = tdsc;
        // set 1 keeps a copy of the state while set 0 (the context's own
        // arrays) is timed; set 0 is then restored from it and set 1 timed last
        bool backed = false;
        if (!rc) backed = (rc = copy_state(pl[1], pl[0])) == 0;
        if (!rc) rc = c.prec == LAT_F64 ? place_time<double>(c, reps, &t[0]) : place_time<float>(c, reps, &t[0]);
        for (int k = 2; k < K && !rc; ++k) {
            if ((rc = copy_state(pl[k], pl[1]))) break;
            use(pl[k]);
            rc = c.prec == LAT_F64 ? place_time<double>(c, reps, &t[k]) : place_time<float>(c, reps, &t[k]);
            use(pl[0]);
        }
        use(pl[0]);
        // set 0 gets the state back from set 1 (once the copy was made),
        // whatever happened since
        if (backed) {
            const int rr = copy_state(pl[0], pl[1]);
            if (!rc) rc = rr;
        }
        if (!rc) {
            use(pl[1]);
            rc = c.prec == LAT_F64 ? place_time<double>(c, reps, &t[1]) : place_time<float>(c, reps, &t[1]);
            use(pl[0]);
        }
        int w = 0;
        for (int k = 1; k < K && !rc; ++k)
            if (t[k] < t[w]) w = k;
        if (!rc && w != 0) rc = copy_state(pl[w], pl[0]);
        if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) rc = -3;
        if (rc) w = 0; // on any failure the context keeps its own (restored) arrays
        c.dsc = dsc0;
        c.prof.on = prof0;
        use(pl[w]);
        (void)hipStreamSynchronize(c.stream);
        for (int k = 0; k < K; ++k) {
            if (k == w) continue;
            for (void *p : {pl[k].fld, pl[k].fldp, pl[k].Fk, pl[k].gwu, pl[k].gwp, pl[k].gwsrc})
                if (p) (void)hipFree(p);
        }
        c.place_ms.resize(K);
        for (int k = 0; k < K; ++k) c.place_ms[k] = t[k] / (float)reps; // per force evaluation
        c.place_pick = w;
    } else {
        c.place_ms.clear(); // no further set could be allocated
        c.place_pick = 0;
    }
    if (tdsc) (void)hipFree(tdsc);
    // the context's device scalars (and with them the cached kinetic sum)
    // were not touched: the timing ran on a copy of them
    if (rc) c.last_error = "placement tuning failed";
    return rc;
}

extern "C" {

int lat_autotune(lat_ctx *h)
{
    CTX_OR_FAIL(h);
    if (c.view) {
        c.last_error = "tune the batch, not a member view";
        return -2;
    }
    return autotune(c, 3);
}

// Placement tuner (no reference counterpart): `candidates` sets of the field,
// momentum and k-space arrays (the context's own and candidates - 1 new
// ones) timed on one force evaluation each; the fastest is kept with the
// state, the others are freed.  Single contexts only; device pointers handed
// out before (lat_device_state) are invalid afterwards.
int lat_place_tune(lat_ctx *h, int candidates, float *ms_out)
{
    CTX_OR_FAIL(h);
    if (c.view || c.B != 1) {
        c.last_error = "placement tuning takes a single context";
        return -2;
    }
    if (candidates < 1 || candidates > 8) return -1;
    if (candidates == 1) { // nothing to compare: the context keeps its arrays, untimed
        if (ms_out) ms_out[0] = -1.f;
        return 0;
    }
    const int rc = place_tune(c, candidates, 3);
    if (!rc && ms_out)
        for (int k = 0; k < candidates; ++k) ms_out[k] = k < (int)c.place_ms.size() ? c.place_ms[k] : -1.f;
    return rc ? rc : c.place_pick;
}

// 1: the field force pipeline keeps its k-space arrays in the blocked row
// order where it is compiled in (cubic 256 and 1024); 0: FFTW order (default)
int lat_set_kspace_order(lat_ctx *h, int blocked)
{
    CTX_OR_FAIL(h);
    if (blocked != 0 && blocked != 1) return -1;
    c.kblk = blocked == 1;
    return 0;
}

// 0: the five-sweep 3-D pipeline; 1: the separable three-sweep Laplacian
// for the field kicks of the mixed form (lat_sep.hip; other forms, the GW
// kick and the reductions keep the 3-D pipeline)
int lat_set_laplacian_form(lat_ctx *h, int form)
{
    CTX_OR_FAIL(h);
    if (form != 0 && form != 1) return -1;
    if (c.view) {
        c.last_error = "set the form on the batch, not a member view";
        return -2;
    }
    if (form == 1 && (c.prec != LAT_MIXED || !sep_supported(c))) {
        c.last_error = "separable Laplacian: mixed form, cubic 128, 256, 512 or 1024 only";
        return -2;
    }
    c.sep = form == 1;
    return 0;
}

int lat_laplacian_form(lat_ctx *h)
{
    CTX_OR_FAIL(h);
    return c.sep ? 1 : 0;
}

int lat_set_tuning(lat_ctx *h, const int *in)
{
    CTX_OR_FAIL(h);
    if (!in || c.view) return -1;
    for (int q = 0; q < 4; ++q)
        if (in[q] != 0 && in[q] != 1) return -1;
    c.tune.r2c = in[0];
    c.tune.coly = in[1];
    c.tune.colz = in[2];
    c.tune.c2r = in[3];
    return 0;
}

int lat_tuning(lat_ctx *h, int *out)
{
    CTX_OR_FAIL(h);
    if (!out) return -1;
    out[0] = c.tune.r2c;
    out[1] = c.tune.coly;
    out[2] = c.tune.colz;
    out[3] = c.tune.c2r;
    return 0;
}

} // extern "C"

// ---------------------------------------------------------------- ensembles
// Size classes of batches sharing one stream.  A Yoshida call over all
// classes is captured once into a hipGraph (kernels only: no memsets, no host
// synchronisation, profiling off) and replayed for every later call with the
// same order, step count and time steps.
struct lat_ensemble {
    std::vector<lat_ctx *> cls;
    std::vector<int> first; // index of each class's first member
    hipStream_t stream = nullptr;
    int device = 0;
    bool use_graph = true;
    bool single_launch = false; // one launch per sweep for all classes (set by ens_tune_launch)
    struct Entry {
        int order, nsteps;
        bool single;
        std::vector<double> dt;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Entry> graphs;
    std::string last_error;
    float tune_ms[2] = {0.f, 0.f}; // sweep timing per class / single launch
};

static int ens_fail(lat_ensemble *e, hipError_t err, const char *where)
{
    e->last_error = std::string(where) + ": " + hipGetErrorString(err);
    return -3;
}

// Lockstep form of run_ops over the size classes of an ensemble.  The op
// sequences of the classes have the same types (same integrator, order and
// step count; only the step sizes differ), so every sweep of a force
// evaluation is one launch for all classes (lat_multi.hip); the scalar,
// reduce and drift kernels stay one launch per class (one workgroup per
// member, or an element-wise sweep).  Per member the arithmetic is that of
// run_ops on the class alone.
template <typename T>
static int force_multi(std::vector<Ctx *> &cs, bool drift, std::vector<int> &npart)
{
    const int K = (int)cs.size(), nf = cs[0]->nfld;
    int rc;
    for (int l = 0; l < nf; ++l)
        if ((rc = launch_x_r2c_multi<T>(cs.data(), K, l, drift ? 1 : 0))) return rc;
    if ((rc = launch_col_multi<T>(cs.data(), K, 1, 0, nf))) return rc;
    if ((rc = launch_col_multi<T>(cs.data(), K, 2, 2, nf))) return rc;
    if ((rc = launch_col_multi<T>(cs.data(), K, 1, 1, nf))) return rc;
    return launch_x_c2r_multi<T>(cs.data(), K, nf, EPI_KICK, npart.data());
}

static int run_ops_multi(std::vector<Ctx *> &cs, const std::vector<std::vector<Op>> &ops)
{
    const size_t K = cs.size();
    const bool f64 = cs[0]->prec == LAT_F64;
    std::vector<int> npart(K, 0);
    for (size_t k = 0; k < K; ++k)
        if (!get_ke_valid(*cs[k])) {
            npart[k] = f64 ? launch_reduce_state<double>(*cs[k]) : launch_reduce_state<float>(*cs[k]);
            if (npart[k] < 0) return npart[k];
        }
    std::vector<ScalarOps> so(K);
    int ndrift = 0;
    bool kick_pending = false;
    auto scalars = [&]() -> int {
        for (size_t k = 0; k < K; ++k) {
            const int rc = launch_scalar(*cs[k], so[k], npart[k]);
            if (rc) return rc;
        }
        return 0;
    };
    auto drifts = [&]() -> int {
        for (size_t k = 0; k < K; ++k) {
            const int rc = f64 ? launch_drift<double>(*cs[k]) : launch_drift<float>(*cs[k]);
            if (rc) return rc;
        }
        return 0;
    };
    auto begin_seg = [&]() {
        for (size_t k = 0; k < K; ++k) {
            so[k] = ScalarOps{};
            if (kick_pending) {
                so[k].type[so[k].n] = OP_FINISH_KICK;
                so[k].h[so[k].n++] = 0.0;
            }
        }
    };
    begin_seg();
    for (size_t i = 0; i < ops[0].size(); ++i) {
        const int type = ops[0][i].type;
        if (type == OP_KICK) {
            for (size_t k = 0; k < K; ++k) {
                so[k].type[so[k].n] = OP_PREP_KICK;
                so[k].h[so[k].n++] = ops[k][i].h;
            }
            int rc = scalars();
            if (rc) return rc;
            rc = f64 ? force_multi<double>(cs, ndrift > 0, npart) : force_multi<float>(cs, ndrift > 0, npart);
            if (rc) return rc;
            ndrift = 0;
            kick_pending = true;
            begin_seg();
        } else {
            if (type == OP_DRIFT && ndrift == 2) {
                // more drifts than the fused sweep applies: settle them now
                int rc = scalars();
                if (rc) return rc;
                std::fill(npart.begin(), npart.end(), 0);
                kick_pending = false;
                if ((rc = drifts())) return rc;
                ndrift = 0;
                begin_seg();
            }
            if (so[0].n >= MAX_OPS - 1) {
                cs[0]->last_error = "op segment overflow";
                return -1;
            }
            for (size_t k = 0; k < K; ++k) {
                so[k].type[so[k].n] = type;
                so[k].h[so[k].n++] = ops[k][i].h;
            }
            if (type == OP_DRIFT) ++ndrift;
        }
    }
    for (size_t k = 0; k < K; ++k)
        if (so[k].n > 0 || npart[k] > 0) {
            const int rc = launch_scalar(*cs[k], so[k], npart[k]);
            if (rc) return rc;
        }
    if (ndrift > 0) {
        const int rc = drifts();
        if (rc) return rc;
    }
    for (Ctx *c : cs) set_ke_valid(*c, true);
    return 0;
}

// Classes that one launch per sweep can carry: spectral Laplacian, no tensor
// modes, cubic lattices of the sizes lat_multi.hip is built for.
static bool ens_single_ok(const lat_ensemble *e)
{
    if (e->cls.empty() || (int)e->cls.size() > MULTI_MAXC) return false;
    const Ctx &c0 = e->cls[0]->c;
    for (const lat_ctx *h : e->cls) {
        const Ctx &c = h->c;
        if (c.cfg.laplacian != LAT_LAP_SPECTRAL || c.gw || c.prec != c0.prec || c.nfld != c0.nfld) return false;
        if (!multi_supported(c.nx) || c.ny != c.nx || c.nz != c.nx) return false;
    }
    return true;
}

static bool ens_single(const lat_ensemble *e) { return e->single_launch && ens_single_ok(e); }

// The sweeps of one force evaluation over all classes, in the non-mutating
// modes the per-class autotuner times (x R2C without drift and the y/z
// passes write only the k-space scratch, the x C2R only reduces).
template <typename T>
static int ens_sweeps(lat_ensemble *e, bool single)
{
    std::vector<Ctx *> cs;
    for (lat_ctx *h : e->cls) cs.push_back(&h->c);
    const int K = (int)cs.size(), nf = cs[0]->nfld;
    int rc;
    if (single) {
        for (int l = 0; l < nf; ++l)
            if ((rc = launch_x_r2c_multi<T>(cs.data(), K, l, 0))) return rc;
        if ((rc = launch_col_multi<T>(cs.data(), K, 1, 0, nf))) return rc;
        if ((rc = launch_col_multi<T>(cs.data(), K, 2, 2, nf))) return rc;
        if ((rc = launch_col_multi<T>(cs.data(), K, 1, 1, nf))) return rc;
        std::vector<int> np(K);
        return launch_x_c2r_multi<T>(cs.data(), K, nf, EPI_REDUCE, np.data());
    }
    for (Ctx *cp : cs) {
        Ctx &c = *cp;
        const cpx<T> *twx = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twx64 : (void *)c.twx32);
        const cpx<T> *twy = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twy64 : (void *)c.twy32);
        const cpx<T> *twz = (const cpx<T> *)(sizeof(T) == 8 ? (void *)c.twz64 : (void *)c.twz32);
        cpx<T> *Fk = (cpx<T> *)c.Fk;
        for (int l = 0; l < nf; ++l)
            if ((rc = launch_x_r2c<T>(c, twx, (T *)c.fld + (size_t)l * c.nvol, nullptr, Fk + (size_t)l * c.nk, 0, l)))
                return rc;
        if ((rc = launch_col<T>(c, twy, Fk, 1, 0, nf))) return rc;
        if ((rc = launch_col<T>(c, twz, Fk, 2, 2, nf))) return rc;
        if ((rc = launch_col<T>(c, twy, Fk, 1, 1, nf))) return rc;
        const int g = launch_x_c2r<T>(c, twx, Fk, nf, EPI_REDUCE, nullptr, (const T *)c.fld, (T *)c.fldp);
        if (g < 0) return g;
    }
    return 0;
}

// Autotuner, ensemble level: times those sweeps with one launch per sweep
// for all classes and with one launch per sweep per class (shapes as tuned
// per class), keeps the faster.  One kernel has one register and LDS
// footprint, that of its largest shape, so the single launch can lose
// occupancy on the smaller classes; which way it goes depends on the sizes.
static int ens_tune_launch(lat_ensemble *e, int reps)
{
    e->single_launch = false;
    if (!ens_single_ok(e)) return 0;
    const bool f64 = e->cls[0]->c.prec == LAT_F64;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -3;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return -3;
    }
    float t[2] = {0.f, 0.f};
    int rc = 0;
    for (int m = 0; m < 2 && !rc; ++m) {
        auto run = [&]() { return f64 ? ens_sweeps<double>(e, m == 1) : ens_sweeps<float>(e, m == 1); };
        if ((rc = run())) break;
        if (hipEventRecord(e0, e->stream) != hipSuccess) {
            rc = -3;
            break;
        }
        for (int r = 0; r < reps && !rc; ++r) rc = run();
        if (rc || hipEventRecord(e1, e->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            rc = rc ? rc : -3;
            break;
        }
        (void)hipEventElapsedTime(&t[m], e0, e1);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (lat_ctx *h : e->cls) set_ke_valid(h->c, false);
    if (rc) {
        e->last_error = "launch-mode timing failed";
        return rc;
    }
    e->single_launch = t[1] < t[0];
    e->tune_ms[0] = t[0];
    e->tune_ms[1] = t[1];
    return 0;
}

// One Yoshida call on every class; each starts from fresh reductions so that
// the captured work is the same on every replay.
static int ens_record(lat_ensemble *e, int order, const double *dt, int nsteps)
{
    if (ens_single(e)) {
        std::vector<Ctx *> cs;
        std::vector<std::vector<Op>> ops(e->cls.size());
        for (size_t k = 0; k < e->cls.size(); ++k) {
            Ctx &c = e->cls[k]->c;
            if (build_symp(c, ops[k], order, dt[k], nsteps)) {
                e->last_error = "symp order must be 2, 4, 6 or 8";
                return -1;
            }
            if (ops[k].size() != ops[0].size()) {
                e->last_error = "classes with different operator sequences";
                return -1;
            }
            for (size_t i = 0; i < ops[k].size(); ++i)
                if (ops[k][i].type != ops[0][i].type) {
                    e->last_error = "classes with different operator sequences";
                    return -1;
                }
            set_ke_valid(c, false);
            cs.push_back(&c);
        }
        std::vector<bool> prof(cs.size());
        for (size_t k = 0; k < cs.size(); ++k) {
            prof[k] = cs[k]->prof.on;
            cs[k]->prof.on = false;
        }
        const int rc = run_ops_multi(cs, ops);
        for (size_t k = 0; k < cs.size(); ++k) cs[k]->prof.on = prof[k];
        if (rc) e->last_error = cs[0]->last_error.empty() ? "multi-size launch failed" : cs[0]->last_error;
        return rc;
    }
    for (size_t k = 0; k < e->cls.size(); ++k) {
        Ctx &c = e->cls[k]->c;
        std::vector<Op> ops;
        if (build_symp(c, ops, order, dt[k], nsteps)) {
            e->last_error = "symp order must be 2, 4, 6 or 8";
            return -1;
        }
        set_ke_valid(c, false);
        const bool prof = c.prof.on;
        c.prof.on = false;
        const int rc = run_ops(c, ops);
        c.prof.on = prof;
        if (rc) {
            e->last_error = c.last_error;
            return rc;
        }
    }
    return 0;
}

extern "C" {

int lat_ensemble_create(const lat_config *base, int nclass, const int *n, const int *count, const double *len,
                        const double *model_params, lat_ensemble **out)
{
    if (!base || !n || !count || !out || nclass < 1) return -1;
    *out = nullptr;
    lat_ensemble *e = new lat_ensemble();
    (void)hipGetDevice(&e->device);
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return -3;
    }
    int total = 0;
    for (int k = 0; k < nclass; ++k) {
        lat_config cfg = *base;
        cfg.nx = cfg.ny = cfg.nz = n[k];
        if (len) cfg.len = len[k];
        cfg.gw = 0;
        lat_ctx *h = nullptr;
        const int rc = create_ctx(&cfg, count[k], model_params ? model_params + 4 * (size_t)total : nullptr, &h);
        if (rc) {
            lat_ensemble_destroy(e);
            return rc;
        }
        Ctx &c = h->c;
        (void)hipStreamSynchronize(c.stream);
        (void)hipStreamDestroy(c.stream);
        c.stream = e->stream;
        c.owns_stream = false;
        // launch shapes, then the fastest of eight placements of the class's
        // arrays (before any graph or member view holds their addresses)
        if (autotune(c, 3) || place_tune(c, 8, 3)) {
            e->last_error = c.last_error;
            lat_destroy(h);
            lat_ensemble_destroy(e);
            return -3;
        }
        e->cls.push_back(h);
        e->first.push_back(total);
        total += count[k];
    }
    if (ens_tune_launch(e, 3)) {
        lat_ensemble_destroy(e);
        return -3;
    }
    *out = e;
    return 0;
}

void lat_ensemble_destroy(lat_ensemble *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto &g : e->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    for (lat_ctx *h : e->cls) lat_destroy(h);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *lat_ensemble_last_error(const lat_ensemble *e) { return e ? e->last_error.c_str() : "null ensemble"; }

int lat_ensemble_size(const lat_ensemble *e)
{
    if (!e || e->cls.empty()) return -1;
    return e->first.back() + e->cls.back()->c.B;
}

int lat_ensemble_member(lat_ensemble *e, int index, lat_ctx **view)
{
    if (!e || !view) return -1;
    for (size_t k = 0; k < e->cls.size(); ++k)
        if (index >= e->first[k] && index < e->first[k] + e->cls[k]->c.B)
            return lat_member(e->cls[k], index - e->first[k], view);
    e->last_error = "member index out of range";
    return -1;
}

int lat_ensemble_use_graph(lat_ensemble *e, int enable)
{
    if (!e) return -1;
    e->use_graph = enable != 0;
    return 0;
}

int lat_ensemble_symp(lat_ensemble *e, int order, const double *dt, int nsteps)
{
    if (!e || !dt) return -1;
    if (hipSetDevice(e->device) != hipSuccess) return -3;
    if (nsteps <= 0) return 0;
    if (!e->use_graph) return ens_record(e, order, dt, nsteps);
    const std::vector<double> key(dt, dt + e->cls.size());
    const bool single = ens_single(e);
    for (auto &g : e->graphs)
        if (g.order == order && g.nsteps == nsteps && g.dt == key && g.single == single) {
            hipError_t err = hipGraphLaunch(g.exec, e->stream);
            return err == hipSuccess ? 0 : ens_fail(e, err, "hipGraphLaunch");
        }
    hipError_t err = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
    if (err != hipSuccess) return ens_fail(e, err, "hipStreamBeginCapture");
    const int rc = ens_record(e, order, dt, nsteps);
    hipGraph_t graph = nullptr;
    err = hipStreamEndCapture(e->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (err != hipSuccess) return ens_fail(e, err, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return ens_fail(e, err, "hipGraphInstantiate");
    }
    e->graphs.push_back({order, nsteps, single, key, graph, exec});
    err = hipGraphLaunch(exec, e->stream);
    return err == hipSuccess ? 0 : ens_fail(e, err, "hipGraphLaunch");
}

// Launch shapes of class k (as lat_set_tuning / lat_tuning); changing them
// drops the captured graphs, which hold the old shapes.
int lat_ensemble_set_tuning(lat_ensemble *e, int k, const int *in)
{
    if (!e || k < 0 || k >= (int)e->cls.size()) return -1;
    const int rc = lat_set_tuning(e->cls[k], in);
    if (rc) return rc;
    (void)hipStreamSynchronize(e->stream);
    for (auto &g : e->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    e->graphs.clear();
    return 0;
}

int lat_ensemble_tuning(const lat_ensemble *e, int k, int *out)
{
    if (!e || k < 0 || k >= (int)e->cls.size()) return -1;
    return lat_tuning(e->cls[k], out);
}

int lat_ensemble_single_launch(lat_ensemble *e, int enable)
{
    if (!e) return -1;
    if (enable >= 0) e->single_launch = enable != 0;
    return ens_single(e) ? 1 : 0;
}

int lat_ensemble_synchronize(lat_ensemble *e)
{
    if (!e) return -1;
    hipError_t err = hipStreamSynchronize(e->stream);
    return err == hipSuccess ? 0 : ens_fail(e, err, "hipStreamSynchronize");
}

} // extern "C"
