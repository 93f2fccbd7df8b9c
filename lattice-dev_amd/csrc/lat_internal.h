// lat_internal.h -- shared definitions of the HIP library (not a public header).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/lattice_c.h"
#include "fft_core.h"

namespace lat {

constexpr int TPB = 256;       // threads per workgroup for all lattice kernels
constexpr int MAX_GRID = 8192; // most reduction partials (blocks) per member of a launch
constexpr int GRID_CAP = 4096; // default cap of grid-stride launches
constexpr int NRED = 8;        // reduction channels per partial

// Reduction channels written by the sweep kernels (FP64 partial per block).
// FP64 form: sums of the full fields.  Mixed form: sums of the fluctuations
// (GE uses dphi, PE is the sum of V(phi) - V(mean)).
enum RedChan {
    RC_GE = 0,   // sum over fields and sites of -phi * lap_raw (lap without a factors)
    RC_PE = 1,   // sum of V(phi) (mixed form: sum of V(phi) - V(mean))
    RC_KE = 2,   // sum over fields of pi^2 (mixed form: of dpi^2)
    RC_S1 = 3,   // sum of pi (dpi) field 0
    RC_S2 = 4,   // sum of pi (dpi) field 1
    RC_M1 = 5,   // sum of phi (dphi) field 0
    RC_M2 = 6,   // sum of phi (dphi) field 1
    RC_X = 7,    // auxiliary (FD: GRAD2 stencil sum)
};

// Background (FRW) scalars and step coefficients, resident in device memory.
struct DevScalars {
    double yscl, ysclp;       // a (conformal) or y = a^{3/2} (cosmic), and its momentum
    double red[NRED];         // last completed reduction (sums, unnormalised)
    double kick_h;            // h of the kick whose p-update is pending
    double hA, hB;            // next kick: pi += hA*lap_raw - hB*V'
    double dc[4];             // pending drift coefficients phi += dc[q]*pi, in order
    int nd;                   // number of pending drifts
    int mixed;                // state holds FP32 fluctuations about the FP64 means below
    // Mixed form: phi_l(x) = mean_phi[l] + dphi_l(x), pi_l(x) = mean_pi[l] + dpi_l(x).
    double mean_phi[2];
    double mean_pi[2];
    double nu[2];             // re-centering: the next drift sweep subtracts nu from dphi
    double mu[2];             // re-centering: the next kick subtracts mu from dpi
    double sdphi[2];          // sum of dphi after the pending drifts (before nu)
    double mp[3];             // model parameters of this member (lat_config.model_params[0..2])
};

// Member strides of a batch context (B realisations of one lattice size,
// stored member-major): kernels launched with gridDim.y = B offset their
// arrays by blockIdx.y times these element counts.
struct MemberStride {
    size_t f;  // fld / fldp elements per member (nfld * nvol)
    size_t k;  // Fk complex elements per member
};

enum ScalarOpType {
    OP_NONE = 0,
    OP_FINISH_KICK = 1, // p -= ... using red[], kick_h
    OP_DRIFT = 2,       // record drift coefficient, p += h K2 / y^3
    OP_GRAV = 3,        // y -= ...
    OP_PREP_KICK = 4,   // hA, hB for a kick of size h; kick_h = h
    OP_CLEAR_DRIFT = 5, // nd = 0
};

constexpr int MAX_OPS = 12;
struct ScalarOps {
    int n;
    int type[MAX_OPS];
    double h[MAX_OPS];
};

struct ModelDev {
    int model;  // LAT_MODEL_*
    int nfld;
    double p0, p1; // phi4: lambda; twofield: m2, g2
    double gwc;    // u_ij source coefficient 2 / M_p^2 (EPI_GWKICK only)
    double p2;     // twofield: m_chi^2 (0 unless lat_config.model_params[2] sets it)
};

// Per-stage HIP-event accounting (lat_profile / lat_profile_read).
struct Prof {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, int>> used; // (stage, first event index of the pair)
    double ms[LAT_NSTAGE] = {};
    long long cnt[LAT_NSTAGE] = {};
};

struct Ctx {
    lat_config cfg;
    Prof prof;
    int B = 1;              // members (realisations) of a batch context
    MemberStride ms{0, 0};
    bool view = false;      // a member view of a batch: borrows every buffer
    // scratch that per-member operations allocate on demand while running on
    // a view (e.g. the k-space arrays of the output spectra when the batch has
    // none): owned by the view and freed by lat_destroy on it
    std::vector<void *> owned;
    Ctx *parent = nullptr;  // the batch a view belongs to
    int member = 0;         // index of a view's member in its batch
    bool owns_stream = true; // false when an ensemble shares its stream
    struct Tune {            // launch-shape variants picked by lat_autotune
        int r2c = 0, coly = 0, colz = 0, c2r = 0;
    } tune;
    std::vector<char> mvalid; // batch: per-member ke_valid
    int nx, ny, nz, nxh, P; // P = pitch (complex elements) of a half-complex row
    size_t nvol, nk;        // nk = P*ny*nz
    double dx, dk;
    int prec;               // LAT_F64 or LAT_MIXED
    int nfld;
    hipStream_t stream;
    int device;

    void *fld, *fldp;       // nfld * nvol of T (state)
    void *Fk;               // nfld * nk of cpx<T>
    void *Fk2;              // nk of cpx<T>, spectra second operand
    double *io64;           // nvol doubles: host-interface staging (FP64)
    void *Fk64;             // nk of cpx<double>: FP64 drop-in transforms
    // metric perturbations (cfg.gw): 6 components xx,xy,xz,yy,yz,zz of T each
    bool gw;
    double gw_coup;         // 2 / M_p^2
    void *gwu, *gwp;        // 6 * nvol of T
    void *gwsrc;            // 2 * nvol of T: S_ij of the component pair being kicked
    void *gwk;              // 6 * nk of cpx<T>: transforms for the GW spectra
    // separable form of the spectral Laplacian (lat_sep.hip): lap holds the
    // running sum D_xx + D_yy of every field between the x and z sweeps
    bool sep = false;
    void *lap = nullptr;    // nfld * nvol of T per member
    DevScalars *dsc;
    double *partials;       // MAX_GRID * NRED
    double *spec_part;      // spectrum block partials
    double *disc_part = nullptr; // Discrete-tree output: block partials of the per-site sums
    double *spec_out;       // 5 * spec_cap reduced channels, then the W cache slots
    int ns;
    int spec_cap;           // bins the histogram buffers hold
    int cus = 0;            // compute units of the device (sizes the binning grid)
    bool kblk = false;      // force pipeline in the blocked k-space order where compiled in (kblk_active)
    std::vector<float> place_ms; // placement tuner: ms per force evaluation per candidate set (0: the arrays it had)
    int place_pick = 0;          // placement tuner: the candidate kept
    // The weight channel W(b) of the binning depends on the lattice geometry,
    // the bin count and the weighting flags only, so launch_bin computes it
    // once per (bin count, flags) and keeps it in one of WCACHE slots of
    // spec_cap doubles after the 5 * spec_cap of spec_out.  A view shares its
    // batch's spec_out and therefore uses its batch's keys.
    static constexpr int WCACHE = 4;
    long long wkey[WCACHE] = {-1, -1, -1, -1};
    int wnext = 0;

    // twiddle tables w[m] = exp(-2 pi i m / n) for n = nx, ny, nz
    cpx<double> *twx64, *twy64, *twz64;
    cpx<float> *twx32, *twy32, *twz32;

    // host mirror (valid after sync_scalars_to_host)
    DevScalars host_sc;
    bool ke_valid; // red[RC_KE] holds sum pi^2 of the current state
    std::string last_error;
};

size_t elem_size(const Ctx &c);

// ---- FFT stage launchers (lat_fft.hip). T = double or float. ----
// x-direction R2C of each real row of `f` (nvol), into half-complex rows of Fk.
// If drift != 0, first applies the pending drifts of dsc to f in place using pi
// (field >= 0: a field, re-centred in the mixed form; field < 0: a tensor
// component, never re-centred).
// layout 1 (x R2C, column passes, x C2R): the blocked k-space row order of the
// force pipeline (sweeps.h), for cubic lattices of a size with it compiled in.
template <typename T>
int launch_x_r2c(Ctx &c, const cpx<T> *tw, T *f, const T *pi, cpx<T> *Fk, int drift, int field = 0, int layout = 0);
// Column FFT along y (axis 1) or z (axis 2) of a half-complex array.
// mode 0: forward; 1: inverse; 2: forward, multiply by -k^2 dk^2/nvol, inverse.
template <typename T>
int launch_col(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int axis, int mode, int n_arrays, int layout = 0);
// x-direction C2R of nfld half-complex arrays followed by an epilogue:
//   EPI_STORE:  write the real rows into out (one array per field)
//   EPI_KICK:   kick pi with the device coefficients; reduce GE, PE, KE'
//   EPI_REDUCE: reduce GE, PE, KE (no state change)
//   EPI_GWKICK: pi_u += hA (lap + gw_coup S) for nf tensor components, with
//               fld = S and fldp = pi_u of those components; no reductions
enum { EPI_STORE = 0, EPI_KICK = 1, EPI_REDUCE = 2, EPI_GWKICK = 3 };
template <typename T>
int launch_x_c2r(Ctx &c, const cpx<T> *tw, cpx<T> *Fk, int nf, int epi, T *out, const T *fld, T *fldp,
                 int layout = 0);

// ---- separable spectral Laplacian (lat_sep.hip), mixed form ----
// Kick of all fields with the three-sweep Laplacian (x: drift + D_xx; y:
// += D_yy; z: + D_zz, kick, reductions).  *npart = reduction partials per
// member.  c.lap must hold nfld * nvol floats per member.
bool sep_supported(const Ctx &c);
template <typename T>
int launch_sep_kick(Ctx &c, int drift, int *npart);

// ---- one launch per sweep over the size classes of an ensemble (lat_multi.hip) ----
// cs[0..ncls) are batches of cubic lattices of the sizes multi_supported()
// accepts, sharing cs[0]'s stream; shapes follow each class's tuning.
// Returns 0, or <0 (unsupported class set / HIP error).
constexpr int MULTI_MAXC = 4;
bool multi_supported(int n);
template <typename T>
int launch_x_r2c_multi(Ctx *const *cs, int ncls, int field, int drift);
template <typename T>
int launch_col_multi(Ctx *const *cs, int ncls, int axis, int mode, int n_arrays);
// EPI_KICK or EPI_REDUCE; npart[k] = reduction partials per member of class k
template <typename T>
int launch_x_c2r_multi(Ctx *const *cs, int ncls, int nf, int epi, int *npart);

// ---- step kernels (lat_step.hip) ----
template <typename T>
int launch_drift(Ctx &c);              // apply pending drifts of dsc to all fields (and u_ij)
template <typename T>
int launch_gw_source(Ctx &c, int c0, int nc, T *dst); // S_ij components c0..c0+nc-1
template <typename T>
int launch_gw_stencil_kick(Ctx &c);    // FD: pi_u += hA (stencil lap u + gw_coup S)
template <typename T>
int launch_reduce_state(Ctx &c);        // partials: PE, KE, means (no Laplacian)
template <typename T>
int launch_stencil(Ctx &c, int epi);    // FD Laplacian + kick / reduce epilogue
int launch_scalar(Ctx &c, const ScalarOps &ops, int npart);
int grid_for(int tiles, int cap = GRID_CAP);

// ---- spectra (lat_spec.hip) ----
// mode 0: |F1|^2; 1: F1 conj(F2); 2: output channels; 3: sum_ij |Lambda U|^2 of
// the six tensor components F1 + c*nk (TT projection).  nbins <= 0: c.ns.
enum { BIN_COUNTS = 1, BIN_HALF = 2, BIN_NO_NYQ = 4 }; // launch_bin flags
// spectrum binning: workgroups (one wave each) and the group size of the
// first stage of the two-stage block reduction; spec_part holds
// BIN_BLOCKS + BIN_BLOCKS / BIN_GROUP histograms
constexpr int BIN_BLOCKS = 4096, BIN_GROUP = 64;
template <typename T>
int launch_bin(Ctx &c, const cpx<T> *F1, const cpx<T> *F2, int mode, int nbins = 0, int flags = 0);
// allocation-free Discrete-tree nofft spectrum (lat_spectrum_discrete_k_noalloc):
// per-device scratch lookup at context creation, then the launch
int noalloc_setup(Ctx &c);
int spectrum_discrete_k_noalloc(Ctx &c, const double *Fk_dev, double *S_dev, int kind);

// ---- Discrete-tree make_output (lat_discrete.hip) ----
// LOG.out line values into vals and the PSD columns into psd (both
// nullable) and, with dir, the LOG.out line and the PSD block appended to
// <dir>/LOG.out and <dir>/PSD
int discrete_output(Ctx &c, double time, double dt, double mpl, int seed, const char *dir, double *vals,
                    double *psd);
// scratch on demand (lat_api.hip): FP64 staging (io64, Fk64), spectrum
// histograms, Discrete-output partials
int ensure_io(Ctx &c);
int ensure_spec(Ctx &c);
int ensure_disc(Ctx &c);

// ---- initial conditions (lat_ic.hip) ----
int sample_to_io(Ctx &c, double gamma, double m2eff, double mpl, double kcut_fac, uint64_t seed, uint64_t stream);
template <typename T>
int store_io_field(Ctx &c, T *dst, double add);

// ---- profiling ----
int prof_begin(Ctx &c, int stage); // returns a pair index or -1
void prof_end(Ctx &c, int idx);

// ---- helpers ----
int hip_fail(Ctx *c, hipError_t e, const char *where);
#define LAT_HIP(ctx, call)                                                     \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #call);               \
    } while (0)

} // namespace lat

struct lat_ctx {
    lat::Ctx c;
};
