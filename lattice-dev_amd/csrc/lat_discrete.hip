// lat_discrete.hip -- the Discrete tree's make_output on a finite-difference
// context (Discrete/Entropy/iofile.f90:317-335): get_variables (:432-582),
// output_reg's PSD block (:388-430 with outputfile.f90:200-216,526-735) and
// output_log_file (:617-647), with the head() header of each file
// (outputfile.f90:1216-1250).
//
// Two sweeps over the state: the first reduces, per site, ln rho and P/rho of
// the local density rho = ekin T + egrad G + epot V and pressure
// P = ekin T - egrad G / 3 - epot V (T = sum pi^2, G = STENCIL(c, GRAD2),
// V = modelv = 2V, homlat.f90:32-44) together with T, G, V and the field
// and momentum sums (FP64 block partials, summed on the host in block
// order); the second writes the two fields output_reg transforms,
// ln rho - <ln rho> and d ln rho / dt (iofile.f90:416-423), into the FP64
// staging arrays.  Both fields then go through the FP64 R2C and one binning
// pass (half space, x-Nyquist plane skipped, two-bin weights: the PSD
// columns of get_fourier_stats and crossspec_nofft).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lat_internal.h"
#include "model.h"

namespace lat {

// neighbour offsets of RANK1..3 in the order of Discrete/Entropy/macros.h:27-29
__constant__ int d_rk[26][3] = {
    {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},   {0, 0, -1}, {0, 0, 1},                          // rank 1
    {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 1, 0},   {-1, 0, -1}, {1, 0, -1},
    {-1, 0, 1},  {1, 0, 1},   {0, -1, -1}, {0, 1, -1},  {0, -1, 1},  {0, 1, 1},                          // rank 2
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}}; // rank 3

struct DiscCoef {
    double c[4];           // c0..c3 of the stencil
    double ekin, egrad, epot;
    double lap_coeff, piphi_coeff;
    double hub, a3;        // H and a^3
    double pr_over_rho;    // prsave / rhoave
    double lnrhoave_tot;   // log(rhoave) + lnrhoave = <ln rho>
};

// The site's full field and momentum values (the mixed form adds its FP64
// homogeneous parts) and, with RANK1..3 over the 26 neighbours,
//   lap[l]   = STENCIL(c, LAPLACE) of field l (c0 times the centre included),
//   grad2    = STENCIL(c, GRAD2), summed over the fields,
//   piphi[l] = STENCIL(c, PI_PHI) of field l,
// each rank summed in the macro's neighbour order.  Differences are taken of
// the stored values (fluctuations in the mixed form, where the means cancel).
template <typename T>
__device__ __forceinline__ void disc_site(const T *__restrict__ fld, const T *__restrict__ fldp, size_t nvol, int nf,
                                          int nx, int ny, int nz, int i, int j, int k, const DevScalars *dsc,
                                          const DiscCoef &cf, double *phi, double *pi, double *lap, double &grad2,
                                          double *piphi)
{
    const size_t o0 = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
    double f0[2] = {0, 0}, p0[2] = {0, 0};
    double rl[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}}, rp[4][2] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    double rg[4] = {0, 0, 0, 0};
    const bool mixed = sizeof(T) == 4;
    for (int l = 0; l < nf; ++l) {
        f0[l] = double(fld[(size_t)l * nvol + o0]);
        p0[l] = double(fldp[(size_t)l * nvol + o0]);
        phi[l] = mixed ? dsc->mean_phi[l] + f0[l] : f0[l];
        pi[l] = mixed ? dsc->mean_pi[l] + p0[l] : p0[l];
        rl[0][l] = f0[l];
    }
    for (int q = 0; q < 26; ++q) {
        const int r = q < 6 ? 1 : (q < 18 ? 2 : 3);
        if ((r == 2 && cf.c[2] == 0.0) || (r == 3 && cf.c[3] == 0.0)) continue;
        int a = i + d_rk[q][0], b = j + d_rk[q][1], cc = k + d_rk[q][2];
        a = a < 0 ? a + nx : (a >= nx ? a - nx : a);
        b = b < 0 ? b + ny : (b >= ny ? b - ny : b);
        cc = cc < 0 ? cc + nz : (cc >= nz ? cc - nz : cc);
        const size_t o = (size_t)a + (size_t)nx * ((size_t)b + (size_t)ny * cc);
        double g = 0.0;
        for (int l = 0; l < nf; ++l) {
            const double fn = double(fld[(size_t)l * nvol + o]);
            const double df = fn - f0[l];
            rl[r][l] += fn;
            g += df * df;
            rp[r][l] += (double(fldp[(size_t)l * nvol + o]) - p0[l]) * df;
        }
        rg[r] += g;
    }
    for (int l = 0; l < nf; ++l) {
        lap[l] = cf.c[0] * rl[0][l] + cf.c[1] * rl[1][l] + cf.c[2] * rl[2][l] + cf.c[3] * rl[3][l];
        piphi[l] = cf.c[0] * 0.0 + cf.c[1] * rp[1][l] + cf.c[2] * rp[2][l] + cf.c[3] * rp[3][l];
    }
    grad2 = cf.c[0] * 0.0 + cf.c[1] * rg[1] + cf.c[2] * rg[2] + cf.c[3] * rg[3];
}

constexpr int DISC_NCH = 9; // ln rho, P/rho, T, G, V, phi_1, phi_2, pi_1, pi_2

template <typename T>
__global__ __launch_bounds__(TPB) void k_disc_stats(const T *__restrict__ fld, const T *__restrict__ fldp, int nx,
                                                    int ny, int nz, int nf, DiscCoef cf, ModelDev md,
                                                    const DevScalars *__restrict__ dsc, double *__restrict__ part)
{
    md.p0 = dsc->mp[0];
    md.p1 = dsc->mp[1];
    md.p2 = dsc->mp[2];
    const size_t nvol = (size_t)nx * ny * nz;
    double acc[DISC_NCH] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t s = blockIdx.x * (size_t)TPB + threadIdx.x; s < nvol; s += (size_t)gridDim.x * TPB) {
        const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
        double phi[2] = {0, 0}, pi[2] = {0, 0}, lap[2], g, pp[2];
        disc_site(fld, fldp, nvol, nf, nx, ny, nz, i, j, k, dsc, cf, phi, pi, lap, g, pp);
        double Tk = 0.0;
        for (int l = 0; l < nf; ++l) Tk += pi[l] * pi[l];
        const double V = 2.0 * model_v64(md, phi);
        const double rho = cf.ekin * Tk + cf.egrad * g + cf.epot * V;
        const double prs = cf.ekin * Tk - (1.0 / 3.0) * cf.egrad * g - cf.epot * V;
        acc[0] += log(rho);
        acc[1] += prs / rho;
        acc[2] += Tk;
        acc[3] += g;
        acc[4] += V;
        acc[5] += phi[0];
        acc[6] += phi[1];
        acc[7] += pi[0];
        acc[8] += pi[1];
    }
    __shared__ double red[DISC_NCH][TPB / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < DISC_NCH; ++c) {
        double x = acc[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) red[c][wid] = x;
    }
    __syncthreads();
    if (threadIdx.x < DISC_NCH) {
        double sum = 0.0;
        for (int w = 0; w < TPB / 64; ++w) sum += red[threadIdx.x][w];
        part[(size_t)blockIdx.x * DISC_NCH + threadIdx.x] = sum;
    }
}

template <typename T>
__global__ __launch_bounds__(TPB) void k_disc_fields(const T *__restrict__ fld, const T *__restrict__ fldp, int nx,
                                                     int ny, int nz, int nf, DiscCoef cf, ModelDev md,
                                                     const DevScalars *__restrict__ dsc, double *__restrict__ lnrho,
                                                     double *__restrict__ dlnrho)
{
    md.p0 = dsc->mp[0];
    md.p1 = dsc->mp[1];
    md.p2 = dsc->mp[2];
    const size_t nvol = (size_t)nx * ny * nz;
    for (size_t s = blockIdx.x * (size_t)TPB + threadIdx.x; s < nvol; s += (size_t)gridDim.x * TPB) {
        const int i = (int)(s % nx), j = (int)((s / nx) % ny), k = (int)(s / ((size_t)nx * ny));
        double phi[2] = {0, 0}, pi[2] = {0, 0}, lap[2], g, pp[2];
        disc_site(fld, fldp, nvol, nf, nx, ny, nz, i, j, k, dsc, cf, phi, pi, lap, g, pp);
        double Tk = 0.0;
        for (int l = 0; l < nf; ++l) Tk += pi[l] * pi[l];
        const double V = 2.0 * model_v64(md, phi);
        const double rho = cf.ekin * Tk + cf.egrad * g + cf.epot * V;
        const double prs = cf.ekin * Tk - (1.0 / 3.0) * cf.egrad * g - cf.epot * V;
        double num = 0.0;
        for (int l = 0; l < nf; ++l) num += pi[l] * (cf.lap_coeff * lap[l]);
        for (int l = 0; l < nf; ++l) num += cf.piphi_coeff * pp[l];
        lnrho[s] = log(rho) - cf.lnrhoave_tot;
        dlnrho[s] = -3.0 * cf.hub * (prs / rho - cf.pr_over_rho) + num / rho / cf.a3;
    }
}

// ---------------------------------------------------------------- Fortran edit descriptors
// Gw.dEe output editing (Fortran 2008 10.7.5.2.2), e = 2: a nonzero value
// whose magnitude rounded to d significant digits r satisfies
// 0.1 <= r < 10^d is written as F(w-4).(d-k) followed by 4 blanks, with
// 10^(k-1) <= r < 10^k; zero as F(w-4).(d-1) and 4 blanks; anything else as
// Ew.dEe, a field of asterisks when the exponent needs more than e digits.
static std::string fortran_g(double x, int w, int d)
{
    char buf[128];
    if (std::isnan(x) || std::isinf(x)) {
        snprintf(buf, sizeof buf, "%*s", w, std::isnan(x) ? "NaN" : (x > 0 ? "Infinity" : "-Infinity"));
        return buf;
    }
    if (x == 0.0) { // the sign of a negative zero is kept
        snprintf(buf, sizeof buf, "%*.*f%4s", w - 4, d - 1, x, "");
        return buf;
    }
    snprintf(buf, sizeof buf, "%.*e", d - 1, std::fabs(x));
    const int X = atoi(strchr(buf, 'e') + 1); // decimal exponent of the rounded magnitude
    if (X >= -1 && X <= d - 1) {
        const int k = X + 1; // '#': the decimal point stays when no digits follow it
        snprintf(buf, sizeof buf, "%#*.*f%4s", w - 4, d - k, x, "");
        return buf;
    }
    // E editing: 0.<d digits>E+ee, the digits those of the rounded value
    const int ex = X + 1;
    if (ex > 99 || ex < -99) return std::string(w, '*');
    char mant[64];
    snprintf(mant, sizeof mant, "%.*e", d - 1, std::fabs(x));
    std::string digits;
    for (const char *p = mant; *p && *p != 'e'; ++p)
        if (*p >= '0' && *p <= '9') digits += *p;
    snprintf(buf, sizeof buf, "%s0.%sE%c%02d", x < 0 ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    std::string f(buf);
    if ((int)f.size() < w) f = std::string(w - f.size(), ' ') + f;
    return f;
}

// '(32g25.16e2)' record of vals (outputfile.f90:21 outformat; iofile.f90:643)
static std::string g25_record(const std::vector<double> &vals)
{
    std::string line;
    for (size_t q = 0; q < vals.size(); ++q) {
        if (q && q % 32 == 0) line += "\n";
        line += fortran_g(vals[q], 25, 16);
    }
    return line + "\n";
}

// Fw.d with w = 0 (minimal width), as modelsummary's f0.5
static std::string fortran_f0(double x, int d)
{
    char buf[64];
    snprintf(buf, sizeof buf, "%.*f", d, x);
    return buf;
}

struct RunInfo {
    int nfld, n;
    double len, dt, mpl;
    int seed;
    double m2phi, m2psi, g2;
};

// head() (outputfile.f90:1216-1250) for one process without OpenMP (the
// reference's Discrete build, Discrete/Entropy/Makefile:30): revision line,
// rank and decomposition lines (ydivs = zdivs = 1, every neighbour rank 0,
// latpar.f90:36-66), modelsummary (model.f90:23-31) and the variable list.
static std::string head(const RunInfo &ri, const std::vector<std::string> &vars)
{
    std::string h = "# This is DEFROST revision 1.0 (" + std::to_string(ri.nfld) + " fields, " + std::to_string(ri.n) +
                    "^3 grid) L=" + fortran_g(ri.len, 32, 15) + " dt = " + fortran_g(ri.dt, 32, 15) +
                    " mpl = " + fortran_g(ri.mpl, 32, 15) + " seed = " + std::to_string(ri.seed) + "\n";
    h += "# MPI rank: 0; total nodes: 1\n";
    h += "# Using 1 y division(s) and 1 z division(s)\n";
    h += "# Local y division: 0; local z division: 0\n";
    h += "# neighbor ranks: 0 <- y -> 0\n";
    h += "# neighbor ranks: 0 <- z -> 0\n";
    h += "# V(phi,psi) = " + fortran_f0(ri.m2phi, 5) + "phi^2/2 + " + fortran_f0(ri.m2psi, 5) + "psi^2/2 + " +
         fortran_f0(ri.g2, 5) + "phi^2*phi^2/2\n";
    // write (buffer,'(a,g12.12",",32(g24.12","))') "OUTPUT:", adjustr(vars):
    // "OUTPUT:" as it is (a), the first name (adjustr to 12) in 12 columns,
    // every later one right-justified in 24, each name followed by ","
    auto rjust = [](const std::string &s, size_t w) { return s.size() >= w ? s : std::string(w - s.size(), ' ') + s; };
    // character(len=12) of the name (truncated), then adjustr: trailing blanks to the front
    auto adjustr12 = [&](std::string v) {
        v = v.substr(0, 12);
        v.erase(v.find_last_not_of(' ') + 1);
        return rjust(v, 12);
    };
    std::string buf = "OUTPUT:";
    for (size_t q = 0; q < vars.size(); ++q) buf += rjust(adjustr12(vars[q]), q == 0 ? 12 : 24) + ",";
    const size_t l = buf.rfind(',') + 1; // index(buffer, ',', .true.)
    h += "# " + std::string(l, '=') + "\n";
    h += "# " + buf.substr(0, l - 1) + ";\n";
    h += "# " + std::string(l, '=') + "\n";
    return h;
}

static int append_text(Ctx &c, const std::string &path, const std::string &hdr, const std::string &body)
{
    FILE *test = fopen(path.c_str(), "r");
    const bool fresh = test == nullptr;
    if (test) fclose(test);
    FILE *f = fopen(path.c_str(), "a");
    if (!f) {
        c.last_error = "cannot open " + path;
        return -5;
    }
    if (fresh) fputs(hdr.c_str(), f);
    fputs(body.c_str(), f);
    fclose(f);
    return 0;
}

template <typename T>
static int discrete_output_t(Ctx &c, double time, double dt, double mpl, int seed, const char *dir, double *vals,
                             double *psd_out)
{
    int rc;
    const double *cs = c.cfg.stencil;
    const DevScalars &hs = c.host_sc;
    const double y = hs.yscl, p = hs.ysclp;
    const bool conf = c.cfg.time_var == LAT_CONFORMAL;
    // get_scale_factor / get_hubble, scafld.f90:56-75
    const double a = conf ? y : std::pow(y, 2.0 / 3.0);
    const double hub = conf ? -p / y / 6.0 / y : -p / y / 4.0;
    DiscCoef cf{};
    for (int q = 0; q < 4; ++q) cf.c[q] = cs[q];
    cf.ekin = 0.5 / std::pow(a, 6);
    cf.egrad = 0.25 / ((a * c.dx) * (a * c.dx)) / cs[4];
    cf.epot = 0.5;
    cf.lap_coeff = 1.0 / ((a * c.dx) * (a * c.dx)) / cs[4];
    cf.piphi_coeff = 0.5 / ((a * c.dx) * (a * c.dx)) / cs[4];
    cf.hub = hub;
    cf.a3 = a * a * a;
    ModelDev md{c.cfg.model, c.nfld, c.cfg.model_params[0], c.cfg.model_params[1], 0.0, c.cfg.model_params[2]};
    // pass 1: block partials of the per-site sums
    const size_t blocks = (c.nvol + TPB - 1) / TPB;
    const int g = grid_for(blocks > (size_t)MAX_GRID ? MAX_GRID : (int)blocks);
    if ((rc = ensure_disc(c))) return rc;
    hipLaunchKernelGGL((k_disc_stats<T>), dim3(g), dim3(TPB), 0, c.stream, (const T *)c.fld, (const T *)c.fldp, c.nx,
                       c.ny, c.nz, c.nfld, cf, md, c.dsc, c.disc_part);
    LAT_HIP(&c, hipGetLastError());
    std::vector<double> part((size_t)g * DISC_NCH);
    LAT_HIP(&c, hipMemcpyAsync(part.data(), c.disc_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                               c.stream));
    LAT_HIP(&c, hipStreamSynchronize(c.stream));
    double sum[DISC_NCH] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < g; ++b)
        for (int q = 0; q < DISC_NCH; ++q) sum[q] += part[(size_t)b * DISC_NCH + q];
    const double n3 = (double)c.nvol;
    // iofile.f90:554-580
    const double GE = cf.egrad * sum[3] / n3, PE = cf.epot * sum[4] / n3, KE = cf.ekin * sum[2] / n3;
    const double eosave = sum[1] / n3, lnrtmp = sum[0] / n3;
    const double rhoave = GE + PE + KE, prsave = KE - PE - (1.0 / 3.0) * GE;
    const double lnrhoave = lnrtmp - std::log(rhoave);
    const double grav = -3.0 * hub * hub;
    std::vector<double> line = {time, a, hub, KE + GE + PE, prsave, lnrhoave, eosave, KE, PE, GE, grav,
                                GE + KE + PE + grav};
    for (int l = 0; l < c.nfld; ++l) line.push_back(sum[5 + l] / n3);
    for (int l = 0; l < c.nfld; ++l) line.push_back(sum[7 + l] / n3 / (a * a * a));
    if (vals)
        for (size_t q = 0; q < line.size(); ++q) vals[q] = line[q];
    // pass 2: the fields of output_reg into the FP64 staging arrays
    if ((rc = ensure_io(c))) return rc;
    cf.pr_over_rho = prsave / rhoave;
    cf.lnrhoave_tot = std::log(rhoave) + lnrhoave;
    double *lnr = c.io64, *dln = c.io64 + c.nvol;
    hipLaunchKernelGGL((k_disc_fields<T>), dim3(g), dim3(TPB), 0, c.stream, (const T *)c.fld, (const T *)c.fldp, c.nx,
                       c.ny, c.nz, c.nfld, cf, md, c.dsc, lnr, dln);
    LAT_HIP(&c, hipGetLastError());
    // FP64 R2C of both, then one binning pass: channels |F(lnrho)|^2,
    // |F(dlnrho)|^2, Re and Im of F(dlnrho) conj F(lnrho), W
    cpx<double> *F = (cpx<double> *)c.Fk64;
    for (int q = 0; q < 2; ++q)
        if ((rc = launch_x_r2c<double>(c, c.twx64, c.io64 + (size_t)q * c.nvol, nullptr, F + (size_t)q * c.nk, 0)))
            return rc;
    if ((rc = launch_col<double>(c, c.twy64, F, 1, 0, 2))) return rc;
    if ((rc = launch_col<double>(c, c.twz64, F, 2, 0, 2))) return rc;
    if ((rc = ensure_spec(c))) return rc;
    const int ns = lat_num_bins_discrete(c.nx);
    if ((rc = launch_bin<double>(c, F + c.nk, F, 2, ns, BIN_HALF | BIN_NO_NYQ)) < 0) return rc;
    std::vector<double> raw(5 * (size_t)ns);
    LAT_HIP(&c, hipMemcpyAsync(raw.data(), c.spec_out, sizeof(double) * raw.size(), hipMemcpyDeviceToHost,
                               c.stream));
    LAT_HIP(&c, hipStreamSynchronize(c.stream));
    const double *W = raw.data() + 4 * (size_t)ns;
    std::vector<double> psdv(4 * (size_t)ns);
    for (int ch = 0; ch < 4; ++ch)
        for (int b = 0; b < ns; ++b) {
            double s = raw[(size_t)ch * ns + b];
            if (W[b] != 0.0) s = s / W[b]; // get_fourier_stats: S / W, no n^6
            psdv[(size_t)ch * ns + b] = s;
        }
    if (psd_out)
        for (size_t q = 0; q < psdv.size(); ++q) psd_out[q] = psdv[q];
    if (!dir) return 0;
    // the files (outputfile.f90:200-216: t, a, H, (k-1) dk, PSD(k, :); two
    // empty records after each block)
    RunInfo ri{c.nfld, c.nx, c.cfg.len, dt, mpl, seed, c.cfg.model_params[0], c.cfg.model_params[2],
               c.cfg.model_params[1]};
    const std::string d(dir);
    std::string psd;
    for (int b = 0; b < ns; ++b) {
        std::vector<double> row = {time, a, hub, (double)b * c.dk};
        for (int ch = 0; ch < 4; ++ch) row.push_back(psdv[(size_t)ch * ns + b]);
        psd += g25_record(row);
    }
    psd += "\n\n";
    if ((rc = append_text(c, d + "/PSD", head(ri, {"t  ", "a  ", "H   ", "k   ", "logrho_ave", "dlnrho", "lnr_dlnr",
                                                   "lnr_dlnr"}), psd)))
        return rc;
    return append_text(c, d + "/LOG.out",
                       head(ri, {"t    ", "a    ", "H    ", "<rho>", "<P>  ", "<lnrho>", "<EOS>", "<E_KIN>", "E_POT>",
                                 "<E_grad>", "GRAVEN", "Ham"}),
                       g25_record(line));
}

int discrete_output(Ctx &c, double time, double dt, double mpl, int seed, const char *dir, double *vals,
                    double *psd)
{
    return c.prec == LAT_F64 ? discrete_output_t<double>(c, time, dt, mpl, seed, dir, vals, psd)
                             : discrete_output_t<float>(c, time, dt, mpl, seed, dir, vals, psd);
}

} // namespace lat
