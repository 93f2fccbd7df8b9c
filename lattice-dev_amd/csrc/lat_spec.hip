// lat_spec.hip -- shell-binned power and cross spectra on gfx950.
//
// Restates the binning loop of spectrum_3d / crossspec_3d
// (Pseudospec/openmp_dev/analysis_spec.f90:84-101,157-177): for every mode,
// p = |k|/dk, l = floor(p), the two-bin weights c = (1-(l-p)^2)^2 and
// (1-(l+1-p)^2)^2 go to bins l and l+1 (0-based).  The reference loops over
// all nx x-indices and re-reads Fk(nx-i) for i > nx/2; on the half-complex
// array that is a multiplicity of 2 for 0 < kx < nx/2.
//
// Work unit: a row class (|ky|, |kz|).  The up to four rows (ky, kz),
// (-ky, kz), (ky, -kz), (-ky, -kz) share every mode's |k|, bin and weights,
// so their values are summed first (in that order) and the binning runs once
// per class.  One wave per workgroup walks its classes; each lane takes 8
// consecutive kx (at 1024^3; ceil((nx/2)/64) in general), and the lane that
// reaches the row's last mode takes that one too.  Along a row p is
// non-decreasing in kx, so a lane sums runs of equal bin in registers and
// adds every run that starts and ends inside it to the wave's own LDS
// histogram; the runs that cross lanes are joined by one segmented DPP scan
// per class.  Within one histogram instruction the lanes' bins are distinct,
// and the instructions run in program order (no atomics, no cross-wave
// ordering; the barriers fence LDS only, so the next class's loads stay in
// flight across them).  The block histograms are reduced in block order by
// k_bin_reduce (groups of 64 blocks, then the groups).  Results are bitwise
// reproducible run to run.
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
constexpr int BIN_LANEMODES = 9;          // modes per lane and row: 8 + the row's last (nx <= 1024)
constexpr size_t BIN_LDS_MAX = 64 * 1024;  // bytes of one block's LDS histogram

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
__global__ __launch_bounds__(64) void k_bin(const cpx<T> *__restrict__ F1, const cpx<T> *__restrict__ F2, int nx,
                                            int ny, int nz, int nxh, int P, int ns, double *__restrict__ part,
                                            int flags, size_t nk)
{
    constexpr int NC = NCH + (WCH ? 1 : 0); // binned channels
    constexpr int NV = NCH > 0 ? NCH : 1;
    constexpr int NOKEY = 0x7fffffff;
    constexpr int LM = BIN_LANEMODES;
    constexpr int LA = MODE <= 2 ? LM : 1;
    // the next class's loads are issued before the current class is binned
    // (FP32 data: two register sets fit; FP64 data loads each class when its
    // turn comes)
    constexpr bool PF = MODE <= 2 && sizeof(T) == 4;
    extern __shared__ double hist[]; // NC * ns
    const int lane = threadIdx.x;
    for (int q = lane; q < NC * ns; q += 64) hist[q] = 0.0;
    lds_sync();
    const bool counts = (flags & BIN_COUNTS) != 0;
    // the lane's modes [s, e) of a row: M consecutive ones, and the row's
    // last mode for the lane whose range reaches it
    const int M = nxh > 1 ? (nxh - 1 + 63) / 64 : 1;
    const int s = lane * M;
    int e = s + M;
    if (e >= nxh - 1) e = nxh;
    if (s > 0 && s >= nxh - 1) e = s;
    const int nyq = ((flags & BIN_NO_NYQ) && nx % 2 == 0) ? nx / 2 : -1;
    const int ncy = ny / 2 + 1, ncz = nz / 2 + 1;
    const int ncls = ncy * ncz;
    const int G = gridDim.x;

    // the row class (jj, kk): rows (j, k) with j in {jj, ny - jj} and k in
    // {kk, nz - kk}, each member once
    struct Cls {
        int jj, kk, nm, mj[4], mk[4];
    };
    auto members = [&](int q) {
        Cls c;
        c.jj = q % ncy;
        c.kk = q / ncy;
        const bool jm = c.jj > 0 && 2 * c.jj != ny, km = c.kk > 0 && 2 * c.kk != nz;
        c.nm = 1 + (jm ? 1 : 0) + (km ? 1 : 0) + (jm && km ? 1 : 0);
        c.mj[0] = c.jj; c.mk[0] = c.kk;
        c.mj[1] = jm ? ny - c.jj : c.jj; c.mk[1] = jm ? c.kk : nz - c.kk;
        c.mj[2] = c.jj; c.mk[2] = nz - c.kk;
        c.mj[3] = ny - c.jj; c.mk[3] = nz - c.kk;
        return c;
    };
    // The class's values, modes sb .. sb + M of each member row: the lane's
    // own (sb = s) or, for a lane without modes, the row's first ones;
    // absent members read member 0.  Every load is in bounds and
    // unpredicated; values of modes or members outside the class are never
    // used.
    const int sb = e > s ? s : 0;
    struct Buf {
        cpx<T> a[LA][4], b[LA][4];
    };
    struct Rows {
        const cpx<T> *r1[4], *r2[4];
    };
    auto rows_of = [&](const Cls &c) {
        Rows r;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int mm = m < c.nm ? m : 0;
            const size_t off = ((size_t)c.mk[mm] * ny + c.mj[mm]) * P + sb;
            r.r1[m] = F1 + off;
            r.r2[m] = MODE != 0 ? F2 + off : F1 + off;
        }
        return r;
    };
    // a whole class: row by row (each row's loads back to back, so the
    // cache lines a lane range spans are reused while they are in L1)
    auto load = [&](Buf &u, int q) {
        if constexpr (MODE <= 2) {
            const Rows r = rows_of(members(q));
#pragma unroll
            for (int m = 0; m < 4; ++m) {
#pragma unroll
                for (int t = 0; t < LM; ++t) {
                    if (t > M) continue; // (uniform)
                    u.a[t][m] = r.r1[m][t];
                }
#pragma unroll
                for (int t = 0; t < LM; ++t) {
                    if (t > M) continue;
                    if constexpr (MODE != 0) u.b[t][m] = r.r2[m][t];
                    else u.b[t][m] = cpx<T>{0, 0};
                }
            }
        }
    };
    auto bin_class = [&](Buf &u, int q) {
        const Cls c = members(q);
        const int rr2 = c.jj * c.jj + c.kk * c.kk;
        // value channels of mode i summed over the members, in member order
        auto values = [&](int i, const cpx<T> (&a_)[4], const cpx<T> (&b_)[4], double (&val)[NV]) {
#pragma unroll
            for (int ch = 0; ch < NV; ++ch) val[ch] = 0.0;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                if (m >= c.nm) break;
                if constexpr (MODE == 3) {
                    const size_t id = ((size_t)c.mk[m] * ny + c.mj[m]) * P + i;
                    const double kv[3] = {double(i), double(c.mj[m] <= ny / 2 ? c.mj[m] : c.mj[m] - ny),
                                          double(c.mk[m] <= nz / 2 ? c.mk[m] : c.mk[m] - nz)};
                    double ur[6], ui[6];
#pragma unroll
                    for (int c6 = 0; c6 < 6; ++c6) {
                        const cpx<T> w = F1[(size_t)c6 * nk + id];
                        ur[c6] = double(w.x);
                        ui[c6] = double(w.y);
                    }
                    val[0] += tt_norm2(kv, ur, ui);
                } else if constexpr (MODE <= 2) {
                    const cpx<T> x = a_[m], y = b_[m];
                    if constexpr (MODE == 0) {
                        val[0] += double(x.x) * x.x + double(x.y) * x.y;
                    } else {
                        const double re = double(x.x) * y.x + double(x.y) * y.y;
                        const double im = double(x.y) * y.x - double(x.x) * y.y;
                        if constexpr (MODE == 1) {
                            val[0] += re;
                            val[1] += im;
                        } else {
                            val[0] += double(y.x) * y.x + double(y.y) * y.y;
                            val[1] += double(x.x) * x.x + double(x.y) * x.y;
                            val[2] += re;
                            val[3] += im;
                        }
                    }
                }
            }
        };
        // Runs of equal bin along the lane's modes.  The first run may have
        // begun in earlier lanes and the last one may go on in later lanes;
        // every run in between is whole and goes straight to the histogram.
        // Held: the first run (fk; fa, fb) once it has closed and the open
        // run (lk; ra, rb), as their bin-l and bin-(l+1) weighted sums.  A
        // mode outside the lane's range keeps the open run's bin and adds
        // zero weight.  Runs are reset and captured by multiplying with 0/1,
        // so a non-finite value can spread into later bins of its row class.
        int fk = NOKEY, lk = NOKEY;
        double fa[NC], fb[NC], ra[NC], rb[NC];
#pragma unroll
        for (int ch = 0; ch < NC; ++ch) fa[ch] = fb[ch] = ra[ch] = rb[ch] = 0.0;
#pragma unroll
        for (int t = 0; t < LM; ++t) {
            if (t > M) continue; // (uniform)
            const int i = s + t;
            const bool valid = i < e && i != nyq;
            const int ic = min(i, nxh - 1);
            const double p = sqrt(double(ic * ic + rr2));
            int l = (int)floor(p);
            const double d1 = double(l) - p, d2 = double(l + 1) - p;
            const double c1 = counts ? 1.0 : (1.0 - d1 * d1) * (1.0 - d1 * d1);
            const double c2 = counts ? 0.0 : (1.0 - d2 * d2) * (1.0 - d2 * d2);
            const double mult = ((flags & BIN_HALF) || ic == 0 || (2 * ic == nx)) ? 1.0 : 2.0;
            double w1 = mult * c1, w2 = mult * c2;
            double val[NV];
            values(ic, u.a[MODE <= 2 ? t : 0], u.b[MODE <= 2 ? t : 0], val);
            if (!valid) {
                l = lk;
                w1 = 0.0;
                w2 = 0.0;
            }
            const bool brk = lk != NOKEY && l != lk;
            const bool whole = brk && fk != NOKEY;
            // a whole run: bin l, then (after every lane's bin-l additions)
            // bin l + 1
            if (whole && lk < ns) {
#pragma unroll
                for (int ch = 0; ch < NC; ++ch) hist[ch * ns + lk] += ra[ch];
            }
            lds_sync();
            if (whole && lk + 1 < ns) {
#pragma unroll
                for (int ch = 0; ch < NC; ++ch) hist[ch * ns + lk + 1] += rb[ch];
            }
            lds_sync();
            const bool first = brk && fk == NOKEY;
            const double kf = first ? 1.0 : 0.0, kr = brk ? 0.0 : 1.0;
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) {
                fa[ch] = fma(ra[ch], kf, fa[ch]);
                fb[ch] = fma(rb[ch], kf, fb[ch]);
            }
            if (first) fk = lk;
#pragma unroll
            for (int ch = 0; ch < NCH; ++ch) {
                ra[ch] = fma(ra[ch], kr, w1 * val[ch]);
                rb[ch] = fma(rb[ch], kr, w2 * val[ch]);
            }
            if constexpr (WCH) {
                ra[NCH] = fma(ra[NCH], kr, w1 * c.nm);
                rb[NCH] = fma(rb[NCH], kr, w2 * c.nm);
            }
            lk = l;
        }
        // Runs across lanes: lanes whose open run has the same bin form one
        // contiguous segment (bins do not decrease along the row), so a
        // segmented scan over the open runs gives, at each lane, the sum of
        // its bin's segment up to the lane's last mode.
        seg_scan<NC>(lk, ra, rb);
        const int head = fk != NOKEY ? fk : lk; // bin of the lane's first mode
        const int nhead = __shfl_down(head, 1, 64);
        const int head_next = lane == 63 ? NOKEY : nhead;
        const int plk = __shfl_up(lk, 1, 64);
        const int lk_prev = lane == 0 ? NOKEY : plk;
        double pa[NC], pb[NC];
#pragma unroll
        for (int ch = 0; ch < NC; ++ch) {
            pa[ch] = __shfl_up(ra[ch], 1, 64);
            pb[ch] = __shfl_up(rb[ch], 1, 64);
        }
        // the first run closes in this lane and takes the segment before it;
        // the open run closes here unless the next lane starts with its bin
        const bool ta = fk != NOKEY;
        const bool tb = lk != NOKEY && head_next != lk;
        if (ta && lk_prev == fk) {
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) {
                fa[ch] += pa[ch];
                fb[ch] += pb[ch];
            }
        }
        if (ta && fk < ns) {
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) hist[ch * ns + fk] += fa[ch];
        }
        if (tb && lk < ns) {
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) hist[ch * ns + lk] += ra[ch];
        }
        lds_sync();
        if (ta && fk + 1 < ns) {
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) hist[ch * ns + fk + 1] += fb[ch];
        }
        if (tb && lk + 1 < ns) {
#pragma unroll
            for (int ch = 0; ch < NC; ++ch) hist[ch * ns + lk + 1] += rb[ch];
        }
        lds_sync();
    };
    if constexpr (PF) {
        Buf u0, u1;
        int q = blockIdx.x;
        if (q < ncls) load(u0, q);
        for (; q < ncls; q += 2 * G) {
            if (q + G < ncls) load(u1, q + G);
            bin_class(u0, q);
            if (q + G >= ncls) break;
            if (q + 2 * G < ncls) load(u0, q + 2 * G);
            bin_class(u1, q + G);
        }
    } else {
        Buf u;
        for (int q = blockIdx.x; q < ncls; q += G) {
            load(u, q);
            bin_class(u, q);
        }
    }
    double *dst = part + (size_t)blockIdx.x * NC * ns;
    for (int q = lane; q < NC * ns; q += 64) dst[q] = hist[q];
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

// Blocks of a binning launch: one wave each, with an LDS histogram of lds
// bytes.  A whole number of rounds of the blocks the device holds at once (at
// most BIN_BLOCKS) and no more blocks than row classes; the block count fixes
// the order of the additions, so every caller of a kernel uses this one.
template <typename K>
static int bin_grid(const Ctx &c, K kern, size_t lds)
{
    const int ncls = (c.ny / 2 + 1) * (c.nz / 2 + 1);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds) != hipSuccess || per_cu < 1) per_cu = 1;
    const int res = per_cu * (c.cus > 0 ? c.cus : 1);
    const int g = res >= BIN_BLOCKS ? BIN_BLOCKS : (BIN_BLOCKS / res) * res;
    return g < ncls ? g : ncls;
}

// The weight channel W of (ns, flags): its cache slot (W at spec_out +
// (5 + slot) * spec_cap), binned by the weight-only pass on a miss.
template <typename T>
static int weight_slot(Ctx &c, int ns, int flags)
{
    Ctx &owner = (c.view && c.parent) ? *c.parent : c;
    const long long key = ((long long)flags << 32) | (unsigned)ns;
    for (int q = 0; q < Ctx::WCACHE; ++q)
        if (owner.wkey[q] == key) return q;
    const int slot = owner.wnext;
    owner.wnext = (owner.wnext + 1) % Ctx::WCACHE;
    owner.wkey[slot] = -1;
    const size_t lw = sizeof(double) * ns;
    const int g = bin_grid(c, k_bin<T, 0, 4, true>, lw);
    // the weight-only pass reads no transform
    const cpx<T> *none = nullptr;
    hipLaunchKernelGGL((k_bin<T, 0, 4, true>), dim3(g), dim3(64), lw, c.stream, none, none, c.nx, c.ny, c.nz, c.nxh,
                       c.P, ns, c.spec_part, flags, c.nk);
    LAT_HIP(&c, hipGetLastError());
    bin_reduce(c, g, ns, c.spec_out + (5 + slot) * (size_t)c.spec_cap);
    LAT_HIP(&c, hipGetLastError());
    owner.wkey[slot] = key;
    return slot;
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
    if (sizeof(double) * nch * ns > BIN_LDS_MAX) {
        c.last_error = "spectrum bin count exceeds the binning kernel's LDS histogram";
        return -1;
    }
    if ((c.nxh - 1 + 63) / 64 + 1 > BIN_LANEMODES) {
        c.last_error = "spectrum rows longer than the binning kernel's row buffer";
        return -1;
    }
    const int slot = weight_slot<T>(c, ns, flags);
    if (slot < 0) return slot;
    const size_t cap = (size_t)c.spec_cap;
    const size_t lds = sizeof(double) * nch * ns;
    int g = 0;
    if (mode == 0) {
        g = bin_grid(c, k_bin<T, 1, 0, false>, lds);
        hipLaunchKernelGGL((k_bin<T, 1, 0, false>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh,
                           c.P, ns, c.spec_part, flags, c.nk);
    } else if (mode == 1) {
        g = bin_grid(c, k_bin<T, 2, 1, false>, lds);
        hipLaunchKernelGGL((k_bin<T, 2, 1, false>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh,
                           c.P, ns, c.spec_part, flags, c.nk);
    } else if (mode == 2) {
        g = bin_grid(c, k_bin<T, 4, 2, false>, lds);
        hipLaunchKernelGGL((k_bin<T, 4, 2, false>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh,
                           c.P, ns, c.spec_part, flags, c.nk);
    } else {
        g = bin_grid(c, k_bin<T, 1, 3, false>, lds);
        hipLaunchKernelGGL((k_bin<T, 1, 3, false>), dim3(g), dim3(64), lds, c.stream, F1, F2, c.nx, c.ny, c.nz, c.nxh,
                           c.P, ns, c.spec_part, flags, c.nk);
    }
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
    const size_t lds = sizeof(double) * ns;
    if (lds > BIN_LDS_MAX) {
        c.last_error = "spectrum bin count exceeds the binning kernel's LDS histogram";
        return -2;
    }
    // the grids of launch_bin for the same kernels (same additions in the
    // same order)
    const int gw = bin_grid(c, k_bin<double, 0, 4, true>, lds);
    const int gv = bin_grid(c, k_bin<double, 1, 0, false>, lds);
    const int g = gw > gv ? gw : gv;
    const int ngrp = (g + BIN_GROUP - 1) / BIN_GROUP;
    const size_t need = ((size_t)g + ngrp + 2) * ns;
    if (need > NOALLOC_DOUBLES) {
        c.last_error = "lattice too large for the allocation-free scratch (" + std::to_string(need * 8) +
                       " bytes needed, " + std::to_string(NOALLOC_DOUBLES * 8) + " held)";
        return -2;
    }
    if ((c.nxh - 1 + 63) / 64 + 1 > BIN_LANEMODES) {
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
    auto reduce = [&](int nb, double *out) {
        const int ng = (nb + BIN_GROUP - 1) / BIN_GROUP;
        hipLaunchKernelGGL(k_bin_reduce, dim3((ns + 255) / 256, ng), dim3(256), 0, c.stream, part, nb, BIN_GROUP, ns,
                           grp);
        hipLaunchKernelGGL(k_bin_reduce, dim3((ns + 255) / 256, 1), dim3(256), 0, c.stream, grp, ng, ng, ns, out);
    };
    // the weight channel, then the value channel, with the kernels, grid and
    // reduction order of launch_bin
    hipLaunchKernelGGL((k_bin<double, 0, 4, true>), dim3(gw), dim3(64), lds, c.stream, F, F, c.nx, c.ny, c.nz, c.nxh,
                       P, ns, part, flags, c.nk);
    reduce(gw, W);
    hipLaunchKernelGGL((k_bin<double, 1, 0, false>), dim3(gv), dim3(64), lds, c.stream, F, F, c.nx, c.ny, c.nz, c.nxh,
                       P, ns, part, flags, c.nk);
    reduce(gv, raw);
    const double n3 = (double)c.nvol;
    hipLaunchKernelGGL(k_spec_finish, dim3((ns + 255) / 256), dim3(256), 0, c.stream, raw, W, ns,
                       kind == 0 ? n3 * n3 : 1.0, S_dev);
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

template int launch_bin<double>(Ctx &, const cpx<double> *, const cpx<double> *, int, int, int);
template int launch_bin<float>(Ctx &, const cpx<float> *, const cpx<float> *, int, int, int);

} // namespace lat
