// fft_core.h -- LDS-resident Stockham FFT building blocks for gfx950.
//
// A workgroup owns `L` independent 1-D lines of length N held in LDS.
// Element n of line b lives at buf[b*LSTR + n*ESTR].  Each radix-R pass
// enumerates L*N/R butterflies; every thread first gathers the R inputs of
// its butterflies into registers (with twiddles), the workgroup barriers,
// then every thread scatters its R outputs to their Stockham (auto-sort)
// positions and barriers again, so the transform runs in place in LDS and
// the result is in natural order.  Twiddles come from a table
// w[m] = exp(-2 pi i m / (N*TWS)) precomputed on the host in long double.
//
// The radix schedule is fixed at compile time: radix 4 while possible, then
// 2, 3, 5.  All loop bounds are compile-time, so the per-thread register
// buffers are statically sized.
#pragma once
#include <hip/hip_runtime.h>

namespace lat {

template <typename T>
struct cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return {a.x, -a.y}; }

// ---------------------------------------------------------------- LDS barrier
// Workgroup barrier for LDS data: the wave's LDS operations complete and
// every wave arrives, with the fences restricted to the LDS address space.
// Unlike __syncthreads(), global memory operations (prefetched loads,
// stores of results) stay in flight across it.  Data that reaches LDS
// through the vector memory path (LDS-direct loads) needs its own vmcnt wait.
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// ---------------------------------------------------------------- schedule
// Largest radix first (fewest LDS round trips), capped at MR: 32, 16, 8, 4,
// 2, then 3, 5.
constexpr int pick_radix(int rem, int mr = 16)
{
    return (mr >= 32 && rem % 32 == 0) ? 32
           : (mr >= 16 && rem % 16 == 0) ? 16
           : (mr >= 8 && rem % 8 == 0) ? 8
           : (rem % 4 == 0) ? 4
           : (rem % 2 == 0) ? 2
           : (rem % 3 == 0) ? 3
           : (rem % 5 == 0) ? 5
                            : rem;
}
constexpr int num_passes(int n, int mr = 16)
{
    int c = 0;
    while (n > 1) { n /= pick_radix(n, mr); ++c; }
    return c;
}
constexpr int radix_at(int n, int p, int mr = 16)
{
    int r = 1;
    for (int i = 0; i <= p; ++i) { r = pick_radix(n, mr); n /= r; }
    return r;
}
constexpr int ns_at(int n, int p, int mr = 16)
{
    int ns = 1;
    for (int i = 0; i < p; ++i) { int r = pick_radix(n, mr); ns *= r; n /= r; }
    return ns;
}
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }
// The same radices in reverse order (REV): a transform whose first radix
// equals the last radix of the forward schedule, so the forward transform's
// last pass and the inverse transform's first pass can share registers.
constexpr int sched_radix(int n, int p, int mr, bool rev)
{
    return rev ? radix_at(n, num_passes(n, mr) - 1 - p, mr) : radix_at(n, p, mr);
}
constexpr int sched_ns(int n, int p, int mr, bool rev)
{
    int ns = 1;
    for (int i = 0; i < p; ++i) ns *= sched_radix(n, i, mr, rev);
    return ns;
}

// ---------------------------------------------------------------- radix kernels
// multiply by e^{SIGN 2 pi i m / 16} for a compile-time m
template <int M, int SIGN, typename T>
__device__ __forceinline__ cpx<T> rot16(cpx<T> a)
{
    constexpr int m = ((M % 16) + 16) % 16;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (m == 4) { // SIGN*i
        return SIGN < 0 ? cpx<T>{a.y, -a.x} : cpx<T>{-a.y, a.x};
    } else if constexpr (m == 8) {
        return {-a.x, -a.y};
    } else if constexpr (m == 12) {
        return SIGN < 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
    } else {
        // cos and sin of 2 pi m / 16 for m = 1..15 (octant-exact constants)
        constexpr double C1 = 0.92387953251128675612818318939679, S1 = 0.38268343236508977172845998403040;
        constexpr double C2 = 0.70710678118654752440084436210485;
        constexpr double cs[16] = {1, C1, C2, S1, 0, -S1, -C2, -C1, -1, -C1, -C2, -S1, 0, S1, C2, C1};
        constexpr double sn[16] = {0, S1, C2, C1, 1, C1, C2, S1, 0, -S1, -C2, -C1, -1, -C1, -C2, -S1};
        const T c = T(cs[m]), s = T(SIGN < 0 ? -sn[m] : sn[m]);
        return {a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_small(cpx<T> *v);

// 16-point DFT as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2.
template <int SIGN, typename T>
__device__ __forceinline__ void dft16(cpx<T> *v)
{
    cpx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cpx<T> t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft_small<4, SIGN>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
    // internal twiddles W16^{n2 k1}
    y[1][1] = rot16<1, SIGN>(y[1][1]);
    y[1][2] = rot16<2, SIGN>(y[1][2]);
    y[1][3] = rot16<3, SIGN>(y[1][3]);
    y[2][1] = rot16<2, SIGN>(y[2][1]);
    y[2][2] = rot16<4, SIGN>(y[2][2]);
    y[2][3] = rot16<6, SIGN>(y[2][3]);
    y[3][1] = rot16<3, SIGN>(y[3][1]);
    y[3][2] = rot16<6, SIGN>(y[3][2]);
    y[3][3] = rot16<9, SIGN>(y[3][3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cpx<T> t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft_small<4, SIGN>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
}

// multiply by e^{SIGN 2 pi i m / 32} for a compile-time m (even m reduce to
// the 16-point rotations)
template <int M, int SIGN, typename T>
__device__ __forceinline__ cpx<T> rot32(cpx<T> a)
{
    constexpr int m = ((M % 32) + 32) % 32;
    if constexpr (m % 2 == 0) {
        return rot16<m / 2, SIGN>(a);
    } else {
        // cos and sin of 2 pi m / 32 for odd m, from the first-octant values
        // cos(pi/16), sin(pi/16), cos(3 pi/16), sin(3 pi/16)
        constexpr double A = 0.98078528040323044912618223613424, B = 0.19509032201612826784828486847702;
        constexpr double C = 0.83146961230254523707878837761791, D = 0.55557023301960222474283081394853;
        constexpr double cs[16] = {A, C, D, B, -B, -D, -C, -A, -A, -C, -D, -B, B, D, C, A};
        constexpr double sn[16] = {B, D, C, A, A, C, D, B, -B, -D, -C, -A, -A, -C, -D, -B};
        const T c = T(cs[m / 2]), s = T(SIGN < 0 ? -sn[m / 2] : sn[m / 2]);
        return {a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

// 32-point DFT as 2 x 16: n = 2 n1 + n2, k = k1 + 16 k2.
template <int SIGN, typename T>
__device__ __forceinline__ void dft32(cpx<T> *v)
{
    cpx<T> a[16], b[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
        a[n1] = v[2 * n1];
        b[n1] = v[2 * n1 + 1];
    }
    dft16<SIGN>(a);
    dft16<SIGN>(b);
#define LAT_R32(K) b[K] = rot32<K, SIGN>(b[K]);
    LAT_R32(1) LAT_R32(2) LAT_R32(3) LAT_R32(4) LAT_R32(5) LAT_R32(6) LAT_R32(7) LAT_R32(8)
    LAT_R32(9) LAT_R32(10) LAT_R32(11) LAT_R32(12) LAT_R32(13) LAT_R32(14) LAT_R32(15)
#undef LAT_R32
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) {
        v[k1] = cadd(a[k1], b[k1]);
        v[k1 + 16] = csub(a[k1], b[k1]);
    }
}

// 8-point DFT as 2 x 4: n = 2 n1 + n2, k = k1 + 4 k2.
template <int SIGN, typename T>
__device__ __forceinline__ void dft8(cpx<T> *v)
{
    cpx<T> a[4] = {v[0], v[2], v[4], v[6]}, b[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4, SIGN>(a);
    dft_small<4, SIGN>(b);
    b[1] = rot16<2, SIGN>(b[1]);
    b[2] = rot16<4, SIGN>(b[2]);
    b[3] = rot16<6, SIGN>(b[3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        v[k1] = cadd(a[k1], b[k1]);
        v[k1 + 4] = csub(a[k1], b[k1]);
    }
}

// SIGN = -1: forward e^{-2 pi i}; +1: backward.
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_small(cpx<T> *v)
{
    if constexpr (R == 32) {
        dft32<SIGN>(v);
    } else if constexpr (R == 16) {
        dft16<SIGN>(v);
    } else if constexpr (R == 8) {
        dft8<SIGN>(v);
    } else if constexpr (R == 2) {
        cpx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        cpx<T> a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
        // rot = (SIGN<0 ? -i : +i) * a3
        cpx<T> rot = (SIGN < 0) ? cpx<T>{a3.y, -a3.x} : cpx<T>{-a3.y, a3.x};
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, rot);
        v[3] = csub(a1, rot);
    } else if constexpr (R == 3) {
        // c = cos(2pi/3) = -1/2, s = sin(2pi/3) = sqrt(3)/2
        const T c = T(-0.5);
        const T s = T(SIGN < 0 ? -0.86602540378443864676372317075294 : 0.86602540378443864676372317075294);
        cpx<T> t1 = cadd(v[1], v[2]);
        cpx<T> t2 = csub(v[1], v[2]);
        cpx<T> m1 = {v[0].x + c * t1.x, v[0].y + c * t1.y};
        cpx<T> m2 = {-s * t2.y, s * t2.x}; // i*s*t2
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m1, m2);
        v[2] = csub(m1, m2);
    } else {
        static_assert(R == 5, "unsupported radix");
        const T c1 = T(0.30901699437494742410229341718282), c2 = T(-0.80901699437494742410229341718282);
        const T s1 = T(SIGN < 0 ? -0.95105651629515357211643933337938 : 0.95105651629515357211643933337938);
        const T s2 = T(SIGN < 0 ? -0.58778525229247312916870595463907 : 0.58778525229247312916870595463907);
        cpx<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cpx<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cpx<T> m1 = {v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y};
        cpx<T> m2 = {v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y};
        cpx<T> n1 = {-(s1 * b1.y + s2 * b2.y), s1 * b1.x + s2 * b2.x};
        cpx<T> n2 = {-(s2 * b1.y - s1 * b2.y), s2 * b1.x - s1 * b2.x};
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
}

// ---------------------------------------------------------------- one pass
// COLMAJOR: butterfly g -> line = g % L, j = g / L (consecutive threads walk
// lines; used when lines are interleaved, ESTR = L, LSTR = 1).  Otherwise
// j = g % (N/R), line = g / (N/R) (rows, ESTR = 1).
template <typename T, int N, int R, int NS, int SIGN, int L, int TPB, bool COLMAJOR, int TWS>
__device__ __forceinline__ void stockham_pass(cpx<T> *buf, int lstr, int estr, const cpx<T> *__restrict__ tw)
{
    constexpr int NR = N / R;
    constexpr int TOTAL = L * NR;
    constexpr int BPT = cdiv(TOTAL, TPB);
    constexpr int TSTEP = (N / (NS * R)) * TWS; // table stride per unit of k*q
    constexpr int TN = N * TWS;
    const int tid = threadIdx.x;
    cpx<T> v[BPT][R];
    int oidx[BPT];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = tid + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
            int line, j;
            if constexpr (COLMAJOR) { line = g % L; j = g / L; }
            else { j = g % NR; line = g / NR; }
            const int k = j % NS;
            cpx<T> *base = buf + line * lstr;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                cpx<T> a = base[(j + q * NR) * estr];
                if (q > 0 && NS > 1) {
                    int m = (k * q * TSTEP) % TN;
                    cpx<T> w = tw[m];
                    if (SIGN > 0) w.y = -w.y;
                    a = cmul(a, w);
                }
                v[b][q] = a;
            }
            dft_small<R, SIGN>(v[b]);
            oidx[b] = line * lstr + ((j / NS) * NS * R + k) * estr;
        }
    }
    lds_sync();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = tid + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
#pragma unroll
            for (int q = 0; q < R; ++q) buf[oidx[b] + q * NS * estr] = v[b][q];
        }
    }
    lds_sync();
}

template <typename T, int N, int SIGN, int L, int TPB, bool COLMAJOR, int TWS, int P = 0>
__device__ __forceinline__ void lds_fft(cpx<T> *buf, int lstr, int estr, const cpx<T> *__restrict__ tw)
{
    if constexpr (P < num_passes(N)) {
        stockham_pass<T, N, radix_at(N, P), ns_at(N, P), SIGN, L, TPB, COLMAJOR, TWS>(buf, lstr, estr, tw);
        lds_fft<T, N, SIGN, L, TPB, COLMAJOR, TWS, P + 1>(buf, lstr, estr, tw);
    }
}

// Passes [P, PEND) in LDS (schedule with max radix MR, reversed with REV).
template <typename T, int N, int SIGN, int L, int TPB, bool COLMAJOR, int TWS, int P, int PEND, int MR = 16,
          bool REV = false>
__device__ __forceinline__ void lds_fft_range(cpx<T> *buf, int lstr, int estr, const cpx<T> *__restrict__ tw)
{
    if constexpr (P < PEND) {
        stockham_pass<T, N, sched_radix(N, P, MR, REV), sched_ns(N, P, MR, REV), SIGN, L, TPB, COLMAJOR, TWS>(
            buf, lstr, estr, tw);
        lds_fft_range<T, N, SIGN, L, TPB, COLMAJOR, TWS, P + 1, PEND, MR, REV>(buf, lstr, estr, tw);
    }
}

// Pass P read from LDS, result handed to st(line, position, value) instead of
// being written back (used for the last pass: results go straight to HBM).
// No trailing barrier: the caller synchronises before LDS is reused.
template <typename T, int N, int P, int SIGN, int L, int TPB, bool COLMAJOR, int TWS, int MR, bool REV = false,
          typename Store>
__device__ __forceinline__ void stockham_pass_out(const cpx<T> *buf, int lstr, int estr,
                                                  const cpx<T> *__restrict__ tw, Store st)
{
    constexpr int R = sched_radix(N, P, MR, REV), NS = sched_ns(N, P, MR, REV);
    constexpr int NR = N / R;
    constexpr int TOTAL = L * NR;
    constexpr int BPT = cdiv(TOTAL, TPB);
    constexpr int TSTEP = (N / (NS * R)) * TWS;
    constexpr int TN = N * TWS;
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = threadIdx.x + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
            int line, j;
            if constexpr (COLMAJOR) { line = g % L; j = g / L; }
            else { j = g % NR; line = g / NR; }
            const int k = j % NS;
            const cpx<T> *base = buf + line * lstr;
            cpx<T> v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                cpx<T> a = base[(j + q * NR) * estr];
                if (q > 0 && NS > 1) {
                    cpx<T> w = tw[(k * q * TSTEP) % TN];
                    if (SIGN > 0) w.y = -w.y;
                    a = cmul(a, w);
                }
                v[q] = a;
            }
            dft_small<R, SIGN>(v);
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) st(line, o + q * NS, v[q]);
        }
    }
}

// The last pass of a forward transform (schedule MR), a pointwise factor
// mul(line, n) on its natural-order outputs n, and the first pass of the
// inverse transform in the reversed schedule, without the LDS round trip in
// between: the last forward pass of butterfly (line, j) yields the elements
// n = j + q N/R, which are exactly the inputs of inverse butterfly (line, j)
// (first pass, radix R, no twiddles).  Column-major lines (ESTR = L).
template <typename T, int N, int L, int TPB, int MR, typename Mul>
__device__ __forceinline__ void fused_last_first(cpx<T> *buf, const cpx<T> *__restrict__ tw, Mul mul)
{
    constexpr int NP = num_passes(N, MR);
    constexpr int R = radix_at(N, NP - 1, MR), NS = ns_at(N, NP - 1, MR);
    static_assert(NS * R == N, "last pass spans the line");
    constexpr int TOTAL = L * NS;
    constexpr int BPT = cdiv(TOTAL, TPB);
    cpx<T> v[BPT][R];
    int pos[BPT];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = threadIdx.x + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
            const int line = g % L, j = g / L;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                cpx<T> a = buf[(j + q * NS) * L + line];
                if (q > 0 && NS > 1) a = cmul(a, tw[(j * q) % N]);
                v[b][q] = a;
            }
            dft_small<R, -1>(v[b]);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const T m = mul(line, j + q * NS);
                v[b][q] = {v[b][q].x * m, v[b][q].y * m};
            }
            dft_small<R, +1>(v[b]);
            pos[b] = j * R * L + line;
        }
    }
    lds_sync();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = threadIdx.x + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
#pragma unroll
            for (int q = 0; q < R; ++q) buf[pos[b] + q * L] = v[b][q];
        }
    }
    lds_sync();
}

// ---------------------------------------------------------------- ping-pong passes
// Out-of-place variants for column-major tiles (ESTR = L): a pass reads one
// LDS buffer and writes the other, so each butterfly stores its outputs as
// soon as they are computed (registers for one butterfly at a time, one
// barrier per pass instead of two).  Twice the LDS of the in-place passes.
template <typename T, int N, int R, int NS, int SIGN, int L, int TPB>
__device__ __forceinline__ void stockham_pass_pp(const cpx<T> *__restrict__ src, cpx<T> *__restrict__ dst,
                                                 const cpx<T> *__restrict__ tw)
{
    constexpr int NR = N / R;
    constexpr int TOTAL = L * NR;
    constexpr int BPT = cdiv(TOTAL, TPB);
    constexpr int TSTEP = N / (NS * R);
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = threadIdx.x + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
            const int line = g % L, j = g / L;
            const int k = j % NS;
            cpx<T> v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                cpx<T> a = src[(j + q * NR) * L + line];
                if (q > 0 && NS > 1) {
                    cpx<T> w = tw[(k * q * TSTEP) % N];
                    if (SIGN > 0) w.y = -w.y;
                    a = cmul(a, w);
                }
                v[q] = a;
            }
            dft_small<R, SIGN>(v);
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) dst[(o + q * NS) * L + line] = v[q];
        }
    }
    lds_sync();
}

// Passes [P, PEND) alternating between b0 and b1; PAR = buffer holding the
// input of pass P.  After the range the data is in buffer PAR ^ ((PEND-P)&1).
template <typename T, int N, int SIGN, int L, int TPB, int P, int PEND, int MR, bool REV, int PAR>
__device__ __forceinline__ void lds_fft_range_pp(cpx<T> *b0, cpx<T> *b1, const cpx<T> *__restrict__ tw)
{
    if constexpr (P < PEND) {
        stockham_pass_pp<T, N, sched_radix(N, P, MR, REV), sched_ns(N, P, MR, REV), SIGN, L, TPB>(
            PAR ? b1 : b0, PAR ? b0 : b1, tw);
        lds_fft_range_pp<T, N, SIGN, L, TPB, P + 1, PEND, MR, REV, PAR ^ 1>(b0, b1, tw);
    }
}

// fused_last_first, out of place
template <typename T, int N, int L, int TPB, int MR, typename Mul>
__device__ __forceinline__ void fused_last_first_pp(const cpx<T> *__restrict__ src, cpx<T> *__restrict__ dst,
                                                    const cpx<T> *__restrict__ tw, Mul mul)
{
    constexpr int NP = num_passes(N, MR);
    constexpr int R = radix_at(N, NP - 1, MR), NS = ns_at(N, NP - 1, MR);
    static_assert(NS * R == N, "last pass spans the line");
    constexpr int TOTAL = L * NS;
    constexpr int BPT = cdiv(TOTAL, TPB);
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = threadIdx.x + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
            const int line = g % L, j = g / L;
            cpx<T> v[R];
#pragma unroll
            for (int q = 0; q < R; ++q) {
                cpx<T> a = src[(j + q * NS) * L + line];
                if (q > 0 && NS > 1) a = cmul(a, tw[(j * q) % N]);
                v[q] = a;
            }
            dft_small<R, -1>(v);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const T m = mul(line, j + q * NS);
                v[q] = {v[q].x * m, v[q].y * m};
            }
            dft_small<R, +1>(v);
#pragma unroll
            for (int q = 0; q < R; ++q) dst[(j * R + q) * L + line] = v[q];
        }
    }
    lds_sync();
}

} // namespace lat
