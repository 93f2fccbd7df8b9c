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

// ---------------------------------------------------------------- schedule
constexpr int pick_radix(int rem)
{
    return (rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2 : (rem % 3 == 0) ? 3 : (rem % 5 == 0) ? 5 : rem;
}
constexpr int num_passes(int n)
{
    int c = 0;
    while (n > 1) { n /= pick_radix(n); ++c; }
    return c;
}
constexpr int radix_at(int n, int p)
{
    int r = 1;
    for (int i = 0; i <= p; ++i) { r = pick_radix(n); n /= r; }
    return r;
}
constexpr int ns_at(int n, int p)
{
    int ns = 1;
    for (int i = 0; i < p; ++i) { int r = pick_radix(n); ns *= r; n /= r; }
    return ns;
}
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------- radix kernels
// SIGN = -1: forward e^{-2 pi i}; +1: backward.
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_small(cpx<T> *v)
{
    if constexpr (R == 2) {
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
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
        const int g = tid + b * TPB;
        if (TOTAL % TPB == 0 || g < TOTAL) {
#pragma unroll
            for (int q = 0; q < R; ++q) buf[oidx[b] + q * NS * estr] = v[b][q];
        }
    }
    __syncthreads();
}

template <typename T, int N, int SIGN, int L, int TPB, bool COLMAJOR, int TWS, int P = 0>
__device__ __forceinline__ void lds_fft(cpx<T> *buf, int lstr, int estr, const cpx<T> *__restrict__ tw)
{
    if constexpr (P < num_passes(N)) {
        stockham_pass<T, N, radix_at(N, P), ns_at(N, P), SIGN, L, TPB, COLMAJOR, TWS>(buf, lstr, estr, tw);
        lds_fft<T, N, SIGN, L, TPB, COLMAJOR, TWS, P + 1>(buf, lstr, estr, tw);
    }
}

} // namespace lat
