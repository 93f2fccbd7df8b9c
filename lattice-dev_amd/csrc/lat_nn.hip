// lat_nn.hip -- the learned corrector: a 3-D convolutional network on MFMA.
//
// Network: 4 layers of 3x3x3 periodic convolutions, 8 -> W -> W -> W -> 4
// channels (W = 32), SiLU between layers, BF16 weights and activations, FP32
// accumulation on the matrix cores (gfx950's v_mfma_f32_16x16x32_bf16).  Input: the
// normalised (phi, chi, pi_phi, pi_chi) fluctuations after one coarse
// leapfrog step and four uniform channels that carry what the normalisation
// removes and the dynamics depend on (the homogeneous phi and pi_phi, the
// coupling, the chi fluctuation amplitude); output: the normalised residual
// to the fine Yoshida-4 trajectory.
//
// Layout: activations are site-major with channels fastest, X[site][C]
// (C a multiple of 4), so the four consecutive reduction indices an MFMA
// lane supplies are one 8-byte load.  A convolution is an implicit GEMM
//   Y[s][co] = sum_k W[co][k] X[nbr(s, tap(k))][ci(k)],  k = tap * CIN + ci,
// with M = output channels (16-row tiles), N = 16 sites per wave tile and
// K = 27 CIN padded to a multiple of 32, 32 reduction indices per MFMA (each
// lane supplies 8 consecutive ones: 8 channels of one tap, or 4 channels of
// two taps when CIN = 4).  The weight matrix of a layer is staged in LDS
// once per workgroup.
//
// Backward: the data gradient is the same convolution with the transposed,
// tap-reversed weights Wt[ci][t * COUT + co] = W[co][(26 - t) * CIN + ci];
// the weight gradient is a GEMM over sites (K = sites), staged through LDS
// per row segment and accumulated per workgroup into partials that are
// reduced in a fixed order (deterministic).
#include <cmath>

#include "lat_internal.h"

namespace lat {
namespace nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

// round-to-nearest-even FP32 -> BF16 (finite inputs)
__host__ __device__ inline bf16 to_bf16(float f)
{
    unsigned u = __builtin_bit_cast(unsigned, f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}
__host__ __device__ inline float from_bf16(bf16 h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

__device__ inline float silu(float v) { return v / (1.0f + expf(-v)); }
__device__ inline float dsilu(float v)
{
    const float s = 1.0f / (1.0f + expf(-v));
    return s * (1.0f + v * (1.0f - s));
}

constexpr int KPAD(int cin) { return (27 * cin + 31) / 32 * 32; }

// D (16 x 16, FP32) += A (16 x 32) B (32 x 16) in BF16; lane l supplies
// A[l % 16][8 (l / 16) .. +7] and B[8 (l / 16) .. +7][l % 16] and holds
// D[4 (l / 16) + e][l % 16], e = 0..3
__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// epilogue modes of the convolution
enum { CONV_SILU = 0,  // Z = acc + b (BF16), Y = silu(Z) (BF16)
       CONV_LINEAR = 1, // Yf = acc + b (FP32)
       CONV_DSILU = 2 };  // Y = acc * silu'(Zin) (BF16): data gradient through the activation

// Neighbour of site (x, y, z) for tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
__device__ inline size_t nbr_site(int x, int y, int z, int t, int n)
{
    int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    int xx = x + dx, yy = y + dy, zz = z + dz;
    xx = xx < 0 ? xx + n : (xx >= n ? xx - n : xx);
    yy = yy < 0 ? yy + n : (yy >= n ? yy - n : yy);
    zz = zz < 0 ? zz + n : (zz >= n ? zz - n : zz);
    return (size_t)xx + (size_t)n * ((size_t)yy + (size_t)n * zz);
}

// One wave per 16-site tile; CONV_WAVES waves per workgroup share the LDS
// copy of the weights; grid-stride over tiles.
// W: [MT*16][KP] BF16 (rows >= COUT and k >= 27*CIN are zero).
constexpr int CONV_WAVES = 8;
template <int CIN, int COUT, int MODE>
__global__ __launch_bounds__(64 * CONV_WAVES) void k_conv(const bf16 *__restrict__ X, const bf16 *__restrict__ W,
                                              const float *__restrict__ bias, bf16 *__restrict__ Y,
                                              float *__restrict__ Yf, bf16 *__restrict__ Z,
                                              const bf16 *__restrict__ Zin, int n)
{
    constexpr int MT = (COUT + 15) / 16;
    constexpr int KP = KPAD(CIN);
    __shared__ __attribute__((aligned(16))) bf16 ws[MT * 16 * KP];
    for (int q = threadIdx.x; q < MT * 16 * KP / 8; q += 64 * CONV_WAVES)
        reinterpret_cast<s16x8 *>(ws)[q] = reinterpret_cast<const s16x8 *>(W)[q];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, kg = lane >> 4;
    const size_t nvol = (size_t)n * n * n, ntile = nvol / 16;
    for (size_t tile = (size_t)blockIdx.x * CONV_WAVES + wave; tile < ntile;
         tile += (size_t)gridDim.x * CONV_WAVES) {
        const size_t s = tile * 16 + j;
        const int x = (int)(s % n), y = (int)((s / n) % n), z = (int)(s / ((size_t)n * n));
        f32x4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < KP; k0 += 32) {
            const int k = k0 + 8 * kg;
            s16x8 b = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if constexpr (CIN % 8 == 0) {
                const int t = k / CIN, ci = k - t * CIN;
                if (t < 27) b = *reinterpret_cast<const s16x8 *>(X + nbr_site(x, y, z, t, n) * CIN + ci);
            } else {
                static_assert(CIN == 4, "4 or a multiple of 8 input channels");
                const int t = k / 4; // taps t and t + 1, channels 0..3 of each
                bf16x4 lo = bf16x4{0, 0, 0, 0}, hi = bf16x4{0, 0, 0, 0};
                if (t < 27) lo = *reinterpret_cast<const bf16x4 *>(X + nbr_site(x, y, z, t, n) * 4);
                if (t + 1 < 27) hi = *reinterpret_cast<const bf16x4 *>(X + nbr_site(x, y, z, t + 1, n) * 4);
                b = s16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            }
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const s16x8 a = *reinterpret_cast<const s16x8 *>(ws + (mt * 16 + j) * KP + k);
                acc[mt] = mfma32(a, b, acc[mt]);
            }
        }
        // lane holds D[co][site] for co = mt*16 + 4*kg + e, site = tile*16 + j
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const int co0 = mt * 16 + 4 * kg;
            if (co0 >= COUT) continue;
            if constexpr (MODE == CONV_LINEAR) {
                float4 o;
                o.x = acc[mt][0] + bias[co0];
                o.y = acc[mt][1] + bias[co0 + 1];
                o.z = acc[mt][2] + bias[co0 + 2];
                o.w = acc[mt][3] + bias[co0 + 3];
                *reinterpret_cast<float4 *>(Yf + s * COUT + co0) = o;
            } else if constexpr (MODE == CONV_SILU) {
                bf16x4 zo, yo;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = acc[mt][e] + bias[co0 + e];
                    const bf16 zb = to_bf16(v);
                    zo[e] = (short)zb;
                    yo[e] = (short)to_bf16(silu(from_bf16(zb)));
                }
                *reinterpret_cast<bf16x4 *>(Z + s * COUT + co0) = zo;
                *reinterpret_cast<bf16x4 *>(Y + s * COUT + co0) = yo;
            } else {
                const bf16x4 zi = *reinterpret_cast<const bf16x4 *>(Zin + s * COUT + co0);
                bf16x4 yo;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    yo[e] = (short)to_bf16(acc[mt][e] * dsilu(from_bf16((bf16)zi[e])));
                *reinterpret_cast<bf16x4 *>(Y + s * COUT + co0) = yo;
            }
        }
    }
}

// Weight gradient partials: one per block, nchunks(nvol) of them at most
// (GCHUNK sites per partial), each [MT*16][KP + 16] FP32 with the bias
// gradient in column KP.
constexpr int GCHUNK = 4096;

// Weight gradient from LDS tiles.  A block of 9 waves walks chunks of CL
// sites along x (one row segment; CL = 64, 32 or 16, dividing n).  Per chunk
// the X rows of the 3 x 3 (dy, dz) neighbour rows with one site of halo on
// each side in x, and the chunk's G, are staged in LDS channel-major
// (xs[ci][row][x], gs[co][x]), so that the 8 consecutive sites an MFMA lane
// supplies are 8 consecutive LDS entries.  Wave w takes the tap row (dy, dz)
// = (w % 3 - 1, w / 3 - 1) and its three dx taps: per 32 sites it loads G
// once (A = G^T, rows co) and X shifted by dx (B, columns ci) and keeps the
// 3 x MT x NTL accumulator tiles of its taps in registers across its chunks.
// The block writes its partial dW (and db in column KP) as [MT*16][KP + 16];
// the partials are summed in block order by k_sum_chunks.
constexpr int DW2_WAVES = 9;
constexpr int DW2_CLMAX = 64;
constexpr int DW2_RP = DW2_CLMAX + 8;   // LDS row pitch of the X tile (halo + 16-byte multiple)
constexpr int DW2_BLOCKS = 512;          // persistent blocks (2 per CU)

__device__ inline int dw2_chunk(int n) { return n % 64 == 0 ? 64 : (n % 32 == 0 ? 32 : 16); }

template <int CIN, int COUT>
__global__ __launch_bounds__(64 * DW2_WAVES) void k_conv_dw2(const bf16 *__restrict__ G, const bf16 *__restrict__ X,
                                                             float *__restrict__ part, int n)
{
    constexpr int MT = (COUT + 15) / 16, NTL = (CIN + 15) / 16;
    constexpr int KP = KPAD(CIN), LD = KP + 16;
    constexpr int NT = 64 * DW2_WAVES;
    // channel rows padded to whole 16-row tiles; rows >= CIN / COUT stay zero
    __shared__ __attribute__((aligned(16))) bf16 xs[NTL * 16][9][DW2_RP];
    __shared__ __attribute__((aligned(16))) bf16 gs[MT * 16][DW2_CLMAX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, kg = lane >> 4;
    const int dy = wave % 3 - 1, dz = wave / 3 - 1;
    const int row = (dz + 1) * 3 + (dy + 1);
    const int CL = dw2_chunk(n);
    const int KS = (CL + 31) / 32;
    const size_t nvol = (size_t)n * n * n;
    const int nchunk = (int)(nvol / CL);
    for (int q = threadIdx.x; q < NTL * 16 * 9 * DW2_RP; q += NT) (&xs[0][0][0])[q] = 0;
    for (int q = threadIdx.x; q < MT * 16 * DW2_CLMAX; q += NT) (&gs[0][0])[q] = 0;
    f32x4 acc[3][MT][NTL];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) acc[d][mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    float db = 0.f; // bias gradient of channel co = threadIdx.x (< COUT)
    __syncthreads();
    for (int ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
        const size_t s0 = (size_t)ch * CL;
        const int x0 = (int)(s0 % n), y = (int)((s0 / n) % n), z = (int)(s0 / ((size_t)n * n));
        // stage X: 9 rows x (CL + 2) sites x CIN channels, 4 channels (8 bytes) per item
        constexpr int C4 = CIN / 4;
        for (int q = threadIdx.x; q < 9 * (CL + 2) * C4; q += NT) {
            const int c4 = q % C4, rest = q / C4;
            const int xi = rest % (CL + 2), r = rest / (CL + 2);
            int xx = x0 + xi - 1, yy = y + r % 3 - 1, zz = z + r / 3 - 1;
            xx = xx < 0 ? xx + n : (xx >= n ? xx - n : xx);
            yy = yy < 0 ? yy + n : (yy >= n ? yy - n : yy);
            zz = zz < 0 ? zz + n : (zz >= n ? zz - n : zz);
            const size_t site = (size_t)xx + (size_t)n * ((size_t)yy + (size_t)n * zz);
            const bf16x4 v = *reinterpret_cast<const bf16x4 *>(X + site * CIN + 4 * c4);
#pragma unroll
            for (int e = 0; e < 4; ++e) xs[4 * c4 + e][r][xi] = (bf16)v[e];
        }
        // stage G: CL sites x COUT channels
        constexpr int O4 = COUT / 4;
        for (int q = threadIdx.x; q < CL * O4; q += NT) {
            const int o4 = q % O4, xi = q / O4;
            const bf16x4 v = *reinterpret_cast<const bf16x4 *>(G + (s0 + xi) * COUT + 4 * o4);
#pragma unroll
            for (int e = 0; e < 4; ++e) gs[4 * o4 + e][xi] = (bf16)v[e];
        }
        __syncthreads();
        if (threadIdx.x < COUT) {
            for (int xi = 0; xi < CL; ++xi) db += from_bf16(gs[threadIdx.x][xi]);
        }
        for (int ks = 0; ks < KS; ++ks) {
            // first of the lane's 8 sites within the chunk; at CL = 16 the upper
            // half of the K step holds no sites and supplies zeros
            const int xk = 32 * ks + 8 * kg;
            s16x8 a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt)
                a[mt] = xk < CL ? *reinterpret_cast<const s16x8 *>(&gs[mt * 16 + j][xk]) : s16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int d = 0; d < 3; ++d) {
#pragma unroll
                for (int nt = 0; nt < NTL; ++nt) {
                    s16x8 b;
                    const bf16 *src = &xs[nt * 16 + j][row][xk + d]; // site x0 + xk + q + (d - 1), halo at 0
#pragma unroll
                    for (int qq = 0; qq < 8; ++qq) b[qq] = xk < CL ? (short)src[qq] : (short)0;
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt) acc[d][mt][nt] = mfma32(a[mt], b, acc[d][mt][nt]);
                }
            }
        }
        __syncthreads();
    }
    // D[row co = mt*16 + 4*kg + e][col ci = nt*16 + j] of tap t = (dz+1)*9 + (dy+1)*3 + d
    float *dst = part + (size_t)blockIdx.x * (MT * 16) * LD;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int t = (dz + 1) * 9 + (dy + 1) * 3 + d;
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int nt = 0; nt < NTL; ++nt) {
                const int ci = nt * 16 + j;
                if (ci >= CIN) continue;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int co = mt * 16 + 4 * kg + e;
                    dst[(size_t)co * LD + t * CIN + ci] = co < COUT ? acc[d][mt][nt][e] : 0.f;
                }
            }
    }
    // padding columns 27 CIN .. KP-1 and the bias block KP .. KP+15
    for (int q = threadIdx.x; q < MT * 16 * (LD - 27 * CIN); q += NT) {
        const int co = q / (LD - 27 * CIN), k = 27 * CIN + q % (LD - 27 * CIN);
        if (!(k == KP && co < COUT)) dst[(size_t)co * LD + k] = 0.f;
    }
    if (threadIdx.x < COUT) dst[(size_t)threadIdx.x * LD + KP] = db;
}

// out[q] = sum over chunks (in order) of part[chunk][q]
__global__ __launch_bounds__(256) void k_sum_chunks(const float *__restrict__ part, int nchunk, int nq,
                                                    float *__restrict__ out)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nq) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * nq + q];
    out[q] = s;
}

// Adam on the FP32 master weights of one layer; grad [MP][KP+16] (bias in
// column KP, row co).  Writes the BF16 copy W and the transposed, tap-reversed
// copy Wt[ci][t*COUT + co] (rows padded to a multiple of 16, K padded to KPt).
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, float *__restrict__ b, float *__restrict__ mw,
                                              float *__restrict__ vw, float *__restrict__ mb,
                                              float *__restrict__ vb, const float *__restrict__ grad, float lr,
                                              float b1, float b2, float eps, float c1, float c2,
                                              bf16 *__restrict__ Wb, bf16 *__restrict__ Wt)
{
    constexpr int KP = KPAD(CIN), LD = KP + 16;
    constexpr int KPt = KPAD(COUT);
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q < COUT * 27 * CIN) {
        const int co = q / (27 * CIN), k = q - co * (27 * CIN);
        const float g = grad[(size_t)co * LD + k];
        const float m = b1 * mw[q] + (1.f - b1) * g;
        const float v = b2 * vw[q] + (1.f - b2) * g * g;
        mw[q] = m;
        vw[q] = v;
        const float wn = w[q] - lr * (m / c1) / (sqrtf(v / c2) + eps);
        w[q] = wn;
        const bf16 hb = to_bf16(wn);
        Wb[(size_t)co * KP + k] = hb;
        const int t = k / CIN, ci = k - t * CIN;
        Wt[(size_t)ci * KPt + (26 - t) * COUT + co] = hb;
    } else if (q < COUT * 27 * CIN + COUT) {
        const int co = q - COUT * 27 * CIN;
        const float g = grad[(size_t)co * LD + KP];
        const float m = b1 * mb[co] + (1.f - b1) * g;
        const float v = b2 * vb[co] + (1.f - b2) * g * g;
        mb[co] = m;
        vb[co] = v;
        b[co] = b[co] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    }
}

} // namespace nn
} // namespace lat

// ======================================================================
// host side: network state, training on pairs from the plain integrator,
// corrected coarse stepping with the plain integrator as fallback
// ======================================================================
#include <cstring>
#include <random>
#include <vector>

namespace lat {
namespace nn {

constexpr int NL = 4;             // layers
constexpr int WID = 32;           // hidden channels
constexpr int NIN = 8;            // input channels: 4 fluctuations, 4 uniform
constexpr int CINS[NL] = {NIN, WID, WID, WID};
constexpr int COUTS[NL] = {WID, WID, WID, 4};
constexpr int CM_BLOCKS = 512;    // blocks of the channel-mean reduction
constexpr int CM_SLOTS = 40;      // doubles of Net::cm

struct Layer {
    int cin, cout, kp, mp, kpt, mpt; // padded shapes of W [mp][kp] and Wt [mpt][kpt]
    float *w = nullptr, *b = nullptr, *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr; // FP32 master
    bf16 *wb = nullptr, *wt = nullptr;                                                           // BF16 copies
    float *grad = nullptr;                                                                       // [mp][kp+16]
};

struct Net {
    Layer L[NL];
    int step = 0;
    // normalisation per channel, from the snapshot at hand (computed on the
    // GPU each time): input (v - <v>) / s_v, target (r - <r>) / (ratio s_v),
    // with <.> the spatial mean and s_v the spatial standard deviation of the
    // coarse state.  The network sees and corrects fluctuations only, at
    // unit scale, so neither the drift of the homogeneous mode nor the growth
    // of the fluctuations over a run leaves the range it was trained on.
    // ratio = RMS of the residual fluctuation / s_v, fixed at the first pair.
    double ratio[4] = {1, 1, 1, 1};
    bool stats = false;
    // activation buffers for lattice size n
    int n = 0;
    bf16 *x0 = nullptr, *a[3] = {nullptr, nullptr, nullptr}, *z[3] = {nullptr, nullptr, nullptr};
    bf16 *g4 = nullptr, *gA = nullptr, *gB = nullptr;
    float *y = nullptr, *tgt = nullptr, *part = nullptr, *lossp = nullptr;
    // channel statistics: [0,4) input mean, [4,8) residual mean, [8,12)
    // output mean, [12,16) input standard deviation, [16,20) residual
    // standard deviation (first training pair), [20,24) zeros, [24,28) RMS
    // of a judged step's difference from the plain step, [28,32) mean and
    // [32,36) standard deviation of the plain step, [36,40) RMS of the
    // leapfrog step's difference from it; block partials
    double *cm = nullptr, *cpart = nullptr;
    void *save_f = nullptr, *save_p = nullptr, *coarse_f = nullptr, *coarse_p = nullptr;
    void *leap_f = nullptr, *leap_p = nullptr; // judged steps: the leapfrog step alone
    DevScalars *save_s = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
};

static size_t layer_wcount(const Layer &l) { return (size_t)l.cout * 27 * l.cin; }

} // namespace nn
} // namespace lat

struct lat_corrector {
    lat::nn::Net net;
};

namespace lat {
namespace nn {

template <typename T>
static int alloc(Net &net, T **p, size_t count, const char *what)
{
    if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess) {
        net.err = std::string("hipMalloc ") + what;
        return -3;
    }
    return 0;
}

static void free_buffers(Net &net)
{
    void *ps[] = {net.x0, net.a[0], net.a[1], net.a[2], net.z[0], net.z[1], net.z[2], net.g4, net.gA, net.gB,
                  net.y, net.tgt, net.part, net.lossp, net.save_f, net.save_p, net.coarse_f, net.coarse_p,
                  net.leap_f, net.leap_p, net.save_s, net.cm, net.cpart};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    net.x0 = net.g4 = net.gA = net.gB = nullptr;
    for (int q = 0; q < 3; ++q) net.a[q] = net.z[q] = nullptr;
    net.y = net.tgt = net.part = net.lossp = nullptr;
    net.cm = net.cpart = nullptr;
    net.save_f = net.save_p = net.coarse_f = net.coarse_p = net.leap_f = net.leap_p = nullptr;
    net.save_s = nullptr;
    net.n = 0;
}

static size_t nchunks(size_t nvol) { return (nvol + GCHUNK - 1) / GCHUNK; }

static int ensure_buffers(Net &net, const Ctx &c)
{
    if (net.n == c.nx) return 0;
    free_buffers(net);
    const size_t nv = c.nvol;
    int rc = 0;
    rc |= alloc(net, &net.x0, nv * NIN, "x0");
    for (int q = 0; q < 3; ++q) {
        rc |= alloc(net, &net.a[q], nv * WID, "act");
        rc |= alloc(net, &net.z[q], nv * WID, "preact");
    }
    rc |= alloc(net, &net.g4, nv * 4, "g4");
    rc |= alloc(net, &net.gA, nv * WID, "gradA");
    rc |= alloc(net, &net.gB, nv * WID, "gradB");
    rc |= alloc(net, &net.y, nv * 4, "out");
    rc |= alloc(net, &net.tgt, nv * 4, "target");
    size_t maxpart = 0;
    for (int l = 0; l < NL; ++l) {
        const size_t q = (size_t)net.L[l].mp * (net.L[l].kp + 16);
        maxpart = q > maxpart ? q : maxpart;
    }
    rc |= alloc(net, &net.part, maxpart * nchunks(nv), "dW partials");
    rc |= alloc(net, &net.lossp, 1024, "loss partials");
    rc |= alloc(net, &net.cm, CM_SLOTS, "channel statistics");
    rc |= alloc(net, &net.cpart, 4 * CM_BLOCKS, "channel-mean partials");
    const size_t es = elem_size(c);
    rc |= alloc(net, (char **)&net.save_f, es * c.ms.f, "saved fields");
    rc |= alloc(net, (char **)&net.save_p, es * c.ms.f, "saved momenta");
    rc |= alloc(net, (char **)&net.coarse_f, es * c.ms.f, "coarse fields");
    rc |= alloc(net, (char **)&net.coarse_p, es * c.ms.f, "coarse momenta");
    rc |= alloc(net, (char **)&net.leap_f, es * c.ms.f, "leapfrog fields");
    rc |= alloc(net, (char **)&net.leap_p, es * c.ms.f, "leapfrog momenta");
    rc |= alloc(net, &net.save_s, 1, "saved scalars");
    if (rc) {
        free_buffers(net);
        return rc;
    }
    const double zero[CM_SLOTS] = {};
    if (hipMemcpy(net.cm, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) {
        free_buffers(net);
        net.err = "channel statistics upload failed";
        return -3;
    }
    net.n = c.nx;
    return 0;
}

// ---------------------------------------------------------------- data kernels
// Spatial channel sums, deterministic: CM_BLOCKS fixed blocks write partials,
// one block adds them in block order.  Source: the FP64 state (f, p), minus
// (cf, cp) when given, or the network output y [site][4]; with sub, the sums
// are of (v - sub)^2 instead (second pass of a standard deviation).
__global__ __launch_bounds__(256) void k_chsum(const double *__restrict__ f, const double *__restrict__ p,
                                               const double *__restrict__ cf, const double *__restrict__ cp,
                                               const float *__restrict__ y, const double *__restrict__ sub,
                                               size_t nvol, double *__restrict__ part)
{
    double a[4] = {0, 0, 0, 0};
    double m[4] = {0, 0, 0, 0};
    if (sub) {
#pragma unroll
        for (int q = 0; q < 4; ++q) m[q] = sub[q];
    }
    for (size_t s = blockIdx.x * (size_t)256 + threadIdx.x; s < nvol; s += (size_t)gridDim.x * 256) {
        double v[4];
        if (y) {
            const float4 t = *reinterpret_cast<const float4 *>(y + s * 4);
            v[0] = t.x;
            v[1] = t.y;
            v[2] = t.z;
            v[3] = t.w;
        } else {
            v[0] = cf ? f[s] - cf[s] : f[s];
            v[1] = cf ? f[nvol + s] - cf[nvol + s] : f[nvol + s];
            v[2] = cp ? p[s] - cp[s] : p[s];
            v[3] = cp ? p[nvol + s] - cp[nvol + s] : p[nvol + s];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] += sub ? (v[q] - m[q]) * (v[q] - m[q]) : v[q];
    }
    __shared__ double red[4][256];
#pragma unroll
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

// out[c] = sum / nvol, or with root the standard deviation sqrt(sum / nvol)
// (root 1: 1 where it is 0, so a channel without fluctuations is left
// unscaled; root 2: the RMS as it is)
__global__ __launch_bounds__(64) void k_chmean(const double *__restrict__ part, int nb, size_t nvol, int root,
                                               double *__restrict__ out)
{
    if (threadIdx.x < 4) {
        double t = 0;
        for (int b = 0; b < nb; ++b) t += part[b * 4 + threadIdx.x];
        t /= (double)nvol;
        if (root) {
            t = sqrt(t);
            if (root == 1 && !(t > 0.0)) t = 1.0;
        }
        out[threadIdx.x] = t;
    }
}

// X0[s][c] = (v_c(s) - <v_c>) / s_c for (phi, chi, pi_phi, pi_chi), then
// the uniform channels <phi>, <pi_phi>, lg2 = log10(g^2/m^2) - 5 and
// log10(s_chi) / 10; st = channel statistics (means at 0, standard
// deviations at 12)
__global__ __launch_bounds__(256) void k_pack(const double *__restrict__ f, const double *__restrict__ p,
                                              size_t nvol, const double *__restrict__ st, float lg2,
                                              bf16 *__restrict__ X)
{
    const double4 mean = {st[0], st[1], st[2], st[3]};
    const double4 isd = {1.0 / st[12], 1.0 / st[13], 1.0 / st[14], 1.0 / st[15]};
    const short u0 = (short)to_bf16((float)st[0]), u1 = (short)to_bf16((float)st[2]);
    const short u2 = (short)to_bf16(lg2), u3 = (short)to_bf16((float)(log10(st[13]) / 10.0));
    for (size_t s = blockIdx.x * (size_t)256 + threadIdx.x; s < nvol; s += (size_t)gridDim.x * 256) {
        s16x8 o;
        o[0] = (short)to_bf16((float)((f[s] - mean.x) * isd.x));
        o[1] = (short)to_bf16((float)((f[nvol + s] - mean.y) * isd.y));
        o[2] = (short)to_bf16((float)((p[s] - mean.z) * isd.z));
        o[3] = (short)to_bf16((float)((p[nvol + s] - mean.w) * isd.w));
        o[4] = u0;
        o[5] = u1;
        o[6] = u2;
        o[7] = u3;
        *reinterpret_cast<s16x8 *>(X + s * NIN) = o;
    }
}

// tgt[s][c] = (r_c(s) - <r_c>) / (ratio_c s_c) with r = fine - coarse
__global__ __launch_bounds__(256) void k_target(const double *__restrict__ ff, const double *__restrict__ fp,
                                                const double *__restrict__ cf, const double *__restrict__ cp,
                                                size_t nvol, const double *__restrict__ st, double4 ratio,
                                                float *__restrict__ tgt)
{
    const double4 m = {st[4], st[5], st[6], st[7]};
    const double4 irsd = {1.0 / (ratio.x * st[12]), 1.0 / (ratio.y * st[13]), 1.0 / (ratio.z * st[14]),
                          1.0 / (ratio.w * st[15])};
    for (size_t s = blockIdx.x * (size_t)256 + threadIdx.x; s < nvol; s += (size_t)gridDim.x * 256) {
        float4 o;
        o.x = (float)((ff[s] - cf[s] - m.x) * irsd.x);
        o.y = (float)((ff[nvol + s] - cf[nvol + s] - m.y) * irsd.y);
        o.z = (float)((fp[s] - cp[s] - m.z) * irsd.z);
        o.w = (float)((fp[nvol + s] - cp[nvol + s] - m.w) * irsd.w);
        *reinterpret_cast<float4 *>(tgt + s * 4) = o;
    }
}

// G4 = 2 (y - tgt) / (4 nvol) in BF16; per-block partial sums of (y - tgt)^2
__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ y, const float *__restrict__ tgt,
                                              size_t nvol, bf16 *__restrict__ g, float *__restrict__ part)
{
    __shared__ float red[256];
    const float scale = 2.0f / (4.0f * (float)nvol);
    float acc = 0.f;
    for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < nvol * 4; q += (size_t)gridDim.x * 256) {
        const float d = y[q] - tgt[q];
        acc += d * d;
        g[q] = to_bf16(d * scale);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// f_c += ratio_c s_c (y_c - <y_c>): the predicted residual fluctuation; the
// homogeneous mode keeps the coarse step's value
__global__ __launch_bounds__(256) void k_apply(double *__restrict__ f, double *__restrict__ p,
                                               const float *__restrict__ y, size_t nvol,
                                               const double *__restrict__ st, double4 ratio)
{
    const double4 m = {st[8], st[9], st[10], st[11]};
    const double4 rsd = {ratio.x * st[12], ratio.y * st[13], ratio.z * st[14], ratio.w * st[15]};
    for (size_t s = blockIdx.x * (size_t)256 + threadIdx.x; s < nvol; s += (size_t)gridDim.x * 256) {
        const float4 v = *reinterpret_cast<const float4 *>(y + s * 4);
        f[s] += rsd.x * (v.x - m.x);
        f[nvol + s] += rsd.y * (v.y - m.y);
        p[s] += rsd.z * (v.z - m.z);
        p[nvol + s] += rsd.w * (v.w - m.w);
    }
}

// ---------------------------------------------------------------- network passes
template <int CIN, int COUT, int MODE>
static void conv(Net &net, const bf16 *X, const bf16 *W, const float *bias, bf16 *Y, float *Yf, bf16 *Z,
                 const bf16 *Zin)
{
    const size_t ntile = (size_t)net.n * net.n * net.n / 16;
    size_t g = (ntile + CONV_WAVES - 1) / CONV_WAVES;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL((k_conv<CIN, COUT, MODE>), dim3((unsigned)g), dim3(64 * CONV_WAVES), 0, net.stream, X, W, bias,
                       Y, Yf, Z, Zin, net.n);
}

static void forward(Net &net)
{
    conv<NIN, WID, CONV_SILU>(net, net.x0, net.L[0].wb, net.L[0].b, net.a[0], nullptr, net.z[0], nullptr);
    conv<WID, WID, CONV_SILU>(net, net.a[0], net.L[1].wb, net.L[1].b, net.a[1], nullptr, net.z[1], nullptr);
    conv<WID, WID, CONV_SILU>(net, net.a[1], net.L[2].wb, net.L[2].b, net.a[2], nullptr, net.z[2], nullptr);
    conv<WID, 4, CONV_LINEAR>(net, net.a[2], net.L[3].wb, net.L[3].b, nullptr, net.y, nullptr, nullptr);
}

template <int CIN, int COUT>
static void weight_grad(Net &net, Layer &l, const bf16 *G, const bf16 *X)
{
    const size_t nv = (size_t)net.n * net.n * net.n;
    const int MT = (COUT + 15) / 16, CT = KPAD(CIN) / 16 + 1;
    // one partial per block: at most the nchunks(nv) partials net.part holds
    size_t nch = nchunks(nv);
    if (nch > DW2_BLOCKS) nch = DW2_BLOCKS;
    (void)MT;
    (void)CT;
    hipLaunchKernelGGL((k_conv_dw2<CIN, COUT>), dim3((unsigned)nch), dim3(64 * DW2_WAVES), 0, net.stream, G, X,
                       net.part, net.n);
    const int nq = l.mp * (l.kp + 16);
    hipLaunchKernelGGL(k_sum_chunks, dim3((nq + 255) / 256), dim3(256), 0, net.stream, net.part, (int)nch, nq,
                       l.grad);
}

// output gradient G4 and the loss partials for the current y / tgt
static void loss_grad(Net &net)
{
    const size_t nv = (size_t)net.n * net.n * net.n;
    hipLaunchKernelGGL(k_loss, dim3(1024), dim3(256), 0, net.stream, net.y, net.tgt, nv, net.g4, net.lossp);
}

// gradients of all layers from G4; returns after launching
static void backward_layers(Net &net)
{
    weight_grad<WID, 4>(net, net.L[3], net.g4, net.a[2]);
    conv<4, WID, CONV_DSILU>(net, net.g4, net.L[3].wt, nullptr, net.gA, nullptr, nullptr, net.z[2]);
    weight_grad<WID, WID>(net, net.L[2], net.gA, net.a[1]);
    conv<WID, WID, CONV_DSILU>(net, net.gA, net.L[2].wt, nullptr, net.gB, nullptr, nullptr, net.z[1]);
    weight_grad<WID, WID>(net, net.L[1], net.gB, net.a[0]);
    conv<WID, WID, CONV_DSILU>(net, net.gB, net.L[1].wt, nullptr, net.gA, nullptr, nullptr, net.z[0]);
    weight_grad<NIN, WID>(net, net.L[0], net.gA, net.x0);
}

static void backward(Net &net)
{
    loss_grad(net);
    backward_layers(net);
}

template <int CIN, int COUT>
static void adam_layer(Net &net, Layer &l, float lr, float c1, float c2)
{
    const int nq = COUT * 27 * CIN + COUT;
    hipLaunchKernelGGL((k_adam<CIN, COUT>), dim3((nq + 255) / 256), dim3(256), 0, net.stream, l.w, l.b, l.mw, l.vw,
                       l.mb, l.vb, l.grad, lr, 0.9f, 0.999f, 1e-8f, c1, c2, l.wb, l.wt);
}

static void adam(Net &net, float lr)
{
    ++net.step;
    const float c1 = 1.f - powf(0.9f, (float)net.step), c2 = 1.f - powf(0.999f, (float)net.step);
    adam_layer<NIN, WID>(net, net.L[0], lr, c1, c2);
    adam_layer<WID, WID>(net, net.L[1], lr, c1, c2);
    adam_layer<WID, WID>(net, net.L[2], lr, c1, c2);
    adam_layer<WID, 4>(net, net.L[3], lr, c1, c2);
}

// Upload FP32 master weights (cout x 27 cin, k = tap*cin + ci) and biases of
// layer l, rebuilding the BF16 copies; the Adam moments are reset.
static int set_layer(Net &net, int li, const float *w, const float *b)
{
    Layer &l = net.L[li];
    const size_t nw = layer_wcount(l);
    std::vector<bf16> wb((size_t)l.mp * l.kp, 0), wt((size_t)l.mpt * l.kpt, 0);
    for (int co = 0; co < l.cout; ++co)
        for (int k = 0; k < 27 * l.cin; ++k) {
            const bf16 h = to_bf16(w[(size_t)co * 27 * l.cin + k]);
            wb[(size_t)co * l.kp + k] = h;
            const int t = k / l.cin, ci = k - t * l.cin;
            wt[(size_t)ci * l.kpt + (26 - t) * l.cout + co] = h;
        }
    std::vector<float> zeros(nw > (size_t)l.cout ? nw : l.cout, 0.f);
    bool ok = hipMemcpy(l.w, w, sizeof(float) * nw, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.b, b, sizeof(float) * l.cout, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.mw, zeros.data(), sizeof(float) * nw, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.vw, zeros.data(), sizeof(float) * nw, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.mb, zeros.data(), sizeof(float) * l.cout, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.vb, zeros.data(), sizeof(float) * l.cout, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.wb, wb.data(), sizeof(bf16) * wb.size(), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(l.wt, wt.data(), sizeof(bf16) * wt.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        net.err = "weight upload failed";
        return -3;
    }
    return 0;
}

} // namespace nn
} // namespace lat

// ---------------------------------------------------------------- C ABI
using namespace lat;
using namespace lat::nn;

static int corr_fail(Net &net, const char *msg)
{
    net.err = msg;
    return -2;
}

// A call borrows its context's stream: it ends with the stream drained and
// the borrow cleared, so the corrector never keeps the stream of a context
// that may be destroyed before it.
struct StreamLease {
    Net &net;
    ~StreamLease()
    {
        if (net.stream) (void)hipStreamSynchronize(net.stream);
        net.stream = nullptr;
    }
};

static int check_ctx(Net &net, Ctx &c)
{
    if (c.B != 1 || c.view) return corr_fail(net, "the corrector runs on single contexts");
    if (c.prec != LAT_F64 || c.nfld != 2) return corr_fail(net, "the corrector needs a two-field FP64 context");
    if (c.nx != c.ny || c.nx != c.nz || c.nx % 16) return corr_fail(net, "the corrector needs a cubic lattice, n % 16 == 0");
    net.stream = c.stream;
    return ensure_buffers(net, c);
}

static double4 d4(const double *v) { return double4{v[0], v[1], v[2], v[3]}; }

static int grid_sites(size_t nvol)
{
    size_t g = (nvol + 255) / 256;
    return (int)(g > 8192 ? 8192 : g);
}

// channel means into net.cm + 4 * slot (see Net::cm)
static void channel_means(Net &net, Ctx &c, const void *f, const void *p, const void *cf, const void *cp,
                          const float *y, int slot)
{
    hipLaunchKernelGGL(k_chsum, dim3(CM_BLOCKS), dim3(256), 0, c.stream, (const double *)f, (const double *)p,
                       (const double *)cf, (const double *)cp, y, (const double *)nullptr, c.nvol, net.cpart);
    hipLaunchKernelGGL(k_chmean, dim3(1), dim3(64), 0, c.stream, net.cpart, CM_BLOCKS, c.nvol, 0,
                       net.cm + 4 * slot);
}

// input means (slot 0) and standard deviations (slot 3) of the state (f, p), then X0
static void pack(Net &net, Ctx &c, const void *f, const void *p)
{
    channel_means(net, c, f, p, nullptr, nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_chsum, dim3(CM_BLOCKS), dim3(256), 0, c.stream, (const double *)f, (const double *)p,
                       (const double *)nullptr, (const double *)nullptr, (const float *)nullptr,
                       (const double *)net.cm, c.nvol, net.cpart);
    hipLaunchKernelGGL(k_chmean, dim3(1), dim3(64), 0, c.stream, net.cpart, CM_BLOCKS, c.nvol, 1, net.cm + 12);
    const double g2 = c.cfg.model_params[1];
    const float lg2 = g2 > 0.0 ? (float)(std::log10(g2) - 5.0) : 0.0f;
    hipLaunchKernelGGL(k_pack, dim3(grid_sites(c.nvol)), dim3(256), 0, c.stream, (const double *)f,
                       (const double *)p, c.nvol, (const double *)net.cm, lg2, net.x0);
}

static int save_state(Net &net, Ctx &c)
{
    const size_t bytes = sizeof(double) * c.ms.f;
    LAT_HIP(&c, hipMemcpyAsync(net.save_f, c.fld, bytes, hipMemcpyDeviceToDevice, c.stream));
    LAT_HIP(&c, hipMemcpyAsync(net.save_p, c.fldp, bytes, hipMemcpyDeviceToDevice, c.stream));
    LAT_HIP(&c, hipMemcpyAsync(net.save_s, c.dsc, sizeof(DevScalars), hipMemcpyDeviceToDevice, c.stream));
    return 0;
}

static int restore_state(Net &net, Ctx &c)
{
    const size_t bytes = sizeof(double) * c.ms.f;
    LAT_HIP(&c, hipMemcpyAsync(c.fld, net.save_f, bytes, hipMemcpyDeviceToDevice, c.stream));
    LAT_HIP(&c, hipMemcpyAsync(c.fldp, net.save_p, bytes, hipMemcpyDeviceToDevice, c.stream));
    LAT_HIP(&c, hipMemcpyAsync(c.dsc, net.save_s, sizeof(DevScalars), hipMemcpyDeviceToDevice, c.stream));
    c.ke_valid = false;
    return 0;
}

extern "C" {

int lat_corrector_create(unsigned long long seed, lat_corrector **out)
{
    if (!out) return -1;
    *out = nullptr;
    lat_corrector *h = new lat_corrector();
    Net &net = h->net;
    std::mt19937_64 rng(seed);
    for (int li = 0; li < NL; ++li) {
        Layer &l = net.L[li];
        l.cin = CINS[li];
        l.cout = COUTS[li];
        l.kp = KPAD(l.cin);
        l.mp = (l.cout + 15) / 16 * 16;
        l.kpt = KPAD(l.cout);
        l.mpt = (l.cin + 15) / 16 * 16;
        const size_t nw = layer_wcount(l);
        int rc = 0;
        rc |= alloc(net, &l.w, nw, "w");
        rc |= alloc(net, &l.mw, nw, "adam m");
        rc |= alloc(net, &l.vw, nw, "adam v");
        rc |= alloc(net, &l.b, l.cout, "b");
        rc |= alloc(net, &l.mb, l.cout, "adam mb");
        rc |= alloc(net, &l.vb, l.cout, "adam vb");
        rc |= alloc(net, &l.wb, (size_t)l.mp * l.kp, "w bf16");
        rc |= alloc(net, &l.wt, (size_t)l.mpt * l.kpt, "wt bf16");
        rc |= alloc(net, &l.grad, (size_t)l.mp * (l.kp + 16), "grad");
        if (rc) {
            lat_corrector_destroy(h);
            return -3;
        }
        // He-normal initialisation; the last layer starts small so that the
        // untrained corrector is close to the plain coarse step
        const double sd = std::sqrt(2.0 / (27.0 * l.cin)) * (li == NL - 1 ? 0.1 : 1.0);
        std::normal_distribution<double> nd(0.0, sd);
        std::vector<float> w(nw), b(l.cout, 0.f);
        for (auto &x : w) x = (float)nd(rng);
        if (set_layer(net, li, w.data(), b.data())) {
            lat_corrector_destroy(h);
            return -3;
        }
    }
    *out = h;
    return 0;
}

void lat_corrector_destroy(lat_corrector *h)
{
    if (!h) return;
    Net &net = h->net;
    if (net.stream) (void)hipStreamSynchronize(net.stream);
    free_buffers(net);
    for (auto &l : net.L) {
        void *ps[] = {l.w, l.b, l.mw, l.vw, l.mb, l.vb, l.wb, l.wt, l.grad};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
    delete h;
}

const char *lat_corrector_last_error(const lat_corrector *h) { return h ? h->net.err.c_str() : "null corrector"; }

int lat_corrector_set_layer(lat_corrector *h, int layer, const float *w, const float *b)
{
    if (!h || !w || !b || layer < 0 || layer >= NL) return -1;
    return set_layer(h->net, layer, w, b);
}

int lat_corrector_get_layer(lat_corrector *h, int layer, float *w, float *b)
{
    if (!h || !w || !b || layer < 0 || layer >= NL) return -1;
    Layer &l = h->net.L[layer];
    if (h->net.stream) (void)hipStreamSynchronize(h->net.stream);
    if (hipMemcpy(w, l.w, sizeof(float) * layer_wcount(l), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b, l.b, sizeof(float) * l.cout, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

int lat_corrector_set_norm(lat_corrector *h, const double *ratio)
{
    if (!h || !ratio) return -1;
    for (int q = 0; q < 4; ++q) {
        if (!(ratio[q] > 0.0)) return -1;
        h->net.ratio[q] = ratio[q];
    }
    h->net.stats = true;
    return 0;
}

int lat_corrector_get_norm(const lat_corrector *h, double *ratio)
{
    if (!h || !ratio) return -1;
    for (int q = 0; q < 4; ++q) {
        ratio[q] = h->net.ratio[q];
    }
    return 0;
}

// Network output (normalised residual, [site][4] FP32) for the current state of ctx.
int lat_corrector_infer(lat_corrector *h, lat_ctx *ctx, float *out)
{
    if (!h || !ctx || !out) return -1;
    Net &net = h->net;
    StreamLease lease{net};
    Ctx &c = ctx->c;
    int rc = check_ctx(net, c);
    if (rc) return rc;
    pack(net, c, c.fld, c.fldp);
    forward(net);
    LAT_HIP(&c, hipMemcpyAsync(out, net.y, sizeof(float) * 4 * c.nvol, hipMemcpyDeviceToHost, c.stream));
    LAT_HIP(&c, hipStreamSynchronize(c.stream));
    return 0;
}

// Loss and gradients for the current state of ctx as input and `target`
// ([site][4] FP32, normalised residual): *loss = mean squared error; grads
// (nullable) receives, per layer, cout x 27 cin weight gradients then cout
// bias gradients.  No weight update.
int lat_corrector_gradients(lat_corrector *h, lat_ctx *ctx, const float *target, double *loss, float *grads)
{
    if (!h || !ctx || !target) return -1;
    Net &net = h->net;
    StreamLease lease{net};
    Ctx &c = ctx->c;
    int rc = check_ctx(net, c);
    if (rc) return rc;
    LAT_HIP(&c, hipMemcpyAsync(net.tgt, target, sizeof(float) * 4 * c.nvol, hipMemcpyHostToDevice, c.stream));
    pack(net, c, c.fld, c.fldp);
    forward(net);
    backward(net);
    std::vector<float> lp(1024);
    LAT_HIP(&c, hipMemcpyAsync(lp.data(), net.lossp, sizeof(float) * 1024, hipMemcpyDeviceToHost, c.stream));
    LAT_HIP(&c, hipStreamSynchronize(c.stream));
    double s = 0;
    for (float v : lp) s += v;
    if (loss) *loss = s / (4.0 * (double)c.nvol);
    if (grads) {
        size_t off = 0;
        for (auto &l : net.L) {
            std::vector<float> g((size_t)l.mp * (l.kp + 16));
            LAT_HIP(&c, hipMemcpy(g.data(), l.grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost));
            for (int co = 0; co < l.cout; ++co)
                for (int k = 0; k < 27 * l.cin; ++k) grads[off++] = g[(size_t)co * (l.kp + 16) + k];
            for (int co = 0; co < l.cout; ++co) grads[off++] = g[(size_t)co * (l.kp + 16) + l.kp];
        }
    }
    return 0;
}

// Train on npairs snapshot pairs generated along the trajectory of ctx by the
// plain integrator: from the current state S, coarse = one leapfrog step of
// coarse_factor * dt, fine = coarse_factor Yoshida-4 steps of dt; input =
// coarse, target = its residual fluctuation over (ratio x the coarse
// state's standard deviation), ratio fixed at the first pair of a fresh
// corrector; one Adam step per pair; the
// trajectory continues from the fine state.  losses (nullable): per pair,
// NaN for a pair whose coarse step was not finite (no update from it).
int lat_corrector_train(lat_corrector *h, lat_ctx *ctx, double dt, int coarse_factor, int npairs, double lr,
                        double *losses)
{
    if (!h || !ctx || coarse_factor < 1 || npairs < 0) return -1;
    Net &net = h->net;
    StreamLease lease{net};
    Ctx &c = ctx->c;
    int rc = check_ctx(net, c);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * c.ms.f;
    for (int it = 0; it < npairs; ++it) {
        if ((rc = save_state(net, c))) return rc;
        if ((rc = lat_symp(ctx, 2, coarse_factor * dt, 1))) return rc;
        LAT_HIP(&c, hipMemcpyAsync(net.coarse_f, c.fld, bytes, hipMemcpyDeviceToDevice, c.stream));
        LAT_HIP(&c, hipMemcpyAsync(net.coarse_p, c.fldp, bytes, hipMemcpyDeviceToDevice, c.stream));
        if ((rc = restore_state(net, c))) return rc;
        if ((rc = lat_symp(ctx, 4, dt, coarse_factor))) return rc;
        pack(net, c, net.coarse_f, net.coarse_p);
        channel_means(net, c, c.fld, c.fldp, net.coarse_f, net.coarse_p, nullptr, 1);
        bool usable = true;
        if (!net.stats) {
            // ratio = residual fluctuation RMS / input standard deviation
            hipLaunchKernelGGL(k_chsum, dim3(CM_BLOCKS), dim3(256), 0, c.stream, (const double *)c.fld,
                               (const double *)c.fldp, (const double *)net.coarse_f, (const double *)net.coarse_p,
                               (const float *)nullptr, (const double *)(net.cm + 4), c.nvol, net.cpart);
            hipLaunchKernelGGL(k_chmean, dim3(1), dim3(64), 0, c.stream, net.cpart, CM_BLOCKS, c.nvol, 1,
                               net.cm + 16);
            double st[20];
            LAT_HIP(&c, hipMemcpyAsync(st, net.cm, sizeof(st), hipMemcpyDeviceToHost, c.stream));
            LAT_HIP(&c, hipStreamSynchronize(c.stream));
            double r[4];
            for (int q = 0; q < 4; ++q) {
                r[q] = st[16 + q] / st[12 + q];
                // a non-finite coarse step (beyond the leapfrog's stability
                // limit) gives no usable scale: wait for a usable pair
                if (!std::isfinite(r[q]) || !std::isfinite(st[q]) || !std::isfinite(st[4 + q])) usable = false;
                if (!(r[q] > 0.0)) r[q] = 1.0;
            }
            if (usable) {
                for (int q = 0; q < 4; ++q) net.ratio[q] = r[q];
                net.stats = true;
            }
        }
        if (!usable) {
            if (losses) losses[it] = std::nan("");
            continue;
        }
        hipLaunchKernelGGL(k_target, dim3(grid_sites(c.nvol)), dim3(256), 0, c.stream, (const double *)c.fld,
                           (const double *)c.fldp, (const double *)net.coarse_f, (const double *)net.coarse_p,
                           c.nvol, (const double *)net.cm, d4(net.ratio), net.tgt);
        forward(net);
        loss_grad(net);
        std::vector<float> lp(1024);
        LAT_HIP(&c, hipMemcpyAsync(lp.data(), net.lossp, sizeof(float) * 1024, hipMemcpyDeviceToHost, c.stream));
        LAT_HIP(&c, hipStreamSynchronize(c.stream));
        double s = 0;
        for (float v : lp) s += v;
        const double loss = s / (4.0 * (double)c.nvol);
        if (losses) losses[it] = loss;
        // a pair whose coarse step left the stability region (non-finite
        // loss) would poison the weights: no update from it
        if (!std::isfinite(loss)) continue;
        backward_layers(net);
        adam(net, (float)lr);
    }
    LAT_HIP(&c, hipGetLastError());
    return 0;
}

// nsteps corrected coarse steps: save S, leapfrog(coarse_factor * dt), add
// the predicted residual, then check the Friedmann constraint: if
// |resid_after - resid_before| > threshold, restore S and take
// coarse_factor Yoshida-4 steps of dt instead (the plain integrator is the
// fallback).  stats (nullable, 3 doubles): fallbacks, max accepted
// constraint change, force evaluations used (a coarse step costs 1, a
// Yoshida-4 step 3).
int lat_corrector_step(lat_corrector *h, lat_ctx *ctx, double dt, int coarse_factor, int nsteps, double threshold,
                       double *stats)
{
    if (!h || !ctx || coarse_factor < 1 || nsteps < 0) return -1;
    Net &net = h->net;
    StreamLease lease{net};
    Ctx &c = ctx->c;
    if (!net.stats) return corr_fail(net, "the corrector has not been trained (no normalisation)");
    int rc = check_ctx(net, c);
    if (rc) return rc;
    double fb = 0, maxv = 0, forces = 0, e[LAT_NE];
    for (int it = 0; it < nsteps; ++it) {
        if ((rc = lat_energies(ctx, e))) return rc;
        const double r0 = e[LAT_E_RESID];
        if ((rc = save_state(net, c))) return rc;
        if ((rc = lat_symp(ctx, 2, coarse_factor * dt, 1))) return rc;
        pack(net, c, c.fld, c.fldp);
        forward(net);
        channel_means(net, c, nullptr, nullptr, nullptr, nullptr, net.y, 2);
        hipLaunchKernelGGL(k_apply, dim3(grid_sites(c.nvol)), dim3(256), 0, c.stream, (double *)c.fld,
                           (double *)c.fldp, net.y, c.nvol, (const double *)net.cm, d4(net.ratio));
        c.ke_valid = false;
        forces += 1;
        if ((rc = lat_energies(ctx, e))) return rc;
        const double v = std::fabs(e[LAT_E_RESID] - r0);
        if (!(v <= threshold)) {
            if ((rc = restore_state(net, c))) return rc;
            if ((rc = lat_symp(ctx, 4, dt, coarse_factor))) return rc;
            fb += 1;
            forces += 3.0 * coarse_factor;
        } else if (v > maxv) {
            maxv = v;
        }
    }
    if (stats) {
        stats[0] = fb;
        stats[1] = maxv;
        stats[2] = forces;
    }
    return 0;
}

// As lat_corrector_step, and every judge_every-th accepted corrected step is
// also taken by the plain integrator (coarse_factor Yoshida-4 steps of dt
// from the same start): the errors of the corrected step and of the leapfrog
// step alone against it, per channel (phi, chi, pi_phi, pi_chi) the RMS
// difference over the plain step's spatial standard deviation, the largest
// over the channels, are recorded and the plain step is kept.  Once a judged
// corrected step's error exceeds err_tol, or that of the leapfrog step
// alone, the corrector is switched off and every remaining step is the plain
// integrator's.  stats (nullable, 7 doubles): fallbacks (constraint), force
// evaluations, judged steps, largest judged error of the corrected step,
// step index at which the corrector was switched off (-1: never), largest
// accepted constraint change, largest judged error of the leapfrog step.
int lat_corrector_step_judged(lat_corrector *h, lat_ctx *ctx, double dt, int coarse_factor, int nsteps,
                              double threshold, int judge_every, double err_tol, double *stats)
{
    if (!h || !ctx || coarse_factor < 1 || nsteps < 0 || judge_every < 1) return -1;
    Net &net = h->net;
    StreamLease lease{net};
    Ctx &c = ctx->c;
    if (!net.stats) return corr_fail(net, "the corrector has not been trained (no normalisation)");
    int rc = check_ctx(net, c);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * c.ms.f;
    double fb = 0, forces = 0, judged = 0, maxerr = 0, maxleap = 0, off_at = -1, maxv = 0, e[LAT_NE];
    int accepted = 0;
    // largest over the channels of RMS(a - plain) / std(plain), the plain
    // step being the current state and its standard deviations at cm + 32
    auto judge_err = [&](const void *af, const void *ap, double *out) -> int {
        hipLaunchKernelGGL(k_chsum, dim3(CM_BLOCKS), dim3(256), 0, c.stream, (const double *)c.fld,
                           (const double *)c.fldp, (const double *)af, (const double *)ap, (const float *)nullptr,
                           (const double *)(net.cm + 20), c.nvol, net.cpart);
        hipLaunchKernelGGL(k_chmean, dim3(1), dim3(64), 0, c.stream, net.cpart, CM_BLOCKS, c.nvol, 2, net.cm + 24);
        double jd[12];
        LAT_HIP(&c, hipMemcpyAsync(jd, net.cm + 24, sizeof(jd), hipMemcpyDeviceToHost, c.stream));
        LAT_HIP(&c, hipStreamSynchronize(c.stream));
        double err = 0;
        for (int q = 0; q < 4; ++q) {
            const double r = jd[q] / jd[8 + q];
            err = (r > err || !(r == r)) ? r : err;
        }
        *out = err;
        return 0;
    };
    for (int it = 0; it < nsteps; ++it) {
        if (off_at >= 0) { // switched off: the plain integrator
            if ((rc = lat_symp(ctx, 4, dt, coarse_factor))) return rc;
            forces += 3.0 * coarse_factor;
            continue;
        }
        if ((rc = lat_energies(ctx, e))) return rc;
        const double r0 = e[LAT_E_RESID];
        if ((rc = save_state(net, c))) return rc;
        if ((rc = lat_symp(ctx, 2, coarse_factor * dt, 1))) return rc;
        const bool judge = accepted % judge_every == 0; // if this step is accepted
        if (judge) {
            LAT_HIP(&c, hipMemcpyAsync(net.leap_f, c.fld, bytes, hipMemcpyDeviceToDevice, c.stream));
            LAT_HIP(&c, hipMemcpyAsync(net.leap_p, c.fldp, bytes, hipMemcpyDeviceToDevice, c.stream));
        }
        pack(net, c, c.fld, c.fldp);
        forward(net);
        channel_means(net, c, nullptr, nullptr, nullptr, nullptr, net.y, 2);
        hipLaunchKernelGGL(k_apply, dim3(grid_sites(c.nvol)), dim3(256), 0, c.stream, (double *)c.fld,
                           (double *)c.fldp, net.y, c.nvol, (const double *)net.cm, d4(net.ratio));
        c.ke_valid = false;
        forces += 1;
        if ((rc = lat_energies(ctx, e))) return rc;
        const double v = std::fabs(e[LAT_E_RESID] - r0);
        if (!(v <= threshold)) {
            if ((rc = restore_state(net, c))) return rc;
            if ((rc = lat_symp(ctx, 4, dt, coarse_factor))) return rc;
            fb += 1;
            forces += 3.0 * coarse_factor;
            continue;
        }
        maxv = v > maxv ? v : maxv;
        ++accepted;
        if (!judge) continue;
        // judge: the plain step from the same start
        LAT_HIP(&c, hipMemcpyAsync(net.coarse_f, c.fld, bytes, hipMemcpyDeviceToDevice, c.stream));
        LAT_HIP(&c, hipMemcpyAsync(net.coarse_p, c.fldp, bytes, hipMemcpyDeviceToDevice, c.stream));
        if ((rc = restore_state(net, c))) return rc;
        if ((rc = lat_symp(ctx, 4, dt, coarse_factor))) return rc;
        forces += 3.0 * coarse_factor;
        judged += 1;
        channel_means(net, c, c.fld, c.fldp, nullptr, nullptr, nullptr, 7);
        hipLaunchKernelGGL(k_chsum, dim3(CM_BLOCKS), dim3(256), 0, c.stream, (const double *)c.fld,
                           (const double *)c.fldp, (const double *)nullptr, (const double *)nullptr,
                           (const float *)nullptr, (const double *)(net.cm + 28), c.nvol, net.cpart);
        hipLaunchKernelGGL(k_chmean, dim3(1), dim3(64), 0, c.stream, net.cpart, CM_BLOCKS, c.nvol, 1, net.cm + 32);
        double ec = 0, el = 0;
        if ((rc = judge_err(net.coarse_f, net.coarse_p, &ec))) return rc;
        if ((rc = judge_err(net.leap_f, net.leap_p, &el))) return rc;
        maxerr = (ec > maxerr || !(ec == ec)) ? ec : maxerr;
        maxleap = (el > maxleap || !(el == el)) ? el : maxleap;
        if (!(ec <= err_tol) || !(ec <= el)) off_at = it;
    }
    if (stats) {
        stats[0] = fb;
        stats[1] = forces;
        stats[2] = judged;
        stats[3] = maxerr;
        stats[4] = off_at;
        stats[5] = maxv;
        stats[6] = maxleap;
    }
    return 0;
}

} // extern "C"
