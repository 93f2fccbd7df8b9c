// model.h -- potential plug-ins of the evolution path.
//
//   LAT_MODEL_PHI4:     V = lambda phi^4 / 4           (hamiltonian_conformal.f90:215-240)
//   LAT_MODEL_TWOFIELD: V = m2 phi^2/2 + g2 phi^2 chi^2/2 + m2chi chi^2/2
//                       (hamiltonian_cosmic.f90:215-232; the chi mass term of
//                       Discrete/Entropy/model.f90:35-55, modelv = 2V, which
//                       the Pseudospec trees leave out: m2chi = 0 there)
//
// Full forms evaluate V and V' at phi (plain FP64 path).  Difference forms
// evaluate dV = V(pb + d) - V(pb) and dV' = V'(pb + d) - V'(pb) for a
// homogeneous part pb and a small fluctuation d, algebraically rearranged so
// that every term carries a factor of d: in FP32 they keep full relative
// precision when |d| << |pb| (the mixed form).
#pragma once
#include "lat_internal.h"

namespace lat {

template <typename T>
__device__ __forceinline__ void model_dv(const ModelDev &m, const T *f, T *dv)
{
    if (m.model == LAT_MODEL_PHI4) {
        dv[0] = T(m.p0) * f[0] * f[0] * f[0];
    } else {
        dv[0] = T(m.p0) * f[0] + T(m.p1) * f[0] * (f[1] * f[1]);
        dv[1] = T(m.p1) * (f[0] * f[0]) * f[1] + T(m.p2) * f[1];
    }
}

template <typename T>
__device__ __forceinline__ double model_v(const ModelDev &m, const T *f)
{
    const double a = f[0];
    if (m.model == LAT_MODEL_PHI4) return 0.25 * m.p0 * (a * a) * (a * a);
    const double b = f[1];
    return 0.5 * m.p0 * (a * a) + 0.5 * m.p1 * (a * a) * (b * b) + 0.5 * m.p2 * (b * b);
}

// dV' for homogeneous part pb[] and fluctuation d[] (FP32).
__device__ __forceinline__ void model_ddv(const ModelDev &m, const float *pb, const float *d, float *ddv)
{
    if (m.model == LAT_MODEL_PHI4) {
        const float p = pb[0] + d[0];
        ddv[0] = float(m.p0) * d[0] * (p * p + p * pb[0] + pb[0] * pb[0]);
    } else {
        const float p = pb[0] + d[0], c = pb[1] + d[1];
        const float g2 = float(m.p1);
        // V'_phi = m2 phi + g2 phi chi^2 ; V'_chi = g2 phi^2 chi
        ddv[0] = float(m.p0) * d[0] + g2 * (d[0] * (c * c) + pb[0] * d[1] * (c + pb[1]));
        ddv[1] = g2 * (d[0] * (p + pb[0]) * c + (pb[0] * pb[0]) * d[1]) + float(m.p2) * d[1];
    }
}

// dV for homogeneous part pb[] and fluctuation d[], accumulated in FP64.
__device__ __forceinline__ double model_dvv(const ModelDev &m, const double *pb, const float *d)
{
    if (m.model == LAT_MODEL_PHI4) {
        const double p = pb[0] + d[0], b = pb[0];
        return 0.25 * m.p0 * double(d[0]) * (p + b) * (p * p + b * b);
    }
    const double p = pb[0] + d[0], c = pb[1] + d[1];
    const double pbb = pb[0], cb = pb[1];
    // m2/2 (p^2 - pb^2) + g2/2 ((p c)^2 - (pb cb)^2), (p c - pb cb) = d0 c + pb d1
    const double dpc = double(d[0]) * c + pbb * double(d[1]);
    return 0.5 * m.p0 * double(d[0]) * (p + pbb) + 0.5 * m.p1 * dpc * (p * c + pbb * cb) +
           0.5 * m.p2 * double(d[1]) * (c + cb);
}

// FP64 V and V' at the homogeneous values (scalar kernel).
__device__ __forceinline__ double model_v64(const ModelDev &m, const double *f)
{
    if (m.model == LAT_MODEL_PHI4) return 0.25 * m.p0 * (f[0] * f[0]) * (f[0] * f[0]);
    return 0.5 * m.p0 * (f[0] * f[0]) + 0.5 * m.p1 * (f[0] * f[0]) * (f[1] * f[1]) + 0.5 * m.p2 * (f[1] * f[1]);
}

__device__ __forceinline__ void model_dv64(const ModelDev &m, const double *f, double *dv)
{
    if (m.model == LAT_MODEL_PHI4) {
        dv[0] = m.p0 * f[0] * f[0] * f[0];
        dv[1] = 0.0;
    } else {
        dv[0] = m.p0 * f[0] + m.p1 * f[0] * (f[1] * f[1]);
        dv[1] = m.p1 * (f[0] * f[0]) * f[1] + m.p2 * f[1];
    }
}

} // namespace lat
