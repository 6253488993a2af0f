// Device helpers shared by the HIP kernels: cutoff functions, integer
// powers, activations and 64-lane wave reductions.  All FP64.
#ifndef HDNNP_DEVICE_COMMON_H
#define HDNNP_DEVICE_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tables.h"

namespace hdnnp {

constexpr int kWave = 64;
constexpr double kPi = 3.141592653589793238462643383279502884;

// sin(pi x) and cos(pi x) for x in [0, 1].  With u = x - 1/2:
// cos(pi x) = -sin(pi u) and sin(pi x) = cos(pi u); both are evaluated from
// their Taylor series in t = u^2 (|u| <= 1/2), truncated after the first
// term below 1e-17.  The coefficients (-1)^k pi^(2k+1)/(2k+1)! and
// (-1)^k pi^(2k)/(2k)! were computed in 60-digit arithmetic and rounded to
// double.  Max abs error against libm over [0, 1]: 4.4e-16.
__device__ __forceinline__ void sincospi01(double x, double& s, double& c)
{
    double const u = x - 0.5;
    double const t = u * u;
    double sp = -1.05184717169320651e-11;
    sp = fma(sp, t, 5.39266466260812895e-10);
    sp = fma(sp, t, -2.29484289972698730e-08);
    sp = fma(sp, t, 7.95205400147551261e-07);
    sp = fma(sp, t, -2.19153534478302173e-05);
    sp = fma(sp, t, 4.66302805767612554e-04);
    sp = fma(sp, t, -7.37043094571435044e-03);
    sp = fma(sp, t, 8.21458866111282326e-02);
    sp = fma(sp, t, -5.99264529320792105e-01);
    sp = fma(sp, t, 2.55016403987734552e+00);
    sp = fma(sp, t, -5.16771278004997026e+00);
    sp = fma(sp, t, 3.14159265358979312e+00);
    double cp = 1.37686472803774139e-12;
    cp = fma(cp, t, -7.70070713060135385e-11);
    cp = fma(cp, t, 3.60473079746250112e-09);
    cp = fma(cp, t, -1.38789524622137714e-07);
    cp = fma(cp, t, 4.30306958703294729e-06);
    cp = fma(cp, t, -1.04638104924845705e-04);
    cp = fma(cp, t, 1.92957430940392314e-03);
    cp = fma(cp, t, -2.58068913900140612e-02);
    cp = fma(cp, t, 2.35330630358893206e-01);
    cp = fma(cp, t, -1.33526276885458950e+00);
    cp = fma(cp, t, 4.05871212641676848e+00);
    cp = fma(cp, t, -4.93480220054467900e+00);
    cp = fma(cp, t, 1.0);
    c = -u * sp;
    s = cp;
}

// 1/sqrt(x) for x > 0: hardware estimate refined by two Newton steps
// y <- y (3 - x y^2) / 2 (each step squares the relative error).
__device__ __forceinline__ double rsqrt_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double const hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}

// CutoffFunction::fdf (CutoffFunction.h:210-220, CutoffFunction.cpp:128-213,
// CoreFunction.h:214-318): value and derivative, zero at and beyond rc.
__device__ __forceinline__ void cutoff_fdf(CutCfg const& c, double r,
                                           double& fc, double& dfc)
{
    if (r >= c.rc) { fc = 0.0; dfc = 0.0; return; }
    int const t = c.type;
    if (t == CT_TANHU || t == CT_TANH)
    {
        double const th = tanh(1.0 - r * c.rcinv);
        double const th2 = th * th;
        // CT_TANH multiplies by tanh(1)^-3 so that f(0) = 1 (host-computed).
        double const pre = (t == CT_TANH) ? c.tanhpre : 1.0;
        fc = pre * th * th2;
        dfc = 3.0 * pre * th2 * (th2 - 1.0) * c.rcinv;
        return;
    }
    if (t == CT_HARD) { fc = 1.0; dfc = 0.0; return; }
    if (r < c.rci) { fc = 1.0; dfc = 0.0; return; }
    double const x = (r - c.rci) * c.iw;
    if (t == CT_COS)
    {
        // f = (cos(pi x) + 1)/2, f' = -iw pi sin(pi x)/2; the reference
        // forms sin(pi x) as sqrt(1 - cos^2), equal on [0, 1].
        double sn, cs;
        sincospi01(x, sn, cs);
        fc = 0.5 * (cs + 1.0);
        dfc = -0.5 * c.iw * kPi * sn;
        return;
    }
    double f, df;
    if (t == CT_POLY1)
    {
        f = (2.0 * x - 3.0) * x * x + 1.0;
        df = x * (6.0 * x - 6.0);
    }
    else if (t == CT_POLY2)
    {
        double const x2 = x * x;
        f = ((15.0 - 6.0 * x) * x - 10.0) * x * x2 + 1.0;
        df = x2 * ((60.0 - 30.0 * x) * x - 30.0);
    }
    else if (t == CT_POLY3)
    {
        double const x2 = x * x;
        f = (x * (x * (20.0 * x - 70.0) + 84.0) - 35.0) * x2 * x2 + 1.0;
        df = x2 * x * (x * (x * (140.0 * x - 420.0) + 420.0) - 140.0);
    }
    else if (t == CT_POLY4)
    {
        double x4 = x * x;
        x4 *= x4;
        f = (x * (x * ((315.0 - 70.0 * x) * x - 540.0) + 420.0) - 126.0)
          * x * x4 + 1.0;
        df = x4 * (x * (x * ((2520.0 - 630.0 * x) * x - 3780.0) + 2520.0)
           - 630.0);
    }
    else  // CT_EXP: f(x) = e^(1 - 1/(1-x^2))
    {
        double const e = 2.718281828459045090795598298427648842334747314453125;
        double const tmp = 1.0 / (x * x - 1.0);
        double const ex = exp(tmp);
        f = e * ex;
        df = -2.0 * e * x * tmp * tmp * ex;
    }
    fc = f;
    dfc = df * c.iw;
}

// Cutoff with the type fixed at compile time (CT >= 0) or read from the
// configuration (CT < 0).  The specialised forms keep the heavy kernels
// free of the other types' code and registers.
template <int CT>
__device__ __forceinline__ void cutoff_fdf_t(CutCfg const& c, double r,
                                             double& fc, double& dfc)
{
    if (CT < 0)
    {
        cutoff_fdf(c, r, fc, dfc);
        return;
    }
    if (r >= c.rc) { fc = 0.0; dfc = 0.0; return; }
    if (CT == CT_TANHU)
    {
        double const th = tanh(1.0 - r * c.rcinv);
        double const th2 = th * th;
        fc = th * th2;
        dfc = 3.0 * th2 * (th2 - 1.0) * c.rcinv;
        return;
    }
    if (r < c.rci) { fc = 1.0; dfc = 0.0; return; }
    double const x = (r - c.rci) * c.iw;
    if (CT == CT_COS)
    {
        double sn, cs;
        sincospi01(x, sn, cs);
        fc = 0.5 * (cs + 1.0);
        dfc = -0.5 * c.iw * kPi * sn;
    }
    else   // CT_POLY2
    {
        double const x2 = x * x;
        fc = ((15.0 - 6.0 * x) * x - 10.0) * x * x2 + 1.0;
        dfc = x2 * ((60.0 - 30.0 * x) * x - 30.0) * c.iw;
    }
}

// Launches KERNEL<CT, ...> for the cutoff types with specialised code and
// KERNEL<-1, ...> (runtime switch) for the others.
#define HDNNP_CUTOFF_DISPATCH(ct, CALL)                                    \
    do                                                                     \
    {                                                                      \
        switch (ct)                                                        \
        {                                                                  \
        case CT_COS: { constexpr int kCT = CT_COS; CALL; } break;           \
        case CT_TANHU: { constexpr int kCT = CT_TANHU; CALL; } break;       \
        case CT_POLY2: { constexpr int kCT = CT_POLY2; CALL; } break;       \
        default: { constexpr int kCT = -1; CALL; } break;                   \
        }                                                                  \
    } while (0)

// Exponentiation by squaring (utility.cpp:292-317 semantics).
__device__ __forceinline__ double pow_int(double x, int n)
{
    unsigned m;
    if (n < 0) { x = 1.0 / x; m = (unsigned)(-n); }
    else m = (unsigned)n;
    double r = 1.0;
    do
    {
        if (m & 1u) r *= x;
        m >>= 1;
        x *= x;
    } while (m);
    return r;
}

// NeuralNetwork::propagateLayer activations (NeuralNetwork.cpp:761-896):
// value and first derivative.  Logistic and softplus clamp at |x| > 35.
__device__ __forceinline__ void activate(int a, double x, double& v, double& d)
{
    switch (a)
    {
    case AF_IDENTITY: v = x; d = 1.0; break;
    case AF_TANH:
    {
        double const t = tanh(x);
        v = t; d = 1.0 - t * t;
        break;
    }
    case AF_LOGISTIC:
        if (x > 35.0) { v = 1.0; d = 0.0; }
        else if (x < -35.0) { v = 0.0; d = 0.0; }
        else
        {
            double const s = 1.0 / (1.0 + exp(-x));
            v = s; d = s * (1.0 - s);
        }
        break;
    case AF_SOFTPLUS:
        if (x > 35.0) { v = x; d = 1.0; }
        else if (x < -35.0) { v = 0.0; d = 0.0; }
        else
        {
            double const ex = exp(x);
            v = log(1.0 + ex);
            d = 1.0 / (1.0 + 1.0 / ex);
        }
        break;
    case AF_RELU:
        if (x > 0.0) { v = x; d = 1.0; } else { v = 0.0; d = 0.0; }
        break;
    case AF_GAUSSIAN:
    {
        double const g = exp(-0.5 * x * x);
        v = g; d = -x * g;
        break;
    }
    case AF_COS: v = cos(x); d = -sin(x); break;
    case AF_REVLOGISTIC:
    {
        double const s = 1.0 / (1.0 + exp(-x));
        v = 1.0 - s; d = s * (s - 1.0);
        break;
    }
    case AF_EXP:
    {
        double const e = exp(-x);
        v = e; d = -e;
        break;
    }
    case AF_HARMONIC: v = x * x; d = 2.0 * x; break;
    default: v = 0.0; d = 0.0; break;
    }
}

// Member parameters of one group, staged in LDS at the start of the group
// and read with wave-uniform addresses (broadcast) inside the pair loops.
struct MParam
{
    double lambda, zl, eta, rs, zeta, w;   // w: per-kernel weight
    int    zint, flags;                    // flags: 1 integer power, 2 new exp
};

__device__ __forceinline__ MParam make_mparam(MemberDesc const& md, double w)
{
    MParam p;
    p.lambda = md.lambda;
    p.zl = md.zl;
    p.eta = md.eta;
    p.rs = md.rs;
    p.zeta = md.zeta;
    p.w = w;
    p.zint = md.zint;
    p.flags = (md.useint ? 1 : 0) | (md.calcexp ? 2 : 0);
    return p;
}

// SymFnc::scale (SymFnc.cpp:169-195).
__device__ __forceinline__ double scale_sf(ScaleDesc const& s, double v)
{
    switch (s.type)
    {
    case ST_NONE: return v;
    case ST_SCALE: return s.smin + s.sfac * (v - s.gmin);
    case ST_CENTER: return v - s.gmean;
    case ST_SCALECENTER:
    case ST_SCALESIGMA: return s.smin + s.sfac * (v - s.gmean);
    }
    return 0.0;
}

// Butterfly sum over the 64 lanes of a wave; every lane gets the total.
// The pairing pattern is fixed, so the result is deterministic.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ int wave_sum_int(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63u); }

// Number of set bits of a 64-lane ballot below this lane.
__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return __popcll(mask & ((1ull << lane_id()) - 1ull));
}

}  // namespace hdnnp

#endif
