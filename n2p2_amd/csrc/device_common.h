// Device helpers shared by the HIP kernels: cutoff functions, integer
// powers, activations and 64-lane wave reductions, templated on the
// arithmetic type (double, or float on the mixed-precision path).
#ifndef HDNNP_DEVICE_COMMON_H
#define HDNNP_DEVICE_COMMON_H

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "tables.h"

namespace hdnnp {

constexpr int kWave = 64;
constexpr double kPi = 3.141592653589793238462643383279502884;

// Entries of the LDS queue of passing neighbour pairs (two batches of 64).
constexpr int kQueue = 128;

// q = p / d, r = p % d for 0 <= p < 2^22 and d > 0 from inv = 1 / (float)d:
// the float quotient is off by at most one, which the correction fixes.
__device__ __forceinline__ void divmod_small(int p, int d, float inv, int& q,
                                             int& r)
{
    q = (int)((float)p * inv);
    r = p - q * d;
    if (r < 0) { --q; r += d; }
    else if (r >= d) { ++q; r -= d; }
}

// Four- and two-component vectors of the arithmetic type R (LDS records).
template <typename R> struct VecT;
template <> struct VecT<float> { typedef float4 V4; typedef float2 V2; };
template <> struct VecT<double> { typedef double4 V4; typedef double2 V2; };
template <typename R> using Vec4 = typename VecT<R>::V4;
template <typename R> using Vec2 = typename VecT<R>::V2;

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

// FP32: the hardware sine and cosine (v_sin_f32 / v_cos_f32 take the
// angle in revolutions, so pi x is x / 2 revolutions).  Max abs error
// against libm over [0, 1]: 1.3e-7 (measured on gfx950 with an in-bounds
// probe over 2^20 points).
__device__ __forceinline__ void sincospi01(float x, float& s, float& c)
{
    float const rev = 0.5f * x;
    s = __builtin_amdgcn_sinf(rev);
    c = __builtin_amdgcn_cosf(rev);
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

// FP32: the hardware estimate (1 ulp) as it is.
__device__ __forceinline__ float rsqrt_nr(float x)
{
    return __builtin_amdgcn_rsqf(x);
}

// e^x.  FP64: the library function.  FP32: the hardware 2^x of x log2(e)
// (relative error below 2e-6 for |x| < 30, the range of the SF exponents).
__device__ __forceinline__ double exp_r(double x) { return exp(x); }
__device__ __forceinline__ float exp_r(float x) { return __expf(x); }

// 1 / x.  FP64: IEEE division.  FP32: the hardware reciprocal (1 ulp).
__device__ __forceinline__ double rcp_r(double x) { return 1.0 / x; }
__device__ __forceinline__ float rcp_r(float x)
{
    return __builtin_amdgcn_rcpf(x);
}

__device__ __forceinline__ double tanh_r(double x) { return tanh(x); }
__device__ __forceinline__ float tanh_r(float x) { return tanhf(x); }

// CutoffFunction::fdf (CutoffFunction.h:210-220, CutoffFunction.cpp:128-213,
// CoreFunction.h:214-318): value and derivative, zero at and beyond rc.
// R is the arithmetic type (double, or float on the mixed-precision path).
template <typename R>
__device__ __forceinline__ void cutoff_fdf(CutCfg const& c, R r, R& fc,
                                           R& dfc)
{
    if (r >= (R)c.rc) { fc = R(0); dfc = R(0); return; }
    int const t = c.type;
    if (t == CT_TANHU || t == CT_TANH)
    {
        R const th = tanh_r(R(1) - r * (R)c.rcinv);
        R const th2 = th * th;
        // CT_TANH multiplies by tanh(1)^-3 so that f(0) = 1 (host-computed).
        R const pre = (t == CT_TANH) ? (R)c.tanhpre : R(1);
        fc = pre * th * th2;
        dfc = R(3) * pre * th2 * (th2 - R(1)) * (R)c.rcinv;
        return;
    }
    if (t == CT_HARD) { fc = R(1); dfc = R(0); return; }
    if (r < (R)c.rci) { fc = R(1); dfc = R(0); return; }
    R const iw = (R)c.iw;
    R const x = (r - (R)c.rci) * iw;
    if (t == CT_COS)
    {
        // f = (cos(pi x) + 1)/2, f' = -iw pi sin(pi x)/2; the reference
        // forms sin(pi x) as sqrt(1 - cos^2), equal on [0, 1].
        R sn, cs;
        sincospi01(x, sn, cs);
        fc = R(0.5) * (cs + R(1));
        dfc = R(-0.5) * iw * (R)kPi * sn;
        return;
    }
    R f, df;
    if (t == CT_POLY1)
    {
        f = (R(2) * x - R(3)) * x * x + R(1);
        df = x * (R(6) * x - R(6));
    }
    else if (t == CT_POLY2)
    {
        R const x2 = x * x;
        f = ((R(15) - R(6) * x) * x - R(10)) * x * x2 + R(1);
        df = x2 * ((R(60) - R(30) * x) * x - R(30));
    }
    else if (t == CT_POLY3)
    {
        R const x2 = x * x;
        f = (x * (x * (R(20) * x - R(70)) + R(84)) - R(35)) * x2 * x2 + R(1);
        df = x2 * x * (x * (x * (R(140) * x - R(420)) + R(420)) - R(140));
    }
    else if (t == CT_POLY4)
    {
        R x4 = x * x;
        x4 *= x4;
        f = (x * (x * ((R(315) - R(70) * x) * x - R(540)) + R(420)) - R(126))
          * x * x4 + R(1);
        df = x4 * (x * (x * ((R(2520) - R(630) * x) * x - R(3780))
                        + R(2520)) - R(630));
    }
    else  // CT_EXP: f(x) = e^(1 - 1/(1-x^2))
    {
        R const e = (R)2.718281828459045090795598298427648842334747314453125;
        R const tmp = R(1) / (x * x - R(1));
        R const ex = exp_r(tmp);
        f = e * ex;
        df = R(-2) * e * x * tmp * tmp * ex;
    }
    fc = f;
    dfc = df * iw;
}

// Cutoff with the type fixed at compile time (CT >= 0) or read from the
// configuration (CT < 0).  The specialised forms keep the heavy kernels
// free of the other types' code and registers.
template <int CT, typename R>
__device__ __forceinline__ void cutoff_fdf_t(CutCfg const& c, R r, R& fc,
                                             R& dfc)
{
    if (CT < 0)
    {
        cutoff_fdf(c, r, fc, dfc);
        return;
    }
    if (r >= (R)c.rc) { fc = R(0); dfc = R(0); return; }
    if (CT == CT_TANHU)
    {
        R const th = tanh_r(R(1) - r * (R)c.rcinv);
        R const th2 = th * th;
        fc = th * th2;
        dfc = R(3) * th2 * (th2 - R(1)) * (R)c.rcinv;
        return;
    }
    if (r < (R)c.rci) { fc = R(1); dfc = R(0); return; }
    R const iw = (R)c.iw;
    R const x = (r - (R)c.rci) * iw;
    if (CT == CT_COS)
    {
        R sn, cs;
        sincospi01(x, sn, cs);
        fc = R(0.5) * (cs + R(1));
        dfc = R(-0.5) * iw * (R)kPi * sn;
    }
    else   // CT_POLY2
    {
        R const x2 = x * x;
        fc = ((R(15) - R(6) * x) * x - R(10)) * x * x2 + R(1);
        dfc = x2 * ((R(60) - R(30) * x) * x - R(30)) * iw;
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
template <typename R>
__device__ __forceinline__ R pow_int(R x, int n)
{
    unsigned m;
    if (n < 0) { x = R(1) / x; m = (unsigned)(-n); }
    else m = (unsigned)n;
    R r = R(1);
    do
    {
        if (m & 1u) r *= x;
        m >>= 1;
        x *= x;
    } while (m);
    return r;
}

// x^N with the products pow_int forms for exponent N, N fixed at compile
// time: straight-line code, bit-identical to pow_int(x, N).
template <int N, typename R>
__device__ __forceinline__ R pow_fixed(R x)
{
    R r = R(1);
    bool first = true;
    int m = N;
#pragma unroll
    for (int b = 0; b < 6; ++b)
    {
        if (m & 1)
        {
            r = first ? x : r * x;
            first = false;
        }
        m >>= 1;
        if (m) x *= x;
    }
    return r;
}

// Powers of x = plambda shared by the members of an (eta, rs, lambda) run
// sorted by zeta: after l steps P = x^(2^l - 1) and X = x^(2^l), the
// running product and square of pow_int's loop, so at(l) for zeta = 2^l is
// bit-identical to pow_int(x, zeta - 1).  l is wave-uniform.
template <typename R>
struct PowChain
{
    R P, X;
    int l;
    __device__ __forceinline__ void reset(R x)
    {
        P = R(1);
        X = x;
        l = 0;
    }
    __device__ __forceinline__ R at(int level)
    {
        while (l < level)
        {
            P *= X;
            X *= X;
            ++l;
        }
        return P;
    }
};

// pow_int(x, n) for a wave-uniform n: a scalar branch to straight-line
// products for 0 <= n <= 16, the loop otherwise.
template <typename R>
__device__ __forceinline__ R pow_uniform(R x, int n)
{
    switch (n)
    {
    case 0: return R(1);
    case 1: return x;
    case 2: return pow_fixed<2>(x);
    case 3: return pow_fixed<3>(x);
    case 4: return pow_fixed<4>(x);
    case 5: return pow_fixed<5>(x);
    case 6: return pow_fixed<6>(x);
    case 7: return pow_fixed<7>(x);
    case 8: return pow_fixed<8>(x);
    case 9: return pow_fixed<9>(x);
    case 10: return pow_fixed<10>(x);
    case 11: return pow_fixed<11>(x);
    case 12: return pow_fixed<12>(x);
    case 13: return pow_fixed<13>(x);
    case 14: return pow_fixed<14>(x);
    case 15: return pow_fixed<15>(x);
    case 16: return pow_fixed<16>(x);
    default: return pow_int(x, n);
    }
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
// Integer fields are made scalar with readfirstlane where they steer
// branches.
enum MemberFlags
{
    MF_INTPOW = 1,   // integer zeta (pow_int)
    MF_NEWEXP = 2,   // first member of an (eta, rs) run: new exponential
    MF_NEWLAM = 4,   // first member of an (eta, rs, lambda) run
    MF_PREFIX = 8    // zeta = 2^level: power from the run's prefix chain
};

template <typename R>
struct MParam
{
    R   lambda, zl, eta, rs, zeta, w;   // w: per-kernel weight
    int zint, flags;                    // flags | level << 8
};

template <typename R>
__device__ __forceinline__ MParam<R> make_mparam(MemberDesc const& md,
                                                 double w)
{
    MParam<R> p;
    p.lambda = (R)md.lambda;
    p.zl = (R)md.zl;
    p.eta = (R)md.eta;
    p.rs = (R)md.rs;
    p.zeta = (R)md.zeta;
    p.w = (R)w;
    p.zint = md.zint;
    p.flags = (md.useint ? MF_INTPOW : 0) | (md.calcexp ? MF_NEWEXP : 0)
            | (md.newlam ? MF_NEWLAM : 0)
            | (md.level >= 0 ? MF_PREFIX | (md.level << 8) : 0);
    return p;
}

__device__ __forceinline__ int uniform(int v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

// Wave-uniform copies of floating-point values (scalar registers).
__device__ __forceinline__ float uniform_r(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

__device__ __forceinline__ double uniform_r(double v)
{
    long long const b = __double_as_longlong(v);
    int const lo = __builtin_amdgcn_readfirstlane((int)b);
    int const hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// exp(-eta s) from the exponent factor of exp_factor(eta): FP64 takes the
// library exp of -eta s; FP32 the hardware 2^x of (-eta log2 e) s.
__device__ __forceinline__ double exp_factor(double eta) { return -eta; }
__device__ __forceinline__ float exp_factor(float eta)
{
    return -eta * 1.44269504088896340736f;
}
__device__ __forceinline__ double exp_scaled(double a) { return exp(a); }
__device__ __forceinline__ float exp_scaled(float a)
{
    return __builtin_amdgcn_exp2f(a);
}

// Cutoff value and derivative for rc > r >= 0 without branches for the
// specialised types: x = (r - rci) iw clamped to [0, 1] gives f = 1, f' = 0
// below rci exactly as the branch does (the cosine form to rounding).  Other
// types take the general form.
template <int CT, typename R>
__device__ __forceinline__ void cutoff_inside(CutCfg const& c, R r, R& fc,
                                              R& dfc)
{
    if (CT == CT_COS)
    {
        R const iw = (R)c.iw;
        R x = (r - (R)c.rci) * iw;
        x = fmin(fmax(x, R(0)), R(1));
        R sn, cs;
        sincospi01(x, sn, cs);
        fc = R(0.5) * (cs + R(1));
        dfc = R(-0.5) * iw * (R)kPi * sn;
    }
    else if (CT == CT_POLY2)
    {
        R const iw = (R)c.iw;
        R x = (r - (R)c.rci) * iw;
        x = fmin(fmax(x, R(0)), R(1));
        R const x2 = x * x;
        fc = ((R(15) - R(6) * x) * x - R(10)) * x * x2 + R(1);
        dfc = x2 * ((R(60) - R(30) * x) * x - R(30)) * iw;
    }
    else cutoff_fdf_t<CT>(c, r, fc, dfc);
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
// The pairing pattern is fixed, so the result is deterministic.  Float
// partials are summed in FP64.
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

// Work index of workgroup b of nb.  The hardware deals workgroups to the 8
// XCDs round-robin (XCD = b % 8); this mapping gives each XCD one contiguous
// range of work indices, so work items that touch the same data (atoms of
// neighbouring cells) share that XCD's L2.  A bijection on [0, nb).
__device__ __forceinline__ int xcd_work_index(int b, int nb)
{
    int const x = b & 7, k = b >> 3;
    int const q = nb >> 3, r = nb & 7;   // XCD x holds q + (x < r) groups
    return x * q + (x < r ? x : r) + k;
}

// Sums M per-lane values over the wave by a transposing butterfly: at each
// of the first log2(M) steps a lane keeps half of its values, sends the
// other half to the partner lane (xor 32, 16, ...) and adds what it
// receives, then the remaining steps are a plain butterfly.  Returns the
// total of member ((lane >> (6 - log2 M)) & (M - 1)); every lane of that
// group of 64/M lanes holds it.  M - 1 + 6 - log2(M) exchanges and adds
// instead of 6 M for M separate sums; the order of the additions is fixed.
template <int M>
__device__ __forceinline__ double wave_sum_transposed(double (&v)[M])
{
    static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16
                  || M == 32, "power of two up to 32");
    int const lane = lane_id();
#pragma unroll
    for (int h = M / 2, o = 32; h >= 1; h >>= 1, o >>= 1)
    {
        bool const upper = (lane & o) != 0;
#pragma unroll
        for (int m = 0; m < h; ++m)
        {
            double const send = upper ? v[m] : v[m + h];
            double const keep = upper ? v[m + h] : v[m];
            v[m] = keep + __shfl_xor(send, o, kWave);
        }
    }
    double t = v[0];
#pragma unroll
    for (int o = 32 / M; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
    return t;
}

// Unordered pairs {j, k} of a group's neighbour blocks in diagonal order.
// Same element (block of n at s0): pair p = c n + r is {r, (r + 1 + c) mod
// n}, at circular distance 1 + c.  The (n - 1) / 2 full diagonals
// (distance below n/2) hold n pairs each; for n even the half diagonal at
// distance n/2 follows with r < n/2.  Every unordered pair appears once
// (n (n - 1) / 2 in all) and one diagonal holds each slot at most twice.
// Different elements, smaller block s0 (ns) and larger block l0 (nl):
// p = c ns + r is {s0 + r, l0 + (r + c) mod nl}, c < nl; one diagonal holds
// each slot at most once.  A lane walks p = lane, lane + 64, ... with an
// incremental (c, r) (no division per step).
struct DiagPairs
{
    bool same;
    int s0, ns, l0, nl, d, npairs;
    int q64, r64;     // 64 = q64 d + r64
    __device__ DiagPairs(bool same_, int a0, int na, int b0, int nb)
        : same(same_)
    {
        if (same)
        {
            s0 = l0 = a0;
            ns = nl = na;
            d = na;
            npairs = na * (na - 1) / 2;
        }
        else
        {
            bool const aSmall = na <= nb;
            s0 = aSmall ? a0 : b0;
            ns = aSmall ? na : nb;
            l0 = aSmall ? b0 : a0;
            nl = aSmall ? nb : na;
            d = ns;
            npairs = na * nb;
        }
        q64 = d > 0 ? 64 / d : 0;
        r64 = 64 - q64 * d;
    }
    // Cursor (c, r) of pair p0 (one division, at the start).
    __device__ void start(int p0, int& c, int& r) const
    {
        c = 0;
        r = 0;
        if (d > 0) divmod_small(p0, d, 1.0f / (float)d, c, r);
    }
    // (c, r) of pair p -> pair p + 64.
    __device__ void advance(int& c, int& r) const
    {
        r += r64;
        c += q64;
        if (r >= d) { r -= d; ++c; }
    }
    // Staged indices of pair (c, r); only for p < npairs.
    __device__ void pair(int c, int r, int& j, int& k) const
    {
        int v = r + c + (same ? 1 : 0);
        // r < d and c < nl (c < n/2 for one block): one subtraction.
        if (v >= nl) v -= nl;
        j = s0 + r;
        k = l0 + v;
    }
};

// rij, rik < rc and (narrow only) rjk^2 < rc^2 from the staged records
// (SymGrpExpAngn.cpp:182-232; the wide type skips rjk), branch free.
template <typename R, typename V4>
__device__ __forceinline__ bool pair_in_range(V4 const& pj, V4 const& pk,
                                              bool wide, R rc, R rc2)
{
    R const dx = pk.x - pj.x;
    R const dy = pk.y - pj.y;
    R const dz = pk.z - pj.z;
    R const d2 = dx * dx + dy * dy + dz * dz;
    return (pj.w < rc) & (pk.w < rc) & (wide | (d2 < rc2));
}

// Orders this wave's LDS accesses around a queue hand-off between lanes,
// where the workgroup is one wave: a wave's LDS operations complete in
// issue order, so a wavefront-scope fence (no wait, unlike the workgroup
// barrier's fence) is enough and later LDS reads no longer wait for the
// queue writes before them to complete.
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Number of set bits of a 64-lane ballot below this lane (two mbcnt
// instructions).
__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return (int)__builtin_amdgcn_mbcnt_hi(
        (unsigned)(mask >> 32),
        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Weight of staged neighbour t for the weighted SF types: the atomic number
// of the element block (neighbours staged partitioned by element) that
// holds it.
__device__ __forceinline__ double staged_weight(ModelTables const& T,
                                                int const* sStart,
                                                int const* sCnt, int t)
{
    double w = 0.0;
#pragma unroll
    for (int q = 0; q < kMaxElements; ++q)
        if (q < T.ne && t >= sStart[q] && t < sStart[q] + sCnt[q])
            w = T.zw[q];
    return w;
}

}  // namespace hdnnp

#endif
