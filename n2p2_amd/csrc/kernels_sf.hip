// Ahead-of-time symmetry-function kernels: the generic forward kernel
// (member loop driven by the group tables), the work counters, the SF
// statistics and extrapolation-warning kernels and the launchers.  The
// forward kernel's device code is in sf_core.h.
#include "sf_core.h"
#include "kernels.h"

namespace hdnnp {

namespace {

template <int MAXM, int CT, bool GENPOW, typename R>
__global__ __launch_bounds__(64) void k_sf_forward(SFArgs a)
{
    sf_forward_body<MAXM, CT, GENPOW, R>(a);
}

// Work counters of the algorithm for one structure (same enumeration and
// filters as the FP64 k_sf_forward, no arithmetic on the members):
//   [0] neighbour-list entries
//   [1] radial member evaluations        [2] narrow pairs tested (rij, rik < rc)
//   [3] wide pairs tested                [4] narrow triplets (rjk < rc)
//   [5] wide triplets                    [6] narrow member evaluations
//   [7] wide member evaluations          [8] multiplies in pow_int(plambda,
//                                            zeta - 1) over all evaluations
//   [9] dE/dG . dG/dr contraction terms of Mode::calculateForces:
//       per atom nsf + sum over slots of numSFD(element of the slot)
__global__ __launch_bounds__(64) void k_sf_count(SFArgs a,
                                                 unsigned long long* out)
{
    extern __shared__ __align__(16) double sm[];
    __shared__ int sStart[kMaxElements], sCnt[kMaxElements];
    Stage<double> S;
    S.P = (double4*)sm;
    S.Q = nullptr;
    int const i = blockIdx.x;
    int const lane = threadIdx.x;
    ModelTables const& T = a.T;
    int const ec = a.elem[i];
    ElemDesc const& ed = T.elem[ec];
    int const s0 = a.off[i];
    int const n = a.cnt[i];
    stage_by_element(a, S, s0, n, sStart, sCnt);
    unsigned long long c[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) c[k] = 0;
    for (int gi = ed.g0; gi < ed.g0 + ed.ng; ++gi)
    {
        GroupDesc const g = T.group[gi];
        bool const weighted = sf_weighted(g.type);
        if (sf_radial(g.type))
        {
            int const a0 = weighted ? 0 : sStart[g.e1];
            int const na = weighted ? n : sCnt[g.e1];
            for (int t = lane; t < na; t += 64)
                if (S.P[a0 + t].w < g.rc) c[1] += g.nmem;
            continue;
        }
        bool const wide = (g.type == 9);
        int powm = 0;   // pow_int multiplies of one evaluation of all members
        for (int m = 0; m < g.nmem; ++m)
        {
            MemberDesc const& md = T.member[g.mem0 + m];
            int e = md.useint ? md.zint - 1 : 0;
            int it = 0, bits = 0;
            do { bits += (e & 1); e >>= 1; ++it; } while (e);
            powm += md.useint ? it + bits : 0;
        }
        DiagPairs const dp = weighted
            ? DiagPairs(true, 0, n, 0, n)
            : DiagPairs(g.e1 == g.e2, sStart[g.e1], sCnt[g.e1],
                        sStart[g.e2], sCnt[g.e2]);
        int pc = 0, pr = 0;
        dp.start(lane, pc, pr);
        for (int base = 0; base < dp.npairs; base += 64)
        {
            int ja, kb;
            dp.pair(pc, pr, ja, kb);
            dp.advance(pc, pr);
            if (base + lane >= dp.npairs) continue;
            if (!(S.P[ja].w < g.rc && S.P[kb].w < g.rc)) continue;
            c[wide ? 3 : 2] += 1;
            if (pair_in_range<double>(S.P[ja], S.P[kb], wide, g.rc, g.rc2))
            {
                c[wide ? 5 : 4] += 1;
                c[wide ? 7 : 6] += g.nmem;
                c[8] += powm;
            }
        }
    }
    if (lane == 0)
    {
        c[0] = n;
        unsigned long long t = ed.nsf;
        for (int q = 0; q < T.ne; ++q)
            t += (unsigned long long)sCnt[q] * ed.nsfd[q];
        c[9] = t;
    }
#pragma unroll
    for (int k = 0; k < 10; ++k)
    {
        unsigned long long v = c[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
        if (lane == 0 && v) atomicAdd(&out[k], v);
    }
}

// Per SF of one element: min, max, sum, sum of squares over its atoms.  One
// block per SF, fixed-order tree reduction.
__global__ __launch_bounds__(256) void k_sf_stats(const double* __restrict__ G,
                                                  int ldG,
                                                  const int* __restrict__ atoms,
                                                  int na,
                                                  double* __restrict__ out)
{
    __shared__ double smin[256], smax[256], ssum[256], ssq[256];
    int const k = blockIdx.x;
    int const t = threadIdx.x;
    double mn = 1.0e300, mx = -1.0e300, su = 0.0, sq = 0.0;
    for (int q = t; q < na; q += 256)
    {
        double const v = G[(size_t)atoms[q] * ldG + k];
        mn = fmin(mn, v);
        mx = fmax(mx, v);
        su += v;
        sq += v * v;
    }
    smin[t] = mn; smax[t] = mx; ssum[t] = su; ssq[t] = sq;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o)
        {
            smin[t] = fmin(smin[t], smin[t + o]);
            smax[t] = fmax(smax[t], smax[t + o]);
            ssum[t] += ssum[t + o];
            ssq[t] += ssq[t + o];
        }
        __syncthreads();
    }
    if (t == 0)
    {
        out[4 * k] = smin[0];
        out[4 * k + 1] = smax[0];
        out[4 * k + 2] = ssum[0];
        out[4 * k + 3] = ssq[0];
    }
}

// Element::updateSymmetryFunctionStatistics (Element.cpp:462-541): count
// SFs whose unscaled value is outside [Gmin, Gmax] by more than 1000 eps.
// Every SF is tested, unscaled ones (ST_NONE) too, as the reference does:
// their Gmin / Gmax come from scaling.data, or are 0 without a scaling
// file (Element::setScalingNone, Element.cpp:344-362).
__global__ void k_ew(ModelTables T, const double* __restrict__ G, int ldG,
                     const int* __restrict__ elem, int n, int* __restrict__ ew)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ElemDesc const& ed = T.elem[elem[i]];
    double const eps = 1000.0 * 2.220446049250313e-16;
    int c = 0;
    for (int k = 0; k < ed.nsf; ++k)
    {
        ScaleDesc const& s = T.scale[ed.sf0 + k];
        double const g = G[(size_t)i * ldG + k];
        double v;
        switch (s.type)   // SymFnc::unscale (SymFnc.cpp:197-223)
        {
        case ST_SCALE: v = (g - s.smin) / s.sfac + s.gmin; break;
        case ST_CENTER: v = g + s.gmean; break;
        case ST_SCALECENTER:
        case ST_SCALESIGMA: v = (g - s.smin) / s.sfac + s.gmean; break;
        default: v = g; break;
        }
        if (v + eps < s.gmin || v - eps > s.gmax) ++c;
    }
    ew[i] = c;
}

template <int MAXM, int CT, typename R>
void launch_fwd(SFArgs const& a, hipStream_t s)
{
    size_t const shm = sf_forward_lds_bytes(a.cap, a.fp32);
    if (a.T.genpow)
        k_sf_forward<MAXM, CT, true, R><<<a.n, 64, shm, s>>>(a);
    else
        k_sf_forward<MAXM, CT, false, R><<<a.n, 64, shm, s>>>(a);
}

template <typename R>
void launch_fwd_r(SFArgs const& a, hipStream_t s)
{
    if (a.T.maxGroupMembers <= 8)
        HDNNP_CUTOFF_DISPATCH(a.T.cutType, (launch_fwd<8, kCT, R>(a, s)));
    else
        HDNNP_CUTOFF_DISPATCH(a.T.cutType,
                              (launch_fwd<kMaxMembers, kCT, R>(a, s)));
}

}  // namespace

size_t sf_forward_lds_bytes(int cap, int fp32)
{
    size_t const r = fp32 ? sizeof(float) : sizeof(double);
    return (size_t)cap * 6 * r + kQueue * sizeof(int);
}

hipError_t launch_sf_forward(SFArgs const& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    if (a.fp32) launch_fwd_r<float>(a, s);
    else launch_fwd_r<double>(a, s);
    return hipGetLastError();
}

hipError_t launch_sf_count(SFArgs const& a, unsigned long long* counters,
                           hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    k_sf_count<<<a.n, 64, sf_forward_lds_bytes(a.cap, 0), s>>>(a, counters);
    return hipGetLastError();
}

hipError_t launch_sf_stats(const double* G, int ldG, const int* atoms, int na,
                           int nsf, double* out, hipStream_t s)
{
    if (nsf <= 0) return hipSuccess;
    k_sf_stats<<<nsf, 256, 0, s>>>(G, ldG, atoms, na, out);
    return hipGetLastError();
}

hipError_t launch_ew(ModelTables T, const double* G, int ldG, const int* elem,
                     int n, int* perAtom, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_ew<<<(n + 255) / 256, 256, 0, s>>>(T, G, ldG, elem, n, perAtom);
    return hipGetLastError();
}

}  // namespace hdnnp
