// Ahead-of-time force kernels: the generic backward kernel (member loop
// driven by the group tables), the per-atom gather and the launchers.  The
// backward kernel's device code is in force_core.h.
#include "force_core.h"
#include "kernels.h"

namespace hdnnp {

namespace {

// FP32 with up to 8 members per group: at most 128 VGPRs so that four
// waves fit per SIMD (measured faster than three waves without the bound).
template <int CT, bool GENPOW, int MAXM, typename R>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 && MAXM <= 8 ? 4 : 1, 8)))
void k_force_backward(ForceArgs a)
{
    force_backward_body<CT, GENPOW, MAXM, R>(a);
}

// F_m = -(cself[m] + sum_q cslot[partner(q)]) over m's slots q; mode 0
// skips slots pointing at m itself (Mode.cpp:1885 starts the unique list
// after the centre atom), mode 1 keeps them.  One wave per atom, fixed
// lane assignment and butterfly: deterministic.  Atoms are taken in cell
// order with the XCD-aware mapping, so the partner entries read by
// neighbouring atoms share an L2.
__global__ __launch_bounds__(256) void k_gather(
    int n, const int* __restrict__ off, const int* __restrict__ cnt,
    const int* __restrict__ nbrJ,
    const int* __restrict__ partner, const double* __restrict__ cslot,
    const double* __restrict__ cself, const int* __restrict__ order,
    int mode, int64_t nslots, double* __restrict__ F)
{
    int const t = xcd_work_index(blockIdx.x, gridDim.x) * 4
                + (int)(threadIdx.x >> 6);
    int const lane = threadIdx.x & 63;
    if (t >= n) return;
    int const i = order[t];
    int const s0 = off[i], s1 = s0 + cnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int q = s0 + lane; q < s1; q += 64)
    {
        int const pq = partner[q];
        // The list is symmetric, so every entry has a partner; a partner
        // outside [0, nslots) would mean an inconsistent list and is
        // skipped rather than read.
        if ((mode == 1 || nbrJ[q] != i) && pq >= 0 && pq < nslots)
        {
            size_t const p = (size_t)pq;
            fx += cslot[3 * p];
            fy += cslot[3 * p + 1];
            fz += cslot[3 * p + 2];
        }
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0)
    {
        F[3 * (size_t)i] = -(cself[3 * (size_t)i] + fx);
        F[3 * (size_t)i + 1] = -(cself[3 * (size_t)i + 1] + fy);
        F[3 * (size_t)i + 2] = -(cself[3 * (size_t)i + 2] + fz);
    }
}

// InterfaceLammps::getForces (InterfaceLammps.cpp:533-621) for a
// caller-given list: f[b] -= every contribution whose target is caller
// atom b, in the reference's loop order (centres in order; per centre its
// self force, then its entries in list order).  src[srcOff[b] ..
// srcOff[b+1]) lists them in that order: src >= 0 is list entry q
// (cslot[q]), src < 0 the self force of centre -src - 1 (cself).  One
// thread per target: each f[b] is the same sequence of subtractions as the
// reference's loop, so equal slot values give bitwise equal forces.
__global__ __launch_bounds__(256) void k_scatter_ext(
    int nt, const int64_t* __restrict__ srcOff, const int* __restrict__ src,
    const double* __restrict__ cslot, const double* __restrict__ cself,
    double* __restrict__ f)
{
    int const b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nt) return;
    double fx = f[3 * (size_t)b], fy = f[3 * (size_t)b + 1],
           fz = f[3 * (size_t)b + 2];
    for (int64_t k = srcOff[b]; k < srcOff[b + 1]; ++k)
    {
        int const q = src[k];
        double const* v = q >= 0 ? cslot + 3 * (size_t)q
                                 : cself + 3 * (size_t)(-(int64_t)q - 1);
        fx -= v[0];
        fy -= v[1];
        fz -= v[2];
    }
    f[3 * (size_t)b] = fx;
    f[3 * (size_t)b + 1] = fy;
    f[3 * (size_t)b + 2] = fz;
}

template <int MAXM, typename R>
void launch_bwd(ForceArgs const& a, size_t shm, hipStream_t s)
{
    if (a.T.genpow)
        HDNNP_CUTOFF_DISPATCH(a.T.cutType,
            (k_force_backward<kCT, true, MAXM, R><<<a.n, 64, shm, s>>>(a)));
    else
        HDNNP_CUTOFF_DISPATCH(a.T.cutType,
            (k_force_backward<kCT, false, MAXM, R><<<a.n, 64, shm, s>>>(a)));
}

template <typename R>
void launch_bwd_r(ForceArgs const& a, size_t shm, hipStream_t s)
{
    if (a.T.maxGroupMembers <= 8) launch_bwd<8, R>(a, shm, s);
    else launch_bwd<kMaxMembers, R>(a, shm, s);
}

}  // namespace

size_t force_backward_lds_bytes(int cap, int fp32)
{
    size_t const r = fp32 ? sizeof(float) : sizeof(double);
    size_t b = 0;
    b += (size_t)cap * 8 * r;                     // P (4) and Q (4)
    b += (size_t)cap * 3 * sizeof(double);        // ACC
    b += (size_t)kQueue * sizeof(int);            // QUEUE
    return b;
}

hipError_t launch_force_backward(ForceArgs const& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    size_t const shm = force_backward_lds_bytes(a.cap, a.fp32);
    if (a.fp32) launch_bwd_r<float>(a, shm, s);
    else launch_bwd_r<double>(a, shm, s);
    return hipGetLastError();
}

hipError_t launch_force_gather(int n, const int* off, const int* cnt,
                               const int* nbrJ, const int* partner,
                               const double* cslot, const double* cself,
                               const int* order, int mode, int64_t nslots,
                               double* F, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_gather<<<(n + 3) / 4, 256, 0, s>>>(n, off, cnt, nbrJ, partner, cslot,
                                         cself, order, mode, nslots, F);
    return hipGetLastError();
}

hipError_t launch_scatter_ext(int nt, const int64_t* srcOff, const int* src,
                              const double* cslot, const double* cself,
                              double* f, hipStream_t s)
{
    if (nt <= 0) return hipSuccess;
    k_scatter_ext<<<(nt + 255) / 256, 256, 0, s>>>(nt, srcOff, src, cslot,
                                                  cself, f);
    return hipGetLastError();
}

}  // namespace hdnnp
