// Force kernels (FP64): contraction dE/dG . dG/dr with dG/dr recomputed,
// then a deterministic per-atom gather.
//
// The reference stores dG/dr per SF and per neighbour slot
// (SymGrp*::calculate with derivatives, Atom::allocate) and contracts them
// in Mode::calculateForces (Mode.cpp:1849-1913) with
// Atom::calculateSelfForceShort / calculatePairForceShort (Atom.cpp:371-401).
// Here nothing per SF is stored: for atom i the kernel produces
//   cself[i]     = sum_s dE_i/dG_s * dG_s/dr_i                 (3 doubles)
//   cslot[i, q]  = sum_s dE_i/dG_s * dG_s/dr_(neighbour slot q) (3 doubles)
// and the gather kernel forms F_m = -(cself[m] + sum over m's slots q
// pointing at another atom j of cslot[j, partner(q)]), partner(q) being the
// slot of j that points back at m with the opposite image.
//
// Backward kernel: one wavefront per atom, neighbours staged in LDS in slot
// order.  Lane L owns slots j = L, L+64, ...; for its slot it evaluates
// every ordered pair (j, k) that passes the angular filters (found with a
// per-lane bitmask so the heavy loop runs only over passing partners), and
// accumulates the slot-j part of the pair derivative (terms p1 and p3 of
// SymGrpExpAngn.cpp:300-337 / SymGrpExpAngw.cpp:288-310) in registers.  The
// partner slot k receives its part from the lane owning k, so every slot is
// written by exactly one lane and no atomics are needed.
#include "device_common.h"
#include "kernels.h"

namespace hdnnp {

namespace {

// Per-member parameters staged in LDS (non-uniform access by lanes whose
// pairs belong to different groups).
struct MemLDS
{
    double lambda, zl, fderiv, fnorm, eta, rs, zeta, w;   // w = dE/dG_sf
    int    zint, useint, calcexp, pad;
};

struct GrpLDS
{
    double rc, rc2;
    int    type, cfg, nmem, mem;   // mem: local member index
};

__global__ __launch_bounds__(64) void k_force_backward(ForceArgs a)
{
    extern __shared__ double sm[];
    int const cap = a.cap;
    int const ncfgMax = a.maxCfg;
    ModelTables const& T = a.T;
    int const ne = T.ne;
    int const i = blockIdx.x;
    int const lane = threadIdx.x;
    int const ec = a.elem[i];
    ElemDesc const& ed = T.elem[ec];
    int const s0 = a.off[i];
    int const n = a.off[i + 1] - s0;
    int const nw = (n + 31) >> 5;

    double* X = sm;
    double* Y = X + cap;
    double* Z = Y + cap;
    double* D = Z + cap;
    double* FC = D + cap;                       // [cfg][cap]
    double* DFC = FC + (size_t)ncfgMax * cap;   // [cfg][cap]
    MemLDS* M = (MemLDS*)(DFC + (size_t)ncfgMax * cap);
    GrpLDS* GR = (GrpLDS*)(M + a.maxMembers);
    double* PR = (double*)(GR + kMaxElements * kMaxElements * kMaxPairGroups);
    double* PJ = PR + kMaxElements * kMaxElements;   // pair rjk^2 limit
    int* E = (int*)(PJ + kMaxElements * kMaxElements);
    uint32_t* MASK = (uint32_t*)(E + cap);           // [word][64]

    // Stage neighbours (slot order) and cutoff values per configuration.
    for (int s = lane; s < n; s += 64)
    {
        double4 const g = a.geo[s0 + s];
        X[s] = g.x;
        Y[s] = g.y;
        Z[s] = g.z;
        D[s] = g.w;
        E[s] = a.elem[a.nbrJ[s0 + s]];
        for (int c = 0; c < ed.ncfg; ++c)
        {
            double f, df;
            cutoff_fdf(T.cfg[ed.cfg0 + c], g.w, f, df);
            FC[c * cap + s] = f;
            DFC[c * cap + s] = df;
        }
    }
    // Members of this element (contiguous) with this atom's dE/dG.
    int const mbase = T.group[ed.g0].mem0;
    int nmTotal = 0;
    for (int gi = ed.g0; gi < ed.g0 + ed.ng; ++gi) nmTotal += T.group[gi].nmem;
    double const* __restrict__ dEdGi = a.dEdG + (size_t)i * a.ldG;
    for (int m = lane; m < nmTotal; m += 64)
    {
        MemberDesc const& md = T.member[mbase + m];
        MemLDS o;
        o.lambda = md.lambda;
        o.zl = md.zl;
        o.fderiv = md.fderiv;
        o.fnorm = md.fnorm;
        o.eta = md.eta;
        o.rs = md.rs;
        o.zeta = md.zeta;
        o.w = dEdGi[md.sf];
        o.zint = md.zint;
        o.useint = md.useint;
        o.calcexp = md.calcexp;
        o.pad = 0;
        M[m] = o;
    }
    // Angular group chunks per (e_j, e_k) and per-pair filter radii.
    for (int pq = lane; pq < ne * ne; pq += 64)
    {
        int const ej = pq / ne, ek = pq % ne;
        int const pb = (ec * ne + ej) * ne + ek;
        int const pc = T.pairCount[pb];
        double rmax = 0.0, rjk2 = 0.0;
        for (int t = 0; t < pc; ++t)
        {
            GroupDesc const& g = T.group[T.pairGroups[pb * kMaxPairGroups + t]];
            GrpLDS o;
            o.rc = g.rc;
            o.rc2 = g.rc2;
            o.type = g.type;
            o.cfg = g.cfg;
            o.nmem = g.nmem;
            o.mem = g.mem0 - mbase;
            GR[pq * kMaxPairGroups + t] = o;
            rmax = fmax(rmax, g.rc);
            rjk2 = fmax(rjk2, g.type == 9 ? 1.0e300 : g.rc2);
        }
        PR[pq] = rmax;      // 0 if no angular group for this pair
        PJ[pq] = rjk2;
    }
    __syncthreads();

    double sx = 0.0, sy = 0.0, sz = 0.0;     // this lane's part of cself
    for (int jr = 0; jr < n; jr += 64)
    {
        int const j = jr + lane;
        bool const act = j < n;
        int const jj = act ? j : 0;
        double const xj = X[jj], yj = Y[jj], zj = Z[jj], dj = D[jj];
        int const ej = E[jj];
        double cx = 0.0, cy = 0.0, cz = 0.0;  // slot-j contribution

        // Radial groups (SymGrpExpRad.cpp:194-233): uniform group loop.
        for (int gi = ed.g0; gi < ed.g0 + ed.nrad; ++gi)
        {
            GroupDesc const g = T.group[gi];
            if (act && g.e1 == ej && dj < g.rc)
            {
                double const fc = FC[g.cfg * cap + jj];
                double const dfc = DFC[g.cfg * cap + jj];
                double P = 0.0;
                for (int m = 0; m < g.nmem; ++m)
                {
                    MemberDesc const& md = T.member[g.mem0 + m];
                    double const pexp = exp(-md.eta * (dj - md.rs) * (dj - md.rs));
                    double const p1 = md.sfac * (dfc - 2.0 * md.eta
                                    * (dj - md.rs) * fc) * pexp / dj;
                    P += M[g.mem0 - mbase + m].w * p1;
                }
                sx += P * xj; sy += P * yj; sz += P * zj;
                cx -= P * xj; cy -= P * yj; cz -= P * zj;
            }
        }

        // Angular: bitmask of partners k passing the pair filters.
        for (int w = 0; w < nw; ++w)
        {
            uint32_t bits = 0;
            int const kmax = min(32, n - w * 32);
            for (int b = 0; b < kmax; ++b)
            {
                int const k = w * 32 + b;
                int const pq = ej * ne + E[k];
                double const rmax = PR[pq];
                double const dx = X[k] - xj, dy = Y[k] - yj, dz = Z[k] - zj;
                double const rjk2 = dx * dx + dy * dy + dz * dz;
                bool const pass = (k != jj) && dj < rmax && D[k] < rmax
                                && rjk2 < PJ[pq];
                bits |= (pass ? 1u : 0u) << b;
            }
            MASK[w * 64 + lane] = act ? bits : 0u;
        }
        if (act)
        {
            int w = 0;
            uint32_t bits = nw > 0 ? MASK[lane] : 0u;
            for (;;)
            {
                while (bits == 0u && w < nw)
                {
                    ++w;
                    if (w < nw) bits = MASK[w * 64 + lane];
                }
                if (w >= nw) break;
                int const b = __builtin_ctz(bits);
                bits &= bits - 1u;
                int const k = w * 32 + b;
                int const ek = E[k];
                int const pq = ej * ne + ek;
                int const pb = (ec * ne + ej) * ne + ek;
                int const pc = T.pairCount[pb];
                double const xk = X[k], yk = Y[k], zk = Z[k], dk = D[k];
                double const dx = xk - xj, dy = yk - yj, dz = zk - zj;
                double const rjk2 = dx * dx + dy * dy + dz * dz;
                for (int t = 0; t < pc; ++t)
                {
                    GrpLDS const g = GR[pq * kMaxPairGroups + t];
                    bool const wide = (g.type == 9);
                    if (!(dj < g.rc && dk < g.rc && (wide || rjk2 < g.rc2)))
                        continue;
                    double const rij = dj, rik = dk;
                    double const fcij = FC[g.cfg * cap + jj];
                    double const dfcij = DFC[g.cfg * cap + jj];
                    double const fcik = FC[g.cfg * cap + k];
                    double const r2ij = rij * rij, r2ik = rik * rik;
                    double pfc, r2sum, pr1, pr3 = 0.0, rjk = 0.0;
                    if (!wide)
                    {
                        rjk = sqrt(rjk2);
                        double fcjk, dfcjk;
                        cutoff_fdf(T.cfg[ed.cfg0 + g.cfg], rjk, fcjk, dfcjk);
                        pfc = fcij * fcik * fcjk;
                        r2sum = r2ij + r2ik + rjk * rjk;
                        pr1 = fcik * fcjk * dfcij / rij;
                        pr3 = fcij * fcik * dfcjk / rjk;
                    }
                    else
                    {
                        pfc = fcij * fcik;
                        r2sum = r2ij + r2ik;
                        pr1 = fcik * dfcij / rij;
                    }
                    double const rinvijik = 1.0 / rij / rik;
                    double const costijk = (xj * xk + yj * yk + zj * zk)
                                         * rinvijik;
                    double vexp = 0.0, rijs = 0.0, rjks = 0.0;
                    double P1 = 0.0, P3 = 0.0;
                    for (int m = 0; m < g.nmem; ++m)
                    {
                        MemLDS const mp = M[g.mem + m];
                        if (mp.calcexp)
                        {
                            if (mp.rs > 0.0)
                            {
                                rijs = rij - mp.rs;
                                double const riks = rik - mp.rs;
                                double s = rijs * rijs + riks * riks;
                                if (!wide)
                                {
                                    rjks = rjk - mp.rs;
                                    s += rjks * rjks;
                                }
                                vexp = exp(-mp.eta * s);
                            }
                            else vexp = exp(-mp.eta * r2sum);
                        }
                        double const plambda = 1.0 + mp.lambda * costijk;
                        double fg = vexp;
                        if (plambda <= 0.0) fg = 0.0;
                        else if (mp.useint) fg *= pow_int(plambda, mp.zint - 1);
                        else fg *= pow(plambda, mp.zeta - 1.0);
                        fg *= mp.fnorm;
                        double const pfczl = pfc * mp.zl;
                        double const p2etapl = plambda * mp.fderiv;
                        double p1, p3;
                        if (mp.rs > 0.0)
                        {
                            p1 = fg * (pfczl * (rinvijik - costijk / r2ij
                               - p2etapl * rijs / rij) + pr1 * plambda);
                            p3 = wide ? fg * pfczl * rinvijik
                                      : fg * (pfczl * (rinvijik + p2etapl
                                        * rjks / rjk) - pr3 * plambda);
                        }
                        else
                        {
                            p1 = fg * (pfczl * (rinvijik - costijk / r2ij
                               - p2etapl) + pr1 * plambda);
                            p3 = wide ? fg * pfczl * rinvijik
                                      : fg * (pfczl * (rinvijik + p2etapl)
                                        - pr3 * plambda);
                        }
                        P1 += mp.w * p1;
                        P3 += mp.w * p3;
                    }
                    sx += P1 * xj; sy += P1 * yj; sz += P1 * zj;
                    cx -= P1 * xj + P3 * dx;
                    cy -= P1 * yj + P3 * dy;
                    cz -= P1 * zj + P3 * dz;
                }
            }
            size_t const q = (size_t)(s0 + j);
            a.cslot[3 * q] = cx;
            a.cslot[3 * q + 1] = cy;
            a.cslot[3 * q + 2] = cz;
        }
        __syncthreads();
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    if (lane == 0)
    {
        a.cself[3 * (size_t)i] = sx;
        a.cself[3 * (size_t)i + 1] = sy;
        a.cself[3 * (size_t)i + 2] = sz;
    }
}

// F_m = -(cself[m] + sum_q cslot[partner(q)]) over m's slots q; mode 0
// skips slots pointing at m itself (Mode.cpp:1885 starts the unique list
// after the centre atom), mode 1 keeps them.  One wave per atom, fixed
// lane assignment and butterfly: deterministic.
__global__ __launch_bounds__(256) void k_gather(
    int n, const int* __restrict__ off, const int* __restrict__ nbrJ,
    const int* __restrict__ partner, const double* __restrict__ cslot,
    const double* __restrict__ cself, int mode, double* __restrict__ F)
{
    int const i = blockIdx.x * 4 + (threadIdx.x >> 6);
    int const lane = threadIdx.x & 63;
    if (i >= n) return;
    int const s0 = off[i], s1 = off[i + 1];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int q = s0 + lane; q < s1; q += 64)
    {
        int const pq = partner[q];
        // The list is symmetric, so every entry has a partner; pq < 0 would
        // mean an inconsistent list and is skipped rather than read.
        if ((mode == 1 || nbrJ[q] != i) && pq >= 0)
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

}  // namespace

size_t force_backward_lds_bytes(int cap, int maxMembers, int maxCfg, int maxnsf)
{
    (void)maxnsf;
    size_t b = 0;
    b += (size_t)cap * 4 * sizeof(double);
    b += (size_t)cap * 2 * maxCfg * sizeof(double);
    b += (size_t)maxMembers * sizeof(MemLDS);
    b += (size_t)kMaxElements * kMaxElements * kMaxPairGroups * sizeof(GrpLDS);
    b += (size_t)kMaxElements * kMaxElements * 2 * sizeof(double);
    b += (size_t)cap * sizeof(int);
    b += (size_t)((cap + 31) / 32) * 64 * sizeof(uint32_t);
    return b;
}

hipError_t launch_force_backward(ForceArgs const& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    size_t const shm = force_backward_lds_bytes(a.cap, a.maxMembers, a.maxCfg,
                                                a.T.maxnsf);
    k_force_backward<<<a.n, 64, shm, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_force_gather(int n, const int* off, const int* nbrJ,
                               const int* partner, const double* cslot,
                               const double* cself, int mode, double* F,
                               hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_gather<<<(n + 3) / 4, 256, 0, s>>>(n, off, nbrJ, partner, cslot, cself,
                                         mode, F);
    return hipGetLastError();
}

}  // namespace hdnnp
