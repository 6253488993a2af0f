// Device code of the force backward kernel, shared by the ahead-of-time
// build (generic member loop reading the group tables) and the
// model-specialised build compiled at run time (HDNNP_PLAN defined and
// plan_backward_members provided before this header).
//
// Force kernels: contraction dE/dG . dG/dr with dG/dr recomputed, then a
// deterministic per-atom gather (the gather lives in kernels_force.hip).
// The contraction runs in FP64 (plain path) or in FP32 (mixed path:
// geometry, member sums and the per-triplet contributions); the per-slot
// accumulators, cslot, cself and the gather are FP64 on both.
//
// The reference stores dG/dr per SF and per neighbour slot
// (SymGrp*::calculate with derivatives, Atom::allocate) and contracts them
// in Mode::calculateForces (Mode.cpp:1849-1913) with
// Atom::calculateSelfForceShort / calculatePairForceShort (Atom.cpp:371-401).
// Here nothing per SF is stored: for atom i the kernel produces
//   cslot[i, q]  = sum_s dE_i/dG_s * dG_s/dr_(neighbour slot q) (3 doubles)
//   cself[i]     = sum_s dE_i/dG_s * dG_s/dr_i = -sum_q cslot[i, q]
// (every G_s depends on relative positions only, so its gradients over the
// centre and all neighbour slots sum to zero), and the gather kernel forms
// F_m = -(cself[m] + sum over m's slots q pointing at another atom j of
// cslot[j, partner(q)]), partner(q) being the slot of j that points back at
// m with the opposite image.
//
// Backward kernel, one wavefront per atom:
//  * neighbours staged in LDS partitioned by element (as the SF kernel), as
//    packed records (x, y, z, r) and (f_c(r), 1/r, f_c'(r)/r) for the
//    current cutoff;
//  * groups run one after another; a group's member parameters, with this
//    atom's dE/dG weights, are held in scalar registers for the pair loop;
//  * angular groups enumerate the unordered pairs {j, k} of their blocks in
//    diagonal order (DiagPairs), test rij, rik < rc and rjk < rc (narrow) and
//    push passing pairs into an LDS queue with a ballot prefix; whenever 64
//    are queued every lane evaluates one triplet, so the work is divergence
//    free and each triplet is evaluated once for both of its slots;
//  * the member sum is contracted before the vector work.  With
//    fg = exp(.) (1 + lambda cos)^(zeta-1) (SymGrpExpAngn.cpp:276-297) the
//    reference's p1 (slot j), its mirror for slot k and p3 (:300-337) are
//      P1j = pfc Aj T1 + pr1j V - pfc Ej,  Aj = 1/(rij rik) - cos/rij^2
//      P1k = pfc Ak T1 + pr1k V - pfc Ek,  Ak = 1/(rij rik) - cos/rik^2
//      P3  = pfc/(rij rik) T1 + pfc E3 - pr3 V  (last two terms narrow only)
//    with T1 = sum w fg zeta lambda, V = sum w fg plambda,
//    Ej = sum w fg plambda 2 eta (rij - rs)/rij, Ek and E3 alike with rik
//    and rjk (rs = 0: all 2 eta), w_m = dE/dG_m * factorNorm_m,
//    pr1j = f'c(rij)/rij fc(rik) fc(rjk) and pr3 = fc(rij) fc(rik)
//    f'c(rjk)/rjk (SymGrpExpAngw.cpp:300-310 for the wide type).  exp(.)
//    is shared by an (eta, rs) run of members and an (eta, rs, lambda) run
//    sorted by zeta shares one chain of powers (PowChain);
//  * slot j receives -(P1j dr_ij + P3 dr_jk), slot k -(P1k dr_ik - P3
//    dr_jk), dr_jk = dr_ik - dr_ij, through LDS atomic adds.  The
//    accumulators belong to this one wave, so the order of the adds is the
//    program order of the wave (lanes of one instruction in hardware
//    order): the result does not depend on timing and repeats bit for bit.
//    The diagonal pair order keeps the slots of a batch of 64 mostly
//    distinct, so the adds rarely collide.
#ifndef HDNNP_FORCE_CORE_H
#define HDNNP_FORCE_CORE_H

#include "device_common.h"
#include "kernel_args.h"

// The model-specialised build defines HDNNP_WEIGHTED 0 for models without
// weighted groups, which removes their code.
#ifndef HDNNP_WEIGHTED
#define HDNNP_WEIGHTED 1
#endif

namespace hdnnp {

namespace {

// LDS of one wave.
template <typename R>
struct BLds
{
    Vec4<R>* P;      // (x, y, z, r) per staged neighbour
    Vec4<R>* Q;      // (f_c(r), 1/r, f_c'(r)/r, 0) for the current cutoff
    double* ACC;     // x, y, z planes of cap doubles per staged neighbour
                     // (FP64 on both paths): lanes on neighbouring slots
                     // hit neighbouring banks
    int cap;
    int* QUEUE;      // kQueue passing pairs, j | k << 16
};

template <typename R>
__device__ __forceinline__ void stage(ForceArgs const& a, BLds<R> const& S,
                                      int s0, int n, int* sStart, int* sCnt)
{
    int const lane = threadIdx.x;
    int const ne = a.T.ne;
    int cnt[kMaxElements];
#pragma unroll
    for (int q = 0; q < kMaxElements; ++q) cnt[q] = 0;
    for (int base = 0; base < n; base += 64)
    {
        int const s = base + lane;
        int const e = s < n ? a.elem[a.nbrJ[s0 + s]] : -1;
#pragma unroll
        for (int q = 0; q < kMaxElements; ++q)
            if (q < ne) cnt[q] += __popcll(__ballot(e == q));
    }
    int start[kMaxElements];
    {
        int acc = 0;
#pragma unroll
        for (int q = 0; q < kMaxElements; ++q)
        {
            start[q] = acc;
            acc += (q < ne) ? cnt[q] : 0;
        }
    }
    if (lane == 0)
    {
#pragma unroll
        for (int q = 0; q < kMaxElements; ++q)
        {
            sStart[q] = start[q];
            sCnt[q] = (q < ne) ? cnt[q] : 0;
        }
    }
    for (int base = 0; base < n; base += 64)
    {
        int const s = base + lane;
        int e = -1;
        double4 g = make_double4(0, 0, 0, 1);
        if (s < n)
        {
            e = a.elem[a.nbrJ[s0 + s]];
            g = a.geo[s0 + s];
        }
#pragma unroll
        for (int q = 0; q < kMaxElements; ++q)
        {
            if (q < ne)
            {
                uint64_t const mask = __ballot(e == q);
                if (e == q)
                {
                    int const pos = start[q] + lanes_below(mask);
                    Vec4<R> p;
                    p.x = (R)g.x;
                    p.y = (R)g.y;
                    p.z = (R)g.z;
                    p.w = (R)g.w;
                    S.P[pos] = p;
                    Vec4<R> qq;
                    qq.x = R(0);
                    qq.y = R(1) / (R)g.w;
                    qq.z = R(0);
                    qq.w = R(0);
                    S.Q[pos] = qq;
                    S.ACC[pos] = 0.0;
                    S.ACC[S.cap + pos] = 0.0;
                    S.ACC[2 * S.cap + pos] = 0.0;
                }
                start[q] += __popcll(mask);
            }
        }
    }
    __syncthreads();
}

// A group's member parameters in wave-uniform (scalar) registers: lambda,
// zeta lambda, the weight and the flags word.  The exponential's parameters
// are read from LDS at the start of each (eta, rs) run.
template <typename R, int MAXM>
struct Members
{
    R lam[MAXM], zl[MAXM], w[MAXM], wz[MAXM];   // wz = w zeta lambda
    int code[MAXM];
};

template <typename R, int MAXM>
__device__ __forceinline__ void load_members(MParam<R> const* MP, int nmem,
                                             Members<R, MAXM>& M)
{
#pragma unroll
    for (int m = 0; m < MAXM; ++m)
    {
        MParam<R> const& p = MP[m < nmem ? m : 0];
        M.lam[m] = uniform_r(p.lambda);
        M.zl[m] = uniform_r(p.zl);
        M.w[m] = uniform_r(p.w);
        M.wz[m] = uniform_r(p.w * p.zl);
        M.code[m] = uniform(p.flags);
    }
}

// Atomic add to an LDS double of this wave's accumulators.
__device__ __forceinline__ void lds_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One triplet {j, k} of angular group gi, all members: adds the slot-j and
// slot-k contributions to the LDS accumulators.
template <int CT, bool GENPOW, int MAXM, typename R>
__device__ __forceinline__ void triplet_backward(
    int gi, BLds<R> const& S, int j, int k, bool wide, CutCfg const& cc,
    Members<R, MAXM> const& M, MParam<R> const* MP, int nmem)
{
    Vec4<R> const pj = S.P[j], pk = S.P[k];
    Vec4<R> const qj = S.Q[j], qk = S.Q[k];
    R const rij = pj.w, rik = pk.w;
    R const fcij = qj.x, invj = qj.y, dfcij = qj.z;
    R const fcik = qk.x, invk = qk.y, dfcik = qk.z;
    R const dx = pk.x - pj.x;
    R const dy = pk.y - pj.y;
    R const dz = pk.z - pj.z;
    R const rinvijik = invj * invk;
    R const costijk = (pj.x * pk.x + pj.y * pk.y + pj.z * pk.z) * rinvijik;
    R pfc, r2sum, pr1j, pr1k, pr3 = R(0), rjk = R(0), invjk = R(0);
    R const fij_ik = fcij * fcik;
    if (!wide)
    {
        R const rjk2 = dx * dx + dy * dy + dz * dz;
        invjk = rsqrt_nr(rjk2);
        rjk = rjk2 * invjk;
        R fcjk, dfcjk;
        cutoff_inside<CT>(cc, rjk, fcjk, dfcjk);
        pfc = fij_ik * fcjk;
        r2sum = rij * rij + rik * rik + rjk2;
        pr1j = fcik * fcjk * dfcij;
        pr1k = fcij * fcjk * dfcik;
        pr3 = fij_ik * dfcjk * invjk;
    }
    else
    {
        pfc = fij_ik;
        r2sum = rij * rij + rik * rik;
        pr1j = fcik * dfcij;
        pr1k = fcij * dfcik;
    }
    R T1 = R(0), V = R(0), Ej = R(0), Ek = R(0), E3 = R(0);
#ifdef HDNNP_PLAN
    // Member sums of the model-specialised build: straight-line code per
    // group chunk with the member constants folded in.
    (void)MP;
    (void)nmem;
    plan_backward_members<R>(gi, wide, costijk, r2sum, rij, rik, rjk, invj,
                             invk, invjk, M.w, M.wz, T1, V, Ej, Ek, E3);
#else
    (void)gi;
    R vexp = R(0), S1 = R(0), U = R(0), tj = R(0), tk = R(0), t3 = R(0);
    R pl = R(0);
    PowChain<R> chain;
    chain.reset(R(0));
#pragma unroll
    for (int m = 0; m < MAXM; ++m)
    {
        if (m < nmem)
        {
            int const fl = M.code[m];
            if (fl & MF_NEWEXP)
            {
                if (m > 0)
                {
                    // Close the previous (eta, rs) run.
                    T1 += vexp * S1;
                    R const vu = vexp * U;
                    V += vu;
                    Ej += vu * tj;
                    Ek += vu * tk;
                    E3 += vu * t3;
                    S1 = U = R(0);
                }
                R const rs = MP[m].rs, eta = MP[m].eta;
                R const neta = exp_factor(eta);
                R const te = R(2) * eta;
                if (rs > R(0))
                {
                    R const rijs = rij - rs;
                    R const riks = rik - rs;
                    R s = rijs * rijs + riks * riks;
                    tj = te * rijs * invj;
                    tk = te * riks * invk;
                    if (!wide)
                    {
                        R const rjks = rjk - rs;
                        s += rjks * rjks;
                        t3 = te * rjks * invjk;
                    }
                    vexp = exp_scaled(neta * s);
                }
                else
                {
                    tj = tk = t3 = te;
                    vexp = exp_scaled(neta * r2sum);
                }
            }
            if (fl & MF_NEWLAM)
            {
                pl = R(1) + M.lam[m] * costijk;
                chain.reset(pl);
            }
            R pw;
            if (fl & MF_PREFIX) pw = chain.at((fl >> 8) & 15);
            else if (!GENPOW || (fl & MF_INTPOW))
                pw = pow_int(pl, MP[m].zint - 1);
            else pw = pow(pl, MP[m].zeta - R(1));
            R const fgw = pl > R(0) ? M.w[m] * pw : R(0);
            S1 += fgw * M.zl[m];
            U += fgw * pl;
        }
    }
    T1 += vexp * S1;
    {
        R const vu = vexp * U;
        V += vu;
        Ej += vu * tj;
        Ek += vu * tk;
        E3 += vu * t3;
    }
#endif
    R const pt = pfc * T1;
    R const P1j = pt * (rinvijik - costijk * invj * invj) + pr1j * V
                - pfc * Ej;
    R const P1k = pt * (rinvijik - costijk * invk * invk) + pr1k * V
                - pfc * Ek;
    R const P3 = pt * rinvijik + (wide ? R(0) : pfc * E3 - pr3 * V);
    // FP64 LDS adds on both paths (the FP32 LDS add is much slower on this
    // hardware, and the FP64 sums keep the mixed path's slot totals exact
    // to FP64 rounding).
    double* const aj = S.ACC + j;
    double* const ak = S.ACC + k;
    int const c = S.cap;
    lds_add(aj, (double)(-(P1j * pj.x + P3 * dx)));
    lds_add(aj + c, (double)(-(P1j * pj.y + P3 * dy)));
    lds_add(aj + 2 * c, (double)(-(P1j * pj.z + P3 * dz)));
    lds_add(ak, (double)(P3 * dx - P1k * pk.x));
    lds_add(ak + c, (double)(P3 * dy - P1k * pk.y));
    lds_add(ak + 2 * c, (double)(P3 * dz - P1k * pk.z));
}

// Body of the backward kernel (one 64-lane workgroup per atom).
template <int CT, bool GENPOW, int MAXM, typename R>
__device__ __forceinline__ void force_backward_body(ForceArgs const& a)
{
    extern __shared__ __align__(16) double sm[];
    __shared__ int sStart[kMaxElements], sCnt[kMaxElements];
    __shared__ MParam<R> MP[kMaxMembers];
    int const cap = a.cap;
    BLds<R> S;
    S.P = (Vec4<R>*)sm;
    S.Q = S.P + cap;
    S.ACC = (double*)(S.Q + cap);
    S.cap = cap;
    S.QUEUE = (int*)(S.ACC + 3 * cap);

    int const i = blockIdx.x;
    int const lane = threadIdx.x;
    ModelTables const& T = a.T;
    int const ec = a.elem[i];
    ElemDesc const& ed = T.elem[ec];
    int const s0 = a.off[i];
    int const n = a.cnt[i];
    double const* __restrict__ dEdGi = a.dEdG + (size_t)i * a.ldG;

    stage(a, S, s0, n, sStart, sCnt);

    // The SF kernel's pair record of this atom, if any (sf_core.h).
    const unsigned short* __restrict__ rec =
        a.pairs ? a.pairs + (size_t)i * (size_t)a.pcap : nullptr;
    int pOff = kRecordGroups;

    int cfgCached = -1;
    for (int gi = ed.g0; gi < ed.g0 + ed.ng; ++gi)
    {
        GroupDesc const g = T.group[gi];
        bool const weighted = HDNNP_WEIGHTED && sf_weighted(g.type);
        // Cache key: the cutoff configuration and whether the values carry
        // the neighbour's atomic number.
        int const cfg = ed.cfg0 + g.cfg;
        int const key = 2 * cfg + (weighted ? 1 : 0);
        CutCfg const cc = T.cfg[cfg];
        if (key != cfgCached)
        {
            // f_c and f_c'/r of every staged neighbour for this cutoff;
            // weighted types: both times the neighbour's atomic number, so
            // every radial term carries Z_j and every triplet term Z_j Z_k
            // (SymGrpExpAngnWeighted.cpp:263-270).
            for (int t = lane; t < n; t += 64)
            {
                R f, df;
                cutoff_fdf_t<CT>(cc, S.P[t].w, f, df);
                if (weighted)
                {
                    R const z = (R)staged_weight(T, sStart, sCnt, t);
                    f *= z;
                    df *= z;
                }
                S.Q[t].x = f;
                S.Q[t].z = df * S.Q[t].y;
            }
            cfgCached = key;
        }
        MemberDesc const* __restrict__ mem = T.member + g.mem0;
        int const nmem = g.nmem;
        // Member parameters with this atom's dE/dG weights, in LDS.
        if (lane < nmem)
        {
            MemberDesc const& md = mem[lane];
            double const w = dEdGi[md.sf]
                           * (sf_radial(g.type) ? md.sfac : md.fnorm);
            MP[lane] = make_mparam<R>(md, w);
        }
        __syncthreads();
        R const grc = (R)g.rc;

        if (sf_radial(g.type))
        {
            // SymGrpExpRad.cpp:194-233: p1 = sf (f' - 2 eta (r-rs) f) e / r;
            // slot j receives -p1 dr_ij.  One lane per slot.
            int const a0 = weighted ? 0 : sStart[g.e1];
            int const na = weighted ? n : sCnt[g.e1];
            for (int t = lane; t < na; t += 64)
            {
                int const j = a0 + t;
                Vec4<R> const pj = S.P[j];
                R const r = pj.w;
                if (!(r < grc)) continue;
                Vec4<R> const qj = S.Q[j];
                R const fcinv = qj.x * qj.y, dfcinv = qj.z;
                R P = R(0);
                for (int m = 0; m < nmem; ++m)
                {
                    R const eta = MP[m].eta, rs = MP[m].rs;
                    R const d = r - rs;
                    R const pexp = exp_r(-eta * d * d);
                    P += MP[m].w * ((dfcinv - R(2) * eta * d * fcinv) * pexp);
                }
                S.ACC[j] -= (double)(P * pj.x);
                S.ACC[cap + j] -= (double)(P * pj.y);
                S.ACC[2 * cap + j] -= (double)(P * pj.z);
            }
            __syncthreads();
            continue;
        }

        bool const wide = (g.type == 9);
        // Weighted type: all unordered pairs of the whole list.
        DiagPairs const dp = weighted
            ? DiagPairs(true, 0, n, 0, n)
            : DiagPairs(g.e1 == g.e2, sStart[g.e1], sCnt[g.e1],
                        sStart[g.e2], sCnt[g.e2]);
        if (dp.npairs == 0)
        {
            __syncthreads();
            continue;
        }
        Members<R, MAXM> M;
        load_members(MP, nmem, M);
        int const nrec = rec && (gi - ed.g0) < kRecordGroups
                       ? (int)rec[gi - ed.g0] : (int)kRecordMissing;
        if (nrec != (int)kRecordMissing)
        {
            // Replay the recorded pairs: the same triplets in the same lanes
            // and order as the walk below would give.  The next batch's
            // entries are loaded while this batch is evaluated.
            int next = lane < nrec ? (int)rec[pOff + lane] : 0;
            for (int base = 0; base < nrec; base += 64)
            {
                int const e = next;
                if (base + 64 + lane < nrec)
                    next = rec[pOff + base + 64 + lane];
                if (base + lane < nrec)
                    triplet_backward<CT, GENPOW, MAXM>(gi, S, e & 0xff,
                                                       e >> 8, wide, cc, M,
                                                       MP, nmem);
            }
            pOff += nrec;
            __syncthreads();
            continue;
        }
        R const grc2 = (R)g.rc2;
        int qn = 0;
        int pc, pr;
        dp.start(lane, pc, pr);
        for (int base = 0; base < dp.npairs; base += 64)
        {
            bool const valid = base + lane < dp.npairs;
            int j, k;
            dp.pair(pc, pr, j, k);
            if (!valid) j = k = dp.s0;
            dp.advance(pc, pr);
            bool const pass = valid
                & pair_in_range<R>(S.P[j], S.P[k], wide, grc, grc2);
            uint64_t const mask = __ballot(pass);
            if (pass) S.QUEUE[qn + lanes_below(mask)] = j | (k << 16);
            qn += __popcll(mask);
            wave_lds_fence();
            if (qn >= 64)
            {
                int const e = S.QUEUE[qn - 64 + lane];
                triplet_backward<CT, GENPOW, MAXM>(gi, S, e & 0xffff, e >> 16,
                                                   wide, cc, M, MP, nmem);
                qn -= 64;
                wave_lds_fence();
            }
        }
        if (lane < qn)
        {
            int const e = S.QUEUE[lane];
            triplet_backward<CT, GENPOW, MAXM>(gi, S, e & 0xffff, e >> 16,
                                               wide, cc, M, MP, nmem);
        }
        __syncthreads();
    }
    // cself = -(sum of all slots), summed per lane in staged order and
    // then by the fixed butterfly.
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int t = lane; t < n; t += 64)
    {
        sx += S.ACC[t];
        sy += S.ACC[cap + t];
        sz += S.ACC[2 * cap + t];
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    if (lane == 0)
    {
        a.cself[3 * (size_t)i] = -sx;
        a.cself[3 * (size_t)i + 1] = -sy;
        a.cself[3 * (size_t)i + 2] = -sz;
    }
    // cslot in FP64, in slot order: each slot's staged position follows
    // from its element as in stage() (no slot map kept in LDS, which leaves
    // room for more workgroups per CU).
    int start[kMaxElements];
#pragma unroll
    for (int q = 0; q < kMaxElements; ++q) start[q] = sStart[q];
    int const ne = T.ne;
    for (int base = 0; base < n; base += 64)
    {
        int const s = base + lane;
        int const e = s < n ? a.elem[a.nbrJ[s0 + s]] : -1;
        int pos = 0;
#pragma unroll
        for (int q = 0; q < kMaxElements; ++q)
        {
            if (q < ne)
            {
                uint64_t const mask = __ballot(e == q);
                if (e == q) pos = start[q] + lanes_below(mask);
                start[q] += __popcll(mask);
            }
        }
        if (s < n)
        {
            size_t const q = (size_t)(s0 + s);
            a.cslot[3 * q] = S.ACC[pos];
            a.cslot[3 * q + 1] = S.ACC[cap + pos];
            a.cslot[3 * q + 2] = S.ACC[2 * cap + pos];
        }
    }
}

}  // namespace
}  // namespace hdnnp

#endif
