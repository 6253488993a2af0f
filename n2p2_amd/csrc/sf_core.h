// Device code of the symmetry-function forward kernel, shared by the
// ahead-of-time build (generic member loop reading the group tables) and
// the model-specialised build compiled at run time (HDNNP_PLAN defined and
// plan_forward_members provided before this header).
//
// Symmetry-function forward kernel: G values of every atom, in FP64 (plain
// path) or with FP32 geometry and member sums (mixed path; lane partials
// summed in FP64, scaling in FP64).
//
// Restates SymGrpExpRad::calculate (SymGrpExpRad.cpp:144-239),
// SymGrpExpAngn::calculate (SymGrpExpAngn.cpp:156-403),
// SymGrpExpAngw::calculate (SymGrpExpAngw.cpp:156-376) and the weighted
// SymGrpExpRadWeighted / SymGrpExpAngnWeighted::calculate
// (SymGrpExpRadWeighted.cpp:144-252, SymGrpExpAngnWeighted.cpp:166-385)
// without derivatives
// (the force kernel recomputes those).  One wavefront per atom:
//   1. the atom's neighbours are staged in LDS, partitioned by element
//      (stable ballot compaction), so every group reads exactly the
//      neighbour blocks it needs;
//   2. groups run one after another; a group's member parameters are staged
//      in LDS and the ones read per member held in scalar registers, its
//      member sums live in registers;
//   3. angular groups enumerate the unordered pairs of their blocks in
//      diagonal order (DiagPairs), test
//      rij, rik < rc and rjk < rc, and push passing pairs into an LDS queue
//      with a ballot prefix; whenever 64 are queued every lane evaluates
//      one, so the expensive triplet work runs without divergence;
//   4. inside a triplet the members are walked as (eta, rs) runs sharing one
//      exponential and (eta, rs, lambda) runs sorted by zeta sharing one
//      chain of powers of plambda (PowChain, bit-identical to pow_int);
//   5. member sums are reduced with a fixed butterfly (deterministic).
#ifndef HDNNP_SF_CORE_H
#define HDNNP_SF_CORE_H

#include "device_common.h"
#include "kernel_args.h"

// The model-specialised build defines HDNNP_WEIGHTED 0 for models without
// weighted groups, which removes their code.
#ifndef HDNNP_WEIGHTED
#define HDNNP_WEIGHTED 1
#endif

namespace hdnnp {

namespace {

// Staged neighbours: (x, y, z, r) and (f_c(r) for the current cutoff, 1/r).
template <typename R>
struct Stage
{
    Vec4<R>* P;
    Vec2<R>* Q;      // null when only geometry is needed
};

// Stages the atom's neighbours in LDS partitioned by element (stable ballot
// compaction); sStart / sCnt receive each element's block.
template <typename R>
__device__ void stage_by_element(SFArgs const& a, Stage<R> const& S, int s0,
                                 int n, int* sStart, int* sCnt)
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
                    if (S.Q) S.Q[pos].y = R(1) / (R)g.w;
                }
                start[q] += __popcll(mask);
            }
        }
    }
    __syncthreads();
}

// A group's member parameters in wave-uniform (scalar) registers: lambda
// and the flags word; the exponential's parameters are read from LDS at the
// start of each (eta, rs) run.
template <typename R, int MAXM>
struct FwdMembers
{
    R lam[MAXM];
    int code[MAXM];
};

// One triplet, all members of the group: acc[m] += fg * plambda * pfc with
// fg = exp(-eta sum r^2) plambda^(zeta - 1) (SymGrpExpAngn.cpp:276-297).
// ST: a staging record set with P (x, y, z, r) and Q (f_c, 1/r, ...).
template <int MAXM, int CT, bool GENPOW, typename R, typename ST>
__device__ __forceinline__ void triplet_forward(
    int gi, ST const& S, int ja, int kb, bool wide, CutCfg const& cc,
    FwdMembers<R, MAXM> const& M, MParam<R> const* MP, int nmem,
    R (&acc)[MAXM])
{
    Vec4<R> const pj = S.P[ja], pk = S.P[kb];
    auto const qj = S.Q[ja];
    auto const qk = S.Q[kb];
    R const rij = pj.w, rik = pk.w;
    R pfc, r2sum, rjk = R(0);
    if (!wide)
    {
        R const dx = pk.x - pj.x, dy = pk.y - pj.y, dz = pk.z - pj.z;
        R const rjk2 = dx * dx + dy * dy + dz * dz;
        rjk = rjk2 * rsqrt_nr(rjk2);
        R fcjk, dfcjk;
        cutoff_inside<CT>(cc, rjk, fcjk, dfcjk);
        pfc = qj.x * qk.x * fcjk;
        r2sum = rij * rij + rik * rik + rjk2;
    }
    else
    {
        pfc = qj.x * qk.x;
        r2sum = rij * rij + rik * rik;
    }
    R const rinvijik = qj.y * qk.y;
    R const costijk = (pj.x * pk.x + pj.y * pk.y + pj.z * pk.z) * rinvijik;
#ifdef HDNNP_PLAN
    // Member sums of the model-specialised build (straight-line code per
    // group chunk with the member constants folded in).
    (void)M;
    (void)MP;
    (void)nmem;
    plan_forward_members<R, MAXM>(gi, wide, costijk, pfc, r2sum, rij, rik,
                                  rjk, acc);
#else
    (void)gi;
    R vp = R(0);             // exp(.) * pfc of the current (eta, rs) run
    R pl = R(0);             // plambda of the current lambda run
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
                R const eta = MP[m].eta, rs = MP[m].rs;
                R s = r2sum;
                if (rs > R(0))
                {
                    R const a = rij - rs, b = rik - rs;
                    s = a * a + b * b;
                    if (!wide)
                    {
                        R const c = rjk - rs;
                        s += c * c;
                    }
                }
                vp = exp_scaled(exp_factor(eta) * s) * pfc;
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
            acc[m] += pl > R(0) ? vp * (pw * pl) : R(0);
        }
    }
#endif
}

// Body of the forward kernel (one 64-lane workgroup per atom).
template <int MAXM, int CT, bool GENPOW, typename R>
__device__ __forceinline__ void sf_forward_body(SFArgs const& a)
{
    static_assert(MAXM == 4 || MAXM == 8 || MAXM == 16,
                  "member chunks of 4, 8 or 16");
    extern __shared__ __align__(16) double sm[];
    __shared__ int sStart[kMaxElements], sCnt[kMaxElements];
    __shared__ MParam<R> MP[kMaxMembers];
    int const cap = a.cap;
    Stage<R> S;
    S.P = (Vec4<R>*)sm;
    S.Q = (Vec2<R>*)(S.P + cap);
    int* Q = (int*)(S.Q + cap);

    int const i = blockIdx.x;
    int const lane = threadIdx.x;
    ModelTables const& T = a.T;
    int const ec = a.elem[i];
    ElemDesc const& ed = T.elem[ec];
    int const s0 = a.off[i];
    int const n = a.cnt[i];

    stage_by_element(a, S, s0, n, sStart, sCnt);

    // Pair record (mixed path): group counts first, all marked missing
    // until their group is done; the pairs follow from kRecordGroups on.
    unsigned short* __restrict__ rec =
        a.pairs ? a.pairs + (size_t)i * (size_t)a.pcap : nullptr;
    bool const record = rec && cap <= 256;
    if (rec)
        for (int g = lane; g < kRecordGroups; g += 64) rec[g] = kRecordMissing;
    int pOff = kRecordGroups;   // next free record entry

    double rcCached = -1.0;
    bool wCached = false;
    double* __restrict__ Gi = a.G + (size_t)i * a.ldG;
    for (int gi = ed.g0; gi < ed.g0 + ed.ng; ++gi)
    {
        GroupDesc const g = T.group[gi];
        CutCfg const cc = T.cfg[ed.cfg0 + g.cfg];
        MemberDesc const* __restrict__ mem = T.member + g.mem0;
        int const nmem = g.nmem;
        bool const weighted = HDNNP_WEIGHTED && sf_weighted(g.type);
        if (lane < nmem) MP[lane] = make_mparam<R>(mem[lane], 0.0);
        if (g.rc != rcCached || weighted != wCached)
        {
            // Weighted types: f_c times the neighbour's atomic number, so
            // radial terms carry Z_j and triplet products Z_j Z_k
            // (SymGrpExpRadWeighted.cpp:200-207,
            // SymGrpExpAngnWeighted.cpp:263-270).
            for (int t = lane; t < n; t += 64)
            {
                R f, df;
                cutoff_fdf_t<CT>(cc, S.P[t].w, f, df);
                if (weighted) f *= (R)staged_weight(T, sStart, sCnt, t);
                S.Q[t].x = f;
            }
            rcCached = g.rc;
            wCached = weighted;
        }
        __syncthreads();
        R acc[MAXM];
#pragma unroll
        for (int m = 0; m < MAXM; ++m) acc[m] = R(0);

        if (sf_radial(g.type))
        {
            int const a0 = weighted ? 0 : sStart[g.e1];
            int const na = weighted ? n : sCnt[g.e1];
            R const rc = (R)g.rc;
            for (int t = lane; t < na; t += 64)
            {
                R const r = S.P[a0 + t].w;
                if (r < rc)
                {
                    R const fc = S.Q[a0 + t].x;
#pragma unroll
                    for (int m = 0; m < MAXM; ++m)
                        if (m < nmem)
                        {
                            R const eta = MP[m].eta, rs = MP[m].rs;
                            acc[m] += exp_r(-eta * (r - rs) * (r - rs)) * fc;
                        }
                }
            }
        }
        else
        {
            bool const wide = (g.type == 9);
            // Weighted type: all unordered pairs of the whole list.
            DiagPairs const dp = weighted
                ? DiagPairs(true, 0, n, 0, n)
                : DiagPairs(g.e1 == g.e2, sStart[g.e1], sCnt[g.e1],
                            sStart[g.e2], sCnt[g.e2]);
            R const grc = (R)g.rc, grc2 = (R)g.rc2;
            FwdMembers<R, MAXM> M;
#pragma unroll
            for (int m = 0; m < MAXM; ++m)
            {
                MParam<R> const& p = MP[m < nmem ? m : 0];
                M.lam[m] = uniform_r(p.lambda);
                M.code[m] = uniform(p.flags);
            }
            // Passing pairs of this group go to the record while they fit,
            // in the order and lane positions the force kernel would meet
            // them (staged indices j | k << 8).
            bool keep = record && (gi - ed.g0) < kRecordGroups;
            int nrec = 0;
            int qn = 0;
            int pc = 0, pr = 0;
            dp.start(lane, pc, pr);
            for (int base = 0; base < dp.npairs; base += 64)
            {
                bool const valid = base + lane < dp.npairs;
                int ja, kb;
                dp.pair(pc, pr, ja, kb);
                if (!valid) ja = kb = dp.s0;
                dp.advance(pc, pr);
                // Invalid lanes test (s0, s0), in bounds: no divergent
                // branch around the LDS reads.
                bool const pass = valid
                    & pair_in_range<R>(S.P[ja], S.P[kb], wide, grc, grc2);
                uint64_t const mask = __ballot(pass);
                if (pass) Q[qn + lanes_below(mask)] = ja | (kb << 16);
                qn += __popcll(mask);
                wave_lds_fence();
                if (qn >= 64)
                {
                    int const e = Q[qn - 64 + lane];
                    triplet_forward<MAXM, CT, GENPOW>(gi, S, e & 0xffff,
                                                      e >> 16, wide, cc, M, MP,
                                                      nmem, acc);
                    if (keep && pOff + nrec + 64 <= a.pcap)
                    {
                        rec[pOff + nrec + lane] = (unsigned short)(
                            (e & 0xff) | ((e >> 16) << 8));
                        nrec += 64;
                    }
                    else keep = false;
                    qn -= 64;
                    wave_lds_fence();
                }
            }
            if (lane < qn)
            {
                int const e = Q[lane];
                triplet_forward<MAXM, CT, GENPOW>(gi, S, e & 0xffff, e >> 16,
                                                  wide, cc, M, MP, nmem, acc);
            }
            if (keep && pOff + nrec + qn <= a.pcap)
            {
                if (lane < qn)
                {
                    int const e = Q[lane];
                    rec[pOff + nrec + lane] = (unsigned short)(
                        (e & 0xff) | ((e >> 16) << 8));
                }
                nrec += qn;
                if (lane == 0) rec[gi - ed.g0] = (unsigned short)nrec;
                pOff += nrec;
            }
        }
        // Member totals by the transposing butterfly: member m ends in the
        // lanes (m << (6 - log2 MAXM)) + t; the first of them writes it.
        double dv[MAXM];
#pragma unroll
        for (int m = 0; m < MAXM; ++m) dv[m] = (double)acc[m];
        double const v = wave_sum_transposed<MAXM>(dv);
        constexpr int kSh = MAXM == 4 ? 4 : MAXM == 8 ? 3 : 2;   // 6 - log2
        int const m = lane >> kSh;
        if ((lane & ((1 << kSh) - 1)) == 0 && m < nmem)
        {
            MemberDesc const& md = mem[m];
            double const raw = sf_radial(g.type) ? v : v * md.gfac;
            Gi[md.sf] = a.raw ? raw : scale_sf(T.scale[ed.sf0 + md.sf], raw);
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace hdnnp

#endif
