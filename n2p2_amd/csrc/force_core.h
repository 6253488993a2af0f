// Device code of the force backward kernel, shared by the ahead-of-time
// build (generic member loop reading the group tables) and the
// model-specialised build compiled at run time (HDNNP_PLAN defined and
// plan_backward_members provided before this header).
//
// Force kernels: contraction dE/dG . dG/dr with dG/dr recomputed, then a
// deterministic per-atom gather (the gather lives in kernels_force.hip).  The contraction runs in FP64 (plain path)
// or in FP32 (mixed path: geometry, member sums, row partials and the
// per-slot sums of a handful of row partials); the centre sums, cslot,
// cself and the gather are FP64 on both.
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
// Backward kernel, one wavefront per atom:
//  * neighbours staged in LDS partitioned by element (as the SF kernel), as
//    packed records (x, y, z, r) and (f_c(r), 1/r);
//  * groups run one after another; a group's member parameters, with this
//    atom's dE/dG weights, are held in scalar registers for the pair loop;
//  * the member sum is contracted before the vector work.  With
//    fg = exp(.) (1 + lambda cos)^(zeta-1) (SymGrpExpAngn.cpp:276-297) the
//    reference's p1 and p3 (:300-337) are
//      p1 = fg [pfc zeta lambda A + plambda c1],  A = 1/(rij rik) - cos/rij^2
//      p3 = fg [pfc zeta lambda / (rij rik) + plambda c3]
//    with c1 = pr1 - 2 eta pfc (rij - rs)/rij and c3 = 2 eta pfc (rjk -
//    rs)/rjk - pr3 (rs > 0), c1 = pr1 - 2 eta pfc and c3 = 2 eta pfc - pr3
//    (rs = 0), c3 = 0 for the wide type (SymGrpExpAngw.cpp:300-310); so
//      P1 = sum_m w_m p1_m = pfc A T1 + T2,  P3 = pfc/(rij rik) T1 + T3,
//      T1 = sum w fg zeta lambda, T2 = sum w fg plambda c1,
//      T3 = sum w fg plambda c3,  w_m = dE/dG_m * factorNorm_m.
//    c1, c3 and exp(.) are shared by an (eta, rs) run of members, so T2 and
//    T3 are formed once per run from U = sum w pw plambda; an (eta, rs,
//    lambda) run sorted by zeta shares one chain of powers (PowChain);
//  * every ordered pair (j, k) of a group is evaluated by the lane that
//    owns slot j and contributes -(P1 dr_ij + P3 dr_jk) to slot j and
//    P1 dr_ij to the centre; slot k's part comes from the pair (k, j).  A
//    lane sums P1 and P3 dr_jk over its run of pairs of one row j and
//    applies dr_ij once per run.  The group's ordered pairs are split into
//    64 equal contiguous chunks, one per lane; a lane first marks its
//    chunk's passing pairs in an LDS bit mask, then walks the set bits with
//    an incremental (row, column) cursor.  It adds the sum of every slot
//    whose pairs start inside its chunk straight into the LDS accumulator
//    and leaves the sum of the slot it inherited from the previous chunk in
//    a per-lane head buffer, merged afterwards in lane order.  No atomics;
//    fixed summation order.
#ifndef HDNNP_FORCE_CORE_H
#define HDNNP_FORCE_CORE_H

#include "device_common.h"
#include "kernel_args.h"

namespace hdnnp {

namespace {

// Pass-bit words per lane: a group has fewer than cap^2 ordered pairs, so a
// lane's chunk holds at most ceil(cap^2 / 64) of them.
__host__ __device__ inline int mask_words(int cap)
{
    int const span = (cap * cap + 63) / 64;
    return (span + 31) / 32;
}

// LDS of one wave.
template <typename R>
struct BLds
{
    Vec4<R>* P;      // (x, y, z, r) per staged neighbour
    Vec2<R>* Q;      // (f_c(r) for the current cutoff, 1/r)
    R* ACC;          // 3 per staged neighbour
    R* HEAD;         // 3 per lane
    int* SLOT;       // original slot of each staged neighbour
    int* HEADJ;      // per lane
    uint32_t* MASK;  // [word][64]
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
                    S.Q[pos].y = R(1) / (R)g.w;
                    S.SLOT[pos] = s;
                    S.ACC[3 * pos] = R(0);
                    S.ACC[3 * pos + 1] = R(0);
                    S.ACC[3 * pos + 2] = R(0);
                }
                start[q] += __popcll(mask);
            }
        }
    }
    __syncthreads();
}

// Ordered pairs (j, k) of a group, q in [0, total): for e1 != e2 first
// j in A x k in B, then j in B x k in A; for e1 == e2 j in A, k in A \ {j}.
struct OrderedPairs
{
    bool same;
    int a0, na, b0, nb, total, split;
    float invA, invB;
    __device__ OrderedPairs(bool s, int a0_, int na_, int b0_, int nb_)
        : same(s), a0(a0_), na(na_), b0(b0_), nb(nb_)
    {
        if (same)
        {
            total = na * (na - 1);
            split = total;
            invA = na > 1 ? 1.0f / (float)(na - 1) : 0.0f;
            invB = 0.0f;
        }
        else
        {
            split = na * nb;
            total = 2 * split;
            invA = nb > 0 ? 1.0f / (float)nb : 0.0f;   // rows of A
            invB = na > 0 ? 1.0f / (float)na : 0.0f;   // rows of B
        }
    }
    // j (staged index of the row), c = position of q in row j, len = the
    // row's length, half = 1 for the B x A part.
    __device__ void decode(int q, int& j, int& c, int& len, int& half) const
    {
        int r;
        if (same)
        {
            int const d = na - 1;
            r = (int)((float)q * invA);
            c = q - r * d;
            if (c < 0) { --r; c += d; } else if (c >= d) { ++r; c -= d; }
            j = a0 + r;
            len = d;
            half = 0;
        }
        else if (q < split)
        {
            r = (int)((float)q * invA);
            c = q - r * nb;
            if (c < 0) { --r; c += nb; } else if (c >= nb) { ++r; c -= nb; }
            j = a0 + r;
            len = nb;
            half = 0;
        }
        else
        {
            int const qq = q - split;
            r = (int)((float)qq * invB);
            c = qq - r * na;
            if (c < 0) { --r; c += na; } else if (c >= na) { ++r; c -= na; }
            j = b0 + r;
            len = na;
            half = 1;
        }
    }
    // The row after row j.
    __device__ void next_row(int& j, int& len, int& half) const
    {
        if (!same && half == 0 && j == a0 + na - 1)
        {
            half = 1;
            j = b0;
            len = na;
        }
        else ++j;
    }
    // Sequential walk: (j, c, len, half) of pair q -> pair q + 1.
    __device__ void next(int& j, int& c, int& len, int& half) const
    {
        if (++c < len) return;
        c = 0;
        next_row(j, len, half);
    }
    // Advances the cursor by d >= 0 pairs.
    __device__ void advance(int d, int& j, int& c, int& len, int& half) const
    {
        c += d;
        while (c >= len)
        {
            c -= len;
            next_row(j, len, half);
        }
    }
    __device__ int partner(int j, int c, int half) const
    {
        if (same) return a0 + (c < j - a0 ? c : c + 1);
        return half == 0 ? b0 + c : a0 + c;
    }
};

// A group's member parameters in wave-uniform (scalar) registers: lambda,
// zeta lambda, the weight and the flags word.  The exponential's parameters
// are read from LDS at the start of each (eta, rs) run.
template <typename R, int MAXM>
struct Members
{
    R lam[MAXM], zl[MAXM], w[MAXM];
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
        M.code[m] = uniform(p.flags);
    }
}

// A lane's partial sums over its run of passing pairs of one row j:
// sum P1 and sum P3 dr_jk.
template <typename R>
struct RowSum
{
    R p1, vx, vy, vz;
    __device__ void clear() { p1 = vx = vy = vz = R(0); }
};

// Ends a lane's run of row j (dr_ij = (xj, yj, zj)): slot j receives
// -(sum P1 dr_ij + sum P3 dr_jk) (straight into the accumulator if the row
// starts in this lane's chunk, else into the lane's head buffer), the
// centre's FP64 sum P1 dr_ij.
template <typename R>
__device__ __forceinline__ void flush(BLds<R> const& S, int j, bool owned,
                                      RowSum<R> const& rs, R xj, R yj, R zj,
                                      double& sx, double& sy, double& sz)
{
    R const ax = rs.p1 * xj;
    R const ay = rs.p1 * yj;
    R const az = rs.p1 * zj;
    sx += (double)ax;
    sy += (double)ay;
    sz += (double)az;
    R const cx = -(ax + rs.vx);
    R const cy = -(ay + rs.vy);
    R const cz = -(az + rs.vz);
    if (owned)
    {
        S.ACC[3 * j] += cx;
        S.ACC[3 * j + 1] += cy;
        S.ACC[3 * j + 2] += cz;
    }
    else
    {
        int const lane = threadIdx.x;
        S.HEADJ[lane] = j;
        S.HEAD[3 * lane] = cx;
        S.HEAD[3 * lane + 1] = cy;
        S.HEAD[3 * lane + 2] = cz;
    }
}

// Quantities of row j held in registers while a lane walks the row.
template <typename R>
struct Row
{
    R x, y, z, r, inv, fc, dfcinv, r2, inv2;
};

// P1 and P3 of one triplet (j first neighbour, k second), all members.
template <int CT, bool GENPOW, int MAXM, typename R>
__device__ __forceinline__ void triplet_backward(
    int gi, Row<R> const& J, Vec4<R> const& pk, Vec2<R> const& qk, bool wide,
    CutCfg const& cc, Members<R, MAXM> const& M, MParam<R> const* MP,
    int nmem, R& P1, R& P3, R& dx, R& dy, R& dz)
{
    R const rik = pk.w, fcik = qk.x, invk = qk.y;
    dx = pk.x - J.x;
    dy = pk.y - J.y;
    dz = pk.z - J.z;
    R const rinvijik = J.inv * invk;
    R const costijk = (J.x * pk.x + J.y * pk.y + J.z * pk.z) * rinvijik;
    R pfc, r2sum, pr1, pr3 = R(0), rjk = R(0), invjk = R(0);
    if (!wide)
    {
        R const rjk2 = dx * dx + dy * dy + dz * dz;
        invjk = rsqrt_nr(rjk2);
        rjk = rjk2 * invjk;
        R fcjk, dfcjk;
        cutoff_inside<CT>(cc, rjk, fcjk, dfcjk);
        R const fij_ik = J.fc * fcik;
        pfc = fij_ik * fcjk;
        r2sum = J.r2 + rik * rik + rjk2;
        pr1 = fcik * fcjk * J.dfcinv;
        pr3 = fij_ik * dfcjk * invjk;
    }
    else
    {
        pfc = J.fc * fcik;
        r2sum = J.r2 + rik * rik;
        pr1 = fcik * J.dfcinv;
    }
    R T1 = R(0), T2 = R(0), T3 = R(0);
#ifdef HDNNP_PLAN
    // Member sums of the model-specialised build: straight-line code per
    // group chunk with the member constants folded in.
    (void)MP;
    (void)nmem;
    plan_backward_members<R>(gi, wide, costijk, pfc, r2sum, pr1, pr3, J.r,
                             rik, rjk, J.inv, invjk, M.w, T1, T2, T3);
#else
    (void)gi;
    R vexp = R(0), c1 = R(0), c3 = R(0), S1 = R(0), U = R(0);
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
                    T2 += vu * c1;
                    T3 += vu * c3;
                    S1 = U = R(0);
                }
                R const rs = MP[m].rs, eta = MP[m].eta;
                R const neta = exp_factor(eta);
                R const te = R(2) * eta * pfc;
                if (rs > R(0))
                {
                    R const rijs = J.r - rs;
                    R const riks = rik - rs;
                    R s = rijs * rijs + riks * riks;
                    c1 = pr1 - te * rijs * J.inv;
                    if (!wide)
                    {
                        R const rjks = rjk - rs;
                        s += rjks * rjks;
                        c3 = te * rjks * invjk - pr3;
                    }
                    vexp = exp_scaled(neta * s);
                }
                else
                {
                    c1 = pr1 - te;
                    if (!wide) c3 = te - pr3;
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
        T2 += vu * c1;
        T3 += vu * c3;
    }
#endif
    R const A = rinvijik - costijk * J.inv2;
    P1 = pfc * A * T1 + T2;
    P3 = pfc * rinvijik * T1 + T3;
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
    S.Q = (Vec2<R>*)(S.P + cap);
    S.ACC = (R*)(S.Q + cap);
    S.HEAD = S.ACC + 3 * cap;
    S.SLOT = (int*)(S.HEAD + 3 * 64);
    S.HEADJ = S.SLOT + cap;
    S.MASK = (uint32_t*)(S.HEADJ + 64);

    int const i = blockIdx.x;
    int const lane = threadIdx.x;
    ModelTables const& T = a.T;
    int const ec = a.elem[i];
    ElemDesc const& ed = T.elem[ec];
    int const s0 = a.off[i];
    int const n = a.off[i + 1] - s0;
    double const* __restrict__ dEdGi = a.dEdG + (size_t)i * a.ldG;

    stage(a, S, s0, n, sStart, sCnt);

    double sx = 0.0, sy = 0.0, sz = 0.0;    // this lane's share of cself
    double rcCached = -1.0;
    for (int gi = ed.g0; gi < ed.g0 + ed.ng; ++gi)
    {
        GroupDesc const g = T.group[gi];
        CutCfg const cc = T.cfg[ed.cfg0 + g.cfg];
        if (g.rc != rcCached)
        {
            for (int t = lane; t < n; t += 64)
            {
                R f, df;
                cutoff_fdf_t<CT>(cc, S.P[t].w, f, df);
                S.Q[t].x = f;
            }
            rcCached = g.rc;
            __syncthreads();
        }
        MemberDesc const* __restrict__ mem = T.member + g.mem0;
        int const nmem = g.nmem;
        // Member parameters with this atom's dE/dG weights, in LDS.
        if (lane < nmem)
        {
            MemberDesc const& md = mem[lane];
            double const w = dEdGi[md.sf]
                           * (g.type == 2 ? md.sfac : md.fnorm);
            MP[lane] = make_mparam<R>(md, w);
        }
        __syncthreads();
        R const grc = (R)g.rc;

        if (g.type == 2)
        {
            // SymGrpExpRad.cpp:194-233: p1 = sf (f' - 2 eta (r-rs) f) e / r.
            int const a0 = sStart[g.e1], na = sCnt[g.e1];
            for (int t = lane; t < na; t += 64)
            {
                int const j = a0 + t;
                Vec4<R> const pj = S.P[j];
                R const r = pj.w;
                if (!(r < grc)) continue;
                R fc, dfc;
                cutoff_fdf_t<CT>(cc, r, fc, dfc);
                R const inv = S.Q[j].y;
                R P = R(0);
                for (int m = 0; m < nmem; ++m)
                {
                    R const eta = MP[m].eta, rs = MP[m].rs;
                    R const pexp = exp_r(-eta * (r - rs) * (r - rs));
                    P += MP[m].w * ((dfc - R(2) * eta * (r - rs) * fc) * pexp
                                    * inv);
                }
                R const px = P * pj.x;
                R const py = P * pj.y;
                R const pz = P * pj.z;
                sx += (double)px; sy += (double)py; sz += (double)pz;
                S.ACC[3 * j] -= px;
                S.ACC[3 * j + 1] -= py;
                S.ACC[3 * j + 2] -= pz;
            }
            __syncthreads();
            continue;
        }

        bool const wide = (g.type == 9);
        OrderedPairs op(g.e1 == g.e2, sStart[g.e1], sCnt[g.e1], sStart[g.e2],
                        sCnt[g.e2]);
        int const total = op.total;
        if (total == 0) continue;
        Members<R, MAXM> M;
        load_members(MP, nmem, M);
        // Chunk of this lane: [q0, q1); the mask buffer holds
        // mask_words(cap) words per lane, enough for the largest chunk.
        {
            int const span = (total + 63) / 64;
            int const q0 = min(total, lane * span);
            int const q1 = min(total, q0 + span);
            int const len = q1 - q0;
            int const nw = (span + 31) >> 5;
            R const grc2 = (R)g.rc2;
            // Cursor at the chunk's first pair.
            int cj = 0, cc0 = 0, clen = 1, chalf = 0;
            if (len > 0) op.decode(q0, cj, cc0, clen, chalf);
            int const j0 = cj;
            bool const firstOwned = (cc0 == 0);
            // Pass bits for the chunk (rij, rik < rc; rjk < rc if narrow),
            // walking the chunk's pairs in order.
            {
                int wj = cj, wc = cc0, wl = clen, wh = chalf;
                for (int wd = 0; wd < nw; ++wd)
                {
                    uint32_t bits = 0;
                    for (int b = 0; b < 32; ++b)
                    {
                        int const t = wd * 32 + b;
                        if (t < len)
                        {
                            int const k = op.partner(wj, wc, wh);
                            Vec4<R> const pj = S.P[wj];
                            Vec4<R> const pk = S.P[k];
                            op.next(wj, wc, wl, wh);
                            bool pass = pj.w < grc && pk.w < grc;
                            if (!wide)
                            {
                                R const dx = pk.x - pj.x;
                                R const dy = pk.y - pj.y;
                                R const dz = pk.z - pj.z;
                                pass = pass
                                    && dx * dx + dy * dy + dz * dz < grc2;
                            }
                            bits |= (pass ? 1u : 0u) << b;
                        }
                    }
                    S.MASK[wd * 64 + lane] = bits;
                }
            }
            // Row of the chunk's last pair, if it starts in this chunk and
            // runs past it: this lane merges its continuation.
            int tailJ = -1, tailEnd = 0;
            if (len > 0)
            {
                int jl, cl, ll, hl;
                op.decode(q1 - 1, jl, cl, ll, hl);
                int const rowStart = q1 - 1 - cl;
                if (rowStart >= q0 && rowStart + ll > q1)
                {
                    tailJ = jl;
                    tailEnd = rowStart + ll;
                }
            }
            S.HEADJ[lane] = -1;
            RowSum<R> rsum;
            rsum.clear();
            bool curHas = false;
            int curJ = -1;
            bool curOwned = true;
            Row<R> J;
            J.x = J.y = J.z = J.r = J.inv = J.fc = J.dfcinv = J.r2 = J.inv2
                = R(0);
            int qc = q0;            // position of the cursor
            int wd = 0;
            uint32_t bits = nw > 0 ? S.MASK[lane] : 0u;
            for (;;)
            {
                while (bits == 0u && wd < nw)
                {
                    ++wd;
                    if (wd < nw) bits = S.MASK[wd * 64 + lane];
                }
                if (wd >= nw) break;
                int const b = __builtin_ctz(bits);
                bits &= bits - 1u;
                int const q = q0 + wd * 32 + b;
                op.advance(q - qc, cj, cc0, clen, chalf);
                qc = q;
                int const j = cj;
                int const k = op.partner(cj, cc0, chalf);
                if (j != curJ)
                {
                    if (curHas)
                        flush(S, curJ, curOwned, rsum, J.x, J.y, J.z, sx, sy,
                              sz);
                    curOwned = (j != j0) || firstOwned;
                    curJ = j;
                    rsum.clear();
                    curHas = false;
                    Vec4<R> const pj = S.P[j];
                    Vec2<R> const qj = S.Q[j];
                    J.x = pj.x;
                    J.y = pj.y;
                    J.z = pj.z;
                    J.r = pj.w;
                    J.fc = qj.x;
                    J.inv = qj.y;
                    J.r2 = J.r * J.r;
                    J.inv2 = J.inv * J.inv;
                    R fcj, dfcj;
                    cutoff_fdf_t<CT>(cc, J.r, fcj, dfcj);
                    J.dfcinv = dfcj * J.inv;
                }
                R P1, P3, dx, dy, dz;
                triplet_backward<CT, GENPOW, MAXM>(gi, J, S.P[k], S.Q[k], wide,
                                                   cc, M, MP, nmem, P1, P3,
                                                   dx, dy, dz);
                rsum.p1 += P1;
                rsum.vx += P3 * dx;
                rsum.vy += P3 * dy;
                rsum.vz += P3 * dz;
                curHas = true;
            }
            if (curHas)
                flush(S, curJ, curOwned, rsum, J.x, J.y, J.z, sx, sy, sz);
            __syncthreads();
            // The lane owning the start of a row that runs past its chunk
            // adds the head partials of the following chunks of that row,
            // in lane order; one writer per row, chunks without passing
            // pairs (HEADJ = -1) are skipped.
            if (tailJ >= 0)
            {
                R hx = R(0), hy = R(0), hz = R(0);
                for (int L = lane + 1; L < 64 && L * span < tailEnd; ++L)
                    if (S.HEADJ[L] == tailJ)
                    {
                        hx += S.HEAD[3 * L];
                        hy += S.HEAD[3 * L + 1];
                        hz += S.HEAD[3 * L + 2];
                    }
                S.ACC[3 * tailJ] += hx;
                S.ACC[3 * tailJ + 1] += hy;
                S.ACC[3 * tailJ + 2] += hz;
            }
            __syncthreads();
        }
    }
    for (int t = lane; t < n; t += 64)
    {
        size_t const q = (size_t)(s0 + S.SLOT[t]);
        a.cslot[3 * q] = S.ACC[3 * t];
        a.cslot[3 * q + 1] = S.ACC[3 * t + 1];
        a.cslot[3 * q + 2] = S.ACC[3 * t + 2];
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

}  // namespace
}  // namespace hdnnp

#endif
