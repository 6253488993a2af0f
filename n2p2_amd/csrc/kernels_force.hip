// Force kernels: contraction dE/dG . dG/dr with dG/dr recomputed, then a
// deterministic per-atom gather.  The contraction runs in FP64 (plain path)
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
//  * neighbours staged in LDS partitioned by element (as the SF kernel);
//  * groups run one after another, so member parameters and the atom's
//    dE/dG weights are wave-uniform;
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
//    T3 are formed once per run from U = sum w pw plambda; the powers come
//    from pow_uniform (straight-line products per uniform exponent);
//  * every ordered pair (j, k) of a group is evaluated by the lane that
//    owns slot j and contributes -(P1 dr_ij + P3 dr_jk) to slot j and
//    P1 dr_ij to the centre; slot k's part comes from the pair (k, j).  A
//    lane sums P1 and P3 dr_jk over its run of pairs of one row j and
//    applies dr_ij once per run.  The group's ordered pairs are split into
//    64 equal contiguous chunks, one per lane; a lane adds the sum of every
//    slot whose pairs start inside its chunk straight into the LDS
//    accumulator and leaves the sum of the slot it inherited from the
//    previous chunk in a per-lane head buffer, merged afterwards in lane
//    order.  No atomics; fixed summation order.
#include "device_common.h"
#include "kernels.h"

namespace hdnnp {

namespace {

// Pass-bit words per lane: a group has fewer than cap^2 ordered pairs, so a
// lane's chunk holds at most ceil(cap^2 / 64) of them.
__host__ __device__ inline int mask_words(int cap)
{
    int const span = (cap * cap + 63) / 64;
    return (span + 31) / 32;
}

// LDS of one wave.  1/r is staged in FP64 only (FP32 takes the hardware
// reciprocal); f_c'(r) is formed per row, not staged.
template <typename R>
struct BLds
{
    R* ACC;          // 3 per staged neighbour
    R* HEAD;         // 3 per lane
    R* X;
    R* Y;
    R* Z;
    R* D;
    R* INV;          // FP64 only
    R* FC;
    int* SLOT;       // original slot of each staged neighbour
    int* HEADJ;      // per lane
    uint32_t* MASK;  // [word][64]
    __device__ R inv(int j) const
    {
        if (sizeof(R) == 8) return INV[j];
        return rcp_r(D[j]);
    }
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
                    S.X[pos] = (R)g.x;
                    S.Y[pos] = (R)g.y;
                    S.Z[pos] = (R)g.z;
                    S.D[pos] = (R)g.w;
                    if (sizeof(R) == 8) S.INV[pos] = R(1) / (R)g.w;
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
    // j, k as staged indices; c = position of q in row j (q - c is the
    // row's first pair); len = the row's length.
    __device__ void decode(int q, int& j, int& k, int& c, int& len) const
    {
        int r;
        if (same)
        {
            int const d = na - 1;
            r = (int)((float)q * invA);
            c = q - r * d;
            if (c < 0) { --r; c += d; } else if (c >= d) { ++r; c -= d; }
            j = a0 + r;
            k = a0 + (c < r ? c : c + 1);
            len = d;
        }
        else if (q < split)
        {
            r = (int)((float)q * invA);
            c = q - r * nb;
            if (c < 0) { --r; c += nb; } else if (c >= nb) { ++r; c -= nb; }
            j = a0 + r;
            k = b0 + c;
            len = nb;
        }
        else
        {
            int const qq = q - split;
            r = (int)((float)qq * invB);
            c = qq - r * na;
            if (c < 0) { --r; c += na; } else if (c >= na) { ++r; c -= na; }
            j = b0 + r;
            k = a0 + c;
            len = na;
        }
    }
    // Sequential walk: (j, c, len, half) of pair q -> pair q + 1.
    __device__ void next(int& j, int& c, int& len, int& half) const
    {
        if (++c < len) return;
        c = 0;
        if (!same && half == 0 && j == a0 + na - 1)
        {
            half = 1;
            j = b0;
            len = na;
        }
        else ++j;
    }
    __device__ int partner(int j, int c, int half) const
    {
        if (same) return a0 + (c < j - a0 ? c : c + 1);
        return half == 0 ? b0 + c : a0 + c;
    }
};

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

template <int CT, bool GENPOW, typename R>
__global__ __launch_bounds__(64) void k_force_backward(ForceArgs a)
{
    extern __shared__ double sm[];
    __shared__ int sStart[kMaxElements], sCnt[kMaxElements];
    __shared__ MParam<R> MP[kMaxMembers];
    int const cap = a.cap;
    BLds<R> S;
    S.ACC = (R*)sm;
    S.HEAD = S.ACC + 3 * cap;
    S.X = S.HEAD + 3 * 64;
    S.Y = S.X + cap;
    S.Z = S.Y + cap;
    S.D = S.Z + cap;
    S.INV = S.D + cap;
    S.FC = S.INV + (sizeof(R) == 8 ? cap : 0);
    S.SLOT = (int*)(S.FC + cap);
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
                cutoff_fdf_t<CT>(cc, S.D[t], f, df);
                S.FC[t] = f;
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
                R const r = S.D[j];
                if (!(r < grc)) continue;
                R fc, dfc;
                cutoff_fdf_t<CT>(cc, r, fc, dfc);
                R const inv = S.inv(j);
                R P = R(0);
                for (int m = 0; m < nmem; ++m)
                {
                    R const eta = MP[m].eta, rs = MP[m].rs;
                    R const pexp = exp_r(-eta * (r - rs) * (r - rs));
                    P += MP[m].w * ((dfc - R(2) * eta * (r - rs) * fc) * pexp
                                    * inv);
                }
                R const px = P * S.X[j];
                R const py = P * S.Y[j];
                R const pz = P * S.Z[j];
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
        // Chunk of this lane: [q0, q1); the mask buffer holds
        // mask_words(cap) words per lane, enough for the largest chunk.
        {
            int const span = (total + 63) / 64;
            int const q0 = min(total, lane * span);
            int const q1 = min(total, q0 + span);
            int const len = q1 - q0;
            int const nw = (span + 31) >> 5;
            R const grc2 = (R)g.rc2;
            // Pass bits for the chunk (rij, rik < rc; rjk < rc if narrow),
            // walking the chunk's pairs in order.
            int wj = 0, wc = 0, wl = 1, wh = 0;
            if (len > 0)
            {
                int k0;
                op.decode(q0, wj, k0, wc, wl);
                wh = (!op.same && q0 >= op.split) ? 1 : 0;
            }
            for (int wd = 0; wd < nw; ++wd)
            {
                uint32_t bits = 0;
                for (int b = 0; b < 32; ++b)
                {
                    int const t = wd * 32 + b;
                    if (t < len)
                    {
                        int const j = wj;
                        int const k = op.partner(wj, wc, wh);
                        op.next(wj, wc, wl, wh);
                        bool pass = S.D[j] < grc && S.D[k] < grc;
                        if (pass && !wide)
                        {
                            R const dx = S.X[k] - S.X[j];
                            R const dy = S.Y[k] - S.Y[j];
                            R const dz = S.Z[k] - S.Z[j];
                            pass = dx * dx + dy * dy + dz * dz < grc2;
                        }
                        bits |= (pass ? 1u : 0u) << b;
                    }
                }
                S.MASK[wd * 64 + lane] = bits;
            }
            // Row of the chunk's first pair (j0); owned if it starts here.
            // Rows starting later are always owned.
            int j0 = -1;
            bool firstOwned = true;
            // Row of the chunk's last pair, if it starts in this chunk and
            // runs past it: this lane merges its continuation.
            int tailJ = -1, tailEnd = 0;
            if (len > 0)
            {
                int k0, c0, l0;
                op.decode(q0, j0, k0, c0, l0);
                firstOwned = (c0 == 0);
                int jl, kl, cl, ll;
                op.decode(q1 - 1, jl, kl, cl, ll);
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
            // Quantities of the current row j, held in registers.
            R rij = R(0), invj = R(0), xj = R(0), yj = R(0), zj = R(0);
            R fcij = R(0), dfcij = R(0);
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
                int j, k, cpos, rlen;
                op.decode(q0 + wd * 32 + b, j, k, cpos, rlen);
                if (j != curJ)
                {
                    if (curHas)
                        flush(S, curJ, curOwned, rsum, xj, yj, zj, sx, sy,
                              sz);
                    curOwned = (j != j0) || firstOwned;
                    curJ = j;
                    rsum.clear();
                    curHas = false;
                    rij = S.D[j];
                    invj = S.inv(j);
                    xj = S.X[j];
                    yj = S.Y[j];
                    zj = S.Z[j];
                    cutoff_fdf_t<CT>(cc, rij, fcij, dfcij);
                }
                // ---- triplet (j as the first neighbour, k second) ----
                R const rik = S.D[k];
                R const invk = S.inv(k);
                R const xk = S.X[k], yk = S.Y[k], zk = S.Z[k];
                R const dx = xk - xj, dy = yk - yj, dz = zk - zj;
                R const fcik = S.FC[k];
                R const rinvijik = invj * invk;
                R const costijk = (xj * xk + yj * yk + zj * zk) * rinvijik;
                R const r2ij = rij * rij, r2ik = rik * rik;
                R pfc, r2sum, pr1, pr3 = R(0), rjk = R(0), invjk = R(0);
                if (!wide)
                {
                    R const rjk2 = dx * dx + dy * dy + dz * dz;
                    invjk = rsqrt_nr(rjk2);
                    rjk = rjk2 * invjk;
                    R fcjk, dfcjk;
                    cutoff_fdf_t<CT>(cc, rjk, fcjk, dfcjk);
                    pfc = fcij * fcik * fcjk;
                    r2sum = r2ij + r2ik + rjk2;
                    pr1 = fcik * fcjk * dfcij * invj;
                    pr3 = fcij * fcik * dfcjk * invjk;
                }
                else
                {
                    pfc = fcij * fcik;
                    r2sum = r2ij + r2ik;
                    pr1 = fcik * dfcij * invj;
                }
                R T1 = R(0), T2 = R(0), T3 = R(0);
                R vexp = R(0), c1 = R(0), c3 = R(0), S1 = R(0), U = R(0);
                for (int m = 0; m < nmem; ++m)
                {
                    int const flags = uniform(MP[m].flags);
                    if (flags & MF_NEWEXP)
                    {
                        // Close the previous (eta, rs) run.
                        T1 += vexp * S1;
                        R const vu = vexp * U;
                        T2 += vu * c1;
                        T3 += vu * c3;
                        S1 = U = R(0);
                        R const eta = MP[m].eta, rs = MP[m].rs;
                        R const te = R(2) * eta * pfc;
                        if (rs > R(0))
                        {
                            R const rijs = rij - rs;
                            R const riks = rik - rs;
                            R s = rijs * rijs + riks * riks;
                            c1 = pr1 - te * rijs * invj;
                            if (!wide)
                            {
                                R const rjks = rjk - rs;
                                s += rjks * rjks;
                                c3 = te * rjks * invjk - pr3;
                            }
                            vexp = exp_r(-eta * s);
                        }
                        else
                        {
                            c1 = pr1 - te;
                            if (!wide) c3 = te - pr3;
                            vexp = exp_r(-eta * r2sum);
                        }
                    }
                    R const pl = R(1) + MP[m].lambda * costijk;
                    if (pl > R(0))
                    {
                        R pw;
                        if (!GENPOW || (flags & MF_INTPOW))
                            pw = pow_uniform(pl, uniform(MP[m].zint) - 1);
                        else pw = pow(pl, MP[m].zeta - R(1));
                        R const fgw = MP[m].w * pw;
                        S1 += fgw * MP[m].zl;
                        U += fgw * pl;
                    }
                }
                T1 += vexp * S1;
                {
                    R const vu = vexp * U;
                    T2 += vu * c1;
                    T3 += vu * c3;
                }
                R const A = rinvijik - costijk * invj * invj;
                R const P1 = pfc * A * T1 + T2;
                R const P3 = pfc * rinvijik * T1 + T3;
                rsum.p1 += P1;
                rsum.vx += P3 * dx;
                rsum.vy += P3 * dy;
                rsum.vz += P3 * dz;
                curHas = true;
            }
            if (curHas)
                flush(S, curJ, curOwned, rsum, xj, yj, zj, sx, sy, sz);
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

template <typename R>
void launch_bwd(ForceArgs const& a, size_t shm, hipStream_t s)
{
    if (a.T.genpow)
        HDNNP_CUTOFF_DISPATCH(a.T.cutType,
            (k_force_backward<kCT, true, R><<<a.n, 64, shm, s>>>(a)));
    else
        HDNNP_CUTOFF_DISPATCH(a.T.cutType,
            (k_force_backward<kCT, false, R><<<a.n, 64, shm, s>>>(a)));
}

}  // namespace

size_t force_backward_lds_bytes(int cap, int fp32)
{
    size_t const r = fp32 ? sizeof(float) : sizeof(double);
    size_t b = 0;
    b += (size_t)cap * 3 * r;                     // ACC
    b += (size_t)64 * 3 * r;                      // HEAD
    b += (size_t)cap * (fp32 ? 5 : 6) * r;        // X Y Z D [INV] FC
    b += (size_t)cap * sizeof(int);               // SLOT
    b += (size_t)64 * sizeof(int);                // HEADJ
    b += (size_t)mask_words(cap) * 64 * sizeof(uint32_t);   // MASK
    return b;
}

hipError_t launch_force_backward(ForceArgs const& a, hipStream_t s)
{
    if (a.n <= 0) return hipSuccess;
    size_t const shm = force_backward_lds_bytes(a.cap, a.fp32);
    if (a.fp32) launch_bwd<float>(a, shm, s);
    else launch_bwd<double>(a, shm, s);
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
