// Compact (polynomial) symmetry functions, types 20-25: SymGrpCompRad
// (SymGrpCompRad.cpp:128-214), SymGrpCompAngn (SymGrpCompAngn.cpp:148-388),
// SymGrpCompAngw and the weighted SymGrpComp*Weighted groups.
//
// These run as their own FP64 pass beside the exponential families: the SF
// kernel writes the compact groups' G after the main SF kernel, and the
// force kernel adds the compact groups' per-slot contributions
// dE/dG . dG/dr to cslot (and their negative sum to cself) after the main
// backward kernel, before the gather.  One wavefront per atom, neighbours
// staged in LDS in list order with their element.
//
// Each member s of a group has an open radial interval (rl, rc) and, for the
// angular types, an open angle interval (al, ar); its radial function is
// the compact function cr(r) = core(|r - center| / width) (asymmetric cores:
// core((2 - y) y), CoreFunction.h:178-191) and its angular function
// ca(theta) alike with theta = acos(cos theta_ijk).  Per pair the members are
// walked in order as the reference's group loop does, including its reuse
// of cr(r_jk) across members whose (rl, rc) equal the previous member's
// (calculateComp), so the values equal the reference's group path.
#include "device_common.h"
#include "kernel_args.h"
#include "kernels.h"

namespace hdnnp {

namespace {

// CompactFunction::fdf (CompactFunction.h:153-160): a = (x - center) scale,
// f = core(|a|), f' = core'(|a|) copysign(scale, a).  The core polynomials
// and the exponential are the cutoff cores on [0, 1) (a configuration with
// rci = 0, iw = 1 and rc beyond 1 returns exactly core(y), core'(y)).
__device__ __forceinline__ void compact_fdf(int core, int asym, double center,
                                            double scale, double x, double& f,
                                            double& df)
{
    CutCfg c{};
    c.type = core;
    c.rc = 2.0;
    c.rcinv = 0.5;
    c.rci = 0.0;
    c.iw = 1.0;
    double const a = (x - center) * scale;
    double const y = fabs(a);
    if (asym)
    {
        cutoff_fdf<double>(c, (2.0 - y) * y, f, df);
        df *= 2.0 - 2.0 * y;
    }
    else cutoff_fdf<double>(c, y, f, df);
    df *= copysign(scale, a);
}

__device__ __forceinline__ void lds_add_d(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Neighbours of atom i in list order: geometry and element.
__device__ __forceinline__ void stage_list(const double4* __restrict__ geo,
                                           const int* __restrict__ nbrJ,
                                           const int* __restrict__ elem,
                                           int s0, int n, double4* P, int* E)
{
    for (int t = threadIdx.x; t < n; t += 64)
    {
        P[t] = geo[s0 + t];
        E[t] = elem[nbrJ[s0 + t]];
    }
    __syncthreads();
}

// Whether the pair (j, k) belongs to the group's element pair.
__device__ __forceinline__ bool pair_elements(CompGroup const& g, int ej,
                                              int ek)
{
    return sf_weighted(g.type) || (g.e1 == ej && g.e2 == ek)
        || (g.e2 == ej && g.e1 == ek);
}

// What angular_pair accumulates.
enum PairMode
{
    PM_VALUE = 0,     // acc[l] += rad * ang
    PM_FORCE = 1,     // P1, P2, P3 += w[l] (p1, p2, p3)
    PM_TANGENT = 2    // acc[l] += p1 t1 + p2 t2 + p3 t3
};

// One pair {j, k} of an angular compact group, all members in order, with
// p1, p2, p3 of SymGrpCompAngn.cpp:296-343 (dG/dr_j = -(p1 dr_ij + p3
// dr_jk), dG/dr_k = -(p2 dr_ik - p3 dr_jk), dG/dr_i = p1 dr_ij + p2
// dr_ik):
//  - PM_VALUE: the member values;
//  - PM_FORCE: with the member weights w[l] (dE/dG), slot j receives
//    -(P1 dr_ij + P3 dr_jk) and slot k -(P2 dr_ik - P3 dr_jk);
//  - PM_TANGENT: sum_i d_i . dG/dr_i for the residual d, given
//    t1 = dr_ij . (d_i - d_j), t2 = dr_ik . (d_i - d_k),
//    t3 = dr_jk . (d_k - d_j).
// z = Z_j Z_k for the weighted types, else 1.
template <int MODE>
__device__ __forceinline__ void angular_pair(
    CompGroup const& g, CompMember const* __restrict__ mem, double4 const pj,
    double4 const pk, double z, double* acc, double const* w, double& P1,
    double& P2, double& P3)
{
    constexpr bool DERIV = MODE != PM_VALUE;
    bool const wide = (g.type == 22 || g.type == 25);
    bool const weighted = sf_weighted(g.type);
    double const rij = pj.w, rik = pk.w;
    // Group range (rmin = 0, SymGrpBaseComp.cpp:26 with
    // SymGrpCompRad.cpp:82-84).
    if (!(rij < g.rmax && rij > 0.0 && rik < g.rmax && rik > 0.0)) return;
    double const d30 = pk.x - pj.x, d31 = pk.y - pj.y, d32 = pk.z - pj.z;
    double rjk = 0.0, rinvjk = 0.0;
    if (!wide)
    {
        double const rjk2 = d30 * d30 + d31 * d31 + d32 * d32;
        if (rjk2 >= g.rmax * g.rmax || rjk2 <= 0.0) return;
        rjk = sqrt(rjk2);
        rinvjk = 1.0 / rjk;
    }
    double const rinvijik = 1.0 / (rij * rik);
    double const costijk = (pj.x * pk.x + pj.y * pk.y + pj.z * pk.z)
                         * rinvijik;
    if (costijk <= -1.0 || costijk >= 1.0) return;
    double const acostijk = acos(costijk);
    double const rinvij = rinvijik * rik;
    double const rinvik = rinvijik * rij;
    double const phiijik0 = rinvij * (rinvik - rinvij * costijk);
    double const phiikij0 = rinvik * (rinvij - rinvik * costijk);
    double const dacostijk = DERIV ? -1.0 / sqrt(1.0 - costijk * costijk)
                                   : 0.0;
    bool skip = false;
    double radjk = 1.0, dradjk = 0.0;
    for (int l = 0; l < g.nmem; ++l)
    {
        CompMember const m = mem[l];
        double radij = 0.0, dradij = 0.0;
        if (rij > m.rl && rij < m.rc)
            compact_fdf(m.core, m.asym, m.rcenter, m.rscale, rij, radij,
                        dradij);
        if (radij == 0.0 || rik <= m.rl || rik >= m.rc
            || (!wide && (rjk <= m.rl || rjk >= m.rc))
            || acostijk <= m.al || acostijk >= m.ar)
        {
            skip = true;
            continue;
        }
        double radik, dradik;
        compact_fdf(m.core, m.asym, m.rcenter, m.rscale, rik, radik, dradik);
        if (!wide && (m.newrad || skip))
        {
            compact_fdf(m.core, m.asym, m.rcenter, m.rscale, rjk, radjk,
                        dradjk);
            skip = false;
        }
        double ang, dang;
        compact_fdf(m.core, 0, m.acenter, m.ascale, acostijk, ang, dang);
        if (weighted) ang *= z;
        double rad = wide ? radij * radik : radij * radik * radjk;
        if (!DERIV)
        {
            acc[l] += rad * ang;
            continue;
        }
        dang *= dacostijk;
        double const phiijik = phiijik0 * dang;
        double const phiikij = phiikij0 * dang;
        double const psiijik = rinvijik * dang;
        double chiij, chiik, chijk;
        if (!wide)
        {
            chiij = rinvij * dradij * radik * radjk;
            chiik = rinvik * radij * dradik * radjk;
            chijk = -rinvjk * radij * radik * dradjk;
        }
        else
        {
            chiij = rinvij * radik * dradij;
            chiik = rinvik * radij * dradik;
            chijk = 0.0;
        }
        double p1, p2, p3;
        if (dang != 0.0 && rad != 0.0)
        {
            rad *= weighted ? m.sfac * z : m.sfac;
            ang *= m.sfac;
            p1 = rad * phiijik + ang * chiij;
            p2 = rad * phiikij + ang * chiik;
            p3 = rad * psiijik + ang * chijk;
        }
        else if (ang != 0.0)
        {
            ang *= m.sfac;
            p1 = ang * chiij;
            p2 = ang * chiik;
            p3 = ang * chijk;
        }
        else continue;
        if (MODE == PM_FORCE)
        {
            P1 += w[l] * p1;
            P2 += w[l] * p2;
            P3 += w[l] * p3;
        }
        else acc[l] += p1 * P1 + p2 * P2 + p3 * P3;   // P1..P3 hold t1..t3
    }
}

// Per-slot accumulators of the force kernel, rounded up to whole 32-byte
// lines so that the staged double4 records that follow stay aligned.
__host__ __device__ constexpr int acc_doubles(int cap)
{
    return (3 * cap + 3) & ~3;
}

// Per-lane member-sum rows (64 x ms doubles), rounded up to whole 32-byte
// lines so that the staged double4 records that follow stay aligned.
__host__ __device__ constexpr int acc_rows(int ms)
{
    return (64 * ms + 3) & ~3;
}

// LDS: per-lane member sums (64 x ms, ms = largest compact group), then the
// staged neighbours.
__global__ __launch_bounds__(64) void k_sf_compact(SFArgs a)
{
    extern __shared__ __align__(16) double sm[];
    ModelTables const& T = a.T;
    int const i = blockIdx.x;
    ElemDesc const& ed = T.elem[a.elem[i]];
    if (ed.ncg == 0) return;
    int const lane = threadIdx.x;
    int const ms = T.maxCompMembers;
    double* ACC = sm;
    double4* P = (double4*)(ACC + acc_rows(ms));
    int* E = (int*)(P + a.cap);
    int const s0 = a.off[i], n = a.cnt[i];
    stage_list(a.geo, a.nbrJ, a.elem, s0, n, P, E);
    double* row = ACC + lane * ms;
    double* __restrict__ Gi = a.G + (size_t)i * a.ldG;
    for (int cg = ed.cg0; cg < ed.cg0 + ed.ncg; ++cg)
    {
        CompGroup const g = T.cgroup[cg];
        CompMember const* __restrict__ mem = T.cmember + g.mem0;
        bool const weighted = sf_weighted(g.type);
        for (int l = 0; l < g.nmem; ++l) row[l] = 0.0;
        if (sf_radial(g.type))
        {
            for (int t = lane; t < n; t += 64)
            {
                if (!weighted && E[t] != g.e1) continue;
                double const r = P[t].w;
                if (!(r < g.rmax && r > 0.0)) continue;
                double const z = weighted ? T.zw[E[t]] : 1.0;
                for (int l = 0; l < g.nmem; ++l)
                {
                    CompMember const& m = mem[l];
                    if (r <= m.rl || r >= m.rc) continue;
                    double rad, drad;
                    compact_fdf(m.core, m.asym, m.rcenter, m.rscale, r, rad,
                                drad);
                    row[l] += weighted ? z * rad : rad;
                }
            }
        }
        else
        {
            DiagPairs const dp(true, 0, n, 0, n);
            int pc = 0, pr = 0;
            dp.start(lane, pc, pr);
            double d1 = 0.0, d2 = 0.0, d3 = 0.0;
            for (int base = 0; base < dp.npairs; base += 64)
            {
                int j, k;
                dp.pair(pc, pr, j, k);
                dp.advance(pc, pr);
                if (base + lane >= dp.npairs) continue;
                if (!pair_elements(g, E[j], E[k])) continue;
                double const z = weighted ? T.zw[E[j]] * T.zw[E[k]] : 1.0;
                angular_pair<PM_VALUE>(g, mem, P[j], P[k], z, row, nullptr,
                                       d1, d2, d3);
            }
        }
        __syncthreads();
        // Member totals: member l sums the 64 lane rows in lane order.
        for (int l = lane; l < g.nmem; l += 64)
        {
            double v = 0.0;
            for (int q = 0; q < 64; ++q) v += ACC[q * ms + l];
            int const sf = mem[l].sf;
            Gi[sf] = a.raw ? v : scale_sf(T.scale[ed.sf0 + sf], v);
        }
        __syncthreads();
    }
}

// LDS: member weights (ms), per-slot accumulators (3 x cap), staged
// neighbours and their elements.
__global__ __launch_bounds__(64) void k_force_compact(ForceArgs a)
{
    extern __shared__ __align__(16) double sm[];
    ModelTables const& T = a.T;
    int const i = blockIdx.x;
    ElemDesc const& ed = T.elem[a.elem[i]];
    if (ed.ncg == 0) return;
    int const lane = threadIdx.x;
    int const cap = a.cap;
    double* W = sm;
    double* ACC = W + ((T.maxCompMembers + 3) & ~3);
    double4* P = (double4*)(ACC + acc_doubles(cap));
    int* E = (int*)(P + cap);
    int const s0 = a.off[i], n = a.cnt[i];
    for (int t = lane; t < 3 * n; t += 64) ACC[t] = 0.0;
    stage_list(a.geo, a.nbrJ, a.elem, s0, n, P, E);
    double const* __restrict__ dEdGi = a.dEdG + (size_t)i * a.ldG;
    for (int cg = ed.cg0; cg < ed.cg0 + ed.ncg; ++cg)
    {
        CompGroup const g = T.cgroup[cg];
        CompMember const* __restrict__ mem = T.cmember + g.mem0;
        bool const weighted = sf_weighted(g.type);
        for (int l = lane; l < g.nmem; l += 64) W[l] = dEdGi[mem[l].sf];
        __syncthreads();
        if (sf_radial(g.type))
        {
            // SymGrpCompRad.cpp:176-200: p = drad sfac / r (times Z_j for
            // the weighted type); slot j receives -p dr_ij.
            for (int t = lane; t < n; t += 64)
            {
                if (!weighted && E[t] != g.e1) continue;
                double4 const pj = P[t];
                double const r = pj.w;
                if (!(r < g.rmax && r > 0.0)) continue;
                double const z = weighted ? T.zw[E[t]] : 1.0;
                double S = 0.0;
                for (int l = 0; l < g.nmem; ++l)
                {
                    CompMember const& m = mem[l];
                    if (r <= m.rl || r >= m.rc) continue;
                    double rad, drad;
                    compact_fdf(m.core, m.asym, m.rcenter, m.rscale, r, rad,
                                drad);
                    if (drad == 0.0) continue;
                    drad *= weighted ? z * m.sfac / r : m.sfac / r;
                    S += W[l] * drad;
                }
                ACC[3 * t] -= S * pj.x;
                ACC[3 * t + 1] -= S * pj.y;
                ACC[3 * t + 2] -= S * pj.z;
            }
        }
        else
        {
            DiagPairs const dp(true, 0, n, 0, n);
            int pc = 0, pr = 0;
            dp.start(lane, pc, pr);
            for (int base = 0; base < dp.npairs; base += 64)
            {
                int j, k;
                dp.pair(pc, pr, j, k);
                dp.advance(pc, pr);
                if (base + lane >= dp.npairs) continue;
                if (!pair_elements(g, E[j], E[k])) continue;
                double const z = weighted ? T.zw[E[j]] * T.zw[E[k]] : 1.0;
                double4 const pj = P[j], pk = P[k];
                double P1 = 0.0, P2 = 0.0, P3 = 0.0;
                angular_pair<PM_FORCE>(g, mem, pj, pk, z, nullptr, W, P1, P2,
                                       P3);
                if (P1 == 0.0 && P2 == 0.0 && P3 == 0.0) continue;
                double const dx = pk.x - pj.x, dy = pk.y - pj.y,
                             dz = pk.z - pj.z;
                lds_add_d(ACC + 3 * j, -(P1 * pj.x + P3 * dx));
                lds_add_d(ACC + 3 * j + 1, -(P1 * pj.y + P3 * dy));
                lds_add_d(ACC + 3 * j + 2, -(P1 * pj.z + P3 * dz));
                lds_add_d(ACC + 3 * k, P3 * dx - P2 * pk.x);
                lds_add_d(ACC + 3 * k + 1, P3 * dy - P2 * pk.y);
                lds_add_d(ACC + 3 * k + 2, P3 * dz - P2 * pk.z);
            }
        }
        __syncthreads();
    }
    // Add to the main kernel's per-slot results; cself = -(sum of slots).
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int t = lane; t < n; t += 64)
    {
        size_t const q = (size_t)(s0 + t);
        a.cslot[3 * q] += ACC[3 * t];
        a.cslot[3 * q + 1] += ACC[3 * t + 1];
        a.cslot[3 * q + 2] += ACC[3 * t + 2];
        sx += ACC[3 * t];
        sy += ACC[3 * t + 1];
        sz += ACC[3 * t + 2];
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    if (lane == 0)
    {
        a.cself[3 * (size_t)i] -= sx;
        a.cself[3 * (size_t)i + 1] -= sy;
        a.cself[3 * (size_t)i + 2] -= sz;
    }
}

// Training: u_as = cF sum_i d_i . dG_s/dr_i for the compact SFs s of atom
// a (the part of the force-loss tangent k_train_u forms for the other
// types), d = F - Fref.  LDS: per-lane member sums (64 x ms), staged
// neighbours, their residuals d_j and elements.
__global__ __launch_bounds__(64) void k_train_u_compact(TrainArgs A, int cap)
{
    extern __shared__ __align__(16) double sm[];
    ModelTables const& T = A.T;
    int const a = blockIdx.x;
    ElemDesc const& ed = T.elem[A.elem[a]];
    if (ed.ncg == 0) return;
    int const lane = threadIdx.x;
    int const ms = T.maxCompMembers;
    double* ACC = sm;
    double4* P = (double4*)(ACC + acc_rows(ms));
    double* D = (double*)(P + cap);
    int* E = (int*)(D + 3 * cap);
    int const s0 = A.off[a], n = A.cnt[a];
    for (int t = lane; t < n; t += 64)
    {
        int const j = A.nbrJ[s0 + t];
        for (int q = 0; q < 3; ++q)
            D[3 * t + q] = A.F[3 * (size_t)j + q] - A.Fref[3 * (size_t)j + q];
    }
    stage_list(A.geo, A.nbrJ, A.elem, s0, n, P, E);
    double da[3];
    for (int q = 0; q < 3; ++q)
        da[q] = A.F[3 * (size_t)a + q] - A.Fref[3 * (size_t)a + q];
    double* row = ACC + lane * ms;
    double* __restrict__ U = A.U + (size_t)a * A.ldG;
    for (int cg = ed.cg0; cg < ed.cg0 + ed.ncg; ++cg)
    {
        CompGroup const g = T.cgroup[cg];
        CompMember const* __restrict__ mem = T.cmember + g.mem0;
        bool const weighted = sf_weighted(g.type);
        for (int l = 0; l < g.nmem; ++l) row[l] = 0.0;
        if (sf_radial(g.type))
        {
            for (int t = lane; t < n; t += 64)
            {
                if (!weighted && E[t] != g.e1) continue;
                double4 const pj = P[t];
                double const r = pj.w;
                if (!(r < g.rmax && r > 0.0)) continue;
                double const z = weighted ? T.zw[E[t]] : 1.0;
                double const t1 = pj.x * (da[0] - D[3 * t])
                                + pj.y * (da[1] - D[3 * t + 1])
                                + pj.z * (da[2] - D[3 * t + 2]);
                for (int l = 0; l < g.nmem; ++l)
                {
                    CompMember const& m = mem[l];
                    if (r <= m.rl || r >= m.rc) continue;
                    double rad, drad;
                    compact_fdf(m.core, m.asym, m.rcenter, m.rscale, r, rad,
                                drad);
                    if (drad == 0.0) continue;
                    drad *= weighted ? z * m.sfac / r : m.sfac / r;
                    row[l] += drad * t1;
                }
            }
        }
        else
        {
            DiagPairs const dp(true, 0, n, 0, n);
            int pc = 0, pr = 0;
            dp.start(lane, pc, pr);
            for (int base = 0; base < dp.npairs; base += 64)
            {
                int j, k;
                dp.pair(pc, pr, j, k);
                dp.advance(pc, pr);
                if (base + lane >= dp.npairs) continue;
                if (!pair_elements(g, E[j], E[k])) continue;
                double const z = weighted ? T.zw[E[j]] * T.zw[E[k]] : 1.0;
                double4 const pj = P[j], pk = P[k];
                double const* dj = D + 3 * j;
                double const* dk = D + 3 * k;
                double t1 = pj.x * (da[0] - dj[0]) + pj.y * (da[1] - dj[1])
                          + pj.z * (da[2] - dj[2]);
                double t2 = pk.x * (da[0] - dk[0]) + pk.y * (da[1] - dk[1])
                          + pk.z * (da[2] - dk[2]);
                double t3 = (pk.x - pj.x) * (dk[0] - dj[0])
                          + (pk.y - pj.y) * (dk[1] - dj[1])
                          + (pk.z - pj.z) * (dk[2] - dj[2]);
                angular_pair<PM_TANGENT>(g, mem, pj, pk, z, row, nullptr, t1,
                                         t2, t3);
            }
        }
        __syncthreads();
        for (int l = lane; l < g.nmem; l += 64)
        {
            double v = 0.0;
            for (int q = 0; q < 64; ++q) v += ACC[q * ms + l];
            U[mem[l].sf] = A.cF * v;
        }
        __syncthreads();
    }
}

}  // namespace

size_t train_u_compact_lds_bytes(ModelTables const& T, int cap)
{
    return (size_t)acc_rows(T.maxCompMembers) * sizeof(double)
         + (size_t)cap * (sizeof(double4) + 3 * sizeof(double) + sizeof(int));
}

hipError_t launch_train_u_compact(TrainArgs const& A, int cap, hipStream_t s)
{
    if (A.n <= 0 || A.T.ncgroup == 0) return hipSuccess;
    k_train_u_compact<<<A.n, 64, train_u_compact_lds_bytes(A.T, cap), s>>>(
        A, cap);
    return hipGetLastError();
}

size_t sf_compact_lds_bytes(ModelTables const& T, int cap)
{
    return (size_t)acc_rows(T.maxCompMembers) * sizeof(double)
         + (size_t)cap * sizeof(double4)
         + (size_t)cap * sizeof(int);
}

size_t force_compact_lds_bytes(ModelTables const& T, int cap)
{
    return (size_t)((T.maxCompMembers + 3) & ~3) * sizeof(double)
         + (size_t)acc_doubles(cap) * sizeof(double)
         + (size_t)cap * sizeof(double4) + (size_t)cap * sizeof(int);
}

hipError_t launch_sf_compact(SFArgs const& a, hipStream_t s)
{
    if (a.n <= 0 || a.T.ncgroup == 0) return hipSuccess;
    k_sf_compact<<<a.n, 64, sf_compact_lds_bytes(a.T, a.cap), s>>>(a);
    return hipGetLastError();
}

hipError_t launch_force_compact(ForceArgs const& a, hipStream_t s)
{
    if (a.n <= 0 || a.T.ncgroup == 0) return hipSuccess;
    k_force_compact<<<a.n, 64, force_compact_lds_bytes(a.T, a.cap), s>>>(a);
    return hipGetLastError();
}

}  // namespace hdnnp
