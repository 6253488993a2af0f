// Training kernels (the distilled surrogate of the multiple-time-step
// integrator, and 2G models with radial and angular SFs): the gradient of
//
//   L = wE ((E - Eref) / N)^2 + wF / (3N) sum_i |F_i - Fref_i|^2
//
// with respect to every network weight, and an Adam step.
//
// With F_i = -sum_a sum_s (dE_a/dG_as)(dG_as/dr_i) the force term's
// gradient is sum_a d/dw [ sum_s u_as dE_a/dG_as ], where
//
//   u_as = -(2 wF / 3N) sum_i (F_i - Fref_i) . dG_as/dr_i,
//
// the derivative of E_a along the direction u_a: a forward-mode tangent
// through the atom's network, then reverse mode through that tangent.
// For a radial G2 member g(r) = exp(-eta (r - rs)^2) fc(r) of neighbour
// element e1 (SymGrpExpRad.cpp:144-239):
//
//   sum_i d_i . dG_as/dr_i = sfac_s sum_j g'(r_aj) (dr_aj . (d_a - d_j)) / r_aj
//
// with dr_aj = r_a - r_j and sfac_s the SF's scaling slope.  The energy
// term is ordinary backpropagation of kappa = 2 wE (E - Eref) / N^2.
// Everything is FP64 and summed in a fixed order.
//
// No counterpart in the reference's 2G prediction path; the reference
// trains with libnnptrain (Training.cpp, NeuralNetwork::calculateDEdc /
// calculateDFdc, NeuralNetwork.cpp:444-719), which this does not mirror.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device_common.h"
#include "kernels.h"

namespace hdnnp {

namespace {

// Value, first and second derivative of the activations training supports.
__device__ __forceinline__ void activate2(int a, double x, double& v,
                                          double& d, double& dd)
{
    switch (a)
    {
    case AF_IDENTITY: v = x; d = 1.0; dd = 0.0; return;
    case AF_TANH:
    {
        double const t = tanh(x);
        v = t; d = 1.0 - t * t; dd = -2.0 * t * d;
        return;
    }
    case AF_LOGISTIC:
    case AF_SOFTPLUS:
    {
        double s1;
        activate(a, x, v, s1);
        if (x > 35.0 || x < -35.0) { d = s1; dd = 0.0; return; }
        double const s = 1.0 / (1.0 + exp(-x));
        d = s1;
        dd = a == AF_SOFTPLUS ? s * (1.0 - s) : s1 * (1.0 - 2.0 * s);
        return;
    }
    default: v = 0.0; d = 0.0; dd = 0.0; return;
    }
}

__device__ __forceinline__ double scale_slope(ScaleDesc const& s)
{
    return (s.type == ST_SCALE || s.type == ST_SCALECENTER
            || s.type == ST_SCALESIGMA) ? s.sfac : 1.0;
}

// u_as for every SF of atom a, and |F_a - Fref_a|^2.  One wave per atom.
__global__ __launch_bounds__(64) void k_train_u(TrainArgs A)
{
    int const a = blockIdx.x;
    if (a >= A.n) return;
    int const lane = threadIdx.x;
    ModelTables const& T = A.T;
    int const e = A.elem[a];
    ElemDesc const& ed = T.elem[e];
    double da[3];
    for (int k = 0; k < 3; ++k)
        da[k] = A.F[3 * (size_t)a + k] - A.Fref[3 * (size_t)a + k];
    if (lane == 0) A.d2[a] = da[0] * da[0] + da[1] * da[1] + da[2] * da[2];
    double* __restrict__ U = A.U + (size_t)a * A.ldG;
    int const o = A.off[a], m = A.cnt[a];
    for (int gi = ed.g0; gi < ed.g0 + ed.nrad; ++gi)
    {
        GroupDesc const g = T.group[gi];
        CutCfg const cc = T.cfg[ed.cfg0 + g.cfg];
        MemberDesc const* __restrict__ mem = T.member + g.mem0;
        bool const weighted = sf_weighted(g.type);
        double acc[kMaxMembers];
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q) acc[q] = 0.0;
        for (int t = lane; t < m; t += 64)
        {
            int const j = A.nbrJ[o + t];
            if (!weighted && A.elem[j] != g.e1) continue;
            double4 const gq = A.geo[o + t];
            double const r = gq.w;
            if (!(r < g.rc)) continue;
            double fc, dfc;
            cutoff_fdf<double>(cc, r, fc, dfc);
            if (weighted)
            {
                // SymGrpExpRadWeighted: every term times Z_j.
                fc *= T.zw[A.elem[j]];
                dfc *= T.zw[A.elem[j]];
            }
            double proj = 0.0;
            proj += gq.x * (da[0] - (A.F[3 * (size_t)j] - A.Fref[3 * (size_t)j]));
            proj += gq.y * (da[1] - (A.F[3 * (size_t)j + 1]
                                     - A.Fref[3 * (size_t)j + 1]));
            proj += gq.z * (da[2] - (A.F[3 * (size_t)j + 2]
                                     - A.Fref[3 * (size_t)j + 2]));
            proj /= r;
#pragma unroll
            for (int q = 0; q < kMaxMembers; ++q)
                if (q < g.nmem)
                {
                    double const eta = mem[q].eta, x = r - mem[q].rs;
                    double const ex = exp(-eta * x * x);
                    acc[q] += (dfc - 2.0 * eta * x * fc) * ex * proj;
                }
        }
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q)
            if (q < g.nmem)
            {
                double const s = wave_sum(acc[q]);
                int const sf = mem[q].sf;
                if (lane == 0)
                    U[sf] = A.cF * scale_slope(T.scale[ed.sf0 + sf]) * s;
            }
    }
    // Angular groups, G3 (SymFncExpAngn.cpp:79-225), G9
    // (SymFncExpAngw.cpp:79-208) and the weighted G3 (type 13,
    // SymFncExpAngnWeighted.cpp:79-228, every pair of the list):
    // for the pair (j, k), j < k in list order,
    // the reference's p1, p2, p3 give
    //   sum_i d_i . dG/dr_i = p1 dr_ij.(d_a - d_j) + p2 dr_ik.(d_a - d_k)
    //                       + p3 dr_jk.(d_k - d_j),   dr_jk = dr_ik - dr_ij.
    for (int gi = ed.g0 + ed.nrad; gi < ed.g0 + ed.ng; ++gi)
    {
        GroupDesc const g = T.group[gi];
        CutCfg const cc = T.cfg[ed.cfg0 + g.cfg];
        MemberDesc const* __restrict__ mem = T.member + g.mem0;
        bool const wide = g.type == 9;
        bool const weighted = sf_weighted(g.type);
        double acc[kMaxMembers];
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q) acc[q] = 0.0;
        for (int j = 0; j + 1 < m; ++j)
        {
            int const aj = A.nbrJ[o + j];
            int const ej = A.elem[aj];
            double4 const gj = A.geo[o + j];
            double const rij = gj.w;
            if (!((weighted || ej == g.e1 || ej == g.e2) && rij < g.rc))
                continue;
            double fcij, dfcij;
            cutoff_fdf<double>(cc, rij, fcij, dfcij);
            double dj[3];
            dj[0] = A.F[3 * (size_t)aj] - A.Fref[3 * (size_t)aj];
            dj[1] = A.F[3 * (size_t)aj + 1] - A.Fref[3 * (size_t)aj + 1];
            dj[2] = A.F[3 * (size_t)aj + 2] - A.Fref[3 * (size_t)aj + 2];
            for (int k = j + 1 + lane; k < m; k += 64)
            {
                int const ak = A.nbrJ[o + k];
                int const ek = A.elem[ak];
                if (!weighted
                    && !((g.e1 == ej && g.e2 == ek) || (g.e2 == ej && g.e1 == ek)))
                    continue;
                double4 const gk = A.geo[o + k];
                double const rik = gk.w;
                if (!(rik < g.rc)) continue;
                double const jx = gk.x - gj.x, jy = gk.y - gj.y,
                             jz = gk.z - gj.z;
                double rjk = jx * jx + jy * jy + jz * jz;
                if (!wide && !(rjk < g.rc2)) continue;
                rjk = sqrt(rjk);
                double fcik, dfcik, fcjk = 1.0, dfcjk = 0.0;
                cutoff_fdf<double>(cc, rik, fcik, dfcik);
                if (!wide) cutoff_fdf<double>(cc, rjk, fcjk, dfcjk);
                double const dk0 = A.F[3 * (size_t)ak] - A.Fref[3 * (size_t)ak];
                double const dk1 = A.F[3 * (size_t)ak + 1]
                                 - A.Fref[3 * (size_t)ak + 1];
                double const dk2 = A.F[3 * (size_t)ak + 2]
                                 - A.Fref[3 * (size_t)ak + 2];
                double const tij = (gj.x * (da[0] - dj[0]) + gj.y * (da[1] - dj[1])
                                    + gj.z * (da[2] - dj[2]));
                double const tik = (gk.x * (da[0] - dk0) + gk.y * (da[1] - dk1)
                                    + gk.z * (da[2] - dk2));
                double const tjk = (jx * (dk0 - dj[0]) + jy * (dk1 - dj[1])
                                    + jz * (dk2 - dj[2]));
                double const rinv = 1.0 / rij / rik;
                double const cost = (gj.x * gk.x + gj.y * gk.y + gj.z * gk.z)
                                  * rinv;
                // Weighted type 13: the cutoff products times Z_j Z_k
                // (SymGrpExpAngnWeighted.cpp:263-270).
                double const z = weighted ? T.zw[ej] * T.zw[ek] : 1.0;
                double const pfc = z * fcij * fcik * fcjk;
#pragma unroll
                for (int q = 0; q < kMaxMembers; ++q)
                    if (q < g.nmem)
                    {
                        MemberDesc const& md = mem[q];
                        double const rijs = rij - md.rs, riks = rik - md.rs,
                                     rjks = rjk - md.rs;
                        double const ex = wide
                            ? exp(-md.eta * (rijs * rijs + riks * riks))
                            : exp(-md.eta * (rijs * rijs + riks * riks
                                             + rjks * rjks));
                        double const pl = 1.0 + md.lambda * cost;
                        if (pl <= 0.0) continue;
                        double fg = ex * (md.useint ? pow_int(pl, md.zint - 1)
                                                    : pow(pl, md.zeta - 1.0));
                        fg *= pow(2.0, 1.0 - md.zeta);
                        double const pzl = md.zeta * md.lambda;
                        double const ri = rinv * pzl, ct = cost * pzl;
                        double const p2e = 2.0 * md.eta * pl;
                        double const p1 = fg * (pfc * (ri - ct / (rij * rij)
                                                - p2e * rijs / rij)
                                                + z * fcik * fcjk * dfcij * pl
                                                / rij);
                        double const p2 = fg * (pfc * (ri - ct / (rik * rik)
                                                - p2e * riks / rik)
                                                + z * fcij * fcjk * dfcik * pl
                                                / rik);
                        double const p3 = wide
                            ? fg * pfc * ri
                            : fg * (pfc * (ri + p2e * rjks / rjk)
                                    - z * fcij * fcik * dfcjk * pl / rjk);
                        acc[q] += md.sfac * (p1 * tij + p2 * tik + p3 * tjk);
                    }
            }
        }
#pragma unroll
        for (int q = 0; q < kMaxMembers; ++q)
            if (q < g.nmem)
            {
                double const s = wave_sum(acc[q]);
                if (lane == 0) U[mem[q].sf] = A.cF * s;
            }
    }
}

// Per-block gradient of sum over the block's atoms of
// kappa E_a + (derivative of E_a along u_a), atoms of element e, apb per
// block.  LDS per atom: the inputs x, tangents t of every layer and the
// adjoints of z and s of every layer, plus two adjoint scratch vectors.
constexpr int kAtoms = 32;

__global__ __launch_bounds__(64) void k_train_grad(TrainArgs A, int e,
                                                   const int* __restrict__ atoms,
                                                   int na, int perAtom,
                                                   int apb,
                                                   double* __restrict__ part,
                                                   int nw)
{
    extern __shared__ double lds[];
    ModelTables const& T = A.T;
    ElemDesc const& ed = T.elem[e];
    int const L = ed.nlayers - 1;
    int const t = threadIdx.x;
    int const b = blockIdx.x;
    int const nat = min(apb, na - b * apb);
    const double* __restrict__ W = T.weights;
    // LDS layout per atom: for l = 1..L: X[l-1] (w_{l-1}), Tg[l-1]
    // (w_{l-1}), Z[l] (w_l), S[l] (w_l); then Xb, Tb (maxwidth each).
    int lo[kMaxLayers];   // offset of layer l's block within an atom
    {
        int p = 0;
        for (int l = 1; l <= L; ++l)
        {
            lo[l] = p;
            p += 2 * ed.width[l - 1] + 2 * ed.width[l];
        }
        lo[0] = p;        // adjoint scratch
    }
    double const kappa = 2.0 * A.wE * (A.energy[0] - A.Eref)
                       / ((double)A.n * (double)A.n);
    if (t < nat)
    {
        int const a = atoms[b * apb + t];
        double* __restrict__ P = lds + (size_t)t * perAtom;
        // inputs: scaled G and its tangent u
        {
            double* X0 = P + lo[1];
            double* T0 = X0 + ed.width[0];
            for (int j = 0; j < ed.width[0]; ++j)
            {
                X0[j] = A.G[(size_t)a * A.ldG + j];
                T0[j] = A.U[(size_t)a * A.ldG + j];
            }
        }
        // forward with tangent
        for (int l = 1; l <= L; ++l)
        {
            int const nin = ed.width[l - 1], nout = ed.width[l];
            double* X = P + lo[l];
            double* Tg = X + nin;
            double* Z = Tg + nin;
            double* S = Z + nout;
            double* Xn = l < L ? P + lo[l + 1] : nullptr;
            double* Tn = l < L ? Xn + nout : nullptr;
            const double* Wl = W + ed.woff[l];
            const double* bl = W + ed.boff[l];
            for (int k = 0; k < nout; ++k)
            {
                double z = bl[k], s = 0.0;
                for (int j = 0; j < nin; ++j)
                {
                    z += Wl[(size_t)k * nin + j] * X[j];
                    s += Wl[(size_t)k * nin + j] * Tg[j];
                }
                Z[k] = z;
                S[k] = s;
                if (l < L)
                {
                    double v, d, dd;
                    activate2(ed.act[l], z, v, d, dd);
                    Xn[k] = v;
                    Tn[k] = d * s;
                }
            }
        }
        // reverse through value and tangent
        double* Xb = P + lo[0];
        double* Tb = Xb + ed.maxwidth;
        Xb[0] = kappa;
        Tb[0] = 1.0;
        for (int l = L; l >= 1; --l)
        {
            int const nin = ed.width[l - 1], nout = ed.width[l];
            double* X = P + lo[l];
            double* Z = X + 2 * nin;
            double* S = Z + nout;
            const double* Wl = W + ed.woff[l];
            for (int k = 0; k < nout; ++k)
            {
                double v, d, dd;
                activate2(ed.act[l], Z[k], v, d, dd);
                double const zb = Xb[k] * d + Tb[k] * S[k] * dd;
                double const sb = Tb[k] * d;
                Z[k] = zb;       // adjoint of z replaces z
                S[k] = sb;       // adjoint of s replaces s
            }
            if (l > 1)
                for (int j = 0; j < nin; ++j)
                {
                    double xb = 0.0, tb = 0.0;
                    for (int k = 0; k < nout; ++k)
                    {
                        xb += Wl[(size_t)k * nin + j] * Z[k];
                        tb += Wl[(size_t)k * nin + j] * S[k];
                    }
                    Xb[j] = xb;
                    Tb[j] = tb;
                }
        }
    }
    __syncthreads();
    // dW_l[k][j] = sum_atoms zb_k x_j + sb_k t_j;  db_l[k] = sum_atoms zb_k
    int const base = ed.woff[1];
    for (int idx = t; idx < nw; idx += 64)
    {
        int const gidx = base + idx;
        int l = 1;
        while (l < L && gidx >= ed.woff[l + 1]) ++l;
        int const nin = ed.width[l - 1], nout = ed.width[l];
        double acc = 0.0;
        if (gidx < ed.boff[l])
        {
            int const r = gidx - ed.woff[l];
            int const k = r / nin, j = r - k * nin;
            for (int q = 0; q < nat; ++q)
            {
                double const* X = lds + (size_t)q * perAtom + lo[l];
                double const* Z = X + 2 * nin;
                acc += Z[k] * X[j] + Z[nout + k] * X[nin + j];
            }
        }
        else
        {
            int const k = gidx - ed.boff[l];
            for (int q = 0; q < nat; ++q)
                acc += (lds + (size_t)q * perAtom + lo[l])[2 * nin + k];
        }
        part[(size_t)b * nw + idx] = acc;
    }
}

// grad[base + idx] = sum over blocks of part (fixed order).
__global__ void k_train_reduce(const double* __restrict__ part, int nb,
                               int nw, double* __restrict__ grad)
{
    int const idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nw) return;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += part[(size_t)b * nw + idx];
    grad[idx] = acc;
}

// Adam (Kingma & Ba 2015, Algorithm 1) on w with gradient g at step t >= 1.
__global__ void k_adam(int n, double* __restrict__ w,
                       float* __restrict__ w32,
                       const double* __restrict__ g, double* __restrict__ m,
                       double* __restrict__ v, double lr, double b1,
                       double b2, double eps, double c1, double c2)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double const gi = g[i];
    double const mi = b1 * m[i] + (1.0 - b1) * gi;
    double const vi = b2 * v[i] + (1.0 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    double const wi = w[i] - lr * (mi / c1) / (sqrt(vi / c2) + eps);
    w[i] = wi;
    w32[i] = (float)wi;
}

__global__ void k_to_f32(int n, const double* __restrict__ w,
                         float* __restrict__ w32)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w32[i] = (float)w[i];
}

// dE_s/dw for every structure s: per block (atoms of one element and one
// structure, at most kAtoms) the forward pass and backpropagation of each
// atom's network (NeuralNetwork::calculateDEdc, NeuralNetwork.cpp:444-490),
// then per weight the sum over the block's atoms in atom order.  LDS per
// atom: for every layer its input x and the adjoint of z, plus one adjoint
// scratch vector.
__global__ __launch_bounds__(64) void k_energy_wgrad(
    ModelTables T, const double* __restrict__ G, int ldG,
    const int* __restrict__ blkAtoms, const int* __restrict__ blk0,
    const int* __restrict__ blkN, const int* __restrict__ blkElem,
    int perAtom, double* __restrict__ part, int maxP)
{
    extern __shared__ double lds[];
    int const b = blockIdx.x;
    int const t = threadIdx.x;
    int const e = blkElem[b];
    ElemDesc const& ed = T.elem[e];
    int const L = ed.nlayers - 1;
    int const nat = blkN[b];
    const double* __restrict__ W = T.weights;
    int lo[kMaxLayers];
    {
        int p = 0;
        for (int l = 1; l <= L; ++l)
        {
            lo[l] = p;
            p += ed.width[l - 1] + ed.width[l];
        }
        lo[0] = p;
    }
    if (t < nat)
    {
        int const a = blkAtoms[blk0[b] + t];
        double* __restrict__ P = lds + (size_t)t * perAtom;
        for (int j = 0; j < ed.width[0]; ++j)
            P[lo[1] + j] = G[(size_t)a * ldG + j];
        for (int l = 1; l <= L; ++l)
        {
            int const nin = ed.width[l - 1], nout = ed.width[l];
            double* X = P + lo[l];
            double* Z = X + nin;
            double* Xn = l < L ? P + lo[l + 1] : nullptr;
            const double* Wl = W + ed.woff[l];
            const double* bl = W + ed.boff[l];
            for (int k = 0; k < nout; ++k)
            {
                double z = bl[k];
                for (int j = 0; j < nin; ++j)
                    z += Wl[(size_t)k * nin + j] * X[j];
                Z[k] = z;
                if (l < L)
                {
                    double v, d;
                    activate(ed.act[l], z, v, d);
                    Xn[k] = v;
                }
            }
        }
        double* Xb = P + lo[0];
        for (int k = 0; k < ed.width[L]; ++k) Xb[k] = 1.0;
        for (int l = L; l >= 1; --l)
        {
            int const nin = ed.width[l - 1], nout = ed.width[l];
            double* X = P + lo[l];
            double* Z = X + nin;
            const double* Wl = W + ed.woff[l];
            for (int k = 0; k < nout; ++k)
            {
                double v, d;
                activate(ed.act[l], Z[k], v, d);
                Z[k] = Xb[k] * d;        // adjoint of z replaces z
            }
            if (l > 1)
                for (int j = 0; j < nin; ++j)
                {
                    double xb = 0.0;
                    for (int k = 0; k < nout; ++k)
                        xb += Wl[(size_t)k * nin + j] * Z[k];
                    Xb[j] = xb;
                }
        }
    }
    __syncthreads();
    int const base = ed.woff[1];
    int nw = 0;
    for (int l = 1; l <= L; ++l) nw += ed.width[l] * ed.width[l - 1] + ed.width[l];
    for (int idx = t; idx < nw; idx += 64)
    {
        int const gidx = base + idx;
        int l = 1;
        while (l < L && gidx >= ed.woff[l + 1]) ++l;
        int const nin = ed.width[l - 1];
        double acc = 0.0;
        if (gidx < ed.boff[l])
        {
            int const r = gidx - ed.woff[l];
            int const k = r / nin, j = r - k * nin;
            for (int q = 0; q < nat; ++q)
            {
                double const* X = lds + (size_t)q * perAtom + lo[l];
                acc += X[nin + k] * X[j];
            }
        }
        else
        {
            int const k = gidx - ed.boff[l];
            for (int q = 0; q < nat; ++q)
                acc += (lds + (size_t)q * perAtom + lo[l])[nin + k];
        }
        part[(size_t)b * maxP + idx] = acc;
    }
}

// J[s][base_e + idx] = sum over the blocks of (structure s, element e), in
// block order.  One workgroup per (s, e); blocks of (s, e) are
// [grp0[s*ne+e], grp0[s*ne+e+1]).
__global__ void k_energy_wgrad_reduce(const double* __restrict__ part,
                                      int maxP, const int* __restrict__ grp0,
                                      const int* __restrict__ base,
                                      const int* __restrict__ nwe, int ne,
                                      int nwTotal, double* __restrict__ J)
{
    int const se = blockIdx.x;
    int const s = se / ne, e = se - s * ne;
    int const b0 = grp0[se], b1 = grp0[se + 1];
    for (int idx = threadIdx.x; idx < nwe[e]; idx += blockDim.x)
    {
        double acc = 0.0;
        for (int b = b0; b < b1; ++b) acc += part[(size_t)b * maxP + idx];
        J[(size_t)s * nwTotal + base[e] + idx] = acc;
    }
}

// Kalman filter, one scalar observation (KalmanFilter.cpp:126-196 with
// sizeObservation 1): X = P h, one wave per row of the symmetric P.
__global__ __launch_bounds__(64) void k_kf_gemv(int nw,
                                                const double* __restrict__ P,
                                                const double* __restrict__ h,
                                                double* __restrict__ X)
{
    int const i = blockIdx.x;
    if (i >= nw) return;
    double acc = 0.0;
    for (int j = threadIdx.x; j < nw; j += 64)
        acc += P[(size_t)i * nw + j] * h[j];
    acc = wave_sum(acc);
    if (threadIdx.x == 0) X[i] = acc;
}

// a = h . X (one workgroup, fixed order).
__global__ __launch_bounds__(256) void k_kf_dot(int nw,
                                                const double* __restrict__ h,
                                                const double* __restrict__ X,
                                                double* __restrict__ a)
{
    __shared__ double s[256];
    int const t = threadIdx.x;
    double acc = 0.0;
    for (int j = t; j < nw; j += 256) acc += h[j] * X[j];
    s[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) a[0] = s[0];
}

// K = X / (a + 1/eta); P -= K X^T; P += q I; w += K xi.  One workgroup
// per row of P, threads over its columns.
__global__ __launch_bounds__(256) void k_kf_update(
    int nw, double* __restrict__ P, const double* __restrict__ X,
    const double* __restrict__ a, double rinv, double q, double xi,
    double* __restrict__ w, float* __restrict__ w32)
{
    int const i = blockIdx.x;
    if (i >= nw) return;
    double const den = a[0] + rinv;
    double const Ki = X[i] / den;
    double const Xi = X[i];
    double* __restrict__ Pi = P + (size_t)i * nw;
    // (X_i X_j) / den: the same operands for (i, j) and (j, i), so P stays
    // exactly symmetric.
    for (int j = threadIdx.x; j < nw; j += 256)
    {
        double v = Pi[j] - (Xi * X[j]) / den;
        if (j == i) v += q;
        Pi[j] = v;
    }
    if (threadIdx.x == 0)
    {
        double const wi = w[i] + Ki * xi;
        w[i] = wi;
        w32[i] = (float)wi;
    }
}

__global__ void k_kf_init(int nw, double* __restrict__ P, double diag)
{
    size_t const k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)nw * nw) return;
    int const i = (int)(k / nw), j = (int)(k - (size_t)i * nw);
    P[k] = i == j ? diag : 0.0;
}

}  // namespace

hipError_t launch_kf_init(int nw, double* P, double diag, hipStream_t s)
{
    size_t const tot = (size_t)nw * nw;
    if (tot == 0) return hipSuccess;
    k_kf_init<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nw, P, diag);
    return hipGetLastError();
}

hipError_t launch_kf_update(int nw, double* P, const double* h, double* X,
                            double* a, double rinv, double q, double xi,
                            double* w, float* w32, hipStream_t s)
{
    if (nw <= 0) return hipSuccess;
    k_kf_gemv<<<nw, 64, 0, s>>>(nw, P, h, X);
    k_kf_dot<<<1, 256, 0, s>>>(nw, h, X, a);
    k_kf_update<<<nw, 256, 0, s>>>(nw, P, X, a, rinv, q, xi, w, w32);
    return hipGetLastError();
}

hipError_t launch_energy_wgrad(ModelTables const& T, const double* G, int ldG,
                               const int* blkAtoms, const int* blk0,
                               const int* blkN, const int* blkElem, int nblk,
                               int perAtom, double* part, int maxP,
                               const int* grp0, const int* base,
                               const int* nwe, int ne, int nstruct,
                               int nwTotal, double* J, hipStream_t s)
{
    if (nblk > 0)
        k_energy_wgrad<<<nblk, 64, (size_t)perAtom * kAtoms * sizeof(double),
                         s>>>(T, G, ldG, blkAtoms, blk0, blkN, blkElem,
                              perAtom, part, maxP);
    if (nstruct * ne > 0)
        k_energy_wgrad_reduce<<<nstruct * ne, 256, 0, s>>>(
            part, maxP, grp0, base, nwe, ne, nwTotal, J);
    return hipGetLastError();
}

int train_atoms_per_block() { return kAtoms; }

// Atoms per block of the loss-gradient kernel: up to kAtoms, fewer for wide
// networks so that their LDS stays within 64 KiB (0 if one atom does not
// fit).
int train_atoms_per_grad_block(ElemDesc const& ed)
{
    size_t const per = train_lds_per_atom(ed) * sizeof(double);
    return (int)std::min<size_t>(kAtoms, 65536 / per);
}

size_t train_lds_per_atom(ElemDesc const& ed)
{
    size_t p = 0;
    for (int l = 1; l < ed.nlayers; ++l)
        p += 2 * ed.width[l - 1] + 2 * ed.width[l];
    return p + 2 * ed.maxwidth;
}

hipError_t launch_train_u(TrainArgs const& A, hipStream_t s)
{
    if (A.n <= 0) return hipSuccess;
    k_train_u<<<A.n, 64, 0, s>>>(A);
    return hipGetLastError();
}

hipError_t launch_train_grad(TrainArgs const& A, ElemDesc const& ed, int e,
                             const int* atoms, int na, double* part, int nw,
                             double* grad, hipStream_t s)
{
    if (na <= 0)
    {
        k_train_reduce<<<(nw + 255) / 256, 256, 0, s>>>(part, 0, nw, grad);
        return hipGetLastError();
    }
    int const per = (int)train_lds_per_atom(ed);
    int const apb = train_atoms_per_grad_block(ed);
    int const nb = (na + apb - 1) / apb;
    size_t const shm = (size_t)per * apb * sizeof(double);
    k_train_grad<<<nb, 64, shm, s>>>(A, e, atoms, na, per, apb, part, nw);
    k_train_reduce<<<(nw + 255) / 256, 256, 0, s>>>(part, nb, nw, grad);
    return hipGetLastError();
}

hipError_t launch_adam(int n, double* w, float* w32, const double* g,
                       double* m, double* v, double lr, double b1, double b2,
                       double eps, int64_t step, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    double const c1 = 1.0 - std::pow(b1, (double)step);
    double const c2 = 1.0 - std::pow(b2, (double)step);
    k_adam<<<(n + 255) / 256, 256, 0, s>>>(n, w, w32, g, m, v, lr, b1, b2,
                                           eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_weights_to_f32(int n, const double* w, float* w32,
                                 hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_to_f32<<<(n + 255) / 256, 256, 0, s>>>(n, w, w32);
    return hipGetLastError();
}

}  // namespace hdnnp
