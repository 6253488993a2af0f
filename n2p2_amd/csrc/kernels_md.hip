// Velocity-Verlet MD on the device with a Verlet-skin neighbour list.
//
// The list is built with rc + skin by the ordinary neighbour builder (same
// order: j, then the image vector).  Every step filters it down to the
// entries within rc at the current positions, in the same order, so the SF
// and force kernels see exactly the list a fresh build would give.  The
// partner of every kept entry is the kept image of its skin partner: |dr|
// is the same for (i, j, b) and (j, i, -b), so both are kept or both
// dropped; the force gather follows it through the skin list.
//
// No counterpart in libnnp: this is the loop LAMMPS runs around
// pair_style hdnnp (fix nve, pair_hdnnp.cpp:81-114) with
// "neigh_modify every 1 delay 0 check yes" (rebuild when an atom has moved
// more than half the skin since the last build).

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "kernels.h"

namespace hdnnp {

namespace {

constexpr int kBias = 64;   // image components biased into 7 bits

// Left-to-right dr of Structure.cpp:319-322 (the neighbour builder's form).
__device__ __forceinline__ void md_image_dr(MDCell const& C, const double* ri,
                                            const double* rj, int b0, int b1,
                                            int b2, double* dr)
{
#pragma clang fp contract(off)
    for (int c = 0; c < 3; ++c)
    {
        double t = ri[c] - rj[c];
        t = t + (double)b0 * C.box[0][c];
        t = t + (double)b1 * C.box[1][c];
        t = t + (double)b2 * C.box[2][c];
        dr[c] = t;
    }
}

// Vec3D::norm2 as the neighbour builder evaluates it.
__device__ __forceinline__ double md_norm2(const double* r)
{
    double t = r[1] * r[1];
    t = fma(r[0], r[0], t);
    return fma(r[2], r[2], t);
}

// v += h * a(F): a = ftoa F / m, per element.
__global__ void k_md_kick(int n, double* __restrict__ v,
                          const double* __restrict__ F,
                          const int* __restrict__ elem,
                          const double* __restrict__ accel, double h)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double const a = h * accel[elem[i]];
    for (int k = 0; k < 3; ++k) v[3 * i + k] += a * F[3 * i + k];
}

// r += dt v; flags[0] = 1 if an atom has moved more than half the skin
// since the list was built (r0); flags[2] = 1 if a coordinate is not
// finite or beyond 1e7 (a run that has become unstable).
__global__ void k_md_drift(int n, double* __restrict__ r,
                           const double* __restrict__ v,
                           const double* __restrict__ r0, double dt,
                           double half2, int* __restrict__ flags)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d2 = 0.0;
    bool bad = false;
    for (int k = 0; k < 3; ++k)
    {
        double const x = r[3 * i + k] + dt * v[3 * i + k];
        r[3 * i + k] = x;
        double const d = x - r0[3 * i + k];
        d2 += d * d;
        bad = bad || !(fabs(x) <= 1e7);
    }
    if (d2 > half2) flags[0] = 1;
    if (bad) flags[2] = 1;
}

__global__ void k_md_clear_flags(int* __restrict__ flags)
{
    if (threadIdx.x < 3) flags[threadIdx.x] = 0;
}

// Atoms that left the cell are shifted back by whole lattice vectors,
// r -= floor(f) box with f = r invbox; atoms inside keep their bits.
// moved[0] = 1 if any atom was shifted.
__global__ void k_md_wrap(int n, double* __restrict__ r, MDCell C,
                          int* __restrict__ moved)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sh[3];
    bool out = false;
    for (int k = 0; k < 3; ++k)
    {
        double const f = r[3 * i] * C.invbox[0][k]
                       + r[3 * i + 1] * C.invbox[1][k]
                       + r[3 * i + 2] * C.invbox[2][k];
        sh[k] = floor(f);
        out = out || sh[k] != 0.0;
    }
    if (!out) return;
    for (int k = 0; k < 3; ++k)
        r[3 * i + k] -= sh[0] * C.box[0][k] + sh[1] * C.box[1][k]
                      + sh[2] * C.box[2][k];
    moved[0] = 1;
}

// Entries of the skin list within rc at the current positions, in order,
// into the slot range of the same atom (the skin list's range bounds the
// count): neighbour index and geometry.  newIdx[s] = kept slot of skin
// slot s (-1 if dropped); the kept entries' keys are not written (the step
// does not read them; k_md_keys fills them on request).  flags[1] = 1 if
// an atom has more than `cap` kept entries.  One wave per atom.
__global__ __launch_bounds__(64) void k_md_filter(
    int n, MDCell C, double rc2, const double* __restrict__ r,
    const int* __restrict__ off, const int* __restrict__ skinCnt,
    const uint64_t* __restrict__ skinKeys, int* __restrict__ cnt,
    int* __restrict__ nbrJ, double4* __restrict__ geo,
    int* __restrict__ newIdx, int cap, int* __restrict__ flags)
{
    int const i = blockIdx.x;
    if (i >= n) return;
    int const lane = threadIdx.x;
    int const o = off[i], m = skinCnt[i];
    double const ri[3] = {r[3 * (size_t)i], r[3 * (size_t)i + 1],
                          r[3 * (size_t)i + 2]};
    int kept = 0;
    for (int base = 0; base < m; base += 64)
    {
        int const t = base + lane;
        bool keep = false;
        int j = 0;
        double dr[3] = {0.0, 0.0, 0.0};
        if (t < m)
        {
            uint64_t const key = skinKeys[o + t];
            j = (int)(key >> 21);
            int const b0 = (int)((key >> 14) & 127) - kBias;
            int const b1 = (int)((key >> 7) & 127) - kBias;
            int const b2 = (int)(key & 127) - kBias;
            double const rj[3] = {r[3 * (size_t)j], r[3 * (size_t)j + 1],
                                  r[3 * (size_t)j + 2]};
            md_image_dr(C, ri, rj, b0, b1, b2, dr);
            keep = md_norm2(dr) <= rc2;
        }
        uint64_t const mask = __ballot(keep);
        if (t < m)
        {
            int const q = o + kept + lanes_below(mask);
            newIdx[o + t] = keep ? q : -1;
            if (keep)
            {
                nbrJ[q] = j;
                geo[q] = make_double4(dr[0], dr[1], dr[2], sqrt(md_norm2(dr)));
            }
        }
        kept += __popcll(mask);
    }
    if (lane == 0)
    {
        cnt[i] = kept;
        if (kept > cap) flags[1] = 1;
    }
}

// Keys of the kept entries: keys[newIdx[s]] = skinKeys[s].
__global__ void k_md_keys(int n, const int* __restrict__ off,
                          const int* __restrict__ skinCnt,
                          const uint64_t* __restrict__ skinKeys,
                          const int* __restrict__ newIdx,
                          uint64_t* __restrict__ keys)
{
    int const i = blockIdx.x;
    if (i >= n) return;
    int const o = off[i], m = skinCnt[i];
    for (int t = threadIdx.x; t < m; t += blockDim.x)
    {
        int const q = newIdx[o + t];
        if (q >= 0) keys[q] = skinKeys[o + t];
    }
}

// The force gather (k_gather) over a filtered list, with each kept entry's
// partner found through the skin list: skin slot s of atom i is kept as
// newIdx[s]; its skin partner skinPartner[s] is kept as newIdx[sp], whose
// contribution cslot[newIdx[sp]] the atom receives.  Walks the skin slots
// in order, so the kept entries are summed in list order.  A skin entry
// without a partner (-1) is counted in misses[0] and skipped.  mode 0
// skips the atom's own images (reference force mode); the neighbour index
// then comes from the skin key.  Four waves per block, one atom each.
__global__ __launch_bounds__(256) void k_gather_md(
    int n, const int* __restrict__ off, const int* __restrict__ skinCnt,
    const int* __restrict__ newIdx, const int* __restrict__ skinPartner,
    const uint64_t* __restrict__ skinKeys, const double* __restrict__ cslot,
    const double* __restrict__ cself, const int* __restrict__ order,
    int mode, int64_t nslots, double* __restrict__ F,
    int* __restrict__ misses)
{
    int const t = xcd_work_index(blockIdx.x, gridDim.x) * 4
                + (int)(threadIdx.x >> 6);
    int const lane = threadIdx.x & 63;
    if (t >= n) return;
    int const i = order[t];
    int const s0 = off[i], s1 = s0 + skinCnt[i];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int s = s0 + lane; s < s1; s += 64)
    {
        if (newIdx[s] < 0) continue;
        if (mode == 0 && (int)(skinKeys[s] >> 21) == i) continue;
        int const sp = skinPartner[s];
        if (sp < 0)
        {
            atomicAdd(misses, 1);
            continue;
        }
        int const pq = newIdx[sp];
        if (pq >= 0 && pq < nslots)
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

// Per-block sums of m v^2 in a fixed order (kinetic energy, twice, in
// mass * velocity^2 units).
__global__ __launch_bounds__(256) void k_md_mv2(
    int n, const double* __restrict__ v, const int* __restrict__ elem,
    const double* __restrict__ mass, double* __restrict__ part)
{
    __shared__ double s[256];
    int const t = threadIdx.x;
    int const i = blockIdx.x * 256 + t;
    double acc = 0.0;
    if (i < n)
    {
        double const* vi = v + 3 * (size_t)i;
        acc = mass[elem[i]] * (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
    }
    s[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = s[0];
}

// Step record: ekin[k] = ekinScale * sum(part), epot[k] = *energy for
// k = counter[0] (if 0 <= k < cap), then counter[0] += 1.  One block.
__global__ __launch_bounds__(256) void k_md_record(
    int nb, const double* __restrict__ part, const double* __restrict__ energy,
    double ekinScale, int* __restrict__ counter, int capRec,
    double* __restrict__ epot, double* __restrict__ ekin)
{
    __shared__ double s[256];
    int const t = threadIdx.x;
    double acc = 0.0;
    for (int b = t; b < nb; b += 256) acc += part[b];
    s[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0)
    {
        int const k = counter[0];
        if (k >= 0 && k < capRec)
        {
            ekin[k] = ekinScale * s[0];
            epot[k] = energy[0];
        }
        counter[0] = k + 1;
    }
}

// v += h a(Ffull - Fsur): the slow-force kick of r-RESPA.
__global__ void k_md_kick_diff(int n, double* __restrict__ v,
                               const double* __restrict__ Ff,
                               const double* __restrict__ Fs,
                               const int* __restrict__ elem,
                               const double* __restrict__ accel, double h)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double const a = h * accel[elem[i]];
    for (int k = 0; k < 3; ++k)
        v[3 * i + k] += a * (Ff[3 * i + k] - Fs[3 * i + k]);
}

// d2[i] = |F1_i - F2_i|^2.
__global__ void k_md_fdiff2(int n, const double* __restrict__ F1,
                            const double* __restrict__ F2,
                            double* __restrict__ d2)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0.0;
    for (int k = 0; k < 3; ++k)
    {
        double const d = F1[3 * i + k] - F2[3 * i + k];
        acc += d * d;
    }
    d2[i] = acc;
}

}  // namespace

hipError_t launch_md_kick_diff(int n, double* v, const double* Ff,
                               const double* Fs, const int* elem,
                               const double* accel, double h, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_kick_diff<<<(n + 255) / 256, 256, 0, s>>>(n, v, Ff, Fs, elem, accel,
                                                   h);
    return hipGetLastError();
}

hipError_t launch_md_fdiff2(int n, const double* F1, const double* F2,
                            double* d2, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_fdiff2<<<(n + 255) / 256, 256, 0, s>>>(n, F1, F2, d2);
    return hipGetLastError();
}

hipError_t launch_md_kick(int n, double* v, const double* F, const int* elem,
                          const double* accel, double h, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_kick<<<(n + 255) / 256, 256, 0, s>>>(n, v, F, elem, accel, h);
    return hipGetLastError();
}

hipError_t launch_md_drift(int n, double* r, const double* v, const double* r0,
                           double dt, double half2, int* flags, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_drift<<<(n + 255) / 256, 256, 0, s>>>(n, r, v, r0, dt, half2, flags);
    return hipGetLastError();
}

hipError_t launch_md_clear_flags(int* flags, hipStream_t s)
{
    k_md_clear_flags<<<1, 64, 0, s>>>(flags);
    return hipGetLastError();
}

hipError_t launch_md_wrap(int n, double* r, MDCell const& C, int* moved,
                          hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_wrap<<<(n + 255) / 256, 256, 0, s>>>(n, r, C, moved);
    return hipGetLastError();
}

hipError_t launch_md_filter(int n, MDCell const& C, double rc2, const double* r,
                            const int* off, const int* skinCnt,
                            const uint64_t* skinKeys, int* cnt, int* nbrJ,
                            double4* geo, int* newIdx, int cap, int* flags,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_filter<<<n, 64, 0, s>>>(n, C, rc2, r, off, skinCnt, skinKeys, cnt,
                                 nbrJ, geo, newIdx, cap, flags);
    return hipGetLastError();
}

hipError_t launch_md_keys(int n, const int* off, const int* skinCnt,
                          const uint64_t* skinKeys, const int* newIdx,
                          uint64_t* keys, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_md_keys<<<n, 64, 0, s>>>(n, off, skinCnt, skinKeys, newIdx, keys);
    return hipGetLastError();
}

hipError_t launch_gather_md(int n, const int* off, const int* skinCnt,
                            const int* newIdx, const int* skinPartner,
                            const uint64_t* skinKeys, const double* cslot,
                            const double* cself, const int* order, int mode,
                            int64_t nslots, double* F, int* misses,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    k_gather_md<<<(n + 3) / 4, 256, 0, s>>>(n, off, skinCnt, newIdx,
                                            skinPartner, skinKeys, cslot,
                                            cself, order, mode, nslots, F,
                                            misses);
    return hipGetLastError();
}

int md_mv2_blocks(int n) { return (n + 255) / 256; }

hipError_t launch_md_record(int n, const double* v, const int* elem,
                            const double* mass, double* part,
                            const double* energy, double ekinScale,
                            int* counter, int capRec, double* epot,
                            double* ekin, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int const nb = md_mv2_blocks(n);
    k_md_mv2<<<nb, 256, 0, s>>>(n, v, elem, mass, part);
    k_md_record<<<1, 256, 0, s>>>(nb, part, energy, ekinScale, counter,
                                  capRec, epot, ekin);
    return hipGetLastError();
}

}  // namespace hdnnp
