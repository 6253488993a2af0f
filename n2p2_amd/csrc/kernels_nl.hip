// Neighbour-list kernels: cell binning, per-atom candidate counting, sorted
// fill and reverse-slot (partner) lookup.
//
// Output matches Structure::calculateNeighborList(rc, false)
// (Structure.cpp:287-399) bit for bit: for each atom i every (j, image bc)
// with |dr|^2 <= rc^2, where dr = r_i - r_j + bc0*a + bc1*b + bc2*c is
// evaluated left to right and |dr|^2 with the contraction the reference
// build uses for Vec3D::norm2 (fma(z, z, fma(x, x, y*y))), ordered by j,
// then bc0, bc1, bc2 ascending, d = sqrt(|dr|^2).
//
// Candidates come from a cell grid in fractional coordinates whose cells
// are at least rc wide, searched over +-m cells with the periodic shift of
// each wrapped cell; hits are sorted per atom in LDS (bitonic) by a key
// (j << 21 | biased bc), which is exactly the reference loop order.
#include "device_common.h"
#include "kernels.h"

namespace hdnnp {

namespace {

// Left-to-right dr of Structure.cpp:319-322 with integer image vector b.
__device__ __forceinline__ void image_dr(NLParams const& p, double const* ri,
                                         double const* rj, int b0, int b1,
                                         int b2, double* dr)
{
#pragma clang fp contract(off)
    for (int c = 0; c < 3; ++c)
    {
        double t = ri[c] - rj[c];
        t = t + (double)b0 * p.box[0][c];
        t = t + (double)b1 * p.box[1][c];
        t = t + (double)b2 * p.box[2][c];
        dr[c] = t;
    }
}

// Vec3D::norm2 as contracted by GCC -O3 -march=native.
__device__ __forceinline__ double norm2_pinned(double const* r)
{
    double t = r[1] * r[1];
    t = fma(r[0], r[0], t);
    return fma(r[2], r[2], t);
}

__device__ __forceinline__ void cell_coords(NLParams const& p,
                                            double const* r, int* ci)
{
#pragma clang fp contract(off)
    for (int c = 0; c < 3; ++c)
    {
        double f;
        if (p.periodic)
        {
            f = r[0] * p.invbox[0][c];
            f = f + r[1] * p.invbox[1][c];
            f = f + r[2] * p.invbox[2][c];
        }
        else
            f = (r[c] - p.origin[c]) * p.invbox[c][c];
        // Clamped in floating point before the conversion, so that a
        // non-finite coordinate still lands in a valid cell (fmax/fmin
        // return the number when the other operand is NaN).
        double const x = fmin(fmax(floor(f * (double)p.nc[c]), 0.0),
                              (double)(p.nc[c] - 1));
        ci[c] = (int)x;
    }
}

constexpr int kBias = 64;   // image components biased into 7 bits

__device__ __forceinline__ uint64_t make_key(int j, int b0, int b1, int b2)
{
    return ((uint64_t)(uint32_t)j << 21)
         | ((uint64_t)(b0 + kBias) << 14)
         | ((uint64_t)(b1 + kBias) << 7)
         | (uint64_t)(b2 + kBias);
}

__device__ __forceinline__ void split_key(uint64_t k, int& j, int& b0, int& b1,
                                          int& b2)
{
    j = (int)(k >> 21);
    b0 = (int)((k >> 14) & 127) - kBias;
    b1 = (int)((k >> 7) & 127) - kBias;
    b2 = (int)(k & 127) - kBias;
}

__global__ void k_bin(NLParams p, const double* __restrict__ r,
                      int* __restrict__ cellOfAtom, int* __restrict__ rank,
                      int* __restrict__ cellCount)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    int ci[3];
    cell_coords(p, r + 3 * (size_t)i, ci);
    int id = (ci[0] * p.nc[1] + ci[1]) * p.nc[2] + ci[2];
    cellOfAtom[i] = id;
    rank[i] = atomicAdd(&cellCount[id], 1);
}

__global__ void k_scatter(NLParams p, const int* __restrict__ cellOfAtom,
                          const int* __restrict__ rank,
                          const int* __restrict__ cellStart,
                          int* __restrict__ cellAtoms)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.n) return;
    cellAtoms[cellStart[cellOfAtom[i]] + rank[i]] = i;
}

// One wave per atom.  Candidate cells are processed 64 at a time: lane c
// owns candidate cell c (its start, count and image shift); a wave prefix
// sum over the counts flattens the candidates so every lane tests one atom
// per step.  MODE 0 counts hits, MODE 1 collects keys into LDS, sorts them
// and writes the atom's slice of the list.
template <int MODE>
__global__ __launch_bounds__(64) void k_nl(
    NLParams p, const double* __restrict__ r, const int* __restrict__ elem,
    const int* __restrict__ cellOfAtom, const int* __restrict__ cellStart,
    const int* __restrict__ cellCount, const int* __restrict__ cellAtoms,
    int* __restrict__ count, int* __restrict__ stats,
    const int* __restrict__ off, int keyCap, uint64_t* __restrict__ keysOut,
    int* __restrict__ nbrJ, double4* __restrict__ geo)
{
    extern __shared__ uint64_t keys[];
    __shared__ int cStart[64], cCnt[64], cPre[64], cShift[64][3];
    int const i = blockIdx.x;
    int const lane = threadIdx.x;
    if (i >= p.n) return;
    double ri[3] = {r[3 * (size_t)i], r[3 * (size_t)i + 1], r[3 * (size_t)i + 2]};
    int const id = cellOfAtom[i];
    int const ci[3] = {id / (p.nc[1] * p.nc[2]), (id / p.nc[2]) % p.nc[1],
                       id % p.nc[2]};
    int const w0 = 2 * p.m[0] + 1, w1 = 2 * p.m[1] + 1, w2 = 2 * p.m[2] + 1;
    int const ncand = w0 * w1 * w2;
    int hits = 0;       // uniform running count (MODE 1: keys stored)
    for (int c0 = 0; c0 < ncand; c0 += 64)
    {
        int const c = c0 + lane;
        int cnt = 0;
        if (c < ncand)
        {
            int d[3] = {c / (w1 * w2) - p.m[0], (c / w2) % w1 - p.m[1],
                        c % w2 - p.m[2]};
            int cw[3], sh[3];
            bool ok = true;
            for (int k = 0; k < 3; ++k)
            {
                int a = ci[k] + d[k];
                int w = ((a % p.nc[k]) + p.nc[k]) % p.nc[k];
                sh[k] = (a - w) / p.nc[k];
                cw[k] = w;
                if (!p.periodic && sh[k] != 0) ok = false;
            }
            if (ok)
            {
                int cid = (cw[0] * p.nc[1] + cw[1]) * p.nc[2] + cw[2];
                cStart[lane] = cellStart[cid];
                cnt = cellCount[cid];
                // The image of j at r_j + sh*box has bc = -sh.
                cShift[lane][0] = -sh[0];
                cShift[lane][1] = -sh[1];
                cShift[lane][2] = -sh[2];
            }
        }
        cCnt[lane] = cnt;
        // inclusive prefix over lanes
        int pre = cnt;
        for (int o = 1; o < 64; o <<= 1)
        {
            int v = __shfl_up(pre, o, 64);
            if (lane >= o) pre += v;
        }
        cPre[lane] = pre - cnt;   // exclusive
        int const total = __shfl(pre, 63, 64);
        __syncthreads();
        for (int t0 = 0; t0 < total; t0 += 64)
        {
            int const t = t0 + lane;
            bool hit = false;
            int j = 0, b0 = 0, b1 = 0, b2 = 0;
            if (t < total)
            {
                // last cell whose exclusive prefix <= t (and count > 0)
                int lo = 0, hi = 63;
                while (lo < hi)
                {
                    int mid = (lo + hi + 1) >> 1;
                    if (cPre[mid] <= t) lo = mid; else hi = mid - 1;
                }
                while (lo > 0 && (cCnt[lo] == 0 || cPre[lo] + cCnt[lo] <= t))
                    --lo;
                j = cellAtoms[cStart[lo] + (t - cPre[lo])];
                b0 = cShift[lo][0];
                b1 = cShift[lo][1];
                b2 = cShift[lo][2];
                if (!(j == i && b0 == 0 && b1 == 0 && b2 == 0))
                {
                    double rj[3] = {r[3 * (size_t)j], r[3 * (size_t)j + 1],
                                    r[3 * (size_t)j + 2]};
                    double dr[3];
                    image_dr(p, ri, rj, b0, b1, b2, dr);
                    hit = norm2_pinned(dr) <= p.rc2;
                }
            }
            uint64_t const mask = __ballot(hit);
            if (MODE == 1 && hit)
            {
                int const pos = hits + lanes_below(mask);
                if (pos < keyCap) keys[pos] = make_key(j, b0, b1, b2);
            }
            hits += __popcll(mask);
        }
        __syncthreads();
    }
    if (MODE == 0)
    {
        if (lane == 0) count[i] = hits;
        return;
    }
    // MODE 1: bitonic sort of keys[0, hits) padded to a power of two.
    if (hits > keyCap)
    {
        if (lane == 0) atomicMax(&stats[1], hits);
        return;
    }
    int P = 1;
    while (P < hits) P <<= 1;
    for (int t = hits + lane; t < P; t += 64) keys[t] = ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
    {
        for (int jj = k >> 1; jj > 0; jj >>= 1)
        {
            for (int t = lane; t < P; t += 64)
            {
                int const q = t ^ jj;
                if (q > t)
                {
                    uint64_t a = keys[t], b = keys[q];
                    bool const up = (t & k) == 0;
                    if ((a > b) == up)
                    {
                        keys[t] = b;
                        keys[q] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    int const o = off[i];
    for (int t = lane; t < hits; t += 64)
    {
        uint64_t const key = keys[t];
        int j, b0, b1, b2;
        split_key(key, j, b0, b1, b2);
        double rj[3] = {r[3 * (size_t)j], r[3 * (size_t)j + 1],
                        r[3 * (size_t)j + 2]};
        double dr[3];
        image_dr(p, ri, rj, b0, b1, b2, dr);
        double const d = sqrt(norm2_pinned(dr));
        keysOut[o + t] = key;
        nbrJ[o + t] = j;
        geo[o + t] = make_double4(dr[0], dr[1], dr[2], d);
    }
    (void)elem;
}

// Cell kernels: one workgroup of four waves per cell.  The atoms of the
// cell's (2m+1)^3 neighbour cells (m = 1 when cells are at least rc wide)
// are staged once in LDS with their image vectors; each wave then takes one
// atom of the cell at a time and tests all staged candidates, 64 per step.
// MODE 0 counts hits.  MODE 1 first sorts the staged candidates by
// (j, image), the reference's loop order, so that the ballot compaction of
// an atom's hits is already in list order and they are written directly.
// The host uses these kernels when at most 64 neighbour cells are searched
// and falls back to k_nl when a cell has more than kCellCand candidates
// (stats[2]).  MODE 1 writes atom i's entries from off[i] on, at most
// limit of them (a fixed slot range per atom), stores the true count in
// count[i] and sets stats[3] when a count exceeds limit (the host then
// rebuilds with a larger range); the largest count comes from a reduction
// over count[] afterwards (launch_max_int).
constexpr int kCellCand = 1024;
constexpr int kCellWaves = 4;

template <int MODE>
__global__ __launch_bounds__(256) void k_nl_cell(
    NLParams p, const double* __restrict__ r,
    const int* __restrict__ cellStart, const int* __restrict__ cellCount,
    const int* __restrict__ cellAtoms, int* __restrict__ count,
    int* __restrict__ stats, const int* __restrict__ off, int limit,
    uint64_t* __restrict__ keysOut, int* __restrict__ nbrJ,
    double4* __restrict__ geo)
{
    __shared__ double cX[kCellCand], cY[kCellCand], cZ[kCellCand];
    // MODE 0: (j << 21 | image).  MODE 1: (j << 31 | image << 10 | t), t the
    // staging slot, sorted.
    __shared__ uint64_t cKey[kCellCand];
    __shared__ int sStart[64], sCnt[64], sPre[64], sImg[64];
    __shared__ int sTotal;
    // Cells in XCD-aware order: each XCD takes a contiguous range of cells,
    // whose candidate regions overlap in its L2.
    int const cell = xcd_work_index(blockIdx.x, gridDim.x);
    int const tid = threadIdx.x;
    int const lane = tid & 63;
    int const w = tid >> 6;
    int const nAt = cellCount[cell];
    if (nAt == 0) return;
    int const ci[3] = {cell / (p.nc[1] * p.nc[2]), (cell / p.nc[2]) % p.nc[1],
                       cell % p.nc[2]};
    int const w0 = 2 * p.m[0] + 1, w1 = 2 * p.m[1] + 1, w2 = 2 * p.m[2] + 1;
    int const ncand = w0 * w1 * w2;
    int const zeroImg = (kBias << 14) | (kBias << 7) | kBias;
    if (tid < 64)
    {
        int cnt = 0, st = 0, img = zeroImg;
        if (tid < ncand)
        {
            int const d[3] = {tid / (w1 * w2) - p.m[0], (tid / w2) % w1 - p.m[1],
                              tid % w2 - p.m[2]};
            int cw[3], sh[3];
            bool ok = true;
            for (int k = 0; k < 3; ++k)
            {
                int const a = ci[k] + d[k];
                int const wr = ((a % p.nc[k]) + p.nc[k]) % p.nc[k];
                sh[k] = (a - wr) / p.nc[k];
                cw[k] = wr;
                if (!p.periodic && sh[k] != 0) ok = false;
            }
            if (ok)
            {
                int const cid = (cw[0] * p.nc[1] + cw[1]) * p.nc[2] + cw[2];
                st = cellStart[cid];
                cnt = cellCount[cid];
                // The image of j at r_j + sh*box has bc = -sh.
                img = ((kBias - sh[0]) << 14) | ((kBias - sh[1]) << 7)
                    | (kBias - sh[2]);
            }
        }
        int pre = cnt;
        for (int o = 1; o < 64; o <<= 1)
        {
            int const v = __shfl_up(pre, o, 64);
            if (lane >= o) pre += v;
        }
        sStart[tid] = st;
        sCnt[tid] = cnt;
        sPre[tid] = pre - cnt;
        sImg[tid] = img;
        if (tid == 63) sTotal = pre;
    }
    __syncthreads();
    int const total = sTotal;
    if (total > kCellCand)
    {
        if (tid == 0) atomicMax(&stats[2], total);
        return;
    }
    int P = 1;
    while (P < total) P <<= 1;
    for (int t = tid; t < (MODE == 1 ? P : total); t += 256)
    {
        if (t >= total)
        {
            cKey[t] = ~0ull;
            continue;
        }
        // last candidate cell whose exclusive prefix <= t
        int lo = 0, hi = 63;
        while (lo < hi)
        {
            int const mid = (lo + hi + 1) >> 1;
            if (sPre[mid] <= t) lo = mid; else hi = mid - 1;
        }
        while (lo > 0 && (sCnt[lo] == 0 || sPre[lo] + sCnt[lo] <= t)) --lo;
        int const j = cellAtoms[sStart[lo] + (t - sPre[lo])];
        cX[t] = r[3 * (size_t)j];
        cY[t] = r[3 * (size_t)j + 1];
        cZ[t] = r[3 * (size_t)j + 2];
        uint64_t const img = (uint64_t)sImg[lo];
        cKey[t] = MODE == 1
                ? ((uint64_t)(uint32_t)j << 31) | (img << 10) | (uint64_t)t
                : ((uint64_t)(uint32_t)j << 21) | img;
    }
    __syncthreads();
    if (MODE == 1)
    {
        // bitonic sort of cKey[0, P)
        for (int k = 2; k <= P; k <<= 1)
        {
            for (int jj = k >> 1; jj > 0; jj >>= 1)
            {
                for (int t = tid; t < P; t += 256)
                {
                    int const q = t ^ jj;
                    if (q > t)
                    {
                        uint64_t const x = cKey[t], y = cKey[q];
                        bool const up = (t & k) == 0;
                        if ((x > y) == up)
                        {
                            cKey[t] = y;
                            cKey[q] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    if (MODE == 1)
    {
        // Candidate coordinates into sorted order (slot u holds the
        // candidate of the u-th key), so that the scan below reads them
        // with consecutive, conflict-free LDS accesses.
        constexpr int kPer = kCellCand / 256;
        double px[kPer], py[kPer], pz[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q)
        {
            int const u = tid + 256 * q;
            if (u < total)
            {
                int const t = (int)(cKey[u] & 1023);
                px[q] = cX[t];
                py[q] = cY[t];
                pz[q] = cZ[t];
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < kPer; ++q)
        {
            int const u = tid + 256 * q;
            if (u < total)
            {
                cX[u] = px[q];
                cY[u] = py[q];
                cZ[u] = pz[q];
            }
        }
    }
    __syncthreads();
    int const a0 = cellStart[cell];
    for (int a = w; a < nAt; a += kCellWaves)
    {
        int const i = cellAtoms[a0 + a];
        double const ri[3] = {r[3 * (size_t)i], r[3 * (size_t)i + 1],
                              r[3 * (size_t)i + 2]};
        int const o = MODE == 1 ? off[i] : 0;
        int hits = 0;
        for (int u0 = 0; u0 < total; u0 += 64)
        {
            int const u = u0 + lane;
            bool hit = false;
            int j = 0, img = 0, t = 0;
            double dr[3] = {0.0, 0.0, 0.0};
            if (u < total)
            {
                uint64_t const ck = cKey[u];
                if (MODE == 1)
                {
                    j = (int)(ck >> 31);
                    img = (int)((ck >> 10) & 0x1fffff);
                    t = u;   // coordinates are in sorted order
                }
                else
                {
                    j = (int)(ck >> 21);
                    img = (int)(ck & 0x1fffff);
                    t = u;
                }
                if (!(j == i && img == zeroImg))
                {
                    int const b0 = ((img >> 14) & 127) - kBias;
                    int const b1 = ((img >> 7) & 127) - kBias;
                    int const b2 = (img & 127) - kBias;
                    double const rj[3] = {cX[t], cY[t], cZ[t]};
                    image_dr(p, ri, rj, b0, b1, b2, dr);
                    hit = norm2_pinned(dr) <= p.rc2;
                }
            }
            uint64_t const mask = __ballot(hit);
            int const h = hits + lanes_below(mask);
            if (MODE == 1 && hit && h < limit)
            {
                int const q = o + h;
                keysOut[q] = ((uint64_t)(uint32_t)j << 21) | (uint64_t)img;
                nbrJ[q] = j;
                geo[q] = make_double4(dr[0], dr[1], dr[2],
                                      sqrt(norm2_pinned(dr)));
            }
            hits += __popcll(mask);
        }
        if (lane == 0)
        {
            count[i] = hits;
            if (MODE == 1 && hits > limit) stats[3] = 1;
        }
    }
}

// Fractional extent of each periodic structure's atoms along its three
// lattice directions (f = r invbox): the bounds the image pruning of
// k_nl_batch uses.  One wave per structure.
__global__ __launch_bounds__(64) void k_batch_extent(
    int nstruct, const double* __restrict__ r,
    const int* __restrict__ structOff, BatchCell* __restrict__ cells)
{
    int const st = blockIdx.x;
    if (st >= nstruct) return;
    BatchCell& C = cells[st];
    if (!C.periodic) return;
    int const lane = threadIdx.x;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = structOff[st] + lane; i < structOff[st + 1]; i += 64)
    {
        double const x = r[3 * (size_t)i], y = r[3 * (size_t)i + 1],
                     z = r[3 * (size_t)i + 2];
        for (int k = 0; k < 3; ++k)
        {
            double const f = x * C.invbox[0][k] + y * C.invbox[1][k]
                           + z * C.invbox[2][k];
            lo[k] = fmin(lo[k], f);
            hi[k] = fmax(hi[k], f);
        }
    }
    for (int k = 0; k < 3; ++k)
    {
        for (int o = 32; o > 0; o >>= 1)
        {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o));
        }
        if (lane == 0)
        {
            C.fmin[k] = lo[k];
            C.fmax[k] = hi[k];
        }
    }
}

// Image range [lo, hi] along one lattice direction that can hold a
// neighbour of an atom at fractional coordinate f.  The image b holds the
// neighbours' copies at fractional [fmin - b, fmax - b]; its distance from
// the atom along the plane normal, (gap in f) * proj, bounds |dr| from
// below.  An image is kept unless that bound exceeds rc by a relative
// margin far above rounding, so the pruned search finds every entry.
__device__ __forceinline__ void image_range(double f, double fmin,
                                            double fmax, double proj,
                                            double rcm, int p, int& lo,
                                            int& hi)
{
    lo = p + 1;
    hi = -p - 1;
    for (int b = -p; b <= p; ++b)
    {
        double gap = 0.0;
        if (f < fmin - b) gap = (fmin - b - f) * proj;
        else if (f > fmax - b) gap = (f - fmax + b) * proj;
        if (gap <= rcm)
        {
            lo = min(lo, b);
            hi = max(hi, b);
        }
    }
}

// Batched structures: atom i of structure s = structOf[i] searches the
// atoms [structOff[s], structOff[s+1]) of its own structure over the
// structure's periodic images |b_k| <= pbc_k (Structure.cpp:287-399); its
// list holds the hits in the reference's loop order (j, then b0, b1, b2).
// Only the images image_range keeps are enumerated, which drops no entry
// and keeps the order.  MODE 0 counts into count[i].  MODE 1 writes at most
// `limit` entries from off[i], stores the count and sets stats[3] when it
// exceeds the limit.  One wave per atom.
//
// At most 64 images (every structure at least about 2 rc wide, and most
// others): lanes take 64 atoms j at a time and walk the images in a
// uniform loop, so each lane computes r_i - r_j once and adds the images'
// shifts to it in image_dr's order and rounding; a lane's hits are kept as
// a bit mask over the images, and an exclusive scan of the lanes' counts
// places each lane's hits, in image order, behind those of the lanes
// (atoms j) before it.  More images: candidates (j, image) in list order,
// 64 per step, and a ballot compaction.
template <int MODE>
__global__ __launch_bounds__(64) void k_nl_batch(
    int n, double rc2, double rcm, const double* __restrict__ r,
    const int* __restrict__ structOf, const int* __restrict__ structOff,
    const BatchCell* __restrict__ cells, int* __restrict__ count,
    int* __restrict__ stats, const int* __restrict__ off, int limit,
    uint64_t* __restrict__ keysOut, int* __restrict__ nbrJ,
    double4* __restrict__ geo)
{
    int const i = blockIdx.x;
    if (i >= n) return;
    int const lane = threadIdx.x;
    int const st = structOf[i];
    BatchCell const& C = cells[st];
    int const a0 = structOff[st], na = structOff[st + 1] - a0;
    NLParams p{};
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) p.box[k][c] = C.box[k][c];
    double const ri[3] = {r[3 * (size_t)i], r[3 * (size_t)i + 1],
                          r[3 * (size_t)i + 2]};
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (C.periodic)
        for (int k = 0; k < 3; ++k)
        {
            double const f = ri[0] * C.invbox[0][k] + ri[1] * C.invbox[1][k]
                           + ri[2] * C.invbox[2][k];
            image_range(f, C.fmin[k], C.fmax[k], C.proj[k], rcm, C.pbc[k],
                        lo[k], hi[k]);
        }
    // The atom's own home image is always within range (gap 0), so every
    // range is non-empty.
    int const w1 = hi[1] - lo[1] + 1, w2 = hi[2] - lo[2] + 1;
    int const w12 = w1 * w2;
    int const nimg = (hi[0] - lo[0] + 1) * w12;
    float const invImg = 1.0f / (float)nimg, invW12 = 1.0f / (float)w12,
                invW2 = 1.0f / (float)w2;
    int const total = na * nimg;
    int const o = MODE == 1 ? off[i] : 0;
    int hits = 0;
    if (nimg <= 64)
    {
        for (int jb = 0; jb < na; jb += 64)
        {
            int const jj = jb + lane;
            bool const live = jj < na;
            int const j = a0 + (live ? jj : 0);
            double const rj[3] = {r[3 * (size_t)j], r[3 * (size_t)j + 1],
                                  r[3 * (size_t)j + 2]};
            uint64_t m = 0;
            {
#pragma clang fp contract(off)
                double const d0 = ri[0] - rj[0], d1 = ri[1] - rj[1],
                             d2 = ri[2] - rj[2];
                int bi = 0;
                for (int b0 = lo[0]; b0 <= hi[0]; ++b0)
                    for (int b1 = lo[1]; b1 <= hi[1]; ++b1)
                        for (int b2 = lo[2]; b2 <= hi[2]; ++b2, ++bi)
                        {
                            double const f0 = (double)b0, f1 = (double)b1,
                                         f2 = (double)b2;
                            double dr[3];
                            dr[0] = d0 + f0 * p.box[0][0];
                            dr[0] = dr[0] + f1 * p.box[1][0];
                            dr[0] = dr[0] + f2 * p.box[2][0];
                            dr[1] = d1 + f0 * p.box[0][1];
                            dr[1] = dr[1] + f1 * p.box[1][1];
                            dr[1] = dr[1] + f2 * p.box[2][1];
                            dr[2] = d2 + f0 * p.box[0][2];
                            dr[2] = dr[2] + f1 * p.box[1][2];
                            dr[2] = dr[2] + f2 * p.box[2][2];
                            bool const self = j == i && b0 == 0 && b1 == 0
                                           && b2 == 0;
                            bool const hit = live && !self
                                          && norm2_pinned(dr) <= rc2;
                            m |= (uint64_t)hit << bi;
                        }
            }
            int const c = __popcll(m);
            int inc = c;
            for (int s = 1; s < 64; s <<= 1)
            {
                int const v = __shfl_up(inc, s, 64);
                if (lane >= s) inc += v;
            }
            if (MODE == 1)
            {
                int h = hits + inc - c;
                while (m)
                {
                    int const bi = __ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    if (h < limit)
                    {
                        int q0, rem, q1, q2;
                        divmod_small(bi, w12, invW12, q0, rem);
                        divmod_small(rem, w2, invW2, q1, q2);
                        int const b0 = lo[0] + q0, b1 = lo[1] + q1,
                                  b2 = lo[2] + q2;
                        double dr[3];
                        image_dr(p, ri, rj, b0, b1, b2, dr);
                        int const q = o + h;
                        keysOut[q] = make_key(j, b0, b1, b2);
                        nbrJ[q] = j;
                        geo[q] = make_double4(dr[0], dr[1], dr[2],
                                              sqrt(norm2_pinned(dr)));
                    }
                    ++h;
                }
            }
            hits += __shfl(inc, 63, 64);
        }
    }
    else
    {
        for (int base = 0; base < total; base += 64)
        {
            int const t = base + lane;
            bool hit = false;
            int j = 0, b0 = 0, b1 = 0, b2 = 0;
            double dr[3] = {0.0, 0.0, 0.0};
            if (t < total)
            {
                int jj, bi, q0, r0, q1, r2;
                divmod_small(t, nimg, invImg, jj, bi);
                divmod_small(bi, w12, invW12, q0, r0);
                divmod_small(r0, w2, invW2, q1, r2);
                j = a0 + jj;
                b0 = lo[0] + q0;
                b1 = lo[1] + q1;
                b2 = lo[2] + r2;
                if (!(j == i && b0 == 0 && b1 == 0 && b2 == 0))
                {
                    double const rj[3] = {r[3 * (size_t)j], r[3 * (size_t)j + 1],
                                          r[3 * (size_t)j + 2]};
                    image_dr(p, ri, rj, b0, b1, b2, dr);
                    hit = norm2_pinned(dr) <= rc2;
                }
            }
            uint64_t const mask = __ballot(hit);
            if (MODE == 1 && hit)
            {
                int const h = hits + lanes_below(mask);
                if (h < limit)
                {
                    int const q = o + h;
                    keysOut[q] = make_key(j, b0, b1, b2);
                    nbrJ[q] = j;
                    geo[q] = make_double4(dr[0], dr[1], dr[2],
                                          sqrt(norm2_pinned(dr)));
                }
            }
            hits += __popcll(mask);
        }
    }
    // The largest count comes from a reduction over count[] afterwards: one
    // atomic per atom on a single word would serialise in L2.
    if (lane == 0)
    {
        count[i] = hits;
        if (MODE == 1 && hits > limit) stats[3] = 1;
    }
}

// off[i] = i * slab: atom i's entries start at a fixed slot range.
__global__ void k_slab_starts(int n, int slab, int* __restrict__ off)
{
    int const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) off[i] = i * slab;
}

// Partner slots: for entry q of atom i pointing at (j, bc) the entry p of
// j pointing at (i, -bc).  Each pair is searched once, from its canonical
// side, the entry whose key exceeds its partner's (j > i, or j == i with a
// positive image), by binary search in j's sorted keys; both slots are
// written, partner[q] = p and partner[p] = q.  The canonical entries of
// atom i are the tail of its list from the first key with j >= i, so the
// lanes work on that range only.  Slots left unwritten (an inconsistent
// list) keep the -1 the launcher fills in.  One wave per atom, atoms in
// cell order (order = atoms sorted by cell) with the XCD-aware mapping: the
// lists searched by neighbouring atoms overlap and stay in one L2.
__global__ __launch_bounds__(64) void k_partner(
    int n, const int* __restrict__ off, const int* __restrict__ cnt,
    const uint64_t* __restrict__ keys, const int* __restrict__ order,
    int* __restrict__ partner)
{
    int const w = xcd_work_index(blockIdx.x, gridDim.x);
    if (w >= n) return;
    int const i = order[w];
    int const s0 = off[i], s1 = s0 + cnt[i];
    uint64_t const first = (uint64_t)(uint32_t)i << 21;
    int a = s0, b = s1;
    while (a < b)
    {
        int const mid = (a + b) >> 1;
        if (keys[mid] < first) a = mid + 1; else b = mid;
    }
    for (int q = a + (int)threadIdx.x; q < s1; q += 64)
    {
        uint64_t const key = keys[q];
        int j, b0, b1, b2;
        split_key(key, j, b0, b1, b2);
        uint64_t const want = make_key(i, -b0, -b1, -b2);
        if (!(key > want)) continue;
        int lo = off[j], hi = off[j] + cnt[j] - 1;
        int found = -1;
        while (lo <= hi)
        {
            int mid = (lo + hi) >> 1;
            uint64_t const v = keys[mid];
            if (v == want) { found = mid; break; }
            if (v < want) lo = mid + 1; else hi = mid - 1;
        }
        partner[q] = found;
        if (found >= 0) partner[found] = q;
    }
}

// out[0] = max(in[0..n)) of non-negative counts; one block, fixed-order
// tree (the per-atom neighbour counts, whose maximum sizes the fill's
// LDS).  `in` is 16-byte aligned (a device allocation): four counts per
// load, four loads in flight per thread.
__global__ __launch_bounds__(1024) void k_max(const int* __restrict__ in,
                                              int n, int* __restrict__ out)
{
    __shared__ int s[1024];
    int const t = threadIdx.x;
    int m = 0;
    int const n4 = n >> 2;
    const int4* __restrict__ in4 = reinterpret_cast<const int4*>(in);
#pragma unroll 4
    for (int q = t; q < n4; q += 1024)
    {
        int4 const v = in4[q];
        m = max(m, max(max(v.x, v.y), max(v.z, v.w)));
    }
    for (int q = 4 * n4 + t; q < n; q += 1024) m = max(m, in[q]);
    s[t] = m;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1)
    {
        if (t < o) s[t] = max(s[t], s[t + o]);
        __syncthreads();
    }
    if (t == 0) out[0] = s[0];
}

// Exclusive scan helpers (block-level, then block offsets).
__global__ void k_scan_block(const int* __restrict__ in, int* __restrict__ out,
                             int* __restrict__ blockSums, int n)
{
    __shared__ int s[1024];
    int const t = threadIdx.x;
    int const g = blockIdx.x * 1024 + t;
    int v = g < n ? in[g] : 0;
    s[t] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1)
    {
        int a = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += a;
        __syncthreads();
    }
    if (g < n) out[g] = s[t] - v;
    if (t == 1023) blockSums[blockIdx.x] = s[t];
}

__global__ void k_scan_sums(int* __restrict__ blockSums, int nb,
                            int* __restrict__ total)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int acc = 0;
    for (int b = 0; b < nb; ++b)
    {
        int v = blockSums[b];
        blockSums[b] = acc;
        acc += v;
    }
    *total = acc;
}

__global__ void k_scan_add(int* __restrict__ out, const int* __restrict__ blockSums,
                           int n, const int* __restrict__ total, int writeTotal)
{
    int const g = blockIdx.x * 1024 + threadIdx.x;
    if (g < n) out[g] += blockSums[blockIdx.x];
    if (writeTotal && g == 0) out[n] = *total;
}

}  // namespace

hipError_t scan_exclusive(const int* in, int* out, int n, int* blockSums,
                          int* total, hipStream_t s)
{
    int const nb = (n + 1023) / 1024;
    if (nb > 0)
    {
        k_scan_block<<<nb, 1024, 0, s>>>(in, out, blockSums, n);
        k_scan_sums<<<1, 64, 0, s>>>(blockSums, nb, total);
        k_scan_add<<<nb, 1024, 0, s>>>(out, blockSums, n, total, 1);
    }
    else
    {
        k_scan_sums<<<1, 64, 0, s>>>(blockSums, 0, total);
        k_scan_add<<<1, 1024, 0, s>>>(out, blockSums, 0, total, 1);
    }
    return hipGetLastError();
}

hipError_t launch_max_int(const int* in, int n, int* out, hipStream_t s)
{
    k_max<<<1, 1024, 0, s>>>(in, n, out);
    return hipGetLastError();
}

hipError_t launch_nl_bin(NLParams const& p, const double* r, int* cellOfAtom,
                         int* rank, int* cellCount, hipStream_t s)
{
    k_bin<<<(p.n + 255) / 256, 256, 0, s>>>(p, r, cellOfAtom, rank, cellCount);
    return hipGetLastError();
}

hipError_t launch_nl_scatter(NLParams const& p, const int* cellOfAtom,
                             const int* rank, const int* cellStart,
                             int* cellAtoms, hipStream_t s)
{
    k_scatter<<<(p.n + 255) / 256, 256, 0, s>>>(p, cellOfAtom, rank, cellStart,
                                                cellAtoms);
    return hipGetLastError();
}

hipError_t launch_nl_count(NLParams const& p, const double* r, const int* elem,
                           const int* cellOfAtom, const int* cellStart,
                           const int* cellCount, const int* cellAtoms,
                           int* count, int* stats, hipStream_t s)
{
    k_nl<0><<<p.n, 64, 0, s>>>(p, r, elem, cellOfAtom, cellStart, cellCount,
                               cellAtoms, count, stats, nullptr, 0, nullptr,
                               nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_nl_fill(NLParams const& p, const double* r, const int* elem,
                          const int* cellOfAtom, const int* cellStart,
                          const int* cellCount, const int* cellAtoms,
                          int* stats, const int* off, int keyCap,
                          uint64_t* keys, int* nbrJ, double4* geo,
                          hipStream_t s)
{
    size_t const shm = (size_t)keyCap * sizeof(uint64_t);
    k_nl<1><<<p.n, 64, shm, s>>>(p, r, elem, cellOfAtom, cellStart, cellCount,
                                 cellAtoms, nullptr, stats, off, keyCap, keys,
                                 nbrJ, geo);
    return hipGetLastError();
}

hipError_t launch_nl_count_cell(NLParams const& p, int ncell, const double* r,
                                const int* cellStart, const int* cellCount,
                                const int* cellAtoms, int* count, int* stats,
                                hipStream_t s)
{
    if (ncell <= 0) return hipSuccess;
    k_nl_cell<0><<<ncell, 256, 0, s>>>(p, r, cellStart, cellCount, cellAtoms,
                                       count, stats, nullptr, 0, nullptr,
                                       nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_nl_fill_cell(NLParams const& p, int ncell, const double* r,
                               const int* cellStart, const int* cellCount,
                               const int* cellAtoms, int* count, int* stats,
                               const int* off, int limit, uint64_t* keys,
                               int* nbrJ, double4* geo, hipStream_t s)
{
    if (ncell <= 0) return hipSuccess;
    k_nl_cell<1><<<ncell, 256, 0, s>>>(p, r, cellStart, cellCount, cellAtoms,
                                       count, stats, off, limit, keys, nbrJ,
                                       geo);
    return hipGetLastError();
}

hipError_t launch_batch_extent(int nstruct, const double* r,
                               const int* structOff, BatchCell* cells,
                               hipStream_t s)
{
    if (nstruct <= 0) return hipSuccess;
    k_batch_extent<<<nstruct, 64, 0, s>>>(nstruct, r, structOff, cells);
    return hipGetLastError();
}

hipError_t launch_nl_batch(int mode, int n, double rc, const double* r,
                           const int* structOf, const int* structOff,
                           const BatchCell* cells, int* count, int* stats,
                           const int* off, int limit, uint64_t* keys,
                           int* nbrJ, double4* geo, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    double const rc2 = rc * rc, rcm = rc * (1.0 + 1e-6);
    if (mode == 0)
        k_nl_batch<0><<<n, 64, 0, s>>>(n, rc2, rcm, r, structOf, structOff,
                                       cells, count, nullptr, nullptr, 0,
                                       nullptr, nullptr, nullptr);
    else
        k_nl_batch<1><<<n, 64, 0, s>>>(n, rc2, rcm, r, structOf, structOff,
                                       cells, count, stats, off, limit, keys,
                                       nbrJ, geo);
    return hipGetLastError();
}

hipError_t launch_slab_starts(int n, int slab, int* off, hipStream_t s)
{
    if (n < 0) return hipSuccess;
    k_slab_starts<<<(n + 256) / 256, 256, 0, s>>>(n, slab, off);
    return hipGetLastError();
}


hipError_t launch_nl_partner(int n, const int* off, const int* cnt,
                             const uint64_t* keys, const int* order,
                             int* partner, int64_t nslots, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipError_t const e = hipMemsetAsync(partner, 0xff,
                                        (size_t)nslots * sizeof(int), s);
    if (e != hipSuccess) return e;
    k_partner<<<n, 64, 0, s>>>(n, off, cnt, keys, order, partner);
    return hipGetLastError();
}

}  // namespace hdnnp
