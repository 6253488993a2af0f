// Atomic neural networks and the energy sum.
//
// NeuralNetwork::propagate (NeuralNetwork.cpp:386-394, :761-896) for all
// atoms of one element at once: thread t handles atom atoms[t]; layer
// weights are wave-uniform (scalar loads); per-atom activations go to a
// scratch buffer laid out [neuron][atom] so neighbouring threads touch
// neighbouring addresses.  dE/dG (NeuralNetwork::calculateDEdG,
// NeuralNetwork.cpp:396-442) is computed in reverse mode,
//   dE/dG = W1^T (g1 .* W2^T (g2 .* ... W_L^T g_L)),  g_l = f'_l / n_{l-1}
// (the division only with normalize_nodes), which is the same product the
// reference forms input by input in forward mode.
#include "device_common.h"
#include "kernels.h"

namespace hdnnp {

namespace {

__global__ __launch_bounds__(256) void k_nn(NNArgs a)
{
    int const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.na) return;
    ModelTables const& T = a.T;
    ElemDesc const& ed = T.elem[a.e];
    int const atom = a.atoms[t];
    int const na = a.na;
    int const L = ed.nlayers;
    double const* __restrict__ Gi = a.G + (size_t)atom * a.ldG;
    // scratch: value, dfdx and delta planes, each sum(width[1..L-1]) x na
    int plane = 0;
    for (int l = 1; l < L; ++l) plane += ed.width[l];
    double* __restrict__ val = a.scratch;
    double* __restrict__ dfd = val + (size_t)plane * na;
    double* __restrict__ del = dfd + (size_t)plane * na;

    int inOff = -1;         // -1: input layer is G
    int outOff = 0;
    double out = 0.0, outd = 1.0;
    for (int l = 1; l < L; ++l)
    {
        int const nin = ed.width[l - 1], nout = ed.width[l];
        double const* __restrict__ W = T.weights + ed.woff[l];
        double const* __restrict__ B = T.weights + ed.boff[l];
        for (int k = 0; k < nout; ++k)
        {
            double acc = 0.0;
            double const* __restrict__ Wk = W + (size_t)k * nin;
            if (inOff < 0)
                for (int j = 0; j < nin; ++j) acc += Wk[j] * Gi[j];
            else
                for (int j = 0; j < nin; ++j)
                    acc += Wk[j] * val[(size_t)(inOff + j) * na + t];
            acc += B[k];
            if (ed.normalize) acc /= nin;
            double v, d;
            activate(ed.act[l], acc, v, d);
            val[(size_t)(outOff + k) * na + t] = v;
            dfd[(size_t)(outOff + k) * na + t] = d;
            out = v;
            outd = d;
        }
        inOff = outOff;
        outOff += nout;
    }
    a.E[atom] = out;
    if (!a.deriv) return;

    // Reverse pass.  Output layer (single neuron) offset = plane - 1.
    int lOff = plane - 1;
    {
        double g = outd;
        if (ed.normalize) g /= ed.width[L - 2];
        del[(size_t)lOff * na + t] = g;
    }
    for (int l = L - 1; l >= 2; --l)
    {
        int const nin = ed.width[l - 1], nout = ed.width[l];
        int const inO = lOff - nin;   // offset of layer l-1
        double const* __restrict__ W = T.weights + ed.woff[l];
        for (int j = 0; j < nin; ++j)
        {
            double s = 0.0;
            for (int k = 0; k < nout; ++k)
                s += W[(size_t)k * nin + j] * del[(size_t)(lOff + k) * na + t];
            double g = dfd[(size_t)(inO + j) * na + t];
            if (ed.normalize) g /= ed.width[l - 2];
            del[(size_t)(inO + j) * na + t] = s * g;
        }
        lOff = inO;
    }
    {
        int const nin = ed.width[0], nout = ed.width[1];
        double const* __restrict__ W = T.weights + ed.woff[1];
        double* __restrict__ out_dEdG = a.dEdG + (size_t)atom * a.ldG;
        for (int j = 0; j < nin; ++j)
        {
            double s = 0.0;
            for (int k = 0; k < nout; ++k)
                s += W[(size_t)k * nin + j] * del[(size_t)k * na + t];
            out_dEdG[j] = s;
        }
    }
}

// Matrix-core (MFMA) form of the same networks: one wavefront per tile of
// 16 atoms of one element; every layer is the GEMM Z = X W^T over the tile
// (K = fan-in), the backward pass the GEMMs Delta_(l-1) = Delta_l W_l
// (K = fan-out).  Activations and derivatives live in LDS per tile, each
// layer's output is written back in the operand layout of the next GEMM.
// MODE 0: v_mfma_f64_16x16x4f64 (plain FP64 path).  MODE 1:
// v_mfma_f32_16x16x4f32 (mixed path).  MODE 2: v_mfma_f32_16x16x16f16,
// FP16 operands with FP32 accumulation.  MODE 3:
// v_mfma_f32_16x16x16bf16_1k, BF16 operands (round to nearest even) in the
// FP16 operand layout, FP32 accumulation.  Modes 2 and 3 split every
// operand into a high and a low 16-bit part (three MFMAs per step, Split4
// below).  Biases, activations and normalisation in FP64 in every mode.
// Operand layouts (16x16 tiles, lane l, v < 4): A[i = l % 16][k = K0 (l /
// 16) + v'], B[k = K0 (l / 16) + v'][j = l % 16] with K0 = 1 (FP64, FP32;
// one value per lane) or 4 (FP16, v' < 4); D[i][j = l % 16] in register v
// with i = l / 16 + 4 v for FP64 and i = 4 (l / 16) + v for the FP32
// accumulators (measured with a lane-layout probe on gfx950).
typedef double nn_d4 __attribute__((ext_vector_type(4)));
typedef float nn_f4 __attribute__((ext_vector_type(4)));
typedef _Float16 nn_h4 __attribute__((ext_vector_type(4)));
typedef short nn_s4 __attribute__((ext_vector_type(4)));

// BF16 bits of x, round to nearest even (finite inputs).
__device__ __forceinline__ short to_bf16(float x)
{
    unsigned const u = __float_as_uint(x);
    unsigned const r = u + 0x7fffu + ((u >> 16) & 1u);
    return (short)(r >> 16);
}

template <int MODE> struct NNMma;

template <> struct NNMma<0>
{
    typedef double Act;
    typedef nn_d4 Acc;
    static constexpr int K = 4;
    // A tile step: X[m][k0 + l/16] (zero beyond nk).
    __device__ static double a(Act const* X, int ld, int nk, int k0, int lane)
    {
        int const m = lane & 15, k = k0 + (lane >> 4);
        return k < nk ? X[m * ld + k] : 0.0;
    }
    // B element (k, j) = W[j * ldw + k] (forward: W^T) or W[k * ldw + j]
    // (backward); zero outside [0, nk) x [0, nj).
    __device__ static double b(ModelTables const& T, int off, bool trans,
                               int ldw, int nk, int nj, int k0, int j0,
                               int lane)
    {
        int const j = j0 + (lane & 15), k = k0 + (lane >> 4);
        if (k >= nk || j >= nj) return 0.0;
        return T.weights[off + (trans ? j * ldw + k : k * ldw + j)];
    }
    __device__ static Acc mma(double a, double b, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <> struct NNMma<1>
{
    typedef float Act;
    typedef nn_f4 Acc;
    static constexpr int K = 4;
    __device__ static float a(Act const* X, int ld, int nk, int k0, int lane)
    {
        int const m = lane & 15, k = k0 + (lane >> 4);
        return k < nk ? X[m * ld + k] : 0.0f;
    }
    __device__ static float b(ModelTables const& T, int off, bool trans,
                              int ldw, int nk, int nj, int k0, int j0,
                              int lane)
    {
        int const j = j0 + (lane & 15), k = k0 + (lane >> 4);
        if (k >= nk || j >= nj) return 0.0f;
        return T.weights32[off + (trans ? j * ldw + k : k * ldw + j)];
    }
    __device__ static Acc mma(float a, float b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// FP16 and BF16 operands split in two: x = hi + lo with hi = round(x) and
// lo = round(x - hi) in the 16-bit type.  A tile step is three MFMAs,
// hi.hi + hi.lo + lo.hi (the lo.lo term is below the FP32 accumulation),
// so the products carry about twice the 16-bit significand: 22 bits for
// FP16, 16 bits for BF16, against 11 and 8 for a single operand.
template <typename H>
struct Split4
{
    H hi, lo;
};

__device__ __forceinline__ void split_f16(float x, _Float16& hi, _Float16& lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

__device__ __forceinline__ float bf16_to_f32(short b)
{
    return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}

__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo)
{
    hi = to_bf16(x);
    lo = to_bf16(x - bf16_to_f32(hi));
}

template <> struct NNMma<2>
{
    typedef float Act;
    typedef nn_f4 Acc;
    typedef Split4<nn_h4> Op;
    static constexpr int K = 16;
    __device__ static Op a(Act const* X, int ld, int nk, int k0, int lane)
    {
        int const m = lane & 15, k = k0 + 4 * (lane >> 4);
        Op v;
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            _Float16 h, l;
            split_f16(k + q < nk ? X[m * ld + k + q] : 0.0f, h, l);
            v.hi[q] = h;
            v.lo[q] = l;
        }
        return v;
    }
    __device__ static Op b(ModelTables const& T, int off, bool trans, int ldw,
                           int nk, int nj, int k0, int j0, int lane)
    {
        int const j = j0 + (lane & 15), k = k0 + 4 * (lane >> 4);
        Op v;
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            float w = 0.0f;
            if (k + q < nk && j < nj)
                w = T.weights32[off + (trans ? j * ldw + k + q
                                             : (k + q) * ldw + j)];
            _Float16 h, l;
            split_f16(w, h, l);
            v.hi[q] = h;
            v.lo[q] = l;
        }
        return v;
    }
    __device__ static Acc mma(Op const& a, Op const& b, Acc c)
    {
        c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.lo, b.hi, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.lo, c, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a.hi, b.hi, c, 0, 0, 0);
    }
    __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <> struct NNMma<3>
{
    typedef float Act;
    typedef nn_f4 Acc;
    typedef Split4<nn_s4> Op;
    static constexpr int K = 16;
    __device__ static Op a(Act const* X, int ld, int nk, int k0, int lane)
    {
        int const m = lane & 15, k = k0 + 4 * (lane >> 4);
        Op v;
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            short h, l;
            split_bf16(k + q < nk ? X[m * ld + k + q] : 0.0f, h, l);
            v.hi[q] = h;
            v.lo[q] = l;
        }
        return v;
    }
    __device__ static Op b(ModelTables const& T, int off, bool trans, int ldw,
                           int nk, int nj, int k0, int j0, int lane)
    {
        int const j = j0 + (lane & 15), k = k0 + 4 * (lane >> 4);
        Op v;
#pragma unroll
        for (int q = 0; q < 4; ++q)
        {
            float w = 0.0f;
            if (k + q < nk && j < nj)
                w = T.weights32[off + (trans ? j * ldw + k + q
                                             : (k + q) * ldw + j)];
            short h, l;
            split_bf16(w, h, l);
            v.hi[q] = h;
            v.lo[q] = l;
        }
        return v;
    }
    __device__ static Acc mma(Op const& a, Op const& b, Acc c)
    {
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.lo, b.hi, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.lo, c, 0, 0, 0);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a.hi, b.hi, c, 0, 0,
                                                         0);
    }
    __device__ static int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// One wave computes R tiles of 16 atoms: every B operand (a weight tile)
// is loaded once and used for the R tiles' MFMAs.
template <int MODE, int R>
__global__ __launch_bounds__(64) void k_nn_mfma(NNArgs a)
{
    typedef NNMma<MODE> MM;
    typedef typename MM::Act Act;
    typedef typename MM::Acc Acc;
    extern __shared__ double nnsm[];
    Act* const act = (Act*)nnsm;
    ModelTables const& T = a.T;
    ElemDesc const& ed = T.elem[a.e];
    int const L = ed.nlayers;
    int const lane = threadIdx.x;
    int const tile0 = blockIdx.x * 16 * R;
    // LDS per tile: value planes of layers 0..L-1, then derivative planes
    // of 1..L-1, each [16][width]; tile r at r * tstride.
    int voff[kMaxLayers], doff[kMaxLayers];
    int tstride;
    {
        int o = 0;
        for (int l = 0; l < L; ++l) { voff[l] = o; o += 16 * ed.width[l]; }
        for (int l = 1; l < L; ++l) { doff[l] = o; o += 16 * ed.width[l]; }
        tstride = o;
    }
    // Input layer: the tiles' scaled SFs (zero rows past the last atom).
    {
        int const nin = ed.width[0];
        for (int q = lane; q < 16 * R * nin; q += 64)
        {
            int const m = q / nin, k = q - m * nin;
            int const t = tile0 + m;
            double g = 0.0;
            if (t < a.na) g = a.G[(size_t)a.atoms[t] * a.ldG + k];
            act[(m >> 4) * tstride + voff[0] + (m & 15) * nin + k] = (Act)g;
        }
    }
    __syncthreads();
    for (int l = 1; l < L; ++l)
    {
        int const nin = ed.width[l - 1], nout = ed.width[l];
        for (int j0 = 0; j0 < nout; j0 += 16)
        {
            Acc acc[R];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[r][v] = 0;
            for (int k0 = 0; k0 < nin; k0 += MM::K)
            {
                auto const b = MM::b(T, ed.woff[l], true, nin, nin, nout, k0,
                                     j0, lane);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    acc[r] = MM::mma(MM::a(act + r * tstride + voff[l - 1],
                                           nin, nin, k0, lane),
                                     b, acc[r]);
            }
            int const j = j0 + (lane & 15);
            if (j < nout)
            {
                double const bias = T.weights[ed.boff[l] + j];
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                    {
                        int const m = MM::row(lane, v);
                        double z = (double)acc[r][v] + bias;
                        if (ed.normalize) z /= nin;
                        double val, d;
                        activate(ed.act[l], z, val, d);
                        Act* const P = act + r * tstride;
                        P[voff[l] + m * nout + j] = (Act)val;
                        P[doff[l] + m * nout + j] = (Act)d;
                        int const t = tile0 + 16 * r + m;
                        if (l == L - 1 && t < a.na)
                            a.E[a.atoms[t]] = val;   // one output neuron
                    }
            }
        }
        __syncthreads();
    }
    if (!a.deriv) return;
    // Output delta (one neuron): f' / fan-in with normalize_nodes.
    for (int q = lane; q < 16 * R; q += 64)
    {
        Act* const P = act + (q >> 4) * tstride + doff[L - 1];
        double g = (double)P[q & 15];
        if (ed.normalize) g /= ed.width[L - 2];
        P[q & 15] = (Act)g;
    }
    __syncthreads();
    for (int l = L - 1; l >= 1; --l)
    {
        int const nin = ed.width[l - 1], nout = ed.width[l];
        for (int j0 = 0; j0 < nin; j0 += 16)
        {
            Acc acc[R];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[r][v] = 0;
            for (int k0 = 0; k0 < nout; k0 += MM::K)
            {
                auto const b = MM::b(T, ed.woff[l], false, nin, nout, nin, k0,
                                     j0, lane);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    acc[r] = MM::mma(MM::a(act + r * tstride + doff[l], nout,
                                           nout, k0, lane),
                                     b, acc[r]);
            }
            int const j = j0 + (lane & 15);
            if (j < nin)
            {
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                    {
                        int const m = MM::row(lane, v);
                        int const t = tile0 + 16 * r + m;
                        if (l >= 2)
                        {
                            int const q = r * tstride + doff[l - 1] + m * nin
                                        + j;
                            double g = (double)act[q];
                            if (ed.normalize) g /= ed.width[l - 2];
                            act[q] = (Act)((double)acc[r][v] * g);
                        }
                        else if (t < a.na)
                            a.dEdG[(size_t)a.atoms[t] * a.ldG + j]
                                = (double)acc[r][v];
                    }
            }
        }
        __syncthreads();
    }
}

// Energy = sum of atomic energies: each thread sums a strided slice in
// order, then a fixed tree; deterministic.
__global__ __launch_bounds__(1024) void k_energy(const double* __restrict__ E,
                                                 int n, double* __restrict__ out)
{
    __shared__ double s[1024];
    int const t = threadIdx.x;
    double acc = 0.0;
    for (int i = t; i < n; i += 1024) acc += E[i];
    s[t] = acc;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1)
    {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) *out = s[0];
}

// Energy of each structure of a batch: one block per structure, the same
// fixed-order tree as k_energy over the structure's atoms.
__global__ __launch_bounds__(256) void k_energy_seg(
    const double* __restrict__ E, const int* __restrict__ structOff,
    double* __restrict__ out)
{
    __shared__ double s[256];
    int const st = blockIdx.x;
    int const t = threadIdx.x;
    int const a0 = structOff[st], a1 = structOff[st + 1];
    double acc = 0.0;
    for (int i = a0 + t; i < a1; i += 256) acc += E[i];
    s[t] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1)
    {
        if (t < o) s[t] += s[t + o];
        __syncthreads();
    }
    if (t == 0) out[st] = s[0];
}

}  // namespace

hipError_t launch_energy_segments(const double* E, const int* structOff,
                                  int nstruct, double* out, hipStream_t s)
{
    if (nstruct <= 0) return hipSuccess;
    k_energy_seg<<<nstruct, 256, 0, s>>>(E, structOff, out);
    return hipGetLastError();
}

size_t nn_mfma_lds_bytes(ElemDesc const& ed, int mode, int tiles)
{
    size_t w = 0;
    for (int l = 0; l < ed.nlayers; ++l) w += ed.width[l];
    for (int l = 1; l < ed.nlayers; ++l) w += ed.width[l];
    return (size_t)tiles * 16 * w
         * (mode >= 1 ? sizeof(float) : sizeof(double));
}

namespace {

template <int MODE>
void launch_nn_tiles(NNArgs const& a, hipStream_t s)
{
    size_t const shm = nn_mfma_lds_bytes(a.ed, a.mode, a.tiles);
    int const nb = (a.na + 16 * a.tiles - 1) / (16 * a.tiles);
    if (a.tiles == 4) k_nn_mfma<MODE, 4><<<nb, 64, shm, s>>>(a);
    else if (a.tiles == 2) k_nn_mfma<MODE, 2><<<nb, 64, shm, s>>>(a);
    else k_nn_mfma<MODE, 1><<<nb, 64, shm, s>>>(a);
}

}  // namespace

hipError_t launch_nn(NNArgs const& a, hipStream_t s)
{
    if (a.na <= 0) return hipSuccess;
    if (a.mode >= 0 && a.tiles != 1 && a.tiles != 2 && a.tiles != 4)
        return hipErrorInvalidValue;
    if (a.mode < 0)
        k_nn<<<(a.na + 255) / 256, 256, 0, s>>>(a);
    else if (a.mode == 3) launch_nn_tiles<3>(a, s);
    else if (a.mode == 2) launch_nn_tiles<2>(a, s);
    else if (a.mode == 1) launch_nn_tiles<1>(a, s);
    else launch_nn_tiles<0>(a, s);
    return hipGetLastError();
}

hipError_t launch_energy_sum(const double* E, int n, double* out, hipStream_t s)
{
    k_energy<<<1, 1024, 0, s>>>(E, n, out);
    return hipGetLastError();
}

}  // namespace hdnnp
