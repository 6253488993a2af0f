// Atomic neural networks (FP64) and the energy sum.
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

}  // namespace

hipError_t launch_nn(NNArgs const& a, hipStream_t s)
{
    if (a.na <= 0) return hipSuccess;
    k_nn<<<(a.na + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_energy_sum(const double* E, int n, double* out, hipStream_t s)
{
    k_energy<<<1, 1024, 0, s>>>(E, n, out);
    return hipGetLastError();
}

}  // namespace hdnnp
