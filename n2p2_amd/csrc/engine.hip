// Engine context and the C ABI declared in include/hdnnp.h.
//
// One context per Mode object: device copies of the model tables, the
// structure, neighbour list, per-atom results and a HIP stream.  Device
// buffers grow on demand and are reused between calls (a structure of the
// same size re-uses everything; nothing is allocated per step).
#include "hdnnp.h"

#include "kernels.h"
#include "model.h"
#include "specialize.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

using namespace hdnnp;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error
{
    int code;
    Error(int c, std::string const& m) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw Error(HDNNP_E_DEVICE, std::string("ERROR: HIP error in ") + what
                    + ": " + hipGetErrorString(e) + "\n");
}

template <class F> int guarded(F&& f)
{
    try
    {
        g_err.clear();
        f();
        return HDNNP_OK;
    }
    catch (Error const& e)
    {
        g_err = e.what();
        return e.code;
    }
    catch (std::exception const& e)
    {
        g_err = e.what();
        std::string const w = e.what();
        if (w.find("open file") != std::string::npos) return HDNNP_E_FILE;
        if (w.find("outside this engine") != std::string::npos)
            return HDNNP_E_UNSUPPORTED;
        return HDNNP_E_SETTINGS;
    }
}

// Grow-only device buffer.
template <class T> struct DBuf
{
    T* p = nullptr;
    size_t cap = 0;
    void ensure(size_t n)
    {
        if (n <= cap) return;
        if (p) hipCheck(hipFree(p), "hipFree");
        p = nullptr;
        size_t const want = std::max<size_t>(n, 1);
        hipCheck(hipMalloc((void**)&p, want * sizeof(T)), "hipMalloc");
        cap = want;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

enum KClass { K_NEIGHBOR = 0, K_SYMFUNC, K_NETWORK, K_FORCE, K_GATHER,
              K_PARTNER, K_NUM };
const char* const kClassNames[K_NUM] = {"neighbor", "symfunc", "network",
                                        "force", "gather", "partner"};

}  // namespace

struct hdnnp_model
{
    Model m;
};

struct hdnnp_ctx
{
    const hdnnp_model* model = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    Model::Flat flat;
    ModelTables T{};
    DBuf<ElemDesc> dElem;
    DBuf<CutCfg> dCfg;
    DBuf<GroupDesc> dGroup;
    DBuf<MemberDesc> dMember;
    DBuf<LamRun> dRun;
    DBuf<ScaleDesc> dScale;
    DBuf<CompGroup> dCompGroup;
    DBuf<CompMember> dCompMember;
    DBuf<int> dPairGroups, dPairCount;
    DBuf<double> dWeights;
    DBuf<float> dWeights32;
    int maxMembers = 0, maxCfg = 1, ld = 1;

    // structure
    int64_t n = 0;
    bool periodic = false;
    double box[3][3] = {{0}};
    double invbox[3][3] = {{0}};
    std::vector<int32_t> hElem;
    std::vector<double> hR;          // host copy of positions (if set from host)
    bool hostPosValid = false;
    std::vector<std::vector<int>> hAtomsOfElem;
    std::vector<DBuf<int>> dAtomsOfElem;
    DBuf<double> dR;
    DBuf<int> dElemAtom;
    // Batch of independent structures (hdnnp_set_structures): atoms of
    // structure s are [structOff[s], structOff[s+1]) of the flat arrays.
    bool batched = false;
    int64_t nstruct = 0;
    std::vector<int> hStructOff;
    std::vector<double> hBoxes;     // 9 per structure (zero if open)
    std::vector<char> hPeriodic;
    std::vector<int> hPbcS;         // 3 per structure
    DBuf<int> dStructOf, dStructOff;
    DBuf<BatchCell> dCells;
    DBuf<double> dEnergySeg;

    // neighbour list
    NLParams nl{};
    int pbc[3] = {0, 0, 0};
    int ncell = 0;
    // Entries of atom i: [off[i], off[i] + count[i]).  Cell builds use a
    // fixed range of `slab` slots per atom (off[i] = i * slab, no count
    // pass); the per-atom fallback builds a compact CSR.  entryCap = slots
    // allocated; M = total entries (summed on demand, Mvalid).
    int slab = 0;
    int slabAtoms = -1;
    bool slabLayout = false;
    int64_t entryCap = 0;
    int64_t M = 0;
    bool Mvalid = false;
    int maxNeigh = 0;
    bool haveNL = false;
    // Neighbour list given by the caller (hdnnp_set_neighbor_list): the
    // caller's index of every centre and of every entry's neighbour, used
    // to scatter the forces.
    bool externalNL = false;
    // Per caller atom b < extTargets, the contributions it receives in the
    // order of InterfaceLammps::getForces (k_scatter_ext): offsets and
    // sources (entry q >= 0, or -(centre + 1) for a centre's self force).
    int64_t extTargets = 0;
    DBuf<int64_t> dScatOff;
    DBuf<int> dScatSrc;
    DBuf<double> dScatF;
    DBuf<int> dCellOfAtom, dRank, dCellCount, dCellStart, dCellAtoms, dCount,
        dOff, dStats, dBlockSums, dTotal;
    DBuf<uint64_t> dKeys;
    DBuf<int> dNbrJ, dPartner;
    DBuf<double4> dGeo;

    // results
    DBuf<double> dG, dE, dEdG, dScratch, dCslot, dCself, dF, dEnergy, dStatOut;
    DBuf<int> dEw;
    bool haveSF = false, haveNN = false, haveNNDeriv = false;
    double energy = 0.0;
    int precision = HDNNP_PRECISION_FP64;
    // Pair record of the mixed path: the SF kernel writes the passing
    // neighbour pairs of every angular group, the force kernel replays them
    // instead of enumerating and testing the pairs again.  pairsValid: the
    // last SF evaluation recorded them for the current list.
    bool pairRecord = pair_record_enabled_by_default();
    bool pairsValid = false;
    DBuf<unsigned short> dPairs;
    // Model-specialised SF / force kernels per arithmetic (0 FP64, 1 FP32),
    // compiled on first use (specialize.hip); generic kernels otherwise.
    bool specialized = plan_enabled_by_default();
    PlanKernels plan[2];
    bool planTried[2] = {false, false};
    // Tuning (hdnnp_ctx_autotune): SF members per group chunk of the device
    // tables, waves-per-SIMD bounds of the specialised kernels, 16-atom
    // tiles per wave of the MFMA networks.
    int chunk = kMaxMembers;
    PlanTuning tune;
    int nnTiles = 1;
    PlanKernels planFor(int fp32)
    {
        if (!specialized) return PlanKernels();
        if (!planTried[fp32])
        {
            plan[fp32] = plan_kernels(flat, T, fp32, device, tune);
            planTried[fp32] = true;
        }
        return plan[fp32];
    }

    // MD with a Verlet-skin list (hdnnp_md_*).  The skin list (rc + skin)
    // lives in skinKeys / skinPartner / skinCnt; every step filters it into
    // the ordinary list buffers.  Graph A: kick, drift, filter, flags to
    // the host.  Graph B: the force evaluation, kick, energy record.
    // MD flags: [0] moved more than half the skin, [1] filtered list over
    // capacity, [2] unstable, [3] partner misses, [4] the last rebuild's
    // wrap shifted atoms.
    static constexpr int kMdFlags = 5;
    struct MDState
    {
        bool on = false;
        double dt = 0.0, skin = 0.0, rc = 0.0, ftoa = 0.0;
        int forceMode = HDNNP_FORCES_CONSISTENT;
        bool useGraph = true;
        int cap = 0;
        int64_t rebuilds = 0, steps = 0;
        // Rebuilds whose wrap shifted atoms by lattice vectors (flags[4]).
        int64_t wrapMoves = 0;
        MDCell cell{};
        DBuf<double> v, r0, accel, mass, part, epot, ekin;
        DBuf<int> flags, counter, skinPartner, skinCnt, newIdx;
        DBuf<uint64_t> skinKeys;
        int* hFlags = nullptr;
        int recCap = 0;
        hipGraph_t gA = nullptr, gB = nullptr;
        hipGraphExec_t eA = nullptr, eB = nullptr;
    } md;
    bool capturing = false;

    // Multiple-time-step (r-RESPA) run driven by this context as the fast
    // (surrogate) model, with `full` as the reference model (hdnnp_mts_*).
    struct MTSState
    {
        hdnnp_ctx* full = nullptr;
        int nInner = 4;
        double dtOuter = 0.0, gate = 0.0;
        bool useSur = true;
        int64_t outer = 0, fallbacks = 0;
        DBuf<double> d2, sum, epot, ekin;
        DBuf<int> counter;
        int recCap = 0;
    } mts;

    // Kalman filter on the weights (hdnnp_kalman_*).
    DBuf<double> kfP, kfX, kfA, kfH;
    double kfEta = 0.0, kfEtaTau = 0.0, kfEtaMax = 0.0, kfQ = 0.0,
           kfQTau = 0.0, kfQMin = 0.0;
    int64_t kfUpdates = -1;

    // Energy weight Jacobian (hdnnp_energy_weight_jacobian).
    DBuf<int> jAtoms, jB0, jBN, jBE, jGrp0, jBase, jNwe;
    DBuf<double> jPart, jOut;

    // Training of a radial-only model (hdnnp_train_*).
    DBuf<double> tU, tD2, tPart, tGrad, tM, tV, tFref, tLoss;
    int64_t adamStep = 0;
    bool haveGrad = false;

    // profiling
    bool prof = false;
    struct Pending { int cls; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> eventPool;
    double profMs[K_NUM] = {0};
    int64_t profCalls[K_NUM] = {0};

    hipEvent_t getEvent()
    {
        if (!eventPool.empty())
        {
            hipEvent_t e = eventPool.back();
            eventPool.pop_back();
            return e;
        }
        hipEvent_t e;
        hipCheck(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    void begin(int cls, hipEvent_t& a, hipStream_t on = nullptr)
    {
        (void)cls;
        if (!prof || capturing) return;
        a = getEvent();
        hipCheck(hipEventRecord(a, on ? on : stream), "hipEventRecord");
    }
    void end(int cls, hipEvent_t a, hipStream_t on = nullptr)
    {
        if (!prof || capturing) return;
        hipEvent_t b = getEvent();
        hipCheck(hipEventRecord(b, on ? on : stream), "hipEventRecord");
        pending.push_back(Pending{cls, a, b});
    }

    // The partner map of a list build runs on a second stream (default
    // priority), concurrently with the SF, network and force kernels that
    // do not read it; the gather, the next build and the MD rebuild wait
    // for it (waitPartner).
    hipStream_t side = nullptr;
    hipEvent_t evListDone = nullptr, evPartnerDone = nullptr;
    bool partnerPending = false;
    void waitPartner()
    {
        if (!partnerPending) return;
        hipCheck(hipStreamWaitEvent(stream, evPartnerDone, 0),
                 "hipStreamWaitEvent");
        partnerPending = false;
    }
    void collect()
    {
        for (auto& p : pending)
        {
            float ms = 0.0f;
            hipCheck(hipEventSynchronize(p.b), "hipEventSynchronize");
            hipCheck(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
            profMs[p.cls] += ms;
            profCalls[p.cls] += 1;
            eventPool.push_back(p.a);
            eventPool.push_back(p.b);
        }
        pending.clear();
    }
    void sync()
    {
        hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (side)
            hipCheck(hipStreamSynchronize(side), "hipStreamSynchronize");
        collect();
    }

    void setDevice() { hipCheck(hipSetDevice(device), "hipSetDevice"); }

    void uploadModel()
    {
        flat = model->m.flatten(chunk);
        planTried[0] = planTried[1] = false;
        auto up = [&](auto& buf, auto const& v) {
            buf.ensure(v.size());
            if (!v.empty())
                hipCheck(hipMemcpy(buf.p, v.data(),
                                   v.size() * sizeof(v[0]),
                                   hipMemcpyHostToDevice), "upload model");
        };
        up(dElem, flat.elem);
        up(dCfg, flat.cfg);
        up(dGroup, flat.group);
        up(dMember, flat.member);
        up(dRun, flat.run);
        up(dScale, flat.scale);
        up(dCompGroup, flat.cgroup);
        up(dCompMember, flat.cmember);
        up(dPairGroups, flat.pairGroups);
        up(dPairCount, flat.pairCount);
        up(dWeights, flat.weights);
        {
            std::vector<float> w32(flat.weights.begin(), flat.weights.end());
            up(dWeights32, w32);
        }
        T.ne = (int)flat.elem.size();
        T.maxnsf = model->m.maxNumSF();
        T.maxCutoff = model->m.maxCutoffRadius;
        for (int k = 0; k < kMaxElements; ++k)
            T.zw[k] = k < T.ne ? (double)model->m.el[k].Z : 0.0;
        T.cutType = model->m.cutoffType;
        T.genpow = 0;
        T.maxGroupMembers = 0;
        T.weighted = 0;
        for (auto const& g : flat.group)
        {
            T.maxGroupMembers = std::max(T.maxGroupMembers, g.nmem);
            if (sf_weighted(g.type)) T.weighted = 1;
            if (!sf_radial(g.type))
                for (int m = g.mem0; m < g.mem0 + g.nmem; ++m)
                    if (!flat.member[m].useint) T.genpow = 1;
        }
        T.elem = dElem.p;
        T.cfg = dCfg.p;
        T.group = dGroup.p;
        T.member = dMember.p;
        T.run = dRun.p;
        T.scale = dScale.p;
        T.pairGroups = dPairGroups.p;
        T.pairCount = dPairCount.p;
        T.weights = dWeights.p;
        T.weights32 = dWeights32.p;
        T.cgroup = dCompGroup.p;
        T.cmember = dCompMember.p;
        T.ncgroup = (int)flat.cgroup.size();
        T.maxCompMembers = 1;
        for (auto const& g : flat.cgroup)
            T.maxCompMembers = std::max(T.maxCompMembers, g.nmem);
        ld = std::max(1, T.maxnsf);
        maxMembers = 0;
        maxCfg = 1;
        for (auto const& e : flat.elem)
        {
            int nm = 0;
            for (int g = e.g0; g < e.g0 + e.ng; ++g) nm += flat.group[g].nmem;
            maxMembers = std::max(maxMembers, nm);
            maxCfg = std::max(maxCfg, e.ncfg);
        }
    }
};

namespace {

void requireStructure(hdnnp_ctx* c)
{
    if (c->n <= 0)
        throw Error(HDNNP_E_INVALID, "ERROR: No structure loaded.\n");
}

void requireNL(hdnnp_ctx* c)
{
    requireStructure(c);
    if (!c->haveNL)
        throw Error(HDNNP_E_INVALID, "ERROR: Neighbor list not calculated.\n");
}

// Structure::calculateInverseBox (Structure.cpp:519-545).
void inverseBox(double const b[3][3], double ib[3][3])
{
    double invdet = b[0][0] * b[1][1] * b[2][2] + b[1][0] * b[2][1] * b[0][2]
                  + b[2][0] * b[0][1] * b[1][2] - b[2][0] * b[1][1] * b[0][2]
                  - b[0][0] * b[2][1] * b[1][2] - b[1][0] * b[0][1] * b[2][2];
    invdet = 1.0 / invdet;
    ib[0][0] = b[1][1] * b[2][2] - b[2][1] * b[1][2];
    ib[0][1] = b[2][1] * b[0][2] - b[0][1] * b[2][2];
    ib[0][2] = b[0][1] * b[1][2] - b[1][1] * b[0][2];
    ib[1][0] = b[2][0] * b[1][2] - b[1][0] * b[2][2];
    ib[1][1] = b[0][0] * b[2][2] - b[2][0] * b[0][2];
    ib[1][2] = b[1][0] * b[0][2] - b[0][0] * b[1][2];
    ib[2][0] = b[1][0] * b[2][1] - b[2][0] * b[1][1];
    ib[2][1] = b[2][0] * b[0][1] - b[0][0] * b[2][1];
    ib[2][2] = b[0][0] * b[1][1] - b[1][0] * b[0][1];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ib[i][j] *= invdet;
}

// Perpendicular widths of the cell (|a . (b x c)| / |b x c| etc.) and the
// image counts of Structure::calculatePbcCopies (Structure.cpp:486-517).
void cellWidths(double const b[3][3], double rc, double proj[3], int pbc[3])
{
    auto cross = [](double const* x, double const* y, double* o) {
        o[0] = x[1] * y[2] - x[2] * y[1];
        o[1] = x[2] * y[0] - x[0] * y[2];
        o[2] = x[0] * y[1] - x[1] * y[0];
    };
    auto normalize = [](double* v) {
        double const n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= n; v[1] /= n; v[2] /= n;
    };
    auto dot = [](double const* x, double const* y) {
        return x[0] * y[0] + x[1] * y[1] + x[2] * y[2];
    };
    double axb[3], axc[3], bxc[3];
    cross(b[0], b[1], axb); normalize(axb);
    cross(b[0], b[2], axc); normalize(axc);
    cross(b[1], b[2], bxc); normalize(bxc);
    proj[0] = std::fabs(dot(b[0], bxc));
    proj[1] = std::fabs(dot(b[1], axc));
    proj[2] = std::fabs(dot(b[2], axb));
    for (int k = 0; k < 3; ++k)
    {
        pbc[k] = 0;
        while (pbc[k] * proj[k] <= rc) pbc[k]++;
    }
}

// Neighbour lists of a batch of structures, one launch per pass.  Each
// structure's image counts follow its own cell (cellWidths, as
// Structure.cpp:287-300); open structures use none.  The list uses fixed
// per-atom slot ranges like the cell path: a rebuild of the same batch is
// one fill pass; the first build (or an overflow) counts first.
// The partner map of the list just built (k_partner), on the side stream
// after the list kernels of the main stream; waitPartner() orders later
// readers after it.
void launchPartner(hdnnp_ctx* c, int n)
{
    if (!c->side)
    {
        hipCheck(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking),
                 "hipStreamCreate");
        hipCheck(hipEventCreateWithFlags(&c->evListDone,
                                         hipEventDisableTiming),
                 "hipEventCreate");
        hipCheck(hipEventCreateWithFlags(&c->evPartnerDone,
                                         hipEventDisableTiming),
                 "hipEventCreate");
    }
    hipCheck(hipEventRecord(c->evListDone, c->stream), "hipEventRecord");
    hipCheck(hipStreamWaitEvent(c->side, c->evListDone, 0),
             "hipStreamWaitEvent");
    hipEvent_t ev = nullptr;
    c->begin(K_PARTNER, ev, c->side);
    hipCheck(launch_nl_partner(n, c->dOff.p, c->dCount.p, c->dKeys.p,
                               c->dCellAtoms.p, c->dPartner.p, c->entryCap,
                               c->side),
             "nl partner");
    c->end(K_PARTNER, ev, c->side);
    hipCheck(hipEventRecord(c->evPartnerDone, c->side), "hipEventRecord");
    c->partnerPending = true;
}

void buildBatchNeighborList(hdnnp_ctx* c, double rc)
{
    c->waitPartner();
    int const n = (int)c->n;
    int64_t const ns = c->nstruct;
    hipStream_t s = c->stream;
    c->hPbcS.assign(3 * ns, 0);
    std::vector<BatchCell> cells((size_t)ns);
    for (int64_t st = 0; st < ns; ++st)
    {
        BatchCell& C = cells[st];
        C = BatchCell{};
        if (!c->hPeriodic[st]) continue;
        C.periodic = 1;
        for (int k = 0; k < 3; ++k)
            for (int q = 0; q < 3; ++q) C.box[k][q] = c->hBoxes[9 * st + 3 * k + q];
        inverseBox(C.box, C.invbox);
        cellWidths(C.box, rc, C.proj, C.pbc);
        int64_t nimg = 1;
        for (int k = 0; k < 3; ++k)
        {
            if (C.pbc[k] > 60)
                throw Error(HDNNP_E_CAPACITY,
                            "ERROR: Cell much smaller than the cutoff "
                            "(more than 60 periodic images per "
                            "direction).\n");
            nimg *= 2 * C.pbc[k] + 1;
            c->hPbcS[3 * st + k] = C.pbc[k];
        }
        int64_t const na = c->hStructOff[st + 1] - c->hStructOff[st];
        if (na * nimg > ((int64_t)1 << 24))
            throw Error(HDNNP_E_CAPACITY,
                        "ERROR: Too many atoms times periodic images in one "
                        "structure of the batch (limit 2^24).\n");
    }
    hipEvent_t ev0 = nullptr;
    c->begin(K_NEIGHBOR, ev0);
    c->dCells.ensure(ns);
    hipCheck(hipMemcpyAsync(c->dCells.p, cells.data(), ns * sizeof(BatchCell),
                            hipMemcpyHostToDevice, s), "cells");
    hipCheck(launch_batch_extent((int)ns, c->dR.p, c->dStructOff.p,
                                 c->dCells.p, s), "batch extent");
    c->dCount.ensure(n);
    c->dOff.ensure(n + 1);
    c->dStats.ensure(4);
    c->dBlockSums.ensure((n + 1023) / 1024 + 1);
    c->dTotal.ensure(1);
    hipCheck(hipMemsetAsync(c->dStats.p, 0, 4 * sizeof(int), s), "memset");
    int hstat[4] = {0, 0, 0, 0};
    auto readStats = [&]() {
        hipCheck(hipMemcpyAsync(hstat, c->dStats.p, 4 * sizeof(int),
                                hipMemcpyDeviceToHost, s), "stats");
        hipCheck(hipStreamSynchronize(s), "sync");
        hipCheck(hipMemsetAsync(c->dStats.p, 0, 4 * sizeof(int), s),
                 "memset");
    };
    auto fill = [&]() {
        if ((int64_t)n * c->slab > (int64_t)2000000000)
            throw Error(HDNNP_E_CAPACITY,
                        "ERROR: Neighbour list too large.\n");
        int64_t const E = std::max<int64_t>((int64_t)n * c->slab, 1);
        c->dKeys.ensure(E);
        c->dNbrJ.ensure(E);
        c->dGeo.ensure(E);
        c->dPartner.ensure(E);
        c->entryCap = E;
        hipCheck(launch_slab_starts(n, c->slab, c->dOff.p, s), "slab starts");
        hipCheck(launch_nl_batch(1, n, rc, c->dR.p, c->dStructOf.p,
                                 c->dStructOff.p, c->dCells.p, c->dCount.p,
                                 c->dStats.p, c->dOff.p, c->slab, c->dKeys.p,
                                 c->dNbrJ.p, c->dGeo.p, s), "batch nl fill");
        hipCheck(launch_max_int(c->dCount.p, n, c->dStats.p, s), "count max");
        readStats();
    };
    bool built = false;
    if (c->slab > 0 && c->slabAtoms == n)
    {
        fill();
        built = hstat[3] == 0;
    }
    if (!built)
    {
        hipCheck(launch_nl_batch(0, n, rc, c->dR.p, c->dStructOf.p,
                                 c->dStructOff.p, c->dCells.p, c->dCount.p,
                                 nullptr, nullptr, 0, nullptr, nullptr,
                                 nullptr, s), "batch nl count");
        hipCheck(launch_max_int(c->dCount.p, n, c->dStats.p, s), "count max");
        readStats();
        int const maxc = hstat[0];
        c->slab = ((maxc + maxc / 8 + 16) + 15) / 16 * 16;
        c->slabAtoms = n;
        fill();
        if (hstat[3] != 0)
            throw Error(HDNNP_E_DEVICE,
                        "ERROR: Neighbour count changed between passes.\n");
    }
    c->end(K_NEIGHBOR, ev0);
    launchPartner(c, n);
    c->slabLayout = true;
    c->Mvalid = false;
    c->maxNeigh = std::min(hstat[0], c->slab);
    c->haveNL = true;
    c->haveSF = c->haveNN = c->haveNNDeriv = false;
}

void buildNeighborList(hdnnp_ctx* c, double rc)
{
    requireStructure(c);
    c->setDevice();
    c->waitPartner();
    if (c->batched)
    {
        buildBatchNeighborList(c, rc);
        return;
    }
    NLParams& p = c->nl;
    p = NLParams{};
    p.n = (int)c->n;
    p.periodic = c->periodic ? 1 : 0;
    p.rc2 = rc * rc;
    double const margin = 1.0 + 1e-6;
    if (c->periodic)
    {
        double proj[3];
        cellWidths(c->box, rc, proj, c->pbc);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                p.box[i][j] = c->box[i][j];
                p.invbox[i][j] = c->invbox[i][j];
            }
        for (int k = 0; k < 3; ++k)
        {
            int nc = (int)std::floor(proj[k] / (rc * margin));
            nc = std::max(1, std::min(nc, 1024));
            p.nc[k] = nc;
            p.m[k] = std::max(1, (int)std::ceil(rc * nc / proj[k] * margin));
            if (p.m[k] > 60)
                throw Error(HDNNP_E_CAPACITY,
                            "ERROR: Cell much smaller than the cutoff "
                            "(more than 60 periodic images per direction).\n");
        }
    }
    else
    {
        if (!c->hostPosValid)
        {
            c->hR.resize(3 * c->n);
            hipCheck(hipMemcpy(c->hR.data(), c->dR.p, 3 * c->n * sizeof(double),
                               hipMemcpyDeviceToHost), "positions to host");
            c->hostPosValid = true;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t i = 0; i < c->n; ++i)
            for (int k = 0; k < 3; ++k)
            {
                lo[k] = std::min(lo[k], c->hR[3 * i + k]);
                hi[k] = std::max(hi[k], c->hR[3 * i + k]);
            }
        c->pbc[0] = c->pbc[1] = c->pbc[2] = 0;
        for (int k = 0; k < 3; ++k)
        {
            double const w = std::max(hi[k] - lo[k], 0.0) + 1e-9 * (1 + std::fabs(hi[k]));
            int nc = (int)std::floor(w / (rc * margin));
            nc = std::max(1, std::min(nc, 1024));
            p.nc[k] = nc;
            p.origin[k] = lo[k];
            p.invbox[k][k] = 1.0 / w;
            p.m[k] = 1;
        }
    }
    int64_t const ncell64 = (int64_t)p.nc[0] * p.nc[1] * p.nc[2];
    if (ncell64 > (int64_t)1 << 27)
        throw Error(HDNNP_E_CAPACITY, "ERROR: Too many cells.\n");
    int const ncell = (int)ncell64;
    c->ncell = ncell;
    int const n = (int)c->n;
    hipStream_t s = c->stream;
    hipEvent_t ev0 = nullptr;
    c->begin(K_NEIGHBOR, ev0);
    c->dCellOfAtom.ensure(n);
    c->dRank.ensure(n);
    c->dCellCount.ensure(ncell);
    c->dCellStart.ensure(ncell + 1);
    c->dCellAtoms.ensure(n);
    c->dCount.ensure(n);
    c->dOff.ensure(n + 1);
    c->dStats.ensure(4);
    c->dBlockSums.ensure((std::max(n, ncell) + 1023) / 1024 + 1);
    c->dTotal.ensure(1);
    hipCheck(hipMemsetAsync(c->dCellCount.p, 0, ncell * sizeof(int), s), "memset");
    hipCheck(hipMemsetAsync(c->dStats.p, 0, 4 * sizeof(int), s), "memset");
    hipCheck(launch_nl_bin(p, c->dR.p, c->dCellOfAtom.p, c->dRank.p,
                           c->dCellCount.p, s), "nl bin");
    hipCheck(scan_exclusive(c->dCellCount.p, c->dCellStart.p, ncell,
                            c->dBlockSums.p, c->dTotal.p, s), "cell scan");
    hipCheck(launch_nl_scatter(p, c->dCellOfAtom.p, c->dRank.p, c->dCellStart.p,
                               c->dCellAtoms.p, s), "nl scatter");
    // Cell-per-workgroup kernels when at most 64 neighbour cells are searched
    // and every cell's candidates fit their LDS; the per-atom kernels
    // otherwise.  Both produce the same list.
    bool cellKernels = (2 * p.m[0] + 1) * (2 * p.m[1] + 1) * (2 * p.m[2] + 1)
                     <= 64;
    int hstat[4] = {0, 0, 0, 0};
    auto readStats = [&]() {
        hipCheck(hipMemcpyAsync(hstat, c->dStats.p, 4 * sizeof(int),
                                hipMemcpyDeviceToHost, s), "stats");
        hipCheck(hipStreamSynchronize(s), "sync");
    };
    auto clearStats = [&]() {
        hipCheck(hipMemsetAsync(c->dStats.p, 0, 4 * sizeof(int), s),
                 "memset");
    };
    auto ensureEntries = [&](int64_t E) {
        if (E < 0 || E > (int64_t)2000000000)
            throw Error(HDNNP_E_CAPACITY,
                        "ERROR: Neighbour list too large.\n");
        c->dKeys.ensure(E);
        c->dNbrJ.ensure(E);
        c->dGeo.ensure(E);
        c->dPartner.ensure(E);
        c->entryCap = E;
    };
    auto slabFor = [](int maxc) {
        // 1/8 headroom over the largest count plus 16, a multiple of 16.
        return ((maxc + maxc / 8 + 16) + 15) / 16 * 16;
    };
    bool built = false;
    if (cellKernels && c->slab > 0 && c->slabAtoms == n)
    {
        // Slab build: one pass, no count kernel.  The counts, their maximum
        // and an overflow flag come back with the fill; on overflow the
        // list is rebuilt below with a range sized by the new maximum.
        ensureEntries((int64_t)n * c->slab);
        hipCheck(launch_slab_starts(n, c->slab, c->dOff.p, s), "slab starts");
        hipCheck(launch_nl_fill_cell(p, ncell, c->dR.p, c->dCellStart.p,
                                     c->dCellCount.p, c->dCellAtoms.p,
                                     c->dCount.p, c->dStats.p, c->dOff.p,
                                     c->slab, c->dKeys.p, c->dNbrJ.p,
                                     c->dGeo.p, s), "nl fill");
        hipCheck(launch_max_int(c->dCount.p, n, c->dStats.p, s), "count max");
        readStats();
        if (hstat[2] > 0) cellKernels = false;    // staging overflow
        else if (hstat[3] == 0) built = true;
        clearStats();
    }
    int maxc = hstat[0];
    if (!built)
    {
        for (int attempt = 0; attempt < 2; ++attempt)
        {
            if (cellKernels)
                hipCheck(launch_nl_count_cell(p, ncell, c->dR.p,
                                              c->dCellStart.p, c->dCellCount.p,
                                              c->dCellAtoms.p, c->dCount.p,
                                              c->dStats.p, s), "nl count");
            else
                hipCheck(launch_nl_count(p, c->dR.p, c->dElemAtom.p,
                                         c->dCellOfAtom.p, c->dCellStart.p,
                                         c->dCellCount.p, c->dCellAtoms.p,
                                         c->dCount.p, c->dStats.p, s),
                         "nl count");
            hipCheck(launch_max_int(c->dCount.p, n, c->dStats.p, s),
                     "count max");
            readStats();
            if (!(cellKernels && hstat[2] > 0)) break;
            cellKernels = false;     // a cell overflowed the staging: redo
            clearStats();
        }
        maxc = hstat[0];
        clearStats();
        if (cellKernels)
        {
            c->slab = slabFor(maxc);
            c->slabAtoms = n;
            if ((int64_t)n * c->slab > (int64_t)2000000000)
                throw Error(HDNNP_E_CAPACITY,
                            "ERROR: Neighbour list too large.\n");
            ensureEntries((int64_t)n * c->slab);
            hipCheck(launch_slab_starts(n, c->slab, c->dOff.p, s),
                     "slab starts");
            hipCheck(launch_nl_fill_cell(p, ncell, c->dR.p, c->dCellStart.p,
                                         c->dCellCount.p, c->dCellAtoms.p,
                                         c->dCount.p, c->dStats.p, c->dOff.p,
                                         c->slab, c->dKeys.p, c->dNbrJ.p,
                                         c->dGeo.p, s), "nl fill");
            hipCheck(launch_max_int(c->dCount.p, n, c->dStats.p, s),
                     "count max");
        }
        else
        {
            hipCheck(scan_exclusive(c->dCount.p, c->dOff.p, n,
                                    c->dBlockSums.p, c->dTotal.p, s),
                     "count scan");
            int total = 0;
            hipCheck(hipMemcpyAsync(&total, c->dTotal.p, sizeof(int),
                                    hipMemcpyDeviceToHost, s), "total");
            hipCheck(hipStreamSynchronize(s), "sync");
            int keyCap = 64;
            while (keyCap < maxc) keyCap <<= 1;
            if (keyCap > 8192)
                throw Error(HDNNP_E_CAPACITY,
                            "ERROR: More than 8192 neighbours for one "
                            "atom.\n");
            ensureEntries(std::max(total, 1));
            hipCheck(launch_nl_fill(p, c->dR.p, c->dElemAtom.p,
                                    c->dCellOfAtom.p, c->dCellStart.p,
                                    c->dCellCount.p, c->dCellAtoms.p,
                                    c->dStats.p, c->dOff.p, keyCap, c->dKeys.p,
                                    c->dNbrJ.p, c->dGeo.p, s), "nl fill");
            c->M = total;
        }
    }
    c->slabLayout = cellKernels;
    c->Mvalid = !cellKernels;
    c->end(K_NEIGHBOR, ev0);
    launchPartner(c, n);
    // Slot capacity of the SF and force kernels: the largest count (the
    // slab bounds it; a rebuild above raised the slab if it was exceeded).
    c->maxNeigh = cellKernels ? std::min(maxc, c->slab) : maxc;
    c->haveNL = true;
    c->haveSF = c->haveNN = c->haveNNDeriv = false;
}

// Total neighbour-list entries (summed from the per-atom counts once per
// build when the slab layout is used).
int64_t totalEntries(hdnnp_ctx* c)
{
    if (!c->Mvalid)
    {
        std::vector<int> cnt((size_t)c->n);
        if (c->n > 0)
            hipCheck(hipMemcpy(cnt.data(), c->dCount.p, c->n * sizeof(int),
                               hipMemcpyDeviceToHost), "counts");
        int64_t t = 0;
        for (int v : cnt) t += v;
        c->M = t;
        c->Mvalid = true;
    }
    return c->M;
}

// Entries per atom of the pair record for an LDS capacity of cap
// neighbours: the group counts, then every unordered pair of the staged
// list up to 4096 (groups beyond that are walked again); none when the
// staged indices do not fit a byte.
int pairRecordCap(int cap)
{
    if (cap > 256) return 0;
    int64_t const all = (int64_t)cap * (cap - 1) / 2;
    return kRecordGroups + (int)std::min<int64_t>(4096, (all + 63) / 64 * 64);
}

// Whether SF evaluations of this context record the passing pairs for the
// force kernel: the mixed path, specialised or generic, of every model
// whose element chunks fit the record's group counts.
bool recordsPairs(hdnnp_ctx* c, bool raw)
{
    if (!c->pairRecord || raw || c->precision == HDNNP_PRECISION_FP64)
        return false;
    for (auto const& e : c->flat.elem)
        if (e.ng > kRecordGroups) return false;
    return pairRecordCap(std::max(c->maxNeigh, 1)) > 0;
}

void computeSF(hdnnp_ctx* c, bool raw)
{
    requireNL(c);
    c->setDevice();
    int const cap = std::max(c->maxNeigh, 1);
    int const fp32 = (!raw && c->precision != HDNNP_PRECISION_FP64) ? 1 : 0;
    if (sf_forward_lds_bytes(cap, fp32) > 65536
        || (c->T.ncgroup > 0 && sf_compact_lds_bytes(c->T, cap) > 65536))
        throw Error(HDNNP_E_CAPACITY,
                    "ERROR: Too many neighbours per atom for the symmetry-"
                    "function kernel's LDS staging.\n");
    c->dG.ensure((size_t)c->n * c->ld);
    SFArgs a{};
    a.T = c->T;
    a.n = (int)c->n;
    a.off = c->dOff.p;
    a.cnt = c->dCount.p;
    a.geo = c->dGeo.p;
    a.nbrJ = c->dNbrJ.p;
    a.elem = c->dElemAtom.p;
    a.G = c->dG.p;
    a.ldG = c->ld;
    a.cap = cap;
    a.raw = raw ? 1 : 0;
    a.fp32 = fp32;
    bool const rec = recordsPairs(c, raw);
    if (rec)
    {
        int const pc = pairRecordCap(cap);
        c->dPairs.ensure(std::max<size_t>((size_t)c->n * (size_t)pc, 1));
        a.pairs = c->dPairs.p;
        a.pcap = pc;
    }
    c->pairsValid = rec;
    hipEvent_t ev = nullptr;
    c->begin(K_SYMFUNC, ev);
    // Raw (unscaled) evaluations use the same kernels as scaled ones, so
    // that statistics and later evaluations of a data set see the same G
    // bit for bit.
    PlanKernels const pk = c->planFor(fp32);
    if (pk.sf && c->n > 0)
    {
        void* params[] = {&a};
        hipCheck(hipModuleLaunchKernel(pk.sf, (unsigned)c->n, 1, 1, 64, 1, 1,
                                       (unsigned)sf_forward_lds_bytes(cap,
                                                                      fp32),
                                       c->stream, params, nullptr),
                 "symmetry functions (specialised)");
    }
    else hipCheck(launch_sf_forward(a, c->stream), "symmetry functions");
    hipCheck(launch_sf_compact(a, c->stream), "compact symmetry functions");
    c->end(K_SYMFUNC, ev);
    c->haveSF = !raw;
    c->haveNN = c->haveNNDeriv = false;
}

void computeNN(hdnnp_ctx* c, bool deriv)
{
    if (!c->haveSF)
        throw Error(HDNNP_E_INVALID,
                    "ERROR: Symmetry functions not calculated.\n");
    c->setDevice();
    int const n = (int)c->n;
    c->dE.ensure(n);
    c->dEdG.ensure((size_t)n * c->ld);
    size_t need = 0;
    for (size_t e = 0; e < c->flat.elem.size(); ++e)
    {
        int plane = 0;
        for (int l = 1; l < c->flat.elem[e].nlayers; ++l)
            plane += c->flat.elem[e].width[l];
        need = std::max(need, (size_t)plane * 3 * c->hAtomsOfElem[e].size());
    }
    c->dScratch.ensure(need);
    hipEvent_t ev = nullptr;
    c->begin(K_NETWORK, ev);
    for (size_t e = 0; e < c->flat.elem.size(); ++e)
    {
        NNArgs a{};
        a.T = c->T;
        a.ed = c->flat.elem[e];
        a.mode = c->precision == HDNNP_PRECISION_MIXED_BF16 ? 3
               : c->precision == HDNNP_PRECISION_MIXED16 ? 2
               : c->precision == HDNNP_PRECISION_MIXED ? 1 : 0;
        a.tiles = c->nnTiles;
        while (a.tiles > 1 && nn_mfma_lds_bytes(a.ed, a.mode, a.tiles) > 65536)
            a.tiles /= 2;
        if (nn_mfma_lds_bytes(a.ed, a.mode, 1) > 65536) a.mode = -1;
        a.e = (int)e;
        a.atoms = c->dAtomsOfElem[e].p;
        a.na = (int)c->hAtomsOfElem[e].size();
        a.G = c->dG.p;
        a.ldG = c->ld;
        a.E = c->dE.p;
        a.dEdG = c->dEdG.p;
        a.scratch = c->dScratch.p;
        a.deriv = deriv ? 1 : 0;
        hipCheck(launch_nn(a, c->stream), "networks");
    }
    c->end(K_NETWORK, ev);
    c->haveNN = true;
    c->haveNNDeriv = deriv;
}

void launchEnergy(hdnnp_ctx* c)
{
    c->dEnergy.ensure(1);
    hipCheck(launch_energy_sum(c->dE.p, (int)c->n, c->dEnergy.p, c->stream),
             "energy");
}

double computeEnergy(hdnnp_ctx* c)
{
    if (!c->haveNN)
        throw Error(HDNNP_E_INVALID, "ERROR: Networks not evaluated.\n");
    c->setDevice();
    launchEnergy(c);
    double e = 0.0;
    hipCheck(hipMemcpyAsync(&e, c->dEnergy.p, sizeof(double),
                            hipMemcpyDeviceToHost, c->stream), "energy copy");
    c->sync();
    c->energy = e;
    return e;
}

void launchGather(hdnnp_ctx* c, int mode);

void computeForces(hdnnp_ctx* c, int mode, bool wait = true)
{
    if (!c->haveNNDeriv)
        throw Error(HDNNP_E_INVALID,
                    "ERROR: dE/dG not available (call the networks with "
                    "derivatives).\n");
    if (mode != HDNNP_FORCES_REFERENCE && mode != HDNNP_FORCES_CONSISTENT)
        throw Error(HDNNP_E_INVALID, "ERROR: Unknown force mode.\n");
    c->setDevice();
    int const cap = std::max(c->maxNeigh, 1);
    int const fp32 = c->precision != HDNNP_PRECISION_FP64 ? 1 : 0;
    size_t const shm = force_backward_lds_bytes(cap, fp32);
    if (shm > 65536
        || (c->T.ncgroup > 0 && force_compact_lds_bytes(c->T, cap) > 65536))
        throw Error(HDNNP_E_CAPACITY,
                    "ERROR: Too many neighbours per atom for the force "
                    "kernel's LDS staging.\n");
    c->dCslot.ensure(3 * (size_t)std::max<int64_t>(c->entryCap, 1));
    c->dCself.ensure(3 * (size_t)c->n);
    c->dF.ensure(3 * (size_t)c->n);
    ForceArgs a{};
    a.T = c->T;
    a.n = (int)c->n;
    a.off = c->dOff.p;
    a.cnt = c->dCount.p;
    a.geo = c->dGeo.p;
    a.nbrJ = c->dNbrJ.p;
    a.elem = c->dElemAtom.p;
    a.dEdG = c->dEdG.p;
    a.ldG = c->ld;
    a.cslot = c->dCslot.p;
    a.cself = c->dCself.p;
    a.cap = cap;
    a.fp32 = fp32;
    if (c->pairsValid && fp32)
    {
        a.pairs = c->dPairs.p;
        a.pcap = pairRecordCap(cap);
    }
    hipEvent_t ev = nullptr;
    c->begin(K_FORCE, ev);
    PlanKernels const pk = c->planFor(fp32);
    if (pk.force && c->n > 0)
    {
        void* params[] = {&a};
        hipCheck(hipModuleLaunchKernel(pk.force, (unsigned)c->n, 1, 1, 64, 1,
                                       1, (unsigned)shm, c->stream, params,
                                       nullptr),
                 "force backward (specialised)");
    }
    else hipCheck(launch_force_backward(a, c->stream), "force backward");
    hipCheck(launch_force_compact(a, c->stream), "compact force backward");
    c->end(K_FORCE, ev);
    launchGather(c, mode);
    if (wait) c->sync();
}

// F = -(cself + partner slots): the gather of the plain list, or of the MD
// list through the skin partners.  Caller-given lists are scattered on
// request (hdnnp_get_scattered_forces) instead.
void launchGather(hdnnp_ctx* c, int mode)
{
    hipEvent_t ev2 = nullptr;
    if (!c->externalNL && c->md.on)
    {
        auto& m = c->md;
        c->begin(K_GATHER, ev2);
        hipCheck(launch_gather_md((int)c->n, c->dOff.p, m.skinCnt.p,
                                  m.newIdx.p, m.skinPartner.p, m.skinKeys.p,
                                  c->dCslot.p, c->dCself.p, c->dCellAtoms.p,
                                  mode, c->entryCap, c->dF.p, m.flags.p + 3,
                                  c->stream), "force gather (MD)");
        c->end(K_GATHER, ev2);
    }
    else if (!c->externalNL)
    {
        c->waitPartner();
        c->begin(K_GATHER, ev2);
        hipCheck(launch_force_gather((int)c->n, c->dOff.p, c->dCount.p,
                                     c->dNbrJ.p, c->dPartner.p, c->dCslot.p,
                                     c->dCself.p, c->dCellAtoms.p, mode,
                                     c->entryCap, c->dF.p, c->stream),
                 "force gather");
        c->end(K_GATHER, ev2);
    }
}


// ------------------------------------------------------------------ MD --

void mdDropGraphs(hdnnp_ctx* c);

// Calls that replace the structure, the positions or the list end an MD
// run (hdnnp_md_init starts a new one).
void mdEnd(hdnnp_ctx* c)
{
    if (!c->md.on) return;
    mdDropGraphs(c);
    c->md.on = false;
}

void mdDropGraphs(hdnnp_ctx* c)
{
    auto& m = c->md;
    if (m.eA) (void)hipGraphExecDestroy(m.eA);
    if (m.eB) (void)hipGraphExecDestroy(m.eB);
    if (m.gA) (void)hipGraphDestroy(m.gA);
    if (m.gB) (void)hipGraphDestroy(m.gB);
    m.eA = m.eB = nullptr;
    m.gA = m.gB = nullptr;
}

// Kick, drift and the filter of the skin list; the two flags (moved more
// than half the skin, filtered list over capacity) to the host.
void mdStepA(hdnnp_ctx* c)
{
    auto& m = c->md;
    int const n = (int)c->n;
    hipStream_t s = c->stream;
    hipCheck(launch_md_clear_flags(m.flags.p, s), "md flags");
    hipCheck(launch_md_kick(n, m.v.p, c->dF.p, c->dElemAtom.p, m.accel.p,
                            0.5 * m.dt, s), "md kick");
    double const half = 0.5 * m.skin;
    hipCheck(launch_md_drift(n, c->dR.p, m.v.p, m.r0.p, m.dt, half * half,
                             m.flags.p, s), "md drift");
    hipCheck(launch_md_filter(n, m.cell, m.rc * m.rc, c->dR.p, c->dOff.p,
                              m.skinCnt.p, m.skinKeys.p, c->dCount.p,
                              c->dNbrJ.p, c->dGeo.p, m.newIdx.p, m.cap,
                              m.flags.p, s), "md filter");
    hipCheck(hipMemcpyAsync(m.hFlags, m.flags.p, 4 * sizeof(int),
                            hipMemcpyDeviceToHost, s), "md flags");
}

// SF, networks, energy, forces on the filtered list (the gather follows
// partners through the skin list).
void mdForces(hdnnp_ctx* c, bool wait)
{
    auto& m = c->md;
    c->haveNL = true;
    computeSF(c, false);
    computeNN(c, true);
    launchEnergy(c);
    computeForces(c, m.forceMode, wait);
}

// Forces, second kick, energy record.
void mdStepB(hdnnp_ctx* c)
{
    auto& m = c->md;
    mdForces(c, false);
    hipCheck(launch_md_kick((int)c->n, m.v.p, c->dF.p, c->dElemAtom.p,
                            m.accel.p, 0.5 * m.dt, c->stream), "md kick");
    hipCheck(launch_md_record((int)c->n, m.v.p, c->dElemAtom.p, m.mass.p,
                              m.part.p, c->dEnergy.p, 0.5 / m.ftoa,
                              m.counter.p, m.recCap, m.epot.p, m.ekin.p,
                              c->stream), "md record");
}

// Wrap the atoms into the cell, build the skin list, filter it at the
// current positions and size the capacity of the filtered list.
void mdRebuild(hdnnp_ctx* c)
{
    auto& m = c->md;
    int const n = (int)c->n;
    hipStream_t s = c->stream;
    mdDropGraphs(c);
    hipCheck(hipMemsetAsync(m.flags.p + 4, 0, sizeof(int), s), "md flags");
    hipCheck(launch_md_wrap(n, c->dR.p, m.cell, m.flags.p + 4, s), "md wrap");
    c->hostPosValid = false;
    buildNeighborList(c, m.rc + m.skin);
    c->waitPartner();   // before any graph is captured on the main stream
    std::swap(c->dKeys, m.skinKeys);
    std::swap(c->dPartner, m.skinPartner);
    std::swap(c->dCount, m.skinCnt);
    size_t const E = (size_t)std::max<int64_t>(c->entryCap, 1);
    c->dKeys.ensure(E);
    c->dPartner.ensure(E);
    c->dCount.ensure(n);
    m.newIdx.ensure(E);
    hipCheck(hipMemcpyAsync(m.r0.p, c->dR.p, 3 * (size_t)n * sizeof(double),
                            hipMemcpyDeviceToDevice, s), "md r0");
    hipCheck(launch_md_clear_flags(m.flags.p, s), "md flags");
    hipCheck(launch_md_filter(n, m.cell, m.rc * m.rc, c->dR.p, c->dOff.p,
                              m.skinCnt.p, m.skinKeys.p, c->dCount.p,
                              c->dNbrJ.p, c->dGeo.p, m.newIdx.p, INT32_MAX,
                              m.flags.p, s), "md filter");
    int maxc = 0;
    hipCheck(launch_max_int(c->dCount.p, n, c->dStats.p, s), "md max");
    hipCheck(hipMemcpyAsync(&maxc, c->dStats.p, sizeof(int),
                            hipMemcpyDeviceToHost, s), "md max");
    hipCheck(hipMemcpyAsync(m.hFlags + 4, m.flags.p + 4, sizeof(int),
                            hipMemcpyDeviceToHost, s), "md wrap flag");
    hipCheck(hipStreamSynchronize(s), "sync");
    if (m.hFlags[4]) m.wrapMoves += 1;
    m.cap = std::min(c->maxNeigh, maxc + maxc / 8 + 8);
    c->maxNeigh = m.cap;
    c->dCslot.ensure(3 * E);
    // The pair record, sized here so that graph captures of the steps
    // allocate nothing.
    if (recordsPairs(c, false))
        c->dPairs.ensure(std::max<size_t>(
            (size_t)n * (size_t)pairRecordCap(m.cap), 1));
    c->Mvalid = false;
    m.rebuilds += 1;
}

// A coordinate went non-finite or ran away: the run has become unstable
// (time step too large for the forces); end it with an error.
void mdCheckStable(hdnnp_ctx* c)
{
    if (!c->md.hFlags[2]) return;
    mdEnd(c);
    throw Error(HDNNP_E_INVALID, "ERROR: MD became unstable (a coordinate "
                                 "is not finite or beyond 1e7); the time "
                                 "step is too large for the forces.\n");
}

template <class F>
void mdReplay(hdnnp_ctx* c, hipGraph_t& g, hipGraphExec_t& e, F&& body)
{
    if (!c->md.useGraph)
    {
        body();
        return;
    }
    if (!e)
    {
        c->capturing = true;
        hipError_t const rc = hipStreamBeginCapture(c->stream,
                                                    hipStreamCaptureModeThreadLocal);
        if (rc != hipSuccess)
        {
            c->capturing = false;
            hipCheck(rc, "hipStreamBeginCapture");
        }
        try { body(); }
        catch (...)
        {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(c->stream, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            c->capturing = false;
            throw;
        }
        c->capturing = false;
        hipCheck(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
        hipCheck(hipGraphInstantiate(&e, g, nullptr, nullptr, 0),
                 "hipGraphInstantiate");
    }
    hipCheck(hipGraphLaunch(e, c->stream), "hipGraphLaunch");
}

void mdStep(hdnnp_ctx* c)
{
    auto& m = c->md;
    mdReplay(c, m.gA, m.eA, [&]() { mdStepA(c); });
    hipCheck(hipStreamSynchronize(c->stream), "sync");
    mdCheckStable(c);
    if (m.hFlags[0] || m.hFlags[1]) mdRebuild(c);
    mdReplay(c, m.gB, m.eB, [&]() { mdStepB(c); });
    m.steps += 1;
}

template <class T>
void copyOut(hdnnp_ctx* c, T* dst, const T* src, size_t count)
{
    if (!dst || !count) return;
    hipCheck(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost,
                            c->stream), "copy to host");
}

}  // namespace

// =========================================================== C ABI
namespace {

// Host copies of the per-atom first entries and counts, and the compact
// CSR offsets (n + 1) of the list in atom order.
struct EntryMap
{
    std::vector<int> off, cnt;
    std::vector<int64_t> csr;
};

EntryMap entryMap(hdnnp_ctx* c)
{
    int const n = (int)c->n;
    EntryMap m;
    m.off.resize(n + 1);
    m.cnt.resize(n);
    copyOut(c, m.off.data(), c->dOff.p, n + 1);
    copyOut(c, m.cnt.data(), c->dCount.p, n);
    c->sync();
    m.csr.resize(n + 1);
    m.csr[0] = 0;
    for (int i = 0; i < n; ++i) m.csr[i + 1] = m.csr[i] + m.cnt[i];
    return m;
}

// Copies a per-entry device array (w values per entry) into compact atom
// order.
template <class T>
void compactEntries(hdnnp_ctx* c, EntryMap const& m, T const* dev, int w,
                    T* out)
{
    std::vector<T> all((size_t)c->entryCap * w);
    copyOut(c, all.data(), dev, all.size());
    c->sync();
    for (size_t i = 0; i < m.cnt.size(); ++i)
        std::copy(all.begin() + (size_t)m.off[i] * w,
                  all.begin() + ((size_t)m.off[i] + m.cnt[i]) * w,
                  out + (size_t)m.csr[i] * w);
}

}  // namespace

extern "C" {

const char* hdnnp_last_error(void) { return g_err.c_str(); }

int hdnnp_model_load(const char* settings_file, const char* scaling_file,
                     const char* weights_fmt, hdnnp_model** out)
{
    return guarded([&]() {
        if (!out || !settings_file || !weights_fmt)
            throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        std::unique_ptr<hdnnp_model> m(new hdnnp_model);
        m->m.load(settings_file, scaling_file ? scaling_file : "",
                  weights_fmt);
        *out = m.release();
    });
}

void hdnnp_model_free(hdnnp_model* m) { delete m; }
int hdnnp_model_num_elements(const hdnnp_model* m) { return m->m.numElements(); }
int hdnnp_model_atomic_number(const hdnnp_model* m, int e)
{
    if (e < 0 || e >= m->m.numElements()) return -1;
    return m->m.el[e].Z;
}
int hdnnp_model_num_sf(const hdnnp_model* m, int e)
{
    if (e < 0 || e >= m->m.numElements()) return -1;
    return (int)m->m.el[e].sf.size();
}
double hdnnp_model_max_cutoff(const hdnnp_model* m)
{
    return m->m.maxCutoffRadius;
}
int hdnnp_model_element_index(const hdnnp_model* m, const char* symbol)
{
    return symbol ? m->m.elementIndex(symbol) : -1;
}
int hdnnp_model_normalization(const hdnnp_model* m, double* me, double* ce,
                              double* cl)
{
    if (me) *me = m->m.meanEnergy;
    if (ce) *ce = m->m.convEnergy;
    if (cl) *cl = m->m.convLength;
    return m->m.normalize ? 1 : 0;
}
double hdnnp_model_energy_offset(const hdnnp_model* m, int e)
{
    if (e < 0 || e >= m->m.numElements()) return 0.0;
    return m->m.el[e].energyOffset;
}
int hdnnp_model_sf_table(const hdnnp_model* m, int e, double* out)
{
    return guarded([&]() {
        if (e < 0 || e >= m->m.numElements())
            throw Error(HDNNP_E_INVALID, "ERROR: Element out of range.\n");
        auto const& sf = m->m.el[e].sf;
        for (size_t i = 0; i < sf.size(); ++i)
        {
            double* o = out + 12 * i;
            o[0] = sf[i].type; o[1] = sf[i].e1; o[2] = sf_radial(sf[i].type) ? -1 : sf[i].e2;
            o[3] = sf[i].eta; o[4] = sf[i].rs; o[5] = sf[i].lambda;
            o[6] = sf[i].zeta; o[7] = sf[i].rc; o[8] = sf[i].cutoffType;
            o[9] = sf[i].cutoffAlpha; o[10] = sf[i].scalingFactor;
            o[11] = (double)sf[i].line;
        }
    });
}
int hdnnp_model_set_scaling(hdnnp_model* m, int e, int type, double smin,
                            double smax, const double* stats)
{
    return guarded([&]() {
        if (e < 0 || e >= m->m.numElements() || type < 0 || type > 4 || !stats)
            throw Error(HDNNP_E_INVALID, "ERROR: Invalid scaling arguments.\n");
        m->m.setScaling(e, type, smin, smax, stats);
    });
}

int hdnnp_ctx_create(const hdnnp_model* m, int device, hdnnp_ctx** out)
{
    return guarded([&]() {
        if (!m || !out) throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        int ndev = 0;
        hipCheck(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
        if (device < 0 || device >= ndev)
            throw Error(HDNNP_E_DEVICE, "ERROR: No such HIP device.\n");
        std::unique_ptr<hdnnp_ctx> c(new hdnnp_ctx);
        c->model = m;
        c->device = device;
        c->setDevice();
        hipCheck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking),
                 "hipStreamCreate");
        c->uploadModel();
        *out = c.release();
    });
}

void hdnnp_ctx_destroy(hdnnp_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending)
    {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (c->side) (void)hipStreamSynchronize(c->side);
    for (auto e : c->eventPool) (void)hipEventDestroy(e);
    if (c->evListDone) (void)hipEventDestroy(c->evListDone);
    if (c->evPartnerDone) (void)hipEventDestroy(c->evPartnerDone);
    c->dElem.release(); c->dCfg.release(); c->dGroup.release();
    c->dMember.release(); c->dRun.release(); c->dScale.release(); c->dPairGroups.release();
    c->dCompGroup.release(); c->dCompMember.release();
    c->dPairCount.release(); c->dWeights.release();
    c->dWeights32.release();
    for (auto& b : c->dAtomsOfElem) b.release();
    c->dR.release(); c->dElemAtom.release();
    c->dCellOfAtom.release(); c->dRank.release(); c->dCellCount.release();
    c->dCellStart.release(); c->dCellAtoms.release(); c->dCount.release();
    c->dOff.release(); c->dStats.release(); c->dBlockSums.release();
    c->dTotal.release(); c->dKeys.release(); c->dNbrJ.release();
    c->dPartner.release(); c->dGeo.release();
    c->dG.release(); c->dE.release(); c->dEdG.release(); c->dScratch.release();
    c->dCslot.release(); c->dCself.release(); c->dF.release();
    c->dEnergy.release(); c->dStatOut.release(); c->dEw.release();
    c->dStructOf.release(); c->dStructOff.release(); c->dCells.release();
    c->dEnergySeg.release();
    mdDropGraphs(c);
    {
        auto& m = c->md;
        m.v.release(); m.r0.release(); m.accel.release(); m.mass.release();
        m.part.release(); m.epot.release(); m.ekin.release();
        m.flags.release(); m.counter.release(); m.skinPartner.release();
        m.skinCnt.release(); m.newIdx.release();
        m.skinKeys.release();
        c->tU.release(); c->tD2.release(); c->tPart.release();
        c->tGrad.release(); c->tM.release(); c->tV.release();
        c->tFref.release(); c->tLoss.release();
        c->jAtoms.release(); c->jB0.release(); c->jBN.release();
        c->jBE.release(); c->jGrp0.release(); c->jBase.release();
        c->jNwe.release(); c->jPart.release(); c->jOut.release();
        c->kfP.release(); c->kfX.release(); c->kfA.release(); c->kfH.release();
        c->mts.d2.release(); c->mts.sum.release(); c->mts.epot.release();
        c->mts.ekin.release(); c->mts.counter.release();
        c->mts.full = nullptr;
        if (m.hFlags) (void)hipHostFree(m.hFlags);
        m.hFlags = nullptr;
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    delete c;
}

int hdnnp_ctx_set_precision(hdnnp_ctx* c, int precision)
{
    return guarded([&]() {
        if (precision != HDNNP_PRECISION_FP64
            && precision != HDNNP_PRECISION_MIXED
            && precision != HDNNP_PRECISION_MIXED16
            && precision != HDNNP_PRECISION_MIXED_BF16)
            throw Error(HDNNP_E_INVALID, "ERROR: Unknown precision.\n");
        if (precision != c->precision)
            c->haveSF = c->haveNN = c->haveNNDeriv = false;
        c->precision = precision;
    });
}

int hdnnp_ctx_set_specialized(hdnnp_ctx* c, int enable)
{
    return guarded([&]() {
        if (!c) throw Error(HDNNP_E_INVALID, "ERROR: Null context.\n");
        if ((enable != 0) != c->specialized)
            c->haveSF = c->haveNN = c->haveNNDeriv = false;
        c->specialized = enable != 0;
    });
}

int hdnnp_ctx_set_pair_record(hdnnp_ctx* c, int enable)
{
    return guarded([&]() {
        if (c->pairRecord != (enable != 0))
        {
            c->pairRecord = enable != 0;
            c->haveSF = c->haveNN = c->haveNNDeriv = false;
            c->pairsValid = false;
        }
    });
}

int hdnnp_ctx_pair_record_active(hdnnp_ctx* c, int* active)
{
    return guarded([&]() {
        if (!active) throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        *active = recordsPairs(c, false) ? 1 : 0;
    });
}

int hdnnp_ctx_specialized_active(hdnnp_ctx* c, int* active)
{
    return guarded([&]() {
        if (!c || !active) throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        c->setDevice();
        PlanKernels const pk = c->planFor(c->precision != HDNNP_PRECISION_FP64);
        *active = (pk.sf && pk.force) ? 1 : 0;
    });
}

int hdnnp_ctx_update_model(hdnnp_ctx* c)
{
    return guarded([&]() {
        c->setDevice();
        hipCheck(hipStreamSynchronize(c->stream), "sync");
        c->uploadModel();
        c->haveSF = c->haveNN = c->haveNNDeriv = false;
    });
}

namespace {

// A tuning choice, kept per (model shape, precision, device) for the
// process: contexts of the same model adopt it without sweeping again.
struct TuneChoice
{
    int chunk, sfWaves, forceWaves, nnTiles;
};
std::mutex g_tuneMutex;
std::map<std::string, TuneChoice> g_tuned;
constexpr int kTuneCols = 8;   // columns of a sweep row

// The model's shape: the generated member code of its default tables (SF
// types, groups, cutoffs) and the network widths, not its weights.
std::string tuneKey(hdnnp_ctx* c)
{
    Model::Flat const F = c->model->m.flatten();
    std::string k = std::to_string(c->device) + ":"
                  + std::to_string(c->precision) + ":";
    for (auto const& e : F.elem)
    {
        for (int l = 0; l < e.nlayers; ++l)
            k += std::to_string(e.width[l]) + ",";
        k += ";";
    }
    return k + plan_source(F);
}

void applyTuning(hdnnp_ctx* c, TuneChoice const& t)
{
    if (t.chunk != c->chunk)
    {
        hipCheck(hipStreamSynchronize(c->stream), "sync");
        c->chunk = t.chunk;
        c->uploadModel();
    }
    if (t.sfWaves != c->tune.sfWaves || t.forceWaves != c->tune.forceWaves)
    {
        c->tune.sfWaves = t.sfWaves;
        c->tune.forceWaves = t.forceWaves;
        c->planTried[0] = c->planTried[1] = false;
    }
    c->nnTiles = t.nnTiles;
    c->haveSF = c->haveNN = c->haveNNDeriv = false;
    c->pairsValid = false;
}

// Mean SF, force and network times (ms) of `reps` evaluations after one
// untimed evaluation (which also compiles the specialised kernels).
void timeEvaluations(hdnnp_ctx* c, int mode, int reps, double* ms)
{
    computeSF(c, false);
    computeNN(c, true);
    computeEnergy(c);
    computeForces(c, mode);
    c->sync();
    for (int k = 0; k < K_NUM; ++k) c->profMs[k] = 0.0, c->profCalls[k] = 0;
    bool const prof = c->prof;
    c->prof = true;
    for (int r = 0; r < reps; ++r)
    {
        computeSF(c, false);
        computeNN(c, true);
        computeEnergy(c);
        computeForces(c, mode);
    }
    c->sync();
    c->prof = prof;
    int const cls[3] = {K_SYMFUNC, K_FORCE, K_NETWORK};
    for (int q = 0; q < 3; ++q)
        ms[q] = c->profCalls[cls[q]]
              ? c->profMs[cls[q]] / c->profCalls[cls[q]] : 0.0;
    for (int k = 0; k < K_NUM; ++k) c->profMs[k] = 0.0, c->profCalls[k] = 0;
}

}  // namespace

int hdnnp_ctx_autotune(hdnnp_ctx* c, int force_mode, int flags, int reps,
                       double* table, int cap, int* nrows)
{
    return guarded([&]() {
        if (!c || !nrows || (cap > 0 && !table))
            throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        if (reps < 1)
            throw Error(HDNNP_E_INVALID, "ERROR: reps must be positive.\n");
        if (force_mode < 0 || force_mode > 1)
            throw Error(HDNNP_E_INVALID, "ERROR: Unknown force mode.\n");
        *nrows = 0;
        c->setDevice();
        mdEnd(c);
        std::string const key = tuneKey(c);
        if (!(flags & HDNNP_TUNE_SWEEP))
        {
            std::lock_guard<std::mutex> lock(g_tuneMutex);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end())
            {
                applyTuning(c, it->second);
                return;
            }
        }
        requireStructure(c);
        if (!c->haveNL && !c->externalNL)
            buildNeighborList(c, c->T.maxCutoff);
        requireNL(c);
        int const fp32 = c->precision != HDNNP_PRECISION_FP64;
        std::vector<double> rows;
        auto addRow = [&](TuneChoice const& t, double const* ms) {
            double const r[kTuneCols] = {
                (double)t.chunk, (double)t.sfWaves, (double)t.forceWaves,
                (double)t.nnTiles, ms[0], ms[1], ms[2], 0.0};
            rows.insert(rows.end(), r, r + kTuneCols);
        };
        // Stage 1: member chunk x waves bounds.  SF and force kernels are
        // timed separately, so candidate q pairs the q-th SF bound with the
        // q-th force bound (one compilation for both).
        int const chunks[3] = {4, 8, 16};
        int const sfW[6] = {0, 4, 5, 6, 7, 8};
        int const fW[6] = {0, 2, 3, 4, 5, 6};
        TuneChoice best{c->chunk, -1, -1, 1};
        double bestTotal = 1e300;
        size_t lastGroups = 0;
        for (int ch : chunks)
        {
            size_t ng = 0;
            try
            {
                ng = c->model->m.flatten(ch).group.size();
            }
            catch (std::exception const&)
            {
                continue;   // more chunks per element pair than supported
            }
            if (ng == lastGroups) continue;   // same partition as before
            lastGroups = ng;
            double bestSf = 1e300, bestF = 1e300;
            int wSf = -1, wF = -1;
            for (int q = 0; q < 6; ++q)
            {
                TuneChoice const t{ch, sfW[q], fW[q], 1};
                applyTuning(c, t);
                if (c->specialized)
                {
                    PlanKernels const pk = c->planFor(fp32);
                    if (!pk.sf || !pk.force)
                    {
                        if (q > 0) break;   // generic kernels: no bounds
                    }
                }
                double ms[3];
                timeEvaluations(c, force_mode, reps, ms);
                addRow(t, ms);
                if (ms[0] < bestSf) bestSf = ms[0], wSf = sfW[q];
                if (ms[1] < bestF) bestF = ms[1], wF = fW[q];
                if (!c->specialized) break;
            }
            if (bestSf + bestF < bestTotal)
            {
                bestTotal = bestSf + bestF;
                best = TuneChoice{ch, wSf, wF, 1};
            }
        }
        // Stage 2: network tiles per wave at the chosen kernels.
        double bestNN = 1e300;
        int tiles = 1;
        for (int nt : {1, 2, 4})
        {
            TuneChoice t = best;
            t.nnTiles = nt;
            applyTuning(c, t);
            double ms[3];
            timeEvaluations(c, force_mode, reps, ms);
            addRow(t, ms);
            if (ms[2] < bestNN) bestNN = ms[2], tiles = nt;
        }
        best.nnTiles = tiles;
        applyTuning(c, best);
        auto same = [&](size_t r) {
            return (int)rows[r] == best.chunk
                && (int)rows[r + 3] == best.nnTiles;
        };
        for (size_t r = 0; r < rows.size(); r += kTuneCols)
        {
            if (same(r) && (int)rows[r + 1] == best.sfWaves)
                rows[r + kTuneCols - 1] += 1.0;
            if (same(r) && (int)rows[r + 2] == best.forceWaves)
                rows[r + kTuneCols - 1] += 2.0;
        }
        {
            std::lock_guard<std::mutex> lock(g_tuneMutex);
            g_tuned[key] = best;
        }
        int const nr = (int)(rows.size() / kTuneCols);
        int const nc = std::min(nr, std::max(cap, 0));
        if (nc > 0)
            std::memcpy(table, rows.data(),
                        (size_t)nc * kTuneCols * sizeof(double));
        *nrows = nr;
    });
}

int hdnnp_ctx_tuning(hdnnp_ctx* c, int* out4)
{
    return guarded([&]() {
        if (!c || !out4) throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        out4[0] = c->chunk;
        out4[1] = c->tune.sfWaves;
        out4[2] = c->tune.forceWaves;
        out4[3] = c->nnTiles;
    });
}

int hdnnp_ctx_kernel_name(hdnnp_ctx* c, int which, char* buf, int cap)
{
    return guarded([&]() {
        if (!c || !buf || cap <= 0)
            throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        if (which != 0 && which != 1)
            throw Error(HDNNP_E_INVALID, "ERROR: Kernel 0 (SF) or 1 (force)."
                                         "\n");
        c->setDevice();
        int const fp32 = c->precision != HDNNP_PRECISION_FP64;
        std::string name = which ? "k_force_backward" : "k_sf_forward";
        PlanKernels const pk = c->planFor(fp32);
        if (pk.sf && pk.force)
            name = plan_kernel_name(c->T, fp32, c->tune, which == 1);
        if ((int)name.size() >= cap)
            throw Error(HDNNP_E_CAPACITY, "ERROR: Name buffer too small.\n");
        std::memcpy(buf, name.c_str(), name.size() + 1);
    });
}

int hdnnp_ctx_set_tuning(hdnnp_ctx* c, const int* in4)
{
    return guarded([&]() {
        if (!c || !in4) throw Error(HDNNP_E_INVALID, "ERROR: Null argument.\n");
        if (in4[0] < 1 || in4[0] > kMaxMembers)
            throw Error(HDNNP_E_INVALID, "ERROR: Member chunk out of range.\n");
        if (in4[1] < -1 || in4[1] > 8 || in4[2] < -1 || in4[2] > 8)
            throw Error(HDNNP_E_INVALID,
                        "ERROR: Waves per SIMD out of range.\n");
        if (in4[3] != 1 && in4[3] != 2 && in4[3] != 4)
            throw Error(HDNNP_E_INVALID,
                        "ERROR: Network tiles per wave must be 1, 2 or 4.\n");
        c->setDevice();
        mdEnd(c);
        // Validate the chunk before anything changes.
        try
        {
            (void)c->model->m.flatten(in4[0]);
        }
        catch (std::exception const& e)
        {
            throw Error(HDNNP_E_INVALID, e.what());
        }
        applyTuning(c, TuneChoice{in4[0], in4[1], in4[2], in4[3]});
    });
}

// Forgets the loaded structure or caller-given list: every compute call
// then fails with "No structure loaded" until a new load succeeds.  Loads
// call this first, so a load rejected by its checks leaves no stale sizes
// behind for the kernels.
static void dropStructure(hdnnp_ctx* c)
{
    c->waitPartner();
    c->sync();   // the side stream no longer reads the list buffers
    mdEnd(c);
    c->n = 0;
    c->M = 0;
    c->Mvalid = false;
    c->externalNL = false;
    c->batched = false;
    c->nstruct = 0;
    c->hostPosValid = false;
    c->haveNL = c->haveSF = c->haveNN = c->haveNNDeriv = false;
}

// Atoms, elements and per-element atom lists of a structure or a batch.
// Every input is checked before any state of the context changes.
static void loadAtoms(hdnnp_ctx* c, int64_t n, const double* r, const int32_t* elem)
{
    dropStructure(c);
    if (n <= 0 || !r || !elem)
        throw Error(HDNNP_E_INVALID, "ERROR: Empty structure.\n");
    if (n > 200000000)
        throw Error(HDNNP_E_CAPACITY, "ERROR: Too many atoms.\n");
    int const ne = c->T.ne;
    std::vector<std::vector<int>> atomsOfElem(ne);
    for (int64_t i = 0; i < n; ++i)
    {
        if (elem[i] < 0 || elem[i] >= ne)
            throw Error(HDNNP_E_INVALID,
                        "ERROR: Element index out of range.\n");
        atomsOfElem[elem[i]].push_back((int)i);
    }
    c->setDevice();
    c->hElem.assign(elem, elem + n);
    c->hAtomsOfElem = std::move(atomsOfElem);
    c->dR.ensure(3 * n);
    c->dElemAtom.ensure(n);
    hipCheck(hipMemcpy(c->dR.p, r, 3 * n * sizeof(double),
                       hipMemcpyHostToDevice), "positions");
    hipCheck(hipMemcpy(c->dElemAtom.p, elem, n * sizeof(int32_t),
                       hipMemcpyHostToDevice), "elements");
    c->hR.assign(r, r + 3 * n);
    c->dAtomsOfElem.resize(ne);
    for (int e = 0; e < ne; ++e)
    {
        c->dAtomsOfElem[e].ensure(c->hAtomsOfElem[e].size());
        if (!c->hAtomsOfElem[e].empty())
            hipCheck(hipMemcpy(c->dAtomsOfElem[e].p,
                               c->hAtomsOfElem[e].data(),
                               c->hAtomsOfElem[e].size() * sizeof(int),
                               hipMemcpyHostToDevice), "element lists");
    }
    c->hostPosValid = true;
    c->n = n;
}

int hdnnp_set_structure(hdnnp_ctx* c, int64_t n, const double* r,
                        const int32_t* elem, const double* box)
{
    return guarded([&]() {
        loadAtoms(c, n, r, elem);
        c->batched = false;
        c->nstruct = 1;
        c->periodic = box != nullptr;
        if (box)
        {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) c->box[i][j] = box[3 * i + j];
            inverseBox(c->box, c->invbox);
        }
    });
}

int hdnnp_set_structures(hdnnp_ctx* c, int64_t nstruct,
                         const int64_t* atom_offsets, const double* r,
                         const int32_t* elem, const double* boxes,
                         const int32_t* periodic)
{
    return guarded([&]() {
        dropStructure(c);
        if (nstruct <= 0 || !atom_offsets)
            throw Error(HDNNP_E_INVALID, "ERROR: Empty batch.\n");
        if (atom_offsets[0] != 0)
            throw Error(HDNNP_E_INVALID,
                        "ERROR: atom_offsets[0] must be 0.\n");
        for (int64_t s = 0; s < nstruct; ++s)
            if (atom_offsets[s + 1] <= atom_offsets[s])
                throw Error(HDNNP_E_INVALID,
                            "ERROR: Every structure of a batch needs at "
                            "least one atom (offsets must increase).\n");
        int64_t const n = atom_offsets[nstruct];
        loadAtoms(c, n, r, elem);
        c->batched = true;
        c->nstruct = nstruct;
        c->periodic = false;
        c->hStructOff.assign(atom_offsets, atom_offsets + nstruct + 1);
        c->hBoxes.assign(9 * nstruct, 0.0);
        c->hPeriodic.assign(nstruct, 0);
        for (int64_t s = 0; s < nstruct; ++s)
        {
            bool const per = boxes && (!periodic || periodic[s]);
            c->hPeriodic[s] = per ? 1 : 0;
            if (per)
                for (int q = 0; q < 9; ++q)
                    c->hBoxes[9 * s + q] = boxes[9 * s + q];
        }
        std::vector<int> structOf((size_t)n), iota((size_t)n);
        for (int64_t s = 0; s < nstruct; ++s)
            for (int64_t i = atom_offsets[s]; i < atom_offsets[s + 1]; ++i)
                structOf[i] = (int)s;
        for (int64_t i = 0; i < n; ++i) iota[i] = (int)i;
        c->dStructOf.ensure(n);
        c->dStructOff.ensure(nstruct + 1);
        c->dCellAtoms.ensure(n);
        hipCheck(hipMemcpy(c->dStructOf.p, structOf.data(), n * sizeof(int),
                           hipMemcpyHostToDevice), "structure index");
        hipCheck(hipMemcpy(c->dStructOff.p, c->hStructOff.data(),
                           (nstruct + 1) * sizeof(int),
                           hipMemcpyHostToDevice), "structure offsets");
        // Atoms are gathered in their own order.
        hipCheck(hipMemcpy(c->dCellAtoms.p, iota.data(), n * sizeof(int),
                           hipMemcpyHostToDevice), "atom order");
    });
}

int64_t hdnnp_num_structures(const hdnnp_ctx* c) { return c ? c->nstruct : 0; }

int hdnnp_get_structure_energies(hdnnp_ctx* c, double* energies)
{
    return guarded([&]() {
        if (!c->haveNN)
            throw Error(HDNNP_E_INVALID, "ERROR: Networks not evaluated.\n");
        c->setDevice();
        if (!c->batched)
        {
            energies[0] = computeEnergy(c);
            return;
        }
        c->dEnergySeg.ensure(c->nstruct);
        hipCheck(launch_energy_segments(c->dE.p, c->dStructOff.p,
                                        (int)c->nstruct, c->dEnergySeg.p,
                                        c->stream), "structure energies");
        copyOut(c, energies, c->dEnergySeg.p, (size_t)c->nstruct);
        c->sync();
    });
}

int hdnnp_get_structure_pbc(hdnnp_ctx* c, int32_t* pbc)
{
    return guarded([&]() {
        requireNL(c);
        if (!c->batched)
        {
            for (int k = 0; k < 3; ++k) pbc[k] = c->pbc[k];
            return;
        }
        for (int64_t q = 0; q < 3 * c->nstruct; ++q) pbc[q] = c->hPbcS[q];
    });
}

int hdnnp_set_positions(hdnnp_ctx* c, const double* r)
{
    return guarded([&]() {
        mdEnd(c);
        requireStructure(c);
        if (c->externalNL)
            throw Error(HDNNP_E_INVALID, "ERROR: No positions: the neighbour "
                                         "list was given by the caller.\n");
        c->setDevice();
        hipCheck(hipMemcpyAsync(c->dR.p, r, 3 * c->n * sizeof(double),
                                hipMemcpyHostToDevice, c->stream), "positions");
        c->hR.assign(r, r + 3 * c->n);
        c->hostPosValid = true;
        c->haveNL = c->haveSF = c->haveNN = c->haveNNDeriv = false;
    });
}

int hdnnp_set_positions_device(hdnnp_ctx* c, const double* r_device)
{
    return guarded([&]() {
        mdEnd(c);
        requireStructure(c);
        if (c->externalNL)
            throw Error(HDNNP_E_INVALID, "ERROR: No positions: the neighbour "
                                         "list was given by the caller.\n");
        c->setDevice();
        hipCheck(hipMemcpyAsync(c->dR.p, r_device, 3 * c->n * sizeof(double),
                                hipMemcpyDeviceToDevice, c->stream),
                 "positions");
        c->hostPosValid = false;
        c->haveNL = c->haveSF = c->haveNN = c->haveNNDeriv = false;
    });
}

int hdnnp_neighbor_list(hdnnp_ctx* c, double cutoff)
{
    return guarded([&]() {
        if (!(cutoff > 0.0))
            throw Error(HDNNP_E_INVALID, "ERROR: Cutoff must be positive.\n");
        if (c && c->externalNL)
            throw Error(HDNNP_E_INVALID, "ERROR: The neighbour list was given "
                                         "by the caller.\n");
        mdEnd(c);
        buildNeighborList(c, cutoff);
    });
}

int hdnnp_symmetry_functions(hdnnp_ctx* c, int derivatives)
{
    (void)derivatives;
    return guarded([&]() { computeSF(c, false); });
}

int hdnnp_atomic_networks(hdnnp_ctx* c, int derivatives)
{
    return guarded([&]() { computeNN(c, derivatives != 0); });
}

int hdnnp_energy(hdnnp_ctx* c, double* energy)
{
    return guarded([&]() {
        double const e = computeEnergy(c);
        if (energy) *energy = e;
    });
}

int hdnnp_forces(hdnnp_ctx* c, int mode)
{
    return guarded([&]() { computeForces(c, mode); });
}

int hdnnp_evaluate(hdnnp_ctx* c, int force_mode, double* energy)
{
    return guarded([&]() {
        mdEnd(c);
        if (!c->externalNL) buildNeighborList(c, c->T.maxCutoff);
        computeSF(c, false);
        computeNN(c, force_mode >= 0);
        double const e = computeEnergy(c);
        if (force_mode >= 0) computeForces(c, force_mode);
        if (energy) *energy = e;
    });
}

int64_t hdnnp_num_atoms(const hdnnp_ctx* c) { return c->n; }
int64_t hdnnp_num_neighbor_entries(const hdnnp_ctx* c)
{
    if (!c || !c->haveNL) return 0;
    int64_t m = 0;
    int const rc = guarded([&]() {
        hdnnp_ctx* cc = const_cast<hdnnp_ctx*>(c);
        cc->setDevice();
        m = totalEntries(cc);
    });
    return rc == HDNNP_OK ? m : rc;
}

int hdnnp_get_pbc(hdnnp_ctx* c, int32_t* pbc)
{
    return guarded([&]() {
        requireNL(c);
        if (c->batched)
            throw Error(HDNNP_E_INVALID, "ERROR: A batch has image counts "
                                         "per structure "
                                         "(hdnnp_get_structure_pbc).\n");
        for (int k = 0; k < 3; ++k) pbc[k] = c->pbc[k];
    });
}


int hdnnp_get_neighbors(hdnnp_ctx* c, int64_t* offsets, int32_t* index,
                        int32_t* image, double* d, double* dr)
{
    return guarded([&]() {
        requireNL(c);
        c->setDevice();
        EntryMap const m = entryMap(c);
        size_t const M = (size_t)m.csr.back();
        if (offsets)
            for (size_t i = 0; i < m.csr.size(); ++i) offsets[i] = m.csr[i];
        if (index) compactEntries(c, m, c->dNbrJ.p, 1, index);
        if (image)
        {
            std::vector<uint64_t> keys(M);
            if (c->md.on)   // the MD filter leaves the kept keys unwritten
                hipCheck(launch_md_keys((int)c->n, c->dOff.p, c->md.skinCnt.p,
                                        c->md.skinKeys.p, c->md.newIdx.p,
                                        c->dKeys.p, c->stream), "md keys");
            compactEntries(c, m, c->dKeys.p, 1, keys.data());
            for (size_t q = 0; q < M; ++q)
            {
                uint64_t const k = keys[q];
                image[3 * q] = (int)((k >> 14) & 127) - 64;
                image[3 * q + 1] = (int)((k >> 7) & 127) - 64;
                image[3 * q + 2] = (int)(k & 127) - 64;
            }
        }
        if (d || dr)
        {
            std::vector<double4> geo(M);
            compactEntries(c, m, c->dGeo.p, 1, geo.data());
            for (size_t q = 0; q < M; ++q)
            {
                if (d) d[q] = geo[q].w;
                if (dr)
                {
                    dr[3 * q] = geo[q].x;
                    dr[3 * q + 1] = geo[q].y;
                    dr[3 * q + 2] = geo[q].z;
                }
            }
        }
    });
}

int hdnnp_get_G(hdnnp_ctx* c, double* G, int ld)
{
    return guarded([&]() {
        if (!c->haveSF)
            throw Error(HDNNP_E_INVALID, "ERROR: No symmetry functions.\n");
        if (ld < c->ld) throw Error(HDNNP_E_INVALID, "ERROR: ld too small.\n");
        c->setDevice();
        std::vector<double> h((size_t)c->n * c->ld);
        copyOut(c, h.data(), c->dG.p, h.size());
        c->sync();
        for (int64_t i = 0; i < c->n; ++i)
        {
            int const nsf = c->flat.elem[c->hElem[i]].nsf;
            for (int k = 0; k < ld; ++k)
                G[(size_t)i * ld + k] = k < nsf ? h[(size_t)i * c->ld + k] : 0.0;
        }
    });
}

int hdnnp_get_dEdG(hdnnp_ctx* c, double* dEdG, int ld)
{
    return guarded([&]() {
        if (!c->haveNNDeriv)
            throw Error(HDNNP_E_INVALID, "ERROR: No dE/dG.\n");
        if (ld < c->ld) throw Error(HDNNP_E_INVALID, "ERROR: ld too small.\n");
        c->setDevice();
        std::vector<double> h((size_t)c->n * c->ld);
        copyOut(c, h.data(), c->dEdG.p, h.size());
        c->sync();
        for (int64_t i = 0; i < c->n; ++i)
        {
            int const nsf = c->flat.elem[c->hElem[i]].nsf;
            for (int k = 0; k < ld; ++k)
                dEdG[(size_t)i * ld + k] = k < nsf ? h[(size_t)i * c->ld + k] : 0.0;
        }
    });
}

int hdnnp_get_atomic_energies(hdnnp_ctx* c, double* e)
{
    return guarded([&]() {
        if (!c->haveNN) throw Error(HDNNP_E_INVALID, "ERROR: No energies.\n");
        c->setDevice();
        copyOut(c, e, c->dE.p, (size_t)c->n);
        c->sync();
    });
}

int hdnnp_get_forces(hdnnp_ctx* c, double* f)
{
    return guarded([&]() {
        if (c->externalNL)
            throw Error(HDNNP_E_INVALID, "ERROR: Forces of a caller-given "
                                         "neighbour list are scattered with "
                                         "hdnnp_get_scattered_forces.\n");
        if (!c->dF.p) throw Error(HDNNP_E_INVALID, "ERROR: No forces.\n");
        c->setDevice();
        copyOut(c, f, c->dF.p, 3 * (size_t)c->n);
        c->sync();
    });
}

const double* hdnnp_forces_device(hdnnp_ctx* c) { return c->dF.p; }

int hdnnp_md_init(hdnnp_ctx* c, const double* masses, double dt, double skin,
                  double ftoa, const double* v, int force_mode, int use_graph)
{
    return guarded([&]() {
        requireStructure(c);
        if (c->externalNL || c->batched || !c->periodic)
            throw Error(HDNNP_E_INVALID, "ERROR: MD needs one periodic "
                                         "structure (hdnnp_set_structure "
                                         "with a cell).\n");
        if (!masses || !(dt > 0.0) || !(skin >= 0.0) || !(ftoa > 0.0))
            throw Error(HDNNP_E_INVALID, "ERROR: MD needs masses, dt > 0, "
                                         "skin >= 0 and ftoa > 0.\n");
        if (force_mode != HDNNP_FORCES_REFERENCE
            && force_mode != HDNNP_FORCES_CONSISTENT)
            throw Error(HDNNP_E_INVALID, "ERROR: Unknown force mode.\n");
        c->setDevice();
        mdEnd(c);
        auto& m = c->md;
        int const n = (int)c->n, ne = c->T.ne;
        std::vector<double> acc(ne), mass(ne);
        for (int e = 0; e < ne; ++e)
        {
            if (!(masses[e] > 0.0))
                throw Error(HDNNP_E_INVALID, "ERROR: Masses must be "
                                             "positive.\n");
            acc[e] = ftoa / masses[e];
            mass[e] = masses[e];
        }
        m.dt = dt;
        m.skin = skin;
        m.ftoa = ftoa;
        m.rc = c->T.maxCutoff;
        m.forceMode = force_mode;
        m.useGraph = use_graph != 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                m.cell.box[i][j] = c->box[i][j];
                m.cell.invbox[i][j] = c->invbox[i][j];
            }
        m.v.ensure(3 * (size_t)n);
        m.r0.ensure(3 * (size_t)n);
        m.accel.ensure(ne);
        m.mass.ensure(ne);
        m.part.ensure(md_mv2_blocks(n));
        m.flags.ensure(hdnnp_ctx::kMdFlags);
        m.counter.ensure(1);
        if (!m.hFlags)
            hipCheck(hipHostMalloc((void**)&m.hFlags,
                                   hdnnp_ctx::kMdFlags * sizeof(int),
                                   hipHostMallocDefault), "hipHostMalloc");
        for (int k = 0; k < hdnnp_ctx::kMdFlags; ++k) m.hFlags[k] = 0;
        m.wrapMoves = 0;
        hipStream_t s = c->stream;
        hipCheck(hipMemsetAsync(m.flags.p, 0,
                                hdnnp_ctx::kMdFlags * sizeof(int), s),
                 "flags");
        // The step record's counter indexes epot / ekin: it must start at
        // zero also for runs that never call hdnnp_md_run (the substeps of
        // hdnnp_mts_run record through it too).
        hipCheck(hipMemsetAsync(m.counter.p, 0, sizeof(int), s), "counter");
        if (v)
            hipCheck(hipMemcpyAsync(m.v.p, v, 3 * (size_t)n * sizeof(double),
                                    hipMemcpyHostToDevice, s), "velocities");
        else
            hipCheck(hipMemsetAsync(m.v.p, 0, 3 * (size_t)n * sizeof(double),
                                    s), "velocities");
        hipCheck(hipMemcpyAsync(m.accel.p, acc.data(), ne * sizeof(double),
                                hipMemcpyHostToDevice, s), "accelerations");
        hipCheck(hipMemcpyAsync(m.mass.p, mass.data(), ne * sizeof(double),
                                hipMemcpyHostToDevice, s), "masses");
        m.steps = 0;
        m.rebuilds = 0;
        m.on = true;
        mdRebuild(c);
        c->dF.ensure(3 * (size_t)n);
        mdForces(c, true);
    });
}

int hdnnp_md_run(hdnnp_ctx* c, int64_t nsteps, double* epot, double* ekin)
{
    return guarded([&]() {
        if (!c->md.on)
            throw Error(HDNNP_E_INVALID, "ERROR: No MD run (hdnnp_md_init; a "
                                         "later structure, position or list "
                                         "call ends the run).\n");
        if (nsteps < 0 || nsteps > 100000000)
            throw Error(HDNNP_E_INVALID, "ERROR: Bad step count.\n");
        c->setDevice();
        auto& m = c->md;
        int const need = (int)std::max<int64_t>(nsteps, 1);
        if (need > m.recCap)
        {
            mdDropGraphs(c);
            m.epot.ensure(need);
            m.ekin.ensure(need);
            m.recCap = need;
        }
        hipCheck(hipMemsetAsync(m.counter.p, 0, sizeof(int), c->stream),
                 "md counter");
        for (int64_t k = 0; k < nsteps; ++k) mdStep(c);
        copyOut(c, epot, m.epot.p, (size_t)nsteps);
        copyOut(c, ekin, m.ekin.p, (size_t)nsteps);
        c->sync();
    });
}

int hdnnp_md_get_state(hdnnp_ctx* c, double* r, double* v)
{
    return guarded([&]() {
        if (!c->md.on) throw Error(HDNNP_E_INVALID, "ERROR: No MD run.\n");
        c->setDevice();
        copyOut(c, r, c->dR.p, 3 * (size_t)c->n);
        copyOut(c, v, c->md.v.p, 3 * (size_t)c->n);
        c->sync();
    });
}

int64_t hdnnp_md_rebuilds(const hdnnp_ctx* c) { return c ? c->md.rebuilds : 0; }

int64_t hdnnp_md_partner_misses(hdnnp_ctx* c)
{
    if (!c || !c->md.on) return 0;
    int v = 0;
    int const rc = guarded([&]() {
        c->setDevice();
        hipCheck(hipMemcpyAsync(&v, c->md.flags.p + 3, sizeof(int),
                                hipMemcpyDeviceToHost, c->stream), "misses");
        c->sync();
    });
    return rc == HDNNP_OK ? v : rc;
}

// ------------------------------------------------------- MTS (r-RESPA) --

// Positions r_src (device, n x 3) into c, displacement check and filter of
// its skin list, a rebuild if needed (or forced), then the forces.
// Returns true if c rebuilt its list (its positions may have been wrapped).
static bool mdEvalAt(hdnnp_ctx* c, const double* r_src, bool forceRebuild)
{
    auto& m = c->md;
    int const n = (int)c->n;
    hipStream_t s = c->stream;
    if (r_src != c->dR.p)
        hipCheck(hipMemcpyAsync(c->dR.p, r_src, 3 * (size_t)n * sizeof(double),
                                hipMemcpyDeviceToDevice, s), "positions");
    bool rebuild = forceRebuild;
    if (!rebuild)
    {
        hipCheck(launch_md_clear_flags(m.flags.p, s), "md flags");
        double const half = 0.5 * m.skin;
        hipCheck(launch_md_drift(n, c->dR.p, m.v.p, m.r0.p, 0.0, half * half,
                                 m.flags.p, s), "md check");
        hipCheck(launch_md_filter(n, m.cell, m.rc * m.rc, c->dR.p, c->dOff.p,
                                  m.skinCnt.p, m.skinKeys.p, c->dCount.p,
                                  c->dNbrJ.p, c->dGeo.p, m.newIdx.p, m.cap,
                                  m.flags.p, s), "md filter");
        hipCheck(hipMemcpyAsync(m.hFlags, m.flags.p, 4 * sizeof(int),
                                hipMemcpyDeviceToHost, s), "md flags");
        hipCheck(hipStreamSynchronize(s), "sync");
        mdCheckStable(c);
        rebuild = m.hFlags[0] || m.hFlags[1];
    }
    if (rebuild) mdRebuild(c);
    mdForces(c, false);
    return rebuild;
}

// RMSE of the surrogate's forces against the full model's (device sum).
static double mtsRmse(hdnnp_ctx* sur)
{
    auto& S = sur->mts;
    int const n = (int)sur->n;
    hipStream_t s = sur->stream;
    hipCheck(hipStreamSynchronize(S.full->stream), "sync");
    hipCheck(launch_md_fdiff2(n, sur->dF.p, S.full->dF.p, S.d2.p, s), "rmse");
    hipCheck(launch_energy_sum(S.d2.p, n, S.sum.p, s), "rmse");
    double v = 0.0;
    hipCheck(hipMemcpyAsync(&v, S.sum.p, sizeof(double),
                            hipMemcpyDeviceToHost, s), "rmse");
    hipCheck(hipStreamSynchronize(s), "sync");
    return std::sqrt(v / (3.0 * (double)n));
}

// Both contexts keep the same positions and lists built at the same
// positions, so that two identical models give forces equal bit for bit.
// When the leader rebuilds inside its substeps, the follower rebuilds its
// list right away at the leader's (wrapped) positions; a wrap of wrapped
// positions shifts nothing, and should rounding shift an atom after all,
// the leader takes the follower's positions and rebuilds as well.
static void mtsRebuildWith(hdnnp_ctx* follower, hdnnp_ctx* leader)
{
    int const n = (int)follower->n;
    hipCheck(hipStreamSynchronize(leader->stream), "sync");
    hipCheck(hipMemcpyAsync(follower->dR.p, leader->dR.p,
                            3 * (size_t)n * sizeof(double),
                            hipMemcpyDeviceToDevice, follower->stream),
             "positions");
    int64_t const wm = follower->md.wrapMoves;
    mdRebuild(follower);
    if (follower->md.wrapMoves == wm) return;
    hipCheck(hipStreamSynchronize(follower->stream), "sync");
    hipCheck(hipMemcpyAsync(leader->dR.p, follower->dR.p,
                            3 * (size_t)n * sizeof(double),
                            hipMemcpyDeviceToDevice, leader->stream),
             "positions");
    mdRebuild(leader);
    mdForces(leader, true);
}

// The leader's substeps, each followed by the follower's rebuild when the
// leader rebuilt.
static void mtsSubsteps(hdnnp_ctx* leader, hdnnp_ctx* follower, int nsub)
{
    for (int k = 0; k < nsub; ++k)
    {
        int64_t const rb = leader->md.rebuilds;
        mdStep(leader);
        if (leader->md.rebuilds != rb) mtsRebuildWith(follower, leader);
    }
}

// After the substeps the follower takes the leader's positions and
// evaluates its forces there.  When its own check rebuilds it, the leader
// rebuilds at the same positions and evaluates again.
static void mtsFollow(hdnnp_ctx* follower, hdnnp_ctx* leader)
{
    int64_t const rb = follower->md.rebuilds;
    mdEvalAt(follower, leader->dR.p, false);
    hipCheck(hipStreamSynchronize(follower->stream), "sync");
    if (follower->md.rebuilds == rb) return;
    mdEvalAt(leader, follower->dR.p, true);
    hipCheck(hipStreamSynchronize(leader->stream), "sync");
}

// One outer step.  r-RESPA (Tuckerman, Berne & Martyna 1992) with the
// surrogate as the fast force and F_full - F_sur as the slow force:
// half kick with the slow force, n_inner velocity-Verlet substeps on the
// surrogate, full evaluation, half kick with the new slow force.  When
// the last measured RMSE exceeded the gate, the step is instead n_inner
// velocity-Verlet substeps on the full model.  The positions of both
// contexts stay identical: whenever one rebuilds (and may wrap atoms into
// the cell) the other takes its positions and rebuilds too.
static void mtsOuter(hdnnp_ctx* sur, double* rmseOut, int* modeOut)
{
    auto& S = sur->mts;
    hdnnp_ctx* full = S.full;
    int const n = (int)sur->n;
    double const h = 0.5 * S.dtOuter;
    if (S.useSur)
    {
        hipCheck(launch_md_kick_diff(n, sur->md.v.p, full->dF.p, sur->dF.p,
                                     sur->dElemAtom.p, sur->md.accel.p, h,
                                     sur->stream), "slow kick");
        mtsSubsteps(sur, full, S.nInner);
        hipCheck(hipStreamSynchronize(sur->stream), "sync");
        mtsFollow(full, sur);
        double const rmse = mtsRmse(sur);
        hipCheck(launch_md_kick_diff(n, sur->md.v.p, full->dF.p, sur->dF.p,
                                     sur->dElemAtom.p, sur->md.accel.p, h,
                                     sur->stream), "slow kick");
        *rmseOut = rmse;
        *modeOut = 0;
    }
    else
    {
        hipCheck(hipStreamSynchronize(sur->stream), "sync");
        hipCheck(hipMemcpyAsync(full->md.v.p, sur->md.v.p,
                                3 * (size_t)n * sizeof(double),
                                hipMemcpyDeviceToDevice, full->stream), "v");
        mtsSubsteps(full, sur, S.nInner);
        hipCheck(hipMemcpyAsync(sur->md.v.p, full->md.v.p,
                                3 * (size_t)n * sizeof(double),
                                hipMemcpyDeviceToDevice, full->stream), "v");
        hipCheck(hipStreamSynchronize(full->stream), "sync");
        mtsFollow(sur, full);
        *rmseOut = mtsRmse(sur);
        *modeOut = 1;
        S.fallbacks += 1;
    }
    S.useSur = *rmseOut <= S.gate;
    hipCheck(hipStreamSynchronize(full->stream), "sync");
    hipCheck(launch_md_record(n, sur->md.v.p, sur->dElemAtom.p,
                              sur->md.mass.p, sur->md.part.p,
                              full->dEnergy.p, 0.5 / sur->md.ftoa,
                              S.counter.p, S.recCap, S.epot.p, S.ekin.p,
                              sur->stream), "mts record");
    S.outer += 1;
}

int hdnnp_mts_init(hdnnp_ctx* full, hdnnp_ctx* sur, const double* masses,
                   double dt_outer, int n_inner, double skin, double ftoa,
                   const double* v, double rmse_gate, int use_graph)
{
    return guarded([&]() {
        if (!full || !sur || full == sur)
            throw Error(HDNNP_E_INVALID, "ERROR: MTS needs two contexts.\n");
        if (full->n != sur->n || full->device != sur->device)
            throw Error(HDNNP_E_INVALID, "ERROR: MTS contexts must hold the "
                                         "same structure on one device.\n");
        if (n_inner < 1 || !(dt_outer > 0.0))
            throw Error(HDNNP_E_INVALID, "ERROR: MTS needs n_inner >= 1 and "
                                         "dt_outer > 0.\n");
        double const dt = dt_outer / n_inner;
        int rc = hdnnp_md_init(sur, masses, dt, skin, ftoa, v,
                               HDNNP_FORCES_CONSISTENT, use_graph);
        if (rc != HDNNP_OK) throw Error(rc, hdnnp_last_error());
        rc = hdnnp_md_init(full, masses, dt, skin, ftoa, nullptr,
                           HDNNP_FORCES_CONSISTENT, use_graph);
        if (rc != HDNNP_OK) throw Error(rc, hdnnp_last_error());
        auto& S = sur->mts;
        S.full = full;
        S.nInner = n_inner;
        S.dtOuter = dt_outer;
        S.gate = rmse_gate;
        S.outer = 0;
        S.fallbacks = 0;
        S.d2.ensure(sur->n);
        S.sum.ensure(1);
        S.counter.ensure(1);
        S.useSur = mtsRmse(sur) <= S.gate;
    });
}

int hdnnp_mts_run(hdnnp_ctx* sur, int64_t nouter, double* epot, double* ekin,
                  double* rmse, int32_t* mode)
{
    return guarded([&]() {
        auto& S = sur->mts;
        if (!S.full || !sur->md.on || !S.full->md.on)
            throw Error(HDNNP_E_INVALID, "ERROR: No MTS run "
                                         "(hdnnp_mts_init).\n");
        if (nouter < 0 || nouter > 100000000)
            throw Error(HDNNP_E_INVALID, "ERROR: Bad step count.\n");
        sur->setDevice();
        int const need = (int)std::max<int64_t>(nouter, 1);
        if (need > S.recCap)
        {
            S.epot.ensure(need);
            S.ekin.ensure(need);
            S.recCap = need;
        }
        hipCheck(hipMemsetAsync(S.counter.p, 0, sizeof(int), sur->stream),
                 "mts counter");
        for (int64_t k = 0; k < nouter; ++k)
        {
            double r = 0.0;
            int md = 0;
            mtsOuter(sur, &r, &md);
            if (rmse) rmse[k] = r;
            if (mode) mode[k] = md;
        }
        copyOut(sur, epot, S.epot.p, (size_t)nouter);
        copyOut(sur, ekin, S.ekin.p, (size_t)nouter);
        sur->sync();
    });
}

int64_t hdnnp_num_weights(const hdnnp_ctx* c)
{
    return c ? (int64_t)c->flat.weights.size() : 0;
}

int hdnnp_get_weights(hdnnp_ctx* c, double* w)
{
    return guarded([&]() {
        c->setDevice();
        copyOut(c, w, c->dWeights.p, c->flat.weights.size());
        c->sync();
    });
}

int hdnnp_set_weights(hdnnp_ctx* c, const double* w)
{
    return guarded([&]() {
        c->setDevice();
        size_t const nw = c->flat.weights.size();
        hipCheck(hipMemcpyAsync(c->dWeights.p, w, nw * sizeof(double),
                                hipMemcpyHostToDevice, c->stream), "weights");
        hipCheck(launch_weights_to_f32((int)nw, c->dWeights.p,
                                       c->dWeights32.p, c->stream), "weights");
        c->sync();
        c->haveNN = c->haveNNDeriv = false;
    });
}

int hdnnp_train_gradient(hdnnp_ctx* c, const double* f_ref, double e_ref,
                         double w_e, double w_f, double* grad, double* loss)
{
    return guarded([&]() {
        requireStructure(c);
        if (c->externalNL || c->batched || !f_ref)
            throw Error(HDNNP_E_INVALID, "ERROR: Training needs one structure "
                                         "from hdnnp_set_structure and its "
                                         "reference forces.\n");
        for (size_t e = 0; e < c->flat.elem.size(); ++e)
        {
            ElemDesc const& ed = c->flat.elem[e];
            bool ok = !ed.normalize
                   && ed.width[ed.nlayers - 1] == 1
                   && train_atoms_per_grad_block(ed) >= 1;
            for (int l = 1; l < ed.nlayers; ++l)
                ok = ok && (ed.act[l] == AF_IDENTITY || ed.act[l] == AF_TANH
                            || ed.act[l] == AF_LOGISTIC
                            || ed.act[l] == AF_SOFTPLUS);
            if (!ok)
                throw Error(HDNNP_E_UNSUPPORTED,
                            "ERROR: Training supports networks of "
                            "identity, tanh, "
                            "logistic or softplus layers with one output and "
                            "no normalize_nodes, of at most a few thousand "
                            "weights.\n");
        }
        c->setDevice();
        mdEnd(c);
        int const n = (int)c->n;
        hipStream_t s = c->stream;
        buildNeighborList(c, c->T.maxCutoff);
        computeSF(c, false);
        computeNN(c, true);
        launchEnergy(c);
        computeForces(c, HDNNP_FORCES_CONSISTENT, false);
        size_t const nw = c->flat.weights.size();
        c->tU.ensure((size_t)n * c->ld);
        c->tD2.ensure(n);
        c->tFref.ensure(3 * (size_t)n);
        c->tGrad.ensure(nw);
        c->tLoss.ensure(1);
        hipCheck(hipMemcpyAsync(c->tFref.p, f_ref, 3 * (size_t)n * sizeof(double),
                                hipMemcpyHostToDevice, s), "reference forces");
        TrainArgs A{};
        A.T = c->T;
        A.n = n;
        A.off = c->dOff.p;
        A.cnt = c->dCount.p;
        A.nbrJ = c->dNbrJ.p;
        A.geo = c->dGeo.p;
        A.elem = c->dElemAtom.p;
        A.G = c->dG.p;
        A.ldG = c->ld;
        A.F = c->dF.p;
        A.Fref = c->tFref.p;
        A.U = c->tU.p;
        A.d2 = c->tD2.p;
        A.energy = c->dEnergy.p;
        A.Eref = e_ref;
        A.wE = w_e;
        A.cF = -2.0 * w_f / (3.0 * (double)n);
        hipCheck(launch_train_u(A, s), "train u");
        {
            int const cap = std::max(c->maxNeigh, 1);
            if (c->T.ncgroup > 0 && train_u_compact_lds_bytes(c->T, cap) > 65536)
                throw Error(HDNNP_E_CAPACITY,
                            "ERROR: Too many neighbours per atom for the "
                            "compact training kernel's LDS staging.\n");
            hipCheck(launch_train_u_compact(A, cap, s), "train u (compact)");
        }
        for (size_t e = 0; e < c->flat.elem.size(); ++e)
        {
            ElemDesc const& ed = c->flat.elem[e];
            int const per = train_atoms_per_grad_block(ed);
            int nwe = 0;
            for (int l = 1; l < ed.nlayers; ++l)
                nwe += ed.width[l] * ed.width[l - 1] + ed.width[l];
            int const na = (int)c->hAtomsOfElem[e].size();
            int const nb = (na + per - 1) / per;
            c->tPart.ensure((size_t)std::max(nb, 1) * nwe);
            hipCheck(launch_train_grad(A, ed, (int)e, c->dAtomsOfElem[e].p, na,
                                       c->tPart.p, nwe,
                                       c->tGrad.p + ed.woff[1], s),
                     "train gradient");
        }
        hipCheck(launch_energy_sum(c->tD2.p, n, c->tLoss.p, s), "train loss");
        double sumd2 = 0.0, energy = 0.0;
        hipCheck(hipMemcpyAsync(&sumd2, c->tLoss.p, sizeof(double),
                                hipMemcpyDeviceToHost, s), "train loss");
        hipCheck(hipMemcpyAsync(&energy, c->dEnergy.p, sizeof(double),
                                hipMemcpyDeviceToHost, s), "energy");
        copyOut(c, grad, c->tGrad.p, nw);
        c->sync();
        c->energy = energy;
        double const de = (energy - e_ref) / (double)n;
        if (loss)
        {
            loss[0] = w_e * de * de;
            loss[1] = w_f * sumd2 / (3.0 * (double)n);
        }
        c->haveGrad = true;
    });
}

int hdnnp_train_adam(hdnnp_ctx* c, double lr, double beta1, double beta2,
                     double eps)
{
    return guarded([&]() {
        if (!c->haveGrad)
            throw Error(HDNNP_E_INVALID, "ERROR: No gradient "
                                         "(hdnnp_train_gradient).\n");
        c->setDevice();
        size_t const nw = c->flat.weights.size();
        if (c->tM.cap < nw || c->adamStep == 0)
        {
            c->tM.ensure(nw);
            c->tV.ensure(nw);
            hipCheck(hipMemsetAsync(c->tM.p, 0, nw * sizeof(double), c->stream),
                     "adam m");
            hipCheck(hipMemsetAsync(c->tV.p, 0, nw * sizeof(double), c->stream),
                     "adam v");
            c->adamStep = 0;
        }
        c->adamStep += 1;
        hipCheck(launch_adam((int)nw, c->dWeights.p, c->dWeights32.p,
                             c->tGrad.p, c->tM.p, c->tV.p, lr, beta1, beta2,
                             eps, c->adamStep, c->stream), "adam");
        c->sync();
        c->haveNN = c->haveNNDeriv = false;
    });
}

int hdnnp_energy_weight_jacobian(hdnnp_ctx* c, double* J, int file_order)
{
    return guarded([&]() {
        requireNL(c);
        int const ne = c->T.ne, n = (int)c->n;
        int const per = train_atoms_per_block();
        int perAtom = 0, maxP = 0;
        std::vector<int> base(ne), nwe(ne);
        for (int e = 0; e < ne; ++e)
        {
            ElemDesc const& ed = c->flat.elem[e];
            if (ed.normalize)
                throw Error(HDNNP_E_UNSUPPORTED, "ERROR: normalize_nodes is "
                                                 "not supported here.\n");
            int pa = ed.maxwidth, nw = 0;
            for (int l = 1; l < ed.nlayers; ++l)
            {
                pa += ed.width[l - 1] + ed.width[l];
                nw += ed.width[l] * ed.width[l - 1] + ed.width[l];
            }
            perAtom = std::max(perAtom, pa);
            maxP = std::max(maxP, nw);
            base[e] = ed.woff[1];
            nwe[e] = nw;
        }
        if ((size_t)perAtom * per * sizeof(double) > 65536)
            throw Error(HDNNP_E_CAPACITY, "ERROR: Networks too wide for the "
                                          "weight-gradient kernel.\n");
        std::vector<int> soff;
        if (c->batched) soff = c->hStructOff;
        else soff = {0, n};
        int const ns = (int)soff.size() - 1;
        std::vector<int> atoms, b0, bn, be, grp0;
        for (int s = 0; s < ns; ++s)
            for (int e = 0; e < ne; ++e)
            {
                grp0.push_back((int)b0.size());
                int cnt = 0;
                for (int i = soff[s]; i < soff[s + 1]; ++i)
                {
                    if (c->hElem[i] != e) continue;
                    if (cnt % per == 0)
                    {
                        b0.push_back((int)atoms.size());
                        bn.push_back(0);
                        be.push_back(e);
                    }
                    atoms.push_back(i);
                    bn.back() += 1;
                    ++cnt;
                }
            }
        grp0.push_back((int)b0.size());
        int const nblk = (int)b0.size();
        c->setDevice();
        hipStream_t st = c->stream;
        auto up = [&](DBuf<int>& d, std::vector<int> const& v) {
            d.ensure(std::max<size_t>(v.size(), 1));
            if (!v.empty())
                hipCheck(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(int),
                                        hipMemcpyHostToDevice, st),
                         "jacobian tables");
        };
        up(c->jAtoms, atoms);
        up(c->jB0, b0);
        up(c->jBN, bn);
        up(c->jBE, be);
        up(c->jGrp0, grp0);
        up(c->jBase, base);
        up(c->jNwe, nwe);
        size_t const nwTot = c->flat.weights.size();
        c->jPart.ensure((size_t)std::max(nblk, 1) * maxP);
        c->jOut.ensure((size_t)ns * nwTot);
        hipCheck(hipMemsetAsync(c->jOut.p, 0, (size_t)ns * nwTot * sizeof(double),
                                st), "jacobian");
        computeSF(c, false);
        hipCheck(launch_energy_wgrad(c->T, c->dG.p, c->ld, c->jAtoms.p,
                                     c->jB0.p, c->jBN.p, c->jBE.p, nblk,
                                     perAtom, c->jPart.p, maxP, c->jGrp0.p,
                                     c->jBase.p, c->jNwe.p, ne, ns,
                                     (int)nwTot, c->jOut.p, st),
                 "energy weight gradient");
        std::vector<double> h((size_t)ns * nwTot);
        copyOut(c, h.data(), c->jOut.p, h.size());
        c->sync();
        if (!file_order)
        {
            std::copy(h.begin(), h.end(), J);
            return;
        }
        // NeuralNetwork::setConnections order (NeuralNetwork.cpp:147-169):
        // per layer, for each input j, for each neuron k: W[k][j]; then b.
        for (int s = 0; s < ns; ++s)
        {
            double const* src = h.data() + (size_t)s * nwTot;
            double* dst = J + (size_t)s * nwTot;
            size_t q = 0;
            for (int e = 0; e < ne; ++e)
            {
                ElemDesc const& ed = c->flat.elem[e];
                for (int l = 1; l < ed.nlayers; ++l)
                {
                    int const nin = ed.width[l - 1], nout = ed.width[l];
                    for (int j = 0; j < nin; ++j)
                        for (int k = 0; k < nout; ++k)
                            dst[q++] = src[ed.woff[l] + (size_t)k * nin + j];
                    for (int k = 0; k < nout; ++k)
                        dst[q++] = src[ed.boff[l] + k];
                }
            }
        }
    });
}

int hdnnp_kalman_init(hdnnp_ctx* c, double epsilon, double q0, double qtau,
                      double qmin, double eta0, double etatau, double etamax)
{
    return guarded([&]() {
        if (!(epsilon > 0.0) || !(eta0 > 0.0))
            throw Error(HDNNP_E_INVALID, "ERROR: Kalman filter needs epsilon "
                                         "> 0 and eta > 0.\n");
        c->setDevice();
        size_t const nw = c->flat.weights.size();
        c->kfP.ensure(nw * nw);
        c->kfX.ensure(nw);
        c->kfA.ensure(1);
        c->kfH.ensure(nw);
        hipCheck(launch_kf_init((int)nw, c->kfP.p, 1.0 / epsilon, c->stream),
                 "kalman init");
        c->kfEta = eta0;
        c->kfEtaTau = etatau;
        c->kfEtaMax = etamax;
        c->kfQ = q0;
        c->kfQTau = qtau;
        c->kfQMin = qmin;
        c->kfUpdates = 0;
        c->sync();
    });
}

int hdnnp_kalman_update(hdnnp_ctx* c, const double* h, double xi)
{
    return guarded([&]() {
        if (c->kfUpdates < 0)
            throw Error(HDNNP_E_INVALID, "ERROR: No Kalman filter "
                                         "(hdnnp_kalman_init).\n");
        c->setDevice();
        size_t const nw = c->flat.weights.size();
        hipCheck(hipMemcpyAsync(c->kfH.p, h, nw * sizeof(double),
                                hipMemcpyHostToDevice, c->stream), "jacobian");
        if (c->kfEta < c->kfEtaMax) c->kfEta *= std::exp(c->kfEtaTau);
        hipCheck(launch_kf_update((int)nw, c->kfP.p, c->kfH.p, c->kfX.p,
                                  c->kfA.p, 1.0 / c->kfEta, c->kfQ, xi,
                                  c->dWeights.p, c->dWeights32.p, c->stream),
                 "kalman update");
        if (c->kfQ > c->kfQMin) c->kfQ *= std::exp(-c->kfQTau);
        c->kfUpdates += 1;
        c->sync();
        c->haveNN = c->haveNNDeriv = false;
    });
}

int hdnnp_train_reset(hdnnp_ctx* c)
{
    return guarded([&]() { c->adamStep = 0; });
}

int hdnnp_set_neighbor_list(hdnnp_ctx* c, int64_t ncentre,
                            const int32_t* centre_index,
                            const int32_t* centre_elem,
                            const int64_t* offsets, const int32_t* nbr_index,
                            const int32_t* nbr_elem, const double* dr)
{
    return guarded([&]() {
        if (!c)
            throw Error(HDNNP_E_INVALID, "ERROR: No context.\n");
        dropStructure(c);
        if (ncentre <= 0 || !centre_elem || !offsets)
            throw Error(HDNNP_E_INVALID, "ERROR: Empty neighbour list.\n");
        if (offsets[0] != 0)
            throw Error(HDNNP_E_INVALID, "ERROR: offsets[0] must be 0.\n");
        int64_t const M = offsets[ncentre];
        if (M < 0 || M > (int64_t)2000000000 || ncentre > 200000000)
            throw Error(HDNNP_E_CAPACITY, "ERROR: Neighbour list too large.\n");
        if (M > 0 && (!nbr_index || !nbr_elem || !dr))
            throw Error(HDNNP_E_INVALID, "ERROR: Missing neighbour data.\n");
        c->setDevice();
        int const ne = c->T.ne;
        int const n = (int)ncentre;
        // Centres 0..n-1; neighbour entry q is the virtual atom n + q.
        std::vector<int32_t> elem((size_t)n + (size_t)M);
        std::vector<int> off(n + 1), cnt(n), nbrJ((size_t)M);
        std::vector<double4> geo((size_t)M);
        std::vector<std::vector<int>> atomsOfElem(ne);
        int maxc = 0;
        for (int i = 0; i < n; ++i)
        {
            if (centre_elem[i] < 0 || centre_elem[i] >= ne)
                throw Error(HDNNP_E_INVALID,
                            "ERROR: Element index out of range.\n");
            elem[i] = centre_elem[i];
            atomsOfElem[centre_elem[i]].push_back(i);
            if (offsets[i + 1] < offsets[i])
                throw Error(HDNNP_E_INVALID, "ERROR: Offsets not ascending.\n");
            off[i] = (int)offsets[i];
            cnt[i] = (int)(offsets[i + 1] - offsets[i]);
            maxc = std::max(maxc, cnt[i]);
        }
        off[n] = (int)M;
        // Caller indices of centres and neighbours (targets of the forces).
        std::vector<int> ctr(n);
        int64_t nt = 0;
        for (int i = 0; i < n; ++i)
        {
            ctr[i] = centre_index ? centre_index[i] : i;
            if (ctr[i] < 0)
                throw Error(HDNNP_E_INVALID, "ERROR: Centre index out of "
                                             "range.\n");
            nt = std::max<int64_t>(nt, (int64_t)ctr[i] + 1);
        }
        for (int64_t q = 0; q < M; ++q)
        {
            if (nbr_index[q] < 0)
                throw Error(HDNNP_E_INVALID, "ERROR: Neighbour index out of "
                                             "range.\n");
            nt = std::max<int64_t>(nt, (int64_t)nbr_index[q] + 1);
        }
        for (int64_t q = 0; q < M; ++q)
        {
            if (nbr_elem[q] < 0 || nbr_elem[q] >= ne)
                throw Error(HDNNP_E_INVALID,
                            "ERROR: Element index out of range.\n");
            elem[n + q] = nbr_elem[q];
            nbrJ[q] = n + (int)q;
            double const x = dr[3 * q], y = dr[3 * q + 1], z = dr[3 * q + 2];
            geo[q] = make_double4(x, y, z, std::sqrt(x * x + y * y + z * z));
        }
        c->hAtomsOfElem = std::move(atomsOfElem);
        c->hElem.assign(centre_elem, centre_elem + n);
        c->periodic = false;
        c->pbc[0] = c->pbc[1] = c->pbc[2] = 0;
        auto up = [&](auto& buf, auto const& v) {
            buf.ensure(std::max<size_t>(v.size(), 1));
            if (!v.empty())
                hipCheck(hipMemcpy(buf.p, v.data(), v.size() * sizeof(v[0]),
                                   hipMemcpyHostToDevice), "neighbour list");
        };
        up(c->dElemAtom, elem);
        up(c->dOff, off);
        up(c->dCount, cnt);
        up(c->dNbrJ, nbrJ);
        up(c->dGeo, geo);
        c->dAtomsOfElem.resize(ne);
        for (int e = 0; e < ne; ++e) up(c->dAtomsOfElem[e], c->hAtomsOfElem[e]);
        c->entryCap = std::max<int64_t>(M, 1);
        c->M = M;
        c->Mvalid = true;
        c->slabLayout = false;
        c->maxNeigh = maxc;
        {
            // Contributions per caller atom, in the reference's loop order
            // (a counting sort that keeps that order).
            std::vector<int64_t> so((size_t)nt + 1, 0);
            for (int i = 0; i < n; ++i)
            {
                ++so[(size_t)ctr[i] + 1];
                for (int64_t q = offsets[i]; q < offsets[i + 1]; ++q)
                    ++so[(size_t)nbr_index[q] + 1];
            }
            for (int64_t b = 0; b < nt; ++b) so[b + 1] += so[b];
            std::vector<int64_t> pos(so.begin(), so.end() - 1);
            std::vector<int> src((size_t)(n + M));
            for (int i = 0; i < n; ++i)
            {
                src[(size_t)pos[ctr[i]]++] = -(i + 1);
                for (int64_t q = offsets[i]; q < offsets[i + 1]; ++q)
                    src[(size_t)pos[nbr_index[q]]++] = (int)q;
            }
            up(c->dScatOff, so);
            up(c->dScatSrc, src);
        }
        c->extTargets = nt;
        c->externalNL = true;
        c->batched = false;
        c->nstruct = 1;
        c->hostPosValid = false;
        c->n = n;
        c->haveNL = true;
        c->haveSF = c->haveNN = c->haveNNDeriv = false;
    });
}

int hdnnp_get_scattered_forces(hdnnp_ctx* c, int64_t nout, double* f)
{
    return guarded([&]() {
        if (!c || !c->externalNL)
            throw Error(HDNNP_E_INVALID, "ERROR: No caller-given neighbour "
                                         "list.\n");
        if (!c->dCself.p || !c->dCslot.p)
            throw Error(HDNNP_E_INVALID, "ERROR: No forces.\n");
        if (nout < c->extTargets)
            throw Error(HDNNP_E_INVALID, "ERROR: Centre or neighbour index "
                                         "out of range (beyond the force "
                                         "array).\n");
        c->setDevice();
        // InterfaceLammps::getForces (InterfaceLammps.cpp:546-581) on the
        // device: only the rows that receive contributions travel, in and
        // out.  Stored sums carry +dE/dG dG/dr, hence the subtractions.
        size_t const nt = (size_t)c->extTargets;
        if (nt == 0) return;
        c->dScatF.ensure(3 * nt);
        hipCheck(hipMemcpyAsync(c->dScatF.p, f, 3 * nt * sizeof(double),
                                hipMemcpyHostToDevice, c->stream), "forces");
        hipCheck(launch_scatter_ext((int)nt, c->dScatOff.p, c->dScatSrc.p,
                                    c->dCslot.p, c->dCself.p, c->dScatF.p,
                                    c->stream), "force scatter");
        copyOut(c, f, c->dScatF.p, 3 * nt);
        c->sync();
    });
}

// ----------------------------------------------- allocation-free scatter --

namespace {

// Staging of hdnnp_get_scattered_forces_noalloc: the forces of up to
// kNoAllocTargets caller atoms, 3 doubles each (12 MiB).  Every GPU the
// library runs on holds its own copy until the process ends.
constexpr int64_t kNoAllocTargets = (int64_t)1 << 19;
static __device__ __align__(256) double scratch[3 * kNoAllocTargets];

// Device address of `scratch` on each device, looked up once per device
// with that device current (the host-side name is not a device address).
std::mutex g_scratchMutex;
std::map<int, double*> g_scratchPtr;

double* scratchFor(int device)
{
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    auto it = g_scratchPtr.find(device);
    if (it != g_scratchPtr.end()) return it->second;
    double* p = nullptr;
    hipCheck(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)),
             "hipGetSymbolAddress");
    g_scratchPtr[device] = p;
    return p;
}

}  // namespace

int hdnnp_get_scattered_forces_noalloc(hdnnp_ctx* c, int64_t nout, double* f)
{
    return guarded([&]() {
        if (!c || !c->externalNL)
            throw Error(HDNNP_E_INVALID, "ERROR: No caller-given neighbour "
                                         "list.\n");
        if (!c->dCself.p || !c->dCslot.p)
            throw Error(HDNNP_E_INVALID, "ERROR: No forces.\n");
        if (nout < c->extTargets)
            throw Error(HDNNP_E_INVALID, "ERROR: Centre or neighbour index "
                                         "out of range (beyond the force "
                                         "array).\n");
        if (c->extTargets > kNoAllocTargets)
            throw Error(HDNNP_E_CAPACITY, "ERROR: More than 524288 caller "
                                          "atoms receive forces: beyond the "
                                          "staging of the allocation-free "
                                          "scatter.\n");
        c->setDevice();
        double* const buf = scratchFor(c->device);
        size_t const nt = (size_t)c->extTargets;
        if (nt == 0) return;
        // The same kernel as hdnnp_get_scattered_forces, staged in `scratch`
        // (every row it reads is copied in first, so nothing needs clearing).
        hipCheck(hipMemcpyAsync(buf, f, 3 * nt * sizeof(double),
                                hipMemcpyHostToDevice, c->stream), "forces");
        hipCheck(launch_scatter_ext((int)nt, c->dScatOff.p, c->dScatSrc.p,
                                    c->dCslot.p, c->dCself.p, buf,
                                    c->stream), "force scatter");
        hipCheck(hipMemcpyAsync(f, buf, 3 * nt * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream), "forces");
        c->sync();
    });
}

int hdnnp_get_pair_forces(hdnnp_ctx* c, double* self_forces,
                          double* pair_forces)
{
    return guarded([&]() {
        if (!c->dF.p || !c->haveNL)
            throw Error(HDNNP_E_INVALID, "ERROR: No forces.\n");
        c->setDevice();
        copyOut(c, self_forces, c->dCself.p, 3 * (size_t)c->n);
        c->sync();
        size_t M = 0;
        if (pair_forces)
        {
            EntryMap const m = entryMap(c);
            M = (size_t)m.csr.back();
            compactEntries(c, m, c->dCslot.p, 3, pair_forces);
        }
        // Stored as +sum dE/dG dG/dr; the reference's self and pair forces
        // carry the minus sign.
        if (self_forces)
            for (size_t k = 0; k < 3 * (size_t)c->n; ++k)
                self_forces[k] = -self_forces[k];
        if (pair_forces)
            for (size_t k = 0; k < 3 * M; ++k) pair_forces[k] = -pair_forces[k];
    });
}

int64_t hdnnp_extrapolation_warnings(hdnnp_ctx* c, int32_t* per_atom)
{
    int64_t total = -1;
    int rc = guarded([&]() {
        if (!c->haveSF)
            throw Error(HDNNP_E_INVALID, "ERROR: No symmetry functions.\n");
        c->setDevice();
        c->dEw.ensure(c->n);
        hipCheck(launch_ew(c->T, c->dG.p, c->ld, c->dElemAtom.p, (int)c->n,
                           c->dEw.p, c->stream), "extrapolation warnings");
        std::vector<int32_t> h(c->n);
        copyOut(c, h.data(), c->dEw.p, (size_t)c->n);
        c->sync();
        total = 0;
        for (int64_t i = 0; i < c->n; ++i)
        {
            total += h[i];
            if (per_atom) per_atom[i] = h[i];
        }
    });
    return rc == HDNNP_OK ? total : rc;
}

int64_t hdnnp_extrapolation_warning_list(hdnnp_ctx* c, int64_t max,
                                         int32_t* atom, int32_t* sf,
                                         double* value, double* gmin,
                                         double* gmax)
{
    int64_t total = -1;
    int rc = guarded([&]() {
        if (!c->haveSF)
            throw Error(HDNNP_E_INVALID, "ERROR: No symmetry functions.\n");
        c->setDevice();
        int const n = (int)c->n;
        c->dEw.ensure(n);
        hipCheck(launch_ew(c->T, c->dG.p, c->ld, c->dElemAtom.p, n,
                           c->dEw.p, c->stream), "extrapolation warnings");
        std::vector<int32_t> cnt(n);
        copyOut(c, cnt.data(), c->dEw.p, (size_t)n);
        c->sync();
        // Atoms with warnings (few, normally): their scaled G rows come back
        // and are unscaled and tested here with the kernel's arithmetic
        // (SymFnc::unscale, SymFnc.cpp:197-223), in the reference's order.
        double const eps = 1000.0 * 2.220446049250313e-16;
        std::vector<double> row((size_t)c->ld);
        total = 0;
        for (int i = 0; i < n; ++i)
        {
            if (cnt[i] == 0) continue;
            copyOut(c, row.data(), c->dG.p + (size_t)i * c->ld, (size_t)c->ld);
            c->sync();
            ElemDesc const& ed = c->flat.elem[c->hElem[i]];
            for (int k = 0; k < ed.nsf; ++k)
            {
                ScaleDesc const& q = c->flat.scale[ed.sf0 + k];
                double const g = row[k];
                double v;
                switch (q.type)
                {
                case ST_SCALE: v = (g - q.smin) / q.sfac + q.gmin; break;
                case ST_CENTER: v = g + q.gmean; break;
                case ST_SCALECENTER:
                case ST_SCALESIGMA: v = (g - q.smin) / q.sfac + q.gmean; break;
                default: v = g; break;
                }
                if (!(v + eps < q.gmin || v - eps > q.gmax)) continue;
                if (total < max)
                {
                    if (atom) atom[total] = i;
                    if (sf) sf[total] = k;
                    if (value) value[total] = v;
                    if (gmin) gmin[total] = q.gmin;
                    if (gmax) gmax[total] = q.gmax;
                }
                ++total;
            }
        }
    });
    return rc == HDNNP_OK ? total : rc;
}

int64_t hdnnp_sf_statistics(hdnnp_ctx* c, int e, double* out)
{
    int64_t na = -1;
    int rc = guarded([&]() {
        requireNL(c);
        if (e < 0 || e >= c->T.ne)
            throw Error(HDNNP_E_INVALID, "ERROR: Element out of range.\n");
        bool const hadSF = c->haveSF;
        computeSF(c, true);
        int const nsf = c->flat.elem[e].nsf;
        na = (int64_t)c->hAtomsOfElem[e].size();
        c->dStatOut.ensure(4 * (size_t)nsf);
        hipCheck(launch_sf_stats(c->dG.p, c->ld, c->dAtomsOfElem[e].p, (int)na,
                                 nsf, c->dStatOut.p, c->stream), "sf stats");
        std::vector<double> h(4 * nsf);
        copyOut(c, h.data(), c->dStatOut.p, h.size());
        c->sync();
        // Dataset.cpp:1074-1085: mean = sum / n,
        // sigma = sqrt((sum2 - sum^2 / n) / (n - 1)).
        double const nd = (double)na;
        for (int k = 0; k < nsf; ++k)
        {
            double const sum = h[4 * k + 2], sum2 = h[4 * k + 3];
            out[4 * k] = h[4 * k];
            out[4 * k + 1] = h[4 * k + 1];
            out[4 * k + 2] = na ? sum / nd : 0.0;
            out[4 * k + 3] = na > 1 ? std::sqrt((sum2 - sum * sum / nd)
                                                / (nd - 1.0)) : 0.0;
        }
        // G now holds unscaled values.
        c->haveSF = false;
        (void)hadSF;
    });
    return rc == HDNNP_OK ? na : rc;
}

int64_t hdnnp_sf_statistics_raw(hdnnp_ctx* c, int e, double* out)
{
    int64_t na = -1;
    int rc = guarded([&]() {
        requireNL(c);
        if (e < 0 || e >= c->T.ne)
            throw Error(HDNNP_E_INVALID, "ERROR: Element out of range.\n");
        computeSF(c, true);
        int const nsf = c->flat.elem[e].nsf;
        na = (int64_t)c->hAtomsOfElem[e].size();
        c->dStatOut.ensure(4 * (size_t)nsf);
        hipCheck(launch_sf_stats(c->dG.p, c->ld, c->dAtomsOfElem[e].p, (int)na,
                                 nsf, c->dStatOut.p, c->stream), "sf stats");
        copyOut(c, out, c->dStatOut.p, 4 * (size_t)nsf);
        c->sync();
        c->haveSF = false;
    });
    return rc == HDNNP_OK ? na : rc;
}

int hdnnp_profile_enable(hdnnp_ctx* c, int enable)
{
    c->prof = enable != 0;
    return HDNNP_OK;
}

double hdnnp_profile_avg_ms(hdnnp_ctx* c, const char* name, int64_t* calls)
{
    try { c->collect(); } catch (...) {}
    for (int k = 0; k < K_NUM; ++k)
        if (name && std::strcmp(name, kClassNames[k]) == 0)
        {
            if (calls) *calls = c->profCalls[k];
            return c->profCalls[k] ? c->profMs[k] / c->profCalls[k] : 0.0;
        }
    return -1.0;
}

int hdnnp_profile_reset(hdnnp_ctx* c)
{
    try { c->collect(); } catch (...) {}
    for (int k = 0; k < K_NUM; ++k)
    {
        c->profMs[k] = 0.0;
        c->profCalls[k] = 0;
    }
    return HDNNP_OK;
}

int hdnnp_work_counts(hdnnp_ctx* c, int64_t* counts)
{
    return guarded([&]() {
        requireNL(c);
        c->setDevice();
        int const cap = std::max(c->maxNeigh, 1);
        if (sf_forward_lds_bytes(cap, 0) > 65536)
            throw Error(HDNNP_E_CAPACITY, "ERROR: Too many neighbours.\n");
        DBuf<unsigned long long> d;
        d.ensure(HDNNP_WORK_COUNTS);
        hipCheck(hipMemsetAsync(d.p, 0, HDNNP_WORK_COUNTS * sizeof(uint64_t),
                                c->stream), "memset");
        SFArgs a{};
        a.T = c->T;
        a.n = (int)c->n;
        a.off = c->dOff.p;
    a.cnt = c->dCount.p;
        a.geo = c->dGeo.p;
        a.nbrJ = c->dNbrJ.p;
        a.elem = c->dElemAtom.p;
        a.ldG = c->ld;
        a.cap = cap;
        hipCheck(launch_sf_count(a, d.p, c->stream), "work counters");
        unsigned long long h[HDNNP_WORK_COUNTS];
        hipCheck(hipMemcpyAsync(h, d.p, sizeof h, hipMemcpyDeviceToHost,
                                c->stream), "counters");
        c->sync();
        d.release();
        for (int k = 0; k < HDNNP_WORK_COUNTS; ++k) counts[k] = (int64_t)h[k];
    });
}

}  // extern "C"
