// Device-side model tables shared by the host setup (model.cpp) and the HIP
// kernels.  Plain structs of doubles and ints, uploaded once per context.
//
// Layout (one flat buffer per table):
//   ElemDesc[ne]        per centre element: SF count, group ranges, NN
//   CutCfg[..]          distinct (cutoff radius) configurations per element
//   GroupDesc[..]       SF groups, split into chunks of <= kMaxMembers
//   MemberDesc[..]      group members (one per SF of a group chunk)
//   ScaleDesc[..]       per SF scaling (SymFnc::scale)
//   pair lists          per (centre, e_j, e_k) the angular group ids
//   NN weights          per element, per layer row-major W[k][j] then b[k]
#ifndef HDNNP_TABLES_H
#define HDNNP_TABLES_H

#ifdef __HIPCC_RTC__
// Runtime-compiled (hipRTC) builds have no C headers.
typedef unsigned int uint32_t;
typedef unsigned long long uint64_t;
typedef long long int64_t;
#else
#include <stdint.h>
#endif

namespace hdnnp {

constexpr int kMaxElements   = 8;
constexpr int kMaxMembers    = 16;   // SF members per device group chunk
constexpr int kMaxLayers     = 8;
constexpr int kMaxCutCfg     = 4;    // distinct cutoff radii per element
constexpr int kMaxPairGroups = 8;    // angular group chunks per element pair
// Pair record of the mixed path: group chunks per element whose passing
// neighbour pairs are recorded, and the count that marks a chunk whose
// pairs were not recorded (the force kernel walks them again).
constexpr int kRecordGroups = 64;
constexpr unsigned short kRecordMissing = 0xffff;

// CutoffFunction.h:42-72 numbering.
enum CutType { CT_HARD = 0, CT_COS, CT_TANHU, CT_TANH, CT_EXP, CT_POLY1,
               CT_POLY2, CT_POLY3, CT_POLY4 };
// SymFnc.h ScalingType numbering.
enum ScaleType { ST_NONE = 0, ST_SCALE, ST_CENTER, ST_SCALECENTER,
                 ST_SCALESIGMA };
// NeuralNetwork.h ActivationFunction numbering (AF_UNSET = 0).
enum ActType { AF_UNSET = 0, AF_IDENTITY, AF_TANH, AF_LOGISTIC, AF_SOFTPLUS,
               AF_RELU, AF_GAUSSIAN, AF_COS, AF_REVLOGISTIC, AF_EXP,
               AF_HARMONIC };

// Symmetry-function types: 2 and 12 are radial, 3, 9 and 13 angular; 12
// and 13 are the element-weighted variants (every neighbour counts, times
// its atomic number; SymFncExpRadWeighted, SymFncExpAngnWeighted).
// Types 20-25 are the compact (polynomial) families, evaluated by their own
// kernels: 20 radial, 21 narrow and 22 wide angular, 23-25 the weighted
// variants.
constexpr bool sf_compact(int type) { return type >= 20 && type <= 25; }
constexpr bool sf_radial(int type)
{
    return type == 2 || type == 12 || type == 20 || type == 23;
}
constexpr bool sf_weighted(int type)
{
    return type == 12 || type == 13 || (type >= 23 && type <= 25);
}

// Members per compact group (SymGrpComp*: all compact SFs of one type and
// centre / neighbour elements).  The compact kernels keep one row of member
// sums per lane in LDS, T.maxCompMembers doubles wide; 96 members fill
// 48 KiB of the 64 KiB a workgroup may use.
constexpr int kMaxCompMembers = 96;

struct CutCfg
{
    int    type;
    int    pad;
    double rc, rcinv, rci, iw;
    double tanhpre;   // ((e + 1/e) / (e - 1/e))^3 (CutoffFunction.cpp:28)
};

struct GroupDesc
{
    int    type;      // 2, 3, 9, 12 or 13
    int    e1, e2;    // neighbour elements (e2 = -1 for radial; both -1 for
                      // the weighted types)
    int    nmem;      // members in this chunk (<= kMaxMembers)
    int    mem0;      // first member in the member table
    int    cfg;       // index into the element's CutCfg range (local)
    int    run0;      // angular: first lambda run in the run table
    int    nrun;      // angular: lambda runs of this chunk
    double rc, rc2;
};

// Highest level l of the shared power chain (zeta = 2^l, i.e. zeta <= 64).
constexpr int kMaxLevel = 6;

// An (eta, rs, lambda) run of an angular group chunk.  Its members are
// stored chain members first, by level (zeta = 2^level, level <= kMaxLevel),
// then the members with any other zeta ("generic", pow_int or pow).
// Member indices are chunk-local.
struct LamRun
{
    double lambda, eta, rs;
    int    newexp;                 // first run of an (eta, rs) run
    int    m0;                     // first member
    int    lmax;                   // highest level with members (-1: none)
    int    mend;                   // one past the last member
    int    lvEnd[kMaxLevel + 1];   // end of the members of each level
    int    pad;
};

struct MemberDesc
{
    double eta, rs, lambda, zeta;
    double zl;        // zeta * lambda
    double fnorm;     // angular: 2^(1-zeta) * scalingFactor (factorNorm
                      // after SymGrpBaseExpAng::setScalingFactors)
    double fderiv;    // angular: 2 eta / zeta / lambda
    double gfac;      // angular: fnorm / scalingFactor (applied to the sum)
    double sfac;      // scalingFactor
    int    zint;      // round(zeta)
    int    useint;    // integer power path (SymFncBaseExpAng.cpp:82-90)
    int    calcexp;   // recompute the exponential for this member
    int    sf;        // SF index within the centre element (NN input slot)
    int    newlam;    // first member of an (eta, rs, lambda) run
    int    level;     // zeta = 2^level (integer power), else -1
};

// A compact group's member: open radial interval (rl, rc) with its compact
// function (x - center) * scale, and for the angular types the open angle
// interval (al, ar) in radians with its compact function
// (CompactFunction.cpp:57-72, SymFncBaseCompAng.cpp:57-110).
struct CompMember
{
    double rl, rc, rcenter, rscale;
    double al, ar, acenter, ascale;
    double sfac;      // scalingFactor
    int    core;      // CT_POLY1..CT_POLY4 or CT_EXP (CoreFunction type)
    int    asym;      // asymmetric radial core (subtype "p?a")
    int    sf;        // SF index within the centre element
    int    newrad;    // (rl, rc) differ from the previous member's
                      // (calculateComp of SymGrpCompAngn.cpp:98-112)
};

struct CompGroup
{
    int    type;      // 20-25
    int    e1, e2;    // neighbour elements (-1 where not used)
    int    nmem, mem0;
    int    pad;
    double rmax;      // largest member rc
};

struct ScaleDesc
{
    int    type;
    int    pad;
    double smin, sfac, gmin, gmax, gmean;
};

struct ElemDesc
{
    int nsf;
    int g0, ng;          // group chunks [g0, g0+ng) (radial first)
    int nrad;            // of which radial chunks: [g0, g0+nrad)
    int cfg0, ncfg;      // cutoff configurations
    int sf0;             // first ScaleDesc of this element
    int cg0, ncg;        // compact groups [cg0, cg0+ncg)
    int nsfd[kMaxElements];   // SF derivatives per neighbour element
                              // (Element::getSymmetryFunctionNumTable)
    // network
    int nlayers;         // including input and output layer
    int width[kMaxLayers];
    int act[kMaxLayers];
    int woff[kMaxLayers];   // offset of layer l's W (l >= 1) in weights
    int boff[kMaxLayers];   // offset of layer l's b
    int normalize;          // normalize_nodes
    int maxwidth;           // widest hidden/output layer
};

struct ModelTables
{
    int ne;
    int maxnsf;
    int cutType;          // cutoff_type of the settings (global in 2G)
    int genpow;           // some angular member has a non-integer zeta
    int maxGroupMembers;  // largest group chunk
    int weighted;         // some group is of a weighted type (12, 13)
    double maxCutoff;
    double zw[kMaxElements];      // atomic number per element (weights of
                                  // the weighted types)
    const ElemDesc*   elem;
    const CutCfg*     cfg;
    const GroupDesc*  group;
    const MemberDesc* member;
    const LamRun*     run;
    const ScaleDesc*  scale;
    // angular group chunks per (centre c, e_j, e_k): block pb = (c*ne+ej)*ne+ek
    const int*        pairGroups;   // [pb * kMaxPairGroups + t], global ids
    const int*        pairCount;    // [pb]
    const double*     weights;
    const float*      weights32;    // the same weights rounded to FP32
    const CompGroup*  cgroup;       // compact groups (may be empty)
    const CompMember* cmember;
    int               ncgroup;      // compact groups over all elements
    int               maxCompMembers;   // largest compact group (>= 1)
};

}  // namespace hdnnp

#endif
