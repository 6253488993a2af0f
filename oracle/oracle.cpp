// Oracle: CPU restatement of n2p2's 2G-HDNNP energy + force path.
//
// TEST INFRASTRUCTURE ONLY -- the checker for the HIP engine, never part of
// the product.  It follows the reference loop-for-loop; each function cites
// the reference source it restates (paths relative to n2p2's src/libnnp
// unless stated otherwise).  Built with g++ -O3 -march=native -fopenmp
// -ffp-contract=off; the one multiply-add pattern the reference build
// contracts on the neighbour-list path (Vec3D::norm2) is written out with
// std::fma below so the neighbour sets match GCC -O3 -march=native builds of
// the reference bit for bit.
//
// Parity status of each function (how it is pinned) is listed in tests/ and
// the design notes; every function here has at least one independent check.

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using std::size_t;
using std::string;
using std::vector;

namespace {

thread_local string g_err;

// ---------------------------------------------------------------- strings --
// utility.cpp: reduce() trims and collapses whitespace, split() on ' '.
string reduce(string const& s)
{
    string out;
    bool sp = false;
    for (char c : s)
    {
        bool w = (c == ' ' || c == '\t' || c == '\r' || c == '\n');
        if (w) { sp = !out.empty(); continue; }
        if (sp) { out.push_back(' '); sp = false; }
        out.push_back(c);
    }
    return out;
}

vector<string> split(string const& s)
{
    vector<string> v;
    std::istringstream is(s);
    string w;
    while (is >> w) v.push_back(w);
    return v;
}

// ------------------------------------------------------------ element map --
// ElementMap.cpp:36-54: symbols registered from the "elements" keyword and
// sorted by atomic number; element index = rank.
char const* const kSymbols[] = {
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al",
    "Si", "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe",
    "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr",
    "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn",
    "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm",
    "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",
    "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn",
    "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf",
    "Es", "Fm", "Md", "No"};

int atomicNumber(string const& sym)
{
    for (size_t i = 0; i < sizeof(kSymbols) / sizeof(*kSymbols); ++i)
        if (sym == kSymbols[i]) return (int)i + 1;
    throw std::runtime_error("ERROR: Element \"" + sym + "\" unknown.\n");
}

// ------------------------------------------------------------------- fma --
// Vec3D.h:299-302 norm2() = x*x + y*y + z*z.  GCC -O3 -march=native emits
// t = y*y; t = fma(x, x, t); t = fma(z, z, t) for it (observed in the
// x86 assembly of an -O3 -march=native build); this is that sequence.
inline double norm2_pinned(double const* r)
{
    double t = r[1] * r[1];
    t = std::fma(r[0], r[0], t);
    return std::fma(r[2], r[2], t);
}

// -------------------------------------------------------- cutoff function --
// CutoffFunction.h:42-72 enum and CutoffFunction.cpp:95-213, CoreFunction.h.
enum CutoffType { CT_HARD, CT_COS, CT_TANHU, CT_TANH, CT_EXP, CT_POLY1,
                  CT_POLY2, CT_POLY3, CT_POLY4 };

struct Cutoff
{
    int    type = CT_HARD;
    double rc = 0.0, rcinv = 0.0, rci = 0.0, alpha = 0.0, iw = 0.0;

    static double PI() { return 4.0 * std::atan(1.0); }
    static double E() { return std::exp(1.0); }
    static double TANH_PRE()
    {
        double e = E();
        return std::pow((e + 1 / e) / (e - 1 / e), 3);
    }

    void set(int t, double r, double a)
    {
        type = t;
        rc = r;
        rcinv = 1.0 / r;                 // CutoffFunction.cpp:95-100
        alpha = a;
        rci = rc * alpha;                // CutoffFunction.cpp:102-112
        iw = 1.0 / (rc - rci);
    }

    // CoreFunction.h:193-318 core polynomials / exp in x in [0,1).
    static void core(int t, double x, double& fx, double& dfx)
    {
        if (t == CT_POLY1)
        {
            fx = (2.0 * x - 3.0) * x * x + 1.0;
            dfx = x * (6.0 * x - 6.0);
        }
        else if (t == CT_POLY2)
        {
            double const x2 = x * x;
            fx = ((15.0 - 6.0 * x) * x - 10.0) * x * x2 + 1.0;
            dfx = x2 * ((60.0 - 30.0 * x) * x - 30.0);
        }
        else if (t == CT_POLY3)
        {
            double const x2 = x * x;
            fx = (x * (x * (20.0 * x - 70.0) + 84.0) - 35.0) * x2 * x2 + 1.0;
            dfx = x2 * x * (x * (x * (140.0 * x - 420.0) + 420.0) - 140.0);
        }
        else if (t == CT_POLY4)
        {
            double x4 = x * x;
            x4 *= x4;
            fx = (x * (x * ((315.0 - 70.0 * x) * x - 540.0) + 420.0) - 126.0)
               * x * x4 + 1.0;
            dfx = x4 * (x * (x * ((2520.0 - 630.0 * x) * x - 3780.0)
                + 2520.0) - 630.0);
        }
        else // CT_EXP
        {
            double const temp = 1.0 / (x * x - 1.0);
            double const temp2 = std::exp(temp);
            fx = E() * temp2;
            dfx = -2.0 * E() * x * temp * temp * temp2;
        }
    }

    // CutoffFunction.h:210-220 fdf(): 0 beyond rc, else per-type.
    void fdf(double r, double& fc, double& dfc) const
    {
        if (r >= rc) { fc = 0.0; dfc = 0.0; return; }
        if (type == CT_HARD) { fc = 1.0; dfc = 0.0; return; }
        if (type == CT_COS)  // CutoffFunction.cpp:128-141
        {
            if (r < rci) { fc = 1.0; dfc = 0.0; return; }
            double const x = (r - rci) * iw;
            double const temp = std::cos(PI() * x);
            fc = 0.5 * (temp + 1.0);
            dfc = -0.5 * iw * PI() * std::sqrt(1.0 - temp * temp);
            return;
        }
        if (type == CT_TANHU)  // CutoffFunction.cpp:155-161
        {
            double const temp = std::tanh(1.0 - r * rcinv);
            double const temp2 = temp * temp;
            fc = temp * temp2;
            dfc = 3.0 * temp2 * (temp2 - 1.0) * rcinv;
            return;
        }
        if (type == CT_TANH)  // CutoffFunction.cpp:177-183
        {
            double const temp = std::tanh(1.0 - r * rcinv);
            double const temp2 = temp * temp;
            fc = TANH_PRE() * temp * temp2;
            dfc = 3.0 * TANH_PRE() * temp2 * (temp2 - 1.0) * rcinv;
            return;
        }
        // CutoffFunction.cpp:199-213 fdfCORE
        if (r < rci) { fc = 1.0; dfc = 0.0; return; }
        double const x = (r - rci) * iw;
        core(type, x, fc, dfc);
        dfc *= iw;
    }
};

// ------------------------------------------------------- symmetry function --
enum ScalingType { ST_NONE, ST_SCALE, ST_CENTER, ST_SCALECENTER,
                   ST_SCALESIGMA };

// Types 2 and 12 are radial; 12 and 13 are the element-weighted variants
// (SymFncExpRadWeighted, SymFncExpAngnWeighted): every neighbour counts,
// weighted by its atomic number.
// Types 20-25 are the compact (polynomial) families: 20 radial, 21 narrow
// and 22 wide angular, 23-25 their weighted variants (SymFncCompRad,
// SymFncCompAngn, SymFncCompAngw, SymFncComp*Weighted).
inline bool isCompact(int type) { return type >= 20 && type <= 25; }
inline bool isRadial(int type)
{
    return type == 2 || type == 12 || type == 20 || type == 23;
}
inline bool isWeighted(int type)
{
    return type == 12 || type == 13 || (type >= 23 && type <= 25);
}

// CompactFunction (CompactFunction.h:139-160, CompactFunction.cpp:42-72)
// with its CoreFunction (CoreFunction.h:193-318): a = (x - center) / width,
// f = core(|a|), f' = core'(|a|) sign(a) / width; the asymmetric core maps
// its argument to (2 - y) y first (CoreFunction.h:178-191).
struct Compact
{
    double center = 0.0, width = 1.0, scale = 1.0;
    int    core = CT_POLY2;
    bool   asym = false;
    void setLeftRight(double left, double right)
    {
        center = (right + left) / 2.0;
        width = (right - left) / 2.0;
        scale = 1.0 / width;
    }
    void fdf(double x, double& f, double& df) const;
};

void Compact::fdf(double x, double& f, double& df) const
{
    double const a = (x - center) * scale;
    double y = std::fabs(a);
    if (asym)
    {
        Cutoff::core(core, (2.0 - y) * y, f, df);
        df *= 2.0 - 2.0 * y;
    }
    else Cutoff::core(core, y, f, df);
    df *= std::copysign(scale, a);
}

struct SF
{
    int    type = 0;
    int    ec = 0, e1 = 0, e2 = 0;
    double eta = 0.0, rs = 0.0, lambda = 0.0, zeta = 0.0, rc = 0.0;
    int    zetaInt = 0;
    bool   useIntegerPow = false;
    int    cutoffType = CT_HARD;
    double cutoffAlpha = 0.0;
    int    index = 0;
    size_t line = 0;
    vector<size_t> indexPerElement;
    // compact types: radial interval (rl, rc), angle interval in degrees,
    // subtype string and the radial / angular compact functions
    double rl = 0.0, angleLeft = 0.0, angleRight = 0.0;
    string subtype;
    Compact cr, ca;
    // SymFnc.cpp:52-98 scaling
    int    scalingType = ST_NONE;
    double Gmin = 0, Gmax = 0, Gmean = 0, Gsigma = 0, Smin = 0, Smax = 0;
    double scalingFactor = 1.0;

    // SymFncExpRad.cpp:230, SymFncBaseExpAng.cpp:179,
    // SymFncExpRadWeighted.cpp:213-216, SymFncExpAngnWeighted.cpp:338-341
    bool relevant(size_t e) const
    {
        if (isWeighted(type)) return true;
        if (type == 2 || type == 20) return (int)e == e1;
        return (int)e == e1 || (int)e == e2;
    }

    // SymFnc.cpp:169-195
    double scale(double v) const
    {
        switch (scalingType)
        {
        case ST_NONE: return v;
        case ST_SCALE: return Smin + scalingFactor * (v - Gmin);
        case ST_CENTER: return v - Gmean;
        case ST_SCALECENTER: return Smin + scalingFactor * (v - Gmean);
        case ST_SCALESIGMA: return Smin + scalingFactor * (v - Gmean);
        }
        return 0.0;
    }
    // SymFnc.cpp:197-223
    double unscale(double v) const
    {
        switch (scalingType)
        {
        case ST_NONE: return v;
        case ST_SCALE: return (v - Smin) / scalingFactor + Gmin;
        case ST_CENTER: return v + Gmean;
        case ST_SCALECENTER: return (v - Smin) / scalingFactor + Gmean;
        case ST_SCALESIGMA: return (v - Smin) / scalingFactor + Gmean;
        }
        return 0.0;
    }
};

// SymFncExpRad.cpp:55-75 and SymFncExpAngn.cpp:51-77 / SymFncExpAngw.cpp
// (identical field order for types 3 and 9); SymFncExpRadWeighted.cpp and
// SymFncExpAngnWeighted.cpp order the same fields without the neighbour
// elements (e1 = e2 = -1 here): strict weak ordering.
bool sfLess(SF const& a, SF const& b)
{
    if (a.ec != b.ec) return a.ec < b.ec;
    if (a.type != b.type) return a.type < b.type;
    if (isCompact(a.type))
    {
        // SymFncCompRad.cpp / SymFncCompAngn.cpp / SymFncCompAngw.cpp and
        // the weighted variants: subtype, [e1, e2], rc, rl, [angles].
        if (a.subtype != b.subtype) return a.subtype < b.subtype;
        if (!isWeighted(a.type))
        {
            if (a.e1 != b.e1) return a.e1 < b.e1;
            if (a.type != 20 && a.e2 != b.e2) return a.e2 < b.e2;
        }
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.rl != b.rl) return a.rl < b.rl;
        if (!isRadial(a.type))
        {
            if (a.angleLeft != b.angleLeft) return a.angleLeft < b.angleLeft;
            if (a.angleRight != b.angleRight)
                return a.angleRight < b.angleRight;
        }
        return false;
    }
    if (a.cutoffType != b.cutoffType) return a.cutoffType < b.cutoffType;
    if (a.cutoffAlpha != b.cutoffAlpha) return a.cutoffAlpha < b.cutoffAlpha;
    if (a.rc != b.rc) return a.rc < b.rc;
    if (a.eta != b.eta) return a.eta < b.eta;
    if (a.rs != b.rs) return a.rs < b.rs;
    if (!isRadial(a.type))
    {
        if (a.zeta != b.zeta) return a.zeta < b.zeta;
        if (a.lambda != b.lambda) return a.lambda < b.lambda;
    }
    if (a.e1 != b.e1) return a.e1 < b.e1;
    if (a.type != 2 && a.e2 != b.e2) return a.e2 < b.e2;
    return false;
}

// Per-group precomputed member tables: SymGrpExpRad.cpp:111-136,
// SymGrpExpAngn.cpp:107-148, SymGrpExpAngw.cpp (same), SymGrpBaseExpAng.cpp:45-57.
struct Group
{
    int    type = 0;
    int    e1 = 0, e2 = 0;
    double rc = 0.0;
    int    cutoffType = CT_HARD;
    double cutoffAlpha = 0.0;
    Cutoff fc;
    vector<int> members;             // indices into element's SF list
    vector<double> factorNorm, factorDeriv, scalingFactors;
    vector<double> weight;           // weighted types: atomic number per
                                     // neighbour element
    double rmin = 0.0, rmax = 0.0;   // compact types: radial range of the
                                     // members (SymGrpCompRad.cpp:82-84)
    vector<char> calculateExp;
};

struct NN
{
    int numLayers = 0;
    vector<int> n;       // neurons per layer
    vector<int> act;     // activation per layer
    bool normalize = false;
    // w[l][k][j]: layer l (1..L-1), neuron k, input j; b[l][k]
    vector<vector<vector<double>>> w;
    vector<vector<double>> b;
};

enum Act { AF_UNSET, AF_IDENTITY, AF_TANH, AF_LOGISTIC, AF_SOFTPLUS, AF_RELU,
           AF_GAUSSIAN, AF_COS, AF_REVLOGISTIC, AF_EXP, AF_HARMONIC };

// NeuralNetwork.cpp:1078-1098
int activationFromString(string const& c)
{
    if (c == "l") return AF_IDENTITY;
    if (c == "t") return AF_TANH;
    if (c == "s") return AF_LOGISTIC;
    if (c == "p") return AF_SOFTPLUS;
    if (c == "r") return AF_RELU;
    if (c == "g") return AF_GAUSSIAN;
    if (c == "c") return AF_COS;
    if (c == "S") return AF_REVLOGISTIC;
    if (c == "e") return AF_EXP;
    if (c == "h") return AF_HARMONIC;
    throw std::runtime_error("ERROR: Unknown activation function.\n");
}

struct ElementData
{
    string symbol;
    int    Z = 0;
    double atomicEnergyOffset = 0.0;
    vector<SF> sfs;
    vector<Group> groups;
    vector<vector<size_t>> table;   // Element.cpp:289-311
    vector<size_t> numTable;
    NN nn;
};

struct Neighbor
{
    size_t index = 0;
    size_t element = 0;
    double d = 0.0;
    double dr[3] = {0, 0, 0};
    vector<double> dGdr;   // 3 * numSFD[element]
};

struct Atom
{
    size_t index = 0;
    size_t element = 0;
    double r[3] = {0, 0, 0};
    double f[3] = {0, 0, 0};
    double energy = 0.0;
    vector<Neighbor> neighbors;
    vector<size_t> neighborsUnique;
    vector<double> G, dEdG, dGdr;   // dGdr: 3 * nsf
};

}  // namespace

struct orc_model
{
    vector<ElementData> el;
    double maxCutoffRadius = 0.0;
    bool   normalize = false;
    double meanEnergy = 0.0, convEnergy = 1.0, convLength = 1.0;
    int    cutoffType = CT_HARD;
    double cutoffAlpha = 0.0;
};

struct orc_struct
{
    vector<Atom> atoms;
    bool   periodic = false;
    double box[3][3] = {{0}};
    double invbox[3][3] = {{0}};
    int    pbc[3] = {0, 0, 0};
    double energy = 0.0;
    size_t numElements = 0;
    vector<size_t> numAtomsPerElement;
};

namespace {

// Structure.cpp:519-545
void calculateInverseBox(orc_struct& s)
{
    auto& box = s.box;
    double invdet = box[0][0] * box[1][1] * box[2][2]
                  + box[1][0] * box[2][1] * box[0][2]
                  + box[2][0] * box[0][1] * box[1][2]
                  - box[2][0] * box[1][1] * box[0][2]
                  - box[0][0] * box[2][1] * box[1][2]
                  - box[1][0] * box[0][1] * box[2][2];
    invdet = 1.0 / invdet;
    auto& ib = s.invbox;
    ib[0][0] = box[1][1] * box[2][2] - box[2][1] * box[1][2];
    ib[0][1] = box[2][1] * box[0][2] - box[0][1] * box[2][2];
    ib[0][2] = box[0][1] * box[1][2] - box[1][1] * box[0][2];
    ib[1][0] = box[2][0] * box[1][2] - box[1][0] * box[2][2];
    ib[1][1] = box[0][0] * box[2][2] - box[2][0] * box[0][2];
    ib[1][2] = box[1][0] * box[0][2] - box[0][0] * box[1][2];
    ib[2][0] = box[1][0] * box[2][1] - box[2][0] * box[1][1];
    ib[2][1] = box[2][0] * box[0][1] - box[0][0] * box[2][1];
    ib[2][2] = box[0][0] * box[1][1] - box[1][0] * box[0][1];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ib[i][j] *= invdet;
}

// Structure.cpp:1223-1247
void remapAtom(orc_struct const& s, Atom& a)
{
    double f[3];
    // Vec3D f = r0*invbox[0] + r1*invbox[1] + r2*invbox[2] (left to right)
    for (int c = 0; c < 3; ++c)
    {
        double t = a.r[0] * s.invbox[0][c];
        t = t + a.r[1] * s.invbox[1][c];
        t = t + a.r[2] * s.invbox[2][c];
        f[c] = t;
    }
    for (int c = 0; c < 3; ++c) if (f[c] > 1.0) f[c] -= (int)f[c];
    for (int c = 0; c < 3; ++c) if (f[c] < 0.0) f[c] += 1.0 - (int)f[c];
    for (int c = 0; c < 3; ++c) if (f[c] == 1.0) f[c] = 0.0;
    for (int c = 0; c < 3; ++c)
    {
        double t = f[0] * s.box[0][c];
        t = t + f[1] * s.box[1][c];
        t = t + f[2] * s.box[2][c];
        a.r[c] = t;
    }
}

// Vec3D cross/normalize and Structure.cpp:486-517 calculatePbcCopies.
void calculatePbcCopies(orc_struct& s, double cutoffRadius)
{
    auto cross = [](double const* a, double const* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto normalize = [](double* v) {
        double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= n; v[1] /= n; v[2] /= n;
    };
    auto dot = [](double const* a, double const* b) {
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    };
    double axb[3], axc[3], bxc[3];
    cross(s.box[0], s.box[1], axb); normalize(axb);
    cross(s.box[0], s.box[2], axc); normalize(axc);
    cross(s.box[1], s.box[2], bxc); normalize(bxc);
    double proj[3] = {std::fabs(dot(s.box[0], bxc)),
                      std::fabs(dot(s.box[1], axc)),
                      std::fabs(dot(s.box[2], axb))};
    for (int k = 0; k < 3; ++k)
    {
        s.pbc[k] = 0;
        while (s.pbc[k] * proj[k] <= cutoffRadius) s.pbc[k]++;
    }
}

inline void pushNeighbor(orc_struct& s, size_t i, size_t j, double const* dr)
{
    Atom& a = s.atoms[i];
    a.neighbors.emplace_back();
    Neighbor& n = a.neighbors.back();
    n.index = j;
    n.element = s.atoms[j].element;
    n.dr[0] = dr[0]; n.dr[1] = dr[1]; n.dr[2] = dr[2];
    n.d = std::sqrt(norm2_pinned(dr));
    if (a.neighborsUnique.back() != j && i != j) a.neighborsUnique.push_back(j);
}

// dr = r_i - r_j + bc0*box0 + bc1*box1 + bc2*box2, evaluated left to right
// (Structure.cpp:319-322).
inline void imageVector(orc_struct const& s, size_t i, size_t j, int b0,
                        int b1, int b2, double* dr)
{
    for (int c = 0; c < 3; ++c)
    {
        double t = s.atoms[i].r[c] - s.atoms[j].r[c];
        t = t + b0 * s.box[0][c];
        t = t + b1 * s.box[1][c];
        t = t + b2 * s.box[2][c];
        dr[c] = t;
    }
}

// Structure.cpp:287-399 (sortByDistance = false).
void neighborListBrute(orc_struct& s, double cutoffRadius)
{
    size_t const N = s.atoms.size();
    for (auto& a : s.atoms)
    {
        a.neighbors.clear();
        a.neighborsUnique.clear();
    }
    if (s.periodic)
    {
        calculatePbcCopies(s, cutoffRadius);
        double const rc2 = cutoffRadius * cutoffRadius;
#pragma omp parallel for schedule(dynamic, 16)
        for (size_t i = 0; i < N; ++i)
        {
            s.atoms[i].neighborsUnique.push_back(i);
            for (size_t j = 0; j < N; ++j)
                for (int b0 = -s.pbc[0]; b0 <= s.pbc[0]; ++b0)
                    for (int b1 = -s.pbc[1]; b1 <= s.pbc[1]; ++b1)
                        for (int b2 = -s.pbc[2]; b2 <= s.pbc[2]; ++b2)
                        {
                            if (i == j && b0 == 0 && b1 == 0 && b2 == 0)
                                continue;
                            double dr[3];
                            imageVector(s, i, j, b0, b1, b2, dr);
                            if (norm2_pinned(dr) <= rc2)
                                pushNeighbor(s, i, j, dr);
                        }
        }
    }
    else
    {
        double const rc2 = cutoffRadius * cutoffRadius;
#pragma omp parallel for schedule(dynamic, 16)
        for (size_t i = 0; i < N; ++i)
        {
            s.atoms[i].neighborsUnique.push_back(i);
            for (size_t j = 0; j < N; ++j)
            {
                if (i == j) continue;
                double dr[3];
                for (int c = 0; c < 3; ++c)
                    dr[c] = s.atoms[i].r[c] - s.atoms[j].r[c];
                if (norm2_pinned(dr) <= rc2) pushNeighbor(s, i, j, dr);
            }
        }
    }
}

// Same set and order as neighborListBrute, found through a cell list: each
// atom's candidates (j, image) come from the 27 surrounding cells, are
// tested with the identical dr / norm2 expression, and are then sorted by
// (j, bc0, bc1, bc2), which is the brute-force loop order.  Used to feed the
// CPU baseline at large sizes (the reference's O(N^2) list is timed
// separately); equivalence with the brute force is a test.
void neighborListCells(orc_struct& s, double cutoffRadius)
{
    if (!s.periodic) { neighborListBrute(s, cutoffRadius); return; }
    size_t const N = s.atoms.size();
    calculatePbcCopies(s, cutoffRadius);
    double const rc2 = cutoffRadius * cutoffRadius;
    // Perpendicular widths -> number of cells per direction (width >= rc).
    double proj[3];
    {
        double axb[3], axc[3], bxc[3];
        auto cross = [](double const* a, double const* b, double* o) {
            o[0] = a[1] * b[2] - a[2] * b[1];
            o[1] = a[2] * b[0] - a[0] * b[2];
            o[2] = a[0] * b[1] - a[1] * b[0];
        };
        cross(s.box[0], s.box[1], axb);
        cross(s.box[0], s.box[2], axc);
        cross(s.box[1], s.box[2], bxc);
        auto pr = [](double const* v, double const* nrm) {
            double n = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1]
                                 + nrm[2] * nrm[2]);
            return std::fabs(v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2])
                 / n;
        };
        proj[0] = pr(s.box[0], bxc);
        proj[1] = pr(s.box[1], axc);
        proj[2] = pr(s.box[2], axb);
    }
    int nc[3];
    for (int k = 0; k < 3; ++k)
    {
        nc[k] = (int)std::floor(proj[k] / (cutoffRadius * (1.0 + 1e-9)));
        if (nc[k] < 1) nc[k] = 1;
    }
    // Cells reach one cell in each direction only when a cell is at least
    // rc wide; with nc == 1 or 2 the +-1 wrap already covers every image
    // within |bc| <= 1, and pbc > 1 directions fall back to brute force.
    if (s.pbc[0] > 1 || s.pbc[1] > 1 || s.pbc[2] > 1)
    {
        neighborListBrute(s, cutoffRadius);
        return;
    }
    vector<int> cellOf(N);
    vector<vector<size_t>> cells((size_t)nc[0] * nc[1] * nc[2]);
    for (size_t j = 0; j < N; ++j)
    {
        double f[3];
        for (int c = 0; c < 3; ++c)
            f[c] = s.atoms[j].r[0] * s.invbox[0][c]
                 + s.atoms[j].r[1] * s.invbox[1][c]
                 + s.atoms[j].r[2] * s.invbox[2][c];
        int ci[3];
        for (int c = 0; c < 3; ++c)
        {
            ci[c] = (int)std::floor(f[c] * nc[c]);
            if (ci[c] < 0) ci[c] = 0;
            if (ci[c] >= nc[c]) ci[c] = nc[c] - 1;
        }
        int id = (ci[0] * nc[1] + ci[1]) * nc[2] + ci[2];
        cellOf[j] = id;
        cells[id].push_back(j);
    }
    for (auto& a : s.atoms)
    {
        a.neighbors.clear();
        a.neighborsUnique.clear();
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < N; ++i)
    {
        int id = cellOf[i];
        int ci[3] = {id / (nc[1] * nc[2]), (id / nc[2]) % nc[1], id % nc[2]};
        struct Hit { size_t j; int b[3]; };
        vector<Hit> hits;
        for (int d0 = -1; d0 <= 1; ++d0)
            for (int d1 = -1; d1 <= 1; ++d1)
                for (int d2 = -1; d2 <= 1; ++d2)
                {
                    int a[3] = {ci[0] + d0, ci[1] + d1, ci[2] + d2};
                    int cw[3], sh[3];
                    for (int c = 0; c < 3; ++c)
                    {
                        int q = a[c];
                        int w = ((q % nc[c]) + nc[c]) % nc[c];
                        sh[c] = (q - w) / nc[c];    // image at r_j + sh*box
                        cw[c] = w;
                    }
                    // Distinct offsets give distinct absolute cells a, and
                    // a determines (cell, shift) uniquely, so no image is
                    // visited twice even when nc is 1 or 2.
                    int cid = (cw[0] * nc[1] + cw[1]) * nc[2] + cw[2];
                    for (size_t j : cells[cid])
                    {
                        int b0 = -sh[0], b1 = -sh[1], b2 = -sh[2];
                        if (i == j && b0 == 0 && b1 == 0 && b2 == 0) continue;
                        double dr[3];
                        imageVector(s, i, j, b0, b1, b2, dr);
                        if (norm2_pinned(dr) <= rc2)
                            hits.push_back(Hit{j, {b0, b1, b2}});
                    }
                }
        std::sort(hits.begin(), hits.end(), [](Hit const& x, Hit const& y) {
            if (x.j != y.j) return x.j < y.j;
            if (x.b[0] != y.b[0]) return x.b[0] < y.b[0];
            if (x.b[1] != y.b[1]) return x.b[1] < y.b[1];
            return x.b[2] < y.b[2];
        });
        s.atoms[i].neighborsUnique.push_back(i);
        for (auto const& h : hits)
        {
            double dr[3];
            imageVector(s, i, h.j, h.b[0], h.b[1], h.b[2], dr);
            pushNeighbor(s, i, h.j, dr);
        }
    }
}

// utility.cpp:292-317
double pow_int(double x, int n)
{
    unsigned int m;
    if (n < 0) { x = 1.0 / x; m = -n; }
    else m = n;
    double result = 1.0;
    do
    {
        if (m & 1) result *= x;
        m >>= 1;
        x *= x;
    }
    while (m);
    return result;
}

// SymGrpExpRad.cpp:144-239; the weighted type 12 as
// SymGrpExpRadWeighted.cpp:144-252 (every neighbour, pexp times its atomic
// number).
void calcRadial(Group const& g, vector<SF> const& sfs, Atom& atom,
                bool derivatives)
{
    size_t const nm = g.members.size();
    bool const weighted = isWeighted(g.type);
    vector<double> result(nm, 0.0);
    for (size_t j = 0; j < atom.neighbors.size(); ++j)
    {
        Neighbor& n = atom.neighbors[j];
        if ((weighted || g.e1 == (int)n.element) && n.d < g.rc)
        {
            double const rij = n.d;
            double pfc, pdfc;
            g.fc.fdf(rij, pfc, pdfc);
            double const* const d1 = n.dr;
            for (size_t k = 0; k < nm; ++k)
            {
                SF const& sf = sfs[g.members[k]];
                double const eta = sf.eta, rs = sf.rs;
                double pexp = std::exp(-eta * (rij - rs) * (rij - rs));
                if (weighted) pexp = g.weight[n.element] * pexp;
                result[k] += pexp * pfc;
                if (!derivatives) continue;
                double const p1 = g.scalingFactors[k] * (pdfc - 2.0 * eta
                                * (rij - rs) * pfc) * pexp / rij;
                double const p1x = p1 * d1[0];
                double const p1y = p1 * d1[1];
                double const p1z = p1 * d1[2];
                size_t ki = sf.index;
                double* dGdr = &atom.dGdr[3 * ki];
                dGdr[0] += p1x; dGdr[1] += p1y; dGdr[2] += p1z;
                ki = sf.indexPerElement[n.element];
                dGdr = &n.dGdr[3 * ki];
                dGdr[0] -= p1x; dGdr[1] -= p1y; dGdr[2] -= p1z;
            }
        }
    }
    for (size_t k = 0; k < nm; ++k)
    {
        SF const& sf = sfs[g.members[k]];
        atom.G[sf.index] = sf.scale(result[k]);
    }
}

// SymGrpExpAngn.cpp:156-403 (narrow, type 3), SymGrpExpAngw.cpp:156-376
// (wide, type 9) and SymGrpExpAngnWeighted.cpp:166-385 (narrow over all
// neighbour pairs, type 13: the cutoff products times Z_j Z_k, the shift rs
// always applied).
void calcAngular(Group const& g, vector<SF> const& sfs, Atom& atom,
                 bool derivatives)
{
    bool const wide = (g.type == 9);
    bool const weighted = isWeighted(g.type);
    size_t const nm = g.members.size();
    vector<double> result(nm, 0.0);
    double const rc = g.rc;
    double const rc2 = rc * rc;
    size_t numNeighbors = atom.neighbors.size();
    if (numNeighbors == 0) numNeighbors = 1;
    int const e1 = g.e1, e2 = g.e2;
    for (size_t j = 0; j < numNeighbors - 1; j++)
    {
        Neighbor& nj = atom.neighbors[j];
        int const nej = (int)nj.element;
        double const rij = nj.d;
        if ((weighted || e1 == nej || e2 == nej) && rij < rc)
        {
            double const r2ij = rij * rij;
            double pfcij, pdfcij;
            g.fc.fdf(rij, pfcij, pdfcij);
            double const* const dr1 = nj.dr;
            for (size_t k = j + 1; k < numNeighbors; k++)
            {
                Neighbor& nk = atom.neighbors[k];
                int const nek = (int)nk.element;
                if (!weighted
                    && !((e1 == nej && e2 == nek) || (e2 == nej && e1 == nek)))
                    continue;
                double const rik = nk.d;
                if (!(rik < rc)) continue;
                double const* const dr2 = nk.dr;
                double dr3[3];
                dr3[0] = dr2[0] - dr1[0];
                dr3[1] = dr2[1] - dr1[1];
                dr3[2] = dr2[2] - dr1[2];
                double rjk = 0.0;
                double pfcjk = 1.0, pdfcjk = 0.0;
                if (!wide)
                {
                    rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
                    if (!(rjk < rc2)) continue;
                }
                double pfcik, pdfcik;
                g.fc.fdf(rik, pfcik, pdfcik);
                if (!wide)
                {
                    rjk = std::sqrt(rjk);
                    g.fc.fdf(rjk, pfcjk, pdfcjk);
                }
                double const rinvijik = 1.0 / rij / rik;
                double const costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1]
                                      + dr1[2] * dr2[2]) * rinvijik;
                double const r2ik = rik * rik;
                double pfc, r2sum, pr1, pr2, pr3 = 0.0;
                if (!wide)
                {
                    pfc = pfcij * pfcik * pfcjk;
                    r2sum = r2ij + r2ik + rjk * rjk;
                    pr1 = pfcik * pfcjk * pdfcij / rij;
                    pr2 = pfcij * pfcjk * pdfcik / rik;
                    pr3 = pfcij * pfcik * pdfcjk / rjk;
                    if (weighted)
                    {
                        double const z = g.weight[nej] * g.weight[nek];
                        pfc = z * pfcij * pfcik * pfcjk;
                        pr1 = z * pfcik * pfcjk * pdfcij / rij;
                        pr2 = z * pfcij * pfcjk * pdfcik / rik;
                        pr3 = z * pfcij * pfcik * pdfcjk / rjk;
                    }
                }
                else
                {
                    pfc = pfcij * pfcik;
                    r2sum = r2ij + r2ik;
                    pr1 = pfcik * pdfcij / rij;
                    pr2 = pfcij * pdfcik / rik;
                }
                double vexp = 0.0, rijs = 0.0, riks = 0.0, rjks = 0.0;
                for (size_t l = 0; l < nm; ++l)
                {
                    SF const& sf = sfs[g.members[l]];
                    bool const shifted = sf.rs > 0.0 || weighted;
                    if (g.calculateExp[l])
                    {
                        if (shifted)
                        {
                            rijs = rij - sf.rs;
                            riks = rik - sf.rs;
                            if (!wide)
                            {
                                rjks = rjk - sf.rs;
                                vexp = std::exp(-sf.eta * (rijs * rijs
                                                + riks * riks + rjks * rjks));
                            }
                            else
                            {
                                vexp = std::exp(-sf.eta * (rijs * rijs
                                                + riks * riks));
                            }
                        }
                        else vexp = std::exp(-sf.eta * r2sum);
                    }
                    double const plambda = 1.0 + sf.lambda * costijk;
                    double fg = vexp;
                    if (plambda <= 0.0) fg = 0.0;
                    else
                    {
                        if (sf.useIntegerPow)
                            fg *= pow_int(plambda, sf.zetaInt - 1);
                        else fg *= std::pow(plambda, sf.zeta - 1.0);
                    }
                    result[l] += fg * plambda * pfc;
                    if (!derivatives) continue;
                    fg *= g.factorNorm[l];
                    double const pfczl = pfc * sf.zeta * sf.lambda;
                    double const p2etapl = plambda * g.factorDeriv[l];
                    double p1, p2, p3;
                    if (!wide)
                    {
                        if (shifted)
                        {
                            p1 = fg * (pfczl * (rinvijik - costijk / r2ij
                               - p2etapl * rijs / rij) + pr1 * plambda);
                            p2 = fg * (pfczl * (rinvijik - costijk / r2ik
                               - p2etapl * riks / rik) + pr2 * plambda);
                            p3 = fg * (pfczl * (rinvijik + p2etapl * rjks
                               / rjk) - pr3 * plambda);
                        }
                        else
                        {
                            p1 = fg * (pfczl * (rinvijik - costijk / r2ij
                               - p2etapl) + pr1 * plambda);
                            p2 = fg * (pfczl * (rinvijik - costijk / r2ik
                               - p2etapl) + pr2 * plambda);
                            p3 = fg * (pfczl * (rinvijik + p2etapl)
                               - pr3 * plambda);
                        }
                    }
                    else
                    {
                        if (sf.rs > 0)
                        {
                            p1 = fg * (pfczl * (rinvijik - costijk / r2ij
                               - p2etapl * rijs / rij) + pr1 * plambda);
                            p2 = fg * (pfczl * (rinvijik - costijk / r2ik
                               - p2etapl * riks / rik) + pr2 * plambda);
                        }
                        else
                        {
                            p1 = fg * (pfczl * (rinvijik - costijk / r2ij
                               - p2etapl) + pr1 * plambda);
                            p2 = fg * (pfczl * (rinvijik - costijk / r2ik
                               - p2etapl) + pr2 * plambda);
                        }
                        p3 = fg * pfczl * rinvijik;
                    }
                    double const p1x = p1 * dr1[0], p1y = p1 * dr1[1],
                                 p1z = p1 * dr1[2];
                    double const p2x = p2 * dr2[0], p2y = p2 * dr2[1],
                                 p2z = p2 * dr2[2];
                    double const p3x = p3 * dr3[0], p3y = p3 * dr3[1],
                                 p3z = p3 * dr3[2];
                    size_t li = sf.index;
                    double* dGdr = &atom.dGdr[3 * li];
                    dGdr[0] += p1x + p2x;
                    dGdr[1] += p1y + p2y;
                    dGdr[2] += p1z + p2z;
                    li = sf.indexPerElement[nej];
                    dGdr = &nj.dGdr[3 * li];
                    dGdr[0] -= p1x + p3x;
                    dGdr[1] -= p1y + p3y;
                    dGdr[2] -= p1z + p3z;
                    li = sf.indexPerElement[nek];
                    dGdr = &nk.dGdr[3 * li];
                    dGdr[0] -= p2x - p3x;
                    dGdr[1] -= p2y - p3y;
                    dGdr[2] -= p2z - p3z;
                }
            }
        }
    }
    for (size_t l = 0; l < nm; ++l)
    {
        SF const& sf = sfs[g.members[l]];
        result[l] *= g.factorNorm[l] / g.scalingFactors[l];
        atom.G[sf.index] = sf.scale(result[l]);
    }
}

// SymGrpCompRad.cpp:128-214 (type 20) and SymGrpCompRadWeighted.cpp
// (type 23: every neighbour, times its atomic number).
void calcCompRad(Group const& g, vector<SF> const& sfs, Atom& atom,
                 bool derivatives)
{
    size_t const nm = g.members.size();
    bool const weighted = isWeighted(g.type);
    vector<double> result(nm, 0.0);
    for (size_t j = 0; j < atom.neighbors.size(); ++j)
    {
        Neighbor& n = atom.neighbors[j];
        double const rij = n.d;
        if ((weighted || g.e1 == (int)n.element) && rij < g.rmax
            && rij > g.rmin)
        {
            double const* const d1 = n.dr;
            double const z = weighted ? g.weight[n.element] : 1.0;
            for (size_t k = 0; k < nm; ++k)
            {
                SF const& sf = sfs[g.members[k]];
                if (rij <= sf.rl || rij >= sf.rc) continue;
                double rad, drad;
                sf.cr.fdf(rij, rad, drad);
                result[k] += weighted ? z * rad : rad;
                if (!derivatives || drad == 0.0) continue;
                drad *= weighted ? z * g.scalingFactors[k] / rij
                                 : g.scalingFactors[k] / rij;
                double const px = drad * d1[0], py = drad * d1[1],
                             pz = drad * d1[2];
                double* dGdr = &atom.dGdr[3 * (size_t)sf.index];
                dGdr[0] += px; dGdr[1] += py; dGdr[2] += pz;
                dGdr = &n.dGdr[3 * sf.indexPerElement[n.element]];
                dGdr[0] -= px; dGdr[1] -= py; dGdr[2] -= pz;
            }
        }
    }
    for (size_t k = 0; k < nm; ++k)
    {
        SF const& sf = sfs[g.members[k]];
        atom.G[sf.index] = sf.scale(result[k]);
    }
}

// SymGrpCompAngn.cpp:148-388 (type 21), SymGrpCompAngw.cpp (type 22, no
// r_jk factor or test) and the weighted SymGrpCompAngnWeighted.cpp /
// SymGrpCompAngwWeighted.cpp (types 24, 25: every pair, the angular part
// times Z_j Z_k).  The radial value at r_jk is recomputed only at a member
// whose (rl, rc) differ from the previous member's or after a skipped
// member, as the reference's group loop does (calculateComp).
void calcCompAng(Group const& g, vector<SF> const& sfs, Atom& atom,
                 bool derivatives)
{
    size_t const nm = g.members.size();
    bool const weighted = isWeighted(g.type);
    bool const wide = (g.type == 22 || g.type == 25);
    vector<double> result(nm, 0.0), radij(nm, 0.0), dradij(nm, 0.0);
    double const r2min = g.rmin * g.rmin, r2max = g.rmax * g.rmax;
    double const PI = 4.0 * std::atan(1.0);
    size_t numNeighbors = atom.neighbors.size();
    if (numNeighbors == 0) numNeighbors = 1;
    for (size_t j = 0; j < numNeighbors - 1; j++)
    {
        Neighbor& nj = atom.neighbors[j];
        int const nej = (int)nj.element;
        double const rij = nj.d;
        if (!((weighted || g.e1 == nej || g.e2 == nej) && rij < g.rmax
              && rij > g.rmin))
            continue;
        for (size_t l = 0; l < nm; ++l)
        {
            SF const& sf = sfs[g.members[l]];
            if (rij > sf.rl && rij < sf.rc)
                sf.cr.fdf(rij, radij[l], dradij[l]);
            else radij[l] = 0.0;
        }
        double const* const dr1 = nj.dr;
        for (size_t k = j + 1; k < numNeighbors; k++)
        {
            Neighbor& nk = atom.neighbors[k];
            int const nek = (int)nk.element;
            if (!weighted
                && !((g.e1 == nej && g.e2 == nek) || (g.e2 == nej && g.e1 == nek)))
                continue;
            double const rik = nk.d;
            if (!(rik < g.rmax && rik > g.rmin)) continue;
            double const* const dr2 = nk.dr;
            double const dr3[3] = {dr2[0] - dr1[0], dr2[1] - dr1[1],
                                   dr2[2] - dr1[2]};
            double rjk = 0.0, rinvjk = 0.0;
            if (!wide)
            {
                rjk = dr3[0] * dr3[0] + dr3[1] * dr3[1] + dr3[2] * dr3[2];
                if (rjk >= r2max || rjk <= r2min) continue;
                rjk = std::sqrt(rjk);
                rinvjk = 1.0 / rjk;
            }
            double const rinvijik = 1.0 / (rij * rik);
            double const costijk = (dr1[0] * dr2[0] + dr1[1] * dr2[1]
                                  + dr1[2] * dr2[2]) * rinvijik;
            if (costijk <= -1.0 || costijk >= 1.0) continue;
            double const acostijk = std::acos(costijk);
            double const rinvij = rinvijik * rik;
            double const rinvik = rinvijik * rij;
            double const phiijik0 = rinvij * (rinvik - rinvij * costijk);
            double const phiikij0 = rinvik * (rinvij - rinvik * costijk);
            double dacostijk = 0.0;
            if (derivatives) dacostijk = -1.0 / std::sqrt(1.0 - costijk * costijk);
            double const z = weighted ? g.weight[nej] * g.weight[nek] : 1.0;
            bool skip = false;
            double radjk = 1.0, dradjk = 0.0;
            for (size_t l = 0; l < nm; ++l)
            {
                SF const& sf = sfs[g.members[l]];
                double const mal = sf.angleLeft * PI / 180.0;
                double const mar = sf.angleRight * PI / 180.0;
                if (radij[l] == 0.0 || rik <= sf.rl || rik >= sf.rc
                    || (!wide && (rjk <= sf.rl || rjk >= sf.rc))
                    || acostijk <= mal || acostijk >= mar)
                {
                    skip = true;
                    continue;
                }
                double radik, dradik;
                sf.cr.fdf(rik, radik, dradik);
                if (!wide && (g.calculateExp[l] || skip))
                {
                    sf.cr.fdf(rjk, radjk, dradjk);
                    skip = false;
                }
                double ang, dang;
                sf.ca.fdf(acostijk, ang, dang);
                if (weighted) ang *= z;
                double rad = wide ? radij[l] * radik
                                  : radij[l] * radik * radjk;
                result[l] += rad * ang;
                if (!derivatives) continue;
                dang *= dacostijk;
                double const phiijik = phiijik0 * dang;
                double const phiikij = phiikij0 * dang;
                double const psiijik = rinvijik * dang;
                double chiij, chiik, chijk;
                if (!wide)
                {
                    chiij = rinvij * dradij[l] * radik * radjk;
                    chiik = rinvik * radij[l] * dradik * radjk;
                    chijk = -rinvjk * radij[l] * radik * dradjk;
                }
                else
                {
                    chiij = rinvij * radik * dradij[l];
                    chiik = rinvik * radij[l] * dradik;
                    chijk = 0.0;
                }
                double const sfac = g.scalingFactors[l];
                double p1, p2, p3;
                if (dang != 0.0 && rad != 0.0)
                {
                    rad *= weighted ? sfac * z : sfac;
                    ang *= sfac;
                    p1 = rad * phiijik + ang * chiij;
                    p2 = rad * phiikij + ang * chiik;
                    p3 = rad * psiijik + ang * chijk;
                }
                else if (ang != 0.0)
                {
                    ang *= sfac;
                    p1 = ang * chiij;
                    p2 = ang * chiik;
                    p3 = ang * chijk;
                }
                else continue;
                double const p1x = p1 * dr1[0], p1y = p1 * dr1[1],
                             p1z = p1 * dr1[2];
                double const p2x = p2 * dr2[0], p2y = p2 * dr2[1],
                             p2z = p2 * dr2[2];
                double const p3x = p3 * dr3[0], p3y = p3 * dr3[1],
                             p3z = p3 * dr3[2];
                double* dGdr = &atom.dGdr[3 * (size_t)sf.index];
                dGdr[0] += p1x + p2x;
                dGdr[1] += p1y + p2y;
                dGdr[2] += p1z + p2z;
                dGdr = &nj.dGdr[3 * sf.indexPerElement[nej]];
                dGdr[0] -= p1x + p3x;
                dGdr[1] -= p1y + p3y;
                dGdr[2] -= p1z + p3z;
                dGdr = &nk.dGdr[3 * sf.indexPerElement[nek]];
                dGdr[0] -= p2x - p3x;
                dGdr[1] -= p2y - p3y;
                dGdr[2] -= p2z - p3z;
            }
        }
    }
    for (size_t l = 0; l < nm; ++l)
    {
        SF const& sf = sfs[g.members[l]];
        atom.G[sf.index] = sf.scale(result[l]);
    }
}

// Mode.cpp:1561-1640 (no SF cache; the cache only avoids recomputing the
// same cutoff values) with Atom::allocate (Atom.cpp:168-240).
void symmetryFunctions(orc_struct& s, orc_model const& m, bool derivatives)
{
    size_t const N = s.atoms.size();
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t i = 0; i < N; ++i)
    {
        Atom& a = s.atoms[i];
        ElementData const& e = m.el[a.element];
        size_t const nsf = e.sfs.size();
        a.G.assign(nsf, 0.0);
        a.dGdr.clear();
        a.dEdG.clear();
        if (derivatives)
        {
            a.dEdG.assign(nsf, 0.0);
            a.dGdr.assign(3 * nsf, 0.0);
        }
        for (auto& n : a.neighbors)
        {
            n.dGdr.clear();
            if (derivatives) n.dGdr.assign(3 * e.numTable[n.element], 0.0);
        }
        for (auto const& g : e.groups)
        {
            if (g.type == 20 || g.type == 23)
                calcCompRad(g, e.sfs, a, derivatives);
            else if (isCompact(g.type))
                calcCompAng(g, e.sfs, a, derivatives);
            else if (isRadial(g.type)) calcRadial(g, e.sfs, a, derivatives);
            else calcAngular(g, e.sfs, a, derivatives);
        }
    }
}

// NeuralNetwork.cpp:761-896 propagateLayer (statistics side effects
// omitted), keeping per-neuron value and dfdx.
void propagate(NN const& nn, double const* input, vector<vector<double>>& val,
               vector<vector<double>>& dfdx)
{
    val.assign(nn.numLayers, {});
    dfdx.assign(nn.numLayers, {});
    val[0].assign(input, input + nn.n[0]);
    for (int l = 1; l < nn.numLayers; ++l)
    {
        int const nl = nn.n[l], np = nn.n[l - 1];
        val[l].assign(nl, 0.0);
        dfdx[l].assign(nl, 0.0);
        for (int i = 0; i < nl; ++i)
        {
            double dtmp = 0.0;
            for (int j = 0; j < np; ++j) dtmp += nn.w[l][i][j] * val[l - 1][j];
            dtmp += nn.b[l][i];
            if (nn.normalize) dtmp /= np;
            double v = 0.0, d = 0.0;
            switch (nn.act[l])
            {
            case AF_IDENTITY: v = dtmp; d = 1.0; break;
            case AF_TANH:
                dtmp = std::tanh(dtmp); v = dtmp; d = 1.0 - dtmp * dtmp; break;
            case AF_LOGISTIC:
                if (dtmp > 35.0) { v = 1.0; d = 0.0; }
                else if (dtmp < -35.0) { v = 0.0; d = 0.0; }
                else
                {
                    dtmp = 1.0 / (1.0 + std::exp(-dtmp));
                    v = dtmp; d = dtmp * (1.0 - dtmp);
                }
                break;
            case AF_SOFTPLUS:
                if (dtmp > 35.0) { v = dtmp; d = 1.0; }
                else if (dtmp < -35.0) { v = 0.0; d = 0.0; }
                else
                {
                    dtmp = std::exp(dtmp);
                    v = std::log(1.0 + dtmp);
                    dtmp = 1.0 / (1.0 + 1.0 / dtmp);
                    d = dtmp;
                }
                break;
            case AF_RELU:
                if (dtmp > 0.0) { v = dtmp; d = 1.0; } else { v = 0; d = 0; }
                break;
            case AF_GAUSSIAN:
            {
                double const t = std::exp(-0.5 * dtmp * dtmp);
                v = t; d = -dtmp * t; break;
            }
            case AF_COS: v = std::cos(dtmp); d = -std::sin(dtmp); break;
            case AF_REVLOGISTIC:
                dtmp = 1.0 / (1.0 + std::exp(-dtmp));
                v = 1.0 - dtmp; d = dtmp * (dtmp - 1.0); break;
            case AF_EXP: dtmp = std::exp(-dtmp); v = dtmp; d = -dtmp; break;
            case AF_HARMONIC: v = dtmp * dtmp; d = 2.0 * dtmp; break;
            default: break;
            }
            val[l][i] = v;
            dfdx[l][i] = d;
        }
    }
}

// NeuralNetwork.cpp:396-442 calculateDEdG (forward mode, one pass per input).
void calculateDEdG(NN const& nn, vector<vector<double>> const& dfdx,
                   double* dEdG)
{
    int const nh = nn.numLayers - 2;
    vector<vector<double>> inner(nh), outer(nh);
    for (int i = 0; i < nh; ++i)
    {
        inner[i].assign(nn.n[i + 1], 0.0);
        outer[i].assign(nn.n[i + 2], 0.0);
    }
    for (int k = 0; k < nn.n[0]; ++k)
    {
        for (int i = 0; i < nn.n[1]; ++i)
        {
            inner[0][i] = nn.w[1][i][k] * dfdx[1][i];
            if (nn.normalize) inner[0][i] /= nn.n[0];
        }
        for (int l = 1; l < nh + 1; ++l)
        {
            for (int i2 = 0; i2 < nn.n[l + 1]; ++i2)
            {
                outer[l - 1][i2] = 0.0;
                for (int i1 = 0; i1 < nn.n[l]; ++i1)
                    outer[l - 1][i2] += nn.w[l + 1][i2][i1] * inner[l - 1][i1];
                outer[l - 1][i2] *= dfdx[l + 1][i2];
                if (nn.normalize) outer[l - 1][i2] /= nn.n[l];
                if (l < nh) inner[l][i2] = outer[l - 1][i2];
            }
        }
        dEdG[k] = outer[nh - 1][0];
    }
}

// Mode.cpp:1642-1663 (serial loop over atoms, as the reference).
void networks(orc_struct& s, orc_model const& m, bool derivatives)
{
    vector<vector<double>> val, dfdx;
    for (auto& a : s.atoms)
    {
        NN const& nn = m.el[a.element].nn;
        propagate(nn, a.G.data(), val, dfdx);
        if (derivatives)
        {
            a.dEdG.assign(a.G.size(), 0.0);
            calculateDEdG(nn, dfdx, a.dEdG.data());
        }
        a.energy = val[nn.numLayers - 1][0];
    }
}

// Mode.cpp:1803-1828
double energy(orc_struct& s)
{
    double e = 0.0;
    for (auto const& a : s.atoms) e += a.energy;
    s.energy = e;
    return e;
}

// Mode.cpp:1849-1913 + Atom.cpp:371-401.  mode 1 adds the atom's own
// self-image slots (absent from the reference gather).
void forces(orc_struct& s, orc_model const& m, int mode)
{
    size_t const N = s.atoms.size();
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t i = 0; i < N; ++i)
    {
        Atom& ai = s.atoms[i];
        double self[3] = {0, 0, 0};
        for (size_t k = 0; k < ai.G.size(); ++k)
            for (int c = 0; c < 3; ++c)
                self[c] -= ai.dEdG[k] * ai.dGdr[3 * k + c];
        for (int c = 0; c < 3; ++c) ai.f[c] = 0.0 + self[c];
        auto addPair = [&](Atom const& aj, Neighbor const& n) {
            auto const& table = m.el[aj.element].table[n.element];
            double pf[3] = {0, 0, 0};
            size_t const nd = n.dGdr.size() / 3;
            for (size_t t = 0; t < nd; ++t)
                for (int c = 0; c < 3; ++c)
                    pf[c] -= aj.dEdG[table[t]] * n.dGdr[3 * t + c];
            for (int c = 0; c < 3; ++c) ai.f[c] += pf[c];
        };
        for (size_t u = 1; u < ai.neighborsUnique.size(); ++u)
        {
            Atom const& aj = s.atoms[ai.neighborsUnique[u]];
            for (auto const& n : aj.neighbors)
                if (n.index == ai.index) addPair(aj, n);
        }
        if (mode == 1)
            for (auto const& n : ai.neighbors)
                if (n.index == ai.index) addPair(ai, n);
    }
}

// ----------------------------------------------------------------- setup --
struct Settings
{
    // keyword -> list of (value, line number)
    std::multimap<string, std::pair<string, size_t>> kv;
    bool exists(string const& k) const { return kv.count(k) > 0; }
    string get(string const& k) const
    {
        auto it = kv.find(k);
        if (it == kv.end())
            throw std::runtime_error("ERROR: Keyword \"" + k
                                     + "\" not found.\n");
        return it->second.first;
    }
};

// Settings.cpp:276-299 readFile, :323-420 parseLines.
Settings readSettings(string const& fileName)
{
    std::ifstream file(fileName);
    if (!file.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + fileName
                                 + "\".\n");
    Settings s;
    string line;
    size_t ln = 0;
    while (std::getline(file, line))
    {
        size_t const thisLine = ln++;
        if (line.empty() || line == "\r") continue;
        if (line.find('#') != string::npos) line.erase(line.find('#'));
        if (line.find('!') != string::npos) line.erase(line.find('!'));
        if (line.find_first_not_of(' ') == string::npos) continue;
        line = reduce(line);
        if (line.empty()) continue;
        size_t const sep = line.find_first_of(' ');
        if (sep == string::npos)
            s.kv.emplace(line, std::make_pair(string(""), thisLine));
        else
            s.kv.emplace(line.substr(0, sep),
                         std::make_pair(line.substr(sep + 1), thisLine));
    }
    return s;
}

vector<double> readColumn0(string const& fileName)
{
    std::ifstream f(fileName);
    if (!f.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + fileName
                                 + "\".\n");
    vector<double> v;
    string line;
    while (std::getline(f, line))
    {
        if (line.empty() || line[0] == '#') continue;
        auto w = split(reduce(line));
        if (w.empty()) continue;
        v.push_back(std::atof(w[0].c_str()));
    }
    return v;
}

string formatZ(string const& fmt, int Z)
{
    char buf[4096];
    std::snprintf(buf, sizeof buf, fmt.c_str(), (size_t)Z);
    return string(buf);
}

orc_model* loadModel(string const& settingsFile, string const& scalingFile,
                     string const& weightsFmt)
{
    std::unique_ptr<orc_model> m(new orc_model);
    Settings st = readSettings(settingsFile);
    // Mode.cpp:238-299 setupNormalization
    bool hm = st.exists("mean_energy"), hce = st.exists("conv_energy"),
         hcl = st.exists("conv_length");
    if (hm && hce && hcl)
    {
        m->normalize = true;
        m->meanEnergy = std::atof(st.get("mean_energy").c_str());
        m->convEnergy = std::atof(st.get("conv_energy").c_str());
        m->convLength = std::atof(st.get("conv_length").c_str());
    }
    else if (hm || hce || hcl)
        throw std::runtime_error("ERROR: Incorrect usage of normalization"
                                 " keywords.\n");
    // Mode.cpp:301-320 setupElementMap (sorted by Z)
    vector<string> syms = split(reduce(st.get("elements")));
    std::sort(syms.begin(), syms.end(), [](string const& a, string const& b) {
        return atomicNumber(a) < atomicNumber(b);
    });
    std::map<string, int> emap;
    for (size_t i = 0; i < syms.size(); ++i) emap[syms[i]] = (int)i;
    auto eidx = [&](string const& sym) {
        auto it = emap.find(sym);
        if (it == emap.end())
            throw std::runtime_error("ERROR: Element \"" + sym
                                     + "\" not in element map.\n");
        return (size_t)it->second;
    };
    size_t const ne = (size_t)std::atoi(st.get("number_of_elements").c_str());
    if (ne != syms.size())
        throw std::runtime_error("ERROR: Inconsistent number of elements.\n");
    m->el.resize(ne);
    for (size_t i = 0; i < ne; ++i)
    {
        m->el[i].symbol = syms[i];
        m->el[i].Z = atomicNumber(syms[i]);
    }
    // Mode.cpp:322-366 atom_energy
    for (auto r = st.kv.equal_range("atom_energy"); r.first != r.second;
         ++r.first)
    {
        auto a = split(reduce(r.first->second.first));
        m->el[eidx(a.at(0))].atomicEnergyOffset = std::atof(a.at(1).c_str());
    }
    // Mode.cpp:520-613 setupCutoff
    {
        auto a = split(st.get("cutoff_type"));
        m->cutoffType = std::atoi(a.at(0).c_str());
        if (a.size() > 1) m->cutoffAlpha = std::atof(a.at(1).c_str());
        if (m->cutoffType < CT_HARD || m->cutoffType > CT_POLY4)
            throw std::invalid_argument("ERROR: Unknown cutoff type.\n");
    }
    // Mode.cpp:615-690 setupSymmetryFunctions + Element.cpp:82-139
    for (auto r = st.kv.equal_range("symfunction_short"); r.first != r.second;
         ++r.first)
    {
        string const& par = r.first->second.first;
        auto a = split(reduce(par));
        SF sf;
        sf.type = std::atoi(a.at(1).c_str());
        sf.ec = (int)eidx(a.at(0));
        sf.line = r.first->second.second;
        if (sf.type == 2)   // SymFncExpRad.cpp:78-97
        {
            sf.e1 = (int)eidx(a.at(2));
            sf.eta = std::atof(a.at(3).c_str());
            sf.rs = std::atof(a.at(4).c_str());
            sf.rc = std::atof(a.at(5).c_str());
        }
        else if (sf.type == 3 || sf.type == 9)  // SymFncBaseExpAng.cpp:50-93
        {
            sf.e1 = (int)eidx(a.at(2));
            sf.e2 = (int)eidx(a.at(3));
            sf.eta = std::atof(a.at(4).c_str());
            sf.lambda = std::atof(a.at(5).c_str());
            sf.zeta = std::atof(a.at(6).c_str());
            sf.rc = std::atof(a.at(7).c_str());
            if (a.size() > 8) sf.rs = std::atof(a.at(8).c_str());
            if (sf.e1 > sf.e2) std::swap(sf.e1, sf.e2);
            sf.zetaInt = (int)std::round(sf.zeta);
            sf.useIntegerPow = std::fabs(sf.zeta - sf.zetaInt)
                             <= std::numeric_limits<double>::min();
        }
        else if (sf.type == 12)   // SymFncExpRadWeighted.cpp:71-90
        {
            sf.e1 = sf.e2 = -1;
            sf.eta = std::atof(a.at(2).c_str());
            sf.rs = std::atof(a.at(3).c_str());
            sf.rc = std::atof(a.at(4).c_str());
        }
        else if (sf.type == 13)   // SymFncExpAngnWeighted.cpp:92-123
        {
            sf.e1 = sf.e2 = -1;
            sf.eta = std::atof(a.at(2).c_str());
            sf.rs = std::atof(a.at(3).c_str());
            sf.lambda = std::atof(a.at(4).c_str());
            sf.zeta = std::atof(a.at(5).c_str());
            sf.rc = std::atof(a.at(6).c_str());
            sf.zetaInt = (int)std::round(sf.zeta);
            sf.useIntegerPow = std::fabs(sf.zeta - sf.zetaInt)
                             <= std::numeric_limits<double>::min();
        }
        else if (isCompact(sf.type))
        {
            // SymFncCompRad.cpp / SymFncBaseCompAng.cpp /
            // SymFncCompRadWeighted.cpp / SymFncBaseCompAngWeighted.cpp
            // setParameters, SymFncBaseComp.cpp:48-100 setCompactFunction.
            size_t p = 2;
            sf.e1 = sf.e2 = -1;
            if (!isWeighted(sf.type))
            {
                sf.e1 = (int)eidx(a.at(p++));
                if (sf.type != 20) sf.e2 = (int)eidx(a.at(p++));
                if (sf.type != 20 && sf.e1 > sf.e2) std::swap(sf.e1, sf.e2);
            }
            sf.rl = std::atof(a.at(p++).c_str());
            sf.rc = std::atof(a.at(p++).c_str());
            if (!isRadial(sf.type))
            {
                sf.angleLeft = std::atof(a.at(p++).c_str());
                sf.angleRight = std::atof(a.at(p++).c_str());
            }
            sf.subtype = a.at(p++);
            if (sf.rl > sf.rc)
                throw std::runtime_error("ERROR: Lower radial boundary >= "
                                         "upper radial boundary.\n");
            string const& st = sf.subtype;
            if (st.size() < 1 || st.size() > 3)
                throw std::runtime_error("ERROR: Invalid compact function "
                                         "type specification.\n");
            if (st[0] == 'p' && st.size() >= 2 && st[1] >= '1' && st[1] <= '4')
            {
                sf.cr.core = CT_POLY1 + (st[1] - '1');
                if (st.size() == 3)
                {
                    if (st[2] != 'a')
                        throw std::runtime_error("ERROR: Invalid polynom "
                                                 "specifier.\n");
                    sf.cr.asym = true;
                }
            }
            else if (st[0] == 'e') sf.cr.core = CT_EXP;
            else
                throw std::runtime_error("ERROR: Unknown compact SF type.\n");
            sf.cr.setLeftRight(sf.rl, sf.rc);
            if (!isRadial(sf.type))
            {
                double const al = sf.angleLeft, ar = sf.angleRight;
                double const center = (al + ar) / 2.0;
                if (al >= ar)
                    throw std::runtime_error("ERROR: Left angle boundary right"
                                             " of or equal to right angle "
                                             "boundary.\n");
                if ((al < 0.0 && center != 0.0)
                    || (ar > 180.0 && center != 180.0))
                    throw std::runtime_error("ERROR: Angle boundary out of "
                                             "[0,180] and center of angular "
                                             "function /= 0 or /= 180.\n");
                if (ar - al > 360.0)
                    throw std::runtime_error("ERROR: Periodic symmetry "
                                             "function cannot spread over "
                                             "domain > 360 degrees\n");
                sf.ca.core = sf.cr.core;
                sf.ca.setLeftRight(al * M_PI / 180.0, ar * M_PI / 180.0);
            }
        }
        else
            throw std::runtime_error("ERROR: Unknown symmetry function type"
                                     " (this engine covers types 2, 3, 9, 12,"
                                     " 13, 20-25).\n");
        sf.indexPerElement.assign(ne, std::numeric_limits<size_t>::max());
        m->el[sf.ec].sfs.push_back(sf);
    }
    m->maxCutoffRadius = 0.0;
    for (auto& e : m->el)
    {
        if (m->normalize)   // changeLengthUnit (SymFncExpRad.cpp:99-110)
            for (auto& sf : e.sfs)
            {
                sf.eta /= m->convLength * m->convLength;
                sf.rs *= m->convLength;
                sf.rc *= m->convLength;
                sf.rl *= m->convLength;   // SymFncCompRad::changeLengthUnit
                if (isCompact(sf.type)) sf.cr.setLeftRight(sf.rl, sf.rc);
            }
        std::sort(e.sfs.begin(), e.sfs.end(), sfLess);   // Element.cpp:154-166
        for (size_t i = 0; i < e.sfs.size(); ++i) e.sfs[i].index = (int)i;
        for (auto& sf : e.sfs)
        {
            m->maxCutoffRadius = std::max(m->maxCutoffRadius, sf.rc);
            sf.cutoffType = m->cutoffType;
            sf.cutoffAlpha = m->cutoffAlpha;
        }
    }
    // Element.cpp:289-311 setupSymmetryFunctionMemory
    for (auto& e : m->el)
    {
        e.table.assign(ne, {});
        for (auto& sf : e.sfs)
            for (size_t i = 0; i < ne; ++i)
                if (sf.relevant(i))
                {
                    sf.indexPerElement[i] = e.table[i].size();
                    e.table[i].push_back(sf.index);
                }
        e.numTable.clear();
        for (size_t i = 0; i < ne; ++i) e.numTable.push_back(e.table[i].size());
    }
    // Element.cpp:194-287 setupSymmetryFunctionGroups + sortMembers
    for (auto& e : m->el)
    {
        for (size_t s = 0; s < e.sfs.size(); ++s)
        {
            SF const& sf = e.sfs[s];
            bool added = false;
            for (auto& g : e.groups)
            {
                SF const& f0 = e.sfs[g.members[0]];
                if (isCompact(sf.type))
                {
                    // SymGrpCompRad.cpp:68-100: centre and neighbour
                    // elements only (none for the weighted types).
                    if (sf.type != f0.type || sf.ec != f0.ec
                        || sf.e1 != f0.e1 || sf.e2 != f0.e2)
                        continue;
                    g.members.push_back((int)s);
                    added = true;
                    break;
                }
                if (sf.type != f0.type || sf.cutoffType != f0.cutoffType
                    || sf.cutoffAlpha != f0.cutoffAlpha || sf.ec != f0.ec
                    || sf.rc != f0.rc || sf.e1 != f0.e1
                    || (sf.type != 2 && sf.e2 != f0.e2))
                    continue;
                g.members.push_back((int)s);
                added = true;
                break;
            }
            if (!added)
            {
                Group g;
                g.type = sf.type;
                g.e1 = sf.e1;
                g.e2 = sf.e2;
                g.rc = sf.rc;
                g.cutoffType = sf.cutoffType;
                g.cutoffAlpha = sf.cutoffAlpha;
                g.fc.set(sf.cutoffType, sf.rc, sf.cutoffAlpha);
                if (isWeighted(sf.type))
                    for (auto const& x : m->el)
                        g.weight.push_back((double)x.Z);
                g.members.push_back((int)s);
                e.groups.push_back(g);
            }
        }
        for (auto& g : e.groups)
        {
            std::stable_sort(g.members.begin(), g.members.end(),
                             [&](int a, int b) {
                                 return sfLess(e.sfs[a], e.sfs[b]);
                             });
            for (size_t i = 0; i < g.members.size(); ++i)
            {
                SF const& sf = e.sfs[g.members[i]];
                if (!isRadial(g.type))
                {
                    g.factorNorm.push_back(std::pow(2.0, 1.0 - sf.zeta));
                    g.factorDeriv.push_back(2.0 * sf.eta / sf.zeta / sf.lambda);
                }
                else
                {
                    g.factorNorm.push_back(1.0);
                    g.factorDeriv.push_back(0.0);
                }
                if (i == 0) g.calculateExp.push_back(1);
                else
                {
                    SF const& p = e.sfs[g.members[i - 1]];
                    if (isCompact(g.type))   // calculateComp
                        g.calculateExp.push_back(p.rc != sf.rc
                                                 || p.rl != sf.rl);
                    else
                        g.calculateExp.push_back(p.eta != sf.eta
                                                 || p.rs != sf.rs);
                }
                if (isCompact(g.type))
                {
                    // SymGrpCompRad.cpp:82-84 with rmin starting at 0
                    // (SymGrpBaseComp.cpp:26): rmin stays 0.
                    if (i == 0) { g.rmin = 0.0; g.rmax = 0.0; }
                    if (sf.rl <= 0.0) g.rmin = 0.0;
                    else g.rmin = std::min(g.rmin, sf.rl);
                    g.rmax = std::max(g.rmax, sf.rc);
                }
            }
        }
    }
    // Mode.cpp:692-840 scaling + SymFnc.cpp:52-98 + group setScalingFactors
    int stype = ST_NONE;
    double Smin = 0.0, Smax = 0.0;
    bool const useScaling = !scalingFile.empty();
    if (useScaling)
    {
        bool sc = st.exists("scale_symmetry_functions");
        bool ce = st.exists("center_symmetry_functions");
        if (sc && !ce) stype = ST_SCALE;
        else if (!sc && ce) stype = ST_CENTER;
        else if (sc && ce) stype = ST_SCALECENTER;
        else if (st.exists("scale_symmetry_functions_sigma"))
            stype = ST_SCALESIGMA;
        if (stype == ST_SCALE || stype == ST_SCALECENTER
            || stype == ST_SCALESIGMA)
        {
            Smin = st.exists("scale_min_short")
                 ? std::atof(st.get("scale_min_short").c_str()) : 0.0;
            Smax = st.exists("scale_max_short")
                 ? std::atof(st.get("scale_max_short").c_str()) : 1.0;
        }
        std::ifstream f(scalingFile);
        if (!f.is_open())
            throw std::runtime_error("ERROR: Could not open file: \""
                                     + scalingFile + "\".\n");
        vector<string> lines;
        string line;
        while (std::getline(f, line))
            if (!line.empty() && line[0] != '#') lines.push_back(line);
        size_t pos = 0;
        for (auto& e : m->el)
            for (auto& sf : e.sfs)
            {
                auto w = split(reduce(lines.at(pos++)));
                sf.scalingType = stype;
                sf.Gmin = std::atof(w.at(2).c_str());
                sf.Gmax = std::atof(w.at(3).c_str());
                sf.Gmean = std::atof(w.at(4).c_str());
                if (w.size() > 5) sf.Gsigma = std::atof(w.at(5).c_str());
                sf.Smin = Smin;
                sf.Smax = Smax;
                if (stype == ST_SCALE || stype == ST_SCALECENTER)
                    sf.scalingFactor = (Smax - Smin) / (sf.Gmax - sf.Gmin);
                else if (stype == ST_SCALESIGMA)
                    sf.scalingFactor = (Smax - Smin) / sf.Gsigma;
                else sf.scalingFactor = 1.0;
            }
    }
    for (auto& e : m->el)
        for (auto& g : e.groups)
        {
            g.scalingFactors.clear();
            for (size_t i = 0; i < g.members.size(); ++i)
            {
                double const sfac = e.sfs[g.members[i]].scalingFactor;
                g.scalingFactors.push_back(sfac);
                if (!isRadial(g.type)) g.factorNorm[i] *= sfac;
            }
        }
    // Mode.cpp:1158-1443 setupNeuralNetwork (2G, "short" suffix)
    {
        size_t globalHidden = 0;
        vector<int> numLayers(ne, 0);
        if (st.exists("global_hidden_layers_short"))
        {
            globalHidden = (size_t)std::atoi(
                st.get("global_hidden_layers_short").c_str());
            for (auto& n : numLayers) n = (int)globalHidden + 2;
        }
        for (auto r = st.kv.equal_range("element_hidden_layers_short");
             r.first != r.second; ++r.first)
        {
            auto a = split(reduce(r.first->second.first));
            numLayers[eidx(a.at(0))] = std::atoi(a.at(1).c_str()) + 2;
        }
        vector<string> gNodes, gAct;
        if (st.exists("global_nodes_short"))
        {
            gNodes = split(reduce(st.get("global_nodes_short")));
            if (gNodes.size() != globalHidden)
                throw std::runtime_error("ERROR: Inconsistent global NN "
                                         "topology keyword.\n");
        }
        if (st.exists("global_activation_short"))
        {
            gAct = split(reduce(st.get("global_activation_short")));
            if (globalHidden != gAct.size() - 1)
                throw std::runtime_error("ERROR: Inconsistent global NN "
                                         "topology keyword.\n");
        }
        for (size_t i = 0; i < ne; ++i)
        {
            NN& nn = m->el[i].nn;
            nn.numLayers = numLayers[i];
            if (nn.numLayers < 3)
                throw std::runtime_error("ERROR: Neural network must have at"
                                         " least three layers.\n");
            nn.n.assign(nn.numLayers, 0);
            nn.act.assign(nn.numLayers, AF_UNSET);
            nn.n[0] = (int)m->el[i].sfs.size();
            nn.act[0] = AF_IDENTITY;
            nn.n[nn.numLayers - 1] = 1;
            if ((size_t)nn.numLayers != globalHidden + 2) continue;
            for (int j = 1; j < nn.numLayers; ++j)
            {
                if (j == nn.numLayers - 1 && !gAct.empty())
                    nn.act[j] = activationFromString(gAct.at(nn.numLayers - 2));
                else
                {
                    if (!gNodes.empty())
                        nn.n[j] = std::atoi(gNodes.at(j - 1).c_str());
                    if (!gAct.empty())
                        nn.act[j] = activationFromString(gAct.at(j - 1));
                }
            }
        }
        for (auto r = st.kv.equal_range("element_nodes_short");
             r.first != r.second; ++r.first)
        {
            auto a = split(reduce(r.first->second.first));
            NN& nn = m->el[eidx(a.at(0))].nn;
            if ((size_t)nn.numLayers != a.size() - 1 + 2)
                throw std::runtime_error("ERROR: Inconsistent per-element NN"
                                         " topology keyword.\n");
            for (int j = 1; j < nn.numLayers - 2; ++j)
                nn.n[j] = std::atoi(a.at(j).c_str());
        }
        for (auto r = st.kv.equal_range("element_activation_short");
             r.first != r.second; ++r.first)
        {
            auto a = split(reduce(r.first->second.first));
            NN& nn = m->el[eidx(a.at(0))].nn;
            if ((size_t)nn.numLayers != a.size() - 1 + 1)
                throw std::runtime_error("ERROR: Inconsistent per-element NN"
                                         " topology keyword.\n");
            for (int j = 1; j < nn.numLayers - 1; ++j)
                nn.act[j] = activationFromString(a.at(j));
        }
        bool const normalizeNeurons = st.exists("normalize_nodes");
        for (size_t i = 0; i < ne; ++i)
        {
            NN& nn = m->el[i].nn;
            for (int j = 0; j < nn.numLayers; ++j)
                if (nn.n[j] == 0 || nn.act[j] == AF_UNSET)
                    throw std::runtime_error("ERROR: NN topology unset.\n");
            nn.normalize = normalizeNeurons;
            // Mode.cpp:2270-2289 + NeuralNetwork.cpp:147-169 setConnections
            // An empty weights format gives zero weights (the SF-only
            // tools: statistics and histograms need no network).
            vector<double> wts;
            if (!weightsFmt.empty())
                wts = readColumn0(formatZ(weightsFmt, m->el[i].Z));
            else
            {
                size_t nw = 0;
                for (int l = 1; l < nn.numLayers; ++l)
                    nw += (size_t)nn.n[l] * (nn.n[l - 1] + 1);
                wts.assign(nw, 0.0);
            }
            nn.w.assign(nn.numLayers, {});
            nn.b.assign(nn.numLayers, {});
            size_t count = 0;
            for (int l = 1; l < nn.numLayers; ++l)
            {
                nn.w[l].assign(nn.n[l], vector<double>(nn.n[l - 1], 0.0));
                nn.b[l].assign(nn.n[l], 0.0);
                for (int j = 0; j < nn.n[l - 1]; ++j)
                    for (int k = 0; k < nn.n[l]; ++k)
                        nn.w[l][k][j] = wts.at(count++);
                for (int j = 0; j < nn.n[l]; ++j) nn.b[l][j] = wts.at(count++);
            }
        }
    }
    return m.release();
}

template <class F> auto guard(F f, decltype(f()) fail) -> decltype(f())
{
    try { g_err.clear(); return f(); }
    catch (std::exception const& e) { g_err = e.what(); return fail; }
}

orc_struct* newStruct(orc_model const& m)
{
    orc_struct* s = new orc_struct;
    s->numElements = m.el.size();
    s->numAtomsPerElement.assign(s->numElements, 0);
    return s;
}

// Structure.cpp:142-271 readFromLines (first configuration of the file).
orc_struct* readStruct(orc_model const& m, string const& path)
{
    std::ifstream f(path);
    if (!f.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + path
                                 + "\".\n");
    std::map<string, int> emap;
    for (size_t i = 0; i < m.el.size(); ++i) emap[m.el[i].symbol] = (int)i;
    std::unique_ptr<orc_struct> s(newStruct(m));
    string line;
    int iBox = 0;
    bool begun = false;
    while (std::getline(f, line))
    {
        auto w = split(reduce(line));
        if (w.empty()) continue;
        if (w[0] == "begin") { begun = true; continue; }
        if (!begun) continue;
        if (w[0] == "comment") continue;
        if (w[0] == "lattice")
        {
            if (iBox > 2)
                throw std::runtime_error("ERROR: Too many box vectors.\n");
            for (int c = 0; c < 3; ++c)
                s->box[iBox][c] = std::atof(w.at(1 + c).c_str());
            iBox++;
            if (iBox == 3)
            {
                s->periodic = true;
                calculateInverseBox(*s);
            }
        }
        else if (w[0] == "atom")
        {
            Atom a;
            a.index = s->atoms.size();
            for (int c = 0; c < 3; ++c) a.r[c] = std::atof(w.at(1 + c).c_str());
            auto it = emap.find(w.at(4));
            if (it == emap.end())
                throw std::runtime_error("ERROR: Unknown element \"" + w[4]
                                         + "\".\n");
            a.element = (size_t)it->second;
            s->numAtomsPerElement[a.element]++;
            s->atoms.push_back(a);
        }
        else if (w[0] == "energy" || w[0] == "charge") continue;
        else if (w[0] == "end")
        {
            if (!(iBox == 0 || iBox == 3))
                throw std::runtime_error("ERROR: Strange number of box "
                                         "vectors.\n");
            break;
        }
        else
            throw std::runtime_error("ERROR: Unexpected file content, unknown"
                                     " keyword \"" + w[0] + "\".\n");
    }
    if (s->periodic)
        for (auto& a : s->atoms) remapAtom(*s, a);
    return s.release();
}

// Structure.cpp:1250-1275 toNormalizedUnits (positions and box).
void toNormalized(orc_struct& s, orc_model const& m)
{
    if (!m.normalize) return;
    double const c = m.convLength;
    if (s.periodic)
    {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
            {
                s.box[i][j] *= c;
                s.invbox[i][j] /= c;
            }
    }
    for (auto& a : s.atoms)
        for (int k = 0; k < 3; ++k) a.r[k] *= c;
}

}  // namespace

// ===================================================================== API
extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

orc_model* orc_model_load(const char* settings_file, const char* scaling_file,
                          const char* weights_fmt)
{
    return guard([&]() {
        return loadModel(settings_file, scaling_file ? scaling_file : "",
                         weights_fmt);
    }, (orc_model*)nullptr);
}

void orc_model_free(orc_model* m) { delete m; }
int orc_model_num_elements(const orc_model* m) { return (int)m->el.size(); }
int orc_model_atomic_number(const orc_model* m, int e) { return m->el.at(e).Z; }
int orc_model_num_sf(const orc_model* m, int e)
{
    return (int)m->el.at(e).sfs.size();
}
double orc_model_max_cutoff(const orc_model* m) { return m->maxCutoffRadius; }
int orc_model_normalize(const orc_model* m, double* me, double* ce, double* cl)
{
    if (me) *me = m->meanEnergy;
    if (ce) *ce = m->convEnergy;
    if (cl) *cl = m->convLength;
    return m->normalize ? 1 : 0;
}

int orc_model_sf_table(const orc_model* m, int e, double* out)
{
    auto const& sfs = m->el.at(e).sfs;
    for (size_t i = 0; i < sfs.size(); ++i)
    {
        SF const& s = sfs[i];
        double* o = out + 12 * i;
        o[0] = s.type; o[1] = s.e1; o[2] = isRadial(s.type) ? -1 : s.e2;
        o[3] = s.eta; o[4] = s.rs; o[5] = s.lambda; o[6] = s.zeta;
        o[7] = s.rc; o[8] = s.cutoffType; o[9] = s.cutoffAlpha;
        o[10] = s.scalingFactor; o[11] = (double)s.line;
    }
    return (int)sfs.size();
}

int orc_model_groups(const orc_model* m, int e, int32_t* out, int cap)
{
    auto const& el = m->el.at(e);
    int pos = 0;
    for (auto const& g : el.groups)
    {
        int need = 4 + (int)g.members.size();
        if (out && pos + need <= cap)
        {
            out[pos] = g.type; out[pos + 1] = g.e1;
            out[pos + 2] = isRadial(g.type) ? -1 : g.e2;
            out[pos + 3] = (int)g.members.size();
            for (size_t i = 0; i < g.members.size(); ++i)
                out[pos + 4 + i] = g.members[i];
        }
        pos += need;
    }
    return (int)el.groups.size();
}

orc_struct* orc_struct_create(const orc_model* m, int64_t n, const double* r,
                              const int32_t* elem, const double* box)
{
    return guard([&]() {
        std::unique_ptr<orc_struct> s(newStruct(*m));
        s->atoms.resize((size_t)n);
        for (int64_t i = 0; i < n; ++i)
        {
            Atom& a = s->atoms[i];
            a.index = (size_t)i;
            a.element = (size_t)elem[i];
            if (a.element >= m->el.size())
                throw std::runtime_error("ERROR: element index out of "
                                         "range.\n");
            for (int c = 0; c < 3; ++c) a.r[c] = r[3 * i + c];
            s->numAtomsPerElement[a.element]++;
        }
        if (box)
        {
            s->periodic = true;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) s->box[i][j] = box[3 * i + j];
            calculateInverseBox(*s);
        }
        return s.release();
    }, (orc_struct*)nullptr);
}

orc_struct* orc_struct_read(const orc_model* m, const char* path)
{
    return guard([&]() { return readStruct(*m, path); }, (orc_struct*)nullptr);
}

void orc_struct_free(orc_struct* s) { delete s; }
int64_t orc_struct_num_atoms(const orc_struct* s)
{
    return (int64_t)s->atoms.size();
}
void orc_struct_remap(orc_struct* s)
{
    if (s->periodic)
        for (auto& a : s->atoms) remapAtom(*s, a);
}
void orc_struct_get_positions(const orc_struct* s, double* r)
{
    for (size_t i = 0; i < s->atoms.size(); ++i)
        for (int c = 0; c < 3; ++c) r[3 * i + c] = s->atoms[i].r[c];
}
void orc_struct_get_elements(const orc_struct* s, int32_t* e)
{
    for (size_t i = 0; i < s->atoms.size(); ++i)
        e[i] = (int32_t)s->atoms[i].element;
}
void orc_struct_get_box(const orc_struct* s, double* box)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) box[3 * i + j] = s->box[i][j];
}
int orc_struct_to_normalized(orc_struct* s, const orc_model* m)
{
    toNormalized(*s, *m);
    return m->normalize ? 1 : 0;
}

int64_t orc_neighbor_list(orc_struct* s, double cutoff, int cells)
{
    if (cells) neighborListCells(*s, cutoff);
    else neighborListBrute(*s, cutoff);
    int64_t total = 0;
    for (auto const& a : s->atoms) total += (int64_t)a.neighbors.size();
    return total;
}

void orc_get_pbc(const orc_struct* s, int32_t* pbc)
{
    for (int k = 0; k < 3; ++k) pbc[k] = s->pbc[k];
}

void orc_get_neighbors(const orc_struct* s, int64_t* off, int32_t* idx,
                       double* d, double* dr)
{
    int64_t p = 0;
    for (size_t i = 0; i < s->atoms.size(); ++i)
    {
        off[i] = p;
        for (auto const& n : s->atoms[i].neighbors)
        {
            if (idx) idx[p] = (int32_t)n.index;
            if (d) d[p] = n.d;
            if (dr)
                for (int c = 0; c < 3; ++c) dr[3 * p + c] = n.dr[c];
            ++p;
        }
    }
    off[s->atoms.size()] = p;
}

int orc_symmetry_functions(orc_struct* s, const orc_model* m, int derivatives)
{
    return guard([&]() { symmetryFunctions(*s, *m, derivatives != 0); return 0; },
                 -1);
}

int orc_networks(orc_struct* s, orc_model* m, int derivatives)
{
    return guard([&]() { networks(*s, *m, derivatives != 0); return 0; }, -1);
}

double orc_energy(orc_struct* s) { return energy(*s); }

int orc_forces(orc_struct* s, const orc_model* m, int mode)
{
    return guard([&]() { forces(*s, *m, mode); return 0; }, -1);
}

int orc_get_G(const orc_struct* s, double* G, int max_nsf)
{
    for (size_t i = 0; i < s->atoms.size(); ++i)
    {
        auto const& a = s->atoms[i];
        for (int k = 0; k < max_nsf; ++k)
            G[(size_t)max_nsf * i + k] = (k < (int)a.G.size()) ? a.G[k] : 0.0;
    }
    return 0;
}

int orc_get_dEdG(const orc_struct* s, double* dEdG, int max_nsf)
{
    for (size_t i = 0; i < s->atoms.size(); ++i)
    {
        auto const& a = s->atoms[i];
        for (int k = 0; k < max_nsf; ++k)
            dEdG[(size_t)max_nsf * i + k]
                = (k < (int)a.dEdG.size()) ? a.dEdG[k] : 0.0;
    }
    return 0;
}

void orc_get_atomic_energies(const orc_struct* s, double* e)
{
    for (size_t i = 0; i < s->atoms.size(); ++i) e[i] = s->atoms[i].energy;
}

void orc_get_forces(const orc_struct* s, double* f)
{
    for (size_t i = 0; i < s->atoms.size(); ++i)
        for (int c = 0; c < 3; ++c) f[3 * i + c] = s->atoms[i].f[c];
}

int orc_get_dGdr_self(const orc_struct* s, int64_t i, double* out)
{
    auto const& a = s->atoms.at((size_t)i);
    std::copy(a.dGdr.begin(), a.dGdr.end(), out);
    return (int)(a.dGdr.size() / 3);
}

int orc_get_dGdr_neighbor(const orc_struct* s, int64_t i, int64_t k,
                          double* out)
{
    auto const& n = s->atoms.at((size_t)i).neighbors.at((size_t)k);
    std::copy(n.dGdr.begin(), n.dGdr.end(), out);
    return (int)(n.dGdr.size() / 3);
}

// Atom.cpp:371-401 for every atom and every neighbour slot.
int orc_get_pair_forces(const orc_struct* s, const orc_model* m,
                        double* self_forces, double* pair_forces)
{
    size_t p = 0;
    for (size_t i = 0; i < s->atoms.size(); ++i)
    {
        Atom const& a = s->atoms[i];
        double sf[3] = {0, 0, 0};
        for (size_t k = 0; k < a.G.size(); ++k)
            for (int c = 0; c < 3; ++c) sf[c] -= a.dEdG[k] * a.dGdr[3 * k + c];
        if (self_forces)
            for (int c = 0; c < 3; ++c) self_forces[3 * i + c] = sf[c];
        for (auto const& n : a.neighbors)
        {
            auto const& table = m->el[a.element].table[n.element];
            double pf[3] = {0, 0, 0};
            size_t const nd = n.dGdr.size() / 3;
            for (size_t t = 0; t < nd; ++t)
                for (int c = 0; c < 3; ++c)
                    pf[c] -= a.dEdG[table[t]] * n.dGdr[3 * t + c];
            if (pair_forces)
                for (int c = 0; c < 3; ++c) pair_forces[3 * p + c] = pf[c];
            ++p;
        }
    }
    return 0;
}

double orc_evaluate(orc_struct* s, orc_model* m, int cells, int force_mode)
{
    if (cells) neighborListCells(*s, m->maxCutoffRadius);
    else neighborListBrute(*s, m->maxCutoffRadius);
    symmetryFunctions(*s, *m, true);
    networks(*s, *m, true);
    double e = energy(*s);
    forces(*s, *m, force_mode);
    return e;
}

double orc_energy_only(orc_struct* s, orc_model* m, int cells)
{
    if (cells) neighborListCells(*s, m->maxCutoffRadius);
    else neighborListBrute(*s, m->maxCutoffRadius);
    symmetryFunctions(*s, *m, false);
    networks(*s, *m, false);
    return energy(*s);
}

// Prediction.cpp:48-76 readStructureFromFile + predict (Mode::evaluateNNP,
// Mode.cpp:1967-2016), then toPhysicalUnits and energy offsets.
int orc_predict_file(orc_model* m, const char* path, double* energy_out,
                     double* forces_out, int64_t cap, int64_t* n_atoms)
{
    return guard([&]() {
        std::unique_ptr<orc_struct> s(readStruct(*m, path));
        toNormalized(*s, *m);
        orc_evaluate(s.get(), m, 0, 0);
        double e = s->energy;
        size_t const N = s->atoms.size();
        if (m->normalize)
            e = e / m->convEnergy + (double)N * m->meanEnergy;
        for (size_t i = 0; i < m->el.size(); ++i)
            e += s->numAtomsPerElement[i] * m->el[i].atomicEnergyOffset;
        *energy_out = e;
        if (n_atoms) *n_atoms = (int64_t)N;
        if ((int64_t)N > cap) throw std::runtime_error("capacity");
        for (size_t i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c)
            {
                double f = s->atoms[i].f[c];
                if (m->normalize) f /= m->convEnergy / m->convLength;
                forces_out[3 * i + c] = f;
            }
        return 0;
    }, -1);
}

// Element::updateSymmetryFunctionStatistics (Element.cpp:462-541) over the
// atoms in order (Mode.cpp:1625-1633): every SF whose unscaled value is
// outside [Gmin - 1000 eps, Gmax + 1000 eps] is one warning record (atom,
// SF index, value); up to cap records are written, the total is returned.
int64_t orc_extrapolation_warnings(const orc_struct* s, const orc_model* m,
                                   int32_t* atom, int32_t* sf, double* value,
                                   int64_t cap)
{
    return guard([&]() -> int64_t {
        int64_t count = 0;
        double const epsilon = 1000.0 * std::numeric_limits<double>::epsilon();
        for (size_t i = 0; i < s->atoms.size(); ++i)
        {
            Atom const& a = s->atoms[i];
            ElementData const& el = m->el.at(a.element);
            if (a.G.size() != el.sfs.size())
                throw std::runtime_error("ERROR: Number of symmetry functions"
                                    " does not match.\n");
            for (size_t k = 0; k < a.G.size(); ++k)
            {
                SF const& f = el.sfs[k];
                double const v = f.unscale(a.G[k]);
                if (v + epsilon < f.Gmin || v - epsilon > f.Gmax)
                {
                    if (count < cap)
                    {
                        atom[count] = (int32_t)i;
                        sf[count] = (int32_t)f.index;
                        value[count] = v;
                    }
                    ++count;
                }
            }
        }
        return count;
    }, (int64_t)-1);
}

int orc_sf_statistics(orc_struct* s, orc_model* m, int e, double* out)
{
    return guard([&]() {
        if (s->atoms.empty() || !s->atoms[0].neighbors.size())
            neighborListCells(*s, m->maxCutoffRadius);
        // Evaluate unscaled SFs: temporarily switch scaling off.
        std::vector<std::pair<int, double>> saved;
        ElementData& el = m->el.at(e);
        for (auto& sf : el.sfs)
        {
            saved.emplace_back(sf.scalingType, sf.scalingFactor);
            sf.scalingType = ST_NONE;
            sf.scalingFactor = 1.0;
        }
        vector<vector<double>> fnSaved, sfSaved;
        for (auto& g : el.groups)
        {
            fnSaved.push_back(g.factorNorm);
            sfSaved.push_back(g.scalingFactors);
            for (size_t i = 0; i < g.members.size(); ++i)
            {
                if (!isRadial(g.type)) g.factorNorm[i] /= g.scalingFactors[i];
                g.scalingFactors[i] = 1.0;
            }
        }
        symmetryFunctions(*s, *m, false);
        size_t const nsf = el.sfs.size();
        vector<double> mn(nsf, std::numeric_limits<double>::max()),
            mx(nsf, -std::numeric_limits<double>::max()), sum(nsf, 0.0),
            sum2(nsf, 0.0);
        size_t cnt = 0;
        for (auto const& a : s->atoms)
        {
            if ((int)a.element != e) continue;
            ++cnt;
            for (size_t k = 0; k < nsf; ++k)
            {
                double v = a.G[k];
                mn[k] = std::min(mn[k], v);
                mx[k] = std::max(mx[k], v);
                sum[k] += v;
                sum2[k] += v * v;
            }
        }
        // Dataset.cpp:1074-1085: mean = sum / n,
        // sigma = sqrt((sum2 - sum^2 / n) / (n - 1)).
        for (size_t k = 0; k < nsf; ++k)
        {
            double const n = (double)cnt;
            double const mean = cnt ? sum[k] / n : 0.0;
            double const sigma = cnt > 1
                ? std::sqrt((sum2[k] - sum[k] * sum[k] / n) / (n - 1)) : 0.0;
            out[4 * k] = mn[k];
            out[4 * k + 1] = mx[k];
            out[4 * k + 2] = mean;
            out[4 * k + 3] = sigma;
        }
        for (size_t i = 0; i < el.sfs.size(); ++i)
        {
            el.sfs[i].scalingType = saved[i].first;
            el.sfs[i].scalingFactor = saved[i].second;
        }
        for (size_t g = 0; g < el.groups.size(); ++g)
        {
            el.groups[g].factorNorm = fnSaved[g];
            el.groups[g].scalingFactors = sfSaved[g];
        }
        return (int)cnt;
    }, -1);
}

// Dataset.cpp:888-950 prepareNumericForces (copy, shift one coordinate by
// +-delta, remap) and nnp-checkf.cpp:265: F = -(E+ - E-) / (2 delta).
int orc_numeric_forces(orc_struct* s, orc_model* m, double delta,
                       const int64_t* list, int64_t nlist, int cells,
                       double* out)
{
    return guard([&]() {
        for (int64_t q = 0; q < nlist; ++q)
        {
            size_t const i = (size_t)list[q];
            for (int c = 0; c < 3; ++c)
            {
                double e[2];
                for (int sgn = 0; sgn < 2; ++sgn)
                {
                    orc_struct t = *s;
                    for (auto& a : t.atoms)
                    {
                        a.neighbors.clear();
                        a.neighborsUnique.clear();
                    }
                    t.atoms[i].r[c] += (sgn == 0 ? delta : -delta);
                    if (t.periodic) remapAtom(t, t.atoms[i]);
                    e[sgn] = orc_energy_only(&t, m, cells);
                }
                out[3 * q + c] = -(e[0] - e[1]) / (2.0 * delta);
            }
        }
        return 0;
    }, -1);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
