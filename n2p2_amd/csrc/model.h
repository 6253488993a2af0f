// Host-side HDNNP model: settings parsing and the symmetry-function /
// network setup of n2p2's Mode, producing the flat tables the kernels use.
#ifndef HDNNP_MODEL_H
#define HDNNP_MODEL_H

#include "tables.h"

#include <map>
#include <string>
#include <utility>
#include <vector>

namespace hdnnp {

// One symmetry function as parsed from a symfunction_short line.
struct SymFn
{
    int    type = 0;
    int    ec = 0, e1 = 0, e2 = -1;
    double eta = 0.0, rs = 0.0, lambda = 0.0, zeta = 0.0, rc = 0.0;
    int    zetaInt = 0;
    bool   useIntegerPow = false;
    int    cutoffType = CT_HARD;
    double cutoffAlpha = 0.0;
    size_t line = 0;
    int    index = 0;                   // position after sorting
    std::vector<int> indexPerElement;   // Element.cpp:289-311
    // compact types (20-25): radial interval, angle interval in degrees,
    // subtype string, core function and asymmetry
    double rl = 0.0, angleLeft = 0.0, angleRight = 0.0;
    std::string subtype;
    int    core = CT_POLY2;
    bool   asym = false;
    int    scalingType = ST_NONE;
    double gmin = 0, gmax = 0, gmean = 0, gsigma = 0, smin = 0, smax = 0;
    double scalingFactor = 1.0;
};

// Settings file contents: keyword -> (value, line number), in file order.
class SettingsFile
{
public:
    void read(std::string const& fileName);
    bool has(std::string const& k) const { return kv_.count(k) > 0; }
    std::string value(std::string const& k) const;
    std::vector<std::pair<std::string, size_t>> all(std::string const& k) const;
    std::vector<std::string> const& lines() const { return lines_; }

private:
    std::multimap<std::string, std::pair<std::string, size_t>> kv_;
    std::vector<std::string> lines_;
};

struct GroupHost
{
    int type = 0, e1 = 0, e2 = -1;
    double rc = 0.0;
    std::vector<int> members;            // SF indices, sorted
};

struct ElementHost
{
    std::string symbol;
    int    Z = 0;
    double energyOffset = 0.0;
    std::vector<SymFn> sf;               // sorted (NN input order)
    std::vector<GroupHost> groups;       // reference grouping
    std::vector<std::vector<int>> table; // per neighbour element
    // network
    std::vector<int> width, act;
    std::vector<double> weights;         // per layer W[k][j] row-major, b[k]
    std::vector<size_t> woff, boff;
};

class Model
{
public:
    void load(std::string const& settingsFile, std::string const& scalingFile,
              std::string const& weightsFmt);
    // Replace scaling statistics of element e (Gmin, Gmax, Gmean, Gsigma).
    void setScaling(int e, int type, double smin, double smax,
                    double const* stats);

    int    numElements() const { return (int)el.size(); }
    int    maxNumSF() const;
    int    elementIndex(std::string const& sym) const;

    // Flat tables for upload.
    struct Flat
    {
        std::vector<ElemDesc>   elem;
        std::vector<CutCfg>     cfg;
        std::vector<GroupDesc>  group;
        std::vector<MemberDesc> member;
        std::vector<LamRun>     run;
        std::vector<ScaleDesc>  scale;
        std::vector<CompGroup>  cgroup;
        std::vector<CompMember> cmember;
        std::vector<int>        pairGroups, pairCount;
        std::vector<double>     weights;
    };
    // chunk: members per device group chunk (at most kMaxMembers).
    Flat flatten(int chunk = kMaxMembers) const;

    std::vector<ElementHost> el;
    double maxCutoffRadius = 0.0;
    bool   normalize = false;
    double meanEnergy = 0.0, convEnergy = 1.0, convLength = 1.0;
    int    cutoffType = CT_HARD;
    double cutoffAlpha = 0.0;
    bool   normalizeNeurons = false;
    SettingsFile settings;

private:
    void setupScaling(std::string const& scalingFile);
};

int atomicNumberOf(std::string const& symbol);
std::string symbolOf(int Z);

}  // namespace hdnnp

#endif
