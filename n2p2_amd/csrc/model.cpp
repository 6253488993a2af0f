// Host-side model setup: the Mode::setupGeneric sequence of n2p2 for a 2G
// HDNNP (settings, element map, cutoff, SF sort, SF memory tables, SF
// groups, scaling, network topology and weights), written to produce the
// flat device tables in tables.h.  Citations are relative to n2p2's
// src/libnnp.
#include "model.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace hdnnp {

namespace {

std::vector<std::string> words(std::string const& s)
{
    std::vector<std::string> v;
    std::istringstream is(s);
    std::string w;
    while (is >> w) v.push_back(w);
    return v;
}

std::string joinWords(std::string const& s)
{
    std::string out;
    for (auto const& w : words(s))
    {
        if (!out.empty()) out += ' ';
        out += w;
    }
    return out;
}

const char* const kTable[] = {
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al",
    "Si", "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe",
    "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr",
    "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn",
    "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm",
    "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",
    "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn",
    "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm", "Bk", "Cf",
    "Es", "Fm", "Md", "No"};

// NeuralNetwork.cpp:1078-1098 letter codes.
int activationCode(std::string const& c)
{
    static const std::map<std::string, int> m = {
        {"l", AF_IDENTITY}, {"t", AF_TANH}, {"s", AF_LOGISTIC},
        {"p", AF_SOFTPLUS}, {"r", AF_RELU}, {"g", AF_GAUSSIAN},
        {"c", AF_COS}, {"S", AF_REVLOGISTIC}, {"e", AF_EXP},
        {"h", AF_HARMONIC}};
    auto it = m.find(c);
    if (it == m.end())
        throw std::runtime_error("ERROR: Unknown activation function.\n");
    return it->second;
}

// Ordering of SymFncExpRad.cpp:55-75 / SymFncExpAngn.cpp:51-77 /
// SymFncExpAngw.cpp (centre element, type, cutoff, rc, eta, rs, [zeta,
// lambda], e1, [e2]); the weighted types (SymFncExpRadWeighted.cpp:39-58,
// SymFncExpAngnWeighted.cpp:41-64) have no neighbour elements (e1 = e2 =
// -1 here).
bool lessSF(SymFn const& a, SymFn const& b)
{
    if (a.ec != b.ec) return a.ec < b.ec;
    if (a.type != b.type) return a.type < b.type;
    if (sf_compact(a.type))
    {
        // SymFncCompRad.cpp / SymFncCompAngn.cpp / SymFncCompAngw.cpp and
        // the weighted variants: subtype, [e1, e2], rc, rl, [angles].
        if (a.subtype != b.subtype) return a.subtype < b.subtype;
        if (!sf_weighted(a.type))
        {
            if (a.e1 != b.e1) return a.e1 < b.e1;
            if (a.type != 20 && a.e2 != b.e2) return a.e2 < b.e2;
        }
        if (a.rc != b.rc) return a.rc < b.rc;
        if (a.rl != b.rl) return a.rl < b.rl;
        if (!sf_radial(a.type))
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
    if (!sf_radial(a.type))
    {
        if (a.zeta != b.zeta) return a.zeta < b.zeta;
        if (a.lambda != b.lambda) return a.lambda < b.lambda;
    }
    if (a.e1 != b.e1) return a.e1 < b.e1;
    if (a.type != 2 && a.e2 != b.e2) return a.e2 < b.e2;
    return false;
}

bool sameGroup(SymFn const& a, SymFn const& b)
{
    // SymGrpExpRad.cpp:72-109 / SymGrpExpAngn.cpp:68-105 addMember checks
    // (the weighted groups, SymGrpExpRadWeighted.cpp:68-99, have no
    // neighbour elements to compare).
    if (sf_compact(a.type))   // SymGrpCompRad.cpp:68-100 and alike
        return a.type == b.type && a.ec == b.ec && a.e1 == b.e1
            && a.e2 == b.e2;
    return a.type == b.type && a.cutoffType == b.cutoffType
        && a.cutoffAlpha == b.cutoffAlpha && a.ec == b.ec && a.rc == b.rc
        && a.e1 == b.e1 && (a.type == 2 || a.e2 == b.e2);
}

}  // namespace

int atomicNumberOf(std::string const& symbol)
{
    for (size_t i = 0; i < sizeof(kTable) / sizeof(*kTable); ++i)
        if (symbol == kTable[i]) return (int)i + 1;
    throw std::runtime_error("ERROR: Element \"" + symbol + "\" unknown.\n");
}

std::string symbolOf(int Z)
{
    if (Z < 1 || Z > (int)(sizeof(kTable) / sizeof(*kTable)))
        throw std::runtime_error("ERROR: Atomic number out of range.\n");
    return kTable[Z - 1];
}

// Settings.cpp:276-420: '#' and '!' start comments; first word = keyword.
void SettingsFile::read(std::string const& fileName)
{
    std::ifstream f(fileName);
    if (!f.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + fileName
                                 + "\".\n");
    kv_.clear();
    lines_.clear();
    std::string line;
    size_t n = 0;
    while (std::getline(f, line))
    {
        lines_.push_back(line);
        size_t const ln = n++;
        auto c = line.find_first_of("#!");
        if (c != std::string::npos) line.erase(c);
        std::string r = joinWords(line);
        if (r.empty()) continue;
        auto sp = r.find(' ');
        if (sp == std::string::npos)
            kv_.emplace(r, std::make_pair(std::string(), ln));
        else
            kv_.emplace(r.substr(0, sp), std::make_pair(r.substr(sp + 1), ln));
    }
}

std::string SettingsFile::value(std::string const& k) const
{
    auto it = kv_.find(k);
    if (it == kv_.end())
        throw std::runtime_error("ERROR: Keyword \"" + k + "\" not found.\n");
    return it->second.first;
}

std::vector<std::pair<std::string, size_t>>
SettingsFile::all(std::string const& k) const
{
    std::vector<std::pair<std::string, size_t>> v;
    auto r = kv_.equal_range(k);
    for (auto it = r.first; it != r.second; ++it) v.push_back(it->second);
    return v;
}

int Model::maxNumSF() const
{
    int m = 0;
    for (auto const& e : el) m = std::max(m, (int)e.sf.size());
    return m;
}

int Model::elementIndex(std::string const& sym) const
{
    for (size_t i = 0; i < el.size(); ++i)
        if (el[i].symbol == sym) return (int)i;
    return -1;
}

void Model::load(std::string const& settingsFile,
                 std::string const& scalingFile,
                 std::string const& weightsFmt)
{
    settings.read(settingsFile);
    SettingsFile const& st = settings;
    el.clear();

    // Mode.cpp:238-299 normalisation keywords (all or none).
    bool const hm = st.has("mean_energy"), he = st.has("conv_energy"),
               hl = st.has("conv_length");
    normalize = hm && he && hl;
    if (!normalize && (hm || he || hl))
        throw std::runtime_error("ERROR: Incorrect usage of normalization "
                                 "keywords.\n");
    if (normalize)
    {
        meanEnergy = std::atof(st.value("mean_energy").c_str());
        convEnergy = std::atof(st.value("conv_energy").c_str());
        convLength = std::atof(st.value("conv_length").c_str());
    }

    // ElementMap.cpp:36-54 (sorted by atomic number), Mode.cpp:322-366.
    std::vector<std::string> syms = words(st.value("elements"));
    std::sort(syms.begin(), syms.end(),
              [](std::string const& a, std::string const& b) {
                  return atomicNumberOf(a) < atomicNumberOf(b);
              });
    int const ne = (int)syms.size();
    if (ne > kMaxElements)
        throw std::runtime_error("ERROR: Too many elements for the engine.\n");
    if (std::atoi(st.value("number_of_elements").c_str()) != ne)
        throw std::runtime_error("ERROR: Inconsistent number of elements.\n");
    el.resize(ne);
    for (int i = 0; i < ne; ++i)
    {
        el[i].symbol = syms[i];
        el[i].Z = atomicNumberOf(syms[i]);
    }
    auto eidx = [&](std::string const& s) {
        int i = elementIndex(s);
        if (i < 0)
            throw std::runtime_error("ERROR: Element \"" + s
                                     + "\" not in element map.\n");
        return i;
    };
    for (auto const& v : st.all("atom_energy"))
    {
        auto a = words(v.first);
        el[eidx(a.at(0))].energyOffset = std::atof(a.at(1).c_str());
    }

    // Mode.cpp:520-613 cutoff_type <type> [alpha].
    {
        auto a = words(st.value("cutoff_type"));
        cutoffType = std::atoi(a.at(0).c_str());
        cutoffAlpha = a.size() > 1 ? std::atof(a.at(1).c_str()) : 0.0;
        if (cutoffType < CT_HARD || cutoffType > CT_POLY4)
            throw std::runtime_error("ERROR: Unknown cutoff type.\n");
        if (cutoffAlpha < 0.0 || cutoffAlpha >= 1.0)
            throw std::runtime_error("ERROR: 0 <= alpha < 1.0 is required.\n");
    }

    // Mode.cpp:615-690 + Element.cpp:82-139 + SymFnc*::setParameters.
    for (auto const& v : st.all("symfunction_short"))
    {
        auto a = words(v.first);
        SymFn s;
        s.ec = eidx(a.at(0));
        s.type = std::atoi(a.at(1).c_str());
        s.line = v.second;
        if (s.type == 2)            // SymFncExpRad.cpp:78-97
        {
            s.e1 = eidx(a.at(2));
            s.eta = std::atof(a.at(3).c_str());
            s.rs = std::atof(a.at(4).c_str());
            s.rc = std::atof(a.at(5).c_str());
        }
        else if (s.type == 3 || s.type == 9)   // SymFncBaseExpAng.cpp:50-93
        {
            s.e1 = eidx(a.at(2));
            s.e2 = eidx(a.at(3));
            s.eta = std::atof(a.at(4).c_str());
            s.lambda = std::atof(a.at(5).c_str());
            s.zeta = std::atof(a.at(6).c_str());
            s.rc = std::atof(a.at(7).c_str());
            if (a.size() > 8) s.rs = std::atof(a.at(8).c_str());
            if (s.e1 > s.e2) std::swap(s.e1, s.e2);
            s.zetaInt = (int)std::round(s.zeta);
            s.useIntegerPow = std::fabs(s.zeta - s.zetaInt)
                           <= std::numeric_limits<double>::min();
        }
        else if (s.type == 12)   // SymFncExpRadWeighted.cpp:71-90
        {
            s.e1 = s.e2 = -1;
            s.eta = std::atof(a.at(2).c_str());
            s.rs = std::atof(a.at(3).c_str());
            s.rc = std::atof(a.at(4).c_str());
        }
        else if (s.type == 13)   // SymFncExpAngnWeighted.cpp:92-123
        {
            s.e1 = s.e2 = -1;
            s.eta = std::atof(a.at(2).c_str());
            s.rs = std::atof(a.at(3).c_str());
            s.lambda = std::atof(a.at(4).c_str());
            s.zeta = std::atof(a.at(5).c_str());
            s.rc = std::atof(a.at(6).c_str());
            s.zetaInt = (int)std::round(s.zeta);
            s.useIntegerPow = std::fabs(s.zeta - s.zetaInt)
                           <= std::numeric_limits<double>::min();
            // The kernels shift the exponent by rs when rs > 0 (the narrow
            // type's rule); type 13 always shifts, which is the same for
            // rs >= 0.
            if (s.rs < 0.0)
                throw std::runtime_error("ERROR: Type 13 symmetry functions "
                                         "with rs < 0 are outside this "
                                         "engine.\n");
        }
        else if (sf_compact(s.type))
        {
            // SymFncCompRad.cpp / SymFncBaseCompAng.cpp and the weighted
            // variants' setParameters; SymFncBaseComp.cpp:48-100.
            size_t p = 2;
            s.e1 = s.e2 = -1;
            if (!sf_weighted(s.type))
            {
                s.e1 = eidx(a.at(p++));
                if (s.type != 20) s.e2 = eidx(a.at(p++));
                if (s.type != 20 && s.e1 > s.e2) std::swap(s.e1, s.e2);
            }
            s.rl = std::atof(a.at(p++).c_str());
            s.rc = std::atof(a.at(p++).c_str());
            if (!sf_radial(s.type))
            {
                s.angleLeft = std::atof(a.at(p++).c_str());
                s.angleRight = std::atof(a.at(p++).c_str());
            }
            s.subtype = a.at(p++);
            if (s.rl > s.rc)
                throw std::runtime_error("ERROR: Lower radial boundary >= "
                                         "upper radial boundary.\n");
            std::string const& st = s.subtype;
            if (st.size() < 1 || st.size() > 3)
                throw std::runtime_error("ERROR: Invalid compact function "
                                         "type specification: \"" + st
                                         + "\".\n");
            if (st[0] == 'p' && st.size() >= 2 && st[1] >= '1' && st[1] <= '4')
            {
                s.core = CT_POLY1 + (st[1] - '1');
                if (st.size() == 3)
                {
                    if (st[2] != 'a')
                        throw std::runtime_error("ERROR: Invalid polynom "
                                                 "specifier: \"" + st
                                                 + "\".\n");
                    s.asym = true;
                }
            }
            else if (st[0] == 'e') s.core = CT_EXP;
            else
                throw std::runtime_error("ERROR: Unknown compact SF type: \""
                                         + st + "\".\n");
            if (!sf_radial(s.type))
            {
                double const al = s.angleLeft, ar = s.angleRight;
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
            }
        }
        else
            throw std::runtime_error(
                "ERROR: Symmetry function type " + a.at(1)
                + " is outside this engine (types 2, 3, 9, 12, 13, 20-25 are "
                "supported).\n");
        s.indexPerElement.assign(ne, -1);
        el[s.ec].sf.push_back(s);
    }

    maxCutoffRadius = 0.0;
    for (auto& e : el)
    {
        if (normalize)        // SymFnc*::changeLengthUnit
            for (auto& s : e.sf)
            {
                s.eta /= convLength * convLength;
                s.rs *= convLength;
                s.rc *= convLength;
                s.rl *= convLength;   // SymFncCompRad::changeLengthUnit
            }
        std::sort(e.sf.begin(), e.sf.end(), lessSF);   // Element.cpp:154-166
        for (size_t i = 0; i < e.sf.size(); ++i)
        {
            e.sf[i].index = (int)i;
            e.sf[i].cutoffType = cutoffType;           // Element.cpp:328-342
            e.sf[i].cutoffAlpha = cutoffAlpha;
            maxCutoffRadius = std::max(maxCutoffRadius, e.sf[i].rc);
        }
        if (e.sf.empty())
            throw std::runtime_error("ERROR: Element " + e.symbol
                                     + " has no symmetry functions.\n");
        // Element.cpp:289-311 per-neighbour-element derivative tables.
        e.table.assign(ne, {});
        for (auto& s : e.sf)
            for (int i = 0; i < ne; ++i)
                if (sf_weighted(s.type) || i == s.e1
                    || (s.type != 2 && s.type != 20 && i == s.e2))
                {
                    s.indexPerElement[i] = (int)e.table[i].size();
                    e.table[i].push_back(s.index);
                }
        // Element.cpp:194-287 groups (first matching group, in SF order).
        e.groups.clear();
        for (auto const& s : e.sf)
        {
            bool added = false;
            for (auto& g : e.groups)
                if (sameGroup(e.sf[g.members.front()], s))
                {
                    g.members.push_back(s.index);
                    added = true;
                    break;
                }
            if (!added)
            {
                GroupHost g;
                g.type = s.type;
                g.e1 = s.e1;
                g.e2 = sf_radial(s.type) ? -1 : s.e2;
                g.rc = s.rc;
                g.members.push_back(s.index);
                e.groups.push_back(g);
            }
        }
    }

    for (auto const& e : el)
        for (auto const& g : e.groups)
            if (sf_compact(g.type) && (int)g.members.size() > kMaxCompMembers)
                throw std::runtime_error("ERROR: More than 96 compact "
                                         "symmetry functions in one group "
                                         "are outside this engine.\n");

    setupScaling(scalingFile);

    // Mode.cpp:1158-1443 network topology ("short" keywords).
    int globalHidden = 0;
    std::vector<int> nlayers(ne, 0);
    if (st.has("global_hidden_layers_short"))
    {
        globalHidden = std::atoi(st.value("global_hidden_layers_short").c_str());
        std::fill(nlayers.begin(), nlayers.end(), globalHidden + 2);
    }
    for (auto const& v : st.all("element_hidden_layers_short"))
    {
        auto a = words(v.first);
        nlayers[eidx(a.at(0))] = std::atoi(a.at(1).c_str()) + 2;
    }
    std::vector<std::string> gNodes, gAct;
    if (st.has("global_nodes_short"))
    {
        gNodes = words(st.value("global_nodes_short"));
        if ((int)gNodes.size() != globalHidden)
            throw std::runtime_error("ERROR: Inconsistent global NN topology "
                                     "keyword \"global_nodes_short\".\n");
    }
    if (st.has("global_activation_short"))
    {
        gAct = words(st.value("global_activation_short"));
        if ((int)gAct.size() - 1 != globalHidden)
            throw std::runtime_error("ERROR: Inconsistent global NN topology "
                                     "keyword \"global_activation_short\".\n");
    }
    for (int i = 0; i < ne; ++i)
    {
        ElementHost& e = el[i];
        int const L = nlayers[i];
        if (L < 3)
            throw std::runtime_error("ERROR: Neural network must have at "
                                     "least three layers\n");
        if (L > kMaxLayers)
            throw std::runtime_error("ERROR: Too many network layers for the "
                                     "engine.\n");
        e.width.assign(L, 0);
        e.act.assign(L, AF_UNSET);
        e.width[0] = (int)e.sf.size();
        e.act[0] = AF_IDENTITY;
        e.width[L - 1] = 1;
        if (L != globalHidden + 2) continue;
        for (int j = 1; j < L; ++j)
        {
            if (j == L - 1 && !gAct.empty())
                e.act[j] = activationCode(gAct.at(L - 2));
            else
            {
                if (!gNodes.empty()) e.width[j] = std::atoi(gNodes.at(j - 1).c_str());
                if (!gAct.empty()) e.act[j] = activationCode(gAct.at(j - 1));
            }
        }
    }
    for (auto const& v : st.all("element_nodes_short"))
    {
        auto a = words(v.first);
        ElementHost& e = el[eidx(a.at(0))];
        if (e.width.size() != a.size() + 1)
            throw std::runtime_error("ERROR: Inconsistent per-element NN "
                                     "topology keyword \"element_nodes_short\".\n");
        for (size_t j = 1; j + 2 < e.width.size(); ++j)
            e.width[j] = std::atoi(a.at(j).c_str());
    }
    for (auto const& v : st.all("element_activation_short"))
    {
        auto a = words(v.first);
        ElementHost& e = el[eidx(a.at(0))];
        if (e.width.size() != a.size())
            throw std::runtime_error("ERROR: Inconsistent per-element NN "
                                     "topology keyword "
                                     "\"element_activation_short\".\n");
        for (size_t j = 1; j + 1 < e.width.size(); ++j)
            e.act[j] = activationCode(a.at(j));
    }
    normalizeNeurons = st.has("normalize_nodes");

    // Mode.cpp:2270-2289 weights, NeuralNetwork.cpp:147-169 file order:
    // per layer, for each input j, for each neuron k: W[k][j]; then biases.
    // An empty weightsFmt leaves the weights zero (the symmetry-function
    // tools, e.g. nnp-scaling, set up no networks).
    for (auto& e : el)
    {
        for (size_t j = 0; j < e.width.size(); ++j)
            if (e.width[j] <= 0 || e.act[j] == AF_UNSET)
                throw std::runtime_error(
                    "ERROR: NN \"short\", element " + e.symbol
                    + ": layer topology or activation unset.\n");
        char buf[4096] = "";
        std::vector<double> file;
        if (!weightsFmt.empty())
        {
            std::snprintf(buf, sizeof buf, weightsFmt.c_str(), (size_t)e.Z);
            std::ifstream f(buf);
            if (!f.is_open())
                throw std::runtime_error(std::string("ERROR: Could not open "
                                         "file: \"") + buf + "\".\n");
            std::string line;
            while (std::getline(f, line))
            {
                if (line.empty() || line[0] == '#') continue;
                auto w = words(line);
                if (!w.empty()) file.push_back(std::atof(w[0].c_str()));
            }
        }
        int const L = (int)e.width.size();
        if (weightsFmt.empty())
        {
            size_t nw = 0;
            for (int l = 1; l < L; ++l)
                nw += (size_t)e.width[l] * e.width[l - 1] + e.width[l];
            file.assign(nw, 0.0);
        }
        e.weights.clear();
        e.woff.assign(L, 0);
        e.boff.assign(L, 0);
        size_t need = 0;
        for (int l = 1; l < L; ++l)
            need += (size_t)e.width[l] * e.width[l - 1] + e.width[l];
        if (file.size() < need)
            throw std::runtime_error(std::string("ERROR: Too few weights in "
                                     "file \"") + buf + "\".\n");
        size_t count = 0;
        for (int l = 1; l < L; ++l)
        {
            int const nin = e.width[l - 1], nout = e.width[l];
            e.woff[l] = e.weights.size();
            e.weights.resize(e.weights.size() + (size_t)nin * nout);
            for (int j = 0; j < nin; ++j)
                for (int k = 0; k < nout; ++k)
                    e.weights[e.woff[l] + (size_t)k * nin + j] = file[count++];
            e.boff[l] = e.weights.size();
            for (int k = 0; k < nout; ++k) e.weights.push_back(file[count++]);
        }
    }
}

// Mode.cpp:712-840 keyword -> scaling type; scaling.data lines (no '#')
// are consumed element by element in SF order.
void Model::setupScaling(std::string const& scalingFile)
{
    SettingsFile const& st = settings;
    int type = ST_NONE;
    double smin = 0.0, smax = 0.0;
    std::vector<std::string> lines;
    if (!scalingFile.empty())
    {
        bool const sc = st.has("scale_symmetry_functions");
        bool const ce = st.has("center_symmetry_functions");
        if (sc && !ce) type = ST_SCALE;
        else if (!sc && ce) type = ST_CENTER;
        else if (sc && ce) type = ST_SCALECENTER;
        else if (st.has("scale_symmetry_functions_sigma")) type = ST_SCALESIGMA;
        if (type == ST_SCALE || type == ST_SCALECENTER || type == ST_SCALESIGMA)
        {
            smin = st.has("scale_min_short")
                 ? std::atof(st.value("scale_min_short").c_str()) : 0.0;
            smax = st.has("scale_max_short")
                 ? std::atof(st.value("scale_max_short").c_str()) : 1.0;
        }
        std::ifstream f(scalingFile);
        if (!f.is_open())
            throw std::runtime_error("ERROR: Could not open file: \""
                                     + scalingFile + "\".\n");
        std::string line;
        while (std::getline(f, line))
            if (!line.empty() && line[0] != '#') lines.push_back(line);
    }
    size_t pos = 0;
    for (int e = 0; e < (int)el.size(); ++e)
    {
        std::vector<double> stats(4 * el[e].sf.size(), 0.0);
        if (!scalingFile.empty())
            for (size_t i = 0; i < el[e].sf.size(); ++i)
            {
                if (pos >= lines.size())
                    throw std::runtime_error("ERROR: Too few lines in scaling "
                                             "file.\n");
                auto w = words(lines[pos++]);
                if (w.size() < 5)
                    throw std::runtime_error("ERROR: Malformed scaling line.\n");
                stats[4 * i] = std::atof(w[2].c_str());
                stats[4 * i + 1] = std::atof(w[3].c_str());
                stats[4 * i + 2] = std::atof(w[4].c_str());
                stats[4 * i + 3] = w.size() > 5 ? std::atof(w[5].c_str()) : 0.0;
            }
        setScaling(e, type, smin, smax, stats.data());
    }
}

// SymFnc.cpp:52-98 scaling factor per type.
void Model::setScaling(int e, int type, double smin, double smax,
                       double const* stats)
{
    for (size_t i = 0; i < el.at(e).sf.size(); ++i)
    {
        SymFn& s = el[e].sf[i];
        s.scalingType = type;
        s.gmin = stats[4 * i];
        s.gmax = stats[4 * i + 1];
        s.gmean = stats[4 * i + 2];
        s.gsigma = stats[4 * i + 3];
        s.smin = smin;
        s.smax = smax;
        if (type == ST_SCALE || type == ST_SCALECENTER)
            s.scalingFactor = (smax - smin) / (s.gmax - s.gmin);
        else if (type == ST_SCALESIGMA)
            s.scalingFactor = (smax - smin) / s.gsigma;
        else
            s.scalingFactor = 1.0;
    }
}

Model::Flat Model::flatten(int chunk) const
{
    if (chunk < 1 || chunk > kMaxMembers)
        throw std::runtime_error("ERROR: Group chunk size out of range.\n");
    Flat F;
    int const ne = (int)el.size();
    F.pairCount.assign((size_t)ne * ne * ne, 0);
    F.pairGroups.assign((size_t)ne * ne * ne * kMaxPairGroups, -1);
    for (int c = 0; c < ne; ++c)
    {
        ElementHost const& e = el[c];
        ElemDesc d{};
        d.nsf = (int)e.sf.size();
        d.sf0 = (int)F.scale.size();
        for (int k = 0; k < ne; ++k) d.nsfd[k] = (int)e.table[k].size();
        for (auto const& s : e.sf)
        {
            ScaleDesc sd{};
            sd.type = s.scalingType;
            sd.smin = s.smin;
            sd.sfac = s.scalingFactor;
            sd.gmin = s.gmin;
            sd.gmax = s.gmax;
            sd.gmean = s.gmean;
            F.scale.push_back(sd);
        }
        // Cutoff configurations: one per distinct rc (type and alpha are
        // global in a 2G settings file).
        d.cfg0 = (int)F.cfg.size();
        std::vector<double> rcs;
        for (auto const& g : e.groups)
            if (!sf_compact(g.type)
                && std::find(rcs.begin(), rcs.end(), g.rc) == rcs.end())
                rcs.push_back(g.rc);
        if ((int)rcs.size() > kMaxCutCfg)
            throw std::runtime_error("ERROR: More distinct cutoff radii per "
                                     "element than the engine supports.\n");
        for (double rc : rcs)
        {
            CutCfg cc{};
            cc.type = cutoffType;
            cc.rc = rc;
            cc.rcinv = 1.0 / rc;                 // CutoffFunction.cpp:95-100
            cc.rci = rc * cutoffAlpha;           // CutoffFunction.cpp:102-112
            cc.iw = 1.0 / (rc - cc.rci);
            double const E = std::exp(1.0);
            cc.tanhpre = std::pow((E + 1 / E) / (E - 1 / E), 3);
            F.cfg.push_back(cc);
        }
        d.ncfg = (int)rcs.size();
        // Group chunks: radial first, then angular.
        d.g0 = (int)F.group.size();
        for (int pass = 0; pass < 2; ++pass)
            for (auto const& g : e.groups)
            {
                if (sf_compact(g.type)) continue;   // own tables below
                if ((pass == 0) != sf_radial(g.type)) continue;
                // Members in SF order (SymGrp*::sortMembers), then, for the
                // angular types, stably ordered by (eta, rs, lambda, zeta):
                // an (eta, rs) run shares one exponential, an (eta, rs,
                // lambda) run one chain of powers of plambda.  Each member
                // keeps its own SF slot, so the order is free.
                std::vector<int> mem = g.members;
                std::sort(mem.begin(), mem.end());
                if (!sf_radial(g.type))
                    std::stable_sort(mem.begin(), mem.end(),
                                     [&](int a, int b) {
                        SymFn const& x = e.sf[a];
                        SymFn const& y = e.sf[b];
                        if (x.eta != y.eta) return x.eta < y.eta;
                        if (x.rs != y.rs) return x.rs < y.rs;
                        if (x.lambda != y.lambda) return x.lambda < y.lambda;
                        return x.zeta < y.zeta;
                    });
                for (size_t c0 = 0; c0 < mem.size(); c0 += (size_t)chunk)
                {
                    GroupDesc gd{};
                    gd.type = g.type;
                    gd.e1 = g.e1;
                    gd.e2 = g.e2;
                    gd.rc = g.rc;
                    gd.rc2 = g.rc * g.rc;
                    gd.cfg = (int)(std::find(rcs.begin(), rcs.end(), g.rc)
                                   - rcs.begin());
                    gd.mem0 = (int)F.member.size();
                    size_t const c1 = std::min(mem.size(),
                                               c0 + (size_t)chunk);
                    gd.nmem = (int)(c1 - c0);
                    // zeta = 2^level: plambda^(zeta-1) is the level-th
                    // prefix product x x^2 x^4 ... of pow_int's squaring
                    // chain; -1 for any other zeta.
                    auto levelOf = [&](SymFn const& s) {
                        if (!s.useIntegerPow) return -1;
                        for (int l = 0; l <= kMaxLevel; ++l)
                            if (s.zetaInt == (1 << l)) return l;
                        return -1;
                    };
                    // Within each (eta, rs, lambda) run: chain members by
                    // level, then generic members.
                    std::vector<int> cm(mem.begin() + c0, mem.begin() + c1);
                    if (!sf_radial(g.type))
                    {
                        size_t r0 = 0;
                        while (r0 < cm.size())
                        {
                            size_t r1 = r0 + 1;
                            SymFn const& f = e.sf[cm[r0]];
                            while (r1 < cm.size()
                                   && e.sf[cm[r1]].eta == f.eta
                                   && e.sf[cm[r1]].rs == f.rs
                                   && e.sf[cm[r1]].lambda == f.lambda)
                                ++r1;
                            std::stable_partition(
                                cm.begin() + r0, cm.begin() + r1,
                                [&](int a) { return levelOf(e.sf[a]) >= 0; });
                            r0 = r1;
                        }
                    }
                    gd.run0 = (int)F.run.size();
                    for (size_t q = 0; q < cm.size(); ++q)
                    {
                        SymFn const& s = e.sf[cm[q]];
                        MemberDesc md{};
                        md.eta = s.eta;
                        md.rs = s.rs;
                        md.lambda = s.lambda;
                        md.zeta = s.zeta;
                        md.zl = s.zeta * s.lambda;
                        md.sfac = s.scalingFactor;
                        if (!sf_radial(g.type))
                        {
                            // SymGrpExpAngn.cpp:107-148 then
                            // SymGrpBaseExpAng.cpp:45-57.
                            md.fnorm = std::pow(2.0, 1.0 - s.zeta);
                            md.fnorm *= s.scalingFactor;
                            md.fderiv = 2.0 * s.eta / s.zeta / s.lambda;
                            md.gfac = md.fnorm / s.scalingFactor;
                        }
                        else
                        {
                            md.fnorm = 1.0;
                            md.fderiv = 0.0;
                            md.gfac = 1.0;
                        }
                        md.zint = s.zetaInt;
                        md.useint = s.useIntegerPow ? 1 : 0;
                        // calculateExp: new eta or rs, or first of a chunk.
                        if (q == 0) md.calcexp = 1;
                        else
                        {
                            SymFn const& p = e.sf[cm[q - 1]];
                            md.calcexp = (p.eta != s.eta || p.rs != s.rs);
                        }
                        md.sf = s.index;
                        md.newlam = md.calcexp
                                 || e.sf[cm[q - 1]].lambda != s.lambda;
                        md.level = levelOf(s);
                        if (!sf_radial(g.type))
                        {
                            if (md.newlam)
                            {
                                LamRun lr{};
                                lr.lambda = s.lambda;
                                lr.eta = s.eta;
                                lr.rs = s.rs;
                                lr.newexp = md.calcexp;
                                lr.m0 = (int)q;
                                lr.lmax = -1;
                                for (int l = 0; l <= kMaxLevel; ++l)
                                    lr.lvEnd[l] = (int)q;
                                F.run.push_back(lr);
                                gd.nrun++;
                            }
                            LamRun& lr = F.run.back();
                            lr.mend = (int)q + 1;
                            if (md.level >= 0)
                            {
                                lr.lmax = std::max(lr.lmax, md.level);
                                for (int l = md.level; l <= kMaxLevel; ++l)
                                    lr.lvEnd[l] = (int)q + 1;
                            }
                        }
                        F.member.push_back(md);
                    }
                    if (pass == 0) d.nrad++;
                    F.group.push_back(gd);
                }
            }
        d.ng = (int)F.group.size() - d.g0;
        // Compact groups: members in SF order (SymGrpComp*::sortMembers).
        d.cg0 = (int)F.cgroup.size();
        for (auto const& g : e.groups)
        {
            if (!sf_compact(g.type)) continue;
            std::vector<int> mem = g.members;
            std::sort(mem.begin(), mem.end());
            CompGroup cg{};
            cg.type = g.type;
            cg.e1 = g.e1;
            cg.e2 = sf_radial(g.type) ? -1 : g.e2;
            cg.nmem = (int)mem.size();
            cg.mem0 = (int)F.cmember.size();
            cg.rmax = 0.0;
            for (size_t q = 0; q < mem.size(); ++q)
            {
                SymFn const& s = e.sf[mem[q]];
                CompMember cm{};
                cm.rl = s.rl;
                cm.rc = s.rc;
                cm.rcenter = (s.rc + s.rl) / 2.0;       // setLeftRight
                cm.rscale = 1.0 / ((s.rc - s.rl) / 2.0);
                // angleLeft * M_PI / 180.0 as SymFncBaseCompAng.cpp:106-108
                cm.al = s.angleLeft * M_PI / 180.0;
                cm.ar = s.angleRight * M_PI / 180.0;
                cm.acenter = (cm.ar + cm.al) / 2.0;
                cm.ascale = 1.0 / ((cm.ar - cm.al) / 2.0);
                cm.sfac = s.scalingFactor;
                cm.core = s.core;
                cm.asym = s.asym ? 1 : 0;
                cm.sf = s.index;
                cm.newrad = (q == 0 || e.sf[mem[q - 1]].rc != s.rc
                             || e.sf[mem[q - 1]].rl != s.rl) ? 1 : 0;
                cg.rmax = std::max(cg.rmax, s.rc);
                F.cmember.push_back(cm);
            }
            F.cgroup.push_back(cg);
        }
        d.ncg = (int)F.cgroup.size() - d.cg0;
        // Angular chunks per neighbour-element pair (both orders).
        for (int gi = d.g0 + d.nrad; gi < d.g0 + d.ng; ++gi)
        {
            GroupDesc const& gd = F.group[gi];
            if (sf_weighted(gd.type)) continue;   // not tied to an element pair
            for (int o = 0; o < 2; ++o)
            {
                int const a = o ? gd.e2 : gd.e1, b = o ? gd.e1 : gd.e2;
                if (o == 1 && a == b) break;
                size_t const pb = ((size_t)c * ne + a) * ne + b;
                if (F.pairCount[pb] >= kMaxPairGroups)
                    throw std::runtime_error("ERROR: Too many angular groups "
                                             "per element pair.\n");
                F.pairGroups[pb * kMaxPairGroups + F.pairCount[pb]++] = gi;
            }
        }
        // Network.
        d.nlayers = (int)e.width.size();
        d.maxwidth = 0;
        for (int l = 0; l < d.nlayers; ++l)
        {
            d.width[l] = e.width[l];
            d.act[l] = e.act[l];
            d.woff[l] = (int)(F.weights.size() + e.woff[l]);
            d.boff[l] = (int)(F.weights.size() + e.boff[l]);
            if (l > 0) d.maxwidth = std::max(d.maxwidth, e.width[l]);
        }
        d.woff[0] = d.boff[0] = 0;
        d.normalize = normalizeNeurons ? 1 : 0;
        F.weights.insert(F.weights.end(), e.weights.begin(), e.weights.end());
        F.elem.push_back(d);
    }
    return F;
}

}  // namespace hdnnp
