// Mode.h and Prediction.h of the nnp layer: the reference's setup and
// compute sequence (Mode.cpp, Prediction.cpp) on top of the HIP engine's
// C ABI.  Errors of the engine are rethrown as std::runtime_error with the
// engine's message (which starts with "ERROR: " as the reference's do).
#include "Mode.h"
#include "Prediction.h"
#include "utility.h"

#include "hdnnp.h"

#include <algorithm>
#include <cinttypes>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace nnp
{

namespace
{

void check(int rc)
{
    if (rc != HDNNP_OK) throw std::runtime_error(hdnnp_last_error());
}

}  // namespace

Mode::Mode()
{
    char const* p = std::getenv("N2P2_AMD_PRECISION");
    if (p && std::strcmp(p, "mixed") == 0) precision = HDNNP_PRECISION_MIXED;
    else if (p && std::strcmp(p, "mixed16") == 0)
        precision = HDNNP_PRECISION_MIXED16;
    else if (p && std::strcmp(p, "mixedbf16") == 0)
        precision = HDNNP_PRECISION_MIXED_BF16;
    char const* d = std::getenv("N2P2_AMD_DEVICE");
    if (d && *d) device = std::atoi(d);
}

Mode::~Mode()
{
    if (ctx) hdnnp_ctx_destroy(ctx);
    if (model) hdnnp_model_free(model);
}

void Mode::initialize()
{
    log << "\n";
    log << "*****************************************"
           "**************************************\n";
    log << "\n";
    log << "n2p2-compatible HDNNP prediction on AMD Instinct GPUs (HIP "
           "engine, 2G-HDNNP energy and forces).\n";
    log << "\n";
    log << "*****************************************"
           "**************************************\n";
}

void Mode::loadSettingsFile(std::string const& fileName)
{
    std::ifstream file(fileName.c_str());
    if (!file.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + fileName
                                 + "\".\n");
    settingsFile = fileName;
    keywords.clear();
    std::string line;
    while (std::getline(file, line))
    {
        std::size_t const hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        std::string const r = reduce(line);
        if (r.empty()) continue;
        std::size_t const sp = r.find(' ');
        std::string const key = r.substr(0, sp);
        std::string const value = sp == std::string::npos ? "" : r.substr(sp + 1);
        keywords.emplace(key, value);
    }
    log << strpr("Settings file: %s (%zu keyword lines).\n", fileName.c_str(),
                 keywords.size());
}

bool Mode::settingsKeywordExists(std::string const& keyword) const
{
    return keywords.count(keyword) > 0;
}

std::string Mode::settingsGetValue(std::string const& keyword) const
{
    auto it = keywords.find(keyword);
    if (it == keywords.end())
        throw std::runtime_error("ERROR: Keyword \"" + keyword
                                 + "\" not found.\n");
    return it->second;
}

void Mode::setupGeneric(std::string const& nnpDir, bool skipNormalize,
                        bool initialHardness)
{
    (void)initialHardness;
    if (!nnpDir.empty() && settingsFile.find('/') == std::string::npos)
        settingsFile = nnpDir + "/" + settingsFile;
    if (!skipNormalize) setupNormalization();
    setupElementMap();
    setupElements();
    setupCutoff();
    setupSymmetryFunctions();
    setupSymmetryFunctionGroups();
    setupNeuralNetwork();
}

// Mode.cpp:238-299: normalisation is used when all three keywords are set.
void Mode::setupNormalization(bool standalone)
{
    (void)standalone;
    if (settingsKeywordExists("mean_energy")
        && settingsKeywordExists("conv_energy")
        && settingsKeywordExists("conv_length"))
    {
        normalize = true;
        meanEnergy = std::atof(settingsGetValue("mean_energy").c_str());
        convEnergy = std::atof(settingsGetValue("conv_energy").c_str());
        convLength = std::atof(settingsGetValue("conv_length").c_str());
        log << strpr("Data set normalization is used: mean energy %24.16E, "
                     "conv_energy %24.16E, conv_length %24.16E.\n",
                     meanEnergy, convEnergy, convLength);
    }
    else
    {
        normalize = false;
        log << "Data set normalization is not used.\n";
    }
}

void Mode::setupElementMap()
{
    if (settingsKeywordExists("nnp_type"))
    {
        int const t = std::atoi(settingsGetValue("nnp_type").c_str());
        if (t != 2)
            throw std::runtime_error("ERROR: This engine evaluates "
                                     "2G-HDNNPs only (nnp_type 2).\n");
    }
    elementMap.deregisterElements();
    numElements = elementMap.registerElements(settingsGetValue("elements"));
    log << strpr("Number of elements is %zu.\n", numElements);
}

void Mode::setupElements()
{
    energyOffset.assign(numElements, 0.0);
    auto range = keywords.equal_range("atom_energy");
    for (auto it = range.first; it != range.second; ++it)
    {
        std::vector<std::string> w = split(reduce(it->second));
        energyOffset.at(elementMap[w.at(0)]) = std::atof(w.at(1).c_str());
    }
}

// The maximum cutoff over the symmetry-function lines (Mode::setupCutoff
// and SymFnc*::setParameters): rc is field 6 of a radial line and field 8
// of an angular one (element, type, neighbours, parameters ...).
void Mode::setupCutoff()
{
    maxCutoffRadius = 0.0;
    auto range = keywords.equal_range("symfunction_short");
    for (auto it = range.first; it != range.second; ++it)
    {
        std::vector<std::string> w = split(reduce(it->second));
        if (w.size() < 2) continue;
        int const type = std::atoi(w[1].c_str());
        std::size_t const field = (type == 2) ? 5 : 7;
        if (w.size() > field)
            maxCutoffRadius = std::max(maxCutoffRadius,
                                       std::atof(w[field].c_str()));
    }
    if (normalize) maxCutoffRadius *= convLength;
}

void Mode::setupSymmetryFunctions() {}
void Mode::setupSymmetryFunctionGroups() {}
void Mode::setupSymmetryFunctionMemory(bool verbose) { (void)verbose; }
void Mode::setupNeuralNetwork() {}

void Mode::setupSymmetryFunctionScalingNone()
{
    scalingNone = true;
    scalingFile.clear();
    if (model) buildModel();
}

void Mode::setupSymmetryFunctionScaling(std::string const& fileName)
{
    scalingNone = false;
    scalingFile = fileName;
    if (model) buildModel();
}

void Mode::setupSymmetryFunctionStatistics(bool collectStatistics,
                                           bool collectExtrapolationWarnings,
                                           bool writeExtrapolationWarnings,
                                           bool stopOnExtrapolationWarnings)
{
    (void)collectStatistics;
    checkExtrapolationWarnings = collectExtrapolationWarnings
                              || writeExtrapolationWarnings
                              || stopOnExtrapolationWarnings;
    this->collectExtrapolationWarnings = collectExtrapolationWarnings;
    this->writeExtrapolationWarnings = writeExtrapolationWarnings;
    this->stopOnExtrapolationWarnings = stopOnExtrapolationWarnings;
}

void Mode::setupNeuralNetworkWeights(
    std::map<std::string, std::string> fileNameFormats)
{
    setupNeuralNetworkWeights("", fileNameFormats);
}

void Mode::setupNeuralNetworkWeights(
    std::string directoryPrefix,
    std::map<std::string, std::string> fileNameFormats)
{
    std::string fmt = "weights.%03zu.data";
    auto it = fileNameFormats.find("short");
    if (it != fileNameFormats.end()) fmt = it->second;
    weightsFormat = directoryPrefix + fmt;
    buildModel();
}

void Mode::buildModel()
{
    if (ctx) hdnnp_ctx_destroy(ctx);
    ctx = nullptr;
    if (model) hdnnp_model_free(model);
    model = nullptr;
    upValid = false;
    check(hdnnp_model_load(settingsFile.c_str(),
                           scalingNone ? nullptr : scalingFile.c_str(),
                           weightsFormat.c_str(), &model));
    numElements = (std::size_t)hdnnp_model_num_elements(model);
    maxCutoffRadius = hdnnp_model_max_cutoff(model);
    double me = 0.0, ce = 1.0, cl = 1.0;
    normalize = hdnnp_model_normalization(model, &me, &ce, &cl) != 0;
    meanEnergy = me;
    convEnergy = ce;
    convLength = cl;
    energyOffset.assign(numElements, 0.0);
    numSF.assign(numElements, 0);
    sfType.assign(numElements, {});
    ewData.assign(numElements, {});
    for (std::size_t e = 0; e < numElements; ++e)
    {
        energyOffset[e] = hdnnp_model_energy_offset(model, (int)e);
        numSF[e] = (std::size_t)hdnnp_model_num_sf(model, (int)e);
        std::vector<double> tab(12 * numSF[e]);
        if (numSF[e]) check(hdnnp_model_sf_table(model, (int)e, tab.data()));
        for (std::size_t k = 0; k < numSF[e]; ++k)
            sfType[e].push_back((std::size_t)tab[12 * k]);
    }
    log << strpr("Model set up on the engine: %zu elements, maximum cutoff "
                 "%.6f.\n", numElements, maxCutoffRadius);
}

void Mode::ensureEngine() const
{
    if (!model)
        throw std::runtime_error("ERROR: Neural network weights are not set "
                                 "up (setupNeuralNetworkWeights).\n");
    if (!ctx)
    {
        Mode* self = const_cast<Mode*>(this);
        check(hdnnp_ctx_create(model, device < 0 ? 0 : device, &self->ctx));
        check(hdnnp_ctx_set_precision(ctx, precision));
    }
}

void Mode::setPrecision(int p)
{
    precision = p;
    if (ctx) check(hdnnp_ctx_set_precision(ctx, p));
}

// Loads the structure into the engine: everything when the atom count,
// elements or cell changed, else the positions only.  The neighbour list
// is then rebuilt with the structure's cutoff.
void Mode::upload(Structure& structure) const
{
    ensureEngine();
    std::size_t const n = structure.atoms.size();
    std::vector<double> r(3 * n);
    std::vector<int32_t> el(n);
    for (std::size_t i = 0; i < n; ++i)
    {
        Atom const& a = structure.atoms[i];
        for (int k = 0; k < 3; ++k) r[3 * i + k] = a.r[k];
        el[i] = (int32_t)a.element;
    }
    double box[9];
    for (int k = 0; k < 3; ++k)
        for (int c = 0; c < 3; ++c) box[3 * k + c] = structure.box[k][c];
    bool same = upValid && n == upAtoms && el == upElem
             && structure.isPeriodic == upPeriodic;
    if (same && structure.isPeriodic)
        same = std::memcmp(box, upBox, sizeof box) == 0;
    if (same) check(hdnnp_set_positions(ctx, r.data()));
    else
    {
        check(hdnnp_set_structure(ctx, (int64_t)n, r.data(), el.data(),
                                  structure.isPeriodic ? box : nullptr));
        upAtoms = n;
        upElem = el;
        upPeriodic = structure.isPeriodic;
        std::memcpy(upBox, box, sizeof box);
        upValid = true;
    }
    double const rc = structure.hasNeighborList && structure.neighborCutoff > 0
                    ? structure.neighborCutoff : maxCutoffRadius;
    check(hdnnp_neighbor_list(ctx, rc));
    if (structure.isPeriodic)
    {
        int32_t p[3];
        check(hdnnp_get_pbc(ctx, p));
        for (int k = 0; k < 3; ++k) structure.pbc[k] = p[k];
    }
}

void Mode::calculateSymmetryFunctions(Structure& structure,
                                      bool const derivatives)
{
    calculateSymmetryFunctionGroups(structure, derivatives);
}

// Mode.cpp:1561-1640: skipped when the structure already holds the results;
// G is written back per atom, the per-neighbour derivatives stay on the
// device for the force kernel.
void Mode::calculateSymmetryFunctionGroups(Structure& structure,
                                           bool const derivatives)
{
    if (structure.hasSymmetryFunctionDerivatives) return;
    if (structure.hasSymmetryFunctions && !derivatives) return;
    upload(structure);
    check(hdnnp_symmetry_functions(ctx, derivatives ? 1 : 0));
    std::size_t ld = 1;
    for (auto v : numSF) ld = std::max(ld, v);
    std::size_t const n = structure.atoms.size();
    std::vector<double> G(n * ld);
    if (n) check(hdnnp_get_G(ctx, G.data(), (int)ld));
    for (std::size_t i = 0; i < n; ++i)
    {
        Atom& a = structure.atoms[i];
        std::size_t const ns = numSF.at(a.element);
        a.numSymmetryFunctions = ns;
        a.G.assign(G.begin() + i * ld, G.begin() + i * ld + ns);
        a.hasSymmetryFunctions = true;
        a.hasSymmetryFunctionDerivatives = derivatives;
    }
    structure.hasSymmetryFunctions = true;
    structure.hasSymmetryFunctionDerivatives = derivatives;
    if (checkExtrapolationWarnings) processExtrapolationWarnings(structure);
}

std::size_t Mode::processExtrapolationWarnings(Structure const& structure)
{
    int64_t const total = hdnnp_extrapolation_warning_list(
        ctx, 0, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (total < 0) check((int)total);
    if (total == 0) return 0;
    std::size_t const m = (std::size_t)total;
    std::vector<int32_t> atom(m), sf(m);
    std::vector<double> value(m), gmin(m), gmax(m);
    int64_t const again = hdnnp_extrapolation_warning_list(
        ctx, total, atom.data(), sf.data(), value.data(), gmin.data(),
        gmax.data());
    if (again < 0) check((int)again);
    // Element.cpp:489-538, record by record in the reference's order.
    for (std::size_t w = 0; w < m; ++w)
    {
        Atom const& a = structure.atoms.at((std::size_t)atom[w]);
        std::size_t const e = a.element;
        std::size_t const k = (std::size_t)sf[w];
        std::size_t const type = sfType.at(e).at(k);
        std::string const symbol = elementMap.symbol(e);
        if (collectExtrapolationWarnings)
        {
            EWContainer& d = ewData.at(e)[k];
            d.countEW++;
            d.Gmin = gmin[w];
            d.Gmax = gmax[w];
            d.type = type;
            d.element = symbol;
            d.valueEW.push_back(value[w]);
            d.indexStructureEW.push_back(a.indexStructure);
            d.indexAtomEW.push_back((std::size_t)a.tag);
        }
        std::string const line = strpr(
            "### NNP EXTRAPOLATION WARNING ### STRUCTURE: %6zu ATOM: %9" PRId64
            " ELEMENT: %2s SYMFUNC: %4zu TYPE: %2zu VALUE: %10.3E MIN: "
            "%10.3E MAX: %10.3E\n", a.indexStructure, a.tag, symbol.c_str(),
            k + 1, type, value[w], gmin[w], gmax[w]);
        if (writeExtrapolationWarnings) std::cerr << line;
        if (stopOnExtrapolationWarnings)
            throw std::out_of_range(line + "ERROR: Symmetry function value "
                                           "out of range.\n");
    }
    return m;
}

std::vector<std::string> Mode::extrapolationWarningLines() const
{
    std::vector<std::string> vs;
    for (auto const& el : ewData)
        for (auto const& kv : el)
        {
            EWContainer const& d = kv.second;
            for (std::size_t i = 0; i < d.valueEW.size(); ++i)
                vs.push_back(strpr(
                    "### NNP EXTRAPOLATION WARNING ### STRUCTURE: %6zu ATOM: "
                    "%9zu ELEMENT: %2s SYMFUNC: %4zu TYPE: %2zu VALUE: %10.3E "
                    "MIN: %10.3E MAX: %10.3E\n", d.indexStructureEW[i],
                    d.indexAtomEW[i], d.element.c_str(), kv.first + 1, d.type,
                    d.valueEW[i], d.Gmin, d.Gmax));
        }
    return vs;
}

// Mode.cpp:1642-1663: atomic energies, and dE/dG with derivatives.
void Mode::calculateAtomicNeuralNetworks(Structure& structure,
                                         bool const derivatives,
                                         std::string id)
{
    (void)id;
    ensureEngine();
    if (!structure.hasSymmetryFunctions)
        throw std::runtime_error("ERROR: Symmetry functions not calculated.\n");
    check(hdnnp_atomic_networks(ctx, derivatives ? 1 : 0));
    std::size_t const n = structure.atoms.size();
    std::vector<double> e(n);
    if (n) check(hdnnp_get_atomic_energies(ctx, e.data()));
    std::vector<double> d;
    std::size_t ld = 1;
    for (auto v : numSF) ld = std::max(ld, v);
    if (derivatives && n)
    {
        d.resize(n * ld);
        check(hdnnp_get_dEdG(ctx, d.data(), (int)ld));
    }
    for (std::size_t i = 0; i < n; ++i)
    {
        Atom& a = structure.atoms[i];
        a.energy = e[i];
        if (derivatives)
            a.dEdG.assign(d.begin() + i * ld,
                          d.begin() + i * ld + a.numSymmetryFunctions);
    }
}

// Mode.cpp:1803-1828.
void Mode::calculateEnergy(Structure& structure) const
{
    ensureEngine();
    double e = 0.0;
    check(hdnnp_energy(ctx, &e));
    structure.energy = e;
}

// Mode.cpp:1849-1913: forces of the reference's gather (an atom's own
// periodic self-image slots skipped, Mode.cpp:1885-1887).
void Mode::calculateForces(Structure& structure) const
{
    ensureEngine();
    check(hdnnp_forces(ctx, HDNNP_FORCES_REFERENCE));
    std::size_t const n = structure.atoms.size();
    std::vector<double> f(3 * n);
    if (n) check(hdnnp_get_forces(ctx, f.data()));
    for (std::size_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) structure.atoms[i].f[k] = f[3 * i + k];
}

// Mode.cpp:1967-2016 for 2G-HDNNPs.
void Mode::evaluateNNP(Structure& structure, bool useForces, bool useDEdG)
{
    useDEdG = useForces || useDEdG;
    structure.calculateNeighborList(maxCutoffRadius, false);
    calculateSymmetryFunctionGroups(structure, useForces);
    calculateAtomicNeuralNetworks(structure, useDEdG);
    calculateEnergy(structure);
    if (useForces) calculateForces(structure);
}

void Mode::evaluateNNPBatch(std::vector<Structure>& structures,
                            bool useForces)
{
    ensureEngine();
    upValid = false;
    std::size_t const ns = structures.size();
    if (ns == 0) return;
    std::vector<int64_t> off(ns + 1, 0);
    std::vector<double> r, box(9 * ns, 0.0);
    std::vector<int32_t> el, per(ns, 0);
    for (std::size_t s = 0; s < ns; ++s)
    {
        Structure const& st = structures[s];
        for (Atom const& a : st.atoms)
        {
            for (int k = 0; k < 3; ++k) r.push_back(a.r[k]);
            el.push_back((int32_t)a.element);
        }
        off[s + 1] = off[s] + (int64_t)st.atoms.size();
        per[s] = st.isPeriodic ? 1 : 0;
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) box[9 * s + 3 * k + c] = st.box[k][c];
    }
    check(hdnnp_set_structures(ctx, (int64_t)ns, off.data(), r.data(),
                               el.data(), box.data(), per.data()));
    double etot = 0.0;
    check(hdnnp_evaluate(ctx, useForces ? HDNNP_FORCES_REFERENCE : -1,
                         &etot));
    std::size_t const n = (std::size_t)off[ns];
    std::size_t ld = 1;
    for (auto v : numSF) ld = std::max(ld, v);
    std::vector<double> es(ns), G(n * ld), e(n), d, f;
    std::vector<int32_t> pbc(3 * ns);
    check(hdnnp_get_structure_energies(ctx, es.data()));
    check(hdnnp_get_structure_pbc(ctx, pbc.data()));
    check(hdnnp_get_G(ctx, G.data(), (int)ld));
    check(hdnnp_get_atomic_energies(ctx, e.data()));
    if (useForces)
    {
        d.resize(n * ld);
        f.resize(3 * n);
        check(hdnnp_get_dEdG(ctx, d.data(), (int)ld));
        check(hdnnp_get_forces(ctx, f.data()));
    }
    for (std::size_t s = 0; s < ns; ++s)
    {
        Structure& st = structures[s];
        st.energy = es[s];
        for (int k = 0; k < 3; ++k) st.pbc[k] = pbc[3 * s + k];
        for (std::size_t q = 0; q < st.atoms.size(); ++q)
        {
            std::size_t const i = (std::size_t)off[s] + q;
            Atom& a = st.atoms[q];
            std::size_t const nsf = numSF.at(a.element);
            a.numSymmetryFunctions = nsf;
            a.G.assign(G.begin() + i * ld, G.begin() + i * ld + nsf);
            a.energy = e[i];
            a.hasSymmetryFunctions = true;
            a.hasSymmetryFunctionDerivatives = useForces;
            if (useForces)
            {
                a.dEdG.assign(d.begin() + i * ld, d.begin() + i * ld + nsf);
                for (int k = 0; k < 3; ++k) a.f[k] = f[3 * i + k];
            }
        }
        st.hasSymmetryFunctions = true;
        st.hasSymmetryFunctionDerivatives = useForces;
    }
}

void Mode::fetchNeighborList(Structure& structure) const
{
    ensureEngine();
    std::size_t const n = structure.atoms.size();
    int64_t const m = hdnnp_num_neighbor_entries(ctx);
    std::vector<int64_t> off(n + 1);
    std::vector<int32_t> idx((std::size_t)m);
    std::vector<double> d((std::size_t)m), dr(3 * (std::size_t)m);
    check(hdnnp_get_neighbors(ctx, off.data(), idx.data(), nullptr, d.data(),
                              dr.data()));
    for (std::size_t i = 0; i < n; ++i)
    {
        Atom& a = structure.atoms[i];
        a.neighbors.clear();
        a.neighborsUnique.assign(1, i);
        a.numNeighborsPerElement.assign(numElements, 0);
        for (int64_t q = off[i]; q < off[i + 1]; ++q)
        {
            Atom::Neighbor nb;
            nb.index = (std::size_t)idx[q];
            nb.tag = (int64_t)nb.index;
            nb.element = structure.atoms[nb.index].element;
            nb.d = d[q];
            nb.dr = Vec3D(dr[3 * q], dr[3 * q + 1], dr[3 * q + 2]);
            a.neighbors.push_back(nb);
            ++a.numNeighborsPerElement[nb.element];
            if (a.neighborsUnique.back() != nb.index
                && std::find(a.neighborsUnique.begin(),
                             a.neighborsUnique.end(), nb.index)
                       == a.neighborsUnique.end())
                a.neighborsUnique.push_back(nb.index);
        }
        a.numNeighbors = a.neighbors.size();
        a.numNeighborsUnique = a.neighborsUnique.size();
    }
}

// Mode.cpp:2018-2067.
void Mode::addEnergyOffset(Structure& structure, bool ref)
{
    double const o = getEnergyOffset(structure);
    if (ref) structure.energyRef += o;
    else structure.energy += o;
}

void Mode::removeEnergyOffset(Structure& structure, bool ref)
{
    double const o = getEnergyOffset(structure);
    if (ref) structure.energyRef -= o;
    else structure.energy -= o;
}

double Mode::getEnergyOffset(Structure const& structure) const
{
    double result = 0.0;
    for (std::size_t e = 0; e < numElements; ++e)
        result += structure.numAtomsPerElement.at(e) * energyOffset.at(e);
    return result;
}

double Mode::getEnergyWithOffset(Structure const& structure, bool ref) const
{
    double const e = ref ? structure.energyRef : structure.energy;
    return e + getEnergyOffset(structure);
}

double Mode::normalized(std::string const& property, double value) const
{
    if (property == "energy") return value * convEnergy;
    if (property == "force") return value * convEnergy / convLength;
    if (property == "charge") return value * convCharge;
    throw std::runtime_error("ERROR: Unknown property to convert to "
                             "normalized units.\n");
}

double Mode::normalizedEnergy(Structure const& structure, bool ref) const
{
    double const e = ref ? structure.energyRef : structure.energy;
    return (e - structure.numAtoms * meanEnergy) * convEnergy;
}

double Mode::physical(std::string const& property, double value) const
{
    if (property == "energy") return value / convEnergy;
    if (property == "force") return value * convLength / convEnergy;
    if (property == "charge") return value / convCharge;
    throw std::runtime_error("ERROR: Unknown property to convert to "
                             "physical units.\n");
}

double Mode::physicalEnergy(Structure const& structure, bool ref) const
{
    double const e = ref ? structure.energyRef : structure.energy;
    return e / convEnergy + structure.numAtoms * meanEnergy;
}

void Mode::convertToNormalizedUnits(Structure& structure) const
{
    structure.toNormalizedUnits(meanEnergy, convEnergy, convLength,
                                convCharge);
}

void Mode::convertToPhysicalUnits(Structure& structure) const
{
    structure.toPhysicalUnits(meanEnergy, convEnergy, convLength, convCharge);
}

// 2G-HDNNPs have no Ewald sum; the reference logs cutoffs for 4G only.
void Mode::logEwaldCutoffs() {}

// Mode.cpp:2155-2178: the collected warnings of all elements; a reset
// empties every SF's record and keeps the SF in the map
// (SymFncStatistics::Container::resetExtrapolationWarnings).
std::size_t Mode::getNumExtrapolationWarnings() const
{
    std::size_t n = 0;
    for (auto const& el : ewData)
        for (auto const& kv : el) n += kv.second.countEW;
    return n;
}

void Mode::resetExtrapolationWarnings()
{
    for (auto& el : ewData)
        for (auto& kv : el) kv.second = EWContainer();
}

// ---------------------------------------------------------- Prediction --

Prediction::Prediction()
    : fileNameSettings("input.nn"),
      fileNameScaling("scaling.data"),
      formatWeightsFilesShort("weights.%03zu.data"),
      formatWeightsFilesCharge("weightse.%03zu.data")
{
}

void Prediction::setup()
{
    initialize();
    loadSettingsFile(fileNameSettings);
    setupGeneric();
    setupSymmetryFunctionScaling(fileNameScaling);
    std::map<std::string, std::string> formats{
        {"short", formatWeightsFilesShort},
        {"elec", formatWeightsFilesCharge}};
    setupNeuralNetworkWeights(formats);
    setupSymmetryFunctionStatistics(false, false, true, false);
}

void Prediction::readStructureFromFile(std::string const& fileName)
{
    structure.reset();
    structure.setElementMap(elementMap);
    structure.readFromFile(fileName);
    removeEnergyOffset(structure);
    if (normalize)
        structure.toNormalizedUnits(meanEnergy, convEnergy, convLength,
                                    convCharge);
}

void Prediction::predict()
{
    evaluateNNP(structure);
    if (normalize)
        structure.toPhysicalUnits(meanEnergy, convEnergy, convLength,
                                  convCharge);
    addEnergyOffset(structure, false);
    addEnergyOffset(structure, true);
}

}  // namespace nnp
