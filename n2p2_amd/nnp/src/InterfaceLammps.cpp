// InterfaceLammps.h of the nnp layer: the pair_style hdnnp object of the
// reference (libnnpif/LAMMPS/InterfaceLammps.cpp) over the engine's
// caller-given neighbour-list path.
#include "InterfaceLammps.h"
#include "utility.h"

#include "hdnnp.h"

#include <algorithm>
#include <cstring>
#include <cmath>
#include <fstream>
#include <limits>
#include <stdexcept>

namespace nnp
{

namespace
{

void check(int rc)
{
    if (rc != HDNNP_OK) throw std::runtime_error(hdnnp_last_error());
}

// Relative difference below which the LAMMPS and SF cutoffs count as equal.
double const kTolCutoff = 1.0e-2;

std::size_t const kNone = std::numeric_limits<std::size_t>::max();

}  // namespace

InterfaceLammps::InterfaceLammps() {}

void InterfaceLammps::initialize(char const* const& directory,
                                 char const* const& emap, bool showew,
                                 bool resetew, int showewsum, int maxew,
                                 double cflength, double cfenergy,
                                 double lammpsCutoff, int lammpsNtypes,
                                 int myRank)
{
    this->emap = emap;
    this->showew = showew;
    this->resetew = resetew;
    this->showewsum = showewsum;
    this->maxew = maxew;
    this->cflength = cflength;
    this->cfenergy = cfenergy;
    this->myRank = myRank;
    log.writeToStdout = false;
    std::string dir(directory);
    if (dir.empty() || dir.back() != '/') dir += '/';
    Mode::initialize();
    loadSettingsFile(dir + "input.nn");
    setupGeneric(dir);
    setupSymmetryFunctionScaling(dir + "scaling.data");
    bool const collectEW = showew || showewsum > 0 || maxew >= 0;
    setupSymmetryFunctionStatistics(false, collectEW, false, false);
    setupNeuralNetworkWeights(dir);

    double const sfCutoff = getMaxCutoffRadius();
    if (lammpsCutoff < sfCutoff)
        throw std::runtime_error("ERROR: LAMMPS cutoff via pair_coeff keyword "
                                 "is smaller than maximum symmetry function "
                                 "cutoff.\n");
    if (std::fabs(sfCutoff - lammpsCutoff) / lammpsCutoff > kTolCutoff)
        log << "WARNING: Maximum symmetry function cutoff differs from the "
               "LAMMPS cutoff by more than 1 %.\n";

    mapTypeToElement.clear();
    mapElementToType.clear();
    if (this->emap.empty())
    {
        if (elementMap.size() != (std::size_t)lammpsNtypes)
            throw std::runtime_error(strpr(
                "ERROR: No element mapping given and number of LAMMPS atom "
                "types (%d) and NNP elements (%zu) does not match.\n",
                lammpsNtypes, elementMap.size()));
        for (int i = 0; i < lammpsNtypes; ++i)
        {
            mapTypeToElement[i + 1] = (std::size_t)i;
            mapElementToType[(std::size_t)i] = i + 1;
        }
    }
    else
    {
        std::vector<std::string> items = split(reduce(trim(this->emap),
                                                      " \t", ""), ',');
        if (elementMap.size() < items.size())
            throw std::runtime_error(strpr(
                "ERROR: Element mapping is inconsistent, NNP elements: %zu, "
                "emap elements: %zu.\n", elementMap.size(), items.size()));
        for (std::string const& it : items)
        {
            std::vector<std::string> ts = split(it, ':');
            if (ts.size() != 2)
                throw std::runtime_error(strpr("ERROR: Invalid element "
                                               "mapping string: \"%s\".\n",
                                               it.c_str()));
            int const t = std::stoi(ts[0]);
            if (t < 1 || t > lammpsNtypes)
                throw std::runtime_error(strpr("ERROR: LAMMPS type %d is not "
                                               "present.\n", t));
            std::size_t const e = elementMap[ts[1]];
            mapTypeToElement[t] = e;
            mapElementToType[e] = t;
        }
    }
    ignoreType.assign((std::size_t)lammpsNtypes + 1, true);
    for (auto const& te : mapTypeToElement) ignoreType[te.first] = false;
    structure.setElementMap(elementMap);
    initialized = true;
}

void InterfaceLammps::setLocalAtoms(int numAtomsLocal,
                                    int const* const atomType)
{
    structure.reset();
    structure.setElementMap(elementMap);
    structure.index = (std::size_t)myRank;
    indexMap.assign((std::size_t)numAtomsLocal, kNone);
    structure.atoms.reserve((std::size_t)numAtomsLocal);
    for (int i = 0; i < numAtomsLocal; ++i)
    {
        int const t = atomType[i];
        if (t < 0 || (std::size_t)t >= ignoreType.size() || ignoreType[t])
            continue;
        indexMap[i] = structure.numAtoms;
        Atom a;
        a.index = (std::size_t)i;
        a.indexStructure = (std::size_t)myRank;
        a.element = mapTypeToElement.at(t);
        a.numNeighborsPerElement.assign(numElements, 0);
        structure.atoms.push_back(a);
        ++structure.numAtoms;
        ++structure.numAtomsPerElement[a.element];
    }
}

void InterfaceLammps::setLocalAtomPositions(double const* const* const atomPos)
{
    for (auto& a : structure.atoms)
        for (int k = 0; k < 3; ++k) a.r[k] = atomPos[a.index][k] * cflength;
}

void InterfaceLammps::setLocalTags(int const* const atomTag)
{
    for (auto& a : structure.atoms) a.tag = (int64_t)atomTag[a.index];
}

void InterfaceLammps::setLocalTags(int64_t const* const atomTag)
{
    for (auto& a : structure.atoms) a.tag = atomTag[a.index];
}

void InterfaceLammps::setBoxVectors(double const* boxlo, double const* boxhi,
                                    double const xy, double const xz,
                                    double const yz)
{
    // The engine works on the neighbour lists; the cell is recorded only.
    structure.isPeriodic = true;
    structure.box[0] = Vec3D(boxhi[0] - boxlo[0], 0.0, 0.0) * cflength;
    structure.box[1] = Vec3D(xy, boxhi[1] - boxlo[1], 0.0) * cflength;
    structure.box[2] = Vec3D(xz, yz, boxhi[2] - boxlo[2]) * cflength;
}

void InterfaceLammps::allocateNeighborlists(int const* const numneigh)
{
    for (auto& a : structure.atoms)
        a.neighbors.reserve((std::size_t)numneigh[a.index]);
}

// InterfaceLammps::addNeighbor (InterfaceLammps.cpp:267-300): LAMMPS units
// to the model's (cflength, then the normalisation length).
void InterfaceLammps::addNeighbor(int i, int j, int64_t tag, int type,
                                  double dx, double dy, double dz, double d2)
{
    if (type < 0 || (std::size_t)type >= ignoreType.size() || ignoreType[type]
        || (std::size_t)i >= indexMap.size() || indexMap[i] == kNone)
        return;
    Atom& a = structure.atoms[indexMap[i]];
    Atom::Neighbor n;
    n.index = (std::size_t)j;
    n.tag = tag;
    n.element = mapTypeToElement.at(type);
    n.dr = Vec3D(dx, dy, dz) * cflength;
    n.d = std::sqrt(d2) * cflength;
    if (normalize)
    {
        n.dr *= convLength;
        n.d *= convLength;
    }
    a.neighbors.push_back(n);
    ++a.numNeighbors;
    ++a.numNeighborsPerElement.at(n.element);
}

// InterfaceLammps::process (InterfaceLammps.cpp:302-327) for 2G-HDNNPs.
void InterfaceLammps::process()
{
    ensureEngine();
    hdnnp_ctx* c = engineContext();
    std::size_t const n = structure.atoms.size();
    std::vector<int32_t> cIndex(n), cElem(n), nIndex, nElem;
    std::vector<int64_t> off(n + 1, 0);
    std::vector<double> dr;
    for (std::size_t i = 0; i < n; ++i)
    {
        Atom const& a = structure.atoms[i];
        cIndex[i] = (int32_t)a.index;
        cElem[i] = (int32_t)a.element;
        for (auto const& nb : a.neighbors)
        {
            nIndex.push_back((int32_t)nb.index);
            nElem.push_back((int32_t)nb.element);
            dr.push_back(nb.dr[0]);
            dr.push_back(nb.dr[1]);
            dr.push_back(nb.dr[2]);
        }
        off[i + 1] = (int64_t)nIndex.size();
    }
    structure.energy = 0.0;
    if (n == 0) return;
    check(hdnnp_set_neighbor_list(c, (int64_t)n, cIndex.data(), cElem.data(),
                                  off.data(), nIndex.data(), nElem.data(),
                                  dr.data()));
    double e = 0.0;
    check(hdnnp_evaluate(c, HDNNP_FORCES_CONSISTENT, &e));
    std::vector<double> ea(n);
    check(hdnnp_get_atomic_energies(c, ea.data()));
    for (std::size_t i = 0; i < n; ++i) structure.atoms[i].energy = ea[i];
    if (checkExtrapolationWarnings) processExtrapolationWarnings(structure);
    structure.energy = e;
    if (normalize) structure.energy = physicalEnergy(structure, false);
    addEnergyOffset(structure, false);
}

double InterfaceLammps::getMaxCutoffRadius() const
{
    if (normalize) return maxCutoffRadius / convLength / cflength;
    return maxCutoffRadius / cflength;
}

double InterfaceLammps::getEnergy() const { return structure.energy / cfenergy; }

double InterfaceLammps::getAtomicEnergy(int index) const
{
    Atom const& a = structure.atoms.at((std::size_t)index);
    double const offset = energyOffset.at(a.element);
    if (normalize)
        return (physical("energy", a.energy) + meanEnergy + offset) / cfenergy;
    return (a.energy + offset) / cfenergy;
}

// InterfaceLammps::getForces (InterfaceLammps.cpp:546-581): self forces of
// the local atoms and pair forces onto their neighbours (local or ghost),
// in LAMMPS units.
void InterfaceLammps::getForces(double* const* const& atomF) const
{
    std::size_t const n = structure.atoms.size();
    if (n == 0) return;
    double const cfforce = cflength / cfenergy;
    double const convForce = normalize ? convLength / convEnergy : 1.0;
    int64_t nout = 0;
    for (auto const& a : structure.atoms)
    {
        nout = std::max(nout, (int64_t)a.index + 1);
        for (auto const& nb : a.neighbors)
            nout = std::max(nout, (int64_t)nb.index + 1);
    }
    std::vector<double> f(3 * (std::size_t)nout, 0.0);
    check(hdnnp_get_scattered_forces(engineContext(), nout, f.data()));
    double const s = cfforce * convForce;
    for (int64_t k = 0; k < nout; ++k)
        for (int c = 0; c < 3; ++c) atomF[k][c] += f[3 * k + c] * s;
}

// 2G-HDNNPs carry no charges.
void InterfaceLammps::getCharges(double* const& atomQ) const
{
    for (auto const& a : structure.atoms) atomQ[a.index] = 0.0;
}

// The extrapolation-warning buffers pair_style hdnnp gathers on rank 0
// (InterfaceLammps.cpp:638-749): per element the number of SF records, then
// per record the SF index, countEW, type, Gmin, Gmax, the element symbol
// (length with the terminating zero, then the characters) and the
// structure indices, atom tags and values of its warnings.  The fields are
// packed in this order in their native representation (what MPI_Pack
// writes for MPI_SIZE_T, MPI_DOUBLE and MPI_CHAR on one host type).
namespace
{
template <class T>
void packBytes(char* buf, long bs, long& p, T const* v, std::size_t n)
{
    long const b = (long)(n * sizeof(T));
    if (p + b > bs)
        throw std::runtime_error("ERROR: Extrapolation-warning buffer too "
                                 "small.\n");
    if (b) std::memcpy(buf + p, v, (std::size_t)b);
    p += b;
}

template <class T>
void unpackBytes(char const* buf, long bs, long& p, T* v, std::size_t n)
{
    long const b = (long)(n * sizeof(T));
    if (p + b > bs)
        throw std::runtime_error("ERROR: Extrapolation-warning buffer too "
                                 "short.\n");
    if (b) std::memcpy(v, buf + p, (std::size_t)b);
    p += b;
}
}  // namespace

long InterfaceLammps::getEWBufferSize() const
{
    long const ss = sizeof(std::size_t), ds = sizeof(double), cs = 1;
    long bs = 0;
    for (auto const& m : ewData)
    {
        bs += ss;
        for (auto const& kv : m)
        {
            EWContainer const& d = kv.second;
            bs += 3 * ss + 2 * ds + ss;
            bs += (long)(d.element.length() + 1) * cs;
            bs += (long)d.countEW * (2 * ss + ds);
        }
    }
    return bs;
}

void InterfaceLammps::fillEWBuffer(char* const& buf, int bs) const
{
    long p = 0;
    for (auto const& m : ewData)
    {
        std::size_t const n = m.size();
        packBytes(buf, bs, p, &n, 1);
        for (auto const& kv : m)
        {
            EWContainer const& d = kv.second;
            packBytes(buf, bs, p, &kv.first, 1);
            packBytes(buf, bs, p, &d.countEW, 1);
            packBytes(buf, bs, p, &d.type, 1);
            packBytes(buf, bs, p, &d.Gmin, 1);
            packBytes(buf, bs, p, &d.Gmax, 1);
            std::size_t const ts = d.element.length() + 1;
            packBytes(buf, bs, p, &ts, 1);
            packBytes(buf, bs, p, d.element.c_str(), ts);
            packBytes(buf, bs, p, d.indexStructureEW.data(), d.countEW);
            packBytes(buf, bs, p, d.indexAtomEW.data(), d.countEW);
            packBytes(buf, bs, p, d.valueEW.data(), d.countEW);
        }
    }
}

void InterfaceLammps::extractEWBuffer(char const* const& buf, int bs)
{
    long p = 0;
    for (auto& m : ewData)
    {
        std::size_t n = 0;
        unpackBytes(buf, bs, p, &n, 1);
        for (std::size_t i = 0; i < n; ++i)
        {
            std::size_t index = 0, countEW = 0, ts = 0;
            unpackBytes(buf, bs, p, &index, 1);
            EWContainer& d = m[index];
            unpackBytes(buf, bs, p, &countEW, 1);
            unpackBytes(buf, bs, p, &d.type, 1);
            unpackBytes(buf, bs, p, &d.Gmin, 1);
            unpackBytes(buf, bs, p, &d.Gmax, 1);
            unpackBytes(buf, bs, p, &ts, 1);
            std::vector<char> el(ts + 1, '\0');
            unpackBytes(buf, bs, p, el.data(), ts);
            d.element = el.data();
            d.indexStructureEW.resize(d.countEW + countEW);
            unpackBytes(buf, bs, p, d.indexStructureEW.data() + d.countEW,
                        countEW);
            d.indexAtomEW.resize(d.countEW + countEW);
            unpackBytes(buf, bs, p, d.indexAtomEW.data() + d.countEW,
                        countEW);
            d.valueEW.resize(d.countEW + countEW);
            unpackBytes(buf, bs, p, d.valueEW.data() + d.countEW, countEW);
            d.countEW += countEW;
        }
    }
}

// InterfaceLammps.cpp:751-765: every collected warning, one line each.
void InterfaceLammps::writeExtrapolationWarnings()
{
    for (std::string const& line : extrapolationWarningLines()) log << line;
}

// InterfaceLammps.cpp:767-776 (SymFncStatistics::clear).
void InterfaceLammps::clearExtrapolationWarnings()
{
    for (auto& m : ewData) m.clear();
}

void InterfaceLammps::writeToFile(std::string const fileName,
                                  bool const append)
{
    structure.writeToFile(fileName, false, append);
}

}  // namespace nnp
