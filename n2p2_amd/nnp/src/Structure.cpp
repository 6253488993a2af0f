// Structure.h of the nnp layer: input.data parsing, cell geometry, unit
// conversion and output, following Structure.cpp of the reference
// (readFromLines :142-271, calculateInverseBox :519-545, remap
// :1214-1247, toNormalizedUnits / toPhysicalUnits :1249-1305,
// writeToFile :1449-1531).  The remap is evaluated in the reference's
// operation order (this file is built without floating-point contraction)
// so that the positions the engine receives match the reference bit for
// bit.
#include "Structure.h"
#include "utility.h"

#include <cmath>
#include <cstdlib>
#include <limits>
#include <stdexcept>

namespace nnp
{

void Structure::setElementMap(ElementMap const& map)
{
    elementMap = map;
    numElements = elementMap.size();
    numAtomsPerElement.resize(numElements, 0);
}

void Structure::addAtom(Atom const& atom, std::string const& element)
{
    atoms.push_back(atom);
    Atom& a = atoms.back();
    a.numNeighborsPerElement.assign(elementMap.size(), 0);
    a.neighbors.clear();
    a.neighborsUnique.clear();
    a.numNeighbors = 0;
    a.hasNeighborList = false;
    a.index = numAtoms;
    a.indexStructure = index;
    a.element = elementMap[element];
    a.numSymmetryFunctions = 0;
    ++numAtoms;
    ++numAtomsPerElement[a.element];
}

void Structure::readFromFile(std::string const fileName)
{
    std::ifstream file(fileName.c_str());
    if (!file.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + fileName
                                 + "\".\n");
    readFromFile(file);
}

void Structure::readFromFile(std::ifstream& file)
{
    std::vector<std::string> lines;
    std::string line;
    bool begun = false;
    while (std::getline(file, line))
    {
        std::vector<std::string> w = split(reduce(line));
        if (w.empty() || w[0].empty()) continue;
        if (w[0] == "begin") begun = true;
        if (!begun) continue;
        lines.push_back(line);
        if (w[0] == "end") break;
    }
    if (lines.empty())
        throw std::runtime_error("ERROR: No structure found in file.\n");
    readFromLines(lines);
}

void Structure::readFromLines(std::vector<std::string> const& lines)
{
    std::vector<std::string> w = split(reduce(lines.at(0)));
    if (w.empty() || w[0] != "begin")
        throw std::runtime_error("ERROR: Unexpected line content, expected "
                                 "\"begin\" keyword.\n");
    std::size_t iBox = 0;
    for (std::string const& line : lines)
    {
        w = split(reduce(line));
        if (w.empty()) continue;
        std::string const& key = w[0];
        if (key == "begin")
        {
            for (std::size_t k = 1; k < w.size(); ++k)
            {
                if (w[k] == "set=train") sampleType = ST_TRAINING;
                else if (w[k] == "set=test") sampleType = ST_TEST;
                else
                    throw std::runtime_error("ERROR: Unknown keyword in "
                                             "structure specification, check "
                                             "\"begin\" arguments.\n");
            }
        }
        else if (key == "comment")
        {
            std::string c = line;
            c.erase(c.find("comment"), key.size() + 1);
            comment = c;
        }
        else if (key == "lattice")
        {
            if (iBox > 2)
                throw std::runtime_error("ERROR: Too many box vectors.\n");
            for (int k = 0; k < 3; ++k)
                box[iBox][k] = std::atof(w.at(1 + k).c_str());
            ++iBox;
            if (iBox == 3)
            {
                isPeriodic = true;
                double const tiny = std::numeric_limits<double>::min();
                isTriclinic = box[0][1] > tiny || box[0][2] > tiny
                           || box[1][0] > tiny || box[1][2] > tiny
                           || box[2][0] > tiny || box[2][1] > tiny;
                calculateInverseBox();
                calculateVolume();
            }
        }
        else if (key == "atom")
        {
            Atom a;
            a.index = numAtoms;
            a.indexStructure = index;
            a.tag = (int64_t)numAtoms;
            for (int k = 0; k < 3; ++k) a.r[k] = std::atof(w.at(1 + k).c_str());
            a.element = elementMap[w.at(4)];
            a.chargeRef = std::atof(w.at(5).c_str());
            for (int k = 0; k < 3; ++k)
                a.fRef[k] = std::atof(w.at(7 + k).c_str());
            a.numNeighborsPerElement.assign(numElements, 0);
            atoms.push_back(a);
            ++numAtoms;
            ++numAtomsPerElement.at(a.element);
        }
        else if (key == "energy")
            energyRef = std::atof(w.at(1).c_str());
        else if (key == "charge")
            chargeRef = std::atof(w.at(1).c_str());
        else if (key == "end")
        {
            if (!(iBox == 0 || iBox == 3))
                throw std::runtime_error("ERROR: Strange number of box "
                                         "vectors.\n");
            if (isPeriodic
                && std::fabs(chargeRef)
                       > 10 * std::numeric_limits<double>::epsilon())
                throw std::runtime_error(
                    "ERROR: In structure with index " + std::to_string(index)
                    + "; if PBCs are applied, the\nsimulation cell has to be "
                      "neutrally charged.\n");
            break;
        }
        else
            throw std::runtime_error("ERROR: Unexpected file content, unknown "
                                     "keyword \"" + key + "\".\n");
    }
    numElementsPresent = 0;
    for (std::size_t e = 0; e < numElements; ++e)
        if (numAtomsPerElement[e] > 0) ++numElementsPresent;
    if (isPeriodic) remap();
}

void Structure::calculatePbcCopies(double cutoffRadius, int (&p)[3]) const
{
    Vec3D const axb = box[0].cross(box[1]).normalize();
    Vec3D const axc = box[0].cross(box[2]).normalize();
    Vec3D const bxc = box[1].cross(box[2]).normalize();
    double const proj[3] = {std::fabs(box[0] * bxc), std::fabs(box[1] * axc),
                            std::fabs(box[2] * axb)};
    for (int k = 0; k < 3; ++k)
    {
        p[k] = 0;
        while (p[k] * proj[k] <= cutoffRadius) ++p[k];
    }
}

void Structure::calculateNeighborList(double cutoffRadius, bool sortByDistance)
{
    neighborCutoff = cutoffRadius;
    NeighborListIsSorted = sortByDistance;
    if (isPeriodic) calculatePbcCopies(cutoffRadius, pbc);
    else pbc[0] = pbc[1] = pbc[2] = 0;
    hasNeighborList = true;
    for (auto& a : atoms) a.hasNeighborList = true;
}

void Structure::calculateInverseBox()
{
    double invdet = box[0][0] * box[1][1] * box[2][2]
                  + box[1][0] * box[2][1] * box[0][2]
                  + box[2][0] * box[0][1] * box[1][2]
                  - box[2][0] * box[1][1] * box[0][2]
                  - box[0][0] * box[2][1] * box[1][2]
                  - box[1][0] * box[0][1] * box[2][2];
    invdet = 1.0 / invdet;
    // Cofactor rows, each scaled by 1/det.
    invbox[0] = Vec3D(box[1][1] * box[2][2] - box[2][1] * box[1][2],
                      box[2][1] * box[0][2] - box[0][1] * box[2][2],
                      box[0][1] * box[1][2] - box[1][1] * box[0][2]) * invdet;
    invbox[1] = Vec3D(box[2][0] * box[1][2] - box[1][0] * box[2][2],
                      box[0][0] * box[2][2] - box[2][0] * box[0][2],
                      box[1][0] * box[0][2] - box[0][0] * box[1][2]) * invdet;
    invbox[2] = Vec3D(box[1][0] * box[2][1] - box[2][0] * box[1][1],
                      box[2][0] * box[0][1] - box[0][0] * box[2][1],
                      box[0][0] * box[1][1] - box[1][0] * box[0][1]) * invdet;
}

void Structure::calculateVolume()
{
    volume = std::fabs(box[0] * box[1].cross(box[2]));
}

void Structure::remap()
{
    for (auto& a : atoms) remap(a);
}

void Structure::remap(Atom& atom)
{
    // Fractional coordinates, wrapped into [0, 1), back to Cartesian.
    Vec3D f = atom.r[0] * invbox[0] + atom.r[1] * invbox[1];
    f = f + atom.r[2] * invbox[2];
    for (int k = 0; k < 3; ++k)
    {
        if (f[k] > 1.0) f[k] -= (int)f[k];
        if (f[k] < 0.0) f[k] += 1.0 - (int)f[k];
        if (f[k] == 1.0) f[k] = 0.0;
    }
    Vec3D r = f[0] * box[0] + f[1] * box[1];
    atom.r = r + f[2] * box[2];
}

void Structure::toNormalizedUnits(double meanEnergy, double convEnergy,
                                  double convLength, double convCharge)
{
    if (isPeriodic)
        for (int k = 0; k < 3; ++k)
        {
            box[k] *= convLength;
            invbox[k] /= convLength;
        }
    energyRef = (energyRef - numAtoms * meanEnergy) * convEnergy;
    energy = (energy - numAtoms * meanEnergy) * convEnergy;
    chargeRef *= convCharge;
    charge *= convCharge;
    volume *= convLength * convLength * convLength;
    neighborCutoff *= convLength;
    for (auto& a : atoms) a.toNormalizedUnits(convEnergy, convLength, convCharge);
}

void Structure::toPhysicalUnits(double meanEnergy, double convEnergy,
                                double convLength, double convCharge)
{
    if (isPeriodic)
        for (int k = 0; k < 3; ++k)
        {
            box[k] /= convLength;
            invbox[k] *= convLength;
        }
    energyRef = energyRef / convEnergy + numAtoms * meanEnergy;
    energy = energy / convEnergy + numAtoms * meanEnergy;
    chargeRef /= convCharge;
    charge /= convCharge;
    volume /= convLength * convLength * convLength;
    neighborCutoff /= convLength;
    for (auto& a : atoms) a.toPhysicalUnits(convEnergy, convLength, convCharge);
}

void Structure::freeAtoms(bool all)
{
    for (auto& a : atoms) a.free(all);
    if (all) hasSymmetryFunctions = false;
    hasSymmetryFunctionDerivatives = false;
}

void Structure::reset()
{
    isPeriodic = isTriclinic = false;
    hasNeighborList = NeighborListIsSorted = false;
    hasSymmetryFunctions = hasSymmetryFunctionDerivatives = false;
    index = numAtoms = numElementsPresent = 0;
    energy = energyRef = energyShort = energyElec = 0.0;
    charge = chargeRef = volume = 0.0;
    neighborCutoff = 0.0;
    sampleType = ST_UNKNOWN;
    comment.clear();
    for (int k = 0; k < 3; ++k)
    {
        pbc[k] = 0;
        box[k] = Vec3D();
        invbox[k] = Vec3D();
    }
    numAtomsPerElement.assign(numElements, 0);
    atoms.clear();
}

void Structure::clearNeighborList()
{
    for (auto& a : atoms)
    {
        a.neighbors.clear();
        a.neighborsUnique.clear();
        a.numNeighbors = a.numNeighborsUnique = 0;
        a.numNeighborsPerElement.assign(numElements, 0);
        a.hasNeighborList = false;
        a.hasSymmetryFunctions = a.hasSymmetryFunctionDerivatives = false;
    }
    hasNeighborList = false;
    hasSymmetryFunctions = hasSymmetryFunctionDerivatives = false;
}

void Structure::writeToFile(std::string const fileName, bool const ref,
                            bool const append) const
{
    std::ofstream file;
    if (append) file.open(fileName.c_str(), std::ofstream::app);
    else file.open(fileName.c_str());
    if (!file.is_open())
        throw std::runtime_error("ERROR: Could not open file: \"" + fileName
                                 + "\".\n");
    writeToFile(&file, ref);
}

// input.data block of this structure: lattice rows if periodic, one atom
// line per atom (position, symbol, charge, 0, force; reference or
// predicted values), energy, charge.
void Structure::writeToFile(std::ofstream* const& file, bool const ref) const
{
    if (!file->is_open())
        throw std::runtime_error("ERROR: Cannot write to file, file is not "
                                 "open.\n");
    std::ofstream& o = *file;
    o << "begin\n";
    o << strpr("comment %s\n", comment.c_str());
    if (isPeriodic)
        for (int k = 0; k < 3; ++k)
            o << strpr("lattice %24.16E %24.16E %24.16E\n", box[k][0],
                       box[k][1], box[k][2]);
    for (auto const& a : atoms)
    {
        Vec3D const& fo = ref ? a.fRef : a.f;
        o << strpr("atom %24.16E %24.16E %24.16E %2s %24.16E %24.16E %24.16E "
                   "%24.16E %24.16E\n",
                   a.r[0], a.r[1], a.r[2], elementMap[a.element].c_str(),
                   ref ? a.chargeRef : a.charge, 0.0, fo[0], fo[1], fo[2]);
    }
    o << strpr("energy %24.16E\n", ref ? energyRef : energy);
    o << strpr("charge %24.16E\n", ref ? chargeRef : charge);
    o << "end\n";
}

std::vector<std::string> Structure::info() const
{
    std::vector<std::string> v;
    v.push_back(strpr("********************************\n"));
    v.push_back(strpr("STRUCTURE                       \n"));
    v.push_back(strpr("********************************\n"));
    std::vector<std::string> vm = elementMap.info();
    v.insert(v.end(), vm.begin(), vm.end());
    v.push_back(strpr("index                          : %zu\n", index));
    v.push_back(strpr("isPeriodic                     : %d\n", (int)isPeriodic));
    v.push_back(strpr("isTriclinic                    : %d\n", (int)isTriclinic));
    v.push_back(strpr("hasNeighborList                : %d\n",
                      (int)hasNeighborList));
    v.push_back(strpr("hasSymmetryFunctions           : %d\n",
                      (int)hasSymmetryFunctions));
    v.push_back(strpr("hasSymmetryFunctionDerivatives : %d\n",
                      (int)hasSymmetryFunctionDerivatives));
    v.push_back(strpr("numAtoms                       : %zu\n", numAtoms));
    v.push_back(strpr("numElements                    : %zu\n", numElements));
    v.push_back(strpr("numElementsPresent             : %zu\n",
                      numElementsPresent));
    v.push_back(strpr("pbc                            : %d %d %d\n", pbc[0],
                      pbc[1], pbc[2]));
    v.push_back(strpr("energy                         : %16.8E\n", energy));
    v.push_back(strpr("energyRef                      : %16.8E\n", energyRef));
    v.push_back(strpr("charge                         : %16.8E\n", charge));
    v.push_back(strpr("chargeRef                      : %16.8E\n", chargeRef));
    v.push_back(strpr("volume                         : %16.8E\n", volume));
    v.push_back(strpr("sampleType                     : %d\n", (int)sampleType));
    v.push_back(strpr("comment                        : %s\n", comment.c_str()));
    for (int k = 0; k < 3; ++k)
        v.push_back(strpr("box[%d]                         : %16.8E %16.8E "
                          "%16.8E\n", k, box[k][0], box[k][1], box[k][2]));
    for (int k = 0; k < 3; ++k)
        v.push_back(strpr("invbox[%d]                      : %16.8E %16.8E "
                          "%16.8E\n", k, invbox[k][0], invbox[k][1],
                          invbox[k][2]));
    v.push_back(strpr("--------------------------------\n"));
    for (std::size_t i = 0; i < atoms.size(); ++i)
    {
        v.push_back(strpr("%29zu  :\n", i));
        std::vector<std::string> va = atoms[i].info();
        v.insert(v.end(), va.begin(), va.end());
    }
    v.push_back(strpr("--------------------------------\n"));
    return v;
}

}  // namespace nnp
