// nnp::Structure (Structure.h of the reference): a configuration read from
// input.data, its cell, and the energies and forces written back by Mode.
// The neighbour list is built by the engine on the device when Mode
// evaluates the structure; calculateNeighborList records the cutoff and
// the periodic image counts.
#ifndef NNP_STRUCTURE_H
#define NNP_STRUCTURE_H

#include "Atom.h"
#include "ElementMap.h"
#include "Vec3D.h"

#include <cstddef>
#include <fstream>
#include <string>
#include <vector>

namespace nnp
{

class Structure
{
public:
    enum SampleType
    {
        ST_UNKNOWN,
        ST_TRAINING,
        ST_VALIDATION,
        ST_TEST
    };

    ElementMap elementMap;
    bool isPeriodic = false;
    bool isTriclinic = false;
    bool hasNeighborList = false;
    bool NeighborListIsSorted = false;
    bool hasSymmetryFunctions = false;
    bool hasSymmetryFunctionDerivatives = false;
    std::size_t index = 0;
    std::size_t numAtoms = 0;
    std::size_t numElements = 0;
    std::size_t numElementsPresent = 0;
    int pbc[3] = {0, 0, 0};
    double energy = 0.0;
    double energyRef = 0.0;
    double energyShort = 0.0;
    double energyElec = 0.0;
    double charge = 0.0;
    double chargeRef = 0.0;
    double volume = 0.0;
    // Cutoff of the last calculateNeighborList call.
    double neighborCutoff = 0.0;
    SampleType sampleType = ST_UNKNOWN;
    std::string comment;
    Vec3D box[3];
    Vec3D invbox[3];
    std::vector<std::size_t> numAtomsPerElement;
    std::vector<Atom> atoms;

    Structure() = default;
    void setElementMap(ElementMap const& elementMap);
    void addAtom(Atom const& atom, std::string const& element);
    void readFromFile(std::string const fileName = "input.data");
    void readFromFile(std::ifstream& file);
    void readFromLines(std::vector<std::string> const& lines);
    // Structure::calculateNeighborList (Structure.cpp:287-399): records the
    // cutoff and the image counts; the list itself is built on the device
    // by the next Mode evaluation.
    void calculateNeighborList(double cutoffRadius,
                               bool sortByDistance = false);
    void calculatePbcCopies(double cutoffRadius, int (&pbc)[3]) const;
    void calculateInverseBox();
    void calculateVolume();
    void remap();
    void remap(Atom& atom);
    void toNormalizedUnits(double meanEnergy, double convEnergy,
                           double convLength, double convCharge);
    void toPhysicalUnits(double meanEnergy, double convEnergy,
                         double convLength, double convCharge);
    void freeAtoms(bool all);
    void reset();
    void clearNeighborList();
    void writeToFile(std::string const fileName = "output.data",
                     bool const ref = true, bool const append = false) const;
    void writeToFile(std::ofstream* const& file, bool const ref = true) const;
    std::vector<std::string> info() const;
};

}  // namespace nnp

#endif
