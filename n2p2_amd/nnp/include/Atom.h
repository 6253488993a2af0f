// nnp::Atom (Atom.h of the reference): per-atom inputs and the results the
// engine writes back (G, dE/dG, energy, force).  Per-neighbour dG/dr is not
// stored: the engine recomputes it on the device inside the force kernel.
#ifndef NNP_ATOM_H
#define NNP_ATOM_H

#include "Vec3D.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace nnp
{

struct Atom
{
    // One neighbour-list entry (Atom::Neighbor).  Filled only by
    // Mode::fetchNeighborList; the engine keeps its list on the device.
    struct Neighbor
    {
        std::size_t index = 0;
        int64_t tag = 0;
        std::size_t element = 0;
        double d = 0.0;
        Vec3D dr;
    };

    bool hasNeighborList = false;
    bool NeighborListIsSorted = false;
    bool hasSymmetryFunctions = false;
    bool hasSymmetryFunctionDerivatives = false;
    bool useChargeNeuron = false;
    std::size_t index = 0;
    std::size_t indexStructure = 0;
    int64_t tag = 0;
    std::size_t element = 0;
    std::size_t numNeighbors = 0;
    std::size_t numNeighborsUnique = 0;
    std::size_t numSymmetryFunctions = 0;
    double energy = 0.0;
    double charge = 0.0;
    double chargeRef = 0.0;
    Vec3D r;
    Vec3D f;
    Vec3D fRef;
    std::vector<std::size_t> neighborsUnique;
    std::vector<std::size_t> numNeighborsPerElement;
    std::vector<double> G;
    std::vector<double> dEdG;
    std::vector<Neighbor> neighbors;

    // Atom::toNormalizedUnits / toPhysicalUnits (Atom.cpp) for the fields
    // held here.
    void toNormalizedUnits(double convEnergy, double convLength,
                           double convCharge);
    void toPhysicalUnits(double convEnergy, double convLength,
                         double convCharge);
    void free(bool all);
    std::vector<std::string> info() const;
};

}  // namespace nnp

#endif
