// nnp::InterfaceLammps (libnnpif/LAMMPS/InterfaceLammps.h of the
// reference): the object pair_style hdnnp drives.  Local atoms and their
// neighbour lists come from LAMMPS (setLocalAtoms, addNeighbor); the energy
// and the forces on local and ghost atoms go back (getEnergy, getForces).
// Backed by the engine's caller-given neighbour-list path
// (hdnnp_set_neighbor_list, hdnnp_get_scattered_forces).
#ifndef NNP_INTERFACELAMMPS_H
#define NNP_INTERFACELAMMPS_H

#include "Mode.h"
#include "Structure.h"

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace nnp
{

class InterfaceLammps : public Mode
{
public:
    InterfaceLammps();
    // InterfaceLammps::initialize (InterfaceLammps.cpp:47-215): NNP files
    // from directory, LAMMPS type -> element map ("1:H,2:O"; empty = types
    // in element order), extrapolation-warning options, length and energy
    // conversion factors, the pair_coeff cutoff (checked against the SF
    // cutoff), the LAMMPS type count and the MPI rank.
    void initialize(char const* const& directory, char const* const& emap,
                    bool showew, bool resetew, int showewsum, int maxew,
                    double cflength, double cfenergy, double lammpsCutoff,
                    int lammpsNtypes, int myRank);
    void setGlobalStructureStatus(bool const status) { globalStructure = status; }
    bool getGlobalStructureStatus() { return globalStructure; }
    void setLocalAtoms(int numAtomsLocal, int const* const atomType);
    void setLocalAtomPositions(double const* const* const atomPos);
    void setLocalTags(int const* const atomTag);
    void setLocalTags(int64_t const* const atomTag);
    void setBoxVectors(double const* boxlo, double const* boxhi,
                       double const xy, double const xz, double const yz);
    void allocateNeighborlists(int const* const numneigh);
    // dx, dy, dz = x_i - x_j in LAMMPS units, d2 their squared length.
    void addNeighbor(int i, int j, int64_t tag, int type, double dx,
                     double dy, double dz, double d2);
    void finalizeNeighborList() {}
    // Symmetry functions, networks, energy and the force contraction on
    // the device.
    void process();
    // Energy with offsets in LAMMPS units.
    double getEnergy() const;
    double getAtomicEnergy(int index) const;
    // Adds the forces of all local and ghost atoms to atomF (LAMMPS units).
    void getForces(double* const* const& atomF) const;
    void getCharges(double* const& atomQ) const;
    bool isInitialized() const { return initialized; }
    // Maximum SF cutoff in LAMMPS units.
    double getMaxCutoffRadius() const;
    double getMaxCutoffRadiusOverall() { return getMaxCutoffRadius(); }
    long getEWBufferSize() const;
    void fillEWBuffer(char* const& buf, int bs) const;
    void extractEWBuffer(char const* const& buf, int bs);
    void writeExtrapolationWarnings();
    void clearExtrapolationWarnings();
    void writeToFile(std::string const fileName, bool const append);

protected:
    int myRank = 0;
    bool initialized = false;
    bool globalStructure = false;
    bool showew = false, resetew = false;
    int showewsum = 0, maxew = 0;
    double cflength = 1.0, cfenergy = 1.0;
    std::string emap;
    std::map<int, std::size_t> mapTypeToElement;
    std::map<std::size_t, int> mapElementToType;
    std::vector<bool> ignoreType;
    std::vector<std::size_t> indexMap;   // LAMMPS local index -> atom
    Structure structure;
};

}  // namespace nnp

#endif
