// nnp::Mode (Mode.h of the reference) for 2G-HDNNPs: the setup sequence and
// the compute methods, backed by the HIP engine through the C ABI of
// hdnnp.h.  The setup methods collect the settings, scaling and weight file
// names; the model is built when the weights are set up.  The compute
// methods upload the structure (positions only when atoms, elements and
// cell are unchanged), run the device kernels and write G, dE/dG, atomic
// energies, the energy and the forces back into the Structure, as the
// reference does; per-neighbour dG/dr stay on the device.
#ifndef NNP_MODE_H
#define NNP_MODE_H

#include "ElementMap.h"
#include "Log.h"
#include "Structure.h"

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

struct hdnnp_model;
struct hdnnp_ctx;

namespace nnp
{

class Mode
{
public:
    enum class NNPType
    {
        HDNNP_2G = 2,
        HDNNP_4G = 4,
        HDNNP_Q = 10
    };

    Mode();
    virtual ~Mode();
    Mode(Mode const&) = delete;
    Mode& operator=(Mode const&) = delete;

    void initialize();
    void loadSettingsFile(std::string const& fileName = "input.nn");
    void setupGeneric(std::string const& nnpDir = "",
                      bool skipNormalize = false,
                      bool initialHardness = false);
    void setupNormalization(bool standalone = true);
    virtual void setupElementMap();
    virtual void setupElements();
    void setupCutoff();
    virtual void setupSymmetryFunctions();
    void setupSymmetryFunctionScalingNone();
    virtual void setupSymmetryFunctionScaling(
        std::string const& fileName = "scaling.data");
    virtual void setupSymmetryFunctionGroups();
    void setupSymmetryFunctionMemory(bool verbose = false);
    void setupSymmetryFunctionStatistics(bool collectStatistics,
                                         bool collectExtrapolationWarnings,
                                         bool writeExtrapolationWarnings,
                                         bool stopOnExtrapolationWarnings);
    virtual void setupNeuralNetwork();
    virtual void setupNeuralNetworkWeights(
        std::map<std::string, std::string> fileNameFormats =
            std::map<std::string, std::string>());
    virtual void setupNeuralNetworkWeights(
        std::string directoryPrefix,
        std::map<std::string, std::string> fileNameFormats =
            std::map<std::string, std::string>());

    void calculateSymmetryFunctions(Structure& structure,
                                    bool const derivatives);
    void calculateSymmetryFunctionGroups(Structure& structure,
                                         bool const derivatives);
    void calculateAtomicNeuralNetworks(Structure& structure,
                                       bool const derivatives,
                                       std::string id = "");
    void calculateEnergy(Structure& structure) const;
    void calculateForces(Structure& structure) const;
    void evaluateNNP(Structure& structure, bool useForces = true,
                     bool useDEdG = true);
    // evaluateNNP over every structure of a data set, as nnp-dataset loops
    // it (nnp-dataset.cpp:206-211), evaluated as one batch on the device
    // (one launch per kernel for all structures).  Writes back what
    // evaluateNNP writes: per structure energy and pbc, per atom G,
    // energy, dEdG and (useForces) f.  No counterpart in the reference.
    void evaluateNNPBatch(std::vector<Structure>& structures,
                          bool useForces = true);
    void addEnergyOffset(Structure& structure, bool ref = true);
    void removeEnergyOffset(Structure& structure, bool ref = true);
    double getEnergyOffset(Structure const& structure) const;
    double getEnergyWithOffset(Structure const& structure,
                               bool ref = true) const;
    double normalized(std::string const& property, double value) const;
    double normalizedEnergy(Structure const& structure, bool ref = true) const;
    double physical(std::string const& property, double value) const;
    double physicalEnergy(Structure const& structure, bool ref = true) const;
    void convertToNormalizedUnits(Structure& structure) const;
    void convertToPhysicalUnits(Structure& structure) const;
    void logEwaldCutoffs();
    std::size_t getNumExtrapolationWarnings() const;
    void resetExtrapolationWarnings();
    NNPType getNnpType() const { return nnpType; }
    double getMeanEnergy() const { return meanEnergy; }
    double getConvEnergy() const { return convEnergy; }
    double getConvLength() const { return convLength; }
    double getConvCharge() const { return convCharge; }
    double getMaxCutoffRadius() const { return maxCutoffRadius; }
    std::size_t getNumElements() const { return numElements; }
    bool useNormalization() const { return normalize; }
    bool settingsKeywordExists(std::string const& keyword) const;
    std::string settingsGetValue(std::string const& keyword) const;

    // Engine controls with no counterpart in the reference.
    // HDNNP_PRECISION_FP64 (default), _MIXED or _MIXED16; also taken from
    // the environment variable N2P2_AMD_PRECISION (fp64, mixed, mixed16).
    void setPrecision(int precision);
    // HIP device of the engine context (default: N2P2_AMD_DEVICE or 0);
    // call before setupNeuralNetworkWeights.
    void setDevice(int device) { this->device = device; }
    // Copies the device neighbour list of the last evaluation into
    // atoms[i].neighbors (index, element, d, dr), numNeighbors and
    // neighborsUnique of the structure evaluated last.
    void fetchNeighborList(Structure& structure) const;
    // The engine context (for callers that keep data on the device).
    hdnnp_ctx* engineContext() const { return ctx; }

    ElementMap elementMap;
    Log log;

protected:
    // One SF's extrapolation warnings (SymFncStatistics::Container,
    // SymFncStatistics.h:28-147, the warning part).
    struct EWContainer
    {
        std::size_t countEW = 0;
        std::size_t type = 0;
        double Gmin = 0.0, Gmax = 0.0;
        std::string element;
        std::vector<std::size_t> indexStructureEW, indexAtomEW;
        std::vector<double> valueEW;
    };

    NNPType nnpType = NNPType::HDNNP_2G;
    bool normalize = false;
    bool checkExtrapolationWarnings = false;
    bool collectExtrapolationWarnings = false;
    bool writeExtrapolationWarnings = false;
    bool stopOnExtrapolationWarnings = false;
    std::size_t numElements = 0;
    double maxCutoffRadius = 0.0;
    double meanEnergy = 0.0;
    double convEnergy = 1.0;
    double convLength = 1.0;
    double convCharge = 1.0;
    std::vector<double> energyOffset;      // per element (atom_energy)
    std::vector<std::size_t> numSF;        // per element
    std::vector<std::vector<std::size_t>> sfType;   // per element and SF
    // Collected warnings per element, keyed by SF index (the reference's
    // Element::statistics.data).
    std::vector<std::map<std::size_t, EWContainer>> ewData;
    // Element::updateSymmetryFunctionStatistics (Element.cpp:462-541) for
    // every atom of the last SF evaluation on the engine, in the
    // reference's order: counts, collects, writes to stderr and stops as
    // the statistics flags say.  Returns the number of warnings.
    std::size_t processExtrapolationWarnings(Structure const& structure);
    // SymFncStatistics::getExtrapolationWarningLines of every element.
    std::vector<std::string> extrapolationWarningLines() const;

protected:
    // Creates the engine context on first use (after the weights are set).
    void ensureEngine() const;

private:
    void upload(Structure& structure) const;
    void buildModel();

    std::string settingsFile;
    std::string scalingFile;
    bool scalingNone = false;
    std::string weightsFormat;
    std::multimap<std::string, std::string> keywords;
    hdnnp_model* model = nullptr;
    hdnnp_ctx* ctx = nullptr;
    int device = -1;
    int precision = 0;
    // Structure the context holds: atom count, elements and cell.
    mutable std::size_t upAtoms = 0;
    mutable std::vector<int32_t> upElem;
    mutable double upBox[9] = {0};
    mutable bool upPeriodic = false;
    mutable bool upValid = false;
};

}  // namespace nnp

#endif
