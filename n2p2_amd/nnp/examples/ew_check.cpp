// Extrapolation warnings through the nnp layer, set up as the reference's
// InterfaceLammps::initialize sets up its Mode (InterfaceLammps.cpp:75-91)
// with the statistics flags of Mode::setupSymmetryFunctionStatistics:
// neighbour list and symmetry functions of the first structure of
// input.data, then the collected warning count and lines on stdout.  The
// immediate warning lines go to stderr (Element.cpp:505-518); with the stop
// flag the out_of_range message is printed after "STOPPED".
//
// Usage: ew_check <nnp directory> <input.data> <collect> <write> <stop>
#include "Mode.h"
#include "Structure.h"
#include "utility.h"

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <stdexcept>
#include <string>

namespace
{
// Mode::extrapolationWarningLines is protected; a small subclass reads it.
class EWMode : public nnp::Mode
{
public:
    std::vector<std::string> lines() const
    {
        return extrapolationWarningLines();
    }
};
}  // namespace

int main(int argc, char* argv[])
{
    if (argc != 6)
    {
        std::fprintf(stderr, "usage: %s <nnp directory> <input.data> "
                             "<collect> <write> <stop>\n", argv[0]);
        return 1;
    }
    std::string const dir = std::string(argv[1]) + "/";
    bool const collect = std::atoi(argv[3]) != 0;
    bool const write = std::atoi(argv[4]) != 0;
    bool const stop = std::atoi(argv[5]) != 0;
    EWMode m;
    m.log.writeToStdout = false;
    nnp::Structure s;
    try
    {
        m.initialize();
        m.loadSettingsFile(dir + "input.nn");
        m.setupGeneric();
        m.setupSymmetryFunctionScaling(dir + "scaling.data");
        m.setupSymmetryFunctionStatistics(false, collect, write, stop);
        m.setupNeuralNetworkWeights(
            std::map<std::string, std::string>{
                {"short", dir + "weights.%03zu.data"}});
        s.setElementMap(m.elementMap);
        s.readFromFile(argv[2]);
        if (m.useNormalization())
            s.toNormalizedUnits(m.getMeanEnergy(), m.getConvEnergy(),
                                m.getConvLength(), m.getConvCharge());
        s.calculateNeighborList(m.getMaxCutoffRadius());
        m.calculateSymmetryFunctionGroups(s, true);
    }
    catch (std::out_of_range const& e)
    {
        std::printf("STOPPED\n%s", e.what());
        return 0;
    }
    catch (std::exception const& e)
    {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
    std::printf("count %zu\n", m.getNumExtrapolationWarnings());
    for (std::string const& l : m.lines()) std::fputs(l.c_str(), stdout);
    return 0;
}
