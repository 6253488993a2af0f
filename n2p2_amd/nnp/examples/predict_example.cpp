// Prediction through the nnp layer, as the nnp-predict application drives
// it: setup from the NNP files of a directory, one structure, energy and
// forces.  Prints "energy E" and one "force i fx fy fz" line per atom with
// 17 significant digits (physical units, offsets included).
//
// Usage: predict_example <directory with input.nn, scaling.data,
//                         weights.%03zu.data and input.data>
#include "Prediction.h"
#include "utility.h"

#include <cstdio>
#include <exception>
#include <string>

int main(int argc, char* argv[])
{
    if (argc != 2)
    {
        std::fprintf(stderr, "usage: %s <nnp directory>\n", argv[0]);
        return 1;
    }
    std::string const dir = std::string(argv[1]) + "/";
    try
    {
        nnp::Prediction p;
        p.log.writeToStdout = false;
        p.fileNameSettings = dir + "input.nn";
        p.fileNameScaling = dir + "scaling.data";
        p.formatWeightsFilesShort = dir + "weights.%03zu.data";
        p.setup();
        p.readStructureFromFile(dir + "input.data");
        p.predict();
        nnp::Structure const& s = p.structure;
        std::printf("energy %.17g\n", s.energy);
        for (auto const& a : s.atoms)
            std::printf("force %zu %.17g %.17g %.17g\n", a.index, a.f[0],
                        a.f[1], a.f[2]);
    }
    catch (std::exception const& e)
    {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
    return 0;
}
