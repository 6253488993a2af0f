// A data set through the nnp layer, as nnp-dataset drives Mode (setup,
// every structure of input.data, evaluateNNP), here with all structures in
// one batch (Mode::evaluateNNPBatch) and, for comparison, one at a time
// (Mode::evaluateNNP).  Prints per structure
//   "structure <k> <atoms> <E batch> <E single> <largest |F batch - F single|>"
// in the model's internal units (17 significant digits).
//
// Usage: dataset_batch <directory with input.nn, scaling.data,
//                       weights.%03zu.data and input.data>
#include "Mode.h"
#include "Structure.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <exception>
#include <fstream>
#include <string>
#include <vector>

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
        nnp::Mode mode;
        mode.log.writeToStdout = false;
        mode.initialize();
        mode.loadSettingsFile(dir + "input.nn");
        mode.setupGeneric();
        mode.setupSymmetryFunctionScaling(dir + "scaling.data");
        mode.setupNeuralNetworkWeights(dir);
        std::vector<nnp::Structure> batch;
        std::ifstream f(dir + "input.data");
        while (f.peek() != EOF)
        {
            nnp::Structure s;
            s.setElementMap(mode.elementMap);
            s.readFromFile(f);
            if (s.atoms.empty()) break;
            if (mode.useNormalization()) mode.convertToNormalizedUnits(s);
            batch.push_back(s);
            f >> std::ws;
        }
        std::vector<nnp::Structure> single = batch;
        mode.evaluateNNPBatch(batch, true);
        for (std::size_t k = 0; k < single.size(); ++k)
        {
            mode.evaluateNNP(single[k], true, true);
            double df = 0.0;
            for (std::size_t i = 0; i < single[k].atoms.size(); ++i)
                for (int c = 0; c < 3; ++c)
                    df = std::max(df, std::fabs(batch[k].atoms[i].f[c]
                                                - single[k].atoms[i].f[c]));
            std::printf("structure %zu %zu %.17g %.17g %.3g\n", k,
                        single[k].atoms.size(), batch[k].energy,
                        single[k].energy, df);
        }
    }
    catch (std::exception const& e)
    {
        std::fprintf(stderr, "%s", e.what());
        return 2;
    }
    return 0;
}
