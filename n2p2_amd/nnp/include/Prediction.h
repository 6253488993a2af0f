// nnp::Prediction (Prediction.h of the reference): the nnp-predict driver
// on top of Mode.
#ifndef NNP_PREDICTION_H
#define NNP_PREDICTION_H

#include "Mode.h"
#include "Structure.h"

#include <string>

namespace nnp
{

class Prediction : public Mode
{
public:
    Prediction();
    // Prediction::setup (Prediction.cpp:34-46).
    void setup();
    // Prediction::readStructureFromFile (Prediction.cpp:48-62).
    void readStructureFromFile(std::string const& fileName = "input.data");
    // Prediction::predict (Prediction.cpp:64-76).
    void predict();

    std::string fileNameSettings;
    std::string fileNameScaling;
    std::string formatWeightsFilesShort;
    std::string formatWeightsFilesCharge;
    Structure structure;
};

}  // namespace nnp

#endif
