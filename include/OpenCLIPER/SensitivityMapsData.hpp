// Coil sensitivity maps, one NDArray per coil (include/OpenCLIPER/SensitivityMapsData.hpp).
#pragma once
#include <OpenCLIPER/Data.hpp>

namespace OpenCLIPER {

class SensitivityMapsData : public Data {
  public:
    SensitivityMapsData(const std::shared_ptr<CLapp>& pCLapp, std::vector<NDArray*>*& pData, numCoilsType nCoils);
    numCoilsType getNCoils() const { return getNumNDArrays(); }
};

}  // namespace OpenCLIPER
