// Device-to-device copy (include/OpenCLIPER/processes/groupwiseRegistration/initialize/CopyDataGPU.hpp).
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class CopyDataGPU : public Process {
  public:
    void launch() override;
    OCL_PROCESS_BOILERPLATE(CopyDataGPU)
};

}  // namespace OpenCLIPER
