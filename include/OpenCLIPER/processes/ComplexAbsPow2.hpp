// out = (|in|^2, 0) (include/OpenCLIPER/processes/ComplexAbsPow2.hpp).
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class ComplexAbsPow2 : public Process {
  public:
    void launch() override;
    OCL_PROCESS_BOILERPLATE(ComplexAbsPow2)
};

}  // namespace OpenCLIPER
