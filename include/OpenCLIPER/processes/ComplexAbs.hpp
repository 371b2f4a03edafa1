// out = (|in|, 0) (include/OpenCLIPER/processes/ComplexAbs.hpp).
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class ComplexAbs : public Process {
  public:
    void launch() override;
    OCL_PROCESS_BOILERPLATE(ComplexAbs)
};

}  // namespace OpenCLIPER
