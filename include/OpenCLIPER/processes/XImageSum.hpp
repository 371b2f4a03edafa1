// x_f = sum_c in[f][c] (include/OpenCLIPER/processes/XImageSum.hpp:44-57).
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class XImageSum : public Process {
  public:
    void launch() override;
    OCL_PROCESS_BOILERPLATE(XImageSum)
};

}  // namespace OpenCLIPER
