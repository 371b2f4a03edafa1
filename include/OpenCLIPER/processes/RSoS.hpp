// Root sum of squares over the coils (include/OpenCLIPER/processes/RSoS.hpp, kernel src/kernels/rss.cl):
// out_f = (sqrt(sum_c |in_{f,c}|^2), 0).  Input KData, output XData.
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class RSoS : public Process {
  public:
    void launch() override;
    OCL_PROCESS_BOILERPLATE(RSoS)
};

}  // namespace OpenCLIPER
