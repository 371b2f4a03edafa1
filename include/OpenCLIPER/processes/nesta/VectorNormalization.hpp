// out = in / max(mu, |in|) (include/OpenCLIPER/processes/nesta/VectorNormalization.hpp:46-61).
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class VectorNormalization : public Process {
  public:
    struct LaunchParameters : Process::LaunchParameters {
        float mu;
        LaunchParameters(float mu) : mu(mu) {}
    };
    void launch() override;
    OCL_PROCESS_BOILERPLATE(VectorNormalization)
};

}  // namespace OpenCLIPER
