// Adjoint motion compensation (include/OpenCLIPER/processes/AdjointMotionCompensation.hpp,
// src/processes/AdjointMotionCompensation.cpp): output = input with the bound box cleared, then the transpose of
// MotionCompensation's resampling scattered into it.  Input and output must be distinct.
#pragma once
#include <OpenCLIPER/Process.hpp>
#include <OpenCLIPER/processes/GroupwiseRegistration.hpp>

namespace OpenCLIPER {

class AdjointMotionCompensation : public Process {
  public:
    struct LaunchParameters : Process::LaunchParameters {
        ArgumentsMotionCompensation* argsMC;
        explicit LaunchParameters(ArgumentsMotionCompensation* args) : argsMC(args) {}
    };
    void launch() override;
    OCL_PROCESS_BOILERPLATE(AdjointMotionCompensation)
};

}  // namespace OpenCLIPER
