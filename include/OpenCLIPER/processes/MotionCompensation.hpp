// Motion compensation of an XData sequence (include/OpenCLIPER/processes/MotionCompensation.hpp,
// src/processes/MotionCompensation.cpp): output = input, then the bound-box pixels resampled bilinearly at the
// deformed positions of argsMC.  Input and output must be distinct.
#pragma once
#include <OpenCLIPER/Process.hpp>
#include <OpenCLIPER/processes/GroupwiseRegistration.hpp>

namespace OpenCLIPER {

class MotionCompensation : public Process {
  public:
    struct LaunchParameters : Process::LaunchParameters {
        ArgumentsMotionCompensation* argsMC;
        explicit LaunchParameters(ArgumentsMotionCompensation* args) : argsMC(args) {}
    };
    void launch() override;
    OCL_PROCESS_BOILERPLATE(MotionCompensation)
};

}  // namespace OpenCLIPER
