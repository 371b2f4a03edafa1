// Zero k-space samples whose mask is 0, in place (include/OpenCLIPER/processes/ApplyMask.hpp:9-26).
#pragma once
#include <OpenCLIPER/SamplingMasksData.hpp>
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class ApplyMask : public Process {
  public:
    struct LaunchParameters : Process::LaunchParameters {
        std::shared_ptr<SamplingMasksData> samplingMasksData;
        explicit LaunchParameters(const std::shared_ptr<SamplingMasksData>& m) : samplingMasksData(m) {}
    };
    void launch() override;
    OCL_PROCESS_BOILERPLATE(ApplyMask)
};

}  // namespace OpenCLIPER
