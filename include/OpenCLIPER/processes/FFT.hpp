// 2-D FFT of every NDArray (include/OpenCLIPER/processes/FFT.hpp:69-142, dim = -1 only): FORWARD unscaled,
// BACKWARD scaled by 1/(H*W), no fftshift; lengths must be products of powers of 2, 3, 5 and 7.
#pragma once
#include <OpenCLIPER/SamplingMasksData.hpp>
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class FFT : public Process {
  public:
    enum Direction { FORWARD = -1, BACKWARD = 1 };
    struct InitParameters : Process::InitParameters {
        int dim;
        std::shared_ptr<SamplingMasksData> samplingMask;
        explicit InitParameters(int d = -1, std::shared_ptr<SamplingMasksData> s = nullptr) : dim(d), samplingMask(s) {}
    };
    struct LaunchParameters : Process::LaunchParameters {
        Direction dir;
        explicit LaunchParameters(Direction d = Direction::FORWARD) : dir(d) {}
    };
    void init() override;
    void launch() override;
    OCL_PROCESS_BOILERPLATE(FFT)
};

}  // namespace OpenCLIPER
