// Fully-sampled multicoil recon x = sum_c conj(S_c) IFFT2(y_c) (include/OpenCLIPER/processes/examples/SimpleMRIRecon.hpp).
// Like the reference, the input k-space is transformed in place.
#pragma once
#include <OpenCLIPER/KData.hpp>
#include <OpenCLIPER/XData.hpp>
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class SimpleMRIRecon : public Process {
  public:
    void init() override;
    void launch() override;
    OCL_PROCESS_BOILERPLATE(SimpleMRIRecon)
};

}  // namespace OpenCLIPER
