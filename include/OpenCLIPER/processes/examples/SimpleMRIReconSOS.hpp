// Fully-sampled multicoil magnitude recon: inverse 2-D FFT of each coil image (in place on the input), then the root
// sum of squares over the coils (include/OpenCLIPER/processes/examples/SimpleMRIReconSOS.hpp).
#pragma once
#include <OpenCLIPER/KData.hpp>
#include <OpenCLIPER/XData.hpp>
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class SimpleMRIReconSOS : public Process {
  public:
    void init() override;
    void launch() override;
    OCL_PROCESS_BOILERPLATE(SimpleMRIReconSOS)
};

}  // namespace OpenCLIPER
