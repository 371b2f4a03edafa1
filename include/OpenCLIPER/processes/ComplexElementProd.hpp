// out[f][c] = in[f][(c)] * (conj?) S[c] (include/OpenCLIPER/processes/ComplexElementProd.hpp:45-96).
#pragma once
#include <OpenCLIPER/SensitivityMapsData.hpp>
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class ComplexElementProd : public Process {
  public:
    enum ConjugateSensMap_t { notConjugate = 0, conjugate = 1 };
    struct LaunchParameters : Process::LaunchParameters {
        ConjugateSensMap_t conjugateSensMap;
        std::shared_ptr<SensitivityMapsData> sensitivityMapsData = nullptr;
        LaunchParameters(ConjugateSensMap_t c, const std::shared_ptr<SensitivityMapsData>& m)
            : conjugateSensMap(c), sensitivityMapsData(m) {}
    };
    void launch() override;
    OCL_PROCESS_BOILERPLATE(ComplexElementProd)
};

}  // namespace OpenCLIPER
