// Temporal total variation (include/OpenCLIPER/processes/nesta/TemporalTV.hpp:45-62): FORWARD is the periodic
// backward difference D, ADJOINT is the reference's adjoint kernel, i.e. -D^T (temporalTV.cl:49-107).
#pragma once
#include <OpenCLIPER/Process.hpp>

namespace OpenCLIPER {

class TemporalTV : public Process {
  public:
    enum Direction { FORWARD = 1, ADJOINT = 0 };
    struct InitParameters : Process::InitParameters {
        Direction dir;
        InitParameters(Direction d) : dir(d) {}
    };
    void init() override;
    void launch() override;
    OCL_PROCESS_BOILERPLATE(TemporalTV)
};

}  // namespace OpenCLIPER
