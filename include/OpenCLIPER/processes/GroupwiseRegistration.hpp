// Arguments of the motion-compensated sparsity operator (include/OpenCLIPER/processes/GroupwiseRegistration.hpp:95-100
// of the reference).  The reference's GroupwiseRegistration process, which estimates them, is not part of this
// library; callers fill the struct themselves:
//   xnData   real (float) Data of 2*F*H*W elements: deformed sample positions, 1-based, [0] column and [1] row for
//            every frame, element [d][f][r][c]
//   xData    uint32 or int32 Data of 2*H*W elements: the original mesh, 1-based, element [d][c][r] = (c + 1, r + 1)
//            (GroupwiseRegistration.cpp:150-164)
//   TData    the registration's transform coefficients (not read by motion compensation)
//   bound_box the 1-based region of interest (GroupwiseRegistration.cpp:485-524), four values
#pragma once
#include <OpenCLIPER/Data.hpp>

namespace OpenCLIPER {

struct ArgumentsMotionCompensation {
    std::shared_ptr<Data> xnData;
    std::shared_ptr<Data> TData;
    std::shared_ptr<Data> xData;
    std::vector<dimIndexType>* bound_box = nullptr;
};

}  // namespace OpenCLIPER
