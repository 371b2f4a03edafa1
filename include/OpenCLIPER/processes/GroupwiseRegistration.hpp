// Groupwise B-spline registration (include/OpenCLIPER/processes/GroupwiseRegistration.hpp,
// src/processes/GroupwiseRegistration.cpp, groupwiseRegistration/{creabspline,optimizer}.cpp of the reference): the
// cardiac motion estimator that fills ArgumentsMotionCompensation for NestaUp with argsMC.  It runs on the device
// through ocl_gwr_* of the C ABI.
//
// ArgumentsMotionCompensation (GroupwiseRegistration.hpp:95-100 of the reference):
//   xnData   real (float) Data of 2*F*H*W elements: deformed sample positions, 1-based, [0] column and [1] row for
//            every frame, element [d][f][r][c]
//   xData    uint32 or int32 Data of 2*H*W elements: the original mesh, 1-based, element [d][c][r] = (c + 1, r + 1)
//            (GroupwiseRegistration.cpp:150-164)
//   TData    the registration's coefficients, real Data of 2*F*Cp*Cp elements, element [d][f][P][Q] (not read by
//            motion compensation; GroupwiseRegistration takes it back as the start of the next motion iteration)
//   bound_box the 1-based region of interest (GroupwiseRegistration.cpp:485-524), four values
// Callers may also fill the struct themselves (MotionCompensation and NestaUp accept any deformation).
#pragma once
#include <OpenCLIPER/Data.hpp>
#include <OpenCLIPER/Process.hpp>

#include <ocl_amd.h>

namespace OpenCLIPER {

struct ArgumentsMotionCompensation {
    std::shared_ptr<Data> xnData;
    std::shared_ptr<Data> TData;
    std::shared_ptr<Data> xData;
    std::vector<dimIndexType>* bound_box = nullptr;
};

/// GroupwiseRegistration (GroupwiseRegistration.hpp:60-260 of the reference).  init() computes the ROI, the control
/// grid and the B-spline tables (the reference's init() and creabspline(), which it runs on the first launch);
/// launch() normalises the input frames and runs the optimizer on the device, then points argsMC at the process's
/// xnData, TData, xData and bound_box (valid until the next launch or the process's destruction, as in the
/// reference).  Supported: square frames, radius > 0, E = 3, Dp[0] == Dp[1], 2 to 64 frames, and a geometry whose
/// control points and supports stay inside the frame; others throw CLError (OCL_E_UNSUPPORTED) with the reason.
class GroupwiseRegistration : public Process {
  public:
    struct InitParameters : ProcessCore::InitParameters {
        float W;
        bool flagW;
        unsigned radius;
        unsigned E;
        int* Dp;
        unsigned nmax;         ///< maximum number of optimizer iterations
        float et;              ///< transformation norm threshold (pixels)
        float eh;              ///< metric variation threshold (fraction of the initial metric)
        std::vector<realType> lambda;   ///< regularisation weights (1st spatial, 2nd spatial, 1st temporal, 2nd temporal)
        InitParameters(float W, bool flagW, unsigned radius, unsigned E, int* Dp, unsigned nmax, float et, float eh,
                       const std::vector<realType>& lambda)
            : W(W), flagW(flagW), radius(radius), E(E), Dp(Dp), nmax(nmax), et(et), eh(eh), lambda(lambda) {}
    };
    struct LaunchParameters : Process::LaunchParameters {
        ArgumentsMotionCompensation* argsMC;
        std::shared_ptr<Data> TData;   ///< starting coefficients (nullptr: zero)
        int currentMotionIter;
        LaunchParameters(ArgumentsMotionCompensation* args, const std::shared_ptr<Data>& TData, int iter)
            : argsMC(args), TData(TData), currentMotionIter(iter) {}
    };
    /// What the reference's optimizer prints per iteration (optimizer.cpp:350-356).
    struct Report {
        int iters = 0;
        float oldmax = 0;                 ///< max |x| over the input frames (the normalisation to 500)
        std::vector<float> H;             ///< cost before the loop and after every iteration
        std::vector<float> Tdif, Hdif, Wn;
    };
    ~GroupwiseRegistration() override;
    void init() override;
    void launch() override;
    const Report& getReport() const { return report_; }
    /// Control points per dimension of the coefficient grid
    int getControlPoints() const { return Cp_; }

  private:
    ocl_gwr* plan_ = nullptr;
    int F_ = 0, N_ = 0, Cp_ = 0;
    std::shared_ptr<Data> xData_, xnData_, TData_;
    std::vector<dimIndexType>* bound_box_ = nullptr;
    Report report_;
    OCL_PROCESS_BOILERPLATE(GroupwiseRegistration)
};

}  // namespace OpenCLIPER
